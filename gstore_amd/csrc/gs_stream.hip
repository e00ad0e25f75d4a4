// gs_stream.hip -- semi-external mode: the tile grid in pinned host memory,
// a resident prefix in HBM, and the rest streamed per sweep through two
// device segment buffers (the copy/sweep pipelining lives in gs::sweep).
//
// The reference streams 256 MB segments from an O_DIRECT file with libaio,
// double-buffered (seg1 is processed while seg2 loads, gstore.cpp:1770-1899,
// 3623-3797), and keeps the first cache_size bytes of each pass cached so the
// next iteration processes them first without IO (gstore.cpp:1491-1768).  Here
// the host side of that pipeline is pinned DRAM, the copy engine is
// hipMemcpyAsync on a side stream, and the cache is the resident prefix of the
// grid: chunks [0, res_chunks) never leave HBM.
//
// External mode (gs_open_external) keeps no host copy: the non-resident part
// of the grid is read from the .grid file in every streamed pass, as the
// reference's libaio driver reads it (io_driver, read_aio_random,
// gstore.cpp:3623-3797; buffer set-up read_aio_init :3044-3119).  Here a pool
// of IO threads issues O_DIRECT preads (buffered reads where the file system
// refuses O_DIRECT) of each segment, split in 4 KiB-aligned parts, into a ring
// of three pinned staging buffers; the copy stream moves a staged segment into
// a device segment buffer and the sweep follows, so reading segment i+2,
// copying i+1 and sweeping i overlap.  The resident HBM prefix is the
// cross-iteration cache (cache_to_cache / new_swap_parts, gstore.cpp:1491-1768):
// it is read from the file once, at open.
#include <errno.h>
#include <fcntl.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "gs_internal.h"

namespace gs {

struct DiskStage {
    static constexpr int NSTAGE = 3;         // staging buffers
    static constexpr int NIO = 4;            // IO threads; a segment is read in NIO parts
    static constexpr uint64_t ALIGN = 4096;  // O_DIRECT offset, length and address alignment
    std::string path;
    int fd_direct = -1, fd_buf = -1;
    uint64_t file_bytes = 0;
    uint64_t cap = 0;                        // bytes per staging buffer
    char *raw[NSTAGE] = {}, *buf[NSTAGE] = {};
    hipEvent_t free_ev[NSTAGE] = {};         // the copy out of the buffer is done
    bool free_rec[NSTAGE] = {};
    struct Slot {
        int64_t seg = -1;                    // segment of the current pass it holds or loads
        int parts_left = 0;
        int err = 0;
    } slot[NSTAGE];
    struct Job {
        int slot;
        uint64_t file_off, len, buf_off;
        bool wait_free;
    };
    std::mutex mu;
    std::condition_variable cv_job, cv_done;
    std::deque<Job> jobs;
    std::vector<std::thread> th;
    bool stop = false;
    uint64_t next_sub = 0;                   // next segment of the pass to submit
    uint64_t nseg = 0;                       // segments per pass
    uint64_t bytes = 0, direct_bytes = 0;    // read from the file (with O_DIRECT)
    double read_ms = 0, wait_ms = 0;         // summed over IO threads; the sweep's waits

    ~DiskStage() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_job.notify_all();
        for (auto &t : th) t.join();
        for (int i = 0; i < NSTAGE; ++i) {
            if (free_ev[i]) (void)hipEventDestroy(free_ev[i]);
            if (raw[i]) (void)hipHostFree(raw[i]);
        }
        if (fd_direct >= 0) close(fd_direct);
        if (fd_buf >= 0) close(fd_buf);
    }

    // How reads reach a staging buffer: 0 O_DIRECT straight into it; 1
    // O_DIRECT into an aligned bounce buffer of the reading thread, then a
    // copy (pinned memory that the GPU driver maps may refuse to be an
    // O_DIRECT target); 2 buffered reads (no O_DIRECT on this file system).
    // Only moves forward, on the first failure of a mode.
    std::atomic<int> mode{0};
    std::atomic<int> last_errno{0};
    static constexpr uint64_t BOUNCE = 16ull << 20;

    static char *bounce_buffer() {
        static thread_local std::unique_ptr<char, void (*)(void *)> b(nullptr, free);
        if (!b) {
            void *q = nullptr;
            if (posix_memalign(&q, ALIGN, BOUNCE) == 0) b.reset((char *)q);
        }
        return b.get();
    }

    // One pread loop of [file_off, file_off + want) into dst with fd; O_DIRECT
    // asks for whole aligned blocks (the tail past EOF reads short).  -1 on
    // an error (errno kept).
    static int64_t pread_all(int fd, bool direct, char *dst, uint64_t file_off, uint64_t want, uint64_t len) {
        uint64_t got = 0;
        while (got < want) {
            const uint64_t ask = direct ? ((len - got + ALIGN - 1) & ~(ALIGN - 1)) : want - got;
            const ssize_t r = pread(fd, dst + got, ask, (off_t)(file_off + got));
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) return -1;
            got += (uint64_t)r;
        }
        return (int64_t)std::min(got, want);
    }

    // Reads [file_off, file_off + len) into dst (file_off, dst and, except at
    // the file's end, len aligned).  Returns the bytes read (len, or less at
    // the end of the file), -1 on failure.
    int64_t read_range(char *dst, uint64_t file_off, uint64_t len, bool *used_direct) {
        const uint64_t want = file_off >= file_bytes ? 0 : std::min(len, file_bytes - file_off);
        for (;;) {
            const int m = fd_direct >= 0 ? mode.load() : 2;
            *used_direct = m < 2;
            int64_t r = -1;
            if (m == 0) {
                r = pread_all(fd_direct, true, dst, file_off, want, len);
            } else if (m == 1) {
                char *b = bounce_buffer();
                r = b ? 0 : -1;
                for (uint64_t o = 0; b && o < want; o += BOUNCE) {
                    const uint64_t piece = std::min(BOUNCE, want - o);
                    const int64_t k = pread_all(fd_direct, true, b, file_off + o, piece, std::min(BOUNCE, len - o));
                    if (k != (int64_t)piece) { r = -1; break; }
                    memcpy(dst + o, b, piece);
                    r += k;
                }
            } else {
                r = pread_all(fd_buf, false, dst, file_off, want, want);
            }
            if (r == (int64_t)want) return r;
            last_errno = errno;
            if (m == 2) return -1;
            int expect = m;
            mode.compare_exchange_strong(expect, m + 1);   // fall back and read the range again
        }
    }

    void io_main() {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_job.wait(lk, [&] { return stop || !jobs.empty(); });
                if (stop) return;
                j = jobs.front();
                jobs.pop_front();
            }
            int err = 0;
            if (j.wait_free && hipEventSynchronize(free_ev[j.slot]) != hipSuccess) err = 1;
            const auto t0 = std::chrono::steady_clock::now();
            bool direct = false;
            int64_t r = err ? -1 : read_range(buf[j.slot] + j.buf_off, j.file_off, j.len, &direct);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            const uint64_t expect = j.file_off >= file_bytes ? 0 : std::min(j.len, file_bytes - j.file_off);
            if (r < 0 || (uint64_t)r != expect) err = 1;
            {
                std::lock_guard<std::mutex> lk(mu);
                read_ms += ms;
                if (r > 0) {
                    bytes += (uint64_t)r;
                    if (direct) direct_bytes += (uint64_t)r;
                }
                slot[j.slot].err |= err;
                --slot[j.slot].parts_left;
            }
            cv_done.notify_all();
        }
    }

    // Queue the reads of segment i (edges [e0, e1)) into its slot.  Lock held.
    void submit(uint64_t i, uint64_t e0, uint64_t e1) {
        const int s = (int)(i % NSTAGE);
        const uint64_t off = e0 * 4, len = (e1 - e0) * 4;
        const uint64_t part = (((len + NIO - 1) / NIO) + ALIGN - 1) & ~(ALIGN - 1);
        slot[s].seg = (int64_t)i;
        slot[s].err = 0;
        slot[s].parts_left = 0;
        for (uint64_t b = 0; b < len; b += part) {
            jobs.push_back(Job{s, off + b, std::min(part, len - b), b, free_rec[s]});
            ++slot[s].parts_left;
        }
        cv_job.notify_all();
    }
};

// Segment i of a pass covers chunks res + i*seg .. : its edge range.
static void seg_range(const gs_graph *g, const StreamState *ss, uint64_t i, uint64_t *e0, uint64_t *e1) {
    const uint64_t c0 = ss->res_chunks + i * ss->seg_chunks;
    const uint64_t cn = std::min(ss->seg_chunks, g->nchunk - c0);
    *e0 = c0 * g->chunk;
    *e1 = std::min(g->E, (c0 + cn) * g->chunk);
}

int disk_pass_begin(StreamState *ss) {
    DiskStage *d = ss->disk;
    std::unique_lock<std::mutex> lk(d->mu);
    // reads queued by a pass that stopped early are dropped (and no longer
    // awaited); those already running must land before the slots are reused
    for (const DiskStage::Job &j : d->jobs) --d->slot[j.slot].parts_left;
    d->jobs.clear();
    d->cv_done.wait(lk, [&] {
        for (auto &s : d->slot)
            if (s.parts_left > 0) return false;
        return true;
    });
    for (auto &s : d->slot) s.seg = -1;
    d->next_sub = 0;
    return GS_OK;
}

int disk_copy(const gs_graph *g, StreamState *ss, uint64_t i, uint64_t e0, uint64_t e1, uint32_t *dst) {
    DiskStage *d = ss->disk;
    const int s = (int)(i % DiskStage::NSTAGE);
    {
        std::unique_lock<std::mutex> lk(d->mu);
        // keep the next segments' reads in flight: up to NSTAGE - 1 ahead.  A
        // slot is reused by segment i + NSTAGE only after segment i's copy
        // was recorded below, so its free event is the one to wait for.
        while (d->next_sub < d->nseg && d->next_sub < i + DiskStage::NSTAGE) {
            uint64_t a, b;
            seg_range(g, ss, d->next_sub, &a, &b);
            d->submit(d->next_sub, a, b);
            ++d->next_sub;
        }
        const auto t0 = std::chrono::steady_clock::now();
        d->cv_done.wait(lk, [&] { return d->slot[s].seg == (int64_t)i && d->slot[s].parts_left == 0; });
        d->wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (d->slot[s].err) {
            set_error("reading %s at byte %llu failed: %s", d->path.c_str(), (unsigned long long)(e0 * 4),
                      strerror(d->last_errno.load()));
            return GS_ERR_IO;
        }
    }
    GS_HIP(hipMemcpyAsync(dst, d->buf[s], (e1 - e0) * 4, hipMemcpyHostToDevice, ss->copy));
    GS_HIP(hipEventRecord(d->free_ev[s], ss->copy));
    d->free_rec[s] = true;
    return GS_OK;
}

// External-mode state for graph g (geometry and nchunk known): open the file,
// upload the resident prefix, allocate the segments and the staging ring.
int disk_setup(gs_graph *g, const std::string &path, uint64_t budget, uint64_t seg_bytes) {
    std::unique_ptr<StreamState> ss(new StreamState());
    std::unique_ptr<DiskStage> d(new DiskStage());
    const uint64_t chunk_bytes = (uint64_t)g->chunk * 4;
    if (!seg_bytes) seg_bytes = std::min<uint64_t>(256ull << 20, budget / 8);   // the reference's 256 MB segments
    ss->seg_chunks = std::max<uint64_t>(1, seg_bytes / chunk_bytes);
    const uint64_t two = 2 * ss->seg_chunks * chunk_bytes;
    if (budget < two + chunk_bytes) {
        set_error("device budget %llu bytes is below two segments (%llu bytes) plus one chunk",
                  (unsigned long long)budget, (unsigned long long)two);
        return GS_ERR_ARG;
    }
    ss->res_chunks = std::min<uint64_t>(g->nchunk, (budget - two) / chunk_bytes);
    d->path = path;
    d->file_bytes = g->E * 4;
    d->fd_buf = open(path.c_str(), O_RDONLY);
    if (d->fd_buf < 0) { set_error("cannot open %s", path.c_str()); return GS_ERR_IO; }
    d->fd_direct = open(path.c_str(), O_RDONLY | O_DIRECT);   // -1 where the file system has no O_DIRECT
    d->cap = ss->seg_chunks * chunk_bytes + DiskStage::ALIGN;
    for (int i = 0; i < DiskStage::NSTAGE; ++i) {
        GS_HIP(hipHostMalloc((void **)&d->raw[i], d->cap + DiskStage::ALIGN, hipHostMallocDefault));
        d->buf[i] = (char *)(((uintptr_t)d->raw[i] + DiskStage::ALIGN - 1) & ~(uintptr_t)(DiskStage::ALIGN - 1));
        GS_HIP(hipEventCreateWithFlags(&d->free_ev[i], hipEventDisableTiming));
    }
    const uint64_t res_edges = std::min<uint64_t>(g->E, ss->res_chunks * g->chunk);
    d->nseg = ss->res_chunks < g->nchunk ? (g->nchunk - ss->res_chunks + ss->seg_chunks - 1) / ss->seg_chunks : 0;
    GS_TRY(g->grid.alloc(res_edges));
    // the resident prefix, read once through staging buffer 0
    for (uint64_t e = 0; e < res_edges;) {
        const uint64_t k = std::min<uint64_t>(res_edges - e, ss->seg_chunks * g->chunk);
        bool direct = false;
        if (d->read_range(d->buf[0], e * 4, k * 4, &direct) != (int64_t)(k * 4)) {
            set_error("reading %s at byte %llu failed: %s", path.c_str(), (unsigned long long)(e * 4),
                      strerror(d->last_errno.load()));
            return GS_ERR_IO;
        }
        GS_HIP(hipMemcpy(g->grid.ptr + e, d->buf[0], k * 4, hipMemcpyHostToDevice));
        e += k;
    }
    GS_TRY(ss->seg[0].alloc(ss->seg_chunks * g->chunk));
    GS_TRY(ss->seg[1].alloc(ss->seg_chunks * g->chunk));
    GS_HIP(hipStreamCreateWithFlags(&ss->copy, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        GS_HIP(hipEventCreateWithFlags(&ss->ev_copied[i], hipEventDisableTiming));
        GS_HIP(hipEventCreateWithFlags(&ss->ev_done[i], hipEventDisableTiming));
        GS_HIP(hipEventRecord(ss->ev_done[i], ss->copy));
    }
    GS_HIP(hipStreamSynchronize(ss->copy));
    DiskStage *dp = d.get();
    for (int i = 0; i < DiskStage::NIO; ++i) d->th.emplace_back([dp] { dp->io_main(); });
    ss->disk = d.release();
    g->ss = ss.release();
    return GS_OK;
}

int disk_read_all(const StreamState *ss, uint32_t *dst) {
    // a caller's buffer: buffered reads (O_DIRECT would need it aligned)
    const uint64_t n = ss->disk->file_bytes;
    if (DiskStage::pread_all(ss->disk->fd_buf, false, (char *)dst, 0, n, n) != (int64_t)n) {
        set_error("reading %s failed", ss->disk->path.c_str());
        return GS_ERR_IO;
    }
    return GS_OK;
}

StreamState::~StreamState() {
    if (copy) (void)hipStreamSynchronize(copy);
    delete disk;   // joins its IO threads (no copy reads its staging buffers any more)
    disk = nullptr;
    for (int i = 0; i < MAXBUF; ++i) {
        if (ev_copied[i]) (void)hipEventDestroy(ev_copied[i]);
        if (ev_done[i]) (void)hipEventDestroy(ev_done[i]);
    }
    if (copy) (void)hipStreamDestroy(copy);
    if (h_grid) (void)hipHostFree(h_grid);
}

// Build the streaming state for graph g whose full grid is in `h_grid`
// (pinned, ownership taken).  Uploads the resident prefix.
int stream_setup(gs_graph *g, uint32_t *h_grid, uint64_t budget, uint64_t seg_bytes) {
    std::unique_ptr<StreamState> ss(new StreamState());  // takes h_grid only on success
    const uint64_t chunk_bytes = (uint64_t)g->chunk * 4;
    if (!seg_bytes) seg_bytes = std::min<uint64_t>(512ull << 20, budget / 8);  // default: 512 MiB or budget/8
    ss->seg_chunks = std::max<uint64_t>(1, seg_bytes / chunk_bytes);
    const uint64_t segb = ss->seg_chunks * chunk_bytes;
    if (budget < 2 * segb + chunk_bytes) {
        set_error("device budget %llu bytes is below two segments (%llu bytes) plus one chunk",
                  (unsigned long long)budget, (unsigned long long)(2 * segb));
        return GS_ERR_ARG;
    }
    // a third segment buffer when the budget holds it and some of the grid is
    // resident beside it: it absorbs segments whose sweep outlasts a copy
    ss->nbuf = budget >= 4 * segb + chunk_bytes ? 3 : 2;
    ss->res_chunks = std::min<uint64_t>(g->nchunk, (budget - ss->nbuf * segb) / chunk_bytes);
    uint64_t res_edges = std::min<uint64_t>(g->E, ss->res_chunks * g->chunk);
    GS_TRY(g->grid.alloc(res_edges));
    if (res_edges) GS_HIP(hipMemcpy(g->grid.ptr, h_grid, res_edges * 4, hipMemcpyHostToDevice));
    for (int i = 0; i < ss->nbuf; ++i) GS_TRY(ss->seg[i].alloc(ss->seg_chunks * g->chunk));
    GS_HIP(hipStreamCreateWithFlags(&ss->copy, hipStreamNonBlocking));
    for (int i = 0; i < ss->nbuf; ++i) {
        GS_HIP(hipEventCreateWithFlags(&ss->ev_copied[i], hipEventDisableTiming));
        GS_HIP(hipEventCreateWithFlags(&ss->ev_done[i], hipEventDisableTiming));
        GS_HIP(hipEventRecord(ss->ev_done[i], ss->copy));  // first waits complete at once
    }
    GS_HIP(hipStreamSynchronize(ss->copy));
    void *dev = nullptr;
    GS_HIP(hipHostGetDevicePointer(&dev, h_grid, 0));
    ss->h_grid_dev = static_cast<const uint32_t *>(dev);
    ss->h_grid = h_grid;
    g->ss = ss.release();
    return GS_OK;
}

}  // namespace gs

using namespace gs;

extern "C" int gs_get_io_stats(gs_graph *g, gs_io_stats *out) {
    if (!g || !out) { set_error("null argument"); return GS_ERR_ARG; }
    *out = gs_io_stats{};
    if (!g->ss || !g->ss->disk) return GS_OK;
    DiskStage *d = g->ss->disk;
    std::lock_guard<std::mutex> lk(d->mu);
    out->disk_bytes = d->bytes;
    out->direct_bytes = d->direct_bytes;
    out->read_ms = d->read_ms;
    out->wait_ms = d->wait_ms;
    out->io_threads = DiskStage::NIO;
    out->staging_buffers = DiskStage::NSTAGE;
    out->external = 1;
    out->read_mode = d->fd_direct >= 0 ? d->mode.load() : 2;
    return GS_OK;
}

extern "C" int gs_set_device_budget(gs_graph *g, uint64_t budget, uint64_t seg_bytes) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (g->ss) { set_error("graph is already in semi-external mode"); return GS_ERR_ARG; }
    if (budget == 0 || g->E * 4 <= budget) return GS_OK;  // fits: stays resident
    GS_HIP(hipDeviceSynchronize());
    uint32_t *h = nullptr;
    GS_HIP(hipHostMalloc((void **)&h, std::max<uint64_t>(g->E, 1) * 4, hipHostMallocDefault));
    if (hipMemcpy(h, g->grid.ptr, g->E * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipHostFree(h);
        set_error("copying the grid to host memory failed");
        return GS_ERR_DEVICE;
    }
    g->grid.release();
    int st = stream_setup(g, h, budget, seg_bytes);
    if (st != GS_OK && !g->ss) {
        // restore the resident grid
        if (g->grid.alloc(g->E) == GS_OK) (void)hipMemcpy(g->grid.ptr, h, g->E * 4, hipMemcpyHostToDevice);
        (void)hipHostFree(h);
    }
    return st;
}
