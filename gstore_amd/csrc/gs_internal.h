// gs_internal.h -- shared state, tile geometry and the edge-chunk sweep used
// by every kernel of the engine.
//
// HBM layout of a resident graph (N vertices, E stored edges, T tile slots):
//   grid     u32[E]        one word per stored edge: v0_local | v1_local << 16,
//                          tiles back to back in .start order (as on disk)
//   tz_beg   u64[nz + 1]   first edge of every non-empty tile, tz_beg[nz] = E
//   tz_rc    u32[nz]       tile row << 16 | tile column (global tile coords)
//   chunk_tz u32[nchunk+1] non-empty tile holding edge c*chunk; last = nz-1
//   degree   i32[N]        the .degree file
// Edges are swept in fixed chunks of `chunk` edges (1024..4096, a tuning
// knob), one workgroup per chunk;
// the tiles a chunk touches are staged in LDS so each lane finds its tile by
// a short search and decodes the 16-bit local ids in registers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/gstore_amd.h"

namespace gs {

constexpr int CHUNK = 4096;       // largest (and default) edges per workgroup sweep
constexpr int PULL_CHUNK = 4096;  // neighbour-list entries per workgroup of the PageRank pull
constexpr int SWEEP_WG = 256;     // threads per sweep workgroup (4 waves)
constexpr int NUM_XCD = 8;        // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr uint32_t ROW_WORDS = 65536 / 32;   // bitmap words per tile row

// Source-blocked copy of the neighbour-list index (PageRank), gs_pb.hip.
constexpr int PB_SHIFT = 15;                 // vertices per block: 2^15, FP32 values fill 128 KiB of LDS
constexpr uint32_t PB_B = 1u << PB_SHIFT;
constexpr uint32_t MS_HEAVY_LOG = 7;         // batched bottom-up: a wave per vertex of degree >= 2^7
constexpr uint32_t MS_HEAVY_DEG = 1u << MS_HEAVY_LOG;
constexpr uint32_t PB_KMAX = 8192;           // blocks at most (vertices with edges <= 2^28)
constexpr uint32_t PB_RUNS = 2048;           // runs staged per gather workgroup at most
constexpr unsigned long long PB_COMPACT = 1ull << 63;   // run stream base: the run's block is compacted
constexpr uint32_t PB_WIN = 16;              // entries per window: every run starts on a window and fills whole ones
constexpr uint32_t PB_UNIT_WIN = 4096;       // gather: windows per unit at most (2^16 entries)
struct PbSlice { uint64_t e0, e1; uint32_t q, pad; };               // build: list entries of block q
struct PbUnit1 { uint64_t w0, w1; uint32_t r0, r1, p, pad; };       // gather: source-major windows of block p
struct PbUnit2 { uint64_t e0, e1; uint32_t q, slot, dense, pad; };  // accum: slot ~0u = the block's only unit
struct PbRed { uint32_t q, slot0, nslot, pad; };                    // split block: its partial slots

void set_error(const char *fmt, ...);
const char *error_text();  // message of the last failure on this thread

#define GS_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            gs::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,              \
                          hipGetErrorString(e_));                                 \
            return GS_ERR_DEVICE;                                                 \
        }                                                                         \
    } while (0)

#define GS_TRY(expr)                 \
    do {                             \
        int s_ = (expr);             \
        if (s_ != GS_OK) return s_;  \
    } while (0)

// ---- tile geometry (gstore.h:109-141 undirected, :176-198 directed) ------
__host__ __device__ inline uint64_t tri_index(uint64_t i, uint64_t j, uint64_t n) {
    return ((i * (2 * n - i - 1)) >> 1) + j;
}
__host__ __device__ inline uint64_t groups_for(uint64_t n) {
    return (n >> 24) + ((n & 0xFFFFFFull) != 0);
}
__host__ __device__ inline uint64_t total_slots(uint64_t p_s, bool directed) {
    return directed ? p_s * p_s : (p_s * (p_s + 1)) >> 1;
}
// .start index of tile (r, c); undirected requires r <= c.
__host__ __device__ inline uint64_t tile_slot(uint64_t r, uint64_t c, uint64_t p, bool directed) {
    uint64_t bi = r >> 8, si = r & 255, bj = c >> 8, sj = c & 255;
    if (directed) return ((bi * p + bj) << 16) + si * 256 + sj;
    if (bi == bj) return ((tri_index(bi, bi, p) - bi) << 16) + ((bi * 257) << 7) + tri_index(si, sj, 256);
    return ((tri_index(bi, bj, p) - bi - 1) << 16) + (((bi + 1) * 257) << 7) + si * 256 + sj;
}

// ---- device buffers --------------------------------------------------------
template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (hipMalloc(&ptr, count * sizeof(T)) != hipSuccess) {
            ptr = nullptr;
            set_error("hipMalloc of %zu bytes failed", count * sizeof(T));
            return GS_ERR_NOMEM;
        }
        n = count;
        return GS_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Two events destroyed with their scope, so an early error return leaks none.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Per-graph timing of kernel classes (HIP events on the graph's stream).
struct KernelTimer {
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    double ms[GS_K_COUNT] = {};
    uint64_t calls[GS_K_COUNT] = {};
    void begin(int k, hipStream_t s, hipEvent_t *out);
    void end(int k, hipStream_t s, hipEvent_t b);
    int collect();  // after a stream sync
    void reset();
};

}  // namespace gs

// State of one in-flight algorithm call: its stream, work arrays and stats.
// A graph owns one default workspace; the serving engine makes more so calls
// on the same graph can run concurrently.
struct gs_ws {
    hipStream_t stream = nullptr;
    gs::DevBuf<float> pr_prior, pr_acc, pr_fac, ppr_rank;
    gs::DevBuf<double> ppr_acc;                                // personalised PageRank accumulates in FP64
    gs::DevBuf<uint8_t> depth;
    gs::DevBuf<uint32_t> bm_vis, bm_front, bm_next, rowact;  // bitmaps, words
    gs::DevBuf<uint32_t> wcc_parent;
    gs::DevBuf<uint32_t> wcc_label;                           // labels by original id (index method)
    gs::DevBuf<uint32_t> wcc_cmin;                            // smallest original id per root (index method)
    gs::DevBuf<uint32_t> wcc_samp;                            // roots of the sampled vertices (index method)
    gs::DevBuf<int32_t> kc_deg;
    gs::DevBuf<uint64_t> ms_vis, ms_front, ms_next, ms_roots;  // batched BFS lane masks
    gs::DevBuf<uint8_t> ms_depth;                             // lane-major depth, lanes x N
    gs::DevBuf<uint8_t> ms_depth_vm;                          // vertex-major depth rows (index method), N x 64
    gs::DevBuf<uint32_t> queue;                               // frontier vertices (direction-optimizing BFS)
    gs::DevBuf<int32_t> qdeg;                                 // their degrees
    gs::DevBuf<uint64_t> qoff;                                // exclusive prefix of qdeg (top-down levels)
    gs::DevBuf<unsigned long long> qsum;                      // per-1024 block sums of qdeg
    gs::DevBuf<uint8_t> depth_r;                              // depth by relabelled id
    gs::DevBuf<unsigned long long> ms_counts;
    gs::DevBuf<unsigned long long> ms_ctl;                    // device-driven batched BFS: control words, level counts
    hipGraphExec_t ms_graph = nullptr;                        // its captured level sequence
    std::vector<const void *> ms_graph_key;                   // what the capture baked in
    gs::DevBuf<unsigned long long> do_ctl;                    // direction-optimizing BFS: control words, level records
    hipGraphExec_t do_graph = nullptr;                        // its captured level sequence
    std::vector<const void *> do_graph_key;
    gs::DevBuf<uint32_t> bin_cnt, bin_off, bin_cur, bin_uoff;  // binned scatter: per bin counts, offsets, cursors, units
    gs::DevBuf<int64_t> bin_pc, bin_pr;                       // per tile of a piece: column / row bin position - first edge
    gs::DevBuf<uint16_t> bin_o;                               // per contribution: destination offset in its bin
    gs::DevBuf<float> bin_v;                                  // per contribution: the source's prior
    gs::DevBuf<unsigned long long> counters;
    unsigned long long *h_counters = nullptr;  // pinned mirror of counters
    gs::KernelTimer timer;
    gs_stats stats{};
    std::vector<gs_level_stat> levels;                        // per-level record of the last BFS
    ~gs_ws();
    size_t device_bytes() const;
};

namespace gs {
// Semi-external mode: the grid lives in pinned host memory; the first
// res_chunks chunks stay resident in HBM and the rest pass through nbuf (2
// or 3) device segment buffers, copied on a side stream while the earlier
// segments are being swept.
struct DiskStage;  // external mode: the grid read from its file every sweep (gs_stream.hip)

struct StreamState {
    DiskStage *disk = nullptr;   // set: no host copy; the non-resident chunks come from the .grid file
    uint32_t *h_grid = nullptr;  // pinned host copy of the whole grid
    const uint32_t *h_grid_dev = nullptr;  // the same memory as the device addresses it (direct sweeps)
    uint64_t res_chunks = 0;     // chunks [0, res_chunks) resident in gs_graph::grid
    uint64_t seg_chunks = 0;     // chunks per segment buffer
    static constexpr int MAXBUF = 3;
    int nbuf = 2;                // segment buffers in use
    DevBuf<uint32_t> seg[MAXBUF];
    hipStream_t copy = nullptr;
    hipEvent_t ev_copied[MAXBUF] = {}, ev_done[MAXBUF] = {};
    ~StreamState();
};
}  // namespace gs

struct gs_graph {
    int device = 0;
    hipStream_t stream = nullptr;            // loading / conversion
    bool directed = false;
    uint64_t n = 0, p = 0, p_s = 0, T = 0, E = 0, nz = 0, nchunk = 0;
    uint32_t chunk = gs::CHUNK;              // edges per sweep chunk (workgroup)
    uint64_t rows = 0;                       // tile rows holding vertices: ceil(N/65536)
    gs::DevBuf<uint32_t> grid;
    gs::DevBuf<uint64_t> tz_beg;
    gs::DevBuf<uint32_t> tz_rc;
    gs::DevBuf<uint32_t> chunk_tz;
    std::vector<uint32_t> h_chunk_tz;        // host copy of chunk_tz (binned scatter), for chunk size h_chunk_tz_of
    uint32_t h_chunk_tz_of = 0;
    gs::DevBuf<int32_t> degree;
    gs::DevBuf<uint32_t> in_touch;           // directed grids: bitmap of vertices with an in-edge
    gs::DevBuf<unsigned long long> counters;
    unsigned long long *h_counters = nullptr;
    gs_ws *ws = nullptr;                     // default workspace of the plain entry points
    gs::StreamState *ss = nullptr;           // non-null in semi-external mode
    int xcd_order = 1;                       // sweep order, see gs_set_sweep_order
    int bfs_mode = 0;                        // 0: tile sweep, 1: direction-optimizing over adj_*
    bool heavy_on = true;                    // PageRank: accumulate the heaviest vertices in LDS
    bool heavy_ready = false;
    uint32_t heavy_dthr = 0xFFFFFFFFu;       // degree threshold of the heavy set
    int nheavy = 0;
    gs::DevBuf<uint32_t> heavy_ids;          // sorted heavy vertex ids
    int persistent_wgs = 1024;               // workgroups of the persistent PageRank scatter
    int do_alpha = 20, do_beta = 24;         // top-down <-> bottom-up switch thresholds
    // neighbour-list index (built on demand), in relabelled ids (descending degree)
    gs::DevBuf<uint32_t> new_id;             // relabelled id of each vertex, N
    gs::DevBuf<uint32_t> old_id;             // original id of each relabelled id, N
    gs::DevBuf<int32_t> deg_r;               // degree by relabelled id, N
    gs::DevBuf<uint64_t> adj_off;            // list offsets by relabelled id, N + 1
    gs::DevBuf<uint32_t> adj_nbr;            // neighbour ids (relabelled), 2E
    gs::DevBuf<uint32_t> adj_first;          // first list entry of each vertex with edges, nzn
    uint64_t nzn = 0, npull = 0;             // vertices with edges (ids 0 .. nzn-1); pull chunks
    uint64_t ms_nhub = 0;                    // ids 0 .. ms_nhub-1 have degree >= MS_HEAVY_DEG
    double index_ms = 0;                     // device time of the index build
    gs::DevBuf<uint32_t> chunkz;             // vertex whose list holds entry c*PULL_CHUNK, npull + 1
    bool adj_ready = false;                  // every index array above built and complete
    int fail_alloc_at = 0;                   // test hook: the k-th allocation of the next index build fails
    int pr_method = 2;                       // 0 tile scatter, 1 pull over the lists, 2 blocked copy (binned scatter
                                             // on directed and semi-external grids for 1 and 2)
    int wcc_method = 1;                      // 0 tile hooking passes, 1 sampled linking over the lists
    int kcore_method = 1;                    // 0 tile scans, 1 peeled lists pushed over the index
    int batch_method = 2;                    // batched BFS: 0 tile sweep, 1 over the index, 2 the same in a hipGraph
    // source-blocked copy of the lists (built on demand), gs_pb.hip
    bool pb_ready = false;
    uint32_t pb_k = 0;                       // blocks; 0 = not available (PageRank pulls)
    uint64_t pb_nu1 = 0, pb_nu2 = 0, pb_nred = 0, pb_nruns = 0;
    double pb_ms = 0;                        // build time
    uint64_t pb_unit1 = 1u << 16;            // gather: entries per workgroup at most
    uint32_t pb_runs = gs::PB_RUNS;          // gather: runs per workgroup at most
    uint64_t pb_unit2 = 1u << 21;            // accum: entries per workgroup at most
    uint32_t pb_dense = 1u << 30;            // accum: stream values per vertex from which a block uses LDS float atomics
    uint64_t pb_entries = 0;                 // 2E plus the padding that fills each run's last window
    uint64_t pb_stream = 0;                  // accumulation stream length (with block alignment)
    uint32_t pb_compact = 800;               // a block is compacted when its segments <= this / 1000 of its entries
    uint32_t pb_grab = 8;                    // gather: units per grab of a resident workgroup (0: a workgroup per unit)
    gs::DevBuf<unsigned long long> pb_ctr;   // gather: next grab
    uint32_t num_cus = 0;                    // compute units of the device (queried once)
    gs::DevBuf<uint16_t> pb_src;             // pb_entries, q-major: offset of u in its block (compacted
                                             // blocks: bit 15 marks a segment's last entry)
    gs::DevBuf<uint32_t> pb_wcoff;           // per q-major window: its first segment, relative to its run
    gs::DevBuf<uint16_t> pb_dst;             // pb_stream: destination offset per stream value
    gs::DevBuf<float> pb_val;                // pb_stream: the stream values, rewritten every iteration
    gs::DevBuf<uint64_t> pb_run_pw, pb_run_qw;  // non-empty runs, source-major: window prefix, q-major window
    gs::DevBuf<uint64_t> pb_run_st;          // their stream base (PB_COMPACT: the block is compacted)
    gs::DevBuf<gs::PbUnit1> pb_u1;
    gs::DevBuf<gs::PbUnit2> pb_u2;
    gs::DevBuf<gs::PbRed> pb_red;
    gs::DevBuf<float> pb_part;               // partial sums of split blocks, PB_B per slot
    gs::DevBuf<double> pb_part64;            // the same in FP64 (personalised PageRank), allocated on use
    std::mutex pb_mu;                        // one blocked PageRank at a time (pb_val, pb_part are shared)
    std::mutex build_mu;                     // lazy builds (index, heavy set) from concurrent workspaces
    ~gs_graph();
};

namespace gs {

int graph_create(gs_graph **out);
int ws_create(gs_ws **out);
int graph_build_meta(gs_graph *g, const uint64_t *h_start);  // tz/chunk lists, from host .start
int build_chunk_map(gs_graph *g, const std::vector<uint64_t> &beg);  // chunk_tz for g->chunk
int ensure_bitmaps(const gs_graph *g, gs_ws *w);
int ensure_rowact(const gs_graph *g, gs_ws *w);
int stream_setup(gs_graph *g, uint32_t *h_grid, uint64_t budget, uint64_t seg_bytes);
// External mode (gs_stream.hip): before a streamed pass, wait for reads left
// from an earlier pass; per segment i of the pass (edges [e0, e1)), wait for
// its file read, then copy it to dst on the copy stream, keeping the next
// segments' reads in flight.
int disk_pass_begin(StreamState *ss);
int disk_copy(const gs_graph *g, StreamState *ss, uint64_t i, uint64_t e0, uint64_t e1, uint32_t *dst);
int disk_setup(gs_graph *g, const std::string &path, uint64_t budget, uint64_t seg_bytes);
int disk_read_all(const StreamState *ss, uint32_t *dst);   // the whole grid from its file
uint64_t bitmap_words(const gs_graph *g);  // rows * ROW_WORDS
int sync_and_collect(gs_ws *w);

// Algorithm bodies on an explicit workspace (the extern "C" entry points use g->ws).
int run_pagerank(const gs_graph *g, gs_ws *w, int iters, int mode, float *rank_out);
int run_ppr(const gs_graph *g, gs_ws *w, uint64_t source, int iters, float *out);
int run_bfs(const gs_graph *g, gs_ws *w, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
            int max_levels, int *levels);
int run_bfs_do(const gs_graph *g, gs_ws *w, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
               int max_levels, int *levels);
uint64_t do_bfs_bytes(const gs_graph *g);  // workspace bytes run_bfs_do allocates
int build_adjacency(gs_graph *g);
int pr_pull(const gs_graph *g, gs_ws *w, const float *prior, float *acc, double *acc64);
int build_blocked(gs_graph *g);
// off[0 .. n] = exclusive prefix of cnt[0 .. n) on stream st (off[n] = the total).
int scan_counts(hipStream_t st, const int32_t *cnt, uint64_t n, uint64_t *off, uint64_t *total);  // the blocked copy (after the lists); pb_k = 0 when too many blocks
int pr_blocked(const gs_graph *g, gs_ws *w, float *prior, const float *fac);
int pr_blocked64(const gs_graph *g, gs_ws *w, const float *prior, double *acc);
size_t pb_bytes(const gs_graph *g);
// values of relabelled ids from `lim` on read as 0
int unpermute_f32(const gs_graph *g, hipStream_t st, const float *x_r, float *out, uint64_t lim);
int unpermute_i32(const gs_graph *g, hipStream_t st, const int32_t *x_r, int32_t *out);
int ensure_queue(const gs_graph *g, gs_ws *w);  // queue, qdeg, qoff, qsum sized for N
int queue_prefix(gs_ws *w, uint64_t nq, std::vector<unsigned long long> &scratch, uint64_t *total);
int run_bfs_batch(const gs_graph *g, gs_ws *w, const uint64_t *roots, int nr, uint8_t *depth_out,
                  uint64_t *level_counts, int max_levels, int *levels,
                  bool keep_depth);
int run_wcc(const gs_graph *g, gs_ws *w, uint32_t *label_out, uint64_t *ncomp);
int run_kcore(const gs_graph *g, gs_ws *w, int k, int32_t *deg_out, uint64_t *count_ne0, uint64_t *count_pos,
              int *rounds);

// Adds every thread's count into *ctr with one atomic per workgroup: an
// atomic per thread on one word serialises (half a million of them cost a
// large bottom-up level at 2^28 vertices about a tenth of its time).  Every
// thread of the workgroup must call it (blockDim a multiple of 64).
__device__ inline void block_count(uint64_t c, unsigned long long *ctr) {
    __shared__ unsigned long long s_tot;
    if (threadIdx.x == 0) s_tot = 0;
    __syncthreads();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_tot, (unsigned long long)c);
    __syncthreads();
    if (threadIdx.x == 0 && s_tot) atomicAdd(ctr, s_tot);
}

// Inclusive prefix sum over the 64 lanes of a wave in six DPP adds: row_shr
// 1, 2, 4, 8 within each row of 16 lanes, then row_bcast:15 (rows 1 and 3 add
// the last lane of the row below) and row_bcast:31 (rows 2 and 3 add lane
// 31).  A lane whose source is out of its row reads 0.  Every lane of the
// wave must be active, or the inactive ones a suffix (sources are lower lanes).
__device__ inline uint32_t wave_incl_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

// OR over the 64 lanes of a wave, returned to every lane: the DPP pattern of
// wave_incl_scan with OR (an inclusive OR scan; lane 63 holds the total), each
// 32-bit half on its own, then read from lane 63.  Twelve VALU operations
// where six 64-bit shuffles took twelve LDS permutes.  All lanes active.
__device__ inline uint32_t wave_or_scan32(uint32_t x) {
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}
__device__ inline unsigned long long wave_or64(unsigned long long x) {
    const uint32_t lo = wave_or_scan32((uint32_t)x), hi = wave_or_scan32((uint32_t)(x >> 32));
    return (unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)hi, 63) << 32 |
           (uint32_t)__builtin_amdgcn_readlane((int)lo, 63);
}

// Largest j in [lo, hi] with qoff[j] <= target, given qoff[lo] <= target
// (qoff non-decreasing, a queue prefix of list lengths), by the 64 lanes of a wave: each step probes 64
// evenly spaced entries and keeps the interval after the last one <= target,
// so a queue of 2^18 entries takes 3 steps of one load per lane (a binary
// search by one thread took 18 dependent loads per chunk, during which the
// rest of the workgroup waited).  Every lane of the wave calls it.
__device__ inline uint64_t wave_upper(const uint64_t *__restrict__ qoff, uint64_t lo, uint64_t hi, uint64_t target) {
    const uint64_t lane = threadIdx.x & 63;
    while (hi > lo) {
        const uint64_t span = hi - lo + 1;
        const uint64_t pos = lo + span * lane / 64;
        const unsigned long long m = __ballot(qoff[pos] <= target);   // lane 0 holds: qoff[lo] <= target
        const uint64_t l = 63 - __clzll(m);
        const uint64_t nlo = lo + span * l / 64;
        hi = l == 63 ? hi : lo + span * (l + 1) / 64 - 1;
        lo = nlo;
    }
    return lo;
}

// XCD-aware chunk order: workgroup b runs on XCD b % 8; give each XCD a
// contiguous run of chunks so its L2 sees consecutive tiles.
__device__ inline uint64_t xcd_chunk(uint64_t bid, uint64_t nchunk) {
    uint64_t per = (nchunk + NUM_XCD - 1) / NUM_XCD;
    return (bid % NUM_XCD) * per + bid / NUM_XCD;
}
inline uint64_t xcd_grid(uint64_t nchunk) {
    uint64_t per = (nchunk + NUM_XCD - 1) / NUM_XCD;
    return per * NUM_XCD;
}

// Tiles of one chunk, staged in LDS.
struct ChunkTiles {
    uint16_t beg[CHUNK + 2];  // first edge of tile k relative to the chunk (clamped), beg[nt] = len
    uint32_t rc[CHUNK + 1];
    uint8_t act[CHUNK + 1];   // per-tile activity flag (frontier/peel sweeps)
    int nt;
    int any;
};

// One sweep launch covers chunks [chunk0, chunk0 + nlaunch); `grid` holds the
// edges of chunks from grid_chunk0 on (the whole grid when resident, one
// staged segment when streaming).
struct SweepArgs {
    const uint32_t *grid;
    const uint64_t *tz_beg;
    const uint32_t *tz_rc;
    const uint32_t *chunk_tz;
    uint64_t E, nchunk;
    uint64_t grid_chunk0, chunk0, nlaunch;
    uint32_t chunk;  // edges per chunk, a multiple of 4 * SWEEP_WG, at most CHUNK
    int xcd_order;   // 1: each XCD sweeps a contiguous eighth; 0: chunks in launch order
};

inline SweepArgs sweep_args(const gs_graph *g, const uint32_t *grid, uint64_t grid_chunk0, uint64_t chunk0,
                            uint64_t nlaunch) {
    return SweepArgs{grid, g->tz_beg.ptr, g->tz_rc.ptr, g->chunk_tz.ptr, g->E, g->nchunk,
                     grid_chunk0, chunk0, nlaunch, g->chunk, g->xcd_order};
}

// Chunk of this workgroup, false if the workgroup is past the launch's range.
__device__ inline bool launch_chunk(const SweepArgs &a, uint64_t &chunk) {
    uint64_t c = a.xcd_order ? xcd_chunk(blockIdx.x, a.nlaunch) : blockIdx.x;
    if (c >= a.nlaunch) return false;
    chunk = a.chunk0 + c;
    return true;
}

__device__ inline bool bit_of(const uint32_t *bm, uint32_t rc_row) {
    return (bm[rc_row >> 5] >> (rc_row & 31)) & 1u;
}

// Stage the chunk's tiles.  ActFn(rc) -> bool gives a tile's activity.
// Returns chunk length (0 if the chunk is out of range).
template <class ActFn>
__device__ inline uint32_t stage_chunk(const SweepArgs &a, uint64_t chunk, ChunkTiles &s, ActFn act_fn) {
    uint64_t base = chunk * (uint64_t)a.chunk;
    uint32_t len = (uint32_t)min<uint64_t>((uint64_t)a.chunk, a.E - base);
    uint32_t t0 = a.chunk_tz[chunk], t1 = a.chunk_tz[chunk + 1];
    int nt = (int)(t1 - t0) + 1;
    if (threadIdx.x == 0) {
        s.nt = nt;
        s.any = 0;
        s.beg[nt] = (uint16_t)len;
    }
    int any = 0;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) {
        uint64_t b = a.tz_beg[t0 + k];
        s.beg[k] = (uint16_t)(b <= base ? 0 : min<uint64_t>(b - base, len));
        uint32_t rc = a.tz_rc[t0 + k];
        s.rc[k] = rc;
        bool ac = act_fn(rc);
        s.act[k] = ac;
        any |= ac;
    }
    __syncthreads();
    if (any) s.any = 1;  // benign same-value race
    __syncthreads();
    return len;
}

// Adds the chunk's edges in active tiles to out[0] and the active tiles that
// start in this chunk to out[1] (each tile counted once over the sweep; one
// atomic each per workgroup).
__device__ inline void count_active_edges(const SweepArgs &a, uint64_t chunk, const ChunkTiles &s,
                                          unsigned long long *out) {
    __shared__ uint32_t acc, tiles;
    if (threadIdx.x == 0) acc = tiles = 0;
    __syncthreads();
    const uint32_t len = s.beg[s.nt];
    uint32_t c = 0, t = 0;
    for (int k = threadIdx.x; k < s.nt; k += blockDim.x)
        if (s.act[k]) {
            c += (uint32_t)s.beg[k + 1] - s.beg[k];
            const bool starts_here = k ? s.beg[k] < len
                                       : len && a.tz_beg[a.chunk_tz[chunk]] == chunk * (uint64_t)a.chunk;
            t += starts_here;
        }
    if (c) atomicAdd(&acc, c);
    if (t) atomicAdd(&tiles, t);
    __syncthreads();
    if (threadIdx.x == 0 && acc) atomicAdd(out, (unsigned long long)acc);
    if (threadIdx.x == 0 && tiles) atomicAdd(out + 1, (unsigned long long)tiles);
}

__device__ inline int find_tile(const ChunkTiles &s, uint32_t off) {
    int lo = 0, hi = s.nt - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (s.beg[mid] <= off) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Visit every edge of the chunk: op(word, rc, tile_k).  Lanes read 16-byte
// quads so a wave streams 1 KiB of contiguous edges per load.  Edges of
// inactive tiles are skipped when skip_inactive is set (quads fully inside
// one inactive tile are not loaded at all).
template <bool SKIP_INACTIVE, class Op>
__device__ inline void sweep_chunk(const SweepArgs &a, uint64_t chunk, const ChunkTiles &s,
                                   uint32_t len, Op &op) {
    const uint32_t *g = a.grid + (chunk - a.grid_chunk0) * (uint64_t)a.chunk;
    const int quads = (int)(a.chunk / (4 * SWEEP_WG));
#pragma unroll 1
    for (int q = 0; q < quads; ++q) {
        uint32_t off = (uint32_t)(q * SWEEP_WG + threadIdx.x) * 4u;
        if (off >= len) break;
        int k = find_tile(s, off);
        if (SKIP_INACTIVE) {
            if (!s.act[k] && s.beg[k + 1] >= off + 4) continue;
        }
        uint32_t w[4];
        if (off + 4 <= len) {
            uint4 v = *reinterpret_cast<const uint4 *>(g + off);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            for (int i = 0; i < 4; ++i) w[i] = (off + i < len) ? g[off + i] : 0u;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint32_t e = off + i;
            if (e >= len) break;
            while (s.beg[k + 1] <= e) ++k;
            if (SKIP_INACTIVE && !s.act[k]) continue;
            op(w[i], s.rc[k]);
        }
    }
}

// As sweep_chunk, in two phases per quad: pre(word, rc) for all four edges
// first -- its loads (e.g. the frontier state of both endpoints) issue
// together -- then op(word, rc, pre_result) for the edges to visit.  pre also
// runs on skipped or out-of-range edges with word 0 and a valid tile, so it
// must only read.
template <bool SKIP_INACTIVE, class Pre, class Op>
__device__ inline void sweep_chunk_pre(const SweepArgs &a, uint64_t chunk, const ChunkTiles &s, uint32_t len,
                                       Pre &pre, Op &op) {
    const uint32_t *g = a.grid + (chunk - a.grid_chunk0) * (uint64_t)a.chunk;
    const int quads = (int)(a.chunk / (4 * SWEEP_WG));
#pragma unroll 1
    for (int q = 0; q < quads; ++q) {
        uint32_t off = (uint32_t)(q * SWEEP_WG + threadIdx.x) * 4u;
        if (off >= len) break;
        int k = find_tile(s, off);
        if (SKIP_INACTIVE) {
            if (!s.act[k] && s.beg[k + 1] >= off + 4) continue;
        }
        uint32_t w[4];
        if (off + 4 <= len) {
            uint4 v = *reinterpret_cast<const uint4 *>(g + off);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            for (int i = 0; i < 4; ++i) w[i] = (off + i < len) ? g[off + i] : 0u;
        }
        uint32_t rc[4];
        bool ok[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t e = off + i;
            ok[i] = e < len;
            while (ok[i] && s.beg[k + 1] <= e) ++k;
            rc[i] = s.rc[k];
            if (SKIP_INACTIVE && !s.act[k]) ok[i] = false;
            if (!ok[i]) w[i] = 0u;
        }
        decltype(pre(0u, 0u)) st[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) st[i] = pre(w[i], rc[i]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (ok[i]) op(w[i], rc[i], st[i]);
    }
}

// PageRank: new prior of a vertex from its sum (pr.cpp:276-290): r = 0.15 +
// 0.85 * acc in FP64 with round-to-nearest steps (no contraction), times the
// degree factor, rounded to FP32 once.
__device__ inline float pr_new_prior(float acc, float fac) {
    const double r = __dadd_rn(0.15, __dmul_rn(0.85, (double)acc));
    return __double2float_rn(__dmul_rn(r, (double)fac));
}

// Global ids of an edge word in tile rc: row << 16 | v0_local, col << 16 | v1_local.
__device__ inline uint32_t gv0(uint32_t w, uint32_t rc) { return (rc & 0xFFFF0000u) | (w & 0xFFFFu); }
__device__ inline uint32_t gv1(uint32_t w, uint32_t rc) { return (rc << 16) | (w >> 16); }

// Launch one sweep over every chunk of the grid: one launch when resident;
// when streaming, the resident prefix first, then each segment after its
// host-to-device copy, through nbuf segment buffers: the copy of segment
// k + nbuf - 1 overlaps the sweep of segment k, and with three buffers a
// segment whose sweep takes longer than a copy does not stall the copy
// engine.
// launch(SweepArgs, grid_size, stream) issues the kernel.  interleave (sweeps
// whose result does not depend on the order of the chunks, such as the
// PageRank scatter): the resident prefix is cut into segment-sized pieces
// launched between the streamed segments, so the copies never wait for it.
template <class F>
int sweep(const gs_graph *g, gs_ws *w, int kclass, F &&launch, bool direct = false, bool interleave = false) {
    if (!g->nchunk) return GS_OK;
    auto run = [&](const uint32_t *grid, uint64_t gc0, uint64_t c0, uint64_t cn) -> int {
        hipEvent_t kb;
        w->timer.begin(kclass, w->stream, &kb);
        launch(sweep_args(g, grid, gc0, c0, cn), (unsigned)(g->xcd_order ? xcd_grid(cn) : cn), w->stream);
        GS_HIP(hipGetLastError());
        w->timer.end(kclass, w->stream, kb);
        return GS_OK;
    };
    StreamState *ss = g->ss;
    if (!ss) return run(g->grid.ptr, 0, 0, g->nchunk);
    const uint64_t nseg = (g->nchunk - ss->res_chunks + ss->seg_chunks - 1) / ss->seg_chunks;
    interleave = interleave && !direct && nseg > 0;
    // resident chunks launched before each streamed segment when interleaving
    const uint64_t res_per_seg =
        interleave ? (ss->res_chunks / ss->seg_chunks + nseg) / nseg * ss->seg_chunks : 0;
    uint64_t res_done = interleave ? 0 : ss->res_chunks;
    if (!interleave && ss->res_chunks) GS_TRY(run(g->grid.ptr, 0, 0, ss->res_chunks));
    // direct: a sweep that reads few tiles (a sparse frontier) reads them from
    // the pinned host grid in place, instead of copying every segment over
    if (direct && ss->h_grid_dev) {   // never in external mode (no host copy)
        if (ss->res_chunks < g->nchunk) GS_TRY(run(ss->h_grid_dev, 0, ss->res_chunks, g->nchunk - ss->res_chunks));
        return GS_OK;
    }
    if (ss->disk && ss->res_chunks < g->nchunk) GS_TRY(disk_pass_begin(ss));
    int k = 0;
    for (uint64_t c0 = ss->res_chunks; c0 < g->nchunk; c0 += ss->seg_chunks, k = (k + 1) % ss->nbuf) {
        uint64_t cn = std::min(ss->seg_chunks, g->nchunk - c0);
        uint64_t e0 = c0 * g->chunk, e1 = std::min(g->E, (c0 + cn) * g->chunk);
        GS_HIP(hipStreamWaitEvent(ss->copy, ss->ev_done[k], 0));
        hipEvent_t cb;
        w->timer.begin(GS_K_H2D, ss->copy, &cb);
        if (ss->disk)
            GS_TRY(disk_copy(g, ss, (c0 - ss->res_chunks) / ss->seg_chunks, e0, e1, ss->seg[k].ptr));
        else
            GS_HIP(hipMemcpyAsync(ss->seg[k].ptr, ss->h_grid + e0, (e1 - e0) * 4, hipMemcpyHostToDevice, ss->copy));
        w->timer.end(GS_K_H2D, ss->copy, cb);
        GS_HIP(hipEventRecord(ss->ev_copied[k], ss->copy));
        if (res_done < ss->res_chunks) {   // interleaved resident chunks, while the copy runs
            const uint64_t rn = std::min(res_per_seg, ss->res_chunks - res_done);
            GS_TRY(run(g->grid.ptr, 0, res_done, rn));
            res_done += rn;
        }
        GS_HIP(hipStreamWaitEvent(w->stream, ss->ev_copied[k], 0));
        GS_TRY(run(ss->seg[k].ptr, c0, c0, cn));
        GS_HIP(hipEventRecord(ss->ev_done[k], w->stream));
        w->stats.h2d_bytes += (e1 - e0) * 4;
    }
    if (res_done < ss->res_chunks) GS_TRY(run(g->grid.ptr, 0, res_done, ss->res_chunks - res_done));
    return GS_OK;
}

}  // namespace gs
