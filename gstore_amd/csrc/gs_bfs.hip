// gs_bfs.hip -- level-synchronous BFS over the resident tile grid.
//
// Restates bfs_t (bfs.cpp:56-338) driven as bfs_mmap (gstore.cpp:2631-2713):
// depth is u8 with 0 = unreached and root = 1; level L reaches v1 from v0
// when d[v0] == L and d[v1] == 0, and v0 from v1 the other way round
// (bfs.cpp:158-177); a tile is swept only when its tile row or tile column
// holds a frontier vertex (gstore.cpp:2687-2698).
//
// Device form: the frontier (depth == L), the visited set (depth != 0) and
// the next frontier are N-bit bitmaps, so the per-edge test reads two bits
// from 8 KiB tile windows that stay in cache; newly reached vertices are set
// with one atomicOr per (vertex word) and their depth bytes are written once
// by the level-end pass, which also rebuilds the tile-row activity bits and
// counts the new frontier exactly.
#include <chrono>

#include "gs_internal.h"

namespace gs {

__global__ void k_bfs_root(uint32_t root, uint8_t *depth, uint32_t *vis, uint32_t *front, uint32_t *rowact) {
    depth[root] = 1;
    vis[root >> 5] |= 1u << (root & 31);
    front[root >> 5] |= 1u << (root & 31);
    uint32_t r = root >> 16;
    rowact[r >> 5] |= 1u << (r & 31);
}

template <bool DIRECTED>
__global__ void __launch_bounds__(SWEEP_WG) k_bfs_expand(SweepArgs a, const uint32_t *__restrict__ front,
                                                         const uint32_t *__restrict__ vis, uint32_t *next,
                                                         const uint32_t *__restrict__ rowact,
                                                         unsigned long long *scanned) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [&](uint32_t rc) {
        return bit_of(rowact, rc >> 16) || (!DIRECTED && bit_of(rowact, rc & 0xFFFFu));
    });
    if (!s.any) return;
    count_active_edges(a, chunk, s, scanned);
    auto reach = [&](uint32_t v) {
        uint32_t wi = v >> 5, b = 1u << (v & 31);
        if ((vis[wi] & b) == 0 && (next[wi] & b) == 0) atomicOr(&next[wi], b);
    };
    // frontier words of both endpoints of the quad's four edges load together
    auto pre = [&](uint32_t w, uint32_t rc) {
        uint2 f;
        f.x = front[gv0(w, rc) >> 5];
        f.y = DIRECTED ? 0u : front[gv1(w, rc) >> 5];
        return f;
    };
    auto op = [&](uint32_t w, uint32_t rc, uint2 f) {
        const uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        if ((f.x >> (v0 & 31)) & 1u) {
            reach(v1);
        } else if (!DIRECTED && ((f.y >> (v1 & 31)) & 1u)) {
            reach(v0);
        }
    };
    sweep_chunk_pre<true>(a, chunk, s, len, pre, op);
}

// One workgroup per tile row (2048 bitmap words).
__global__ void __launch_bounds__(256) k_bfs_level_end(uint32_t *vis, uint32_t *front, uint32_t *next,
                                                       uint32_t *rowact, uint8_t *depth, uint32_t newdepth,
                                                       unsigned long long *count) {
    const uint32_t r = blockIdx.x;
    const uint64_t base = (uint64_t)r * ROW_WORDS;
    uint32_t cnt = 0;
    for (uint32_t w = threadIdx.x; w < ROW_WORDS; w += blockDim.x) {
        uint64_t gw = base + w;
        uint32_t nx = next[gw];
        front[gw] = nx;
        if (nx) {
            next[gw] = 0;
            vis[gw] |= nx;
            cnt += __popc(nx);
            uint32_t m = nx;
            while (m) {
                uint32_t b = __ffs(m) - 1;
                m &= m - 1;
                depth[gw * 32 + b] = (uint8_t)newdepth;
            }
        }
    }
    int any = __syncthreads_or(cnt != 0);
    // block sum of cnt
    __shared__ unsigned long long tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    if (cnt) atomicAdd(&tot, (unsigned long long)cnt);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (any) atomicOr(&rowact[r >> 5], 1u << (r & 31));
        else atomicAnd(&rowact[r >> 5], ~(1u << (r & 31)));
        if (tot) atomicAdd(count, tot);
        if (any) atomicAdd(count + 3, 1ull);  // tile rows holding the new frontier
    }
}

}  // namespace gs

using namespace gs;

int gs::run_bfs(const gs_graph *g, gs_ws *w, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
                 int max_levels, int *levels) {
    if (!g || root >= g->n || max_levels < 0) {
        set_error("gs_bfs: bad arguments (root %llu, N %llu)", (unsigned long long)root,
                  g ? (unsigned long long)g->n : 0ull);
        return GS_ERR_ARG;
    }
    GS_TRY(ensure_bitmaps(g, w));
    if (w->depth.n != g->n) GS_TRY(w->depth.alloc(g->n));
    const uint64_t words = bitmap_words(g);
    w->timer.reset();
    w->stats = gs_stats{};
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    GS_HIP(hipMemsetAsync(w->depth.ptr, 0, g->n, w->stream));
    GS_HIP(hipMemsetAsync(w->bm_vis.ptr, 0, words * 4, w->stream));
    GS_HIP(hipMemsetAsync(w->bm_front.ptr, 0, words * 4, w->stream));
    GS_HIP(hipMemsetAsync(w->bm_next.ptr, 0, words * 4, w->stream));
    GS_HIP(hipMemsetAsync(w->rowact.ptr, 0, w->rowact.bytes(), w->stream));
    GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 32, w->stream));
    k_bfs_root<<<1, 1, 0, w->stream>>>((uint32_t)root, w->depth.ptr, w->bm_vis.ptr, w->bm_front.ptr, w->rowact.ptr);
    GS_HIP(hipGetLastError());
    int nlev = 0, status = GS_OK;
    w->levels.clear();
    EventPair ev_la;
    GS_HIP(hipEventCreate(&ev_la.a));
    GS_HIP(hipEventCreate(&ev_la.b));
    hipEvent_t la = ev_la.a, lb = ev_la.b;
    uint64_t prev_e = 0, prev_t = 0, front_n = 1, rows_on = 1;
    for (uint32_t level = 1;; ++level) {
        GS_HIP(hipEventRecord(la, w->stream));
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 8, w->stream));
        GS_HIP(hipMemsetAsync(w->counters.ptr + 3, 0, 8, w->stream));
        hipEvent_t kb;
        // semi-external: with the frontier in at most 1/16 of the tile rows the
        // sweep reads about 1/8 of the tiles, which it reads from host memory in
        // place rather than streaming the whole grid through the segments
        const bool direct = g->ss && rows_on * 16 <= g->rows;
        GS_TRY(sweep(g, w, GS_K_BFS_EXPAND, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            if (g->directed)
                k_bfs_expand<true><<<grid, SWEEP_WG, 0, st>>>(a, w->bm_front.ptr, w->bm_vis.ptr, w->bm_next.ptr,
                                                             w->rowact.ptr, w->counters.ptr + 1);
            else
                k_bfs_expand<false><<<grid, SWEEP_WG, 0, st>>>(a, w->bm_front.ptr, w->bm_vis.ptr, w->bm_next.ptr,
                                                              w->rowact.ptr, w->counters.ptr + 1);
        }, direct));
        w->timer.begin(GS_K_BFS_LEVEL_END, w->stream, &kb);
        k_bfs_level_end<<<(unsigned)g->rows, 256, 0, w->stream>>>(w->bm_vis.ptr, w->bm_front.ptr, w->bm_next.ptr,
                                                                   w->rowact.ptr, w->depth.ptr, level + 1,
                                                                   w->counters.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_BFS_LEVEL_END, w->stream, kb);
        GS_HIP(hipEventRecord(lb, w->stream));
        GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 32, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        uint64_t cnt = w->h_counters[0];
        rows_on = w->h_counters[3];
        {
            float lms = 0.f;
            GS_HIP(hipEventElapsedTime(&lms, la, lb));
            w->levels.push_back(gs_level_stat{level, 0, front_n, w->h_counters[1] - prev_e, w->h_counters[2] - prev_t,
                                              cnt, lms});
            prev_e = w->h_counters[1];
            prev_t = w->h_counters[2];
            front_n = cnt;
        }
        if (level_counts && nlev < max_levels) level_counts[nlev] = cnt;
        ++nlev;
        if (cnt == 0) break;
        if (level + 1 > 255) {
            set_error("BFS from %llu needs depth > 255, beyond the u8 depth encoding", (unsigned long long)root);
            status = GS_ERR_DEPTH_OVERFLOW;
            break;
        }
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)nlev;
    w->stats.edges_scanned = w->h_counters[1];
    w->stats.tiles_scanned = w->h_counters[2];
    if (levels) *levels = nlev;
    if (status != GS_OK) return status;
    if (depth_out) {
        GS_HIP(hipMemcpyAsync(depth_out, w->depth.ptr, g->n, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

// ---- BFS that allocates nothing ------------------------------------------------
// The same tile-sweep levels with every work array -- level counters, tile-row
// activity bits, the three bitmaps and the depth bytes -- carved out of one
// file-scope device array instead of the workspace.  Each GPU holds its own copy
// of the array for the life of the process; its device address is looked up
// once per device with hipGetSymbolAddress (the array's name in host code is
// only a stand-in and must never reach a kernel).  Calls on one GPU share the
// array, so calls that overlap on one GPU race; each call clears what it uses.
namespace gs {

constexpr uint64_t NOALLOC_BYTES = 32ull << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex scratch_mu;
static unsigned char *scratch_dev[64];   // per device id: scratch's device address

static int scratch_base(unsigned char **out) {
    int dev = 0;
    GS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) { set_error("device id %d out of range", dev); return GS_ERR_DEVICE; }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return GS_ERR_DEVICE;
        }
        scratch_dev[dev] = (unsigned char *)p;
    }
    *out = scratch_dev[dev];
    return GS_OK;
}

__global__ void k_zero16(uint4 *p, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

__global__ void k_level_reset(unsigned long long *c) {
    c[0] = 0;  // vertices reached this level
    c[3] = 0;  // tile rows holding them
}

static uint64_t up256(uint64_t b) { return (b + 255) & ~255ull; }

// Scratch bytes the no-allocation BFS needs on graph g.
static uint64_t noalloc_bytes(const gs_graph *g) {
    return 256 + up256(4 * ((g->p_s + 31) / 32)) + 3 * up256(4 * bitmap_words(g)) + up256(g->n);
}

}  // namespace gs

extern "C" int gs_bfs_noalloc(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
                              int max_levels, int *levels) {
    if (!g || root >= g->n || max_levels < 0) {
        set_error("gs_bfs_noalloc: bad arguments");
        return GS_ERR_ARG;
    }
    if (g->ss) {
        set_error("gs_bfs_noalloc: the grid is in semi-external mode, whose sweeps stream through segments");
        return GS_ERR_UNSUPPORTED;
    }
    const uint64_t need = noalloc_bytes(g);
    if (need > NOALLOC_BYTES) {
        set_error("gs_bfs_noalloc: N = %llu needs %llu bytes of work space, beyond the %llu-byte static array "
                  "(about 24 M vertices at most)", (unsigned long long)g->n, (unsigned long long)need,
                  (unsigned long long)NOALLOC_BYTES);
        return GS_ERR_ARG;
    }
    unsigned char *base = nullptr;
    GS_TRY(scratch_base(&base));
    gs_ws *w = g->ws;
    hipStream_t st = w->stream;
    const uint64_t words = bitmap_words(g), raw = (g->p_s + 31) / 32;
    unsigned long long *counters = (unsigned long long *)base;
    uint32_t *rowact = (uint32_t *)(base + 256);
    uint32_t *vis = (uint32_t *)((unsigned char *)rowact + up256(4 * raw));
    uint32_t *front = (uint32_t *)((unsigned char *)vis + up256(4 * words));
    uint32_t *next = (uint32_t *)((unsigned char *)front + up256(4 * words));
    uint8_t *depth = (uint8_t *)next + up256(4 * words);
    const auto t0 = std::chrono::steady_clock::now();
    k_zero16<<<(unsigned)std::min<uint64_t>((need / 16 + 255) / 256, 4096), 256, 0, st>>>((uint4 *)base, need / 16);
    GS_HIP(hipGetLastError());
    k_bfs_root<<<1, 1, 0, st>>>((uint32_t)root, depth, vis, front, rowact);
    GS_HIP(hipGetLastError());
    int nlev = 0, status = GS_OK;
    const unsigned grid = (unsigned)(g->xcd_order ? xcd_grid(g->nchunk) : g->nchunk);
    for (uint32_t level = 1;; ++level) {
        k_level_reset<<<1, 1, 0, st>>>(counters);
        if (g->nchunk) {
            const SweepArgs a = sweep_args(g, g->grid.ptr, 0, 0, g->nchunk);
            if (g->directed)
                k_bfs_expand<true><<<grid, SWEEP_WG, 0, st>>>(a, front, vis, next, rowact, counters + 1);
            else
                k_bfs_expand<false><<<grid, SWEEP_WG, 0, st>>>(a, front, vis, next, rowact, counters + 1);
        }
        k_bfs_level_end<<<(unsigned)g->rows, 256, 0, st>>>(vis, front, next, rowact, depth, level + 1, counters);
        GS_HIP(hipGetLastError());
        GS_HIP(hipMemcpyAsync(w->h_counters, counters, 32, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
        const uint64_t cnt = w->h_counters[0];
        if (level_counts && nlev < max_levels) level_counts[nlev] = cnt;
        ++nlev;
        if (cnt == 0) break;
        if (level + 1 > 255) {
            set_error("BFS from %llu needs depth > 255, beyond the u8 depth encoding", (unsigned long long)root);
            status = GS_ERR_DEPTH_OVERFLOW;
            break;
        }
    }
    w->stats = gs_stats{};
    w->stats.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    w->stats.iterations = (uint64_t)nlev;
    w->stats.edges_scanned = w->h_counters[1];
    w->stats.tiles_scanned = w->h_counters[2];
    if (levels) *levels = nlev;
    if (status != GS_OK) return status;
    if (depth_out) {
        GS_HIP(hipMemcpyAsync(depth_out, depth, g->n, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
    }
    return GS_OK;
}

extern "C" int gs_bfs(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *level_counts, int max_levels,
                      int *levels) {
    if (!g) { set_error("gs_bfs: null graph"); return GS_ERR_ARG; }
    if (g->bfs_mode == 1) return run_bfs_do(g, g->ws, root, depth_out, level_counts, max_levels, levels);
    return run_bfs(g, g->ws, root, depth_out, level_counts, max_levels, levels);
}
