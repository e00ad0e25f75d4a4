// gs_bfs_batch.hip -- up to 64 BFS roots as bit lanes of one sweep.
//
// Each vertex carries three 64-bit masks (visited, frontier, next); lane i is
// the BFS from roots[i].  One level sweeps the tiles whose row or column holds
// a frontier vertex of any lane and applies, per stored edge,
//   next[v1] |= front[v0] & ~vis[v1],  next[v0] |= front[v1] & ~vis[v0]
// which is the single-root rule of bfs.cpp:158-177 for all lanes at once, so
// every lane's depth array equals the single-root result bit for bit.  The
// level-end pass writes lane i's depth (u8, 0 = unreached, root = 1) into
// depth[i * N + v] and counts each lane's new frontier exactly.
#include "gs_internal.h"

namespace gs {

__global__ void k_msbfs_roots(const uint64_t *__restrict__ roots, int nr, uint64_t n, uint8_t *depth,
                              unsigned long long *vis, unsigned long long *front, uint32_t *rowact) {
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
        uint64_t r = roots[i];
        atomicOr(&vis[r], 1ull << i);
        atomicOr(&front[r], 1ull << i);
        depth[(uint64_t)i * n + r] = 1;
        uint32_t row = (uint32_t)(r >> 16);
        atomicOr(&rowact[row >> 5], 1u << (row & 31));
    }
}

template <bool DIRECTED>
__global__ void __launch_bounds__(SWEEP_WG) k_msbfs_expand(SweepArgs a, const unsigned long long *__restrict__ front,
                                                           const unsigned long long *__restrict__ vis,
                                                           unsigned long long *next,
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
    auto push = [&](unsigned long long f, uint32_t v) {
        unsigned long long m = f & ~vis[v];
        if (m && (m & ~next[v])) atomicOr(&next[v], m);
    };
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        unsigned long long f0 = front[v0];
        if (f0) push(f0, v1);
        if (!DIRECTED) {
            unsigned long long f1 = front[v1];
            if (f1) push(f1, v0);
        }
    };
    sweep_chunk<true>(a, chunk, s, len, op);
}

// One workgroup per tile row; lane counts accumulate in LDS.
__global__ void __launch_bounds__(256) k_msbfs_level_end(unsigned long long *vis, unsigned long long *front,
                                                         unsigned long long *next, uint32_t *rowact, uint8_t *depth,
                                                         uint64_t n, uint32_t newdepth,
                                                         unsigned long long *lane_counts) {
    __shared__ uint32_t cnt[64];
    if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t r = blockIdx.x;
    const uint64_t base = (uint64_t)r << 16;
    const int lane = threadIdx.x & 63;
    uint32_t my = 0;  // count for lane index == this thread's lane
    int any = 0;
    for (uint32_t step = 0; step < 256; ++step) {
        uint64_t v = base + step * 256 + threadIdx.x;
        unsigned long long nx = 0;
        if (v < n) {
            nx = next[v];
            if (nx) {
                next[v] = 0;
                vis[v] |= nx;
            }
            front[v] = nx;
        }
        if (__ballot(nx != 0) == 0) continue;  // wave-uniform
        any = 1;
        unsigned long long m = nx;
        while (m) {
            int l = __ffsll(m) - 1;
            m &= m - 1;
            depth[(uint64_t)l * n + v] = (uint8_t)newdepth;
        }
        for (int l = 0; l < 64; ++l) {
            uint32_t c = (uint32_t)__popcll(__ballot((nx >> l) & 1ull));
            if (lane == l) my += c;
        }
    }
    if (my) atomicAdd(&cnt[lane], my);
    int blk_any = __syncthreads_or(any);
    if (threadIdx.x < 64 && cnt[threadIdx.x]) atomicAdd(&lane_counts[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (blk_any) atomicOr(&rowact[r >> 5], 1u << (r & 31));
        else atomicAnd(&rowact[r >> 5], ~(1u << (r & 31)));
    }
}

}  // namespace gs

using namespace gs;

int gs::run_bfs_batch(const gs_graph *g, gs_ws *w, const uint64_t *roots, int nr, uint8_t *depth_out,
                       uint64_t *level_counts, int max_levels, int *levels) {
    if (!g || !roots || nr < 1 || nr > 64 || max_levels < 0) {
        set_error("gs_bfs_batch: bad arguments (need 1..64 roots)");
        return GS_ERR_ARG;
    }
    for (int i = 0; i < nr; ++i)
        if (roots[i] >= g->n) {
            set_error("gs_bfs_batch: root %llu >= N", (unsigned long long)roots[i]);
            return GS_ERR_ARG;
        }
    GS_TRY(ensure_rowact(g, w));
    if (w->ms_vis.n != g->n) {
        GS_TRY(w->ms_vis.alloc(g->n));
        GS_TRY(w->ms_front.alloc(g->n));
        GS_TRY(w->ms_next.alloc(g->n));
    }
    if (w->ms_depth.n < (size_t)nr * g->n) GS_TRY(w->ms_depth.alloc((size_t)nr * g->n));
    if (w->ms_roots.n < 64) GS_TRY(w->ms_roots.alloc(64));
    if (w->ms_counts.n < 64) GS_TRY(w->ms_counts.alloc(64));
    w->timer.reset();
    w->stats = gs_stats{};
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventCreate(&t1));
    GS_HIP(hipEventRecord(t0, w->stream));
    GS_HIP(hipMemcpyAsync(w->ms_roots.ptr, roots, nr * 8, hipMemcpyHostToDevice, w->stream));
    GS_HIP(hipMemsetAsync(w->ms_vis.ptr, 0, g->n * 8, w->stream));
    GS_HIP(hipMemsetAsync(w->ms_front.ptr, 0, g->n * 8, w->stream));
    GS_HIP(hipMemsetAsync(w->ms_next.ptr, 0, g->n * 8, w->stream));
    GS_HIP(hipMemsetAsync(w->ms_depth.ptr, 0, (size_t)nr * g->n, w->stream));
    GS_HIP(hipMemsetAsync(w->rowact.ptr, 0, w->rowact.bytes(), w->stream));
    GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 16, w->stream));
    k_msbfs_roots<<<1, 64, 0, w->stream>>>(w->ms_roots.ptr, nr, g->n, w->ms_depth.ptr,
                                           (unsigned long long *)w->ms_vis.ptr,
                                           (unsigned long long *)w->ms_front.ptr, w->rowact.ptr);
    GS_HIP(hipGetLastError());
    std::vector<unsigned long long> lc(64);
    int nlev = 0, status = GS_OK;
    for (uint32_t level = 1;; ++level) {
        GS_HIP(hipMemsetAsync(w->ms_counts.ptr, 0, 64 * 8, w->stream));
        hipEvent_t kb;
        auto *fr = (unsigned long long *)w->ms_front.ptr, *vi = (unsigned long long *)w->ms_vis.ptr,
             *nx = (unsigned long long *)w->ms_next.ptr;
        GS_TRY(sweep(g, w, GS_K_BFS_EXPAND, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            if (g->directed)
                k_msbfs_expand<true><<<grid, SWEEP_WG, 0, st>>>(a, fr, vi, nx, w->rowact.ptr, w->counters.ptr);
            else
                k_msbfs_expand<false><<<grid, SWEEP_WG, 0, st>>>(a, fr, vi, nx, w->rowact.ptr, w->counters.ptr);
        }));
        w->timer.begin(GS_K_BFS_LEVEL_END, w->stream, &kb);
        k_msbfs_level_end<<<(unsigned)g->rows, 256, 0, w->stream>>>(
            (unsigned long long *)w->ms_vis.ptr, (unsigned long long *)w->ms_front.ptr,
            (unsigned long long *)w->ms_next.ptr, w->rowact.ptr, w->ms_depth.ptr, g->n, level + 1,
            w->ms_counts.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_BFS_LEVEL_END, w->stream, kb);
        GS_HIP(hipMemcpyAsync(lc.data(), w->ms_counts.ptr, 64 * 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        unsigned long long tot = 0;
        for (int i = 0; i < nr; ++i) {
            if (level_counts && nlev < max_levels) level_counts[(size_t)i * max_levels + nlev] = lc[i];
            tot += lc[i];
        }
        ++nlev;
        if (tot == 0) break;
        if (level + 1 > 255) {
            set_error("batched BFS needs depth > 255, beyond the u8 depth encoding");
            status = GS_ERR_DEPTH_OVERFLOW;
            break;
        }
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 16, hipMemcpyDeviceToHost, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)nlev;
    w->stats.edges_scanned = w->h_counters[0];
    w->stats.tiles_scanned = w->h_counters[1];
    if (levels) *levels = nlev;
    if (status != GS_OK) return status;
    if (depth_out) {
        GS_HIP(hipMemcpyAsync(depth_out, w->ms_depth.ptr, (size_t)nr * g->n, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

extern "C" int gs_bfs_batch(gs_graph *g, const uint64_t *roots, int nr, uint8_t *depth_out, uint64_t *level_counts,
                            int max_levels, int *levels) {
    if (!g) { set_error("gs_bfs_batch: null graph"); return GS_ERR_ARG; }
    return run_bfs_batch(g, g->ws, roots, nr, depth_out, level_counts, max_levels, levels);
}
