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
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventCreate(&t1));
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
    hipEvent_t la, lb;
    GS_HIP(hipEventCreate(&la));
    GS_HIP(hipEventCreate(&lb));
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
    (void)hipEventDestroy(la);
    (void)hipEventDestroy(lb);
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
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

extern "C" int gs_bfs(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *level_counts, int max_levels,
                      int *levels) {
    if (!g) { set_error("gs_bfs: null graph"); return GS_ERR_ARG; }
    if (g->bfs_mode == 1) return run_bfs_do(g, g->ws, root, depth_out, level_counts, max_levels, levels);
    return run_bfs(g, g->ws, root, depth_out, level_counts, max_levels, levels);
}
