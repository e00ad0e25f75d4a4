// gs_pagerank.hip -- PageRank over the resident tile grid.
//
// Restates pr_t (pr.cpp:82-303) with the resident iteration structure of
// pagerank_mmap (gstore.cpp:2572-2629):
//   init      prior = (float)(1.0/N); factor = (int32)(1.0/deg) (compat) or
//             (float)(1.0/deg) (intended); acc = 0            pr.cpp:82-117
//   scatter   for every stored edge: acc[v1] += prior[v0], and for the
//             undirected grid acc[v0] += prior[v1]             pr.cpp:119-154
//   finalize  prior = (float)((0.15 + 0.85*(double)acc) * factor); acc = 0
//                                                              pr.cpp:281-302
// The scatter is one FP32 hardware atomic add per direction (global_atomic_
// add_f32 on coarse-grained HBM, no CAS loop).  Finalize uses explicit
// round-to-nearest double ops so the compiler cannot fuse 0.85*x + 0.15.
#include "gs_internal.h"

namespace gs {

__global__ void k_pr_init(const int32_t *__restrict__ deg, uint64_t n, int mode, float inv_n, float *prior,
                          float *acc, float *fac) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        int32_t d = deg[v];
        float f = 0.0f;
        if (d != 0) {
            double q = 1.0 / (double)d;
            f = mode == GS_PR_COMPAT ? (float)(int32_t)q : (float)q;
        }
        fac[v] = f;
        prior[v] = inv_n;
        acc[v] = 0.0f;
    }
}

template <bool DIRECTED>
__global__ void __launch_bounds__(SWEEP_WG) k_pr_scatter(SweepArgs a, const float *__restrict__ prior, float *acc) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        float p0 = prior[v0];
        if (!DIRECTED) {
            float p1 = prior[v1];
            unsafeAtomicAdd(&acc[v0], p1);
        }
        unsafeAtomicAdd(&acc[v1], p0);
    };
    sweep_chunk<false>(a, chunk, s, len, op);
}

__global__ void k_pr_finalize(uint64_t n, float *__restrict__ prior, float *__restrict__ acc,
                              const float *__restrict__ fac) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        double r = __dadd_rn(0.15, __dmul_rn(0.85, (double)acc[v]));
        prior[v] = __double2float_rn(__dmul_rn(r, (double)fac[v]));
        acc[v] = 0.0f;
    }
}

// Personalised PageRank from one source s (no reference counterpart; same
// scatter as above): r0 = e_s, prior = r * (float)(1/deg),
// r' = 0.15 [v == s] + 0.85 * sum of prior over in-edges.
__global__ void k_ppr_init(const int32_t *__restrict__ deg, uint64_t n, uint64_t src, float *prior, float *acc,
                           float *fac) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        int32_t d = deg[v];
        float f = d != 0 ? (float)(1.0 / (double)d) : 0.0f;
        fac[v] = f;
        prior[v] = v == src ? f : 0.0f;
        acc[v] = 0.0f;
    }
}

__global__ void k_ppr_finalize(uint64_t n, uint64_t src, float *__restrict__ prior, float *__restrict__ acc,
                               const float *__restrict__ fac, float *__restrict__ rank) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        float r = __double2float_rn(__dadd_rn(v == src ? 0.15 : 0.0, __dmul_rn(0.85, (double)acc[v])));
        prior[v] = __fmul_rn(r, fac[v]);
        acc[v] = 0.0f;
        if (rank) rank[v] = r;
    }
}

static unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

}  // namespace gs

using namespace gs;

int gs::run_pagerank(const gs_graph *g, gs_ws *w, int iters, int mode, float *rank_out) {
    if (!g || iters < 0 || (mode != GS_PR_COMPAT && mode != GS_PR_INTENDED)) {
        set_error("gs_pagerank: bad arguments");
        return GS_ERR_ARG;
    }
    if (w->pr_prior.n != g->n) {
        GS_TRY(w->pr_prior.alloc(g->n));
        GS_TRY(w->pr_acc.alloc(g->n));
        GS_TRY(w->pr_fac.alloc(g->n));
    }
    w->timer.reset();
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventCreate(&t1));
    GS_HIP(hipEventRecord(t0, w->stream));
    k_pr_init<<<vgrid(g->n), 256, 0, w->stream>>>(g->degree.ptr, g->n, mode, (float)(1.0 / (double)g->n),
                                                  w->pr_prior.ptr, w->pr_acc.ptr, w->pr_fac.ptr);
    GS_HIP(hipGetLastError());
    for (int it = 0; it < iters; ++it) {
        hipEvent_t kb;
        GS_TRY(sweep(g, w, GS_K_PR_SCATTER, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            if (g->directed) k_pr_scatter<true><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, w->pr_acc.ptr);
            else k_pr_scatter<false><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, w->pr_acc.ptr);
        }));
        w->timer.begin(GS_K_PR_FINALIZE, w->stream, &kb);
        k_pr_finalize<<<vgrid(g->n), 256, 0, w->stream>>>(g->n, w->pr_prior.ptr, w->pr_acc.ptr, w->pr_fac.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PR_FINALIZE, w->stream, kb);
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    w->stats = gs_stats{};
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)iters;
    w->stats.edges_scanned = g->E * (uint64_t)iters;
    w->stats.tiles_scanned = g->nz * (uint64_t)iters;
    if (rank_out) {
        GS_HIP(hipMemcpyAsync(rank_out, w->pr_prior.ptr, g->n * 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

int gs::run_ppr(const gs_graph *g, gs_ws *w, uint64_t src, int iters, float *out) {
    if (!g || iters < 1 || src >= g->n) {
        set_error("gs_ppr: bad arguments");
        return GS_ERR_ARG;
    }
    if (w->pr_prior.n != g->n) {
        GS_TRY(w->pr_prior.alloc(g->n));
        GS_TRY(w->pr_acc.alloc(g->n));
        GS_TRY(w->pr_fac.alloc(g->n));
    }
    if (w->ppr_rank.n != g->n) GS_TRY(w->ppr_rank.alloc(g->n));
    w->timer.reset();
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventCreate(&t1));
    GS_HIP(hipEventRecord(t0, w->stream));
    k_ppr_init<<<vgrid(g->n), 256, 0, w->stream>>>(g->degree.ptr, g->n, src, w->pr_prior.ptr, w->pr_acc.ptr,
                                                   w->pr_fac.ptr);
    GS_HIP(hipGetLastError());
    for (int it = 0; it < iters; ++it) {
        hipEvent_t kb;
        GS_TRY(sweep(g, w, GS_K_PR_SCATTER, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            if (g->directed) k_pr_scatter<true><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, w->pr_acc.ptr);
            else k_pr_scatter<false><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, w->pr_acc.ptr);
        }));
        w->timer.begin(GS_K_PR_FINALIZE, w->stream, &kb);
        k_ppr_finalize<<<vgrid(g->n), 256, 0, w->stream>>>(g->n, src, w->pr_prior.ptr, w->pr_acc.ptr, w->pr_fac.ptr,
                                                           it + 1 == iters ? w->ppr_rank.ptr : nullptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PR_FINALIZE, w->stream, kb);
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    w->stats = gs_stats{};
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)iters;
    w->stats.edges_scanned = g->E * (uint64_t)iters;
    if (out) {
        GS_HIP(hipMemcpyAsync(out, w->ppr_rank.ptr, g->n * 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

extern "C" int gs_ppr(gs_graph *g, uint64_t source, int iters, float *rank_out) {
    if (!g) { set_error("gs_ppr: null graph"); return GS_ERR_ARG; }
    return run_ppr(g, g->ws, source, iters, rank_out);
}

extern "C" int gs_pagerank(gs_graph *g, int iters, int mode, float *rank_out) {
    if (!g) { set_error("gs_pagerank: null graph"); return GS_ERR_ARG; }
    return run_pagerank(g, g->ws, iters, mode, rank_out);
}
