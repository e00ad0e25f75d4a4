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
#include <algorithm>

#include "gs_internal.h"

namespace gs {

// Heavy vertices (degree >= dthr, the few thousand largest) carry the sign
// bit on fac and prior: the scatter then knows from the prior it already
// reads that the vertex's adds go to the workgroup's LDS table instead of a
// contended global atomic.  Values themselves are never negative.
__device__ inline float mark(float x, bool heavy) { return heavy ? __int_as_float(__float_as_int(x) | 0x80000000) : x; }
__device__ inline bool is_heavy(float x) { return __float_as_int(x) < 0; }
__device__ inline float unmark(float x) { return __int_as_float(__float_as_int(x) & 0x7FFFFFFF); }

__global__ void k_pr_init(const int32_t *__restrict__ deg, uint64_t n, int mode, float inv_n, uint32_t dthr,
                          float *prior, float *acc, float *fac) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        int32_t d = deg[v];
        float f = 0.0f;
        if (d != 0) {
            double q = 1.0 / (double)d;
            f = mode == GS_PR_COMPAT ? (float)(int32_t)q : (float)q;
        }
        bool h = (uint32_t)d >= dthr;
        fac[v] = mark(f, h);
        prior[v] = mark(inv_n, h);
        acc[v] = 0.0f;
    }
}

template <bool DIRECTED, class A>
__global__ void __launch_bounds__(SWEEP_WG) k_pr_scatter(SweepArgs a, const float *__restrict__ prior, A *acc) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        float p0 = unmark(prior[v0]);
        if (!DIRECTED) {
            float p1 = unmark(prior[v1]);
            unsafeAtomicAdd(&acc[v0], (A)p1);
        }
        unsafeAtomicAdd(&acc[v1], (A)p0);
    };
    sweep_chunk<false>(a, chunk, s, len, op);
}

// Undirected scatter with heavy vertices accumulated in LDS.  Persistent:
// gridDim.x workgroups, each sweeping a contiguous run of chunks (runs ordered
// so the workgroups of one XCD hold one contiguous region), then adding its
// LDS table to the heavy vertices' acc once.
constexpr int HEAVY_MAX = 2048;

template <class A>
__global__ void __launch_bounds__(SWEEP_WG) k_pr_scatter_heavy(SweepArgs a, const float *__restrict__ prior,
                                                               A *acc, const uint32_t *__restrict__ heavy,
                                                               int nheavy) {
    __shared__ ChunkTiles s;
    __shared__ uint32_t hid[HEAVY_MAX];
    __shared__ A hacc[HEAVY_MAX];
    for (int i = threadIdx.x; i < nheavy; i += blockDim.x) {
        hid[i] = heavy[i];
        hacc[i] = A(0);
    }
    const uint64_t G = gridDim.x;
    const uint64_t rank = (blockIdx.x % NUM_XCD) * ((G + NUM_XCD - 1) / NUM_XCD) + blockIdx.x / NUM_XCD;
    const uint64_t per = (a.nlaunch + G - 1) / G;
    const uint64_t lo = min(a.nlaunch, rank * per), hi = min(a.nlaunch, lo + per);
    auto add = [&](uint32_t v, float x, bool h) {
        if (h) {
            int l = 0, r = nheavy - 1;
            while (l < r) {
                int m = (l + r) >> 1;
                if (hid[m] < v) l = m + 1; else r = m;
            }
            atomicAdd(&hacc[l], (A)x);
        } else {
            unsafeAtomicAdd(&acc[v], (A)x);
        }
    };
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        float p0 = prior[v0], p1 = prior[v1];
        add(v1, unmark(p0), is_heavy(p1));
        add(v0, unmark(p1), is_heavy(p0));
    };
    for (uint64_t c = lo; c < hi; ++c) {
        uint64_t chunk = a.chunk0 + c;
        __syncthreads();  // previous chunk's tiles are no longer read
        uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
        sweep_chunk<false>(a, chunk, s, len, op);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nheavy; i += blockDim.x)
        if (hacc[i] != A(0)) unsafeAtomicAdd(&acc[hid[i]], hacc[i]);
}

// Degree threshold selection: histogram of floor(log2(degree)).
__global__ void k_deg_log_hist(const int32_t *__restrict__ deg, uint64_t n, unsigned long long *hist) {
    __shared__ uint32_t h[33];
    if (threadIdx.x < 33) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t d = (uint32_t)deg[v];
        if (d) atomicAdd(&h[31 - __clz(d)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 33 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void k_collect_heavy(const int32_t *__restrict__ deg, uint64_t n, uint32_t dthr, uint32_t *out,
                                unsigned long long *count) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        if ((uint32_t)deg[v] >= dthr) out[atomicAdd(count, 1ull)] = (uint32_t)v;
}

__global__ void k_pr_finalize(uint64_t n, float *__restrict__ prior, float *__restrict__ acc,
                              const float *__restrict__ fac, int last) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        float f = fac[v];
        float x = pr_new_prior(acc[v], unmark(f));
        prior[v] = last ? x : mark(x, is_heavy(f));
        acc[v] = 0.0f;
    }
}

// Personalised PageRank from one source s (no reference counterpart; same
// scatter as above): r0 = e_s, prior = r * (float)(1/deg),
// r' = 0.15 [v == s] + 0.85 * sum of prior over in-edges.  Ranks and priors
// are FP32; the per-vertex sums accumulate in FP64 so the L1 error against an
// FP64 recurrence stays below 1e-6 whatever the summation order.
__global__ void k_ppr_init(const int32_t *__restrict__ deg, uint64_t n, uint64_t src, uint32_t dthr, float *prior,
                           double *acc, float *fac) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        int32_t d = deg[v];
        float f = d != 0 ? (float)(1.0 / (double)d) : 0.0f;
        bool h = (uint32_t)d >= dthr;
        fac[v] = mark(f, h);
        prior[v] = mark(v == src ? f : 0.0f, h);
        acc[v] = 0.0;
    }
}

__global__ void k_ppr_finalize(uint64_t n, uint64_t src, float *__restrict__ prior, double *__restrict__ acc,
                               const float *__restrict__ fac, float *__restrict__ rank) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        float f = fac[v];
        float r = __double2float_rn(__dadd_rn(v == src ? 0.15 : 0.0, __dmul_rn(0.85, acc[v])));
        prior[v] = mark(__fmul_rn(r, unmark(f)), is_heavy(f));
        acc[v] = 0.0;
        if (rank) rank[v] = r;
    }
}

static unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

static bool use_pull(const gs_graph *g);

// Heavy set: vertices of degree >= 2^b for the smallest b leaving at most
// HEAVY_MAX of them (none below degree 256).  Built once per graph.
// blocked: the caller sums in FP32 and may use the blocked copy of the lists.
int ensure_heavy(gs_graph *g, hipStream_t st, bool blocked) {
    if (use_pull(g)) {
        g->heavy_dthr = 0xFFFFFFFFu;  // the pull has no atomics to relieve
        g->nheavy = 0;
        g->heavy_ready = false;
        return blocked && g->pr_method == 2 ? build_blocked(g) : build_adjacency(g);
    }
    std::lock_guard<std::mutex> lk(g->build_mu);
    if (g->heavy_ready) return GS_OK;
    g->heavy_ready = true;
    g->heavy_dthr = 0xFFFFFFFFu;
    g->nheavy = 0;
    if (g->directed || !g->heavy_on) return GS_OK;
    DevBuf<unsigned long long> hist;
    GS_TRY(hist.alloc(34));
    GS_HIP(hipMemsetAsync(hist.ptr, 0, 34 * 8, st));
    k_deg_log_hist<<<vgrid(g->n), 256, 0, st>>>(g->degree.ptr, g->n, hist.ptr);
    GS_HIP(hipGetLastError());
    unsigned long long h[34];
    GS_HIP(hipMemcpyAsync(h, hist.ptr, 34 * 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    unsigned long long above = 0;
    int b = 32;
    while (b > 8 && above + h[b - 1] <= (unsigned long long)HEAVY_MAX) above += h[--b];
    if (above == 0) return GS_OK;
    const uint32_t dthr = 1u << b;
    GS_TRY(g->heavy_ids.alloc(HEAVY_MAX));
    GS_HIP(hipMemsetAsync(hist.ptr, 0, 8, st));
    k_collect_heavy<<<vgrid(g->n), 256, 0, st>>>(g->degree.ptr, g->n, dthr, g->heavy_ids.ptr, hist.ptr);
    GS_HIP(hipGetLastError());
    std::vector<uint32_t> ids(above);
    GS_HIP(hipMemcpyAsync(ids.data(), g->heavy_ids.ptr, above * 4, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    std::sort(ids.begin(), ids.end());
    GS_HIP(hipMemcpyAsync(g->heavy_ids.ptr, ids.data(), above * 4, hipMemcpyHostToDevice, st));
    GS_HIP(hipStreamSynchronize(st));
    g->heavy_dthr = dthr;
    g->nheavy = (int)above;
    return GS_OK;
}

static bool use_pull(const gs_graph *g) { return g->pr_method >= 1 && !g->directed && !g->ss; }

// One accumulation pass: pull over the neighbour lists when the graph has
// them, else the tile scatter (with the heavy-vertex kernel when there is a
// heavy set).
template <class A>
static int scatter_t(const gs_graph *g, gs_ws *w, A *acc) {
    if (g->nheavy && !g->directed) {
        return sweep(g, w, GS_K_PR_SCATTER, [&](const SweepArgs &a, unsigned, hipStream_t st) {
            unsigned grid = (unsigned)std::min<uint64_t>(a.nlaunch, (uint64_t)g->persistent_wgs);
            k_pr_scatter_heavy<A><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, acc, g->heavy_ids.ptr, g->nheavy);
        });
    }
    return sweep(g, w, GS_K_PR_SCATTER, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
        if (g->directed) k_pr_scatter<true, A><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, acc);
        else k_pr_scatter<false, A><<<grid, SWEEP_WG, 0, st>>>(a, w->pr_prior.ptr, acc);
    });
}

// acc64 non-null: accumulate in FP64 (personalised PageRank).
static int scatter(const gs_graph *g, gs_ws *w, double *acc64 = nullptr) {
    if (use_pull(g)) return pr_pull(g, w, w->pr_prior.ptr, w->pr_acc.ptr, acc64);
    if (acc64) return scatter_t<double>(g, w, acc64);
    return scatter_t<float>(g, w, w->pr_acc.ptr);
}

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
    // the blocked method's value and partial buffers belong to the graph: calls
    // from several workspaces take turns (held until the call's stream is drained)
    std::unique_lock<std::mutex> pb_lock(const_cast<gs_graph *>(g)->pb_mu, std::defer_lock);
    if (use_pull(g) && g->pr_method == 2) pb_lock.lock();
    GS_TRY(ensure_heavy(const_cast<gs_graph *>(g), w->stream, true));
    w->timer.reset();
    w->stats = gs_stats{};
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventCreate(&t1));
    GS_HIP(hipEventRecord(t0, w->stream));
    // the pull works in the index's relabelled ids; results are mapped back at the end
    const bool pull = use_pull(g);
    k_pr_init<<<vgrid(g->n), 256, 0, w->stream>>>(pull ? g->deg_r.ptr : g->degree.ptr, g->n, mode,
                                                  (float)(1.0 / (double)g->n), iters ? g->heavy_dthr : 0xFFFFFFFFu,
                                                  w->pr_prior.ptr, w->pr_acc.ptr, w->pr_fac.ptr);
    GS_HIP(hipGetLastError());
    // the blocked passes write the new priors of the vertices with edges themselves
    const bool blocked = pull && g->pr_method == 2 && g->pb_k;
    for (int it = 0; it < iters; ++it) {
        if (blocked) {
            GS_TRY(pr_blocked(g, w, w->pr_prior.ptr, w->pr_fac.ptr));
            continue;
        }
        hipEvent_t kb;
        GS_TRY(scatter(g, w));
        w->timer.begin(GS_K_PR_FINALIZE, w->stream, &kb);
        k_pr_finalize<<<vgrid(g->n), 256, 0, w->stream>>>(g->n, w->pr_prior.ptr, w->pr_acc.ptr, w->pr_fac.ptr,
                                                          it + 1 == iters);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PR_FINALIZE, w->stream, kb);
    }
    if (blocked && iters > 0 && g->nzn < g->n) {  // vertices without edges: sum 0
        hipEvent_t kb;
        w->timer.begin(GS_K_PR_FINALIZE, w->stream, &kb);
        k_pr_finalize<<<vgrid(g->n - g->nzn), 256, 0, w->stream>>>(g->n - g->nzn, w->pr_prior.ptr + g->nzn,
                                                                    w->pr_acc.ptr + g->nzn, w->pr_fac.ptr + g->nzn, 1);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PR_FINALIZE, w->stream, kb);
    }
    if (pull) {  // prior (relabelled) -> acc (original order) -> prior
        GS_TRY(unpermute_f32(g, w->stream, w->pr_prior.ptr, w->pr_acc.ptr));
        GS_HIP(hipMemcpyAsync(w->pr_prior.ptr, w->pr_acc.ptr, g->n * 4, hipMemcpyDeviceToDevice, w->stream));
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
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
    if (w->ppr_acc.n != g->n) GS_TRY(w->ppr_acc.alloc(g->n));
    // the blocked method sums in FP64 too; its value buffer belongs to the graph (calls take turns)
    std::unique_lock<std::mutex> pb_lock(const_cast<gs_graph *>(g)->pb_mu, std::defer_lock);
    if (use_pull(g) && g->pr_method == 2) pb_lock.lock();
    GS_TRY(ensure_heavy(const_cast<gs_graph *>(g), w->stream, true));
    w->timer.reset();
    w->stats = gs_stats{};
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventCreate(&t1));
    GS_HIP(hipEventRecord(t0, w->stream));
    const bool pull = use_pull(g);
    uint64_t s_id = src;
    if (pull) {
        uint32_t r = 0;
        GS_HIP(hipMemcpyAsync(&r, g->new_id.ptr + src, 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        s_id = r;
    }
    k_ppr_init<<<vgrid(g->n), 256, 0, w->stream>>>(pull ? g->deg_r.ptr : g->degree.ptr, g->n, s_id, g->heavy_dthr,
                                                   w->pr_prior.ptr, w->ppr_acc.ptr, w->pr_fac.ptr);
    GS_HIP(hipGetLastError());
    const bool blocked = pull && g->pr_method == 2 && g->pb_k;
    for (int it = 0; it < iters; ++it) {
        hipEvent_t kb;
        if (blocked) GS_TRY(pr_blocked64(g, w, w->pr_prior.ptr, w->ppr_acc.ptr));
        else GS_TRY(scatter(g, w, w->ppr_acc.ptr));
        w->timer.begin(GS_K_PR_FINALIZE, w->stream, &kb);
        k_ppr_finalize<<<vgrid(g->n), 256, 0, w->stream>>>(g->n, s_id, w->pr_prior.ptr, w->ppr_acc.ptr,
                                                           w->pr_fac.ptr, it + 1 == iters ? w->ppr_rank.ptr : nullptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PR_FINALIZE, w->stream, kb);
    }
    if (pull) {  // rank (relabelled) -> acc (original order) -> rank
        GS_TRY(unpermute_f32(g, w->stream, w->ppr_rank.ptr, w->pr_acc.ptr));
        GS_HIP(hipMemcpyAsync(w->ppr_rank.ptr, w->pr_acc.ptr, g->n * 4, hipMemcpyDeviceToDevice, w->stream));
    }
    if (pull) {  // prior (relabelled) -> acc (original order) -> prior
        GS_TRY(unpermute_f32(g, w->stream, w->pr_prior.ptr, w->pr_acc.ptr));
        GS_HIP(hipMemcpyAsync(w->pr_prior.ptr, w->pr_acc.ptr, g->n * 4, hipMemcpyDeviceToDevice, w->stream));
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
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
