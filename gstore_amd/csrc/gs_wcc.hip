// gs_wcc.hip -- weakly connected components over the resident tile grid.
//
// The reference (wcc2_t, wcc.cpp:57-395; wcc_mmap gstore.cpp:2502-2570)
// propagates atomic-counter group ids through a cid[] map and pointer-jumps
// it between full passes; its numbering depends on thread timing and its
// printed "WCC count" covers components with a stored edge (wcc.cpp:324-337).
// The device form computes the same partition as min-label propagation over
// a parent forest: every stored edge hooks the larger of its two roots under
// the smaller with atomicMin, finds use pointer jumping (path halving), and a
// compress pass makes every vertex point at its root.  Because a vertex is
// only ever hooked under a smaller id, each root is its component's minimum
// vertex id, which is the canonical label returned.  A check pass counts
// edges whose endpoints still differ and repeats the hook pass until none do.
#include "gs_internal.h"

namespace gs {

__device__ inline uint32_t ld_par(const uint32_t *p, uint32_t x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_par(uint32_t *p, uint32_t x, uint32_t v) {
    __hip_atomic_store(p + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; parent[y] <= y holds throughout.
__device__ inline uint32_t wcc_find(uint32_t *par, uint32_t x) {
    for (;;) {
        uint32_t p = ld_par(par, x);
        if (p == x) return x;
        uint32_t gp = ld_par(par, p);
        if (gp == p) return p;
        st_par(par, x, gp);
        x = gp;
    }
}

__global__ void k_wcc_init(uint32_t *par, uint64_t n) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        par[v] = (uint32_t)v;
}

__global__ void __launch_bounds__(SWEEP_WG) k_wcc_hook(SweepArgs a, uint32_t *par) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t ru = wcc_find(par, gv0(w, rc)), rv = wcc_find(par, gv1(w, rc));
        while (ru != rv) {
            uint32_t hi = max(ru, rv), lo = min(ru, rv);
            uint32_t old = atomicMin(&par[hi], lo);
            if (old == hi) break;  // hi was a root and now hangs under lo
            // hi had been hooked to old meanwhile: join old's tree with lo's
            ru = wcc_find(par, old);
            rv = wcc_find(par, lo);
        }
    };
    sweep_chunk<false>(a, chunk, s, len, op);
}

__global__ void k_wcc_compress(uint32_t *par, uint64_t n) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t r = wcc_find(par, (uint32_t)v);
        if (ld_par(par, (uint32_t)v) != r) st_par(par, (uint32_t)v, r);
    }
}

__global__ void __launch_bounds__(SWEEP_WG) k_wcc_check(SweepArgs a, const uint32_t *__restrict__ par,
                                                        unsigned long long *bad) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    uint32_t nb = 0;
    auto op = [&](uint32_t w, uint32_t rc) { nb += par[gv0(w, rc)] != par[gv1(w, rc)]; };
    sweep_chunk<false>(a, chunk, s, len, op);
    if (nb) atomicAdd(bad, (unsigned long long)nb);
}

// label = root for vertices with a stored edge, 0xFFFFFFFF otherwise; count roots.
// A vertex has a stored edge when its .degree is non-zero (undirected: endpoint
// count) or, on a directed grid (.degree = out-degree), when it has an in-edge.
__global__ void k_wcc_label(uint32_t *par, const int32_t *__restrict__ deg, const uint32_t *__restrict__ in_touch,
                            uint64_t n, unsigned long long *ncomp) {
    uint32_t c = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        bool touched = deg[v] != 0 || (in_touch && ((in_touch[v >> 5] >> (v & 31)) & 1u));
        if (!touched) par[v] = 0xFFFFFFFFu;
        else c += (par[v] == (uint32_t)v);
    }
    if (c) atomicAdd(ncomp, (unsigned long long)c);
}

static unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

}  // namespace gs

using namespace gs;

int gs::run_wcc(const gs_graph *g, gs_ws *w, uint32_t *label_out, uint64_t *ncomp) {
    if (!g) { set_error("gs_wcc: null graph"); return GS_ERR_ARG; }
    if (w->wcc_parent.n != g->n) GS_TRY(w->wcc_parent.alloc(g->n));
    w->timer.reset();
    w->stats = gs_stats{};
    hipEvent_t t0, t1;
    GS_HIP(hipEventCreate(&t0));
    GS_HIP(hipEventRecord(t0, w->stream));
    k_wcc_init<<<vgrid(g->n), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->n);
    GS_HIP(hipGetLastError());
    uint64_t passes = 0;
    for (;;) {
        if (passes >= 64) {
            set_error("gs_wcc: no convergence after %llu hook passes", (unsigned long long)passes);
            return GS_ERR_DEVICE;
        }
        hipEvent_t kb;
        GS_TRY(sweep(g, w, GS_K_WCC_HOOK, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            k_wcc_hook<<<grid, SWEEP_WG, 0, st>>>(a, w->wcc_parent.ptr);
        }));
        ++passes;
        w->timer.begin(GS_K_WCC_COMPRESS, w->stream, &kb);
        k_wcc_compress<<<vgrid(g->n), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->n);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_WCC_COMPRESS, w->stream, kb);
        if (!g->nchunk) break;
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 8, w->stream));
        GS_TRY(sweep(g, w, GS_K_WCC_CHECK, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            k_wcc_check<<<grid, SWEEP_WG, 0, st>>>(a, w->wcc_parent.ptr, w->counters.ptr);
        }));
        GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        if (w->h_counters[0] == 0) break;
    }
    GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 8, w->stream));
    k_wcc_label<<<vgrid(g->n), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->degree.ptr,
                                                    g->directed ? g->in_touch.ptr : nullptr, g->n, w->counters.ptr);
    GS_HIP(hipGetLastError());
    GS_HIP(hipEventCreate(&t1));
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 8, hipMemcpyDeviceToHost, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    w->stats.ms = ms;
    w->stats.iterations = passes;
    w->stats.edges_scanned = g->E * passes * 2;
    if (ncomp) *ncomp = w->h_counters[0];
    if (label_out) {
        GS_HIP(hipMemcpyAsync(label_out, w->wcc_parent.ptr, g->n * 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

extern "C" int gs_wcc(gs_graph *g, uint32_t *label_out, uint64_t *ncomp) {
    if (!g) { set_error("gs_wcc: null graph"); return GS_ERR_ARG; }
    return run_wcc(g, g->ws, label_out, ncomp);
}
