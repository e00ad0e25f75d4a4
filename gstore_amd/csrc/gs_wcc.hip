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
// On undirected resident grids the default method links over the
// neighbour-list index instead (sampled linking, below), in the index's
// relabelled ids with the forest ordered by original id: the same roots, so
// the same labels, with far fewer random accesses.
#include <algorithm>
#include <vector>

#include "gs_internal.h"

namespace gs {

__device__ inline uint32_t ld_par(const uint32_t *p, uint32_t x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_par(uint32_t *p, uint32_t x, uint32_t v) {
    __hip_atomic_store(p + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (halving keeps every vertex under the same root;
// the links keep id order: original ids over the tiles, relabelled ids over
// the index).
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

// Joins the trees of u and v: the larger root is hooked under the smaller one
// with atomicMin; if it was hooked elsewhere meanwhile, join that tree instead.
__device__ inline void wcc_link(uint32_t *par, uint32_t u, uint32_t v) {
    uint32_t ru = wcc_find(par, u), rv = wcc_find(par, v);
    while (ru != rv) {
        uint32_t hi = max(ru, rv), lo = min(ru, rv);
        uint32_t old = atomicMin(&par[hi], lo);
        if (old == hi) break;  // hi was a root and now hangs under lo
        ru = wcc_find(par, old);
        rv = wcc_find(par, lo);
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
    auto op = [&](uint32_t w, uint32_t rc) { wcc_link(par, gv0(w, rc), gv1(w, rc)); };
    sweep_chunk<false>(a, chunk, s, len, op);
}

// (with `only_if`: nothing unless *only_if is set)
__global__ void k_wcc_compress(uint32_t *par, uint64_t n, const unsigned long long *only_if = nullptr) {
    if (only_if && !*only_if) return;
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
__global__ void __launch_bounds__(256) k_wcc_label(uint32_t *par, const int32_t *__restrict__ deg, const uint32_t *__restrict__ in_touch,
                            uint64_t n, unsigned long long *ncomp) {
    uint32_t c = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        bool touched = deg[v] != 0 || (in_touch && ((in_touch[v >> 5] >> (v & 31)) & 1u));
        if (!touched) par[v] = 0xFFFFFFFFu;
        else c += (par[v] == (uint32_t)v);
    }
    block_count(c, ncomp);
}

static unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

// ---- sampled linking over the neighbour-list index (Sutton, Ben-Nun and Barak,
// "Optimizing parallel graph connectivity computation via subgraph sampling",
// IPDPS 2018).  The forest is kept in relabelled ids (the index's own), so a
// vertex's parent entry sits next to its neighbours' in the order the lists
// are walked and the heaviest vertices' entries stay cached.  It is ordered by
// relabelled id (wcc_link: a root hangs under a smaller one), so the heaviest
// vertex of a component becomes its root: a vertex linked to its heaviest
// neighbour hangs itself under it, one compare on its own entry, where an
// order unrelated to degree would make every neighbour of a hub race to hang
// the hub under itself.  The reference's label, the component's smallest
// original id, is found afterwards (k_aff_cmin) and written by k_aff_label.
constexpr uint32_t AFF_ROUNDS = 2;     // neighbours linked for every vertex
constexpr uint32_t AFF_SAMPLES = 1024; // vertices sampled to find the largest component

// Round r links every vertex with its r-th list entry; round 0 takes it from
// first[] (every vertex below nzn has one), without the offsets.  Round 0
// runs on the fresh forest (every vertex its own root): the first entries
// that go to a smaller id form chains of falling ids, and each vertex follows
// its chain to the end (a vertex whose first entry is not smaller, a root) and
// hangs itself there with a plain store to its own parent word.  No other
// thread writes that word and none reads the forest in this round, so the
// links need no find and no atomic.  On Kronecker graphs the chains are a few
// steps and the forest is flat afterwards; a walk is cut after AFF_WALK steps
// (a path-like graph would make the chains as long as the graph), hanging the
// vertex on the chain vertex reached, and then *cut is set so the compress
// pass after the round runs.  A first entry above its vertex (the vertex is
// heavier than every neighbour) is linked in round 1 instead.
constexpr int AFF_WALK = 32;
__global__ void __launch_bounds__(256) k_aff_round(uint32_t *par, const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                            const uint32_t *__restrict__ first, uint64_t nzn, uint32_t r,
                            unsigned long long *linked, unsigned long long *cut) {
    uint32_t c = 0;
    bool cut_here = false;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nzn; x += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t f = first[x];
        if (r == 0) {
            uint32_t y = (uint32_t)x, z = f;
            for (int s = 0; s < AFF_WALK && z < y; ++s) {
                y = z;
                z = first[y];
            }
            cut_here |= z < y;   // y is on the chain, not at its end
            if (y != (uint32_t)x) par[x] = y;
            ++c;
            continue;
        }
        if (r == 1 && f > (uint32_t)x) wcc_link(par, (uint32_t)x, f);
        const uint64_t b = off[x];
        if (off[x + 1] - b > r) {
            wcc_link(par, (uint32_t)x, nbr[b + r]);
            ++c;
        }
    }
    if (cut_here) atomicOr(cut, 1ull);
    block_count(c, linked);
}

__global__ void k_aff_sample(uint32_t *par, uint64_t n, uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= AFF_SAMPLES) return;
    uint64_t z = (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;  // splitmix64 finaliser
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[i] = wcc_find(par, (uint32_t)(z % n));
}

// Remaining neighbours of every vertex outside the component rooted at skip.
__global__ void __launch_bounds__(256) k_aff_final(uint32_t *par, const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                            uint64_t nzn, uint32_t skip, unsigned long long *linked) {
    uint64_t c = 0;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nzn; x += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = off[x] + AFF_ROUNDS, e = off[x + 1];
        if (b >= e) continue;
        const uint32_t u = (uint32_t)x;
        if (wcc_find(par, u) == skip) continue;
        for (uint64_t k = b; k < e; ++k) wcc_link(par, u, nbr[k]);
        c += e - b;
    }
    block_count(c, linked);
}

// cmin[r] = the root's own original id, for every root r of the compressed
// forest (par[x] is x's root).
__global__ void k_aff_cmin_init(const uint32_t *__restrict__ par, const uint32_t *__restrict__ old_id, uint64_t nzn,
                                uint32_t *cmin) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nzn; x += (uint64_t)gridDim.x * blockDim.x)
        if (par[x] == (uint32_t)x) cmin[x] = old_id[x];
}

// cmin[r] = the smallest original id of r's component.  The largest component
// (the one holding the sampled vertex `skip`, nearly every vertex at scale)
// is reduced in registers and once per workgroup; the others take one
// atomicMin per vertex on their root.
__global__ void __launch_bounds__(256) k_aff_cmin(const uint32_t *__restrict__ par, const uint32_t *__restrict__ old_id,
                                                  uint64_t nzn, uint32_t skip, uint32_t *cmin) {
    __shared__ uint32_t wmin[4];
    const uint32_t big = skip < nzn ? par[skip] : 0xFFFFFFFFu;   // its root, after the last links
    uint32_t m = 0xFFFFFFFFu;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < nzn; x += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = par[x];
        if (r == (uint32_t)x) continue;   // the root's own id is already in cmin
        const uint32_t o = old_id[x];
        if (r == big) m = min(m, o);
        else atomicMin(&cmin[r], o);
    }
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) m = min(m, (uint32_t)__shfl_xor(m, s));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t b = min(min(wmin[0], wmin[1]), min(wmin[2], wmin[3]));
        if (b != 0xFFFFFFFFu) atomicMin(&cmin[big], b);   // b is set only when big is a root below nzn
    }
}

// Labels by original id from the compressed forest over relabelled ids: the
// component's smallest original id (cmin of the root) for vertices with
// edges (relabelled ids below nzn), 0xFFFFFFFF for the others; counts the
// roots.
__global__ void __launch_bounds__(256) k_aff_label(const uint32_t *__restrict__ par, const uint32_t *__restrict__ cmin,
                            const uint32_t *__restrict__ new_id, uint64_t nzn, uint64_t n, uint32_t *label,
                            unsigned long long *ncomp) {
    uint32_t c = 0;
    // four vertices per thread: a 16-byte load of new_id and a 16-byte label store
    for (uint64_t o0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) * 4; o0 < n;
         o0 += (uint64_t)gridDim.x * blockDim.x * 4) {
        if (o0 + 4 <= n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(new_id + o0);
            const uint32_t x[4] = {q.x, q.y, q.z, q.w};
            uint32_t r[4], l[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = x[k] < nzn ? par[x[k]] : 0u;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                l[k] = x[k] < nzn ? cmin[r[k]] : 0xFFFFFFFFu;
                c += x[k] < nzn && r[k] == x[k];
            }
            *reinterpret_cast<uint4 *>(label + o0) = uint4{l[0], l[1], l[2], l[3]};
        } else {
            for (uint64_t o = o0; o < n; ++o) {
                const uint32_t x = new_id[o];
                if (x < nzn) {
                    const uint32_t r = par[x];
                    label[o] = cmin[r];
                    c += r == x;
                } else {
                    label[o] = 0xFFFFFFFFu;
                }
            }
        }
    }
    block_count(c, ncomp);
}

}  // namespace gs

using namespace gs;

int gs::run_wcc(const gs_graph *g, gs_ws *w, uint32_t *label_out, uint64_t *ncomp) {
    if (!g) { set_error("gs_wcc: null graph"); return GS_ERR_ARG; }
    if (w->wcc_parent.n != g->n) GS_TRY(w->wcc_parent.alloc(g->n));
    const bool by_index = g->wcc_method == 1 && !g->directed && !g->ss;
    if (by_index) {
        GS_TRY(build_adjacency(const_cast<gs_graph *>(g)));
        if (w->wcc_label.n != g->n) GS_TRY(w->wcc_label.alloc(g->n));
        if (w->wcc_cmin.n != std::max<uint64_t>(g->nzn, 1)) GS_TRY(w->wcc_cmin.alloc(std::max<uint64_t>(g->nzn, 1)));
    }
    w->timer.reset();
    w->stats = gs_stats{};
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    // over the index only the vertices with edges have a parent entry that is read
    k_wcc_init<<<vgrid(g->n), 256, 0, w->stream>>>(w->wcc_parent.ptr, by_index ? g->nzn : g->n);
    GS_HIP(hipGetLastError());
    uint64_t passes = 0;
    if (by_index) {
        // link AFF_ROUNDS neighbours of every vertex, compressing after each round (after
        // the first only when a chain walk was cut: it leaves a flat forest otherwise)
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 8, w->stream));
        GS_HIP(hipMemsetAsync(w->counters.ptr + 3, 0, 8, w->stream));
        hipEvent_t kb;
        for (uint32_t r = 0; r < AFF_ROUNDS && g->nzn; ++r) {
            w->timer.begin(GS_K_WCC_HOOK, w->stream, &kb);
            k_aff_round<<<vgrid(g->nzn), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->adj_off.ptr, g->adj_nbr.ptr,
                                                              g->adj_first.ptr, g->nzn, r, w->counters.ptr,
                                                              w->counters.ptr + 3);
            GS_HIP(hipGetLastError());
            w->timer.end(GS_K_WCC_HOOK, w->stream, kb);
            w->timer.begin(GS_K_WCC_COMPRESS, w->stream, &kb);
            k_wcc_compress<<<vgrid(g->nzn), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->nzn,
                                                                 r == 0 ? w->counters.ptr + 3 : nullptr);
            GS_HIP(hipGetLastError());
            w->timer.end(GS_K_WCC_COMPRESS, w->stream, kb);
        }
        // the most frequent root among sampled vertices: its members skip their other neighbours
        uint32_t skip = 0xFFFFFFFFu;
        if (g->nzn) {
            if (w->wcc_samp.n != AFF_SAMPLES) GS_TRY(w->wcc_samp.alloc(AFF_SAMPLES));
            k_aff_sample<<<(AFF_SAMPLES + 255) / 256, 256, 0, w->stream>>>(w->wcc_parent.ptr, g->nzn, w->wcc_samp.ptr);
            GS_HIP(hipGetLastError());
            std::vector<uint32_t> h(AFF_SAMPLES);
            GS_HIP(hipMemcpyAsync(h.data(), w->wcc_samp.ptr, AFF_SAMPLES * 4, hipMemcpyDeviceToHost, w->stream));
            GS_HIP(hipStreamSynchronize(w->stream));
            std::sort(h.begin(), h.end());
            uint32_t best = 0;
            for (uint32_t i = 0, j; i < AFF_SAMPLES; i = j) {
                for (j = i; j < AFF_SAMPLES && h[j] == h[i]; ++j) {}
                if (j - i > best) { best = j - i; skip = h[i]; }
            }
            w->timer.begin(GS_K_WCC_HOOK, w->stream, &kb);
            k_aff_final<<<vgrid(g->nzn), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->adj_off.ptr, g->adj_nbr.ptr,
                                                              g->nzn, skip, w->counters.ptr);
            GS_HIP(hipGetLastError());
            w->timer.end(GS_K_WCC_HOOK, w->stream, kb);
        }
        if (g->nzn) {   // relabelled ids from nzn on have no edges: their own roots
            w->timer.begin(GS_K_WCC_COMPRESS, w->stream, &kb);
            k_wcc_compress<<<vgrid(g->nzn), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->nzn);
            GS_HIP(hipGetLastError());
            // each component's smallest original id, by root
            k_aff_cmin_init<<<vgrid(g->nzn), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->old_id.ptr, g->nzn,
                                                                  w->wcc_cmin.ptr);
            GS_HIP(hipGetLastError());
            k_aff_cmin<<<vgrid(g->nzn), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->old_id.ptr, g->nzn, skip,
                                                             w->wcc_cmin.ptr);
            GS_HIP(hipGetLastError());
            w->timer.end(GS_K_WCC_COMPRESS, w->stream, kb);
        }
        GS_HIP(hipMemcpyAsync(w->h_counters + 1, w->counters.ptr, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        w->stats.edges_scanned = w->h_counters[1];
        passes = 1;
    }
    for (; !by_index;) {
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
    if (by_index) {
        k_aff_label<<<vgrid((g->n + 3) / 4), 256, 0, w->stream>>>(w->wcc_parent.ptr, w->wcc_cmin.ptr, g->new_id.ptr, g->nzn, g->n,
                                                        w->wcc_label.ptr, w->counters.ptr);
    } else {
        k_wcc_label<<<vgrid(g->n), 256, 0, w->stream>>>(w->wcc_parent.ptr, g->degree.ptr,
                                                        g->directed ? g->in_touch.ptr : nullptr, g->n, w->counters.ptr);
    }
    GS_HIP(hipGetLastError());
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 8, hipMemcpyDeviceToHost, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    w->stats.ms = ms;
    w->stats.iterations = passes;
    if (!by_index) w->stats.edges_scanned = g->E * passes * 2;
    if (ncomp) *ncomp = w->h_counters[0];
    if (label_out) {
        GS_HIP(hipMemcpyAsync(label_out, by_index ? w->wcc_label.ptr : w->wcc_parent.ptr, g->n * 4,
                              hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

extern "C" int gs_wcc(gs_graph *g, uint32_t *label_out, uint64_t *ncomp) {
    if (!g) { set_error("gs_wcc: null graph"); return GS_ERR_ARG; }
    return run_wcc(g, g->ws, label_out, ncomp);
}
