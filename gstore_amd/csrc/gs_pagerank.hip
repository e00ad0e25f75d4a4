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

// ---- Binned scatter (semi-external grids and directed resident grids) ----
// The tile scatter above adds into acc with global atomics whose addresses,
// on a grid that streams through HBM, spread over the whole accumulator (a
// segment spans a few tile rows but every tile column), so each add misses
// the caches.  The binned scatter takes the grid a piece at a time (one
// streamed segment, or a segment-sized run of the resident prefix) and:
//   1. counts the piece's contributions per destination bin -- a tile column
//      (and, undirected, a tile row): 2^16 destinations (k_bin_count);
//   2. scans the counts into bin offsets (k_bin_scan);
//   3. reserves each tile's range in its column bin and row bin: an atomic
//      per tile for the column (a wave's tiles mostly differ there), one per
//      wave for each row it holds (k_bin_pos);
//   4. writes every edge's contributions -- (destination offset in the bin,
//      prior of the source) -- at its tile's place (k_bin_write, a sweep);
//   5. sums each bin in LDS, a part of 2^15 (FP32) or 2^14 (FP64)
//      destinations per workgroup, and adds the part into acc with plain
//      loads and stores: no other workgroup touches those destinations in
//      this step (k_bin_acc).
// Work memory per workspace: 6 bytes per contribution of a piece (two per
// undirected edge), bin tables and 16 bytes per tile of a piece.

// atomicAdd(&ctr[key], n) for each active lane, with each run of lanes that
// hold the same key (consecutive lanes) combined into one atomic by the run's
// first lane; returns each active lane's reserved position: the run's old
// value plus the n of the run's lanes below it.  The tiles of a wave come in
// .start order, so equal tile rows sit in consecutive lanes; a key that
// recurs after another is a run of its own (one more atomic, still correct).
// A constant number of steps whatever the number of distinct keys: a loop
// over them (one scan and one atomic per key) made the per-tile reservations
// of a sparse piece, whose wave spans dozens of tile rows, 40 times slower.
// Every lane of the wave calls it, all lanes active.
__device__ inline uint32_t wave_reserve(uint32_t *ctr, uint32_t key, uint32_t n, bool active) {
    const int lane = threadIdx.x & 63;
    const uint32_t x = active ? n : 0u;
    const uint32_t prev = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    const uint32_t incl = wave_incl_scan(x), excl = incl - x;
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);   // lanes 0 .. lane
    const int head = 63 - __clzll(heads & upto);                               // this lane's run starts here
    const unsigned long long later = heads & ~upto;
    const int last = later ? __ffsll((long long)later) - 2 : 63;               // for a head: its run's last lane
    const uint32_t run_incl = __shfl(incl, last);
    uint32_t base = 0;
    if (lane == head && run_incl != excl) base = atomicAdd(&ctr[key], run_incl - excl);
    return __shfl(base, head) + excl - __shfl(excl, head);
}

// Edges of tile tz inside the piece [e0, e1) and the first of them.
__device__ inline uint32_t piece_tile(const uint64_t *__restrict__ tz_beg, uint64_t tz, uint64_t e0, uint64_t e1,
                                      uint64_t &first) {
    const uint64_t b = max(tz_beg[tz], e0), e = min(tz_beg[tz + 1], e1);
    first = b;
    return e > b ? (uint32_t)(e - b) : 0u;
}

// cnt[bin] += the piece's contributions to it; tiles t0 .. t0 + nt - 1.
__global__ void __launch_bounds__(256) k_bin_count(const uint64_t *__restrict__ tz_beg,
                                                   const uint32_t *__restrict__ tz_rc, uint64_t t0, uint64_t nt,
                                                   uint64_t e0, uint64_t e1, int directed, uint32_t *cnt) {
    for (uint64_t b = blockIdx.x * 256ull; b < nt; b += (uint64_t)gridDim.x * 256) {   // wave-uniform bound
        const uint64_t i = b + threadIdx.x;
        uint64_t first = 0;
        const uint32_t n = i < nt ? piece_tile(tz_beg, t0 + i, e0, e1, first) : 0u;
        const uint32_t rc = i < nt ? tz_rc[t0 + i] : 0u;
        // a wave's tiles span a few tile rows but many columns: rows combined, columns one atomic each
        if (n) atomicAdd(&cnt[rc & 0xFFFFu], n);
        if (!directed) (void)wave_reserve(cnt, rc >> 16, n, n != 0);
    }
}

// off[0 .. nb] = exclusive prefix of cnt (off[nb] = the piece's contributions);
// cur = off[0 .. nb) (the reservation cursors); uoff[0 .. nb] = exclusive
// prefix of the bins' work units (a non-empty bin has ceil(cnt / BIN_SLICE)).
// One workgroup.
constexpr uint32_t BIN_SLICE = 1u << 18;   // contributions per work unit of k_bin_acc at most

__device__ inline uint32_t bin_units(uint32_t c) { return (c + BIN_SLICE - 1) / BIN_SLICE; }

__global__ void __launch_bounds__(1024) k_bin_scan(const uint32_t *__restrict__ cnt, uint32_t nb, uint32_t *off,
                                                   uint32_t *cur, uint32_t *uoff) {
    __shared__ uint32_t part[1024], upart[1024];
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b0 = min(threadIdx.x * per, nb), b1 = min(b0 + per, nb);
    uint32_t t = 0, tu = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        t += cnt[b];
        tu += bin_units(cnt[b]);
    }
    part[threadIdx.x] = t;
    upart[threadIdx.x] = tu;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t y = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
        const uint32_t yu = (int)threadIdx.x >= o ? upart[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += y;
        upart[threadIdx.x] += yu;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - t, urun = upart[threadIdx.x] - tu;
    for (uint32_t b = b0; b < b1; ++b) {
        off[b] = run;
        cur[b] = run;
        uoff[b] = urun;
        run += cnt[b];
        urun += bin_units(cnt[b]);
    }
    if (threadIdx.x == 1023) {
        off[nb] = part[1023];
        uoff[nb] = upart[1023];
    }
}

// Per tile of the piece: where its contributions start in the column bin
// (pc) and the row bin (pr), as position minus the tile's first edge of the
// piece, so an edge's position is pc + its global edge index.
__global__ void __launch_bounds__(256) k_bin_pos(const uint64_t *__restrict__ tz_beg,
                                                 const uint32_t *__restrict__ tz_rc, uint64_t t0, uint64_t nt,
                                                 uint64_t e0, uint64_t e1, int directed, uint32_t *cur, int64_t *pc,
                                                 int64_t *pr) {
    for (uint64_t b = blockIdx.x * 256ull; b < nt; b += (uint64_t)gridDim.x * 256) {
        const uint64_t i = b + threadIdx.x;
        uint64_t first = 0;
        const uint32_t n = i < nt ? piece_tile(tz_beg, t0 + i, e0, e1, first) : 0u;
        const uint32_t rc = i < nt ? tz_rc[t0 + i] : 0u;
        if (n) pc[i] = (int64_t)atomicAdd(&cur[rc & 0xFFFFu], n) - (int64_t)first;
        if (!directed) {
            const uint32_t r = wave_reserve(cur, rc >> 16, n, n != 0);
            if (n) pr[i] = (int64_t)r - (int64_t)first;
        }
    }
}

// Every edge (v0, v1) of the piece's chunks: (v1 mod 2^16, prior[v0]) into
// v1's column bin and, undirected, (v0 mod 2^16, prior[v1]) into v0's row bin.
template <bool DIRECTED>
__global__ void __launch_bounds__(SWEEP_WG) k_bin_write(SweepArgs a, uint64_t t0, const int64_t *__restrict__ pc,
                                                        const int64_t *__restrict__ pr,
                                                        const float *__restrict__ prior, uint16_t *bo, float *bv) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    const uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    const uint64_t cbase = chunk * (uint64_t)a.chunk, tb = a.chunk_tz[chunk] - t0;
    const uint32_t *g = a.grid + (chunk - a.grid_chunk0) * (uint64_t)a.chunk;
    const int quads = (int)(a.chunk / (4 * SWEEP_WG));
#pragma unroll 1
    for (int q = 0; q < quads; ++q) {
        const uint32_t off = (uint32_t)(q * SWEEP_WG + threadIdx.x) * 4u;
        if (off >= len) break;
        int k = find_tile(s, off);
        uint32_t w[4];
        if (off + 4 <= len) {
            const uint4 v = *reinterpret_cast<const uint4 *>(g + off);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            for (int i = 0; i < 4; ++i) w[i] = (off + i < len) ? g[off + i] : 0u;
        }
        int kl = -1;
        int64_t bc = 0, br = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t e = off + i;
            if (e >= len) break;
            while (s.beg[k + 1] <= e) ++k;
            if (k != kl) {
                bc = pc[tb + k];
                if (!DIRECTED) br = pr[tb + k];
                kl = k;
            }
            const uint32_t rc = s.rc[k], v0 = gv0(w[i], rc), v1 = gv1(w[i], rc);
            const int64_t ge = (int64_t)(cbase + e);
            const uint64_t p1 = (uint64_t)(bc + ge);
            bo[p1] = (uint16_t)(v1 & 0xFFFFu);
            bv[p1] = unmark(prior[v0]);
            if (!DIRECTED) {
                const uint64_t p0 = (uint64_t)(br + ge);
                bo[p0] = (uint16_t)(v0 & 0xFFFFu);
                bv[p0] = unmark(prior[v1]);
            }
        }
    }
}

__device__ inline void lds_add(float *p, float x) {
    unsigned *a = reinterpret_cast<unsigned *>(p);
    unsigned old = *a, assumed;
    do {
        assumed = old;
        old = atomicCAS(a, assumed, __float_as_uint(__uint_as_float(assumed) + x));
    } while (old != assumed);
}
__device__ inline void lds_add(double *p, double x) {
    unsigned long long *a = reinterpret_cast<unsigned long long *>(p);
    unsigned long long old = *a, assumed;
    do {
        assumed = old;
        old = atomicCAS(a, assumed, (unsigned long long)__double_as_longlong(__longlong_as_double((long long)assumed) + x));
    } while (old != assumed);
}

// Workgroup (unit, part): a unit is a slice of at most BIN_SLICE of one bin's
// contributions; the slice's contributions to its part of the destinations
// are summed in LDS (eight loads in flight per thread), then added into acc:
// with plain loads and stores when the bin is one slice (no other workgroup
// touches those destinations in this step), else with atomics for the
// non-zero sums (the few bins of a segment's own tile rows).  Workgroups
// past the piece's units leave at once.
template <class A>
__global__ void __launch_bounds__(1024) k_bin_acc(const uint32_t *__restrict__ off, const uint32_t *__restrict__ uoff,
                                                  uint32_t nb, const uint16_t *__restrict__ bo,
                                                  const float *__restrict__ bv, uint64_t n, A *acc) {
    constexpr uint32_t SH = sizeof(A) == 4 ? 15 : 14, PS = 1u << SH, PARTS = 1u << (16 - SH);
    __shared__ A sum[PS];
    const uint32_t u = blockIdx.x / PARTS, part = blockIdx.x % PARTS;
    if (u >= uoff[nb]) return;   // the same in every thread of the workgroup
    uint32_t lo = 0, hi = nb - 1;   // the bin: largest b with uoff[b] <= u
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (uoff[mid] <= u) lo = mid; else hi = mid - 1;
    }
    const uint32_t bin = lo, slice = u - uoff[bin], whole = uoff[bin + 1] - uoff[bin] == 1;
    const uint32_t i0 = off[bin] + slice * BIN_SLICE, i1 = min(off[bin + 1], i0 + BIN_SLICE);
    const uint64_t vbase = ((uint64_t)bin << 16) + (uint64_t)part * PS;
    if (vbase >= n) return;
    for (uint32_t i = threadIdx.x; i < PS; i += 1024) sum[i] = A(0);
    __syncthreads();
    for (uint32_t i = i0 + threadIdx.x; i < i1; i += 8 * 1024) {
        uint32_t o[8];
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t x = i + j * 1024;
            o[j] = x < i1 ? (uint32_t)bo[x] : 0xFFFFFFFFu;
            v[j] = x < i1 ? bv[x] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((o[j] >> SH) == part) atomicAdd(&sum[o[j] & (PS - 1)], (A)v[j]);
    }
    __syncthreads();
    const uint32_t m = (uint32_t)min<uint64_t>(PS, n - vbase);
    for (uint32_t j = threadIdx.x; j < m; j += 1024) {
        const A x = sum[j];
        if (x == A(0)) continue;
        if (whole) acc[vbase + j] += x;
        else unsafeAtomicAdd(&acc[vbase + j], x);
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
static bool use_binned(const gs_graph *g) { return g->pr_method >= 1 && (g->directed || g->ss); }

// Chunks per piece of the binned scatter: a streamed segment, else 2^25 edges.
static uint64_t bin_piece(const gs_graph *g) {
    return g->ss ? g->ss->seg_chunks : std::max<uint64_t>(1, (1ull << 25) / g->chunk);
}

// The binned scatter's host copy of chunk_tz and its work buffers, sized for
// the largest piece the sweep hands it (the resident prefix and the segments
// are cut at the same piece boundaries).
static int ensure_bins(gs_graph *g, gs_ws *w) {
    const uint64_t piece = bin_piece(g);
    {
        std::lock_guard<std::mutex> lk(g->build_mu);
        if (g->h_chunk_tz_of != g->chunk || g->h_chunk_tz.size() != g->nchunk + 1) {
            g->h_chunk_tz.resize(g->nchunk + 1);
            GS_HIP(hipMemcpy(g->h_chunk_tz.data(), g->chunk_tz.ptr, (g->nchunk + 1) * 4, hipMemcpyDeviceToHost));
            g->h_chunk_tz_of = g->chunk;
        }
    }
    uint64_t tiles = 1;
    auto span = [&](uint64_t c0, uint64_t c1) {
        for (uint64_t c = c0; c < c1; c += piece) {
            const uint64_t ce = std::min(c1, c + piece);
            tiles = std::max<uint64_t>(tiles, (uint64_t)g->h_chunk_tz[ce] - g->h_chunk_tz[c] + 1);
        }
    };
    const uint64_t res = g->ss ? g->ss->res_chunks : g->nchunk;
    span(0, res);
    span(res, g->nchunk);
    const uint64_t contrib = (g->directed ? 1 : 2) * std::min<uint64_t>(piece * g->chunk, g->E);
    if (contrib >= (1ull << 32)) {
        set_error("binned scatter: %llu contributions per piece exceed 2^32", (unsigned long long)contrib);
        return GS_ERR_UNSUPPORTED;
    }
    const uint64_t nb = g->p_s;
    if (w->bin_cnt.n != nb) {
        GS_TRY(w->bin_cnt.alloc(nb));
        GS_TRY(w->bin_off.alloc(nb + 1));
        GS_TRY(w->bin_cur.alloc(nb));
        GS_TRY(w->bin_uoff.alloc(nb + 1));
    }
    if (w->bin_pc.n < tiles) {
        GS_TRY(w->bin_pc.alloc(tiles));
        if (!g->directed) GS_TRY(w->bin_pr.alloc(tiles));
    }
    if (w->bin_o.n < contrib) {
        GS_TRY(w->bin_o.alloc(contrib));
        GS_TRY(w->bin_v.alloc(contrib));
    }
    return GS_OK;
}

template <class A>
static int scatter_binned(const gs_graph *gc, gs_ws *w, A *acc) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    GS_TRY(ensure_bins(g, w));
    const uint64_t piece = bin_piece(g), nb = g->p_s;
    constexpr uint32_t PARTS = sizeof(A) == 4 ? 2 : 4;
    hipError_t err = hipSuccess;
    GS_TRY(sweep(g, w, GS_K_PR_SCATTER, [&](const SweepArgs &a, unsigned, hipStream_t st) {
        for (uint64_t c = a.chunk0; c < a.chunk0 + a.nlaunch; c += piece) {
            const uint64_t cn = std::min(piece, a.chunk0 + a.nlaunch - c);
            const uint64_t e0 = c * g->chunk, e1 = std::min(g->E, (c + cn) * g->chunk);
            const uint64_t t0 = g->h_chunk_tz[c], nt = (uint64_t)g->h_chunk_tz[c + cn] - t0 + 1;
            const unsigned tg = (unsigned)std::min<uint64_t>((nt + 255) / 256, 8192);
            if (err == hipSuccess) err = hipMemsetAsync(w->bin_cnt.ptr, 0, nb * 4, st);
            k_bin_count<<<tg, 256, 0, st>>>(a.tz_beg, a.tz_rc, t0, nt, e0, e1, (int)g->directed, w->bin_cnt.ptr);
            k_bin_scan<<<1, 1024, 0, st>>>(w->bin_cnt.ptr, (uint32_t)nb, w->bin_off.ptr, w->bin_cur.ptr,
                                           w->bin_uoff.ptr);
            k_bin_pos<<<tg, 256, 0, st>>>(a.tz_beg, a.tz_rc, t0, nt, e0, e1, (int)g->directed, w->bin_cur.ptr,
                                          w->bin_pc.ptr, w->bin_pr.ptr);
            SweepArgs pa = a;
            pa.chunk0 = c;
            pa.nlaunch = cn;
            const unsigned wg = (unsigned)(g->xcd_order ? xcd_grid(cn) : cn);
            if (g->directed)
                k_bin_write<true><<<wg, SWEEP_WG, 0, st>>>(pa, t0, w->bin_pc.ptr, w->bin_pr.ptr, w->pr_prior.ptr,
                                                           w->bin_o.ptr, w->bin_v.ptr);
            else
                k_bin_write<false><<<wg, SWEEP_WG, 0, st>>>(pa, t0, w->bin_pc.ptr, w->bin_pr.ptr, w->pr_prior.ptr,
                                                            w->bin_o.ptr, w->bin_v.ptr);
            // units: at most one per bin plus one per BIN_SLICE contributions
            const uint64_t units = nb + (2 * (e1 - e0) + BIN_SLICE - 1) / BIN_SLICE;
            k_bin_acc<A><<<(unsigned)(units * PARTS), 1024, 0, st>>>(w->bin_off.ptr, w->bin_uoff.ptr, (uint32_t)nb,
                                                                     w->bin_o.ptr, w->bin_v.ptr, g->n, acc);
        }
    }, false, true));
    if (err != hipSuccess) {
        set_error("binned scatter: %s", hipGetErrorString(err));
        return GS_ERR_DEVICE;
    }
    return GS_OK;
}

// One accumulation pass: pull over the neighbour lists when the graph has
// them, else the tile scatter (with the heavy-vertex kernel when there is a
// heavy set).
template <class A>
static int scatter_t(const gs_graph *g, gs_ws *w, A *acc) {
    if (use_binned(g)) return scatter_binned(g, w, acc);
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
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    // the pull works in the index's relabelled ids; results are mapped back at the end
    const bool pull = use_pull(g);
    // the blocked passes write the new priors of the vertices with edges themselves and
    // read no other; a vertex without edges ends at 0 (its factor is 0), which the
    // unpermute below gives it without a value in the arrays
    const bool blocked = pull && g->pr_method == 2 && g->pb_k;
    const uint64_t ninit = blocked && iters > 0 ? g->nzn : g->n;
    k_pr_init<<<vgrid(ninit), 256, 0, w->stream>>>(pull ? g->deg_r.ptr : g->degree.ptr, ninit, mode,
                                                   (float)(1.0 / (double)g->n), iters ? g->heavy_dthr : 0xFFFFFFFFu,
                                                   w->pr_prior.ptr, w->pr_acc.ptr, w->pr_fac.ptr);
    GS_HIP(hipGetLastError());
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
    if (pull) {  // prior (relabelled) -> acc (original order), then the two trade places
        GS_TRY(unpermute_f32(g, w->stream, w->pr_prior.ptr, w->pr_acc.ptr, ninit));
        std::swap(w->pr_prior.ptr, w->pr_acc.ptr);   // both hold n values
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
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
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
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
    if (pull) {  // rank (relabelled) -> acc (original order), then the two trade places
        GS_TRY(unpermute_f32(g, w->stream, w->ppr_rank.ptr, w->pr_acc.ptr, g->n));
        std::swap(w->ppr_rank.ptr, w->pr_acc.ptr);   // both hold n values
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
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
