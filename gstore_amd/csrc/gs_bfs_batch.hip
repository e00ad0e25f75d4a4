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
#include <algorithm>
#include <vector>

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
    // frontier masks of both endpoints of the quad's four edges load together
    auto pre = [&](uint32_t w, uint32_t rc) {
        ulonglong2 f;
        f.x = front[gv0(w, rc)];
        f.y = DIRECTED ? 0ull : front[gv1(w, rc)];
        return f;
    };
    auto op = [&](uint32_t w, uint32_t rc, ulonglong2 f) {
        if (f.x) push(f.x, gv1(w, rc));
        if (!DIRECTED && f.y) push(f.y, gv0(w, rc));
    };
    sweep_chunk_pre<true>(a, chunk, s, len, pre, op);
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

// ---- multi-source BFS over the neighbour-list index (relabelled ids) -----------
// The same per-lane rule as the tile sweep above, with Beamer's direction
// switch on the union frontier: top-down levels push each frontier vertex's
// lane mask along its list (balanced over edges, as the single-root top-down);
// bottom-up levels give every vertex that still misses lanes a thread that ORs
// its neighbours' frontier masks, heaviest neighbours first, and stops once all
// missing lanes are found.  Depths, when kept, are written vertex-major (one
// 64-byte row per vertex) and mapped back to vertex order at the end.
constexpr uint32_t MS_CHUNK = 2048;

// One wave.  Roots without edges go to the end of the first queue: zero-width
// entries there are never inside a top-down chunk's range, so a chunk stages
// at most one entry per edge position (plus one).  Depth rows exist only for
// the vertices with edges (ids below nzn); the lanes of roots without edges
// get their depth in k_ms_unpermute.  Their mask words (above nzn, never
// cleared) are read only as a queue entry with an empty list.
__global__ void __launch_bounds__(64) k_ms_roots(const uint64_t *__restrict__ roots, int nr,
                                                 const uint32_t *__restrict__ new_id, const int32_t *__restrict__ deg,
                                                 uint64_t nzn, unsigned long long *vis, unsigned long long *front,
                                                 uint32_t *queue, int32_t *qdeg, uint8_t *depth_vm) {
    const int i = threadIdx.x;
    const bool in = i < nr;
    uint32_t r = 0;
    int32_t d = 0;
    if (in) {
        r = new_id[roots[i]];
        d = deg[r];
        atomicOr(&vis[r], 1ull << i);
        atomicOr(&front[r], 1ull << i);
        if (depth_vm && r < nzn) depth_vm[(uint64_t)r * 64 + i] = 1;
    }
    const unsigned long long with_edges = __ballot(in && d > 0), without = __ballot(in && d <= 0);
    if (!in) return;
    const unsigned long long below = (1ull << i) - 1;
    const int pos = d > 0 ? __popcll(with_edges & below) : __popcll(with_edges) + __popcll(without & below);
    queue[pos] = r;
    qdeg[pos] = d;
}

__global__ void __launch_bounds__(256) k_ms_td(const uint32_t *__restrict__ queue, const uint64_t *__restrict__ qoff,
                                               uint64_t nq, uint64_t m, const uint64_t *__restrict__ off,
                                               const uint32_t *__restrict__ nbr, const unsigned long long *__restrict__ front,
                                               const unsigned long long *__restrict__ vis, unsigned long long *next) {
    __shared__ uint32_t sbeg[MS_CHUNK + 2];
    __shared__ uint64_t sadj[MS_CHUNK + 1];
    __shared__ unsigned long long sfr[MS_CHUNK + 1];
    __shared__ uint64_t j_first;
    __shared__ int nstaged;
    const uint64_t c0 = blockIdx.x * (uint64_t)MS_CHUNK;
    if (c0 >= m) return;
    const uint32_t len = (uint32_t)min<uint64_t>(MS_CHUNK, m - c0);
    if (threadIdx.x < 64) {   // the first wave searches the queue prefix (wave_upper)
        const uint64_t ja = wave_upper(qoff, 0, nq - 1, c0);
        const uint64_t jb = wave_upper(qoff, ja, nq - 1, c0 + len - 1);
        if (threadIdx.x == 0) {
            j_first = ja;
            nstaged = (int)(jb - ja + 1);
        }
    }
    __syncthreads();
    const uint64_t ja = j_first;
    const int ns = nstaged;
    for (int k = threadIdx.x; k < ns; k += blockDim.x) {
        const uint64_t q = qoff[ja + k];
        const uint32_t v = queue[ja + k];
        sbeg[k] = q <= c0 ? 0u : (uint32_t)(q - c0);
        sadj[k] = off[v] + (q < c0 ? c0 - q : 0);
        sfr[k] = front[v];
    }
    if (threadIdx.x == 0) sbeg[ns] = len;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) {
        int lo = 0, hi = ns - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (sbeg[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const uint32_t u = nbr[sadj[lo] + (t - sbeg[lo])];
        const unsigned long long mm = sfr[lo] & ~vis[u];
        if (mm && (mm & ~next[u])) atomicOr(&next[u], mm);
    }
}

// Bottom-up, one thread per vertex of ids [v0, v1) (degree < MS_HEAVY_DEG):
// OR the neighbours' frontier masks, the first list entry from first[] (no
// offset loads), then four at a time, until every lane of `all` the vertex
// misses is found.  The finds are or-ed into next: the top-down
// kernel of the same level, run before, owns the other lanes' bits.
__device__ inline void ms_bu_light(uint64_t v0, uint64_t v1, unsigned long long all, const uint64_t *__restrict__ off,
                                   const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ first,
                                   const unsigned long long *__restrict__ vis,
                                   const unsigned long long *__restrict__ front, unsigned long long *next,
                                   unsigned long long *checked) {
    uint32_t ck = 0;
    for (uint64_t v = v0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < v1; v += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long need = all & ~vis[v];
        if (!need) continue;
        // the first list entry (the heaviest neighbour) without the offsets; the rest if lanes remain
        unsigned long long acc = front[first[v]] & need;
        ++ck;
        if (acc == need) {
            next[v] |= acc;
            continue;
        }
        const uint64_t b = off[v] + 1, e = off[v + 1];
        for (uint64_t k = b; k < e && acc != need; k += 4) {
            uint32_t u[4];
            unsigned long long f[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) u[j] = nbr[min(k + j, e - 1)];
#pragma unroll
            for (int j = 0; j < 4; ++j) f[j] = front[u[j]];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (k + j < e) acc |= f[j] & need;
            ck += (uint32_t)min<uint64_t>(4, e - k);
        }
        if (acc) next[v] |= acc;   // a top-down kernel of the same level may have set other lanes' bits
    }
    block_count(ck, checked);
}

// Bottom-up for the heavy vertices (ids [0, nheavy), degree >= MS_HEAVY_DEG):
// one wave per vertex, lanes stride its list 256 entries per step, and the
// wave ORs its lanes' finds after each step, stopping once every missing lane
// is found.  A thread per vertex here would walk a hub's whole list alone
// while the rest of the grid waits.
__device__ inline void ms_bu_heavy(uint64_t nheavy, unsigned long long all, const uint64_t *__restrict__ off,
                                   const uint32_t *__restrict__ nbr, const unsigned long long *__restrict__ vis,
                                   const unsigned long long *__restrict__ front, unsigned long long *next,
                                   unsigned long long *checked) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long ck = 0;
    for (uint64_t v = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; v < nheavy; v += nw) {
        const unsigned long long need = all & ~vis[v];
        if (!need) continue;
        const uint64_t b = off[v], e = off[v + 1];
        unsigned long long acc = 0;
        for (uint64_t k0 = b; k0 < e; k0 += 256) {
            uint32_t u[4];
            unsigned long long f[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) u[j] = nbr[min(k0 + j * 64 + lane, e - 1)];
#pragma unroll
            for (int j = 0; j < 4; ++j) f[j] = front[u[j]];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (k0 + j * 64 + lane < e) acc |= f[j] & need;
            ck += min<uint64_t>(256, e - k0);
            acc = wave_or64(acc);
            if (acc == need) break;
        }
        if (lane == 0 && acc) next[v] |= acc;
    }
    if (lane == 0 && ck) atomicAdd(checked, ck);
}

__global__ void __launch_bounds__(256) k_ms_bu(uint64_t v0, uint64_t nzn, unsigned long long all,
                                               const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                               const uint32_t *__restrict__ first,
                                               const unsigned long long *__restrict__ vis,
                                               const unsigned long long *__restrict__ front, unsigned long long *next,
                                               unsigned long long *checked) {
    ms_bu_light(v0, nzn, all, off, nbr, first, vis, front, next, checked);
}

__global__ void __launch_bounds__(256) k_ms_bu_heavy(uint64_t nheavy, unsigned long long all,
                                                     const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                     const unsigned long long *__restrict__ vis,
                                                     const unsigned long long *__restrict__ front,
                                                     unsigned long long *next, unsigned long long *checked) {
    ms_bu_heavy(nheavy, all, off, nbr, vis, front, next, checked);
}

// Level end over relabelled ids 0 .. nzn-1: next becomes the frontier; per-lane
// counts of newly reached vertices; frontier size and degree sum (count[0],
// count[1]) and the degree sum of vertices still missing lanes (count[3]).
// BUILD lists the new frontier (queue, qdeg; cursor count[2]), eight vertices
// per thread per step so a workgroup reserves its queue slots with one atomic
// per 2048 vertices (one per 256 had serialised ~470K atomics on the cursor
// at Kron-28: 5.7 ms for the level that lists a 100M-vertex frontier); DEPTH
// writes depth rows.
template <bool BUILD, bool DEPTH>
__device__ inline void ms_level_end(uint64_t nzn, unsigned long long all, unsigned long long *vis,
                                    unsigned long long *front, unsigned long long *next, const int32_t *__restrict__ deg,
                                    uint8_t *depth_vm, uint32_t newdepth, uint32_t *queue, int32_t *qdeg,
                                    unsigned long long *count, unsigned long long *lane_counts) {
    constexpr int VPT = BUILD ? 8 : 1;
    __shared__ uint32_t cnt[64];
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long bbase, bsum[3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
    if (threadIdx.x < 3) bsum[threadIdx.x] = 0;
    uint32_t my = 0;
    unsigned long long nf = 0, mf = 0, mu = 0;
    for (uint64_t base = blockIdx.x * (256ull * VPT); base < nzn; base += (uint64_t)gridDim.x * 256 * VPT) {
        unsigned long long nxs[VPT], vss[VPT];
        int32_t dvs[VPT];
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < VPT; ++k) {   // all loads first: the stores below may not pass them
            const uint64_t v = base + k * 256ull + threadIdx.x;
            nxs[k] = v < nzn ? next[v] : 0ull;
            vss[k] = v < nzn ? vis[v] : 0ull;
            dvs[k] = v < nzn ? deg[v] : 0;
        }
#pragma unroll
        for (int k = 0; k < VPT; ++k) {
            const uint64_t v = base + k * 256ull + threadIdx.x;
            const unsigned long long nx = nxs[k];
            const int32_t dv = dvs[k];
            if (v < nzn) {
                front[v] = nx;
                const unsigned long long vs = vss[k];
                if (nx) {
                    next[v] = 0;
                    vis[v] = vs | nx;
                    nf += 1;
                    mf += (uint32_t)dv;
                    if (DEPTH) {
                        unsigned long long mm = nx;
                        while (mm) {
                            const int l = __ffsll((long long)mm) - 1;
                            mm &= mm - 1;
                            depth_vm[v * 64 + l] = (uint8_t)newdepth;
                        }
                    }
                }
                if (((vs | nx) & all) != all) mu += (uint32_t)dv;
            }
            if (__ballot(nx != 0)) {
                for (int l = 0; l < 64; ++l) {
                    const uint32_t cl = (uint32_t)__popcll(__ballot((nx >> l) & 1ull));
                    if (lane == l) my += cl;
                }
            }
            c += nx != 0;
        }
        if (BUILD) {
            uint32_t incl = c;   // this thread's slots after the lanes below it
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            if (lane == 63) wsum[wv] = incl;
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint32_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
                bbase = t ? atomicAdd(&count[2], (unsigned long long)t) : 0ull;
            }
            __syncthreads();
            uint64_t pos = bbase + incl - c;
            for (int j = 0; j < wv; ++j) pos += wsum[j];
#pragma unroll
            for (int k = 0; k < VPT; ++k)
                if (nxs[k]) {
                    queue[pos] = (uint32_t)(base + k * 256ull + threadIdx.x);
                    qdeg[pos] = dvs[k];
                    ++pos;
                }
            __syncthreads();
        }
    }
    __syncthreads();
    if (my) atomicAdd(&cnt[lane], my);
    if (nf) atomicAdd(&bsum[0], nf);
    if (mf) atomicAdd(&bsum[1], mf);
    if (mu) atomicAdd(&bsum[2], mu);
    __syncthreads();
    if (threadIdx.x < 64 && cnt[threadIdx.x]) atomicAdd(&lane_counts[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (bsum[0]) atomicAdd(&count[0], bsum[0]);
        if (bsum[1]) atomicAdd(&count[1], bsum[1]);
        if (bsum[2]) atomicAdd(&count[3], bsum[2]);
    }
}

template <bool BUILD, bool DEPTH>
__global__ void __launch_bounds__(256) k_ms_level_end(uint64_t nzn, unsigned long long all, unsigned long long *vis,
                                                      unsigned long long *front, unsigned long long *next,
                                                      const int32_t *__restrict__ deg, uint8_t *depth_vm,
                                                      uint32_t newdepth, uint32_t *queue, int32_t *qdeg,
                                                      unsigned long long *count, unsigned long long *lane_counts) {
    ms_level_end<BUILD, DEPTH>(nzn, all, vis, front, next, deg, depth_vm, newdepth, queue, qdeg, count, lane_counts);
}

// Lists the current frontier (vertices with a non-zero mask) for a top-down level.
__global__ void __launch_bounds__(256) k_ms_list(uint64_t nzn, const unsigned long long *__restrict__ front,
                                                 const int32_t *__restrict__ deg, uint32_t *queue, int32_t *qdeg,
                                                 unsigned long long *count) {
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long bbase;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint64_t base = blockIdx.x * 256ull; base < nzn; base += (uint64_t)gridDim.x * 256) {
        const uint64_t v = base + threadIdx.x;
        const bool in = v < nzn && front[v] != 0;
        const unsigned long long mask = __ballot(in);
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
            bbase = t ? atomicAdd(&count[0], (unsigned long long)t) : 0ull;
        }
        __syncthreads();
        if (in) {
            uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1));
            for (int j = 0; j < wv; ++j) before += wsum[j];
            queue[bbase + before] = (uint32_t)v;
            qdeg[bbase + before] = deg[v];
        }
        __syncthreads();
    }
}

// depth (lane-major, vertex order) from the vertex-major rows of relabelled
// ids; a vertex without edges (no row) is unreached in every lane but its own
// root's, set afterwards by k_ms_root_depth.
__global__ void k_ms_unpermute(const uint8_t *__restrict__ depth_vm, const uint32_t *__restrict__ new_id, uint64_t n,
                               uint64_t nzn, int nr, uint8_t *depth) {
    for (uint64_t o = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; o < n; o += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = new_id[o];
        uint4 q[4] = {};
        if (r < nzn) {
            const uint4 *row = reinterpret_cast<const uint4 *>(depth_vm + (uint64_t)r * 64);
#pragma unroll
            for (int k = 0; k < 4; ++k) q[k] = row[k];
        }
        const uint8_t *b = reinterpret_cast<const uint8_t *>(q);
        for (int l = 0; l < nr; ++l) depth[(uint64_t)l * n + o] = b[l];
    }
}

// Depth 1 at the root of every lane whose root has no edges (one thread per lane).
__global__ void k_ms_root_depth(const uint64_t *__restrict__ roots, int nr, const uint32_t *__restrict__ new_id,
                                uint64_t n, uint64_t nzn, uint8_t *depth) {
    const int l = threadIdx.x;
    if (l < nr && new_id[roots[l]] >= nzn) depth[(uint64_t)l * n + roots[l]] = 1;
}

static unsigned ms_grid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }
// four waves per workgroup, one heavy vertex per wave at a time
static unsigned ms_heavy_grid(uint64_t nh) { return (unsigned)std::min<uint64_t>((nh + 3) / 4, 8192); }

// ---- the same levels, driven from the device (one hipGraph per workspace) -------
// The host loop above reads three counters after every level to choose one
// direction for the batch and to stop.  Here those decisions are kernels: per
// level, a one-wave kernel picks each lane's direction from its frontier count
// (k_msg_decide) and raises flags; the list, queue-prefix, top-down, bottom-up
// and level-end kernels return at once when their flag is down (or the search
// is done); a 64-thread kernel closes the level.  MSG_LEVELS levels are captured once into
// a hipGraph, which the host launches until the done flag is up, so a batch
// costs one host round trip per MSG_LEVELS levels instead of one per level.
// Control words, u64, in w->ms_ctl (the first four are the level-end counters
// in the order k_ms_level_end adds them):
enum : int {
    C_NF = 0,      // frontier vertices after the last level end
    C_MF = 1,      // their degree sum
    C_BUILD = 2,   // queue cursor of the level end
    C_MU = 3,      // degree sum of vertices still missing lanes
    C_CHECKED = 4, // entries examined, all levels
    C_M = 5,       // top-down: the queue's degree sum (edges this level)
    C_LEVEL = 6,   // current level (1 = roots)
    C_FLAGS = 7,
    C_ACTIVE = 8,  // lanes whose frontier is not empty
    C_ALL = 9,     // lanes in use
    C_NQ = 10,     // queue length this level
    C_BUL = 11,    // lanes searched bottom-up this level
    C_TDL = 12,    // lanes pushed top-down this level
    C_NLEV = 13,   // levels done
    C_LIST = 14,   // queue cursor of the frontier listing
    C_ALPHA = 15,
    C_BETA = 16,
    C_LANES = 18,  // per-level lane counts, 64 words per level
    C_WORDS = C_LANES + 64 * 256
};
enum : unsigned long long { F_BU = 1, F_HAVEQ = 2, F_DONE = 4, F_TD = 8, F_LIST = 16, F_OVERFLOW = 32 };
constexpr int MSG_LEVELS = 4;
constexpr unsigned MSG_TD_GRID = 2048;
constexpr unsigned MSG_GRID = 2048;
constexpr uint64_t MS_TD_DIV = 16384;   // in a bottom-up level, lanes with frontiers below nzn / 16384 stay top-down

__global__ void k_msg_init(unsigned long long *ctl, int nr, const int32_t *__restrict__ qdeg, uint64_t m_u,
                           uint32_t alpha, uint32_t beta) {
    if (threadIdx.x != 0) return;
    unsigned long long mf = 0;
    for (int i = 0; i < nr; ++i) mf += (uint32_t)qdeg[i];
    const unsigned long long all = nr == 64 ? ~0ull : ((1ull << nr) - 1);
    ctl[C_NF] = (unsigned long long)nr;
    ctl[C_MF] = mf;
    ctl[C_MU] = m_u;
    ctl[C_LEVEL] = 1;
    ctl[C_FLAGS] = F_HAVEQ;
    ctl[C_ACTIVE] = all;
    ctl[C_ALL] = all;
    ctl[C_ALPHA] = alpha;
    ctl[C_BETA] = beta;
}

// The level's direction.  For the batch as a whole Beamer's rule on the
// summed frontier (as the host loop).  When that says bottom-up, the lanes
// whose own frontiers are still small (below nzn / MS_TD_DIV: a lane's first
// levels or its tail) are pushed top-down in the same level instead: searched
// bottom-up, every vertex they miss would scan its whole list for them, where
// the other lanes' finds stop the scan at the first neighbours.  A 64-root
// batch took 140 -> 80 ms at Kron-28, 8.9 -> 7.7 ms at Kron-24 and 2.9 -> 2.5
// ms at Kron-20 (MS_TD_DIV from a sweep of 2^10 .. 2^18 at Kron-26 and -28).
// One wave, a lane per BFS lane.
__global__ void __launch_bounds__(64) k_msg_decide(unsigned long long *ctl, uint64_t nzn) {
    const unsigned long long f = ctl[C_FLAGS];
    if (f & F_DONE) return;
    const uint64_t level = ctl[C_LEVEL];
    const unsigned long long act = ctl[C_ACTIVE], mf = ctl[C_MF], mu = ctl[C_MU];
    bool bu = f & F_BU;
    if (!bu && mf > mu / ctl[C_ALPHA]) bu = true;
    else if (bu && ctl[C_NF] < nzn / ctl[C_BETA]) bu = false;
    const int l = threadIdx.x;
    // the frontier of lane l: its count of vertices reached by the previous level (the root before)
    const unsigned long long nfl = level > 1 ? ctl[C_LANES + 64ull * (level - 2) + l] : 1ull;
    const unsigned long long bul = __ballot(bu && ((act >> l) & 1ull) && nfl * MS_TD_DIV >= nzn);
    if (l != 0) return;
    const unsigned long long tdl = act & ~bul, nf = ctl[C_NF];
    const bool td = tdl != 0 && nf != 0;
    ctl[C_BUL] = bul;
    ctl[C_TDL] = tdl;
    ctl[C_NQ] = nf;   // the queue the last level end wrote (all frontier vertices), if it wrote one
    ctl[C_NF] = ctl[C_MF] = ctl[C_MU] = 0;
    ctl[C_BUILD] = ctl[C_LIST] = ctl[C_M] = 0;
    // F_BU also when no lane is left for it: the batch stays in its bottom-up phase.  After a
    // top-down level the level end lists the new frontier (F_HAVEQ), which serves a top-down
    // level of every lane; after a bottom-up level, or when some lanes go bottom-up (their
    // frontiers are the big ones), the top-down lanes' frontier is listed first (F_LIST).
    const bool list = td && (!(f & F_HAVEQ) || bul);
    ctl[C_FLAGS] = (bu ? F_BU : F_HAVEQ) | (td ? F_TD : 0) | (list ? F_LIST : 0);
}

// Queue length of a top-down level: the roots, or the listed frontier.
__device__ inline uint64_t msg_nq(const unsigned long long *ctl) {
    return (ctl[C_FLAGS] & F_LIST) ? ctl[C_LIST] : ctl[C_NQ];
}

__global__ void __launch_bounds__(256) k_msg_list(const unsigned long long *ctl, uint64_t nzn,
                                                  const unsigned long long *__restrict__ front,
                                                  const int32_t *__restrict__ deg, uint32_t *queue, int32_t *qdeg,
                                                  unsigned long long *cursor) {
    if ((ctl[C_FLAGS] & (F_LIST | F_DONE)) != F_LIST) return;
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long bbase;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long tdl = ctl[C_TDL];
    for (uint64_t base = blockIdx.x * 256ull; base < nzn; base += (uint64_t)gridDim.x * 256) {
        const uint64_t v = base + threadIdx.x;
        const bool in = v < nzn && (front[v] & tdl) != 0;
        const unsigned long long mask = __ballot(in);
        if (lane == 0) wsum[wv] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
            bbase = t ? atomicAdd(cursor, (unsigned long long)t) : 0ull;
        }
        __syncthreads();
        if (in) {
            uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1));
            for (int j = 0; j < wv; ++j) before += wsum[j];
            queue[bbase + before] = (uint32_t)v;
            qdeg[bbase + before] = deg[v];
        }
        __syncthreads();
    }
}

// qsum[b] = degree sum of queue entries [1024 b, 1024 b + 1024)
__global__ void __launch_bounds__(256) k_msg_block_sums(const unsigned long long *ctl, const int32_t *__restrict__ qdeg,
                                                        unsigned long long *qsum) {
    if ((ctl[C_FLAGS] & (F_TD | F_DONE)) != F_TD) return;
    const uint64_t nq = msg_nq(ctl);
    __shared__ unsigned long long s;
    for (uint64_t base = (uint64_t)blockIdx.x * 1024; base < nq; base += (uint64_t)gridDim.x * 1024) {
        if (threadIdx.x == 0) s = 0;
        __syncthreads();
        unsigned long long c = 0;
        for (int i = threadIdx.x; i < 1024; i += blockDim.x)
            if (base + i < nq) c += (uint32_t)qdeg[base + i];
        atomicAdd(&s, c);
        __syncthreads();
        if (threadIdx.x == 0) qsum[base / 1024] = s;
        __syncthreads();
    }
}

// Exclusive prefix of qsum[0 .. nb) in place (one workgroup); the total is the
// level's top-down edge count.
__global__ void __launch_bounds__(1024) k_msg_scan(unsigned long long *ctl, unsigned long long *qsum) {
    if ((ctl[C_FLAGS] & (F_TD | F_DONE)) != F_TD) return;
    __shared__ unsigned long long part[1024];
    const uint64_t nb = (msg_nq(ctl) + 1023) / 1024;
    const uint64_t per = (nb + 1023) / 1024, b0 = threadIdx.x * per, b1 = std::min<uint64_t>(b0 + per, nb);
    unsigned long long s = 0;
    for (uint64_t b = b0; b < b1; ++b) s += qsum[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int step = 1; step < 1024; step <<= 1) {
        const unsigned long long v = threadIdx.x >= (unsigned)step ? part[threadIdx.x - step] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = part[threadIdx.x] - s;   // exclusive
    for (uint64_t b = b0; b < b1; ++b) {
        const unsigned long long c = qsum[b];
        qsum[b] = run;
        run += c;
    }
    if (threadIdx.x == 1023) ctl[C_M] = part[1023];
}

// qoff[j] = exclusive prefix of qdeg over the queue (blocks of 1024, as k_deg_scan)
__global__ void __launch_bounds__(256) k_msg_deg_scan(const unsigned long long *ctl, const int32_t *__restrict__ qdeg,
                                                      const unsigned long long *__restrict__ bstart, uint64_t *qoff) {
    if ((ctl[C_FLAGS] & (F_TD | F_DONE)) != F_TD) return;
    const uint64_t nq = msg_nq(ctl);
    __shared__ uint32_t d[1024];
    __shared__ unsigned long long part[256];
    for (uint64_t base = (uint64_t)blockIdx.x * 1024; base < nq; base += (uint64_t)gridDim.x * 1024) {
        for (int i = threadIdx.x; i < 1024; i += blockDim.x) d[i] = base + i < nq ? (uint32_t)qdeg[base + i] : 0;
        __syncthreads();
        unsigned long long loc[4], run = 0;
        for (int j = 0; j < 4; ++j) {
            loc[j] = run;
            run += d[threadIdx.x * 4 + j];
        }
        part[threadIdx.x] = run;
        __syncthreads();
        for (int step = 1; step < 256; step <<= 1) {
            const unsigned long long v = threadIdx.x >= (unsigned)step ? part[threadIdx.x - step] : 0;
            __syncthreads();
            part[threadIdx.x] += v;
            __syncthreads();
        }
        const unsigned long long before = (threadIdx.x ? part[threadIdx.x - 1] : 0) + bstart[base / 1024];
        for (int j = 0; j < 4; ++j) {
            const uint64_t v = base + threadIdx.x * 4 + j;
            if (v < nq) qoff[v] = before + loc[j];
        }
        __syncthreads();   // d and part are rewritten by the next step
    }
}

// Top-down over the queue's edges, MS_CHUNK per workgroup step (grid-stride:
// the grid is fixed when the level sequence is captured).
__global__ void __launch_bounds__(256) k_msg_td(const unsigned long long *ctl, const uint32_t *__restrict__ queue,
                                                const uint64_t *__restrict__ qoff, const uint64_t *__restrict__ off,
                                                const uint32_t *__restrict__ nbr,
                                                const unsigned long long *__restrict__ front,
                                                const unsigned long long *__restrict__ vis, unsigned long long *next) {
    if ((ctl[C_FLAGS] & (F_TD | F_DONE)) != F_TD) return;
    __shared__ uint32_t sbeg[MS_CHUNK + 2];
    __shared__ uint64_t sadj[MS_CHUNK + 1];
    __shared__ unsigned long long sfr[MS_CHUNK + 1];
    __shared__ uint64_t j_first;
    __shared__ int nstaged;
    const uint64_t nq = msg_nq(ctl), m = ctl[C_M];
    const unsigned long long tdl = ctl[C_TDL];
    for (uint64_t c0 = blockIdx.x * (uint64_t)MS_CHUNK; c0 < m; c0 += (uint64_t)gridDim.x * MS_CHUNK) {
        const uint32_t len = (uint32_t)min<uint64_t>(MS_CHUNK, m - c0);
        if (threadIdx.x < 64) {   // the first wave searches the queue prefix (wave_upper)
            const uint64_t ja = wave_upper(qoff, 0, nq - 1, c0);
            const uint64_t jb = wave_upper(qoff, ja, nq - 1, c0 + len - 1);
            if (threadIdx.x == 0) {
                j_first = ja;
                nstaged = (int)(jb - ja + 1);
            }
        }
        __syncthreads();
        const uint64_t ja = j_first;
        const int ns = nstaged;
        for (int k = threadIdx.x; k < ns; k += blockDim.x) {
            const uint64_t q = qoff[ja + k];
            const uint32_t v = queue[ja + k];
            sbeg[k] = q <= c0 ? 0u : (uint32_t)(q - c0);
            sadj[k] = off[v] + (q < c0 ? c0 - q : 0);
            sfr[k] = front[v] & tdl;   // the top-down lanes only
        }
        if (threadIdx.x == 0) sbeg[ns] = len;
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) {
            int lo = 0, hi = ns - 1;
            while (lo < hi) {
                int mid = (lo + hi + 1) >> 1;
                if (sbeg[mid] <= t) lo = mid; else hi = mid - 1;
            }
            const uint32_t u = nbr[sadj[lo] + (t - sbeg[lo])];
            const unsigned long long mm = sfr[lo] & ~vis[u];
            if (mm && (mm & ~next[u])) atomicOr(&next[u], mm);
        }
        __syncthreads();   // the staged tables are rewritten by the next step
    }
}

template <bool HEAVY>
__global__ void __launch_bounds__(256) k_msg_bu(unsigned long long *ctl, uint64_t v0, uint64_t v1,
                                                const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                const uint32_t *__restrict__ first,
                                                const unsigned long long *__restrict__ vis,
                                                const unsigned long long *__restrict__ front,
                                                unsigned long long *next) {
    if ((ctl[C_FLAGS] & (F_BU | F_DONE)) != F_BU || !ctl[C_BUL]) return;
    if (HEAVY) ms_bu_heavy(v1, ctl[C_BUL], off, nbr, vis, front, next, ctl + C_CHECKED);
    else ms_bu_light(v0, v1, ctl[C_BUL], off, nbr, first, vis, front, next, ctl + C_CHECKED);
}

template <bool DEPTH, bool BUILD>
__global__ void __launch_bounds__(256) k_msg_level_end(unsigned long long *ctl, uint64_t nzn, unsigned long long *vis,
                                                       unsigned long long *front, unsigned long long *next,
                                                       const int32_t *__restrict__ deg, uint8_t *depth_vm,
                                                       uint32_t *queue, int32_t *qdeg) {
    const unsigned long long f = ctl[C_FLAGS];
    // the listing and the plain level end are separate kernels (the listing
    // one needs more registers); the one that does not match the level returns
    if ((f & F_DONE) || ((f & F_HAVEQ) != 0) != BUILD) return;
    const uint32_t level = (uint32_t)ctl[C_LEVEL];
    const unsigned long long all = ctl[C_ACTIVE];   // as the host loop: the lanes still searching
    unsigned long long *lanes = ctl + C_LANES + 64ull * (level - 1);
    ms_level_end<BUILD, DEPTH>(nzn, all, vis, front, next, deg, depth_vm, level + 1, queue, qdeg, ctl, lanes);
}

// Closes a level: lanes still active, entries examined, and the stop test
// (no lane reached a new vertex, or the next depth would not fit in u8).
__global__ void __launch_bounds__(64) k_msg_finish(unsigned long long *ctl) {
    const unsigned long long f = ctl[C_FLAGS];
    if (f & F_DONE) return;
    const uint32_t level = (uint32_t)ctl[C_LEVEL];
    const unsigned long long c = ctl[C_LANES + 64ull * (level - 1) + threadIdx.x];
    const unsigned long long active = __ballot(c != 0);
    if (threadIdx.x != 0) return;
    ctl[C_ACTIVE] = active;
    if (f & F_TD) ctl[C_CHECKED] += ctl[C_M];
    ctl[C_NLEV] += 1;
    if (!active) ctl[C_FLAGS] = f | F_DONE;
    else if (level + 2 > 256) ctl[C_FLAGS] = f | F_DONE | F_OVERFLOW;
    else ctl[C_LEVEL] = level + 1;
}

}  // namespace gs

using namespace gs;

// Multi-source BFS over the index (undirected resident grids).  keep_depth: write
// the per-lane depths into w->ms_depth (lane-major, vertex order).
static int run_bfs_batch_index(const gs_graph *gc, gs_ws *w, const uint64_t *roots, int nr, uint64_t *level_counts,
                               int max_levels, int *levels, bool keep_depth) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    const uint64_t n = g->n, nzn = g->nzn;
    const unsigned long long all = nr == 64 ? ~0ull : ((1ull << nr) - 1);
    unsigned long long *vis = (unsigned long long *)w->ms_vis.ptr, *front = (unsigned long long *)w->ms_front.ptr,
                       *next = (unsigned long long *)w->ms_next.ptr;
    uint8_t *dvm = keep_depth ? w->ms_depth_vm.ptr : nullptr;
    GS_HIP(hipMemcpyAsync(w->ms_roots.ptr, roots, nr * 8, hipMemcpyHostToDevice, w->stream));
    // the vertices with edges only (the others keep no state; see k_ms_roots)
    GS_HIP(hipMemsetAsync(vis, 0, nzn * 8, w->stream));
    GS_HIP(hipMemsetAsync(front, 0, nzn * 8, w->stream));
    GS_HIP(hipMemsetAsync(next, 0, nzn * 8, w->stream));
    if (dvm) GS_HIP(hipMemsetAsync(dvm, 0, nzn * 64, w->stream));
    k_ms_roots<<<1, 64, 0, w->stream>>>(w->ms_roots.ptr, nr, g->new_id.ptr, g->deg_r.ptr, nzn, vis, front, w->queue.ptr,
                                        w->qdeg.ptr, dvm);
    GS_HIP(hipGetLastError());
    std::vector<unsigned long long> lc(64), scratch;
    uint64_t n_f = (uint64_t)nr, m_u = 2 * g->E, checked = 0;
    bool bottom_up = false, have_queue = true;
    int nlev = 0, status = GS_OK;
    uint64_t m_f = 0;
    {   // the roots' degree sum
        std::vector<int32_t> d(nr);
        GS_HIP(hipMemcpyAsync(d.data(), w->qdeg.ptr, nr * 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        for (int i = 0; i < nr; ++i) m_f += (uint32_t)d[i];
    }
    // lanes whose frontier is not empty: a finished lane can reach nothing more,
    // so bottom-up stops looking for it
    unsigned long long active = all;
    for (uint32_t level = 1;; ++level) {
        if (!bottom_up && m_f > m_u / (uint64_t)g->do_alpha) bottom_up = true;
        else if (bottom_up && n_f < nzn / (uint64_t)g->do_beta) bottom_up = false;
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 5 * 8, w->stream));
        GS_HIP(hipMemsetAsync(w->ms_counts.ptr, 0, 64 * 8, w->stream));
        hipEvent_t kb;
        const int kclass = bottom_up ? GS_K_BFS_BOTTOMUP : GS_K_BFS_TOPDOWN;
        w->timer.begin(kclass, w->stream, &kb);
        if (bottom_up) {
            const uint64_t nh = g->ms_nhub;
            if (nh)
                k_ms_bu_heavy<<<ms_heavy_grid(nh), 256, 0, w->stream>>>(nh, active, g->adj_off.ptr, g->adj_nbr.ptr,
                                                                        vis, front, next, w->counters.ptr + 4);
            if (nzn > nh)
                k_ms_bu<<<ms_grid(nzn - nh), 256, 0, w->stream>>>(nh, nzn, active, g->adj_off.ptr, g->adj_nbr.ptr,
                                                                  g->adj_first.ptr, vis, front, next,
                                                                  w->counters.ptr + 4);
        } else if (n_f && m_f) {
            uint64_t nq = n_f;
            if (!have_queue) {
                GS_HIP(hipMemsetAsync(w->counters.ptr + 5, 0, 8, w->stream));
                k_ms_list<<<ms_grid(nzn), 256, 0, w->stream>>>(nzn, front, g->deg_r.ptr, w->queue.ptr, w->qdeg.ptr,
                                                               w->counters.ptr + 5);
                GS_HIP(hipGetLastError());
            }
            uint64_t m = 0;
            GS_TRY(queue_prefix(w, nq, scratch, &m));
            if (m)
                k_ms_td<<<(unsigned)((m + MS_CHUNK - 1) / MS_CHUNK), 256, 0, w->stream>>>(
                    w->queue.ptr, w->qoff.ptr, nq, m, g->adj_off.ptr, g->adj_nbr.ptr, front, vis, next);
            checked += m;
        }
        GS_HIP(hipGetLastError());
        w->timer.end(kclass, w->stream, kb);
        have_queue = !bottom_up;
        w->timer.begin(GS_K_BFS_LEVEL_END, w->stream, &kb);
        if (nzn) {
#define MS_LE(B, D) k_ms_level_end<B, D><<<ms_grid(nzn), 256, 0, w->stream>>>(nzn, active, vis, front, next, g->deg_r.ptr, \
                        dvm, level + 1, w->queue.ptr, w->qdeg.ptr, w->counters.ptr, w->ms_counts.ptr)
            if (have_queue && dvm) MS_LE(true, true);
            else if (have_queue) MS_LE(true, false);
            else if (dvm) MS_LE(false, true);
            else MS_LE(false, false);
#undef MS_LE
        }
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_BFS_LEVEL_END, w->stream, kb);
        GS_HIP(hipMemcpyAsync(lc.data(), w->ms_counts.ptr, 64 * 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 5 * 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        n_f = w->h_counters[0];
        m_f = w->h_counters[1];
        m_u = w->h_counters[3];
        if (bottom_up) checked += w->h_counters[4];
        unsigned long long tot = 0;
        active = 0;
        for (int i = 0; i < nr; ++i) {
            if (level_counts && nlev < max_levels) level_counts[(size_t)i * max_levels + nlev] = lc[i];
            tot += lc[i];
            if (lc[i]) active |= 1ull << i;
        }
        ++nlev;
        if (tot == 0) break;
        if (level + 1 > 255) {
            set_error("batched BFS needs depth > 255, beyond the u8 depth encoding");
            status = GS_ERR_DEPTH_OVERFLOW;
            break;
        }
    }
    if (status == GS_OK && dvm) {
        k_ms_unpermute<<<ms_grid(n), 256, 0, w->stream>>>(dvm, g->new_id.ptr, n, nzn, nr, w->ms_depth.ptr);
        k_ms_root_depth<<<1, 64, 0, w->stream>>>(w->ms_roots.ptr, nr, g->new_id.ptr, n, nzn, w->ms_depth.ptr);
        GS_HIP(hipGetLastError());
    }
    w->stats.edges_scanned = checked;
    if (levels) *levels = nlev;
    return status;
}

// The index method with its level loop on the device: the MSG_LEVELS-level
// sequence is captured once per workspace (again when a buffer it names has
// moved) and launched until the done flag is up.
static int run_bfs_batch_graph(const gs_graph *gc, gs_ws *w, const uint64_t *roots, int nr, uint64_t *level_counts,
                               int max_levels, int *levels, bool keep_depth) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    const uint64_t n = g->n, nzn = g->nzn;
    unsigned long long *vis = (unsigned long long *)w->ms_vis.ptr, *front = (unsigned long long *)w->ms_front.ptr,
                       *next = (unsigned long long *)w->ms_next.ptr;
    uint8_t *dvm = keep_depth ? w->ms_depth_vm.ptr : nullptr;
    if (w->ms_ctl.n != C_WORDS) GS_TRY(w->ms_ctl.alloc(C_WORDS));
    unsigned long long *ctl = w->ms_ctl.ptr;
    GS_HIP(hipMemcpyAsync(w->ms_roots.ptr, roots, nr * 8, hipMemcpyHostToDevice, w->stream));
    // the vertices with edges only (the others keep no state; see k_ms_roots)
    GS_HIP(hipMemsetAsync(vis, 0, nzn * 8, w->stream));
    GS_HIP(hipMemsetAsync(front, 0, nzn * 8, w->stream));
    GS_HIP(hipMemsetAsync(next, 0, nzn * 8, w->stream));
    if (dvm) GS_HIP(hipMemsetAsync(dvm, 0, nzn * 64, w->stream));
    GS_HIP(hipMemsetAsync(ctl, 0, C_WORDS * 8, w->stream));
    k_ms_roots<<<1, 64, 0, w->stream>>>(w->ms_roots.ptr, nr, g->new_id.ptr, g->deg_r.ptr, nzn, vis, front, w->queue.ptr,
                                        w->qdeg.ptr, dvm);
    GS_HIP(hipGetLastError());
    k_msg_init<<<1, 64, 0, w->stream>>>(ctl, nr, w->qdeg.ptr, 2 * g->E, (uint32_t)g->do_alpha, (uint32_t)g->do_beta);
    GS_HIP(hipGetLastError());
    // what the captured kernels were given (the thresholds travel in ctl)
    const std::vector<const void *> key = {g, (const void *)(uintptr_t)nzn, (const void *)(uintptr_t)g->ms_nhub,
                                           vis, front, next, dvm, ctl,
                                           w->queue.ptr, w->qdeg.ptr, w->qoff.ptr, w->qsum.ptr, g->adj_off.ptr,
                                           g->adj_nbr.ptr, g->adj_first.ptr, g->deg_r.ptr};
    if (!w->ms_graph || w->ms_graph_key != key) {
        if (w->ms_graph) {
            GS_HIP(hipGraphExecDestroy(w->ms_graph));
            w->ms_graph = nullptr;
        }
        GS_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
        // grid-stride kernels, at most MSG_GRID workgroups: enough to fill the GPU, and a kernel
        // whose flag is down for the level costs the launch of fewer empty workgroups
        auto cap = [](unsigned x) { return std::max(1u, std::min(x, MSG_GRID)); };
        const unsigned gv = cap(ms_grid(nzn)), gb = cap((unsigned)std::min<uint64_t>((nzn + 1023) / 1024, MSG_GRID));
        const uint64_t nh = g->ms_nhub;
        const unsigned hgrid = cap(ms_heavy_grid(nh)), glight = cap(ms_grid(nzn - nh));
        for (int l = 0; l < MSG_LEVELS; ++l) {
            k_msg_decide<<<1, 64, 0, w->stream>>>(ctl, nzn);
            k_msg_list<<<gv, 256, 0, w->stream>>>(ctl, nzn, front, g->deg_r.ptr, w->queue.ptr, w->qdeg.ptr,
                                                  ctl + C_LIST);
            k_msg_block_sums<<<gb, 256, 0, w->stream>>>(ctl, w->qdeg.ptr, w->qsum.ptr);
            k_msg_scan<<<1, 1024, 0, w->stream>>>(ctl, w->qsum.ptr);
            k_msg_deg_scan<<<gb, 256, 0, w->stream>>>(ctl, w->qdeg.ptr, w->qsum.ptr, w->qoff.ptr);
            k_msg_td<<<MSG_TD_GRID, 256, 0, w->stream>>>(ctl, w->queue.ptr, w->qoff.ptr, g->adj_off.ptr, g->adj_nbr.ptr,
                                                         front, vis, next);
            if (nh)
                k_msg_bu<true><<<hgrid, 256, 0, w->stream>>>(ctl, 0, nh, g->adj_off.ptr, g->adj_nbr.ptr, nullptr, vis,
                                                             front, next);
            if (nzn > nh)
                k_msg_bu<false><<<glight, 256, 0, w->stream>>>(ctl, nh, nzn, g->adj_off.ptr, g->adj_nbr.ptr,
                                                                          g->adj_first.ptr, vis, front, next);
            if (dvm) {
                k_msg_level_end<true, true><<<gv, 256, 0, w->stream>>>(ctl, nzn, vis, front, next, g->deg_r.ptr, dvm,
                                                                       w->queue.ptr, w->qdeg.ptr);
                k_msg_level_end<true, false><<<gv, 256, 0, w->stream>>>(ctl, nzn, vis, front, next, g->deg_r.ptr,
                                                                        dvm, w->queue.ptr, w->qdeg.ptr);
            } else {
                k_msg_level_end<false, true><<<gv, 256, 0, w->stream>>>(ctl, nzn, vis, front, next, g->deg_r.ptr,
                                                                        dvm, w->queue.ptr, w->qdeg.ptr);
                k_msg_level_end<false, false><<<gv, 256, 0, w->stream>>>(ctl, nzn, vis, front, next, g->deg_r.ptr,
                                                                         dvm, w->queue.ptr, w->qdeg.ptr);
            }
            k_msg_finish<<<1, 64, 0, w->stream>>>(ctl);
        }
        const hipError_t launch = hipGetLastError();
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(w->stream, &graph);
        if (launch != hipSuccess || end != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            set_error("batched BFS: capture of the level sequence failed: %s",
                      hipGetErrorString(launch != hipSuccess ? launch : end));
            return GS_ERR_DEVICE;
        }
        const hipError_t inst = hipGraphInstantiate(&w->ms_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (inst != hipSuccess) {
            w->ms_graph = nullptr;
            set_error("batched BFS: hipGraphInstantiate: %s", hipGetErrorString(inst));
            return GS_ERR_DEVICE;
        }
        w->ms_graph_key = key;
    }
    unsigned long long *h = w->h_counters;   // pinned: flags, levels done, entries examined
    for (;;) {
        GS_HIP(hipGraphLaunch(w->ms_graph, w->stream));
        GS_HIP(hipMemcpyAsync(&h[0], ctl + C_FLAGS, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipMemcpyAsync(&h[1], ctl + C_NLEV, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipMemcpyAsync(&h[2], ctl + C_CHECKED, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        if (h[0] & F_DONE) break;
    }
    const unsigned long long flags = h[0], checked = h[2];
    const int nlev = (int)h[1];
    if (level_counts && max_levels > 0) {
        const int rows = std::min(nlev, max_levels);
        std::vector<unsigned long long> lc((size_t)rows * 64);
        if (rows) {
            GS_HIP(hipMemcpyAsync(lc.data(), ctl + C_LANES, lc.size() * 8, hipMemcpyDeviceToHost, w->stream));
            GS_HIP(hipStreamSynchronize(w->stream));
        }
        for (int l = 0; l < rows; ++l)
            for (int i = 0; i < nr; ++i) level_counts[(size_t)i * max_levels + l] = lc[(size_t)l * 64 + i];
    }
    w->stats.edges_scanned = checked;
    if (levels) *levels = nlev;
    if (flags & F_OVERFLOW) {
        set_error("batched BFS needs depth > 255, beyond the u8 depth encoding");
        return GS_ERR_DEPTH_OVERFLOW;
    }
    if (dvm) {
        k_ms_unpermute<<<ms_grid(n), 256, 0, w->stream>>>(dvm, g->new_id.ptr, n, nzn, nr, w->ms_depth.ptr);
        k_ms_root_depth<<<1, 64, 0, w->stream>>>(w->ms_roots.ptr, nr, g->new_id.ptr, n, nzn, w->ms_depth.ptr);
        GS_HIP(hipGetLastError());
    }
    return GS_OK;
}

int gs::run_bfs_batch(const gs_graph *g, gs_ws *w, const uint64_t *roots, int nr, uint8_t *depth_out,
                       uint64_t *level_counts, int max_levels, int *levels, bool keep_depth) {
    if (!g || !roots || nr < 1 || nr > 64 || max_levels < 0) {
        set_error("gs_bfs_batch: bad arguments (need 1..64 roots)");
        return GS_ERR_ARG;
    }
    for (int i = 0; i < nr; ++i)
        if (roots[i] >= g->n) {
            set_error("gs_bfs_batch: root %llu >= N", (unsigned long long)roots[i]);
            return GS_ERR_ARG;
        }
    const bool by_index = g->batch_method >= 1 && !g->directed && !g->ss;
    keep_depth = keep_depth || depth_out;
    if (by_index) {
        GS_TRY(build_adjacency(const_cast<gs_graph *>(g)));
        GS_TRY(ensure_queue(g, w));
        if (keep_depth && w->ms_depth_vm.n != g->n * 64) GS_TRY(w->ms_depth_vm.alloc(g->n * 64));
    } else {
        GS_TRY(ensure_rowact(g, w));
    }
    if (w->ms_vis.n != g->n) {
        GS_TRY(w->ms_vis.alloc(g->n));
        GS_TRY(w->ms_front.alloc(g->n));
        GS_TRY(w->ms_next.alloc(g->n));
    }
    if (keep_depth || !by_index)
        if (w->ms_depth.n < (size_t)nr * g->n) GS_TRY(w->ms_depth.alloc((size_t)nr * g->n));
    if (w->ms_roots.n < 64) GS_TRY(w->ms_roots.alloc(64));
    if (w->ms_counts.n < 64) GS_TRY(w->ms_counts.alloc(64));
    w->timer.reset();
    w->stats = gs_stats{};
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    if (by_index) {
        int nlev = 0;
        const int status = g->batch_method == 2
                               ? run_bfs_batch_graph(g, w, roots, nr, level_counts, max_levels, &nlev, keep_depth)
                               : run_bfs_batch_index(g, w, roots, nr, level_counts, max_levels, &nlev, keep_depth);
        GS_HIP(hipEventRecord(t1, w->stream));
        GS_TRY(sync_and_collect(w));
        float ms = 0.f;
        GS_HIP(hipEventElapsedTime(&ms, t0, t1));
        w->stats.ms = ms;
        w->stats.iterations = (uint64_t)nlev;
        if (levels) *levels = nlev;
        if (status != GS_OK) return status;
        if (depth_out) {
            GS_HIP(hipMemcpyAsync(depth_out, w->ms_depth.ptr, (size_t)nr * g->n, hipMemcpyDeviceToHost, w->stream));
            GS_HIP(hipStreamSynchronize(w->stream));
        }
        return GS_OK;
    }
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
    return run_bfs_batch(g, g->ws, roots, nr, depth_out, level_counts, max_levels, levels, depth_out != nullptr);
}
