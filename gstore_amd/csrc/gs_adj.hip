// gs_adj.hip -- neighbour-list index built from the tile grid, and the
// direction-optimizing BFS and PageRank pull that use it.
//
// The tile grid stores each undirected edge once, in the tile of (min, max);
// a vertex's neighbours are spread over its whole tile row and tile column,
// so a BFS level over tiles has to read every edge of every active tile, and
// a PageRank iteration over tiles needs an atomic add per edge endpoint.  The
// index is built once from the grid on the device and kept with the graph:
//   * vertices are relabelled in descending degree order (log-spaced degree
//     buckets), so the per-vertex values of the heavily shared vertices sit in
//     few cache lines that stay in L2 / the Infinity Cache; vertices without
//     edges come last;
//   * off_r (u64, N+1) is the exclusive scan of the relabelled degrees and
//     nbr_r (u32, 2E) holds every stored edge in both directions, as
//     relabelled ids, filled by one sweep over the grid.
// Results are mapped back to the original ids at the end of each call.  BFS
// keeps the level structure, frontier bitmaps and depth encoding of the tile
// BFS (bfs.cpp:56-224), so depths are bit-identical in either mode.
#include <vector>

#include "gs_internal.h"

namespace gs {

// ---- index build -------------------------------------------------------------

// Per-block sums of degree (blocks of 1024 vertices).
__global__ void k_deg_block_sums(const int32_t *__restrict__ deg, uint64_t n, unsigned long long *bsum) {
    __shared__ unsigned long long s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * 1024;
    unsigned long long c = 0;
    for (int i = threadIdx.x; i < 1024; i += blockDim.x)
        if (base + i < n) c += (uint32_t)deg[base + i];
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

// off[v] = bstart[block] + exclusive prefix of degree within the block of 1024
// (4 entries per thread, then a Hillis-Steele scan of the 256 thread sums).
__global__ void k_deg_scan(const int32_t *__restrict__ deg, uint64_t n, const unsigned long long *__restrict__ bstart,
                           uint64_t *off) {
    __shared__ uint32_t d[1024];
    uint64_t base = (uint64_t)blockIdx.x * 1024;
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) d[i] = base + i < n ? (uint32_t)deg[base + i] : 0;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 entries with 256 threads (4 per thread)
    __shared__ unsigned long long part[256];
    unsigned long long loc[4], run = 0;
    for (int j = 0; j < 4; ++j) {
        loc[j] = run;
        run += d[threadIdx.x * 4 + j];
    }
    part[threadIdx.x] = run;
    __syncthreads();
    for (int step = 1; step < 256; step <<= 1) {
        unsigned long long v = threadIdx.x >= step ? part[threadIdx.x - step] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long before = (threadIdx.x ? part[threadIdx.x - 1] : 0) + bstart[blockIdx.x];
    for (int j = 0; j < 4; ++j) {
        uint64_t v = base + threadIdx.x * 4 + j;
        if (v < n) off[v] = before + loc[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) off[n] = before + run;
}

__global__ void __launch_bounds__(SWEEP_WG) k_adj_scatter(SweepArgs a, const uint32_t *__restrict__ new_id,
                                                          const uint64_t *__restrict__ off, uint32_t *cur,
                                                          uint32_t *nbr) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t p0 = new_id[gv0(w, rc)], p1 = new_id[gv1(w, rc)];
        nbr[off[p0] + atomicAdd(&cur[p0], 1u)] = p1;
        nbr[off[p1] + atomicAdd(&cur[p1], 1u)] = p0;
    };
    sweep_chunk<false>(a, chunk, s, len, op);
}

// ---- direction-optimizing BFS kernels ------------------------------------------

// The levels run from the device: per level, k_do_begin picks the direction
// from the previous level's counters and raises flags; the listing, queue
// prefix, top-down, bottom-up and level-end kernels return at once unless
// their flag is up (and the search is not done); k_do_end closes the level.
// DO_LEVELS levels are captured once per workspace into a hipGraph, which the
// host launches until the done flag is up: one host round trip per DO_LEVELS
// levels.  Control words, u64, in w->do_ctl; the first four are the counters
// the level end and bottom-up kernels add into, in this order.
enum : int {
    DO_NF = 0,        // level end: vertices of the new frontier
    DO_MF = 1,        //            their degree sum
    DO_CUR = 2,       //            queue cursor
    DO_CHK = 3,       // bottom-up: list entries checked
    DO_LIST = 4,      // frontier listing: count, degree sum, cursor (4 .. 6)
    DO_ULN = 8,       // unreached-list length
    DO_NQ = 9,        // frontier at the level's start (queue length)
    DO_MQ = 10,       // its degree sum (the top-down level's edges)
    DO_MU = 11,       // degree sum of the vertices in no frontier yet
    DO_NFB = 12,      // frontier one level earlier
    DO_REACHED = 13,  // vertices with edges reached
    DO_FLAGS = 14,
    DO_LEVEL = 15,    // current level (1 = the root's)
    DO_NLEV = 16,     // levels done
    DO_CHECKED = 17,  // entries examined, all levels
    DO_ALPHA = 18,
    DO_BETA = 19,
    DO_REC = 32,      // per level: direction, frontier, entries, reached, start and end clock
    DO_RECW = 6,
    DO_WORDS = DO_REC + DO_RECW * 256
};
enum : unsigned long long {
    DF_BU = 1,        // bottom-up level
    DF_HAVEQ = 2,     // the queue holds the frontier (the root's level)
    DF_DONE = 4,
    DF_TD = 8,        // top-down kernels run
    DF_LIST = 16,     // the frontier is listed first (every top-down level after the root's)
    DF_OVERFLOW = 32, // the next depth would not fit in u8
    DF_ULIST = 64,    // bottom-up over the unreached list
    DF_BUFULL = 128,  // bottom-up over all vertices with edges
    DF_LEVEL = 256,   // a level runs (its level end)
    DF_TDNV = 512,    // top-down without the visited test (few vertices reached yet)
    DF_TDH = 1024     // a hub level: marks of the heaviest TD_HUB_IDS vertices gathered in LDS first
};
constexpr int DO_LEVELS = 4;
constexpr unsigned DO_TD_GRID = 256;   // TD_WG-thread workgroups: one per CU (LDS: staging + hub bitmap)
constexpr unsigned DO_SCAN_GRID = 1024;
constexpr unsigned DO_GRID = 2048;

// true when the kernel's flag is up and the search is not done
__device__ inline bool do_gate(const unsigned long long *ctl, unsigned long long flag) {
    return (ctl[DO_FLAGS] & (flag | DF_DONE)) == flag;
}

// Top-down, balanced over edges: qoff is the exclusive prefix of the frontier
// queue's degrees, and each workgroup step takes TD_CHUNK consecutive frontier
// edges (grid-stride: the grid is fixed when the levels are captured).  The
// queue entries its range touches are staged in LDS (offset into the range
// and list start), each lane finds its entry by a search there.  A neighbour
// not yet visited or marked is marked with atomicOr.  While few vertices are
// reached (DF_TDNV: the levels after the root's, where a frontier of hubs
// sends hundreds of millions of edges to unvisited vertices) the visited test
// is left out -- one random read per edge fewer -- and the level end drops
// the visited vertices from the marks.
constexpr uint32_t TD_CHUNK = 4096;
constexpr int TD_EPT = 4;   // edges per thread whose loads are in flight together
constexpr unsigned TD_WG = 1024;                 // threads per top-down workgroup: one TD_CHUNK step is one batch
constexpr uint32_t TD_HUB_IDS = 1u << 19;        // hub level: relabelled ids whose marks go through LDS
constexpr uint32_t TD_HUB_WORDS = TD_HUB_IDS / 32;

// On a hub level (DF_TDH: hundreds of millions of frontier edges while few
// vertices are reached) about half of the edges land on the TD_HUB_IDS
// heaviest vertices.  Their marks are or-ed into a bitmap in LDS instead of
// the global one (no random global read, no atomic per edge), and each
// workgroup ors its LDS words into next once at the end.  hub_words is
// min(TD_HUB_WORDS, the bitmap words of the vertices with edges).
__global__ void __launch_bounds__(TD_WG) k_td_edges(const unsigned long long *ctl, const uint32_t *__restrict__ queue,
                                                    const uint64_t *__restrict__ qoff, const uint64_t *__restrict__ off,
                                                    const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ vis,
                                                    uint32_t *next, uint32_t hub_words) {
    if (!do_gate(ctl, DF_TD)) return;
    const unsigned long long fl = ctl[DO_FLAGS];
    const bool test_vis = !(fl & DF_TDNV), hub = (fl & DF_TDH) && hub_words;
    const uint32_t hub_ids = hub ? hub_words * 32 : 0u;
    __shared__ uint32_t sbeg[TD_CHUNK + 2];  // first edge of staged entry k, relative to the range (clamped)
    __shared__ uint64_t sadj[TD_CHUNK + 1];  // list start of staged entry k
    __shared__ uint32_t hb[TD_HUB_WORDS];    // hub level: marks of ids below hub_ids
    __shared__ uint64_t j_first;
    __shared__ int nstaged;
    if (hub) {
        for (uint32_t i = threadIdx.x; i < hub_words; i += blockDim.x) hb[i] = 0;
        __syncthreads();
    }
    const uint64_t nq = ctl[DO_NQ], m = ctl[DO_MQ];
    for (uint64_t c0 = blockIdx.x * (uint64_t)TD_CHUNK; c0 < m; c0 += (uint64_t)gridDim.x * TD_CHUNK) {
        const uint32_t len = (uint32_t)min<uint64_t>(TD_CHUNK, m - c0);
        if (threadIdx.x < 64) {
            // largest j < nq with qoff[j] <= c0, then with qoff[j] <= c0 + len - 1
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
            sbeg[k] = q <= c0 ? 0u : (uint32_t)(q - c0);
            sadj[k] = off[queue[ja + k]] + (q < c0 ? c0 - q : 0);  // list position of the range's first edge of entry k
        }
        if (threadIdx.x == 0) sbeg[ns] = len;
        __syncthreads();
        // TD_EPT edges per thread at a time, their loads issued together: a
        // thread walking its edges one by one had one load in flight
        for (uint32_t t0 = threadIdx.x; t0 < len; t0 += TD_EPT * blockDim.x) {
            uint32_t v[TD_EPT], nw[TD_EPT], vw[TD_EPT];
#pragma unroll
            for (int j = 0; j < TD_EPT; ++j) {
                const uint32_t t = t0 + j * blockDim.x;
                v[j] = 0xFFFFFFFFu;
                if (t < len) {
                    int lo = 0, hi = ns - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (sbeg[mid] <= t) lo = mid; else hi = mid - 1;
                    }
                    v[j] = nbr[sadj[lo] + (t - sbeg[lo])];
                }
            }
#pragma unroll
            for (int j = 0; j < TD_EPT; ++j) {
                if (v[j] < hub_ids) {   // (never on other levels: hub_ids is 0 there)
                    atomicOr(&hb[v[j] >> 5], 1u << (v[j] & 31));
                    v[j] = 0xFFFFFFFFu;
                }
            }
#pragma unroll
            for (int j = 0; j < TD_EPT; ++j) {
                nw[j] = v[j] != 0xFFFFFFFFu ? next[v[j] >> 5] : ~0u;
                vw[j] = test_vis && v[j] != 0xFFFFFFFFu ? vis[v[j] >> 5] : 0u;
            }
#pragma unroll
            for (int j = 0; j < TD_EPT; ++j) {
                if (v[j] == 0xFFFFFFFFu) continue;
                const uint32_t bit = 1u << (v[j] & 31);
                if (!((nw[j] | vw[j]) & bit)) atomicOr(&next[v[j] >> 5], bit);
            }
        }
        __syncthreads();  // the staged tables are rewritten by the next step
    }
    if (hub) {   // (the loop ended with a barrier)
        for (uint32_t i = threadIdx.x; i < hub_words; i += blockDim.x)
            if (hb[i]) atomicOr(&next[i], hb[i]);
    }
}

// Bottom-up: BU_VPL vertices per lane (v, v + 64, ... of the wave's 64 x BU_VPL),
// so each lane keeps all its vertices' loads in flight; an unreached vertex
// with neighbours scans them until it meets a frontier vertex.  The first
// list entry comes from first[] (one coalesced load, no offsets): short lists
// are sorted, so it is the heaviest neighbour and on the big levels nearly
// always in the frontier.  The vertices it leaves unresolved (about a tenth on
// the big levels) are packed into LDS per wave and scanned one per lane from
// the second entry: four loads of ids and four of frontier words per round for
// the wave, where scanning each lane's four vertices side by side issued 32
// loads per round whatever few lanes still had work (the texture address unit
// was 90% busy at s28 level 4).  Without LIST each wave owns 64 x BU_VPL
// consecutive vertices 0 .. n-1, skips the reached ones, gathers its finds in
// LDS words and writes its next-bitmap words without atomics.  With LIST the
// vertices come from a list of the unreached ones (k_unreached_list), so a
// late level touches only them; found vertices are or-ed into next.
constexpr int BU_VPL = 4;
constexpr int BU_LIST_DIV = 32;  // bottom-up from the unreached list once at most nzn / 32 are left

// Orders a wave's LDS writes before its lanes read each other's entries.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool LIST>
__global__ void __launch_bounds__(256) k_bu_expand(unsigned long long *ctl, uint64_t n, const uint32_t *__restrict__ list,
                                                   const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                   const uint32_t *__restrict__ first, const uint32_t *__restrict__ vis,
                                                   const uint32_t *__restrict__ front, uint32_t *next) {
    if (!do_gate(ctl, LIST ? DF_ULIST : DF_BUFULL)) return;
    __shared__ uint32_t s_u[4][64 * BU_VPL];   // per wave: its unresolved vertices
    __shared__ uint32_t s_fw[4][2 * BU_VPL];   // per wave: next-bitmap words of its vertices (full mode)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t checked = 0;
    const uint64_t cnt = LIST ? ctl[DO_ULN] : n;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * BU_VPL;
    for (uint64_t base = ((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) & ~63ull) * BU_VPL; base < cnt;
         base += stride) {
        uint64_t v[BU_VPL];
        bool found[BU_VPL], open[BU_VPL], any = false;
        uint32_t f0[BU_VPL];
#pragma unroll
        for (int h = 0; h < BU_VPL; ++h) {
            const uint64_t i = base + 64 * h + lane;
            found[h] = false;
            if (LIST) {
                open[h] = i < cnt;
                v[h] = open[h] ? list[i] : 0;
            } else {
                v[h] = i;
                open[h] = i < n && !((vis[i >> 5] >> (i & 31)) & 1u);
            }
            any = any || open[h];
        }
        if (!__ballot(any)) continue;  // the wave's vertices are all reached (late levels: nearly every wave)
#pragma unroll
        for (int h = 0; h < BU_VPL; ++h) f0[h] = first[open[h] ? v[h] : 0];
#pragma unroll
        for (int h = 0; h < BU_VPL; ++h) {
            const uint32_t fw = front[f0[h] >> 5];
            if (open[h]) {
                ++checked;
                found[h] = (fw >> (f0[h] & 31)) & 1u;
            }
        }
        // the vertices the first entry left unresolved, packed per wave (the
        // wave's first-entry finds go to LDS words first in the full mode)
        uint32_t nu = 0;
#pragma unroll
        for (int h = 0; h < BU_VPL; ++h) {
            if (!LIST) {
                const unsigned long long m = __ballot(found[h]);
                if (lane == 0) {
                    s_fw[wv][2 * h] = (uint32_t)m;
                    s_fw[wv][2 * h + 1] = (uint32_t)(m >> 32);
                }
            } else if (found[h]) {
                atomicOr(&next[v[h] >> 5], 1u << (v[h] & 31));
            }
            const bool un = open[h] && !found[h];
            const unsigned long long um = __ballot(un);
            if (un) s_u[wv][nu + __popcll(um & ((1ull << lane) - 1))] = (uint32_t)v[h];
            nu += (uint32_t)__popcll(um);
        }
        wave_sync();
        // one unresolved vertex per lane, four neighbours per round trip: ids, then their
        // frontier words (8 loads per round for the wave, where four vertices per lane took 32)
        for (uint32_t t = lane; t < nu; t += 64) {
            const uint32_t x = s_u[wv][t];
            uint64_t kk = off[x] + 1;
            const uint64_t ee = off[x + 1];
            bool fx = false;
            while (kk < ee && !fx) {
                uint32_t u[4], fw[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) u[j] = nbr[min(kk + j, ee - 1)];
#pragma unroll
                for (int j = 0; j < 4; ++j) fw[j] = front[u[j] >> 5];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (!fx && kk + j < ee) {
                        ++checked;
                        fx = (fw[j] >> (u[j] & 31)) & 1u;
                    }
                kk += 4;
            }
            if (fx) {
                if (LIST) atomicOr(&next[x >> 5], 1u << (x & 31));
                else atomicOr(&s_fw[wv][(uint32_t)(x - base) >> 5], 1u << ((x - base) & 31));
            }
        }
        if (!LIST) {
            wave_sync();
            // the wave owns the 8 next-bitmap words of its 256 vertices
            if (lane < 8 && s_fw[wv][lane]) next[(base >> 5) + lane] = s_fw[wv][lane];
            wave_sync();   // s_fw and s_u are rewritten by the next step
        } else {
            wave_sync();   // s_u is rewritten by the next step
        }
    }
    block_count(checked, ctl + DO_CHK);
}

// The unreached vertices with edges (clear bits of vis below nzn), listed in
// list[0 .. ctl[DO_ULN]): four words per thread, slots reserved once per workgroup
// and step (wave scan of the thread counts, one global atomic).
__global__ void __launch_bounds__(256) k_unreached_list(unsigned long long *ctl, const uint32_t *__restrict__ vis,
                                                        uint64_t nzn, uint32_t *list) {
    if (!do_gate(ctl, DF_ULIST)) return;
    unsigned long long *count = ctl + DO_ULN;
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long bbase;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t nwords = (nzn + 31) / 32;
    for (uint64_t base = blockIdx.x * 1024ull; base < nwords; base += (uint64_t)gridDim.x * 1024) {
        const uint64_t g0 = base + threadIdx.x * 4;
        uint32_t m[4], c = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint64_t gw = g0 + k;
            m[k] = 0;
            if (gw < nwords) {
                m[k] = ~vis[gw];
                if (gw * 32 + 32 > nzn) m[k] &= (1u << (uint32_t)(nzn - gw * 32)) - 1u;
            }
            c += __popc(m[k]);
        }
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
            bbase = t ? atomicAdd(count, (unsigned long long)t) : 0ull;
        }
        __syncthreads();
        unsigned long long pos = bbase + incl - c;
        for (int w = 0; w < wv; ++w) pos += wsum[w];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            uint32_t x = m[k];
            while (x) {
                const uint32_t b = __ffs(x) - 1;
                x &= x - 1;
                list[pos++] = (uint32_t)((g0 + k) * 32 + b);
            }
        }
        __syncthreads();  // wsum and bbase are rewritten by the next step
    }
}

// Level end for the direction-optimizing BFS: as k_bfs_level_end, plus the
// degree sum of the new frontier and its vertex queue (with the degrees) for
// a top-down level.
// LEVEL_END: next becomes the frontier (bitmaps and depths updated); else the
// current frontier is only listed.  BUILD: write the queue and its degrees.
// Each thread takes LE_WPT consecutive words per step.  With BUILD, queue
// slots are reserved once per workgroup and step (block scan of the thread
// counts): one global atomic per 32K vertices, where one per 8K serialised on
// the cursor.  Without BUILD there is no scan and no reservation.  Counts and
// degree sums are added once per workgroup.
constexpr int LE_WPT = 4;

template <bool LEVEL_END, bool BUILD>
__device__ inline void do_level_word(uint32_t nx, uint64_t gw, uint64_t n, uint8_t *depth, uint32_t newdepth,
                                     const int32_t *__restrict__ deg, uint32_t *queue, int32_t *qdeg,
                                     unsigned long long &pos, unsigned long long &msum) {
    const uint64_t v0 = gw * 32;
    if (nx && v0 + 32 <= n) {
        // the word's 32 degrees and depth bytes as 16-byte loads, updated in registers
        int4 dq[8];
        uint4 bq[2];
#pragma unroll
        for (int k = 0; k < 8; ++k) dq[k] = reinterpret_cast<const int4 *>(deg + v0)[k];
#pragma unroll
        for (int k = 0; k < 2; ++k) bq[k] = LEVEL_END ? reinterpret_cast<const uint4 *>(depth + v0)[k] : uint4{};
#pragma unroll
        for (int b = 0; b < 32; ++b) {
            if (!((nx >> b) & 1u)) continue;
            const int4 d4 = dq[b >> 2];
            const int32_t dv = (b & 3) == 0 ? d4.x : (b & 3) == 1 ? d4.y : (b & 3) == 2 ? d4.z : d4.w;
            if (LEVEL_END) {
                uint4 &bw = bq[b >> 4];
                uint32_t &word = ((b >> 2) & 3) == 0 ? bw.x : ((b >> 2) & 3) == 1 ? bw.y
                               : ((b >> 2) & 3) == 2 ? bw.z : bw.w;
                word = (word & ~(0xFFu << ((b & 3) * 8))) | (newdepth << ((b & 3) * 8));
            }
            msum += (uint32_t)dv;
            if (BUILD) {
                qdeg[pos] = dv;
                queue[pos++] = (uint32_t)(v0 + b);
            }
        }
        if (LEVEL_END) {
#pragma unroll
            for (int k = 0; k < 2; ++k) reinterpret_cast<uint4 *>(depth + v0)[k] = bq[k];
        }
    } else {
        uint32_t m = nx;
        while (m) {
            const uint32_t bit = __ffs(m) - 1;
            m &= m - 1;
            const uint64_t v = v0 + bit;
            if (LEVEL_END) depth[v] = (uint8_t)newdepth;
            const int32_t dv = deg[v];
            msum += (uint32_t)dv;
            if (BUILD) {
                qdeg[pos] = dv;
                queue[pos++] = (uint32_t)v;
            }
        }
    }
}

template <bool LEVEL_END, bool BUILD>
__global__ void __launch_bounds__(256) k_do_level_end(const unsigned long long *ctl, unsigned long long flag,
                                                      uint32_t *vis, uint32_t *front, uint32_t *next, uint8_t *depth,
                                                      const int32_t *__restrict__ deg, uint64_t n, uint64_t nwords,
                                                      uint32_t *queue, int32_t *qdeg, unsigned long long *count) {
    if (!do_gate(ctl, flag)) return;
    const uint32_t newdepth = (uint32_t)ctl[DO_LEVEL] + 1;
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long bbase, bcnt, bdeg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) bcnt = bdeg = 0;
    // before any wave adds into them: without the listing variant's barriers a
    // wave could finish its words and add before the clear, losing its count
    __syncthreads();
    unsigned long long cnt = 0, msum = 0;
    for (uint64_t base = blockIdx.x * (256ull * LE_WPT); base < nwords; base += (uint64_t)gridDim.x * 256 * LE_WPT) {
        const uint64_t g0 = base + threadIdx.x * LE_WPT;
        uint32_t nx[LE_WPT];
#pragma unroll
        for (int k = 0; k < LE_WPT; ++k) nx[k] = 0;
        if (g0 + LE_WPT <= nwords) {
            if (LEVEL_END) {
                // marks of visited vertices (left by a top-down level without the
                // visited test) are dropped here
                uint4 q = *reinterpret_cast<const uint4 *>(next + g0);
                if (q.x | q.y | q.z | q.w) {
                    *reinterpret_cast<uint4 *>(next + g0) = uint4{0, 0, 0, 0};
                    uint4 vv = *reinterpret_cast<const uint4 *>(vis + g0);
                    q.x &= ~vv.x; q.y &= ~vv.y; q.z &= ~vv.z; q.w &= ~vv.w;
                    vv.x |= q.x; vv.y |= q.y; vv.z |= q.z; vv.w |= q.w;
                    *reinterpret_cast<uint4 *>(vis + g0) = vv;
                }
                nx[0] = q.x; nx[1] = q.y; nx[2] = q.z; nx[3] = q.w;
                *reinterpret_cast<uint4 *>(front + g0) = q;
            } else {
                const uint4 q = *reinterpret_cast<const uint4 *>(front + g0);
                nx[0] = q.x; nx[1] = q.y; nx[2] = q.z; nx[3] = q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < LE_WPT; ++k) {
                const uint64_t gw = g0 + k;
                if (gw >= nwords) break;
                if (LEVEL_END) {
                    nx[k] = next[gw];
                    if (nx[k]) {
                        next[gw] = 0;
                        nx[k] &= ~vis[gw];
                        vis[gw] |= nx[k];
                    }
                    front[gw] = nx[k];
                } else {
                    nx[k] = front[gw];
                }
            }
        }
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < LE_WPT; ++k) c += __popc(nx[k]);
        cnt += c;
        unsigned long long pos = 0;
        if (BUILD) {
            uint32_t incl = c;
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
            uint32_t before = 0;
            for (int k = 0; k < wv; ++k) before += wsum[k];
            pos = bbase + before + incl - c;
        }
#pragma unroll
        for (int k = 0; k < LE_WPT; ++k)
            if (nx[k])
                do_level_word<LEVEL_END, BUILD>(nx[k], g0 + k, n, depth, newdepth, deg, queue, qdeg, pos, msum);
        if (BUILD) __syncthreads();  // wsum and bbase are rewritten by the next step
    }
    if (cnt) {
        atomicAdd(&bcnt, cnt);
        atomicAdd(&bdeg, msum);
    }
    __syncthreads();
    if (threadIdx.x == 0 && bcnt) {
        atomicAdd(&count[0], bcnt);
        atomicAdd(&count[1], bdeg);
    }
}

// The root's level: depth 1, visited, the frontier queue holds it, and ctl
// starts the heuristic (Beamer et al.): m_u = degree sum of the vertices in no
// frontier yet.
__global__ void k_do_root(uint64_t root, const uint32_t *__restrict__ new_id, const int32_t *__restrict__ deg,
                          uint64_t nzn, uint64_t m_total, uint32_t alpha, uint32_t beta, unsigned long long *ctl,
                          uint8_t *depth, uint32_t *vis, uint32_t *front, uint32_t *queue, int32_t *qdeg) {
    const uint32_t r = new_id[root];
    const int32_t d = deg[r];
    depth[r] = 1;
    vis[r >> 5] |= 1u << (r & 31);
    front[r >> 5] |= 1u << (r & 31);
    queue[0] = r;
    qdeg[0] = d;
    ctl[DO_NF] = 1;
    ctl[DO_MF] = (uint32_t)d;
    ctl[DO_MU] = m_total - (uint32_t)d;
    ctl[DO_REACHED] = r < nzn ? 1 : 0;
    ctl[DO_FLAGS] = DF_HAVEQ;
    ctl[DO_LEVEL] = 1;
    ctl[DO_ALPHA] = alpha;
    ctl[DO_BETA] = beta;
}

// Picks the level's direction (Beamer et al.) and kernels from the previous
// level's counters, then clears the counters.  Top-down to bottom-up only
// while the frontier grows: a shrinking tail with few unchecked edges left
// stays top-down, where a bottom-up pass would still walk every vertex's
// visited bit.  Back to top-down once the frontier is small, unless its edges
// outnumber all the unreached vertices' edges (a bottom-up level checks at
// most those).
__global__ void k_do_begin(unsigned long long *ctl, uint64_t n, uint64_t nzn) {
    const unsigned long long f = ctl[DO_FLAGS];
    if (f & DF_DONE) return;
    const unsigned long long nf = ctl[DO_NF], mf = ctl[DO_MF], mu = ctl[DO_MU];
    bool bu = f & DF_BU;
    if (!bu && mf > mu / ctl[DO_ALPHA] && nf > ctl[DO_NFB]) bu = true;
    else if (bu && nf < n / ctl[DO_BETA] && mf <= mu) bu = false;
    const unsigned long long reached = ctl[DO_REACHED];
    const unsigned long long unreached = nzn - (reached < nzn ? reached : nzn);
    // the level end never lists the new frontier: the next level is as likely
    // bottom-up (s28 level 3 -> 4), and listing 4.8M vertices there cost more
    // than the listing passes the small top-down levels now make
    unsigned long long nfl = DF_LEVEL | (bu ? DF_BU : 0);
    if (bu) {
        if (unreached && unreached <= nzn / BU_LIST_DIV) nfl |= DF_ULIST;
        else if (unreached) nfl |= DF_BUFULL;
    } else if (nf && mf) {
        nfl |= DF_TD | (f & DF_HAVEQ ? 0 : DF_LIST);
        if (reached < nzn / 64) nfl |= DF_TDNV | (mf >= nzn / 8 ? DF_TDH : 0);
    }
    ctl[DO_NFB] = nf;
    ctl[DO_NQ] = nf;
    ctl[DO_MQ] = mf;
    for (int i = 0; i <= DO_ULN; ++i) ctl[i] = 0;
    unsigned long long *rec = ctl + DO_REC + DO_RECW * (ctl[DO_LEVEL] - 1);
    rec[0] = bu ? 2 : 1;
    rec[1] = nf;
    rec[4] = (unsigned long long)wall_clock64();
    ctl[DO_FLAGS] = nfl;
}

// qsum[b] = degree sum of queue entries [1024 b, 1024 b + 1024) (grid-stride)
__global__ void __launch_bounds__(256) k_do_block_sums(const unsigned long long *ctl, const int32_t *__restrict__ qdeg,
                                                       unsigned long long *qsum) {
    if (!do_gate(ctl, DF_TD)) return;
    __shared__ unsigned long long s;
    const uint64_t nq = ctl[DO_NQ];
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

// Exclusive prefix of the block sums in place (one workgroup: each thread adds
// a contiguous slice, the slice sums are scanned in LDS, each thread rewrites
// its slice).
__global__ void __launch_bounds__(1024) k_do_scan(const unsigned long long *ctl, unsigned long long *qsum) {
    if (!do_gate(ctl, DF_TD)) return;
    __shared__ unsigned long long part[1024];
    const uint64_t nb = (ctl[DO_NQ] + 1023) / 1024;
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t b0 = std::min<uint64_t>(threadIdx.x * per, nb), b1 = std::min<uint64_t>(b0 + per, nb);
    unsigned long long t = 0;
    for (uint64_t b = b0; b < b1; ++b) t += qsum[b];
    part[threadIdx.x] = t;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned long long y = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0ull;
        __syncthreads();
        part[threadIdx.x] += y;
        __syncthreads();
    }
    unsigned long long run = part[threadIdx.x] - t;
    for (uint64_t b = b0; b < b1; ++b) {
        const unsigned long long c = qsum[b];
        qsum[b] = run;
        run += c;
    }
}

// qoff[j] = exclusive prefix of qdeg over the queue (blocks of 1024 as
// k_deg_scan, grid-stride)
__global__ void __launch_bounds__(256) k_do_deg_scan(const unsigned long long *ctl, const int32_t *__restrict__ qdeg,
                                                     const unsigned long long *__restrict__ bstart, uint64_t *qoff) {
    if (!do_gate(ctl, DF_TD)) return;
    __shared__ uint32_t d[1024];
    __shared__ unsigned long long part[256];
    const uint64_t nq = ctl[DO_NQ];
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
        for (int o = 1; o < 256; o <<= 1) {
            const unsigned long long y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0ull;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        const unsigned long long before = (threadIdx.x ? part[threadIdx.x - 1] : 0ull) + bstart[base / 1024];
        for (int j = 0; j < 4; ++j) {
            const uint64_t v = base + threadIdx.x * 4 + j;
            if (v < nq) qoff[v] = before + loc[j];
        }
        __syncthreads();  // d and part are rewritten by the next step
    }
}

// Closes the level: entries examined, the level's record, the heuristic's
// state, and the stop test (nothing new reached, or the next depth would not
// fit in u8).
__global__ void k_do_end(unsigned long long *ctl) {
    const unsigned long long f = ctl[DO_FLAGS];
    if (f & DF_DONE) return;
    const unsigned long long nf = ctl[DO_NF], mf = ctl[DO_MF], level = ctl[DO_LEVEL];
    const unsigned long long edges = (f & DF_BU) ? ctl[DO_CHK] : ctl[DO_MQ];
    unsigned long long *rec = ctl + DO_REC + DO_RECW * (level - 1);
    rec[2] = edges;
    rec[3] = nf;
    rec[5] = (unsigned long long)wall_clock64();
    ctl[DO_CHECKED] += edges;
    const unsigned long long mu = ctl[DO_MU];
    ctl[DO_MU] = mu - (mf < mu ? mf : mu);
    ctl[DO_REACHED] += nf;
    ctl[DO_NLEV] = level;
    if (nf == 0) ctl[DO_FLAGS] = f | DF_DONE;
    else if (level + 1 > 255) ctl[DO_FLAGS] = f | DF_DONE | DF_OVERFLOW;
    else ctl[DO_LEVEL] = level + 1;
}

// first[v] = the first entry of v's list (every vertex below nzn has one).
__global__ void k_first_entry(const uint64_t *__restrict__ off, const uint32_t *__restrict__ nbr, uint64_t nzn,
                              uint32_t *first) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nzn; v += (uint64_t)gridDim.x * blockDim.x)
        first[v] = nbr[off[v]];
}

// Sorts the lists of at most 64 entries ascending (one wave per list, bitonic
// network over the lanes).  With ids relabelled by descending degree, the
// bottom-up BFS then meets the heaviest neighbours -- the likeliest frontier
// members -- first.
__global__ void __launch_bounds__(256) k_sort_short_lists(const uint64_t *__restrict__ off, uint64_t v0, uint64_t v1,
                                                          uint32_t *nbr) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t v = v0 + ((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6); v < v1; v += nw) {
        const uint64_t b = off[v], d = off[v + 1] - b;
        if (d < 2 || d > 64) continue;  // wave-uniform
        uint32_t x = (uint64_t)lane < d ? nbr[b + lane] : 0xFFFFFFFFu;
#pragma unroll
        for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const uint32_t y = __shfl_xor(x, j);
                const bool up = (lane & k) == 0, lower = (lane & j) == 0;
                x = (lower == up) ? min(x, y) : max(x, y);
            }
        if ((uint64_t)lane < d) nbr[b + lane] = x;
    }
}

// Sorts the lists of 65 to MID_LIST entries ascending: one workgroup per list,
// a bitonic network over the list padded to a power of two in LDS.  Vertices
// v0 .. v1-1 are those whose degree bucket may hold such lists; the others
// are skipped (workgroup-uniform).  First entries then are the heaviest
// neighbours for these lists too, which the bottom-up first-entry check tests.
constexpr uint32_t MID_LIST = 4096;
__global__ void __launch_bounds__(256) k_sort_mid_lists(const uint64_t *__restrict__ off, uint64_t v0, uint64_t v1,
                                                        uint32_t *nbr) {
    __shared__ uint32_t s[MID_LIST];
    for (uint64_t v = v0 + blockIdx.x; v < v1; v += gridDim.x) {
        const uint64_t b = off[v];
        const uint32_t d = (uint32_t)(off[v + 1] - b);
        if (d <= 64 || d > MID_LIST) continue;
        uint32_t p = 128;
        while (p < d) p <<= 1;
        for (uint32_t i = threadIdx.x; i < p; i += 256) s[i] = i < d ? nbr[b + i] : 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= p; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < p; i += 256) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const uint32_t x = s[i], y = s[l];
                        if ((x > y) == ((i & k) == 0)) {
                            s[i] = y;
                            s[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        for (uint32_t i = threadIdx.x; i < d; i += 256) nbr[b + i] = s[i];
        __syncthreads();   // s is refilled for the next list
    }
}

// ---- relabelling by descending degree ------------------------------------------------

// Bucket of a degree: 8 sub-buckets per power of two; 0 for no edges.  Higher
// bucket = earlier position.
__host__ __device__ inline uint32_t deg_bucket(uint32_t d) {
    if (!d) return 0;
    uint32_t lg = 31 - __builtin_clz(d);
    uint32_t sub = lg >= 3 ? (d >> (lg - 3)) & 7 : 0;
    return 1 + lg * 8 + sub;  // 1 .. 256
}
constexpr int NBUCKET = 257;

__global__ void k_bucket_hist(const int32_t *__restrict__ deg, uint64_t n, unsigned long long *hist) {
    __shared__ uint32_t h[NBUCKET];
    for (int i = threadIdx.x; i < NBUCKET; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[deg_bucket((uint32_t)deg[v])], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < NBUCKET; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], (unsigned long long)h[i]);
}

// Each workgroup takes 4096 vertices: local ranks per bucket in LDS, one global
// reservation per (workgroup, bucket), then new_id / old_of.
__global__ void __launch_bounds__(256) k_relabel(const int32_t *__restrict__ deg, uint64_t n,
                                                 unsigned long long *cursor, uint32_t *new_id, uint32_t *old_of) {
    __shared__ uint32_t h[NBUCKET];
    __shared__ unsigned long long base[NBUCKET];
    for (int i = threadIdx.x; i < NBUCKET; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t v0 = (uint64_t)blockIdx.x * 4096;
    uint32_t key[16], rank[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        uint64_t v = v0 + i * 256 + threadIdx.x;
        key[i] = v < n ? deg_bucket((uint32_t)deg[v]) : 0;
        rank[i] = v < n ? atomicAdd(&h[key[i]], 1u) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NBUCKET; i += blockDim.x)
        if (h[i]) base[i] = atomicAdd(&cursor[i], (unsigned long long)h[i]);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        uint64_t v = v0 + i * 256 + threadIdx.x;
        if (v < n) {
            uint64_t p = base[key[i]] + rank[i];
            new_id[v] = (uint32_t)p;
            old_of[p] = (uint32_t)v;
        }
    }
}

__global__ void k_gather_deg(const int32_t *__restrict__ deg, const uint32_t *__restrict__ old_of, uint64_t n,
                             int32_t *deg_r) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n; p += (uint64_t)gridDim.x * blockDim.x)
        deg_r[p] = deg[old_of[p]];
}

// out[v] = x_r[new_id[v]] for 4-byte values, four vertices per thread (a
// 16-byte load of new_id, four gathers, a 16-byte store; out and new_id are
// device allocations, so 16-byte aligned).
template <class T>
__global__ void k_unpermute(const T *__restrict__ x_r, const uint32_t *__restrict__ new_id, uint64_t n, uint64_t lim,
                            T *out) {
    static_assert(sizeof(T) == 4, "4-byte values");
    // relabelled ids from lim on read as 0 (k-core: only the vertices with edges hold degrees)
    auto at = [&](uint32_t r) { return r < lim ? x_r[r] : T(0); };
    for (uint64_t v0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) * 4; v0 < n;
         v0 += (uint64_t)gridDim.x * blockDim.x * 4) {
        if (v0 + 4 <= n) {
            const uint4 q = *reinterpret_cast<const uint4 *>(new_id + v0);
            const T a = at(q.x), b = at(q.y), c = at(q.z), d = at(q.w);
            T *o = out + v0;
            *reinterpret_cast<uint4 *>(o) = uint4{__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b),
                                                  __builtin_bit_cast(uint32_t, c), __builtin_bit_cast(uint32_t, d)};
        } else {
            for (uint64_t v = v0; v < n; ++v) out[v] = at(new_id[v]);
        }
    }
}

// BFS depths by original id: 16 vertices per thread (four 16-byte loads of
// new_id, one 16-byte store).  Vertices without edges (relabelled ids from
// nzn on) are unreached unless one is the root, so only the others gather.
__global__ void __launch_bounds__(256) k_unpermute_depth(const uint8_t *__restrict__ d_r,
                                                         const uint32_t *__restrict__ new_id, uint64_t n, uint64_t nzn,
                                                         uint64_t root, uint8_t *out) {
    for (uint64_t v0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) * 16; v0 < n;
         v0 += (uint64_t)gridDim.x * blockDim.x * 16) {
        if (v0 + 16 <= n) {
            uint32_t id[16];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 q = reinterpret_cast<const uint4 *>(new_id + v0)[k];
                id[4 * k] = q.x; id[4 * k + 1] = q.y; id[4 * k + 2] = q.z; id[4 * k + 3] = q.w;
            }
            uint32_t wd[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                wd[k] = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t r = id[4 * k + j];
                    const uint32_t d = r < nzn ? d_r[r] : (v0 + 4 * k + j == root ? 1u : 0u);
                    wd[k] |= d << (8 * j);
                }
            }
            *reinterpret_cast<uint4 *>(out + v0) = uint4{wd[0], wd[1], wd[2], wd[3]};
        } else {
            for (uint64_t v = v0; v < n; ++v) {
                const uint32_t r = new_id[v];
                out[v] = r < nzn ? d_r[r] : (uint8_t)(v == root ? 1 : 0);
            }
        }
    }
}

// chunkz[c] = index (in nzv) of the vertex whose neighbour list holds entry c*PULL_CHUNK.
__global__ void k_pull_chunk_map(const uint64_t *__restrict__ offz, uint64_t nzn, uint32_t *chunkz) {
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < nzn; k += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t b = offz[k], e = offz[k + 1];
        for (uint64_t c = (b + PULL_CHUNK - 1) / PULL_CHUNK; c * PULL_CHUNK < e; ++c) chunkz[c] = (uint32_t)k;
    }
}

static unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

// Frees every array of the index; a failed build leaves none of it behind.
static void release_adjacency(gs_graph *g) {
    g->adj_ready = false;
    g->new_id.release();
    g->old_id.release();
    g->deg_r.release();
    g->adj_off.release();
    g->adj_nbr.release();
    g->adj_first.release();
    g->chunkz.release();
    g->nzn = g->npull = g->ms_nhub = 0;
}

static int build_adjacency_locked(gs_graph *g);

// Every device allocation of the index build goes through here, so a test can
// make the k-th one fail (gs_debug_fail_alloc) and check the clean-up.
template <class T>
static int index_alloc(gs_graph *g, DevBuf<T> &b, size_t count) {
    if (g->fail_alloc_at > 0 && --g->fail_alloc_at == 0) {
        set_error("device allocation %zu bytes failed (injected by gs_debug_fail_alloc)", count * sizeof(T));
        return GS_ERR_NOMEM;
    }
    return b.alloc(count);
}

// The index counts as built only once every step has finished on the device
// (adj_ready); any failure on the way (an allocation, the degree-sum check, a
// launch) releases what was allocated, so the next call builds again.
int build_adjacency(gs_graph *g) {
    std::lock_guard<std::mutex> lk(g->build_mu);
    if (g->adj_ready) return GS_OK;
    if (g->directed) { set_error("the adjacency index is built for undirected grids"); return GS_ERR_UNSUPPORTED; }
    const int rc = build_adjacency_locked(g);
    if (rc != GS_OK) {
        (void)hipStreamSynchronize(g->ws->stream);   // no kernel may still write the arrays being freed
        release_adjacency(g);
        return rc;
    }
    g->adj_ready = true;
    return GS_OK;
}

static int build_adjacency_locked(gs_graph *g) {
    gs_ws *w = g->ws;
    hipStream_t st = w->stream;
    const uint64_t n = g->n;
    EventPair ev;
    GS_HIP(hipEventCreate(&ev.a));
    GS_HIP(hipEventCreate(&ev.b));
    hipEvent_t t0 = ev.a, t1 = ev.b;
    GS_HIP(hipEventRecord(t0, st));
    // 1. relabel: histogram of degree buckets, descending bucket order
    DevBuf<unsigned long long> cur64;
    GS_TRY(index_alloc(g, cur64, NBUCKET));
    GS_HIP(hipMemsetAsync(cur64.ptr, 0, NBUCKET * 8, st));
    k_bucket_hist<<<vgrid(n), 256, 0, st>>>(g->degree.ptr, n, cur64.ptr);
    GS_HIP(hipGetLastError());
    std::vector<unsigned long long> h(NBUCKET);
    GS_HIP(hipMemcpyAsync(h.data(), cur64.ptr, NBUCKET * 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    g->nzn = n - h[0];
    g->ms_nhub = 0;   // degree >= 2^MS_HEAVY_LOG: buckets from 1 + 8 * MS_HEAVY_LOG up (deg_bucket)
    for (int b = 1 + 8 * (int)MS_HEAVY_LOG; b < NBUCKET; ++b) g->ms_nhub += h[b];
    unsigned long long run = 0;
    for (int b = NBUCKET - 1; b >= 0; --b) { unsigned long long c = h[b]; h[b] = run; run += c; }
    // ids of the buckets that hold degrees 65 .. MID_LIST (h[b] = first id of bucket b)
    const uint64_t mid_lo = h[deg_bucket(MID_LIST)], mid_hi = h[deg_bucket(65) - 1];
    GS_HIP(hipMemcpyAsync(cur64.ptr, h.data(), NBUCKET * 8, hipMemcpyHostToDevice, st));
    GS_TRY(index_alloc(g, g->new_id, n));
    GS_TRY(index_alloc(g, g->old_id, n));
    k_relabel<<<(unsigned)((n + 4095) / 4096), 256, 0, st>>>(g->degree.ptr, n, cur64.ptr, g->new_id.ptr,
                                                             g->old_id.ptr);
    GS_HIP(hipGetLastError());
    GS_TRY(index_alloc(g, g->deg_r, n));
    k_gather_deg<<<vgrid(n), 256, 0, st>>>(g->degree.ptr, g->old_id.ptr, n, g->deg_r.ptr);
    GS_HIP(hipGetLastError());
    // 2. offsets = exclusive scan of the relabelled degrees
    const uint64_t nb = (n + 1023) / 1024;
    DevBuf<unsigned long long> bsum;
    GS_TRY(index_alloc(g, bsum, nb));
    GS_TRY(index_alloc(g, g->adj_off, n + 1));
    GS_TRY(index_alloc(g, g->adj_nbr, std::max<uint64_t>(2 * g->E, 1)));
    k_deg_block_sums<<<(unsigned)nb, 256, 0, st>>>(g->deg_r.ptr, n, bsum.ptr);
    GS_HIP(hipGetLastError());
    std::vector<unsigned long long> hb(nb);
    GS_HIP(hipMemcpyAsync(hb.data(), bsum.ptr, nb * 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    run = 0;
    for (uint64_t i = 0; i < nb; ++i) { unsigned long long c = hb[i]; hb[i] = run; run += c; }
    if (run != 2 * g->E) {
        set_error("degree sum %llu differs from twice the stored edges %llu", run, (unsigned long long)(2 * g->E));
        return GS_ERR_FORMAT;
    }
    GS_HIP(hipMemcpyAsync(bsum.ptr, hb.data(), nb * 8, hipMemcpyHostToDevice, st));
    k_deg_scan<<<(unsigned)nb, 256, 0, st>>>(g->deg_r.ptr, n, bsum.ptr, g->adj_off.ptr);
    GS_HIP(hipGetLastError());
    // 3. neighbour lists, relabelled, by one sweep over the grid
    DevBuf<uint32_t> cur;
    GS_TRY(index_alloc(g, cur, n));
    GS_HIP(hipMemsetAsync(cur.ptr, 0, n * 4, st));
    GS_TRY(sweep(g, w, GS_K_CONVERT, [&](const SweepArgs &a, unsigned grid, hipStream_t s2) {
        k_adj_scatter<<<grid, SWEEP_WG, 0, s2>>>(a, g->new_id.ptr, g->adj_off.ptr, cur.ptr, g->adj_nbr.ptr);
    }));
    cur.release();
    if (g->nzn) {
        k_sort_short_lists<<<8192, 256, 0, st>>>(g->adj_off.ptr, 0, g->nzn, g->adj_nbr.ptr);
        GS_HIP(hipGetLastError());
        if (mid_hi > mid_lo) {
            k_sort_mid_lists<<<(unsigned)std::min<uint64_t>(mid_hi - mid_lo, 8192), 256, 0, st>>>(
                    g->adj_off.ptr, mid_lo, mid_hi, g->adj_nbr.ptr);
            GS_HIP(hipGetLastError());
        }
        GS_TRY(index_alloc(g, g->adj_first, g->nzn));
        k_first_entry<<<vgrid(g->nzn), 256, 0, st>>>(g->adj_off.ptr, g->adj_nbr.ptr, g->nzn, g->adj_first.ptr);
        GS_HIP(hipGetLastError());
    }
    // 4. pull chunk map over the vertices with edges (relabelled ids 0 .. nzn-1)
    const uint64_t total = 2 * g->E;
    g->npull = (total + PULL_CHUNK - 1) / PULL_CHUNK;
    GS_TRY(index_alloc(g, g->chunkz, g->npull + 1));
    if (g->nzn) {
        k_pull_chunk_map<<<vgrid(g->nzn), 256, 0, st>>>(g->adj_off.ptr, g->nzn, g->chunkz.ptr);
        GS_HIP(hipGetLastError());
        uint32_t last = (uint32_t)(g->nzn - 1);
        GS_HIP(hipMemcpyAsync(g->chunkz.ptr + g->npull, &last, 4, hipMemcpyHostToDevice, st));
    }
    GS_HIP(hipEventRecord(t1, st));
    GS_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    g->index_ms = ms;
    w->timer.collect();
    w->timer.reset();
    w->stats = gs_stats{};
    return GS_OK;
}

// ---- PageRank pull over the neighbour lists ----------------------------------------
// acc[v] = sum of prior over v's neighbours: one workgroup per PULL_CHUNK
// consecutive list entries; lanes read 16 entries each (4 x 16-byte loads), find
// their vertex by a search over the chunk's segment starts in LDS, and add
// runs of entries of one vertex in registers, then into an LDS slot per vertex.
// A vertex whose list lies inside the chunk is stored; one crossing a chunk
// edge is added atomically (at most two per chunk).  No per-edge atomics.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <class A>  // accumulator type: float (PageRank) or double (personalised PageRank)
struct PullTiles {
    uint16_t beg[PULL_CHUNK + 2];
    uint32_t vid[PULL_CHUNK + 1];
    A sum[PULL_CHUNK + 1];
    uint8_t full[PULL_CHUNK + 1];
    int nt;
};

template <class A>
__global__ void __launch_bounds__(256) k_pr_pull(const uint32_t *__restrict__ nbr, uint64_t nent,
                                                 const uint64_t *__restrict__ offz,
                                                 const uint32_t *__restrict__ chunkz, uint64_t nchunk,
                                                 const float *__restrict__ prior, A *acc) {
    __shared__ PullTiles<A> s;
    const uint64_t chunk = blockIdx.x;
    if (chunk >= nchunk) return;
    const uint64_t base = chunk * PULL_CHUNK;
    const uint32_t len = (uint32_t)min<uint64_t>(PULL_CHUNK, nent - base);
    const uint32_t k0 = chunkz[chunk], k1 = chunkz[chunk + 1];
    const int nt = (int)(k1 - k0) + 1;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) {
        uint64_t b = offz[k0 + k], e = offz[k0 + k + 1];
        s.beg[k] = (uint16_t)(b <= base ? 0 : min<uint64_t>(b - base, len));
        s.vid[k] = k0 + k;
        s.sum[k] = A(0);
        s.full[k] = b >= base && e <= base + len;
    }
    if (threadIdx.x == 0) {
        s.nt = nt;
        s.beg[nt] = (uint16_t)len;
    }
    __syncthreads();
    const uint32_t e0 = threadIdx.x * 16;
    if (e0 < len) {
        uint32_t idx[16];
        const uint32_t *p = nbr + base + e0;
        if (e0 + 16 <= len) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // streamed once per iteration: non-temporal, so it does not evict the gathered values
                u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p) + q);
                idx[4 * q] = v[0]; idx[4 * q + 1] = v[1]; idx[4 * q + 2] = v[2]; idx[4 * q + 3] = v[3];
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; ++i) idx[i] = e0 + i < len ? p[i] : 0u;
        }
        float val[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) val[i] = prior[idx[i]];  // idx is 0 past the end: in bounds, all loads in flight
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (e0 + i >= len) val[i] = 0.0f;
        int lo = 0, hi = nt - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (s.beg[mid] <= e0) lo = mid; else hi = mid - 1;
        }
        int k = lo;
        A run = A(0);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            uint32_t e = e0 + i;
            if (e >= len) break;
            while (s.beg[k + 1] <= e) {
                if (run != A(0)) atomicAdd(&s.sum[k], run);
                run = A(0);
                ++k;
            }
            run += (A)val[i];
        }
        if (run != A(0)) atomicAdd(&s.sum[k], run);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nt; k += blockDim.x) {
        A x = s.sum[k];
        if (s.full[k]) acc[s.vid[k]] = x;
        else if (x != A(0)) unsafeAtomicAdd(&acc[s.vid[k]], x);
    }
}

// w->qoff = exclusive prefix of w->qdeg[0 .. nq) (blocks of 1024, block starts
// on the host); *total (if given) = the sum.
int queue_prefix(gs_ws *w, uint64_t nq, std::vector<unsigned long long> &hsum, uint64_t *total) {
    const uint64_t nb = (nq + 1023) / 1024;
    k_deg_block_sums<<<(unsigned)nb, 256, 0, w->stream>>>(w->qdeg.ptr, nq, w->qsum.ptr);
    GS_HIP(hipGetLastError());
    hsum.resize(nb);
    GS_HIP(hipMemcpyAsync(hsum.data(), w->qsum.ptr, nb * 8, hipMemcpyDeviceToHost, w->stream));
    GS_HIP(hipStreamSynchronize(w->stream));
    unsigned long long run = 0;
    for (uint64_t i = 0; i < nb; ++i) { const unsigned long long c = hsum[i]; hsum[i] = run; run += c; }
    GS_HIP(hipMemcpyAsync(w->qsum.ptr, hsum.data(), nb * 8, hipMemcpyHostToDevice, w->stream));
    k_deg_scan<<<(unsigned)nb, 256, 0, w->stream>>>(w->qdeg.ptr, nq, w->qsum.ptr, w->qoff.ptr);
    GS_HIP(hipGetLastError());
    if (total) *total = run;
    return GS_OK;
}

int scan_counts(hipStream_t st, const int32_t *cnt, uint64_t n, uint64_t *off, uint64_t *total) {
    if (!n) {
        GS_HIP(hipMemsetAsync(off, 0, 8, st));
        if (total) *total = 0;
        return GS_OK;
    }
    const uint64_t nb = (n + 1023) / 1024;
    DevBuf<unsigned long long> bsum;
    GS_TRY(bsum.alloc(nb));
    k_deg_block_sums<<<(unsigned)nb, 256, 0, st>>>(cnt, n, bsum.ptr);
    GS_HIP(hipGetLastError());
    std::vector<unsigned long long> hb(nb);
    GS_HIP(hipMemcpyAsync(hb.data(), bsum.ptr, nb * 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    unsigned long long run = 0;
    for (uint64_t i = 0; i < nb; ++i) { const unsigned long long c = hb[i]; hb[i] = run; run += c; }
    GS_HIP(hipMemcpyAsync(bsum.ptr, hb.data(), nb * 8, hipMemcpyHostToDevice, st));
    k_deg_scan<<<(unsigned)nb, 256, 0, st>>>(cnt, n, bsum.ptr, off);
    GS_HIP(hipGetLastError());
    GS_HIP(hipStreamSynchronize(st));   // bsum and hb die here
    if (total) *total = run;
    return GS_OK;
}

int ensure_queue(const gs_graph *g, gs_ws *w) {
    if (w->queue.n != g->n) GS_TRY(w->queue.alloc(g->n));
    if (w->qdeg.n != g->n) GS_TRY(w->qdeg.alloc(g->n));
    if (w->qoff.n != g->n + 1) GS_TRY(w->qoff.alloc(g->n + 1));
    if (w->qsum.n != g->n / 1024 + 1) GS_TRY(w->qsum.alloc(g->n / 1024 + 1));
    return GS_OK;
}

int unpermute_i32(const gs_graph *g, hipStream_t st, const int32_t *x_r, int32_t *out) {
    k_unpermute<int32_t><<<vgrid((g->n + 3) / 4), 256, 0, st>>>(x_r, g->new_id.ptr, g->n, g->nzn, out);
    GS_HIP(hipGetLastError());
    return GS_OK;
}

int unpermute_f32(const gs_graph *g, hipStream_t st, const float *x_r, float *out, uint64_t lim) {
    k_unpermute<float><<<vgrid((g->n + 3) / 4), 256, 0, st>>>(x_r, g->new_id.ptr, g->n, lim, out);
    GS_HIP(hipGetLastError());
    return GS_OK;
}

int pr_pull(const gs_graph *g, gs_ws *w, const float *prior, float *acc, double *acc64) {
    if (!g->npull) return GS_OK;
    hipEvent_t kb;
    w->timer.begin(GS_K_PR_PULL, w->stream, &kb);
    if (acc64)
        k_pr_pull<double><<<(unsigned)g->npull, 256, 0, w->stream>>>(g->adj_nbr.ptr, 2 * g->E, g->adj_off.ptr,
                                                                     g->chunkz.ptr, g->npull, prior, acc64);
    else
        k_pr_pull<float><<<(unsigned)g->npull, 256, 0, w->stream>>>(g->adj_nbr.ptr, 2 * g->E, g->adj_off.ptr,
                                                                    g->chunkz.ptr, g->npull, prior, acc);
    GS_HIP(hipGetLastError());
    w->timer.end(GS_K_PR_PULL, w->stream, kb);
    return GS_OK;
}

}  // namespace gs

using namespace gs;

// Rate of the device's constant clock (wall_clock64), kHz.
static double wall_clock_khz() {
    static double khz = 0.0;
    if (khz == 0.0) {
        int dev = 0, rate = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev) == hipSuccess && rate > 0)
            khz = rate;
        else
            khz = 100000.0;   // the documented 100 MHz of gfx9 parts
    }
    return khz;
}

uint64_t gs::do_bfs_bytes(const gs_graph *g) {
    const uint64_t n = g->n;
    return 3 * 4 * bitmap_words(g) + 4 * ((g->p_s + 31) / 32) +  // bitmaps and tile-row bits (ensure_bitmaps)
           2 * n +                                               // depth, depth_r
           4 * n + 4 * n + 8 * (n + 1) + 8 * (n / 1024 + 1) +   // queue, qdeg, qoff, qsum (ensure_queue)
           8 * (uint64_t)DO_WORDS;                               // control block
}

int gs::run_bfs_do(const gs_graph *gc, gs_ws *w, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
                   int max_levels, int *levels) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    if (!g || root >= g->n || max_levels < 0) {
        set_error("gs_bfs: bad arguments");
        return GS_ERR_ARG;
    }
    GS_TRY(build_adjacency(g));
    GS_TRY(ensure_bitmaps(g, w));
    if (w->depth.n != g->n) GS_TRY(w->depth.alloc(g->n));
    if (w->depth_r.n != g->n) GS_TRY(w->depth_r.alloc(g->n));
    GS_TRY(ensure_queue(g, w));
    // bitmap words that hold vertices with edges: only they change after the root
    const uint64_t act_words = (g->nzn + 31) / 32;
    w->timer.reset();
    w->stats = gs_stats{};
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    // only the vertices with edges: every kernel of the search reads and writes
    // depths and bitmap words below nzn (act_words), and the unpermute gives the
    // others depth 0 (the root 1) without reading depth_r; what a root without
    // edges leaves above them is never read
    GS_HIP(hipMemsetAsync(w->depth_r.ptr, 0, g->nzn, w->stream));
    GS_HIP(hipMemsetAsync(w->bm_vis.ptr, 0, act_words * 4, w->stream));
    GS_HIP(hipMemsetAsync(w->bm_front.ptr, 0, act_words * 4, w->stream));
    GS_HIP(hipMemsetAsync(w->bm_next.ptr, 0, act_words * 4, w->stream));
    if (w->do_ctl.n != DO_WORDS) GS_TRY(w->do_ctl.alloc(DO_WORDS));
    unsigned long long *ctl = w->do_ctl.ptr;
    GS_HIP(hipMemsetAsync(ctl, 0, DO_WORDS * 8, w->stream));
    k_do_root<<<1, 1, 0, w->stream>>>(root, g->new_id.ptr, g->deg_r.ptr, g->nzn, 2 * g->E, (uint32_t)g->do_alpha,
                                      (uint32_t)g->do_beta, ctl, w->depth_r.ptr, w->bm_vis.ptr, w->bm_front.ptr,
                                      w->queue.ptr, w->qdeg.ptr);
    GS_HIP(hipGetLastError());
    // what the captured kernels were given (the thresholds travel in ctl)
    const std::vector<const void *> key = {g, (const void *)(uintptr_t)g->n, (const void *)(uintptr_t)g->nzn, ctl,
                                           w->bm_vis.ptr, w->bm_front.ptr, w->bm_next.ptr, w->depth_r.ptr, w->queue.ptr,
                                           w->qdeg.ptr, w->qoff.ptr, w->qsum.ptr, g->adj_off.ptr, g->adj_nbr.ptr,
                                           g->adj_first.ptr, g->deg_r.ptr};
    if (!w->do_graph || w->do_graph_key != key) {
        if (w->do_graph) {
            GS_HIP(hipGraphExecDestroy(w->do_graph));
            w->do_graph = nullptr;
        }
        const uint64_t nzn = g->nzn;
        // grid-stride kernels, at most DO_GRID workgroups: enough to fill the GPU, and a
        // kernel whose flag is down for the level costs the launch of fewer empty workgroups
        auto cap = [](unsigned x) { return std::max(1u, std::min(x, DO_GRID)); };
        const unsigned gle = cap(vgrid((act_words + LE_WPT - 1) / LE_WPT));
        const unsigned gul = cap(vgrid((act_words + 3) / 4));
        const unsigned gfull = cap(vgrid((nzn + BU_VPL - 1) / BU_VPL));
        const unsigned glist = cap(vgrid(nzn / BU_LIST_DIV / BU_VPL + 1));
        const unsigned gsc = (unsigned)std::min<uint64_t>(DO_SCAN_GRID, std::max<uint64_t>((nzn + 1023) / 1024, 1));
        uint32_t *vis = w->bm_vis.ptr, *front = w->bm_front.ptr, *next = w->bm_next.ptr;
        GS_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
        for (int l = 0; l < DO_LEVELS; ++l) {
            k_do_begin<<<1, 1, 0, w->stream>>>(ctl, g->n, nzn);
            // top-down: list the frontier if a bottom-up level made it, prefix of its degrees, edges
            k_do_level_end<false, true><<<gle, 256, 0, w->stream>>>(ctl, DF_LIST, vis, front, next, w->depth_r.ptr,
                                                                   g->deg_r.ptr, g->n, act_words, w->queue.ptr,
                                                                   w->qdeg.ptr, ctl + DO_LIST);
            k_do_block_sums<<<gsc, 256, 0, w->stream>>>(ctl, w->qdeg.ptr, w->qsum.ptr);
            k_do_scan<<<1, 1024, 0, w->stream>>>(ctl, w->qsum.ptr);
            k_do_deg_scan<<<gsc, 256, 0, w->stream>>>(ctl, w->qdeg.ptr, w->qsum.ptr, w->qoff.ptr);
            k_td_edges<<<DO_TD_GRID, TD_WG, 0, w->stream>>>(ctl, w->queue.ptr, w->qoff.ptr, g->adj_off.ptr,
                                                           g->adj_nbr.ptr, vis, next,
                                                           (uint32_t)std::min<uint64_t>(TD_HUB_WORDS, act_words));
            // bottom-up: over the unreached list (few left; the queue is free then) or over all
            k_unreached_list<<<gul, 256, 0, w->stream>>>(ctl, vis, nzn, w->queue.ptr);
            k_bu_expand<true><<<glist, 256, 0, w->stream>>>(ctl, nzn, w->queue.ptr, g->adj_off.ptr, g->adj_nbr.ptr,
                                                           g->adj_first.ptr, vis, front, next);
            k_bu_expand<false><<<gfull, 256, 0, w->stream>>>(ctl, nzn, nullptr, g->adj_off.ptr, g->adj_nbr.ptr,
                                                            g->adj_first.ptr, vis, front, next);
            k_do_level_end<true, false><<<gle, 256, 0, w->stream>>>(ctl, DF_LEVEL, vis, front, next, w->depth_r.ptr,
                                                                   g->deg_r.ptr, g->n, act_words, w->queue.ptr,
                                                                   w->qdeg.ptr, ctl);
            k_do_end<<<1, 1, 0, w->stream>>>(ctl);
        }
        const hipError_t launch = hipGetLastError();
        hipGraph_t graph = nullptr;
        const hipError_t end = hipStreamEndCapture(w->stream, &graph);
        if (launch != hipSuccess || end != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            set_error("BFS: capture of the level sequence failed: %s",
                      hipGetErrorString(launch != hipSuccess ? launch : end));
            return GS_ERR_DEVICE;
        }
        const hipError_t inst = hipGraphInstantiate(&w->do_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (inst != hipSuccess) {
            w->do_graph = nullptr;
            set_error("BFS: hipGraphInstantiate: %s", hipGetErrorString(inst));
            return GS_ERR_DEVICE;
        }
        w->do_graph_key = key;
    }
    unsigned long long *h = w->h_counters;   // pinned
    for (;;) {
        GS_HIP(hipGraphLaunch(w->do_graph, w->stream));
        GS_HIP(hipMemcpyAsync(&h[0], ctl + DO_FLAGS, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        if (h[0] & DF_DONE) break;
    }
    k_unpermute_depth<<<vgrid((g->n + 15) / 16), 256, 0, w->stream>>>(w->depth_r.ptr, g->new_id.ptr, g->n, g->nzn, root,
                                                                       w->depth.ptr);
    GS_HIP(hipGetLastError());
    GS_HIP(hipEventRecord(t1, w->stream));
    std::vector<unsigned long long> hc(DO_WORDS);
    GS_HIP(hipMemcpyAsync(hc.data(), ctl, DO_WORDS * 8, hipMemcpyDeviceToHost, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    const int nlev = (int)hc[DO_NLEV];
    const unsigned long long checked = hc[DO_CHECKED];
    const double khz = wall_clock_khz();
    w->levels.clear();
    for (int l = 0; l < nlev; ++l) {
        const unsigned long long *rec = hc.data() + DO_REC + DO_RECW * l;
        w->levels.push_back(gs_level_stat{(uint32_t)(l + 1), (int32_t)rec[0], rec[1], rec[2], 0, rec[3],
                                          (double)(rec[5] - rec[4]) / khz});
        if (level_counts && l < max_levels) level_counts[l] = rec[3];
    }
    int status = GS_OK;
    if (hc[DO_FLAGS] & DF_OVERFLOW) {
        set_error("BFS from %llu needs depth > 255, beyond the u8 depth encoding", (unsigned long long)root);
        status = GS_ERR_DEPTH_OVERFLOW;
    }
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)nlev;
    w->stats.edges_scanned = checked;
    if (levels) *levels = nlev;
    if (status != GS_OK) return status;
    if (depth_out) {
        GS_HIP(hipMemcpyAsync(depth_out, w->depth.ptr, g->n, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}
