// gs_pb.hip -- PageRank over a source-blocked copy of the neighbour-list index.
//
// The pull over neighbour lists (gs_adj.hip) reads prior[u] for every list
// entry at a random address; on a Kronecker graph of 2^28 vertices nearly
// every such read misses L2 and moves a whole cache line for 4 useful bytes.
// Propagation blocking (Beamer, Asanovic and Patterson, IPDPS 2017) turns the
// random reads into two streamed passes.  The relabelled ids with edges
// (0 .. nzn-1) are cut into blocks of PB_B = 2^15 vertices, whose FP32 values
// fill 128 KiB of LDS.  Every list entry (v <- u) is copied once into an array
// ordered by destination block q = v / PB_B, then source block p = u / PB_B
// (the run (p, q)), keeping the 15-bit source offset.  Every run starts on a
// window of PB_WIN = 16 entries and fills whole windows (on a Kronecker graph
// of 2^28 vertices the padding adds about 1.2% of entries: runs average ~900
// entries), and a run's entries are sorted by destination, so a window splits
// into segments of equal destination.  The accumulation reads a stream laid
// out per destination block: one value per entry (a plain block), or one per
// segment (a compacted block, chosen when its segments are few; bit 15 of a
// source offset marks a segment's last entry).
//   gather, per source block p: stage prior[p*PB_B ..] in LDS and, one lane
//           per window of the runs (p, *), write prior[u] per entry or the
//           sum per segment into the stream -- 2 bytes read per entry, 16-byte
//           loads, vector stores;
//   accum,  per destination block q: add the stream values into an LDS array
//           over the block at their destination offsets (one LDS atomic per
//           run of equal offsets in a lane) and write the new priors.
// A destination block with more values than one workgroup takes is split
// into units whose sums go to a partial buffer, added by k_pb_reduce.  The
// order of the FP32 additions is the only difference from the other methods.
#include <algorithm>
#include <vector>

#include "gs_internal.h"

#include <hipcub/block/block_radix_sort.hpp>

namespace gs {

namespace {

constexpr int PB_WG = 1024;            // threads per workgroup of the three per-iteration kernels
constexpr int PB_PASS = 4096;          // build: list entries staged per pass
constexpr uint64_t PB_SLICE = 65536;   // build: list entries per workgroup

unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

// ---- build -------------------------------------------------------------------

__global__ void k_pb_bin_begin(const uint64_t *__restrict__ off, uint64_t nzn, uint32_t K, uint64_t *beg) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q <= K; q += gridDim.x * blockDim.x)
        beg[q] = off[std::min<uint64_t>((uint64_t)q << PB_SHIFT, nzn)];
}

// Entries per run: cnt[q*K + p] over one slice of the lists of block q.
__global__ void __launch_bounds__(PB_WG) k_pb_count(const PbSlice *__restrict__ sl, const uint32_t *__restrict__ nbr,
                                                    uint32_t K, unsigned long long *cnt) {
    __shared__ uint32_t hist[PB_KMAX];
    const PbSlice s = sl[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < K; k += PB_WG) hist[k] = 0;
    __syncthreads();
    for (uint64_t e = s.e0 + threadIdx.x; e < s.e1; e += PB_WG) atomicAdd(&hist[nbr[e] >> PB_SHIFT], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += PB_WG)
        if (hist[k]) atomicAdd(&cnt[(uint64_t)s.q * K + k], (unsigned long long)hist[k]);
}

// Copy one slice of the lists of block q into the runs (p, q): each workgroup
// reserves a contiguous stretch of every run it feeds (one global atomic per
// run), then places its entries in list order, so within a stretch a run's
// entries are sorted by destination and a destination's entries from this
// slice are adjacent (the accumulation then adds them in registers).  Per pass
// of PB_PASS entries: a stable block radix sort by source block gives each
// entry its rank among the pass's entries of the same run.  cur[q*K + p]
// starts at the run's first position.
typedef hipcub::BlockRadixSort<uint32_t, PB_WG, PB_PASS / PB_WG, uint32_t> PbPassSort;

__global__ void __launch_bounds__(PB_WG) k_pb_fill(const PbSlice *__restrict__ sl, const uint64_t *__restrict__ off,
                                                   const uint32_t *__restrict__ nbr, uint64_t nzn, uint32_t K,
                                                   unsigned long long *cur, uint16_t *__restrict__ src_out,
                                                   uint16_t *__restrict__ dst_out) {
    constexpr int IPT = PB_PASS / PB_WG;   // entries per thread per pass, consecutive (blocked arrangement)
    constexpr uint32_t NONE = PB_KMAX;     // sort key of the slots past the slice's end
    __shared__ uint32_t head[PB_KMAX];     // slice counts, then per pass: first sorted slot of each run
    __shared__ unsigned long long base[PB_KMAX];
    __shared__ union {
        uint64_t voff[PB_PASS + 2];
        typename PbPassSort::TempStorage sort;
        uint16_t key[PB_PASS];
    } sh;
    __shared__ uint64_t v_first;
    __shared__ int nvs;
    const PbSlice s = sl[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < K; k += PB_WG) head[k] = 0;
    __syncthreads();
    for (uint64_t e = s.e0 + threadIdx.x; e < s.e1; e += PB_WG) atomicAdd(&head[nbr[e] >> PB_SHIFT], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < K; k += PB_WG) {
        uint32_t c = head[k];
        base[k] = c ? atomicAdd(&cur[(uint64_t)s.q * K + k], (unsigned long long)c) : 0ull;
    }
    const uint64_t vlo = (uint64_t)s.q << PB_SHIFT, vhi = std::min<uint64_t>(vlo + PB_B, nzn);
    for (uint64_t a = s.e0; a < s.e1; a += PB_PASS) {
        const uint64_t b = std::min<uint64_t>(a + PB_PASS, s.e1);
        __syncthreads();
        if (threadIdx.x == 0) {
            // largest v in [vlo, vhi) with off[v] <= a, then with off[v] <= b - 1
            uint64_t lo = vlo, hi = vhi - 1;
            while (lo < hi) {
                uint64_t mid = (lo + hi + 1) >> 1;
                if (off[mid] <= a) lo = mid; else hi = mid - 1;
            }
            uint64_t v0 = lo;
            hi = vhi - 1;
            while (lo < hi) {
                uint64_t mid = (lo + hi + 1) >> 1;
                if (off[mid] <= b - 1) lo = mid; else hi = mid - 1;
            }
            v_first = v0;
            nvs = (int)(lo - v0 + 1);
        }
        __syncthreads();
        const int nv = nvs;
        const uint64_t v0 = v_first;
        for (int k = threadIdx.x; k <= nv; k += PB_WG) sh.voff[k] = off[v0 + k];
        __syncthreads();
        uint32_t key[IPT], val[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint64_t e = a + threadIdx.x * IPT + j;
            key[j] = NONE;
            val[j] = 0;
            if (e < b) {
                int lo = 0, hi = nv - 1;
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (sh.voff[mid] <= e) lo = mid; else hi = mid - 1;
                }
                const uint32_t u = nbr[e];
                key[j] = u >> PB_SHIFT;
                val[j] = (u & (PB_B - 1)) | (uint32_t)((v0 + lo) & (PB_B - 1)) << 16;
            }
        }
        __syncthreads();   // voff is dead; its storage becomes the sort's
        PbPassSort(sh.sort).Sort(key, val, 0, 14);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) sh.key[threadIdx.x * IPT + j] = (uint16_t)key[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t i = threadIdx.x * IPT + j;
            if (key[j] != NONE && (i == 0 || sh.key[i - 1] != key[j])) head[key[j]] = i;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (key[j] == NONE) continue;
            const uint32_t i = threadIdx.x * IPT + j;
            const unsigned long long pos = base[key[j]] + (i - head[key[j]]);
            src_out[pos] = (uint16_t)(val[j] & 0xFFFFu);
            dst_out[pos] = (uint16_t)(val[j] >> 16);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            const uint32_t i = threadIdx.x * IPT + j;
            if (key[j] != NONE && (i + 1 == PB_PASS || sh.key[i + 1] != key[j])) base[key[j]] += i - head[key[j]] + 1;
        }
    }
}

// After the fill, cur[i] is the end of run i's entries; the rest of its last
// window becomes padding: source offset PB_B (a zero in the gather's LDS) and
// the destination offset of the run's last entry.
__global__ void k_pb_pad(const unsigned long long *__restrict__ cur, uint64_t nrun, uint16_t *__restrict__ src,
                         uint16_t *__restrict__ dst) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nrun; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t end = cur[i];
        if ((end & (PB_WIN - 1)) == 0) continue;
        const uint16_t d = dst[end - 1];
        for (uint64_t j = end; j & (PB_WIN - 1); ++j) {
            src[j] = (uint16_t)PB_B;
            dst[j] = d;
        }
    }
}

// Per q-major window: a bit for the last entry of each segment of equal
// destinations (padding, source offset PB_B, never ends one; a segment never
// crosses a window) and the segment count.
__global__ void k_pb_wmask(const uint16_t *__restrict__ src, const uint16_t *__restrict__ dst, uint64_t nwin,
                           uint16_t *__restrict__ wmask, int32_t *__restrict__ wcnt) {
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwin; w += (uint64_t)gridDim.x * blockDim.x) {
        uint16_t s[PB_WIN], d[PB_WIN];
#pragma unroll
        for (int i = 0; i < PB_WIN; ++i) {
            s[i] = src[w * PB_WIN + i];
            d[i] = dst[w * PB_WIN + i];
        }
        uint32_t m = 0;
#pragma unroll
        for (int i = 0; i < PB_WIN; ++i) {
            const bool real = s[i] != (uint16_t)PB_B;
            const bool last = i + 1 == PB_WIN || s[i + 1] == (uint16_t)PB_B || d[i + 1] != d[i];
            if (real && last) m |= 1u << i;
        }
        wmask[w] = (uint16_t)m;
        wcnt[w] = __popc(m);
    }
}

// out[i] = a[idx[i]]
__global__ void k_pb_pick(const uint64_t *__restrict__ a, const uint64_t *__restrict__ idx, uint64_t n,
                          uint64_t *__restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = a[idx[i]];
}

// wcoff[w] = segments before window w within its run (a wave per run).
__global__ void k_pb_wcoff(const uint64_t *__restrict__ run_pw, const uint64_t *__restrict__ run_qw, uint64_t nruns,
                           const uint64_t *__restrict__ cpos, uint32_t *__restrict__ wcoff) {
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t r = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; r < nruns; r += nw) {
        const uint64_t q0 = run_qw[r], n = run_pw[r + 1] - run_pw[r], base = cpos[q0];
        for (uint64_t j = lane; j < n; j += 64) wcoff[q0 + j] = (uint32_t)(cpos[q0 + j] - base);
    }
}

// The stream's destination offsets, per q-major window: the window's 16
// offsets for a plain block, one offset per segment for a compacted block.
// blk_w[q] = first window of block q (K + 1 entries), blk_sb[q] = its stream
// base (PB_COMPACT set when compacted), blk_c0[q] = segments before it.
__global__ void k_pb_stream_dst(const uint16_t *__restrict__ dst, const uint16_t *__restrict__ wmask, uint64_t nwin,
                                const uint64_t *__restrict__ cpos, const uint64_t *__restrict__ blk_w,
                                const uint64_t *__restrict__ blk_sb, const uint64_t *__restrict__ blk_c0, uint32_t K,
                                uint16_t *__restrict__ out) {
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwin; w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = K - 1;   // largest q with blk_w[q] <= w
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (blk_w[mid] <= w) lo = mid; else hi = mid - 1;
        }
        const uint64_t sb = blk_sb[lo];
        if (sb & PB_COMPACT) {
            uint16_t *o = out + (sb & ~PB_COMPACT) + (cpos[w] - blk_c0[lo]);
            const uint32_t m = wmask[w];
            uint32_t k = 0;
            for (int i = 0; i < PB_WIN; ++i)
                if ((m >> i) & 1u) o[k++] = dst[w * PB_WIN + i];
        } else {
            uint16_t *o = out + sb + (w - blk_w[lo]) * PB_WIN;
            for (int i = 0; i < PB_WIN; ++i) o[i] = dst[w * PB_WIN + i];
        }
    }
}

// In the windows of compacted blocks, bit 15 of a source offset marks the last
// entry of a segment and padding becomes offset 0 without the mark (it follows
// the window's last mark, so the gather never stores it).
__global__ void k_pb_mark(uint16_t *__restrict__ src, const uint16_t *__restrict__ wmask, uint64_t nwin,
                          const uint64_t *__restrict__ blk_w, const uint64_t *__restrict__ blk_sb, uint32_t K) {
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < nwin; w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = K - 1;   // largest q with blk_w[q] <= w
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (blk_w[mid] <= w) lo = mid; else hi = mid - 1;
        }
        if (!(blk_sb[lo] & PB_COMPACT)) continue;
        const uint32_t m = wmask[w];
        for (int i = 0; i < PB_WIN; ++i) {
            const uint16_t s = src[w * PB_WIN + i];
            src[w * PB_WIN + i] = s == (uint16_t)PB_B ? (uint16_t)0 : (uint16_t)(s | (((m >> i) & 1u) << 15));
        }
    }
}

// ---- per iteration -----------------------------------------------------------

constexpr int PB_WPL = 4;   // gather: windows per lane per step, all loads issued first
constexpr uint32_t PB_CHUNKS = PB_UNIT_WIN / 64;   // gather: 64-window chunks of a unit

// prior[pbase .. pbase + pn) into LDS with 16-byte loads, two per thread in
// flight at a time (pbase is a multiple of 2^15, so the loads are aligned).
// Staging happens once per source block a workgroup meets, so it keeps few
// registers live beside the window loop rather than all its loads in flight.
__device__ inline void stage_block(const float *__restrict__ prior, uint64_t pbase, uint32_t pn, float4 *pri4) {
    const float4 *p4 = reinterpret_cast<const float4 *>(prior + pbase);
    const uint32_t n4 = pn >> 2;
    constexpr uint32_t PER = PB_B / 4 / PB_WG, IN = 2;
#pragma unroll
    for (uint32_t j0 = 0; j0 < PER; j0 += IN) {
        float4 v[IN];
#pragma unroll
        for (uint32_t j = 0; j < IN; ++j) {
            const uint32_t i = threadIdx.x + (j0 + j) * PB_WG;
            v[j] = i < n4 ? p4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (uint32_t j = 0; j < IN; ++j) {
            const uint32_t i = threadIdx.x + (j0 + j) * PB_WG;
            if (i < n4) pri4[i] = v[j];
        }
    }
    float *pri = reinterpret_cast<float *>(pri4);
    for (uint32_t i = (n4 << 2) + threadIdx.x; i < pn; i += PB_WG) pri[i] = prior[pbase + i];
}

typedef unsigned int pb_u32x4 __attribute__((ext_vector_type(4)));
typedef float pb_f32x4 __attribute__((ext_vector_type(4)));

// Segment sums of one compacted window (at most S segments) into out[0 ..
// segments): the running sum of each segment lands in register slot `seg` by
// selects (no indexing), then leaves in 16-, 8- and 4-byte stores at 4-byte
// alignment.
template <int S>
__device__ inline void pb_segments(const pb_u32x4 (&sq)[2], const float *pri, float *out) {
    float acc = 0.0f, sl[S];
#pragma unroll
    for (int k = 0; k < S; ++k) sl[k] = 0.0f;
    uint32_t seg = 0;
#pragma unroll
    for (int i = 0; i < PB_WIN; ++i) {
        const uint32_t s = (sq[i >> 3][(i >> 1) & 3] >> ((i & 1) * 16)) & 0xFFFFu;
        acc += pri[s & 0x7FFFu];
        const bool e = s >> 15;
#pragma unroll
        for (int k = 0; k < S && k <= i; ++k) sl[k] = (e && seg == (uint32_t)k) ? acc : sl[k];
        seg += e;
        acc = e ? 0.0f : acc;
    }
    const uint32_t n = seg;
#pragma unroll
    for (int g4 = 0; g4 < S / 4; ++g4)
        if (n >= 4u * (g4 + 1)) {
            const pb_f32x4 a = {sl[4 * g4], sl[4 * g4 + 1], sl[4 * g4 + 2], sl[4 * g4 + 3]};
            __builtin_memcpy(out + 4 * g4, &a, 16);
        }
    const uint32_t b = n & ~3u, r = n & 3;
    float t0 = sl[0], t1 = sl[1], t2 = S > 2 ? sl[S > 2 ? 2 : 0] : 0.0f;
#pragma unroll
    for (int g4 = 1; g4 < S / 4; ++g4)
        if (b == 4u * g4) {
            t0 = sl[4 * g4];
            t1 = sl[4 * g4 + 1];
            t2 = sl[4 * g4 + 2];
        }
    if (r >= 2) {
        const float2 a = {t0, t1};
        __builtin_memcpy(out + b, &a, 8);
    }
    if (r & 1) out[b + (r & 2)] = (r & 2) ? t2 : t0;
}

// The same for windows with up to 16 segments, a quarter (4 entries) at a
// time.  The running sums a0..a3 (reset after each mark) are formed first;
// the quarter's sums, in order, are then picked by the marks themselves (7
// selects per quarter, where a slot per segment count took 40 selects and
// compares per window) and leave at the quarter's place, after the segments
// of the quarters before it.  The running sum carries over quarters, so a
// segment may span them; each value is the same FP32 sum in the same order as
// one accumulator over the window gives.
__device__ inline void pb_segments_q(const pb_u32x4 (&sq)[2], const float *pri, float *out) {
    float acc = 0.0f;
    uint32_t base = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t w0 = sq[q >> 1][(q & 1) * 2], w1 = sq[q >> 1][(q & 1) * 2 + 1];
        const uint32_t s0 = w0 & 0xFFFFu, s1 = w0 >> 16, s2 = w1 & 0xFFFFu, s3 = w1 >> 16;
        const bool e0 = (s0 >> 15) != 0, e1 = (s1 >> 15) != 0, e2 = (s2 >> 15) != 0, e3 = (s3 >> 15) != 0;
        const float a0 = acc + pri[s0 & 0x7FFFu];
        const float a1 = (e0 ? 0.0f : a0) + pri[s1 & 0x7FFFu];
        const float a2 = (e1 ? 0.0f : a1) + pri[s2 & 0x7FFFu];
        const float a3 = (e2 ? 0.0f : a2) + pri[s3 & 0x7FFFu];
        acc = e3 ? 0.0f : a3;
        // k-th marked sum of the quarter: o0 first, o1 second, o2 third, a3 fourth
        const float y = e2 ? a2 : a3;
        const float x = e1 ? a1 : y;
        const float o0 = e0 ? a0 : x;
        const float o1 = e0 ? x : (e1 ? y : a3);
        const float o2 = (e0 && e1) ? y : a3;
        const uint32_t seg = (uint32_t)e0 + (uint32_t)e1 + (uint32_t)e2 + (uint32_t)e3;
        float *o = out + base;
        if (seg == 4) {
            const pb_f32x4 v = {o0, o1, o2, a3};
            __builtin_memcpy(o, &v, 16);
        } else {
            if (seg & 2) {
                const float2 v = {o0, o1};
                __builtin_memcpy(o, &v, 8);
            }
            if (seg & 1) o[seg & 2] = (seg & 2) ? o2 : o0;
        }
        base += seg;
    }
}

// Values of the windows [w0, w1) of source block p (counted in source-major
// order) into the accumulation stream; the unit's runs are staged in LDS.
// Lane i of the workgroup takes windows w0 + i, w0 + i + PB_WG, ... (adjacent
// lanes mostly share a run).  A wave's 64 windows form one 64-window chunk of
// the unit: a window's run is the run holding the chunk's first window plus
// the runs starting in the chunk at or before it, a popcount of the chunk's
// run-start bits under the lane's mask (an 11-step search over the staged run
// starts had cost a fifth of the kernel's VALU work).  A run of a plain destination block writes
// val[e] = prior[p*PB_B + src[e]] for each of its 16 entries (four 16-byte
// stores); a run of a compacted block writes one sum per segment of equal
// destinations (bit 15 of a source offset marks each segment's last entry,
// padding follows the last mark) at the window's place in the compacted stream.
// A window's sums collect in registers and leave in vector stores (one scalar
// store per segment made the gather 30% slower at Kron-28): pb_segments with
// two slots when no window of the wave has more segments (the heaviest
// blocks), else pb_segments_q a quarter window at a time.
//
// Work split: with grab == 0 each workgroup takes one unit; otherwise one
// workgroup per CU stays resident and takes `grab` consecutive units at a time
// from a counter (ctr, zero at launch) until none are left, staging a prior
// block only when its source block changes -- consecutive units mostly share
// one (a heavy source block spans many units), so most units skip the 128 KiB
// staging and the barrier behind it.
__global__ void __launch_bounds__(PB_WG) k_pb_gather(const PbUnit1 *__restrict__ units, uint32_t nunits,
                                                     uint32_t grab, unsigned long long *ctr,
                                                     const uint64_t *__restrict__ run_pw,
                                                     const uint64_t *__restrict__ run_qw,
                                                     const uint64_t *__restrict__ run_st,
                                                     const uint16_t *__restrict__ src,
                                                     const uint32_t *__restrict__ wcoff,
                                                     const float *__restrict__ prior, uint64_t nzn,
                                                     float *__restrict__ val) {
    __shared__ float4 pri4[PB_B / 4 + 1];   // + the zero read by padding entries (offset PB_B)
    // per staged run: first window in the unit (< PB_UNIT_WIN), first q-major
    // window (< 2^32, checked at build) and stream offset
    __shared__ uint16_t rs[PB_RUNS];
    __shared__ uint32_t rq[PB_RUNS];
    __shared__ uint64_t rt[PB_RUNS];
    __shared__ unsigned long long cbits[PB_CHUNKS];   // runs starting inside each 64-window chunk (bit 0 unused)
    __shared__ uint16_t ck0[PB_CHUNKS];               // the run holding each chunk's first window
    __shared__ uint32_t s_grab;
    if (threadIdx.x == 0) pri4[PB_B / 4] = make_float4(0.f, 0.f, 0.f, 0.f);
    const float *pri = reinterpret_cast<const float *>(pri4);
    const uint32_t lane = threadIdx.x & 63;   // window w of the unit sits in lane w mod 64 of some wave
    const unsigned long long lmask = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    uint32_t ui = grab ? 0 : blockIdx.x, uend = grab ? 0 : blockIdx.x + 1;
    uint32_t staged = 0xFFFFFFFFu;
    for (;;) {
        if (ui >= uend) {
            if (!grab) break;
            __syncthreads();   // every lane has read the previous grab
            if (threadIdx.x == 0) s_grab = (uint32_t)atomicAdd(ctr, 1ull);
            __syncthreads();
            // the same in every lane: read it into a scalar register, so the unit
            // records below load through the scalar cache and hold no vector registers
            const uint64_t first = (uint64_t)__builtin_amdgcn_readfirstlane(s_grab) * grab;
            if (first >= nunits) break;   // the same value in every lane: the whole workgroup leaves
            ui = (uint32_t)first;
            uend = (uint32_t)std::min<uint64_t>(nunits, first + grab);
        }
        const PbUnit1 u = units[ui++];
        __syncthreads();   // the previous unit's reads of the prior block and run tables are done
        if (u.p != staged) {
            const uint64_t pbase = (uint64_t)u.p << PB_SHIFT;
            stage_block(prior, pbase, (uint32_t)std::min<uint64_t>(PB_B, nzn - pbase), pri4);
            staged = u.p;
        }
        const int nr = (int)(u.r1 - u.r0);
        for (int k = threadIdx.x; k < nr; k += PB_WG) {
            const uint64_t r0 = run_pw[u.r0 + k];
            const uint64_t s = r0 > u.w0 ? r0 : u.w0;
            const uint64_t t = run_st[u.r0 + k];
            rs[k] = (uint16_t)(s - u.w0);
            rq[k] = (uint32_t)(run_qw[u.r0 + k] + (s - r0));
            // plain runs: the stream follows the windows; compacted: wcoff is relative to the run
            rt[k] = (t & PB_COMPACT) ? t : t + (s - r0) * PB_WIN;
        }
        const uint32_t len = (uint32_t)(u.w1 - u.w0);   // <= PB_CHUNKS * 64 (the unit tables are built so)
        for (uint32_t c = threadIdx.x; c < PB_CHUNKS; c += PB_WG) cbits[c] = 0;
        __syncthreads();
        // runs are non-empty and adjacent, rs[0] = 0: run k holds windows [rs[k], rs[k+1])
        for (int k = threadIdx.x; k < nr; k += PB_WG) {
            const uint32_t s = rs[k], e = k + 1 < nr ? rs[k + 1] : len;
            if (s & 63) atomicOr(&cbits[s >> 6], 1ull << (s & 63));
            for (uint32_t c = (s + 63) >> 6; (c << 6) < e; ++c) ck0[c] = (uint16_t)k;
        }
        __syncthreads();
        for (uint32_t wb = threadIdx.x; wb < len; wb += PB_WPL * PB_WG) {
            uint64_t qw[PB_WPL], st[PB_WPL];
            pb_u32x4 sq[PB_WPL][2];
            uint32_t co[PB_WPL];
#pragma unroll
            for (int j = 0; j < PB_WPL; ++j) {
                // past the unit, a lane reloads its first window and stores nothing
                const uint32_t w = wb + j * PB_WG < len ? wb + j * PB_WG : wb;
                const int k = ck0[w >> 6] + __popcll(cbits[w >> 6] & lmask);
                qw[j] = (uint64_t)rq[k] + (w - rs[k]);
                st[j] = rt[k] & PB_COMPACT ? rt[k] : rt[k] + (uint64_t)(w - rs[k]) * PB_WIN;
                // plain loads: nontemporal ones made the gather 11% slower at Kron-28
                const pb_u32x4 *s4 = reinterpret_cast<const pb_u32x4 *>(src + qw[j] * PB_WIN);
                sq[j][0] = s4[0];
                sq[j][1] = s4[1];
                co[j] = st[j] & PB_COMPACT ? wcoff[qw[j]] : 0u;   // plain windows need no offset
            }
#pragma unroll
            for (int j = 0; j < PB_WPL; ++j) {
                if (wb + j * PB_WG >= len) break;
                if (st[j] & PB_COMPACT) {
                    float *out = val + (st[j] & ~PB_COMPACT) + co[j];
                    uint32_t n = 0;   // segments: the marks of the window
#pragma unroll
                    for (int h = 0; h < 8; ++h) n += __popc(sq[j][h >> 2][h & 3] & 0x80008000u);
                    // two register slots when the wave's windows allow (wave-uniform branch)
                    if (__ballot(n > 2)) pb_segments_q(sq[j], pri, out);
                    else pb_segments<2>(sq[j], pri, out);
                } else {
                    float x[PB_WIN];
#pragma unroll
                    for (int h = 0; h < 8; ++h) {
                        const uint32_t a = sq[j][h >> 2][h & 3];
                        x[2 * h] = pri[a & 0xFFFFu];
                        x[2 * h + 1] = pri[a >> 16];
                    }
                    // plain stores: nontemporal ones measured four times slower here
                    pb_f32x4 *v4 = reinterpret_cast<pb_f32x4 *>(val + st[j]);
#pragma unroll
                    for (int h = 0; h < 4; ++h) v4[h] = pb_f32x4{x[4 * h], x[4 * h + 1], x[4 * h + 2], x[4 * h + 3]};
                }
            }
        }
    }
}

// FP32 add into LDS by a 32-bit compare-and-swap loop.  On this GPU the LDS
// float atomic costs about four times a streamed entry when the addresses of a
// wave are spread (blocks of low-degree vertices), while a CAS that rarely
// retries costs about as much as an integer atomic; where a wave's lanes share
// addresses (blocks of the heaviest vertices, whose runs are sorted by
// destination) the CAS retries and the hardware float atomic is used.
__device__ inline void lds_add_cas(float *p, float x) {
    unsigned *a = reinterpret_cast<unsigned *>(p);
    unsigned old = *a, assumed;
    do {
        assumed = old;
        old = atomicCAS(a, assumed, __float_as_uint(__uint_as_float(assumed) + x));
    } while (old != assumed);
}

__global__ void __launch_bounds__(PB_WG) k_pb_accum(const PbUnit2 *__restrict__ units,
                                                    const uint16_t *__restrict__ dst,
                                                    const float *__restrict__ val, uint64_t nzn,
                                                    const float *__restrict__ fac, float *__restrict__ prior,
                                                    float *__restrict__ part) {
    __shared__ float sum[PB_B];
    const PbUnit2 u = units[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < PB_B; i += PB_WG) sum[i] = 0.0f;
    __syncthreads();
    uint32_t cur = 0xFFFFFFFFu;
    float run = 0.0f;
    auto flush = [&]() {
        if (cur == 0xFFFFFFFFu) return;
        if (u.dense) atomicAdd(&sum[cur], run);
        else lds_add_cas(&sum[cur], run);
    };
    auto add = [&](uint32_t d, float x) {
        if (d != cur) {
            flush();
            cur = d;
            run = x;
        } else {
            run += x;
        }
    };
    // 16 consecutive values per lane: runs of one destination (the
    // heaviest vertices' segments continue over windows) add once.  Plain
    // loads: nontemporal ones made this kernel half as slow again at Kron-28.
    for (uint64_t sb = u.e0 & ~15ull; sb < u.e1; sb += 16 * PB_WG) {
        const uint64_t e0 = sb + threadIdx.x * 16;
        if (e0 >= u.e1) break;
        pb_u32x4 dq[2], vq[4];
#pragma unroll
        for (int k = 0; k < 2; ++k) dq[k] = *(reinterpret_cast<const pb_u32x4 *>(dst + e0) + k);
#pragma unroll
        for (int k = 0; k < 4; ++k) vq[k] = *(reinterpret_cast<const pb_u32x4 *>(val + e0) + k);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const uint64_t e = e0 + i;
            if (e < u.e0 || e >= u.e1) continue;
            add((dq[i >> 3][(i >> 1) & 3] >> ((i & 1) * 16)) & 0xFFFFu, __uint_as_float(vq[i >> 2][i & 3]));
        }
    }
    flush();
    __syncthreads();
    const uint64_t qbase = (uint64_t)u.q << PB_SHIFT;
    const uint32_t qn = (uint32_t)std::min<uint64_t>(PB_B, nzn - qbase);
    if (u.slot == 0xFFFFFFFFu) {  // the block's only unit: its sums are final
        for (uint32_t i = threadIdx.x; i < qn; i += PB_WG) prior[qbase + i] = pr_new_prior(sum[i], fac[qbase + i]);
    } else {
        float *pp = part + (uint64_t)u.slot * PB_B;
        for (uint32_t i = threadIdx.x; i < qn; i += PB_WG) pp[i] = sum[i];
    }
}

// FP64 variant for personalised PageRank: the same entries, summed in FP64
// into an LDS array over half a destination block (2^14 doubles) per pass, so
// each unit reads its entries twice.
__device__ inline void lds_add_cas64(double *p, double x) {
    unsigned long long *a = reinterpret_cast<unsigned long long *>(p);
    unsigned long long old = *a, assumed;
    do {
        assumed = old;
        old = atomicCAS(a, assumed, (unsigned long long)__double_as_longlong(__longlong_as_double((long long)assumed) + x));
    } while (old != assumed);
}

__global__ void __launch_bounds__(PB_WG) k_pb_accum64(const PbUnit2 *__restrict__ units,
                                                      const uint16_t *__restrict__ dst,
                                                      const float *__restrict__ val, uint64_t nzn,
                                                      double *__restrict__ acc, double *__restrict__ part) {
    constexpr uint32_t H = PB_B / 2;
    __shared__ double sum[H];
    const PbUnit2 u = units[blockIdx.x];
    const uint64_t qbase = (uint64_t)u.q << PB_SHIFT;
    const uint32_t qn = (uint32_t)std::min<uint64_t>(PB_B, nzn - qbase);
    for (uint32_t lo = 0; lo < qn; lo += H) {
        for (uint32_t i = threadIdx.x; i < H; i += PB_WG) sum[i] = 0.0;
        __syncthreads();
        for (uint64_t sb = u.e0 & ~15ull; sb < u.e1; sb += 16 * PB_WG) {
            const uint64_t e0 = sb + threadIdx.x * 16;
            if (e0 >= u.e1) break;
            pb_u32x4 dq[2], vq[4];
#pragma unroll
            for (int k = 0; k < 2; ++k)
                dq[k] = *(reinterpret_cast<const pb_u32x4 *>(dst + e0) + k);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                vq[k] = *(reinterpret_cast<const pb_u32x4 *>(val + e0) + k);
            uint32_t cur = 0xFFFFFFFFu;
            double run = 0.0;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const uint64_t e = e0 + i;
                const uint32_t d = (dq[i >> 3][(i >> 1) & 3] >> ((i & 1) * 16)) & 0xFFFFu;
                if (e < u.e0 || e >= u.e1 || d < lo || d >= lo + H) continue;
                const double x = (double)__uint_as_float(vq[i >> 2][i & 3]);
                if (d != cur) {
                    if (cur != 0xFFFFFFFFu) {
                        if (u.dense) atomicAdd(&sum[cur - lo], run);
                        else lds_add_cas64(&sum[cur - lo], run);
                    }
                    cur = d;
                    run = x;
                } else {
                    run += x;
                }
            }
            if (cur != 0xFFFFFFFFu) {
                if (u.dense) atomicAdd(&sum[cur - lo], run);
                else lds_add_cas64(&sum[cur - lo], run);
            }
        }
        __syncthreads();
        const uint32_t hn = std::min(H, qn - lo);
        if (u.slot == 0xFFFFFFFFu) {
            for (uint32_t i = threadIdx.x; i < hn; i += PB_WG) acc[qbase + lo + i] = sum[i];
        } else {
            double *pp = part + (uint64_t)u.slot * PB_B + lo;
            for (uint32_t i = threadIdx.x; i < hn; i += PB_WG) pp[i] = sum[i];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_pb_reduce64(const PbRed *__restrict__ red, const double *__restrict__ part,
                                                     uint64_t nzn, double *__restrict__ acc) {
    const PbRed r = red[blockIdx.y];
    const uint64_t qbase = (uint64_t)r.q << PB_SHIFT;
    const uint32_t qn = (uint32_t)std::min<uint64_t>(PB_B, nzn - qbase);
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= qn) return;
    const double *pp = part + (uint64_t)r.slot0 * PB_B + i;
    double s = 0.0;
    for (uint32_t k = 0; k < r.nslot; ++k) s += pp[(uint64_t)k * PB_B];
    acc[qbase + i] = s;
}

// New priors of a split destination block from the sum of its units' partials.
__global__ void __launch_bounds__(256) k_pb_reduce(const PbRed *__restrict__ red, const float *__restrict__ part,
                                                   uint64_t nzn, const float *__restrict__ fac,
                                                   float *__restrict__ prior) {
    const PbRed r = red[blockIdx.y];
    const uint64_t qbase = (uint64_t)r.q << PB_SHIFT;
    const uint32_t qn = (uint32_t)std::min<uint64_t>(PB_B, nzn - qbase);
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= qn) return;
    const float *pp = part + (uint64_t)r.slot0 * PB_B + i;
    float s = 0.0f;
    for (uint32_t k = 0; k < r.nslot; ++k) s += pp[(uint64_t)k * PB_B];
    prior[qbase + i] = pr_new_prior(s, fac[qbase + i]);
}

}  // namespace

// The unit tables must tile the entry ranges exactly and stay within the
// LDS tables of the kernels: checked on the host before the first launch.
static int check_units(const std::vector<PbUnit1> &u1, const std::vector<PbUnit2> &u2,
                       const std::vector<PbRed> &red, const std::vector<uint64_t> &rpw, uint64_t total,
                       uint32_t runs_max, uint32_t slots, const std::vector<uint64_t> &sb,
                       const std::vector<uint64_t> &slen) {
    uint64_t pw = 0;
    const uint64_t nr = rpw.size() - 1;
    for (const PbUnit1 &x : u1) {
        const bool ok = x.w0 == pw && x.w1 > x.w0 && x.r0 < x.r1 && x.r1 <= nr && x.r1 - x.r0 <= runs_max &&
                        x.r1 - x.r0 <= PB_RUNS && rpw[x.r0] <= x.w0 && x.w0 < rpw[x.r0 + 1] &&
                        rpw[x.r1 - 1] < x.w1 && x.w1 <= rpw[x.r1] && x.w1 - x.w0 <= (1ull << 16);
        if (!ok) { set_error("blocked index: bad gather unit at window %llu", (unsigned long long)pw); return GS_ERR_FORMAT; }
        pw = x.w1;
    }
    if (pw * PB_WIN != total) {
        set_error("blocked index: gather units end at window %llu", (unsigned long long)pw);
        return GS_ERR_FORMAT;
    }
    // accumulation units: in block order, tiling each block's stream range
    size_t i = 0;
    for (uint32_t q = 0; q < (uint32_t)sb.size(); ++q) {
        if (!slen[q]) continue;
        uint64_t e = sb[q] & ~PB_COMPACT;
        const uint64_t end = e + slen[q];
        while (e < end) {
            if (i >= u2.size() || u2[i].q != q || u2[i].e0 != e || u2[i].e1 <= e || u2[i].e1 > end ||
                (u2[i].slot != 0xFFFFFFFFu && u2[i].slot >= slots)) {
                set_error("blocked index: bad accum unit for block %u at %llu", q, (unsigned long long)e);
                return GS_ERR_FORMAT;
            }
            e = u2[i++].e1;
        }
    }
    if (i != u2.size()) { set_error("blocked index: %zu accum units left over", u2.size() - i); return GS_ERR_FORMAT; }
    for (const PbRed &r : red)
        if (r.nslot < 2 || r.slot0 + r.nslot > slots) { set_error("blocked index: bad partial slots"); return GS_ERR_FORMAT; }
    return GS_OK;
}

size_t pb_bytes(const gs_graph *g) {
    return g->pb_src.bytes() + g->pb_wcoff.bytes() + g->pb_dst.bytes() + g->pb_val.bytes() +
           g->pb_run_pw.bytes() + g->pb_run_qw.bytes() + g->pb_run_st.bytes() + g->pb_u1.bytes() + g->pb_u2.bytes() + g->pb_red.bytes() + g->pb_part.bytes() +
           g->pb_part64.bytes();
}

// Builds the blocked copy from the neighbour-list index (built first if
// needed).  Leaves pb_k = 0 -- PageRank then pulls over the lists -- when the
// graph has more than PB_KMAX blocks of vertices with edges.
int build_blocked(gs_graph *g) {
    GS_TRY(build_adjacency(g));
    std::lock_guard<std::mutex> lk(g->build_mu);
    if (g->pb_ready) return GS_OK;
    const uint64_t nzn = g->nzn;
    const uint64_t K64 = (nzn + PB_B - 1) >> PB_SHIFT;
    if (K64 == 0 || K64 > PB_KMAX) {
        g->pb_k = 0;
        g->pb_ready = true;
        return GS_OK;
    }
    const uint32_t K = (uint32_t)K64;
    gs_ws *w = g->ws;
    hipStream_t st = w->stream;
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, st));
    // 1. list range of every destination block, cut into build slices
    std::vector<uint64_t> beg(K + 1);
    {
        DevBuf<uint64_t> d_beg;
        GS_TRY(d_beg.alloc(K + 1));
        k_pb_bin_begin<<<vgrid(K + 1), 256, 0, st>>>(g->adj_off.ptr, nzn, K, d_beg.ptr);
        GS_HIP(hipGetLastError());
        GS_HIP(hipMemcpyAsync(beg.data(), d_beg.ptr, (K + 1) * 8, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
    }
    std::vector<PbSlice> slices;
    for (uint32_t q = 0; q < K; ++q)
        for (uint64_t a = beg[q]; a < beg[q + 1]; a += PB_SLICE)
            slices.push_back(PbSlice{a, std::min(a + PB_SLICE, beg[q + 1]), q, 0});
    DevBuf<PbSlice> d_sl;
    GS_TRY(d_sl.alloc(slices.size()));
    GS_HIP(hipMemcpyAsync(d_sl.ptr, slices.data(), slices.size() * sizeof(PbSlice), hipMemcpyHostToDevice, st));
    // 2. entries per run (q-major table)
    const uint64_t KK = (uint64_t)K * K;
    DevBuf<unsigned long long> cnt;
    GS_TRY(cnt.alloc(KK));
    GS_HIP(hipMemsetAsync(cnt.ptr, 0, KK * 8, st));
    if (!slices.empty()) {
        k_pb_count<<<(unsigned)slices.size(), PB_WG, 0, st>>>(d_sl.ptr, g->adj_nbr.ptr, K, cnt.ptr);
        GS_HIP(hipGetLastError());
    }
    std::vector<unsigned long long> c(KK);
    GS_HIP(hipMemcpyAsync(c.data(), cnt.ptr, KK * 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    // 3. run starts: q-major positions (whole windows), and the non-empty runs
    //    in source-major order with their window prefix and q-major window
    std::vector<unsigned long long> qm(KK);
    unsigned long long run = 0, real = 0;
    for (uint64_t i = 0; i < KK; ++i) {
        qm[i] = run;
        run += (c[i] + PB_WIN - 1) & ~(unsigned long long)(PB_WIN - 1);
        real += c[i];
    }
    if (real != 2 * g->E) {
        set_error("blocked index: %llu entries, expected %llu", real, (unsigned long long)(2 * g->E));
        return GS_ERR_FORMAT;
    }
    std::vector<uint64_t> rpw, rqw;
    std::vector<uint32_t> prun(K + 1), rblk;   // rblk: destination block of each run
    {
        // source-major walk of a q-major table, in 64 x 64 tiles for the host caches
        std::vector<uint64_t> tmp_c(KK), tmp_q(KK);
        constexpr uint32_t TB = 64;
        for (uint32_t q0 = 0; q0 < K; q0 += TB)
            for (uint32_t p0 = 0; p0 < K; p0 += TB)
                for (uint32_t q = q0; q < std::min(K, q0 + TB); ++q)
                    for (uint32_t p = p0; p < std::min(K, p0 + TB); ++p) {
                        tmp_c[(uint64_t)p * K + q] = c[(uint64_t)q * K + p];
                        tmp_q[(uint64_t)p * K + q] = qm[(uint64_t)q * K + p];
                    }
        uint64_t pw = 0;
        for (uint32_t p = 0; p < K; ++p) {
            prun[p] = (uint32_t)rpw.size();
            for (uint32_t q = 0; q < K; ++q) {
                const uint64_t n_ = tmp_c[(uint64_t)p * K + q];
                if (!n_) continue;
                if (n_ >= (1ull << 32)) {
                    set_error("blocked index: a run of %llu entries", (unsigned long long)n_);
                    return GS_ERR_FORMAT;
                }
                rpw.push_back(pw);
                rqw.push_back(tmp_q[(uint64_t)p * K + q] / PB_WIN);
                rblk.push_back(q);
                pw += (n_ + PB_WIN - 1) / PB_WIN;
            }
        }
        prun[K] = (uint32_t)rpw.size();
        rpw.push_back(pw);
    }
    const uint64_t nruns = rqw.size();
    const uint64_t unit1 = std::min<uint64_t>(std::max<uint64_t>(g->pb_unit1 / PB_WIN, 1), PB_UNIT_WIN);
    const uint64_t unit2 = g->pb_unit2;
    const uint32_t runs_max = g->pb_runs;
    // 4. gather units: source-major window ranges within one source block, at
    //    most pb_unit1 entries and pb_runs runs each
    std::vector<PbUnit1> u1;
    for (uint32_t p = 0; p < K; ++p) {
        uint32_t r = prun[p];
        const uint32_t rend = prun[p + 1];
        uint64_t pw = rpw[r];
        const uint64_t pend = rpw[rend];
        while (pw < pend) {
            PbUnit1 x{};
            x.p = p;
            x.w0 = pw;
            x.r0 = r;
            uint64_t lim = std::min(pw + unit1, pend);
            if (r + runs_max < rend && rpw[r + runs_max] < lim) lim = rpw[r + runs_max];
            x.w1 = lim;
            // runs overlapping [pw, lim): r .. the run holding lim - 1
            uint32_t rl = r;
            while (rl + 1 < rend && rpw[rl + 1] <= lim - 1) ++rl;
            x.r1 = rl + 1;
            u1.push_back(x);
            pw = lim;
            r = (rl + 1 < rend && rpw[rl + 1] <= pw) ? rl + 1 : rl;
        }
    }
    // 5. the copy in q-major order (runs sorted by destination), the padding of
    //    each run's last window, and the segment masks of every window
    const uint64_t nwin = run / PB_WIN;
    if (nwin >= (1ull << 32)) {  // the gather stages q-major window indices as u32
        set_error("blocked index: %llu windows exceed 2^32", (unsigned long long)nwin);
        return GS_ERR_UNSUPPORTED;
    }
    GS_TRY(g->pb_src.alloc(run + PB_WIN));
    DevBuf<uint16_t> qdst;   // q-major destination offsets, until the stream is laid out
    GS_TRY(qdst.alloc(run + PB_WIN));
    GS_HIP(hipMemcpyAsync(cnt.ptr, qm.data(), KK * 8, hipMemcpyHostToDevice, st));
    if (!slices.empty()) {
        k_pb_fill<<<(unsigned)slices.size(), PB_WG, 0, st>>>(d_sl.ptr, g->adj_off.ptr, g->adj_nbr.ptr, nzn, K,
                                                             cnt.ptr, g->pb_src.ptr, qdst.ptr);
        GS_HIP(hipGetLastError());
        k_pb_pad<<<vgrid(KK), 256, 0, st>>>(cnt.ptr, KK, g->pb_src.ptr, qdst.ptr);
        GS_HIP(hipGetLastError());
    }
    GS_HIP(hipMemsetAsync(g->pb_src.ptr + run, 0, PB_WIN * 2, st));
    DevBuf<uint16_t> wmask;   // per q-major window: the last entry of each segment
    GS_TRY(wmask.alloc(std::max<uint64_t>(nwin, 1)));
    GS_TRY(g->pb_wcoff.alloc(std::max<uint64_t>(nwin, 1)));
    DevBuf<uint64_t> cpos;   // segments before each q-major window (nwin + 1)
    GS_TRY(cpos.alloc(nwin + 1));
    uint64_t nseg = 0;
    {
        DevBuf<int32_t> wcnt;
        GS_TRY(wcnt.alloc(std::max<uint64_t>(nwin, 1)));
        if (nwin) {
            k_pb_wmask<<<vgrid(nwin), 256, 0, st>>>(g->pb_src.ptr, qdst.ptr, nwin, wmask.ptr, wcnt.ptr);
            GS_HIP(hipGetLastError());
        }
        GS_TRY(scan_counts(st, wcnt.ptr, nwin, cpos.ptr, &nseg));
    }
    // 6. per block: plain or compacted, and its place in the stream (starts
    //    16-aligned, so plain blocks keep 16-byte stores)
    std::vector<uint64_t> bw(K + 1), bc(K + 1);
    for (uint32_t q = 0; q < K; ++q) bw[q] = qm[(uint64_t)q * K] / PB_WIN;
    bw[K] = nwin;
    {
        DevBuf<uint64_t> idx, out;
        GS_TRY(idx.alloc(K + 1));
        GS_TRY(out.alloc(K + 1));
        GS_HIP(hipMemcpyAsync(idx.ptr, bw.data(), (K + 1) * 8, hipMemcpyHostToDevice, st));
        k_pb_pick<<<vgrid(K + 1), 256, 0, st>>>(cpos.ptr, idx.ptr, K + 1, out.ptr);
        GS_HIP(hipGetLastError());
        GS_HIP(hipMemcpyAsync(bc.data(), out.ptr, (K + 1) * 8, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
    }
    std::vector<uint64_t> sb(K), slen(K);
    uint64_t stream = 0;
    for (uint32_t q = 0; q < K; ++q) {
        const uint64_t ent = (bw[q + 1] - bw[q]) * PB_WIN, seg = bc[q + 1] - bc[q];
        const bool comp = ent && seg * 1000 <= (uint64_t)g->pb_compact * ent;
        sb[q] = stream | (comp ? PB_COMPACT : 0);
        slen[q] = comp ? seg : ent;
        stream = (stream + slen[q] + PB_WIN - 1) & ~(uint64_t)(PB_WIN - 1);
    }
    // 7. stream base of every run: its block's base plus, for a compacted
    //    block, the segments before the run, else its entries before it
    std::vector<uint64_t> rst(nruns);
    if (nruns) {
        DevBuf<uint64_t> idx, out;
        GS_TRY(idx.alloc(nruns));
        GS_TRY(out.alloc(nruns));
        GS_HIP(hipMemcpyAsync(idx.ptr, rqw.data(), nruns * 8, hipMemcpyHostToDevice, st));
        k_pb_pick<<<vgrid(nruns), 256, 0, st>>>(cpos.ptr, idx.ptr, nruns, out.ptr);
        GS_HIP(hipGetLastError());
        GS_HIP(hipMemcpyAsync(rst.data(), out.ptr, nruns * 8, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
        for (uint64_t r = 0; r < nruns; ++r) {
            const uint32_t q = rblk[r];
            rst[r] = (sb[q] & PB_COMPACT) ? sb[q] + (rst[r] - bc[q]) : sb[q] + (rqw[r] - bw[q]) * PB_WIN;
        }
    }
    GS_TRY(g->pb_run_pw.alloc(rpw.size()));
    GS_TRY(g->pb_run_qw.alloc(std::max<uint64_t>(nruns, 1)));
    GS_TRY(g->pb_run_st.alloc(std::max<uint64_t>(nruns, 1)));
    GS_HIP(hipMemcpyAsync(g->pb_run_pw.ptr, rpw.data(), rpw.size() * 8, hipMemcpyHostToDevice, st));
    if (nruns) {
        GS_HIP(hipMemcpyAsync(g->pb_run_qw.ptr, rqw.data(), nruns * 8, hipMemcpyHostToDevice, st));
        GS_HIP(hipMemcpyAsync(g->pb_run_st.ptr, rst.data(), nruns * 8, hipMemcpyHostToDevice, st));
        k_pb_wcoff<<<(unsigned)std::min<uint64_t>((nruns + 3) / 4, 8192), 256, 0, st>>>(
            g->pb_run_pw.ptr, g->pb_run_qw.ptr, nruns, cpos.ptr, g->pb_wcoff.ptr);
        GS_HIP(hipGetLastError());
    }
    // 8. the stream's destination offsets; the q-major copy is then dropped
    GS_TRY(g->pb_dst.alloc(stream + PB_WIN));
    GS_TRY(g->pb_val.alloc(stream + PB_WIN));
    GS_HIP(hipMemsetAsync(g->pb_dst.ptr, 0, (stream + PB_WIN) * 2, st));
    if (nwin) {
        DevBuf<uint64_t> d_bw, d_sb, d_bc;
        GS_TRY(d_bw.alloc(K + 1));
        GS_TRY(d_sb.alloc(K));
        GS_TRY(d_bc.alloc(K + 1));
        GS_HIP(hipMemcpyAsync(d_bw.ptr, bw.data(), (K + 1) * 8, hipMemcpyHostToDevice, st));
        GS_HIP(hipMemcpyAsync(d_sb.ptr, sb.data(), K * 8, hipMemcpyHostToDevice, st));
        GS_HIP(hipMemcpyAsync(d_bc.ptr, bc.data(), (K + 1) * 8, hipMemcpyHostToDevice, st));
        k_pb_stream_dst<<<vgrid(nwin), 256, 0, st>>>(qdst.ptr, wmask.ptr, nwin, cpos.ptr, d_bw.ptr, d_sb.ptr,
                                                     d_bc.ptr, K, g->pb_dst.ptr);
        GS_HIP(hipGetLastError());
        k_pb_mark<<<vgrid(nwin), 256, 0, st>>>(g->pb_src.ptr, wmask.ptr, nwin, d_bw.ptr, d_sb.ptr, K);
        GS_HIP(hipGetLastError());
        GS_HIP(hipStreamSynchronize(st));
    }
    qdst.release();
    cpos.release();
    wmask.release();
    // 9. accum units: each block's stream range in equal parts of at most
    //    pb_unit2 values; split blocks get partial slots
    std::vector<PbUnit2> u2;
    std::vector<PbRed> red;
    uint32_t slots = 0;
    for (uint32_t q = 0; q < K; ++q) {
        const uint64_t e0 = sb[q] & ~PB_COMPACT, n_ = slen[q], e1 = e0 + n_;
        if (!n_) continue;
        const uint64_t parts = (n_ + unit2 - 1) / unit2;
        // stream values per vertex of the block: above pb_dense, lanes of a wave mostly share destinations
        const uint64_t qn = std::min<uint64_t>(PB_B, nzn - ((uint64_t)q << PB_SHIFT));
        const uint32_t dense = n_ >= (uint64_t)g->pb_dense * qn;
        if (parts == 1) {
            u2.push_back(PbUnit2{e0, e1, q, 0xFFFFFFFFu, dense, 0});
            continue;
        }
        red.push_back(PbRed{q, slots, (uint32_t)parts, 0});
        for (uint64_t k = 0; k < parts; ++k)
            u2.push_back(PbUnit2{e0 + n_ * k / parts, e0 + n_ * (k + 1) / parts, q, slots++, dense, 0});
    }
    GS_TRY(check_units(u1, u2, red, rpw, run, runs_max, slots, sb, slen));
    GS_TRY(g->pb_u1.alloc(u1.size()));
    GS_TRY(g->pb_u2.alloc(u2.size()));
    GS_TRY(g->pb_red.alloc(red.size()));
    GS_TRY(g->pb_part.alloc((uint64_t)slots * PB_B));
    if (!u1.empty())
        GS_HIP(hipMemcpyAsync(g->pb_u1.ptr, u1.data(), u1.size() * sizeof(PbUnit1), hipMemcpyHostToDevice, st));
    if (!u2.empty())
        GS_HIP(hipMemcpyAsync(g->pb_u2.ptr, u2.data(), u2.size() * sizeof(PbUnit2), hipMemcpyHostToDevice, st));
    if (!red.empty())
        GS_HIP(hipMemcpyAsync(g->pb_red.ptr, red.data(), red.size() * sizeof(PbRed), hipMemcpyHostToDevice, st));
    g->pb_stream = stream;
    GS_HIP(hipEventRecord(t1, st));
    GS_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    g->pb_k = K;
    g->pb_nu1 = u1.size();
    g->pb_nu2 = u2.size();
    g->pb_nred = red.size();
    g->pb_nruns = nruns;
    g->pb_entries = run;
    g->pb_ms = ms;
    g->pb_ready = true;
    return GS_OK;
}

__global__ void k_pb_ctr_zero(unsigned long long *ctr) { *ctr = 0; }

static int pb_gather(const gs_graph *gc, gs_ws *w, const float *prior) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    if (!g->pb_nu1) return GS_OK;
    const uint32_t grab = g->pb_grab;
    unsigned grid = (unsigned)g->pb_nu1;
    if (grab) {
        if (!g->pb_ctr.ptr) GS_TRY(g->pb_ctr.alloc(1));
        if (!g->num_cus) {
            int dev = 0, cus = 0;
            GS_HIP(hipGetDevice(&dev));
            GS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            g->num_cus = (uint32_t)std::max(cus, 1);
        }
        // one resident workgroup per CU (the kernel's LDS admits one)
        grid = (unsigned)std::min<uint64_t>(g->num_cus, (g->pb_nu1 + grab - 1) / grab);
        k_pb_ctr_zero<<<1, 1, 0, w->stream>>>(g->pb_ctr.ptr);
        GS_HIP(hipGetLastError());
    }
    hipEvent_t kb;
    w->timer.begin(GS_K_PB_GATHER, w->stream, &kb);
    k_pb_gather<<<grid, PB_WG, 0, w->stream>>>(g->pb_u1.ptr, (uint32_t)g->pb_nu1, grab, g->pb_ctr.ptr,
                                               g->pb_run_pw.ptr, g->pb_run_qw.ptr, g->pb_run_st.ptr, g->pb_src.ptr,
                                               g->pb_wcoff.ptr, prior, g->nzn, g->pb_val.ptr);
    GS_HIP(hipGetLastError());
    w->timer.end(GS_K_PB_GATHER, w->stream, kb);
    return GS_OK;
}

// One FP64 accumulation (personalised PageRank): acc[v] = sum of prior over v's
// list for the relabelled ids 0 .. nzn-1, stored (not added).
int pr_blocked64(const gs_graph *gc, gs_ws *w, const float *prior, double *acc) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    if (g->pb_nred && g->pb_part64.n != g->pb_part.n) GS_TRY(g->pb_part64.alloc(g->pb_part.n));
    GS_TRY(pb_gather(g, w, prior));
    hipEvent_t kb;
    if (g->pb_nu2) {
        w->timer.begin(GS_K_PB_ACCUM, w->stream, &kb);
        k_pb_accum64<<<(unsigned)g->pb_nu2, PB_WG, 0, w->stream>>>(g->pb_u2.ptr, g->pb_dst.ptr, g->pb_val.ptr, g->nzn,
                                                                   acc, g->pb_part64.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PB_ACCUM, w->stream, kb);
    }
    if (g->pb_nred) {
        w->timer.begin(GS_K_PB_REDUCE, w->stream, &kb);
        k_pb_reduce64<<<dim3(PB_B / 256, (unsigned)g->pb_nred), 256, 0, w->stream>>>(g->pb_red.ptr, g->pb_part64.ptr,
                                                                                    g->nzn, acc);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PB_REDUCE, w->stream, kb);
    }
    return GS_OK;
}

// One PageRank iteration for the relabelled ids 0 .. nzn-1: the sums over
// the lists and, fused into the accumulation, the new priors (the gather has
// read every prior it needs before the accumulation starts).
int pr_blocked(const gs_graph *g, gs_ws *w, float *prior, const float *fac) {
    GS_TRY(pb_gather(g, w, prior));
    hipEvent_t kb;
    if (g->pb_nu2) {
        w->timer.begin(GS_K_PB_ACCUM, w->stream, &kb);
        k_pb_accum<<<(unsigned)g->pb_nu2, PB_WG, 0, w->stream>>>(g->pb_u2.ptr, g->pb_dst.ptr, g->pb_val.ptr, g->nzn,
                                                                 fac, prior, g->pb_part.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PB_ACCUM, w->stream, kb);
    }
    if (g->pb_nred) {
        w->timer.begin(GS_K_PB_REDUCE, w->stream, &kb);
        k_pb_reduce<<<dim3(PB_B / 256, (unsigned)g->pb_nred), 256, 0, w->stream>>>(g->pb_red.ptr, g->pb_part.ptr,
                                                                                  g->nzn, fac, prior);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_PB_REDUCE, w->stream, kb);
    }
    return GS_OK;
}

}  // namespace gs
