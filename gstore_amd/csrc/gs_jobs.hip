// gs_jobs.hip -- the reference's remaining jobs and tool on the device:
//   -j 30  asynchronous BFS        (bfs2_t, bfs2.cpp:82-179; grid::bfs2, gstore.cpp:1900-2042)
//   -j 4   single colour spreading (traverse_t, traverse.cpp:56-211; grid::traverse, gstore.cpp:2169-2177)
//   -c 4   text edge list -> u32 pairs (text_to_bin, util.cpp:58-168)
//
// The two jobs are tile sweeps repeated until a sweep changes nothing.  Their
// updates land in place while the sweep runs (a vertex set early in a sweep
// can pass its value on later in the same sweep), as the reference's
// parallel sweeps do, so the number of sweeps and the per-sweep counts depend
// on the order workgroups run; the final arrays do not.  A tile is swept only
// when its row (or, for the BFS on an undirected grid, its column) had an
// update in the previous sweep: an edge can change something only when an
// endpoint changed since the tile was last swept.
//
// The text conversion keeps the whole text in HBM: a pass marks token starts
// (the reference's scanner: a token at the start of the text and after every
// run of ' ', '\n', '\t'), a scan of per-block counts places them, and a
// second pass parses each token as atoi does and tracks the minimum (from
// 999999) and maximum id.
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <vector>

#include "gs_internal.h"

namespace gs {
namespace {

constexpr uint32_t INVALID_CID = 0xFFFFFFFFu;   // invalid_cid (traverse.cpp:54, wcc.cpp:54)

// Lowers the u8 depth of v to nd when v is unreached (0) or deeper than nd,
// by compare-and-swap on the aligned word that holds its byte.  True when it
// stored.  The depth buffer is padded to a whole word.
__device__ inline bool lower_depth(uint8_t *depth, uint32_t v, uint32_t nd) {
    uint32_t *wp = reinterpret_cast<uint32_t *>(depth + (v & ~3u));
    const uint32_t sh = (v & 3u) * 8u;
    uint32_t old = *reinterpret_cast<volatile uint32_t *>(wp);
    for (;;) {
        const uint32_t cur = (old >> sh) & 0xFFu;
        if (cur != 0 && cur <= nd) return false;
        const uint32_t prev = atomicCAS(wp, old, (old & ~(0xFFu << sh)) | (nd << sh));
        if (prev == old) return true;
        old = prev;
    }
}

__device__ inline void mark_row(uint32_t *bm, uint32_t row) {
    const uint32_t b = 1u << (row & 31);
    if (!(bm[row >> 5] & b)) atomicOr(&bm[row >> 5], b);
}

// Workgroup sum of per-lane counts into out (one atomic per workgroup).
__device__ inline void wg_add(uint32_t x, unsigned long long *out) {
    __shared__ unsigned long long tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    if (x) atomicAdd(&tot, (unsigned long long)x);
    __syncthreads();
    if (threadIdx.x == 0 && tot) atomicAdd(out, tot);
}

// One sweep of the asynchronous BFS (bfs2.cpp:103-129): l0 = d[v0], l1 =
// d[v1]; if l0 != 0 and (l1 == 0 or l1 - l0 > 1) then d[v1] = l0 + 1, else
// (undirected) if l1 != 0 and (l0 == 0 or l0 - l1 > 1) then d[v0] = l1 + 1.
// cnt[0] += stores made; cnt[1] = 1 when a depth beyond 255 was needed.  A
// forward store marks the tile column's row bit for the next sweep, a
// backward one the tile row's.
template <bool DIRECTED>
__global__ void __launch_bounds__(SWEEP_WG) k_bfs2_sweep(SweepArgs a, uint8_t *depth, const uint32_t *act,
                                                          uint32_t *act_next, int all, unsigned long long *cnt) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    const uint32_t len = stage_chunk(a, chunk, s, [&](uint32_t rc) {
        return all || bit_of(act, rc >> 16) || (!DIRECTED && bit_of(act, rc & 0xFFFFu));
    });
    if (!s.any) return;
    uint32_t upd = 0, over = 0;
    auto op = [&](uint32_t w, uint32_t rc) {
        const uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        const uint32_t l0 = depth[v0], l1 = depth[v1];
        if (l0 != 0 && (l1 == 0 || l1 > l0 + 1)) {
            if (l0 == 255) { over = 1; return; }
            if (lower_depth(depth, v1, l0 + 1)) { ++upd; mark_row(act_next, rc & 0xFFFFu); }
        } else if (!DIRECTED && l1 != 0 && (l0 == 0 || l0 > l1 + 1)) {
            if (l1 == 255) { over = 1; return; }
            if (lower_depth(depth, v0, l1 + 1)) { ++upd; mark_row(act_next, rc >> 16); }
        }
    };
    sweep_chunk<true>(a, chunk, s, len, op);
    wg_add(upd, cnt);
    if (over) atomicOr(cnt + 1, 1ull);
}

// One sweep of the colour spreading (traverse.cpp:178-186): cid[v1] = color
// when cid[v0] == color and cid[v1] is invalid, stored edges v0 -> v1 only.
// Tiles whose row got no colour in the previous sweep are skipped (only v0,
// in the tile row, passes colour on).
__global__ void __launch_bounds__(SWEEP_WG) k_trav_sweep(SweepArgs a, uint32_t *cid, uint32_t color,
                                                          const uint32_t *act, uint32_t *act_next, int all,
                                                          unsigned long long *cnt) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    const uint32_t len = stage_chunk(a, chunk, s, [&](uint32_t rc) { return all || bit_of(act, rc >> 16); });
    if (!s.any) return;
    uint32_t upd = 0;
    auto op = [&](uint32_t w, uint32_t rc) {
        const uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        if (cid[v0] == color && cid[v1] == INVALID_CID && atomicCAS(&cid[v1], INVALID_CID, color) == INVALID_CID) {
            ++upd;
            mark_row(act_next, rc & 0xFFFFu);
        }
    };
    sweep_chunk<true>(a, chunk, s, len, op);
    wg_add(upd, cnt);
}

__global__ void k_set_byte(uint8_t *p, uint64_t i, uint8_t v) { p[i] = v; }
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- text edge list ------------------------------------------------------------
constexpr int TXT_WG = 256;
constexpr int TXT_BPT = 16;                      // bytes per thread
constexpr uint64_t TXT_BLOCK = TXT_WG * TXT_BPT;  // bytes per workgroup

__device__ inline bool tdelim(uint32_t c) { return c == ' ' || c == '\n' || c == '\t'; }
__device__ inline bool tspace(uint32_t c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

// Start flags of this thread's 16 bytes, bit k = byte i0 + k starts a token.
__device__ inline uint32_t txt_starts(const uint8_t *t, uint64_t size, uint64_t i0) {
    if (i0 >= size) return 0;
    uint8_t b[TXT_BPT];
    if (i0 + TXT_BPT <= size) {
        const uint4 v = *reinterpret_cast<const uint4 *>(t + i0);
        __builtin_memcpy(b, &v, 16);
    } else {
        for (int k = 0; k < TXT_BPT; ++k) b[k] = i0 + k < size ? t[i0 + k] : 0;
    }
    uint32_t prev = i0 ? t[i0 - 1] : 0, m = 0;
#pragma unroll
    for (int k = 0; k < TXT_BPT; ++k) {
        const uint64_t i = i0 + k;
        const bool st = i < size && (i == 0 || (tdelim(prev) && !tdelim(b[k])));
        m |= (uint32_t)st << k;
        prev = b[k];
    }
    return m;
}

__global__ void __launch_bounds__(TXT_WG) k_txt_count(const uint8_t *t, uint64_t size, uint32_t *blk) {
    const uint64_t i0 = blockIdx.x * TXT_BLOCK + (uint64_t)threadIdx.x * TXT_BPT;
    const uint32_t c = __popc(txt_starts(t, size, i0));
    __shared__ uint32_t tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    if (c) atomicAdd(&tot, c);
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// atoi from byte i: leading white space, optional sign, decimal digits; the
// strtol value (saturated to 64 bits) cut to 32 bits.
__device__ inline uint32_t txt_atoi(const uint8_t *t, uint64_t size, uint64_t i) {
    while (i < size && tspace(t[i])) ++i;
    bool neg = false;
    if (i < size && (t[i] == '+' || t[i] == '-')) { neg = t[i] == '-'; ++i; }
    uint64_t v = 0;
    bool sat = false;
    while (i < size && t[i] >= '0' && t[i] <= '9') {
        const uint64_t d = t[i] - '0';
        if (!sat && v > (0x7FFFFFFFFFFFFFFFull - d) / 10) sat = true;
        if (!sat) v = v * 10 + d;
        ++i;
    }
    const int64_t r = sat ? (neg ? (int64_t)(-0x7FFFFFFFFFFFFFFFll - 1) : 0x7FFFFFFFFFFFFFFFll)
                          : (neg ? -(int64_t)v : (int64_t)v);
    return (uint32_t)(uint64_t)r;
}

// Token values at their places (block offset + rank within the block), and
// the minimum / maximum into mm[0] / mm[1].
__global__ void __launch_bounds__(TXT_WG) k_txt_parse(const uint8_t *t, uint64_t size, const uint64_t *blk_off,
                                                      uint32_t *tok, uint32_t *mm) {
    const uint64_t i0 = blockIdx.x * TXT_BLOCK + (uint64_t)threadIdx.x * TXT_BPT;
    uint32_t m = txt_starts(t, size, i0);
    __shared__ uint32_t pre[TXT_WG];
    pre[threadIdx.x] = __popc(m);
    __syncthreads();
    for (int d = 1; d < TXT_WG; d <<= 1) {   // inclusive scan
        const uint32_t x = threadIdx.x >= (unsigned)d ? pre[threadIdx.x - d] : 0;
        __syncthreads();
        pre[threadIdx.x] += x;
        __syncthreads();
    }
    uint64_t o = blk_off[blockIdx.x] + pre[threadIdx.x] - __popc(m);
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    while (m) {
        const int k = __ffs(m) - 1;
        m &= m - 1;
        const uint32_t v = txt_atoi(t, size, i0 + k);
        tok[o++] = v;
        mn = min(mn, v);
        mx = max(mx, v);
    }
    if (mn != 0xFFFFFFFFu || mx) {
        atomicMin(&mm[0], mn);
        atomicMax(&mm[1], mx);
    }
}

__global__ void k_txt_shift(uint32_t *tok, uint64_t n, const uint32_t *mm) {
    const uint32_t mn = mm[0];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        tok[i] -= mn;
}

unsigned vgrid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

// Host side of the two sweeping jobs: run sweeps until one changes nothing.
// launch(SweepArgs, grid, stream, act, act_next, all) issues one sweep.
template <class Launch>
int sweep_until_quiet(const gs_graph *g, gs_ws *w, int kclass, uint64_t small, uint64_t *sweep_counts, int max_sweeps,
                      int *sweeps, uint64_t *total, bool *overflow, Launch launch) {
    const uint64_t rw = (g->p_s + 31) / 32;
    DevBuf<uint32_t> act, act_next;
    GS_TRY(act.alloc(rw));
    GS_TRY(act_next.alloc(rw));
    GS_HIP(hipMemsetAsync(act.ptr, 0, rw * 4, w->stream));
    GS_HIP(hipMemsetAsync(act_next.ptr, 0, rw * 4, w->stream));
    int all = 1, ns = 0;
    *overflow = false;
    *total = 0;
    for (;;) {
        if (ns >= (1 << 24)) { set_error("no fixed point after 2^24 sweeps"); return GS_ERR_DEVICE; }
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 16, w->stream));
        GS_TRY(sweep(g, w, kclass, [&](const SweepArgs &sa, unsigned grid, hipStream_t st) {
            launch(sa, grid, st, act.ptr, act_next.ptr, all);
        }));
        GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 16, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        const uint64_t c = w->h_counters[0];
        if (w->h_counters[1]) { *overflow = true; break; }
        if (sweep_counts && ns < max_sweeps) sweep_counts[ns] = c;
        *total += c;
        ++ns;
        if (c == 0) break;
        std::swap(act.ptr, act_next.ptr);
        GS_HIP(hipMemsetAsync(act_next.ptr, 0, rw * 4, w->stream));
        if (c <= small) all = 0;
    }
    *sweeps = ns;
    return GS_OK;
}

}  // namespace
}  // namespace gs

using namespace gs;

extern "C" int gs_bfs_async(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *sweep_counts, int max_sweeps,
                            int *sweeps) {
    if (!g || root >= g->n || max_sweeps < 0) { set_error("gs_bfs_async: bad arguments"); return GS_ERR_ARG; }
    gs_ws *w = g->ws;
    w->timer.reset();
    w->stats = gs_stats{};
    DevBuf<uint8_t> depth;
    GS_TRY(depth.alloc((g->n + 3) & ~3ull));
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    GS_HIP(hipMemsetAsync(depth.ptr, 0, depth.n, w->stream));
    k_set_byte<<<1, 1, 0, w->stream>>>(depth.ptr, root, 1);
    GS_HIP(hipGetLastError());
    // every tile is swept until a sweep stores at most 16 p depths (undirected;
    // 4 p_s directed), as the reference's ioskip (bfs2.cpp:161-172)
    const uint64_t small = g->directed ? 4 * g->p_s : 16 * g->p;
    int ns = 0;
    bool over = false;
    uint64_t total = 0;
    GS_TRY(sweep_until_quiet(g, w, GS_K_BFS_EXPAND, small, sweep_counts, max_sweeps, &ns, &total, &over,
                             [&](const SweepArgs &sa, unsigned grid, hipStream_t st, const uint32_t *act,
                                 uint32_t *act_next, int all) {
                                 if (g->directed)
                                     k_bfs2_sweep<true><<<grid, SWEEP_WG, 0, st>>>(sa, depth.ptr, act, act_next, all,
                                                                                  w->counters.ptr);
                                 else
                                     k_bfs2_sweep<false><<<grid, SWEEP_WG, 0, st>>>(sa, depth.ptr, act, act_next, all,
                                                                                   w->counters.ptr);
                             }));
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)ns;
    if (sweeps) *sweeps = ns;
    if (over) {
        set_error("asynchronous BFS from %llu needs depth > 255, beyond the u8 depth encoding", (unsigned long long)root);
        return GS_ERR_DEPTH_OVERFLOW;
    }
    if (depth_out) {
        GS_HIP(hipMemcpyAsync(depth_out, depth.ptr, g->n, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

extern "C" int gs_traverse(gs_graph *g, uint64_t pivot, uint32_t color, uint32_t *cid_out, uint64_t *sweep_counts,
                           int max_sweeps, int *sweeps, uint64_t *colored) {
    if (!g || pivot >= g->n || color == INVALID_CID || max_sweeps < 0) {
        set_error("gs_traverse: bad arguments (pivot < N, colour != 0xFFFFFFFF)");
        return GS_ERR_ARG;
    }
    gs_ws *w = g->ws;
    w->timer.reset();
    w->stats = gs_stats{};
    DevBuf<uint32_t> cid;
    GS_TRY(cid.alloc(g->n));
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    k_fill_u32<<<vgrid(g->n), 256, 0, w->stream>>>(cid.ptr, g->n, INVALID_CID);
    GS_HIP(hipGetLastError());
    GS_HIP(hipMemcpyAsync(cid.ptr + pivot, &color, 4, hipMemcpyHostToDevice, w->stream));
    int ns = 0;
    bool over = false;
    uint64_t total = 0;
    GS_TRY(sweep_until_quiet(g, w, GS_K_BFS_EXPAND, 0, sweep_counts, max_sweeps, &ns, &total, &over,
                             [&](const SweepArgs &sa, unsigned grid, hipStream_t st, const uint32_t *act,
                                 uint32_t *act_next, int all) {
                                 k_trav_sweep<<<grid, SWEEP_WG, 0, st>>>(sa, cid.ptr, color, act, act_next, all,
                                                                         w->counters.ptr);
                             }));
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)ns;
    if (sweeps) *sweeps = ns;
    if (colored) *colored = 1 + total;   // each colouring is counted once (compare-and-swap)
    if (cid_out) {
        GS_HIP(hipMemcpyAsync(cid_out, cid.ptr, g->n * 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

// Leading lines that start with '%' (util.cpp:71-77): their length.
static uint64_t text_header(const char *t, uint64_t size) {
    uint64_t h = 0;
    while (h < size && t[h] == '%') {
        while (h < size && t[h] != '\n') ++h;
        if (h < size) ++h;
    }
    return h;
}

// Tokens of text[0 .. size) (no header) on the device: values into tok (u32,
// allocated here), minimum / maximum.
static int text_tokens(hipStream_t st, const uint8_t *dtext, uint64_t size, DevBuf<uint32_t> &tok, uint64_t *ntok,
                       uint32_t *vmin, uint32_t *vmax) {
    const uint64_t nb = std::max<uint64_t>((size + TXT_BLOCK - 1) / TXT_BLOCK, 1);
    if (nb >= (1ull << 31)) { set_error("text of %llu bytes is too long", (unsigned long long)size); return GS_ERR_ARG; }
    DevBuf<uint32_t> blk;
    DevBuf<uint64_t> boff;
    DevBuf<uint32_t> mm;
    GS_TRY(blk.alloc(nb));
    GS_TRY(boff.alloc(nb));
    GS_TRY(mm.alloc(2));
    k_txt_count<<<(unsigned)nb, TXT_WG, 0, st>>>(dtext, size, blk.ptr);
    GS_HIP(hipGetLastError());
    std::vector<uint32_t> hb(nb);
    std::vector<uint64_t> ho(nb);
    GS_HIP(hipMemcpyAsync(hb.data(), blk.ptr, nb * 4, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    uint64_t run = 0;
    for (uint64_t i = 0; i < nb; ++i) { ho[i] = run; run += hb[i]; }
    *ntok = run;
    const uint32_t init[2] = {999999u, 0u};   // v_min "as infinity" and v_max (util.cpp:86-87)
    GS_HIP(hipMemcpyAsync(mm.ptr, init, 8, hipMemcpyHostToDevice, st));
    GS_HIP(hipMemcpyAsync(boff.ptr, ho.data(), nb * 8, hipMemcpyHostToDevice, st));
    GS_TRY(tok.alloc(std::max<uint64_t>(run, 1)));
    k_txt_parse<<<(unsigned)nb, TXT_WG, 0, st>>>(dtext, size, boff.ptr, tok.ptr, mm.ptr);
    GS_HIP(hipGetLastError());
    const uint64_t keep = run / 2 * 2;    // tokens/2 pairs (util.cpp:105, 158)
    if (keep) {
        k_txt_shift<<<vgrid(keep), 256, 0, st>>>(tok.ptr, keep, mm.ptr);
        GS_HIP(hipGetLastError());
    }
    uint32_t hm[2];
    GS_HIP(hipMemcpyAsync(hm, mm.ptr, 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));   // the host vectors and buffers die here
    *vmin = hm[0];
    *vmax = hm[1];
    return GS_OK;
}

extern "C" int gs_text_to_bin_mem(const char *text, uint64_t size, uint32_t *out, uint64_t cap_pairs, uint64_t *pairs,
                                  uint32_t *vmin, uint32_t *vmax, uint64_t *tokens) {
    if ((!text && size) || !pairs || !vmin || !vmax) { set_error("gs_text_to_bin_mem: null argument"); return GS_ERR_ARG; }
    const uint64_t h = text_header(text, size);
    const uint64_t n = size - h;
    hipStream_t st;
    GS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = GS_OK;
    {
        DevBuf<uint8_t> dtext;
        DevBuf<uint32_t> tok;
        uint64_t ntok = 0;
        rc = dtext.alloc(n + 16);
        if (rc == GS_OK && n && hipMemcpyAsync(dtext.ptr, text + h, n, hipMemcpyHostToDevice, st) != hipSuccess) {
            set_error("gs_text_to_bin_mem: copy to the device failed");
            rc = GS_ERR_DEVICE;
        }
        if (rc == GS_OK) rc = text_tokens(st, dtext.ptr, n, tok, &ntok, vmin, vmax);
        if (rc == GS_OK) {
            *pairs = ntok / 2;
            if (tokens) *tokens = ntok;
            const uint64_t k = std::min(cap_pairs, *pairs);
            if (out && k && (hipMemcpyAsync(out, tok.ptr, k * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                             hipStreamSynchronize(st) != hipSuccess)) {
                set_error("gs_text_to_bin_mem: copy to the host failed");
                rc = GS_ERR_DEVICE;
            }
        }
        (void)hipStreamSynchronize(st);
    }
    (void)hipStreamDestroy(st);
    return rc;
}

extern "C" int gs_text_to_bin(const char *text_path, const char *out_path, uint64_t *pairs, uint32_t *vmin,
                              uint32_t *vmax, uint64_t *tokens) {
    if (!text_path || !pairs || !vmin || !vmax) { set_error("gs_text_to_bin: null argument"); return GS_ERR_ARG; }
    struct stat sb;
    if (stat(text_path, &sb) != 0) { set_error("cannot stat %s", text_path); return GS_ERR_IO; }
    const uint64_t size = (uint64_t)sb.st_size;
    FILE *f = fopen(text_path, "rb");
    if (!f) { set_error("cannot open %s", text_path); return GS_ERR_IO; }
    // the whole text goes to HBM (tokens may cross any boundary); the host
    // reads it through a pinned buffer
    char *pin = nullptr;
    const uint64_t chunk = std::min<uint64_t>(size ? size : 1, 256ull << 20);
    if (hipHostMalloc((void **)&pin, chunk, hipHostMallocDefault) != hipSuccess) {
        fclose(f);
        set_error("pinned host buffer of %llu bytes", (unsigned long long)chunk);
        return GS_ERR_NOMEM;
    }
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipHostFree(pin);
        fclose(f);
        set_error("stream creation failed");
        return GS_ERR_DEVICE;
    }
    int rc = GS_OK;
    {
        DevBuf<uint8_t> dtext;
        DevBuf<uint32_t> tok;
        rc = dtext.alloc(size + 16);
        // header lines are read on the host from the first chunk
        uint64_t h = 0, done = 0;
        bool first = true;
        while (rc == GS_OK && done < size) {
            const uint64_t k = std::min(chunk, size - done);
            if (fread(pin, 1, k, f) != k) { set_error("short read of %s", text_path); rc = GS_ERR_IO; break; }
            uint64_t skip = 0;
            if (first) {
                skip = text_header(pin, k);
                if (skip == k && k < size) {   // a header longer than a chunk: not worth supporting
                    set_error("%s: header lines longer than %llu bytes", text_path, (unsigned long long)chunk);
                    rc = GS_ERR_FORMAT;
                    break;
                }
                h = skip;
                first = false;
            }
            if (hipMemcpyAsync(dtext.ptr + (done + skip - h), pin + skip, k - skip, hipMemcpyHostToDevice, st) !=
                    hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                set_error("copy to the device failed");
                rc = GS_ERR_DEVICE;
                break;
            }
            done += k;
        }
        uint64_t ntok = 0;
        if (rc == GS_OK) rc = text_tokens(st, dtext.ptr, size - h, tok, &ntok, vmin, vmax);
        if (rc == GS_OK) {
            *pairs = ntok / 2;
            if (tokens) *tokens = ntok;
            const std::string out = out_path ? std::string(out_path) : std::string(text_path) + ".edge";
            FILE *o = fopen(out.c_str(), "wb");
            if (!o) { set_error("cannot create %s", out.c_str()); rc = GS_ERR_IO; }
            const uint64_t bytes = *pairs * 8;
            for (uint64_t off = 0; rc == GS_OK && off < bytes; off += chunk) {
                const uint64_t k = std::min(chunk, bytes - off);
                if (hipMemcpyAsync(pin, (const uint8_t *)tok.ptr + off, k, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess) {
                    set_error("copy to the host failed");
                    rc = GS_ERR_DEVICE;
                } else if (fwrite(pin, 1, k, o) != k) {
                    set_error("short write of %s", out.c_str());
                    rc = GS_ERR_IO;
                }
            }
            if (o && fclose(o) != 0 && rc == GS_OK) { set_error("closing %s failed", out.c_str()); rc = GS_ERR_IO; }
        }
        (void)hipStreamSynchronize(st);
    }
    (void)hipStreamDestroy(st);
    (void)hipHostFree(pin);
    fclose(f);
    return rc;
}
