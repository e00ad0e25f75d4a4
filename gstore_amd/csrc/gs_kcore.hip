// gs_kcore.hip -- k-core peeling over the resident tile grid.
//
// Restates kcore_t (kcore.cpp:55-217) driven as kcore_mmap
// (gstore.cpp:2716-2804):
//   peel   m = (int16)deg[v]; if 0 < m < k: deg[v] = 0, v active
//          (kcore.cpp:186-196); the first peel runs before any edge pass
//   scan   for every stored edge in a tile whose row or column holds an
//          active vertex: if active(v0) --deg[v1]; if active(v1) --deg[v0]
//          (kcore.cpp:69-103, tile skip gstore.cpp:2768-2781)
//   until a peel removes nothing; result count(deg != 0) (gstore.cpp:2794-2800)
// Decrements are integer atomics, so the final degree array is independent
// of order and bit-exact against the CPU restatement.  The active set is an
// N-bit bitmap built with one wave ballot per 64 vertices.
#include <algorithm>
#include <vector>

#include "gs_internal.h"

namespace gs {

// One workgroup per tile row (65536 vertices, 256 steps of 256).
__global__ void __launch_bounds__(256) k_kc_peel(int32_t *deg, uint64_t n, int k, uint32_t *act, uint32_t *rowact,
                                                 unsigned long long *count) {
    const uint32_t r = blockIdx.x;
    const uint64_t rbase = (uint64_t)r << 16;
    uint32_t cnt = 0;
    const int lane = threadIdx.x & 63;
    for (uint32_t step = 0; step < 256; ++step) {
        uint64_t v = rbase + step * 256 + threadIdx.x;
        bool pred = false;
        if (v < n) {
            int16_t m = (int16_t)deg[v];
            if (m > 0 && m < k) {
                deg[v] = 0;
                pred = true;
            }
        }
        unsigned long long mask = __ballot(pred);
        if (lane == 0) {
            uint64_t wv = v >> 5;  // v is the wave's first vertex here, a multiple of 64
            act[wv] = (uint32_t)mask;
            act[wv + 1] = (uint32_t)(mask >> 32);
            cnt += __popcll(mask);
        }
    }
    int any = __syncthreads_or(cnt != 0);
    __shared__ unsigned long long tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    if (cnt) atomicAdd(&tot, (unsigned long long)cnt);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (any) atomicOr(&rowact[r >> 5], 1u << (r & 31));
        else atomicAnd(&rowact[r >> 5], ~(1u << (r & 31)));
        if (tot) atomicAdd(count, tot);
    }
}

__global__ void __launch_bounds__(SWEEP_WG) k_kc_scan(SweepArgs a, const uint32_t *__restrict__ act, int32_t *deg,
                                                      const uint32_t *__restrict__ rowact,
                                                      unsigned long long *scanned) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [&](uint32_t rc) {
        return bit_of(rowact, rc >> 16) || bit_of(rowact, rc & 0xFFFFu);
    });
    if (!s.any) return;
    count_active_edges(a, chunk, s, scanned);
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        if ((act[v0 >> 5] >> (v0 & 31)) & 1u) atomicSub(&deg[v1], 1);
        if ((act[v1 >> 5] >> (v1 & 31)) & 1u) atomicSub(&deg[v0], 1);
    };
    sweep_chunk<true>(a, chunk, s, len, op);
}

// count(deg != 0) and count(deg > 0): four degrees per 16-byte load, summed
// over the wave and the workgroup, one global atomic per workgroup and count
// (an atomic per thread on the same two words had made this pass 0.9 ms at
// 2^28 vertices, most of which have a non-zero final degree).
__global__ void __launch_bounds__(256) k_kc_count(const int32_t *__restrict__ deg, uint64_t n,
                                                  unsigned long long *out) {
    __shared__ uint32_t s_ne, s_pos;
    if (threadIdx.x == 0) s_ne = s_pos = 0;
    __syncthreads();
    uint32_t ne = 0, pos = 0;
    const uint64_t n4 = n / 4;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x) {
        const int4 d = reinterpret_cast<const int4 *>(deg)[i];
        ne += (d.x != 0) + (d.y != 0) + (d.z != 0) + (d.w != 0);
        pos += (d.x > 0) + (d.y > 0) + (d.z > 0) + (d.w > 0);
    }
    if (blockIdx.x == 0 && threadIdx.x < n - n4 * 4) {
        const int32_t d = deg[n4 * 4 + threadIdx.x];
        ne += d != 0;
        pos += d > 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ne += __shfl_xor(ne, o);
        pos += __shfl_xor(pos, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&s_ne, ne);
        atomicAdd(&s_pos, pos);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_ne) atomicAdd(&out[0], (unsigned long long)s_ne);
        if (s_pos) atomicAdd(&out[1], (unsigned long long)s_pos);
    }
}

// ---- peeling over the neighbour-list index (relabelled ids) -----------------
// The same rule; the removed vertices are listed (with their list lengths)
// instead of marked, and their lists are pushed: one decrement per list entry,
// which is one per stored edge with an active endpoint, as the tile scan does.

// Peel of relabelled ids 0 .. nzn-1 (the others have no edges), appending the
// removed ones to queue / qlen; count[0] = how many.  Each thread takes
// KC_VPT consecutive vertices (two 16-byte degree loads), so a workgroup
// reserves queue slots once per 2048 vertices: one reservation per 256
// serialised on the cursor when most vertices peel in the first round.
constexpr int KC_VPT = 8;

__global__ void __launch_bounds__(256) k_kc_peel_list(int32_t *deg, uint64_t nzn, int k,
                                                      const uint64_t *__restrict__ off, uint32_t *queue,
                                                      int32_t *qlen, unsigned long long *count) {
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long bbase;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (uint64_t base = blockIdx.x * (256ull * KC_VPT); base < nzn; base += (uint64_t)gridDim.x * 256 * KC_VPT) {
        const uint64_t x0 = base + threadIdx.x * KC_VPT;
        int32_t d[KC_VPT];
        const bool whole = x0 + KC_VPT <= nzn;
        if (whole) {
            const int4 a = reinterpret_cast<const int4 *>(deg + x0)[0], b = reinterpret_cast<const int4 *>(deg + x0)[1];
            d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
        } else {
#pragma unroll
            for (int i = 0; i < KC_VPT; ++i) d[i] = x0 + i < nzn ? deg[x0 + i] : 0;
        }
        uint32_t pm = 0;
#pragma unroll
        for (int i = 0; i < KC_VPT; ++i) {
            const int16_t m = (int16_t)d[i];
            if (m > 0 && m < k) {
                pm |= 1u << i;
                d[i] = 0;
            }
        }
        if (pm) {
            if (whole) {
                reinterpret_cast<int4 *>(deg + x0)[0] = int4{d[0], d[1], d[2], d[3]};
                reinterpret_cast<int4 *>(deg + x0)[1] = int4{d[4], d[5], d[6], d[7]};
            } else {
#pragma unroll
                for (int i = 0; i < KC_VPT; ++i)
                    if ((pm >> i) & 1u) deg[x0 + i] = 0;
            }
        }
        const uint32_t c = __popc(pm);
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
            bbase = t ? atomicAdd(&count[0], (unsigned long long)t) : 0ull;
        }
        __syncthreads();
        if (pm) {
            uint32_t before = incl - c;
            for (int j = 0; j < wv; ++j) before += wsum[j];
            unsigned long long pos = bbase + before;
#pragma unroll
            for (int i = 0; i < KC_VPT; ++i)
                if ((pm >> i) & 1u) {
                    queue[pos] = (uint32_t)(x0 + i);
                    qlen[pos++] = (int32_t)(off[x0 + i + 1] - off[x0 + i]);
                }
        }
        __syncthreads();   // wsum and bbase are rewritten by the next step
    }
}

// One decrement per list entry of the listed vertices.  The KC_HUB heaviest
// vertices (relabelled ids below KC_HUB) collect their decrements in LDS
// first, and the workgroups (at most KC_GRID) stride over the queue, so a
// hub's count is flushed with one global atomic per workgroup.  2^15 hub
// counters fill 128 KiB of LDS, one 1024-thread workgroup per CU: at Kron-28
// k = 8 the pushes took 5.8 ms per run, against 6.7 with 4096 counters in
// 256-thread workgroups (16384: 6.1, 8192: 6.4).
constexpr uint32_t KC_HUB = 32768;
constexpr uint32_t KC_GRID = 512;
constexpr int KC_WG = 1024;

// A lane per listed vertex pushes a list of at most KC_LANE_MAX entries
// alone; longer lists (a vertex whose degree fell below k keeps its whole
// list, and the (int16) test can list a hub) are pushed by the whole wave, 64
// entries per step.  No prefix of the list lengths and no search per entry
// (the edge-balanced push with both took 7.6 ms per run at s28, this one 6.8);
// *pushed counts the entries.
constexpr uint32_t KC_LANE_MAX = 64;

__global__ void __launch_bounds__(KC_WG) k_kc_push(const uint32_t *__restrict__ queue, const int32_t *__restrict__ qlen,
                                                 uint64_t nq, const uint64_t *__restrict__ off,
                                                 const uint32_t *__restrict__ nbr, int32_t *deg,
                                                 unsigned long long *pushed) {
    __shared__ uint32_t hub[KC_HUB];
    for (uint32_t i = threadIdx.x; i < KC_HUB; i += blockDim.x) hub[i] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    auto push = [&](uint32_t y) {
        if (y < KC_HUB) atomicAdd(&hub[y], 1u);
        else atomicSub(&deg[y], 1);
    };
    unsigned long long cnt = 0;
    for (uint64_t w0 = ((blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6) * 64; w0 < nq; w0 += nw * 64) {
        const uint64_t i = w0 + lane;
        uint32_t v = 0, len = 0;
        if (i < nq) {
            v = queue[i];
            len = (uint32_t)qlen[i];
        }
        if (len && len <= KC_LANE_MAX) {
            const uint64_t b = off[v];
            for (uint32_t k = 0; k < len; ++k) push(nbr[b + k]);
        }
        cnt += len;
        unsigned long long heavy = __ballot(len > KC_LANE_MAX);
        while (heavy) {   // wave-uniform
            const int l = __ffsll((long long)heavy) - 1;
            heavy &= heavy - 1;
            const uint32_t hv = __shfl(v, l), hl = __shfl(len, l);
            const uint64_t b = off[hv];
            for (uint32_t k = lane; k < hl; k += 64) push(nbr[b + k]);
        }
    }
    block_count(cnt, pushed);
    for (uint32_t i = threadIdx.x; i < KC_HUB; i += blockDim.x)
        if (hub[i]) atomicSub(&deg[i], (int32_t)hub[i]);
}

static unsigned kc_grid(uint64_t n) { return (unsigned)std::min<uint64_t>((n + 255) / 256, 8192); }

}  // namespace gs

using namespace gs;

// k-core over the index: degrees in relabelled ids, unpermuted at the end.
static int run_kcore_index(const gs_graph *gc, gs_ws *w, int k, int *rounds_out) {
    gs_graph *g = const_cast<gs_graph *>(gc);
    // the vertices with edges only: the count and the unpermute read no degree above nzn
    GS_HIP(hipMemcpyAsync(w->kc_deg.ptr, g->deg_r.ptr, g->nzn * 4, hipMemcpyDeviceToDevice, w->stream));
    GS_HIP(hipMemsetAsync(w->counters.ptr + 1, 0, 8, w->stream));   // entries pushed, all rounds
    int nr = 0;
    for (;;) {
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 8, w->stream));
        hipEvent_t kb;
        w->timer.begin(GS_K_KCORE_PEEL, w->stream, &kb);
        if (g->nzn)
            k_kc_peel_list<<<kc_grid((g->nzn + KC_VPT - 1) / KC_VPT), 256, 0, w->stream>>>(w->kc_deg.ptr, g->nzn, k, g->adj_off.ptr,
                                                                   w->queue.ptr, w->qdeg.ptr, w->counters.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_KCORE_PEEL, w->stream, kb);
        GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        ++nr;
        const uint64_t nq = w->h_counters[0];
        if (nq == 0) break;
        if (nr >= 1000000) {
            set_error("gs_kcore: still removing vertices after %d rounds", nr);
            return GS_ERR_DEVICE;
        }
        w->timer.begin(GS_K_KCORE_SCAN, w->stream, &kb);
        k_kc_push<<<(unsigned)std::min<uint64_t>((nq + KC_WG - 1) / KC_WG, KC_GRID), KC_WG, 0, w->stream>>>(
            w->queue.ptr, w->qdeg.ptr, nq, g->adj_off.ptr, g->adj_nbr.ptr, w->kc_deg.ptr, w->counters.ptr + 1);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_KCORE_SCAN, w->stream, kb);
    }
    GS_HIP(hipMemcpyAsync(w->h_counters + 1, w->counters.ptr + 1, 8, hipMemcpyDeviceToHost, w->stream));
    GS_HIP(hipStreamSynchronize(w->stream));
    w->stats.edges_scanned = w->h_counters[1];
    *rounds_out = nr;
    return GS_OK;
}

int gs::run_kcore(const gs_graph *g, gs_ws *w, int k, int32_t *deg_out, uint64_t *count_ne0,
                   uint64_t *count_pos, int *rounds) {
    if (!g) { set_error("gs_kcore: null graph"); return GS_ERR_ARG; }
    if (g->directed) {
        set_error("k-core is not applicable to directed graphs");
        return GS_ERR_UNSUPPORTED;
    }
    GS_TRY(ensure_bitmaps(g, w));
    if (w->kc_deg.n != g->n) GS_TRY(w->kc_deg.alloc(g->n));
    if (g->kcore_method == 1 && !g->ss) {
        GS_TRY(build_adjacency(const_cast<gs_graph *>(g)));
        GS_TRY(ensure_queue(g, w));
    }
    w->timer.reset();
    w->stats = gs_stats{};
    EventPair ev_t0;
    GS_HIP(hipEventCreate(&ev_t0.a));
    GS_HIP(hipEventCreate(&ev_t0.b));
    hipEvent_t t0 = ev_t0.a, t1 = ev_t0.b;
    GS_HIP(hipEventRecord(t0, w->stream));
    const bool by_index = g->kcore_method == 1 && !g->ss;
    int nr = 0;
    if (by_index) {
        GS_TRY(run_kcore_index(g, w, k, &nr));
    } else {
        GS_HIP(hipMemcpyAsync(w->kc_deg.ptr, g->degree.ptr, g->n * 4, hipMemcpyDeviceToDevice, w->stream));
        GS_HIP(hipMemsetAsync(w->rowact.ptr, 0, w->rowact.bytes(), w->stream));
        GS_HIP(hipMemsetAsync(w->counters.ptr + 2, 0, 16, w->stream));
    }
    for (; !by_index;) {
        GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 8, w->stream));
        hipEvent_t kb;
        w->timer.begin(GS_K_KCORE_PEEL, w->stream, &kb);
        k_kc_peel<<<(unsigned)g->rows, 256, 0, w->stream>>>(w->kc_deg.ptr, g->n, k, w->bm_front.ptr, w->rowact.ptr,
                                                            w->counters.ptr);
        GS_HIP(hipGetLastError());
        w->timer.end(GS_K_KCORE_PEEL, w->stream, kb);
        GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 8, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
        ++nr;
        if (w->h_counters[0] == 0 || !g->nchunk) break;
        if (nr >= 1000000) {
            set_error("gs_kcore: still removing vertices after %d rounds", nr);
            return GS_ERR_DEVICE;
        }
        GS_TRY(sweep(g, w, GS_K_KCORE_SCAN, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            k_kc_scan<<<grid, SWEEP_WG, 0, st>>>(a, w->bm_front.ptr, w->kc_deg.ptr, w->rowact.ptr, w->counters.ptr + 2);
        }));
    }
    GS_HIP(hipMemsetAsync(w->counters.ptr, 0, 16, w->stream));
    const uint64_t ncount = by_index ? g->nzn : g->n;   // over the index: degrees of the vertices with edges
    k_kc_count<<<(unsigned)std::min<uint64_t>((ncount / 4 + 255) / 256 + 1, 2048), 256, 0, w->stream>>>(
        w->kc_deg.ptr, ncount, w->counters.ptr);
    GS_HIP(hipGetLastError());
    // the final degrees in vertex order: relabelled ones are mapped back into qdeg (N entries)
    const int32_t *final_deg = w->kc_deg.ptr;
    if (by_index) {
        GS_TRY(unpermute_i32(g, w->stream, w->kc_deg.ptr, w->qdeg.ptr));
        final_deg = w->qdeg.ptr;
    }
    GS_HIP(hipEventRecord(t1, w->stream));
    GS_HIP(hipMemcpyAsync(w->h_counters, w->counters.ptr, 32, hipMemcpyDeviceToHost, w->stream));
    GS_TRY(sync_and_collect(w));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, t0, t1));
    w->stats.ms = ms;
    w->stats.iterations = (uint64_t)nr;
    if (!by_index) {
        w->stats.edges_scanned = w->h_counters[2];
        w->stats.tiles_scanned = w->h_counters[3];
    }
    if (count_ne0) *count_ne0 = w->h_counters[0];
    if (count_pos) *count_pos = w->h_counters[1];
    if (rounds) *rounds = nr;
    if (deg_out) {
        GS_HIP(hipMemcpyAsync(deg_out, final_deg, g->n * 4, hipMemcpyDeviceToHost, w->stream));
        GS_HIP(hipStreamSynchronize(w->stream));
    }
    return GS_OK;
}

extern "C" int gs_kcore(gs_graph *g, int k, int32_t *deg_out, uint64_t *count_ne0, uint64_t *count_pos,
                        int *rounds) {
    if (!g) { set_error("gs_kcore: null graph"); return GS_ERR_ARG; }
    return run_kcore(g, g->ws, k, deg_out, count_ne0, count_pos, rounds);
}
