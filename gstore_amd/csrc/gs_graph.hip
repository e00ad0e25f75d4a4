// gs_graph.hip -- graph handle: loading tile files into HBM, device-side
// Kronecker generation and tile conversion, tile metadata, saving, errors.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <memory>

#include "gs_internal.h"

namespace gs {

static thread_local std::string g_err;

// Every failure path records its message here.  A failed HIP call also
// leaves HIP's per-thread last error set until it is read; it is cleared
// here, so that the failure is reported once and a later launch check on
// this thread (hipGetLastError after a kernel) does not report it again.
void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    (void)hipGetLastError();
}

const char *error_text() { return g_err.c_str(); }

// ---- kernel timer ------------------------------------------------------
void KernelTimer::begin(int k, hipStream_t s, hipEvent_t *out) {
    *out = nullptr;
    if (!on) return;
    if (hipEventCreate(out) != hipSuccess) { *out = nullptr; return; }
    (void)hipEventRecord(*out, s);
    (void)k;
}
void KernelTimer::end(int k, hipStream_t s, hipEvent_t b) {
    if (!on || !b) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, s);
    pending.push_back({k, {b, e}});
}
int KernelTimer::collect() {
    for (auto &p : pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, p.second.first, p.second.second) == hipSuccess) {
            ms[p.first] += t;
            calls[p.first] += 1;
        }
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    pending.clear();
    return GS_OK;
}
void KernelTimer::reset() {
    for (int i = 0; i < GS_K_COUNT; ++i) { ms[i] = 0; calls[i] = 0; }
}

int sync_and_collect(gs_ws *w) {
    GS_HIP(hipStreamSynchronize(w->stream));
    return w->timer.collect();
}

uint64_t bitmap_words(const gs_graph *g) { return g->rows * ROW_WORDS; }

int ensure_rowact(const gs_graph *g, gs_ws *w) {
    uint64_t rw = (g->p_s + 31) / 32;
    if (w->rowact.n != rw) GS_TRY(w->rowact.alloc(rw));
    return GS_OK;
}

int ensure_bitmaps(const gs_graph *g, gs_ws *w) {
    uint64_t wd = bitmap_words(g);
    if (w->bm_vis.n != wd) {
        GS_TRY(w->bm_vis.alloc(wd));
        GS_TRY(w->bm_front.alloc(wd));
        GS_TRY(w->bm_next.alloc(wd));
    }
    return ensure_rowact(g, w);
}

int ws_create(gs_ws **out) {
    std::unique_ptr<gs_ws> w(new gs_ws());
    GS_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    GS_TRY(w->counters.alloc(16));
    GS_HIP(hipHostMalloc((void **)&w->h_counters, 16 * sizeof(unsigned long long), hipHostMallocDefault));
    *out = w.release();
    return GS_OK;
}

int graph_create(gs_graph **out) {
    std::unique_ptr<gs_graph> g(new gs_graph());
    GS_HIP(hipGetDevice(&g->device));
    GS_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    GS_TRY(g->counters.alloc(16));
    GS_HIP(hipHostMalloc((void **)&g->h_counters, 16 * sizeof(unsigned long long), hipHostMallocDefault));
    GS_TRY(ws_create(&g->ws));
    *out = g.release();
    return GS_OK;
}

}  // namespace gs

gs_ws::~gs_ws() {
    if (stream) (void)hipStreamSynchronize(stream);
    timer.collect();
    if (ms_graph) (void)hipGraphExecDestroy(ms_graph);
    if (do_graph) (void)hipGraphExecDestroy(do_graph);
    if (h_counters) (void)hipHostFree(h_counters);
    if (stream) (void)hipStreamDestroy(stream);
}

size_t gs_ws::device_bytes() const {
    return pr_prior.bytes() + pr_acc.bytes() + pr_fac.bytes() + ppr_rank.bytes() + ppr_acc.bytes() + depth.bytes() + bm_vis.bytes() + bm_front.bytes() +
           bm_next.bytes() + rowact.bytes() + wcc_parent.bytes() + wcc_label.bytes() + wcc_cmin.bytes() + wcc_samp.bytes() + kc_deg.bytes() + ms_vis.bytes() +
           ms_front.bytes() + ms_next.bytes() + ms_roots.bytes() + ms_depth.bytes() + ms_depth_vm.bytes() + ms_counts.bytes() +
           queue.bytes() + qdeg.bytes() + qoff.bytes() + qsum.bytes() + depth_r.bytes() + counters.bytes() + ms_ctl.bytes() + do_ctl.bytes() +
           bin_cnt.bytes() + bin_off.bytes() + bin_cur.bytes() + bin_uoff.bytes() + bin_pc.bytes() + bin_pr.bytes() + bin_o.bytes() + bin_v.bytes();
}

gs_graph::~gs_graph() {
    delete ws;
    delete ss;
    if (stream) (void)hipStreamSynchronize(stream);
    if (h_counters) (void)hipHostFree(h_counters);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace gs {

static int set_geometry(gs_graph *g, uint64_t n, bool directed) {
    if (n == 0 || n > (1ull << 32)) {
        set_error("vertex count %llu out of range (1 .. 2^32)", (unsigned long long)n);
        return GS_ERR_ARG;
    }
    g->n = n;
    g->directed = directed;
    g->p = groups_for(n);
    g->p_s = 256 * g->p;
    g->T = total_slots(g->p_s, directed);
    g->rows = (n + 65535) >> 16;
    return GS_OK;
}

// Non-empty tile list and chunk map from the host copy of .start.
int graph_build_meta(gs_graph *g, const uint64_t *st) {
    std::vector<uint64_t> beg;
    std::vector<uint32_t> rc;
    uint64_t t = 0;
    const uint64_t p = g->p;
    for (uint64_t bi = 0; bi < p; ++bi)
        for (uint64_t bj = g->directed ? 0 : bi; bj < p; ++bj)
            for (uint64_t si = 0; si < 256; ++si)
                for (uint64_t sj = (!g->directed && bi == bj) ? si : 0; sj < 256; ++sj, ++t) {
                    if (st[t + 1] == st[t]) continue;
                    beg.push_back(st[t]);
                    rc.push_back((uint32_t)(((bi * 256 + si) << 16) | (bj * 256 + sj)));
                }
    if (t != g->T) {
        set_error("internal: tile walk visited %llu slots, expected %llu", (unsigned long long)t,
                  (unsigned long long)g->T);
        return GS_ERR_FORMAT;
    }
    g->nz = beg.size();
    beg.push_back(g->E);
    GS_TRY(g->tz_beg.alloc(beg.size()));
    GS_TRY(g->tz_rc.alloc(std::max<size_t>(rc.size(), 1)));
    GS_HIP(hipMemcpy(g->tz_beg.ptr, beg.data(), beg.size() * 8, hipMemcpyHostToDevice));
    if (!rc.empty()) GS_HIP(hipMemcpy(g->tz_rc.ptr, rc.data(), rc.size() * 4, hipMemcpyHostToDevice));
    return build_chunk_map(g, beg);
}

// chunk_tz[c] = non-empty tile holding edge c * g->chunk; chunk_tz[nchunk] = nz - 1.
int build_chunk_map(gs_graph *g, const std::vector<uint64_t> &beg) {
    g->nchunk = (g->E + g->chunk - 1) / g->chunk;
    std::vector<uint32_t> ct(g->nchunk + 1);
    uint64_t k = 0;
    for (uint64_t c = 0; c < g->nchunk; ++c) {
        uint64_t e = c * g->chunk;
        while (beg[k + 1] <= e) ++k;
        ct[c] = (uint32_t)k;
    }
    ct[g->nchunk] = g->nz ? (uint32_t)(g->nz - 1) : 0;
    GS_TRY(g->chunk_tz.alloc(ct.size()));
    GS_HIP(hipMemcpy(g->chunk_tz.ptr, ct.data(), ct.size() * 4, hipMemcpyHostToDevice));
    return GS_OK;
}

// ---- validation: ids < N and degree file == endpoint counts -----------------
__global__ void __launch_bounds__(SWEEP_WG) k_validate(SweepArgs a, uint32_t n, int directed,
                                                       int32_t *gdeg, uint32_t *in_touch, unsigned long long *bad) {
    __shared__ ChunkTiles s;
    uint64_t chunk;
    if (!launch_chunk(a, chunk)) return;
    uint32_t len = stage_chunk(a, chunk, s, [](uint32_t) { return true; });
    unsigned long long nbad = 0;
    auto op = [&](uint32_t w, uint32_t rc) {
        uint32_t v0 = gv0(w, rc), v1 = gv1(w, rc);
        if (v0 >= n || v1 >= n) { ++nbad; return; }
        atomicAdd(&gdeg[v0], 1);
        if (!directed) atomicAdd(&gdeg[v1], 1);
        else if (!((in_touch[v1 >> 5] >> (v1 & 31)) & 1u)) atomicOr(&in_touch[v1 >> 5], 1u << (v1 & 31));
    };
    sweep_chunk<false>(a, chunk, s, len, op);
    if (nbad) atomicAdd(bad, nbad);
}

__global__ void k_count_mismatch(const int32_t *x, const int32_t *y, uint64_t n, unsigned long long *bad) {
    unsigned long long c = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        c += (x[i] != y[i]);
    if (c) atomicAdd(bad, c);
}

static int validate(gs_graph *g) {
    DevBuf<int32_t> gdeg;
    GS_TRY(gdeg.alloc(g->n));
    if (g->directed) {
        // vertices with an in-edge: with .degree (out-degree) they are the vertices with a stored edge
        GS_TRY(g->in_touch.alloc(bitmap_words(g)));
        GS_HIP(hipMemsetAsync(g->in_touch.ptr, 0, g->in_touch.bytes(), g->stream));
    }
    GS_HIP(hipMemsetAsync(gdeg.ptr, 0, g->n * 4, g->stream));
    GS_HIP(hipMemsetAsync(g->counters.ptr, 0, 16 * 8, g->stream));
    {
        // the validation sweep runs on the default workspace's stream
        gs_ws *w = g->ws;
        GS_HIP(hipStreamSynchronize(g->stream));
        uint32_t n32 = (uint32_t)std::min<uint64_t>(g->n, 0xFFFFFFFFull);
        GS_TRY(sweep(g, w, GS_K_CONVERT, [&](const SweepArgs &a, unsigned grid, hipStream_t st) {
            k_validate<<<grid, SWEEP_WG, 0, st>>>(a, n32, g->directed, gdeg.ptr, g->in_touch.ptr, g->counters.ptr);
        }));
        GS_HIP(hipStreamSynchronize(w->stream));
        w->timer.collect();
        w->timer.reset();
    }
    k_count_mismatch<<<1024, 256, 0, g->stream>>>(gdeg.ptr, g->degree.ptr, g->n, g->counters.ptr + 1);
    GS_HIP(hipGetLastError());
    GS_HIP(hipMemcpyAsync(g->h_counters, g->counters.ptr, 16, hipMemcpyDeviceToHost, g->stream));
    GS_HIP(hipStreamSynchronize(g->stream));
    if (g->h_counters[0]) {
        set_error("%llu stored edges have a vertex id >= N=%llu", g->h_counters[0], (unsigned long long)g->n);
        return GS_ERR_FORMAT;
    }
    if (g->h_counters[1]) {
        set_error("degree array disagrees with the grid's endpoint counts at %llu vertices", g->h_counters[1]);
        return GS_ERR_FORMAT;
    }
    return GS_OK;
}

static int check_start(const gs_graph *g, const uint64_t *st) {
    if (st[0] != 0) { set_error(".start[0] = %llu, expected 0", (unsigned long long)st[0]); return GS_ERR_FORMAT; }
    for (uint64_t t = 0; t < g->T; ++t)
        if (st[t + 1] < st[t]) {
            set_error(".start not monotone at slot %llu", (unsigned long long)t);
            return GS_ERR_FORMAT;
        }
    return GS_OK;
}

static int open_common(gs_graph *g, const uint64_t *h_start) {
    GS_TRY(check_start(g, h_start));
    GS_TRY(graph_build_meta(g, h_start));
    return validate(g);
}

static int read_file(const std::string &path, void *dst, uint64_t bytes) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { set_error("cannot open %s", path.c_str()); return GS_ERR_IO; }
    uint64_t got = 0;
    const uint64_t step = 1ull << 28;
    while (got < bytes) {
        size_t want = (size_t)std::min<uint64_t>(step, bytes - got);
        size_t r = fread((char *)dst + got, 1, want, f);
        if (r == 0) break;
        got += r;
    }
    fclose(f);
    if (got != bytes) {
        set_error("short read on %s: %llu of %llu bytes", path.c_str(), (unsigned long long)got, (unsigned long long)bytes);
        return GS_ERR_IO;
    }
    return GS_OK;
}

static int file_size(const std::string &path, uint64_t *sz) {
    struct stat st;
    if (stat(path.c_str(), &st) != 0) { set_error("cannot stat %s", path.c_str()); return GS_ERR_IO; }
    *sz = (uint64_t)st.st_size;
    return GS_OK;
}

static int write_file(const std::string &path, const void *src, uint64_t bytes) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { set_error("cannot create %s", path.c_str()); return GS_ERR_IO; }
    uint64_t put = 0;
    const uint64_t step = 1ull << 28;
    while (put < bytes) {
        size_t want = (size_t)std::min<uint64_t>(step, bytes - put);
        size_t w = fwrite((const char *)src + put, 1, want, f);
        if (w != want) break;
        put += w;
    }
    int rc = fclose(f);
    if (put != bytes || rc != 0) { set_error("short write on %s", path.c_str()); return GS_ERR_IO; }
    return GS_OK;
}

// ---- Kronecker generator (graph500-generator/graph_generator.c:41-198) ------
// MRG32k5 state transition z' = A z with A's first row (x, 0, 0, 0, y) and a
// shifted identity below (splittable_mrg.c:25-57).  Tuple e starts from
// A^(e * 2^64) * seed; the byte tables hold A^(v * 256^b * 2^64).
constexpr uint64_t MP = 0x7FFFFFFFull;
__host__ __device__ inline uint32_t mmod(uint64_t v) {
    v = (v & MP) + (v >> 31);
    v = (v & MP) + (v >> 31);
    return (uint32_t)(v >= MP ? v - MP : v);
}

struct M5 { uint32_t m[5][5]; };

static M5 m5_mul(const M5 &a, const M5 &b) {
    M5 r;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            uint64_t s = 0;
            for (int k = 0; k < 5; ++k) s = mmod(s + mmod((uint64_t)a.m[i][k] * b.m[k][j]));
            r.m[i][j] = (uint32_t)s;
        }
    return r;
}
static M5 m5_eye() {
    M5 r{};
    for (int i = 0; i < 5; ++i) r.m[i][i] = 1;
    return r;
}
static M5 m5_pow(M5 b, uint64_t e) {
    M5 r = m5_eye();
    while (e) {
        if (e & 1) r = m5_mul(r, b);
        b = m5_mul(b, b);
        e >>= 1;
    }
    return r;
}
static void m5_apply(const M5 &a, uint32_t z[5]) {
    uint32_t o[5];
    for (int i = 0; i < 5; ++i) {
        uint64_t s = 0;
        for (int k = 0; k < 5; ++k) s = mmod(s + mmod((uint64_t)a.m[i][k] * z[k]));
        o[i] = (uint32_t)s;
    }
    memcpy(z, o, sizeof o);
}
__host__ __device__ inline uint32_t mrg_draw(uint32_t z[5]) {
    uint32_t nz = mmod((uint64_t)mmod(107374182ull * z[0]) + mmod(104480ull * z[4]));
    z[4] = z[3]; z[3] = z[2]; z[2] = z[1]; z[1] = z[0]; z[0] = nz;
    return nz;
}

struct KronParams {
    uint32_t seed[5];
    uint64_t k0, k1;
    int lg;
};

static void kron_setup(int lg, uint64_t s1, uint64_t s2, KronParams *kp, std::vector<uint32_t> *tables) {
    kp->lg = lg;
    kp->seed[0] = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
    kp->seed[1] = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
    kp->seed[2] = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
    kp->seed[3] = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
    kp->seed[4] = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
    M5 a{};
    a.m[0][0] = 107374182u;
    a.m[0][4] = 104480u;
    for (int i = 1; i < 5; ++i) a.m[i][i - 1] = 1;
    M5 a64 = a;
    for (int i = 0; i < 64; ++i) a64 = m5_mul(a64, a64);
    M5 a128 = a64;
    for (int i = 0; i < 64; ++i) a128 = m5_mul(a128, a128);
    uint32_t z[5];
    memcpy(z, kp->seed, sizeof z);
    m5_apply(m5_pow(a64, 7), z);
    m5_apply(m5_pow(a128, 50), z);
    uint64_t k0 = mrg_draw(z); k0 *= 0xFFFFFFFFull; k0 += mrg_draw(z);
    uint64_t k1 = mrg_draw(z); k1 *= 0xFFFFFFFFull; k1 += mrg_draw(z);
    kp->k0 = k0;
    kp->k1 = k1;
    tables->assign(8 * 256 * 25, 0);
    M5 base = a64;
    for (int b = 0; b < 8; ++b) {
        M5 cur = m5_eye();
        for (int v = 0; v < 256; ++v) {
            memcpy(&(*tables)[(b * 256 + v) * 25], cur.m, 25 * 4);
            cur = m5_mul(cur, base);
        }
        base = m5_pow(base, 256);
    }
}

__device__ inline uint64_t scramble_dev(uint64_t v, int lg, uint64_t k0, uint64_t k1) {
    v += k0 + k1;
    v *= (k0 | 0x4519840211493211ull);
    v = __brevll(v) >> (64 - lg);
    v *= (k1 | 0x3050852102C843A5ull);
    v = __brevll(v) >> (64 - lg);
    return v;
}

__global__ void __launch_bounds__(256) k_kron(KronParams kp, const uint32_t *__restrict__ tab, int64_t e0, int64_t e1,
                                              uint2 *__restrict__ out) {
    const uint32_t limit = 0xFFFFFFFFu % 10000u;
    for (int64_t e = e0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < e1;
         e += (int64_t)gridDim.x * blockDim.x) {
        uint32_t z[5] = {kp.seed[0], kp.seed[1], kp.seed[2], kp.seed[3], kp.seed[4]};
        uint64_t ee = (uint64_t)e;
        for (int b = 0; ee; ++b, ee >>= 8) {
            uint32_t v = (uint32_t)(ee & 255);
            if (!v) continue;
            const uint32_t *m = tab + (b * 256 + v) * 25;
            uint32_t o[5];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                uint64_t s = 0;
#pragma unroll
                for (int k = 0; k < 5; ++k) s += mmod((uint64_t)m[i * 5 + k] * z[k]);
                o[i] = mmod(s);
            }
#pragma unroll
            for (int i = 0; i < 5; ++i) z[i] = o[i];
        }
        uint64_t src = 0, dst = 0, half = 1ull << kp.lg;
        for (int l = 0; l < kp.lg; ++l) {
            uint32_t v = mrg_draw(z);
            while (v < limit) v = mrg_draw(z);
            v %= 10000u;
            int q = v < 1900 ? 1 : v < 3800 ? 2 : v < 9500 ? 0 : 3;
            int so = q >> 1, to = q & 1;
            if (src == dst && so > to) { so = 0; to = 1; }
            half >>= 1;
            src += half * so;
            dst += half * to;
        }
        out[e - e0] = make_uint2((uint32_t)scramble_dev(src, kp.lg, kp.k0, kp.k1),
                                 (uint32_t)scramble_dev(dst, kp.lg, kp.k0, kp.k1));
    }
}

static int kron_device(int lg, uint64_t s1, uint64_t s2, int64_t e0, int64_t e1, uint2 *d_out, hipStream_t st) {
    if (lg < 1 || lg > 32) { set_error("scale %d out of range 1..32", lg); return GS_ERR_ARG; }
    KronParams kp;
    std::vector<uint32_t> tab;
    kron_setup(lg, s1, s2, &kp, &tab);
    DevBuf<uint32_t> d_tab;
    GS_TRY(d_tab.alloc(tab.size()));
    GS_HIP(hipMemcpyAsync(d_tab.ptr, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, st));
    int64_t cnt = e1 - e0;
    if (cnt > 0) {
        uint64_t blocks = std::min<uint64_t>((cnt + 255) / 256, 1u << 20);
        k_kron<<<(unsigned)blocks, 256, 0, st>>>(kp, d_tab.ptr, e0, e1, d_out);
        GS_HIP(hipGetLastError());
    }
    GS_HIP(hipStreamSynchronize(st));
    return GS_OK;
}

// ---- converter (gstore.cpp:1245-1458) ---------------------------------------
__global__ void k_conv_count(const uint2 *__restrict__ ed, uint64_t m, uint32_t n, uint64_t p, int directed,
                             unsigned long long *cnt, int32_t *deg, unsigned long long *bad) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        uint2 e = ed[i];
        if (e.x == e.y) continue;
        if (e.x >= n || e.y >= n) { atomicAdd(bad, 1ull); continue; }
        uint32_t v0 = e.x, v1 = e.y;
        if (!directed && v0 > v1) { v0 = e.y; v1 = e.x; }
        atomicAdd(&cnt[tile_slot(v0 >> 16, v1 >> 16, p, directed)], 1ull);
        atomicAdd(&deg[e.x], 1);
        if (!directed) atomicAdd(&deg[e.y], 1);
    }
}

__global__ void k_conv_scatter(const uint2 *__restrict__ ed, uint64_t m, uint64_t p, int directed,
                               unsigned long long *cur, uint32_t *grid) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        uint2 e = ed[i];
        if (e.x == e.y) continue;
        uint32_t v0 = e.x, v1 = e.y;
        if (!directed && v0 > v1) { v0 = e.y; v1 = e.x; }
        unsigned long long pos = atomicAdd(&cur[tile_slot(v0 >> 16, v1 >> 16, p, directed)], 1ull);
        grid[pos] = (v0 & 0xFFFFu) | (v1 << 16);
    }
}

// Tuples already on the device -> grid/degree on the device, start on host.
static int convert_device(gs_graph *g, const uint2 *d_ed, uint64_t m, std::vector<uint64_t> *h_start) {
    DevBuf<unsigned long long> cnt;
    GS_TRY(cnt.alloc(g->T + 1));
    GS_TRY(g->degree.alloc(g->n));
    GS_HIP(hipMemsetAsync(cnt.ptr, 0, (g->T + 1) * 8, g->stream));
    GS_HIP(hipMemsetAsync(g->degree.ptr, 0, g->n * 4, g->stream));
    GS_HIP(hipMemsetAsync(g->counters.ptr, 0, 8, g->stream));
    unsigned blocks = (unsigned)std::min<uint64_t>((m + 255) / 256 + 1, 1u << 20);
    uint32_t n32 = (uint32_t)std::min<uint64_t>(g->n, 0xFFFFFFFFull);
    if (m) k_conv_count<<<blocks, 256, 0, g->stream>>>(d_ed, m, n32, g->p, g->directed, cnt.ptr, g->degree.ptr, g->counters.ptr);
    GS_HIP(hipGetLastError());
    h_start->assign(g->T + 1, 0);
    GS_HIP(hipMemcpyAsync(h_start->data(), cnt.ptr, (g->T + 1) * 8, hipMemcpyDeviceToHost, g->stream));
    GS_HIP(hipMemcpyAsync(g->h_counters, g->counters.ptr, 8, hipMemcpyDeviceToHost, g->stream));
    GS_HIP(hipStreamSynchronize(g->stream));
    if (g->h_counters[0]) {
        set_error("%llu tuples have a vertex id >= N=%llu", g->h_counters[0], (unsigned long long)g->n);
        return GS_ERR_ARG;
    }
    uint64_t s = 0;
    for (uint64_t t = 0; t <= g->T; ++t) { uint64_t c = (*h_start)[t]; (*h_start)[t] = s; s += c; }
    g->E = s;
    GS_TRY(g->grid.alloc(g->E));
    GS_HIP(hipMemcpyAsync(cnt.ptr, h_start->data(), (g->T + 1) * 8, hipMemcpyHostToDevice, g->stream));
    if (m) k_conv_scatter<<<blocks, 256, 0, g->stream>>>(d_ed, m, g->p, g->directed, cnt.ptr, g->grid.ptr);
    GS_HIP(hipGetLastError());
    GS_HIP(hipStreamSynchronize(g->stream));
    return GS_OK;
}

}  // namespace gs

using namespace gs;

extern "C" {

size_t gs_last_error(char *buf, size_t len) {
    if (buf && len) {
        size_t c = std::min(len - 1, g_err.size());
        memcpy(buf, g_err.data(), c);
        buf[c] = 0;
    }
    return g_err.size();
}

const char *gs_kernel_name(int k) {
    static const char *names[GS_K_COUNT] = {"pr_scatter", "pr_finalize", "bfs_expand", "bfs_level_end",
                                            "wcc_hook", "wcc_compress", "wcc_check", "kcore_peel",
                                            "kcore_scan", "kron_gen", "convert", "h2d_copy", "pr_pull",
                                            "index_build", "pb_gather", "pb_accum", "pb_reduce",
                                            "bfs_topdown", "bfs_bottomup"};
    return (k >= 0 && k < GS_K_COUNT) ? names[k] : "?";
}

void gs_close(gs_graph *g) { delete g; }

int gs_open_arrays(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, const int32_t *degree,
                   gs_graph **out) {
    if (!start || !degree || !out) { set_error("null argument"); return GS_ERR_ARG; }
    gs_graph *g = nullptr;
    GS_TRY(graph_create(&g));
    std::unique_ptr<gs_graph, void (*)(gs_graph *)> guard(g, gs_close);
    GS_TRY(set_geometry(g, n, directed != 0));
    g->E = start[g->T];
    if (g->E && !grid) { set_error("null grid with %llu edges", (unsigned long long)g->E); return GS_ERR_ARG; }
    GS_TRY(g->grid.alloc(g->E));
    GS_TRY(g->degree.alloc(n));
    if (g->E) GS_HIP(hipMemcpy(g->grid.ptr, grid, g->E * 4, hipMemcpyHostToDevice));
    GS_HIP(hipMemcpy(g->degree.ptr, degree, n * 4, hipMemcpyHostToDevice));
    GS_TRY(open_common(g, start));
    *out = guard.release();
    return GS_OK;
}

static int open_files(const char *prefix, uint64_t n, int directed, uint64_t budget, bool external, uint64_t seg_bytes,
                      gs_graph **out);

int gs_open(const char *prefix, uint64_t n, int directed, uint64_t budget, gs_graph **out) {
    return open_files(prefix, n, directed, budget, false, 0, out);
}

int gs_open_external(const char *prefix, uint64_t n, int directed, uint64_t budget, uint64_t seg_bytes,
                     gs_graph **out) {
    return open_files(prefix, n, directed, budget, true, seg_bytes, out);
}

// Loads .start and .degree, and the grid: into HBM when it fits the budget;
// else into pinned host memory with a resident HBM prefix (semi-external), or,
// external, left in its file with only the prefix read (gs_stream.hip).
static int open_files(const char *prefix, uint64_t n, int directed, uint64_t budget, bool external, uint64_t seg_bytes,
                      gs_graph **out) {
    if (!prefix || !out) { set_error("null argument"); return GS_ERR_ARG; }
    gs_graph *g = nullptr;
    GS_TRY(graph_create(&g));
    std::unique_ptr<gs_graph, void (*)(gs_graph *)> guard(g, gs_close);
    GS_TRY(set_geometry(g, n, directed != 0));
    std::string pre(prefix);
    uint64_t sz;
    GS_TRY(file_size(pre + ".start", &sz));
    if (sz != (g->T + 1) * 8) {
        set_error("%s.start holds %llu bytes; N=%llu needs (T+1)*8 = %llu", prefix, (unsigned long long)sz,
                  (unsigned long long)n, (unsigned long long)((g->T + 1) * 8));
        return GS_ERR_FORMAT;
    }
    std::vector<uint64_t> st(g->T + 1);
    GS_TRY(read_file(pre + ".start", st.data(), sz));
    g->E = st[g->T];
    g->nchunk = (g->E + g->chunk - 1) / g->chunk;
    GS_TRY(file_size(pre + ".grid", &sz));
    if (sz != g->E * 4) {
        set_error("%s.grid holds %llu bytes; .start sentinel says %llu edges", prefix, (unsigned long long)sz,
                  (unsigned long long)g->E);
        return GS_ERR_FORMAT;
    }
    GS_TRY(file_size(pre + ".degree", &sz));
    if (sz != n * 4) {
        set_error("%s.degree holds %llu bytes, expected 4*N = %llu", prefix, (unsigned long long)sz,
                  (unsigned long long)(n * 4));
        return GS_ERR_FORMAT;
    }
    const bool streaming = budget && g->E * 4 > budget;
    GS_TRY(g->degree.alloc(n));
    if (streaming && external) {
        GS_TRY(disk_setup(g, pre + ".grid", budget, seg_bytes));
    } else if (streaming) {
        // semi-external: the grid goes to pinned host memory, a prefix to HBM
        uint32_t *h = nullptr;
        GS_HIP(hipHostMalloc((void **)&h, std::max<uint64_t>(g->E, 1) * 4, hipHostMallocDefault));
        int rs = read_file(pre + ".grid", h, g->E * 4);
        if (rs == GS_OK) rs = stream_setup(g, h, budget, 0);
        if (rs != GS_OK) {
            (void)hipHostFree(h);
            return rs;
        }
    } else {
        GS_TRY(g->grid.alloc(g->E));
    }
    // stage through pinned host memory in 256 MiB pieces
    {
        const uint64_t piece = 1ull << 28;
        char *pin = nullptr;
        GS_HIP(hipHostMalloc((void **)&pin, piece, hipHostMallocDefault));
        std::unique_ptr<char, void (*)(char *)> pg(pin, [](char *q) { (void)hipHostFree(q); });
        auto stage = [&](const std::string &path, void *dst, uint64_t bytes) -> int {
            FILE *f = fopen(path.c_str(), "rb");
            if (!f) { set_error("cannot open %s", path.c_str()); return GS_ERR_IO; }
            uint64_t done = 0;
            while (done < bytes) {
                size_t want = (size_t)std::min(piece, bytes - done);
                if (fread(pin, 1, want, f) != want) {
                    fclose(f);
                    set_error("short read on %s", path.c_str());
                    return GS_ERR_IO;
                }
                if (hipMemcpy((char *)dst + done, pin, want, hipMemcpyHostToDevice) != hipSuccess) {
                    fclose(f);
                    set_error("hipMemcpy failed while loading %s", path.c_str());
                    return GS_ERR_DEVICE;
                }
                done += want;
            }
            fclose(f);
            return GS_OK;
        };
        if (!streaming) GS_TRY(stage(pre + ".grid", g->grid.ptr, g->E * 4));
        GS_TRY(stage(pre + ".degree", g->degree.ptr, n * 4));
    }
    GS_TRY(open_common(g, st.data()));
    *out = guard.release();
    return GS_OK;
}

int gs_kron_edges(int scale, uint64_t s1, uint64_t s2, int64_t e0, int64_t e1, uint32_t *out) {
    if (!out || e1 < e0) { set_error("bad arguments"); return GS_ERR_ARG; }
    DevBuf<uint2> d;
    GS_TRY(d.alloc((size_t)(e1 - e0)));
    GS_TRY(kron_device(scale, s1, s2, e0, e1, d.ptr, 0));
    if (e1 > e0) GS_HIP(hipMemcpy(out, d.ptr, (size_t)(e1 - e0) * 8, hipMemcpyDeviceToHost));
    return GS_OK;
}

int gs_open_kron(int scale, int ef, uint64_t s1, uint64_t s2, int directed, gs_graph **out) {
    if (!out || ef < 1 || scale < 1 || scale > 32) { set_error("bad arguments"); return GS_ERR_ARG; }
    gs_graph *g = nullptr;
    GS_TRY(graph_create(&g));
    std::unique_ptr<gs_graph, void (*)(gs_graph *)> guard(g, gs_close);
    GS_TRY(set_geometry(g, 1ull << scale, directed != 0));
    uint64_t m = (uint64_t)ef << scale;
    std::vector<uint64_t> st;
    {
        DevBuf<uint2> ed;
        GS_TRY(ed.alloc(m));
        GS_TRY(kron_device(scale, s1, s2, 0, (int64_t)m, ed.ptr, g->stream));
        GS_TRY(convert_device(g, ed.ptr, m, &st));
    }
    GS_TRY(open_common(g, st.data()));
    *out = guard.release();
    return GS_OK;
}

int gs_convert(const uint32_t *edges, uint64_t m, uint64_t n, int directed, const char *prefix) {
    if ((!edges && m) || !prefix) { set_error("null argument"); return GS_ERR_ARG; }
    gs_graph *g = nullptr;
    GS_TRY(graph_create(&g));
    std::unique_ptr<gs_graph, void (*)(gs_graph *)> guard(g, gs_close);
    GS_TRY(set_geometry(g, n, directed != 0));
    std::vector<uint64_t> st;
    {
        DevBuf<uint2> ed;
        GS_TRY(ed.alloc(m));
        if (m) GS_HIP(hipMemcpy(ed.ptr, edges, m * 8, hipMemcpyHostToDevice));
        GS_TRY(convert_device(g, ed.ptr, m, &st));
    }
    std::vector<uint32_t> grid(g->E);
    std::vector<int32_t> deg(n);
    if (g->E) GS_HIP(hipMemcpy(grid.data(), g->grid.ptr, g->E * 4, hipMemcpyDeviceToHost));
    GS_HIP(hipMemcpy(deg.data(), g->degree.ptr, n * 4, hipMemcpyDeviceToHost));
    std::string pre(prefix);
    GS_TRY(write_file(pre + ".grid", grid.data(), g->E * 4));
    GS_TRY(write_file(pre + ".start", st.data(), (g->T + 1) * 8));
    GS_TRY(write_file(pre + ".degree", deg.data(), n * 4));
    return GS_OK;
}

int gs_get_info(gs_graph *g, gs_info *info) {
    if (!g || !info) { set_error("null argument"); return GS_ERR_ARG; }
    info->vert_count = g->n;
    info->edge_count = g->E;
    info->tile_slots = g->T;
    info->nonempty_tiles = g->nz;
    info->groups = g->p;
    info->tiles_per_side = g->p_s;
    info->directed = g->directed;
    info->resident = g->ss ? 0 : 1;
    info->resident_edges = g->ss ? std::min<uint64_t>(g->E, g->ss->res_chunks * g->chunk) : g->E;
    info->segment_edges = g->ss ? g->ss->seg_chunks * g->chunk : 0;
    info->device_bytes = g->grid.bytes() + g->tz_beg.bytes() + g->tz_rc.bytes() + g->chunk_tz.bytes() +
                         g->degree.bytes() + g->ws->device_bytes() + g->adj_off.bytes() + g->adj_nbr.bytes() + g->adj_first.bytes() +
                         g->new_id.bytes() + g->old_id.bytes() + g->deg_r.bytes() + g->chunkz.bytes() + gs::pb_bytes(g) +
                         (g->ss ? g->ss->seg[0].bytes() + g->ss->seg[1].bytes() + g->ss->seg[2].bytes() : 0);
    const bool built = g->adj_ready;
    info->index_entries = built ? 2 * g->E : 0;
    info->index_vertices = built ? g->nzn : 0;
    info->index_ms = built ? g->index_ms + g->pb_ms : 0.0;
    return GS_OK;
}

// .start rebuilt from the non-empty tile list (empty tiles repeat the offset).
static int rebuild_start(gs_graph *g, std::vector<uint64_t> *st) {
    std::vector<uint64_t> beg(g->nz + 1);
    std::vector<uint32_t> rc(g->nz);
    GS_HIP(hipMemcpy(beg.data(), g->tz_beg.ptr, (g->nz + 1) * 8, hipMemcpyDeviceToHost));
    if (g->nz) GS_HIP(hipMemcpy(rc.data(), g->tz_rc.ptr, g->nz * 4, hipMemcpyDeviceToHost));
    st->assign(g->T + 1, 0);
    uint64_t t = 0;
    for (uint64_t k = 0; k < g->nz; ++k) {
        uint64_t sk = tile_slot(rc[k] >> 16, rc[k] & 0xFFFF, g->p, g->directed);
        for (; t <= sk; ++t) (*st)[t] = beg[k];
    }
    for (; t <= g->T; ++t) (*st)[t] = g->E;
    return GS_OK;
}

int gs_get_arrays(gs_graph *g, uint64_t *start, uint32_t *grid, int32_t *degree) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (start) {
        std::vector<uint64_t> st;
        GS_TRY(rebuild_start(g, &st));
        memcpy(start, st.data(), (g->T + 1) * 8);
    }
    if (grid && g->E) {
        if (g->ss && g->ss->disk) GS_TRY(disk_read_all(g->ss, grid));
        else if (g->ss) memcpy(grid, g->ss->h_grid, g->E * 4);
        else GS_HIP(hipMemcpy(grid, g->grid.ptr, g->E * 4, hipMemcpyDeviceToHost));
    }
    if (degree) GS_HIP(hipMemcpy(degree, g->degree.ptr, g->n * 4, hipMemcpyDeviceToHost));
    return GS_OK;
}

int gs_get_host_grid(gs_graph *g, const uint32_t **grid, uint64_t *edges) {
    if (!g || !grid) { set_error("null argument"); return GS_ERR_ARG; }
    if (!g->ss || g->ss->disk || !g->ss->h_grid) {
        set_error("gs_get_host_grid: the grid has no pinned host copy (resident or external mode)");
        return GS_ERR_UNSUPPORTED;
    }
    *grid = g->ss->h_grid;
    if (edges) *edges = g->E;
    return GS_OK;
}

int gs_save(gs_graph *g, const char *prefix) {
    if (!g || !prefix) { set_error("null argument"); return GS_ERR_ARG; }
    std::vector<uint64_t> st;
    GS_TRY(rebuild_start(g, &st));
    const bool host_copy = g->ss && !g->ss->disk;
    std::vector<uint32_t> grid(host_copy ? 0 : g->E);
    std::vector<int32_t> deg(g->n);
    if (g->E && !g->ss) GS_HIP(hipMemcpy(grid.data(), g->grid.ptr, g->E * 4, hipMemcpyDeviceToHost));
    if (g->E && g->ss && g->ss->disk) GS_TRY(disk_read_all(g->ss, grid.data()));
    GS_HIP(hipMemcpy(deg.data(), g->degree.ptr, g->n * 4, hipMemcpyDeviceToHost));
    std::string pre(prefix);
    GS_TRY(write_file(pre + ".grid", host_copy ? g->ss->h_grid : grid.data(), g->E * 4));
    GS_TRY(write_file(pre + ".start", st.data(), (g->T + 1) * 8));
    GS_TRY(write_file(pre + ".degree", deg.data(), g->n * 4));
    return GS_OK;
}

int gs_get_stats(gs_graph *g, gs_stats *st) {
    if (!g || !st) { set_error("null argument"); return GS_ERR_ARG; }
    *st = g->ws->stats;
    for (int i = 0; i < GS_K_COUNT; ++i) {
        st->kern_ms[i] = g->ws->timer.ms[i];
        st->kern_calls[i] = g->ws->timer.calls[i];
    }
    return GS_OK;
}

int gs_get_level_stats(gs_graph *g, gs_level_stat *out, int max, int *n) {
    if (!g || (max > 0 && !out) || max < 0) { set_error("bad arguments"); return GS_ERR_ARG; }
    const auto &lv = g->ws->levels;
    const int cnt = (int)lv.size();
    for (int i = 0; i < cnt && i < max; ++i) out[i] = lv[i];
    if (n) *n = cnt;
    return GS_OK;
}

int gs_set_chunk(gs_graph *g, uint32_t edges) {
    if (!g || edges < 4 * SWEEP_WG || edges > (uint32_t)CHUNK || edges % (4 * SWEEP_WG)) {
        set_error("chunk must be a multiple of %d in [%d, %d]", 4 * SWEEP_WG, 4 * SWEEP_WG, CHUNK);
        return GS_ERR_ARG;
    }
    if (g->ss) { set_error("the chunk size of a streaming graph is fixed"); return GS_ERR_UNSUPPORTED; }
    if (edges == g->chunk) return GS_OK;
    std::vector<uint64_t> beg(g->nz + 1);
    GS_HIP(hipMemcpy(beg.data(), g->tz_beg.ptr, (g->nz + 1) * 8, hipMemcpyDeviceToHost));
    g->chunk = edges;
    return build_chunk_map(g, beg);
}

int gs_set_bfs_mode(gs_graph *g, int mode, int alpha, int beta) {
    if (!g || mode < 0 || mode > 1 || alpha < 0 || beta < 0) { set_error("gs_set_bfs_mode: bad arguments"); return GS_ERR_ARG; }
    if (mode == 1 && g->directed) { set_error("direction-optimizing BFS needs an undirected grid"); return GS_ERR_UNSUPPORTED; }
    g->bfs_mode = mode;
    if (alpha) g->do_alpha = alpha;  // 0 keeps the current (possibly tuned) value
    if (beta) g->do_beta = beta;
    return mode == 1 ? build_adjacency(g) : GS_OK;
}

int gs_set_pagerank_heavy(gs_graph *g, int on) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    g->heavy_on = on != 0;
    g->heavy_ready = false;
    return GS_OK;
}

int gs_set_pagerank_method(gs_graph *g, int method) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (method < 0 || method > 2) { set_error("PageRank method %d: expected 0, 1 or 2", method); return GS_ERR_ARG; }
    g->pr_method = method;
    g->heavy_ready = false;
    return GS_OK;
}

int gs_set_pagerank_blocking(gs_graph *g, uint64_t gather_entries, uint32_t gather_runs, uint64_t accum_entries,
                             uint32_t dense_entries, uint32_t compact_permille) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (gather_entries < 1 || gather_entries > (1ull << 20) || gather_runs < 1 || gather_runs > gs::PB_RUNS ||
        accum_entries < 1 || accum_entries > (1ull << 30) || compact_permille > 1000) {
        set_error("blocking limits out of range");
        return GS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lp(g->pb_mu);  // not while a blocked PageRank runs
    std::lock_guard<std::mutex> lk(g->build_mu);
    g->pb_unit1 = gather_entries;
    g->pb_runs = gather_runs;
    g->pb_unit2 = accum_entries;
    g->pb_dense = dense_entries;
    g->pb_compact = compact_permille;
    g->pb_ready = false;  // tables rebuilt at the next PageRank
    return GS_OK;
}

int gs_set_pagerank_gather(gs_graph *g, uint32_t units_per_grab) {
    if (!g || units_per_grab > 4096) { set_error("gs_set_pagerank_gather: 0..4096 units per grab"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lp(g->pb_mu);
    g->pb_grab = units_per_grab;
    return GS_OK;
}

int gs_debug_fail_alloc(gs_graph *g, int nth) {
    if (!g || nth < 0) { set_error("bad argument"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lk(g->build_mu);
    g->fail_alloc_at = nth;
    return GS_OK;
}

int gs_set_batch_bfs_method(gs_graph *g, int method) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (method < 0 || method > 2) { set_error("batched BFS method %d: expected 0, 1 or 2", method); return GS_ERR_ARG; }
    g->batch_method = method;
    return GS_OK;
}

int gs_set_kcore_method(gs_graph *g, int method) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (method < 0 || method > 1) { set_error("k-core method %d: expected 0 or 1", method); return GS_ERR_ARG; }
    g->kcore_method = method;
    return GS_OK;
}

int gs_set_wcc_method(gs_graph *g, int method) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (method < 0 || method > 1) { set_error("WCC method %d: expected 0 or 1", method); return GS_ERR_ARG; }
    g->wcc_method = method;
    return GS_OK;
}

int gs_set_sweep_order(gs_graph *g, int xcd_contiguous) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    g->xcd_order = xcd_contiguous ? 1 : 0;
    return GS_OK;
}

int gs_set_kernel_timing(gs_graph *g, int on) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    g->ws->timer.on = on != 0;
    return GS_OK;
}

}  // extern "C"
