/*
 * gs_oracle.c -- CPU restatement of the G-Store tile-grid hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine in
 * gstore_amd/ is compared against, and the CPU baseline that bench.py times.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it; the engine never links or calls it.
 *
 * Every function cites the reference (the-data-lab/gstore) file:line whose
 * behaviour it restates.  Nothing here is copied: the loops are written from
 * the described semantics, and the Kronecker generator is restated from the
 * Graph500 specification (MRG32k5 of L'Ecuyer-Blouin-Couture 1993, table 3,
 * m = 2^31-1, k = 5) with its skip-ahead matrices recomputed by exponentiation.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   or_kron_range      -- reference generator built from source (oracle/_ref)
 *                         and the edge listing in graph500-generator/README.md
 *   or_convert_*       -- invariants (multiset of (min,max) non-self-loop edges,
 *                         degree = endpoint count, start monotone, sentinel)
 *   or_pagerank compat -- closed form (only degree-1 vertices are non-zero)
 *   or_pagerank intd.  -- float64 scipy recurrence in tests (tolerance stated)
 *   or_pagerank64      -- the same float64 scipy recurrence (1e-12), and the
 *                         FP32 or_pagerank within FP32 rounding of it
 *   or_ppr             -- float64 scipy recurrence from two sources (1e-12)
 *   or_kron_convert_*  -- equal to or_convert(or_kron_range) over the whole list
 *   or_grid_checksum   -- invariant under in-tile permutation, changed by one
 *                         moved or altered edge
 *   or_*_mt            -- equal to the serial forms (tests), which are pinned
 *   or_bfs2            -- depths equal scipy breadth-first distances + 1
 *   or_traverse        -- coloured set = scipy reachability from the pivot over
 *                         the edges oriented v0 -> v1
 *   or_text_to_bin     -- hand-written texts with known pairs, and a round trip
 *                         of random pairs printed with several separators
 *   or_bfs             -- scipy.sparse.csgraph breadth-first distances
 *   or_wcc             -- scipy.sparse.csgraph.connected_components
 *   or_kcore           -- independent numpy peeling in tests + brute force
 *   or_pr_step64_sample -- step k -> k+1 of the scipy float64 recurrence at the
 *                         sampled vertices (1e-12); fails on the iterate of
 *                         another step and on a directed grid read undirected
 *   or_bfs_check       -- accepts scipy distances + 1 on every test graph;
 *                         rejects each planted mistake (one depth raised or
 *                         lowered, a vertex dropped, one of another component
 *                         marked reached, a second depth-1 vertex)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Tile addressing (gstore.h:63-141, HALF_GRID branch; full grid :176-198)   */
/* ------------------------------------------------------------------------ */

/* Upper-triangle row-major index of (i, j), i <= j, in an n x n triangle
 * (gstore.h:109-113). */
static uint64_t tri_index(uint64_t i, uint64_t j, uint64_t n)
{
    return ((i * (2 * n - i - 1)) >> 1) + j;
}

/* Number of tile slots for p_s tiles per side (gstore.h:121-125). */
uint64_t or_total_tiles(uint64_t p_s, int directed)
{
    return directed ? p_s * p_s : (p_s * (p_s + 1)) >> 1;
}

/* Physical groups per side: ceil(N / 2^24) (gstore.cpp:152-157). */
uint64_t or_groups(uint64_t n)
{
    return (n >> 24) + ((n & 0xFFFFFFull) != 0);
}

/* Index in .start of tile (r, c), where r, c are global tile coordinates
 * (tile = 2^16 vertices, group = 256 tiles).  Undirected: r <= c, groups in
 * upper-triangle order, a diagonal group holds its 32896 upper-triangle tiles,
 * an off-diagonal group 65536 tiles (gstore.h:128-141, matrix::get :365-369,
 * matrix_f::get :494-497).  Directed: groups p x p row-major, tiles row-major
 * (gstore.h:176-198). */
uint64_t or_tile_slot(uint64_t r, uint64_t c, uint64_t p, int directed)
{
    uint64_t bi = r >> 8, si = r & 255, bj = c >> 8, sj = c & 255;
    if (directed)
        return ((bi * p + bj) << 16) + si * 256 + sj;
    if (bi == bj) {
        uint64_t base = ((tri_index(bi, bi, p) - bi) << 16) + ((bi * 257) << 7);
        return base + tri_index(si, sj, 256);
    }
    uint64_t base = ((tri_index(bi, bj, p) - bi - 1) << 16) + (((bi + 1) * 257) << 7);
    return base + si * 256 + sj;
}

/* Inverse of or_tile_slot over a whole grid: fills row[t], col[t] for every
 * slot t in .start order.  Used by the tile walks below. */
void or_slot_coords(uint64_t p, int directed, uint32_t *row, uint32_t *col)
{
    uint64_t t = 0;
    for (uint64_t bi = 0; bi < p; ++bi)
        for (uint64_t bj = directed ? 0 : bi; bj < p; ++bj)
            for (uint64_t si = 0; si < 256; ++si)
                for (uint64_t sj = (!directed && bi == bj) ? si : 0; sj < 256; ++sj) {
                    row[t] = (uint32_t)(bi * 256 + si);
                    col[t] = (uint32_t)(bj * 256 + sj);
                    ++t;
                }
}

/* ------------------------------------------------------------------------ */
/* Graph500 Kronecker generator (graph500-generator/graph_generator.c:41-198, */
/* splittable_mrg.c:116-285, utils.c:51-57)                                 */
/* ------------------------------------------------------------------------ */

#define MRG_P 0x7FFFFFFFull
#define MRG_X 107374182ull
#define MRG_Y 104480ull

static uint64_t mrg_mod(uint64_t v)
{
    v = (v & MRG_P) + (v >> 31);
    v = (v & MRG_P) + (v >> 31);
    return v >= MRG_P ? v - MRG_P : v;
}

typedef struct { uint64_t m[5][5]; } mat5;

static void mat5_mul(const mat5 *a, const mat5 *b, mat5 *out)
{
    mat5 r;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            uint64_t s = 0;
            for (int k = 0; k < 5; ++k) s = mrg_mod(s + mrg_mod(a->m[i][k] * b->m[k][j]));
            r.m[i][j] = s;
        }
    *out = r;
}

static void mat5_identity(mat5 *a)
{
    memset(a, 0, sizeof *a);
    for (int i = 0; i < 5; ++i) a->m[i][i] = 1;
}

/* One MRG step on the column state (z1..z5): z1' = x*z1 + y*z5, shift down. */
static void mat5_step(mat5 *a)
{
    memset(a, 0, sizeof *a);
    a->m[0][0] = MRG_X;
    a->m[0][4] = MRG_Y;
    for (int i = 1; i < 5; ++i) a->m[i][i - 1] = 1;
}

static void mat5_pow2k(const mat5 *a, int k, mat5 *out) /* a^(2^k) */
{
    mat5 r = *a;
    for (int i = 0; i < k; ++i) mat5_mul(&r, &r, &r);
    *out = r;
}

static void mat5_pow(const mat5 *a, uint64_t e, mat5 *out)
{
    mat5 r, b = *a;
    mat5_identity(&r);
    while (e) {
        if (e & 1) mat5_mul(&r, &b, &r);
        mat5_mul(&b, &b, &b);
        e >>= 1;
    }
    *out = r;
}

static void mat5_apply(const mat5 *a, uint64_t z[5])
{
    uint64_t o[5];
    for (int i = 0; i < 5; ++i) {
        uint64_t s = 0;
        for (int k = 0; k < 5; ++k) s = mrg_mod(s + mrg_mod(a->m[i][k] * z[k]));
        o[i] = s;
    }
    memcpy(z, o, sizeof o);
}

static uint32_t mrg_next(uint64_t z[5])
{
    uint64_t nz = mrg_mod(mrg_mod(MRG_X * z[0]) + mrg_mod(MRG_Y * z[4]));
    z[4] = z[3]; z[3] = z[2]; z[2] = z[1]; z[1] = z[0]; z[0] = nz;
    return (uint32_t)nz;
}

static uint64_t bitrev64(uint64_t x)
{
    uint64_t r = 0;
    for (int i = 0; i < 64; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* Vertex scramble (graph_generator.c:120-130). */
static uint64_t kron_scramble(uint64_t v, int lg, uint64_t k0, uint64_t k1)
{
    v += k0 + k1;
    v *= (k0 | 0x4519840211493211ull);
    v = bitrev64(v) >> (64 - lg);
    v *= (k1 | 0x3050852102C843A5ull);
    v = bitrev64(v) >> (64 - lg);
    return v;
}

/* Generate tuples [e0, e1) of the scale-lg Kronecker graph for user seeds
 * (s1, s2) into out[2*(e-e0)+{0,1}] (u32 pairs, the gedge_t layout of
 * gstore.h:254-279).  Tuple e starts from the seed state advanced by e*2^64
 * steps (graph_generator.c:189-197); each of the lg levels draws one
 * quadrant with A=.57, B=C=.19, D=.05 by rejection of values below
 * 2^32-1 mod 10000 (graph_generator.c:41-68), clip-and-flip on the diagonal
 * (:141-148), then both endpoints are scrambled (:155-156). */
void or_kron_range(int lg, uint64_t s1, uint64_t s2, int64_t e0, int64_t e1, uint32_t *out)
{
    uint64_t seed[5] = {
        (s1 & 0x3FFFFFFF) + 1, ((s1 >> 30) & 0x3FFFFFFF) + 1,
        (s2 & 0x3FFFFFFF) + 1, ((s2 >> 30) & 0x3FFFFFFF) + 1,
        ((s2 >> 60) << 4) + (s1 >> 60) + 1 };
    mat5 a1, a64, a128, t;
    mat5_step(&a1);
    mat5_pow2k(&a1, 64, &a64);
    mat5_pow2k(&a64, 64, &a128);
    /* scramble keys: state advanced by 50*2^128 + 7*2^64 steps */
    uint64_t z[5];
    memcpy(z, seed, sizeof z);
    mat5_pow(&a64, 7, &t);   mat5_apply(&t, z);
    mat5_pow(&a128, 50, &t); mat5_apply(&t, z);
    uint64_t k0 = mrg_next(z); k0 *= 0xFFFFFFFFull; k0 += mrg_next(z);
    uint64_t k1 = mrg_next(z); k1 *= 0xFFFFFFFFull; k1 += mrg_next(z);

    /* byte tables: bt[b][v] = a64^(v * 256^b) */
    mat5 *bt = (mat5 *)malloc(sizeof(mat5) * 8 * 256);
    mat5 base = a64;
    for (int b = 0; b < 8; ++b) {
        mat5_identity(&bt[b * 256]);
        for (int v = 1; v < 256; ++v) mat5_mul(&bt[b * 256 + v - 1], &base, &bt[b * 256 + v]);
        mat5_pow(&base, 256, &base);
    }
    const uint32_t limit = 0xFFFFFFFFu % 10000u;
    #pragma omp parallel for schedule(static)
    for (int64_t e = e0; e < e1; ++e) {
        uint64_t st[5];
        memcpy(st, seed, sizeof st);
        uint64_t ee = (uint64_t)e;
        for (int b = 0; ee; ++b, ee >>= 8)
            if (ee & 255) mat5_apply(&bt[b * 256 + (ee & 255)], st);
        uint64_t src = 0, dst = 0, half = 1ull << lg;
        for (int l = 0; l < lg; ++l) {
            uint32_t v = mrg_next(st);
            while (v < limit) v = mrg_next(st);
            v %= 10000u;
            int q;
            if (v < 1900) q = 1;
            else if (v < 3800) q = 2;
            else if (v < 9500) q = 0;
            else q = 3;
            int so = q >> 1, to = q & 1;
            if (src == dst && so > to) { so = 0; to = 1; }
            half >>= 1;
            src += half * so;
            dst += half * to;
        }
        out[2 * (e - e0)]     = (uint32_t)kron_scramble(src, lg, k0, k1);
        out[2 * (e - e0) + 1] = (uint32_t)kron_scramble(dst, lg, k0, k1);
    }
    free(bt);
}

/* ------------------------------------------------------------------------ */
/* Converter -c 1 (gstore.cpp:1245-1458): drop self-loops, store (min,max), */
/* per-tile count -> prefix -> scatter, degree += 1 on both endpoints.       */
/* The reference's in-tile order depends on thread timing; this one keeps    */
/* input order inside a tile.                                                */
/* ------------------------------------------------------------------------ */

/* Pass 1: fills start[0..T] (prefix offsets, start[T] = E) and degree[0..N).
 * Returns E, or -1 if a vertex id is >= n. */
int64_t or_convert_count(const uint32_t *edges, uint64_t m, uint64_t n, int directed,
                         uint64_t *start, int32_t *degree)
{
    uint64_t p = or_groups(n), T = or_total_tiles(256 * p, directed);
    memset(start, 0, sizeof(uint64_t) * (T + 1));
    memset(degree, 0, sizeof(int32_t) * n);
    for (uint64_t k = 0; k < m; ++k) {
        uint64_t a = edges[2 * k], b = edges[2 * k + 1];
        if (a == b) continue;
        if (a >= n || b >= n) return -1;
        uint64_t v0 = a, v1 = b;
        if (!directed && v0 > v1) { v0 = b; v1 = a; }
        start[or_tile_slot(v0 >> 16, v1 >> 16, p, directed)]++;
        degree[a]++;
        if (!directed) degree[b]++;
    }
    uint64_t s = 0;
    for (uint64_t t = 0; t <= T; ++t) { uint64_t c = start[t]; start[t] = s; s += c; }
    return (int64_t)s;
}

/* Pass 2: scatter into grid[0..E) as u32 words (v0 & 0xFFFF) | (v1 << 16)
 * (edge_t {u16 v0, v1}, gstore.h:311-334). */
void or_convert_scatter(const uint32_t *edges, uint64_t m, uint64_t n, int directed,
                        const uint64_t *start, uint32_t *grid)
{
    uint64_t p = or_groups(n), T = or_total_tiles(256 * p, directed);
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * T);
    memcpy(cur, start, sizeof(uint64_t) * T);
    for (uint64_t k = 0; k < m; ++k) {
        uint64_t a = edges[2 * k], b = edges[2 * k + 1];
        if (a == b) continue;
        uint64_t v0 = a, v1 = b;
        if (!directed && v0 > v1) { v0 = b; v1 = a; }
        uint64_t t = or_tile_slot(v0 >> 16, v1 >> 16, p, directed);
        grid[cur[t]++] = (uint32_t)(v0 & 0xFFFF) | ((uint32_t)(v1 & 0xFFFF) << 16);
    }
    free(cur);
}

/* Generate-and-convert in chunks, for scales whose tuple list does not fit in
 * host memory beside the grid (M = ef * 2^lg tuples; at scale 28 the list is
 * 34 GB and the grid 17 GB).  The tuples are generated twice, `chunk` at a
 * time: or_kron_convert_count fills start[0..T] and degree[0..N) exactly as
 * or_convert_count does for the whole list, or_kron_convert_scatter fills the
 * grid as or_convert_scatter does (input order within a tile).  The two calls
 * together give the same arrays as or_convert(or_kron_range(0, M)). */
int64_t or_kron_convert_count(int lg, int ef, uint64_t s1, uint64_t s2, int directed, uint64_t chunk,
                              uint64_t *start, int32_t *degree)
{
    uint64_t n = 1ull << lg, m = (uint64_t)ef << lg;
    uint64_t p = or_groups(n), T = or_total_tiles(256 * p, directed);
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * 2 * chunk);
    if (!buf) return -1;
    memset(start, 0, sizeof(uint64_t) * (T + 1));
    memset(degree, 0, sizeof(int32_t) * n);
    for (uint64_t e0 = 0; e0 < m; e0 += chunk) {
        uint64_t c = m - e0 < chunk ? m - e0 : chunk;
        or_kron_range(lg, s1, s2, (int64_t)e0, (int64_t)(e0 + c), buf);
        for (uint64_t k = 0; k < c; ++k) {
            uint64_t a = buf[2 * k], b = buf[2 * k + 1];
            if (a == b) continue;
            uint64_t v0 = a, v1 = b;
            if (!directed && v0 > v1) { v0 = b; v1 = a; }
            start[or_tile_slot(v0 >> 16, v1 >> 16, p, directed)]++;
            degree[a]++;
            if (!directed) degree[b]++;
        }
    }
    free(buf);
    uint64_t s = 0;
    for (uint64_t t = 0; t <= T; ++t) { uint64_t c = start[t]; start[t] = s; s += c; }
    return (int64_t)s;
}

int or_kron_convert_scatter(int lg, int ef, uint64_t s1, uint64_t s2, int directed, uint64_t chunk,
                            const uint64_t *start, uint32_t *grid)
{
    uint64_t n = 1ull << lg, m = (uint64_t)ef << lg;
    uint64_t p = or_groups(n), T = or_total_tiles(256 * p, directed);
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * 2 * chunk);
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * T);
    if (!buf || !cur) { free(buf); free(cur); return -1; }
    memcpy(cur, start, sizeof(uint64_t) * T);
    for (uint64_t e0 = 0; e0 < m; e0 += chunk) {
        uint64_t c = m - e0 < chunk ? m - e0 : chunk;
        or_kron_range(lg, s1, s2, (int64_t)e0, (int64_t)(e0 + c), buf);
        for (uint64_t k = 0; k < c; ++k) {
            uint64_t a = buf[2 * k], b = buf[2 * k + 1];
            if (a == b) continue;
            uint64_t v0 = a, v1 = b;
            if (!directed && v0 > v1) { v0 = b; v1 = a; }
            uint64_t t = or_tile_slot(v0 >> 16, v1 >> 16, p, directed);
            grid[cur[t]++] = (uint32_t)(v0 & 0xFFFF) | ((uint32_t)(v1 & 0xFFFF) << 16);
        }
    }
    free(buf);
    free(cur);
    return 0;
}

/* Order-independent checksum of a grid: the sum (mod 2^64) over stored edges
 * of a 64-bit mix of (tile slot << 32 | edge word).  Equal for two grids that
 * hold the same multiset of edges per tile, whatever their in-tile order
 * (which the reference's converter leaves to thread timing,
 * gstore.cpp:1442-1446). */
static inline uint64_t mix64(uint64_t x)
{
    x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27; x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

uint64_t or_grid_checksum(uint64_t T, const uint64_t *start, const uint32_t *grid)
{
    uint64_t s = 0;
    #pragma omp parallel for reduction(+:s) schedule(dynamic, 4096)
    for (uint64_t t = 0; t < T; ++t)
        for (uint64_t k = start[t]; k < start[t + 1]; ++k)
            s += mix64((t << 32) | grid[k]);
    return s;
}

/* ------------------------------------------------------------------------ */
/* Common tile walk: every tile slot in .start order with its coordinates.   */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t n, p, T;
    int directed;
    const uint64_t *start;
    const uint32_t *grid;
    uint32_t *row, *col;
} walk_t;

static int walk_init(walk_t *w, uint64_t n, int directed, const uint64_t *start, const uint32_t *grid)
{
    w->n = n;
    w->directed = directed;
    w->p = or_groups(n);
    w->T = or_total_tiles(256 * w->p, directed);
    w->start = start;
    w->grid = grid;
    w->row = (uint32_t *)malloc(sizeof(uint32_t) * w->T);
    w->col = (uint32_t *)malloc(sizeof(uint32_t) * w->T);
    if (!w->row || !w->col) return -1;
    or_slot_coords(w->p, directed, w->row, w->col);
    return 0;
}

static void walk_free(walk_t *w)
{
    free(w->row);
    free(w->col);
}

static inline float atomic_add_f32(float *a, float v)
{
    /* CAS loop, like qthread_dincr (pr.cpp:60-80) */
    union { float f; uint32_t u; } o, nw;
    o.u = __atomic_load_n((uint32_t *)a, __ATOMIC_RELAXED);
    do { nw.f = o.f + v; } while (!__atomic_compare_exchange_n((uint32_t *)a, &o.u, nw.u, 1,
                                                               __ATOMIC_RELAXED, __ATOMIC_RELAXED));
    return o.f;
}

/* ------------------------------------------------------------------------ */
/* PageRank (pr.cpp:82-154, 229-303; driver gstore.cpp:2572-2629)           */
/* ------------------------------------------------------------------------ */

/* mode 0 (compat): factor = (int32)(1.0/deg) as the reference stores it in
 * its int32 degree array (pr.cpp:112-114) -> 1 for degree 1, else 0.
 * mode 1 (intended): factor = (float)(1.0/deg).
 * prior0 = (float)(1.0/N) (pr.cpp:110); each iteration scatters
 * acc[v1] += prior[v0], acc[v0] += prior[v1] for every stored edge
 * (pr.cpp:141-152), then prior = (float)((0.15 + 0.85*(double)acc) * factor)
 * (pr.cpp:286-290) with acc cleared.  Output is prior after `iters`
 * iterations.  nthreads == 1 gives a fixed summation order (tile order,
 * edge order); nthreads > 1 sweeps tile rows in parallel with CAS adds like
 * the reference.  Build with -ffp-contract=off so 0.15 + 0.85*x is not fused.
 * directed != 0: full grid, acc[v1] += prior[v0] only (pr.cpp without
 * HALF_GRID). */
int or_pagerank(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid,
                const int32_t *degree, int iters, int mode, int nthreads, float *out)
{
    walk_t w;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    float *prior = out, *acc = (float *)calloc(n, sizeof(float));
    float *fac = (float *)malloc(sizeof(float) * n);
    double inv_n = 1.0 / (double)n;
    for (uint64_t v = 0; v < n; ++v) {
        prior[v] = (float)inv_n;
        int32_t d = degree[v];
        if (d == 0) fac[v] = 0.0f;
        else if (mode == 0) fac[v] = (float)(int32_t)(1.0 / (double)d);
        else fac[v] = (float)(1.0 / (double)d);
    }
    for (int it = 0; it < iters; ++it) {
        if (nthreads <= 1) {
            for (uint64_t t = 0; t < w.T; ++t) {
                uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
                for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                    uint64_t v0 = o0 + (grid[k] & 0xFFFF), v1 = o1 + (grid[k] >> 16);
                    acc[v1] += prior[v0];
                    if (!directed) acc[v0] += prior[v1];
                }
            }
        } else {
            #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
            for (uint64_t t = 0; t < w.T; ++t) {
                uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
                for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                    uint64_t v0 = o0 + (grid[k] & 0xFFFF), v1 = o1 + (grid[k] >> 16);
                    atomic_add_f32(&acc[v1], prior[v0]);
                    if (!directed) atomic_add_f32(&acc[v0], prior[v1]);
                }
            }
        }
        #pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
        for (uint64_t v = 0; v < n; ++v) {
            double r = 0.15 + 0.85 * (double)acc[v];
            prior[v] = (float)(r * (double)fac[v]);
            acc[v] = 0.0f;
        }
    }
    free(acc);
    free(fac);
    walk_free(&w);
    return 0;
}

static inline double atomic_add_f64(double *a, double v)
{
    union { double f; uint64_t u; } o, nw;
    o.u = __atomic_load_n((uint64_t *)a, __ATOMIC_RELAXED);
    do { nw.f = o.f + v; } while (!__atomic_compare_exchange_n((uint64_t *)a, &o.u, nw.u, 1,
                                                               __ATOMIC_RELAXED, __ATOMIC_RELAXED));
    return o.f;
}

/* One FP64 scatter over every stored edge (pr.cpp:141-152): acc[v1] +=
 * prior[v0], and acc[v0] += prior[v1] on undirected grids.  nthreads > 1
 * sweeps tiles in parallel with CAS adds (FP64 summation order then varies,
 * far below the FP32 tolerances this oracle is used for). */
static void scatter64(const walk_t *w, const double *prior, double *acc, int nthreads)
{
    const uint32_t *grid = w->grid;
    int directed = w->directed;
    if (nthreads <= 1) {
        for (uint64_t t = 0; t < w->T; ++t) {
            uint64_t o0 = (uint64_t)w->row[t] << 16, o1 = (uint64_t)w->col[t] << 16;
            for (uint64_t k = w->start[t]; k < w->start[t + 1]; ++k) {
                uint64_t v0 = o0 + (grid[k] & 0xFFFF), v1 = o1 + (grid[k] >> 16);
                acc[v1] += prior[v0];
                if (!directed) acc[v0] += prior[v1];
            }
        }
        return;
    }
    #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (uint64_t t = 0; t < w->T; ++t) {
        uint64_t o0 = (uint64_t)w->row[t] << 16, o1 = (uint64_t)w->col[t] << 16;
        for (uint64_t k = w->start[t]; k < w->start[t + 1]; ++k) {
            uint64_t v0 = o0 + (grid[k] & 0xFFFF), v1 = o1 + (grid[k] >> 16);
            atomic_add_f64(&acc[v1], prior[v0]);
            if (!directed) atomic_add_f64(&acc[v0], prior[v1]);
        }
    }
}

/* PageRank of the intended form in FP64 (the float degree factor of
 * GS_PR_INTENDED taken exactly): prior0 = 1/N, acc = scatter of prior,
 * prior = (0.15 + 0.85 acc) / deg (0 for deg 0), as pr.cpp:82-117, 286-290
 * with the int32 truncation of pr.cpp:112-114 left out.  Output is prior after
 * `iters` iterations (the post-multiplied form of pr.cpp, Q2).  This is the
 * FP64 reference that FP32 PageRank is held to. */
int or_pagerank64(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid,
                  const int32_t *degree, int iters, int nthreads, double *out)
{
    walk_t w;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    double *prior = out, *acc = (double *)calloc(n, sizeof(double));
    if (!acc) { walk_free(&w); return -1; }
    for (uint64_t v = 0; v < n; ++v) prior[v] = 1.0 / (double)n;
    for (int it = 0; it < iters; ++it) {
        scatter64(&w, prior, acc, nthreads);
        #pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
        for (uint64_t v = 0; v < n; ++v) {
            double r = 0.15 + 0.85 * acc[v];
            prior[v] = degree[v] ? r / (double)degree[v] : 0.0;
            acc[v] = 0.0;
        }
    }
    free(acc);
    walk_free(&w);
    return 0;
}

/* Personalised PageRank in FP64 (no reference counterpart; uses the tile
 * scatter of pr.cpp:141-152): r_0 = e_src, prior = r / deg,
 * r_{t+1}[v] = 0.15 [v == src] + 0.85 * sum of prior over in-edges. */
int or_ppr(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, const int32_t *degree,
           uint64_t src, int iters, int nthreads, double *out)
{
    walk_t w;
    if (src >= n) return -1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    double *prior = (double *)calloc(n, sizeof(double)), *acc = (double *)calloc(n, sizeof(double));
    if (!prior || !acc) { free(prior); free(acc); walk_free(&w); return -1; }
    prior[src] = degree[src] ? 1.0 / degree[src] : 0.0;
    for (int it = 0; it < iters; ++it) {
        scatter64(&w, prior, acc, nthreads);
        #pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
        for (uint64_t v = 0; v < n; ++v) {
            double r = (v == src ? 0.15 : 0.0) + 0.85 * acc[v];
            out[v] = r;
            prior[v] = degree[v] ? r / degree[v] : 0.0;
            acc[v] = 0.0;
        }
    }
    if (iters < 1) for (uint64_t v = 0; v < n; ++v) out[v] = v == src;
    free(prior);
    free(acc);
    walk_free(&w);
    return 0;
}

/* One FP64 PageRank step (pr.cpp:141-152 scatter, pr.cpp:286-290 update, in
 * the intended form of or_pagerank64) evaluated only at the sampled vertices
 * idx[0 .. ns) (sorted, distinct): out[i] = (0.15 + 0.85 * sum of prior[u]
 * over the stored edges (u, idx[i]) -- both directions on undirected grids)
 * / deg(idx[i]), 0 for degree 0.  prior is any PageRank iterate in output
 * form, so a FP32 engine's iterate k+1 can be held against this step of its
 * own iterate k at sizes where the whole FP64 run does not fit the host.
 * One pass over the grid; per-thread sums over the sample, no atomics. */
static int64_t sample_pos(const uint32_t *idx, uint64_t ns, uint64_t v)
{
    uint64_t lo = 0, hi = ns;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (idx[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo < ns && idx[lo] == v ? (int64_t)lo : -1;
}

int or_pr_step64_sample(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid,
                        const int32_t *degree, const double *prior, const uint32_t *idx, uint64_t ns,
                        int nthreads, double *out)
{
    walk_t w;
    if (nthreads < 1) nthreads = 1;
    for (uint64_t i = 0; i < ns; ++i)
        if (idx[i] >= n || (i && idx[i] <= idx[i - 1])) return -1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    uint8_t *mark = (uint8_t *)calloc((n + 7) / 8, 1);
    double *part = (double *)calloc((size_t)nthreads * (ns ? ns : 1), sizeof(double));
    if (!mark || !part) { free(mark); free(part); walk_free(&w); return -1; }
    for (uint64_t i = 0; i < ns; ++i) mark[idx[i] >> 3] |= (uint8_t)(1u << (idx[i] & 7));
    #pragma omp parallel num_threads(nthreads)
    {
        double *acc = part + (size_t)omp_get_thread_num() * ns;
        #pragma omp for schedule(dynamic, 64)
        for (uint64_t t = 0; t < w.T; ++t) {
            uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
            for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                uint64_t v0 = o0 + (grid[k] & 0xFFFF), v1 = o1 + (grid[k] >> 16);
                if ((mark[v1 >> 3] >> (v1 & 7)) & 1) acc[sample_pos(idx, ns, v1)] += prior[v0];
                if (!directed && ((mark[v0 >> 3] >> (v0 & 7)) & 1)) acc[sample_pos(idx, ns, v0)] += prior[v1];
            }
        }
    }
    for (uint64_t i = 0; i < ns; ++i) {
        double a = 0.0;
        for (int th = 0; th < nthreads; ++th) a += part[(size_t)th * ns + i];
        int32_t d = degree[idx[i]];
        out[i] = d ? (0.15 + 0.85 * a) / (double)d : 0.0;
    }
    free(mark);
    free(part);
    walk_free(&w);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* BFS depth check (the conditions that fix BFS distances)                   */
/* ------------------------------------------------------------------------ */

/* Returns the number of violations of the conditions that make depth[] the
 * BFS depth array from root in the u8 encoding of bfs.cpp (0 unreached, root
 * 1): depth[root] == 1; over every stored edge, undirected grids: both ends
 * reached or neither, and reached ends differ by at most 1; directed grids
 * (v0 -> v1): v0 reached implies v1 reached with depth[v1] <= depth[v0] + 1;
 * and every reached vertex but the root has a reached neighbour (an
 * in-neighbour on directed grids) one level closer.  Together these hold
 * only for the true distances + 1: the parent chain bounds a depth from
 * below, the edge rule bounds it from above along a shortest path, and the
 * reached set is closed under the edges.  One pass over the grid plus one
 * over the vertices, so it serves sizes where the oracle's BFS is too slow. */
int64_t or_bfs_check(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint64_t root,
                     const uint8_t *depth, int nthreads)
{
    walk_t w;
    if (root >= n) return -1;
    if (nthreads < 1) nthreads = 1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    uint8_t *has_parent = (uint8_t *)calloc(n, 1);
    if (!has_parent) { walk_free(&w); return -1; }
    int64_t bad = depth[root] == 1 ? 0 : 1;
    #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(+ : bad)
    for (uint64_t t = 0; t < w.T; ++t) {
        uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
        for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
            uint64_t v0 = o0 + (grid[k] & 0xFFFF), v1 = o1 + (grid[k] >> 16);
            int a = depth[v0], b = depth[v1];
            if (directed) {
                if (a && (!b || b > a + 1)) ++bad;
                if (a && b == a + 1) __atomic_store_n(&has_parent[v1], 1, __ATOMIC_RELAXED);
            } else {
                if ((a == 0) != (b == 0) || (a && (a > b + 1 || b > a + 1))) ++bad;
                if (a && b && b == a + 1) __atomic_store_n(&has_parent[v1], 1, __ATOMIC_RELAXED);
                if (a && b && a == b + 1) __atomic_store_n(&has_parent[v0], 1, __ATOMIC_RELAXED);
            }
        }
    }
    #pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : bad)
    for (uint64_t v = 0; v < n; ++v)
        if (v != root && depth[v] && !has_parent[v]) ++bad;
    free(has_parent);
    walk_free(&w);
    return bad;
}

/* ------------------------------------------------------------------------ */
/* BFS (bfs.cpp:56-224; resident driver gstore.cpp:2631-2713)                */
/* ------------------------------------------------------------------------ */

/* depth is u8, 0 = unreached, root = 1 (bfs.cpp:54, 70-71).  Level L visits
 * every stored edge once: if d[v0] == L and d[v1] == 0 then d[v1] = L+1,
 * else if d[v1] == L and d[v0] == 0 then d[v0] = L+1 (bfs.cpp:158-177).  The
 * run stops at the first level that reaches nothing (bfs.cpp:196-214).  The
 * reference skips tiles whose row and column hold no frontier vertex
 * (gstore.cpp:2687-2698); such tiles cannot produce an update, so visiting
 * all tiles gives the same depths.  level_counts[L-1] receives the exact
 * number of vertices reached at level L (the reference's printed count can
 * over-count under races).  Returns the number of levels run, or -2 when a
 * level reaches vertices at depth 256, which the u8 depth of the reference
 * cannot hold.  directed != 0:
 * full grid, v0 -> v1 only (bfs.cpp without HALF_GRID). */
int or_bfs(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint64_t root,
           uint8_t *depth, uint64_t *level_counts, int max_levels)
{
    walk_t w;
    if (root >= n) return -1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    memset(depth, 0, n);
    depth[root] = 1;
    int level = 1, nlev = 0;
    for (;;) {
        uint64_t cnt = 0;
        for (uint64_t t = 0; t < w.T; ++t) {
            uint8_t *d0 = depth + ((uint64_t)w.row[t] << 16);
            uint8_t *d1 = depth + ((uint64_t)w.col[t] << 16);
            for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                uint32_t a = grid[k] & 0xFFFF, b = grid[k] >> 16;
                if (d0[a] == level && d1[b] == 0) { d1[b] = (uint8_t)(level + 1); ++cnt; }
                else if (!directed && d1[b] == level && d0[a] == 0) { d0[a] = (uint8_t)(level + 1); ++cnt; }
            }
        }
        if (nlev < max_levels && level_counts) level_counts[nlev] = cnt;
        ++nlev;
        if (cnt == 0) break;
        if (level + 1 > 255) { walk_free(&w); return -2; }  /* reached vertices got depth 256 */
        ++level;
    }
    walk_free(&w);
    return nlev;
}

/* ------------------------------------------------------------------------ */
/* WCC (wcc.cpp:57-395; resident driver gstore.cpp:2502-2570)                */
/* ------------------------------------------------------------------------ */

static uint32_t uf_find(uint32_t *par, uint32_t x)
{
    while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
    return x;
}

/* The reference labels components with atomic-counter group ids
 * (wcc.cpp:204) and prints the number of distinct ids, which counts only
 * components with at least one stored edge (wcc.cpp:324-337).  Its numbering
 * depends on thread timing, so the canonical form here is: label[v] = the
 * minimum vertex id of v's component for every vertex with a stored edge,
 * 0xFFFFFFFF (the reference's invalid_cid, wcc.cpp:54) otherwise.
 * *ncomp = number of components with at least one stored edge.  A directed
 * grid gives its weakly connected components. */
int or_wcc(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint32_t *label,
           uint64_t *ncomp)
{
    walk_t w;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint8_t *touched = (uint8_t *)calloc(n, 1);
    for (uint64_t v = 0; v < n; ++v) par[v] = (uint32_t)v;
    for (uint64_t t = 0; t < w.T; ++t) {
        uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
        for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
            uint32_t a = (uint32_t)(o0 + (grid[k] & 0xFFFF)), b = (uint32_t)(o1 + (grid[k] >> 16));
            touched[a] = touched[b] = 1;
            uint32_t ra = uf_find(par, a), rb = uf_find(par, b);
            if (ra < rb) par[rb] = ra;
            else if (rb < ra) par[ra] = rb;
        }
    }
    uint64_t c = 0;
    for (uint64_t v = 0; v < n; ++v) {
        if (!touched[v]) { label[v] = 0xFFFFFFFFu; continue; }
        label[v] = uf_find(par, (uint32_t)v);
        c += (label[v] == v);
    }
    *ncomp = c;
    free(par);
    free(touched);
    walk_free(&w);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* k-core (kcore.cpp:55-217; drivers gstore.cpp:2346-2500, 2716-2804)        */
/* ------------------------------------------------------------------------ */

/* deg starts as the .degree array (multigraph degree without self-loops).
 * Each round tests m = (int16)deg[v] (sdegree_t, kcore.cpp:189) and removes
 * v when 0 < m < k by setting deg[v] = 0 and marking it active
 * (kcore.cpp:190-195).  The first round runs before any edge pass
 * (gstore.cpp:2355-2357).  An edge pass then applies, for every stored edge,
 * if active(v0) --deg[v1]; if active(v1) --deg[v0] (kcore.cpp:69-103; removed
 * vertices keep being decremented).  Tiles with no active endpoint are
 * skipped by the reference (gstore.cpp:2768-2781), which changes nothing.
 * Stops at the first round that removes nothing.  *count_ne0 = number of
 * vertices with deg != 0, the reference's printed ">= k vertices count"
 * (gstore.cpp:2794-2800); *count_pos = number with deg > 0 (the true core).
 * Returns the number of removal rounds (including the first), or -1. */
int or_kcore(uint64_t n, const uint64_t *start, const uint32_t *grid, const int32_t *degree,
             int k, int32_t *deg, uint64_t *count_ne0, uint64_t *count_pos)
{
    walk_t w;
    if (walk_init(&w, n, 0, start, grid)) return -1;
    uint8_t *act = (uint8_t *)calloc(n, 1);
    memcpy(deg, degree, sizeof(int32_t) * n);
    int rounds = 0;
    for (;;) {
        uint64_t f = 0;
        for (uint64_t v = 0; v < n; ++v) {
            int16_t m = (int16_t)deg[v];
            act[v] = 0;
            if (m < k && m > 0) { deg[v] = 0; act[v] = 1; ++f; }
        }
        ++rounds;
        if (f == 0) break;
        for (uint64_t t = 0; t < w.T; ++t) {
            uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
            for (uint64_t e = w.start[t]; e < w.start[t + 1]; ++e) {
                uint64_t a = o0 + (grid[e] & 0xFFFF), b = o1 + (grid[e] >> 16);
                if (act[a]) deg[b]--;
                if (act[b]) deg[a]--;
            }
        }
    }
    uint64_t ne0 = 0, pos = 0;
    for (uint64_t v = 0; v < n; ++v) { ne0 += (deg[v] != 0); pos += (deg[v] > 0); }
    *count_ne0 = ne0;
    *count_pos = pos;
    free(act);
    walk_free(&w);
    return rounds;
}

/* ------------------------------------------------------------------------ */
/* Threaded forms (the CPU baseline's timed runs).  They parallelise the     */
/* tile sweeps over tiles with atomics, as the reference's OpenMP jobs do    */
/* (bfs.cpp:158-177 sets depths with plain stores; kcore.cpp:69-103 and      */
/* pr.cpp:60-80 use atomic updates), and give the serial functions' results. */
/* ------------------------------------------------------------------------ */

/* or_bfs with the level sweep split over tiles.  A vertex reached by two
 * threads in one level gets the same depth from both; the level count is
 * taken afterwards from the depth array, so it is exact. */
int or_bfs_mt(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint64_t root,
              uint8_t *depth, uint64_t *level_counts, int max_levels, int nthreads)
{
    walk_t w;
    if (root >= n) return -1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    memset(depth, 0, n);
    depth[root] = 1;
    int level = 1, nlev = 0;
    for (;;) {
        int any = 0;
        #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(|:any)
        for (uint64_t t = 0; t < w.T; ++t) {
            uint8_t *d0 = depth + ((uint64_t)w.row[t] << 16);
            uint8_t *d1 = depth + ((uint64_t)w.col[t] << 16);
            for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                uint32_t a = grid[k] & 0xFFFF, b = grid[k] >> 16;
                uint8_t da = __atomic_load_n(&d0[a], __ATOMIC_RELAXED), db = __atomic_load_n(&d1[b], __ATOMIC_RELAXED);
                if (da == level && db == 0) { __atomic_store_n(&d1[b], (uint8_t)(level + 1), __ATOMIC_RELAXED); any = 1; }
                else if (!directed && db == level && da == 0) {
                    __atomic_store_n(&d0[a], (uint8_t)(level + 1), __ATOMIC_RELAXED);
                    any = 1;
                }
            }
        }
        uint64_t cnt = 0;
        if (any) {
            #pragma omp parallel for num_threads(nthreads) reduction(+:cnt) schedule(static)
            for (uint64_t v = 0; v < n; ++v) cnt += depth[v] == level + 1;
        }
        if (nlev < max_levels && level_counts) level_counts[nlev] = cnt;
        ++nlev;
        if (cnt == 0) break;
        if (level + 1 > 255) { walk_free(&w); return -2; }
        ++level;
    }
    walk_free(&w);
    return nlev;
}

/* WCC by parallel hooking over the tiles: every stored edge hangs the larger
 * of its endpoints' roots under the smaller (compare-and-swap on the root's
 * parent entry), then every parent pointer jumps to its root; repeated until
 * a pass hooks nothing.  A root only ever hangs under a smaller id, so each
 * final root is its component's smallest vertex: the same labels and count
 * as or_wcc. */
static uint32_t mt_root(uint32_t *par, uint32_t x)
{
    uint32_t p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
    while (p != x) { x = p; p = __atomic_load_n(&par[x], __ATOMIC_RELAXED); }
    return x;
}

int or_wcc_mt(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint32_t *label,
              uint64_t *ncomp, int nthreads)
{
    walk_t w;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    uint32_t *par = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint8_t *touched = (uint8_t *)calloc(n, 1);
    if (!par || !touched) { free(par); free(touched); walk_free(&w); return -1; }
    #pragma omp parallel for num_threads(nthreads) schedule(static)
    for (uint64_t v = 0; v < n; ++v) par[v] = (uint32_t)v;
    for (int changed = 1; changed;) {
        changed = 0;
        #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(|:changed)
        for (uint64_t t = 0; t < w.T; ++t) {
            uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
            for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                uint32_t a = (uint32_t)(o0 + (grid[k] & 0xFFFF)), b = (uint32_t)(o1 + (grid[k] >> 16));
                touched[a] = touched[b] = 1;
                for (;;) {
                    uint32_t ra = mt_root(par, a), rb = mt_root(par, b);
                    if (ra == rb) break;
                    uint32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra, exp = hi;
                    if (__atomic_compare_exchange_n(&par[hi], &exp, lo, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                        changed = 1;
                        break;
                    }
                }
            }
        }
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for (uint64_t v = 0; v < n; ++v) par[v] = mt_root(par, (uint32_t)v);
    }
    uint64_t c = 0;
    #pragma omp parallel for num_threads(nthreads) reduction(+:c) schedule(static)
    for (uint64_t v = 0; v < n; ++v) {
        if (!touched[v]) { label[v] = 0xFFFFFFFFu; continue; }
        label[v] = par[v];
        c += (par[v] == v);
    }
    *ncomp = c;
    free(par);
    free(touched);
    walk_free(&w);
    return 0;
}

/* or_kcore with the peel and the edge pass split over threads; degree
 * decrements are atomic (kcore.cpp:69-103 uses atomic subtraction).  The
 * edge pass skips tiles whose tile row and column hold no removed vertex, as
 * the reference does (gstore.cpp:2768-2781). */
int or_kcore_mt(uint64_t n, const uint64_t *start, const uint32_t *grid, const int32_t *degree,
                int k, int32_t *deg, uint64_t *count_ne0, uint64_t *count_pos, int nthreads,
                uint64_t *scanned)
{
    uint64_t scan = 0;
    walk_t w;
    if (walk_init(&w, n, 0, start, grid)) return -1;
    uint8_t *act = (uint8_t *)calloc(n, 1);
    uint64_t rows = (n + 65535) >> 16;
    uint8_t *rowact = (uint8_t *)calloc(rows, 1);
    if (!act || !rowact) { free(act); free(rowact); walk_free(&w); return -1; }
    memcpy(deg, degree, sizeof(int32_t) * n);
    int rounds = 0;
    for (;;) {
        uint64_t f = 0;
        memset(rowact, 0, rows);
        #pragma omp parallel for num_threads(nthreads) reduction(+:f) schedule(static, 65536)
        for (uint64_t v = 0; v < n; ++v) {
            int16_t m = (int16_t)deg[v];
            act[v] = 0;
            if (m < k && m > 0) { deg[v] = 0; act[v] = 1; ++f; rowact[v >> 16] = 1; }
        }
        ++rounds;
        if (f == 0) break;
        #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64) reduction(+:scan)
        for (uint64_t t = 0; t < w.T; ++t) {
            if (!rowact[w.row[t]] && !rowact[w.col[t]]) continue;
            scan += w.start[t + 1] - w.start[t];
            uint64_t o0 = (uint64_t)w.row[t] << 16, o1 = (uint64_t)w.col[t] << 16;
            for (uint64_t e = w.start[t]; e < w.start[t + 1]; ++e) {
                uint64_t a = o0 + (grid[e] & 0xFFFF), b = o1 + (grid[e] >> 16);
                if (act[a]) __atomic_fetch_sub(&deg[b], 1, __ATOMIC_RELAXED);
                if (act[b]) __atomic_fetch_sub(&deg[a], 1, __ATOMIC_RELAXED);
            }
        }
    }
    uint64_t ne0 = 0, pos = 0;
    #pragma omp parallel for num_threads(nthreads) reduction(+:ne0,pos) schedule(static)
    for (uint64_t v = 0; v < n; ++v) { ne0 += (deg[v] != 0); pos += (deg[v] > 0); }
    *count_ne0 = ne0;
    *count_pos = pos;
    if (scanned) *scanned = scan;   /* stored edges in the tiles the edge passes visited */
    free(act);
    free(rowact);
    walk_free(&w);
    return rounds;
}

/* ------------------------------------------------------------------------ */
/* Remaining jobs and tools: -j 30 (bfs2), -j 4 (traverse), -c 4 (text).     */
/* ------------------------------------------------------------------------ */

/* Asynchronous BFS, -j 30 (bfs2_t, bfs2.cpp:82-179; driver grid::bfs2,
 * gstore.cpp:1900-2042).  depth u8, 0 = unreached, root = 1.  A sweep visits
 * the tiles in .start order and relaxes every stored edge in place:
 * l0 = d[v0], l1 = d[v1]; if l0 != 0 and (l1 == 0 or l1 - l0 > 1) then
 * d[v1] = l0 + 1, else (undirected) if l1 != 0 and (l0 == 0 or l0 - l1 > 1)
 * then d[v0] = l1 + 1 (bfs2.cpp:103-118).  A forward update marks the tile
 * column's row bit in the current and the next read set, a backward one the
 * tile row's bit in the next set (bfs2.cpp:121-129).  Until a sweep's update
 * count drops to 16 p (undirected; 4 p_s directed) every tile is visited
 * (ioskip, bfs2.cpp:161-172); afterwards only tiles whose row -- or, on
 * undirected grids, column -- bit is set (bfs2.cpp:251-275).  The run stops
 * at the first sweep without updates.  sweep_counts[s] = updates in sweep s.
 * The fixed point is the BFS depth of or_bfs.  Returns the number of sweeps,
 * -2 if a depth beyond 255 would be stored (the reference's u8 wraps). */
int or_bfs2(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint64_t root,
            uint8_t *depth, uint64_t *sweep_counts, int max_sweeps)
{
    walk_t w;
    if (root >= n) return -1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    uint64_t ps = 256 * w.p;
    uint8_t *cur = (uint8_t *)calloc(ps, 1), *nxt = (uint8_t *)calloc(ps, 1);
    if (!cur || !nxt) { free(cur); free(nxt); walk_free(&w); return -1; }
    memset(depth, 0, n);
    depth[root] = 1;
    int ioskip = 1, sweeps = 0, rc = 0;
    const uint64_t small = directed ? 4 * ps : 16 * w.p;
    for (;;) {
        uint64_t total = 0;
        for (uint64_t t = 0; t < w.T && !rc; ++t) {
            uint32_t r = w.row[t], c = w.col[t];
            if (!(ioskip || cur[r] || (!directed && cur[c]))) continue;
            uint8_t *d0 = depth + ((uint64_t)r << 16), *d1 = depth + ((uint64_t)c << 16);
            uint64_t fwd = 0, back = 0;
            for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                uint32_t a = grid[k] & 0xFFFF, b = grid[k] >> 16;
                int l0 = d0[a], l1 = d1[b];
                if (l0 != 0 && (l1 == 0 || l1 - l0 > 1)) {
                    if (l0 + 1 > 255) { rc = -2; break; }
                    d1[b] = (uint8_t)(l0 + 1);
                    ++fwd;
                } else if (!directed && l1 != 0 && (l0 == 0 || l0 - l1 > 1)) {
                    if (l1 + 1 > 255) { rc = -2; break; }
                    d0[a] = (uint8_t)(l1 + 1);
                    ++back;
                }
            }
            if (fwd) { nxt[c] = 1; cur[c] = 1; }
            if (back) nxt[r] = 1;
            total += fwd + back;
        }
        if (rc) break;
        if (sweeps < max_sweeps && sweep_counts) sweep_counts[sweeps] = total;
        ++sweeps;
        if (total == 0) break;
        uint8_t *tmp = cur; cur = nxt; nxt = tmp;
        memset(nxt, 0, ps);
        if (total <= small) ioskip = 0;
    }
    free(cur);
    free(nxt);
    walk_free(&w);
    return rc ? rc : sweeps;
}

/* Single colour propagation, -j 4 (traverse_t, traverse.cpp:56-211; driver
 * grid::traverse, gstore.cpp:2169-2177, and do_algo_prop).  cid starts at
 * 0xFFFFFFFF (invalid_cid; the reference's calloc path, traverse.cpp:66-69)
 * with cid[pivot] = color; a sweep visits every tile and, for each stored
 * edge, sets cid[v1] = color when cid[v0] == color and cid[v1] is invalid
 * (traverse.cpp:178-186) -- one direction only, also on undirected grids,
 * where v0 < v1.  Stops at the first sweep that colours nothing.
 * sweep_counts[s] = vertices coloured in sweep s.  Returns the sweeps run,
 * -1 on bad arguments (color may not be 0xFFFFFFFF). */
int or_traverse(uint64_t n, int directed, const uint64_t *start, const uint32_t *grid, uint64_t pivot,
                uint32_t color, uint32_t *cid, uint64_t *sweep_counts, int max_sweeps)
{
    walk_t w;
    if (pivot >= n || color == 0xFFFFFFFFu) return -1;
    if (walk_init(&w, n, directed, start, grid)) return -1;
    for (uint64_t v = 0; v < n; ++v) cid[v] = 0xFFFFFFFFu;
    cid[pivot] = color;
    int sweeps = 0;
    for (;;) {
        uint64_t total = 0;
        for (uint64_t t = 0; t < w.T; ++t) {
            uint32_t *c0 = cid + ((uint64_t)w.row[t] << 16), *c1 = cid + ((uint64_t)w.col[t] << 16);
            for (uint64_t k = w.start[t]; k < w.start[t + 1]; ++k) {
                uint32_t a = grid[k] & 0xFFFF, b = grid[k] >> 16;
                if (c0[a] == color && c1[b] == 0xFFFFFFFFu) { c1[b] = color; ++total; }
            }
        }
        if (sweeps < max_sweeps && sweep_counts) sweep_counts[sweeps] = total;
        ++sweeps;
        if (total == 0) break;
    }
    walk_free(&w);
    return sweeps;
}

/* Text edge list to u32 pairs, -c 4 (text_to_bin, util.cpp:58-168).  Lines
 * at the very start that begin with '%' are skipped (util.cpp:71-77).  The
 * scanner takes a token at the start of the remaining text and after every
 * run of ' ', '\n', '\t' (util.cpp:91-103), so text that starts with one of
 * those yields its first number twice.  Each token's value is atoi of its
 * text: leading white space skipped, optional sign, decimal digits, the
 * strtol result (saturated to the 64-bit range) cut to 32 bits.  v_min starts
 * at 999999 and v_max at 0 (util.cpp:86-96, Q14); there are tokens/2 pairs
 * (v0 - v_min, v1 - v_min) (util.cpp:130-161).  Fills out[0 .. 2*pairs) when
 * out is given (cap_pairs pairs at most) and returns the number of pairs;
 * *ntok = tokens. */
static int txt_delim(char c) { return c == ' ' || c == '\n' || c == '\t'; }
static int txt_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

static uint32_t txt_atoi(const char *s, uint64_t i, uint64_t size)
{
    while (i < size && txt_space(s[i])) ++i;
    int neg = 0;
    if (i < size && (s[i] == '+' || s[i] == '-')) { neg = s[i] == '-'; ++i; }
    uint64_t v = 0;
    int sat = 0;
    while (i < size && s[i] >= '0' && s[i] <= '9') {
        uint64_t d = (uint64_t)(s[i] - '0');
        if (!sat && v > (0x7FFFFFFFFFFFFFFFull - d) / 10) sat = 1;
        if (!sat) v = v * 10 + d;
        ++i;
    }
    int64_t r;
    if (sat) r = neg ? (int64_t)(-0x7FFFFFFFFFFFFFFFll - 1) : 0x7FFFFFFFFFFFFFFFll;
    else r = neg ? -(int64_t)v : (int64_t)v;
    return (uint32_t)(uint64_t)r;
}

int64_t or_text_to_bin(const char *text, uint64_t size, uint32_t *out, uint64_t cap_pairs, uint32_t *vmin,
                       uint32_t *vmax, uint64_t *ntok)
{
    uint64_t h = 0;
    while (h < size && text[h] == '%') {
        while (h < size && text[h] != '\n') ++h;
        if (h < size) ++h;
    }
    const char *ss = text + h;
    uint64_t sz = size - h, tok = 0;
    uint32_t mn = 999999u, mx = 0;
    for (uint64_t i = 0; i < sz; ++i) {
        if (i != 0 && !(txt_delim(ss[i - 1]) && !txt_delim(ss[i]))) continue;
        uint32_t a = txt_atoi(ss, i, sz);
        if (a > mx) mx = a;
        if (a < mn) mn = a;
        ++tok;
    }
    uint64_t pairs = tok / 2, t = 0;
    if (out) {
        for (uint64_t i = 0; i < sz && t < 2 * pairs; ++i) {
            if (i != 0 && !(txt_delim(ss[i - 1]) && !txt_delim(ss[i]))) continue;
            if (t < 2 * cap_pairs) out[t] = txt_atoi(ss, i, sz) - mn;
            ++t;
        }
    }
    *vmin = mn;
    *vmax = mx;
    *ntok = tok;
    return (int64_t)pairs;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
