/*
 * gstore_amd.h -- C ABI of the MI355X G-Store engine.
 *
 * The reference (the-data-lab/gstore) has no library interface: its drop-in
 * surface is the gstoreu/gstored command line and the .grid/.start/.degree
 * tile files, with algorithms plugged in through algo_t (algo.h:38-44).  Each
 * entry point below replaces one reference function or job, cited by
 * file:line.  Conventions: integer status codes (GS_OK = 0), no exceptions
 * across the ABI, caller-owned host output buffers (a NULL output pointer
 * leaves the result on the device), device memory owned by the gs_graph
 * handle.  One call at a time per handle; handles are independent.
 */
#ifndef GSTORE_AMD_H
#define GSTORE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gs_graph gs_graph;

enum gs_status {
    GS_OK = 0,
    GS_ERR_ARG = 1,            /* bad argument */
    GS_ERR_IO = 2,             /* file missing, short read/write */
    GS_ERR_FORMAT = 3,         /* tile files inconsistent with each other or with N */
    GS_ERR_NOMEM = 4,          /* device or host allocation failed */
    GS_ERR_DEVICE = 5,         /* HIP runtime error */
    GS_ERR_UNSUPPORTED = 6,    /* e.g. k-core on a directed grid (gstore.cpp:2348-2351) */
    GS_ERR_DEPTH_OVERFLOW = 7  /* BFS would need depth > 255 (u8 depth, gstore.h:49) */
};

enum gs_pr_mode {
    GS_PR_COMPAT = 0,   /* degree factor (int32)(1.0/deg) as pr.cpp:112-114 stores it */
    GS_PR_INTENDED = 1  /* degree factor (float)(1.0/deg) */
};

/* Kernel classes whose device time is recorded when kernel timing is on. */
enum gs_kernel {
    GS_K_PR_SCATTER = 0, GS_K_PR_FINALIZE, GS_K_BFS_EXPAND, GS_K_BFS_LEVEL_END,
    GS_K_WCC_HOOK, GS_K_WCC_COMPRESS, GS_K_WCC_CHECK, GS_K_KCORE_PEEL, GS_K_KCORE_SCAN,
    GS_K_GEN, GS_K_CONVERT, GS_K_H2D, GS_K_PR_PULL, GS_K_INDEX, GS_K_PB_GATHER, GS_K_PB_ACCUM,
    GS_K_PB_REDUCE, GS_K_BFS_TOPDOWN, GS_K_BFS_BOTTOMUP, GS_K_COUNT
};

typedef struct gs_info {
    uint64_t vert_count;      /* N */
    uint64_t edge_count;      /* E = stored edges = start[T] */
    uint64_t tile_slots;      /* T (calc_total_part, gstore.h:121-125) */
    uint64_t nonempty_tiles;
    uint64_t groups;          /* p, physical groups per side */
    uint64_t tiles_per_side;  /* p_s = 256 p */
    uint64_t device_bytes;    /* device memory held by the handle */
    uint64_t resident_edges;  /* edges kept in HBM (all of them when resident) */
    uint64_t segment_edges;   /* edges per streamed segment (0 when resident) */
    int32_t directed;
    int32_t resident;         /* 1: whole grid in HBM */
    uint64_t index_entries;   /* neighbour-list index entries (2E once built, else 0) */
    uint64_t index_vertices;  /* vertices with at least one stored edge (index built) */
    double index_ms;          /* build time of the index and its blocked copy (0 until built) */
} gs_info;

typedef struct gs_stats {
    double ms;                     /* device time of the last algorithm call */
    uint64_t iterations;           /* PR iterations, BFS levels, WCC passes, k-core rounds */
    uint64_t edges_scanned;        /* stored edges visited by edge sweeps (all iterations) */
    uint64_t tiles_scanned;
    uint64_t h2d_bytes;            /* grid bytes streamed from host memory (semi-external mode) */
    double kern_ms[GS_K_COUNT];    /* per-class device time (kernel timing on) */
    uint64_t kern_calls[GS_K_COUNT];
} gs_stats;

/* Open a converted graph: PREFIX.grid, PREFIX.start, PREFIX.degree.  Replaces
 * read_grid_in_mem / read_start_in_mem / read_degree_in_mem
 * (gstore.cpp:2953-3011, 3799-3813) and the -v/-s geometry
 * (gstore.cpp:143-161): vert_count is N (the -v value).  The whole grid is
 * placed in HBM when it fits device_budget_bytes (0 = no limit).  The files
 * are checked against each other (sizes, start monotone, sentinel, degree
 * sums, local ids < N). */
int gs_open(const char *prefix, uint64_t vert_count, int directed,
            uint64_t device_budget_bytes, gs_graph **out);

/* Same as gs_open from host arrays already in memory (start: T+1 u64,
 * grid: E u32 words {u16 v0, u16 v1}, degree: N i32). */
int gs_open_arrays(uint64_t vert_count, int directed, const uint64_t *start,
                   const uint32_t *grid, const int32_t *degree, gs_graph **out);

/* Generate the Graph500 Kronecker graph of 2^scale vertices and
 * edge_factor * 2^scale tuples with user seeds (seed1, seed2) on the device
 * and convert it to the tile grid there.  Replaces make_graph
 * (graph500-generator/make_graph.c:42-59) followed by the -c 1 converter
 * proc_grid (gstore.cpp:1329-1458). */
int gs_open_kron(int scale, int edge_factor, uint64_t seed1, uint64_t seed2, int directed,
                 gs_graph **out);

/* Generator alone: tuples [e0, e1) into out (u32 pairs), computed on the
 * device (graph500-generator/graph_generator.c:163-198). */
int gs_kron_edges(int scale, uint64_t seed1, uint64_t seed2, int64_t e0, int64_t e1,
                  uint32_t *out);

/* Converter alone (-c 1, gstore.cpp:1245-1458, 2818-2935): tuples (u32
 * pairs) -> PREFIX.grid/.start/.degree written to disk.  Self-loops are
 * dropped, duplicates kept, undirected edges stored as (min, max). */
int gs_convert(const uint32_t *edges, uint64_t tuples, uint64_t vert_count, int directed,
               const char *out_prefix);

/* Write an open graph's tile files (save_grid, gstore.cpp:2818-2935; the
 * unused .sdegree/.bdegree are not written). */
int gs_save(gs_graph *g, const char *prefix);

int gs_get_info(gs_graph *g, gs_info *info);

/* Cap the HBM used by the tile grid (semi-external mode, replacing the
 * reference's libaio segment streaming, gstore.cpp:3044-3946, and its
 * cross-iteration cache, :1491-1768).  If the grid is larger than
 * budget_bytes it moves to pinned host memory; the first
 * budget - 2 * segment_bytes stay resident in HBM (swept first, never
 * copied again) and the rest is copied through two segment buffers on a
 * side stream, each copy overlapping the sweep of the previous segment.
 * segment_bytes 0 = min(512 MiB, budget / 8).  gs_open with a budget below the
 * grid size opens in this mode directly.  A streaming graph serves one call
 * at a time. */
int gs_set_device_budget(gs_graph *g, uint64_t budget_bytes, uint64_t segment_bytes);

/* Open with the grid left in its file (external mode; the SSD-streaming
 * path of the reference: io_driver / read_aio_random, gstore.cpp:3623-3797,
 * buffers read_aio_init :3044-3119, cross-iteration cache :1491-1768).  When
 * the grid is larger than device_budget_bytes, the first budget - 2 segments
 * of it are read once into HBM (the cache, swept first in every pass); every
 * streamed pass reads the rest from PREFIX.grid with O_DIRECT preads
 * (through an aligned bounce buffer where the pinned buffers cannot be
 * O_DIRECT targets, buffered where the file system refuses O_DIRECT) by 4 IO
 * threads into 3
 * pinned staging buffers, and copies each staged segment to one of two device
 * segment buffers while the previous one is swept.  No host copy of the grid
 * is kept.  segment_bytes 0 = min(256 MiB, budget / 8).  A grid that fits the
 * budget opens resident as with gs_open. */
int gs_open_external(const char *prefix, uint64_t vert_count, int directed, uint64_t device_budget_bytes,
                     uint64_t segment_bytes, gs_graph **out);

/* File reads of an external-mode graph since it was opened (zeros
 * otherwise): bytes read, of them with O_DIRECT, read time summed over IO
 * threads, and the time sweeps waited for a segment's reads. */
typedef struct gs_io_stats {
    uint64_t disk_bytes;
    uint64_t direct_bytes;
    double read_ms;
    double wait_ms;
    int32_t io_threads;
    int32_t staging_buffers;
    int32_t external;
    int32_t read_mode;     /* 0 O_DIRECT into pinned buffers, 1 O_DIRECT through a bounce buffer, 2 buffered */
} gs_io_stats;
int gs_get_io_stats(gs_graph *g, gs_io_stats *out);

/* Copy the tile arrays back to host buffers (start: T+1, grid: E, degree: N). */
int gs_get_arrays(gs_graph *g, uint64_t *start, uint32_t *grid, int32_t *degree);

/* Semi-external mode only: the pinned host copy of the .grid array (E
 * edges), read-only, owned by the graph and valid until gs_close or the next
 * gs_set_device_budget.  Lets a host reader of a grid too large for a second
 * copy (a checker, a writer) use it in place.  GS_ERR_UNSUPPORTED for
 * resident and external graphs.  No counterpart in the reference (its grid
 * lives in the process's own memory). */
int gs_get_host_grid(gs_graph *g, const uint32_t **grid, uint64_t *edges);

/* PageRank, `iters` iterations (pr_t, pr.cpp:82-303; resident driver
 * pagerank_mmap gstore.cpp:2572-2629; -j 1/11/21 -a iters, default 5). */
int gs_pagerank(gs_graph *g, int iters, int mode, float *rank_out);

/* Personalised PageRank from one source (no reference counterpart; same
 * tile scatter as gs_pagerank): r_0 = e_source, prior = r * (float)(1/deg),
 * r_{t+1} = 0.15 [v == source] + 0.85 * sum of prior over neighbours.
 * rank_out[v] = r after `iters` iterations (FP32 ranks, FP64 per-vertex
 * sums). */
int gs_ppr(gs_graph *g, uint64_t source, int iters, float *rank_out);

/* BFS from root (bfs_t, bfs.cpp:56-338; driver bfs_mmap gstore.cpp:2631-2713;
 * -j 0/10/20 -a root).  depth_out: u8, 0 = unreached, root = 1.
 * level_counts[L-1] = vertices reached at level L (up to max_levels).
 * *levels = number of levels run (including the last, empty one). */
int gs_bfs(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
           int max_levels, int *levels);

/* BFS that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture
 * setup, tight serving loops, harnesses that forbid hipMalloc).  Same results
 * and encodings as gs_bfs with the tile-sweep method; its work arrays (level
 * counters, tile-row bits, three N-bit bitmaps, N depth bytes: about 1.375 N
 * bytes) live in one static 32 MiB device array per GPU, so it serves graphs
 * of up to about 24 M vertices and returns GS_ERR_ARG beyond that, and
 * GS_ERR_UNSUPPORTED on semi-external graphs.  It calls no hipMalloc/hipFree.
 * Every call on one GPU shares that array: calls that overlap on one GPU race
 * (each call clears what it uses before it starts). */
int gs_bfs_noalloc(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *level_counts,
                   int max_levels, int *levels);

/* Batched BFS: up to 64 roots run as bit lanes of one multi-source sweep
 * (no reference counterpart; each lane applies the rule of bfs.cpp:158-177).
 * depth_out: n_roots x N bytes, lane i at depth_out + i*N, same encoding as
 * gs_bfs and bit-identical to gs_bfs(roots[i]).  level_counts: n_roots x
 * max_levels (lane-major).  *levels = levels run for the whole batch. */
int gs_bfs_batch(gs_graph *g, const uint64_t *roots, int n_roots, uint8_t *depth_out,
                 uint64_t *level_counts, int max_levels, int *levels);

/* Weakly connected components (wcc2_t, wcc.cpp:57-395; driver wcc_mmap
 * gstore.cpp:2502-2570; -j 3/13/23).  label_out[v] = smallest vertex id of
 * v's component for vertices with a stored edge, 0xFFFFFFFF otherwise.
 * *ncomp = components with at least one stored edge ("WCC count"). */
int gs_wcc(gs_graph *g, uint32_t *label_out, uint64_t *ncomp);

/* k-core peeling (kcore_t, kcore.cpp:55-300; driver kcore_mmap
 * gstore.cpp:2716-2804; -j 2/12/22 -a k, default 10).  deg_out = final
 * degree array (removed vertices <= 0).  *count_ne0 = #(deg != 0), the
 * reference's printed count; *count_pos = #(deg > 0).  *rounds = removal
 * rounds run, including the first. */
int gs_kcore(gs_graph *g, int k, int32_t *deg_out, uint64_t *count_ne0, uint64_t *count_pos,
             int *rounds);

/* Asynchronous BFS, -j 30 (bfs2_t, bfs2.cpp:82-179; driver grid::bfs2,
 * gstore.cpp:1900-2042; root -a, default 1).  Sweeps relax every stored edge
 * in place -- d[v1] = d[v0] + 1 when v1 is unreached or deeper by more than
 * one, and the reverse on undirected grids -- until a sweep changes nothing;
 * after the first sweep that stores at most 16 p depths (4 p_s directed),
 * only tiles whose row (or column) changed in the previous sweep are swept.
 * depth_out: u8, identical to gs_bfs.  sweep_counts[s] = depths stored in
 * sweep s (the order of stores, hence the sweep count and these counts, varies
 * between runs as in the reference).  *sweeps = sweeps run, the last one
 * storing nothing.  Works on semi-external grids too. */
int gs_bfs_async(gs_graph *g, uint64_t root, uint8_t *depth_out, uint64_t *sweep_counts, int max_sweeps,
                 int *sweeps);

/* Single colour propagation, -j 4 (traverse_t, traverse.cpp:56-211; driver
 * grid::traverse, gstore.cpp:2169-2177; pivot -a, default 1, colour 0).
 * cid starts at 0xFFFFFFFF except cid[pivot] = color; each sweep sets
 * cid[v1] = color for every stored edge (v0, v1) with cid[v0] == color and
 * cid[v1] unset -- v0 to v1 only, so on an undirected grid colour spreads
 * along increasing vertex ids -- until a sweep colours nothing.  cid_out: N
 * u32; sweep_counts[s] = vertices coloured in sweep s (order-dependent like
 * gs_bfs_async's); *colored = vertices with the colour.  color may not be
 * 0xFFFFFFFF. */
int gs_traverse(gs_graph *g, uint64_t pivot, uint32_t color, uint32_t *cid_out, uint64_t *sweep_counts,
                int max_sweeps, int *sweeps, uint64_t *colored);

/* Text edge list to u32 pairs, -c 4 (text_to_bin, util.cpp:58-168): leading
 * lines starting with '%' skipped; a token at the text's start and after every
 * run of ' ', '\n', '\t'; each token's value as atoi reads it; pairs of
 * consecutive tokens with the minimum id subtracted, the minimum starting at
 * 999999 as the reference's does.  Parsed on the device.  Writes out_path
 * (NULL: text_path + ".edge"), *pairs = tokens / 2, *tokens, *vmin, *vmax. */
int gs_text_to_bin(const char *text_path, const char *out_path, uint64_t *pairs, uint32_t *vmin, uint32_t *vmax,
                   uint64_t *tokens);
/* The same from a text in host memory; out (cap_pairs pairs) receives the
 * first min(cap_pairs, *pairs) pairs. */
int gs_text_to_bin_mem(const char *text, uint64_t size, uint32_t *out, uint64_t cap_pairs, uint64_t *pairs,
                       uint32_t *vmin, uint32_t *vmax, uint64_t *tokens);

/* ---- Serving (no reference counterpart) ------------------------------------
 * Several resident graphs, queries submitted at any time, a pool of worker
 * threads each with its own stream and per-graph workspaces so queries overlap
 * on the device.  Continuous batching: a free worker takes the oldest pending
 * query; a BFS takes every pending BFS of its graph along (up to max_batch,
 * default 64) into one gs_bfs_batch sweep.  A query starts only when the
 * workspace memory it needs fits hbm_budget_bytes (0 = unlimited) beside the
 * memory other workspaces hold; idle workspaces are freed to make room.  A BFS
 * that finds fewer waiting BFS of its graph than the graph's batch_min runs
 * alone on the single-root direction-optimizing path instead.
 * Results: `out` (optional, caller-owned, must stay valid until the query
 * completes) receives the BFS depth (N bytes), the PPR / PageRank ranks or the
 * WCC labels (4N bytes); the result record carries latency stamps. */
enum gs_query_kind { GS_Q_BFS = 0, GS_Q_PPR = 1, GS_Q_WCC = 2, GS_Q_PAGERANK = 3 };

typedef struct gs_query_result {
    uint64_t id;
    uint64_t arg;        /* BFS root / PPR source / PageRank mode */
    uint64_t value;      /* BFS: reached vertices; WCC: component count */
    uint64_t aux;        /* BFS: levels of its batch */
    uint64_t reserved;   /* 0 */
    double t_submit_ms, t_start_ms, t_done_ms;  /* server clock */
    int32_t graph, kind, status, batch_size;
} gs_query_result;

typedef struct gs_server gs_server;

int gs_server_create(uint64_t hbm_budget_bytes, int workers, gs_server **out);
/* The server does not own the graph; keep it open until gs_server_destroy.
 * On undirected grids it builds the graph's neighbour-list index and blocked
 * PageRank copy, and measures the BFS batching rule (gs_server_bfs_policy),
 * before returning. */
int gs_server_add_graph(gs_server *s, gs_graph *g, int *graph_id);
/* max_batch 1..64 BFS lanes; iterations of PPR and PageRank queries (default 20). */
int gs_server_set(gs_server *s, int max_batch, int ppr_iters, int pr_iters);
int gs_server_submit(gs_server *s, int graph_id, int kind, uint64_t arg, void *out, uint64_t *id);
/* Completed results (up to max), waiting up to wait_ms for the first. */
int gs_server_poll(gs_server *s, gs_query_result *out, int max, int wait_ms, int *n);
/* Block until nothing is pending or running. */
int gs_server_drain(gs_server *s);
int gs_server_info(gs_server *s, uint64_t *held_bytes, uint64_t *pending, uint64_t *inflight);
/* Device memory: workspace bytes the queries hold (what hbm_budget_bytes
 * bounds) and the graphs' own bytes (grid, index, blocked copy). */
int gs_server_memory(gs_server *s, uint64_t *workspace_bytes, uint64_t *graph_bytes);
/* The BFS batching rule of one graph, measured when it was added: a batch
 * forms once batch_min BFS of the graph are waiting (65: never); fewer run one
 * at a time on the single-root direction-optimizing path.  Also returned: the
 * measured ms of one single-root BFS and of batches of 1 and 64 lanes. */
int gs_server_bfs_policy(gs_server *s, int graph, int *batch_min, double *single_ms, double *batch1_ms,
                         double *batch64_ms);
/* Override batch_min (1 = always batch, 65 = never). */
int gs_server_set_batch_min(gs_server *s, int graph, int batch_min);
/* Message of the last query that failed inside a worker (status in its
 * result record); returns its length. */
size_t gs_server_last_error(gs_server *s, char *buf, size_t len);
void gs_server_destroy(gs_server *s);

/* BFS method for gs_bfs: 0 (default) sweeps the active tiles of the grid each
 * level; 1 is direction-optimizing: it builds (once, kept with the graph) a
 * neighbour-list index from the grid -- 8(N+1) + 8E bytes of HBM -- and runs
 * each level top-down over the frontier's lists or bottom-up (each unreached
 * vertex stops at its first frontier neighbour), switching to bottom-up when
 * the frontier's edges exceed the unreached side's edges / alpha and back when
 * the frontier holds fewer than N / beta vertices (defaults 20, 24; 0 keeps
 * the current values, e.g. those gs_tune chose).  Depths and level counts are
 * identical in both modes.  Undirected grids only. */
int gs_set_bfs_mode(gs_graph *g, int mode, int alpha, int beta);

/* PageRank method on undirected resident grids:
 * 2 (default) blocked: a copy of the neighbour-list index ordered by
 *   destination block then source block (2^15 vertices each) turns every
 *   iteration into two streamed passes -- stage a block of priors in LDS and
 *   write each entry's value; add the values into an LDS array over the
 *   destination block (falls back to 1 above 2^28 vertices with edges);
 * 1 pulls over the neighbour-list index (built once from the grid, as for
 *   direction-optimizing BFS): each vertex's sum is formed in registers and
 *   LDS and stored, with no per-edge atomics;
 * 0 scatters over the tile grid with FP32 atomics, the reference's scheme
 *   (pr.cpp:119-154).
 * Personalised PageRank sums in FP64 with either (the blocked accumulation
 * then covers half a destination block per pass).  Directed and
 * semi-external grids scatter over the tiles: 1 and 2 select the binned
 * scatter -- per piece of the grid (a streamed segment), each edge's
 * contributions are written into bins of 2^16 destinations (its tile's row
 * and column) at places reserved per tile, and each bin is summed in LDS and
 * added to the ranks without global atomics -- and 0 the atomic scatter.
 * Only the FP32 summation order differs. */
int gs_set_pagerank_method(gs_graph *g, int method);

/* WCC method on undirected resident grids: 1 (default) links over the
 * neighbour-list index -- two neighbours of every vertex, then, outside the
 * largest component found by sampling, the remaining neighbours (the Afforest
 * scheme); 0 repeats hook passes over the tile grid until no stored edge joins
 * two trees.  Labels and count are identical (minimum vertex id per component,
 * wcc.cpp:324-337 count).  Directed and semi-external grids use 0. */
int gs_set_wcc_method(gs_graph *g, int method);

/* Batched BFS method on undirected resident grids: 1 runs the roots as bit
 * lanes over the neighbour-list index, direction-optimizing on the union
 * frontier (top-down pushes of lane masks, bottom-up ORs of the neighbours'
 * masks with a wave per vertex of degree >= 128), with a host round trip per
 * level; 2 (default) runs the levels with the direction choice and the stop
 * test in kernels, four levels per launch of a hipGraph captured once per
 * workspace, and in bottom-up levels pushes the lanes whose own frontiers are
 * still small top-down in the same level; 0 sweeps the tiles with lane masks.
 * Depths and level counts are identical.  Directed and semi-external grids
 * use 0. */
int gs_set_batch_bfs_method(gs_graph *g, int method);

/* k-core method on undirected resident grids: 1 (default) lists the vertices
 * each peel removes and pushes one decrement per entry of their neighbour
 * lists (the index), the heaviest vertices' decrements gathered per
 * workgroup; 0 rescans the tiles whose row or column holds a removed vertex
 * (kcore.cpp:69-103 with the skip of gstore.cpp:2768-2781).  Final degrees,
 * counts and rounds are identical.  Semi-external grids use 0. */
int gs_set_kcore_method(gs_graph *g, int method);

/* Work split of the blocked PageRank: source-major entries (accepted up to
 * 2^20, used up to 2^16) and runs (<= 2048) per gather workgroup, stream values per accumulation
 * workgroup (<= 2^30), the stream values per vertex from which a destination
 * block adds with LDS float atomics rather than compare-and-swap, and the
 * compaction threshold in permille (0..1000): a destination block whose
 * segments of equal destinations (per 16-entry window) number at most that
 * share of its entries gets one stream value per segment, summed by the
 * gather, instead of one per entry (0: never, 1000: always).  Defaults 2^16,
 * 2048, 2^21, 2^30 (compare-and-swap throughout), 800.  Small values
 * exercise the split paths on small graphs; results change only in FP32
 * (FP64 for personalised PageRank) summation order.  No counterpart in the
 * reference. */
int gs_set_pagerank_blocking(gs_graph *g, uint64_t gather_entries, uint32_t gather_runs, uint64_t accum_entries,
                             uint32_t dense_entries, uint32_t compact_permille);

/* Work split of the blocked PageRank's gather: 0 launches one workgroup per
 * unit (each stages its source block's priors); n > 0 (default 8) keeps one
 * workgroup per CU resident, taking n consecutive units at a time and
 * staging a source block only when it changes.  Results do not change.  No
 * counterpart in the reference. */
int gs_set_pagerank_gather(gs_graph *g, uint32_t units_per_grab);

/* PageRank / personalised PageRank on undirected grids: 1 (default) adds
 * into the (at most 2048) highest-degree vertices through a per-workgroup LDS
 * table flushed once per workgroup, instead of one contended global atomic
 * per edge; 0 uses plain global atomics for every vertex.  Only the FP32
 * summation order of those vertices changes. */
int gs_set_pagerank_heavy(gs_graph *g, int on);

/* Edges per workgroup in every edge sweep: a multiple of 1024 up to 4096
 * (default 4096).  Results do not depend on it; gs_tune picks it. */
int gs_set_chunk(gs_graph *g, uint32_t edges);

/* Order in which workgroups take edge chunks: 1 (default) gives each of the
 * 8 XCDs a contiguous eighth of the grid (its L2 sees consecutive tiles);
 * 0 deals chunks to XCDs round-robin (all XCDs work in one region of the
 * grid at a time, sharing the last-level cache).  Results do not depend on
 * it; gs_tune picks the faster one. */
int gs_set_sweep_order(gs_graph *g, int xcd_contiguous);

/* Autotuner: times the tile methods of WCC and k-core (full and
 * frontier-skipping sweeps) for each edges-per-workgroup value (1024, 2048,
 * 4096) and both sweep orders, and, on undirected resident grids, the
 * direction-optimizing BFS from eight roots at evenly spaced degree ranks for
 * each top-down/bottom-up switch pair (alpha, beta) in {10, 14, 20, 30, 45,
 * 60} x {24, 64}, keeping a pair only when it is at least 2% faster;
 * applies the fastest, and appends it to cache_path (optional) keyed by N, E,
 * non-empty tiles and directedness; a cached entry is applied without timing. */
typedef struct gs_tune_result {
    uint32_t chunk;
    int32_t xcd_order;
    int32_t from_cache;
    int32_t ncand;
    double best_ms;           /* ms of the timed sweeps for the chosen setting */
    uint32_t cand_chunk[6];
    int32_t cand_order[6];
    double cand_ms[6];
    int32_t do_alpha, do_beta;  /* chosen switch thresholds (0 when not tuned) */
    double do_ms;               /* ms of the two timed BFS with them */
} gs_tune_result;

int gs_tune(gs_graph *g, const char *cache_path, gs_tune_result *out);

/* Stats of the last algorithm call on this handle. */
int gs_get_stats(gs_graph *g, gs_stats *st);

/* Per-level record of the last BFS call on the graph's default workspace. */
typedef struct gs_level_stat {
    uint32_t level;       /* 1 = expansion of the root's level */
    int32_t direction;    /* 0 tile sweep, 1 top-down, 2 bottom-up */
    uint64_t frontier;    /* vertices expanded at this level */
    uint64_t edges;       /* tile sweep: edges in scanned tiles; top-down: frontier list entries;
                             bottom-up: list entries checked */
    uint64_t tiles;       /* tile sweep: non-empty tiles scanned (0 otherwise) */
    uint64_t reached;     /* vertices newly reached */
    double ms;            /* device time of the level */
} gs_level_stat;

/* Copies up to max records of the last BFS into out; *n = levels recorded.
 * No counterpart in the reference (bfs.cpp prints only the per-level count). */
int gs_get_level_stats(gs_graph *g, gs_level_stat *out, int max, int *n);

/* 1: record per-kernel-class device time with HIP events (adds a little
 * launch overhead); 0: off (default). */
int gs_set_kernel_timing(gs_graph *g, int on);

/* Test hook (no reference counterpart): the nth device allocation of the
 * next neighbour-list index build fails with GS_ERR_NOMEM (0 = off).  A failed
 * build frees every index array it allocated; the next call builds again. */
int gs_debug_fail_alloc(gs_graph *g, int nth);

/* Name of a gs_kernel class. */
const char *gs_kernel_name(int k);

void gs_close(gs_graph *g);

/* Message of the last failed call on this thread; returns its length. */
size_t gs_last_error(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* GSTORE_AMD_H */
