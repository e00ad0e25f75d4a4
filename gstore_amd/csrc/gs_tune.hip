// gs_tune.hip -- autotuner for the sweep knobs of one graph.
//
// Knobs: edges per workgroup (the tile chunk each workgroup stages, 1024 /
// 2048 / 4096) and the chunk-to-XCD order (contiguous eighths vs round-robin),
// timed on the tile methods of WCC (hook and check passes are full sweeps) and,
// for undirected graphs, k-core with k = 8 (sweeps that skip inactive tiles);
// the smallest sum is applied.  On undirected resident grids, the
// direction-optimizing BFS switch thresholds (alpha, beta) are then timed on
// eight roots at evenly spaced degree ranks (what uniformly drawn roots with
// edges look like); a setting replaces the current one only when at least 2%
// faster, so timing noise does not move it.  With a cache path, results are kept per
// graph signature (N, E, non-empty tiles, directed) so a later open of the
// same graph applies them without timing.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "gs_internal.h"

using namespace gs;

namespace {

std::string signature(const gs_graph *g) {
    char buf[128];
    snprintf(buf, sizeof buf, "n=%llu e=%llu nz=%llu d=%d", (unsigned long long)g->n, (unsigned long long)g->E,
             (unsigned long long)g->nz, (int)g->directed);
    return buf;
}

bool cache_lookup(const char *path, const std::string &sig, uint32_t *chunk, int *order, int *alpha, int *beta) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool hit = false;
    while (fgets(line, sizeof line, f)) {
        // format: <signature> | chunk=<c> order=<o>
        char *bar = strchr(line, '|');
        if (!bar) continue;
        std::string key(line, bar - line);
        while (!key.empty() && key.back() == ' ') key.pop_back();
        unsigned c;
        int o, al = 0, be = 0;
        const int got = sscanf(bar + 1, " chunk=%u order=%d alpha=%d beta=%d", &c, &o, &al, &be);
        if (key == sig && got >= 2) {
            *chunk = c;
            *order = o;
            *alpha = got == 4 ? al : 0;
            *beta = got == 4 ? be : 0;
            hit = true;  // last entry wins
        }
    }
    fclose(f);
    return hit;
}

}  // namespace

extern "C" int gs_tune(gs_graph *g, const char *cache_path, gs_tune_result *out) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    gs_tune_result r{};
    const std::string sig = signature(g);
    uint32_t c;
    int o, al, be;
    if (cache_path && cache_lookup(cache_path, sig, &c, &o, &al, &be)) {
        if (!g->ss) GS_TRY(gs_set_chunk(g, c));
        g->xcd_order = o;
        if (al > 0 && be > 0) {
            g->do_alpha = al;
            g->do_beta = be;
        }
        r.chunk = g->chunk;
        r.xcd_order = g->xcd_order;
        r.do_alpha = al;
        r.do_beta = be;
        r.from_cache = 1;
        if (out) *out = r;
        return GS_OK;
    }
    // the sweep knobs act on the tile methods: time those
    const int wm = g->wcc_method, km = g->kcore_method;
    g->wcc_method = 0;
    g->kcore_method = 0;
    const uint32_t chunks[3] = {1024, 2048, 4096};
    double best = 1e30;
    uint32_t best_c = g->chunk;
    int best_o = g->xcd_order, k = 0;
    const bool timing = g->ws->timer.on;
    g->ws->timer.on = false;
    for (int ci = 0; ci < 3; ++ci) {
        if (g->ss && chunks[ci] != g->chunk) continue;  // streaming graphs keep their chunk
        if (!g->ss) GS_TRY(gs_set_chunk(g, chunks[ci]));
        for (int order = 0; order < 2; ++order, ++k) {
            g->xcd_order = order;
            GS_TRY(run_wcc(g, g->ws, nullptr, nullptr));  // warm
            GS_TRY(run_wcc(g, g->ws, nullptr, nullptr));
            double ms = g->ws->stats.ms;
            if (!g->directed) {
                GS_TRY(run_kcore(g, g->ws, 8, nullptr, nullptr, nullptr, nullptr));
                ms += g->ws->stats.ms;
            }
            if (k < 6) {
                r.cand_chunk[k] = chunks[ci];
                r.cand_order[k] = order;
                r.cand_ms[k] = ms;
            }
            if (ms < best) {
                best = ms;
                best_c = chunks[ci];
                best_o = order;
            }
        }
    }
    g->wcc_method = wm;
    g->kcore_method = km;
    if (!g->ss) GS_TRY(gs_set_chunk(g, best_c));
    g->xcd_order = best_o;
    r.chunk = g->chunk;
    r.xcd_order = g->xcd_order;
    r.ncand = k < 6 ? k : 6;
    r.best_ms = best;
    // direction-optimizing BFS thresholds, on roots at evenly spaced degree ranks
    if (!g->directed && !g->ss) {
        GS_TRY(build_adjacency(g));
        if (g->nzn >= 16) {
            constexpr int NR = 8;
            uint32_t roots[NR];
            for (int i = 0; i < NR; ++i)
                GS_HIP(hipMemcpy(&roots[i], g->old_id.ptr + g->nzn * (2 * i + 1) / (2 * NR), 4, hipMemcpyDeviceToHost));
            const int alphas[6] = {10, 14, 20, 30, 45, 60}, betas[2] = {24, 64};
            auto time_roots = [&](double *ms) -> int {
                *ms = 0;
                for (int ri = 0; ri < NR; ++ri) {
                    GS_TRY(run_bfs_do(g, g->ws, roots[ri], nullptr, nullptr, 0, nullptr));
                    *ms += g->ws->stats.ms;
                }
                return GS_OK;
            };
            int ba = g->do_alpha, bb = g->do_beta;
            GS_TRY(run_bfs_do(g, g->ws, roots[0], nullptr, nullptr, 0, nullptr));  // warm
            double best_do = 0;
            GS_TRY(time_roots(&best_do));   // the current setting is the one to beat
            for (int ai = 0; ai < 6; ++ai)
                for (int bi = 0; bi < 2; ++bi) {
                    if (alphas[ai] == ba && betas[bi] == bb) continue;
                    g->do_alpha = alphas[ai];
                    g->do_beta = betas[bi];
                    double ms = 0;
                    GS_TRY(time_roots(&ms));
                    if (ms < 0.98 * best_do) {
                        best_do = ms;
                        ba = alphas[ai];
                        bb = betas[bi];
                    }
                }
            g->do_alpha = ba;
            g->do_beta = bb;
            r.do_alpha = ba;
            r.do_beta = bb;
            r.do_ms = best_do;
        }
    }
    g->ws->timer.on = timing;
    if (cache_path) {
        FILE *f = fopen(cache_path, "a");
        if (f) {
            fprintf(f, "%s | chunk=%u order=%d alpha=%d beta=%d sweep_ms=%.3f\n", sig.c_str(), g->chunk,
                    g->xcd_order, r.do_alpha, r.do_beta, best);
            fclose(f);
        }
    }
    if (out) *out = r;
    return GS_OK;
}
