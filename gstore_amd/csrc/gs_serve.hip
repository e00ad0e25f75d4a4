// gs_serve.hip -- serving loop with continuous batching (no reference
// counterpart; the reference runs one job per process, gstore.cpp:235-333).
//
// Several graphs stay resident; queries (BFS, personalised PageRank, WCC,
// PageRank) are submitted at any time and served by a pool of worker threads,
// each with its own HIP stream and one workspace per graph, so the work of
// different queries overlaps on the device.  A worker that comes free takes
// the oldest pending query; if it is a BFS it also takes every other pending
// BFS on the same graph, up to 64, and runs them as the bit lanes of one
// batched sweep -- queries that arrived while earlier batches were on the GPU
// join the next batch.  A batch forms only when enough BFS of one graph are
// waiting for the batched sweep to beat running them one root at a time: that
// count is measured per graph when the graph is added (one direction-optimizing
// single-root BFS against batches of 1 and 64 lanes).  Below it, a worker takes
// one BFS and runs the single-root direction-optimizing BFS, and other workers
// take the others concurrently.  A query is admitted only when the workspace
// memory it needs fits the HBM budget next to the memory already held; idle
// workspaces give their memory back when a query needs room.  Graph-level
// structures (the neighbour-list index and the blocked PageRank copy) are
// built when a graph is added, never inside a query.
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "gs_internal.h"

namespace gs {

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Query {
    uint64_t id;
    int graph, kind;
    uint64_t arg;
    void *out;
    double t_submit;
};

struct Worker;

// Internal query kind: a BFS run alone by the single-root direction-optimizing path.
enum { KIND_BFS1 = 100 };

// What the server measured for one graph when it was added.
struct GraphPolicy {
    int batch_min = 2;            // pending BFS from which a batched sweep runs
    double single_ms = 0, batch1_ms = 0, batch64_ms = 0;
    uint64_t device_bytes = 0;    // the graph's own device memory (grid, index, blocked copy)
};

}  // namespace gs

struct gs_server {
    uint64_t budget = 0;       // bytes of workspace memory allowed
    uint64_t held = 0;         // bytes held by all workspaces
    int max_batch = 64;
    int ppr_iters = 20, pr_iters = 20;
    std::vector<gs_graph *> graphs;
    std::vector<gs::GraphPolicy> policy;
    std::string last_err;      // message of the last failed query
    std::deque<gs::Query> pending;
    std::vector<gs_query_result> done;
    std::vector<std::unique_ptr<gs::Worker>> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    uint64_t next_id = 1, inflight = 0;
    bool stop = false;
    double t0 = 0;
};

namespace gs {

struct Worker {
    gs_server *srv;
    int index;
    std::map<int, std::unique_ptr<gs_ws>> ws;  // per graph
    std::thread th;
    bool busy = false;
};

// Workspace bytes a query of `kind` needs on graph g: exactly what run_*
// allocates (plus the 128-byte counter block every workspace has).
// BFS on an undirected resident grid runs over the neighbour-list index (lane
// masks, frontier queue and its prefix; depths only when a query asks for
// them); otherwise over the tiles (lane masks, lane-major depths, row bits).
// The index itself belongs to the graph and is not counted here.
static bool bfs_by_index(const gs_graph *g) { return g->batch_method >= 1 && !g->directed && !g->ss; }

static uint64_t need_bytes(const gs_graph *g, int kind, int lanes, bool keep_depth) {
    const uint64_t n = g->n, rowact = (g->p_s + 31) / 32 * 4, base = 16 * 8;
    switch (kind) {
        case GS_Q_BFS:
            if (bfs_by_index(g))
                return base + 24 * n + 64 * 8 * 2 + 8 * n + 8 * (n + 1) + 8 * (n / 1024 + 1) +
                       (keep_depth ? (uint64_t)lanes * n + 64 * n : 0) +
                       (g->batch_method == 2 ? 8 * (18 + 64 * 256) : 0);   // device-driven levels: control words
            return base + 24 * n + (uint64_t)lanes * n + rowact + 64 * 8 * 2;
        case KIND_BFS1: return base + do_bfs_bytes(g);
        case GS_Q_PPR: return base + 24 * n;  // prior, acc, factor, rank (f32) and the FP64 sums
        case GS_Q_PAGERANK: return base + 12 * n;
        case GS_Q_WCC: return base + (g->wcc_method == 1 && !g->directed && !g->ss ? 12 * n + 4 * 1024 : 4 * n);
        default: return base;
    }
}

static uint64_t ws_bytes(const gs_ws *w) { return w->device_bytes(); }

// Bytes of the buffers a query of `kind` reuses.
static uint64_t kind_bytes(const gs_ws *w, int kind) {
    uint64_t b = w->counters.bytes();
    switch (kind) {
        case GS_Q_BFS:
            return b + w->ms_vis.bytes() + w->ms_front.bytes() + w->ms_next.bytes() + w->ms_depth.bytes() +
                   w->ms_depth_vm.bytes() + w->rowact.bytes() + w->ms_roots.bytes() + w->ms_counts.bytes() +
                   w->queue.bytes() + w->qdeg.bytes() + w->qoff.bytes() + w->qsum.bytes() + w->ms_ctl.bytes();
        case KIND_BFS1:
            return b + w->depth.bytes() + w->depth_r.bytes() + w->bm_vis.bytes() + w->bm_front.bytes() +
                   w->bm_next.bytes() + w->rowact.bytes() + w->queue.bytes() + w->qdeg.bytes() + w->qoff.bytes() +
                   w->qsum.bytes() + w->do_ctl.bytes();
        case GS_Q_PPR:
            return b + w->pr_prior.bytes() + w->pr_acc.bytes() + w->pr_fac.bytes() + w->ppr_rank.bytes() +
                   w->ppr_acc.bytes();
        case GS_Q_PAGERANK: return b + w->pr_prior.bytes() + w->pr_acc.bytes() + w->pr_fac.bytes();
        case GS_Q_WCC: return b + w->wcc_parent.bytes() + w->wcc_label.bytes() + w->wcc_cmin.bytes() + w->wcc_samp.bytes();
        default: return b;
    }
}

// Release the buffers a query of `kind` does not use.
static void trim(gs_ws *w, int kind) {
    const bool bfs = kind == GS_Q_BFS, bfs1 = kind == KIND_BFS1;
    if (!bfs) {
        w->ms_vis.release(); w->ms_front.release(); w->ms_next.release(); w->ms_depth.release();
        w->ms_depth_vm.release(); w->ms_roots.release(); w->ms_counts.release(); w->ms_ctl.release();
    }
    if (!bfs && !bfs1) {
        w->queue.release(); w->qdeg.release(); w->qoff.release(); w->qsum.release(); w->rowact.release();
    }
    if (!bfs1) {
        w->depth.release(); w->depth_r.release(); w->bm_vis.release(); w->bm_front.release(); w->bm_next.release();
        w->do_ctl.release();
    }
    if (kind != GS_Q_PPR && kind != GS_Q_PAGERANK) {
        w->pr_prior.release(); w->pr_acc.release(); w->pr_fac.release();
        w->bin_cnt.release(); w->bin_off.release(); w->bin_cur.release(); w->bin_uoff.release(); w->bin_pc.release(); w->bin_pr.release();
        w->bin_o.release(); w->bin_v.release();
    }
    if (kind != GS_Q_PPR) {
        w->ppr_rank.release();
        w->ppr_acc.release();
    }
    if (kind != GS_Q_WCC) {
        w->wcc_parent.release();
        w->wcc_label.release();
        w->wcc_cmin.release();
        w->wcc_samp.release();
    }
    w->kc_deg.release();
}

// Free idle workspaces (not the caller's) until `want` more bytes fit.  Lock held.
static void make_room(gs_server *s, Worker *self, uint64_t want) {
    for (auto &wk : s->workers) {
        if (s->held + want <= s->budget) return;
        if (wk.get() == self || wk->busy) continue;
        for (auto &kv : wk->ws) {
            uint64_t b = ws_bytes(kv.second.get());
            kv.second.reset();
            s->held -= std::min(s->held, b);
        }
        wk->ws.clear();
    }
}

static void worker_main(Worker *wk) {
    gs_server *s = wk->srv;
    std::vector<uint64_t> roots;
    std::vector<uint64_t> counts;
    for (;;) {
        std::vector<Query> batch;
        int run_kind = 0;
        uint64_t need = 0, have_all = 0;
        gs_ws *w = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                if (!s->pending.empty()) {
                    // oldest query first; a BFS collects the pending BFS of its graph
                    const Query &q = s->pending.front();
                    gs_graph *g = s->graphs[q.graph];
                    int lanes = 1, ikind = q.kind;
                    bool keep = false;  // some query of the batch wants its depths
                    if (q.kind == GS_Q_BFS) {
                        lanes = 0;
                        for (auto &p : s->pending)
                            if (p.kind == GS_Q_BFS && p.graph == q.graph && lanes < s->max_batch) ++lanes;
                        if (bfs_by_index(g) && lanes < s->policy[q.graph].batch_min) {
                            // too few waiting for a batch to pay: this one alone, single-root path
                            lanes = 1;
                            ikind = KIND_BFS1;
                        } else {
                            for (auto &p : s->pending)
                                if (p.kind == GS_Q_BFS && p.graph == q.graph) keep = keep || p.out;
                            while (lanes > 1 && need_bytes(g, GS_Q_BFS, lanes, keep) > s->budget) --lanes;
                            keep = false;
                            int k = 0;
                            for (auto &p : s->pending)
                                if (p.kind == GS_Q_BFS && p.graph == q.graph && k < lanes) {
                                    ++k;
                                    keep = keep || p.out;
                                }
                        }
                    }
                    auto it = wk->ws.find(q.graph);
                    const gs_ws *cur = it != wk->ws.end() ? it->second.get() : nullptr;
                    have_all = cur ? ws_bytes(cur) : 0;
                    uint64_t have = cur ? kind_bytes(cur, ikind) : 0;
                    uint64_t want = need_bytes(g, ikind, lanes, keep);
                    need = want > have ? want - have : 0;
                    if (s->held + need > s->budget) make_room(s, wk, need);
                    if (s->held + need <= s->budget) {
                        s->held += need;
                        int gid = q.graph;
                        if (ikind == GS_Q_BFS) {
                            for (auto p = s->pending.begin(); p != s->pending.end() && (int)batch.size() < lanes;) {
                                if (p->kind == GS_Q_BFS && p->graph == gid) {
                                    batch.push_back(*p);
                                    p = s->pending.erase(p);
                                } else {
                                    ++p;
                                }
                            }
                        } else {
                            batch.push_back(q);
                            s->pending.pop_front();
                        }
                        run_kind = ikind;
                        ++s->inflight;
                        wk->busy = true;
                        break;
                    }
                }
                s->cv_work.wait(lk);
            }
        }
        const int gid = batch[0].graph, kind = batch[0].kind;
        const gs_graph *g = s->graphs[gid];
        std::string err;
        auto &slot = wk->ws[gid];
        int status = GS_OK;
        if (!slot) {
            gs_ws *nw = nullptr;
            status = ws_create(&nw);
            slot.reset(nw);
        }
        w = slot.get();
        if (w) trim(w, run_kind);
        double t_start = now_ms();
        std::vector<gs_query_result> res(batch.size());
        for (size_t i = 0; i < batch.size(); ++i) {
            res[i] = gs_query_result{};
            res[i].id = batch[i].id;
            res[i].graph = gid;
            res[i].kind = kind;
            res[i].arg = batch[i].arg;
            res[i].batch_size = (int)batch.size();
            res[i].t_submit_ms = batch[i].t_submit - s->t0;
            res[i].t_start_ms = t_start - s->t0;
        }
        if (status == GS_OK) {
            switch (run_kind) {
                case KIND_BFS1: {
                    counts.assign(256, 0);
                    int levels = 0;
                    status = run_bfs_do(g, w, batch[0].arg, (uint8_t *)batch[0].out, counts.data(), 256, &levels);
                    if (status == GS_OK) {
                        uint64_t reached = 1;
                        for (int l = 0; l < levels && l < 256; ++l) reached += counts[l];
                        res[0].value = reached;
                        res[0].aux = (uint64_t)levels;
                    }
                    break;
                }
                case GS_Q_BFS: {
                    const int nr = (int)batch.size();
                    roots.resize(nr);
                    for (int i = 0; i < nr; ++i) roots[i] = batch[i].arg;
                    counts.assign((size_t)nr * 256, 0);
                    int levels = 0;
                    bool keep = false;
                    for (int i = 0; i < nr; ++i) keep = keep || batch[i].out;
                    status = run_bfs_batch(g, w, roots.data(), nr, nullptr, counts.data(), 256, &levels, keep);
                    for (int i = 0; i < nr && status == GS_OK; ++i) {
                        uint64_t reached = 1;
                        for (int l = 0; l < levels && l < 256; ++l) reached += counts[(size_t)i * 256 + l];
                        res[i].value = reached;
                        res[i].aux = (uint64_t)levels;
                        if (batch[i].out &&
                            (hipMemcpyAsync(batch[i].out, w->ms_depth.ptr + (size_t)i * g->n, g->n,
                                            hipMemcpyDeviceToHost, w->stream) != hipSuccess ||
                             hipStreamSynchronize(w->stream) != hipSuccess))
                            status = GS_ERR_DEVICE;
                    }
                    break;
                }
                case GS_Q_PPR:
                    status = run_ppr(g, w, batch[0].arg, s->ppr_iters, (float *)batch[0].out);
                    break;
                case GS_Q_PAGERANK:
                    status = run_pagerank(g, w, s->pr_iters, (int)batch[0].arg, (float *)batch[0].out);
                    break;
                case GS_Q_WCC: {
                    uint64_t nc = 0;
                    status = run_wcc(g, w, (uint32_t *)batch[0].out, &nc);
                    res[0].value = nc;
                    break;
                }
                default:
                    status = GS_ERR_ARG;
            }
        }
        if (status != GS_OK) err = error_text();
        double t_done = now_ms();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            // replace the reservation and the workspace's old size by what it holds now
            uint64_t after = w ? ws_bytes(w) : 0;
            uint64_t drop = std::min(s->held, need + have_all);
            s->held = s->held - drop + after;
            if (status != GS_OK) s->last_err = err;
            for (auto &r : res) {
                r.status = status;
                r.t_done_ms = t_done - s->t0;
                s->done.push_back(r);
            }
            --s->inflight;
            wk->busy = false;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

// Per-graph BFS batching threshold: time the single-root direction-optimizing
// BFS (mean of 8 roots) and batched sweeps of 1 and 64 lanes on a scratch
// workspace, from roots at evenly spread degree ranks, and set batch_min to
// the smallest k whose batched cost (interpolated between 1 and 64 lanes)
// is below k single roots (65: never batch).
static int calibrate_bfs(gs_graph *g, GraphPolicy *pol) {
    if (g->nzn == 0) return GS_OK;
    gs_ws *wp = nullptr;
    GS_TRY(ws_create(&wp));
    std::unique_ptr<gs_ws> w(wp);
    std::vector<uint32_t> ids(64);
    for (int i = 0; i < 64; ++i)
        GS_HIP(hipMemcpy(&ids[i], g->old_id.ptr + (uint64_t)i * g->nzn / 64, 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> roots(ids.begin(), ids.end());
    std::vector<uint64_t> counts(64 * 256);
    int levels = 0;
    auto single = [&](uint64_t r) { return run_bfs_do(g, w.get(), r, nullptr, nullptr, 0, nullptr); };
    auto batch = [&](int nr) {
        return run_bfs_batch(g, w.get(), roots.data(), nr, nullptr, counts.data(), 256, &levels, false);
    };
    GS_TRY(single(roots[0]));   // first calls capture their level graphs
    GS_TRY(batch(1));
    GS_TRY(batch(64));
    double t = now_ms();
    for (int i = 0; i < 8; ++i) GS_TRY(single(roots[i * 8 + 4]));
    pol->single_ms = (now_ms() - t) / 8;
    t = now_ms();
    GS_TRY(batch(1));
    pol->batch1_ms = now_ms() - t;
    t = now_ms();
    GS_TRY(batch(64));
    pol->batch64_ms = now_ms() - t;
    pol->batch_min = 65;
    for (int k = 2; k <= 64; ++k) {
        const double tb = pol->batch1_ms + (pol->batch64_ms - pol->batch1_ms) * (k - 1) / 63.0;
        if (tb < k * pol->single_ms) { pol->batch_min = k; break; }
    }
    return GS_OK;
}

}  // namespace gs

using namespace gs;

extern "C" {

int gs_server_create(uint64_t budget, int workers, gs_server **out) {
    if (!out || workers < 1 || workers > 16) { set_error("gs_server_create: 1..16 workers"); return GS_ERR_ARG; }
    std::unique_ptr<gs_server> s(new gs_server());
    s->budget = budget ? budget : ~0ull;
    s->t0 = now_ms();
    for (int i = 0; i < workers; ++i) {
        s->workers.emplace_back(new Worker());
        s->workers.back()->srv = s.get();
        s->workers.back()->index = i;
    }
    for (auto &wk : s->workers) wk->th = std::thread(worker_main, wk.get());
    *out = s.release();
    return GS_OK;
}

int gs_server_add_graph(gs_server *s, gs_graph *g, int *graph_id) {
    if (!s || !g || !graph_id) { set_error("null argument"); return GS_ERR_ARG; }
    if (g->ss) { set_error("the server keeps graphs resident; this one is in semi-external mode"); return GS_ERR_UNSUPPORTED; }
    GraphPolicy pol;
    // the neighbour-list index serves BFS, PPR and WCC queries on undirected
    // graphs, the blocked copy PageRank and PPR: build both now rather than
    // inside the first query that needs them
    if (!g->directed) {
        GS_TRY(build_adjacency(g));
        if (g->pr_method == 2) GS_TRY(build_blocked(g));
        if (bfs_by_index(g)) GS_TRY(calibrate_bfs(g, &pol));
    }
    gs_info info;
    GS_TRY(gs_get_info(g, &info));
    pol.device_bytes = info.device_bytes;
    std::lock_guard<std::mutex> lk(s->mu);
    s->graphs.push_back(g);
    s->policy.push_back(pol);
    *graph_id = (int)s->graphs.size() - 1;
    return GS_OK;
}

int gs_server_bfs_policy(gs_server *s, int graph, int *batch_min, double *single_ms, double *batch1_ms,
                         double *batch64_ms) {
    if (!s) { set_error("null server"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph < 0 || graph >= (int)s->graphs.size()) { set_error("unknown graph %d", graph); return GS_ERR_ARG; }
    const GraphPolicy &p = s->policy[graph];
    if (batch_min) *batch_min = p.batch_min;
    if (single_ms) *single_ms = p.single_ms;
    if (batch1_ms) *batch1_ms = p.batch1_ms;
    if (batch64_ms) *batch64_ms = p.batch64_ms;
    return GS_OK;
}

int gs_server_set_batch_min(gs_server *s, int graph, int batch_min) {
    if (!s || batch_min < 1 || batch_min > 65) { set_error("gs_server_set_batch_min: 1..65"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph < 0 || graph >= (int)s->graphs.size()) { set_error("unknown graph %d", graph); return GS_ERR_ARG; }
    s->policy[graph].batch_min = batch_min;
    return GS_OK;
}

int gs_server_memory(gs_server *s, uint64_t *workspace_bytes, uint64_t *graph_bytes) {
    if (!s) { set_error("null server"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    uint64_t gb = 0;
    for (auto &p : s->policy) gb += p.device_bytes;
    if (workspace_bytes) *workspace_bytes = s->held;
    if (graph_bytes) *graph_bytes = gb;
    return GS_OK;
}

size_t gs_server_last_error(gs_server *s, char *buf, size_t len) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    if (buf && len) {
        size_t k = std::min(len - 1, s->last_err.size());
        memcpy(buf, s->last_err.data(), k);
        buf[k] = 0;
    }
    return s->last_err.size();
}

int gs_server_set(gs_server *s, int max_batch, int ppr_iters, int pr_iters) {
    if (!s || max_batch < 1 || max_batch > 64 || ppr_iters < 1 || pr_iters < 0) {
        set_error("gs_server_set: bad arguments");
        return GS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->max_batch = max_batch;
    s->ppr_iters = ppr_iters;
    s->pr_iters = pr_iters;
    return GS_OK;
}

int gs_server_submit(gs_server *s, int graph, int kind, uint64_t arg, void *out, uint64_t *id) {
    if (!s) { set_error("null server"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (graph < 0 || graph >= (int)s->graphs.size()) { set_error("unknown graph %d", graph); return GS_ERR_ARG; }
    const gs_graph *g = s->graphs[graph];
    if (kind < GS_Q_BFS || kind > GS_Q_PAGERANK) { set_error("unknown query kind %d", kind); return GS_ERR_ARG; }
    if ((kind == GS_Q_BFS || kind == GS_Q_PPR) && arg >= g->n) {
        set_error("query source %llu >= N", (unsigned long long)arg);
        return GS_ERR_ARG;
    }
    if (need_bytes(g, kind, 1, out != nullptr) > s->budget) {
        set_error("query needs more workspace memory than the server budget");
        return GS_ERR_NOMEM;
    }
    Query q{s->next_id++, graph, kind, arg, out, now_ms()};
    s->pending.push_back(q);
    if (id) *id = q.id;
    s->cv_work.notify_one();
    return GS_OK;
}

int gs_server_poll(gs_server *s, gs_query_result *out, int max, int wait_ms, int *n) {
    if (!s || !n || (max > 0 && !out)) { set_error("null argument"); return GS_ERR_ARG; }
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    int k = std::min<int>(max, (int)s->done.size());
    for (int i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return GS_OK;
}

int gs_server_drain(gs_server *s) {
    if (!s) { set_error("null server"); return GS_ERR_ARG; }
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
    return GS_OK;
}

int gs_server_info(gs_server *s, uint64_t *held_bytes, uint64_t *pending, uint64_t *inflight) {
    if (!s) { set_error("null server"); return GS_ERR_ARG; }
    std::lock_guard<std::mutex> lk(s->mu);
    if (held_bytes) *held_bytes = s->held;
    if (pending) *pending = s->pending.size();
    if (inflight) *inflight = s->inflight;
    return GS_OK;
}

void gs_server_destroy(gs_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &wk : s->workers)
        if (wk->th.joinable()) wk->th.join();
    delete s;
}

}  // extern "C"
