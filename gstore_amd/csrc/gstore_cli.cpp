// gstore_cli.cpp -- gstoreu / gstored command line over the C ABI.
//
// Keeps the reference's flags and job numbers (grid::init, gstore.cpp:132-348):
//   -s S      scale, N = 2^S           -v N   vertex count
//   -i PATH   input prefix (jobs) or edge-list file (-c)
//   -o PATH   output prefix (-c, -g)   -j J   job   -a A   job argument
//   -c C      convert (1: binary u32 edge list -> tile files)
//   -m MB     device budget for the tile grid in MiB (the reference's -m
//             sizes its streaming buffer)
// Jobs (code order, gstore.cpp:235-333): 0/10/20 BFS (-a root), 1/11/21
// PageRank (-a iterations), 2/12/22 k-core (-a k), 3/13/23 WCC, 30
// asynchronous BFS (-a root, default 1), 4 single colour propagation (-a
// pivot, default 1).  The x, 1x and 2x variants all run the resident HIP
// engine.  Conversions: -c 1 binary u32 pairs -> tile files, -c 4 text edge
// list -> binary pairs in EDGES.edge (util.cpp:58-168).  Defaults follow the
// reference: root 0 (-j 0; -j 10/20 use root 1, gstore.cpp:2634), 5
// iterations, k = 10.
// Methods: on an undirected grid resident in HBM the jobs run the engine's
// fastest methods (direction-optimizing BFS, blocked PageRank, index WCC and
// k-core), whose results equal the tile methods'; -b 0 selects the tile-sweep
// BFS of the reference.  These methods first build a neighbour-list index
// (and, for PageRank, its blocked copy) once per process; like the load, the
// build is outside the job's timer and printed on its own line
// ("Index build time").
// Extensions: -g S generates the Kronecker graph of scale S on the device
// (edge factor -e, default 16; seeds 1, 2) and writes it to -o; -r 1 selects
// the float (intended) PageRank degree factor; -p FILE writes the job's
// per-vertex result as raw binary.
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <chrono>
#include <string>
#include <vector>

#include "gstore_amd.h"

#ifndef GS_CLI_DIRECTED
#define GS_CLI_DIRECTED 0
#endif

static void die(const char *what) {
    char buf[1024];
    gs_last_error(buf, sizeof buf);
    fprintf(stderr, "%s failed: %s\n", what, buf);
    exit(EXIT_FAILURE);
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void write_raw(const char *path, const void *p, size_t bytes) {
    if (!path) return;
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(EXIT_FAILURE);
    }
    fclose(f);
}

int main(int argc, char **argv) {
    uint64_t n = 0;
    int job = 0, conv = 0, gen_scale = 0, ef = 16, pr_mode = GS_PR_COMPAT, bfs_method = 1;
    long arg = -1;
    uint64_t budget_mb = 0;
    std::string in, out;
    const char *result_path = nullptr;
    int o;
    while ((o = getopt(argc, argv, "s:o:hi:j:c:m:v:a:g:e:r:p:b:")) != -1) {
        switch (o) {
            case 's': n = 1ull << atoi(optarg); break;
            case 'v': n = strtoull(optarg, nullptr, 10); break;
            case 'i': in = optarg; break;
            case 'o': out = optarg; break;
            case 'j': job = atoi(optarg); break;
            case 'c': conv = atoi(optarg); break;
            case 'm': budget_mb = strtoull(optarg, nullptr, 10); break;
            case 'a': arg = atol(optarg); break;
            case 'g': gen_scale = atoi(optarg); break;
            case 'e': ef = atoi(optarg); break;
            case 'r': pr_mode = atoi(optarg); break;
            case 'p': result_path = optarg; break;
            case 'b': bfs_method = atoi(optarg); break;
            case 'h':
            default:
                fprintf(stderr,
                        "usage: %s -s S|-v N -i PREFIX -j {0,1,2,3,4,10..13,20..23,30} [-a ARG] [-m MB] [-r MODE] [-b 0|1] [-p OUT]\n"
                        "       %s -s S|-v N -i EDGES -o PREFIX -c 1    |    %s -i TEXT -c 4\n"
                        "       %s -g S [-e EF] -o PREFIX\n",
                        argv[0], argv[0], argv[0], argv[0]);
                return o == 'h' ? 0 : 1;
        }
    }
    const int directed = GS_CLI_DIRECTED;
    if (gen_scale) {
        if (out.empty()) { fprintf(stderr, "-g needs -o PREFIX\n"); return 1; }
        gs_graph *g;
        double t = now_s();
        if (gs_open_kron(gen_scale, ef, 1, 2, directed, &g)) die("gs_open_kron");
        gs_info info;
        gs_get_info(g, &info);
        printf("Total edges = %llu\n", (unsigned long long)info.edge_count);
        if (gs_save(g, out.c_str())) die("gs_save");
        printf("GStore Time = %f\n", now_s() - t);
        gs_close(g);
        return 0;
    }
    uint64_t p = (n >> 24) + ((n & 0xFFFFFF) != 0);
    if (conv == 4) {   // text -> binary needs no vertex count (text_to_bin, util.cpp:58-168)
        printf("Total Physcial Group Count (in One Direction) = %llu\n", (unsigned long long)p);
        printf("Total Tiles Count (in One Direction) = %llu\n", (unsigned long long)(256 * p));
        uint64_t pairs = 0, tokens = 0;
        uint32_t vmin = 0, vmax = 0;
        if (gs_text_to_bin(in.c_str(), out.empty() ? nullptr : out.c_str(), &pairs, &vmin, &vmax, &tokens))
            die("gs_text_to_bin");
        fprintf(stderr, "max vertex id: %u\nmin vertex id: %u\nedge count: %llu\nvert count: %llu\n", vmax, vmin,
                (unsigned long long)pairs, (unsigned long long)((uint64_t)(uint32_t)(vmax - vmin) + 1));
        return 0;
    }
    if (n == 0) { fprintf(stderr, "give the vertex count with -s or -v\n"); return 1; }
    printf("Total Physcial Group Count (in One Direction) = %llu\n", (unsigned long long)p);
    printf("Total Tiles Count (in One Direction) = %llu\n", (unsigned long long)(256 * p));
    if (conv) {
        if (conv != 1) { fprintf(stderr, "-c %d: -c 1 and -c 4 are provided\n", conv); return 1; }
        struct stat st;
        if (stat(in.c_str(), &st) != 0 || st.st_size % 8) { fprintf(stderr, "bad edge file %s\n", in.c_str()); return 1; }
        std::vector<uint32_t> edges(st.st_size / 4);
        FILE *f = fopen(in.c_str(), "rb");
        if (!f || fread(edges.data(), 1, st.st_size, f) != (size_t)st.st_size) { fprintf(stderr, "read failed\n"); return 1; }
        fclose(f);
        double t = now_s();
        if (gs_convert(edges.data(), edges.size() / 2, n, directed, out.c_str())) die("gs_convert");
        printf("GStore Time = %f\n", now_s() - t);
        return 0;
    }
    gs_graph *g;
    if (gs_open(in.c_str(), n, directed, budget_mb << 20, &g)) die("gs_open");
    gs_info info;
    if (gs_get_info(g, &info)) die("gs_get_info");
    if (!directed && info.resident && bfs_method == 1 && gs_set_bfs_mode(g, 1, 0, 0)) die("gs_set_bfs_mode");
    gs_stats st;
    if (job == 30) {   // grid::bfs2 (gstore.cpp:1900-2042), root -a, default 1
        const uint64_t root = arg == -1 ? 1 : (uint64_t)arg;
        printf("Root = %llu\n", (unsigned long long)root);
        std::vector<uint8_t> depth(n);
        std::vector<uint64_t> counts(1 << 16);
        int sweeps = 0;
        if (gs_bfs_async(g, root, depth.data(), counts.data(), (int)counts.size(), &sweeps)) die("gs_bfs_async");
        gs_get_stats(g, &st);
        for (int l = 0; l < sweeps && l < (int)counts.size(); ++l)
            printf("level: %d frontier count: %llu\n", l + 1, (unsigned long long)counts[l]);
        printf("Half BFS time = %f\n", st.ms / 1e3);
        write_raw(result_path, depth.data(), n);
        gs_close(g);
        return 0;
    }
    if (job == 4) {    // grid::traverse (gstore.cpp:2169-2177), pivot -a, default 1, colour 0
        const uint64_t pivot = arg == -1 ? 1 : (uint64_t)arg;
        std::vector<uint32_t> cid(n);
        std::vector<uint64_t> counts(1 << 16);
        int sweeps = 0;
        uint64_t colored = 0;
        if (gs_traverse(g, pivot, 0, cid.data(), counts.data(), (int)counts.size(), &sweeps, &colored))
            die("gs_traverse");
        gs_get_stats(g, &st);
        for (int l = 0; l < sweeps && l < (int)counts.size(); ++l)
            printf("level: %d frontier count: %llu\n\n", l, (unsigned long long)counts[l]);
        printf("Single color propagation time = %f\n", st.ms / 1e3);
        write_raw(result_path, cid.data(), n * 4);
        gs_close(g);
        return 0;
    }
    const int base = job % 10;
    if (base > 3 || (job != base && job != base + 10 && job != base + 20)) {
        printf("Wrong value for -j argument\n");
        gs_close(g);
        return 1;
    }
    switch (base) {
        case 0: {
            uint64_t root = arg == -1 ? (job == 0 ? 0 : 1) : (uint64_t)arg;
            if (job != 0) root = 1;  // -j 10/20 hard-code root 1 (gstore.cpp:2634)
            printf("Root = %llu\n", (unsigned long long)root);
            std::vector<uint8_t> depth(n);
            std::vector<uint64_t> counts(256);
            int levels = 0;
            if (gs_bfs(g, root, depth.data(), counts.data(), 256, &levels)) die("gs_bfs");
            gs_get_stats(g, &st);
            for (int l = 0; l < levels && l < 256; ++l)
                printf("level: %d frontier count: %llu\n", l + 1, (unsigned long long)counts[l]);
            printf("Half BFS time = %f\n", st.ms / 1e3);
            write_raw(result_path, depth.data(), n);
            break;
        }
        case 1: {
            int iters = arg == -1 ? 5 : (int)arg;
            std::vector<float> rank(n);
            if (gs_pagerank(g, iters, pr_mode, rank.data())) die("gs_pagerank");
            gs_get_stats(g, &st);
            printf("Page-rank%s time = %f\n", job == 1 ? "" : "_mmap", st.ms / 1e3);
            write_raw(result_path, rank.data(), n * 4);
            break;
        }
        case 2: {
            int k = arg == -1 ? 10 : (int)arg;
            std::vector<int32_t> deg(n);
            uint64_t ne0 = 0, pos = 0;
            int rounds = 0;
            int s = gs_kcore(g, k, deg.data(), &ne0, &pos, &rounds);
            if (s == GS_ERR_UNSUPPORTED) { printf("Kcore not applicable for directed graphs\n"); break; }
            if (s) die("gs_kcore");
            gs_get_stats(g, &st);
            printf(">= %dvertices count = %llu\n", k, (unsigned long long)ne0);
            printf("Kcore time = %f\n", st.ms / 1e3);
            write_raw(result_path, deg.data(), n * 4);
            break;
        }
        case 3: {
            std::vector<uint32_t> label(n);
            uint64_t nc = 0;
            if (gs_wcc(g, label.data(), &nc)) die("gs_wcc");
            gs_get_stats(g, &st);
            printf("WCC count = %llu\n", (unsigned long long)nc);
            printf("WCC time = %f\n", st.ms / 1e3);
            write_raw(result_path, label.data(), n * 4);
            break;
        }
        default:
            printf("Wrong value for -j argument\n");
            gs_close(g);
            return 1;
    }
    if (gs_get_info(g, &info)) die("gs_get_info");
    if (info.index_ms > 0) printf("Index build time = %f\n", info.index_ms / 1e3);
    gs_close(g);
    return 0;
}
