// gs_stream.hip -- semi-external mode: the tile grid in pinned host memory,
// a resident prefix in HBM, and the rest streamed per sweep through two
// device segment buffers (the copy/sweep pipelining lives in gs::sweep).
//
// The reference streams 256 MB segments from an O_DIRECT file with libaio,
// double-buffered (seg1 is processed while seg2 loads, gstore.cpp:1770-1899,
// 3623-3797), and keeps the first cache_size bytes of each pass cached so the
// next iteration processes them first without IO (gstore.cpp:1491-1768).  Here
// the host side of that pipeline is pinned DRAM, the copy engine is
// hipMemcpyAsync on a side stream, and the cache is the resident prefix of the
// grid: chunks [0, res_chunks) never leave HBM.
#include <string.h>

#include <memory>

#include "gs_internal.h"

namespace gs {

StreamState::~StreamState() {
    if (copy) (void)hipStreamSynchronize(copy);
    for (int i = 0; i < 2; ++i) {
        if (ev_copied[i]) (void)hipEventDestroy(ev_copied[i]);
        if (ev_done[i]) (void)hipEventDestroy(ev_done[i]);
    }
    if (copy) (void)hipStreamDestroy(copy);
    if (h_grid) (void)hipHostFree(h_grid);
}

// Build the streaming state for graph g whose full grid is in `h_grid`
// (pinned, ownership taken).  Uploads the resident prefix.
int stream_setup(gs_graph *g, uint32_t *h_grid, uint64_t budget, uint64_t seg_bytes) {
    std::unique_ptr<StreamState> ss(new StreamState());  // takes h_grid only on success
    const uint64_t chunk_bytes = (uint64_t)g->chunk * 4;
    if (!seg_bytes) seg_bytes = std::min<uint64_t>(512ull << 20, budget / 8);  // default: 512 MiB or budget/8
    ss->seg_chunks = std::max<uint64_t>(1, seg_bytes / chunk_bytes);
    uint64_t two = 2 * ss->seg_chunks * chunk_bytes;
    if (budget < two + chunk_bytes) {
        set_error("device budget %llu bytes is below two segments (%llu bytes) plus one chunk",
                  (unsigned long long)budget, (unsigned long long)two);
        return GS_ERR_ARG;
    }
    ss->res_chunks = std::min<uint64_t>(g->nchunk, (budget - two) / chunk_bytes);
    uint64_t res_edges = std::min<uint64_t>(g->E, ss->res_chunks * g->chunk);
    GS_TRY(g->grid.alloc(res_edges));
    if (res_edges) GS_HIP(hipMemcpy(g->grid.ptr, h_grid, res_edges * 4, hipMemcpyHostToDevice));
    GS_TRY(ss->seg[0].alloc(ss->seg_chunks * g->chunk));
    GS_TRY(ss->seg[1].alloc(ss->seg_chunks * g->chunk));
    GS_HIP(hipStreamCreateWithFlags(&ss->copy, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        GS_HIP(hipEventCreateWithFlags(&ss->ev_copied[i], hipEventDisableTiming));
        GS_HIP(hipEventCreateWithFlags(&ss->ev_done[i], hipEventDisableTiming));
        GS_HIP(hipEventRecord(ss->ev_done[i], ss->copy));  // first waits complete at once
    }
    GS_HIP(hipStreamSynchronize(ss->copy));
    void *dev = nullptr;
    GS_HIP(hipHostGetDevicePointer(&dev, h_grid, 0));
    ss->h_grid_dev = static_cast<const uint32_t *>(dev);
    ss->h_grid = h_grid;
    g->ss = ss.release();
    return GS_OK;
}

}  // namespace gs

using namespace gs;

extern "C" int gs_set_device_budget(gs_graph *g, uint64_t budget, uint64_t seg_bytes) {
    if (!g) { set_error("null graph"); return GS_ERR_ARG; }
    if (g->ss) { set_error("graph is already in semi-external mode"); return GS_ERR_ARG; }
    if (budget == 0 || g->E * 4 <= budget) return GS_OK;  // fits: stays resident
    GS_HIP(hipDeviceSynchronize());
    uint32_t *h = nullptr;
    GS_HIP(hipHostMalloc((void **)&h, std::max<uint64_t>(g->E, 1) * 4, hipHostMallocDefault));
    if (hipMemcpy(h, g->grid.ptr, g->E * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipHostFree(h);
        set_error("copying the grid to host memory failed");
        return GS_ERR_DEVICE;
    }
    g->grid.release();
    int st = stream_setup(g, h, budget, seg_bytes);
    if (st != GS_OK && !g->ss) {
        // restore the resident grid
        if (g->grid.alloc(g->E) == GS_OK) (void)hipMemcpy(g->grid.ptr, h, g->E * 4, hipMemcpyHostToDevice);
        (void)hipHostFree(h);
    }
    return st;
}
