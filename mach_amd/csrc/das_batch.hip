// Batched delay-and-sum, host side: many requests of different shapes, dtypes and interpolation types
// in one submission of persistent kernels (das_batch_tc1.hip, das_batch_tcn.hip), replayed from a
// captured hipGraph.
//
// A batch is a table of BatchItem descriptors (one per request), each carrying the request's
// DasParams, its kernel variant and the size of its work grid (voxel groups x frame chunks, the
// same blocks the plain kernel would launch). Work items are numbered consecutively across requests
// (`first_block` is the prefix sum). A persistent grid of workgroups pulls work items from one
// device counter until the batch is drained, so small requests fill the GPU together and a large
// request spreads over every CU. Every workgroup exits once the counter passes the total.
//
// Host side: two slots, each with pinned staging, a device descriptor table, a counter and an
// instantiated graph [H2D copy of the table -> counter reset kernel -> persistent kernels]. A slot is
// reused only after the event recorded behind its previous submission has completed, so the host
// never rewrites staging that a pending copy still reads.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <new>

#include "das_batch.h"
#include "das_batch_types.h"

namespace mach {

using namespace batch;

namespace {

int layout_index(DasLayout l) {
    for (int i = 0; i < kNumBatchLayouts; ++i)
        if (kBatchLayouts[i][0] == l.lanes_per_voxel && kBatchLayouts[i][1] == l.frames_per_lane) return i;
    return -1;
}

int ilog2(int v) {
    int r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}

}  // namespace

// The plain kernel's layout when it is compiled into the batched kernels; otherwise the compiled
// variant nearest to it in frames per voxel group (lanes x frames per lane, in octaves) whose frames
// per lane do not exceed the wanted ones (whose divisibility and alignment have been checked). Ties go
// to more frames per lane. Band layouts run as the general kernel's heuristic choice, lateral layouts
// as the 64-lane layout with their frames per lane.
DasLayout batch_layout(const DasParams &p, int store, DasLayout want) {
    if (want.lanes_per_voxel == 0) want = choose_layout(p, store);
    want.chunks_per_lane = 1;  // the batched kernels walk one frame group per lane
    want.voxels_per_lane = 1;  // and have no lateral groups: a lateral layout runs as its 64-lane layout
    if (layout_index(want) >= 0) return want;
    const int target = ilog2(want.lanes_per_voxel * want.frames_per_lane);
    int best = kNumBatchLayouts - 1, best_d = 1 << 30;
    for (int i = 0; i < kNumBatchLayouts; ++i) {
        const int lf = kBatchLayouts[i][0], k = kBatchLayouts[i][1];
        if (k > want.frames_per_lane) continue;
        const int d = ilog2(lf * k) > target ? ilog2(lf * k) - target : target - ilog2(lf * k);
        if (d < best_d || (d == best_d && k > kBatchLayouts[best][1])) {
            best = i;
            best_d = d;
        }
    }
    return DasLayout{kBatchLayouts[best][0], kBatchLayouts[best][1], 1};
}

bool batch_layout_supported(DasLayout l) { return layout_index(l) >= 0; }

struct BatchSlot {
    void *h_buf = nullptr;  // pinned: header + items
    void *d_buf = nullptr;  // device: header + items
    uint32_t *d_counter = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct BatchState {
    uint32_t max_items = 0;
    bool use_graph = true;
    size_t table_bytes = 0;
    int grid[kSections] = {0, 0};  // persistent grid per section: CUs x resident workgroups per CU
    int next = 0;
    BatchSlot slot[2];
};

namespace {

size_t items_offset() { return (sizeof(BatchHeader) + 255) & ~size_t(255); }

// `sections` is a bit mask of the sections to launch: both when captured into a graph (the graph is
// replayed for every batch; an empty section's workgroups exit at once), only non-empty ones otherwise.
hipError_t enqueue_direct(BatchState *b, BatchSlot &s, hipStream_t stream, unsigned sections) {
    hipError_t e = hipMemcpyAsync(s.d_buf, s.h_buf, b->table_bytes, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = launch_reset_counters(stream, s.d_counter);
    const BatchHeader *hdr = (const BatchHeader *)s.d_buf;
    const BatchItem *items = (const BatchItem *)((char *)s.d_buf + items_offset());
    if (e == hipSuccess && (sections & 1u)) e = launch_section0((unsigned)b->grid[0], stream, hdr, items, s.d_counter);
    if (e == hipSuccess && (sections & 2u)) e = launch_section1((unsigned)b->grid[1], stream, hdr, items, s.d_counter);
    return e;
}

hipError_t capture(BatchState *b, BatchSlot &s) {
    hipStream_t cap;
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t le = enqueue_direct(b, s, cap, 3u);
        e = hipStreamEndCapture(cap, &s.graph);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0);
    hipStreamDestroy(cap);
    return e;
}

}  // namespace

hipError_t batch_create(uint32_t max_items, bool use_graph, BatchState **out) {
    *out = nullptr;
    BatchState *b = new (std::nothrow) BatchState();
    if (b == nullptr) return hipErrorOutOfMemory;
    b->max_items = max_items;
    b->use_graph = use_graph;
    b->table_bytes = items_offset() + (size_t)max_items * sizeof(BatchItem);
    int dev = 0, cus = 0, per_cu[kSections] = {0, 0};
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = section_occupancy(0, &per_cu[0]);
    if (e == hipSuccess) e = section_occupancy(1, &per_cu[1]);
    for (int i = 0; i < kSections; ++i) b->grid[i] = (per_cu[i] > 0 ? per_cu[i] : 1) * (cus > 0 ? cus : 1);
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        BatchSlot &s = b->slot[i];
        e = hipHostMalloc(&s.h_buf, b->table_bytes, hipHostMallocDefault);
        if (e == hipSuccess) std::memset(s.h_buf, 0, b->table_bytes);
        if (e == hipSuccess) e = hipMalloc(&s.d_buf, b->table_bytes);
        if (e == hipSuccess) e = hipMalloc((void **)&s.d_counter, 256);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        batch_destroy(b);
        return e;
    }
    *out = b;
    return hipSuccess;
}

void batch_destroy(BatchState *b) {
    if (b == nullptr) return;
    for (BatchSlot &s : b->slot) {
        if (s.done) hipEventSynchronize(s.done);
        if (s.exec) hipGraphExecDestroy(s.exec);
        if (s.graph) hipGraphDestroy(s.graph);
        if (s.done) hipEventDestroy(s.done);
        if (s.d_counter) hipFree(s.d_counter);
        if (s.d_buf) hipFree(s.d_buf);
        if (s.h_buf) hipHostFree(s.h_buf);
    }
    delete b;
}

uint32_t batch_capacity(const BatchState *b) { return b->max_items; }

hipError_t batch_submit(BatchState *b, const DasParams *params, const int *stores, const int *interps,
                        const DasLayout *layouts, uint32_t n, hipStream_t stream) {
    BatchSlot &s = b->slot[b->next];
    b->next ^= 1;
    if (s.pending) {
        const hipError_t e = hipEventSynchronize(s.done);  // previous use of this slot has finished
        if (e != hipSuccess) return e;
        s.pending = false;
    }
    BatchHeader *hdr = (BatchHeader *)s.h_buf;
    BatchItem *items = (BatchItem *)((char *)s.h_buf + items_offset());
    uint32_t kept = 0;
    unsigned sections = 0;
    for (int sec = 0; sec < kSections; ++sec) {
        hdr->first_item[sec] = kept;
        uint64_t total = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const DasParams &p = params[i];
            if (p.n_vox == 0 || p.n_frames == 0 || (p.n_tx > 1 ? 1 : 0) != sec) continue;
            const DasLayout l = layouts[i];
            const int idx = layout_index(l);
            if (idx < 0) return hipErrorInvalidValue;
            const int lv = 64 / l.lanes_per_voxel;
            const int64_t bx = ((p.n_vox + lv - 1) / lv + 3) / 4;
            const int64_t by = (p.n_frames + (int64_t)l.lanes_per_voxel * l.frames_per_lane - 1) /
                               ((int64_t)l.lanes_per_voxel * l.frames_per_lane);
            BatchItem &it = items[kept++];
            it.p = p;
            it.store = stores[i];
            it.interp = interps[i];
            it.layout = idx;
            it.blocks_x = (uint32_t)bx;
            it.blocks_y = (uint32_t)by;
            it.first_block = (uint32_t)total;
            total += (uint64_t)bx * (uint64_t)by;
            if (total > 0xffffffffull) return hipErrorInvalidValue;
        }
        hdr->n_items[sec] = kept - hdr->first_item[sec];
        hdr->total_blocks[sec] = (uint32_t)total;
        if (total > 0) sections |= 1u << sec;
    }
    if (sections == 0) return hipSuccess;
    hipError_t e;
    if (b->use_graph) {
        if (s.exec == nullptr) {
            e = capture(b, s);
            if (e != hipSuccess) return e;
        }
        e = hipGraphLaunch(s.exec, stream);
    } else {
        e = enqueue_direct(b, s, stream, sections);
    }
    if (e != hipSuccess) return e;
    e = hipEventRecord(s.done, stream);
    if (e == hipSuccess) s.pending = true;
    return e;
}

}  // namespace mach
