// Batched delay-and-sum: many requests of different shapes, dtypes and interpolation types in one
// persistent-kernel launch, replayed from a captured hipGraph.
//
// A batch is a table of BatchItem descriptors (one per request), each carrying the request's
// DasParams, its kernel variant and the size of its work grid (voxel groups x frame chunks, the
// same blocks the plain kernel would launch). Work items are numbered consecutively across requests
// (`first_block` is the prefix sum). A persistent grid of workgroups pulls work items from one
// device counter until the batch is drained, so small requests fill the GPU together and a large
// request spreads over every CU. Every workgroup exits once the counter passes the total.
//
// Host side: two slots, each with pinned staging, a device descriptor table, a counter and an
// instantiated graph [H2D copy of the table -> counter reset kernel -> persistent kernel]. A slot is
// reused only after the event recorded behind its previous submission has completed, so the host
// never rewrites staging that a pending copy still reads.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <new>

#include "das_batch.h"
#include "das_device.h"

namespace mach {

namespace {

using namespace dev;

// (lanes per voxel, frames per lane) variants compiled into the batched kernels: every layout the
// plain kernel's heuristic picks for ensembles (16x4, 32x2, 64x1) and for short frame counts, the
// autotuner's winners for the benchmark workloads (32x2 and 16x4 ensembles, 2x4 volumetric, 4x1 single
// frame, 1x1 compounding), and fallbacks for every frame count.
constexpr int kBatchLayouts[][2] = {{16, 4}, {32, 2}, {64, 1}, {8, 4}, {4, 4}, {2, 4}, {1, 4},
                                    {16, 2}, {4, 2}, {1, 2}, {16, 1}, {4, 1}, {1, 1}};
constexpr int kNumBatchLayouts = sizeof(kBatchLayouts) / sizeof(kBatchLayouts[0]);

// A batch has two sections, drained by two persistent kernels launched back to back: single-transmit
// requests (section 0, no transmit chunking) and multi-transmit requests (section 1, chunks of
// kTxChunk transmits). Each kernel's register budget, and so its occupancy, is set by its own
// variants only: the single-transmit variants need about half the registers of the chunked ones.
constexpr int kSections = 2;

struct BatchHeader {
    uint32_t n_items[kSections];
    uint32_t total_blocks[kSections];
    uint32_t first_item[kSections];
    uint32_t pad[2];
};

struct BatchItem {
    DasParams p;
    int32_t store;
    int32_t interp;
    int32_t layout;  // index into kBatchLayouts
    uint32_t blocks_x;
    uint32_t blocks_y;
    uint32_t first_block;  // within the item's section
    uint32_t pad[2];
};

template <int ST, int INTERP, int TC>
__device__ __forceinline__ void run_layout(const BatchItem &it, uint32_t bx, uint32_t by) {
    switch (it.layout) {
        case 0: das_block<ST, INTERP, 16, 4, TC>(it.p, bx, by); break;
        case 1: das_block<ST, INTERP, 32, 2, TC>(it.p, bx, by); break;
        case 2: das_block<ST, INTERP, 64, 1, TC>(it.p, bx, by); break;
        case 3: das_block<ST, INTERP, 8, 4, TC>(it.p, bx, by); break;
        case 4: das_block<ST, INTERP, 4, 4, TC>(it.p, bx, by); break;
        case 5: das_block<ST, INTERP, 2, 4, TC>(it.p, bx, by); break;
        case 6: das_block<ST, INTERP, 1, 4, TC>(it.p, bx, by); break;
        case 7: das_block<ST, INTERP, 16, 2, TC>(it.p, bx, by); break;
        case 8: das_block<ST, INTERP, 4, 2, TC>(it.p, bx, by); break;
        case 9: das_block<ST, INTERP, 1, 2, TC>(it.p, bx, by); break;
        case 10: das_block<ST, INTERP, 16, 1, TC>(it.p, bx, by); break;
        case 11: das_block<ST, INTERP, 4, 1, TC>(it.p, bx, by); break;
        default: das_block<ST, INTERP, 1, 1, TC>(it.p, bx, by); break;
    }
}

template <int ST, int TC>
__device__ __forceinline__ void run_interp(const BatchItem &it, uint32_t bx, uint32_t by) {
    switch (it.interp) {
        case 0: run_layout<ST, 0, TC>(it, bx, by); break;
        case 2: run_layout<ST, 2, TC>(it, bx, by); break;
        default: run_layout<ST, 1, TC>(it, bx, by); break;
    }
}

__global__ void batch_reset_counter(uint32_t *counter) {
    if (threadIdx.x < kSections) counter[threadIdx.x] = 0u;
}

// Persistent kernel for one section. TC is the transmit chunk the plain kernel uses for the same
// requests (1 for single-transmit requests, kTxChunk otherwise), so a request's result is identical
// in both paths.
template <int TC>
__global__ __launch_bounds__(256) void das_batch_kernel(const BatchHeader *__restrict__ header,
                                                        const BatchItem *__restrict__ all_items,
                                                        uint32_t *counters) {
    constexpr int kSec = TC > 1 ? 1 : 0;
    __shared__ uint32_t s_work;
    const uint32_t n_items = header->n_items[kSec];
    const uint32_t total = header->total_blocks[kSec];
    const BatchItem *__restrict__ items = all_items + header->first_item[kSec];
    uint32_t *counter = counters + kSec;
    for (;;) {
        __syncthreads();  // everyone has read the previous s_work
        if (threadIdx.x == 0) s_work = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t w = s_work;
        if (w >= total) break;  // block-uniform: the whole workgroup leaves together
        uint32_t lo = 0, hi = n_items - 1;  // last item with first_block <= w (n_items >= 1 here)
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (items[mid].first_block <= w) lo = mid; else hi = mid - 1;
        }
        const BatchItem &it = items[lo];
        const uint32_t local = w - it.first_block;
        const uint32_t bx = local % it.blocks_x, by = local / it.blocks_x;
        switch (it.store) {
            case kStoreF32: run_interp<kStoreF32, TC>(it, bx, by); break;
            case kStoreC32H: run_interp<kStoreC32H, TC>(it, bx, by); break;
            default: run_interp<kStoreC64, TC>(it, bx, by); break;
        }
    }
}

int layout_index(DasLayout l) {
    for (int i = 0; i < kNumBatchLayouts; ++i)
        if (kBatchLayouts[i][0] == l.lanes_per_voxel && kBatchLayouts[i][1] == l.frames_per_lane) return i;
    return -1;
}

}  // namespace

namespace {

int ilog2(int v) {
    int r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}

}  // namespace

// The plain kernel's layout when it is compiled into the batched kernels; otherwise the compiled
// variant nearest to it in frames per voxel group (lanes x frames per lane, in octaves) whose frames
// per lane do not exceed the wanted ones (whose divisibility and alignment have been checked). Ties go
// to more frames per lane. Band layouts run as the general kernel's heuristic choice.
DasLayout batch_layout(const DasParams &p, int store, DasLayout want) {
    if (want.lanes_per_voxel == 0) want = choose_layout(p, store);
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
    return DasLayout{kBatchLayouts[best][0], kBatchLayouts[best][1]};
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
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_reset_counter, dim3(1), dim3(64), 0, stream, s.d_counter);
    const BatchHeader *hdr = (const BatchHeader *)s.d_buf;
    const BatchItem *items = (const BatchItem *)((char *)s.d_buf + items_offset());
    if (sections & 1u)
        hipLaunchKernelGGL(das_batch_kernel<1>, dim3(b->grid[0]), dim3(256), 0, stream, hdr, items, s.d_counter);
    if (sections & 2u)
        hipLaunchKernelGGL(das_batch_kernel<kTxChunk>, dim3(b->grid[1]), dim3(256), 0, stream, hdr, items,
                           s.d_counter);
    return hipGetLastError();
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
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[0], das_batch_kernel<1>, 256, 0);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu[1], das_batch_kernel<kTxChunk>, 256, 0);
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
