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

// (lanes per voxel, frames per lane) variants compiled into the batched kernel.
constexpr int kBatchLayouts[][2] = {{64, 4}, {64, 1}, {16, 1}, {4, 2}, {4, 1}, {1, 4}, {1, 1}};
constexpr int kNumBatchLayouts = sizeof(kBatchLayouts) / sizeof(kBatchLayouts[0]);

struct BatchHeader {
    uint32_t n_items;
    uint32_t total_blocks;
    uint32_t pad[2];
};

struct BatchItem {
    DasParams p;
    int32_t store;
    int32_t interp;
    int32_t layout;  // index into kBatchLayouts
    uint32_t blocks_x;
    uint32_t blocks_y;
    uint32_t first_block;
    uint32_t pad[2];
};

template <int ST, int INTERP, int TC>
__device__ __forceinline__ void run_layout_tc(const BatchItem &it, uint32_t bx, uint32_t by) {
    switch (it.layout) {
        case 0: das_block<ST, INTERP, 64, 4, TC>(it.p, bx, by); break;
        case 1: das_block<ST, INTERP, 64, 1, TC>(it.p, bx, by); break;
        case 2: das_block<ST, INTERP, 16, 1, TC>(it.p, bx, by); break;
        case 3: das_block<ST, INTERP, 4, 2, TC>(it.p, bx, by); break;
        case 4: das_block<ST, INTERP, 4, 1, TC>(it.p, bx, by); break;
        case 5: das_block<ST, INTERP, 1, 4, TC>(it.p, bx, by); break;
        default: das_block<ST, INTERP, 1, 1, TC>(it.p, bx, by); break;
    }
}

// Same transmit chunking as the plain kernel, so a request gives identical results in both.
template <int ST, int INTERP>
__device__ __forceinline__ void run_layout(const BatchItem &it, uint32_t bx, uint32_t by) {
    if (it.p.n_tx > 1) run_layout_tc<ST, INTERP, kTxChunk>(it, bx, by);
    else run_layout_tc<ST, INTERP, 1>(it, bx, by);
}

template <int ST>
__device__ __forceinline__ void run_interp(const BatchItem &it, uint32_t bx, uint32_t by) {
    switch (it.interp) {
        case 0: run_layout<ST, 0>(it, bx, by); break;
        case 2: run_layout<ST, 2>(it, bx, by); break;
        default: run_layout<ST, 1>(it, bx, by); break;
    }
}

__global__ void batch_reset_counter(uint32_t *counter) {
    if (threadIdx.x == 0) counter[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void das_batch_kernel(const BatchHeader *__restrict__ header,
                                                        const BatchItem *__restrict__ items,
                                                        uint32_t *counter) {
    __shared__ uint32_t s_work;
    const uint32_t n_items = header->n_items;
    const uint32_t total = header->total_blocks;
    for (;;) {
        __syncthreads();  // everyone has read the previous s_work
        if (threadIdx.x == 0) s_work = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t w = s_work;
        if (w >= total) break;  // block-uniform: the whole workgroup leaves together
        uint32_t lo = 0, hi = n_items - 1;  // last item with first_block <= w
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (items[mid].first_block <= w) lo = mid; else hi = mid - 1;
        }
        const BatchItem &it = items[lo];
        const uint32_t local = w - it.first_block;
        const uint32_t bx = local % it.blocks_x, by = local / it.blocks_x;
        switch (it.store) {
            case kStoreF32: run_interp<kStoreF32>(it, bx, by); break;
            case kStoreC32H: run_interp<kStoreC32H>(it, bx, by); break;
            default: run_interp<kStoreC64>(it, bx, by); break;
        }
    }
}

int layout_index(DasLayout l) {
    for (int i = 0; i < kNumBatchLayouts; ++i)
        if (kBatchLayouts[i][0] == l.lanes_per_voxel && kBatchLayouts[i][1] == l.frames_per_lane) return i;
    return -1;
}

}  // namespace

// Closest compiled variant to the plain kernel's choice, never exceeding its frames-per-lane (whose
// alignment and divisibility choose_layout has checked).
DasLayout batch_layout(const DasParams &p, int store, DasLayout want) {
    if (layout_index(want) >= 0) return want;
    const int64_t F = p.n_frames;
    if (F >= 64) return DasLayout{64, want.frames_per_lane >= 4 ? 4 : 1};
    if (F >= 16) return DasLayout{16, 1};
    if (F >= 4) return DasLayout{4, want.frames_per_lane >= 2 ? 2 : 1};
    return DasLayout{1, want.frames_per_lane >= 4 ? 4 : 1};
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
    int grid = 0;
    int next = 0;
    BatchSlot slot[2];
};

namespace {

size_t items_offset() { return (sizeof(BatchHeader) + 255) & ~size_t(255); }

hipError_t enqueue_direct(BatchState *b, BatchSlot &s, hipStream_t stream) {
    hipError_t e = hipMemcpyAsync(s.d_buf, s.h_buf, b->table_bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_reset_counter, dim3(1), dim3(64), 0, stream, s.d_counter);
    const BatchHeader *hdr = (const BatchHeader *)s.d_buf;
    const BatchItem *items = (const BatchItem *)((char *)s.d_buf + items_offset());
    hipLaunchKernelGGL(das_batch_kernel, dim3(b->grid), dim3(256), 0, stream, hdr, items, s.d_counter);
    return hipGetLastError();
}

hipError_t capture(BatchState *b, BatchSlot &s) {
    hipStream_t cap;
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t le = enqueue_direct(b, s, cap);
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
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, das_batch_kernel, 256, 0);
    b->grid = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
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
    uint64_t total = 0;
    uint32_t kept = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const DasParams &p = params[i];
        if (p.n_vox == 0 || p.n_frames == 0) continue;
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
    hdr->n_items = kept;
    hdr->total_blocks = (uint32_t)total;
    if (kept == 0) return hipSuccess;
    hipError_t e;
    if (b->use_graph) {
        if (s.exec == nullptr) {
            e = capture(b, s);
            if (e != hipSuccess) return e;
        }
        e = hipGraphLaunch(s.exec, stream);
    } else {
        e = enqueue_direct(b, s, stream);
    }
    if (e != hipSuccess) return e;
    e = hipEventRecord(s.done, stream);
    if (e == hipSuccess) s.pending = true;
    return e;
}

}  // namespace mach
