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

#include <atomic>
#include <cstring>
#include <mutex>
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
    // requests whose transmits split the block's waves run as 1 x 1 (the only batched layout compiled
    // with the split; every layout of theirs sums in the same order)
    if (tx_split_waves(p.n_tx, p.n_frames) != 1) return DasLayout{1, 1};
    if (want.lanes_per_voxel == 0) want = choose_general_layout(p, store);
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

constexpr size_t items_offset() { return (sizeof(BatchHeader) + 255) & ~size_t(255); }

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

namespace {

// Fills the descriptor table (header and items, the items of each section consecutive) for the n
// requests; requests with no voxels or frames get no item. Returns the non-empty sections as a bit mask
// in *sections and the number of items written in *n_items.
hipError_t pack_table(const DasParams *params, const int *stores, const int *interps, const DasLayout *layouts,
                      uint32_t n, BatchHeader *hdr, BatchItem *items, unsigned *sections, uint32_t *n_items) {
    uint32_t kept = 0;
    *sections = 0;
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
            // one voxel group per block where the block's waves split the transmits (run_fpl1)
            const int64_t per_block = 4 / tx_split_waves(p.n_tx, p.n_frames);
            const int64_t bx = ((p.n_vox + lv - 1) / lv + per_block - 1) / per_block;
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
        if (total > 0) *sections |= 1u << sec;
    }
    *n_items = kept;
    return hipSuccess;
}

}  // namespace

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
    unsigned sections = 0;
    uint32_t kept = 0;
    hipError_t e = pack_table(params, stores, interps, layouts, n, hdr, items, &sections, &kept);
    if (e != hipSuccess) return e;
    if (sections == 0) return hipSuccess;
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

// ---------------------------------------------------------------------------------------------------
// Allocation-free batches (mach_batch_submit_noalloc). The descriptor table and the two work counters
// live in `scratch`, a file-scope device array of fixed size: each GPU holds its own copy from module
// load until the process ends. Host code never uses the array's name as an address (on the host it is
// a stand-in symbol): the device address comes from hipGetSymbolAddress, once per device, at the
// device's first call. The table reaches the device through kernel arguments (write_table), which the
// runtime copies at launch, so the call needs no pinned staging and no allocation, and the caller may
// reuse its request array as soon as the call returns. Every call on one GPU shares that GPU's array:
// calls on one stream are ordered, calls that overlap on one GPU (other streams, other threads) race.

constexpr size_t kNoAllocCountersOffset = 128;  // two uint32 work counters after the 32-byte header
static_assert(sizeof(BatchHeader) <= kNoAllocCountersOffset, "header overlaps the counters");
static_assert(kNoAllocCountersOffset + kSections * sizeof(uint32_t) <= items_offset(), "counters overlap items");
constexpr size_t kNoAllocBytes = items_offset() + (size_t)kNoAllocMaxRequests * sizeof(BatchItem);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

// Items one write_table launch carries in its arguments (kept under 4 KiB of kernel arguments).
constexpr uint32_t kItemsPerWrite = 14;

struct TableWrite {
    BatchHeader header;
    uint32_t first;  // table index of items[0]
    uint32_t count;  // items to write
    uint32_t reset;  // 1: also write the header and zero the work counters
    uint32_t pad;
    BatchItem items[kItemsPerWrite];
};
static_assert(sizeof(TableWrite) <= 4096, "table write exceeds 4 KiB of kernel arguments");
static_assert(sizeof(BatchItem) % 4 == 0 && sizeof(BatchHeader) % 4 == 0, "table words");

// Copies `w` (kernel arguments) into the table at `table` word by word.
__global__ __launch_bounds__(256) void write_table(const TableWrite w, unsigned char *table) {
    const uint32_t tid = threadIdx.x;
    constexpr uint32_t kItemWords = sizeof(BatchItem) / 4;
    const uint32_t n_words = w.count * kItemWords;
    uint32_t *dst = (uint32_t *)(table + items_offset()) + (size_t)w.first * kItemWords;
    const uint32_t *src = (const uint32_t *)w.items;
    for (uint32_t i = tid; i < n_words; i += blockDim.x) dst[i] = src[i];
    if (w.reset) {
        constexpr uint32_t kHeaderWords = sizeof(BatchHeader) / 4;
        if (tid < kHeaderWords) ((uint32_t *)table)[tid] = ((const uint32_t *)&w.header)[tid];
        if (tid < kSections) ((uint32_t *)(table + kNoAllocCountersOffset))[tid] = 0u;
    }
}

// Per-device state of the allocation-free path, set up at the device's first call.
struct NoAllocDevice {
    std::atomic<int> ready{0};
    unsigned char *table = nullptr;  // device address of this device's `scratch`
    int grid[kSections] = {0, 0};    // persistent grid per section
};
constexpr int kNoAllocMaxDevices = 64;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];
std::mutex g_noalloc_mutex;

hipError_t noalloc_device(NoAllocDevice **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    NoAllocDevice &d = g_noalloc[dev];
    if (!d.ready.load(std::memory_order_acquire)) {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        if (!d.ready.load(std::memory_order_relaxed)) {
            void *ptr = nullptr;
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // this device's copy of the array
            if (e != hipSuccess) return e;
            int cus = 0, per_cu[kSections] = {0, 0};
            e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (e == hipSuccess) e = section_occupancy(0, &per_cu[0]);
            if (e == hipSuccess) e = section_occupancy(1, &per_cu[1]);
            if (e != hipSuccess) return e;
            d.table = (unsigned char *)ptr;
            for (int i = 0; i < kSections; ++i) d.grid[i] = (per_cu[i] > 0 ? per_cu[i] : 1) * (cus > 0 ? cus : 1);
            d.ready.store(1, std::memory_order_release);
        }
    }
    *out = &d;
    return hipSuccess;
}

}  // namespace

hipError_t batch_submit_noalloc(const DasParams *params, const int *stores, const int *interps,
                                const DasLayout *layouts, uint32_t n, hipStream_t stream) {
    if (n > kNoAllocMaxRequests) return hipErrorInvalidValue;
    BatchHeader hdr{};
    BatchItem items[kNoAllocMaxRequests];
    unsigned sections = 0;
    uint32_t kept = 0;
    hipError_t e = pack_table(params, stores, interps, layouts, n, &hdr, items, &sections, &kept);
    if (e != hipSuccess || sections == 0) return e;  // nothing to run: no device state is touched
    NoAllocDevice *d = nullptr;
    e = noalloc_device(&d);
    if (e != hipSuccess) return e;
    TableWrite w;
    for (uint32_t first = 0; first < kept; first += kItemsPerWrite) {
        std::memset(&w, 0, sizeof(w));
        w.header = hdr;
        w.first = first;
        w.count = kept - first < kItemsPerWrite ? kept - first : kItemsPerWrite;
        w.reset = first == 0 ? 1u : 0u;
        std::memcpy(w.items, items + first, (size_t)w.count * sizeof(BatchItem));
        hipLaunchKernelGGL(write_table, dim3(1), dim3(256), 0, stream, w, d->table);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const BatchHeader *dh = (const BatchHeader *)d->table;
    const BatchItem *di = (const BatchItem *)(d->table + items_offset());
    uint32_t *dc = (uint32_t *)(d->table + kNoAllocCountersOffset);
    if (sections & 1u) e = launch_section0((unsigned)d->grid[0], stream, dh, di, dc);
    if (e == hipSuccess && (sections & 2u)) e = launch_section1((unsigned)d->grid[1], stream, dh, di, dc);
    return e;
}

}  // namespace mach
