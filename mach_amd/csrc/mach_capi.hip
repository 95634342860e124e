// C ABI of the beamformer (declared in include/mach_hip.h): argument validation, error reporting and
// dispatch to the delay-and-sum kernels. Validation messages keep the reference's wording
// (src/mach/kernel.cu:740-769) so callers that match on them keep working.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <climits>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <string>

#include <vector>

#include "../../include/mach_hip.h"
#include "das_band.h"
#include "das_batch.h"
#include "das_hilbert.h"
#include "das_kernel.h"
#include "das_post.h"

struct mach_batch {
    mach::BatchState *state;
};

static_assert(sizeof(mach_wave) == sizeof(mach::WaveDesc), "mach_wave and WaveDesc must have one layout");
static_assert(offsetof(mach_wave, vector_m) == offsetof(mach::WaveDesc, vec), "mach_wave layout");
static_assert(offsetof(mach_wave, time_offset_s) == offsetof(mach::WaveDesc, offset), "mach_wave layout");

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(MACH_ERR_HIP, std::string("HIP error in ") + what + ": " + hipGetErrorName(e) + " (" +
                                  hipGetErrorString(e) + ")");
}

// Block-to-XCD remap granularity (xcd_remap): MACH_AMD_XCD_GROUP=<n> forces contiguous eighths (0) or
// runs of n blocks for every launch, for measurements; unset, each launch picks (mach::xcd_group_for).
// Read once per process.
int32_t xcd_group_override() {
    static const int32_t g = [] {
        const char *env = std::getenv("MACH_AMD_XCD_GROUP");
        if (env == nullptr || *env == '\0') return (int32_t)-1;
        const long v = std::strtol(env, nullptr, 10);
        return (int32_t)(v < 0 ? 0 : (v > 4096 ? 4096 : v));
    }();
    return g;
}

// Opt-in timing (MACH_AMD_PROFILE=1, read once per process), the counterpart of the reference's
// CUDA_PROFILE build (kernel.cu:31-65: host section timers; 805-890: CUDA events around the kernel).
// A timed call records HIP events around its launches on its stream, waits for them, prints the kernel
// time and the call's host wall time (which then includes the wait) to stderr and keeps both for
// mach_last_timing. Off, it costs one branch per call.
bool profile_enabled() {
    static const bool on = [] {
        const char *env = std::getenv("MACH_AMD_PROFILE");
        return env != nullptr && *env != '\0' && std::strcmp(env, "0") != 0;
    }();
    return on;
}

thread_local float g_last_kernel_ms = -1.f, g_last_call_ms = -1.f;

class CallTimer {
  public:
    CallTimer(const char *what, hipStream_t stream) : what_(what), stream_(stream), on_(profile_enabled()) {
        if (!on_) return;
        t0_ = std::chrono::steady_clock::now();
        ok_ = hipEventCreate(&start_) == hipSuccess && hipEventCreate(&stop_) == hipSuccess;
    }
    // Marks the start of the call's device work.
    void launches_begin() {
        if (on_ && ok_) ok_ = hipEventRecord(start_, stream_) == hipSuccess;
    }
    // Marks the end of the call's device work, waits for it and reports; `detail` describes the launch.
    void launches_end(const std::string &detail) {
        if (!on_) return;
        float ms = -1.f;
        if (ok_ && hipEventRecord(stop_, stream_) == hipSuccess && hipEventSynchronize(stop_) == hipSuccess)
            hipEventElapsedTime(&ms, start_, stop_);
        const double call = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
        g_last_kernel_ms = ms;
        g_last_call_ms = (float)call;
        std::fprintf(stderr, "mach_amd profile: %s kernel %.4f ms, call %.4f ms; %s\n", what_, ms, call, detail.c_str());
    }
    ~CallTimer() {
        if (start_ != nullptr) hipEventDestroy(start_);
        if (stop_ != nullptr) hipEventDestroy(stop_);
    }

  private:
    const char *what_;
    hipStream_t stream_;
    bool on_;
    bool ok_ = false;
    hipEvent_t start_ = nullptr, stop_ = nullptr;
    std::chrono::steady_clock::time_point t0_;
};

std::string describe(const mach::DasParams &p, int data_type, int interp, const mach::DasLayout &l, bool band) {
    static const char *kStore[] = {"f32", "c64", "fp16-iq"};
    static const char *kInterp[] = {"nearest", "linear", "quadratic"};
    return std::string(band ? "band kernel" : (l.voxels_per_lane > 1 ? "lateral kernel" : "general kernel")) + " " +
           kStore[data_type] + " " + kInterp[interp] + " lanes/voxel=" + std::to_string(l.lanes_per_voxel) +
           " frames/lane=" + std::to_string(l.frames_per_lane) + " groups=" + std::to_string(l.chunks_per_lane) +
           " voxels/lane=" + std::to_string(l.voxels_per_lane) + " n_tx=" + std::to_string(p.n_tx) +
           " n_rx=" + std::to_string(p.n_rx) + " n_samples=" + std::to_string(p.n_samples) +
           " n_frames=" + std::to_string(p.n_frames) + " n_vox=" + std::to_string(p.n_vox);
}

int validate_and_build(const mach_beamform_args *a, mach::DasParams *p) {
    if (a == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: args is NULL");
    if (a->data_type < MACH_DATA_F32 || a->data_type > MACH_DATA_C32H)
        return fail(MACH_ERR_UNSUPPORTED, "Error: unsupported data_type " + std::to_string(a->data_type));
    if (a->interp_type < MACH_INTERP_NEAREST || a->interp_type > MACH_INTERP_QUADRATIC)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: unsupported interp_type " + std::to_string(a->interp_type));
    // kernel.cu:745-748
    if (!(a->tukey_alpha >= 0.0f && a->tukey_alpha <= 1.0f))
        return fail(MACH_ERR_INVALID_ARGUMENT,
                    "Error: tukey_alpha must be in range [0, 1], but got " + std::to_string(a->tukey_alpha));
    if (a->n_samples > (uint64_t)INT32_MAX)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: n_samples exceeds " + std::to_string(INT32_MAX));
    if (a->n_tx > (uint64_t)INT64_MAX / 4 || a->n_rx > (uint64_t)INT32_MAX || a->n_vox > (uint64_t)INT64_MAX / 64 ||
        a->n_frames > (uint64_t)INT32_MAX)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: array dimensions exceed the supported range");
    // 64-bit element offsets replace the reference's 32-bit indexing cap (kernel.cu:738-743); the
    // product still has to fit int64.
    const long double total = (long double)a->n_tx * a->n_rx * a->n_samples * a->n_frames;
    if (total > (long double)INT64_MAX / 4)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: channel data size exceeds the 64-bit indexing limit");
    if (a->output < MACH_OUTPUT_SIGNAL || a->output > MACH_OUTPUT_ENVELOPE)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: unsupported output kind " + std::to_string(a->output));
    if (a->output == MACH_OUTPUT_ENVELOPE && a->accumulate)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: the envelope output cannot be accumulated (|a + b| != |a| + |b|)");
    if (a->output == MACH_OUTPUT_SIGNAL && a->peak != nullptr)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: peak applies to the power and envelope outputs only");
    // the power output has n_vox values even when there are no frames (a sum over nothing is 0)
    const bool need_data = a->n_vox > 0 && (a->n_frames > 0 || a->output == MACH_OUTPUT_POWER);
    if (need_data && a->out == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: out is NULL");
    if (need_data && a->scan_coords_m == nullptr)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: scan_coords_m is NULL");
    if (a->tx_waves != nullptr && a->tx_wave_arrivals_s != nullptr)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: give tx_wave_arrivals_s or tx_waves, not both");
    const bool need_inputs = need_data && a->n_tx > 0 && a->n_rx > 0;
    if (need_inputs && (a->rx_coords_m == nullptr || (a->tx_wave_arrivals_s == nullptr && a->tx_waves == nullptr)))
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: rx_coords_m / tx_wave_arrivals_s is NULL");
    if (need_inputs && a->n_samples > 0 && a->n_frames > 0 && a->channel_data == nullptr)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: channel_data is NULL");

    p->data = a->channel_data;
    p->rx = a->rx_coords_m;
    p->vox = a->scan_coords_m;
    p->arrivals = a->tx_wave_arrivals_s;
    p->waves = (const mach::WaveDesc *)a->tx_waves;
    p->out = a->out;
    p->order = a->voxel_order;
    p->n_tx = (int64_t)a->n_tx;
    p->n_rx = (int64_t)a->n_rx;
    p->n_samples = (int64_t)a->n_samples;
    p->n_frames = (int64_t)a->n_frames;
    p->n_vox = (int64_t)a->n_vox;
    p->tap_step = a->n_samples > 1 ? (int64_t)a->n_frames : 0;
    p->two_f_number = 2.0 * (double)a->f_number;
    p->inv_c = 1.0 / (double)a->sound_speed_m_s;
    p->t0 = (double)a->rx_start_s;
    p->fs = (double)a->sampling_freq_hz;
    p->fmod = (double)a->modulation_freq_hz;
    p->flat_r2 = (1.0 - (double)a->tukey_alpha) * (1.0 - (double)a->tukey_alpha);
    p->lin_tap_max = a->n_samples > 2 ? (double)(a->n_samples - 2) : 0.0;
    p->quarter_over_alpha = a->tukey_alpha > 0.0f ? (float)(0.25 / (double)a->tukey_alpha) : 0.f;
    p->rotate = (a->data_type != MACH_DATA_F32 && a->modulation_freq_hz != 0.0f) ? 1 : 0;
    p->accumulate = a->accumulate ? 1 : 0;
    p->apod = a->tukey_alpha > 0.0f ? 1 : 0;  // kernel.cu:749
    const uint64_t elem = a->data_type == MACH_DATA_C64 ? 8u : 4u;
    p->bytes32 = ((long double)a->n_samples * a->n_frames * elem) < 4294967296.0L ? 1 : 0;
    p->output = a->output;
    p->peak = a->peak;
    p->xcd_group = xcd_group_override();
    return MACH_OK;
}

}  // namespace

extern "C" {

// Kernel choice for one request: the LDS-staged band kernel (frame ensembles with a depth-major voxel
// order) or the general kernel with a frame-lane layout. `layout` 0 chooses the general kernel's
// heuristic layout; a value with lanes_per_voxel == 0 requests the band kernel with that many frames
// per lane (the autotuner records it where it wins).
struct Plan {
    bool band;
    int band_fpl;
    mach::DasLayout layout;
};

static int resolve_plan(const mach_beamform_args *a, const mach::DasParams &p, Plan *out) {
    out->band = false;
    out->band_fpl = 0;
    if ((a->layout & 0xffff) == 0) {  // the band kernel is chosen only through a (tuned) layout
        out->layout = mach::choose_layout(p, a->data_type);
        return MACH_OK;
    }
    const mach::DasLayout l = mach::decode_layout(a->layout);
    // epilogues and on-device arrivals live in the general kernel
    if (l.lanes_per_voxel == 0 && (p.output != mach::kOutSignal || p.waves != nullptr)) {
        out->layout = mach::choose_layout(p, a->data_type);
        return MACH_OK;
    }
    if (l.lanes_per_voxel == 0) {
        const int k = mach::band_frames_per_lane(p, a->data_type);
        if (l.chunks_per_lane != 1)
            return fail(MACH_ERR_INVALID_ARGUMENT, "Error: layout " + std::to_string(a->layout) +
                                                       " (band kernel) takes no frame-group count");
        if (k == 0 || (l.frames_per_lane != 1 && l.frames_per_lane != 2 && l.frames_per_lane != 4) ||
            p.n_frames % (64 * l.frames_per_lane) != 0 || l.frames_per_lane > k)
            return fail(MACH_ERR_INVALID_ARGUMENT, "Error: layout " + std::to_string(a->layout) +
                                                       " (band kernel) does not apply to this request");
        out->band = true;
        out->band_fpl = l.frames_per_lane;
        out->layout = mach::choose_general_layout(p, a->data_type);
        return MACH_OK;
    }
    if (!mach::layout_valid(p, a->data_type, l))
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: layout " + std::to_string(a->layout) +
                                                   " is not valid for this request (lanes_per_voxel must be a power of"
                                                   " two <= 64, frames_per_lane in {1, 2, 4} dividing n_frames and"
                                                   " keeping vector accesses aligned, chunks_per_lane in {2, 4} only"
                                                   " for 16x4, 32x2, 64x1 and 8x4; lateral groups of 2 or 4 voxels"
                                                   " only for 64x2 and 64x4 and a single transmit)");
    out->layout = l;
    return MACH_OK;
}

// Debug builds (MACH_DEBUG): after each call's launches, wait for them and collect the kernels' failed
// bounds checks (counted on the device instead of trapping). The counterpart of the reference's
// CUDA_DEBUG device asserts (kernel.cu:85-92, used at 123, 127, 165-176, 218-230). Release builds
// return `st` unchanged and add no synchronisation.
static int debug_check(int st, hipStream_t stream, const char *what) {
#ifdef MACH_DEBUG
    if (st != MACH_OK) return st;
    hipError_t e = hipStreamSynchronize(stream);
    uint32_t total = 0;
    for (hipError_t (*read)(uint32_t *) : {mach::debug_violations_f32, mach::debug_violations_c64,
                                           mach::debug_violations_c32h, mach::debug_violations_batch0,
                                           mach::debug_violations_batch1, mach::debug_violations_band,
                                           mach::debug_violations_lateral}) {
        uint32_t c = 0;
        if (e == hipSuccess) e = read(&c);
        total += c;
    }
    if (e != hipSuccess) return hip_fail(e, what);
    if (total != 0)
        return fail(MACH_ERR_DEBUG_CHECK, std::string("Error (debug build): ") + std::to_string(total) +
                                              " failed bounds checks in the kernels of " + what);
    return MACH_OK;
#else
    (void)stream;
    (void)what;
    return st;
#endif
}

// True when a validated request has no term to sum: no voxels, no frames, no transmits or no receive
// elements. Such a request launches no delay-and-sum kernel and reads none of its inputs (validation
// lets them be NULL then); the power image still has one (zero) value per voxel.
static bool empty_sum(const mach::DasParams &p) {
    return p.n_vox == 0 || p.n_frames == 0 || p.n_tx == 0 || p.n_rx == 0;
}

// Result of an empty sum: overwrite mode writes zeros to every output value, accumulate mode leaves
// `out` unchanged; `peak` (if given) is set to 0.
static hipError_t enqueue_empty_sum(const mach::DasParams &p, int data_type, hipStream_t stream) {
    hipError_t e = hipSuccess;
    uint64_t n_out = 0;  // float32 values of `out`
    if (p.output == mach::kOutPower)
        n_out = (uint64_t)p.n_vox;
    else
        n_out = (uint64_t)p.n_vox * (uint64_t)p.n_frames *
                ((p.output == mach::kOutSignal && data_type != MACH_DATA_F32) ? 2u : 1u);
    if (!p.accumulate && n_out > 0) e = mach::launch_fill((float *)p.out, n_out, 0.f, stream);
    if (e == hipSuccess && p.peak != nullptr) e = mach::launch_fill(p.peak, 1, 0.f, stream);
    return e;
}

// Validates every request of a batch and resolves its kernel variant into caller-provided arrays of
// n entries. All-or-nothing: the first invalid request fails the whole batch (its index prefixes the
// message) before anything is enqueued, so a caller can resubmit the valid requests.
static int build_batch(const mach_beamform_args *requests, uint32_t n, mach::DasParams *params, int *stores,
                       int *interps, mach::DasLayout *layouts) {
    for (uint32_t i = 0; i < n; ++i) {
        int st = validate_and_build(&requests[i], &params[i]);
        if (st == MACH_OK && params[i].output != mach::kOutSignal)
            st = fail(MACH_ERR_UNSUPPORTED, "Error: batches run the signal output only");
        if (st == MACH_OK && !params[i].bytes32)
            st = fail(MACH_ERR_UNSUPPORTED, "Error: channel blocks of 4 GiB or more (n_samples * n_frames * element "
                                            "bytes) run through mach_beamform only");
        Plan plan{false, 0, mach::DasLayout{1, 1}};
        if (st == MACH_OK && !empty_sum(params[i])) st = resolve_plan(&requests[i], params[i], &plan);
        if (st != MACH_OK) return fail(st, "request " + std::to_string(i) + ": " + g_last_error);
        layouts[i] = mach::batch_layout(params[i], requests[i].data_type, plan.layout);
        stores[i] = requests[i].data_type;
        interps[i] = requests[i].interp_type;
    }
    return MACH_OK;
}

// Empty sums of a validated batch get their fill (enqueue_empty_sum) and are taken out of the work
// table (n_vox = 0), so the persistent kernels never read their (possibly NULL) inputs.
static hipError_t enqueue_empty_sums(const mach_beamform_args *requests, mach::DasParams *params, uint32_t n,
                                     hipStream_t stream) {
    for (uint32_t i = 0; i < n; ++i) {
        if (!empty_sum(params[i])) continue;
        const hipError_t e = enqueue_empty_sum(params[i], requests[i].data_type, stream);
        if (e != hipSuccess) return e;
        params[i].n_vox = 0;
    }
    return hipSuccess;
}

int mach_beamform(const mach_beamform_args *args, void *stream) {
    g_last_error.clear();
    mach::DasParams p;
    int st = validate_and_build(args, &p);
    if (st != MACH_OK) return st;
    if (empty_sum(p)) {
        const hipError_t e = enqueue_empty_sum(p, args->data_type, (hipStream_t)stream);
        return e == hipSuccess ? MACH_OK : hip_fail(e, "beamform empty-sum fill");
    }
    Plan plan;
    st = resolve_plan(args, p, &plan);
    if (st != MACH_OK) return st;
    // a remap fixed by the layout code (the autotuner's choice), unless the environment forces one
    if (p.xcd_group < 0) p.xcd_group = mach::layout_xcd_group(args->layout);
    CallTimer timer("mach_beamform", (hipStream_t)stream);
    timer.launches_begin();
    const hipError_t e =
        plan.band ? mach::launch_das_band(p, args->data_type, args->interp_type, plan.band_fpl, (hipStream_t)stream)
                  : mach::launch_das(p, args->data_type, args->interp_type, plan.layout, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "beamform kernel launch");
    if (profile_enabled()) timer.launches_end(describe(p, args->data_type, args->interp_type, plan.layout, plan.band));
    return debug_check(MACH_OK, (hipStream_t)stream, "mach_beamform");
}

int mach_batch_create(uint32_t max_requests, int32_t use_graph, mach_batch **batch) {
    g_last_error.clear();
    if (batch == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: batch is NULL");
    *batch = nullptr;
    if (max_requests == 0 || max_requests > (1u << 20))
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: max_requests must be in [1, 1048576]");
    mach::BatchState *s = nullptr;
    const hipError_t e = mach::batch_create(max_requests, use_graph != 0, &s);
    if (e != hipSuccess) return hip_fail(e, "mach_batch_create");
    *batch = new mach_batch{s};
    return MACH_OK;
}

int mach_batch_submit(mach_batch *batch, const mach_beamform_args *requests, uint32_t n_requests, void *stream) {
    g_last_error.clear();
    if (batch == nullptr || batch->state == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: batch is NULL");
    if (n_requests > mach::batch_capacity(batch->state))
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: " + std::to_string(n_requests) +
                                                   " requests exceed the batch capacity of " +
                                                   std::to_string(mach::batch_capacity(batch->state)));
    if (n_requests > 0 && requests == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: requests is NULL");
    std::vector<mach::DasParams> params(n_requests);
    std::vector<int> stores(n_requests), interps(n_requests);
    std::vector<mach::DasLayout> layouts(n_requests);
    const int st = build_batch(requests, n_requests, params.data(), stores.data(), interps.data(), layouts.data());
    if (st != MACH_OK) return st;
    CallTimer timer("mach_batch_submit", (hipStream_t)stream);
    timer.launches_begin();
    hipError_t e = enqueue_empty_sums(requests, params.data(), n_requests, (hipStream_t)stream);
    if (e == hipSuccess)
        e = mach::batch_submit(batch->state, params.data(), stores.data(), interps.data(), layouts.data(), n_requests,
                               (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "mach_batch_submit");
    if (profile_enabled()) timer.launches_end(std::to_string(n_requests) + " requests");
    return debug_check(MACH_OK, (hipStream_t)stream, "mach_batch_submit");
}

int mach_batch_submit_noalloc(const mach_beamform_args *requests, uint32_t n_requests, void *stream) {
    g_last_error.clear();
    static_assert(MACH_NOALLOC_MAX_REQUESTS == mach::kNoAllocMaxRequests, "allocation-free table size");
    if (n_requests > MACH_NOALLOC_MAX_REQUESTS)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: " + std::to_string(n_requests) +
                                                   " requests exceed the allocation-free maximum of " +
                                                   std::to_string(MACH_NOALLOC_MAX_REQUESTS));
    if (n_requests > 0 && requests == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: requests is NULL");
    if (n_requests == 0) return MACH_OK;
    // fixed-size arrays on the stack: nothing is allocated on the success path
    mach::DasParams params[MACH_NOALLOC_MAX_REQUESTS];
    int stores[MACH_NOALLOC_MAX_REQUESTS], interps[MACH_NOALLOC_MAX_REQUESTS];
    mach::DasLayout layouts[MACH_NOALLOC_MAX_REQUESTS];
    const int st = build_batch(requests, n_requests, params, stores, interps, layouts);
    if (st != MACH_OK) return st;
    CallTimer timer("mach_batch_submit_noalloc", (hipStream_t)stream);
    timer.launches_begin();
    hipError_t e = enqueue_empty_sums(requests, params, n_requests, (hipStream_t)stream);
    if (e == hipSuccess)
        e = mach::batch_submit_noalloc(params, stores, interps, layouts, n_requests, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "mach_batch_submit_noalloc");
    if (profile_enabled()) timer.launches_end(std::to_string(n_requests) + " requests");
    return debug_check(MACH_OK, (hipStream_t)stream, "mach_batch_submit_noalloc");
}

int mach_batch_destroy(mach_batch *batch) {
    g_last_error.clear();
    if (batch == nullptr) return MACH_OK;
    mach::batch_destroy(batch->state);
    delete batch;
    return MACH_OK;
}

static int beamform_single(int data_type, const float *channel_data, const float *rx_coords_m,
                           const float *scan_coords_m, const float *tx_wave_arrivals_s, float *out, uint64_t n_rx,
                           uint64_t n_samples, uint64_t n_frames, uint64_t n_vox, float f_number, float rx_start_s,
                           float sampling_freq_hz, float sound_speed_m_s, float modulation_freq_hz,
                           float tukey_alpha, int32_t interp_type, void *stream) {
    mach_beamform_args a;
    std::memset(&a, 0, sizeof(a));
    a.channel_data = channel_data;
    a.rx_coords_m = rx_coords_m;
    a.scan_coords_m = scan_coords_m;
    a.tx_wave_arrivals_s = tx_wave_arrivals_s;
    a.out = out;
    a.n_tx = 1;
    a.n_rx = n_rx;
    a.n_samples = n_samples;
    a.n_frames = n_frames;
    a.n_vox = n_vox;
    a.data_type = data_type;
    a.interp_type = interp_type;
    a.f_number = f_number;
    a.rx_start_s = rx_start_s;
    a.sampling_freq_hz = sampling_freq_hz;
    a.sound_speed_m_s = sound_speed_m_s;
    a.modulation_freq_hz = modulation_freq_hz;
    a.tukey_alpha = tukey_alpha;
    a.accumulate = 1;
    return mach_beamform(&a, stream);
}

int mach_beamform_c64(const float *channel_data, const float *rx_coords_m, const float *scan_coords_m,
                      const float *tx_wave_arrivals_s, float *out, uint64_t n_rx, uint64_t n_samples,
                      uint64_t n_frames, uint64_t n_vox, float f_number, float rx_start_s, float sampling_freq_hz,
                      float sound_speed_m_s, float modulation_freq_hz, float tukey_alpha, int32_t interp_type,
                      void *stream) {
    return beamform_single(MACH_DATA_C64, channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, n_rx,
                           n_samples, n_frames, n_vox, f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s,
                           modulation_freq_hz, tukey_alpha, interp_type, stream);
}

int mach_beamform_f32(const float *channel_data, const float *rx_coords_m, const float *scan_coords_m,
                      const float *tx_wave_arrivals_s, float *out, uint64_t n_rx, uint64_t n_samples,
                      uint64_t n_frames, uint64_t n_vox, float f_number, float rx_start_s, float sampling_freq_hz,
                      float sound_speed_m_s, float modulation_freq_hz, float tukey_alpha, int32_t interp_type,
                      void *stream) {
    return beamform_single(MACH_DATA_F32, channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, n_rx,
                           n_samples, n_frames, n_vox, f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s,
                           modulation_freq_hz, tukey_alpha, interp_type, stream);
}

int mach_tx_arrivals(const mach_wave *tx_waves, uint64_t n_tx, const float *scan_coords_m, uint64_t n_vox,
                     float *out, void *stream) {
    g_last_error.clear();
    if (n_tx == 0 || n_vox == 0) return MACH_OK;
    if (tx_waves == nullptr || scan_coords_m == nullptr || out == nullptr)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: tx_waves / scan_coords_m / out is NULL");
    if (n_tx > (uint64_t)INT32_MAX || n_vox > (uint64_t)INT64_MAX / 4 / n_tx)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: array dimensions exceed the supported range");
    const hipError_t e = mach::launch_tx_arrivals((const mach::WaveDesc *)tx_waves, n_tx, scan_coords_m, n_vox, out,
                                                  (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "mach_tx_arrivals");
    return MACH_OK;
}

int mach_analytic_signal(const float *rf, uint64_t n_samples, uint64_t n_rx, uint64_t n_tx, uint64_t n_frames,
                         void *out, int32_t out_type, void *stream) {
    g_last_error.clear();
    if (out_type != MACH_DATA_C64 && out_type != MACH_DATA_C32H)
        return fail(MACH_ERR_UNSUPPORTED, "Error: analytic signal output must be MACH_DATA_C64 or MACH_DATA_C32H");
    if (n_samples == 0 || n_rx == 0 || n_tx == 0 || n_frames == 0) return MACH_OK;
    mach::HilbertPlan plan;
    if (!mach::hilbert_plan((int64_t)n_samples, &plan))
        return fail(MACH_ERR_UNSUPPORTED, "Error: n_samples = " + std::to_string(n_samples) +
                                              " is not a product of 2, 3 and 5 up to " +
                                              std::to_string(mach::kHilbertMaxSamples));
    if (rf == nullptr || out == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: rf / out is NULL");
    if (n_rx > (uint64_t)INT32_MAX || n_tx > (uint64_t)INT32_MAX || n_frames > (uint64_t)INT32_MAX ||
        (long double)n_samples * n_rx * n_tx * n_frames > (long double)INT64_MAX / 8)
        return fail(MACH_ERR_INVALID_ARGUMENT, "Error: array dimensions exceed the supported range");
    const hipError_t e = mach::launch_analytic_signal(rf, (int64_t)n_samples, (int64_t)n_rx, (int64_t)n_tx,
                                                      (int64_t)n_frames, out, out_type == MACH_DATA_C32H,
                                                      (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "mach_analytic_signal");
    return MACH_OK;
}

int mach_to_db(float *x, uint64_t n, const float *peak, int32_t power_mode, void *stream) {
    g_last_error.clear();
    if (n == 0) return MACH_OK;
    if (x == nullptr || peak == nullptr) return fail(MACH_ERR_INVALID_ARGUMENT, "Error: x / peak is NULL");
    const hipError_t e = mach::launch_to_db(x, n, peak, power_mode != 0, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "mach_to_db");
    return MACH_OK;
}

const char *mach_last_error(void) { return g_last_error.c_str(); }

int mach_last_timing(float *kernel_ms, float *call_ms) {
    if (kernel_ms != nullptr) *kernel_ms = g_last_kernel_ms;
    if (call_ms != nullptr) *call_ms = g_last_call_ms;
    return profile_enabled() ? MACH_OK : MACH_ERR_UNSUPPORTED;
}

int32_t mach_abi_version(void) { return MACH_ABI_VERSION; }

const char *mach_hip_runtime_version(void) {
    static thread_local char buf[64];
    int v = 0;
    if (hipRuntimeGetVersion(&v) != hipSuccess) {
        std::snprintf(buf, sizeof(buf), "%d.%d.%d", HIP_VERSION_MAJOR, HIP_VERSION_MINOR, HIP_VERSION_PATCH);
        return buf;
    }
    // HIP encodes the runtime version as major * 10000000 + minor * 100000 + patch
    std::snprintf(buf, sizeof(buf), "%d.%d.%d", v / 10000000, (v / 100000) % 100, v % 100000);
    return buf;
}

int mach_device_arch(char *arch, size_t len) {
    g_last_error.clear();
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(MACH_ERR_NO_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return fail(MACH_ERR_NO_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    if (arch != nullptr && len > 0) {
        std::strncpy(arch, prop.gcnArchName, len - 1);
        arch[len - 1] = '\0';
    }
    return MACH_OK;
}

}  // extern "C"
