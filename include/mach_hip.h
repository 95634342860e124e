/*
 * mach_hip.h - C ABI of the MI355X (gfx950) delay-and-sum beamformer.
 *
 * Every pointer argument named as "device" must point to device memory of the current HIP device.
 * Arrays are C-contiguous. Coordinates are float32 [n][3] (x, y, z in metres). Complex data is
 * interleaved (re, im) float32 (complex64). FP16 IQ data is interleaved (re, im) IEEE binary16.
 * Scalars are float32, exactly the type the reference binding passes to its kernel.
 * Functions return MACH_OK (0) or a mach_status code; mach_last_error() then describes the failure
 * (thread-local; valid until the next call on the same thread). No function calls exit().
 * Work is enqueued on `stream` (a hipStream_t; NULL = the default stream) and is not synchronised.
 */
#ifndef MACH_HIP_H
#define MACH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MACH_ABI_VERSION 2

/* Interpolation selector; same integer values as the reference's
 * `enum class InterpolationType` (src/mach/kernel.cu:25-29, exported at kernel.cu:1222-1226). */
typedef enum mach_interp_type {
    MACH_INTERP_NEAREST = 0,
    MACH_INTERP_LINEAR = 1,
    MACH_INTERP_QUADRATIC = 2
} mach_interp_type;

/* Storage type of channel data and the matching output type. The reference supports float (RF)
 * and float2 (IQ) only (src/mach/kernel.cu:543-545); FP16 IQ storage is new here. */
typedef enum mach_data_type {
    MACH_DATA_F32 = 0,   /* RF float32 in, float32 out                       */
    MACH_DATA_C64 = 1,   /* IQ complex64 in, complex64 out                   */
    MACH_DATA_C32H = 2   /* IQ as (binary16 re, binary16 im) in, complex64 out */
} mach_data_type;

/* What a beamform call writes to `out`. The reference returns the beamformed signal and leaves
 * post-processing to the caller: power Doppler as np.square(np.abs(y)).sum(axis=-1)
 * (examples/doppler.py:252, tests/compare/test_pymust.py:317-318) and B-mode as 20 log10 |y|
 * normalised to its maximum (src/mach/_vis.py:17-40). The epilogues compute these inside the kernel,
 * so the per-frame signal never reaches HBM. */
typedef enum mach_output_kind {
    MACH_OUTPUT_SIGNAL = 0,   /* out [n_vox][n_frames]: complex64 for IQ, float32 for RF (reference) */
    MACH_OUTPUT_POWER = 1,    /* out float32 [n_vox]: sum over frames of |y|^2 (power Doppler)       */
    MACH_OUTPUT_ENVELOPE = 2  /* out float32 [n_vox][n_frames]: |y| (B-mode envelope)              */
} mach_output_kind;

/* Transmit wavefront kinds for on-device arrival generation (mach_wave). */
typedef enum mach_wave_kind {
    MACH_WAVE_PLANE = 0,     /* distance = (p - origin) . direction (src/mach/wavefront.py:8-50)       */
    MACH_WAVE_SPHERICAL = 1  /* distance = |origin - focus| sign(focus_z - origin_z)
                                         - |focus - p| sign(focus_z - p_z) (src/mach/wavefront.py:53-116) */
} mach_wave_kind;

/* One transmit's wavefront. The arrival of voxel p is float32(distance(p) / sound_speed_m_s +
 * time_offset_s), evaluated in float64 from the float32 voxel coordinates -- the value the reference's
 * callers compute on the host (tests/compare/test_pymust.py:69-95; src/mach/io/uff.py:37-67 and
 * 156-158, where time_offset_s is the per-transmit sequence delay) and pass as tx_wave_arrivals_s. */
typedef struct mach_wave {
    int32_t kind;            /* mach_wave_kind                                                  */
    int32_t reserved;        /* 0                                                               */
    double origin_m[3];      /* plane: wave origin; spherical: transmitting element / reference */
    double vector_m[3];      /* plane: unit propagation direction; spherical: focus (virtual source
                                behind the origin for a diverging wave)                          */
    double sound_speed_m_s;
    double time_offset_s;
} mach_wave;

typedef enum mach_status {
    MACH_OK = 0,
    MACH_ERR_INVALID_ARGUMENT = 1, /* reference: std::runtime_error (kernel.cu:740-769)       */
    MACH_ERR_DIMENSION = 2,        /* reference: "Dimension mismatch ..." (kernel.cu:914-951) */
    MACH_ERR_HIP = 3,              /* reference: checkCudaErrors -> exit(1); here a status    */
    MACH_ERR_NO_DEVICE = 4,
    MACH_ERR_UNSUPPORTED = 5,
    MACH_ERR_DEBUG_CHECK = 6       /* debug build (libmach_hip_debug.so) only: a kernel bounds check
                                      failed (the reference's CUDA_DEBUG asserts, kernel.cu:85-92);
                                      the access was skipped, the result is not valid               */
} mach_status;

/* General beamform request. Replaces, in one call, the reference's per-transmit loop
 * `for t: kernel.beamform(channel_data[t], ..., tx_wave_arrivals_s[t], out=out)`
 * (src/mach/experimental.py:80-97) on top of `beamform<T>` (src/mach/kernel.cu:1033-1209):
 * transmits are compounded inside the kernel instead of by n_tx launches. */
typedef struct mach_beamform_args {
    const void *channel_data;         /* device [n_tx][n_rx][n_samples][n_frames] of data_type  */
    const float *rx_coords_m;         /* device [n_rx][3]                                       */
    const float *scan_coords_m;       /* device [n_vox][3]                                      */
    const float *tx_wave_arrivals_s;  /* device [n_tx][n_vox]                                   */
    void *out;                        /* device [n_vox][n_frames]; complex64 for IQ, f32 for RF */
    uint64_t n_tx;
    uint64_t n_rx;
    uint64_t n_samples;
    uint64_t n_frames;
    uint64_t n_vox;
    int32_t data_type;                /* mach_data_type                                         */
    int32_t interp_type;              /* mach_interp_type                                       */
    float f_number;
    float rx_start_s;
    float sampling_freq_hz;
    float sound_speed_m_s;
    float modulation_freq_hz;         /* ignored for RF                                         */
    float tukey_alpha;                /* [0, 1]; 0 disables apodization                         */
    int32_t accumulate;               /* 1: out += result (reference device-array semantics,
                                         kernel.cu:674-680); 0: out = result                    */
    int32_t layout;                   /* 0: automatic. Otherwise (lanes_per_voxel << 4) |
                                         frames_per_lane, e.g. 0x104: 16 lanes share a voxel,
                                         4 frames each (the autotuner's choice; rejected if it
                                         does not divide n_frames or break vector alignment).
                                         lanes_per_voxel 0 (0x001, 0x002, 0x004) selects the
                                         LDS-staged kernel for frame ensembles, which needs a
                                         voxel_order and n_frames a multiple of 64 x
                                         frames_per_lane. 0x8000 | (log2(voxels_per_lane)
                                         << 12) | 0x400 | frames_per_lane (0x9402, 0x9404,
                                         0xA402, 0xA404) selects lateral groups: each lane
                                         holds its frames of 2 or 4 voxels consecutive in
                                         voxel_order, read from shared sample windows (single
                                         transmit only). Every layout gives bit-identical
                                         signal and envelope outputs for single-transmit
                                         requests; general-kernel layouts also for several
                                         transmits. The power image is not layout-independent:
                                         each lane sums |y|^2 over its own frames before the
                                         lanes of a voxel are added, so its rounding depends on
                                         lanes per voxel and frames per lane. Bits 16-19 may fix
                                         the launch's block order over the 8 XCDs: 0 = chosen
                                         per launch, 1 = each XCD a contiguous eighth,
                                         2 <= v <= 10 = runs of 2^(v+1) blocks dealt
                                         round-robin, 15 = the hardware order (block bx of
                                         every row on XCD bx % 8; for a voxel_order that gives
                                         each XCD its own voxels, such as the package's XCD
                                         wedges of a volume). Results do not depend on it;
                                         batches ignore it.                                     */
    const int32_t *voxel_order;       /* optional device [n_vox] permutation of 0..n_vox-1: the
                                         order voxels are processed in (NULL = 0, 1, 2, ...).
                                         Results still land at out[v]; a depth-major order
                                         (sorted by z, then x, then y) keeps the sample windows
                                         of neighbouring waves in the same L2.                   */
    int32_t output;                   /* mach_output_kind (0 = the reference's signal output).
                                         POWER accepts accumulate (adds to out); ENVELOPE needs
                                         accumulate = 0. Epilogues run in the general kernel
                                         only (not the LDS-staged one, not in batches).          */
    float *peak;                      /* optional device float for POWER / ENVELOPE: receives the
                                         maximum value written to out (set to 0 first, then
                                         raised with an order-independent atomic max). Calls
                                         that share one `peak` must not overlap.                 */
    const mach_wave *tx_waves;        /* optional device [n_tx] wavefronts: the kernel computes each
                                         voxel's transmit arrival instead of reading
                                         tx_wave_arrivals_s, which must then be NULL.            */
} mach_beamform_args;

/* Complex64 (IQ) entry. Replaces the nanobind overload `_cuda_impl.beamform` for complex64
 * (src/mach/kernel.cu:1229-1242, implemented by beamform<std::complex<float>> at 1033-1209 and
 * _beamform_impl at 715-891): same arguments in the same order, plus explicit sizes and a stream.
 * `out` is accumulated into, as the reference does for device arrays. */
int mach_beamform_c64(const float *channel_data, const float *rx_coords_m, const float *scan_coords_m,
                      const float *tx_wave_arrivals_s, float *out, uint64_t n_rx, uint64_t n_samples,
                      uint64_t n_frames, uint64_t n_vox, float f_number, float rx_start_s,
                      float sampling_freq_hz, float sound_speed_m_s, float modulation_freq_hz,
                      float tukey_alpha, int32_t interp_type, void *stream);

/* Float32 (RF) entry. Replaces the nanobind overload `_cuda_impl.beamform` for float32
 * (src/mach/kernel.cu:1244-1255, beamform<float>). `out` is accumulated into. */
int mach_beamform_f32(const float *channel_data, const float *rx_coords_m, const float *scan_coords_m,
                      const float *tx_wave_arrivals_s, float *out, uint64_t n_rx, uint64_t n_samples,
                      uint64_t n_frames, uint64_t n_vox, float f_number, float rx_start_s,
                      float sampling_freq_hz, float sound_speed_m_s, float modulation_freq_hz,
                      float tukey_alpha, int32_t interp_type, void *stream);

/* General entry (transmit compounding, FP16 IQ storage, overwrite or accumulate). Replaces
 * src/mach/experimental.py:23-106 over src/mach/kernel.cu:1033-1209.
 * A request with no term to sum (n_vox, n_frames, n_tx or n_rx equal to 0) launches no kernel and
 * reads none of its inputs, which may then be NULL: overwrite mode writes zeros to `out` (n_vox values
 * for the power image), accumulate mode leaves it unchanged, and `peak` is set to 0. */
int mach_beamform(const mach_beamform_args *args, void *stream);

/* Materialises the transmit arrivals of `n_tx` device wavefronts for `n_vox` device voxel
 * coordinates into device float32 out[n_tx][n_vox], with the same values the kernel computes from
 * mach_beamform_args.tx_waves. Replaces the host-side wavefront.plane / wavefront.spherical / c
 * step of the reference's callers (src/mach/wavefront.py:8-116, src/mach/io/uff.py:37-67). */
int mach_tx_arrivals(const mach_wave *tx_waves, uint64_t n_tx, const float *scan_coords_m, uint64_t n_vox,
                     float *out, void *stream);

/* Analytic signal of real RF channel data, the step before beamforming IQ data: the device
 * counterpart of the reference's UFF preprocessing
 * (src/mach/io/uff.py:70-103), which takes scipy.signal.hilbert along the sample axis and rearranges
 * (n_samples, n_elements, n_waves, n_frames) to (n_waves, n_elements, n_samples, n_frames).
 * rf: device float32 [n_samples][n_rx][n_tx][n_frames]; out: device [n_tx][n_rx][n_samples][n_frames]
 * of out_type MACH_DATA_C64 (complex64) or MACH_DATA_C32H (binary16 re, im pairs; values must fit
 * binary16). n_samples must be 2^a 3^b 5^c and at most 4096 (MACH_ERR_UNSUPPORTED otherwise).
 * Beamform the result with modulation_freq_hz = 0 (the analytic signal keeps its carrier). */
int mach_analytic_signal(const float *rf, uint64_t n_samples, uint64_t n_rx, uint64_t n_tx, uint64_t n_frames,
                         void *out, int32_t out_type, void *stream);

/* In-place logarithmic compression of a POWER or ENVELOPE output: x[i] = k log10(a(x[i])) -
 * k log10(a(*peak)) with a(x) = |x|, or float64 machine epsilon where x == 0, and k = 10 for power
 * (power_mode != 0) or 20 for amplitude. Restates `db_zero` of src/mach/_vis.py:17-40 (which
 * normalises by the maximum of the converted image) for a device image whose maximum is `*peak`.
 * `peak` is a device pointer; n may be 0. No reference counterpart as a library function. */
int mach_to_db(float *x, uint64_t n, const float *peak, int32_t power_mode, void *stream);

/* Batched execution: many requests (any mix of shapes, data types, interpolation types and
 * transmit counts) run in one submission of two persistent kernels (single-transmit requests, then
 * multi-transmit ones) that drain a device work queue, replayed from a captured hipGraph. No reference counterpart: the reference runs one synchronous launch per
 * request and transmit (src/mach/kernel.cu:820-874, src/mach/experimental.py:80-97).
 * The batch object owns two descriptor slots, so two submissions can be in flight; a third waits
 * for the oldest. Requests' arrays must stay valid until the submission's work has completed on
 * `stream`. Results are identical to mach_beamform on each request alone (same per-voxel order). */
typedef struct mach_batch mach_batch;

/* Allocates the batch object (pinned + device descriptor tables for up to `max_requests` each).
 * use_graph != 0 replays a captured hipGraph (copy descriptors -> reset queues -> persistent
 * kernels). */
int mach_batch_create(uint32_t max_requests, int32_t use_graph, mach_batch **batch);

/* Enqueues all `n_requests` requests on `stream` as one submission. The band layouts run as the
 * general kernel here. All or nothing: every request is validated before anything is enqueued, and
 * the first invalid one fails the call (mach_last_error() starts with "request <index>: "), so the
 * caller can drop it and resubmit the others. */
int mach_batch_submit(mach_batch *batch, const mach_beamform_args *requests, uint32_t n_requests,
                      void *stream);

/* Waits for the batch's pending work and frees it. */
int mach_batch_destroy(mach_batch *batch);

/* Requests one mach_batch_submit_noalloc call takes at most. */
#define MACH_NOALLOC_MAX_REQUESTS 64

/* Allocation-free variant of mach_batch_submit, with no counterpart in the reference: for callers that
 * cannot allocate (code under stream or graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It needs no mach_batch object and allocates nothing, on the host or the device: the
 * descriptor table and work counters are a fixed device array of the library (about 15 KiB, held by
 * each GPU the library runs on until the process ends), written from kernel arguments, and the same
 * persistent kernels run the requests without graph replay. Results equal mach_batch_submit's and
 * mach_beamform's on each request. At most MACH_NOALLOC_MAX_REQUESTS requests (MACH_ERR_INVALID_ARGUMENT
 * above that); validation is all or nothing, as for mach_batch_submit. Every call on one GPU shares its
 * table: calls on one stream are ordered, but calls that overlap on one GPU (other streams, other
 * threads) race. The request array may be reused as soon as the call returns; the arrays it points to
 * must stay valid until the work has completed on `stream`. Only error paths allocate (the message). */
int mach_batch_submit_noalloc(const mach_beamform_args *requests, uint32_t n_requests, void *stream);

/* Message of the last failure on this thread ("" if none). Replaces the text of the reference's
 * std::runtime_error messages (src/mach/kernel.cu:740-769, 926-950). */
const char *mach_last_error(void);

/* Opt-in timing: with MACH_AMD_PROFILE=1 in the environment (read once per process), every
 * mach_beamform / mach_batch_submit / mach_batch_submit_noalloc call records HIP events around its
 * launches on its stream, waits for them, and prints the kernel time and the call's wall time to
 * stderr. This reads the last timed call's figures on this thread (-1 before any) and returns MACH_OK,
 * or MACH_ERR_UNSUPPORTED when timing is off. Replaces the reference's compile-time CUDA_PROFILE
 * section timers and kernel events (src/mach/kernel.cu:31-65, 805-890). */
int mach_last_timing(float *kernel_ms, float *call_ms);

/* ABI version (MACH_ABI_VERSION). */
int32_t mach_abi_version(void);

/* HIP runtime version as "major.minor.patch". Replaces the module attribute __nvcc_version__
 * (src/mach/kernel.cu:1215). */
const char *mach_hip_runtime_version(void);

/* Writes the current device's gcnArchName into `arch` (NUL-terminated, truncated to `len`) and
 * returns MACH_OK, or MACH_ERR_NO_DEVICE. Replaces the import-time capability checks
 * checkCudaDriverCompatibility / checkComputeCapability (src/mach/kernel.cu:357-423). */
int mach_device_arch(char *arch, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* MACH_HIP_H */
