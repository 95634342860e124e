"""Throughput benchmark of the delay-and-sum hot path on one MI355X.

Default workload: the Doppler ensemble (128 receive channels x 1024 samples x 256 frames, complex64
IQ, 256x256 voxels, quadratic interpolation, f-number 1.7, per-frame output), synthetic rotating-disk
data from a fixed seed. One step = one beamform of the whole ensemble into a preallocated output
(one kernel launch, inputs resident in HBM). Prints ONE JSON line on stdout:

  value        points/s in Gpts/s, points = voxels x channels x frames x transmits (nominal)
  roofline     HBM roofline of the kernel on algorithmic bytes (channel data at its stored dtype +
               coordinates + arrivals + output written) / measured kernel time (HIP events on the
               launch stream); plus the FP32-VALU figure under roofline_valu
  roofline_valu FP32 FMA throughput from the committed instruction counters of the same kernel
               (SQ_INSTS_VALU_FMA_F32 x 64 lanes x 2 flop / kernel time) against the FP32 vector peak
  call         the reference's timed closure (tests/compare/test_pymust.py:277-297): `out` zeroed, the
               drop-in entry called (accumulate), and the device synchronised, wall clock per call
  served       the mixed request stream through the continuous-batching server against one plain launch
               per request: at 300 QPS (latency; arrival_bound when the plain path keeps up with the
               arrivals) and offered at 20000 QPS (throughput_ratio: batched vs plain points/s)
  cpu_baseline an FP32 delay-and-sum of the same algorithm (oracle/das_cpu32.c, OpenMP over the host
               cores, built -O3 -march=native on the host it runs on) timed on a seeded voxel sample of
               the same workload; the FP64 oracle's time on the same sample is reported beside it

Only one GPU is used: ``--gpus`` above 1 exits with a message.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np
import torch

HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3
FLOP_PER_POINT = {0: 8, 1: 16, 2: 20}  # complex NN / linear / quadratic (per point, setup amortised)
SIMDS = 1024  # 256 CUs x 4 SIMDs
ENGINE_CLOCK_GHZ = 2.4
WARMUP_SECONDS = 0.5  # launches before the counted warm-up, so clocks have ramped up


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(acq, storage: str, accumulate: bool = False, output: str = "signal") -> int:
    T, E, S, F, V = acq.n_tx, acq.n_rx, acq.n_samples, acq.n_frames, acq.n_vox
    data = T * E * S * F * (4 if storage == "fp16" else 8)
    coords = (E + V) * 3 * 4
    arrivals = T * V * 4
    out = {"signal": V * F * 8, "power": V * 4, "envelope": V * F * 4}[output] * (2 if accumulate else 1)
    return data + coords + arrivals + out


def _counters_summary(workload: str, storage: str, output: str):
    """(text, repo-relative path) of the committed counters summary of this workload, or (None, None)."""
    suffix = "" if output == "signal" else f"_{output}"
    root = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(root, "profiles", f"das_kernel_{workload}_{storage}{suffix}_counters.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        return fh.read(), os.path.relpath(path, root)


def measured_traffic(workload: str, storage: str, output: str):
    """HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE counters (calibrated on a copy kernel)
    for this workload, from its committed counters summary; None when there is none."""
    text, path = _counters_summary(workload, storage, output)
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", text) if text else None
    return (int(m.group(1)), path) if m else (None, None)


def counter_value(text: str, name: str):
    """Per-launch value of counter `name` from a counters summary table row (`| NAME | 1,234 |`)."""
    m = re.search(r"\|\s*" + re.escape(name) + r"\s*\|\s*([\d,]+)\s*\|", text or "")
    return int(m.group(1).replace(",", "")) if m else None


def valu_roofline(workload: str, storage: str, output: str, kernel_ms: float, nominal_tflops: float):
    """FP32 VALU figures from the committed instruction counters of the same kernel: FMA throughput
    (each SQ_INSTS_VALU_FMA_F32 wave instruction is 64 lanes x 2 flop) and VALU busy (one wave64 VALU
    instruction per 4 clocks per SIMD) over this run's kernel time. The nominal flop model, which counts
    the pairs the kernel skips, is kept beside it for comparison."""
    text, path = _counters_summary(workload, storage, output)
    fma = counter_value(text, "SQ_INSTS_VALU_FMA_F32")
    valu = counter_value(text, "SQ_INSTS_VALU")
    t = kernel_ms * 1e-3
    res = {"peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s", "counters_source": path,
           "nominal_model_tflops": round(nominal_tflops, 2)}
    if fma is None or valu is None:
        res.update({"achieved": None, "frac": None, "valu_busy": None})
        return res
    tflops = fma * 64 * 2 / t / 1e12
    res.update({"achieved": round(tflops, 2), "frac": round(tflops / FP32_VECTOR_PEAK_TFLOPS, 4),
                "valu_busy": round(valu * 4 / (SIMDS * ENGINE_CLOCK_GHZ * 1e9) / t, 4),
                "fma_share_of_valu": round(fma / valu, 4)})
    return res


def run_gpu(acq, storage, steps, warmup, output="signal", warmup_seconds=WARMUP_SECONDS):
    import mach_amd as mach
    from mach_amd import _hip_impl, postprocess

    p = dict(acq.params)
    cd = acq.fp16_channel_data() if storage == "fp16" else acq.channel_data
    interp = mach.InterpolationType(p["interp_type"])
    if output == "signal":
        out = torch.empty((acq.n_vox, acq.n_frames), dtype=torch.complex64, device=cd.device)

        def step():
            _hip_impl.beamform_compound(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, out,
                                        p["f_number"], p["rx_start_s"], p["sampling_freq_hz"], p["sound_speed_m_s"],
                                        p["modulation_freq_hz"], p["tukey_alpha"], interp, accumulate=False)
    else:
        fn = postprocess.power_doppler if output == "power" else postprocess.envelope
        shape = (acq.n_vox,) if output == "power" else (acq.n_vox, acq.n_frames)
        out = torch.empty(shape, dtype=torch.float32, device=cd.device)
        kw = {k: p[k] for k in ("f_number", "rx_start_s", "sampling_freq_hz", "sound_speed_m_s",
                                "modulation_freq_hz", "tukey_alpha")}

        def step():
            fn(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, out, interp_type=interp, **kw)

    # launch until warmup_seconds have passed (the GPU's clocks ramp up under load), then the counted
    # warm-up; the queue is drained every 64 launches
    t_end = time.perf_counter() + warmup_seconds
    n = 0
    while time.perf_counter() < t_end:
        step()
        n += 1
        if n % 64 == 0:
            torch.cuda.synchronize()
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / steps
    return wall / steps * 1e3, kernel_ms, out


def _time_rate(fn, pts, min_seconds, max_runs=50):
    """Median points/s of repeated calls of fn (after one warm-up call), until min_seconds have run."""
    fn()
    rates, total = [], 0.0
    while (total < min_seconds or len(rates) < 3) and len(rates) < max_runs:
        t0 = time.perf_counter()
        fn()
        dt = time.perf_counter() - t0
        total += dt
        rates.append(pts / dt / 1e9)
    return float(np.median(rates)), len(rates), total


def cpu_baseline(acq, storage, n_voxels, min_seconds=10.0, seed=0):
    """The FP32 CPU delay-and-sum (oracle/das_cpu32.c, -O3 -march=native, OpenMP over the host cores
    this process may use) timed over a seeded voxel sample of the same workload and storage, repeated
    until at least ``min_seconds`` of CPU work has run; the median rate over the repetitions. The FP64 C
    oracle is timed on the same sample for a few seconds and reported beside it, as the oracle."""
    from oracle import das_c, das_cpu32

    V = acq.n_vox
    n = min(n_voxels, V) if n_voxels > 0 else V
    idx = np.sort(np.random.default_rng(seed).choice(V, n, replace=False))
    cd = (acq.fp16_channel_data() if storage == "fp16" else acq.channel_data).cpu().numpy()
    rx = acq.rx_coords_m.cpu().numpy()
    scan = np.ascontiguousarray(acq.scan_coords_m.cpu().numpy()[idx])
    arr = np.ascontiguousarray(acq.tx_wave_arrivals_s.cpu().numpy()[:, idx])
    threads = len(os.sched_getaffinity(0))
    pts = len(idx) * acq.n_rx * acq.n_frames * acq.n_tx
    _, build = das_cpu32.load(native=True)
    rate, runs, total = _time_rate(
        lambda: das_cpu32.das_cpu32(cd, rx, scan, arr, **acq.params, n_threads=threads), pts, min_seconds, 1000)
    orate, oruns, ototal = _time_rate(
        lambda: das_c.das_c(cd, rx, scan, arr, **acq.params, n_threads=threads), pts, min(min_seconds, 3.0), 5)
    return {
        "value": rate,
        "unit": "Gpts/s",
        "cores": min(threads, das_cpu32.max_threads()),
        "kind": "port",
        "cpu_model": das_cpu32.cpu_model(),
        "build": build + " (OMP threads = the cores this process may use)",
        "sample": f"{len(idx)} of {V} voxels (seeded), all {acq.n_rx} ch x {acq.n_frames} frames x {acq.n_tx} tx, "
                  f"{storage} storage; FP32 delay-and-sum (oracle/das_cpu32.c), median of {runs} runs, "
                  f"{total:.1f} s total",
        "fp64_oracle": {"value": orate, "unit": "Gpts/s", "runs": oruns, "seconds": round(ototal, 2),
                        "what": "oracle/das_oracle.c (FP64, OpenMP), the checker, timed on the same sample"},
    }


def call_time(acq, storage, reps=50):
    """The reference's timed closure (tests/compare/test_pymust.py:277-297) through this package's
    drop-in: zero `out`, call the entry (nb_beamform for one transmit, experimental.beamform for
    several; both accumulate into a device `out`), synchronise the device as the reference's binding
    does after every call (kernel.cu:874). Median wall clock per call over `reps` calls."""
    import mach_amd as mach
    from mach_amd import experimental

    p = dict(acq.params)
    interp = mach.InterpolationType(p["interp_type"])
    cd = acq.fp16_channel_data() if storage == "fp16" else acq.channel_data
    out = torch.empty((acq.n_vox, acq.n_frames), dtype=torch.complex64, device=cd.device)
    kw = {k: p[k] for k in ("f_number", "rx_start_s", "sampling_freq_hz", "sound_speed_m_s", "modulation_freq_hz",
                            "tukey_alpha")}
    if acq.n_tx == 1:
        def call():
            out.zero_()
            mach.kernel.nb_beamform(cd[0], acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s[0], out,
                                    interp_type=interp, **kw)
            torch.cuda.synchronize()
        entry = "mach_amd.kernel.nb_beamform"
    else:
        def call():
            out.zero_()
            experimental.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, out,
                                  interp_type=interp, **kw)
            torch.cuda.synchronize()
        entry = "mach_amd.experimental.beamform"
    for _ in range(5):
        call()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        call()
        times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    return {"ms": round(ms, 4), "gpts_per_s": round(acq.points / (ms * 1e-3) / 1e9, 3), "entry": entry,
            "calls": reps, "what": "out zeroed + drop-in call (accumulate) + device synchronisation, median"}


def _serve_pair(reqs, dev):
    from mach_amd import serving

    plain = serving.serve_plain(reqs, dev)
    batched = serving.serve_batched(reqs, dev)
    return {
        "batched": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in batched.as_dict().items()},
        "plain": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in plain.as_dict().items()},
        "throughput_ratio": round(batched.gpts_per_s / plain.gpts_per_s, 3),
    }


def served_stream(n_requests, qps, saturated_qps=None, seed=4):
    """Mixed request stream through the continuous-batching server vs one plain launch per request,
    at ``qps`` (latency under Poisson arrivals); with ``saturated_qps`` the same requests again with
    their arrival times compressed to that rate, an offered load above what the plain path can serve.

    ``throughput_ratio`` at the top is the saturated stream's (batched vs plain points/s over the
    makespan): at ``qps`` the stream is arrival-bound whenever the plain path keeps up with the
    arrivals (both paths then finish just after the last request arrives, whatever the server does),
    which ``arrival_bound`` records for that stream."""
    from mach_amd import serving

    dev = torch.device("cuda", 0)
    reqs = serving.make_stream(dev, n_requests=n_requests, qps=qps, seed=seed)
    # warm both paths (graph capture, template caches) on a short prefix
    warm = reqs[:8]
    serving.serve_plain(warm, dev)
    serving.serve_batched(warm, dev)
    res = {
        "workload": f"{n_requests} requests, Poisson {qps:g} QPS (seed {seed}), drawn from plane_wave/compound/"
                    "doppler/volumetric with grid x0.5-1.0 and 1-256 frames (plane_wave, doppler)",
        "budget_gb": 200,
        "last_arrival_s": round(max(r.arrival_s for r in reqs), 4),
    }
    res.update(_serve_pair(reqs, dev))
    # arrival-bound: the plain path, one request at a time, ends within 10 % of the last arrival
    res["arrival_bound"] = bool(res["plain"]["makespan_s"] <= 1.1 * res["last_arrival_s"])
    res["poisson_throughput_ratio"] = res.pop("throughput_ratio")
    res["throughput_ratio"] = res["poisson_throughput_ratio"]
    res["throughput_ratio_basis"] = f"Poisson {qps:g} QPS stream"
    if saturated_qps:
        for r in reqs:
            r.arrival_s *= qps / saturated_qps
        sat = {"qps": saturated_qps, "last_arrival_s": round(max(r.arrival_s for r in reqs), 4)}
        sat.update(_serve_pair(reqs, dev))
        res["saturated"] = sat
        res["throughput_ratio"] = sat["throughput_ratio"]
        res["throughput_ratio_basis"] = (f"the same requests offered at {saturated_qps:g} QPS, above the plain "
                                         "path's capacity (points/s over the makespan, batched vs plain)")
    return res


def copy_calibration(mib: int, reps: int = 5):
    """A plain device-to-device copy of known size, run so that a counter pass over this process can
    scale its FETCH_SIZE / WRITE_SIZE readings (bytes moved = mib MiB read + mib MiB written)."""
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel (a DMA copy would not show in counters)
    torch.cuda.synchronize()
    del src, dst


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--warmup-seconds", type=float, default=WARMUP_SECONDS,
                    help="untimed launches for this long before the counted warm-up (clock ramp)")
    ap.add_argument("--workload", default="doppler",
                    choices=["doppler", "plane_wave", "compound", "volumetric", "pymust_disk"])
    ap.add_argument("--storage", default=None, choices=["c64", "fp16"])
    ap.add_argument("--output", default="signal", choices=["signal", "power", "envelope"],
                    help="what the kernel writes: the per-frame signal (the reference's output, default), the "
                         "power Doppler image or the envelope (fused epilogues)")
    ap.add_argument("--cpu-voxels", type=int, default=0, help="voxel sample for the CPU baseline (0 = all)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the served mixed-stream measurement")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-qps", type=float, default=300.0)
    ap.add_argument("--serve-saturated-qps", type=float, default=20000.0,
                    help="also serve the same requests arriving at this rate (0 = skip)")
    ap.add_argument("--no-call-time", action="store_true", help="skip the reference-style call timing")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain copy of this many MiB (counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the delay-and-sum path is "
              "sharded across GPUs by voxel tiles, which this build does not run.", file=sys.stderr)
        return 2
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 2

    from mach_amd import phantoms

    dev = torch.device("cuda", 0)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
    storage = args.storage or ("fp16" if args.workload == "volumetric" else "c64")
    log(f"bench: synthesizing {args.workload} (seeded) on {dev}")
    acq = phantoms.make(args.workload, device=dev)
    torch.cuda.synchronize()
    log(f"bench: shape T,E,S,F={tuple(acq.channel_data.shape)} voxels={acq.n_vox} storage={storage}")
    ms_step, kernel_ms, _ = run_gpu(acq, storage, args.steps, args.warmup, args.output, args.warmup_seconds)
    pts = acq.points
    value = pts / (ms_step * 1e-3) / 1e9
    b_alg = algorithmic_bytes(acq, storage, output=args.output)
    traffic, traffic_src = measured_traffic(args.workload, storage, args.output)
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    interp = int(acq.params["interp_type"])
    nominal_tflops = FLOP_PER_POINT[interp] * pts / (kernel_ms * 1e-3) / 1e12
    log(f"bench: {ms_step:.3f} ms/step, kernel {kernel_ms:.3f} ms, {value:.1f} Gpts/s, {achieved:.1f} GB/s alg")
    result = {
        "metric": "Gpts/s (voxel·ch·frame), % HBM roofline; served p50/p99 ms vs plain",
        "value": round(value, 3),
        "unit": "Gpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage == "c64" else "f16-storage/f32-compute",
        "data": "synthetic (seeded closed-form echoes)",
        "config": {
            "workload": {
                "doppler": "Doppler ensemble: 128 ch x 1024 samples x 256 frames complex64 IQ, 256x256 voxels, "
                           "quadratic, f#1.7, per-frame output",
                "plane_wave": "single 0deg plane wave: 128 ch x 2048 samples x 1 frame, 256x256 voxels, linear, f#1.7",
                "compound": "75-angle plane-wave compounding: 75 tx x 128 ch x 2048 samples, 387x609 voxels, linear",
                "volumetric": "32x32 matrix, 5 diverging tx x 1024 ch x 1536 samples x 8 frames, 128^3 voxels, linear, f#1.5",
                "pymust_disk": "shape of the reference's published benchmark: 128 ch x 1024 samples x 32 frames "
                               "complex64 IQ, 251x251 voxels, linear, f#1.0, no taper (synthetic rotating disk)",
            }[args.workload],
            "storage": storage,
            "output": args.output,
            "points_per_step": pts,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes": b_alg,
            "kernel_ms": round(kernel_ms, 4),
        },
        "roofline_valu": valu_roofline(args.workload, storage, args.output, kernel_ms, nominal_tflops),
    }
    if not args.no_call_time and args.output == "signal":
        result["call"] = call_time(acq, storage)
        log(f"bench: drop-in call (zero out + call + sync) {result['call']['ms']} ms")
    if not args.no_serve:
        log("bench: served mixed stream (batched persistent kernel vs plain per-request)")
        result["served"] = served_stream(args.serve_requests, args.serve_qps, args.serve_saturated_qps or None)
        srv = result["served"]
        log(f"bench: served at {args.serve_qps:g} QPS {srv['poisson_throughput_ratio']}x plain"
            f"{' (arrival-bound)' if srv['arrival_bound'] else ''}; "
            f"p50/p99 {srv['batched']['p50_ms']}/{srv['batched']['p99_ms']} ms")
        if "saturated" in result["served"]:
            sat = result["served"]["saturated"]
            log(f"bench: saturated ({sat['qps']:g} QPS) {sat['throughput_ratio']}x plain; "
                f"p50/p99 {sat['batched']['p50_ms']}/{sat['batched']['p99_ms']} ms")
    if not args.no_cpu_baseline:
        log("bench: timing CPU baseline")
        result["cpu_baseline"] = cpu_baseline(acq, storage, args.cpu_voxels, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
