"""Throughput benchmark of the photon transport hot path on one MI355X.

A step is one full transport of the workload (every exposure, every
history) on the FP32 persistent path with the world, tables and tally
resident in HBM. The default workload is the CT configuration the metric is
quoted on: 120 kVp CT spiral over a 512x512x300 x 0.8 mm anthropomorphic
phantom, 3600 exposures, 1e9 histories.

Printed (one JSON line): histories/s of the FP32 path (value), the plain
FP64 path's histories/s on a fixed subset of the exposures (fp64 block),
the FP32-vs-FP64 dose agreement on that subset, the roofline of the
transport kernel (algorithmic bytes 6 per in-world Woodcock step + 40 per
scoring event, from device counters, over the kernel time measured with HIP
events on the kernel's stream) and the CPU baseline (the oracle, a CPU
restatement of the reference method, on the host cores on a bounded sample).
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", default="ct_body", choices=["ct_body", "dx_thorax", "water_pencil", "whole_body"])
    p.add_argument("--fp64-stride", type=int, default=1, help="FP64 path runs every n-th exposure (1: the whole workload)")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--block-size", type=int, default=0)
    p.add_argument("--blocks-per-cu", type=int, default=0)
    p.add_argument("--threshold", type=int, default=0, help="FP32 regeneration threshold (dead lanes per wave), 0: default")
    p.add_argument("--interaction-threshold", type=int, default=0, help="FP32 pending interactions per wave before they run, 0: default")
    p.add_argument("--queue-depth", type=int, default=0, help="FP32 histories per pool refill, 0: default")
    p.add_argument("--waves-per-simd", type=int, default=0, help="FP32 kernel variant (5, 6, 7), 0: default (7 on worlds of 1-byte codes or palette indices and 6 on others for L=0, 1; 5 for L=2)")
    p.add_argument("--lds-tables", type=int, default=0, help="FP32 tables in LDS: 1 none, 2 all but the LUT, 3 all; 0: default")
    p.add_argument("--block-bounds", type=int, default=0, help="FP32 on 1-byte code worlds: 1 block code ranges first, 2 voxel codes only; 0: default")
    p.add_argument("--low-energy-correction", type=int, default=1)
    p.add_argument("--no-autotune", action="store_true", help="skip the autotuner (defaults or the fields given)")
    p.add_argument("--no-serving", action="store_true", help="skip the request-stream (serving) measurement")
    p.add_argument("--serving-requests", type=int, default=256)
    p.add_argument("--serving-slots", type=int, default=4)
    p.add_argument("--serving-prep-threads", type=int, default=min(16, os.cpu_count() or 4))
    return p.parse_args()


def make_workload(name):
    from dxmclib_amd import workloads
    if name == "ct_body":
        return workloads.ct_body()
    if name == "dx_thorax":
        return workloads.dx_thorax()
    if name == "water_pencil":
        return workloads.water_pencil()
    return workloads.whole_body()


def subset(exposures, stride):
    from dxmclib_amd import capi
    idx = list(range(0, len(exposures), stride))
    out = (capi.Exposure * len(idx))()
    for j, i in enumerate(idx):
        out[j] = exposures[i]
    return out


def run_once(dev, w, t, exposures, n, precision, L, tuning):
    """Transport into a fresh tally and download it (parity comparisons)."""
    tally = dev.tally(n)
    try:
        stats = dev.run(w, t, exposures, tally, precision, L, seed=42, tuning=tuning)
        dose, _, var = dev.download(tally, n)
    finally:
        dev.destroy_tally(tally)
    return stats, dose, var


def device_step(dev, w, t, tally, exposures, histories, precision, L, tuning):
    """One job on the device: clear the tally, transport every history,
    normalise to dose. Inputs and the result stay resident in HBM."""
    from dxmclib_amd import capi
    dev.reset_tally(tally)
    stats = dev.run(w, t, exposures, tally, precision, L, seed=42, tuning=tuning)
    dev.normalize(tally, w, capi.OUTPUT_DOSE, histories)
    dev.synchronize()
    return stats


def cpu_baseline(wl, tables, exposures, ext, L, seconds):
    """Oracle on the host cores over a bounded sample: every k-th exposure
    with a reduced history count, sized to about `seconds` of CPU time."""
    import oracle
    maxd = [float(wl.density[wl.material == i].max()) if np.any(wl.material == i) else 0.0 for i in range(len(wl.names))]
    max_e = wl.spec.photon_energy if wl.spec.photon_energy > 0 else wl.spec.tube_voltage
    orc = oracle.Oracle(wl.names, maxd, max_e, 1.0)
    src = tables.source_arrays()
    spectra = []
    if src.get("spectrum_offset") is not None:
        spectra = [spectrum_of(wl.spec)]
    orc.add_source_tables(src, spectra)
    world = oracle.world_view(wl.material, wl.density.astype(np.float64), wl.spacing, ext)
    threads = min(os.cpu_count() or 1, 16)
    stride = max(1, len(exposures) // 36)
    sample = subset(exposures, stride)
    # calibrate on a small probe, then size the sample
    for x in sample:
        x.n_histories = 2000
    t0 = time.perf_counter()
    orc.transport(world, sample, L=L, threads=threads)
    probe = time.perf_counter() - t0
    rate = len(sample) * 2000 / max(probe, 1e-3)
    per = int(max(2000, min(10**6, rate * seconds / len(sample))))
    for x in sample:
        x.n_histories = per
    t0 = time.perf_counter()
    r = orc.transport(world, sample, L=L, threads=threads)
    dt = time.perf_counter() - t0
    # the FP32 restatement on the same sample
    t0 = time.perf_counter()
    r32 = orc.transport(world, sample, L=L, f32=True, threads=threads)
    dt32 = time.perf_counter() - t0
    return {"value": r["histories"] / dt, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} of {len(exposures)} exposures (every {stride}th) x {per} histories, FP64, L={L}, {dt:.1f} s",
            "fp32_value": r32["histories"] / dt32, "host_cpu": host_cpu(), "host_threads": os.cpu_count()}


def host_cpu():
    """Model name of the host CPU (first /proc/cpuinfo entry)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def serving(n_requests, rate, slots, prep_threads):
    """The request stream of the serving configuration: n_requests Poisson
    arrivals (rate per second) of water cubes with a bone sphere (32-256
    voxels a side), DX or CT axial at 80-140 kVp, 1e6-1e8 histories. Each is
    submitted to the request server at its arrival time; latency is arrival
    to result ready."""
    import concurrent.futures as cf
    from dxmclib_amd import backend, capi, workloads
    reqs = workloads.serving_requests(n=n_requests, rate=rate)

    def make(r):
        mat, dens, names = workloads.cube_phantom(r["side"], r["spacing"], r["seed"])
        return backend.host_world(mat, dens, (r["spacing"],) * 3, names), workloads.request_spec(r)

    with backend.Server(slots=slots, prep_threads=prep_threads) as srv, cf.ThreadPoolExecutor(8) as pool:
        made = [pool.submit(make, r) for r in reqs]
        # worlds are generated ahead of their arrival times
        for f in made[:min(16, len(made))]:
            f.result()
        t0 = time.perf_counter()
        ids = []
        for r, f in zip(reqs, made):
            world, spec = f.result()
            delay = r["arrival"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(world, spec, seed=r["seed"]))
        infos = []
        for i in ids:
            srv.wait(i)
            infos.append(srv.collect(i)["info"])
    lat = np.array([x.finished - x.submitted for x in infos])
    wait = np.array([x.admitted - x.prepared for x in infos])
    prep = np.array([x.prepared - x.submitted for x in infos])
    prep_wait = np.array([x.prep_started - x.submitted for x in infos])
    world_build = np.array([x.world_built - x.prep_started for x in infos])
    source_build = np.array([x.source_ready - x.world_built for x in infos])
    tables_build = np.array([x.prepared - x.source_ready for x in infos])
    run = np.array([x.finished - x.admitted for x in infos])
    upload = np.array([x.uploaded - x.admitted for x in infos])
    transport = np.array([x.transported - x.uploaded for x in infos])
    download = np.array([x.downloaded - x.transported for x in infos])
    release = np.array([x.released - x.downloaded for x in infos])
    publish = np.array([x.finished - x.released for x in infos])
    hist = sum(x.histories for x in infos)
    span = max(x.finished for x in infos) - min(x.submitted for x in infos)
    # what the slowest 1 % of requests spend their time on
    tail = lat >= np.percentile(lat, 99)
    sides = [r["side"] for r in reqs]
    tail_phases = {"requests": int(tail.sum()), "voxels_mean": float(np.mean([sides[i] ** 3 for i in np.nonzero(tail)[0]])),
                   "histories_mean": float(np.mean([infos[i].histories for i in np.nonzero(tail)[0]])),
                   "prepare_s": float(prep[tail].mean()), "prepare_wait_s": float(prep_wait[tail].mean()),
                   "world_build_s": float(world_build[tail].mean()), "source_s": float(source_build[tail].mean()),
                   "tables_s": float(tables_build[tail].mean()), "admission_wait_s": float(wait[tail].mean()), "upload_s": float(upload[tail].mean()),
                   "transport_s": float(transport[tail].mean()), "download_s": float(download[tail].mean()),
                   "release_s": float(release[tail].mean()), "publish_s": float(publish[tail].mean())}
    return {"requests": len(infos), "rate_per_s": rate, "slots": slots, "prep_threads": prep_threads, "p99_tail": tail_phases,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()), "prepare_mean_s": float(prep.mean()), "prepare_wait_mean_s": float(prep_wait.mean()),
            "world_build_mean_s": float(world_build.mean()), "source_mean_s": float(source_build.mean()), "tables_mean_s": float(tables_build.mean()),
            "admission_wait_mean_s": float(wait.mean()),
            "device_mean_s": float(run.mean()), "upload_mean_s": float(upload.mean()), "transport_mean_s": float(transport.mean()),
            "download_mean_s": float(download.mean()), "release_mean_s": float(release.mean()),
            "publish_mean_s": float(publish.mean()), "histories": int(hist), "makespan_s": span, "histories_per_s": hist / span,
            "failed": int(sum(1 for x in infos if x.status != 0))}


def measured_traffic(workload, histories_per_launch):
    """HBM bytes per launch (GB) of the transport kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes of this workload (scaled by the
    copy-kernel calibration), per history times this run's launch size, with
    where the figure comes from; None when no profile of the workload is
    committed."""
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "transport_fp32_*.json"))):
        with open(path) as f:
            prof = json.load(f)
        if prof.get("workload") != workload:
            continue
        per = prof["fetch_bytes_per_history"] + prof["write_bytes_per_history"]
        name = os.path.basename(path)
        return per * histories_per_launch / 1e9, f"rocprofv3 FETCH_SIZE+WRITE_SIZE per history (profiles/{name}) x histories per launch"
    return None


def spectrum_of(spec):
    import ctypes as C
    from dxmclib_amd import capi
    n = C.c_uint64()
    lib = capi.host()
    lib.dxmc_host_source_spectrum(C.byref(spec), None, None, 0, C.byref(n))
    e = np.zeros(n.value)
    w = np.zeros(n.value)
    lib.dxmc_host_source_spectrum(C.byref(spec), e.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                  C.byref(n))
    return w, e


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (on several GPUs the transport would shard "
              "by histories, with one tally reduction at the end)", file=sys.stderr)
        sys.exit(2)
    from dxmclib_amd import backend, capi

    t0 = time.perf_counter()
    wl = make_workload(args.config)
    ext = backend.world_extent(wl.material.shape[::-1], wl.spacing)
    maxd = [float(wl.density[wl.material == i].max()) if np.any(wl.material == i) else 0.0 for i in range(len(wl.names))]
    max_e = wl.spec.photon_energy if wl.spec.type == capi.SOURCE_PENCIL else wl.spec.tube_voltage
    tables = backend.HostTables(wl.names, maxd, max_e, 1.0)
    exposures = tables.exposures(wl.spec)
    setup_s = time.perf_counter() - t0
    n = wl.voxels
    L = args.low_energy_correction
    tuning = capi.RunCfg()
    tuning.block_size = args.block_size
    tuning.blocks_per_cu = args.blocks_per_cu
    tuning.regeneration_threshold = args.threshold
    tuning.interaction_threshold = args.interaction_threshold
    tuning.queue_depth = args.queue_depth
    tuning.waves_per_simd = args.waves_per_simd
    tuning.lds_tables = args.lds_tables
    tuning.block_bounds = args.block_bounds
    tuning.flags = 1

    with backend.Device(0) as dev:
        w = dev.upload_world(wl.material, wl.density, wl.spacing, ext, len(wl.names))
        t = dev.upload_tables(tables.desc())
        tune_report = None
        if not args.no_autotune:
            t1 = time.perf_counter()
            # each trial is one launch of the size the timed run launches
            trial = int(min(max(wl.histories, 1 << 20), 1 << 27))
            tuning, rep = dev.autotune(w, t, exposures, capi.PRECISION_FP32, L, trial_histories=trial, base=tuning)
            tuning.flags = 1
            tune_report = {"seconds": time.perf_counter() - t1, "trials": rep.trials, "default_rate": rep.default_rate,
                           "best_rate": rep.best_rate, "regeneration_threshold": tuning.regeneration_threshold,
                           "interaction_threshold": tuning.interaction_threshold, "queue_depth": tuning.queue_depth,
                           "blocks_per_cu": tuning.blocks_per_cu, "waves_per_simd": tuning.waves_per_simd, "lds_tables": tuning.lds_tables,
                           "block_bounds": tuning.block_bounds}
            print(f"autotune: {tune_report}", file=sys.stderr, flush=True)
        tally = dev.tally(n)
        for i in range(args.warmup):
            device_step(dev, w, t, tally, exposures, wl.histories, capi.PRECISION_FP32, L, tuning)
            print(f"warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
        dev.synchronize()
        times, kernel_ms, launches, steps, events = [], 0.0, 0, 0, 0
        for i in range(args.steps):
            dev.synchronize()
            t1 = time.perf_counter()
            stats = device_step(dev, w, t, tally, exposures, wl.histories, capi.PRECISION_FP32, L, tuning)
            times.append(time.perf_counter() - t1)
            kernel_ms += stats.kernel_ms
            launches += stats.launches
            steps += stats.steps
            events += stats.events
            print(f"step {i + 1}/{args.steps}: {times[-1]:.3f} s, {stats.histories / times[-1]:.4e} histories/s", file=sys.stderr, flush=True)
        # the result handed to a host caller crosses PCIe: reported beside value
        t1 = time.perf_counter()
        dev.download(tally, n)
        download_s = time.perf_counter() - t1
        dev.destroy_tally(tally)
        elapsed = sum(times)
        histories = wl.histories * args.steps
        value = histories / elapsed
        alg_bytes = 6.0 * steps + 40.0 * events
        achieved = alg_bytes / (kernel_ms * 1e-3)
        result = {
            "metric": "histories/s at 120kVp CT on 1 GPU (FP32 vs FP64 path); p50/p99 served latency",
            "value": value,
            "unit": "histories/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": 1e3 * elapsed / args.steps,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32 transport, f64 tally",
            "data": "synthetic",
            "config": {"workload": wl.name, "histories": wl.histories, "exposures": len(exposures), "voxels": n,
                       "low_energy_correction": L, "description": wl.description},
            # bound: the roofline the kernel is priced against (algorithmic HBM bytes); limiter: what the counters
            # show holds it below that roof (vector-ALU issue, not bandwidth)
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                         "traffic": None, "traffic_unit": "GB per launch", "traffic_source": None, "limiter": "valu_issue_and_gather_latency"},
            "counters": {"steps_per_history": steps / histories, "events_per_history": events / histories,
                         "kernel_ms_per_launch": kernel_ms / max(launches, 1), "launches_per_step": launches / args.steps},
            "setup_seconds": setup_s,
            "autotune": tune_report,
            "host_download_seconds": download_s,
            "pcie_inclusive_value": wl.histories / (sum(times) / args.steps + download_s),
        }
        result["roofline"]["achieved"] = achieved / 1e9
        result["roofline"]["peak"] = HBM_PEAK / 1e9
        traffic = measured_traffic(wl.name, histories / max(launches, 1))
        if traffic is not None:
            result["roofline"]["traffic"], result["roofline"]["traffic_source"] = traffic
        if not args.no_fp64:
            sub = subset(exposures, args.fp64_stride)
            sub_hist = sum(x.n_histories for x in sub)
            tally = dev.tally(n)
            t1 = time.perf_counter()
            device_step(dev, w, t, tally, sub, sub_hist, capi.PRECISION_FP64, L, tuning)
            dt64 = time.perf_counter() - t1
            dev.destroy_tally(tally)
            # dose maps of both paths on identical streams
            s64, d64, v64 = run_once(dev, w, t, sub, n, capi.PRECISION_FP64, L, tuning)
            s32, d32, v32 = run_once(dev, w, t, sub, n, capi.PRECISION_FP32, L, tuning)
            hot = d64 > 0.1 * d64.max()
            err = np.abs(d32[hot] - d64[hot]) / d64[hot]
            sig = np.sqrt(v64[hot]) / d64[hot]
            result["fp64"] = {"value": sub_hist / dt64, "unit": "histories/s", "histories": sub_hist,
                              "sample": "all exposures" if args.fp64_stride == 1 else f"every {args.fp64_stride}th exposure",
                              "fp32_speedup": value / (sub_hist / dt64),
                              "fp32_vs_fp64_total_rel": float(abs(d32.sum() - d64.sum()) / d64.sum()),
                              "hot_voxels": int(hot.sum()),
                              "fp32_vs_fp64_hot_rel_mean": float(err.mean()) if err.size else None,
                              "fp32_vs_fp64_hot_rel_p99": float(np.percentile(err, 99)) if err.size else None,
                              "fp32_vs_fp64_hot_rel_max": float(err.max()) if err.size else None,
                              "fp32_vs_fp64_hot_frac_below_1e-3": float(np.mean(err < 1e-3)) if err.size else None,
                              "fp64_hot_rel_sigma_mean": float(sig.mean()) if sig.size else None,
                              # the FP32 difference in units of the FP64 statistical uncertainty
                              "fp32_vs_fp64_hot_err_over_sigma_p99": float(np.percentile(err / sig, 99)) if err.size else None}
            print(f"fp64: {sub_hist / dt64:.4e} histories/s", file=sys.stderr, flush=True)
        dev.destroy_tables(t)
        dev.destroy_world(w)

    if not args.no_serving:
        result["serving"] = serving(args.serving_requests, 20.0, args.serving_slots, args.serving_prep_threads)
        print(f"serving: {result['serving']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(wl, tables, exposures, ext, L, args.cpu_seconds)
        print(f"cpu: {result['cpu_baseline']['value']:.4e} histories/s", file=sys.stderr, flush=True)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
