#!/usr/bin/env python
"""Likelihood evaluations per second of the MI355X engine on one GPU.

Workload (one "step" = one likelihood evaluation, the plain single-model
GPU emission configuration): 100 levels log-spaced 1e2..1e-6 bar, poly 6 1
temperature profile, 6 absorbers with 30 T x 9 p log10 cross-section tables
plus H2-H2 / H2-He CIA and H2 Rayleigh, 400,000 wavenumbers at 0.025 cm^-1,
300 photometric bins each with a seeded 50-point bandpass.  Tables are
synthetic (seeded Lorentzian line lists), resident in HBM before timing.

Prints one JSON line with the throughput, the roofline of the dominant
kernel (fused opacity + radiative transfer, timed with HIP events on its own
stream) and a CPU baseline timed on this host's cores.
"""

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "likelihood evals/sec (batched+adjoint vs plain GPU) at ≤10 ppm spectral error"
HBM_PEAK_GBS = 8000.0
FULL_N = 400_000
SAMPLE_N = 20_000


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(n_wavenumbers, backend):
    from bear_amd import synth
    return synth.emission_c2(backend=backend, n_wavenumbers=n_wavenumbers)


def open_with_injected_data(case, scratch):
    """Retrieval whose observed data are the engine's own model at the case's
    truth parameters plus 2 % noise; returns (retrieval, nearby live point)."""
    from bear_amd import synth
    r = synth.open_retrieval_in_memory(case, scratch)
    truth = case.parameters.copy()
    synth.inject_data(case, r.computeModel(truth, False).spectrum_obs)
    r.close()
    r = synth.open_retrieval_in_memory(case, scratch)
    return r, synth.perturbed(truth, case, seed=5)


def isothermal_check(r, params):
    """Size-independent property at full size: an isothermal atmosphere emits
    pi B(T) 1e-3 scale at every wavenumber (returns the max relative error)."""
    from bear_amd import synth
    p = params.copy()
    nb_mix = 6
    p[3 + nb_mix + 1:3 + nb_mix + 7] = 1.0
    out = r.computeModel(p, True)
    nu, lam = r.spectral_grid()
    h, c, kb = synth.PLANCK_H, synth.LIGHT_C, synth.BOLTZMANN_K
    t = p[3 + nb_mix]
    b = 2 * h * c * c * nu ** 3 / np.expm1(h * c * nu / (kb * t))
    scale = (synth.RADIUS_JUPITER / p[2]) ** 2 * p[1]
    expected = np.pi * b * 1e-3 * scale / lam ** 2 * 1e4
    return float(np.max(np.abs(out.spectrum - expected) / expected))


def traffic_from_profiles(kernel="emissionKernelCached"):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, KiB per dispatch,
    scaled by the copy-kernel calibration recorded with them)."""
    path = os.path.join(ROOT, "profiles", "emission_c2_counters.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        for name, v in d.items():
            if name.startswith(kernel):
                return float(v["hbm_bytes_per_launch"])
    except Exception:
        pass
    return None


def batched_leg(r, points, batch_steps, plain_rate):
    """Batches of live points per launch sequence at each table precision:
    throughput, speedup over the single-point path, and the error against the
    FP64 single-point path (pinned to the oracle by the parity tests) on the
    first four points: high-resolution spectrum and bins (relative),
    log-likelihood (absolute)."""
    import torch
    pts = np.array(points)
    ref = [r.computeModel(p, True) for p in pts[:4]]
    ref_ll = [r.computeLikelihood(p) for p in pts[:4]]
    out = {"batch": len(pts)}
    for precision in ("fp64", "fp32", "fp16", "bf16"):
        r.set_table_precision(precision)
        r.computeLikelihoodBatch(pts)                   # tuning + graph capture
        r.computeLikelihoodBatch(pts)
        r.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(batch_steps):
            ll = r.computeLikelihoodBatch(pts)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        st = r.batch_stats()
        r.set_profiling(False)
        spec, bins, _ = r.computeModelBatch(pts[:4], True)
        err_spec = max(float(np.max(np.abs(spec[k] - ref[k].spectrum) / np.abs(ref[k].spectrum))) for k in range(4))
        err_bins = max(float(np.max(np.abs(bins[i][k] - ref[k].spectrum_obs[i]) / np.abs(ref[k].spectrum_obs[i])))
                       for k in range(4) for i in range(len(bins)))
        err_ll = max(abs(float(ll[k]) - ref_ll[k]) for k in range(4))
        rate = batch_steps * len(pts) / elapsed
        out[precision] = {"evals_per_s": round(rate, 1), "ms_per_batch": round(elapsed / batch_steps * 1e3, 3),
                          "gpu_ms_per_batch": round(st["gpu_ms_total"] / batch_steps, 3),
                          "speedup_vs_single": round(rate / plain_rate, 2),
                          "max_spectrum_rel_err": err_spec, "max_bin_rel_err": err_bins, "max_abs_dloglike": err_ll,
                          "launch": {k: st[k] for k in ("block", "tile", "grid", "used_graph")}}
        log(f"batch {precision}: {rate:.1f} evals/s, spectrum err {err_spec:.2e}, |dlogL| {err_ll:.2e}")
    r.set_table_precision("fp64")
    # value and gradient (hand-written adjoint) of the same batch
    r.gradientBatch(pts)
    r.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(batch_steps):
        ll, grad = r.gradientBatch(pts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = r.batch_stats()
    r.set_profiling(False)
    rate = batch_steps * len(pts) / elapsed
    # spot check at full size: central differences (step 1e-4 relative) of
    # three components of the first point through the single-point path
    p0 = pts[0]
    fd_rel = {}
    for k, label in ((0, "log_g"), (3, "first_mixing_ratio"), (9, "first_temperature")):
        h = 1e-4 * abs(p0[k])
        e = np.zeros_like(p0)
        e[k] = h
        fd = (r.computeLikelihood(p0 + e) - r.computeLikelihood(p0 - e)) / (2 * h)
        fd_rel[label] = abs(grad[0][k] - fd) / abs(fd)
    out["adjoint"] = {"evals_per_s": round(rate, 1), "ms_per_batch": round(elapsed / batch_steps * 1e3, 3),
                      "gpu_ms_per_batch": round(st["gpu_ms_total"] / batch_steps, 3),
                      "cost_vs_fp64_forward": round(rate and out["fp64"]["evals_per_s"] / rate, 3),
                      "points_per_chunk": st["adjoint_chunk"],
                      "table_dtype": "f64", "fd_check_rel_diff": fd_rel}
    log(f"adjoint: {rate:.1f} value+gradient/s, {out['adjoint']['cost_vs_fp64_forward']}x the FP64 forward, "
        f"FD check {fd_rel}")
    return out


def cpu_baseline(threads, scratch):
    """The reference's own CPU path (oracle/_ref build) on a bounded sample:
    the same configuration at 20,000 wavenumbers, scaled to 400,000 by the
    per-wavenumber cost; the oracle restatement stands in when the reference
    build is absent."""
    from bear_amd import synth
    from oracle import api as oracle_api
    case = build_case(SAMPLE_N, "numpy")
    folder = synth.write_case(case, os.path.join(scratch, "cpu_sample"))
    truth = case.parameters.copy()
    o = oracle_api.EmissionOracle(folder)
    synth.inject_data(case, o.evaluate(truth)["bins"])
    folder = synth.write_case(case, folder)
    params = synth.perturbed(truth, case, seed=5)
    with open(os.path.join(folder, "case.txt")) as f:
        lines = [l for l in f.read().splitlines() if not l.startswith("parameters ")]
    lines.append(f"parameters {params.size} " + " ".join(repr(float(v)) for v in params))
    with open(os.path.join(folder, "case.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    n_evals = 12
    if os.path.exists(oracle_api.REF_BIN):
        out = oracle_api.run_reference(folder, os.path.join(scratch, "cpu_out"), n_timed=n_evals, threads=threads)
        t = float(np.median(out["timings"]))
        kind = "reference"
    else:
        os.environ["OMP_NUM_THREADS"] = str(threads)
        o = oracle_api.EmissionOracle(folder)
        o.evaluate(params)
        ts = []
        for _ in range(n_evals):
            t0 = time.perf_counter()
            o.evaluate(params)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        kind = "port"
    n_sample = o.nb_points if kind == "port" else oracle_api.EmissionOracle(folder).nb_points
    return {
        "value": (1.0 / t) * n_sample / FULL_N,
        "unit": "evals/s",
        "cores": threads,
        "kind": kind,
        "sample": f"same emission configuration at {n_sample} wavenumbers (1000-1500 cm^-1), median of "
                  f"{n_evals} evaluations = {t * 1e3:.1f} ms, scaled by {n_sample}/{FULL_N} wavenumbers",
    }


CONFIG_WORKLOADS = {
    # name: (builder, live points, table precisions of the batched legs, what a result is)
    "transmission_c3": ("transmission_c3", 256, ("fp64", "fp32", "fp16"), "ppm of transit depth"),
    "emission_c4": ("emission_c4", 64, ("fp64",), "W m^-2 micron^-1"),
    "eclipse_c5": ("eclipse_c5", 512, ("fp64", "fp16", "bf16"), "ppm of eclipse depth"),
}


def config_workload(name, batch, steps):
    """The other configurations at their full sizes (one JSON line): the FP64
    single-point rate, each table precision's batched rate and its error
    against the FP64 single-point path on the first four points (bins as
    absolute differences in the result's unit, high-resolution spectrum
    relative, log-likelihood absolute), and the FP64 value + gradient rate
    (hand-written adjoint) with its central-difference check on one point."""
    import torch
    from bear_amd import synth
    builder, default_batch, precisions, unit = CONFIG_WORKLOADS[name]
    nb = batch or default_batch
    with tempfile.TemporaryDirectory() as scratch:
        log(f"generating the {name} tables")
        t0 = time.time()
        case = getattr(synth, builder)(backend="torch")
        log(f"tables ready in {time.time() - t0:.1f} s; building the engine")
        r, params = open_with_injected_data(case, os.path.join(scratch, "gpu"))
        points = np.array([params] + [synth.perturbed(params, case, seed=2000 + k) for k in range(nb - 1)])
        del case
        st = r.stats()
        out = {"workload": name, "points": nb, "nb_spectral_points": st["nb_spectral_points"],
               "levels": st["nb_levels"], "tabulated_species": st["nb_species"], "bins": st["nb_bins"],
               "result_unit": unit, "data": "synthetic"}
        for _ in range(3):
            r.computeLikelihood(params)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            r.computeLikelihood(params)
        torch.cuda.synchronize()
        single = 20 / (time.perf_counter() - t0)
        out["single_fp64_evals_per_s"] = round(single, 2)
        log(f"single point FP64: {single:.1f} evals/s")
        ref = [r.computeModel(p, True) for p in points[:4]]
        ref_ll = [r.computeLikelihood(p) for p in points[:4]]
        legs = {}
        for precision in precisions:
            r.set_table_precision(precision)
            r.computeLikelihoodBatch(points)                 # tuning + graph capture
            r.computeLikelihoodBatch(points)
            r.set_profiling(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                ll = r.computeLikelihoodBatch(points)
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t0
            bst = r.batch_stats()
            r.set_profiling(False)
            spec, bins, _ = r.computeModelBatch(points[:4], True)
            err_spec = max(float(np.max(np.abs(spec[k] - ref[k].spectrum) / np.abs(ref[k].spectrum)))
                           for k in range(4))
            err_bins = max(float(np.max(np.abs(bins[i][k] - ref[k].spectrum_obs[i])))
                           for k in range(4) for i in range(len(bins)))
            err_ll = max(abs(float(ll[k]) - ref_ll[k]) for k in range(4))
            rate = steps * nb / elapsed
            legs[precision] = {"evals_per_s": round(rate, 1), "gpu_ms_per_batch": round(bst["gpu_ms_total"] / steps, 3),
                               "speedup_vs_single_fp64": round(rate / single, 2),
                               "max_spectrum_rel_err": err_spec, "max_bin_abs_err": err_bins,
                               "max_abs_dloglike": err_ll,
                               "launch": {k: bst[k] for k in ("block", "tile", "grid", "point_group", "used_graph")}}
            log(f"{name} batch {precision}: {rate:.1f} evals/s, bins err {err_bins:.3e} {unit}, |dlogL| {err_ll:.2e}")
        out["batched"] = legs
        if "fp64" in precisions:
            r.set_table_precision("fp64")
            r.gradientBatch(points)
            r.set_profiling(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                ll, grad = r.gradientBatch(points)
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t0
            bst = r.batch_stats()
            r.set_profiling(False)
            rate = steps * nb / elapsed
            p0 = points[0]
            fd_rel = {}
            for k in range(p0.size):
                h = 1e-4 * (abs(p0[k]) if p0[k] != 0 else 1.0)
                e = np.zeros_like(p0)
                e[k] = h
                fd = (r.computeLikelihood(p0 + e) - r.computeLikelihood(p0 - e)) / (2 * h)
                fd_rel[k] = float(abs(grad[0][k] - fd) / max(abs(fd), 1e-300))
            out["adjoint"] = {"evals_per_s": round(rate, 1), "gpu_ms_per_batch": round(bst["gpu_ms_total"] / steps, 3),
                              "cost_vs_fp64_forward": round(legs["fp64"]["evals_per_s"] / rate, 3),
                              "points_per_chunk": bst["adjoint_chunk"],
                              "fd_rel_diff_per_parameter": fd_rel}
            log(f"{name} value + gradient: {rate:.1f}/s, {out['adjoint']['cost_vs_fp64_forward']}x the forward; "
                f"max FD rel diff {max(fd_rel.values()):.2e}")
        r.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=128, help="live points per batch in the batched leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="emission_c2", choices=["emission_c2"] + sorted(CONFIG_WORKLOADS),
                    help="emission_c2: the headline line; the others print their own line (batch size "
                         "from --batch, default theirs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the likelihood path on a single GPU "
              "(the multi-GPU sharding is a design, not a run); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    from bear_amd import api
    if args.workload != "emission_c2":
        explicit = any(a == "--batch" or a.startswith("--batch=") for a in sys.argv[1:])
        config_workload(args.workload, args.batch if explicit else 0, max(1, args.batch_steps // 4))
        return

    with tempfile.TemporaryDirectory() as scratch:
        log("generating the 400,000-wavenumber tables")
        t0 = time.time()
        case = build_case(FULL_N, "torch")
        log(f"tables ready in {time.time() - t0:.1f} s; building the engine")
        r, params = open_with_injected_data(case, os.path.join(scratch, "gpu"))
        from bear_amd import synth
        points = [params] + [synth.perturbed(params, case, seed=1000 + k) for k in range(max(args.batch, 1) - 1)]
        del case
        iso_err = isothermal_check(r, params)
        log(f"isothermal identity at full size: max rel err {iso_err:.2e}")

        for _ in range(args.warmup):
            r.computeLikelihood(params)
        r.set_profiling(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ll = r.computeLikelihood(params)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        st = r.stats()
        r.set_profiling(False)
        log(f"{args.steps} evaluations in {elapsed:.3f} s, log L = {ll:.6f}")

        n = st["nb_spectral_points"]
        kernel_s = st["kernel_ms_total"] / 1e3 / max(st["kernel_launches"], 1)
        # compulsory bytes of one launch: every distinct table row once, the
        # wavenumbers, the flux written; per-level inputs are a few tens of KB
        algorithmic = st["distinct_table_rows"] * n * st["table_value_bytes"] + n * 8 + n * 8 \
            + st["nb_levels"] * (32 + st["nb_species"] * 48)
        achieved = algorithmic / kernel_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                    "kernel": "emissionKernelCached<8>", "kernel_ms": round(kernel_s * 1e3, 4),
                    "algorithmic_bytes": int(algorithmic), "distinct_table_rows": st["distinct_table_rows"]}

        batched = None
        if args.batch > 0:
            batched = batched_leg(r, points[:args.batch], args.batch_steps, args.steps / elapsed)

        cpu = None
        if not args.no_cpu_baseline:
            threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
            log(f"CPU baseline on {threads} threads")
            try:
                cpu = cpu_baseline(threads, scratch)
            except Exception as e:  # the GPU figure stands without it
                log(f"CPU baseline failed: {e}")

    value = args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "emission_c2: one emission likelihood evaluation, 100 levels, 6 absorbers + "
                               "H2-H2/H2-He CIA + H2 Rayleigh, 400,000 wavenumbers, 30T x 9p tables, "
                               "300 bandpass bins", "nb_spectral_points": n, "batch": 1,
                   "table_dtype": "f64", "levels": st["nb_levels"], "bins": st["nb_bins"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "checks": {"isothermal_max_rel_err": iso_err},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
