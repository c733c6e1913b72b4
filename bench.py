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
import hashlib
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "likelihood evals/sec (batched+adjoint vs plain GPU) at ≤10 ppm spectral error"
HBM_PEAK_GBS = 8000.0
FULL_N = 400_000
CPU_THREADS = 16            # the GPU box's CPU share per GPU
CPU_EVALS = 20
SPECTRUM_BAR = 1e-5         # relative, emission spectra and bins
LOGLIKE_BAR = 1e-6          # absolute
# the sources of the single-point kernel: a counter profile is valid for these only
KERNEL_SOURCES = ("bear_amd/csrc/engine.hip", "bear_amd/csrc/device_math.h", "bear_amd/csrc/engine.h")


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


def transcendentals(st, transmission, evals_per_s):
    """Algorithmic transcendental count of one evaluation and the rate it
    implies: one 10^x per (level, tabulated species, wavenumber) for the
    opacity interpolation, and per (level, wavenumber) two exponentials
    (the short-characteristics attenuation of both sweeps) and one Planck
    exponential for emission, or one exp(-tau) per (tangent path,
    wavenumber) for a transit spectrum."""
    nl, ns, n = st["nb_levels"], st["nb_species"], st["nb_spectral_points"]
    exp10 = nl * ns * n
    rt = (nl - 1) * n if transmission else 2 * nl * n
    planck = 0 if transmission else nl * n
    total = exp10 + rt + planck
    return {"per_eval": int(total), "exp10": int(exp10), "rt_exp": int(rt), "planck": int(planck),
            "rate_per_s": float(f"{total * evals_per_s:.4g}")}


def kernel_source_digest():
    """sha256 over the single-point kernel's source files (their contents)."""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def traffic_from_profiles(kernel="emissionKernelCached"):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, KiB per dispatch,
    scaled by the copy-kernel calibration recorded with them), provided the
    summary was taken on these kernel sources; (bytes or None, note)."""
    path = os.path.join(ROOT, "profiles", "emission_c2_counters.json")
    if not os.path.exists(path):
        return None, "no counter profile"
    d = json.load(open(path))
    digest = kernel_source_digest()
    if d.get("kernel_source_digest") != digest:
        return None, (f"counter profile taken on kernel sources {d.get('kernel_source_digest')}, "
                      f"these are {digest}: not attached")
    for name, v in d.items():
        if name.startswith(kernel):
            return float(v["hbm_bytes_per_launch"]), f"counters of kernel sources {digest}"
    return None, "kernel not in the counter profile"


def _counter_rows(folder):
    """(kernel name without its arguments, counter, value, dispatch id) of
    every rocprofv3 counter_collection.csv under folder."""
    import csv
    import glob
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as fh:
            for row in csv.DictReader(fh):
                name = row["Kernel_Name"].replace("(anonymous namespace)::", "").split("(")[0].strip()
                if name.startswith("void "):
                    name = name[5:]
                if name.startswith("bear_amd::"):
                    name = name[len("bear_amd::"):]
                yield name, row["Counter_Name"], float(row["Counter_Value"]), row.get("Dispatch_Id", "")


def _per_dispatch(folder, counter):
    """kernel -> (dispatches, mean counter value per dispatch)."""
    acc = {}
    for name, c, v, d in _counter_rows(folder):
        if c != counter:
            continue
        e = acc.setdefault(name, {})
        e[d] = e.get(d, 0.0) + v                 # per-XCD / per-channel rows of one dispatch summed
    return {k: (len(v), sum(v.values()) / len(v)) for k, v in acc.items()}


def summarize_counters(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes):
    """The counter summary bench.py attaches as roofline.traffic: per kernel
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) from two rocprofv3 --pmc
    runs of the bench, scaled by what the same counters read for a plain
    device-to-device copy of copy_bytes, with the kernel sources' digest."""
    import subprocess
    copy_fetch = _per_dispatch(copy_fetch_dir, "FETCH_SIZE")
    copy_write = _per_dispatch(copy_write_dir, "WRITE_SIZE")
    ck = [k for k in copy_fetch if "copyBuffer" in k or "copy" in k.lower()]
    if not ck:
        raise SystemExit(f"no copy kernel in {copy_fetch_dir}: {sorted(copy_fetch)}")
    f_kib, w_kib = copy_fetch[ck[0]][1], copy_write[ck[0]][1]
    fetch_factor, write_factor = copy_bytes / (f_kib * 1024), copy_bytes / (w_kib * 1024)
    try:
        commit = subprocess.run(["git", "rev-parse", "HEAD"], cwd=ROOT, capture_output=True, text=True).stdout.strip()
    except OSError:
        commit = "unknown"
    out = {"kernel_source_digest": kernel_source_digest(), "commit_before_profile": commit,
           "calibration": {"copy_kernel": ck[0], "copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes,
                           "FETCH_SIZE_KiB": f_kib, "WRITE_SIZE_KiB": w_kib,
                           "fetch_factor": fetch_factor, "write_factor": write_factor}}
    fetch, write = _per_dispatch(fetch_dir, "FETCH_SIZE"), _per_dispatch(write_dir, "WRITE_SIZE")
    for k in sorted(fetch, key=lambda k: -fetch[k][1]):
        if k not in write or "Kernel" not in k:         # this library's kernels only
            continue
        rd, wr = fetch[k][1] * 1024 * fetch_factor, write[k][1] * 1024 * write_factor
        out[k] = {"dispatches": fetch[k][0], "FETCH_SIZE_KiB": fetch[k][1], "WRITE_SIZE_KiB": write[k][1],
                  "hbm_bytes_read": rd, "hbm_bytes_written": wr, "hbm_bytes_per_launch": rd + wr}
    return out


def batched_leg(r, points, batch_steps, plain_rate):
    """Batches of live points per launch sequence at each table precision:
    throughput, speedup over the single-point path, and the error against the
    FP64 single-point path (equal to the oracle within 5e-11 relative at this
    size, tests/test_gpu_full_size.py) on the first four points:
    high-resolution spectrum and bins (relative), log-likelihood (absolute),
    with whether each leg meets the spectral bar (1e-5 relative) and the
    log-likelihood bar (1e-6 absolute)."""
    import torch
    pts = np.array(points)
    ref = [r.computeModel(p, True) for p in pts[:4]]
    ref_ll = [r.computeLikelihood(p) for p in pts[:4]]
    out = {"batch": len(pts)}
    for precision in ("fp64", "fp32", "fp16", "bf16", "int16"):
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
                          "meets_spectral_bar": bool(err_spec <= SPECTRUM_BAR and err_bins <= SPECTRUM_BAR),
                          "meets_loglike_bar": bool(err_ll <= LOGLIKE_BAR),
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed_reference(folder, scratch, threads, n_evals):
    """Median seconds of one evaluation on the reference build (the oracle
    restatement where the build is absent): table rows preloaded by an
    untimed first evaluation; (seconds, kind)."""
    from oracle import api as oracle_api
    if os.path.exists(oracle_api.REF_BIN):
        out = oracle_api.run_reference(folder, os.path.join(scratch, f"cpu_out_{threads}"), n_timed=n_evals,
                                       threads=threads)
        return float(np.median(out["timings"])), "reference"
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = oracle_api.EmissionOracle(folder)
    params = np.array(json.load(open(os.path.join(folder, "case.json")))["parameters"])
    o.evaluate(params)
    ts = []
    for _ in range(n_evals):
        t0 = time.perf_counter()
        o.evaluate(params)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), "port"


def write_cpu_case(case, params, folder):
    """The case in the reference's file formats with `params` as its
    parameter vector (the reference build evaluates case.txt's)."""
    from bear_amd import synth
    keep = case.parameters
    case.parameters = np.asarray(params, dtype=np.float64)
    try:
        return synth.write_case(case, folder)
    finally:
        case.parameters = keep


def cpu_baseline(case, params, scratch):
    """The reference's own CPU path (oracle/_ref build, compiled here from
    the reference's sources) on this workload in full: every wavenumber, the
    same tables, observations and live point as the GPU line, CPU_THREADS
    threads, the median of CPU_EVALS evaluations (no scaling).  Beside it the
    plain CPU configuration (emission_c1: 60 levels, 2 absorbers, 20,001
    wavenumbers, 100 bins) at CPU_THREADS threads and at one thread."""
    from bear_amd import synth
    folder = write_cpu_case(case, params, os.path.join(scratch, "cpu_c2"))
    log(f"CPU baseline: full workload on {CPU_THREADS} threads, {CPU_EVALS} evaluations")
    t, kind = timed_reference(folder, scratch, CPU_THREADS, CPU_EVALS)
    c1 = synth.emission_c1()
    c1_folder = synth.write_case(c1, os.path.join(scratch, "cpu_c1"))
    t1_many, _ = timed_reference(c1_folder, scratch, CPU_THREADS, CPU_EVALS)
    t1_one, _ = timed_reference(c1_folder, scratch, 1, CPU_EVALS)
    log(f"CPU baseline: {t * 1e3:.1f} ms per evaluation ({kind}); emission_c1 {t1_many * 1e3:.2f} ms on "
        f"{CPU_THREADS} threads, {t1_one * 1e3:.2f} ms on one")
    return {
        "value": round(1.0 / t, 4),
        "unit": "evals/s",
        "cores": CPU_THREADS,
        "kind": kind,
        "sample": f"the whole workload (all {FULL_N} wavenumbers, the same tables, observations and live point), "
                  f"median of {CPU_EVALS} evaluations = {t * 1e3:.1f} ms, {CPU_THREADS} threads "
                  f"(the box's CPU share per GPU; nproc {os.cpu_count()}), {cpu_model()}",
        "emission_c1": {"evals_per_s_threads": round(1.0 / t1_many, 2), "evals_per_s_one_thread": round(1.0 / t1_one, 2),
                        "threads": CPU_THREADS, "evaluations": CPU_EVALS},
    }


CONFIG_WORKLOADS = {
    # name: (builder, live points, table precisions of the batched legs, what a result is)
    "transmission_c3": ("transmission_c3", 256, ("fp64", "fp32", "fp16", "int16"), "ppm of transit depth"),
    "emission_c4": ("emission_c4", 64, ("fp64",), "W m^-2 micron^-1"),
    "eclipse_c5": ("eclipse_c5", 512, ("fp64", "fp16", "bf16", "int16"), "ppm of eclipse depth"),
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
        best = max(v["evals_per_s"] for v in legs.values())
        out["transcendentals"] = transcendentals(st, name.startswith("transmission"), best)
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
    ap.add_argument("--workload", default="emission_c2", choices=["emission_c2"] + sorted(CONFIG_WORKLOADS),
                    help="emission_c2: the headline line; the others print their own line (batch size "
                         "from --batch, default theirs)")
    ap.add_argument("--summarize-counters", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="write the counter summary (JSON, stdout) from rocprofv3 --pmc output folders; no GPU")
    ap.add_argument("--copy-bytes", type=int, default=2 ** 31, help="bytes of the calibration copy")
    args = ap.parse_args()
    if args.summarize_counters:
        print(json.dumps(summarize_counters(*args.summarize_counters, args.copy_bytes), indent=1))
        return
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
        traffic, traffic_note = traffic_from_profiles()
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_note": traffic_note,
                    "kernel": "emissionKernelCached<8>", "kernel_ms": round(kernel_s * 1e3, 4),
                    "algorithmic_bytes": int(algorithmic), "distinct_table_rows": st["distinct_table_rows"],
                    "kernel_source_digest": kernel_source_digest()}

        batched = None
        if args.batch > 0:
            batched = batched_leg(r, points[:args.batch], args.batch_steps, args.steps / elapsed)

        r.close()
        cpu = None
        if not args.no_cpu_baseline:
            try:
                cpu = cpu_baseline(case, params, scratch)
            except Exception as e:  # the GPU figure stands without it
                log(f"CPU baseline failed: {e}")
        del case

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
        "transcendentals": transcendentals(st, False, value),
        "checks": {"isothermal_max_rel_err": iso_err},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
