"""Writes the golden fixtures of the emission path from the reference's own
CPU code (the build of oracle/ref.mk, compiled from /root/reference/src).

For each seeded case (tests/cases.py) the observed data are injected from
the reference build's own model at the truth parameters, then the reference
build evaluates every parameter vector of the case.  The fixtures hold the
inputs that matter (parameter vectors, injected data, a digest of the table
files) and the reference's outputs: sampled grid, band-edge and convolution
indices, atmosphere, spectra, bins and log-likelihoods.

    make -f oracle/ref.mk && python tests/golden/make_golden.py [case ... | direct]

(all cases but the wide-band ones by default; "direct": the direct models)
"""

import hashlib
import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import api  # noqa: E402
import cases  # noqa: E402


def table_digest(folder):
    h = hashlib.sha256()
    for dirpath, _, files in sorted(os.walk(os.path.join(folder, "opacity"))):
        for f in sorted(files):
            with open(os.path.join(dirpath, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()


def reference_bins(folder, params, work):
    import json
    meta = json.load(open(os.path.join(folder, "case.json")))
    meta_txt = open(os.path.join(folder, "case.txt")).read().splitlines()
    meta_txt = [l for l in meta_txt if not l.startswith("parameters ")]
    meta_txt.append(f"parameters {len(params)} " + " ".join(repr(float(v)) for v in params))
    with open(os.path.join(folder, "case.txt"), "w") as f:
        f.write("\n".join(meta_txt) + "\n")
    out = api.run_reference(folder, work)
    bins = []
    for i in range(len(meta["observations"])):
        b = out[f"obs{i}_bins"]
        bins.append(b[::-1].copy() if out[f"obs{i}_ascending"][0] else b)
    return out, bins


def main():
    if not os.path.exists(api.REF_BIN):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    import json
    tmp = tempfile.mkdtemp()
    try:
        names = sys.argv[1:] or [n for n in cases.CASES if not n.endswith("_wide")]
        for name in names:
            folder = os.path.join(tmp, name)
            work = os.path.join(tmp, name + "_out")
            transmission = name.startswith("transmission")
            # the reference build has no transmission integral: its data are
            # injected from the oracle, and its fixtures hold the reference's
            # atmosphere and opacity (the integral is pinned by closed forms)
            model_fn = None if transmission else (lambda f, p: reference_bins(f, p, work)[1])
            folder, params = cases.build(name, folder, model_fn=model_fn)
            fixture = {"parameters": np.array(params), "table_digest": np.array(table_digest(folder))}
            meta = json.load(open(os.path.join(folder, "case.json")))
            for k, p in enumerate(params):
                shutil.rmtree(work, ignore_errors=True)
                out, _ = reference_bins(folder, p, work) if not transmission else (None, None)
                if transmission:
                    shutil.rmtree(work, ignore_errors=True)
                    txt = [l for l in open(os.path.join(folder, "case.txt")).read().splitlines()
                           if not l.startswith("parameters ")]
                    txt.append(f"parameters {len(p)} " + " ".join(repr(float(v)) for v in p))
                    open(os.path.join(folder, "case.txt"), "w").write("\n".join(txt) + "\n")
                    out = api.run_reference(folder, work)
                pre = f"p{k}_"
                for key in ("temperature", "altitude"):
                    fixture[pre + key] = out[key]
                if "cloud_optical_depth" in out:
                    # the clouds: scale heights, layer optical depths and (transmission)
                    # level extinctions on every 10th wavenumber
                    fixture[pre + "scale_height"] = out["scale_height"]
                    fixture[pre + "cloud_optical_depth"] = out["cloud_optical_depth"][:, ::10]
                    if "cloud_extinction" in out:
                        fixture[pre + "cloud_extinction"] = out["cloud_extinction"][:, ::10]
                if transmission:
                    if name == "transmission_plain" and k == 0:
                        # extinction inputs on every 5th wavenumber
                        fixture[pre + "absorption_coeff"] = out["absorption_coeff"][::5]
                        fixture[pre + "scattering_coeff"] = out["scattering_coeff"][::5]
                    continue
                fixture[pre + "loglike"] = out["loglike"]
                # the high-resolution spectrum in full for the first vector, every 10th point otherwise
                fixture[pre + "spectrum_flux"] = out["spectrum_flux"] if k == 0 else out["spectrum_flux"][::10]
                fixture[pre + "spectrum"] = out["spectrum"][::10]
                for i in range(len(meta["observations"])):
                    fixture[pre + f"obs{i}_bins"] = out[f"obs{i}_bins"]
            for key in ("wavenumber", "index_list", "pressure", "number_densities"):
                fixture[key] = out[key]
            for i in range(len(meta["observations"])):
                for key in ("edge_indices", "index_range", "conv_intervals", "sigma", "filter",
                            "filter_norm", "data", "error", "weight", "ascending"):
                    if f"obs{i}_{key}" in out:
                        fixture[f"obs{i}_{key}"] = out[f"obs{i}_{key}"]
            path = os.path.join(ROOT, "tests", "golden", f"{name if not name.startswith('emission_') else name}.npz")
            if name.startswith("tiny"):
                path = os.path.join(ROOT, "tests", "golden", "emission_" + name + ".npz")
            np.savez_compressed(path, **fixture)
            print(path, os.path.getsize(path), "bytes")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def make_direct():
    """Fixtures of pyBeAR's direct models: per model, the calcSpectrum
    arguments of cases.direct_case and the reference build's atmosphere,
    coefficients (every 7th wavenumber) and, for emission and eclipse, the
    spectrum."""
    tmp = tempfile.mkdtemp()
    try:
        for model in cases.DIRECT_MODELS:
            folder, inputs = cases.direct_case(model, os.path.join(tmp, model))
            fixture = {"table_digest": np.array(table_digest(folder))}
            for k, d in enumerate(inputs):
                cases.write_direct_input(folder, d)
                out = api.run_reference(folder, os.path.join(tmp, f"{model}_{k}"))
                pre = f"d{k}_"
                for key in ("g", "a", "b", "variable_gravity"):
                    fixture[pre + key] = np.array(d[key])
                fixture[pre + "pressure"], fixture[pre + "temperature"] = d["p"], d["t"]
                fixture[pre + "mixing_ratios"] = d["x"]
                for key in ("altitude", "number_densities"):
                    fixture[pre + key] = out[key]
                fixture[pre + "absorption_coeff"] = out["absorption_coeff"][::7]
                fixture[pre + "scattering_coeff"] = out["scattering_coeff"][::7]
                if "spectrum" in out:
                    fixture[pre + "spectrum"] = out["spectrum"]
                if "cloud_extinction" in out and model == "transmission":
                    fixture[pre + "cloud_extinction"] = out["cloud_extinction"][:, ::7]
            fixture["wavenumber"] = out["wavenumber"]
            path = os.path.join(ROOT, "tests", "golden", f"direct_{model}.npz")
            np.savez_compressed(path, **fixture)
            print(path, os.path.getsize(path), "bytes")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    if sys.argv[1:] == ["direct"]:
        make_direct()
    else:
        main()
