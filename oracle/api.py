"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: imported only by the tests, the smoke check and the
CPU-baseline leg of bench.py.  Also runs the reference build
(oracle/_ref/bear_ref_emission) when it exists.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "bear_ref_emission")
NB_SPECIES_IDS = 46

_lib = None


def build():
    """Compile the oracle library (and nothing of the product)."""
    root = os.path.dirname(HERE)
    subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=root, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        vp, sz, dp, ip = C.c_void_p, C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_emission_create.restype = vp
        L.oracle_emission_create.argtypes = [C.c_char_p]
        L.oracle_emission_destroy.argtypes = [vp]
        for f in ("oracle_nb_points", "oracle_nb_levels", "oracle_nb_parameters", "oracle_nb_observations"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp]
        for f in ("oracle_nb_bins", "oracle_nb_conv"):
            getattr(L, f).restype = sz
            getattr(L, f).argtypes = [vp, sz]
        L.oracle_grid.argtypes = [vp, dp, dp, ip]
        L.oracle_observation.restype = C.c_int
        L.oracle_observation.argtypes = [vp, sz, ip, ip, dp, ip, dp, dp, dp, dp, dp, dp, C.POINTER(C.c_int)]
        L.oracle_emission_eval.restype = C.c_int
        L.oracle_emission_eval.argtypes = [vp, dp, sz, dp, dp, dp, dp, C.POINTER(C.c_int)]
        L.oracle_atmosphere.argtypes = [vp, dp, dp, dp, dp, dp]
        L.oracle_coefficients.argtypes = [vp, dp, dp]
        L.oracle_planck.restype = C.c_double
        L.oracle_planck.argtypes = [C.c_double, C.c_double]
        i32 = C.POINTER(C.c_int32)
        L.oracle_register_table.argtypes = [C.c_char_p, C.c_double, C.c_int, sz, dp, dp, sz, C.POINTER(C.c_float)]
        L.oracle_transmission_record_stops.argtypes = [vp, C.c_int]
        L.oracle_transmission_stops.argtypes = [vp, i32]
        L.oracle_transmission_freeze_stops.argtypes = [vp, i32]
        L.oracle_transmission_paths.argtypes = [vp, sz, dp, i32]
        L.oracle_direct_spectrum.restype = C.c_int
        L.oracle_direct_spectrum.argtypes = [vp, C.c_double, C.c_double, C.c_double, dp, dp, C.c_char_p, dp, C.c_int,
                                             dp, dp]
        L.oracle_cloud_optical_depths.restype = C.c_int
        L.oracle_cloud_optical_depths.argtypes = [vp, dp]
        L.oracle_scale_height.argtypes = [vp, dp]
        L.oracle_cloud_extinction.restype = C.c_int
        L.oracle_cloud_extinction.argtypes = [vp, dp]
        L.oracle_tangent_sum.restype = C.c_double
        L.oracle_tangent_sum.argtypes = [dp, dp, sz, C.c_int32, i32]
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


# arrays of the in-memory tables, kept alive while registered
_REGISTERED = {}


def register_table(folder, temperatures, pressures, rows, species_mass=0.0, log_storage=False):
    """Hold one species' table in memory under ``folder`` (in place of
    opacity/<folder>/filelist.dat and its row files): rows [nb_rows,
    nb_points] float32 over the full wavenumber list with the (T, p) of each
    row.  The oracle references the arrays; they stay alive here until
    clear_tables()."""
    rows = np.ascontiguousarray(rows, dtype=np.float32).reshape(len(temperatures), -1)
    t = np.ascontiguousarray(temperatures, dtype=np.float64)
    p = np.ascontiguousarray(pressures, dtype=np.float64)
    _REGISTERED[folder] = (rows, t, p)
    lib().oracle_register_table(folder.encode(), float(species_mass), int(log_storage), t.size, _p(t), _p(p),
                                rows.shape[1], _p(rows, C.c_float))


def clear_tables():
    lib().oracle_clear_tables()
    _REGISTERED.clear()


class EmissionOracle:
    """The oracle's emission model, loaded from a retrieval folder."""

    def __init__(self, folder: str):
        self._L = lib()
        self._h = self._L.oracle_emission_create(folder.encode())
        if not self._h:
            raise RuntimeError(self._L.oracle_last_error().decode())
        self.nb_points = self._L.oracle_nb_points(self._h)
        self.nb_levels = self._L.oracle_nb_levels(self._h)
        self.nb_parameters = self._L.oracle_nb_parameters(self._h)
        self.nb_observations = self._L.oracle_nb_observations(self._h)
        self.nb_bins = [self._L.oracle_nb_bins(self._h, i) for i in range(self.nb_observations)]

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.oracle_emission_destroy(self._h)
            self._h = None

    def grid(self):
        n = self.nb_points
        nu, lam, idx = np.empty(n), np.empty(n), np.empty(n, np.int64)
        self._L.oracle_grid(self._h, _p(nu), _p(lam), _p(idx, C.c_int64))
        return nu, lam, idx

    def observation(self, i):
        n, nb, nc = self.nb_points, self.nb_bins[i], self._L.oracle_nb_conv(self._h, i)
        edges = np.empty((nb, 2), np.int64)
        rng = np.empty(2, np.int64)
        sigma, fil, fw = np.empty(n), np.empty(n), np.empty(n)
        conv = np.empty((max(nc, 1), 2), np.int64)
        norm = np.empty(1)
        data, err, w = np.empty(nb), np.empty(nb), np.empty(nb)
        asc = C.c_int(0)
        self._L.oracle_observation(self._h, i, _p(edges, C.c_int64), _p(rng, C.c_int64), _p(sigma),
                                   _p(conv, C.c_int64), _p(fil), _p(fw), _p(norm), _p(data), _p(err),
                                   _p(w), C.byref(asc))
        return dict(edge_indices=edges, index_range=rng, sigma=sigma, conv=conv[:nc], filter=fil,
                    filter_weight=fw, filter_norm=norm[0], data=data, error=err, weight=w,
                    ascending=bool(asc.value))

    def evaluate(self, params):
        params = np.ascontiguousarray(params, dtype=np.float64)
        n = self.nb_points
        flux, spec = np.empty(n), np.empty(n)
        bins = np.empty(sum(self.nb_bins))
        ll = C.c_double(0)
        neg = C.c_int(0)
        rc = self._L.oracle_emission_eval(self._h, _p(params), params.size, _p(flux), _p(spec), _p(bins),
                                          C.byref(ll), C.byref(neg))
        if rc != 0:
            raise RuntimeError(self._L.oracle_last_error().decode())
        out, k = [], 0
        for nb in self.nb_bins:
            out.append(bins[k:k + nb].copy())
            k += nb
        return dict(spectrum_flux=flux, spectrum=spec, bins=out, loglike=ll.value, neglect=bool(neg.value))

    def direct_spectrum(self, surface_gravity, a, b, pressure, temperature, species, mixing_ratios,
                        use_variable_gravity=False, cloud_optical_depth=None):
        """The direct models' calcSpectrum (a case written with a direct
        range): emission a = radius, b = distance; transmission a = planet
        radius, b = radius ratio; eclipse a = radius ratio (b unused).
        mixing_ratios [levels][species]; cloud_optical_depth [levels - 1][points]
        or None."""
        p = np.ascontiguousarray(pressure, dtype=np.float64)
        t = np.ascontiguousarray(temperature, dtype=np.float64)
        x = np.ascontiguousarray(mixing_ratios, dtype=np.float64)
        cloud = None
        if cloud_optical_depth is not None:
            cloud = np.ascontiguousarray(cloud_optical_depth, dtype=np.float64)
            assert cloud.size == (p.size - 1) * self.nb_points
        out = np.empty(self.nb_points)
        rc = self._L.oracle_direct_spectrum(self._h, float(surface_gravity), float(a), float(b), _p(p), _p(t),
                                            "\n".join(species).encode(), _p(x), int(use_variable_gravity),
                                            None if cloud is None else _p(cloud), _p(out))
        if rc != 0:
            raise RuntimeError(self._L.oracle_last_error().decode())
        return out

    # ---- transmission: stopping segments of the tangent paths
    def record_stops(self, enable=True):
        self._L.oracle_transmission_record_stops(self._h, int(enable))

    def stops(self):
        """Stopping segment of every (wavenumber, tangent path) of the last
        evaluation, [N, L-1] (recording on)."""
        out = np.empty((self.nb_points, self.nb_levels - 1), np.int32)
        if self._L.oracle_transmission_stops(self._h, _p(out, C.c_int32)) != 0:
            raise RuntimeError(self._L.oracle_last_error().decode())
        return out

    def freeze_stops(self, stops):
        """Hold every path's stopping segment fixed in the following
        evaluations (None: back to the tau > 10 rule)."""
        if stops is None:
            self._L.oracle_transmission_freeze_stops(self._h, None)
            return
        s = np.ascontiguousarray(stops, dtype=np.int32)
        assert s.shape == (self.nb_points, self.nb_levels - 1)
        self._L.oracle_transmission_freeze_stops(self._h, _p(s, C.c_int32))

    def paths(self, w):
        """(tau [L-1], stop [L-1]) of the tangent paths at wavenumber index w
        in the last evaluation."""
        tau = np.empty(self.nb_levels - 1)
        stop = np.empty(self.nb_levels - 1, np.int32)
        if self._L.oracle_transmission_paths(self._h, int(w), _p(tau), _p(stop, C.c_int32)) != 0:
            raise RuntimeError(self._L.oracle_last_error().decode())
        return tau, stop

    def atmosphere(self):
        L = self.nb_levels
        t, z, p, mmw = np.empty(L), np.empty(L), np.empty(L), np.empty(L)
        nd = np.empty((L, NB_SPECIES_IDS))
        self._L.oracle_atmosphere(self._h, _p(t), _p(z), _p(p), _p(nd), _p(mmw))
        return dict(temperature=t, altitude=z, pressure=p, number_densities=nd, mean_molecular_weight=mmw)

    def coefficients(self):
        k = np.empty((self.nb_levels, self.nb_points))
        s = np.empty_like(k)
        self._L.oracle_coefficients(self._h, _p(k), _p(s))
        return k, s

    def cloud_optical_depths(self):
        """The last evaluation's cloud layer optical depths [levels - 1, points],
        or None without a cloud."""
        out = np.empty((self.nb_levels - 1, self.nb_points))
        return out if self._L.oracle_cloud_optical_depths(self._h, _p(out)) else None

    def cloud_extinction(self):
        """The last transmission evaluation's cloud extinction [levels, points]
        (1/cm), or None without a cloud."""
        out = np.empty((self.nb_levels, self.nb_points))
        return out if self._L.oracle_cloud_extinction(self._h, _p(out)) else None

    def scale_height(self):
        """The last evaluation's scale height per level (cm)."""
        out = np.empty(self.nb_levels)
        self._L.oracle_scale_height(self._h, _p(out))
        return out


def tangent_sum(path, extinction, frozen=-1):
    """One tangent path's optical depth on given segment lengths [n] and
    level extinctions [n + 1]: (tau, last segment summed)."""
    path = np.ascontiguousarray(path, dtype=np.float64)
    ext = np.ascontiguousarray(extinction, dtype=np.float64)
    assert ext.size == path.size + 1
    last = C.c_int32(0)
    tau = lib().oracle_tangent_sum(_p(path), _p(ext), path.size, int(frozen), C.byref(last))
    return tau, last.value


def planck(t, nu):
    return lib().oracle_planck(float(t), float(nu))


def run_reference(case_folder: str, out_folder: str, n_timed: int = 0, threads: int | None = None):
    """Run the reference build on a case folder; returns its outputs as a dict
    of arrays, or None when the reference build is absent."""
    if not os.path.exists(REF_BIN):
        return None
    os.makedirs(out_folder, exist_ok=True)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    res = subprocess.run([REF_BIN, case_folder, out_folder, str(n_timed)], capture_output=True, text=True, env=env)
    if res.returncode != 0:
        raise RuntimeError(f"reference build failed ({res.returncode}): {res.stderr[-2000:]}")
    out = {f[:-4]: np.load(os.path.join(out_folder, f)) for f in os.listdir(out_folder) if f.endswith(".npy")}
    out["stdout"] = res.stdout
    return out
