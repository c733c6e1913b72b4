"""The C-ABI library: it loads, exports every entry point the header
declares, and reports errors as status codes (no GPU needed)."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "bear_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bear_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    from bear_amd import api
    lib = api.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_empty_error():
    from bear_amd import api
    assert "gfx950" in api.version()
    assert api.lib().bear_last_error() == b""


def test_cpu_path_is_refused():
    from bear_amd import api
    cfg = api.GlobalConfig(use_gpu=False)
    mc = api.ModelConfig(10, 100.0, 1e-5, "const", ["iso H2O", "bg H2He"], [("H2O", "H2O")])
    with pytest.raises(api.BearError) as e:
        api.Retrieval(cfg, mc, observation_files=["x.dat"])
    assert e.value.code == -1 and "GPU" in str(e.value)


def test_bad_inputs_are_status_codes(tmp_path):
    from bear_amd import api
    cfg = api.GlobalConfig(retrieval_folder_path=str(tmp_path))
    mc = api.ModelConfig(10, 100.0, 1e-5, "const", ["iso H2O", "bg H2He"], [("H2O", "H2O")])
    with pytest.raises(api.BearError) as e:
        api.Retrieval(cfg, mc, observation_files=["missing.dat"])
    assert e.value.code == -2                                     # FileNotFound
    with pytest.raises(api.BearError) as e:
        api.Retrieval(cfg, api.ModelConfig(10, 100.0, 1e-5, "guillot", [], []), observation_files=["missing.dat"])
    assert e.value.code in (-1, -2)
    with pytest.raises(api.BearError) as e:
        api.register_opacity_data("x", np.zeros(0), np.zeros(0), np.zeros((0, 4), np.float32))
    assert e.value.code == -1


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_missing_device_is_a_device_error(case_factory):
    from bear_amd import synth, api
    folder, _ = case_factory("tiny")
    with pytest.raises(api.BearError) as e:
        synth.open_retrieval(folder)
    assert e.value.code == -3


def test_batch_entry_points_refuse_null_arguments():
    from bear_amd import api
    L = api.lib()
    assert L.bear_retrieval_set_table_precision(None, 0) == -1
    assert L.bear_retrieval_set_batch_options(None, 1, 1) == -1
    assert L.bear_retrieval_set_device_staging(None, 1) == -1
    assert L.bear_retrieval_compute_likelihood_batch(None, None, 1, 4, None) == -1
    assert L.bear_retrieval_compute_model_batch(None, None, 1, 4, None, None, None) == -1
    assert L.bear_retrieval_batch_stats(None, None) == -1
    assert L.bear_retrieval_set_adjoint_memory(None, 0) == -1
    assert b"null" in L.bear_last_error()
