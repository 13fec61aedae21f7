"""CPU checks of the C ABI boundary: the library loads and exports every symbol the header declares."""
import ctypes
import os
import re

import pytest

from fastmultipole_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "fmmhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fmmhip_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    fns = header_functions()
    assert set(fns) == set(_lib.EXPORTED), set(fns) ^ set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_options_defaults_mirror_fmm_kwargs():
    o = _lib.default_options()
    # src/fmm.jl:862-923 defaults: expansion_order 5, leaf size 20, MAC 0.4, SelfTuningTreeStop, shrink on
    assert (o.expansion_order, o.leaf_size_source, o.leaf_size_target) == (5, 20, 20)
    assert o.multipole_acceptance == pytest.approx(0.4)
    assert o.list_method == _lib.SELFTUNING_TREESTOP and o.shrink_recenter == 1
    assert (o.scalar_potential, o.gradient, o.hessian) == (0, 1, 0)
    assert (o.farfield, o.nearfield, o.self_induced) == (1, 1, 1)


def test_no_gpu_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.fmmhip_ctx_create(0, ctypes.byref(h)) == _lib.FMMHIP_EDEVICE
