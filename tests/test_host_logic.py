"""CPU checks of host-side logic: the tune=true leaf-size update and the request server's argument
and no-GPU error paths (no compute call)."""
import ctypes
from types import SimpleNamespace

import pytest

from fastmultipole_amd import _lib
from fastmultipole_amd.core import _tuned_leaf_size


def _stats(n_m2l, ms_m2l, n_pairs, ms_p2p):
    return SimpleNamespace(n_m2l=n_m2l, ms_m2l=ms_m2l, n_p2p_pairs=n_pairs, ms_p2p=ms_p2p)


def test_tuned_leaf_size_balances_m2l_and_pair_costs():
    # src/fmm.jl:1073-1083: t_direct * leaf^2 = t_m2l. One M2L costs as much as 100 pairs -> leaf 10.
    assert _tuned_leaf_size(_stats(1000, 2.0, 50_000, 1.0), 20) == 10
    # rounding is upward: a ratio of 101 pairs per M2L gives 11
    assert _tuned_leaf_size(_stats(1000, 1.01, 100_000, 1.0), 20) == 11
    # the result does not depend on the old leaf size
    assert _tuned_leaf_size(_stats(10, 4.0, 10, 1.0), 3) == _tuned_leaf_size(_stats(10, 4.0, 10, 1.0), 300) == 2


def test_tuned_leaf_size_empty_lists():
    # src/fmm.jl:1084-1087: no M2L -> halve the leaf size, never below 1
    assert _tuned_leaf_size(_stats(0, 0.0, 100, 1.0), 20) == 10
    assert _tuned_leaf_size(_stats(0, 0.0, 100, 1.0), 1) == 1
    # no body pair evaluated (nearfield off): the ratio is undefined, the leaf size is kept
    assert _tuned_leaf_size(_stats(50, 1.0, 0, 0.0), 20) == 20


def test_server_arguments_and_no_gpu_path():
    import torch
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.fmmhip_server_create(0, 0, 1e9, ctypes.byref(h)) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_server_create(0, 17, 1e9, ctypes.byref(h)) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_server_create(0, 2, 0.0, ctypes.byref(h)) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_server_submit(None, None, None) == _lib.FMMHIP_EINVAL
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    # workers cannot create their contexts without a GPU: creation fails loudly and cleans up
    assert lib.fmmhip_server_create(0, 2, 1e9, ctypes.byref(h)) == _lib.FMMHIP_EDEVICE
