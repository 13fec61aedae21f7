"""Several independent systems in one evaluation (fmmhip_eval_set_batch, core.fmm_batch): one tree per
system in a forest, interaction lists within each system, every pass of all systems in the same launches.

Bar: each system's result is bit-identical to its own fmm!(system) through one evaluation (every pass is
deterministic and sums a target's terms in the order the single call does; the single calls are pinned
against the oracle by the parity tests), and the batch's branch, leaf and list counts are the sums of the
single calls'. Cases cover the error-controlled path with its influence estimate (relative tolerance),
fixed orders with Hessians, all three list methods, both precision forms, and systems of one and two
bodies."""
import ctypes as C

import numpy as np
import pytest

import fastmultipole_amd as F
from fastmultipole_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu


class _Rows:
    def __init__(self, kernel, rows):
        self.fmm_kernel = kernel
        self.rows = np.ascontiguousarray(rows, dtype=np.float64)

    def source_buffer(self):
        return self.rows


def _systems(kernel, sizes, seed):
    out = []
    for i, n in enumerate(sizes):
        if kernel == _lib.POINT_VORTEX:
            rows = W.vortex_ring(n, seed=seed + i) if i % 2 == 0 else W.uniform_vortices(n, seed=seed + i)
        else:
            rows = W.uniform_sources(n, seed=seed + i) if i % 2 == 0 else W.plummer_sources(n, seed=seed + i)
        out.append(_Rows(kernel, rows))
    return out


CASES = [
    ("vortex-relative-gradient", _lib.POINT_VORTEX, [3000, 1, 12000, 2, 700, 40000, 64, 9000],
     dict(leaf_size=32, expansion_order=20, multipole_acceptance=0.5, error_tolerance=F.PowerRelativeGradient(1e-8))),
    ("sources-fixed-hessian-selftuning", _lib.POINT_SOURCE, [5000, 20000, 2, 900, 1, 15000],
     dict(leaf_size=24, expansion_order=6, multipole_acceptance=0.45, scalar_potential=True, hessian=True,
          interaction_list_method=F.SelfTuning())),
    ("vortex-barba-mixed", _lib.POINT_VORTEX, [8000, 300, 25000, 1500],
     dict(leaf_size=40, expansion_order=8, multipole_acceptance=0.4, interaction_list_method=F.Barba(),
          precision="mixed")),
    ("sources-absolute-power", _lib.POINT_SOURCE, [4000, 30000, 50],
     dict(leaf_size=20, expansion_order=16, multipole_acceptance=0.5, scalar_potential=True,
          error_tolerance=F.PowerAbsolutePotential(1e-7))),
]


def _single(system, kw):
    ev = F.Evaluation()
    _, ev, *_ = F.fmm(system, evaluation=ev, update_target_systems=False, **kw)
    out = ev.targets()
    st = ev.last_stats
    tv = ev.tree(0)
    counts = (tv.n_branches, len(tv.leaf_index), st.n_m2l, st.n_direct, st.pmax, st.error_success)
    ev.close()
    return out, counts


@pytest.mark.parametrize("name,kernel,sizes,kw", CASES, ids=[c[0] for c in CASES])
def test_batch_matches_single_calls(name, kernel, sizes, kw):
    print(f"[gpu] batch {name}: sizes {sizes}", flush=True)
    systems = _systems(kernel, sizes, seed=700)
    singles = [_single(s, kw) for s in systems]
    outputs, ev = F.fmm_batch(systems, update_target_systems=False, **kw)
    st = ev.last_stats
    for i, (ref, _) in enumerate(singles):
        assert outputs[i].shape == ref.shape
        np.testing.assert_array_equal(outputs[i], ref, err_msg=f"system {i} (n={sizes[i]})")
    tv = ev.tree(0)
    assert tv.n_branches == sum(c[0] for _, c in singles)
    assert len(tv.leaf_index) == sum(c[1] for _, c in singles)
    assert st.n_m2l == sum(c[2] for _, c in singles)
    assert st.n_direct == sum(c[3] for _, c in singles)
    assert st.pmax == max(c[4] for _, c in singles)
    assert st.error_success == min(c[5] for _, c in singles)
    pm, ok = ev.batch_stats()
    np.testing.assert_array_equal(pm, [c[4] for _, c in singles])
    np.testing.assert_array_equal(ok, [bool(c[5]) for _, c in singles])
    ev.close()


def test_batch_reuses_evaluation_across_sizes():
    """One evaluation handle runs a batch, then a smaller batch, then a single system: each result equals
    its single call (buffers sized for the larger batch must not leak into the next)."""
    kw = dict(leaf_size=30, expansion_order=10, multipole_acceptance=0.5, error_tolerance=F.PowerRelativeGradient(1e-8))
    ev = F.Evaluation()
    for sizes in ([20000, 5000, 800, 12000], [3000, 2]):
        systems = _systems(_lib.POINT_VORTEX, sizes, seed=900 + len(sizes))
        outputs, _ = F.fmm_batch(systems, evaluation=ev, update_target_systems=False, **kw)
        for s, o in zip(systems, outputs):
            np.testing.assert_array_equal(o, _single(s, kw)[0])
    one = _systems(_lib.POINT_VORTEX, [6000], seed=950)[0]
    _, ev2, *_ = F.fmm(one, evaluation=ev, update_target_systems=False, **kw)
    np.testing.assert_array_equal(ev2.targets(), _single(one, kw)[0])
    ev.close()


def test_batch_rejects_bad_input():
    ctx = F.default_context()
    lib = ctx.lib
    ev = F.Evaluation()
    rows = W.uniform_sources(100, seed=1)
    counts = np.array([60, 0, 40], dtype=np.int64)
    st = lib.fmmhip_eval_set_batch(ev._h, _lib.POINT_SOURCE, 3, counts.ctypes.data, rows.ctypes.data, 0)
    assert st == _lib.FMMHIP_EINVAL
    st = lib.fmmhip_eval_set_batch(ev._h, _lib.POINT_SOURCE, 0, counts.ctypes.data, rows.ctypes.data, 0)
    assert st == _lib.FMMHIP_EINVAL
    counts = np.array([60, 40], dtype=np.int64)
    assert lib.fmmhip_eval_set_batch(ev._h, _lib.POINT_SOURCE, 2, counts.ctypes.data, rows.ctypes.data, 0) == 0
    opt = F.make_options(4, 20)
    assert lib.fmmhip_fgs_setup(ev._h, C.byref(opt), _lib.INFLUENCE_POTENTIAL) == _lib.FMMHIP_EINVAL
    ev.close()
