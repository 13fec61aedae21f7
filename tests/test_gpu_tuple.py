"""Several systems in one tree (fmm!((sys1, sys2, ...)), fmmhip_eval_set_tuple) against the oracle's tuple run.

Bars: the tree is bit-exact (every branch field, body order); potential, gradient and Hessian within 1e-11 of
the oracle relative to the largest target value of each slot, and the expansions within 1e-11 of the largest
coefficient (same expansion order, leaf size, MAC); with an error tolerance, pmax and error_success equal the
oracle's, every M2L order decision that is not within 1e-9 (relative) of the tolerance equals the oracle's,
and the targets match to 1e-11 away from flipped decisions. A tuple of one system equals the plain
evaluation of that system bit for bit. Cases cover point and panel kernels, a leaf size per system, and
SelfTuningTreeStop and Barba lists.
"""
import ctypes
import sys

import numpy as np
import pytest

import fastmultipole_amd as F
from fastmultipole_amd import _lib, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_ORC = {_lib.POINT_SOURCE: O.KIND_SOURCE, _lib.POINT_VORTEX: O.KIND_VORTEX, _lib.POINT_DIPOLE: O.KIND_DIPOLE,
        _lib.VORTEX_GAUSSIAN: O.KIND_VORTEX_GAUSS, _lib.PANEL_SOURCE_TRI: O.KIND_PANEL_SOURCE_TRI,
        _lib.PANEL_DIPOLE_TRI: O.KIND_PANEL_DIPOLE_TRI, _lib.PANEL_SOURCE_QUAD: O.KIND_PANEL_SOURCE_QUAD,
        _lib.PANEL_DIPOLE_QUAD: O.KIND_PANEL_DIPOLE_QUAD, _lib.POINT_SOURCE_VORTEX: O.KIND_SOURCE_VORTEX,
        _lib.PANEL_SOURCE_DIPOLE_TRI: O.KIND_PANEL_SOURCE_DIPOLE_TRI,
        _lib.PANEL_SOURCE_DIPOLE_QUAD: O.KIND_PANEL_SOURCE_DIPOLE_QUAD}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Rows:
    def __init__(self, kernel, rows):
        self.fmm_kernel = kernel
        self.rows = np.ascontiguousarray(rows, dtype=np.float64)

    def source_buffer(self):
        return self.rows


def _dipoles(n, seed, shift):
    rng = np.random.default_rng(seed)
    d = np.zeros((n, 7))
    d[:, :3] = rng.random((n, 3)) + shift
    d[:, 4:7] = rng.normal(size=(n, 3)) * (0.05 / n)
    return d


def _gauss(n, seed, shift):
    d = W.uniform_vortices(n, seed=seed).copy()
    d[:, :3] += shift
    d[:, 3] = 0.004
    return d


def _src(n, seed, shift):
    d = W.uniform_sources(n, seed=seed).copy()
    d[:, :3] += shift
    d[:, 4] *= 10.0
    return d


def _panels(quads):
    def gen(n, seed, shift):   # panels tiling a sphere of radius 1.6 about the origin, about n of them
        m = max(2, int(round((n / (6 if quads else 12)) ** 0.5)))
        rows, _ = W.sphere_panels(m, quads=quads, seed=seed, radius=1.6)
        rows = rows.copy()
        rows[:, 4] *= 30.0
        return rows
    return gen


def _ring(n, seed, shift):
    d = W.vortex_ring(n, seed=seed).copy()
    d[:, :3] += shift
    return d


CASES = [
    ("source+vortex-p10-hessian", [(_lib.POINT_SOURCE, _src, 6000), (_lib.POINT_VORTEX, _ring, 9000)],
     dict(expansion_order=10, leaf_size=32, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("vortex+dipole+gauss+source-p8", [(_lib.POINT_VORTEX, _ring, 5000), (_lib.POINT_DIPOLE, _dipoles, 3000),
                                       (_lib.VORTEX_GAUSSIAN, _gauss, 2000), (_lib.POINT_SOURCE, _src, 4000)],
     dict(expansion_order=8, leaf_size=24, multipole_acceptance=0.45, hessian=True)),
    ("source+dipole-p12-potential", [(_lib.POINT_SOURCE, _src, 7000), (_lib.POINT_DIPOLE, _dipoles, 2500)],
     dict(expansion_order=12, leaf_size=40, multipole_acceptance=0.5, scalar_potential=True)),
    # the reference's motivating tuple: panels and vortex particles (17- and 14-row panel buffers, points padded)
    ("quads+vortex-p10", [(_lib.PANEL_SOURCE_QUAD, _panels(True), 2400), (_lib.POINT_VORTEX, _ring, 6000)],
     dict(expansion_order=10, leaf_size=24, multipole_acceptance=0.5, hessian=True)),
    ("dipole-tris+gauss+source-p8", [(_lib.PANEL_DIPOLE_TRI, _panels(False), 2400), (_lib.VORTEX_GAUSSIAN, _gauss, 3000),
                                     (_lib.POINT_SOURCE, _src, 2000)],
     dict(expansion_order=8, leaf_size=20, multipole_acceptance=0.5, scalar_potential=True)),
    ("dipole-quads+source-tris-p9", [(_lib.PANEL_DIPOLE_QUAD, _panels(True), 1500),
                                     (_lib.PANEL_SOURCE_TRI, _panels(False), 1200)],
     dict(expansion_order=9, leaf_size=16, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    # a leaf size per system (the reference's SVector leaf_size): the leaf test counts each system
    ("source+vortex-leaf-sizes", [(_lib.POINT_SOURCE, _src, 6000), (_lib.POINT_VORTEX, _ring, 9000)],
     dict(expansion_order=9, leaf_size=[40, 12], multipole_acceptance=0.5, hessian=True)),
    ("quads+gauss-barba-leaf-sizes", [(_lib.PANEL_SOURCE_QUAD, _panels(True), 2400), (_lib.VORTEX_GAUSSIAN, _gauss, 4000)],
     dict(expansion_order=8, leaf_size=[10, 30], multipole_acceptance=0.5, interaction_list_method=F.Barba())),
    ("source-vortex+quads-p8", [(_lib.POINT_SOURCE_VORTEX, lambda n, seed, shift: W.source_vortex_points(n, seed=seed), 6000),
                                (_lib.PANEL_SOURCE_QUAD, _panels(True), 1500)],
     dict(expansion_order=8, leaf_size=[24, 16], multipole_acceptance=0.5, hessian=True)),
    ("srcdip-quads+vortex-p8", [(_lib.PANEL_SOURCE_DIPOLE_QUAD,
                                 lambda n, seed, shift: W.sphere_source_dipole_panels(16, seed=seed, radius=1.6)[0], 1536),
                                (_lib.POINT_VORTEX, _ring, 5000)],
     dict(expansion_order=8, leaf_size=20, multipole_acceptance=0.5, hessian=True)),
    ("vortex+dipole+source-barba", [(_lib.POINT_VORTEX, _ring, 5000), (_lib.POINT_DIPOLE, _dipoles, 3000),
                                    (_lib.POINT_SOURCE, _src, 3000)],
     dict(expansion_order=8, leaf_size=24, multipole_acceptance=0.5, interaction_list_method=F.Barba())),
]


def _make(spec, seed):
    return [_Rows(k, gen(n, seed + i, 0.15 * i)) for i, (k, gen, n) in enumerate(spec)]


def _oracle(systems, kw):
    ls = kw["leaf_size"]
    per_system = np.ndim(ls) > 0
    okw = dict(p=kw["expansion_order"], leaf_size_source=int(min(ls)) if per_system else ls,
               leaf_sizes=list(ls) if per_system else None, multipole_acceptance=kw["multipole_acceptance"],
               scalar_potential=kw.get("scalar_potential", False), gradient=True, hessian=kw.get("hessian", False))
    if kw.get("interaction_list_method") is not None:
        okw["method"] = kw["interaction_list_method"].code
    em = kw.get("error_tolerance")
    if em is not None:
        okw.update(error_kind=em.kind, relative=em.relative, eps=em.eps, BE=em.BE)
    return O.Run.tuple([_ORC[s.fmm_kernel] for s in systems], [s.rows for s in systems], **okw).fmm()


def _compare_tree(ev, run, which):
    g = ev.tree(which)
    ints, reals = run.branches(which)
    assert g.n_branches == len(ints)
    gi, oi = g.ints.copy(), ints.copy()
    leaves = oi[:, 2] == 0
    gi[leaves, 3] = oi[leaves, 3] = -1
    np.testing.assert_array_equal(gi, oi)
    assert np.array_equal(g.reals[:, :15], reals[:, :15]), "branch geometry differs bitwise"
    np.testing.assert_array_equal(g.sort_index, run.sort_index(which))


def _slots(kw):
    out = [(4, 7)]
    if kw.get("scalar_potential"):
        out.append((3, 4))
    if kw.get("hessian"):
        out.append((7, 16))
    return out


@pytest.mark.parametrize("name,spec,kw", CASES, ids=[c[0] for c in CASES])
def test_tuple_matches_oracle(name, spec, kw):
    _say(f"[gpu] tuple {name}")
    systems = _make(spec, seed=1200)
    _, ev, *_rest, ok = F.fmm(tuple(systems), update_target_systems=False, **kw)
    run = _oracle(systems, kw)
    got = np.concatenate(ev.tuple_outputs)
    ref = run.targets()
    assert ev.last_stats.pmax == run.pmax and ok == run.error_success
    _compare_tree(ev, run, 0)
    _compare_tree(ev, run, 1)
    assert ev.last_stats.n_m2l == len(run.list(0)) > 0
    for lo, hi in _slots(kw):
        scale = np.abs(ref[:, lo:hi]).max()
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / scale
        assert err < 1e-11, (name, lo, hi, err)
    p = kw["expansion_order"]
    for which in (1, 0):
        e_g, e_o = ev.tree(which).expansions(p), run.expansions(which)
        assert np.abs(e_g - e_o).max() <= 1e-11 * np.abs(e_o).max(), which
    ev.close()


def test_tuple_error_controlled_matches_oracle():
    _say("[gpu] tuple error-controlled vortex+source")
    kw = dict(expansion_order=20, leaf_size=32, multipole_acceptance=0.5, error_tolerance=F.PowerRelativeGradient(1e-7))
    systems = _make([(_lib.POINT_VORTEX, _ring, 8000), (_lib.POINT_SOURCE, _src, 5000)], seed=1300)
    _, ev, *_rest, ok = F.fmm(tuple(systems), update_target_systems=False, **kw)
    run = _oracle(systems, kw)
    assert ev.last_stats.pmax == run.pmax and ok == run.error_success
    np.testing.assert_allclose(ev.tree(0).reals[:, 14], run.branches(0)[1][:, 14], rtol=1e-11, atol=0)
    Pg, Po, margin = ev.m2l_orders(), run.m2l_orders(), run.m2l_margins()
    assert len(Pg) == len(Po)
    flipped = Pg != Po
    assert np.all(np.abs(margin[flipped]) < 1e-9), "an order decision away from the tolerance differs"
    got, ref = np.concatenate(ev.tuple_outputs), run.targets()
    if not flipped.any():
        err = np.abs(got[:, 4:7] - ref[:, 4:7]).max() / np.abs(ref[:, 4:7]).max()
        assert err < 1e-11, err
    ev.close()


@pytest.mark.parametrize("kernel,gen", [(_lib.POINT_VORTEX, _ring), (_lib.POINT_SOURCE, _src)])
def test_tuple_of_one_system_is_the_plain_evaluation(kernel, gen):
    _say(f"[gpu] tuple of one system, kernel {kernel}")
    opt = F.make_options(9, 30, multipole_acceptance=0.5, scalar_potential=kernel == _lib.POINT_SOURCE, hessian=True)
    rows = gen(12000, 1400, 0.0)
    ev1 = F.Evaluation()
    ev1.set_systems(kernel, None, rows)
    ev1.fmm(opt)
    ev2 = F.Evaluation()
    ev2.set_tuple([kernel], [rows])
    ev2.fmm(opt)
    np.testing.assert_array_equal(ev2.targets(), ev1.targets())
    ev1.close()
    ev2.close()


def test_tuple_rejects_bad_input():
    ctx = F.default_context()
    lib = ctx.lib
    ev = F.Evaluation()
    rows = np.zeros((100, 7))
    rows[:, :3] = np.random.default_rng(3).random((100, 3))
    counts = np.array([60, 40], dtype=np.int64)
    bad = np.array([_lib.POINT_SOURCE, 11], dtype=np.int32)   # no kernel 11
    assert lib.fmmhip_eval_set_tuple(ev._h, 2, bad.ctypes.data, counts.ctypes.data, rows.ctypes.data, 0) == _lib.FMMHIP_EINVAL
    good = np.array([_lib.POINT_SOURCE, _lib.POINT_VORTEX], dtype=np.int32)
    neg = np.array([60, -1], dtype=np.int64)
    assert lib.fmmhip_eval_set_tuple(ev._h, 2, good.ctypes.data, neg.ctypes.data, rows.ctypes.data, 0) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_eval_set_tuple(ev._h, 0, good.ctypes.data, counts.ctypes.data, rows.ctypes.data, 0) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_eval_set_tuple(ev._h, 2, good.ctypes.data, counts.ctypes.data, rows.ctypes.data, 0) == 0
    opt = F.make_options(4, 20, interaction_list_method=F.SelfTuning())
    assert lib.fmmhip_eval_build_trees(ev._h, ctypes.byref(opt)) == 0
    assert lib.fmmhip_eval_build_lists(ev._h, ctypes.byref(opt)) == _lib.FMMHIP_EINVAL
    bad_ls = np.array([10, 0], dtype=np.int64)
    assert lib.fmmhip_eval_set_tuple_leaf_sizes(ev._h, 2, bad_ls.ctypes.data) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_eval_set_tuple_leaf_sizes(ev._h, 3, counts.ctypes.data) == _lib.FMMHIP_EINVAL
    assert lib.fmmhip_eval_set_tuple_leaf_sizes(ev._h, 2, counts.ctypes.data) == 0
    assert lib.fmmhip_fgs_setup(ev._h, ctypes.byref(opt), _lib.INFLUENCE_POTENTIAL) == _lib.FMMHIP_EINVAL
    ev.close()
