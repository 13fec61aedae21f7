"""GPU parity tests: libfmmhip (through the C ABI) against the CPU oracle.

Bars:
* tree and interaction lists: bit-exact (every branch field, body order, list order);
* FP64 evaluation: potential/gradient/hessian within 1e-11 of the oracle's FMM
  relative to the largest target value (same expansion order, leaf size, MAC);
* error-controlled expansion order: the order chosen per M2L interaction equals
  the oracle's for every interaction whose error estimate is not within 1e-9
  (relative) of the tolerance (only such a decision can flip under last-bit
  differences of the rotated coefficients); pmax and error_success equal the
  oracle's; targets under no flipped interaction match the oracle to 1e-11
  relative, the others within twice the oracle's own error against direct!;
* mixed precision: the stated error bound against the FP64 O(N^2) direct sum.
"""
import math
import sys

import numpy as np
import pytest

import fastmultipole_amd as F
from fastmultipole_amd import _lib, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _opts(**kw):
    return F.make_options(**kw)


def _gpu_eval(kernel, sbuf, txyz=None, **kw):
    ev = F.Evaluation()
    ev.set_systems(kernel, txyz, sbuf)
    opt = _opts(**kw)
    return ev, opt


def _oracle_run(kind, sbuf, txyz=None, **kw):
    pos = sbuf[:, :3] if txyz is None else txyz
    m = {"interaction_list_method": "method"}
    okw = {}
    for k, v in kw.items():
        if k == "interaction_list_method":
            okw["method"] = v.code
        elif k == "error_tolerance":
            if v is not None:
                okw.update(error_kind=v.kind, relative=v.relative, eps=v.eps, BE=v.BE)
        elif k == "precision":
            pass
        else:
            okw[k] = v
    return O.Run(kind, pos, sbuf, **okw)


def _compare_trees(ev, run, which):
    g = ev.tree(which)
    ints, reals = run.branches(which)
    assert g.n_branches == len(ints), (g.n_branches, len(ints))
    gi = g.ints.copy()
    oi = ints.copy()
    leaves = oi[:, 2] == 0
    gi[leaves, 3] = oi[leaves, 3] = -1          # child range start is meaningless for leaves
    np.testing.assert_array_equal(gi, oi)
    assert np.array_equal(g.reals[:, :15], reals[:, :15]), "branch geometry differs bitwise"
    np.testing.assert_array_equal(g.levels_index, run.levels(which))
    np.testing.assert_array_equal(g.leaf_index, run.leaf_index(which))
    np.testing.assert_array_equal(g.sort_index, run.sort_index(which))


TREE_CASES = [
    ("uniform-20k-ls20", lambda: W.uniform_sources(20000, seed=11), O.KIND_SOURCE, 20, O.METHOD_SELFTUNING_TREESTOP, True),
    ("uniform-5k-ls1-selftuning", lambda: W.uniform_sources(5000, seed=12), O.KIND_SOURCE, 1, O.METHOD_SELFTUNING, False),
    ("ring-40k-ls64", lambda: W.vortex_ring(40000, seed=13), O.KIND_VORTEX, 64, O.METHOD_SELFTUNING_TREESTOP, True),
    ("plummer-30k-barba", lambda: W.plummer_sources(30000, seed=14), O.KIND_SOURCE, 30, O.METHOD_BARBA, True),
]


@pytest.mark.parametrize("name,gen,kind,ls,method,shrink", TREE_CASES, ids=[c[0] for c in TREE_CASES])
def test_tree_and_lists_bit_exact(name, gen, kind, ls, method, shrink):
    _say(f"[gpu] tree+lists {name}")
    sbuf = gen()
    if kind == O.KIND_SOURCE and name.startswith("plummer"):
        sbuf[:, 3] = 0.001 * np.abs(np.sin(np.arange(len(sbuf))))   # nonzero body radii exercise the source boxes
    lm = {O.METHOD_BARBA: F.Barba(), O.METHOD_SELFTUNING: F.SelfTuning(), O.METHOD_SELFTUNING_TREESTOP: F.SelfTuningTreeStop()}[method]
    ev, opt = _gpu_eval(kind, sbuf, expansion_order=4, leaf_size_source=ls, multipole_acceptance=0.5,
                        interaction_list_method=lm, shrink_recenter=shrink)
    ev.build_trees(opt)
    ev.build_lists(opt)
    run = O.Run(kind, sbuf[:, :3], sbuf, p=4, leaf_size_source=ls, method=method, shrink_recenter=shrink,
                multipole_acceptance=0.5).trees().lists()
    _compare_trees(ev, run, 0)
    _compare_trees(ev, run, 1)
    for which in (0, 1):
        np.testing.assert_array_equal(ev.list(which), run.list(which))


def _with_duplicates(b, k, seed):
    """b with k more bodies at the positions of k of its bodies (other radii and strengths): the near
    field then holds coincident pairs of distinct bodies, which contribute nothing."""
    rng = np.random.default_rng(seed)
    extra = b[rng.choice(len(b), size=k, replace=False)].copy()
    extra[:, 3:] = b[rng.choice(len(b), size=k, replace=False), 3:]
    return np.ascontiguousarray(np.vstack([b, extra]))


FMM_CASES = [
    ("laplace-dup-3k-p6", lambda: _with_duplicates(W.uniform_sources(3000, seed=26), 300, 1), O.KIND_SOURCE,
     dict(expansion_order=6, leaf_size_source=30, multipole_acceptance=0.5, scalar_potential=True, hessian=False)),
    ("vortex-dup-3k-p8", lambda: _with_duplicates(W.uniform_vortices(3000, seed=27), 300, 2), O.KIND_VORTEX,
     dict(expansion_order=8, leaf_size_source=30, multipole_acceptance=0.5)),
    ("dipole-dup-3k-p6", lambda: _with_duplicates(W.uniform_dipoles(3000, seed=28), 300, 3), O.KIND_DIPOLE,
     dict(expansion_order=6, leaf_size_source=30, multipole_acceptance=0.5, scalar_potential=True)),
    ("laplace-4k-p8", lambda: W.uniform_sources(4000, seed=21), O.KIND_SOURCE, dict(expansion_order=8, leaf_size_source=50,
                                                                                    multipole_acceptance=0.5,
                                                                                    scalar_potential=True, hessian=False)),
    ("laplace-3k-p6-noshrink", lambda: W.uniform_sources(3000, seed=22), O.KIND_SOURCE,
     dict(expansion_order=6, leaf_size_source=20, multipole_acceptance=0.4, shrink_recenter=False, scalar_potential=True,
          hessian=True)),
    ("vortex-4k-p8", lambda: W.uniform_vortices(4000, seed=23), O.KIND_VORTEX,
     dict(expansion_order=8, leaf_size_source=30, multipole_acceptance=0.5, hessian=True)),
    ("ring-6k-p10", lambda: W.vortex_ring(6000, seed=24), O.KIND_VORTEX,
     dict(expansion_order=10, leaf_size_source=64, multipole_acceptance=0.4, hessian=True)),
    # SelfTuning pairs branches that are not leaves in the direct list, so direct targets nest (an ancestor's
    # bodies include its descendants'); Barba's lists (leaf pairs only) for comparison
    ("laplace-5k-p6-selftuning", lambda: W.uniform_sources(5000, seed=700), O.KIND_SOURCE,
     dict(expansion_order=6, leaf_size_source=24, multipole_acceptance=0.45, scalar_potential=True, hessian=True,
          interaction_list_method=F.SelfTuning())),
    ("vortex-4k-p8-selftuning-ls8", lambda: W.uniform_vortices(4000, seed=701), O.KIND_VORTEX,
     dict(expansion_order=8, leaf_size_source=8, multipole_acceptance=0.5, hessian=True,
          interaction_list_method=F.SelfTuning())),
    ("vortex-4k-p8-barba", lambda: W.uniform_vortices(4000, seed=702), O.KIND_VORTEX,
     dict(expansion_order=8, leaf_size_source=30, multipole_acceptance=0.5, interaction_list_method=F.Barba())),
    # orders <= 4 take the lane-per-interaction M2L kernel
    ("vortex-5k-p3-leaf5", lambda: W.uniform_vortices(5000, seed=25), O.KIND_VORTEX,
     dict(expansion_order=3, leaf_size_source=5, multipole_acceptance=0.6, scalar_potential=True, hessian=True)),
    ("laplace-4k-p4", lambda: W.uniform_sources(4000, seed=26), O.KIND_SOURCE,
     dict(expansion_order=4, leaf_size_source=10, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("laplace-2k-p1", lambda: W.uniform_sources(2000, seed=27), O.KIND_SOURCE,
     dict(expansion_order=1, leaf_size_source=8, multipole_acceptance=0.5, scalar_potential=True)),
    # order 20: two back-rotation chunks in the wavefront engine
    ("ring-3k-p20", lambda: W.vortex_ring(3000, seed=28), O.KIND_VORTEX,
     dict(expansion_order=20, leaf_size_source=40, multipole_acceptance=0.5, hessian=True)),
    # point dipoles (P2M of src/bodytomultipole.jl:431-463) and Gaussian-core vortices
    ("dipole-4k-p8", lambda: W.uniform_dipoles(4000, seed=29), O.KIND_DIPOLE,
     dict(expansion_order=8, leaf_size_source=30, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("dipole-3k-p3", lambda: W.uniform_dipoles(3000, seed=30), O.KIND_DIPOLE,
     dict(expansion_order=3, leaf_size_source=10, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("gauss-ring-6k-p10", lambda: W.gaussian_vortex_ring(6000, seed=31), O.KIND_VORTEX_GAUSS,
     dict(expansion_order=10, leaf_size_source=64, multipole_acceptance=0.4, hessian=True)),
    # the precision-sweep distribution: Plummer sphere (scale 1, truncated at r = 20), N(0,1) strengths, at
    # the sweep's orders 4 / 8 / 16 (deep, unbalanced trees: the core sits many levels below the halo)
    ("plummer-30k-p4", lambda: W.plummer_sources(30000, seed=81), O.KIND_SOURCE,
     dict(expansion_order=4, leaf_size_source=40, multipole_acceptance=0.5, scalar_potential=True)),
    ("plummer-30k-p8", lambda: W.plummer_sources(30000, seed=82), O.KIND_SOURCE,
     dict(expansion_order=8, leaf_size_source=40, multipole_acceptance=0.5, scalar_potential=True)),
    ("plummer-30k-p16", lambda: W.plummer_sources(30000, seed=83), O.KIND_SOURCE,
     dict(expansion_order=16, leaf_size_source=40, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    # Point{SourceVortex}: a source and a vortex on every body
    ("source-vortex-20k-p10", lambda: W.source_vortex_points(20000, seed=84), O.KIND_SOURCE_VORTEX,
     dict(expansion_order=10, leaf_size_source=32, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    # constant-strength panels tiling a sphere (quadrilaterals, and triangles from splitting them)
    ("panel-src-tri-3k-p8", lambda: W.sphere_panels(16, quads=False)[0], O.KIND_PANEL_SOURCE_TRI,
     dict(expansion_order=8, leaf_size_source=20, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("panel-dip-tri-3k-p8", lambda: W.sphere_panels(16, quads=False)[0], O.KIND_PANEL_DIPOLE_TRI,
     dict(expansion_order=8, leaf_size_source=20, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("panel-src-quad-1.5k-p6", lambda: W.sphere_panels(16)[0], O.KIND_PANEL_SOURCE_QUAD,
     dict(expansion_order=6, leaf_size_source=16, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("panel-dip-quad-1.5k-p10", lambda: W.sphere_panels(16)[0], O.KIND_PANEL_DIPOLE_QUAD,
     dict(expansion_order=10, leaf_size_source=16, multipole_acceptance=0.4, scalar_potential=True, hessian=True)),
    # Panel{SourceDipole}: a source and a dipole strength on every panel
    ("panel-srcdip-tri-3k-p8", lambda: W.sphere_source_dipole_panels(16, quads=False)[0], O.KIND_PANEL_SOURCE_DIPOLE_TRI,
     dict(expansion_order=8, leaf_size_source=20, multipole_acceptance=0.5, scalar_potential=True, hessian=True)),
    ("panel-srcdip-quad-1.5k-p9", lambda: W.sphere_source_dipole_panels(16)[0], O.KIND_PANEL_SOURCE_DIPOLE_QUAD,
     dict(expansion_order=9, leaf_size_source=16, multipole_acceptance=0.45, scalar_potential=True, hessian=True)),
]


@pytest.mark.parametrize("name,gen,kind,kw", FMM_CASES, ids=[c[0] for c in FMM_CASES])
def test_fmm_fp64_matches_oracle(name, gen, kind, kw):
    _say(f"[gpu] fmm fp64 {name}")
    sbuf = gen()
    ev, opt = _gpu_eval(kind, sbuf, **kw)
    st = ev.fmm(opt)
    got = ev.targets()
    run = _oracle_run(kind, sbuf, p=kw["expansion_order"], **{k: v for k, v in kw.items() if k != "expansion_order"}).fmm()
    ref = run.targets()
    assert st.pmax == run.pmax and bool(st.error_success) == run.error_success
    if name.startswith("plummer"):
        _compare_trees(ev, run, 0)
        _compare_trees(ev, run, 1)
    for lo, hi in ((3, 4), (4, 7), (7, 16)):
        scale = np.abs(ref[:, lo:hi]).max()
        if scale == 0:
            continue
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / scale
        assert err < 1e-11, (name, lo, hi, err)
    # expansions themselves (multipoles of the source tree, locals of the target tree)
    p = kw["expansion_order"]
    for which in (1, 0):
        e_g = ev.tree(which).expansions(p)
        e_o = run.expansions(which)
        sc = np.abs(e_o).max()
        assert np.abs(e_g - e_o).max() <= 1e-11 * sc


ERR_CASES = [
    # (name, generator, kernel, error method, leaf size[, multipole acceptance])
    ("laplace-rotated-abs", lambda: W.uniform_sources(5000, seed=31, strength_scale=1 / 5000), O.KIND_SOURCE,
     F.RotatedCoefficientsAbsoluteGradient(1e-5, False), 10),
    ("vortex-power-abs", lambda: W.uniform_vortices(5000, seed=32), O.KIND_VORTEX, F.PowerAbsoluteGradient(1e-5, False), 10),
    ("laplace-power-potential-abs", lambda: W.uniform_sources(4000, seed=33, strength_scale=1 / 4000), O.KIND_SOURCE,
     F.PowerAbsolutePotential(1e-6), 16),
    ("ring-power-rel", lambda: W.vortex_ring(8000, seed=34), O.KIND_VORTEX, F.PowerRelativeGradient(1e-6), 64),
    ("laplace-rotated-rel", lambda: W.uniform_sources(5000, seed=35), O.KIND_SOURCE,
     F.RotatedCoefficientsRelativeGradient(1e-6), 20),
    # PowerRelativePotential: estimate_influence! keeps the potential per target (src/fmm.jl:1185-1187)
    ("laplace-power-potential-rel", lambda: W.uniform_sources(5000, seed=40), O.KIND_SOURCE,
     F.PowerRelativePotential(1e-7), 20),
    ("source-vortex-power-rel", lambda: W.source_vortex_points(6000, seed=45), O.KIND_SOURCE_VORTEX,
     F.PowerRelativeGradient(1e-7), 24),
    # tolerances the order cap of 20 cannot meet everywhere (error_success false, pmax 21): orders 19-20
    # come from the deferred second M2L launch, and capped interactions keep the full order
    ("ring-power-rel-1e-10-mac0.55", lambda: W.vortex_ring(8000, seed=38), O.KIND_VORTEX,
     F.PowerRelativeGradient(1e-10), 64, 0.55),
    ("ring-power-rel-1e-12-mac0.55", lambda: W.vortex_ring(6000, seed=39), O.KIND_VORTEX,
     F.PowerRelativeGradient(1e-12), 48, 0.55),
]

# an order decision whose estimate lies this close (relative) to the tolerance may flip under
# last-bit differences of the rotated coefficients; every other interaction must pick the oracle's order
MARGIN = 1e-9


def _bodies_under(run, branches):
    """Original indices of the target bodies inside the given target branches (and so under every
    L2L from them)."""
    ints, _ = run.branches(0)
    si = run.sort_index(0)
    out = np.zeros(len(si), dtype=bool)
    for b in set(int(x) for x in branches):
        out[si[ints[b, 1]:ints[b, 1] + ints[b, 0]]] = True
    return out


@pytest.mark.parametrize("case", ERR_CASES, ids=[c[0] for c in ERR_CASES])
def test_error_controlled_order_matches_oracle(case):
    name, gen, kind, em, ls = case[:5]
    mac = case[5] if len(case) > 5 else 0.5
    _say(f"[gpu] error-controlled {name}")
    sbuf = gen()
    potential = em.kind == _lib.ERR_POWER_POTENTIAL
    kw = dict(expansion_order=20, leaf_size_source=ls, multipole_acceptance=mac, error_tolerance=em,
              scalar_potential=potential)
    ev, opt = _gpu_eval(kind, sbuf, **kw)
    st = ev.fmm(opt)
    run = _oracle_run(kind, sbuf, p=20, **{k: v for k, v in kw.items() if k != "expansion_order"}).fmm()
    if em.relative:   # estimate_influence!: max_influence per target branch
        mi_g = ev.tree(0).reals[:, 14]
        mi_o = run.branches(0)[1][:, 14]
        np.testing.assert_allclose(mi_g, mi_o, rtol=1e-11, atol=0)
    Pg, Po, margin = ev.m2l_orders(), run.m2l_orders(), run.m2l_margins()
    assert len(Pg) == len(Po)
    flipped = Pg != Po
    match = 1.0 - float(np.mean(flipped)) if len(Po) else 1.0
    _say(f"    same order for {match * 100:.4f}% of {len(Po)} M2L interactions ({int(flipped.sum())} differ, "
         f"{int((margin < MARGIN).sum())} within {MARGIN:g} of the tolerance); orders >= 19: {int((Po >= 19).sum())}; "
         f"pmax gpu {st.pmax} oracle {run.pmax}; error_success gpu {st.error_success} oracle {int(run.error_success)}")
    assert not np.any(flipped & (margin >= MARGIN)), "an order decision away from the tolerance differs"
    assert match >= 0.999
    assert st.pmax == run.pmax and bool(st.error_success) == run.error_success
    # outputs: bodies under no flipped interaction's target branch match the oracle to 1e-11; the others
    # differ by at most the truncation error, bounded by twice the oracle's own error against direct!
    got, ref = ev.targets(), run.targets()
    lo, hi = (3, 4) if potential else (4, 7)
    scale = np.abs(ref[:, lo:hi]).max()
    moved = _bodies_under(run, run.list(0)[flipped, 0])
    if np.any(~moved):
        err = np.abs(got[~moved, lo:hi] - ref[~moved, lo:hi]).max() / scale
        _say(f"    {int((~moved).sum())} targets under unflipped interactions: max rel diff {err:.3e}")
        assert err < 1e-11, err
    if np.any(moved):
        d = O.direct(kind, sbuf[:, :3], sbuf, scalar_potential=potential)
        bound = 2 * np.abs(ref[:, lo:hi] - d[:, lo:hi]).max() + 1e-11 * scale
        dev = np.abs(got[moved, lo:hi] - ref[moved, lo:hi]).max()
        _say(f"    {int(moved.sum())} targets under flipped interactions: max diff {dev:.3e} (bound {bound:.3e})")
        assert dev <= bound


def test_relative_tolerance_separate_targets_matches_oracle():
    """estimate_influence! with a target system of its own (two tree builds for the estimate, the
    main target tree of other points): max_influence per target branch and the per-interaction
    orders as in the oracle."""
    _say("[gpu] error-controlled ring-power-rel with separate targets")
    sbuf = W.vortex_ring(6000, seed=36)
    txyz = np.ascontiguousarray(W.vortex_ring(3000, seed=37)[:, :3] * 1.02)
    kw = dict(expansion_order=20, leaf_size_source=48, multipole_acceptance=0.5,
              error_tolerance=F.PowerRelativeGradient(1e-6))
    ev, opt = _gpu_eval(O.KIND_VORTEX, sbuf, txyz=txyz, **kw)
    ev.fmm(opt)
    run = _oracle_run(O.KIND_VORTEX, sbuf, txyz=txyz, p=20, **{k: v for k, v in kw.items() if k != "expansion_order"}).fmm()
    np.testing.assert_allclose(ev.tree(0).reals[:, 14], run.branches(0)[1][:, 14], rtol=1e-11, atol=0)
    Pg, Po, margin = ev.m2l_orders(), run.m2l_orders(), run.m2l_margins()
    assert len(Pg) == len(Po) and not np.any((Pg != Po) & (margin >= MARGIN))
    _say(f"    same order for {float(np.mean(Pg == Po)) * 100:.4f}% of {len(Po)} M2L interactions")
    ref = run.targets()
    err = np.abs(ev.targets()[:, 4:7] - ref[:, 4:7]).max() / np.abs(ref[:, 4:7]).max()
    if np.all(Pg == Po):
        assert err < 1e-11, err


MFMA_FMM = [c for c in FMM_CASES if c[0] in ("laplace-4k-p8", "vortex-4k-p8", "ring-6k-p10", "ring-3k-p20", "panel-src-tri-3k-p8")]
MFMA_ERR = [c for c in ERR_CASES if c[0] in ("laplace-rotated-abs", "vortex-power-abs", "ring-power-rel")]


@pytest.mark.parametrize("name,gen,kind,kw", MFMA_FMM, ids=[c[0] for c in MFMA_FMM])
def test_fmm_fp64_matrix_core_m2l(name, gen, kind, kw, monkeypatch):
    """The matrix-core M2L (FMMHIP_M2L_KERNEL=mfma) meets the same bars as the default kernel."""
    monkeypatch.setenv("FMMHIP_M2L_KERNEL", "mfma")
    test_fmm_fp64_matches_oracle(name, gen, kind, kw)


@pytest.mark.parametrize("case", MFMA_ERR, ids=[c[0] for c in MFMA_ERR])
def test_error_controlled_matrix_core_m2l(case, monkeypatch):
    monkeypatch.setenv("FMMHIP_M2L_KERNEL", "mfma")
    test_error_controlled_order_matches_oracle(case)


def test_direct_fp64_matches_oracle():
    _say("[gpu] direct fp64")
    for kind, sbuf in ((O.KIND_SOURCE, W.uniform_sources(3000, seed=41)), (O.KIND_VORTEX, W.vortex_ring(3000, seed=42)),
                       (O.KIND_DIPOLE, W.uniform_dipoles(3000, seed=43)),
                       (O.KIND_VORTEX_GAUSS, W.gaussian_vortex_ring(3000, seed=44)),
                       (O.KIND_SOURCE_VORTEX, W.source_vortex_points(3000, seed=46)),
                       (O.KIND_PANEL_SOURCE_TRI, W.sphere_panels(10, quads=False)[0]),
                       (O.KIND_PANEL_DIPOLE_QUAD, W.sphere_panels(12)[0]),
                       (O.KIND_PANEL_SOURCE_DIPOLE_TRI, W.sphere_source_dipole_panels(10, quads=False)[0]),
                       (O.KIND_PANEL_SOURCE_DIPOLE_QUAD, W.sphere_source_dipole_panels(12)[0])):
        ctx = F.default_context()
        txyz = np.ascontiguousarray(sbuf[:, :3])
        out = np.zeros((len(sbuf), 16))
        out[:, :3] = txyz
        ctx.check(ctx.lib.fmmhip_direct(ctx._h, kind, len(sbuf), txyz.ctypes.data, len(sbuf),
                                        np.ascontiguousarray(sbuf).ctypes.data, 1, 1, 1, _lib.FP64, out.ctypes.data, 0))
        ref = O.direct(kind, txyz, sbuf, scalar_potential=True, gradient=True, hessian=True)
        for lo, hi in ((3, 4), (4, 7), (7, 16)):
            sc = np.abs(ref[:, lo:hi]).max()
            if sc:
                assert np.abs(out[:, lo:hi] - ref[:, lo:hi]).max() / sc < 1e-13


def test_fmm_api_mirror_and_edge_cases():
    _say("[gpu] api mirror + edge cases")
    # fmm!(system) through the user-system interface, vs direct! on the GPU
    g = W.uniform_sources(2000, seed=51)
    sysA = F.Gravitational(np.column_stack([g, np.zeros((len(g), 3))]))
    out = F.fmm(sysA, expansion_order=12, leaf_size=20, multipole_acceptance=0.5, scalar_potential=True)
    assert len(out) == 8 and out[7] is True
    sysB = F.Gravitational(np.column_stack([g, np.zeros((len(g), 3))]))
    F.direct(sysB, scalar_potential=True)
    rel = np.abs(sysA.potential[:, 0] - sysB.potential[:, 0]).max() / np.abs(sysB.potential[:, 0]).max()
    assert rel < 1e-8
    # a single body, two coincident bodies, a root leaf
    for n in (1, 2, 5):
        b = np.zeros((n, 5))
        b[:, 4] = 1.0
        if n == 5:
            b[:, :3] = np.random.default_rng(5).random((5, 3))
        ev, opt = _gpu_eval(O.KIND_SOURCE, b, expansion_order=4, leaf_size_source=20, scalar_potential=True)
        ev.fmm(opt)
        got = ev.targets()
        ref = O.direct(O.KIND_SOURCE, b[:, :3], b, scalar_potential=True)
        np.testing.assert_allclose(got[:, 3:7], ref[:, 3:7], rtol=1e-12, atol=1e-14)
    # an empty target set and an empty source set: nothing to evaluate, no error (src/fmm.jl:1154-1158)
    b = W.uniform_sources(50, seed=7)
    ev, opt = _gpu_eval(O.KIND_SOURCE, b, txyz=np.zeros((0, 3)), expansion_order=4, scalar_potential=True)
    st = ev.fmm(opt)
    assert ev.targets().shape == (0, 16) and st.n_m2l == 0 and st.n_direct == 0 and st.error_success == 1
    tx = np.random.default_rng(8).random((40, 3))
    ev, opt = _gpu_eval(O.KIND_SOURCE, np.zeros((0, 5)), txyz=tx, expansion_order=4, scalar_potential=True)
    st = ev.fmm(opt)
    got = ev.targets()
    np.testing.assert_array_equal(got[:, :3], tx)
    assert not np.any(got[:, 3:]) and st.n_m2l == 0 and st.n_direct == 0
    # unsupported expansion order is an error, not a fallback
    ev, opt = _gpu_eval(O.KIND_SOURCE, W.uniform_sources(100, seed=6), expansion_order=25)
    with pytest.raises(_lib.FmmHipError):
        ev.fmm(opt)


@pytest.mark.parametrize("kind,gen", [(O.KIND_DIPOLE, lambda: W.uniform_dipoles(20000, seed=51)),
                                      (O.KIND_VORTEX_GAUSS, lambda: W.gaussian_vortex_ring(20000, seed=52)),
                                      (O.KIND_SOURCE_VORTEX, lambda: W.source_vortex_points(20000, seed=53))],
                         ids=["dipole", "gauss-ring", "source-vortex"])
def test_mixed_precision_new_kernels(kind, gen):
    """Mixed form of the dipole and Gaussian-core kernels against the FP64 form of the same
    evaluation: the FP32 nearfield's error stays below 1e-4 (max) and 1e-6 (median) per target."""
    _say(f"[gpu] mixed {kind}")
    sbuf = gen()
    res = {}
    for prec in ("fp64", "mixed"):
        ev, opt = _gpu_eval(kind, sbuf, expansion_order=8, leaf_size_source=32, multipole_acceptance=0.4, hessian=True,
                            precision=prec)
        ev.fmm(opt)
        res[prec] = ev.targets()
    e = W.relative_error(res["mixed"][:, 4:7], res["fp64"][:, 4:7])
    _say(f"    mixed vs fp64: max {e[0]:.3e} p99 {e[1]:.3e} median {e[2]:.3e}")
    assert e[0] < 1e-4 and e[2] < 1e-6, e


def test_mixed_precision_ring_error():
    """Mixed form (FP32 P2P with compensated sums) on a 20k vortex ring at p=16 vs FP64 direct."""
    _say("[gpu] mixed precision ring")
    sbuf = W.vortex_ring(20000, seed=61)
    res = {}
    for prec in ("fp64", "mixed"):
        ev, opt = _gpu_eval(O.KIND_VORTEX, sbuf, expansion_order=16, leaf_size_source=64, multipole_acceptance=0.4,
                            precision=prec)
        ev.fmm(opt)
        res[prec] = ev.targets()
    ctx = F.default_context()
    txyz = np.ascontiguousarray(sbuf[:, :3])
    ref = np.zeros((len(sbuf), 16))
    ref[:, :3] = txyz
    ctx.check(ctx.lib.fmmhip_direct(ctx._h, O.KIND_VORTEX, len(sbuf), txyz.ctypes.data, len(sbuf),
                                    np.ascontiguousarray(sbuf).ctypes.data, 0, 1, 0, _lib.FP64, ref.ctypes.data, 0))
    e64 = W.relative_error(res["fp64"][:, 4:7], ref[:, 4:7])
    e32 = W.relative_error(res["mixed"][:, 4:7], ref[:, 4:7])
    _say(f"    fp64 form rel err max {e64[0]:.3e} p99 {e64[1]:.3e};  mixed max {e32[0]:.3e} p99 {e32[1]:.3e}")
    assert e64[1] < 1e-6
    assert e32[1] < 1e-5


@pytest.mark.xfail(strict=False, reason="the mixed form's FP32 nearfield sets an error floor above 1e-6: the "
                                        "1M-ring bench measures max 1.44e-6 (velocity) and 1.81e-6 (gradient)")
def test_mixed_form_meets_bench_bar():
    """The bench's accuracy bar for the mixed form at its settings (PowerRelativeGradient(1e-9), p <= 20,
    leaf 64, MAC 0.55) on a 100k ring: max relative error of velocity and gradient <= 1e-6 against the FP64
    direct sum on 4,000 sampled targets. Expected to fail until the FP32 pair terms change."""
    _say("[gpu] mixed form at the bench bar")
    sbuf = W.vortex_ring(100000, seed=62)
    ev, opt = _gpu_eval(O.KIND_VORTEX, sbuf, expansion_order=20, leaf_size_source=64, multipole_acceptance=0.55,
                        error_tolerance=F.PowerRelativeGradient(1e-9), hessian=True, precision="mixed")
    ev.fmm(opt)
    got = ev.targets()
    idx = np.sort(np.random.default_rng(5).choice(len(sbuf), 4000, replace=False))
    txyz = np.ascontiguousarray(sbuf[idx, :3])
    ref = np.zeros((len(idx), 16))
    ref[:, :3] = txyz
    ctx = F.default_context()
    ctx.check(ctx.lib.fmmhip_direct(ctx._h, O.KIND_VORTEX, len(idx), txyz.ctypes.data, len(sbuf),
                                    np.ascontiguousarray(sbuf).ctypes.data, 0, 1, 1, _lib.FP64, ref.ctypes.data, 0))
    ev_v = W.relative_error(got[idx, 4:7], ref[:, 4:7])
    ev_g = W.relative_error(got[idx, 7:16], ref[:, 7:16])
    _say(f"    mixed form: velocity max {ev_v[0]:.3e}, gradient max {ev_g[0]:.3e}")
    assert ev_v[0] <= 1e-6 and ev_g[0] <= 1e-6, (ev_v[0], ev_g[0])


def test_direct_on_tuples_sums_every_source_system():
    """direct!((t1, t2), (s1, s2)) (src/direct.jl:43-75): every source system acts on every target system, each
    target system with its own switches; against the oracle's direct sums."""
    _say("[gpu] direct on tuples")
    src = F.Gravitational(W.uniform_sources(2000, seed=47))
    vb = W.vortex_ring(1500, seed=48)
    vor = F.VortexParticles(vb[:, :3], vb[:, 4:7], vb[:, 3])
    outs = F.direct((src, vor), (src, vor), scalar_potential=[True, False], gradient=True, hessian=[False, True])
    assert isinstance(outs, list) and len(outs) == 2
    for i, t in enumerate((src, vor)):
        ref = np.zeros((t.get_n_bodies(), 16))
        for s in (src, vor):
            kind = O.KIND_SOURCE if s is src else O.KIND_VORTEX
            ref[:, 3:] += O.direct(kind, t.positions(), s.source_buffer(), scalar_potential=i == 0, gradient=True,
                                   hessian=i == 1)[:, 3:]
        for lo, hi in ((3, 4), (4, 7), (7, 16)):
            sc = np.abs(ref[:, lo:hi]).max()
            if sc:
                assert np.abs(outs[i][:, lo:hi] - ref[:, lo:hi]).max() / sc < 1e-13, (i, lo)
    # no Hessian where it was not asked for (the point-source kernel writes its potential whatever the switches,
    # test/gravitational.jl:92-98, so the second target system's potential slot is not empty)
    assert np.abs(outs[0][:, 7:16]).max() == 0 and np.abs(outs[1][:, 7:16]).max() > 0
