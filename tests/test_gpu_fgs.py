"""GPU FastGaussSeidel (fmmhip_fgs_*) against the oracle's restatement of src/solve.jl.

Bars: influence matrices within 1e-13 relative of the oracle (same kernels, unit strengths);
solved strengths within 1e-9 of the oracle's after the same number of sweeps (the GPU solves
each block with a precomputed inverse, the reference and the oracle factorise: rounding only);
mean-squared-residual history within 1e-6 relative; the reference's own full-solve check."""
import numpy as np
import pytest

import fastmultipole_amd as F
from fastmultipole_amd import _lib, workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _gpu(kind, rows, **kw):
    influence = kw.pop("influence", _lib.INFLUENCE_POTENTIAL)
    ev = F.Evaluation()
    ev.set_systems(kind, None, rows)
    opt = F.make_options(expansion_order=kw.get("expansion_order", 4), leaf_size_source=kw.get("leaf_size", 30),
                         multipole_acceptance=kw.get("multipole_acceptance", 0.5),
                         shrink_recenter=kw.get("shrink_recenter", True), interaction_list_method=F.Barba())
    ev.ctx.check(ev.ctx.lib.fmmhip_fgs_setup(ev._h, opt, influence))
    return ev


def _leaf_info(ev):
    n = ev.ctx.lib.fmmhip_fgs_n_leaves(ev._h)
    out = np.zeros((n, 4), dtype=np.int64)
    ev.ctx.check(ev.ctx.lib.fmmhip_fgs_leaf_info(ev._h, out.ctypes.data))
    return out


def _matrix(ev, which, k, info):
    rows = int(info[k, 1] if which == 0 else info[k, 2])
    buf = np.zeros(max(1, rows * int(info[k, 1])))
    ev.ctx.check(ev.ctx.lib.fmmhip_fgs_get_matrix(ev._h, which, k, buf.ctypes.data))
    return buf[:rows * int(info[k, 1])].reshape(int(info[k, 1]), rows).T


def _solve(ev, ext, max_it, tol):
    import ctypes as C
    n = ev.ns
    st = np.zeros(n)
    hist = np.zeros(max(1, max_it))
    sw = C.c_int32(0)
    e = np.ascontiguousarray(ext, dtype=np.float64)
    ev.ctx.check(ev.ctx.lib.fmmhip_fgs_solve(ev._h, e.ctypes.data, 0, max_it, tol, st.ctypes.data, C.byref(sw),
                                             hist.ctypes.data))
    return st, ev.targets(), sw.value, hist


CASES = [
    ("sources-1k-ls100", O.KIND_SOURCE, lambda: W.uniform_sources(1000, seed=1234), dict(leaf_size=100)),
    ("quads-sphere-potential", None, lambda: W.sphere_panels(10, quads=True), dict(leaf_size=30, expansion_order=6)),
    ("tris-sphere-normal", None, lambda: W.sphere_panels(8, quads=False),
     dict(leaf_size=20, expansion_order=6, influence=_lib.INFLUENCE_NORMAL_GRADIENT)),
    # leaves of more than 101 bodies: self blocks inverted in device memory, panels staged in chunks
    ("quads-sphere-ls400", None, lambda: W.sphere_panels(20, quads=True), dict(leaf_size=400, expansion_order=6)),
    ("tris-sphere-normal-ls350", None, lambda: W.sphere_panels(14, quads=False),
     dict(leaf_size=350, expansion_order=6, influence=_lib.INFLUENCE_NORMAL_GRADIENT)),
]


def test_fgs_large_leaf_cases_have_large_leaves():
    for name, kind, gen, kw in CASES[3:]:
        _, rows = _make(kind, gen)
        og = O.GaussSeidel(_make(kind, gen)[0], rows, expansion_order=kw["expansion_order"], multipole_acceptance=0.5,
                           leaf_size=kw["leaf_size"], influence=kw.get("influence", 0))
        assert og.leaves()[:, 1].max() > 101, name


def _make(kind, gen):
    data = gen()
    if kind is None:
        rows, kind = data
    else:
        rows = data
    return kind, rows


@pytest.mark.parametrize("name,kind,gen,kw", CASES, ids=[c[0] for c in CASES])
def test_fgs_matrices_match_oracle(name, kind, gen, kw):
    print(f"[gpu] fgs matrices {name}", flush=True)
    kind, rows = _make(kind, gen)
    ev = _gpu(kind, rows, **dict(kw))
    og = O.GaussSeidel(kind, rows, expansion_order=kw.get("expansion_order", 4), multipole_acceptance=0.5,
                       leaf_size=kw.get("leaf_size", 30), influence=kw.get("influence", 0))
    gi, oi = _leaf_info(ev), og.leaves()
    np.testing.assert_array_equal(gi, oi)
    for k in range(len(gi)):
        A, B = _matrix(ev, 0, k, gi), og.self_matrix(k)
        assert np.abs(A - B).max() <= 1e-13 * max(np.abs(B).max(), 1e-300), (name, k)
        if gi[k, 2]:
            A, B = _matrix(ev, 1, k, gi), og.nonself_matrix(k)
            assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max(), (name, k)
            rows_g = np.zeros(int(gi[k, 2]), dtype=np.int64)
            ev.ctx.check(ev.ctx.lib.fmmhip_fgs_nonself_rows(ev._h, k, rows_g.ctypes.data))
            np.testing.assert_array_equal(rows_g, og.nonself_rows(k))


@pytest.mark.parametrize("name,kind,gen,kw", CASES[1:], ids=[c[0] for c in CASES[1:]])
def test_fgs_solve_matches_oracle(name, kind, gen, kw):
    print(f"[gpu] fgs solve {name}", flush=True)
    kind, rows = _make(kind, gen)
    n = len(rows)
    ext = -np.random.default_rng(3).standard_normal(n) if kw.get("influence") else -np.ones(n)
    ev = _gpu(kind, rows, **dict(kw))
    og = O.GaussSeidel(kind, rows, expansion_order=kw.get("expansion_order", 4), multipole_acceptance=0.5,
                       leaf_size=kw.get("leaf_size", 30), influence=kw.get("influence", 0))
    for max_it in (0, 1, 4):
        sg, tg, swg, hg = _solve(ev, ext, max_it, 1e-14)
        so, to, swo, ho = og.solve(ext, max_iterations=max_it, tolerance=1e-14)
        assert swg == swo
        np.testing.assert_allclose(hg[:max_it], ho[:max_it], rtol=1e-6)
        assert np.abs(sg - so).max() <= 1e-9 * np.abs(so).max(), (name, max_it)
        for lo, hi in ((3, 4), (4, 7)):
            sc = np.abs(to[:, lo:hi]).max()
            assert np.abs(tg[:, lo:hi] - to[:, lo:hi]).max() <= 1e-9 * sc, (name, max_it, lo)


def test_fgs_full_solve_recovers_strengths():
    """test/solve_test.jl:309-350 through the host mirror: one leaf, perturbed strengths."""
    print("[gpu] fgs full solve", flush=True)
    sb = W.uniform_sources(100, seed=123)
    s_true = sb[:, 4].copy()
    phi = O.direct(O.KIND_SOURCE, sb[:, :3], sb, scalar_potential=True, gradient=False)[:, 3]
    rng = np.random.default_rng(5)
    sb2 = sb.copy()
    sb2[:, 4] = s_true + np.round(s_true, 1) * 100 * (rng.random(100) - 0.5)
    system = F.Gravitational(sb2)
    system.potential[:, 0] = -phi
    fgs = F.FastGaussSeidel(system, expansion_order=4, multipole_acceptance=0.5, leaf_size=100, shrink_recenter=False)
    st, tg, sweeps, hist = F.solve(system, fgs, max_iterations=10, tolerance=1e-3)
    assert sweeps == 1
    np.testing.assert_allclose(system.strength, s_true, rtol=0, atol=1e-10)
    np.testing.assert_allclose(system.potential[:, 0], phi, rtol=1e-10)


def test_fgs_rejects_unsupported_systems():
    ev = F.Evaluation()
    ev.set_systems(_lib.POINT_VORTEX, None, W.vortex_ring(200, seed=1))
    opt = F.make_options(expansion_order=4, leaf_size_source=30)
    assert ev.ctx.lib.fmmhip_fgs_setup(ev._h, opt, 0) == _lib.FMMHIP_EINVAL
    ev2 = F.Evaluation()
    ev2.set_systems(_lib.POINT_SOURCE, None, W.uniform_sources(200, seed=2))
    assert ev2.ctx.lib.fmmhip_fgs_setup(ev2._h, opt, _lib.INFLUENCE_NORMAL_GRADIENT) == _lib.FMMHIP_EINVAL


def test_fgs_graph_replays_match_direct_launches(monkeypatch):
    """Solve iterations after the first replay one captured graph (sweep, farfield pass, residual);
    FMMHIP_GRAPH=0 launches the same kernels one by one. Eight iterations (six replays, each followed
    by the host's residual read-back): the residual history and the solved strengths agree to
    rounding (the sweep's non-self updates are floating-point atomic adds), and so do the outputs."""
    import time
    rows, kind = W.sphere_panels(12, quads=True)
    n = len(rows)
    ext = -np.random.default_rng(5).standard_normal(n)
    out = {}
    for graph in ("1", "0"):
        monkeypatch.setenv("FMMHIP_GRAPH", graph)
        ev = _gpu(kind, rows, leaf_size=30, expansion_order=6, influence=_lib.INFLUENCE_NORMAL_GRADIENT)
        _solve(ev, ext, 8, 1e-30)   # warm-up (tables, buffers)
        t0 = time.perf_counter()
        out[graph] = _solve(ev, ext, 8, 1e-30)
        print(f"[gpu] fgs 8 iterations, graph={graph}: {1e3 * (time.perf_counter() - t0):.2f} ms", flush=True)
    (sg, tg, swg, hg), (sd, td, swd, hd) = out["1"], out["0"]
    assert swg == swd == 8
    # near convergence the residual is at the rounding floor of the atomic sweep: relative bar plus a floor
    np.testing.assert_allclose(hg, hd, rtol=1e-6, atol=1e-9 * hd[0])
    assert np.abs(sg - sd).max() <= 1e-10 * np.abs(sd).max()
    sc = np.abs(td[:, 4:7]).max()
    assert np.abs(tg[:, 4:7] - td[:, 4:7]).max() <= 1e-10 * sc
