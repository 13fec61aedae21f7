"""Pins the oracle's point-dipole and Gaussian-core vortex kernels.

The reference has no direct! kernel for either (only the dipole's multipole
expansion, src/bodytomultipole.jl:431-463, and an analytic potential check,
test/bodytomultipole_test.jl:508-514), so the kernels are pinned by mathematics:

* point dipole: the direct kernel must equal the reference's own expansion chain
  P2M(dipole) -> M2L -> L2P (each link pinned by golden vectors in
  test_oracle_golden.py) for the potential, its gradient slot and its Hessian
  slot, and the potential must equal the reference test's closed form q.d/(4 pi r^3);
* Gaussian-core vortex: the core factor g must equal the chi distribution's CDF
  with 3 degrees of freedom (scipy), sigma = 0 and r >> sigma must reduce to the
  singular kernel of test/vortex.jl, and the velocity-gradient slot must equal
  central finite differences of the velocity.
"""
import math

import numpy as np
import pytest
from scipy.stats import chi

from fastmultipole_amd import workloads as W
from oracle import oracle as O


def _direct(kind, xt, src, **kw):
    return O.direct(kind, np.atleast_2d(xt), np.atleast_2d(src), **kw)[0]


def test_dipole_direct_matches_expansion_chain():
    rng = np.random.default_rng(5)
    p = 24
    xs = np.array([0.1, -0.2, 0.3])
    q = np.array([0.7, -0.4, 1.1])
    cs = xs + np.array([0.05, -0.08, 0.02])
    M = O.p2m(O.KIND_DIPOLE, xs - cs, q, p)
    for _ in range(4):
        xt = xs + rng.normal(size=3) * 3 + np.array([6.0, 0, 0])
        tc = xt + rng.normal(size=3) * 0.05
        L, _, _ = O.m2l(O.geom(tc), O.geom(cs), M, p)
        phi, v, h = O.l2p(xt - tc, L, p)
        ref = _direct(O.KIND_DIPOLE, xt, np.array([*xs, 0.0, *q]), scalar_potential=True, gradient=True, hessian=True)
        d = xt - xs
        r = np.linalg.norm(d)
        assert phi == pytest.approx(np.dot(q, d) / r ** 3 / 4 / math.pi, rel=1e-10)
        assert ref[3] == pytest.approx(phi, rel=1e-10)
        np.testing.assert_allclose(ref[4:7], v, rtol=1e-9, atol=1e-12 * np.abs(v).max())
        np.testing.assert_allclose(ref[7:16], np.asarray(h).reshape(-1), rtol=1e-8, atol=1e-11 * np.abs(h).max())


def test_dipole_gradient_slot_is_minus_grad_phi():
    xs, q = np.array([0.0, 0.0, 0.0]), np.array([0.3, 0.5, -0.2])
    src = np.array([*xs, 0.0, *q])
    xt = np.array([0.7, -0.4, 0.9])
    out = _direct(O.KIND_DIPOLE, xt, src, scalar_potential=True, gradient=True, hessian=True)
    h = 1e-5
    for a in range(3):
        e = np.zeros(3); e[a] = h
        fp = _direct(O.KIND_DIPOLE, xt + e, src, scalar_potential=True, gradient=True, hessian=True)
        fm = _direct(O.KIND_DIPOLE, xt - e, src, scalar_potential=True, gradient=True, hessian=True)
        assert -(fp[3] - fm[3]) / (2 * h) == pytest.approx(out[4 + a], rel=1e-7)
        # hessian slot = d(gradient slot)/dx_a (symmetric)
        np.testing.assert_allclose((fp[4:7] - fm[4:7]) / (2 * h), out[7 + 3 * a:10 + 3 * a], rtol=1e-6, atol=1e-9)


def test_gaussian_core_factor_is_chi3_cdf():
    xs, G = np.zeros(3), np.array([0.2, -0.7, 0.4])
    for sig in (0.05, 0.3, 1.0):
        for r in (0.01, 0.2, 0.9, 2.5):
            xt = np.array([r, 0.0, 0.0]) @ np.array([[0.6, 0.8, 0.0], [-0.8, 0.6, 0.0], [0, 0, 1.0]])
            sg = _direct(O.KIND_VORTEX_GAUSS, xt, np.array([*xs, sig, *G]))
            s0 = _direct(O.KIND_VORTEX, xt, np.array([*xs, 0.0, *G]))
            np.testing.assert_allclose(sg[4:7], chi(3).cdf(r / sig) * s0[4:7], rtol=1e-12, atol=1e-300)


def test_gaussian_core_limits():
    rng = np.random.default_rng(6)
    src = np.column_stack([rng.normal(size=(20, 3)), np.zeros(20), rng.normal(size=(20, 3))])
    xt = rng.normal(size=(15, 3)) * 2
    a = O.direct(O.KIND_VORTEX_GAUSS, xt, src, gradient=True, hessian=True)
    b = O.direct(O.KIND_VORTEX, xt, src, gradient=True, hessian=True)
    np.testing.assert_array_equal(a[:, 4:16], b[:, 4:16])          # sigma = 0: singular kernel
    far = src.copy(); far[:, 3] = 1e-3                                  # r >> sigma: g = 1 to rounding
    c = O.direct(O.KIND_VORTEX_GAUSS, xt, far, gradient=True, hessian=True)
    np.testing.assert_allclose(c[:, 4:16], b[:, 4:16], rtol=1e-12, atol=1e-14)


def test_gaussian_core_gradient_is_finite_difference():
    src = np.array([0.1, -0.3, 0.2, 0.4, 0.5, -0.2, 0.8])
    xt = np.array([0.35, -0.1, 0.5])
    out = _direct(O.KIND_VORTEX_GAUSS, xt, src, gradient=True, hessian=True)
    h = 1e-6
    for a in range(3):
        e = np.zeros(3); e[a] = h
        fp = _direct(O.KIND_VORTEX_GAUSS, xt + e, src, gradient=True, hessian=False)
        fm = _direct(O.KIND_VORTEX_GAUSS, xt - e, src, gradient=True, hessian=False)
        np.testing.assert_allclose((fp[4:7] - fm[4:7]) / (2 * h), out[7 + 3 * a:10 + 3 * a], rtol=1e-6, atol=1e-9)
    assert abs(out[7] + out[11] + out[15]) < 1e-9 * np.abs(out[7:16]).max()   # divergence-free


@pytest.mark.parametrize("kind", [O.KIND_DIPOLE, O.KIND_VORTEX_GAUSS])
def test_fmm_converges_to_direct(kind):
    rng = np.random.default_rng(7)
    n = 1500
    src = np.zeros((n, 7))
    src[:, :3] = rng.random((n, 3))
    src[:, 3] = 0.002 if kind == O.KIND_VORTEX_GAUSS else 0.0
    src[:, 4:7] = rng.normal(size=(n, 3)) / n
    run = O.Run(kind, src[:, :3], src, p=16, leaf_size_source=20, multipole_acceptance=0.4, scalar_potential=True,
                gradient=True, hessian=True).fmm()
    got = run.targets()
    ref = O.direct(kind, src[:, :3], src, scalar_potential=True, gradient=True, hessian=True)
    for lo, hi in ((4, 7), (7, 16)):
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / np.abs(ref[:, lo:hi]).max()
        assert err < 1e-6, (kind, lo, err)


# ------------------------------------------------------------------ panels
def _panel_row(verts, strength):
    verts = [np.asarray(v, float) for v in verts]
    cen = sum(verts[:3]) * 0.33333333333333333 if len(verts) == 3 else sum(verts) / len(verts)
    rad = max(np.linalg.norm(v - cen) for v in verts)
    return np.concatenate([cen, [rad, strength], *verts])


_TRI = [np.array([0.0, 0.29, 0.0]), np.array([0.0, 0.59, 0.0]), np.array([-0.5, 0.59, 0.0])]   # test/bodytomultipole_test.jl:905-907


@pytest.mark.parametrize("kind", [O.KIND_PANEL_SOURCE_TRI, O.KIND_PANEL_DIPOLE_TRI])
def test_panel_expansion_chain_matches_induced(kind):
    """test/bodytomultipole_test.jl:901-1095: panel P2M -> M2L -> L2P equals induced() (velocity atol 1e-12)."""
    row = _panel_row(_TRI, 1.0)
    xt = np.array([0.2, 0.3, 5.7])
    p = 10
    center = row[:3] + np.array([0.1, 0.2, 0.08])
    M = O.p2m_panel(kind, row, center, p)
    tc = xt + np.array([0.0001, -0.0002, 0.003])
    L, _, _ = O.m2l(O.geom(tc), O.geom(center), M, p)
    phi, v, h = O.l2p(xt - tc, L, p)
    ref = O.direct(kind, xt[None, :], row[None, :], scalar_potential=True, gradient=True, hessian=True)[0]
    np.testing.assert_allclose(v, ref[4:7], atol=1e-12)
    assert phi == pytest.approx(ref[3], rel=1e-10)
    np.testing.assert_allclose(np.asarray(h).reshape(-1), ref[7:16], atol=1e-12)


@pytest.mark.parametrize("kind", [O.KIND_PANEL_SOURCE_QUAD, O.KIND_PANEL_DIPOLE_QUAD])
def test_quad_panel_expansion_chain_matches_induced(kind):
    quad = [np.array([0.1, 0.2, 0.0]), np.array([0.5, 0.25, 0.02]), np.array([0.45, 0.6, 0.04]), np.array([0.05, 0.55, 0.02])]
    # make the quadrilateral planar: project the 4th vertex onto the plane of the first three
    n = np.cross(quad[1] - quad[0], quad[2] - quad[0]); n /= np.linalg.norm(n)
    quad[3] = quad[3] - np.dot(quad[3] - quad[0], n) * n
    row = _panel_row(quad, -0.7)
    xt = np.array([1.9, -2.2, 3.1])
    p = 14
    center = row[:3] + np.array([-0.05, 0.03, 0.02])
    M = O.p2m_panel(kind, row, center, p)
    tc = xt + np.array([0.01, 0.02, -0.03])
    L, _, _ = O.m2l(O.geom(tc), O.geom(center), M, p)
    phi, v, h = O.l2p(xt - tc, L, p)
    ref = O.direct(kind, xt[None, :], row[None, :], scalar_potential=True, gradient=True, hessian=True)[0]
    assert phi == pytest.approx(ref[3], rel=1e-9)
    np.testing.assert_allclose(v, ref[4:7], rtol=1e-8, atol=1e-12 * np.abs(ref[4:7]).max())
    np.testing.assert_allclose(np.asarray(h).reshape(-1), ref[7:16], rtol=1e-7, atol=1e-10 * np.abs(ref[7:16]).max())


@pytest.mark.parametrize("kind", [O.KIND_PANEL_SOURCE_TRI, O.KIND_PANEL_DIPOLE_TRI])
def test_panel_far_field_and_derivatives(kind):
    """Far away a panel is a point source of strength q*area (or a dipole q*area*n); near it the
    velocity slot is -grad(phi) and the Hessian slot is d(velocity slot), by finite differences."""
    row = _panel_row(_TRI, 0.8)
    area = 0.5 * np.linalg.norm(np.cross(_TRI[1] - _TRI[0], _TRI[2] - _TRI[0]))
    n = np.cross(_TRI[1] - _TRI[0], _TRI[2] - _TRI[0]); n /= np.linalg.norm(n)
    xt = row[:3] + np.array([300.0, -200.0, 500.0])
    got = O.direct(kind, xt[None, :], row[None, :], scalar_potential=True, gradient=True, hessian=True)[0]
    if kind == O.KIND_PANEL_SOURCE_TRI:
        pt = O.direct(O.KIND_SOURCE, xt[None, :], np.array([[*row[:3], 0.0, 0.8 * area]]), scalar_potential=True)[0]
    else:
        pt = O.direct(O.KIND_DIPOLE, xt[None, :], np.array([[*row[:3], 0.0, *(0.8 * area * n)]]), scalar_potential=True,
                      gradient=True)[0]
    assert got[3] == pytest.approx(pt[3], rel=1e-5)
    np.testing.assert_allclose(got[4:7], pt[4:7], rtol=1e-4, atol=1e-6 * np.abs(pt[4:7]).max())
    xt = np.array([0.05, 0.41, 0.13])
    out = O.direct(kind, xt[None, :], row[None, :], scalar_potential=True, gradient=True, hessian=True)[0]
    h = 1e-6
    for a in range(3):
        e = np.zeros(3); e[a] = h
        fp = O.direct(kind, (xt + e)[None, :], row[None, :], scalar_potential=True, gradient=True)[0]
        fm = O.direct(kind, (xt - e)[None, :], row[None, :], scalar_potential=True, gradient=True)[0]
        assert -(fp[3] - fm[3]) / (2 * h) == pytest.approx(out[4 + a], rel=1e-6, abs=1e-9)
        np.testing.assert_allclose((fp[4:7] - fm[4:7]) / (2 * h), out[7 + 3 * a:10 + 3 * a], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("kind", [O.KIND_PANEL_SOURCE_TRI, O.KIND_PANEL_DIPOLE_QUAD])
def test_panel_fmm_converges_to_direct(kind):
    rng = np.random.default_rng(8)
    n = 800
    nv = 4 if kind >= O.KIND_PANEL_SOURCE_QUAD else 3
    rows = []
    for _ in range(n):
        c = rng.random(3)
        u = rng.normal(size=3); u /= np.linalg.norm(u)
        w = np.cross(u, rng.normal(size=3)); w /= np.linalg.norm(w)
        s = 0.01 + 0.01 * rng.random()
        ang = np.sort(rng.random(nv)) * 2 * math.pi if nv == 3 else np.array([0.1, 1.6, 3.2, 4.7])
        verts = [c + s * (math.cos(t) * u + math.sin(t) * w) for t in ang]
        rows.append(_panel_row(verts, rng.normal() / n))
    src = np.array(rows)
    run = O.Run(kind, src[:, :3], src, p=14, leaf_size_source=20, multipole_acceptance=0.4, scalar_potential=True,
                gradient=True, hessian=True).fmm()
    got = run.targets()
    ref = O.direct(kind, src[:, :3], src, scalar_potential=True, gradient=True, hessian=True)
    for lo, hi in ((3, 4), (4, 7), (7, 16)):
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / np.abs(ref[:, lo:hi]).max()
        assert err < 1e-6, (kind, lo, err)


# ------------------------------------------------------------ Point{SourceVortex}
def _source_vortex(n, seed):
    rng = np.random.default_rng(seed)
    d = np.zeros((n, 8))
    d[:, :3] = rng.random((n, 3))
    d[:, 4] = rng.normal(size=n) / n
    d[:, 5:8] = rng.normal(size=(n, 3)) / n
    return d


def test_source_vortex_direct_is_source_plus_vortex():
    """The Point{SourceVortex} nearfield: the point source's test kernel plus the singular vortex's, with the
    source's Hessian (which the source test kernel leaves out) equal to central differences of its gradient."""
    d = _source_vortex(400, 31)
    xt = np.random.default_rng(32).random((300, 3))
    got = O.direct(O.KIND_SOURCE_VORTEX, xt, d, scalar_potential=True, gradient=True, hessian=True)
    src = O.direct(O.KIND_SOURCE, xt, d[:, :5], scalar_potential=True, gradient=True, hessian=True)
    vor = O.direct(O.KIND_VORTEX, xt, np.column_stack([d[:, :4], d[:, 5:8]]), scalar_potential=True, gradient=True,
                   hessian=True)
    sc = np.abs(got[:, 3:7]).max()
    np.testing.assert_allclose(got[:, 3:7], src[:, 3:7] + vor[:, 3:7], rtol=1e-13, atol=1e-13 * sc)
    h_src = got[:, 7:16] - vor[:, 7:16]
    h = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        gp = O.direct(O.KIND_SOURCE, xt + e, d[:, :5], gradient=True)[:, 4:7]
        gm = O.direct(O.KIND_SOURCE, xt - e, d[:, :5], gradient=True)[:, 4:7]
        # slot (i, j) holds d(gradient_j)/dx_i
        np.testing.assert_allclose(h_src[:, 3 * a:3 * a + 3], (gp - gm) / (2 * h), rtol=1e-5,
                                   atol=1e-6 * np.abs(h_src).max())


def test_source_vortex_fmm_converges_to_direct():
    """Expansions of Point{SourceVortex} (src/bodytomultipole.jl:491-496: the source's, then the vortex's)."""
    d = _source_vortex(3000, 33)
    run = O.Run(O.KIND_SOURCE_VORTEX, d[:, :3], d, p=16, leaf_size_source=12, multipole_acceptance=0.5,
                scalar_potential=True, gradient=True, hessian=True).fmm()
    assert len(run.list(0)) > 0
    got = run.targets()
    ref = O.direct(O.KIND_SOURCE_VORTEX, d[:, :3], d, scalar_potential=True, gradient=True, hessian=True)
    # velocity and its gradient; the far field's scalar potential is not the source's once a vector potential
    # is present (the reference warns of it, src/fmm.jl:783)
    for lo, hi in ((4, 7), (7, 16)):
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / np.abs(ref[:, lo:hi]).max()
        assert err < 1e-6, (lo, err)


# ------------------------------------------------------------ Panel{SourceDipole}
def _source_dipole_rows(quads, m=8):
    """Source-dipole panels of the sphere's panels: the source strength, then a dipole strength."""
    rows, kind = W.sphere_panels(m, quads=quads)
    sd = np.insert(rows, 5, np.random.default_rng(34).normal(size=len(rows)) / len(rows), axis=1)
    return rows, kind, sd, O.KIND_PANEL_SOURCE_DIPOLE_QUAD if quads else O.KIND_PANEL_SOURCE_DIPOLE_TRI


@pytest.mark.parametrize("quads", [False, True])
def test_source_dipole_panel_is_source_plus_dipole(quads):
    """Panel{SourceDipole} (test/panels.jl:323-429): the source panel's closed form with the first strength plus
    the dipole panel's with the second."""
    rows, kind, sd, sdk = _source_dipole_rows(quads)
    dip = np.delete(sd, 4, axis=1)
    dip_kind = kind + 1   # the dipole panel kind of the same shape
    xt = np.random.default_rng(35).random((200, 3)) * 2.4 - 1.2
    got = O.direct(sdk, xt, sd, scalar_potential=True, gradient=True, hessian=True)
    ref = (O.direct(kind, xt, rows, scalar_potential=True, gradient=True, hessian=True)[:, 3:]
           + O.direct(dip_kind, xt, dip, scalar_potential=True, gradient=True, hessian=True)[:, 3:])
    np.testing.assert_allclose(got[:, 3:], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("quads", [False, True])
def test_source_dipole_panel_fmm_is_source_plus_dipole_fmm(quads):
    """Expansions of Panel{SourceDipole} (src/bodytomultipole.jl:717-720: the source's, then the dipole's): the
    source-dipole FMM equals the source-panel FMM plus the dipole-panel FMM on the same panels (same trees and
    lists, linear passes), and it converges to the direct sum."""
    rows, kind, sd, sdk = _source_dipole_rows(quads, m=16)
    dip = np.delete(sd, 4, axis=1)
    kw = dict(p=14, leaf_size_source=12, multipole_acceptance=0.5, scalar_potential=True, gradient=True)
    run = O.Run(sdk, sd[:, :3], sd, **kw).fmm()
    assert len(run.list(0)) > 0
    got = run.targets()
    parts = (O.Run(kind, rows[:, :3], rows, **kw).fmm().targets()[:, 3:7]
             + O.Run(kind + 1, dip[:, :3], dip, **kw).fmm().targets()[:, 3:7])
    np.testing.assert_allclose(got[:, 3:7], parts, rtol=1e-11, atol=1e-12 * np.abs(parts).max())
    ref = O.direct(sdk, sd[:, :3], sd, scalar_potential=True, gradient=True)
    for lo, hi in ((3, 4), (4, 7)):
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / np.abs(ref[:, lo:hi]).max()
        assert err < 1e-6, (lo, err)
