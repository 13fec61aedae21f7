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
