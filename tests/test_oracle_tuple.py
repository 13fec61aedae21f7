"""Pins the oracle's fmm! on a tuple of systems (one tree over the bodies of every system, each source
system's nearfield and P2M with its own kernel; src/fmm.jl:3-40, src/fmm.jl:331-345, src/tree.jl:521-549).

* a tuple of one system is the plain one-system run, bit for bit (the per-system ranges cover each leaf,
  and a 7-row buffer only changes the row stride);
* a tuple of a source, a vortex and a dipole system converges to the sum of the three direct sums over
  every body (a dropped system, a kernel applied to the wrong range or a wrong Lamb-Helmholtz switch fails
  this by orders of magnitude);
* the same for quadrilateral panels (source and dipole) with vortex particles, the reference's motivating
  case, whose buffer rows differ per system;
* the error-controlled path with its influence estimate runs on the tuple and meets its tolerance.
"""
import numpy as np
import pytest

from oracle import oracle as O


# strength scales that give each system a like share of the velocity (a dipole's field falls off as 1/r^3)
_SCALE = {O.KIND_SOURCE: 10.0, O.KIND_VORTEX: 1.0, O.KIND_DIPOLE: 0.05, O.KIND_VORTEX_GAUSS: 1.0}


def _system(kind, n, rng, shift):
    d = np.zeros((n, O.data_per_body(kind)))
    d[:, :3] = rng.random((n, 3)) * 2.0 + shift
    if kind == O.KIND_VORTEX_GAUSS:
        d[:, 3] = 0.004
    if kind == O.KIND_SOURCE:
        d[:, 4] = rng.normal(size=n) / n * _SCALE[kind]
    else:
        d[:, 4:7] = rng.normal(size=(n, 3)) / n * _SCALE[kind]
    return d


@pytest.mark.parametrize("kind", [O.KIND_SOURCE, O.KIND_VORTEX])
def test_tuple_of_one_system_is_the_plain_run(kind):
    rng = np.random.default_rng(21)
    d = _system(kind, 1800, rng, 0.0)
    kw = dict(p=8, leaf_size_source=24, multipole_acceptance=0.5, scalar_potential=True, gradient=True, hessian=True)
    one = O.Run(kind, d[:, :3], d, **kw).fmm().targets()
    tup = O.Run.tuple([kind], [d], **kw).fmm().targets()
    np.testing.assert_array_equal(tup, one)


def _direct_sum(kinds, datas, xyz, **kw):
    out = 0.0
    for k, d in zip(kinds, datas):
        out = out + O.direct(k, xyz, d, **kw)
    return out


@pytest.mark.parametrize("kinds,slots", [
    ((O.KIND_SOURCE, O.KIND_VORTEX, O.KIND_DIPOLE), ((4, 7),)),
    # the point-source kernel of test/gravitational.jl writes no Hessian, so Hessians are compared without it
    ((O.KIND_VORTEX, O.KIND_DIPOLE, O.KIND_VORTEX_GAUSS), ((4, 7), (7, 16))),
])
def test_mixed_tuple_converges_to_direct_sums(kinds, slots):
    rng = np.random.default_rng(22)
    datas = [_system(k, n, rng, s) for k, n, s in zip(kinds, (900, 1400, 700), (0.0, 0.3, 0.6))]
    run = O.Run.tuple(kinds, datas, p=16, leaf_size_source=20, multipole_acceptance=0.4, gradient=True,
                      hessian=True).fmm()
    assert len(run.list(0)) > 0   # the far field takes part
    got = run.targets()
    xyz = np.concatenate([d[:, :3] for d in datas])
    ref = _direct_sum(kinds, datas, xyz, gradient=True, hessian=True)
    for lo, hi in slots:
        err = np.abs(got[:, lo:hi] - ref[:, lo:hi]).max() / np.abs(ref[:, lo:hi]).max()
        assert err < 1e-6, (lo, err)
    # each system matters: leaving one out moves the result far beyond the expansion error
    for drop in range(len(kinds)):
        part = _direct_sum([k for i, k in enumerate(kinds) if i != drop], [d for i, d in enumerate(datas) if i != drop],
                           xyz, gradient=True)
        rel = np.linalg.norm(part[:, 4:7] - ref[:, 4:7], axis=1) / np.linalg.norm(ref[:, 4:7], axis=1)
        assert np.median(rel) > 0.05, (drop, np.median(rel))


def _sphere_quads(m, rng):
    """Source or dipole quadrilaterals tiling a cube's surface at distance 1 (rows of the panel kernels)."""
    out = []
    t = np.linspace(-1.0, 1.0, m + 1)
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a, b = [k for k in range(3) if k != axis]
            for i in range(m):
                for j in range(m):
                    v = np.zeros((4, 3))
                    for q, (di, dj) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
                        v[q, axis] = sgn
                        v[q, a] = t[i + di]
                        v[q, b] = t[j + dj]
                    if sgn < 0:
                        v = v[::-1]
                    cen = v.mean(axis=0)
                    rad = np.linalg.norm(v - cen, axis=1).max()
                    out.append(np.concatenate([cen, [rad, rng.normal() / (6 * m * m)], v.ravel()]))
    return np.array(out)


@pytest.mark.parametrize("panel", [O.KIND_PANEL_SOURCE_QUAD, O.KIND_PANEL_DIPOLE_QUAD])
def test_panels_and_vortices_converge_to_direct_sums(panel):
    """The reference's motivating tuple: panels and vortex particles in one tree (16-row panel buffer,
    points padded)."""
    rng = np.random.default_rng(24)
    quads = _sphere_quads(12, rng)
    quads[:, 4] *= 30.0   # a like share of the velocity for each system
    vort = _system(O.KIND_VORTEX, 1500, rng, -0.5)
    vort[:, :3] = rng.random((1500, 3)) * 1.4 - 0.7
    kinds, datas = [panel, O.KIND_VORTEX], [quads, vort]
    run = O.Run.tuple(kinds, datas, p=16, leaf_size_source=10, multipole_acceptance=0.5, gradient=True).fmm()
    assert len(run.list(0)) > 0
    got = run.targets()
    xyz = np.concatenate([d[:, :3] for d in datas])
    ref = _direct_sum(kinds, datas, xyz, gradient=True)
    err = np.abs(got[:, 4:7] - ref[:, 4:7]).max() / np.abs(ref[:, 4:7]).max()
    assert err < 1e-6, err
    for drop in range(2):
        part = O.direct(kinds[1 - drop], xyz, datas[1 - drop], gradient=True)
        rel = np.linalg.norm(part[:, 4:7] - ref[:, 4:7], axis=1) / np.linalg.norm(ref[:, 4:7], axis=1)
        assert np.median(rel) > 0.05, (drop, np.median(rel))


def test_tuple_error_controlled_meets_tolerance():
    rng = np.random.default_rng(23)
    kinds = [O.KIND_VORTEX, O.KIND_SOURCE]
    datas = [_system(k, n, rng, s) for k, n, s in zip(kinds, (1600, 1100), (0.0, 0.2))]
    run = O.Run.tuple(kinds, datas, p=20, leaf_size_source=24, multipole_acceptance=0.5, gradient=True,
                      error_kind=O.ERR_POWER_GRADIENT, relative=True, eps=1e-6).fmm()
    got = run.targets()
    xyz = np.concatenate([d[:, :3] for d in datas])
    ref = _direct_sum(kinds, datas, xyz, gradient=True)
    v = np.linalg.norm(ref[:, 4:7], axis=1)
    err = np.linalg.norm(got[:, 4:7] - ref[:, 4:7], axis=1)
    assert run.error_success
    assert np.max(err / v) < 1e-4


# --------------------------------------------------------------- the leaf rule, per system
def _rule(counts, ls, method):
    """src/tree.jl:492-569: does a branch with these per-system counts split?"""
    if method == O.METHOD_BARBA:
        return any(c > l for c, l in zip(counts, ls))
    frac = 0.0
    for c, l in zip(counts, ls):
        frac += c / (l * l)
    frac *= min(ls)
    thr = 0.5 if method == O.METHOD_SELFTUNING else 2.8284271247461903
    return frac > thr and sum(counts) > 1


@pytest.mark.parametrize("method", [O.METHOD_SELFTUNING_TREESTOP, O.METHOD_SELFTUNING, O.METHOD_BARBA])
@pytest.mark.parametrize("leaf_sizes", [None, (30, 6, 14)])
def test_every_branch_follows_the_leaf_rule(method, leaf_sizes):
    """Every branch that has children passes the reference's leaf test on its per-system counts, and every
    leaf fails it (for one system and for three with their own leaf sizes)."""
    rng = np.random.default_rng(25)
    kinds = [O.KIND_SOURCE, O.KIND_VORTEX, O.KIND_DIPOLE]
    datas = [_system(k, n, rng, s) for k, n, s in zip(kinds, (1500, 900, 1200), (0.0, 0.4, 0.8))]
    ls = list(leaf_sizes) if leaf_sizes else [12, 12, 12]
    runs = [("tuple", O.Run.tuple(kinds, datas, p=4, leaf_size_source=12, leaf_sizes=leaf_sizes, method=method), None)]
    if leaf_sizes is None:
        runs.append(("one", O.Run(O.KIND_VORTEX, datas[1][:, :3], datas[1], p=4, leaf_size_source=12, method=method), 1))
    for name, run, only in runs:
        run.trees()
        for which in (0, 1):
            ints, _ = run.branches(which)
            si = run.sort_index(which)
            off = np.concatenate([[0], np.cumsum([len(d) for d in datas])]) if only is None else np.array([0, len(si)])
            sys_of = np.searchsorted(off, si, side="right") - 1
            n_split = 0
            for b in ints:
                members = sys_of[b[1]:b[1] + b[0]]
                counts = [int(np.sum(members == i)) for i in range(len(off) - 1)]
                rule_ls = ls if only is None else [12]
                assert _rule(counts, rule_ls, method) == (b[2] > 0), (name, which, b, counts)
                n_split += b[2] > 0
            assert n_split > 3
