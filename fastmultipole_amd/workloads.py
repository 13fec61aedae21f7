"""Seeded synthetic particle sets for the benchmark configurations.

All inputs are generated here (numpy PCG64), never from Julia's RNG, so the
GPU path and the oracle consume identical data.
"""
import math

import numpy as np


def uniform_sources(n, seed=1, strength_scale=None):
    """N point sources, positions U[0,1]^3, strengths U[0,1] (scaled by strength_scale if given).

    Returns an (n, 5) source buffer [x, y, z, radius=0, q]."""
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 5))
    b[:, :3] = rng.random((n, 3))
    b[:, 4] = rng.random(n) * (1.0 if strength_scale is None else strength_scale)
    return b


def uniform_vortices(n, seed=1):
    """test/vortex.jl:32-41 distribution: positions U[0,1]^3, strengths (2U-1-0.5)*2/n. (n, 7) buffer."""
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 7))
    b[:, :3] = rng.random((n, 3))
    b[:, 4:7] = (2.0 * rng.random((n, 3)) - 1.0 - 0.5) * 2.0 / n
    return b


def source_vortex_points(n, seed=6):
    """Point{SourceVortex} bodies: positions U[0,1]^3, source strengths N(0,1)/n, vortex strengths as
    uniform_vortices. (n, 8) buffer [x, y, z, sigma, q, Gamma]."""
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8))
    b[:, :3] = rng.random((n, 3))
    b[:, 4] = rng.normal(size=n) / n
    b[:, 5:8] = (2.0 * rng.random((n, 3)) - 1.0 - 0.5) * 2.0 / n
    return b


def vortex_ring(n, seed=2, radius=1.0, core_sigma=0.1, noise=1e-3):
    """Vortex ring of radius `radius`: angle U[0, 2pi), cross-section offsets N(0, sigma^2) in the
    radial and axial directions (a Gaussian core), vector strength tangential with magnitude
    2*pi*radius/n (unit total circulation) plus relative N(0,1) noise of size `noise` per component.
    Particles are points (buffer radius 0) and interact through the singular vortex kernel of
    test/vortex.jl. Returns an (n, 7) buffer [x, y, z, 0, Gamma]."""
    rng = np.random.default_rng(seed)
    phi = rng.random(n) * 2.0 * math.pi
    a = rng.normal(0.0, core_sigma, n)
    b = rng.normal(0.0, core_sigma, n)
    rr = radius + a
    out = np.zeros((n, 7))
    out[:, 0] = rr * np.cos(phi)
    out[:, 1] = rr * np.sin(phi)
    out[:, 2] = b
    g = 2.0 * math.pi * radius / n
    t = np.stack([-np.sin(phi), np.cos(phi), np.zeros(n)], axis=1)
    out[:, 4:7] = g * (t + noise * rng.normal(0.0, 1.0, (n, 3)))
    return out


def plummer_sources(n, seed=3, scale=1.0, rmax=20.0):
    """Plummer sphere: r = a / sqrt(u^(-2/3) - 1) (rejected beyond rmax), isotropic direction, q ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    r = np.empty(0)
    while r.size < n:
        u = rng.random(2 * n)
        rs = scale / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = np.concatenate([r, rs[rs <= rmax]])
    r = r[:n]
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    b = np.zeros((n, 5))
    b[:, :3] = v * r[:, None]
    b[:, 4] = rng.normal(size=n)
    return b


def relative_error(v, v_ref):
    """Per-target ||v - v_ref|| / ||v_ref|| (guarded by max ||v_ref||); returns (max, p99, median)."""
    num = np.linalg.norm(v - v_ref, axis=1)
    den = np.linalg.norm(v_ref, axis=1)
    den = np.maximum(den, 1e-300 * max(den.max(), 1e-300))
    e = num / den
    return float(e.max()), float(np.percentile(e, 99)), float(np.median(e))


def uniform_dipoles(n, seed=1):
    """Point dipoles: positions U[0,1]^3, moments N(0,1)/n per component. (n, 7) buffer [x, y, z, 0, q]."""
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 7))
    b[:, :3] = rng.random((n, 3))
    b[:, 4:7] = rng.normal(size=(n, 3)) / n
    return b


def gaussian_vortex_ring(n, seed=2, radius=1.0, core_sigma=0.1, noise=1e-3, sigma=None):
    """vortex_ring() with every particle given a Gaussian core of width `sigma` (buffer row 4);
    the default sigma is 1.5x the mean particle spacing along the ring's core."""
    b = vortex_ring(n, seed=seed, radius=radius, core_sigma=core_sigma, noise=noise)
    if sigma is None:
        vol = 2.0 * math.pi * radius * 2.0 * math.pi * core_sigma ** 2
        sigma = 1.5 * (vol / n) ** (1.0 / 3.0)
    b[:, 3] = sigma
    return b


def panel_rows(vertices, strength):
    """Panel buffer rows [centroid(3), radius, strength, v1(3), v2(3), v3(3), (v4(3))] from an
    (n, nv, 3) vertex array (counter-clockwise about the normal) and (n,) strengths. The
    centroid is the vertex mean; the radius is the largest centroid-vertex distance."""
    v = np.asarray(vertices, dtype=np.float64)
    cen = v.mean(axis=1)
    rad = np.linalg.norm(v - cen[:, None, :], axis=2).max(axis=1)
    return np.column_stack([cen, rad, np.asarray(strength, dtype=np.float64), v.reshape(len(v), -1)])


def sphere_source_dipole_panels(m, quads=True, seed=5, radius=1.0):
    """sphere_panels with a dipole strength N(0,1)/n beside each source strength (Panel{SourceDipole} rows
    [centroid(3), radius, source, dipole, vertices]). Returns (rows, kind)."""
    from . import _lib as L
    rows, _ = sphere_panels(m, quads=quads, seed=seed, radius=radius)
    dip = np.random.default_rng(seed + 1000).normal(size=len(rows)) / len(rows)
    kind = L.PANEL_SOURCE_DIPOLE_QUAD if quads else L.PANEL_SOURCE_DIPOLE_TRI
    return np.ascontiguousarray(np.insert(rows, 5, dip, axis=1)), kind


def sphere_panels(m, quads=True, seed=5, radius=1.0):
    """Panels tiling a sphere: each cube face split into an m x m grid projected onto the sphere.
    Quadrilaterals are made planar (the 4th vertex is projected onto the plane of the first three);
    with quads=False each quadrilateral is split into two triangles. Strengths N(0,1)/n_panels.
    Returns (rows, kind)."""
    rng = np.random.default_rng(seed)
    t = np.linspace(-1.0, 1.0, m + 1)
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    corners = [(ii, jj), (ii + 1, jj), (ii + 1, jj + 1), (ii, jj + 1)]
    blocks = []
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            a, b = [k for k in range(3) if k != axis]
            quad = np.zeros((m * m, 4, 3))
            for c, (ci, cj) in enumerate(corners):
                quad[:, c, axis] = sgn
                quad[:, c, a] = t[ci].ravel()
                quad[:, c, b] = t[cj].ravel()
            blocks.append(quad)
    faces = np.concatenate(blocks)
    faces = radius * faces / np.linalg.norm(faces, axis=2, keepdims=True)
    n = np.cross(faces[:, 1] - faces[:, 0], faces[:, 2] - faces[:, 0])
    flip = np.einsum("ij,ij->i", n, faces.mean(axis=1)) < 0      # counter-clockwise about the outward normal
    faces[flip] = faces[flip, ::-1]
    # the plane of the (reordered) first three vertices, onto which the fourth is projected
    n = np.cross(faces[:, 1] - faces[:, 0], faces[:, 2] - faces[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    faces[:, 3] -= np.einsum("ij,ij->i", faces[:, 3] - faces[:, 0], n)[:, None] * n
    if quads:
        q = rng.normal(size=len(faces)) / len(faces)
        return panel_rows(faces, q), 6
    tris = np.concatenate([faces[:, [0, 1, 2]], faces[:, [0, 2, 3]]], axis=0)
    q = rng.normal(size=len(tris)) / len(tris)
    return panel_rows(tris, q), 4
