"""Seeded synthetic vorton / tracer distributions for the velocity path.

Each generator returns float32 arrays: positions (N,3), angular velocity
Omega = vorticity / 2 (N,3) (Vorton stores half the vorticity,
vorton.h:316-332) and the vorton radius.  The reference builds its scenes
from grid-cell-centred vortons (vorticityDistribution.cpp:90-190); these
generators instead draw positions from seeded PCG64 streams so a workload of
any size is reproducible from (name, N, seed).
"""
from __future__ import annotations

import numpy as np


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _radius_for(volume: float, n: int) -> float:
    # vortons spaced like cell centres of cells of volume V/N, with the
    # reference's radius rule radius = cbrt(V_cell) * 32^(-1/6)
    # (vorticityDistribution.cpp:126)
    return float(np.cbrt(volume / n) * 32.0 ** (-1.0 / 6.0))


def _ring(rng, n, R, r, center, axis, strength_sigma=0.05, sign=1.0):
    theta = rng.uniform(0.0, 2.0 * np.pi, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    rho = r * np.sqrt(rng.uniform(0.0, 1.0, n))
    axis = _unit(np.asarray(axis, np.float64))
    # orthonormal frame (e1, e2, axis)
    t = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = _unit(np.cross(axis, t))
    e2 = np.cross(axis, e1)
    radial = np.cos(theta)[:, None] * e1 + np.sin(theta)[:, None] * e2
    tangent = -np.sin(theta)[:, None] * e1 + np.cos(theta)[:, None] * e2
    pos = (center + (R + rho * np.cos(phi))[:, None] * radial
           + (rho * np.sin(phi))[:, None] * axis)
    mag = 1.0 + rng.normal(0.0, strength_sigma, n)
    vort = sign * mag[:, None] * tangent
    return pos, vort


def vortex_ring(n=4096, seed=1, R=1.0, r=0.25, radius=0.05):
    """Single ring: torus R, tube r, vorticity tangent to the ring with
    |w| = 1 + N(0, 0.05); vorton radius 0.05 (size 0.1)."""
    rng = np.random.default_rng(seed)
    pos, vort = _ring(rng, n, R, r, np.zeros(3), [0.0, 0.0, 1.0])
    return pos.astype(np.float32), (0.5 * vort).astype(np.float32), float(radius)


def colliding_rings(n=65536, seed=2, R=1.0, separation=1.5, r=0.2):
    """Two coaxial rings, opposite circulation, axial separation 1.5."""
    rng = np.random.default_rng(seed)
    h = n // 2
    p1, w1 = _ring(rng, h, R, r, np.array([0.0, 0.0, -0.5 * separation]), [0, 0, 1], sign=1.0)
    p2, w2 = _ring(rng, n - h, R, r, np.array([0.0, 0.0, 0.5 * separation]), [0, 0, 1], sign=-1.0)
    pos = np.concatenate([p1, p2])
    vort = np.concatenate([w1, w2])
    volume = 2.0 * (2.0 * np.pi ** 2 * R * r * r)
    return pos.astype(np.float32), (0.5 * vort).astype(np.float32), _radius_for(volume, n)


def jet_plume(n=262144, seed=3, sigma=0.2, length=4.0):
    """Buoyant jet plume: Gaussian radial profile, axial length 4,
    azimuthal vorticity with |w| ~ lognormal(0, 0.5)."""
    rng = np.random.default_rng(seed)
    xy = rng.normal(0.0, sigma, (n, 2))
    z = rng.uniform(0.0, length, n)
    pos = np.column_stack([xy, z])
    rho = np.linalg.norm(xy, axis=1) + 1e-12
    az = np.column_stack([-xy[:, 1] / rho, xy[:, 0] / rho, np.zeros(n)])
    mag = rng.lognormal(0.0, 0.5, n)
    vort = mag[:, None] * az
    volume = np.pi * (2.0 * sigma) ** 2 * length
    return pos.astype(np.float32), (0.5 * vort).astype(np.float32), _radius_for(volume, n)


def random_rings(n=1 << 20, seed=5, rings=64, box=1.5):
    """Turbulent mix of random rings: R ~ U[0.3, 1.2], random axes and
    centres in [-box, box]^3, tube radius 0.2 R."""
    rng = np.random.default_rng(seed)
    counts = np.full(rings, n // rings)
    counts[: n - counts.sum()] += 1
    ps, ws = [], []
    volume = 0.0
    for k in range(rings):
        R = rng.uniform(0.3, 1.2)
        r = 0.2 * R
        axis = _unit(rng.normal(size=3))
        center = rng.uniform(-box, box, 3)
        p, w = _ring(rng, int(counts[k]), R, r, center, axis)
        ps.append(p)
        ws.append(w)
        volume += 2.0 * np.pi ** 2 * R * r * r
    pos = np.concatenate(ps)
    vort = np.concatenate(ws)
    return pos.astype(np.float32), (0.5 * vort).astype(np.float32), _radius_for(volume, n)


def tracers_in_box(n, mn, mx, seed=7):
    rng = np.random.default_rng(seed)
    mn = np.asarray(mn, np.float64)
    mx = np.asarray(mx, np.float64)
    return (mn + (mx - mn) * rng.uniform(0.0, 1.0, (n, 3))).astype(np.float32)


def tracers_emitted(n, mn, mx, per_cell=2, seed=7):
    """n tracers as PclOpEmit::Emit places them (pclOpEmit.cpp:77-172): the
    box is cut into cells (about cubic, enough for n), each cell holds
    per_cell^3 tracers, one per sub-cell at the sub-cell centre plus a
    uniform jitter of up to half a sub-cell, emitted cell by cell in z, y, x
    order and inside a cell sub-cell by sub-cell in z, y, x order; the first
    n are kept.  Memory order follows space, as in the reference's tracer
    arrays."""
    rng = np.random.default_rng(seed)
    mn = np.asarray(mn, np.float64)
    mx = np.asarray(mx, np.float64)
    ext = np.maximum(mx - mn, 1e-30)
    cells = n / per_cell ** 3
    side = (cells / np.prod(ext)) ** (1.0 / 3.0)
    nc = np.maximum(1, np.ceil(ext * side)).astype(np.int64)
    while np.prod(nc) * per_cell ** 3 < n:
        nc[np.argmin(nc / ext)] += 1
    h = ext / nc
    sub = h / per_cell
    cz, cy, cx, sz, sy, sx = np.meshgrid(np.arange(nc[2]), np.arange(nc[1]), np.arange(nc[0]),
                                         np.arange(per_cell), np.arange(per_cell),
                                         np.arange(per_cell), indexing="ij")
    cell = np.stack([cx, cy, cz], -1).reshape(-1, 3)[:n]
    sc = np.stack([sx, sy, sz], -1).reshape(-1, 3)[:n]
    jitter = rng.uniform(-0.5, 0.5, (n, 3)) * 0.9999
    pos = mn + cell * h + (sc + 0.5 + jitter) * sub
    return pos.astype(np.float32)


SCENARIOS = {
    "vortex_ring": vortex_ring,
    "colliding_rings": colliding_rings,
    "jet_plume": jet_plume,
    "random_rings": random_rings,
}
