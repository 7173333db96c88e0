"""Seeded synthetic point clouds with analytic normals for the reconstruction workloads.

Every generator returns float32 (N, 3) positions and unit normals, sampled uniformly by surface
area (rejection on the area element), from numpy.random.default_rng(seed).
"""
import math

import numpy as np


def torus(n, R=0.35, r=0.15, center=(0.5, 0.5, 0.5), seed=0):
    """Torus around the z axis; area element (R + r cos v) du dv."""
    rng = np.random.default_rng(seed)
    us, vs = [], []
    need = n
    while need > 0:
        m = int(need * 1.6) + 16
        u = rng.uniform(0, 2 * np.pi, m)
        v = rng.uniform(0, 2 * np.pi, m)
        keep = rng.uniform(0, 1, m) < (R + r * np.cos(v)) / (R + r)
        us.append(u[keep])
        vs.append(v[keep])
        need -= int(keep.sum())
    u = np.concatenate(us)[:n]
    v = np.concatenate(vs)[:n]
    nrm = np.stack([np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v)], 1)
    ring = np.stack([R * np.cos(u), R * np.sin(u), np.zeros_like(u)], 1)
    xyz = ring + r * nrm + np.asarray(center)
    return xyz.astype(np.float32), nrm.astype(np.float32)


def _ellipsoid_surface(rng, n, axes):
    a, b, c = axes
    gmax = max(b * c, a * c, a * b)
    out = []
    need = n
    while need > 0:
        m = int(need * 2.0) + 16
        d = rng.standard_normal((m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        g = np.sqrt((b * c * d[:, 0]) ** 2 + (a * c * d[:, 1]) ** 2 + (a * b * d[:, 2]) ** 2)
        keep = rng.uniform(0, 1, m) < g / gmax
        out.append(d[keep])
        need -= int(keep.sum())
    d = np.concatenate(out)[:n]
    p = d * np.asarray(axes)
    nrm = p / np.asarray(axes) ** 2
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return p, nrm


def ellipsoid_area(axes, p=1.6075):
    """Knud Thomsen's approximation of an ellipsoid's surface area."""
    a, b, c = axes
    return 4 * np.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def random_rotation(rng):
    """Haar-random rotation (QR of a Gaussian matrix with sign fix)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def vmf_perturb(rng, normals, kappa):
    """Von Mises-Fisher noise around each unit normal (Wood 1994, exact inversion for S^2)."""
    n = normals.shape[0]
    u = rng.uniform(0, 1, n)
    w = 1 + np.log(u + (1 - u) * np.exp(-2 * kappa)) / kappa
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(np.maximum(0, 1 - w * w))
    # orthonormal frame around each normal
    a = np.where(np.abs(normals[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    t1 = np.cross(normals, a)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(normals, t1)
    out = w[:, None] * normals + s[:, None] * (np.cos(phi)[:, None] * t1 + np.sin(phi)[:, None] * t2)
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def ellipsoids(n, k=16, axis_range=(0.05, 0.25), center_range=(0.2, 0.8), sigma=0.0, kappa=None, seed=0,
               return_params=False):
    """Union of k random ellipsoids, points allocated by (approximate) area. With return_params, also
    (semi-axes (k, 3), rotations (k, 3, 3), centres (k, 3))."""
    rng = np.random.default_rng(seed)
    axes = rng.uniform(*axis_range, (k, 3))
    rots = [random_rotation(rng) for _ in range(k)]
    cents = rng.uniform(*center_range, (k, 3))
    area = np.array([ellipsoid_area(a) for a in axes])
    counts = np.floor(n * area / area.sum()).astype(np.int64)
    counts[: n - counts.sum()] += 1
    P, N = [], []
    for i in range(k):
        p, nm = _ellipsoid_surface(rng, int(counts[i]), axes[i])
        P.append(p @ rots[i].T + cents[i])
        N.append(nm @ rots[i].T)
    xyz, nrm = np.concatenate(P), np.concatenate(N)
    if sigma > 0:
        xyz = xyz + rng.normal(0, sigma, xyz.shape)
    if kappa is not None:
        nrm = vmf_perturb(rng, nrm, kappa)
    if return_params:
        return xyz.astype(np.float32), nrm.astype(np.float32), (axes, np.stack(rots), cents)
    return xyz.astype(np.float32), nrm.astype(np.float32)


def ellipsoids_sdf(p, params):
    """Signed distance (approximate: the normalised radial distance times the smallest semi-axis,
    exact on spheres) to the union of the ellipsoids of ellipsoids(..., return_params=True)."""
    axes, rots, cents = params
    d = np.full(p.shape[0], np.inf)
    for a, R, c in zip(axes, rots, cents):
        local = (np.asarray(p, np.float64) - c) @ R  # rotated into the ellipsoid's frame
        d = np.minimum(d, (np.linalg.norm(local / a, axis=1) - 1.0) * a.min())
    return d


def torus_sdf(p, R=0.35, r=0.15, center=(0.5, 0.5, 0.5)):
    q = np.asarray(p, np.float64) - np.asarray(center)
    return np.sqrt((np.sqrt(q[:, 0] ** 2 + q[:, 1] ** 2) - R) ** 2 + q[:, 2] ** 2) - r


def room(n, size=(8.0, 6.0, 3.0), n_boxes=40, box_range=(0.2, 1.5), sigma=0.0, seed=0, scale=1 / 8.0):
    """Axis-aligned room (floor, ceiling, 4 walls; normals point into the room) with boxes resting
    on the floor (outward normals, no bottom face), scaled by `scale` and centred in the unit cube."""
    rng = np.random.default_rng(seed)
    size = np.asarray(size, np.float64)
    boxes = []
    for _ in range(n_boxes):
        s = rng.uniform(*box_range, 3)
        s[2] = min(s[2], size[2] * 0.9)
        lo = np.array([rng.uniform(0, size[0] - s[0]), rng.uniform(0, size[1] - s[1]), 0.0])
        boxes.append((lo, lo + s))
    faces = []  # (area, axis, plane, lo, hi, normal)
    for bi, (lo, hi) in enumerate([(np.zeros(3), size)] + boxes):
        for ax in range(3):
            o1, o2 = [a for a in range(3) if a != ax]
            for plane, sgn in ((lo[ax], -1.0), (hi[ax], 1.0)):
                if bi > 0 and ax == 2 and plane == 0.0:
                    continue
                nv = np.zeros(3)
                nv[ax] = -sgn if bi == 0 else sgn
                faces.append(((hi[o1] - lo[o1]) * (hi[o2] - lo[o2]), ax, plane, lo, hi, nv))
    areas = np.array([f[0] for f in faces])
    which = rng.choice(len(faces), size=n, p=areas / areas.sum())
    xyz = np.zeros((n, 3))
    nrm = np.zeros((n, 3))
    for i, (_, ax, plane, lo, hi, nv) in enumerate(faces):
        sel = np.nonzero(which == i)[0]
        o1, o2 = [a for a in range(3) if a != ax]
        xyz[sel, ax] = plane
        xyz[sel, o1] = rng.uniform(lo[o1], hi[o1], sel.shape[0])
        xyz[sel, o2] = rng.uniform(lo[o2], hi[o2], sel.shape[0])
        nrm[sel] = nv
    if sigma > 0:
        xyz = xyz + rng.normal(0, sigma, xyz.shape)
    xyz = xyz * scale + (1 - size * scale) / 2
    return xyz.astype(np.float32), nrm.astype(np.float32)


def sphere(n, radius=0.3, center=(0.5, 0.5, 0.5), seed=0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * radius + np.asarray(center)).astype(np.float32), d.astype(np.float32)


def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
