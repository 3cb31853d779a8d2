"""Independent FP64 ray casting against a scene description, for checking the oracle's and the GPU's
Scene::Intersect / Occluded (scene.hxx:42-74, scene_graph.hxx:33-245).

Triangles by Moller-Trumbore (double-sided, as the reference's edge-sign test is), spheres by the
quadratic formula; both in float64 from the float32 scene data.  Besides the closest hit, each ray
gets a margin: how far (relative) it is from changing its answer, through a barycentric
coordinate, the discriminant of a sphere, the tmin / tmax bound, or a second primitive at nearly
the same distance.  Rays with a small margin may legitimately differ between two float32
implementations and are excluded from exact comparisons.
"""
import numpy as np


def scene_arrays(sb):
    d = sb.desc
    tris = np.ctypeslib.as_array(d.triangles, shape=(d.num_triangles, 9)).astype(np.float64).reshape(-1, 3, 3) \
        if d.num_triangles else np.zeros((0, 3, 3))
    tmat = np.ctypeslib.as_array(d.triangle_mat, shape=(d.num_triangles,)).copy() if d.num_triangles else np.zeros(0, int)
    sph = np.ctypeslib.as_array(d.spheres, shape=(d.num_spheres, 4)).astype(np.float64) if d.num_spheres \
        else np.zeros((0, 4))
    smat = np.ctypeslib.as_array(d.sphere_mat, shape=(d.num_spheres,)).copy() if d.num_spheres else np.zeros(0, int)
    return tris, tmat, sph, smat


def cast(sb, org, dir, tmin, tmax=None):
    """Closest hit in (tmin, tmax) of each ray: (t (inf: miss), material (-1), normal (n, 3), margin)."""
    tris, tmat, sph, smat = scene_arrays(sb)
    o = np.asarray(org, np.float64)
    d = np.asarray(dir, np.float64)
    n = len(o)
    tmin = np.broadcast_to(np.asarray(tmin, np.float64), (n,))
    tmax = np.full(n, np.inf) if tmax is None else np.broadcast_to(np.asarray(tmax, np.float64), (n,))
    cand_t, cand_m, cand_n, cand_margin, raw_t = [], [], [], [], []
    for k in range(len(tris)):
        p0, p1, p2 = tris[k]
        e1, e2 = p1 - p0, p2 - p0
        pv = np.cross(d, e2)
        det = pv @ e1
        inv = 1.0 / np.where(det == 0, np.nan, det)
        s = o - p0
        u = (s * pv).sum(1) * inv
        qv = np.cross(s, e1)
        v = (d * qv).sum(1) * inv
        t = (qv @ e2) * inv
        w = 1.0 - u - v
        inside = (u >= 0) & (v >= 0) & (w >= 0)
        # distance to changing the inside test, in barycentric units: inside, the smallest
        # coordinate; outside, the largest violation (every negative coordinate must reach 0)
        bmargin = np.where(inside, np.minimum(np.minimum(u, v), w),
                           np.maximum(np.maximum(-np.minimum(u, 0), -np.minimum(v, 0)), -np.minimum(w, 0)))
        nrm = np.cross(e1, e2)
        nrm = nrm / np.linalg.norm(nrm)
        valid = inside & np.isfinite(t) & (t > tmin) & (t < tmax)
        raw_t.append(np.where(inside & np.isfinite(t), t, np.inf))
        cand_t.append(np.where(valid, t, np.inf))
        cand_m.append(np.full(n, tmat[k]))
        cand_n.append(np.broadcast_to(nrm, (n, 3)))
        # rays whose hit/miss a float32 edge test may decide differently
        cand_margin.append(np.where(np.isfinite(t) & (t > 0), bmargin, np.inf))
    for k in range(len(sph)):
        c, r = sph[k, :3], sph[k, 3]
        oc = o - c
        A = (d * d).sum(1)
        B = 2 * (d * oc).sum(1)
        C = (oc * oc).sum(1) - r * r
        disc = B * B - 4 * A * C
        sq = np.sqrt(np.maximum(disc, 0))
        t0 = (-B - sq) / (2 * A)
        t1 = (-B + sq) / (2 * A)
        # the reference keeps the second root only when |q| = |B + sign(B) sqrt(disc)| / 2 > 0.1
        # and otherwise repeats the first, q / A (scene_graph.hxx:202-214)
        q = np.where(B >= 0, (-B - sq) / 2, (-B + sq) / 2)
        first = q / A
        with np.errstate(divide="ignore", invalid="ignore"):
            second = np.where(np.abs(q) > 0.1, C / q, first)
        t0, t1 = np.minimum(first, second), np.maximum(first, second)
        ok0 = (disc >= 0) & (t0 > tmin) & (t0 < tmax)
        ok1 = (disc >= 0) & (t1 > tmin) & (t1 < tmax)
        t = np.where(ok0, t0, np.where(ok1, t1, np.inf))
        raw_t.append(np.where(disc >= 0, t0, np.inf))
        raw_t.append(np.where(disc >= 0, t1, np.inf))
        p = oc + t[:, None] * d
        with np.errstate(invalid="ignore", divide="ignore"):
            nrm = p / np.linalg.norm(p, axis=1, keepdims=True)
        cand_t.append(t)
        cand_m.append(np.full(n, smat[k]))
        cand_n.append(nrm)
        cand_margin.append(np.minimum(np.abs(disc) / np.maximum(B * B + np.abs(4 * A * C), 1e-30),
                                      np.where(disc >= 0, np.abs(np.abs(q) - 0.1) / 0.1, np.inf)))
    T = np.stack(cand_t, 1)
    M = np.stack(cand_m, 1)
    N = np.stack(cand_n, 1)
    G = np.stack(cand_margin, 1)
    best = np.argmin(T, 1)
    rows = np.arange(n)
    t = T[rows, best]
    hit = np.isfinite(t)
    mat = np.where(hit, M[rows, best], -1)
    nrm = np.where(hit[:, None], N[rows, best], 0.0)
    margin = G.min(1)
    # a second primitive at nearly the same distance, or a hit close to tmin / tmax
    Ts = np.sort(T, 1)
    if T.shape[1] > 1:
        with np.errstate(invalid="ignore"):
            gap = np.where(np.isfinite(Ts[:, 1]), (Ts[:, 1] - Ts[:, 0]) / np.maximum(Ts[:, 1], 1e-30), np.inf)
        margin = np.minimum(margin, np.where(hit, gap, np.inf))
    # every candidate distance near the bounds (inside or just outside them): a float32 t may land
    # on the other side
    R = np.stack(raw_t, 1)
    with np.errstate(invalid="ignore"):
        for bound in (tmin, tmax):
            b = np.where(np.isfinite(bound), bound, 0.0)[:, None]
            near = np.abs(R - b) / np.maximum(np.abs(b), 1e-3)
            ok = np.isfinite(R) & np.isfinite(bound)[:, None]
            margin = np.minimum(margin, np.where(ok, near, np.inf).min(1))
    return np.where(hit, t, np.inf), mat, nrm, margin


def random_rays(rng, n, lo=0.02, hi=0.98):
    org = rng.uniform(lo, hi, (n, 3))
    dir = rng.normal(size=(n, 3))
    dir /= np.linalg.norm(dir, axis=1, keepdims=True)
    return org.astype(np.float32), dir.astype(np.float32)
