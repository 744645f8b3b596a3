"""Pins of the oracle's superquadric contact detection (checkSurfaceIntersect,
surface_model_superquadric.h:108-205) against what geometry fixes:

  obb_intersect          a separating-axis test written here from scratch over
                         the 15 candidate axes (rotations from scipy), and the
                         "last separating axis first" bookkeeping
  surface_line_intersection
                         ellipsoids: the quadratic's roots (numpy); blocky and
                         non-integer shapes: bisection on an independent shape
                         function along the ray
  overlap                two spheres: ri + rj - d, normal from j to i, surface
                         points on the line of centres; two equal aligned
                         superquadrics of any blockiness apart along x:
                         2a - d
  state machine          far / close / intersecting, contact point carried in
                         the history, a rigidly translated pair's next point
                         predicted exactly from its velocity
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as orc
import superquadrics as sq


def pose(center, quat, shape, blk):
    return np.array([*center, *quat, *shape, *blk], dtype=np.float64)


def rotation(q):
    # scipy takes (x, y, z, w)
    return Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()


def sat_separating(a, b, koef=1.0001):
    """Indices (in the reference's order: A's axes, B's axes, A_i x B_j) of
    the axes that separate the two boxes of half-axes grown by koef."""
    Ra, Rb = rotation(a[3:7]), rotation(b[3:7])
    ha, hb = a[7:10] * koef, b[7:10] * koef
    d = a[0:3] - b[0:3]
    axes = [Ra[:, 0], Ra[:, 1], Ra[:, 2], Rb[:, 0], Rb[:, 1], Rb[:, 2]]
    axes += [np.cross(Ra[:, i], Rb[:, j]) for i in range(3) for j in range(3)]
    out, margin = [], np.inf
    for k, L in enumerate(axes):
        r = abs(L @ d)
        r0 = sum(ha[m] * abs(L @ Ra[:, m]) for m in range(3))
        r1 = sum(hb[m] * abs(L @ Rb[:, m]) for m in range(3))
        if abs(r0 + r1) < 1e-14 and r < 1e-14:
            continue
        margin = min(margin, abs(r - (r0 + r1)) / max(r0 + r1, 1e-300))
        if r > r0 + r1:
            out.append(k)
    return out, margin


def shape_f(p, x):
    """Superquadric inside-outside function (no scaling) at global x. As in
    the reference, blockiness within 1e-2 of whole numbers (both of them) is
    used as those whole numbers in the powers of the half-axis terms."""
    R = rotation(p[3:7])
    loc = R.T @ (x - p[0:3])
    a, b, c = p[7:10]
    n1, n2 = p[10:12]
    e1, e2 = n1, n2
    near = [abs(n - np.floor(n + 1e-2)) < 1e-2 for n in (n1, n2)]
    if all(near):
        e1, e2 = np.floor(n1 + 1e-2), np.floor(n2 + 1e-2)
    xy = np.abs(loc[0] / a) ** e2 + np.abs(loc[1] / b) ** e2
    if not (all(near) and e1 == e2):
        xy = xy ** (n1 / n2)
    return xy + np.abs(loc[2] / c) ** e1 - 1.0


def test_obb_matches_separating_axis_test():
    rng = np.random.default_rng(11)
    q = sq.random_quats(rng, 4000)
    agree = skipped = hits = 0
    for k in range(2000):
        sa = rng.uniform(0.5, 2.0, 3)
        sb = rng.uniform(0.5, 2.0, 3)
        d = rng.normal(size=3)
        d *= rng.uniform(0.5, 5.0) / np.linalg.norm(d)
        a = pose([0, 0, 0], q[2 * k], sa, [2, 2])
        b = pose(d, q[2 * k + 1], sb, [2, 2])
        sep, margin = sat_separating(a, b)
        if margin < 1e-9:
            skipped += 1
            continue
        hit, start = orc.sq_obb(a, b, 0)
        assert hit == (len(sep) == 0), k
        if not hit:
            assert start in sep
            # an axis that separates is tried first and kept
            s0 = sep[-1]
            hit2, start2 = orc.sq_obb(a, b, s0)
            assert not hit2 and start2 == s0
        else:
            assert start == 0
            hits += 1
        agree += 1
    assert agree > 1900 and skipped < 50
    # both outcomes are well represented
    assert 0.2 * agree < hits < 0.8 * agree


def test_obb_axis_aligned_boxes_are_interval_overlap():
    ident = [1, 0, 0, 0]
    a = pose([0, 0, 0], ident, [1.0, 2.0, 3.0], [2, 2])
    rng = np.random.default_rng(3)
    for _ in range(500):
        c = rng.uniform(-5, 5, 3)
        s = rng.uniform(0.2, 2.0, 3)
        b = pose(c, ident, s, [2, 2])
        gap = np.abs(c) - (a[7:10] + s) * 1.0001
        if np.min(np.abs(gap)) < 1e-9:
            continue
        hit, _ = orc.sq_obb(a, b, 0)
        assert hit == bool(np.all(gap < 0))


@pytest.mark.parametrize("inside", [True, False])
def test_line_intersection_ellipsoid_roots(inside):
    rng = np.random.default_rng(5 + inside)
    q = sq.random_quats(rng, 300)
    for k in range(300):
        shape = rng.uniform(0.5, 2.0, 3)
        p = pose(rng.normal(size=3), q[k], shape, [2, 2])
        R = rotation(q[k])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        if inside:
            start = p[0:3] + R @ (rng.uniform(-0.5, 0.5, 3) * shape)
            direction = rng.normal(size=3) * rng.uniform(0.5, 2.0)
        else:
            # outside on one side, pointing through the centre region
            target = p[0:3] + R @ (rng.uniform(-0.3, 0.3, 3) * shape)
            start = target + u * 4.0 * shape.max()
            direction = (target - start) * rng.uniform(0.1, 1.0)
        # roots of |S^-1 R^T (start + t dir - c)|^2 = 1
        s0 = (R.T @ (start - p[0:3])) / shape
        dd = (R.T @ direction) / shape
        roots = np.sort(np.roots([dd @ dd, 2 * s0 @ dd, s0 @ s0 - 1]).real)
        if inside:
            want = roots[1]  # the crossing ahead
        else:
            want = roots[0]  # the nearer crossing ahead
        alpha, res = orc.sq_line(p, start, direction)
        assert alpha == pytest.approx(want, rel=1e-10, abs=1e-12)
        assert np.allclose(res, start + alpha * direction, rtol=0, atol=1e-12)


@pytest.mark.parametrize("blk", [(4.0, 4.0), (6.0, 3.0), (3.5, 2.5), (8.0, 8.0), (2.0, 5.0)])
def test_line_intersection_blocky_bisection(blk):
    rng = np.random.default_rng(int(blk[0] * 10 + blk[1]))
    q = sq.random_quats(rng, 200)
    errs = []
    for k in range(200):
        shape = rng.uniform(1.0, 3.0, 3) * 1e-3
        p = pose(rng.normal(size=3) * 1e-2, q[k], shape, blk)
        R = rotation(q[k])
        start = p[0:3] + R @ (rng.uniform(-0.4, 0.4, 3) * shape)
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)

        def crossing(sign):
            lo, hi = 0.0, 2.0 * np.linalg.norm(shape)
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if shape_f(p, start + sign * mid * direction) < 0:
                    lo = mid
                else:
                    hi = mid
            return sign * lo

        ahead, behind = crossing(1.0), crossing(-1.0)
        alpha, res = orc.sq_line(p, start, direction)
        # Newton from alpha = 0 converges to one of the two crossings of the
        # line (the nearer one in the direction the shape function falls)
        errs.append(min(abs(alpha - ahead), abs(alpha - behind)))
        assert np.allclose(res, start + alpha * direction, rtol=0, atol=1e-15)
    errs = np.array(errs)
    assert (errs < 1e-12 * 1e-3).mean() >= 0.97, np.sort(errs)[-10:]


def _pair_state(A, B, hist=None, vel=None, mass=None, dt=0.0):
    return orc.sq_check_batch(A[None, :], B[None, :], hist, vel, mass, dt)


def test_two_spheres_overlap_closed_form():
    rng = np.random.default_rng(9)
    q = sq.random_quats(rng, 200)
    for k in range(100):
        ri, rj = rng.uniform(0.5, 2.0, 2) * 1e-3
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = (ri + rj) * rng.uniform(0.85, 0.999)
        xi = rng.normal(size=3) * 1e-2
        xj = xi + d * u
        A = pose(xi, q[2 * k], [ri] * 3, [2, 2])
        B = pose(xj, q[2 * k + 1], [rj] * 3, [2, 2])
        hist, contact, en, dn, cpa, cpb, it = _pair_state(A, B)
        assert contact[0] == 1 and hist[0, 0] == orc.SQ_INTERSECT
        assert dn[0] == pytest.approx(ri + rj - d, rel=1e-9, abs=1e-15)
        assert np.allclose(en[0], -u, atol=1e-9)  # from j to i
        assert np.allclose(cpa[0], xi + ri * u, atol=1e-12)
        assert np.allclose(cpb[0], xj - rj * u, atol=1e-12)
        assert np.allclose(hist[0, 5:7], [ri - (hist[0, 2:5] - xi) @ u, rj - (xj - hist[0, 2:5]) @ u], atol=1e-12)


@pytest.mark.parametrize("blk", [(2.0, 2.0), (4.0, 4.0), (6.0, 3.0), (3.5, 2.5)])
def test_aligned_equal_superquadrics_overlap(blk):
    shape = np.array([2.0, 1.2, 1.5]) * 1e-3
    for frac in (0.999, 0.99, 0.95, 0.9):
        d = 2 * shape[0] * frac
        A = pose([0, 0, 0], [1, 0, 0, 0], shape, blk)
        B = pose([d, 0, 0], [1, 0, 0, 0], shape, blk)
        hist, contact, en, dn, cpa, cpb, it = _pair_state(A, B)
        assert contact[0] == 1
        assert dn[0] == pytest.approx(2 * shape[0] - d, rel=1e-8)
        assert np.allclose(en[0], [-1, 0, 0], atol=1e-8)
        assert np.allclose(hist[0, 2:5], [d / 2, 0, 0], atol=1e-12)
        assert cpa[0] == pytest.approx([shape[0], 0, 0], abs=1e-12)
        assert cpb[0] == pytest.approx([d - shape[0], 0, 0], abs=1e-12)


def test_random_overlaps_points_on_surfaces_along_normal():
    A, B = sq.overlapping_pairs(31, 300)
    hist, contact, en, dn, cpa, cpb, it = orc.sq_check_batch(A, B)
    c = contact == 1
    assert c.mean() > 0.9
    cp = hist[:, 2:5]
    for k in np.nonzero(c)[0]:
        assert abs(shape_f(A[k], cpa[k])) < 1e-8
        assert abs(shape_f(B[k], cpb[k])) < 1e-8
        assert np.linalg.norm(en[k]) == pytest.approx(1.0, abs=1e-12)
        # both surface points lie on the line through the contact point along en
        for p in (cpa[k], cpb[k]):
            v = p - cp[k]
            assert np.linalg.norm(np.cross(v, en[k])) <= 1e-9 * max(np.linalg.norm(v), 1e-12)
        # A's point lies towards B (against en), B's towards A (along en)
        assert (cpa[k] - cp[k]) @ en[k] <= 1e-15 and (cpb[k] - cp[k]) @ en[k] >= -1e-15
        assert dn[k] == pytest.approx(np.linalg.norm(cpb[k] - cpa[k]), rel=1e-9)
    assert np.all(hist[c, 0] == orc.SQ_INTERSECT)


def test_states_far_close_intersect():
    s = [1e-3] * 3
    ident = [1, 0, 0, 0]
    # boxes apart: far, no Newton iterations, the separating axis kept
    A = pose([0, 0, 0], ident, s, [2, 2])
    B = pose([0, 3e-3, 0], ident, s, [2, 2])
    hist, contact, *_, it = _pair_state(A, B)
    assert contact[0] == 0 and hist[0, 0] == orc.SQ_FAR and it[0] == 0 and hist[0, 1] == 1
    # boxes overlap on the diagonal, spheres do not: close, point stored
    B = pose([1.8e-3, 1.8e-3, 0], ident, s, [2, 2])
    hist, contact, *_, it = _pair_state(A, B)
    assert contact[0] == 0 and hist[0, 0] == orc.SQ_CLOSE
    assert np.allclose(hist[0, 2:5], [0.9e-3, 0.9e-3, 0], atol=1e-12)
    # from the stored point the next step's solve is immediate
    hist2, contact, *_, it2 = _pair_state(A, B, hist.copy())
    assert hist2[0, 0] == orc.SQ_CLOSE and it2[0] <= 1
    # moved into contact: intersecting, and the OBB test is skipped after that
    B = pose([1.3e-3, 1.3e-3, 0], ident, s, [2, 2])
    hist3, contact, *_ = _pair_state(A, B, hist2.copy())
    assert contact[0] == 1 and hist3[0, 0] == orc.SQ_INTERSECT
    hist3[0, 1] = 7.0
    hist4, contact, *_ = _pair_state(A, B, hist3.copy())
    assert contact[0] == 1 and hist4[0, 1] == 7.0


def test_translated_pair_point_predicted_from_velocity():
    A, B = sq.overlapping_pairs(33, 100)
    hist, contact, *_ = orc.sq_check_batch(A, B)
    rng = np.random.default_rng(4)
    v = rng.normal(size=(100, 3)) * 0.5
    dt = 1e-4
    A2, B2 = A.copy(), B.copy()
    A2[:, 0:3] += v * dt
    B2[:, 0:3] += v * dt
    vel = np.concatenate([v, v, np.zeros((100, 6))], axis=1)
    mass = rng.uniform(0.5, 2.0, (100, 2))
    h_moving, c1, _, dn1, _, _, it1 = orc.sq_check_batch(A2, B2, hist.copy(), vel, mass, dt)
    h_still, c0, _, dn0, _, _, it0 = orc.sq_check_batch(A2, B2, hist.copy())
    ok = contact == 1
    # the velocity estimate lands on the translated point: no Newton step
    # needed (a pair whose first solve stopped on the small-progress test
    # rather than the merit tolerance goes on from there)
    zero = ok & (it1 == 0)
    assert zero.sum() >= 0.9 * ok.sum()
    assert it0[ok].mean() > 3 * it1[ok].mean()
    assert np.allclose(h_moving[zero, 2:5], hist[zero, 2:5] + v[zero] * dt, rtol=0, atol=1e-15)
    # both solves stop at the merit tolerance, whose points differ by ~1e-10 m
    assert np.allclose(dn1[zero], dn0[zero], rtol=1e-5)


def test_mass_weighted_contact_velocity():
    """Only j moves; the estimate moves by m_j / (m_i + m_j) of j's step."""
    A, B = sq.overlapping_pairs(34, 50)
    hist, contact, *_ = orc.sq_check_batch(A, B)
    v = np.array([0.3, -0.2, 0.1])
    dt = 1e-4
    mass = np.tile([1.0, 3.0], (50, 1))
    vel = np.zeros((50, 12))
    vel[:, 3:6] = v
    est = hist[:, 2:5] + 0.75 * v * dt
    h1, *_, it1 = orc.sq_check_batch(A, B, hist.copy(), vel, mass, dt)
    p_given, _, _, it_given, _ = orc.sq_contact_batch(A, B, est)
    ok = contact == 1
    assert np.array_equal(it1[ok], it_given[ok])
    assert np.allclose(h1[ok, 2:5], p_given[ok], rtol=0, atol=0)
