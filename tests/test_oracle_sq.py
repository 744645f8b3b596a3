"""Pins the oracle's superquadric contact-point solve (oracle/sq_oracle.c)
against what the geometry fixes independently:

  shape function      zero on the parametric superquadric surface; the
                      global form equals the body-frame form under the
                      rotation of an independent quaternion convention
  derivatives         gradient and Hessian equal central differences of the
                      shape function
  spheres             the contact point is the radius-weighted point on the
                      line of centres, found without a Newton iteration
  general shapes      the solution is the minimiser of max(F_A, F_B) (both
                      functions are convex): an independent constrained
                      optimiser finds the same point; gradients antiparallel
                      and F_A = F_B there
  fresh contacts      the homotopy from spheres takes at least its 11
                      stages; a start at the solution returns at once
"""
import math

import numpy as np
import pytest
from scipy.optimize import minimize
from scipy.spatial.transform import Rotation

import pyoracle as orc
import superquadrics as sq


def pose(center, quat, shape, blk):
    return np.concatenate([center, quat, shape, blk]).astype(np.float64)


def sgnpow(t, e):
    return np.sign(t) * np.abs(t) ** e


def test_shape_function_zero_on_parametric_surface():
    rng = np.random.default_rng(0)
    for n1, n2 in [(2.0, 2.0), (4.0, 4.0), (3.0, 6.0), (5.5, 2.7), (8.0, 3.0)]:
        shape = rng.uniform(0.001, 0.003, 3)
        p = pose(np.zeros(3), [1, 0, 0, 0], shape, [n1, n2])
        for _ in range(40):
            eta = rng.uniform(-math.pi / 2, math.pi / 2)
            om = rng.uniform(-math.pi, math.pi)
            x = np.array([shape[0] * sgnpow(math.cos(eta), 2 / n1) * sgnpow(math.cos(om), 2 / n2),
                          shape[1] * sgnpow(math.cos(eta), 2 / n1) * sgnpow(math.sin(om), 2 / n2),
                          shape[2] * sgnpow(math.sin(eta), 2 / n1)])
            f, _, _ = orc.sq_props(p, x)
            assert abs(f) < 1e-12, (n1, n2, f)
        # inside / outside
        assert orc.sq_props(p, np.zeros(3))[0] < 0
        assert orc.sq_props(p, 2 * shape)[0] > 0


def test_global_form_and_quaternion_convention():
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = sq.random_quats(rng, 1)[0]
        c = rng.uniform(-0.01, 0.01, 3)
        shape = rng.uniform(0.001, 0.003, 3)
        blk = rng.uniform(2, 8, 2)
        xl = rng.uniform(-0.003, 0.003, 3)
        R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()  # scalar-last there
        fg, gg, hg = orc.sq_props(pose(c, q, shape, blk), c + R @ xl)
        fl, gl, hl = orc.sq_props(pose(np.zeros(3), [1, 0, 0, 0], shape, blk), xl)
        assert fg == pytest.approx(fl, rel=1e-12, abs=1e-14)
        assert np.allclose(gg, R @ gl, rtol=1e-10, atol=1e-10 * np.abs(gl).max())
        assert np.allclose(hg, R @ hl @ R.T, rtol=1e-9, atol=1e-9 * np.abs(hl).max())


@pytest.mark.parametrize("blk", [(2.0, 2.0), (4.0, 4.0), (6.0, 3.0), (3.3, 5.7)])
def test_gradient_and_hessian_are_derivatives(blk):
    rng = np.random.default_rng(2)
    shape = np.array([0.0012, 0.0025, 0.0018])
    p = pose([0.001, -0.002, 0.0005], sq.random_quats(rng, 1)[0], shape, blk)
    for _ in range(20):
        x = p[:3] + rng.uniform(-0.003, 0.003, 3)
        f, g, H = orc.sq_props(p, x)
        h = 1e-7
        gn = np.zeros(3)
        Hn = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fp, gp, _ = orc.sq_props(p, x + e)
            fm, gm, _ = orc.sq_props(p, x - e)
            gn[k] = (fp - fm) / (2 * h)
            Hn[:, k] = (gp - gm) / (2 * h)
        assert np.allclose(g, gn, rtol=1e-5, atol=1e-6 * np.abs(g).max())
        assert np.allclose(H, Hn, rtol=1e-4, atol=1e-5 * np.abs(H).max())


def test_spheres_contact_on_line_of_centres():
    rng = np.random.default_rng(3)
    for _ in range(30):
        ri, rj = rng.uniform(0.001, 0.003, 2)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        A = pose(np.zeros(3), sq.random_quats(rng, 1)[0], [ri] * 3, [2, 2])
        B = pose(u * (ri + rj) * 0.95, sq.random_quats(rng, 1)[0], [rj] * 3, [2, 2])
        pt, fi, fj, it, fl = orc.sq_contact_batch(A[None], B[None])
        expect = B[:3] * (ri / (ri + rj))
        assert np.allclose(pt[0], expect, rtol=0, atol=1e-15)
        assert fi[0] < 0 and fj[0] < 0 and fl[0] == 0
        assert it[0] == 0  # every homotopy stage starts at the solution


def _minimax(Ap, Bp, x0):
    """Independent solve: min t s.t. F_A(x) <= t, F_B(x) <= t."""
    scale = 1e3  # metres -> millimetres for the optimiser

    def FA(z):
        return orc.sq_props(Ap, z[:3] / scale, False)[0]

    def FB(z):
        return orc.sq_props(Bp, z[:3] / scale, False)[0]

    cons = [{"type": "ineq", "fun": lambda z: z[3] - FA(z),
             "jac": lambda z: np.r_[-orc.sq_props(Ap, z[:3] / scale, False)[1] / scale, 1.0]},
            {"type": "ineq", "fun": lambda z: z[3] - FB(z),
             "jac": lambda z: np.r_[-orc.sq_props(Bp, z[:3] / scale, False)[1] / scale, 1.0]}]
    z0 = np.r_[x0 * scale, max(FA(np.r_[x0 * scale, 0]), FB(np.r_[x0 * scale, 0]))]
    res = minimize(lambda z: z[3], z0, jac=lambda z: np.r_[0, 0, 0, 1.0], constraints=cons,
                   method="SLSQP", options={"ftol": 1e-15, "maxiter": 500})
    return res.x[:3] / scale


def test_contact_point_is_minimax_of_shape_functions():
    A, B = sq.overlapping_pairs(5, 60)
    pt, fi, fj, it, fl = orc.sq_contact_batch(A, B)
    n_checked = 0
    for k in range(len(A)):
        if fl[k]:
            continue
        fa, ga, _ = orc.sq_props(A[k], pt[k], False)
        fb, gb, _ = orc.sq_props(B[k], pt[k], False)
        sin2 = np.sum(np.cross(ga, gb) ** 2) / ((ga @ ga) * (gb @ gb))
        if not (ga @ gb < 0 and sin2 < 1e-10 and abs(fa - fb) <= 1e-5 * max(abs(fa), abs(fb))):
            continue  # the reference's relative-change stop left this one short
        n_checked += 1
        x = _minimax(A[k], B[k], 0.5 * (A[k, :3] + B[k, :3]))
        size = min(A[k, 7:10].min(), B[k, 7:10].min())
        assert np.linalg.norm(x - pt[k]) < 1e-4 * size, k
        assert max(fa, fb) < 0  # overlapping pairs by construction
    assert n_checked >= 50


def test_fresh_contacts_take_the_homotopy_and_warm_starts_return():
    A, B = sq.overlapping_pairs(6, 200)
    pt, fi, fj, it, fl = orc.sq_contact_batch(A, B)
    assert fl.sum() == 0
    assert it.min() >= 11 - 1  # 11 stages, the first one (spheres) needs none
    assert it.mean() > 15
    pt2, _, _, it2, fl2 = orc.sq_contact_batch(A, B, pt)
    assert (it2 <= 1).mean() > 0.95
    assert np.abs(pt2 - pt).max() < 1e-3
