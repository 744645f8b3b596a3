"""Pins of the oracle's superquadric DEM step against what mathematics fixes:

  mass properties      ellipsoid / sphere closed forms; an n = 4 superquadric
                       against midpoint quadrature of its inside-outside test
  bounding radius      never below the farthest sampled surface point, and
                       within 1 % of it (n >= 2.5)
  pair contact         two equal spheres written as superquadrics give the
                       sphere Hertz / tangential-history chain's forces,
                       torques and shear (sphere chain: orc_contact_pair)
  plane contact        a sphere: 4/3 Y* sqrt(r d) d; a rotated ellipsoid: the
                       support-function depth and support point; an aligned
                       superquadric of any blockiness: c - z
  integrator           free flight under gravity exact; torque-free
                       asymmetric top keeps its global angular momentum and
                       its kinetic energy; a spherical top keeps omega and
                       turns by omega t (to the splitting error)
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as orc
import workloads as wl

MATS = wl.materials_single()


def _mats():
    return orc.materials(MATS)


def test_mass_props_ellipsoid_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(20):
        a, b, c = rng.uniform(0.5, 2.0, 3)
        rho = rng.uniform(500, 3000)
        V, I = orc.sq_mass_props([a, b, c], [2.0, 2.0], rho)
        assert V == pytest.approx(4.0 / 3.0 * np.pi * a * b * c, rel=1e-13)
        m = V * rho
        assert I == pytest.approx([m * (b * b + c * c) / 5, m * (a * a + c * c) / 5, m * (a * a + b * b) / 5],
                                  rel=1e-13)


def test_mass_props_blocky_quadrature():
    a, b, c, n1, n2 = 1.0, 1.5, 0.8, 4.0, 6.0
    m = 240
    xs = (np.arange(m) + 0.5) / m * 2 - 1
    X, Y, Z = np.meshgrid(xs * a, xs * b, xs * c, indexing="ij")
    inside = ((np.abs(X / a) ** n2 + np.abs(Y / b) ** n2) ** (n1 / n2) + np.abs(Z / c) ** n1) <= 1.0
    dV = (2 * a / m) * (2 * b / m) * (2 * c / m)
    Vq = inside.sum() * dV
    Iq = np.array([((Y ** 2 + Z ** 2) * inside).sum(), ((X ** 2 + Z ** 2) * inside).sum(),
                   ((X ** 2 + Y ** 2) * inside).sum()]) * dV
    V, I = orc.sq_mass_props([a, b, c], [n1, n2], 1.0)
    assert V == pytest.approx(Vq, rel=3e-3)
    assert I == pytest.approx(Iq, rel=5e-3)


def test_bounding_radius_covers_surface():
    rng = np.random.default_rng(2)
    eta = np.linspace(-np.pi / 2, np.pi / 2, 201)
    om = np.linspace(-np.pi, np.pi, 401)
    E, O = np.meshgrid(eta, om, indexing="ij")

    def spow(v, e):
        return np.sign(v) * np.abs(v) ** e

    for _ in range(40):
        shape = rng.uniform(1.0, 3.0, 3)
        blk = rng.uniform(2.5, 8.0, 2)
        x = shape[0] * spow(np.cos(E), 2 / blk[0]) * spow(np.cos(O), 2 / blk[1])
        y = shape[1] * spow(np.cos(E), 2 / blk[0]) * spow(np.sin(O), 2 / blk[1])
        z = shape[2] * spow(np.sin(E), 2 / blk[0])
        far = np.sqrt(x * x + y * y + z * z).max()
        rb = orc.sq_bounding_radius(shape, blk)
        assert rb >= far * (1 - 1e-9)
        assert rb <= far * 1.01
    # ellipsoids: the largest half-axis
    assert orc.sq_bounding_radius([1.0, 2.5, 1.5], [2.0, 2.0]) == pytest.approx(2.5 * 1.0001, rel=1e-15)


def _pose(center, quat, r):
    return np.array([*center, *quat, r, r, r, 2.0, 2.0])


def test_equal_spheres_match_sphere_contact_chain():
    rng = np.random.default_rng(3)
    mats = _mats()
    q = rng.normal(size=(400, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    params = {"tangential_damping": 1}
    for k in range(200):
        r = rng.uniform(1e-3, 2e-3)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = 2 * r * rng.uniform(0.97, 0.9999)
        xi = rng.normal(size=3) * 1e-2
        xj = xi - d * u  # del = xi - xj = d u
        vi, vj = rng.normal(size=3) * 0.1, rng.normal(size=3) * 0.1
        wi, wj = rng.normal(size=3) * 20, rng.normal(size=3) * 20
        V, _ = orc.sq_mass_props([r, r, r], [2, 2], 2500.0)
        m = V * 2500.0
        rv = np.cbrt(0.75 * V / np.pi)
        shear = rng.normal(size=3) * 1e-6
        shear -= (shear @ u) * u
        dt = 1e-5
        upd = k % 2
        t1, F1, Ti1, Tj1, sh1 = orc.contact_pair(params, mats, 1, 1, xi - xj, r, r, vi, vj, wi, wj, m, m, dt, upd,
                                                 shear)
        hist = np.zeros(orc.SQS_HIST)
        hist[7:10] = shear
        t2, F2, Ti2, Tj2, h2 = orc.sqs_contact_pair(mats, _pose(xi, q[2 * k], r), _pose(xj, q[2 * k + 1], r),
                                                    vi, vj, wi, wj, m, m, rv, rv, dt, upd, hist)
        assert t1 and t2
        scale_f = np.linalg.norm(F1)
        scale_t = max(np.linalg.norm(Ti1), np.linalg.norm(Tj1), 1e-30)
        assert np.allclose(F2, F1, rtol=0, atol=1e-8 * scale_f), k
        assert np.allclose(Ti2, Ti1, rtol=0, atol=1e-7 * scale_t + 1e-8 * scale_f * r), k
        assert np.allclose(Tj2, Tj1, rtol=0, atol=1e-7 * scale_t + 1e-8 * scale_f * r), k
        assert np.allclose(h2[7:10], sh1, rtol=0, atol=1e-8 * max(np.linalg.norm(sh1), 1e-12)), k


def _one(x, quat, shape, blk, v=(0, 0, 0), omega=(0, 0, 0), density=2500.0):
    return {"x": np.array([x], dtype=float), "v": np.array([v], dtype=float),
            "omega": np.array([omega], dtype=float), "quat": np.array([quat], dtype=float),
            "shape": np.array([shape], dtype=float), "blockiness": np.array([blk], dtype=float),
            "density": np.array([density]), "type": np.array([1], dtype=np.int32)}


def _params(grav=None):
    p = {"dt": 1e-5, "skin": 1e-3, "boxlo": [-1, -1, -1], "boxhi": [1, 1, 1], "periodic": [0, 0, 0]}
    if grav is not None:
        p["grav_acc"] = grav
    return p


def test_sphere_on_plane_hertz_closed_form():
    mats = _mats()
    Ystar = 1.0 / (2 * (1 - 0.45 ** 2) / 5e6)
    for r, delta in [(1e-3, 1e-6), (2e-3, 5e-6), (1.5e-3, 2e-7)]:
        P = _one([0, 0, r - delta], [1, 0, 0, 0], [r, r, r], [2, 2])
        s = orc.SqsSystem(_params(), mats, P, walls=[(2, 0.0, 1)])
        s.setup()
        st = s.state()
        rv = s.props()["rvol"][0]
        want = 4.0 / 3.0 * Ystar * np.sqrt(rv * delta) * delta
        assert st["f"][0] == pytest.approx([0, 0, want], rel=1e-6, abs=1e-12 * want)
        assert np.abs(st["torque"][0]).max() <= 1e-9 * want * r
        s.close()


def test_rotated_ellipsoid_on_plane_support_function():
    rng = np.random.default_rng(4)
    mats = _mats()
    Ystar = 1.0 / (2 * (1 - 0.45 ** 2) / 5e6)
    for _ in range(20):
        shape = rng.uniform(1e-3, 3e-3, 3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
        M = R @ np.diag(shape ** 2) @ R.T
        n = np.array([0.0, 0.0, -1.0])
        h = np.sqrt(n @ M @ n)  # support function along -z
        delta = h * rng.uniform(1e-4, 5e-3)
        z = h - delta
        P = _one([0.001, -0.002, z], q, shape, [2, 2])
        s = orc.SqsSystem(_params(), mats, P, walls=[(2, 0.0, 1)])
        s.setup()
        st = s.state()
        rv = s.props()["rvol"][0]
        want = 4.0 / 3.0 * Ystar * np.sqrt(rv * delta) * delta
        assert st["f"][0][2] == pytest.approx(want, rel=1e-6)
        assert abs(st["f"][0][0]) < 1e-12 * want and abs(st["f"][0][1]) < 1e-12 * want
        # the support point's lever arm gives the torque
        arm = M @ n / h
        assert st["torque"][0] == pytest.approx(np.cross(arm, [0, 0, want]), rel=1e-6, abs=1e-9 * want * h)
        s.close()


@pytest.mark.parametrize("blk", [(2.0, 2.0), (4.0, 4.0), (7.3, 3.1)])
def test_aligned_superquadric_on_plane(blk):
    mats = _mats()
    Ystar = 1.0 / (2 * (1 - 0.45 ** 2) / 5e6)
    shape = [2e-3, 1.5e-3, 1e-3]
    delta = 3e-6
    P = _one([0, 0, shape[2] - delta], [1, 0, 0, 0], shape, blk)
    s = orc.SqsSystem(_params(), mats, P, walls=[(2, 0.0, 1)])
    s.setup()
    st = s.state()
    rv = s.props()["rvol"][0]
    want = 4.0 / 3.0 * Ystar * np.sqrt(rv * delta) * delta
    assert st["f"][0] == pytest.approx([0, 0, want], rel=1e-6, abs=1e-12 * want)
    s.close()


def test_free_flight_under_gravity_exact():
    g = [0.0, 0.0, -9.81]
    P = _one([0.0, 0.0, 0.5], [1, 0, 0, 0], [1e-3, 2e-3, 1.5e-3], [3.0, 5.0], v=[0.1, -0.2, 0.3])
    s = orc.SqsSystem(_params(g), _mats(), P)
    s.setup()
    n = 2000
    s.run(n)
    t = n * 1e-5
    st = s.state()
    assert st["x"][0] == pytest.approx([0.1 * t, -0.2 * t, 0.5 + 0.3 * t - 0.5 * 9.81 * t * t], rel=1e-12)
    assert st["v"][0] == pytest.approx([0.1, -0.2, 0.3 - 9.81 * t], rel=1e-12)
    s.close()


def test_torque_free_asymmetric_top_conserves_momentum_and_energy():
    q = np.array([0.3, -0.5, 0.7, 0.2])
    q /= np.linalg.norm(q)
    P = _one([0, 0, 0], q, [1e-3, 2e-3, 3e-3], [2.0, 2.0], omega=[30.0, -50.0, 80.0])
    s = orc.SqsSystem(_params(), _mats(), P)
    s.setup()
    L0 = s.state()["angmom"][0].copy()
    e0 = 0.5 * s.state()["omega"][0] @ L0
    for _ in range(10):
        s.run(500)
        st = s.state()
        assert st["angmom"][0] == pytest.approx(L0, rel=1e-10, abs=1e-10 * np.linalg.norm(L0))  # rounding only
        e = 0.5 * st["omega"][0] @ st["angmom"][0]
        assert e == pytest.approx(e0, rel=1e-6)
        assert np.linalg.norm(st["quat"][0]) == pytest.approx(1.0, abs=1e-14)
    s.close()


def test_spherical_top_spins_at_constant_rate():
    """omega constant; after n steps the orientation is the rotation by
    omega * t (the splitting error of no-squish vanishes for equal moments
    up to O(dt^3) per step)."""
    w = np.array([40.0, -20.0, 10.0])
    P = _one([0, 0, 0], [1, 0, 0, 0], [2e-3, 2e-3, 2e-3], [2.0, 2.0], omega=w)
    s = orc.SqsSystem(_params(), _mats(), P)
    s.setup()
    n = 1000
    s.run(n)
    st = s.state()
    assert st["omega"][0] == pytest.approx(w, rel=1e-12)
    rot = Rotation.from_rotvec(w * n * 1e-5)
    qs = Rotation.from_quat([*st["quat"][0][1:], st["quat"][0][0]])
    # splitting error bound n (|omega| dt)^3 ~ 1e-7 rad against a 0.46 rad turn
    assert (rot.inv() * qs).magnitude() < n * (np.linalg.norm(w) * 1e-5) ** 3
    s.close()
