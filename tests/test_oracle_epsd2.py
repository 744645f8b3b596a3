"""Pins the oracle's rolling_friction epsd2 (rolling_model_epsd2.h:98-221)
against what the model's mechanics fixes independently of the oracle's code:

  spring      from an empty history one step adds dt kr w_rel with
              kr = kt R*^2, kt = 8 G* sqrt(R* delta) the Hertz tangential
              stiffness; opposite spins of two equal spheres roll without
              slip, so the rolling torque is the whole torque
  torsion     the relative spin about the contact normal is dropped unless
              torsionTorque is on
  cap         the accumulated torque stops at mu_r |Fn| R*, along w_rel
  history     shearupdate off leaves the history; the partner's view of a
              contact (newtonflag 1) negates it and gets the opposite torque
  wall        the plane contact spins at (R - delta/2)/(R - delta) omega, the
              surface rate of the sphere, with R* = R
  mesh        a tessellated flat floor gives the plane's state and histories
              while the spheres roll across its triangles' edges
"""
import math

import numpy as np
import pytest

import pyoracle as orc

R, RHO, Y, NU, MUR = 0.002, 2500.0, 5e6, 0.45, 0.3
MASS = RHO * 4.0 / 3.0 * math.pi * R ** 3
MATS = {"ntypes": 1, "youngs": [Y], "poisson": [NU], "restitution": [[0.5]],
        "friction": [[0.5]], "rolling_friction": [[MUR]]}
EPSD2 = {"rolling_epsd2": 1}


def _hertz(reff, delta):
    ystar = Y / (2.0 * (1.0 - NU * NU))
    gstar = Y / (4.0 * (2.0 - NU) * (1.0 + NU))
    sq = math.sqrt(reff * delta)
    return 4.0 / 3.0 * ystar * sq, 8.0 * gstar * sq  # kn, kt


def _pair(model, wi, wj, rh, dt=1e-6, delta=1e-5, shearupdate=1):
    del_ = (2 * R - delta, 0.0, 0.0)
    return orc.contact_pair_rolling(model, orc.materials(MATS), 1, 1, del_, R, R, (0, 0, 0),
                                    (0, 0, 0), wi, wj, MASS, MASS, dt, shearupdate, (0, 0, 0), rh)


def test_spring_one_step_rolls_without_slip():
    dt, delta = 1e-6, 1e-5
    touch, dF, dTi, dTj, sh, rh = _pair(EPSD2, (0, 10, 0), (0, -10, 0), (0, 0, 0), dt, delta)
    assert touch
    kn, kt = _hertz(R / 2, delta)
    rt = np.array([0.0, 20.0, 0.0]) * dt * kt * (R / 2) ** 2
    assert np.linalg.norm(rt) < kn * delta * (R / 2) * MUR  # below the cap
    np.testing.assert_allclose(sh, 0.0, atol=1e-30)  # no sliding at the contact point
    np.testing.assert_allclose(dF, [kn * delta, 0, 0], rtol=1e-12)
    np.testing.assert_allclose(rh, rt, rtol=1e-12)
    np.testing.assert_allclose(dTi, -rt, rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(dTj, rt, rtol=1e-12, atol=1e-30)


def test_no_rolling_torque_without_the_model():
    _, _, dTi, dTj, _, rh = _pair({}, (0, 10, 0), (0, -10, 0), (1, 2, 3))
    np.testing.assert_allclose(dTi, 0.0, atol=1e-30)
    np.testing.assert_allclose(dTj, 0.0, atol=1e-30)
    np.testing.assert_array_equal(rh, [1, 2, 3])


@pytest.mark.parametrize("torsion", [0, 1])
def test_torsion_about_the_normal(torsion):
    dt, delta = 1e-6, 1e-5
    _, _, dTi, dTj, _, rh = _pair({**EPSD2, "torsion_torque": torsion}, (5, 0, 0), (-5, 0, 0),
                                  (0, 0, 0), dt, delta)
    _, kt = _hertz(R / 2, delta)
    want = np.array([10.0, 0, 0]) * dt * kt * (R / 2) ** 2 * torsion
    np.testing.assert_allclose(dTj, want, rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(dTi, -want, rtol=1e-12, atol=1e-30)


def test_cap_at_rolling_friction_times_normal_force():
    dt, delta = 1e-5, 1e-5
    w = np.array([0.0, 30.0, 40.0])
    kn, kt = _hertz(R / 2, delta)
    rmax = kn * delta * (R / 2) * MUR
    rh = np.zeros(3)
    steps = 0
    while True:
        _, _, dTi, dTj, _, rh = _pair(EPSD2, 0.5 * w, -0.5 * w, rh, dt, delta)
        steps += 1
        if np.linalg.norm(rh) >= rmax * (1 - 1e-12):
            break
        np.testing.assert_allclose(rh, w * dt * kt * (R / 2) ** 2 * steps, rtol=1e-10)
    assert steps > 1
    np.testing.assert_allclose(np.linalg.norm(rh), rmax, rtol=1e-12)
    np.testing.assert_allclose(rh / np.linalg.norm(rh), w / np.linalg.norm(w), rtol=1e-12)
    np.testing.assert_allclose(dTj, rh, rtol=1e-12)
    # held at the cap on further steps
    _, _, _, _, _, rh2 = _pair(EPSD2, 0.5 * w, -0.5 * w, rh, dt, delta)
    np.testing.assert_allclose(rh2, rh, rtol=1e-12)


def test_history_kept_without_shearupdate_and_negated_for_the_partner():
    h = np.array([1e-9, -2e-9, 3e-9])
    _, _, dTi, dTj, _, rh = _pair(EPSD2, (0, 10, 0), (0, -10, 0), h, shearupdate=0)
    np.testing.assert_array_equal(rh, h)
    # the same contact seen from j: del, spins and history negated
    _, _, dTi2, dTj2, _, rh2 = orc.contact_pair_rolling(
        EPSD2, orc.materials(MATS), 1, 1, (-(2 * R - 1e-5), 0.0, 0.0), R, R, (0, 0, 0), (0, 0, 0),
        (0, -10, 0), (0, 10, 0), MASS, MASS, 1e-6, 0, (0, 0, 0), -h)
    np.testing.assert_allclose(dTi2, dTj, rtol=1e-12)
    np.testing.assert_allclose(dTj2, dTi, rtol=1e-12)


def test_wall_rolls_at_the_surface_rate():
    dt, delta = 1e-6, 1e-5
    om = np.array([0.0, 7.0, 3.0])  # floor below (en = -z): spin part along z is torsion
    mats = orc.materials(MATS)
    args = (mats, 1, 1, delta, (0, 0, -(R - delta)), R, (0, 0, 0), om, (0, 0, 0), MASS, dt, 1,
            (0, 0, 0))
    dF0, dT0, sh0 = orc.contact_wall({}, *args)
    dF, dT, sh, rh = orc.contact_wall_rolling(EPSD2, *args, (0, 0, 0))
    _, kt = _hertz(R, delta)
    wr = (R - 0.5 * delta) / (R - delta) * np.array([0.0, 7.0, 0.0])
    rt = wr * dt * kt * R * R
    np.testing.assert_allclose(dF, dF0, rtol=0, atol=0)
    np.testing.assert_allclose(sh, sh0, rtol=0, atol=0)
    np.testing.assert_allclose(rh, rt, rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(dT - dT0, -rt, rtol=1e-9, atol=1e-25)


def test_epsd2_on_a_flat_mesh_equals_the_plane():
    """A tessellated flat floor with rolling_friction epsd2 is the zplane wall
    at the same height: spheres pressed onto it and rolling across triangle
    edges (the slot history handed to the coplanar neighbour) keep the same
    state, shear and rolling torque as on the plane, to rounding."""
    import meshes as msh
    rng = np.random.default_rng(5)
    n = 9
    xy = np.stack(np.meshgrid(np.linspace(0.004, 0.016, 3), np.linspace(0.004, 0.016, 3)), -1).reshape(-1, 2)
    P = {"x": np.column_stack([xy, np.full(n, R - 2.7e-6)]),  # the static overlap under gravity
         "v": np.column_stack([rng.normal(0, 1.0, (n, 2)), np.zeros(n)]),
         "omega": rng.normal(0, 200.0, (n, 3)),
         "radius": np.full(n, R), "rmass": np.full(n, MASS), "type": np.ones(n, dtype=np.int32),
         "tag": np.arange(1, n + 1, dtype=np.int32)}
    params = {"dt": 1e-6, "skin": 2e-4, "boxlo": (-0.01, -0.01, -0.01), "boxhi": (0.03, 0.03, 0.03),
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1}
    floor = msh.grid(-0.01, 0.03, -0.01, 0.03, 0.0, 8, 8)
    a = orc.OracleSystem(params, MATS, P, walls=[(2, 0.0, 1, EPSD2)])
    b = orc.OracleSystem(params, MATS, P, meshes=[(floor, 1, EPSD2, None)])
    a.run(2000)
    b.run(2000)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega", "f", "torque"):
        scale = np.abs(sa[k]).max()
        assert np.abs(sa[k] - sb[k]).max() <= 1e-10 * scale, k
    # the spheres crossed triangle edges, and the rolling torque is not small
    assert np.abs(sa["x"][:, :2] - P["x"][:, :2]).max() > 0.001
    tri, hist = b.mesh_history(0)
    wh = a.wall_history()
    touching = 0
    for i in range(n):
        occ = tri[i] >= 0
        assert occ.sum() <= 1
        if occ.sum():
            touching += 1
            np.testing.assert_allclose(hist[i][occ][0], wh[i, 0], rtol=1e-8, atol=1e-14)
    assert touching >= 5
    a.close()
    b.close()
