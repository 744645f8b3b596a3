"""Pins the oracle's sibling contact models (pair_style gran model hooke,
hertz/stiffness, hooke/stiffness; tangential no_history; cohesion sjkr)
against what mechanics fixes independently of the oracle's code:

  hooke             undamped impact at the characteristic velocity reaches the
                    Hertz maximum overlap (the stiffness is defined that way),
                    lasts pi * delta_max / v, and the damped impact returns the
                    input restitution coefficient
  hooke/stiffness   linear spring-dashpot impact: restitution
                    exp(-pi zeta / sqrt(1 - zeta^2)) and duration pi / omega_d,
                    with the damping per unit mass or absolute
  hertz/stiffness   static force kn sqrt(R* delta) delta
  ktToKnUser        the tangential spring is 2/7 of the normal one
  no_history        viscous friction gammat |v_t| below the Coulomb cap and
                    mu |Fn| above it; a sliding sphere rolls at 5/7 v0
  sjkr              the cohesive force is the energy density times the area of
                    the circle where the two spheres (or sphere and plane)
                    intersect, attractive; a sphere resting on a floor settles
                    where Hertz repulsion balances weight plus cohesion;
                    tangential_reduce raises the Coulomb cap by the cohesion
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import pyoracle as orc
import workloads as wl

R, RHO = 0.002, 2500.0


def _mats(**kw):
    m = wl.materials_single(**{k: kw.pop(k) for k in ("restitution", "friction", "youngs", "poisson")
                               if k in kw})
    for k, v in kw.items():
        m[k] = v if k == "characteristic_velocity" else [[v]]
    return m


def _ystar(Y=5e6, nu=0.45):
    return Y / (2.0 * (1.0 - nu * nu))


def _impact(model, mats, v0, dt, max_steps=20000):
    """Two equal spheres meeting head-on at relative speed v0 along x: returns
    (restitution, contact duration, maximum overlap)."""
    rad = np.array([R, R])
    mass = wl.sphere_mass(rad, RHO)
    x = np.array([[0.5 - R, 0.5, 0.5], [0.5 + R, 0.5, 0.5]])  # just touching
    v = np.array([[0.5 * v0, 0, 0], [-0.5 * v0, 0, 0]])
    params = {"dt": dt, "skin": 1e-4, "boxlo": (0, 0, 0), "boxhi": (1, 1, 1), "periodic": (0, 0, 0),
              **model}
    P = {"x": x, "v": v, "omega": np.zeros((2, 3)), "radius": rad, "rmass": mass,
         "type": np.array([1, 1]), "tag": np.array([1, 2])}
    s = orc.OracleSystem(params, mats, P)
    s.setup()
    dmax, steps = 0.0, 0
    while True:
        s.run_continue(1)
        steps += 1
        st = s.state()
        gap = np.linalg.norm(st["x"][1] - st["x"][0]) - 2 * R
        dmax = max(dmax, -gap)
        if gap >= 0:
            break
        assert steps < max_steps
    vrel = st["v"][1, 0] - st["v"][0, 0]
    s.close()
    return vrel / v0, steps * dt, dmax


def test_hooke_stiffness_is_set_by_the_characteristic_velocity():
    """Undamped (e = 1) hooke impact at v0 = characteristicVelocity: maximum
    overlap = Hertz's (15 m* v^2 / (16 Y* sqrt(R*)))^(2/5), duration
    pi delta_max / v0 (harmonic half period with k = m* v0^2 / delta_max^2)."""
    cv = 0.4
    m = wl.sphere_mass(np.array([R]), RHO)[0]
    meff, reff = m / 2, R / 2
    dmax_hertz = (15 * meff * cv ** 2 / (16 * _ystar() * math.sqrt(reff))) ** 0.4
    tc = math.pi * dmax_hertz / cv
    cor, t, dmax = _impact({"normal": "hooke"}, _mats(restitution=1.0, characteristic_velocity=cv), cv,
                           tc / 3000)
    assert dmax == pytest.approx(dmax_hertz, rel=1e-5)
    assert t == pytest.approx(tc, rel=1e-3)
    assert cor == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("e", [0.3, 0.6, 0.9])
def test_hooke_returns_the_input_restitution(e):
    """The hooke damping makes a linear oscillator whose restitution is exactly
    the input e, at any impact speed."""
    cv = 0.4
    for v0 in (0.1, 1.0):
        cor, _, _ = _impact({"normal": "hooke"}, _mats(restitution=e, characteristic_velocity=cv), v0, 2e-7)
        assert cor == pytest.approx(e, rel=2e-3)


@pytest.mark.parametrize("absolute", [0, 1])
def test_hooke_stiffness_damped_oscillator(absolute):
    kn, gn = 100.0, 500.0
    m = wl.sphere_mass(np.array([R]), RHO)[0]
    meff = m / 2
    gamma = gn * meff
    gn_in = gamma if absolute else gn  # gamman_abs is the damping itself
    zeta = gamma / (2 * math.sqrt(kn * meff))
    e_th = math.exp(-math.pi * zeta / math.sqrt(1 - zeta * zeta))
    wd = math.sqrt(kn / meff) * math.sqrt(1 - zeta * zeta)
    mats = _mats(kn=kn, kt=kn, gamman=gn_in, gammat=gn_in)
    cor, t, _ = _impact({"normal": "hooke/stiffness", "absolute_damping": absolute}, mats, 0.2,
                        math.pi / wd / 4000)
    assert cor == pytest.approx(e_th, rel=1e-3)
    assert t == pytest.approx(math.pi / wd, rel=1e-3)


def _pair(model, mats, del_, ri=R, rj=R, vi=(0, 0, 0), vj=(0, 0, 0), mi=1e-4, mj=1e-4,
          shear=(0, 0, 0), su=1, dt=1e-6):
    return orc.contact_pair(model, orc.materials(mats), 1, 1, del_, ri, rj, vi, vj, (0, 0, 0),
                            (0, 0, 0), mi, mj, dt, su, shear)


def test_hertz_stiffness_static_force():
    kn = 3e8
    mats = _mats(kn=kn, kt=kn, gamman=10.0, gammat=10.0)
    for ri, rj, d in ((0.002, 0.002, 1e-6), (0.0015, 0.001, 3e-5)):
        rs = ri * rj / (ri + rj)
        touch, dF, dTi, _, _ = _pair({"normal": "hertz/stiffness"}, mats, (ri + rj - d, 0, 0), ri, rj)
        assert touch
        assert dF[0] == pytest.approx(kn * math.sqrt(rs * d) * d, rel=1e-12)
        assert dF[1] == 0 and dF[2] == 0 and not dTi.any()


@pytest.mark.parametrize("kt_to_kn", [0, 1])
def test_kt_to_kn_scales_the_tangential_spring(kt_to_kn):
    """hooke with tangential history: Fn = kn delta at rest; a small sticking
    shear s gives |Ft| = kt s with kt = kn (or 2/7 kn with ktToKnUser)."""
    mats = _mats(restitution=0.5, friction=0.9, characteristic_velocity=0.5)
    d, s = 1e-5, 1e-8
    model = {"normal": "hooke", "kt_to_kn": kt_to_kn}
    _, dF0, _, _, _ = _pair(model, mats, (2 * R - d, 0, 0), su=0)
    kn = dF0[0] / d
    _, dF, _, _, _ = _pair(model, mats, (2 * R - d, 0, 0), shear=(0, s, 0), su=0)
    ratio = -dF[1] / s / kn
    assert ratio == pytest.approx(0.285714286 if kt_to_kn else 1.0, rel=1e-9)


def test_no_history_friction_viscous_then_coulomb():
    """hooke + no_history: gammat = gamman (tangential damping on); at low
    sliding speed Ft = -gamman v_t, above the cap |Ft| = mu |Fn|."""
    mats = _mats(restitution=0.5, friction=0.3, characteristic_velocity=0.5)
    d = 1e-5
    model = {"normal": "hooke", "tangential": "no_history"}
    _, dFs, _, _, _ = _pair(model, mats, (2 * R - d, 0, 0), su=0)
    Fn = dFs[0]
    # gamman from the normal damping: Fn(vn) - Fn(0) = -gamman vn
    _, dFd, _, _, _ = _pair(model, mats, (2 * R - d, 0, 0), vi=(1e-3, 0, 0), su=0)
    gamman = -(dFd[0] - Fn) / 1e-3
    vslow = 1e-3 * 0.3 * Fn / gamman
    _, dF, dTi, dTj, sh = _pair(model, mats, (2 * R - d, 0, 0), vi=(0, vslow, 0))
    assert dF[1] == pytest.approx(-gamman * vslow, rel=1e-10)
    assert not sh.any()  # nothing is stored
    # torque on both sides: -cr (en x Ft), cr = R - d/2
    assert dTi[2] == pytest.approx(-(R - d / 2) * dF[1], rel=1e-12)
    assert dTj[2] == pytest.approx(dTi[2], rel=1e-12)
    _, dF, _, _, _ = _pair(model, mats, (2 * R - d, 0, 0), vi=(0, 0, 5.0))
    assert dF[2] == pytest.approx(-0.3 * Fn, rel=1e-12)


def test_no_history_sliding_sphere_rolls_at_five_sevenths():
    g, v0, mu = 9.81, 0.2, 0.5
    m = wl.sphere_mass(np.array([R]), RHO)
    Ys = _ystar()
    d_eq = (m[0] * g / (4.0 / 3.0 * Ys * math.sqrt(R))) ** (2.0 / 3.0)
    params = {"dt": 2e-6, "skin": 0.001, "boxlo": (-0.05, -0.05, 0.0), "boxhi": (0.05, 0.05, 0.05),
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -g), "tangential": "no_history"}
    P = {"x": np.array([[0.0, 0.0, R - d_eq]]), "v": np.array([[v0, 0.0, 0.0]]),
         "omega": np.zeros((1, 3)), "radius": np.array([R]), "rmass": m, "type": np.array([1]),
         "tag": np.array([1])}
    s = orc.OracleSystem(params, wl.materials_single(friction=mu), P, walls=[(2, 0.0, 1)])
    t_roll = 2.0 * v0 / (7.0 * mu * g)
    s.run(int(2.0 * t_roll / params["dt"]))
    st = s.state()
    assert st["v"][0, 0] == pytest.approx(5.0 / 7.0 * v0, rel=2e-3)
    assert st["omega"][0, 1] * R == pytest.approx(st["v"][0, 0], rel=2e-3)
    s.close()


def _circle_area(ri, rj, r):
    """area of the circle where spheres ri, rj at distance r intersect"""
    x = (r * r + ri * ri - rj * rj) / (2 * r)
    return math.pi * (ri * ri - x * x)


def test_sjkr_pair_force_is_energy_density_times_contact_area():
    ced = 2.0e5
    for ri, rj, d in ((0.002, 0.002, 2e-5), (0.0015, 0.001, 1e-4), (0.001, 0.003, 3e-6)):
        base = _mats(restitution=0.5)
        _, dF0, _, _, _ = _pair({}, base, (ri + rj - d, 0, 0), ri, rj)
        _, dF1, _, _, _ = _pair({"cohesion": "sjkr"}, _mats(restitution=0.5, cohesion_energy_density=ced),
                                (ri + rj - d, 0, 0), ri, rj)
        # en = +x points from j to i: cohesion pulls i towards j, along -x
        assert dF1[0] - dF0[0] == pytest.approx(-ced * _circle_area(ri, rj, ri + rj - d), rel=1e-8)


def test_sjkr_tangential_reduce_raises_the_coulomb_cap():
    ced, mu, d = 5.0e5, 0.3, 2e-5
    mats = _mats(restitution=0.5, friction=mu, cohesion_energy_density=ced)
    A = _circle_area(R, R, 2 * R - d)
    for reduce in (0, 1):
        model = {"cohesion": "sjkr", "tangential_reduce": reduce}
        _, dFs, _, _, _ = _pair({}, _mats(restitution=0.5, friction=mu), (2 * R - d, 0, 0), su=0)
        Fn = dFs[0]  # Hertz part at rest
        # a large shear saturates the Coulomb cap
        _, dF, _, _, _ = _pair(model, mats, (2 * R - d, 0, 0), shear=(0, 1e-3, 0), su=0)
        cap = mu * abs(Fn - ced * A) if reduce else mu * abs(Fn)
        assert -dF[1] == pytest.approx(cap, rel=1e-9)


def test_sjkr_sphere_on_floor_settles_where_forces_balance():
    """Weight plus wall cohesion ced pi (R^2 - h^2) (h: centre height) against
    the Hertz wall force 4/3 Y* sqrt(R) delta^(3/2)."""
    g, ced = 9.81, 1.0e5
    m = wl.sphere_mass(np.array([R]), RHO)[0]
    Ys = _ystar()  # the wall has the particles' type
    f = lambda d: 4.0 / 3.0 * Ys * math.sqrt(R) * d ** 1.5 - m * g - ced * math.pi * (R * R - (R - d) ** 2)
    d_eq = brentq(f, 1e-12, 1e-4, xtol=1e-16)
    params = {"dt": 2e-6, "skin": 0.001, "boxlo": (-0.05, -0.05, 0.0), "boxhi": (0.05, 0.05, 0.05),
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -g), "cohesion": "sjkr"}
    P = {"x": np.array([[0.0, 0.0, R - 0.5 * d_eq]]), "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)),
         "radius": np.array([R]), "rmass": np.array([m]), "type": np.array([1]), "tag": np.array([1])}
    s = orc.OracleSystem(params, _mats(restitution=0.3, cohesion_energy_density=ced), P, walls=[(2, 0.0, 1)])
    s.run(20000)
    st = s.state()
    assert R - st["x"][0, 2] == pytest.approx(d_eq, rel=1e-4)
    # without cohesion the overlap is smaller
    assert d_eq > (m * g / (4.0 / 3.0 * Ys * math.sqrt(R))) ** (2.0 / 3.0) * 1.05
    s.close()
