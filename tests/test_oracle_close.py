"""Pins of two parts of the oracle's tangential history that the other
oracle tests leave open:

1. The surfacesClose branch (pair_gran_base.h:414-417), reachable with
   neigh_modify contact_distance_factor > 1: a pair within cdf * radsum that
   does not touch has its shear zeroed (tangential_model_history.h:358-369),
   and its contact flag, which decides whether the pair is carried over at the
   next build, stays set only if the normal model sets its bit there
   (hertz, hooke, hooke/stiffness) or the build carried the pair over
   (hertz/stiffness). A plane wall in the band gets zero history.
2. The re-projection of the shear onto the tangent plane under a rotating
   normal (shear -= (shear . en) en, tangential_model_history.h:166-176):
   for two spheres in rigid co-rotation the surface velocity is zero, so the
   shear only follows the normal, and |shear| shrinks by exactly cos(w dt) per
   step while staying perpendicular to en.
"""
import math

import numpy as np
import pytest

import pyoracle as orc
import workloads as wl

R = 0.002
RHO = 2500.0
STIFF = {"kn": [[4.0e6]], "kt": [[5.0e6]], "gamman": [[2.0e5]], "gammat": [[1.0e5]]}


def _two(params_extra, x, v, omega=None, mats=None, walls=()):
    params = {"dt": 1e-5, "skin": 0.001, "cdf": 1.0, "boxlo": (-0.02, -0.02, -0.02),
              "boxhi": (0.02, 0.02, 0.02), "periodic": (0, 0, 0), "gravity_first": 1,
              "rolling_cdt": 0}
    params.update(params_extra)
    n = len(x)
    r = np.full(n, R)
    P = {"x": np.array(x, float), "v": np.array(v, float),
         "omega": np.zeros((n, 3)) if omega is None else np.array(omega, float),
         "radius": r, "rmass": wl.sphere_mass(r, RHO),
         "type": np.ones(n, dtype=np.int32), "tag": np.arange(1, n + 1, dtype=np.int32)}
    m = wl.materials_single() if mats is None else mats
    return orc.OracleSystem(params, m, P, walls), params, m, P


def _oblique_pair(extra, mats=None):
    """Two spheres overlapping by 1e-3 of their diameter, sliding past each
    other at 1 cm/s: the Hertz push separates them after ~80 steps with a
    shear built up, and they drift apart through the band for thousands of
    steps (no gravity, no walls, no rebuild in the first 400 steps)."""
    d0 = 2 * R * (1 - 1e-3)
    return _two(extra, [(0, 0, 0), (d0, 0, 0)], [(0, 0.005, 0), (0, -0.005, 0)], mats=mats)


def _run_until_apart(o, extra_steps=20, limit=400):
    for k in range(limit):
        o.run_continue(1)
        x = o.state()["x"]
        if np.linalg.norm(x[1] - x[0]) >= 2 * R:
            o.run_continue(extra_steps)
            x = o.state()["x"]
            gap = np.linalg.norm(x[1] - x[0])
            return gap
    raise AssertionError("the pair did not separate")


@pytest.mark.parametrize("normal", ["hertz", "hooke", "hertz/stiffness"])
def test_band_zeroes_shear_and_keeps_flag_by_model(normal):
    extra = {"cdf": 1.1, "normal": normal}
    mats = wl.materials_single()
    if normal == "hooke":
        mats = dict(mats, characteristic_velocity=1.0)
    if normal == "hertz/stiffness":
        mats = dict(mats, **STIFF)
    o, *_ = _oblique_pair(extra, mats)
    o.setup()
    o.run_continue(30)  # still touching: shear built up
    t, f, h = o.pairs()
    assert len(t) == 1 and f[0] == 1 and np.linalg.norm(h[0]) > 1e-9
    gap = _run_until_apart(o)
    assert 2 * R < gap < 1.1 * 2 * R  # in the band
    assert o.nrebuilds() == 1
    t, f, h = o.pairs()
    assert len(t) == 1
    assert np.array_equal(h[0], np.zeros(3))  # surfacesClose zeroed it
    # the pair touched after the setup build (not carried): only the models
    # whose normal part sets its flag bit in the band keep the flag
    assert f[0] == (0 if normal == "hertz/stiffness" else 1)
    o.close()


def test_without_band_the_stale_shear_stays():
    """cdf = 1: the same trajectory leaves the shear of the last touching
    step and the flag untouched until the next build (the branch is
    unreachable)."""
    o, *_ = _oblique_pair({"cdf": 1.0})
    o.setup()
    last = None
    for _ in range(400):
        o.run_continue(1)
        x = o.state()["x"]
        if np.linalg.norm(x[1] - x[0]) >= 2 * R:
            break
        last = o.pairs()[2][0].copy()
    o.run_continue(20)
    t, f, h = o.pairs()
    assert f[0] == 1 and np.linalg.norm(last) > 0
    assert np.array_equal(h[0], last)
    o.close()


def test_band_flag_of_a_carried_hertz_stiffness_pair():
    """With a build every step (neigh_modify every 1 check no) a touching
    pair is carried over at each build, which stores its flag as 1 (the
    normal bit, neigh_gran.cpp:592). hertz/stiffness never clears that bit,
    so in the band the flag stays set, the pair keeps being carried while it
    is within cdf * radsum, and its history stays zero."""
    mats = dict(wl.materials_single(), **STIFF)
    o, *_ = _oblique_pair({"cdf": 1.1, "normal": "hertz/stiffness", "neigh_every": 1,
                           "neigh_check": False}, mats)
    o.setup()
    _run_until_apart(o, extra_steps=5)
    t, f, h = o.pairs()
    assert len(t) == 1 and f[0] == 1
    assert np.array_equal(h[0], np.zeros(3))
    o.close()


def _half_step(st, rmass, dt):
    """Velocities and positions the next step's forces see (initial
    integrate of fix nve/sphere from the state after a step)."""
    dtf = 0.5 * dt
    m = rmass[:, None]
    v = st["v"] + dtf * st["f"] / m
    x = st["x"] + dt * v
    w = st["omega"] + (dtf / 0.4) * st["torque"] / (m * R * R)
    return x, v, w


@pytest.mark.parametrize("cdf", [1.0, 1.1])
def test_retouch_after_the_band_starts_from_zero_shear(cdf):
    """A sphere dropped obliquely onto a sphere resting on a floor bounces
    (e = 0.3) to a small height inside the band and lands again. At the
    landing step its shear must equal one contact update from zero shear
    (cdf 1.1: the band zeroed it) or from the stale shear of the first impact
    (cdf 1: nothing zeroed it), each computed by the single-contact routine
    from the pre-step state."""
    dt = 1e-5
    extra = {"cdf": cdf, "grav_acc": (0.0, 0.0, -9.81), "dt": dt}
    # lower sphere resting on the floor, upper one 0.2 mm above contact, 0.5 mm off axis
    xb = (0.0, 0.0, R)
    off = 0.0005
    zc = math.sqrt((2 * R) ** 2 - off ** 2)
    xa = (off, 0.0, R + zc + 0.0002)
    o, params, mats, P = _two(extra, [xb, xa], [(0, 0, 0), (0.01, 0, -0.05)],
                              walls=[(2, 0.0, 1)])
    o.setup()
    phase, stale, landed = 0, None, False
    for k in range(6000):
        prev = o.state()
        prev_pairs = o.pairs()
        o.run_continue(1)
        st = o.state()
        d = np.linalg.norm(st["x"][1] - st["x"][0])
        if phase == 0 and d < 2 * R:
            phase = 1  # first impact
        elif phase == 1 and d >= 2 * R:
            phase = 2  # bouncing, inside the band
            assert d < 2 * R * 1.1
        elif phase == 2:
            assert d < 2 * R * 1.1, "the bounce left the band"
            if d < 2 * R:
                landed = True
                break
    assert landed
    assert o.nrebuilds() == 1  # no build between the impacts
    tags, flags, hist = prev_pairs
    sel = [q for q, (a, b) in enumerate(tags) if {a, b} == {1, 2}]
    assert len(sel) == 1
    q = sel[0]
    stale = hist[q] * (1 if tuple(tags[q]) == (1, 2) else -1)  # oriented from tag 1 to tag 2
    if cdf > 1:
        assert np.array_equal(stale, np.zeros(3))
    else:
        assert np.linalg.norm(stale) > 0
    x, v, w = _half_step(prev, P["rmass"], dt)
    del_ = x[0] - x[1]
    touch, dF, dTi, dTj, sh = orc.contact_pair(params, orc.materials(mats), 1, 1, del_, R, R, v[0], v[1],
                                               w[0], w[1], P["rmass"][0], P["rmass"][1], dt, 1, stale)
    assert touch
    tags, flags, hist = o.pairs()
    q = [q for q, (a, b) in enumerate(tags) if {a, b} == {1, 2}][0]
    got = hist[q] * (1 if tuple(tags[q]) == (1, 2) else -1)
    assert np.linalg.norm(got - sh) <= 1e-12 * np.linalg.norm(sh)
    o.close()


def test_wall_band_zeroes_history():
    """A sphere leaving a plane wall into the band (deltan < (cdf - 1) r)
    has zero wall history, as one beyond it has."""
    o, *_ = _two({"cdf": 1.2}, [(0, 0, R * (1 - 1e-3))], [(0.01, 0, 0.002)],
                 walls=[(2, 0.0, 1)])
    o.setup()
    o.run_continue(20)
    assert np.linalg.norm(o.wall_history()[0, 0]) > 0
    for _ in range(400):
        o.run_continue(1)
        z = o.state()["x"][0, 2]
        if z > R:
            break
    o.run_continue(5)
    z = o.state()["x"][0, 2]
    assert R < z < 1.2 * R
    assert np.array_equal(o.wall_history()[0, 0], np.zeros(3))
    o.close()


def test_shear_follows_a_rotating_normal():
    """Rigid co-rotation of two touching spheres about their midpoint: the
    relative surface velocity vanishes, so each step only re-projects the
    shear onto the plane normal to the new en. Starting from a shear along
    the in-plane tangent, after N steps |shear| = s0 cos(w dt)^N, it lies
    along the tangent of step N, and shear . en = 0. A dropped projection
    keeps |shear| = s0, a flipped sign grows it, a magnitude-preserving
    rotation keeps s0 as well."""
    mats = orc.materials(wl.materials_single(friction=0.9))
    params = {"rolling_cdt": 0}
    m = wl.sphere_mass(R, RHO)
    delta = 2e-6
    half = R - 0.5 * delta  # centre to midpoint
    w, dt, N = 2000.0, 1e-5, 400
    s0 = 5e-7  # kt s0 stays below mu Fn: no Coulomb cap
    th0 = 0.3
    t0 = np.array([-math.sin(th0), math.cos(th0), 0.0])
    sh = s0 * t0
    omega = np.array([0.0, 0.0, w])
    for k in range(1, N + 1):
        th = th0 + k * w * dt
        en = np.array([math.cos(th), math.sin(th), 0.0])
        xi, xj = half * en, -half * en
        vi, vj = np.cross(omega, xi), np.cross(omega, xj)
        touch, dF, dTi, dTj, sh = orc.contact_pair(params, mats, 1, 1, xi - xj, R, R, vi, vj, omega, omega,
                                                   m, m, dt, 1, sh)
        assert touch
    th = th0 + N * w * dt
    en = np.array([math.cos(th), math.sin(th), 0.0])
    tn = np.array([-math.sin(th), math.cos(th), 0.0])
    expect = s0 * math.cos(w * dt) ** N
    assert abs(np.dot(sh, en)) <= 1e-12 * s0
    assert abs(sh[2]) <= 1e-12 * s0
    assert np.dot(sh, tn) == pytest.approx(expect, rel=1e-9)
    assert expect < 0.99 * s0  # the shrink is visible at this N
