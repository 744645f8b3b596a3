"""Pins of the oracle's non-periodic boundary faces against free flight,
where the velocity-Verlet positions are exact (x_n = x_0 + n v dt with no
force), so the neighbor builds (every time the atom has moved more than
skin/2 since the last one) and the positions the builds see are known in
closed form:

  f face      the atom is lost at the first build that sees it at or beyond
              the face (x >= hi, or x < lo; Comm::exchange, comm.cpp:945-1003);
              not before, and an atom exactly at lo is kept
  s face      nothing is lost; the box face sits at the atom extent of the last
              build plus 1e-4 of the created box length (Domain::reset_box,
              domain.cpp:381-449)
  m face      as s, but never inside its created position
"""
import numpy as np
import pytest

import pyoracle as orc
import workloads as wl

R = 0.001
DT = 1e-5
SKIN = 0.001


def _one(vx, x0, faces=(0, 0), extra=None):
    params = {"dt": DT, "skin": SKIN, "cdf": 1.0, "boxlo": (0.0, 0.0, 0.0), "boxhi": (0.01, 0.01, 0.01),
              "periodic": (0, 0, 0), "gravity_first": 1, "rolling_cdt": 0,
              "shrink_faces": faces[0], "shrink_min": faces[1]}
    params.update(extra or {})
    P = {"x": np.array([[x0, 0.005, 0.005]]), "v": np.array([[vx, 0.0, 0.0]]), "omega": np.zeros((1, 3)),
         "radius": np.array([R]), "rmass": wl.sphere_mass(np.array([R]), 2500.0),
         "type": np.ones(1, dtype=np.int32), "tag": np.ones(1, dtype=np.int32)}
    return orc.OracleSystem(params, wl.materials_single(), P, ()), params


def _builds(x0, vx, nsteps):
    """Closed form: (step, x seen) of every build of a run of nsteps from
    setup (step 0). A build happens at step n when |x_n - x_last| > skin/2."""
    out = [(0, x0)]
    xl = x0
    for n in range(1, nsteps + 1):
        xn = x0 + n * vx * DT
        if (xn - xl) ** 2 > 0.25 * SKIN * SKIN:
            out.append((n, xn))
            xl = xn
    return out


@pytest.mark.parametrize("vx,x0", [(1.7, 0.0081), (-1.7, 0.0019)])
def test_f_face_loses_the_atom_at_the_first_build_beyond_it(vx, x0):
    o, _ = _one(vx, x0)
    o.run(200)
    builds = _builds(x0, vx, 200)
    lost_at = next(n for n, x in builds if (x >= 0.01 if vx > 0 else x < 0.0))
    nlost, first = o.lost_atoms()
    assert first == lost_at
    # not one build earlier, although the atom was already past the face
    xs_before = [x for n, x in builds if n < lost_at]
    assert all((x < 0.01 if vx > 0 else x >= 0.0) for x in xs_before)
    # it crossed the face between builds: the loss waits for the build
    cross = next(n for n in range(1, 201) if (x0 + n * vx * DT >= 0.01 if vx > 0 else x0 + n * vx * DT < 0.0))
    assert cross < lost_at
    o.close()


def test_atom_exactly_at_the_lower_face_is_kept():
    o, _ = _one(0.0, 0.0)
    o.run(3)
    assert o.lost_atoms() == (0, -1)
    o.close()


@pytest.mark.parametrize("style", ["s", "m"])
def test_shrink_wrapped_face_follows_the_atom(style):
    faces = (1 << 1, (1 << 1) if style == "m" else 0)  # upper x face
    vx, x0 = 1.7, 0.0081
    o, _ = _one(vx, x0, faces)
    o.run(200)
    assert o.lost_atoms() == (0, -1)
    builds = _builds(x0, vx, 200)
    xlast = builds[-1][1]
    lo, hi = o.box()
    small = 1e-4 * 0.01
    expect = xlast + small
    if style == "m":
        expect = max(expect, 0.01)
    assert hi[0] == pytest.approx(expect, rel=0, abs=1e-15)
    assert lo[0] == 0.0 and hi[1] == 0.01  # f faces stay where created
    o.close()


def test_m_face_does_not_move_inside_its_created_position():
    faces = (1 << 1, 1 << 1)
    o, _ = _one(-1.7, 0.0060, faces)  # moving away from the upper face
    o.run(50)
    lo, hi = o.box()
    assert hi[0] == 0.01
    o2, _ = _one(-1.7, 0.0060, (1 << 1, 0))  # the same face as s shrinks onto the atom
    o2.run(50)
    builds = _builds(0.0060, -1.7, 50)
    assert o2.box()[1][0] == pytest.approx(builds[-1][1] + 1e-6, abs=1e-15)
    o.close()
    o2.close()
