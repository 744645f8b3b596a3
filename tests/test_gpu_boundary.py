"""GPU checks of the non-periodic boundary faces against the oracle:

  f face   a sphere thrown out of a loose pack through the upper x face:
           the run stops with "Lost atoms: original N current N-1" at the
           step whose neighbor build sees it beyond the face, the oracle's
           first loss step; the state before that step equals the oracle's
  s / m    the same throw with the x faces shrink-wrapped: nothing is lost,
           the box at the last build equals the oracle's (atom extent plus
           1e-4 of the created length; m never inside its created face), and
           the state after the run equals the oracle's
  script   `boundary f f f` stops the run with the "Lost atoms" message;
           `boundary m m m` runs and thermo `vol` is the shrink-wrapped box
"""
import sys

import numpy as np
import pytest

import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _throw(faces=(0, 0)):
    # lattice spacing above the largest diameter: no pair overlaps at the
    # start, so only the thrown sphere reaches an x face (a compressed pack
    # would push its outer layers through them)
    params, mats, P, walls = wl.dense_pack(seed=51, n=1000, rolling=True, periodic=(0, 1, 0), overlap=-0.25)
    P = {k: v.copy() for k, v in P.items()}
    lo, hi = params["boxlo"], params["boxhi"]
    # the particle nearest the upper x face leaves at 3 m/s
    k = int(np.argmax(P["x"][:, 0]))
    P["v"][k] = (3.0, 0.0, 0.0)
    params = dict(params, shrink_faces=faces[0], shrink_min=faces[1])
    return params, mats, P, walls, k


def test_f_face_loses_the_atom_at_the_oracles_step():
    say("[gpu] boundary f: lost atom")
    params, mats, P, walls, k = _throw()
    g = pydem.GpuSystem(params, mats, P, walls)
    o = orc.OracleSystem(params, mats, P, walls)
    o.run(400)
    nlost, first = o.lost_atoms()
    say(f"  oracle: {nlost} lost, first at step {first}")
    assert nlost == 1 and first > 1
    g.run(1)
    step = 1
    err = None
    prev = g.state()
    while step < 400:
        try:
            g.run_continue(1)
        except pydem.DemError as e:
            err = str(e)
            break
        step += 1
        prev = g.state()
    say(f"  gpu: error at step {step + 1}: {err}")
    assert err is not None and f"Lost atoms: original {len(P['tag'])} current {len(P['tag']) - 1}" in err
    assert step + 1 == first
    # the state up to the step before agrees with the oracle's
    o2 = orc.OracleSystem(params, mats, P, walls)
    o2.run(step)
    gs, os_ = prev, o2.state()
    for key in ("x", "v", "omega"):
        d = np.linalg.norm(gs[key] - os_[key], axis=1).max() / np.linalg.norm(os_[key], axis=1).max()
        say(f"  {key} per-particle error {d:.2e}")
        assert d <= 1e-8
    g.close()
    o.close()
    o2.close()


@pytest.mark.parametrize("style", ["s", "m"])
def test_shrink_wrapped_faces_follow_the_atoms(style):
    say(f"[gpu] boundary {style}: shrink-wrapped x faces")
    bits = (1 << 0) | (1 << 1)
    params, mats, P, walls, k = _throw((bits, bits if style == "m" else 0))
    g = pydem.GpuSystem(params, mats, P, walls)
    o = orc.OracleSystem(params, mats, P, walls)
    g.run(300)
    o.run(300)
    assert o.lost_atoms() == (0, -1)
    assert g.stats()["nrebuilds"] == o.nrebuilds()
    glo, ghi = g.box()
    olo, ohi = o.box()
    say(f"  box gpu {glo} {ghi} oracle {olo} {ohi}")
    assert np.allclose(glo, olo, rtol=0, atol=1e-12) and np.allclose(ghi, ohi, rtol=0, atol=1e-12)
    assert ghi[0] > params["boxhi"][0]  # the thrown sphere carried the face out
    gs, os_ = g.state(), o.state()
    for key in ("x", "v", "omega", "f", "torque"):
        d = np.linalg.norm(gs[key] - os_[key]) / max(np.linalg.norm(os_[key]), 1e-300)
        assert d <= 1e-8, key
    g.close()
    o.close()


def _script(tmp_path, boundary, data):
    return f"""atom_style granular
atom_modify map array
boundary {boundary}
newton off
communicate single vel yes
units si
region reg block 0 0.02 0 0.02 0 0.02 units box
create_box 1 reg
neighbor 0.001 bin
neigh_modify delay 0
fix m1 all property/global youngsModulus peratomtype 5.e6
fix m2 all property/global poissonsRatio peratomtype 0.45
fix m3 all property/global coefficientRestitution peratomtypepair 1 0.3
fix m4 all property/global coefficientFriction peratomtypepair 1 0.5
pair_style gran model hertz tangential history
pair_coeff * *
timestep 0.00001
read_data {data}
fix integr all nve/sphere
thermo_style custom step atoms ke vol
thermo_modify lost ignore norm no
thermo 100
run 400
"""


def test_script_boundary_f_and_m(tmp_path):
    say("[gpu] script: boundary f f f loses, boundary m m m wraps")
    n = 8
    r = np.full(n, 0.001)
    x = np.stack([0.002 + 0.0022 * np.arange(n), np.full(n, 0.01), np.full(n, 0.01)], axis=1)
    v = np.zeros((n, 3))
    v[-1] = (2.0, 0.0, 0.0)  # leaves through the upper x face after ~130 steps
    P = {"x": x, "v": v, "omega": np.zeros((n, 3)), "radius": r, "rmass": wl.sphere_mass(r, 2500.0),
         "type": np.ones(n, dtype=np.int32), "tag": np.arange(1, n + 1, dtype=np.int32)}
    params = {"boxlo": (0, 0, 0), "boxhi": (0.02, 0.02, 0.02)}
    data = tmp_path / "d.data"
    wl.write_data_file(str(data), params, P)
    s1 = tmp_path / "in.f"
    s1.write_text(_script(tmp_path, "f f f", data))
    log1 = tmp_path / "log.f"
    rc = pydem.run_script(str(s1), str(log1))
    t1 = log1.read_text()
    say(t1[-400:])
    assert rc != 0 and "Lost atoms: original 8 current 7" in t1
    s2 = tmp_path / "in.m"
    s2.write_text(_script(tmp_path, "m m m", data))
    log2 = tmp_path / "log.m"
    rc = pydem.run_script(str(s2), str(log2))
    t2 = log2.read_text()
    say(t2[-400:])
    assert rc == 0
    rows = [l.split() for l in t2.splitlines() if l.strip() and l.split()[0].isdigit()]
    last = rows[-1]
    assert int(last[0]) == 400 and int(last[1]) == 8
    # the sphere left the created box: the m face followed it out
    assert float(last[3]) > 0.02 ** 3
