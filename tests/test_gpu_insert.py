"""GPU parity of runs with particle insertion (fix insert/pack): particles
placed by the library's inserter join a system on the GPU through
dem_insert_particles and the same particles join the CPU oracle through
orc_insert, at the first step of a run, after the initial integrate of the
particles already there and before a forced neighbor build.

  one step after an insertion   pair sets bit-exact, flags exact, x / v / omega /
                                f / torque / history within 1e-10 (FP64)
  later steps                   state within 1e-8 (FP64; summation order only)
  mixed precision               forces and torques within 1e-5 of the FP64 oracle
"""
import sys

import numpy as np
import pytest

import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    den = np.linalg.norm(b, axis=1).max()
    return np.linalg.norm(a - b, axis=1).max() / den if den > 0 else np.abs(a).max()


def compare(g, o, tol, keys=("x", "v", "omega", "f", "torque"), pairs=True):
    gs, os_ = g.state(), o.state()
    assert len(gs["x"]) == len(os_["x"])
    for k in keys:
        e = rel(gs[k], os_[k])
        say(f"  {k}: {e:.3e}")
        assert e <= tol, f"{k} differs by {e:.3e}"
    if pairs:
        gt, gf, gh = g.pairs()
        ot, of, oh = o.pairs()
        gm = {(int(a), int(b)): (int(f), h) for (a, b), f, h in zip(gt, gf, gh)}
        om = {(int(a), int(b)): (int(f), h) for (a, b), f, h in zip(ot, of, oh)}
        assert set(gm) == set(om), f"pair sets differ: {len(set(gm) ^ set(om))}"
        hs = max(np.abs(oh).max(), 1e-300)
        for k, (f, h) in om.items():
            assert gm[k][0] == f
            assert np.abs(gm[k][1] - h).max() <= tol * hs
        say(f"  pairs: {len(om)}, flagged {sum(f for f, _ in om.values())}")


BOX = ((-0.02, -0.02, 0.0), (0.02, 0.02, 0.06))


def closed_box():
    """in.packing's setting in a smaller box: six plane walls, Hertz/history,
    CDT rolling friction, gravity so that the packing settles."""
    lo, hi = BOX
    params = {"dt": 1e-5, "skin": 0.0005, "cdf": 1.0, "boxlo": lo, "boxhi": hi, "periodic": (0, 0, 0),
              "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1, "rolling_cdt": 1}
    mats = wl.materials_single(rolling_friction=0.1)
    walls = [(a, (lo if s == 0 else hi)[a], 1) for a in range(3) for s in (0, 1)]
    return params, mats, walls


def inserter(**over):
    kw = dict(templates=[(15485863, 1, 0.001, 2500.0), (15485867, 1, 0.0016, 2500.0)], weights=[0.3, 0.7],
              dist_seed=32452843, seed=32452867, maxattempt=200, all_in=1, volumefraction_region=0.3,
              region=BOX, box=BOX, vel_style=1, v_fluct=(0.05, 0.05, 0.05), v_insert=(0.0, 0.0, -0.2))
    kw.update(over)
    return pydem.Inserter(**kw)


def as_particles(out, tag0):
    m = len(out["x"])
    return {"x": out["x"], "v": out["v"], "omega": out["omega"], "radius": out["radius"],
            "rmass": out["rmass"], "type": out["type"], "tag": np.arange(tag0, tag0 + m, dtype=np.int32)}


EMPTY = {"x": np.zeros((0, 3)), "v": np.zeros((0, 3)), "omega": np.zeros((0, 3)), "radius": np.zeros(0),
         "rmass": np.zeros(0), "type": np.zeros(0, dtype=np.int32), "tag": np.zeros(0, dtype=np.int32)}


def test_insert_into_empty_system_fp64():
    params, mats, walls = closed_box()
    ins = inserter()
    out, req = ins.step()
    P = as_particles(out, 1)
    say(f"inserted {len(out['x'])} of {req}")
    assert len(out["x"]) == req > 1000
    g = pydem.GpuSystem(params, mats, None, walls)
    o = orc.OracleSystem(params, mats, EMPTY, walls)
    rmax = ins.info()["max_radius"]
    g.set_max_radius(rmax)
    o.set_max_radius(rmax)
    g.insert(P)
    o.insert(P)
    g.run(1)
    o.run(1)
    say("step 1 (insertion step):")
    compare(g, o, FP64_TOL)
    g.run_continue(300)
    o.run_continue(300)
    say("step 301:")
    compare(g, o, 1e-8, pairs=False)
    assert g.stats()["ncontacts"] > 0
    g.close()
    o.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_insert_mid_run_with_history_fp64(graphs):
    """A compressed pack with contact history runs 40 steps; the inserter fills
    the empty top of its box (the present particles count and block), the new
    particles join both systems, and the arrays grow past their capacity."""
    params, mats, P, walls = wl.dense_pack(seed=3, n=1500, rolling=True, periodic=(0, 0, 0))
    # side faces shrink-wrapped (boundary m m f): the compressed pack spreads
    # sideways, and an f face would lose those spheres
    params = dict(params, shrink_faces=0b1111, shrink_min=0b1111)
    lo, hi = params["boxlo"], params["boxhi"]
    g = pydem.GpuSystem(params, mats, P, walls)
    if graphs:
        g.set_graph_capture(True)
    o = orc.OracleSystem(params, mats, P, walls)
    g.run(40)
    o.run(40)
    top = ((lo[0], lo[1], 0.7 * hi[2]), hi)
    ins = inserter(templates=[(15485863, 1, 0.0012, 2500.0)], weights=[1.0], region=top, box=(lo, hi),
                   volumefraction_region=0.0, particles_in_region=150, vel_style=0, v_insert=(0.0, 0.0, -0.5))
    gs = g.state()
    out, req = ins.step(gs["x"], P["radius"], P["rmass"])
    say(f"inserted {len(out['x'])} of {req}")
    assert len(out["x"]) > 50
    Q = as_particles(out, 1501)
    g.insert(Q)
    o.insert(Q)
    g.run_continue(1)
    o.run_continue(1)
    say("insertion step:")
    compare(g, o, FP64_TOL)
    g.run_continue(100)
    o.run_continue(100)
    say("100 steps later:")
    compare(g, o, 1e-8, pairs=False)
    g.close()
    o.close()


def test_repeated_insertion_mixed():
    """Three insertions 30 steps apart in mixed precision: one step after each,
    forces and torques within 1e-5 of the FP64 oracle fed the same particles."""
    params, mats, walls = closed_box()
    g = pydem.GpuSystem(params, mats, None, walls, precision=pydem.DEM_MIXED)
    o = orc.OracleSystem(params, mats, EMPTY, walls)
    region = ((BOX[0][0], BOX[0][1], 0.03), BOX[1])
    ins = inserter(region=region, volumefraction_region=0.0, particles_in_region=400)
    g.set_max_radius(ins.info()["max_radius"])
    o.set_max_radius(ins.info()["max_radius"])
    ntag = 1
    x = r = m = None
    for k in range(3):
        if x is not None:
            x = o.state()["x"]
        out, _ = ins.step(x, r, m)
        Q = as_particles(out, ntag)
        ntag += len(out["x"])
        r = Q["radius"] if r is None else np.concatenate([r, Q["radius"]])
        m = Q["rmass"] if m is None else np.concatenate([m, Q["rmass"]])
        x = np.zeros((0, 3))
        g.insert(Q)
        o.insert(Q)
        (g.run if k == 0 else g.run_continue)(1)
        (o.run if k == 0 else o.run_continue)(1)
        say(f"insertion {k}: {len(out['x'])} particles, {g.n} in all")
        compare(g, o, 1e-5, keys=("f", "torque"), pairs=False)
        g.run_continue(29)
        o.run_continue(29)
    assert g.n == o.state()["x"].shape[0] == ntag - 1
    g.close()
    o.close()
