"""Pins of the oracle's compute wall/gran/local (orc_wall_local):

  forces    on a pack resting on a tessellated floor without gravity, the
            rows' forces and torques, added to each row's particle, plus the
            pair contributions of compute pair/gran/local (force +-F, torque
            of the second sphere the first one's times cr_j / cr_i without
            pair rolling friction) rebuild every particle's force and torque of
            the force pass with shearupdate off (orc_compute_forces(0))
  slots     at the end of a step the rows are exactly the particles' occupied
            history slots (the step's own contact pass touched every one), each
            with its slot's history
  geometry  |cp - x|^2 is the row's r^2, the unit normal is (cp - x) / |cp - x|
            and the overlap r - |cp - x| is positive for every row with a force
            (the floor is flat: cp lies on its plane)
  motion    on a floor rotating about the vertical through `origin` with
            period T, the wall velocity at the contact point is the rigid
            rotation's omega x (cp - origin)
  passive   a second call and the state and histories afterwards are identical
"""
import numpy as np

import meshes as msh
import pyoracle as orc
import workloads as wl


def _floor_case(rotate=False, seed=43, n=900):
    params, mats, P, _ = wl.dense_pack(seed=seed, n=n, rolling=False, periodic=(0, 0, 0), speed=0.2)
    params = dict(params)
    params.pop("grav_acc")
    lo, hi = params["boxlo"], params["boxhi"]
    floor = msh.grid(lo[0] - 0.002, hi[0] + 0.002, lo[1] - 0.002, hi[1] + 0.002, lo[2], 7, 6)
    origin = (0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.0)
    rot = (origin, (0.0, 0.0, 1.0), 0.5) if rotate else None
    mats = wl.materials_two(rolling_friction=0.1)
    return params, mats, P, [(floor, 2, {"rolling_cdt": 1}, rot)], origin, lo[2]


def test_wall_rows_and_pairs_rebuild_the_force_pass():
    params, mats, P, meshes, _, _ = _floor_case()
    o = orc.OracleSystem(params, mats, P, [], meshes=meshes)
    o.run(25)
    ptags, pd = o.pair_local()
    wids, wd = o.wall_local()
    o.lib.orc_compute_forces(o.s, 0)
    st = o.state()
    idx = {int(t): k for k, t in enumerate(P["tag"])}
    f = np.zeros_like(st["f"])
    t = np.zeros_like(st["torque"])
    x, r = st["x"], P["radius"]
    for k in range(len(ptags)):
        a, b = idx[int(ptags[k, 0])], idx[int(ptags[k, 1])]
        f[a] += pd[k, 12:15]
        f[b] -= pd[k, 12:15]
        t[a] += pd[k, 15:18]
        dn = r[a] + r[b] - np.linalg.norm(x[a] - x[b])
        if dn > 0:
            t[b] += pd[k, 15:18] * (r[b] - 0.5 * dn) / (r[a] - 0.5 * dn)
    assert len(wids) > 50
    for k in range(len(wids)):
        i = idx[int(wids[k, 2])]
        f[i] += wd[k, 12:15]
        t[i] += wd[k, 15:18]
    fs, ts = np.abs(st["f"]).max(), np.abs(st["torque"]).max()
    assert fs > 0 and ts > 0
    assert np.abs(f - st["f"]).max() <= 1e-12 * fs
    assert np.abs(t - st["torque"]).max() <= 1e-12 * ts
    o.close()


def test_rows_are_the_occupied_slots_with_their_history():
    params, mats, P, meshes, _, z0 = _floor_case(seed=44)
    o = orc.OracleSystem(params, mats, P, [], meshes=meshes)
    o.run(30)
    before = o.state()
    tri, hist = o.mesh_history()
    wids, wd = o.wall_local()
    wids2, wd2 = o.wall_local()
    assert np.array_equal(wids, wids2) and np.array_equal(wd, wd2)
    idx = {int(t): k for k, t in enumerate(P["tag"])}
    got = {}
    for k in range(len(wids)):
        assert wids[k, 0] == 0
        got[(idx[int(wids[k, 2])], int(wids[k, 1]))] = wd[k, 23:26]
    want = {}
    for i in range(tri.shape[0]):
        for s in range(tri.shape[1]):
            if tri[i, s] >= 0:
                want[(i, int(tri[i, s]))] = hist[i, s]
    assert set(got) == set(want) and len(want) > 50
    for key, h in want.items():
        assert np.array_equal(got[key], h)
    assert np.abs(wd[:, 23:26]).max() > 0  # some contacts carry shear
    # geometry of each row
    cp, x, nrm, rsq, rad = wd[:, 0:3], wd[:, 3:6], wd[:, 18:21], wd[:, 21], wd[:, 22]
    d = cp - x
    dist = np.linalg.norm(d, axis=1)
    assert np.allclose(dist * dist, rsq, rtol=1e-12, atol=0)
    assert np.allclose(nrm, d / dist[:, None], rtol=0, atol=1e-12)
    assert np.allclose(cp[:, 2], z0, rtol=0, atol=1e-15)
    assert np.array_equal(rad, P["radius"][[idx[int(t)] for t in wids[:, 2]]])
    forced = np.abs(wd[:, 12:15]).sum(axis=1) > 0
    assert forced.all() and np.all(rad - dist > 0)
    # resting floor: zero wall velocity
    assert np.all(wd[:, 6:9] == 0.0)
    after = o.state()
    for k in ("x", "v", "omega"):
        assert np.array_equal(before[k], after[k])
    tri2, hist2 = o.mesh_history()
    assert np.array_equal(tri, tri2) and np.array_equal(hist, hist2)
    o.close()


def test_rotating_floor_wall_velocity_is_the_rigid_rotation():
    params, mats, P, meshes, origin, _ = _floor_case(rotate=True, seed=45)
    o = orc.OracleSystem(params, mats, P, [], meshes=meshes)
    o.run(20)
    wids, wd = o.wall_local()
    assert len(wids) > 50
    w = np.array([0.0, 0.0, 2 * np.pi / 0.5])
    cp = wd[:, 0:3]
    want = np.cross(w, cp - np.array(origin))
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(wd[:, 6:9] - want).max() <= 1e-9 * scale
    o.close()
