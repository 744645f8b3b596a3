"""GPU parity of compute wall/gran/local (dem_compute_wall_local) against the
oracle's restatement (orc_wall_local) on the same state: the same rows
(mesh, triangle, tag), contact points, positions, wall and particle
velocities, unit normals and r^2 to 1e-12; force, torque and history within
1e-10 relative in FP64 (1e-5 in the mixed mode); the evaluation changes
nothing; and the script's compute wall/gran/local + dump local equals the
library's rows."""
import sys

import numpy as np
import pytest

import meshes as msh
import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _floor_case(seed, n, rotate):
    params, mats, P, _ = wl.dense_pack(seed=seed, n=n, rolling=True, periodic=(0, 0, 0))
    lo, hi = params["boxlo"], params["boxhi"]
    floor = msh.grid(lo[0] - 0.002, hi[0] + 0.002, lo[1] - 0.002, hi[1] + 0.002, lo[2], 7, 6)
    rot = ((0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.0), (0.0, 0.0, 1.0), 0.5) if rotate else None
    mats = wl.materials_two(rolling_friction=0.1)
    return params, mats, P, [], [(floor, 2, {"rolling_cdt": 1}, rot)]


def _sorted(ids, data):
    key = (ids[:, 0].astype(np.int64) * 1_000_000 + ids[:, 1]) * 10_000_000 + ids[:, 2]
    o = np.argsort(key, kind="stable")
    return key[o], data[o]


@pytest.mark.parametrize("rotate", [False, True])
@pytest.mark.parametrize("prec", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_wall_local_matches_oracle(rotate, prec):
    say(f"[gpu] compute wall/gran/local, rotating floor {rotate}, precision {prec}")
    params, mats, P, walls, meshes = _floor_case(71, 2000, rotate)
    g = pydem.GpuSystem(params, mats, P, walls, precision=prec, meshes=meshes)
    o = orc.OracleSystem(params, mats, P, walls, meshes=meshes)
    # mixed mode integrates FP32 forces, so over many steps a contact at the
    # edge of touching may differ from the FP64 oracle's: one step there
    fp64 = prec == pydem.DEM_FP64
    nsteps = 30 if fp64 else 1
    g.run(nsteps)
    o.run(nsteps)
    gi, gd = g.wall_local()
    oi, od = o.wall_local()
    assert gd.shape[1] == od.shape[1] == 26
    gk, gd = _sorted(gi, gd)
    ok, od = _sorted(oi, od)
    assert np.array_equal(gk, ok), "contact rows differ"
    tol = 1e-10 if fp64 else 1e-5
    gtol = 1e-12 if fp64 else 1e-9
    for name, c0, c1, t in (("contact point", 0, 3, gtol), ("x", 3, 6, gtol), ("v_wall", 6, 9, gtol),
                            ("v", 9, 12, gtol), ("force", 12, 15, tol), ("torque", 15, 18, tol),
                            ("normal", 18, 21, gtol), ("rsq", 21, 22, gtol), ("radius", 22, 23, 0.0),
                            ("history", 23, 26, tol)):
        a, b = gd[:, c0:c1], od[:, c0:c1]
        scale = max(np.abs(b).max(), 1e-300)
        err = np.abs(a - b).max() / scale
        say(f"  {name}: max error / scale {err:.3e}")
        if name == "v" and not fp64:
            t = 1e-5  # velocities integrate the FP32 forces
        assert err <= t, name
    say(f"  rows {len(gk)}")
    assert len(gk) > 50
    if rotate:
        assert np.abs(gd[:, 6:9]).max() > 0
    g.close()
    o.close()


def test_wall_local_changes_nothing():
    say("[gpu] compute wall/gran/local leaves the run unchanged")
    params, mats, P, walls, meshes = _floor_case(72, 1500, True)
    a = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    b = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    a.run(20)
    b.run(20)
    i1, d1 = b.wall_local()
    i2, d2 = b.wall_local()
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    a.run_continue(20)
    b.run_continue(20)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sb[k]), k
    ta, ha = a.mesh_history()
    tb, hb = b.mesh_history()
    assert np.array_equal(ta, tb) and np.array_equal(ha, hb)
    a.close()
    b.close()


def test_script_wall_local_dump_equals_the_api(tmp_path):
    """compute wall/gran/local (default columns: pos, vel, id, force, torque,
    history, contactArea) and one with derived columns (force_normal +
    force_tangential = force, delta = r - |cp - x|, contactArea = (r^2 -
    rsq) pi), written by dump local, equal the library's rows of the same
    system driven through the C ABI."""
    say("[gpu] script: compute wall/gran/local + dump local")
    params, mats, P, walls, meshes = _floor_case(73, 1200, False)
    data = tmp_path / "wl.data"
    wl.write_data_file(str(data), params, P, ntypes=2)
    stl = tmp_path / "floor.stl"
    msh.write_stl(str(stl), meshes[0][0])
    lo, hi = params["boxlo"], params["boxhi"]
    script = tmp_path / "in.wl"
    script.write_text(f"""atom_style granular
atom_modify map array
boundary f f f
newton off
communicate single vel yes
units si
region reg block {lo[0] - 0.003!r} {hi[0] + 0.003!r} {lo[1] - 0.003!r} {hi[1] + 0.003!r} {lo[2] - 0.001!r} {hi[2]!r} units box
create_box 2 reg
neighbor {params['skin']!r} bin
neigh_modify delay 0
fix m1 all property/global youngsModulus peratomtype 5.e6 5.e6
fix m2 all property/global poissonsRatio peratomtype 0.45 0.45
fix m3 all property/global coefficientRestitution peratomtypepair 2 0.3 0.3 0.3 0.3
fix m4 all property/global coefficientFriction peratomtypepair 2 0.5 0.4 0.4 0.4
fix m5 all property/global coefficientRollingFriction peratomtypepair 2 0.1 0.1 0.1 0.1
pair_style gran model hertz tangential history rolling_friction cdt
pair_coeff * *
timestep {params['dt']!r}
fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
fix cad all mesh/surface file {stl} type 2
fix walls all wall/gran model hertz tangential history rolling_friction cdt mesh n_meshes 1 meshes cad
read_data {data}
fix integr all nve/sphere
compute cw all wall/gran/local
compute cn all wall/gran/local id force_normal force_tangential contactArea delta
dump d1 all local 100 {tmp_path}/wl_*.dump {' '.join(f'c_cw[{k}]' for k in range(1, 26))}
dump d2 all local 100 {tmp_path}/wn_*.dump index c_cn[1] c_cn[2] c_cn[3] c_cn[4] c_cn[5] c_cn[6] c_cn[7] c_cn[8] c_cn[9] c_cn[10] c_cn[11]
thermo 100
run 200
""")
    log = tmp_path / "log.wl"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-400:])
    assert rc == 0, text
    # boundary f f f: a fixed box, no shrink-wrapped faces
    params = dict(params, boxlo=(lo[0] - 0.003, lo[1] - 0.003, lo[2] - 0.001),
                  boxhi=(hi[0] + 0.003, hi[1] + 0.003, hi[2]), shrink_faces=0, shrink_min=0)
    g = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    g.run(200)
    gi, gd = g.wall_local()
    lines = (tmp_path / "wl_200.dump").read_text().splitlines()
    assert lines[0] == "ITEM: TIMESTEP" and lines[1] == "200" and lines[2] == "ITEM: NUMBER OF ENTRIES"
    assert int(lines[3]) == len(gi) and lines[8].startswith("ITEM: ENTRIES c_cw[1] c_cw[2]")
    rows = np.loadtxt(tmp_path / "wl_200.dump", skiprows=9, ndmin=2)
    assert rows.shape == (len(gi), 25)
    # id columns: mesh index in the fix, triangle, tag
    ks = (rows[:, 12].astype(np.int64) * 1_000_000 + rows[:, 13].astype(np.int64)) * 10_000_000 + rows[:, 14].astype(np.int64)
    rows = rows[np.argsort(ks)]
    gk, gd = _sorted(gi, gd)
    assert np.array_equal(np.sort(ks), gk)

    def close(a, b):  # %g keeps six significant digits
        scale = np.abs(b).max(axis=0, keepdims=True) + 1e-300
        return np.all(np.abs(a - b) <= 2e-5 * np.abs(b) + 1e-6 * scale)

    assert close(rows[:, 0:12], gd[:, 0:12])      # contact point, x, v_wall, v
    assert close(rows[:, 15:21], gd[:, 12:18])    # force, torque
    assert close(rows[:, 21:24], gd[:, 23:26])    # history
    area = (gd[:, 22] ** 2 - gd[:, 21]) * np.pi
    assert close(rows[:, 24:25], area[:, None])
    say(f"  {len(gi)} wall contacts")
    assert len(gi) > 50
    rn = np.loadtxt(tmp_path / "wn_200.dump", skiprows=9, ndmin=2)
    assert rn.shape == (len(gi), 12) and np.array_equal(rn[:, 0], np.arange(1, len(gi) + 1))
    kn = (rn[:, 1].astype(np.int64) * 1_000_000 + rn[:, 2].astype(np.int64)) * 10_000_000 + rn[:, 3].astype(np.int64)
    rn = rn[np.argsort(kn)]
    assert close(rn[:, 4:7] + rn[:, 7:10], gd[:, 12:15])
    nrm = gd[:, 18:21]
    fnm = (gd[:, 12:15] * nrm).sum(axis=1)
    assert close(rn[:, 4:7], nrm * fnm[:, None])
    assert close(rn[:, 10:11], area[:, None])
    assert close(rn[:, 11:12], (gd[:, 22] - np.sqrt(gd[:, 21]))[:, None])
    g.close()
