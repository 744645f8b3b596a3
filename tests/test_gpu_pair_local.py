"""GPU parity of compute pair/gran/local (dem_compute_pair_local) against the
oracle's restatement (orc_pair_local) on the same state: the same half-list
pairs, positions remapped into the box, velocities, the periodic-image flags
and histories equal; the per-pair force and torque on the first particle
within 1e-10 relative in FP64 (1e-5 in the mixed mode); and the evaluation
changes nothing (the run continues bitwise as without it)."""
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


def _sorted(tags, data):
    key = tags[:, 0].astype(np.int64) * 10_000_000 + tags[:, 1]
    o = np.argsort(key, kind="stable")
    return key[o], data[o]


def _remap(x, params):
    x = x.copy()
    lo = np.array(params["boxlo"], float)
    L = np.array(params["boxhi"], float) - lo
    for d in range(3):
        if params["periodic"][d]:
            for c in (d, 3 + d):
                x[:, c] = np.where(x[:, c] < lo[d], x[:, c] + L[d], x[:, c])
                x[:, c] = np.where(x[:, c] >= lo[d] + L[d], x[:, c] - L[d], x[:, c])
    return x


CASES = {
    "hertz_cdt_walls": lambda: wl.dense_pack(seed=61, n=2000, rolling=True),
    "hertz_epsd2": lambda: (lambda p, m, P, w: (dict(p, rolling_cdt=0, rolling_epsd2=1),
                                                dict(m), P, w))(*wl.dense_pack(seed=62, n=1500, rolling=True)),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("prec", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_pair_local_matches_oracle(case, prec):
    say(f"[gpu] compute pair/gran/local, {case}, precision {prec}")
    params, mats, P, walls = CASES[case]()
    g = pydem.GpuSystem(params, mats, P, walls, precision=prec)
    o = orc.OracleSystem(params, mats, P, walls)
    g.run(30)
    o.run(30)
    gt, gd = g.pair_local()
    ot, od = o.pair_local()
    assert gd.shape[1] == od.shape[1] == 21 + (6 if params.get("rolling_epsd2") else 3)
    gk, gd = _sorted(gt, gd)
    ok, od = _sorted(ot, _remap(od, params))
    assert np.array_equal(gk, ok), "pair sets differ"
    fp64 = prec == pydem.DEM_FP64
    tol = 1e-10 if fp64 else 1e-5
    for name, c0, c1 in (("x", 0, 6), ("v", 6, 12), ("force", 12, 15), ("torque", 15, 18), ("radii", 19, 21), ("history", 21, None)):
        a, b = gd[:, c0:c1], od[:, c0:c1]
        scale = max(np.abs(b).max(), 1e-300)
        err = np.abs(a - b).max() / scale
        say(f"  {name}: max error / scale {err:.3e}")
        assert err <= (1e-12 if name in ("x", "v", "radii") and fp64 else tol), name
    assert np.array_equal(gd[:, 18], od[:, 18])
    touching = np.abs(od[:, 12:15]).sum(axis=1) > 0
    say(f"  pairs {len(gk)}, touching {int(touching.sum())}, periodic {int(od[:, 18].sum())}")
    assert touching.sum() > 200 and (~touching).sum() > 200
    assert np.all(gd[~touching, 12:18] == 0.0)
    g.close()
    o.close()


def test_pair_local_changes_nothing():
    say("[gpu] compute pair/gran/local leaves the run unchanged")
    params, mats, P, walls = wl.dense_pack(seed=63, n=1500, rolling=True)
    a = pydem.GpuSystem(params, mats, P, walls)
    b = pydem.GpuSystem(params, mats, P, walls)
    a.run(20)
    b.run(20)
    b.pair_local()
    a.run_continue(20)
    b.run_continue(20)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sb[k]), k
    ta, _, ha = a.pairs()
    tb, _, hb = b.pairs()
    assert np.array_equal(ta, tb) and np.array_equal(ha, hb)
    a.close()
    b.close()


def test_script_dump_local_equals_the_api(tmp_path):
    """compute pair/gran/local (default columns: pos, vel, id, force, torque,
    history, contactArea) written by dump local every 40 steps equals the
    library's pair records of the same system driven through the C ABI, and a
    keyword list selects and orders the derived columns as the reference does
    (force_normal + force_tangential = force, delta = r_i + r_j - |x_i - x_j|)."""
    say("[gpu] script: compute pair/gran/local + dump local")
    params, mats, P, walls = wl.settling_4k(seed=4, n=1000)
    P = {k: v.copy() for k, v in P.items()}
    # random speeds of 0.3 m/s close the 0.4 mm lattice gaps within the run
    P["v"] = P["v"] + np.random.default_rng(9).normal(0.0, 0.3, P["v"].shape)
    data = tmp_path / "pl.data"
    wl.write_data_file(str(data), params, P)
    script = tmp_path / "in.pl"
    script.write_text(f"""atom_style granular
atom_modify map array
boundary p p f
newton off
communicate single vel yes
units si
region reg block 0 0.1 0 0.1 0 0.2 units box
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
fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
fix zwall all wall/gran model hertz tangential history primitive type 1 zplane 0.0
read_data {data}
fix integr all nve/sphere
compute cp all pair/gran/local
compute cn all pair/gran/local id force_normal force_tangential delta
dump d1 all local 400 {tmp_path}/pl_*.dump {' '.join(f'c_cp[{k}]' for k in range(1, 26))}
dump d2 all local 400 {tmp_path}/pn_*.dump index c_cn[1] c_cn[2] c_cn[4] c_cn[5] c_cn[6] c_cn[7] c_cn[8] c_cn[9] c_cn[10]
thermo 400
run 800
""")
    log = tmp_path / "log.pl"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-400:])
    assert rc == 0, text
    g = pydem.GpuSystem(params, mats, P, walls)
    g.run(800)
    gt, gd = g.pair_local()
    lines = (tmp_path / "pl_800.dump").read_text().splitlines()
    assert lines[0] == "ITEM: TIMESTEP" and lines[1] == "800" and lines[2] == "ITEM: NUMBER OF ENTRIES"
    assert int(lines[3]) == len(gt) and lines[8].startswith("ITEM: ENTRIES c_cp[1] c_cp[2]")
    rows = np.loadtxt(tmp_path / "pl_800.dump", skiprows=9, ndmin=2)
    assert rows.shape == (len(gt), 25)
    key_s = rows[:, 12].astype(np.int64) * 10_000_000 + rows[:, 13].astype(np.int64)
    rows = rows[np.argsort(key_s)]
    gk, gd = _sorted(gt, gd)
    assert np.array_equal(np.sort(key_s), gk)
    # %g keeps six significant digits
    def close(a, b):
        scale = np.abs(b).max(axis=0, keepdims=True) + 1e-300
        return np.all(np.abs(a - b) <= 2e-5 * np.abs(b) + 1e-6 * scale)
    assert close(rows[:, 0:12], gd[:, 0:12])      # pos, vel
    assert close(rows[:, 14:15], gd[:, 18:19])    # periodic flag
    assert close(rows[:, 15:21], gd[:, 12:18])    # force, torque
    assert close(rows[:, 21:24], gd[:, 21:24])    # history
    touching = np.abs(gd[:, 12:15]).sum(axis=1) > 0
    say(f"  {len(gt)} pairs, {int(touching.sum())} touching")
    assert touching.sum() > 50
    # the second compute's derived columns
    rn = np.loadtxt(tmp_path / "pn_800.dump", skiprows=9, ndmin=2)
    assert rn.shape == (len(gt), 10) and np.array_equal(rn[:, 0], np.arange(1, len(gt) + 1))
    kn = rn[:, 1].astype(np.int64) * 10_000_000 + rn[:, 2].astype(np.int64)
    rn = rn[np.argsort(kn)]
    fsum = rn[:, 3:6] + rn[:, 6:9]
    assert close(fsum, gd[:, 12:15])
    xi, xj = gd[:, 0:3], gd[:, 3:6]
    L = np.array(params["boxhi"]) - np.array(params["boxlo"])
    dl = xi - xj
    dl[:, :2] -= np.round(dl[:, :2] / L[:2]) * L[:2]
    delta = gd[:, 19] + gd[:, 20] - np.linalg.norm(dl, axis=1)
    assert close(rn[:, 9:10], delta[:, None])
    g.close()
