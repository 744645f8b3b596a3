"""GPU checks of particle insertion in input scripts: the command sequence of
the reference's packing tutorial (insert/pack by volume fraction into an
empty box with m boundaries) and repeated insertion into a region that
already holds particles (particles_in_region, insert_every, run ... upto)
give the same particles and the same state as the same system driven through
the C ABI (dem_inserter_* places the particles, dem_insert_particles adds
them at the right step)."""
import sys

import numpy as np
import pytest

import pydem
import workloads as wl
from test_script import PACKING_SCRIPT

pytestmark = pytest.mark.gpu

LO, HI = (-0.02, -0.02, 0.0), (0.02, 0.02, 0.06)
TEMPLATES = [(15485863, 1, 0.001, 2500.0), (15485867, 1, 0.0016, 2500.0)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def api_system():
    params = {"dt": 1e-5, "skin": 0.0005, "cdf": 1.0, "boxlo": LO, "boxhi": HI, "periodic": (0, 0, 0),
              "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 0, "rolling_cdt": 0,
              "shrink_faces": 63, "shrink_min": 63}
    walls = [(0, LO[0], 1), (0, HI[0], 1), (1, LO[1], 1), (1, HI[1], 1), (2, LO[2], 1), (2, HI[2], 1)]
    return params, wl.materials_single(), walls


def read_dump(path):
    rows = np.loadtxt(path, skiprows=9, ndmin=2)
    return rows[np.argsort(rows[:, 0])]


def compare_dump(rows, st, tags):
    order = np.argsort(tags)
    assert np.array_equal(rows[:, 0].astype(int), tags[order])
    for c, (key, k) in enumerate((("x", 0), ("x", 1), ("x", 2), ("v", 0), ("v", 1), ("v", 2)), start=1):
        # dump custom writes %g (6 significant digits)
        assert np.allclose(rows[:, c], st[key][order, k], rtol=2e-5, atol=1e-9), key


def test_packing_script_equals_the_api(tmp_path):
    say("[gpu] script: packing insertion into an empty box")
    script = tmp_path / "in.packing"
    script.write_text(PACKING_SCRIPT + f"""fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
dump d1 all custom 201 {tmp_path}/pk_*.dump id x y z vx vy vz radius
run 1
unfix ins
run 200
""")
    log = tmp_path / "log.packing"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-800:])
    assert rc == 0, text
    params, mats, walls = api_system()
    ins = pydem.Inserter(templates=TEMPLATES, weights=[0.3, 0.7], dist_seed=32452843, seed=32452867,
                         maxattempt=200, all_in=1, volumefraction_region=0.25, region=(LO, HI), box=(LO, HI),
                         box_change=1)
    out, req = ins.step()
    n = len(out["x"])
    say(f"API inserter: {n} of {req}")
    assert n == req and f"inserted {n} particle templates" in text
    P = {k: out[k] for k in ("x", "v", "omega", "radius", "rmass", "type")}
    P["tag"] = np.arange(1, n + 1, dtype=np.int32)
    g = pydem.GpuSystem(params, mats, None, walls)
    g.set_max_radius(0.0016)
    g.insert(P)
    g.run(1)
    g.run(200)
    st = g.state()
    rows = read_dump(tmp_path / "pk_201.dump")
    compare_dump(rows, st, P["tag"])
    assert np.allclose(rows[:, 7], P["radius"], rtol=1e-5)
    g.close()


def test_repeated_insertion_script_equals_the_api(tmp_path):
    """insert_every 500 with particles_in_region topping a region up (the
    spheres leave it downwards at 1 m/s between insertions), velocity
    constant, through `run ... upto`: insertions at steps 1, 501, 1001, 1501."""
    say("[gpu] script: repeated insertion, particles_in_region, run upto")
    head = PACKING_SCRIPT.split("variable alphastart")[0].replace(
        "fix pdd1 all particledistribution/discrete 32452843 2 pts1 0.3 pts2 0.7",
        "fix pdd1 all particledistribution/discrete 32452843 1 pts2 1.0")
    script = tmp_path / "in.stream"
    script.write_text(head + f"""region bc block -0.018 0.018 -0.018 0.018 0.045 0.058 units box
fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
fix ins all insert/pack seed 32452867 distributiontemplate pdd1 vel constant 0. 0. -1.0 &
    insert_every 500 overlapcheck yes all_in yes particles_in_region 150 region bc
fix integr all nve/sphere
thermo_style custom step atoms ke
thermo 500
dump d1 all custom 1510 {tmp_path}/st_*.dump id x y z vx vy vz
run 1510 upto
""")
    log = tmp_path / "log.stream"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-1200:])
    assert rc == 0, text
    params, mats, walls = api_system()
    reg = ((-0.018, -0.018, 0.045), (0.018, 0.018, 0.058))
    ins = pydem.Inserter(templates=[TEMPLATES[1]], weights=[1.0], dist_seed=32452843, seed=32452867,
                         all_in=1, particles_in_region=150, region=reg, box=(LO, HI), box_change=1,
                         v_insert=(0.0, 0.0, -1.0))
    g = pydem.GpuSystem(params, mats, None, walls)
    g.set_max_radius(0.0016)
    g.run(0)
    r = np.zeros(0)
    m = np.zeros(0)
    tags = np.zeros(0, dtype=np.int32)
    counts = []
    for s in (1, 501, 1001, 1501):
        st = g.state()
        box = g.box() if len(r) else (LO, HI)
        out, req = ins.step(st["x"], r, m, box=box)
        k = len(out["x"])
        counts.append(k)
        P = {kk: out[kk] for kk in ("x", "v", "omega", "radius", "rmass", "type")}
        P["tag"] = np.arange(len(tags) + 1, len(tags) + k + 1, dtype=np.int32)
        tags = np.concatenate([tags, P["tag"]])
        r = np.concatenate([r, P["radius"]])
        m = np.concatenate([m, P["rmass"]])
        g.insert(P)
        g.run_continue(500 if s < 1501 else 10)
    say(f"API insertions: {counts}")
    assert counts[0] == 150 and sum(counts) > 150
    for c in counts:
        assert f"inserted {c} particle templates" in text
    compare_dump(read_dump(tmp_path / "st_1510.dump"), g.state(), tags)
    g.close()
