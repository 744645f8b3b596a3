"""GPU check of fix mesh/surface ... generate in an input script: a bed of
spheres falling onto a generated square floor (moved into place with the
move keyword) gives the same state as the same system driven through the C
ABI with the generator's triangles (dem_mesh_generate) passed to
dem_add_mesh, and the generated cylinder with closed caps keeps every sphere
inside it."""
import sys

import numpy as np
import pytest

import pydem
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


HEAD = """atom_style granular
atom_modify map array
boundary f f f
newton off
communicate single vel yes
units si
region reg block 0 0.1 0 0.1 -0.01 0.1 units box
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
"""


def _bed(n=600, z0=0.00215):
    """Spheres r = 2 mm on a 4.5 mm lattice; by default the lowest layer starts
    0.05-0.25 mm above z = 0 and lands there within the run."""
    rng = np.random.default_rng(8)
    r = np.full(n, 0.002)
    x = wl.lattice(n, (0.02, 0.02, 0.0), (0.08, 0.08, 0.1), 0.0045, 0.0001, rng, z0=z0)
    P = {"x": x, "v": np.zeros((n, 3)), "omega": np.zeros((n, 3)), "radius": r,
         "rmass": wl.sphere_mass(r, 2500.0), "type": np.ones(n, dtype=np.int32),
         "tag": np.arange(1, n + 1, dtype=np.int32)}
    return P


def test_generated_plane_floor_equals_the_api_mesh(tmp_path):
    say("[gpu] script: fix mesh/surface generate plane")
    P = _bed()
    params = {"dt": 1e-5, "skin": 0.001, "cdf": 1.0, "boxlo": (0.0, 0.0, -0.01), "boxhi": (0.1, 0.1, 0.1),
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1, "rolling_cdt": 0}
    data = tmp_path / "bed.data"
    wl.write_data_file(str(data), params, P)
    script = tmp_path / "in.gen"
    script.write_text(HEAD + f"""fix floor all mesh/surface generate plane 0.1 type 1 move 0.05 0.05 0.0
fix wall all wall/gran model hertz tangential history mesh n_meshes 1 meshes floor
read_data {data}
fix integr all nve/sphere
dump d1 all custom 800 {tmp_path}/gen_*.dump id x y z vx vy vz
thermo_style custom step atoms ke
thermo 400
run 800
""")
    log = tmp_path / "log.gen"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-500:])
    assert rc == 0 and "plane mesh generated: 2 triangels" in text
    tris = pydem.mesh_generate("plane", [0.1]) + np.array([0.05, 0.05, 0.0])
    g = pydem.GpuSystem(params, wl.materials_single(), P, [], meshes=[(tris, 1, None, None)])
    g.run(800)
    st = g.state()
    rows = np.loadtxt(tmp_path / "gen_800.dump", skiprows=9)
    order = np.argsort(P["tag"])
    assert np.array_equal(rows[:, 0].astype(int), P["tag"][order])
    for c, key, k in ((1, "x", 0), (2, "x", 1), (3, "x", 2), (4, "v", 0), (5, "v", 1), (6, "v", 2)):
        assert np.allclose(rows[:, c], st[key][order, k], rtol=1e-5, atol=1e-9), key
    assert g.stats()["nwall_contacts"] > 50  # the bed rests on the generated floor
    g.close()


def test_generated_closed_cylinder_holds_the_bed(tmp_path):
    say("[gpu] script: fix mesh/surface generate cylinder (closed)")
    # starting 2 mm above the bottom cap: near the cap's centre every one of
    # its 48 fan triangles would be a candidate (the limit is 32 per particle)
    P = _bed(400, z0=0.004)
    params = {"boxlo": (0.0, 0.0, -0.01), "boxhi": (0.1, 0.1, 0.1)}
    data = tmp_path / "bed.data"
    wl.write_data_file(str(data), params, P)
    script = tmp_path / "in.cyl"
    script.write_text(HEAD + f"""fix can all mesh/surface generate cylinder 0.045 0.09 48 7 type 1 move 0.05 0.05 0.045
fix wall all wall/gran model hertz tangential history mesh n_meshes 1 meshes can
read_data {data}
fix integr all nve/sphere
dump d1 all custom 1500 {tmp_path}/cyl_*.dump id x y z
thermo_style custom step atoms ke
thermo 500
run 1500
""")
    log = tmp_path / "log.cyl"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-400:])
    assert rc == 0 and "cylinder mesh generated: 192 triangels" in text
    rows = np.loadtxt(tmp_path / "cyl_1500.dump", skiprows=9)
    rad = np.hypot(rows[:, 1] - 0.05, rows[:, 2] - 0.05)
    assert rad.max() < 0.045 and rows[:, 3].min() > 0.0
