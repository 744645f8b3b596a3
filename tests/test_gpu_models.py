"""GPU parity of the sibling contact models (pair_style gran model hooke,
hertz/stiffness, hooke/stiffness; tangential no_history; cohesion sjkr with
tangential_reduce; ktToKnUser; absolute_damping; rolling_friction epsd2) and of
walls and meshes whose fix wall/gran selects its own model, against the CPU
oracle:

  pair sets, flags            exact
  forces / torques / history  1e-10 relative in FP64 (setup and one step)
  mixed mode                  forces and torques 1e-5 relative to the FP64 oracle
  300 steps (chaotic)         kinetic energy and contact count within 1 %
  epsd2 over 10 steps         state and rolling histories 1e-9 relative in FP64
                              (round-off differences grow over the steps)
  epsd2 restart               continuation bitwise equal to the uninterrupted run
  missing properties          the run refuses to start, naming what is missing
"""
import sys

import numpy as np
import pytest

import meshes as msh
import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _stiff(mats, kn, kt, gn, gt):
    """the stiffness models' coefficients, and no elastic properties"""
    m = {k: v for k, v in mats.items() if k not in ("youngs", "poisson", "restitution")}
    m.update(kn=[[kn]], kt=[[kt]], gamman=[[gn]], gammat=[[gt]])
    return m


def _oracle_mats(mats):
    """the oracle's mixing needs elastic values even where no model reads them"""
    if "youngs" in mats:
        return mats
    return dict(mats, youngs=[5e6], poisson=[0.3], restitution=[[0.5]])


def _case(model, mats_fn, wall_model=None, seed=31, n=2000):
    params, mats, P, walls = wl.dense_pack(seed=seed, n=n, rolling=bool(model.get("rolling_cdt", 1)),
                                           speed=0.2)
    params = dict(params, **model)
    mats = mats_fn(mats)
    if wall_model is not None:
        walls = [w[:3] + (dict(wall_model),) for w in walls]
    return params, mats, P, walls


CASES = {
    "hooke_ktToKn_cdt": lambda: _case({"normal": "hooke", "kt_to_kn": 1, "rolling_cdt": 1},
                                      lambda m: dict(m, characteristic_velocity=0.5)),
    "hooke_nohistory_sjkr_reduce": lambda: _case(
        {"normal": "hooke", "tangential": "no_history", "cohesion": "sjkr", "tangential_reduce": 1,
         "rolling_cdt": 0},
        lambda m: dict(m, characteristic_velocity=0.5, cohesion_energy_density=[[3e5]])),
    "hertz_stiffness_cdt": lambda: _case({"normal": "hertz/stiffness", "rolling_cdt": 1},
                                         lambda m: _stiff(m, 4e6, 5e6, 2e5, 1e5)),
    "hooke_stiffness_absolute_notdamp": lambda: _case(
        {"normal": "hooke/stiffness", "absolute_damping": 1, "tangential_damping": 0, "rolling_cdt": 0},
        lambda m: _stiff(m, 400.0, 300.0, 0.01, 0.005)),
    "hertz_sjkr_wall_hooke_nohistory": lambda: _case(
        {"cohesion": "sjkr", "rolling_cdt": 1},
        lambda m: dict(m, characteristic_velocity=0.5, cohesion_energy_density=[[2e5]]),
        wall_model={"normal": "hooke", "tangential": "no_history", "rolling_cdt": 0}),
    "hertz_epsd2": lambda: _case({"rolling_epsd2": 1, "rolling_cdt": 0},
                                 lambda m: dict(m, rolling_friction=[[0.1]])),
    "hooke_epsd2_torsion_wall_cdt": lambda: _case(
        {"normal": "hooke", "rolling_epsd2": 1, "rolling_cdt": 0, "torsion_torque": 1},
        lambda m: dict(m, characteristic_velocity=0.5, rolling_friction=[[0.2]]),
        wall_model={"rolling_cdt": 1}),
}


def build(params, mats, P, walls, precision=pydem.DEM_FP64, meshes=(), max_neighbors=0):
    g = pydem.GpuSystem(params, mats, P, walls, precision=precision, meshes=meshes,
                        max_neighbors=max_neighbors)
    o = orc.OracleSystem(params, _oracle_mats(mats), P, walls, meshes=meshes)
    return g, o


def rel(a, b):
    den = max(np.linalg.norm(b), 1e-300)
    per = np.linalg.norm(a - b, axis=1).max() / max(np.linalg.norm(b, axis=1).max(), 1e-300)
    return np.linalg.norm(a - b) / den, per


def assert_close(name, a, b, tol):
    nw, per = rel(a, b)
    say(f"  {name}: normwise {nw:.3e} per-particle {per:.3e}")
    assert nw <= tol and per <= tol, f"{name}: {nw:.3e} / {per:.3e} > {tol}"


def compare_pairs(g, o, tol=FP64_TOL):
    gt, gf, gh = g.pairs()
    ot, of, oh = o.pairs()
    gm = {(int(a), int(b)): (int(f), h) for (a, b), f, h in zip(gt, gf, gh)}
    om = {(int(a), int(b)): (int(f), h) for (a, b), f, h in zip(ot, of, oh)}
    assert set(gm) == set(om), f"pair sets differ: {len(set(gm) ^ set(om))} pairs"
    scale = max(np.abs(oh).max(), 1e-300)
    for k, (f, h) in om.items():
        assert gm[k][0] == f, f"flag differs for pair {k}"
        assert np.abs(gm[k][1] - h).max() <= tol * scale, f"history differs for pair {k}"
    return len(om), sum(f for f, _ in om.values())


@pytest.mark.parametrize("case", sorted(CASES))
def test_setup_and_one_step_fp64(case):
    say(f"[gpu] contact models, setup + one step fp64, case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls)
    g.setup()
    o.setup()
    npairs, nflag = compare_pairs(g, o)
    say(f"  pairs={npairs} touching={nflag}")
    assert nflag > 500
    gs, os_ = g.state(), o.state()
    assert_close("setup f", gs["f"], os_["f"], FP64_TOL)
    assert_close("setup torque", gs["torque"], os_["torque"], FP64_TOL)
    g.run(1)
    o.run(1)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], FP64_TOL)
    compare_pairs(g, o)
    gw, ow = g.wall_history(), o.wall_history()
    assert np.abs(gw - ow).max() <= FP64_TOL * max(np.abs(ow).max(), 1e-300)
    st = g.stats()
    assert st["ncontacts"] == o.contact_counts()[0]
    assert st["nwall_contacts"] == o.contact_counts()[1] > 0
    g.close()
    o.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_mixed_forces_against_fp64_oracle(case):
    say(f"[gpu] contact models, mixed precision, case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls, precision=pydem.DEM_MIXED)
    g.setup()
    o.setup()
    gs, os_ = g.state(), o.state()
    assert_close("f", gs["f"], os_["f"], MIXED_TOL)
    assert_close("torque", gs["torque"], os_["torque"], MIXED_TOL)
    g.close()
    o.close()


def _ke(st, m):
    return 0.5 * (m[:, None] * st["v"] ** 2).sum()


@pytest.mark.parametrize("case", ["hooke_nohistory_sjkr_reduce", "hertz_stiffness_cdt", "hertz_epsd2"])
def test_three_hundred_steps_statistics(case):
    say(f"[gpu] contact models, 300 steps, case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls, max_neighbors=64)  # cohesion pulls the pack together
    g.run(300)
    o.run(300)
    gs, os_ = g.state(), o.state()
    kg, ko = _ke(gs, P["rmass"]), _ke(os_, P["rmass"])
    cg, co = g.stats()["ncontacts"], o.contact_counts()[0]
    say(f"  KE {kg:.6e} / {ko:.6e}  contacts {cg} / {co}")
    assert abs(kg - ko) <= 0.01 * ko
    assert abs(cg - co) <= 0.01 * co
    g.close()
    o.close()


def test_mesh_wall_with_its_own_model_fp64():
    """A triangle-mesh floor whose fix wall/gran uses hooke + sjkr while the
    pair style is hertz: one step equal to the oracle, mesh history included."""
    say("[gpu] contact models on a mesh floor")
    params, mats, P, _ = wl.dense_pack(seed=41, n=2000, rolling=True, periodic=(0, 0, 0), speed=0.2)
    lo, hi = params["boxlo"], params["boxhi"]
    floor = msh.grid(lo[0] - 0.002, hi[0] + 0.002, lo[1] - 0.002, hi[1] + 0.002, lo[2], 7, 6)
    mats = dict(wl.materials_two(rolling_friction=0.1), characteristic_velocity=0.5,
                cohesion_energy_density=[[2e5, 2e5], [2e5, 2e5]])
    meshes = [(floor, 2, {"normal": "hooke", "cohesion": "sjkr", "rolling_cdt": 1}, None)]
    g, o = build(params, mats, P, [], meshes=meshes)
    g.run(1)
    o.run(1)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], FP64_TOL)
    gt, gh = g.mesh_history()
    ot, oh = o.mesh_history()
    assert np.array_equal(gt[:, :8], ot[:, :8])
    assert (ot >= 0).sum() > 50
    assert np.abs(gh[:, :8] - oh[:, :8]).max() <= FP64_TOL * max(np.abs(oh).max(), 1e-300)
    assert g.stats()["nwall_contacts"] > 50  # the oracle's counter covers plane walls only
    g.close()
    o.close()


@pytest.mark.parametrize("prec", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_mesh_wall_epsd2(prec):
    """rolling_friction epsd2 on a triangle-mesh floor (the rolling torque kept
    in the mesh contact slots beside the shear) under an epsd2 pair style: 40
    steps against the oracle, state and mesh shear histories 1e-9 relative in
    FP64, one step at the mixed tolerance."""
    say(f"[gpu] epsd2 on a mesh floor, precision {prec}")
    params, mats, P, _ = wl.dense_pack(seed=42, n=2000, rolling=False, periodic=(0, 0, 0), speed=0.2)
    params = dict(params, rolling_epsd2=1)
    lo, hi = params["boxlo"], params["boxhi"]
    floor = msh.grid(lo[0] - 0.002, hi[0] + 0.002, lo[1] - 0.002, hi[1] + 0.002, lo[2], 7, 6)
    mats = wl.materials_two(rolling_friction=0.2)
    meshes = [(floor, 2, {"rolling_epsd2": 1}, None)]
    g, o = build(params, mats, P, [], precision=prec, meshes=meshes)
    fp64 = prec == pydem.DEM_FP64
    nsteps = 40 if fp64 else 1
    g.run(nsteps)
    o.run(nsteps)
    tol = 1e-9 if fp64 else MIXED_TOL
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], tol)
    gt, gh = g.mesh_history()
    ot, oh = o.mesh_history()
    assert np.array_equal(gt[:, :8], ot[:, :8])
    assert (ot >= 0).sum() > 50
    assert np.abs(gh[:, :8] - oh[:, :8]).max() <= tol * max(np.abs(oh).max(), 1e-300)
    g.close()
    o.close()


@pytest.mark.parametrize("model,mats_fn,msg", [
    ({"normal": "hertz/stiffness"}, lambda m: m, "needs kn"),
    ({"normal": "hooke"}, lambda m: m, "characteristicVelocity"),
    ({"cohesion": "sjkr"}, lambda m: m, "cohesionEnergyDensity"),
    ({"normal": "hertz"}, lambda m: _stiff(m, 1.0, 1.0, 1.0, 1.0), "youngsModulus"),
])
def test_missing_properties_are_refused(model, mats_fn, msg):
    say(f"[gpu] missing property for {model}")
    params, mats, P, walls = _case(model, mats_fn, n=200)
    g = pydem.GpuSystem(params, mats, P, walls)
    with pytest.raises(pydem.DemError, match=msg):
        g.run(1)
    g.close()


def test_script_with_sibling_models_matches_api(tmp_path):
    """An input script with pair_style gran model hooke tangential no_history
    cohesion sjkr and a hertz/stiffness floor (fix wall/gran with its own
    model) gives the kinetic energy of the same system driven through the C ABI."""
    say("[gpu] script front end with sibling contact models")
    params, mats, P, walls = wl.settling_4k(seed=5, n=800)
    data = tmp_path / "s.data"
    wl.write_data_file(str(data), params, P)
    script = tmp_path / "in.models"
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
fix m5 all property/global characteristicVelocity scalar 1.0
fix m6 all property/global cohesionEnergyDensity peratomtypepair 1 50000
fix m7 all property/global kn peratomtypepair 1 5e6
fix m8 all property/global kt peratomtypepair 1 4e6
fix m9 all property/global gamman peratomtypepair 1 3e4
fix m10 all property/global gammat peratomtypepair 1 2e4
pair_style gran model hooke tangential no_history cohesion sjkr
pair_coeff * *
timestep 0.00001
fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
fix zwall all wall/gran model hertz/stiffness tangential history primitive type 1 zplane 0.0
read_data {data}
fix integr all nve/sphere
thermo_style custom step atoms ke
thermo 500
run 1500
""")
    log = tmp_path / "log.models"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-600:])
    assert rc == 0
    rows = [l.split() for l in text.splitlines() if l.strip() and l.split()[0].isdigit()]
    last = rows[-1]
    assert int(last[0]) == 1500 and int(last[1]) == 800
    params = dict(params, normal="hooke", tangential="no_history", cohesion="sjkr", rolling_cdt=0)
    mats = dict(mats, characteristic_velocity=1.0, cohesion_energy_density=[[5e4]], kn=[[5e6]],
                kt=[[4e6]], gamman=[[3e4]], gammat=[[2e4]])
    walls = [w[:3] + ({"normal": "hertz/stiffness"},) for w in walls]
    g = pydem.GpuSystem(params, mats, P, walls)
    g.run(1500)
    st = g.state()
    ke = 0.5 * (P["rmass"] @ (st["v"] ** 2).sum(1))
    assert float(last[2]) == pytest.approx(ke, rel=1e-6)
    assert g.stats()["nwall_contacts"] > 0
    g.close()


def compare_rolling(g, o, tol):
    """pair rolling histories keyed by the pair's tags, and the plane walls'"""
    gt, _, _ = g.pairs()
    ot, _, _ = o.pairs()
    gr, orr = g.pair_rolling(), o.pair_rolling()
    gm = {(int(a), int(b)): h for (a, b), h in zip(gt, gr)}
    om = {(int(a), int(b)): h for (a, b), h in zip(ot, orr)}
    assert set(gm) == set(om)
    scale = max(np.abs(orr).max(), 1e-300)
    worst = max(np.abs(gm[k] - h).max() for k, h in om.items()) / scale
    say(f"  pair rolling history: max |r| {scale:.3e}, worst {worst:.3e}")
    assert scale > 0 and worst <= tol
    gw, ow = g.wall_rolling(), o.wall_rolling()
    wscale = max(np.abs(ow).max(), 1e-300)
    say(f"  wall rolling history: max |r| {wscale:.3e}")
    assert np.abs(gw - ow).max() <= tol * wscale


@pytest.mark.parametrize("case", ["hertz_epsd2", "hooke_epsd2_torsion_wall_cdt"])
def test_epsd2_ten_steps_fp64(case):
    say(f"[gpu] epsd2 rolling, 10 steps fp64, case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls)
    g.run(10)
    o.run(10)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-9)
    compare_pairs(g, o, tol=1e-9)
    compare_rolling(g, o, 1e-9)
    if case == "hooke_epsd2_torsion_wall_cdt":
        assert not np.any(g.wall_rolling())  # the cdt walls keep no rolling torque
    else:
        assert np.abs(o.wall_rolling()).max() > 0
    g.close()
    o.close()


def test_epsd2_restart_continues_bitwise(tmp_path):
    say("[gpu] epsd2 restart continuation")
    params, mats, P, walls = CASES["hertz_epsd2"]()
    a = pydem.GpuSystem(params, mats, P, walls)
    a.run(40)
    a.run(20)
    b = pydem.GpuSystem(params, mats, P, walls)
    b.run(40)
    f = tmp_path / "e.restart"
    b.write_restart(f)
    b.close()
    c = pydem.GpuSystem(params, mats, P, walls)
    c.read_restart(f)
    c.run(20)
    sa, sc = a.state(), c.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sc[k]), k
    assert np.array_equal(a.pair_rolling(), c.pair_rolling())
    assert np.array_equal(a.wall_rolling(), c.wall_rolling())
    # a system without epsd2 cannot take the file
    d = pydem.GpuSystem(dict(params, rolling_epsd2=0), mats, P, [w[:3] + ({},) for w in walls])
    with pytest.raises(pydem.DemError):
        d.read_restart(f)
    a.close()
    c.close()
    d.close()
