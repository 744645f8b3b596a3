"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on
the same seeded inputs.

  neighbor pair sets          bit-exact (same FP64 predicate, minimum image)
  contact flags               exact
  forces / torques / history  1e-10 relative in FP64 mode (norm-wise over all
                              particles, and per particle against the largest
                              force of the system)
  mixed mode forces           1e-5 relative against the FP64 oracle
  multi-step runs (chaotic)   kinetic energy, contact count and packing
                              fraction within 1 %
"""
import os
import sys

import numpy as np
import pytest

import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MIXED_TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def build(params, mats, P, walls, precision=pydem.DEM_FP64):
    g = pydem.GpuSystem(params, mats, P, walls, precision=precision)
    o = orc.OracleSystem(params, mats, P, walls)
    return g, o


def rel_err(a, b):
    """(norm-wise relative error, worst per-particle error / largest magnitude)."""
    den = np.linalg.norm(b)
    if den == 0.0:
        return np.linalg.norm(a), np.abs(a).max()
    per = np.linalg.norm(a - b, axis=1).max() / np.linalg.norm(b, axis=1).max()
    return np.linalg.norm(a - b) / den, per


def assert_close(name, a, b, tol):
    nw, per = rel_err(a, b)
    say(f"  {name}: normwise {nw:.3e} per-particle {per:.3e}")
    assert nw <= tol and per <= tol, f"{name}: {nw:.3e} / {per:.3e} > {tol}"


def pair_maps(tags, flags, hist):
    return {(int(a), int(b)): (int(f), h) for (a, b), f, h in zip(tags, flags, hist)}


def compare_pairs(g, o, hist_tol=FP64_TOL):
    gt, gf, gh = g.pairs()
    ot, of, oh = o.pairs()
    gm, om = pair_maps(gt, gf, gh), pair_maps(ot, of, oh)
    assert len(gm) == len(gt) and len(om) == len(ot)
    assert set(gm) == set(om), f"pair sets differ: {len(set(gm) ^ set(om))} pairs"
    nflag = 0
    hscale = max(np.abs(oh).max(), 1e-300)
    for k, (f, h) in om.items():
        gf_, gh_ = gm[k]
        assert gf_ == f, f"flag differs for pair {k}"
        nflag += f
        assert np.abs(gh_ - h).max() <= hist_tol * hscale, f"history differs for pair {k}"
    return len(om), nflag


def _walled_pack():
    """Closed box with a floor and two side walls whose fix wall/gran models
    differ from the pair style (no rolling friction, no tangential damping)."""
    params, mats, P, walls = wl.dense_pack(seed=14, n=2000, rolling=True, periodic=(0, 0, 0))
    lo, hi = params["boxlo"], params["boxhi"]
    walls = [(2, lo[2], 1, {"rolling_cdt": 0, "tangential_damping": 0}),
             (0, lo[0], 1, {"rolling_cdt": 1, "tangential_damping": 1}),
             (0, hi[0], 1, {"rolling_cdt": 0, "tangential_damping": 1, "limit_force": 1})]
    return params, mats, P, walls


CASES = {
    "walls_own_models": _walled_pack,
    "dense_pbc_xy": lambda: wl.dense_pack(seed=11, n=3000, rolling=True),
    "dense_pbc_xyz_norolling": lambda: wl.dense_pack(seed=12, n=2500, rolling=False,
                                                      periodic=(1, 1, 1)),
    "dense_closed": lambda: wl.dense_pack(seed=13, n=2000, rolling=True, periodic=(0, 0, 0)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_setup_pairs_and_forces_fp64(case):
    say(f"[gpu] setup parity fp64 case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls)
    g.setup()
    o.setup()
    npairs, nflag = compare_pairs(g, o)
    say(f"  pairs={npairs} touching={nflag}")
    assert nflag > 1000
    gs, os_ = g.state(), o.state()
    assert_close("f", gs["f"], os_["f"], FP64_TOL)
    assert_close("torque", gs["torque"], os_["torque"], FP64_TOL)
    st = g.stats()
    assert st["npairs"] == npairs
    assert st["ncontacts"] == o.contact_counts()[0]
    assert st["nwall_contacts"] == o.contact_counts()[1]
    g.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_one_step_state_history_fp64(case):
    say(f"[gpu] one-step parity fp64 case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls)
    g.run(1)
    o.run(1)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], FP64_TOL)
    compare_pairs(g, o)
    gw, ow = g.wall_history(), o.wall_history()
    assert np.abs(gw - ow).max() <= FP64_TOL * max(np.abs(ow).max(), 1e-300)
    assert np.abs(ow).max() > 0
    g.close()


def test_history_carried_across_rebuilds_fp64():
    """Twenty steps with a forced rebuild (the second run's setup) and the
    displacement-triggered rebuilds in between: history and pair sets stay
    equal to the oracle's."""
    say("[gpu] history carry across rebuilds")
    params, mats, P, walls = wl.dense_pack(seed=21, n=2000, rolling=True, speed=0.5)
    params = dict(params, skin=0.00002)  # tiny skin: rebuild every few steps
    g, o = build(params, mats, P, walls)
    g.run(10)
    o.run(10)
    g.run(10)
    o.run(10)
    say(f"  rebuilds gpu={g.stats()['nrebuilds']} oracle={o.nrebuilds()}")
    assert g.stats()["nrebuilds"] == o.nrebuilds()
    assert o.nrebuilds() >= 4
    n, nflag = compare_pairs(g, o, hist_tol=1e-8)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-8)
    g.close()


STIFF = {"kn": [[4.0e6]], "kt": [[5.0e6]], "gamman": [[2.0e5]], "gammat": [[1.0e5]]}


@pytest.mark.parametrize("normal", ["hertz", "hertz/stiffness"])
def test_contact_distance_factor_band_fp64(normal):
    """neigh_modify contact_distance_factor 1.1 on an agitated pack with a
    small skin (builds every few steps): pairs leaving contact pass through
    the surfacesClose band, where the shear is zeroed and the contact flag
    follows the normal model's rule (hertz keeps it, hertz/stiffness only for
    a pair carried over at the last build). Pair sets, flags, histories,
    build counts and state equal the oracle's after 40 steps."""
    say(f"[gpu] contact_distance_factor 1.1 band, {normal}")
    params, mats, P, walls = wl.dense_pack(seed=23, n=2000, rolling=normal == "hertz", speed=0.5)
    params = dict(params, skin=0.00004, cdf=1.1, normal=normal)
    if normal == "hertz/stiffness":
        mats = dict(mats, **STIFF)
    g, o = build(params, mats, P, walls)
    g.run(1)
    o.run(1)
    compare_pairs(g, o)
    g.run_continue(39)
    o.run_continue(39)
    say(f"  rebuilds gpu={g.stats()['nrebuilds']} oracle={o.nrebuilds()}")
    assert g.stats()["nrebuilds"] == o.nrebuilds() >= 3
    n, nflag = compare_pairs(g, o, hist_tol=1e-8)
    ot, of, oh = o.pairs()
    # pairs in the band: apart, within 1.1 radsum
    st = o.state()
    idx = {int(t): k for k, t in enumerate(P["tag"])}
    a = np.array([idx[int(t)] for t in ot[:, 0]])
    b = np.array([idx[int(t)] for t in ot[:, 1]])
    d = st["x"][a] - st["x"][b]
    L = np.array(params["boxhi"]) - np.array(params["boxlo"])
    per = np.array(params["periodic"], dtype=bool)
    d[:, per] -= np.round(d[:, per] / L[per]) * L[per]
    rs = P["radius"][a] + P["radius"][b]
    dist = np.linalg.norm(d, axis=1)
    band = (dist >= rs) & (dist < 1.1 * rs)
    say(f"  pairs {n}, flagged {nflag}, in the band {int(band.sum())} "
        f"(flagged {int(of[band].sum())}, max |shear| there {np.abs(oh[band]).max():.1e})")
    assert band.sum() > 100
    assert np.abs(oh[band]).max() == 0.0
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-8)
    g.close()


@pytest.mark.parametrize("every,check", [(3, True), (2, False), (5, False)])
def test_neigh_modify_every_check_fp64(every, check):
    """neigh_modify every N check yes|no: the same build schedule as the
    oracle (build counts equal, across a `run pre no` boundary) and the same
    state and history to 1e-8 after 60 steps."""
    say(f"[gpu] neigh_modify every {every} check {check}")
    params, mats, P, walls = wl.dense_pack(seed=22, n=1500, rolling=True, speed=0.5)
    params = dict(params, skin=0.00004, neigh_every=every, neigh_check=check)
    g, o = build(params, mats, P, walls)
    g.run(37)
    o.run(37)
    g.run_continue(23)
    o.run_continue(23)
    say(f"  rebuilds gpu={g.stats()['nrebuilds']} oracle={o.nrebuilds()}")
    assert g.stats()["nrebuilds"] == o.nrebuilds()
    if not check:
        assert o.nrebuilds() == 1 + 60 // every
    compare_pairs(g, o, hist_tol=1e-8)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-8)
    g.close()


def test_step_shapes_are_bitwise_equal():
    """Every tile shape of the step kernel (threads per block, neighbor loads
    in flight, persistent grid) and an autotuned run give bitwise the same
    state: they only change which thread walks which particles."""
    say("[gpu] step kernel shapes")
    params, mats, P, walls = wl.dense_pack(seed=33, n=3000, rolling=True, speed=0.3)
    ref = {}
    shapes = [(v, f) for v in range(11) for f in (0, 4)] + [(-1, 0)]
    for v, f in shapes:
        for prec in (pydem.DEM_FP64, pydem.DEM_MIXED):
            g = pydem.GpuSystem(params, mats, P, walls, precision=prec)
            g.set_step_shape(v, f)
            g.run(160 if v == -1 else 40)
            if v == -1:
                chosen = g.step_shape()
                say(f"  autotuned: variant {chosen[0]} persist {chosen[1]}, "
                    f"{sum(1 for c in chosen[2] if c[2] > 0)} candidates timed")
                g.close()
                continue
            st = g.state()
            ref.setdefault(prec, st)
            for k in ("x", "v", "omega", "f", "torque"):
                assert np.array_equal(st[k], ref[prec][k]), (v, f, prec, k)
            g.close()
    # the autotuned run against the default shape over the same 160 steps
    a = pydem.GpuSystem(params, mats, P, walls)
    a.run(160)
    b = pydem.GpuSystem(params, mats, P, walls)
    b.set_step_shape(-1, 0)
    b.run(160)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_sibling_models_wave_compaction_is_bitwise(prec):
    """The sibling contact models' kernel (hooke + no_history + sjkr pairs,
    a hertz/stiffness + epsd2 floor) in its per-lane, wavefront-compacted and row-queued
    forms: bitwise the same state and pair histories after 60 steps."""
    say("[gpu] sibling models: per-lane vs wave-compacted step kernel")
    params, mats, P, walls = wl.dense_pack(seed=34, n=2500, rolling=False, speed=0.3)
    params = dict(params, normal="hooke", tangential="no_history", cohesion="sjkr")
    mats = dict(mats, characteristic_velocity=1.0, cohesion_energy_density=[[3.0e4]],
                rolling_friction=[[0.1]], **STIFF)
    walls = [(2, params["boxlo"][2], 1, {"normal": "hertz/stiffness", "rolling_epsd2": 1})]
    out = []
    for v in (0, 5, 7):
        g = pydem.GpuSystem(params, mats, P, walls, precision=prec)
        g.set_step_shape(v, 0)
        g.run(60)
        out.append((g.state(), g.pairs(), g.wall_history()))
        g.close()
    for o in out[1:]:
        for k in ("x", "v", "omega", "f", "torque"):
            assert np.array_equal(out[0][0][k], o[0][k]), k
        for a, b in zip(out[0][1], o[1]):
            assert np.array_equal(a, b)
        assert np.array_equal(out[0][2], o[2])


def test_run_continue_is_bitwise_one_run():
    say("[gpu] run + run_continue == one run (bitwise)")
    params, mats, P, walls = wl.dense_pack(seed=31, n=1500, rolling=True, speed=0.3)
    a = pydem.GpuSystem(params, mats, P, walls)
    a.run(40)
    b = pydem.GpuSystem(params, mats, P, walls)
    b.run(15)
    b.run_continue(25)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sb[k]), k
    # and the library is deterministic run to run
    c = pydem.GpuSystem(params, mats, P, walls)
    c.run(40)
    sc = c.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sc[k]), k
    for s in (a, b, c):
        s.close()


@pytest.mark.parametrize("case", sorted(CASES))
def test_mixed_precision_forces_vs_fp64_oracle(case):
    say(f"[gpu] mixed precision vs fp64 oracle case={case}")
    params, mats, P, walls = CASES[case]()
    g, o = build(params, mats, P, walls, precision=pydem.DEM_MIXED)
    g.run(1)
    o.run(1)
    gs, os_ = g.state(), o.state()
    for k in ("f", "torque"):
        assert_close(k + " (mixed)", gs[k], os_[k], MIXED_TOL)
    # x and v integrate in FP64: their one-step change carries only the force error
    assert_close("dv (mixed)", gs["v"] - P["v"], os_["v"] - P["v"], MIXED_TOL)
    assert_close("dx (mixed)", gs["x"] - P["x"], os_["x"] - P["x"], MIXED_TOL)
    gt, gf, gh = g.pairs()
    ot, of, oh = o.pairs()
    assert {tuple(t) for t in gt} == {tuple(t) for t in ot}
    g.close()


def _stats(state, P, params):
    ke = 0.5 * (P["rmass"] @ (state["v"] ** 2).sum(1))
    area = (params["boxhi"][0] - params["boxlo"][0]) * (params["boxhi"][1] - params["boxlo"][1])
    vol = (4.0 / 3.0 * np.pi * P["radius"] ** 3).sum()
    phi = vol / (area * 2.0 * state["x"][:, 2].mean())
    return ke, phi


@pytest.mark.parametrize("precision", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_settling_statistics_within_one_percent(precision):
    """Config-0 settling (4,000 spheres) over 3,000 steps: the bed hits the
    floor after ~600 steps, so contacts, walls and rebuilds are all active."""
    say(f"[gpu] settling statistics precision={precision}")
    params, mats, P, walls = wl.settling_4k(seed=0)
    g, o = build(params, mats, P, walls, precision=precision)
    checkpoints = [1000, 1000, 1000]
    g.run(0)
    o.run(0)
    for k, n in enumerate(checkpoints):
        g.run_continue(n)
        o.run_continue(n)
        gs, os_ = g.state(), o.state()
        gke, gphi = _stats(gs, P, params)
        oke, ophi = _stats(os_, P, params)
        gc = g.stats()["ncontacts"] + g.stats()["nwall_contacts"]
        oc = sum(o.contact_counts())
        say(f"  step {sum(checkpoints[:k+1])}: KE {gke:.6e}/{oke:.6e} phi {gphi:.5f}/{ophi:.5f} "
            f"contacts {gc}/{oc}")
        assert gke == pytest.approx(oke, rel=0.01)
        assert gphi == pytest.approx(ophi, rel=0.01)
        assert gc == pytest.approx(oc, rel=0.01, abs=3)
    assert oc > 1000
    g.close()


def test_script_front_end_matches_api(tmp_path):
    """An unchanged-style LIGGGHTS input script run by dem_run_script gives the
    same kinetic energy as the same system driven through the C ABI."""
    say("[gpu] script front end")
    params, mats, P, walls = wl.settling_4k(seed=3, n=1000)
    data = tmp_path / "c0.data"
    wl.write_data_file(str(data), params, P)
    script = tmp_path / "in.c0"
    script.write_text(f"""# 1,000-sphere settling on a zplane floor
atom_style granular
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
thermo_style custom step atoms ke
thermo 500
run 1500
""")
    log = tmp_path / "log.c0"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-600:])
    assert rc == 0
    rows = [l.split() for l in text.splitlines() if l.strip() and l.split()[0].isdigit()]
    last = rows[-1]
    assert int(last[0]) == 1500 and int(last[1]) == 1000
    g = pydem.GpuSystem(params, mats, P, walls)
    g.run(1500)
    st = g.state()
    ke = 0.5 * (P["rmass"] @ (st["v"] ** 2).sum(1))
    assert float(last[2]) == pytest.approx(ke, rel=1e-6)
    g.close()


# --------------------------------------------------------------------------
# triangle-mesh walls
# --------------------------------------------------------------------------
import meshes as msh  # noqa: E402


@pytest.mark.parametrize("name", ["grid", "cube", "drum"])
def test_mesh_topology_matches_oracle(name):
    say(f"[gpu] mesh topology {name}")
    tris = {"grid": msh.grid(-0.05, 0.05, -0.05, 0.05, 0.0, 5, 4),
            "cube": msh.cube((-0.02, -0.02, -0.02), (0.02, 0.02, 0.02)),
            "drum": msh.drum(0.15, 0.1, 64, 16)}[name]
    params = {"dt": 1e-5, "skin": 0.001, "boxlo": (-0.2, -0.2, -0.2), "boxhi": (0.2, 0.2, 0.2),
              "periodic": (0, 0, 0)}
    P = {"x": np.array([[0.001, 0.002, 0.003]]), "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)),
         "radius": np.array([0.001]), "rmass": np.array([1e-5]), "type": np.array([1]),
         "tag": np.array([1])}
    mats = wl.materials_two()
    g = pydem.GpuSystem(params, mats, P, meshes=[(tris, 2, None, None)])
    o = orc.OracleSystem(params, mats, P, meshes=[(tris, 2, None, None)])
    for a, b in zip(g.mesh_topology(), o.mesh_topology()):
        assert np.array_equal(a, b)
    g.close()


def _mesh_floor_pack(seed=41, n=2000, moving=False):
    """moving: False, True (fix move/mesh rotate about the vertical axis),
    "linear" (a floor moving up and sideways) or "wiggle" (a vertical shaker)"""
    params, mats, P, walls = wl.dense_pack(seed=seed, n=n, rolling=True, periodic=(0, 0, 0))
    lo, hi = params["boxlo"], params["boxhi"]
    floor = msh.grid(lo[0] - 0.002, hi[0] + 0.002, lo[1] - 0.002, hi[1] + 0.002, lo[2], 7, 6)
    if moving == "linear":
        rot = {"linear": (0.2, -0.1, 0.05)}
    elif moving == "wiggle":
        rot = {"wiggle": ((0.0001, 0.0, 0.0003), 0.02)}
    elif moving == "viblin":
        rot = {"viblin": ((0.2, 0.0, 1.0), (0.0002, 0.0001), (0.0, 1.0), (0.02, 0.007))}
    elif moving == "vibrot":
        rot = {"vibrot": ((0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.0), (0.0, 0.3, 1.0), (0.4,), (0.2,),
                          (0.01,))}
    elif moving:
        rot = ((0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.0), (0.0, 0.0, 1.0), 0.5)
    else:
        rot = None
    mats = wl.materials_two(rolling_friction=0.1)
    return params, mats, P, [], [(floor, 2, {"rolling_cdt": 1}, rot)]


def _build_mesh_case(case, precision=pydem.DEM_FP64):
    params, mats, P, walls, meshes = case
    g = pydem.GpuSystem(params, mats, P, walls, precision=precision, meshes=meshes)
    o = orc.OracleSystem(params, mats, P, walls, meshes=meshes)
    return g, o


def _compare_mesh_history(g, o, tol):
    gt, gh = g.mesh_history()
    ot, oh = o.mesh_history()
    # slot order may differ only where the reference's order is irrelevant; here
    # both follow the same slot rules, so compare slot by slot
    assert np.array_equal(gt[:, :8], ot[:, :8])
    assert (ot[:, 8:] == -1).all()
    scale = max(np.abs(oh).max(), 1e-300)
    assert np.abs(gh[:, :8] - oh[:, :8]).max() <= tol * scale
    return int((ot >= 0).sum())


@pytest.mark.parametrize("moving", [False, True, "linear", "wiggle", "viblin", "vibrot"])
def test_mesh_floor_one_step_fp64(moving):
    say(f"[gpu] mesh floor one step moving={moving}")
    g, o = _build_mesh_case(_mesh_floor_pack(moving=moving))
    g.run(1)
    o.run(1)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], FP64_TOL)
    nslots = _compare_mesh_history(g, o, FP64_TOL)
    say(f"  mesh contacts with history: {nslots}")
    assert nslots > 50
    gn, gv = g.mesh_nodes()
    on, ov = o.mesh_nodes()
    assert np.array_equal(gn, on) and np.array_equal(gv, ov)
    g.close()


def test_mesh_floor_many_steps_fp64():
    """200 steps on a static tessellated floor: particles slide across
    triangle boundaries (history copied between coplanar faces) and the
    neighbor list is rebuilt several times."""
    say("[gpu] mesh floor 200 steps")
    case = _mesh_floor_pack(seed=42, n=1500)
    params = dict(case[0], skin=0.00005)
    case = (params,) + case[1:]
    g, o = _build_mesh_case(case)
    g.run(200)
    o.run(200)
    say(f"  rebuilds gpu={g.stats()['nrebuilds']} oracle={o.nrebuilds()}")
    assert g.stats()["nrebuilds"] == o.nrebuilds() >= 3
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-8)
    _compare_mesh_history(g, o, 1e-8)
    g.close()


def test_mesh_floor_contact_distance_factor_fp64():
    """A tessellated floor with contact_distance_factor 1.2: particles
    rising off the floor within 0.2 r keep their slot with zeroed shear
    (surfacesClose through compute_force), as the oracle's slots do."""
    say("[gpu] mesh floor, contact_distance_factor 1.2")
    case = _mesh_floor_pack(seed=43, n=1500)
    # side faces shrink-wrapped: the compressed pack ejects a few spheres sideways
    params = dict(case[0], skin=0.00005, cdf=1.2, shrink_faces=0b1111, shrink_min=0b1111)
    P = dict(case[2])
    P["v"] = P["v"] + np.array([0.0, 0.0, 0.05])  # lift off the floor
    case = (params, case[1], P) + case[3:]
    g, o = _build_mesh_case(case)
    g.run(120)
    o.run(120)
    assert g.stats()["nrebuilds"] == o.nrebuilds()
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-8)
    nslots = _compare_mesh_history(g, o, 1e-8)
    ot, oh = o.mesh_history()
    zero = ((ot[:, :8] >= 0) & (np.abs(oh[:, :8]).max(axis=-1) == 0.0)).sum()
    say(f"  mesh slots {nslots}, of them with zero shear {int(zero)}")
    g.close()


@pytest.mark.parametrize("motion", ["linear", "wiggle", "viblin", "vibrot"])
def test_translating_floor_many_steps_fp64(motion):
    """fix move/mesh linear / wiggle / viblin / vibrot under a packed bed, 500 steps (two runs,
    the second continuing the mover's time): node positions and velocities
    bitwise equal to the oracle, the same rebuilds, state and mesh history
    within 1e-8."""
    say(f"[gpu] translating floor ({motion}) 500 steps")
    g, o = _build_mesh_case(_mesh_floor_pack(seed=43, n=1500, moving=motion))
    g.run(300)
    o.run(300)
    assert _compare_mesh_history(g, o, 1e-8) >= 5
    g.run(200)
    o.run(200)
    gn, gv = g.mesh_nodes()
    on, ov = o.mesh_nodes()
    assert np.array_equal(gn, on) and np.array_equal(gv, ov)
    say(f"  rebuilds gpu={g.stats()['nrebuilds']} oracle={o.nrebuilds()} forced={o.mesh_forced_rebuilds()}")
    assert g.stats()["nrebuilds"] == o.nrebuilds() >= 3
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert_close(k, gs[k], os_[k], 1e-8)
    _compare_mesh_history(g, o, 1e-8)
    g.close()


def _small_drum():
    params, mats, P, walls, meshes = wl.drum(seed=5, radius=0.03, length=0.02, n_az=16, n_ax=4,
                                            period=0.1, skin=0.0005)
    P = dict(P)
    P["v"] = P["v"].copy()
    P["v"][:, 2] = -0.5
    return params, mats, P, walls, meshes


@pytest.mark.parametrize("precision", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_rotating_drum(precision):
    """A small rotating drum (160 triangles, ~2,000 spheres thrown at its
    wall): mesh rotation bitwise equal to the oracle, the same number of
    mesh-forced rebuilds, and the state after 600 steps equal to 1e-8 (FP64)
    or kinetic energy and wall-contact count within 1% (mixed)."""
    say(f"[gpu] rotating drum precision={precision}")
    g, o = _build_mesh_case(_small_drum(), precision)
    g.run(600)
    o.run(600)
    gn, gv = g.mesh_nodes()
    on, ov = o.mesh_nodes()
    assert np.array_equal(gn, on) and np.array_equal(gv, ov)
    st = g.stats()
    say(f"  rebuilds gpu={st['nrebuilds']} oracle={o.nrebuilds()} forced={o.mesh_forced_rebuilds()} "
        f"wall contacts gpu={st['nwall_contacts']}")
    assert o.mesh_forced_rebuilds() >= 3
    gs, os_ = g.state(), o.state()
    P = _small_drum()[2]
    if precision == pydem.DEM_FP64:
        assert st["nrebuilds"] == o.nrebuilds()
        for k in ("x", "v", "omega"):
            assert_close(k, gs[k], os_[k], 1e-8)
        n = _compare_mesh_history(g, o, 1e-8)
        assert n >= 30
    else:
        gke = 0.5 * (P["rmass"] @ (gs["v"] ** 2).sum(1))
        oke = 0.5 * (P["rmass"] @ (os_["v"] ** 2).sum(1))
        assert gke == pytest.approx(oke, rel=0.01)
        gt, _ = g.mesh_history()
        ot, _ = o.mesh_history()
        assert (gt >= 0).sum() == pytest.approx((ot >= 0).sum(), rel=0.01, abs=3)
    g.close()


def test_script_rotating_drum_matches_api(tmp_path):
    """The rotating drum written as an input script (fix mesh/surface from an
    STL file, fix wall/gran ... mesh, fix move/mesh ... rotate) runs to the
    same kinetic energy as the system built through the C ABI."""
    say("[gpu] script rotating drum")
    params, mats, P, walls, meshes = _small_drum()
    data = tmp_path / "drum.data"
    wl.write_data_file(str(data), params, P, ntypes=2)
    stl = tmp_path / "drum.stl"
    msh.write_stl(str(stl), meshes[0][0])
    lo, hi = params["boxlo"], params["boxhi"]
    script = tmp_path / "in.drum"
    script.write_text(f"""atom_style granular
atom_modify map array
boundary f f f
newton off
communicate single vel yes
units si
region reg block {lo[0]!r} {hi[0]!r} {lo[1]!r} {hi[1]!r} {lo[2]!r} {hi[2]!r} units box
create_box 2 reg
neighbor {params['skin']!r} bin
neigh_modify delay 0
fix m1 all property/global youngsModulus peratomtype 5.e6 5.e6
fix m2 all property/global poissonsRatio peratomtype 0.45 0.45
fix m3 all property/global coefficientRestitution peratomtypepair 2 0.3 0.3 0.3 0.3
fix m4 all property/global coefficientFriction peratomtypepair 2 0.5 0.4 0.4 0.4
pair_style gran model hertz tangential history
pair_coeff * *
timestep {params['dt']!r}
fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
fix cad all mesh/surface file {stl} type 2
fix walls all wall/gran model hertz tangential history mesh n_meshes 1 meshes cad
fix rot all move/mesh mesh cad rotate origin 0. 0. 0. axis 0. 1. 0. period {meshes[0][3][2]!r}
read_data {data}
fix integr all nve/sphere
thermo_style custom step atoms ke
thermo 200
run 400
""")
    log = tmp_path / "log.drum"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-600:])
    assert rc == 0
    rows = [l.split() for l in text.splitlines() if l.strip() and l.split()[0].isdigit()]
    last = rows[-1]
    assert int(last[0]) == 400 and int(last[1]) == len(P["radius"])
    g = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    g.run(400)
    st = g.state()
    ke = 0.5 * (P["rmass"] @ (st["v"] ** 2).sum(1))
    assert float(last[2]) == pytest.approx(ke, rel=1e-6)
    g.close()


def test_script_wiggling_floor_matches_api(tmp_path):
    """A shaker written as an input script (fix mesh/surface/stress floor from
    an STL file, fix move/mesh ... wiggle amplitude ... period ..., two runs so
    the mover's time runs on; thermo with compute erotate/sphere and the
    floor's f_cad[3] / f_cad[5]) gives the kinetic and rotational energy and
    the mesh force and torque of the same system built through the C ABI;
    thermo vol / time give the box volume and step * dt, and dump mesh/stl
    writes the moved floor (ASCII and binary) at steps 0, 150 and 300."""
    say("[gpu] script wiggling floor")
    params, mats, P, walls, meshes = _mesh_floor_pack(seed=45, n=1200, moving="wiggle")
    data = tmp_path / "shake.data"
    wl.write_data_file(str(data), params, P, ntypes=2)
    stl = tmp_path / "floor.stl"
    msh.write_stl(str(stl), meshes[0][0])
    lo, hi = params["boxlo"], params["boxhi"]
    amp, period = meshes[0][3]["wiggle"]
    script = tmp_path / "in.shake"
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
fix cad all mesh/surface/stress file {stl} type 2 reference_point 0.01 0.0 0.0
fix walls all wall/gran model hertz tangential history rolling_friction cdt mesh n_meshes 1 meshes cad
fix shake all move/mesh mesh cad wiggle amplitude {amp[0]!r} {amp[1]!r} {amp[2]!r} period {period!r}
read_data {data}
fix integr all nve/sphere
compute er all erotate/sphere
thermo_style custom step atoms ke c_er f_cad[3] f_cad[5] vol time
thermo 150
dump stl all mesh/stl 150 {tmp_path}/floor*.stl
dump stlb all mesh/stl 300 {tmp_path}/floorb*.stl binary cad
run 150
run 150
""")
    log = tmp_path / "log.shake"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-600:])
    assert rc == 0
    rows = [l.split() for l in text.splitlines() if l.strip() and l.split()[0].isdigit()]
    last = rows[-1]
    assert int(last[0]) == 300 and int(last[1]) == len(P["radius"])
    params = dict(params, boxlo=(lo[0] - 0.003, lo[1] - 0.003, lo[2] - 0.001),
                  boxhi=(hi[0] + 0.003, hi[1] + 0.003, hi[2]))
    meshes = [tuple(meshes[0][:4]) + ((0.01, 0.0, 0.0),)]
    g = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    g.run(150)
    g.run(150)
    st = g.state()
    ke = 0.5 * (P["rmass"] @ (st["v"] ** 2).sum(1))
    er = 0.5 * (0.4 * P["rmass"] * P["radius"] ** 2 @ (st["omega"] ** 2).sum(1))
    F, T = g.mesh_force()
    assert float(last[2]) == pytest.approx(ke, rel=1e-6)
    assert float(last[3]) == pytest.approx(er, rel=1e-6)
    # the mesh totals are atomically summed, so equal to rounding only
    assert float(last[4]) == pytest.approx(F[2], rel=1e-6) and F[2] < 0
    assert float(last[5]) == pytest.approx(T[1], rel=1e-6)
    vol = (hi[0] - lo[0] + 0.006) * (hi[1] - lo[1] + 0.006) * (hi[2] - lo[2] + 0.001)
    assert float(last[6]) == pytest.approx(vol, rel=1e-7)
    assert float(last[7]) == pytest.approx(300 * params["dt"], rel=1e-7)
    nodes, _ = g.mesh_nodes()
    for stepno in (0, 150, 300):
        assert (tmp_path / f"floor{stepno}.stl").exists()
    lines = (tmp_path / "floor300.stl").read_text().splitlines()
    assert lines[0] == "solid LIGGGHTS_STL_EXPORT" and lines[-1] == "endsolid LIGGGHTS_STL_EXPORT"
    verts = np.array([[float(w) for w in l.split()[1:]] for l in lines if l.strip().startswith("vertex")])
    normals = np.array([[float(w) for w in l.split()[2:]] for l in lines if l.strip().startswith("facet")])
    assert verts.shape == (3 * len(nodes), 3) and normals.shape == (len(nodes), 3)
    np.testing.assert_allclose(verts, nodes.reshape(-1, 3), rtol=1e-5, atol=1e-9)  # %g: 6 digits
    cr = np.cross(nodes[:, 1] - nodes[:, 0], nodes[:, 2] - nodes[:, 1])
    np.testing.assert_allclose(normals, cr / np.linalg.norm(cr, axis=1)[:, None], rtol=1e-5, atol=1e-6)
    raw = (tmp_path / "floorb300.stl").read_bytes()
    assert raw[:28] == b"STL File written by LIGGGHTS" and raw[28:80] == b" " * 52
    ntri = int(np.frombuffer(raw[80:84], dtype=np.uint32)[0])
    assert ntri == len(nodes) and len(raw) == 84 + 50 * ntri
    rec = np.frombuffer(raw[84:], dtype=np.dtype([("v", "<f4", 12), ("a", "<u2")]))
    np.testing.assert_array_equal(rec["v"][:, 3:], nodes.reshape(-1, 9).astype(np.float32))
    assert not (tmp_path / "floorb150.stl").exists()
    g.close()


def test_rotating_drum_run_in_pieces():
    """`run N pre no` in uneven pieces on a rotating drum: mesh-forced rebuilds
    that fall on a piece boundary carry over to the next run, so the pieces
    give bitwise the state of one run, and the oracle driven in the same
    pieces agrees to 1e-8 with the same rebuild count."""
    say("[gpu] rotating drum in pieces")
    case = _small_drum()
    pieces = [37, 1, 64, 2, 150, 46]
    g1, o = _build_mesh_case(case)
    g1.run(sum(pieces))
    g2 = pydem.GpuSystem(case[0], case[1], case[2], case[3], meshes=case[4])
    g2.run(pieces[0])
    o.run(pieces[0])
    for k in pieces[1:]:
        g2.run_continue(k)
        o.run_continue(k)
    s1, s2, so = g1.state(), g2.state(), o.state()
    say(f"  rebuilds one-run={g1.stats()['nrebuilds']} pieces={g2.stats()['nrebuilds']} "
        f"oracle={o.nrebuilds()} forced={o.mesh_forced_rebuilds()}")
    assert g1.stats()["nrebuilds"] == g2.stats()["nrebuilds"] == o.nrebuilds()
    for k in ("x", "v", "omega"):
        assert np.array_equal(s1[k], s2[k]), k
        assert_close(k, s2[k], so[k], 1e-8)
    n1, v1 = g1.mesh_nodes()
    n2, v2 = g2.mesh_nodes()
    on, ov = o.mesh_nodes()
    assert np.array_equal(n1, n2) and np.array_equal(n2, on) and np.array_equal(v2, ov)
    g1.close()
    g2.close()


@pytest.mark.parametrize("case", ["pack", "drum", "wiggle", "vibrot"])
def test_graph_capture_is_bitwise_eager(case):
    """Steps replayed from captured HIP graphs (32 per replay, rebuilds and
    other eager work between replays, several `run pre no` pieces) give
    bitwise the state of step-by-step launches."""
    say(f"[gpu] graph capture {case}")
    if case == "pack":
        params, mats, P, walls = wl.dense_pack(seed=34, n=3000, rolling=True, speed=0.3)
        meshes = ()
        pieces = [150, 70, 180]
    elif case == "drum":
        params, mats, P, walls, meshes = _small_drum()
        pieces = [200, 33, 300]
    else:  # the time-dependent movers read their motion table inside the captured steps
        params, mats, P, walls, meshes = _mesh_floor_pack(seed=44, n=1500, moving=case)
        pieces = [150, 70, 180]
    a = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    b = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    b.set_graph_capture(True)
    a.run(pieces[0])
    b.run(pieces[0])
    for k in pieces[1:]:
        a.run_continue(k)
        b.run_continue(k)
    sa, sb = a.state(), b.state()
    say(f"  rebuilds eager={a.stats()['nrebuilds']} graph={b.stats()['nrebuilds']}")
    assert a.stats()["nrebuilds"] == b.stats()["nrebuilds"]
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sb[k]), k
    if meshes:
        assert np.array_equal(a.mesh_nodes()[0], b.mesh_nodes()[0])
    a.close()
    b.close()


def test_mesh_stress_totals_match_oracle():
    """fix mesh/surface/stress on the small rotating drum: the total force and
    the torque about the drum axis that the particles exert on the mesh equal
    the oracle's at setup (1e-10) and after 400 steps (1e-8, where the state
    agrees to 1e-8)."""
    say("[gpu] mesh stress totals")
    params, mats, P, walls, meshes = _small_drum()
    meshes = [tuple(meshes[0][:4]) + ((0.0, 0.0, 0.0),)]
    g, o = _build_mesh_case((params, mats, P, walls, meshes))
    g.setup()
    o.setup()
    for tol in (1e-10, 1e-8):
        gF, gT = g.mesh_force()
        oF, oT = o.mesh_force()
        say(f"  F gpu {gF} oracle {oF}  T gpu {gT} oracle {oT}")
        assert np.linalg.norm(gF - oF) <= tol * max(np.linalg.norm(oF), 1e-300)
        assert np.linalg.norm(gT - oT) <= tol * max(np.linalg.norm(oT), 1e-300)
        if tol == 1e-10:
            g.run(400)
            o.run(400)
    assert np.linalg.norm(oF) > 0 and np.linalg.norm(oT) > 0
    g.close()


def _read_dump(path):
    """snapshots of a dump custom text file: [(step, columns, rows)]"""
    snaps, lines, k = [], open(path).read().splitlines(), 0
    while k < len(lines):
        assert lines[k] == "ITEM: TIMESTEP"
        step = int(lines[k + 1])
        assert lines[k + 2] == "ITEM: NUMBER OF ATOMS"
        n = int(lines[k + 3])
        assert lines[k + 4].startswith("ITEM: BOX BOUNDS ")
        cols = lines[k + 8].split()[2:]
        rows = np.array([[float(v) for v in l.split()] for l in lines[k + 9:k + 9 + n]])
        snaps.append((step, cols, rows, lines[k + 4][17:]))
        k += 9 + n
    return snaps


def test_script_dump_custom_matches_api(tmp_path):
    """dump custom every 100 steps (one file, and one file per snapshot with
    '*'), across two runs: the snapshots sit at the right steps, in the
    reference's layout (box bounds string, %g values, atoms by ID), and equal
    the state of the same system driven through the C ABI."""
    say("[gpu] script dump custom")
    params, mats, P, walls = wl.settling_4k(seed=7, n=500)
    data = tmp_path / "d.data"
    wl.write_data_file(str(data), params, P)
    d1 = tmp_path / "all.txt"
    script = tmp_path / "in.dump"
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
dump d1 all custom 100 {d1} id type x y z vx vy vz omegax fz radius mass
dump d2 all custom 250 {tmp_path}/snap*.txt id x y z
thermo_style custom step atoms ke
thermo 1000
run 300
run 200
""")
    rc = pydem.run_script(str(script), str(tmp_path / "log"))
    assert rc == 0, (tmp_path / "log").read_text()[-800:]
    snaps = _read_dump(d1)
    assert [s[0] for s in snaps] == [0, 100, 200, 300, 400, 500]
    assert sorted(p.name for p in tmp_path.glob("snap*.txt")) == ["snap0.txt", "snap250.txt", "snap500.txt"]
    step, cols, rows, bounds = snaps[-1]
    assert bounds == "pp pp ff"
    assert cols == ["id", "type", "x", "y", "z", "vx", "vy", "vz", "omegax", "fz", "radius", "mass"]
    assert np.array_equal(rows[:, 0], np.arange(1, 501))
    g = pydem.GpuSystem(params, mats, P, walls)
    g.run(300)
    g.run(200)
    st = g.state()
    order = np.argsort(P["tag"])
    for j, (key, comp) in enumerate((("x", 0), ("x", 1), ("x", 2), ("v", 0), ("v", 1), ("v", 2),
                                     ("omega", 0), ("f", 2))):
        ref = st[key][order, comp]
        assert np.allclose(rows[:, 2 + j], ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max() + 1e-300), key
    assert np.allclose(rows[:, 10], P["radius"][order], rtol=1e-5)
    assert np.allclose(rows[:, 11], P["rmass"][order], rtol=1e-5)
    s250 = _read_dump(tmp_path / "snap250.txt")
    assert len(s250) == 1 and s250[0][0] == 250
    g.close()


def _read_vtk_polydata(path):
    """A legacy ASCII VTK polydata file: (points (n,3), vertex cells, {name:
    (ncomp, dtype, values)}) in file order."""
    tok = path.read_text().split("\n")
    assert tok[0] == "# vtk DataFile Version 3.0" and tok[2] == "ASCII" and tok[3] == "DATASET POLYDATA"
    words = " ".join(tok[4:]).split()
    assert words[0] == "POINTS" and words[2] == "double"
    n = int(words[1])
    pts = np.array(words[3:3 + 3 * n], dtype=float).reshape(n, 3)
    k = 3 + 3 * n
    assert words[k] == "VERTICES" and int(words[k + 1]) == n and int(words[k + 2]) == 2 * n
    cells = np.array(words[k + 3:k + 3 + 2 * n], dtype=int).reshape(n, 2)
    k += 3 + 2 * n
    assert words[k] == "POINT_DATA" and int(words[k + 1]) == n and words[k + 2:k + 4] == ["FIELD", "FieldData"]
    narr = int(words[k + 4])
    k += 5
    arrays = {}
    for _ in range(narr):
        name, ncomp, nt, typ = words[k], int(words[k + 1]), int(words[k + 2]), words[k + 3]
        assert nt == n and typ in ("int", "double")
        vals = np.array(words[k + 4:k + 4 + ncomp * n], dtype=float).reshape(n, ncomp)
        arrays[name] = (ncomp, typ, vals)
        k += 4 + ncomp * n
    assert k == len(words)
    return pts, cells, arrays


def test_script_dump_custom_vtk(tmp_path):
    """dump custom/vtk (the packing tutorial's column list without image
    counts, a periodic box) beside dump custom at the same steps: the legacy
    VTK polydata holds the atoms by ID as points with one vertex each and the
    columns as point-data arrays in the reference's attribute order (id, type,
    mass, v, f, radius, omega; three components of a vector in one array), the
    same numbers as the text dump, and a bounding-box grid file with the box."""
    say("[gpu] script dump custom/vtk")
    params, mats, P, walls = wl.settling_4k(seed=8, n=400)
    data = tmp_path / "d.data"
    wl.write_data_file(str(data), params, P)
    script = tmp_path / "in.vtk"
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
dump dmp all custom/vtk 150 {tmp_path}/post/packing_*.vtk id type type x y z vx vy vz fx fy fz omegax omegay omegaz radius mass
dump txt all custom 150 {tmp_path}/post/t*.txt id type x y z vx vy vz fx fy fz omegax omegay omegaz radius mass
thermo 1000
run 300
""")
    (tmp_path / "post").mkdir()
    rc = pydem.run_script(str(script), str(tmp_path / "log"))
    assert rc == 0, (tmp_path / "log").read_text()[-800:]
    names = sorted(p.name for p in (tmp_path / "post").glob("packing_*.vtk"))
    assert names == ["packing_0.vtk", "packing_150.vtk", "packing_300.vtk",
                     "packing_boundingBox_0.vtk", "packing_boundingBox_150.vtk", "packing_boundingBox_300.vtk"]
    for st in (0, 150, 300):
        pts, cells, arrays = _read_vtk_polydata(tmp_path / "post" / f"packing_{st}.vtk")
        assert list(arrays) == ["id", "type", "mass", "v", "f", "radius", "omega"]
        assert [arrays[k][0] for k in arrays] == [1, 1, 1, 3, 3, 1, 3]
        assert [arrays[k][1] for k in arrays] == ["int", "int", "double", "double", "double", "double", "double"]
        assert np.array_equal(cells, np.stack([np.ones(400, int), np.arange(400)], axis=1))
        _, cols, rows, _ = _read_dump(tmp_path / "post" / f"t{st}.txt")[0]
        assert np.array_equal(arrays["id"][2][:, 0], rows[:, 0]) and np.array_equal(arrays["type"][2][:, 0], rows[:, 1])

        def close(a, b):  # %g keeps six significant digits, %.11lg eleven
            return np.allclose(a, b, rtol=1e-5, atol=1e-5 * np.abs(b).max() + 1e-300)

        assert close(pts, rows[:, 2:5])
        assert close(arrays["v"][2], rows[:, 5:8]) and close(arrays["f"][2], rows[:, 8:11])
        assert close(arrays["omega"][2], rows[:, 11:14])
        assert close(arrays["radius"][2][:, 0], rows[:, 14]) and close(arrays["mass"][2][:, 0], rows[:, 15])
        box = (tmp_path / "post" / f"packing_boundingBox_{st}.vtk").read_text().split("\n")
        assert box[3] == "DATASET RECTILINEAR_GRID" and box[4] == "DIMENSIONS 2 2 2"
        got = [float(v) for line in (box[6], box[8], box[10]) for v in line.split()]
        assert np.allclose(got[:4], [0.0, 0.1, 0.0, 0.1])
    # the vtk file's own numbers keep eleven digits: the last snapshot against the API state
    g = pydem.GpuSystem(params, mats, P, walls)
    g.run(300)
    st = g.state()
    order = np.argsort(P["tag"])
    pts, _, arrays = _read_vtk_polydata(tmp_path / "post" / "packing_300.vtk")
    assert np.allclose(pts, st["x"][order], rtol=1e-10, atol=1e-13)
    assert np.allclose(arrays["v"][2], st["v"][order], rtol=1e-9, atol=1e-12)
    g.close()


@pytest.mark.parametrize("case", ["pack_walls", "drum"])
def test_restart_continues_bitwise(tmp_path, case):
    """write_restart after 200 steps, read_restart into a fresh system built
    the same way, 100 more steps: bitwise the state, pair set, histories and
    mesh nodes of the uninterrupted 200 + 100 steps (both second runs set up
    again, as `run` does after read_restart)."""
    say(f"[gpu] restart continuation {case}")
    if case == "drum":
        params, mats, P, walls, meshes = _small_drum()
    else:
        params, mats, P, walls = _walled_pack()
        meshes = ()
    a = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    a.run(200)
    a.run(100)
    b = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    b.run(200)
    f = tmp_path / "r.restart"
    b.write_restart(f)
    b.close()
    hd, P2 = pydem.restart_particles(f)
    assert hd["n"] == len(P["radius"]) and hd["ntimestep"] == 200
    assert np.array_equal(P2["radius"], P["radius"]) and np.array_equal(P2["tag"], P["tag"])
    c = pydem.GpuSystem(params, mats, P, walls, meshes=meshes)
    c.read_restart(f)
    c.run(100)
    sa, sc = a.state(), c.state()
    for k in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(sa[k], sc[k]), k
    pa, pc = a.pairs(), c.pairs()
    for u, w in zip(pa, pc):
        assert np.array_equal(u, w)
    assert a.stats()["ntimestep"] == c.stats()["ntimestep"] == 300
    if meshes:
        assert np.array_equal(a.mesh_nodes()[0], c.mesh_nodes()[0])
        ta, ha = a.mesh_history()
        tc, hc = c.mesh_history()
        assert np.array_equal(ta, tc) and np.array_equal(ha, hc)
    else:
        assert np.array_equal(a.wall_history(), c.wall_history())
    a.close()
    c.close()


def test_script_restart_matches_one_run(tmp_path):
    """A script that runs 300 steps equals one that runs 200, writes a
    restart, and a second script that reads it, redefines the physics and runs
    100 more: the same thermo line at step 300."""
    say("[gpu] script restart")
    params, mats, P, walls = wl.settling_4k(seed=9, n=600)
    data = tmp_path / "d.data"
    wl.write_data_file(str(data), params, P)
    physics = """neighbor 0.001 bin
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
"""
    head = """atom_style granular
atom_modify map array
boundary p p f
newton off
communicate single vel yes
units si
"""
    box = "region reg block 0 0.1 0 0.1 0 0.2 units box\ncreate_box 1 reg\n"
    tail = "fix integr all nve/sphere\nthermo_style custom step atoms ke\nthermo 100\n"
    rst = tmp_path / "s.restart"
    scripts = {
        "one": head + box + physics + f"read_data {data}\n" + tail + "run 200\nrun 100\n",
        "first": head + box + physics + f"read_data {data}\n" + tail + f"run 200\nwrite_restart {rst}\n",
        "second": head + f"read_restart {rst}\n" + physics + tail + "run 100\n",
    }
    last = {}
    for name in ("one", "first", "second"):
        p = tmp_path / f"in.{name}"
        p.write_text(scripts[name])
        log = tmp_path / f"log.{name}"
        rc = pydem.run_script(str(p), str(log))
        text = log.read_text()
        assert rc == 0, text[-800:]
        rows = [l.split() for l in text.splitlines() if l.strip() and l.split()[0].isdigit()]
        last[name] = rows[-1]
    say(f"  one {last['one']}  restarted {last['second']}")
    assert last["second"][0] == "300" and last["second"] == last["one"]
