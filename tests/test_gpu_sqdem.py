"""GPU checks of the superquadric DEM step (dem_sqdem_*) against the oracle.

  particle properties     mass, volume-equivalent and bounding radius,
                          principal inertia equal to the oracle's (1e-13)
  setup and one step      from a colliding state: the pair set bit-exact, touch
                          flags and contact states equal, forces / torques /
                          shear history to 1e-10 relative, contact points to
                          1e-12 m; after one step positions, velocities,
                          orientations likewise
  2000 steps              kinetic energy, contact count and mean height within
                          1 % of the oracle's (the run is chaotic)
  warm start              with a trained network the same contacts form and
                          the fresh solves take fewer Newton iterations
"""
import sys

import numpy as np
import pytest

import pydem
import pyoracle as orc
import superquadrics as sq
import workloads as wl

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gas(seed=2, n=500):
    """Superquadrics (half-axes 1-2 mm, blockiness 2-8) in a 0.04 x 0.04 m
    periodic column above a floor, moving and spinning at random."""
    params, mats, P, walls = wl.superquadric_settling(seed, n, base=(0.04, 0.04), axes=(0.001, 0.002),
                                                      gap=0.0003)
    rng = np.random.default_rng(seed + 100)
    P["v"] = rng.normal(0, 0.3, (n, 3))
    P["omega"] = rng.normal(0, 30, (n, 3))
    return params, mats, P, walls


_SNAP = {}


def _snapshot(steps=1500):
    """The oracle's state after `steps` steps of the colliding gas."""
    if steps not in _SNAP:
        params, mats, P, walls = _gas()
        s = orc.SqsSystem(params, orc.materials(mats), P, walls)
        s.setup()
        s.run(steps)
        st = s.state()
        P2 = dict(P)
        P2.update(x=st["x"], v=st["v"], omega=st["omega"], quat=st["quat"])
        _SNAP[steps] = (params, mats, P2, walls)
        s.close()
    return _SNAP[steps]


def _close(a, b, rel, floor):
    """per-row |a - b| <= rel |b| + floor"""
    d = np.linalg.norm(a - b, axis=-1)
    return d <= rel * np.linalg.norm(b, axis=-1) + floor


def test_particle_properties_match_oracle():
    params, mats, P, walls = _gas()
    g = pydem.SqDemSystem(params, mats, P, walls)
    o = orc.SqsSystem(params, orc.materials(mats), P, walls)
    pg, po = g.props(), o.props()
    for k in ("mass", "rvol", "rbound", "inertia"):
        assert np.allclose(pg[k], po[k], rtol=1e-13, atol=0), k
    g.close()
    o.close()


def _compare_pairs(g, o, tag):
    ijg, flg, hg = g.pairs()
    ijo, flo, ho = o.pairs()
    assert np.array_equal(ijg, ijo), f"{tag}: pair sets differ"
    assert np.array_equal(flg, flo), f"{tag}: touch flags differ"
    assert np.array_equal(hg[:, 0], ho[:, 0]), f"{tag}: contact states differ"
    assert np.array_equal(hg[:, 1], ho[:, 1]), f"{tag}: OBB axes differ"
    near = ho[:, 0] > 0
    dcp = np.abs(hg[near, 2:5] - ho[near, 2:5]).max(initial=0)
    assert dcp <= 1e-12, f"{tag}: contact points differ by {dcp:.2e} m"
    touch = ho[:, 0] == 2
    assert np.abs(hg[touch, 5:7] - ho[touch, 5:7]).max(initial=0) <= 1e-12, tag
    sh_ok = _close(hg[:, 7:10], ho[:, 7:10], 1e-10, 1e-10 * np.abs(ho[:, 7:10]).max(initial=1e-30))
    assert sh_ok.all(), f"{tag}: shear history differs"
    return int(touch.sum()), int(len(ijo))


def test_setup_and_one_step_match_oracle():
    say("[gpu] superquadric DEM: setup and one step vs the oracle")
    params, mats, P, walls = _snapshot()
    g = pydem.SqDemSystem(params, mats, P, walls)
    o = orc.SqsSystem(params, orc.materials(mats), P, walls)
    g.run(0)
    o.setup()
    for step in (0, 1):
        if step:
            g.run(1)
            o.run(1)
        sg, so = g.state(), o.state()
        fscale = np.abs(so["f"]).max()
        tscale = np.abs(so["torque"]).max()
        fok = _close(sg["f"], so["f"], 1e-10, 1e-10 * fscale)
        tok = _close(sg["torque"], so["torque"], 1e-10, 1e-10 * tscale)
        ntouch, npairs = _compare_pairs(g, o, f"step {step}")
        df = np.abs(sg["f"] - so["f"]).max()
        say(f"  step {step}: pairs {npairs} touching {ntouch} wall {g.stats()['wall_touch']}  "
            f"max |df| {df:.2e} N (scale {fscale:.2e})  rows off f {int((~fok).sum())} t {int((~tok).sum())}")
        assert ntouch > 10
        assert fok.all() and tok.all()
        for k in ("x", "v", "omega", "angmom", "quat"):
            scale = np.abs(so[k]).max()
            assert _close(sg[k], so[k], 1e-12, 1e-12 * scale).all(), k
        wg, wo = g.wall_hist(), o.wall_hist()
        assert np.allclose(wg, wo, rtol=1e-10, atol=1e-10 * max(np.abs(wo).max(), 1e-30))
    st = g.stats()
    assert st["pair_touch"] == o.stats()["pair_touch"]
    assert st["wall_touch"] == o.stats()["wall_touch"]
    g.close()
    o.close()


def _ke(state, props):
    return 0.5 * (props["mass"][:, None] * state["v"] ** 2).sum() + 0.5 * (state["omega"] * state["angmom"]).sum()


def test_two_thousand_steps_statistics_match_oracle():
    say("[gpu] superquadric DEM: 2000 steps vs the oracle")
    params, mats, P, walls = _snapshot()
    g = pydem.SqDemSystem(params, mats, P, walls)
    o = orc.SqsSystem(params, orc.materials(mats), P, walls)
    g.run(0)
    o.setup()
    kg, ko, cg, co, zg, zo = [], [], [], [], [], []
    for k in range(8):
        g.run(250)
        o.run(250)
        sg, so = g.state(), o.state()
        kg.append(_ke(sg, g.props()))
        ko.append(_ke(so, o.props()))
        stg, sto = g.stats(), o.stats()
        cg.append(stg["pair_touch"] + stg["wall_touch"])
        co.append(sto["pair_touch"] + sto["wall_touch"])
        zg.append(sg["x"][:, 2].mean())
        zo.append(so["x"][:, 2].mean())
        say(f"  step {250 * (k + 1)}: KE {kg[-1]:.4e} / {ko[-1]:.4e}  contacts {cg[-1]} / {co[-1]}  "
            f"z {zg[-1]:.5f} / {zo[-1]:.5f}  rebuilds {stg['rebuilds']} / {sto['rebuilds']}")
    assert abs(kg[-1] - ko[-1]) <= 0.01 * ko[-1]
    # contact counts of single steps fluctuate; compare the mean over the run
    assert abs(np.mean(cg) - np.mean(co)) <= 0.01 * np.mean(co) + 1
    assert abs(zg[-1] - zo[-1]) <= 0.01 * zo[-1]
    g.close()
    o.close()


def test_warm_start_forms_the_same_contacts_with_fewer_iterations():
    say("[gpu] superquadric DEM: network warm start")
    solver = pydem.SqSolver()
    A, B = sq.overlapping_pairs(24, 40000, axes=(0.001, 0.002))
    solver.set_pairs(A, B)
    solver.solve(pydem.DEM_SQ_FRESH)
    solver.net_init(seed=1)
    curve, ms = solver.net_train(steps=1200, batch=4096, lr=2e-3, seed=2)
    params, mats, P, walls = _snapshot()
    runs = {}
    for mode in ("plain", "warm"):
        g = pydem.SqDemSystem(params, mats, P, walls)
        if mode == "warm":
            g.set_warm_start(solver)
        g.run(200)
        runs[mode] = (g.stats(), g.state(), g.pairs())
        g.close()
    sp, sw = runs["plain"][0], runs["warm"][0]
    ip = sp["fresh_iters"] / max(sp["fresh"], 1)
    iw = sw["fresh_iters"] / max(sw["fresh"], 1)
    say(f"  fresh solves {sp['fresh']} / {sw['fresh']}, iterations per fresh solve {ip:.1f} -> {iw:.1f}, "
        f"fallbacks {sw['fallbacks']}, contacts {sp['pair_touch']} / {sw['pair_touch']}")
    assert sp["fresh"] > 50
    assert iw < ip
    assert abs(sp["pair_touch"] - sw["pair_touch"]) <= max(2, 0.05 * sp["pair_touch"])
    solver.close()


def test_script_front_end_matches_api(tmp_path):
    """A superquadric input script (atom_style superquadric, read_data,
    set group all quat/random, surface superquadric, fix nve/superquadric)
    run by dem_run_script gives the kinetic energy of the same system driven
    through dem_sqdem_*, with orientations from the restated quat/random."""
    say("[gpu] superquadric script front end")
    from test_script_superquadric import HEADER, _gas as _sgas, quat_random
    params, mats, P, walls = _sgas(300)
    data = tmp_path / "sq.data"
    wl.write_superquadric_data(str(data), params, P)
    script = tmp_path / "in.sq"
    script.write_text(HEADER + f"read_data {data}\nset group all quat/random 4711\n"
                      "fix integr all nve/superquadric\nthermo_style custom step atoms ke\nthermo 100\nrun 300\n")
    log = tmp_path / "log.sq"
    rc = pydem.run_script(str(script), str(log))
    text = log.read_text()
    say(text[-500:])
    assert rc == 0
    rows = [l.split() for l in text.splitlines()
            if len(l.split()) == 3 and l.split()[0].isdigit() and l.split()[1].isdigit()]
    assert int(rows[-1][0]) == 300 and int(rows[-1][1]) == 300
    P2 = dict(P)
    P2["quat"] = quat_random(4711, P["x"])
    g = pydem.SqDemSystem(params, mats, P2, walls)
    kes = []
    for _ in range(3):
        g.run(100)
        st = g.state()
        kes.append(0.5 * (g.props()["mass"] @ (st["v"] ** 2).sum(1)))
    for row, ke in zip(rows[1:], kes):
        assert float(row[2]) == pytest.approx(ke, rel=1e-6)
    g.close()
