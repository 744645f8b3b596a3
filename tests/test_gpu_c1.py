"""The 1M-sphere configurations at full size, through the C ABI, against the
CPU oracle on the same seeded inputs:

  settling (1M polydisperse spheres, Hertz/history + CDT rolling, dilute
  jittered lattice, periodic xy, floor, dt 5e-6 s): FP64 setup + one step with
  the half-list pair set equal by tags, flags equal, and per-particle force,
  torque, shear, x, v, omega within 1e-10 relative; then 100 more steps (a
  forced build at the second run's setup) with the state within 1e-10; and
  the whole 2,000-step run in both GPU modes, whose kinetic energy, contact
  count and packing fraction agree within 1 %.

  packed bed (the same 1M spheres from a compressed lattice, about three
  touching neighbors each): FP64 setup + one step as above, and the mixed
  FP32/FP64 mode's per-particle force and torque within 1e-5 of the FP64
  oracle (measured values printed).

Errors are per particle: max_i |a_i - b_i| over max_i |b_i| (the force scale
of the system), and norm-wise.
"""
import sys

import numpy as np
import pytest

import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu

N = 1_000_000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def errs(a, b):
    den = np.linalg.norm(b)
    scale = np.linalg.norm(b, axis=1).max()
    d = np.linalg.norm(a - b, axis=1).max()
    if den == 0.0:
        return float(np.linalg.norm(a)), float(d)
    return float(np.linalg.norm(a - b) / den), float(d / scale)


def check(name, a, b, tol):
    nw, per = errs(a, b)
    say(f"  {name}: normwise {nw:.3e} per-particle {per:.3e}")
    assert nw <= tol and per <= tol, f"{name}: {nw:.3e} / {per:.3e} > {tol}"
    return per


def sorted_pairs(sysobj, n):
    t, f, h = sysobj.pairs()
    key = t[:, 0].astype(np.int64) * (n + 1) + t[:, 1].astype(np.int64)
    order = np.argsort(key, kind="stable")
    return key[order], f[order], h[order]


def compare_pairs(g, o, n, hist_tol):
    gk, gf, gh = sorted_pairs(g, n)
    ok, of, oh = sorted_pairs(o, n)
    say(f"  pairs gpu={len(gk)} oracle={len(ok)} flagged={int(of.sum())}")
    assert len(gk) == len(ok) and np.array_equal(gk, ok), "pair sets differ"
    assert np.unique(gk).size == len(gk)
    assert np.array_equal(gf, of), f"{int((gf != of).sum())} flags differ"
    hs = max(np.abs(oh).max(), 1e-300)
    herr = float(np.abs(gh - oh).max() / hs)
    say(f"  shear history: max error / max |shear| = {herr:.3e}")
    assert herr <= hist_tol
    return len(gk), int(of.sum())


def test_settling_1m_fp64_setup_one_step_and_100_steps():
    say("[gpu] settling 1M fp64: setup + 1 step, then 100 steps")
    params, mats, P, walls = wl.settling_1m(seed=0, n=N)
    g = pydem.GpuSystem(params, mats, P, walls, precision=pydem.DEM_FP64)
    o = orc.OracleSystem(params, mats, P, walls)
    g.run(1)
    o.run(1)
    npairs, nflag = compare_pairs(g, o, N, 1e-10)
    assert npairs > 2_000_000
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        check(k, gs[k], os_[k], 1e-10)
    st = g.stats()
    assert st["npairs"] == npairs
    assert (st["ncontacts"], st["nwall_contacts"]) == o.contact_counts()
    # a second run: its setup rebuilds (forced) and 100 steps follow
    g.run(100)
    o.run(100)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "omega", "f", "torque"):
        check(k, gs[k], os_[k], 1e-10)
    assert g.stats()["nrebuilds"] == o.nrebuilds()
    assert (g.stats()["ncontacts"], g.stats()["nwall_contacts"]) == o.contact_counts()
    say(f"  contacts after 101 steps: pairs {o.contact_counts()[0]} walls {o.contact_counts()[1]}")
    compare_pairs(g, o, N, 1e-10)
    g.close()
    o.close()


def test_packed_1m_fp64_and_mixed_vs_oracle():
    say("[gpu] packed 1M: fp64 parity and mixed error vs the fp64 oracle")
    params, mats, P, walls = wl.packed_1m(seed=0, n=N)
    o = orc.OracleSystem(params, mats, P, walls)
    o.run(1)
    os_ = o.state()
    g = pydem.GpuSystem(params, mats, P, walls, precision=pydem.DEM_FP64)
    g.run(1)
    npairs, nflag = compare_pairs(g, o, N, 1e-10)
    # about three touching neighbours per sphere (1.6M touching half-list pairs)
    assert nflag > 1_500_000
    gs = g.state()
    for k in ("x", "v", "omega", "f", "torque"):
        check(k, gs[k], os_[k], 1e-10)
    g.close()
    m = pydem.GpuSystem(params, mats, P, walls, precision=pydem.DEM_MIXED)
    m.run(1)
    ms = m.state()
    ef = check("f (mixed)", ms["f"], os_["f"], 1e-5)
    et = check("torque (mixed)", ms["torque"], os_["torque"], 1e-5)
    say(f"  mixed vs fp64 oracle at 1M packed: force {ef:.3e} torque {et:.3e}")
    mk, mf, _ = sorted_pairs(m, N)
    ok, of, _ = sorted_pairs(o, N)
    assert np.array_equal(mk, ok) and np.array_equal(mf, of)
    m.close()
    o.close()


def _stats(state, P, params):
    ke = 0.5 * float(P["rmass"] @ (state["v"] ** 2).sum(1))
    area = (params["boxhi"][0] - params["boxlo"][0]) * (params["boxhi"][1] - params["boxlo"][1])
    vol = float((4.0 / 3.0 * np.pi * P["radius"] ** 3).sum())
    phi = vol / (area * 2.0 * float(state["x"][:, 2].mean()))
    return ke, phi


def test_settling_1m_2000_steps_mixed_vs_fp64_statistics():
    """The whole configured run (2,000 steps) in FP64 and in the mixed mode:
    kinetic energy, contact count (pairs + floor) and packing fraction within
    1 % of each other."""
    say("[gpu] settling 1M, 2000 steps: mixed vs fp64 statistics")
    params, mats, P, walls = wl.settling_1m(seed=0, n=N)
    out = {}
    for name, prec in (("fp64", pydem.DEM_FP64), ("mixed", pydem.DEM_MIXED)):
        g = pydem.GpuSystem(params, mats, P, walls, precision=prec)
        g.run(2000)
        ke, phi = _stats(g.state(), P, params)
        st = g.stats()
        out[name] = (ke, st["ncontacts"] + st["nwall_contacts"], phi)
        say(f"  {name}: KE {ke:.6e} contacts {out[name][1]} (pairs {st['ncontacts']}, "
            f"floor {st['nwall_contacts']}) phi {phi:.6f} rebuilds {st['nrebuilds']}")
        g.close()
    (k0, c0, p0), (k1, c1, p1) = out["fp64"], out["mixed"]
    assert abs(k1 - k0) <= 0.01 * k0
    assert abs(c1 - c0) <= 0.01 * max(c0, 1)
    assert abs(p1 - p0) <= 0.01 * p0
