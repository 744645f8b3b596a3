"""Pins of the oracle's compute pair/gran/local (orc_pair_local):

  forces   on a fully periodic pack without walls or gravity, the per-pair
           force on the first particle, applied with its reaction to both
           particles in the list's order, rebuilds every particle's force of
           the force pass with shearupdate off (orc_compute_forces(0)) exactly
  torques  without rolling friction the contact's torques on the two spheres
           are -cr (en x Ft) with cr = r - delta/2 each, so the second
           particle's torque is the first one's times cr_j / cr_i; summed,
           they rebuild every particle's torque (1e-12)
  rows     one per half-list pair, in the order of orc_get_pairs with its tags
           and history; pairs that do not touch carry zero force and torque
  passive  the evaluation changes nothing: a second call and the state and
           histories afterwards are identical
"""
import numpy as np

import pyoracle as orc
import workloads as wl


def _periodic_pack(rolling):
    params, mats, P, _ = wl.dense_pack(seed=17, n=800, rolling=rolling, periodic=(1, 1, 1), speed=0.3)
    params = dict(params)
    params.pop("grav_acc")
    return params, mats, P


def test_pair_forces_rebuild_the_force_pass():
    params, mats, P = _periodic_pack(rolling=False)
    o = orc.OracleSystem(params, mats, P, [])
    o.run(30)
    tags, d = o.pair_local()
    o.lib.orc_compute_forces(o.s, 0)
    st = o.state()
    idx = {int(t): k for k, t in enumerate(P["tag"])}
    # the oracle's pair order is its list order: owner i, then its entries;
    # rows are oriented from the lower tag, so recover each owner from the
    # pairs' order is not needed: both particles get +-F in the row order
    f = np.zeros_like(st["f"])
    a = np.array([idx[int(t)] for t in tags[:, 0]])
    b = np.array([idx[int(t)] for t in tags[:, 1]])
    for k in range(len(tags)):
        f[a[k]] += d[k, 12:15]
        f[b[k]] -= d[k, 12:15]
    scale = np.abs(st["f"]).max()
    assert scale > 0
    assert np.abs(f - st["f"]).max() <= 1e-12 * scale
    # torques: the second sphere's is the first one's scaled by cr_j / cr_i
    x, r = st["x"], P["radius"]
    L = np.array(params["boxhi"]) - np.array(params["boxlo"])
    t = np.zeros_like(st["torque"])
    for k in range(len(tags)):
        del_ = x[a[k]] - x[b[k]]
        del_ -= np.round(del_ / L) * L
        dist = np.linalg.norm(del_)
        dn = r[a[k]] + r[b[k]] - dist
        t[a[k]] += d[k, 15:18]
        if dn > 0:
            t[b[k]] += d[k, 15:18] * (r[b[k]] - 0.5 * dn) / (r[a[k]] - 0.5 * dn)
    tscale = np.abs(st["torque"]).max()
    assert tscale > 0
    assert np.abs(t - st["torque"]).max() <= 1e-12 * tscale
    o.close()


def test_rows_follow_the_pair_list_and_do_not_change_the_state():
    params, mats, P = _periodic_pack(rolling=True)
    o = orc.OracleSystem(params, mats, P, [])
    o.run(25)
    before = o.state()
    pt, pf, ph = o.pairs()
    tags, d = o.pair_local()
    tags2, d2 = o.pair_local()
    assert np.array_equal(tags, pt) and np.array_equal(tags, tags2) and np.array_equal(d, d2)
    assert np.array_equal(d[:, 21:24], ph)
    assert np.array_equal(d[:, 19], P['radius'][[{int(t): k for k, t in enumerate(P['tag'])}[int(t)] for t in tags[:, 0]]])
    after = o.state()
    for k in ("x", "v", "omega"):
        assert np.array_equal(before[k], after[k])
    assert np.array_equal(o.pairs()[2], ph)
    # pairs apart: no force, no torque; touching pairs: a force
    x = after["x"]
    idx = {int(t): k for k, t in enumerate(P["tag"])}
    a = np.array([idx[int(t)] for t in tags[:, 0]])
    b = np.array([idx[int(t)] for t in tags[:, 1]])
    L = np.array(params["boxhi"]) - np.array(params["boxlo"])
    del_ = x[a] - x[b]
    del_ -= np.round(del_ / L) * L
    apart = np.linalg.norm(del_, axis=1) >= P["radius"][a] + P["radius"][b]
    assert apart.sum() > 100 and (~apart).sum() > 100
    assert np.all(d[apart, 12:18] == 0.0)
    assert np.all(np.abs(d[~apart, 12:15]).sum(axis=1) > 0)
    # the periodic flag marks the pairs whose partner is an image
    raw = x[a] - x[b]
    assert np.array_equal(d[:, 18] == 1.0, np.any(np.abs(raw) > 0.5 * L, axis=1))
    o.close()
