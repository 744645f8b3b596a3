"""GPU checks of the superquadric contact-point solve and its warm-start
network, through the C ABI (dem_sq_*):

  plain solve (fresh contacts)   contact points equal to the oracle's to 1e-12 m
                                 where both converged; iteration counts equal
  given start points             the same against the oracle's solve from the
                                 same points
  network inputs                 the device encoding equals superquadrics.py
  network forward / gradient     equal to the numpy checker with the same BF16
                                 rounding points (1e-3 norm-wise)
  contact detection              checkSurfaceIntersect over an approach
                                 sequence (far -> close -> intersecting):
                                 states, OBB axes and contact flags equal the
                                 oracle's; points, normals, overlaps and
                                 surface points to 1e-12 m (1e-9 normals)
  training + warm start          the trained network cuts Newton iterations
                                 per fresh contact at least 2x; warm-started
                                 points agree with the plain solve (median
                                 1e-9 m under the reference's tolerance; 99 %
                                 within 1e-9 m when both take the same extra
                                 Newton refinement)
"""
import sys

import numpy as np
import pytest

import mlp_oracle as mo
import pydem
import pyoracle as orc
import superquadrics as sq

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _merit(A, B, p):
    out = np.zeros(len(A))
    for k in range(len(A)):
        fa, ga, _ = orc.sq_props(A[k], p[k], False)
        fb, gb, _ = orc.sq_props(B[k], p[k], False)
        v = np.cross(ga, gb)
        s = abs(v @ v / ((ga @ ga) * (gb @ gb)))
        aa = abs(fa - fb) / (max(abs(fa), abs(fb)) + 1e-16)
        out[k] = max(s, aa * aa)
    return out


def test_plain_solve_matches_oracle():
    say("[gpu] superquadric plain solve vs oracle")
    A, B = sq.overlapping_pairs(21, 600)
    s = pydem.SqSolver()
    s.set_pairs(A, B)
    r = s.solve(pydem.DEM_SQ_FRESH)
    op, ofi, ofj, oit, ofl = orc.sq_contact_batch(A, B)
    d = np.linalg.norm(r["point"] - op, axis=1)
    same_it = (r["iters"] == oit).mean()
    say(f"  |dp| max {d.max():.3e}  median {np.median(d):.3e}  iterations equal {same_it:.3f}  "
        f"mean iters gpu {r['iters'].mean():.2f} oracle {oit.mean():.2f}  solve {r['solve_ms']:.2f} ms")
    assert (r["flags"] & 1).sum() == ofl.sum() == 0
    assert same_it >= 0.95
    assert (d <= 1e-12).mean() >= 0.95
    s.close()


def test_given_start_matches_oracle():
    say("[gpu] superquadric solve from given points vs oracle")
    A, B = sq.overlapping_pairs(22, 400)
    op, *_ = orc.sq_contact_batch(A, B)
    rng = np.random.default_rng(0)
    L = sq.equiv_radius(A[:, 7:10]) + sq.equiv_radius(B[:, 7:10])
    init = op + rng.normal(size=op.shape) * 0.05 * L[:, None]
    s = pydem.SqSolver()
    s.set_pairs(A, B)
    r = s.solve(pydem.DEM_SQ_GIVEN, init)
    gp, _, _, git, gfl = orc.sq_contact_batch(A, B, init)
    # the GPU adds the fresh fallback where the solve misses the tolerance
    direct = (r["flags"] & 2) == 0
    d = np.linalg.norm(r["point"][direct] - gp[direct], axis=1)
    say(f"  fallbacks {int((~direct).sum())}  |dp| max {d.max():.3e}  mean iters {r['iters'].mean():.2f}")
    assert direct.mean() > 0.95
    assert (d <= 1e-12).mean() >= 0.95
    assert (r["iters"][direct] == git[direct]).mean() >= 0.95
    s.close()


def test_device_features_match_encoding():
    A, B = sq.overlapping_pairs(23, 500)
    s = pydem.SqSolver()
    s.set_pairs(A, B)
    assert np.allclose(s.features(), sq.features(A, B), rtol=1e-6, atol=1e-6)
    s.close()


def test_network_forward_and_gradient_match_checker():
    say("[gpu] network forward / gradient vs the numpy checker")
    rng = np.random.default_rng(3)
    s = pydem.SqSolver()
    s.net_init(seed=5)
    p = s.net_params()
    # move the biases off zero so every parameter block is exercised
    p = p + rng.normal(0, 0.05, p.shape).astype(np.float32)
    s.set_net_params(p)
    X = rng.normal(0, 1, (1000, 18)).astype(np.float32)
    T = rng.normal(0, 0.5, (1000, 3)).astype(np.float32)
    Y = s.net_forward(X)
    Yc, _ = mo.forward(p, X)
    ey = np.linalg.norm(Y - Yc) / np.linalg.norm(Yc)
    g, loss = s.net_gradient(X, T)
    gc, lc = mo.gradient(p, X, T)
    eg = np.linalg.norm(g - gc) / np.linalg.norm(gc)
    say(f"  forward rel {ey:.2e}  gradient rel {eg:.2e}  loss {loss:.6f} / {lc:.6f}")
    assert ey <= 1e-3
    assert eg <= 1e-3
    assert loss == pytest.approx(lc, rel=1e-4)
    # the same gradient, block by block
    for name, blk in mo.unpack(g).items():
        ref = mo.unpack(gc)[name]
        assert np.linalg.norm(blk - ref) <= 2e-3 * np.linalg.norm(ref) + 1e-9, name
    s.close()


def _approach(seed, n, steps):
    """Pose pairs closing in along their line of centres, B turning: the
    poses of each step and the velocities that move them."""
    A, B = sq.overlapping_pairs(seed, n)
    rng = np.random.default_rng(seed)
    u = B[:, 0:3] - A[:, 0:3]
    dt = 1e-4
    spin = rng.normal(0, 20.0, (n, 3))  # rad/s
    seq = []
    for t in range(steps):
        Bt = B.copy()
        Bt[:, 0:3] = A[:, 0:3] + u * (1.3 - 0.35 * t / (steps - 1))
        ang = spin * dt * t
        th = np.linalg.norm(ang, axis=1, keepdims=True)
        axis = ang / np.maximum(th, 1e-300)
        dq = np.concatenate([np.cos(th / 2), axis * np.sin(th / 2)], axis=1)
        Bt[:, 3:7] = sq.quat_mul(dq, B[:, 3:7])
        vel = np.zeros((n, 12))
        vel[:, 3:6] = -0.35 / (steps - 1) * u / dt
        vel[:, 9:12] = spin
        seq.append((A, Bt, vel))
    mass = np.stack([np.prod(A[:, 7:10], axis=1), np.prod(B[:, 7:10], axis=1)], axis=1) * 2500.0
    return seq, mass, dt


def test_contact_detection_matches_oracle():
    say("[gpu] superquadric contact detection over an approach sequence vs oracle")
    n, steps = 500, 8
    seq, mass, dt = _approach(41, n, steps)
    s = pydem.SqSolver()
    hg = np.zeros((n, 7))
    ho = np.zeros((n, 7))
    seen = set()
    for t, (A, B, vel) in enumerate(seq):
        s.set_pairs(A, B)
        r = s.check(hg, vel, mass, dt)
        ho, oc, oen, odn, ocpa, ocpb, oit = orc.sq_check_batch(A, B, ho, vel, mass, dt)
        seen.update(np.unique(hg[:, 0]).tolist())
        assert np.array_equal(hg[:, 0], ho[:, 0]), t
        assert np.array_equal(hg[:, 1], ho[:, 1]), t
        assert np.array_equal(r["contact"], oc), t
        c = oc == 1
        dp = np.abs(hg[:, 2:5] - ho[:, 2:5]).max(axis=1)
        same = dp <= 1e-12
        say(f"  step {t}: far {int((ho[:, 0] == 0).sum())} close {int((ho[:, 0] == 1).sum())} "
            f"intersecting {int(c.sum())}  |dp| max {dp.max():.2e}  iterations equal "
            f"{(r['iters'] == oit).mean():.3f}  solve {r['solve_ms']:.2f} ms")
        assert same.mean() >= 0.98
        cs = c & same
        assert np.abs(r["deltan"][cs] - odn[cs]).max(initial=0) <= 1e-12
        assert np.abs(r["en"][cs] - oen[cs]).max(initial=0) <= 1e-9
        assert np.abs(r["cp_i"][cs] - ocpa[cs]).max(initial=0) <= 1e-12
        assert np.abs(r["cp_j"][cs] - ocpb[cs]).max(initial=0) <= 1e-12
        assert np.abs(hg[cs, 5:7] - ho[cs, 5:7]).max(initial=0) <= 1e-12
        assert (r["iters"] == oit).mean() >= 0.95
        assert np.all(((r["flags"] & 4) != 0) == (ho[:, 0] == 0))
        # continue both from the oracle's history so one near-tie cannot fork the sequences
        hg = ho.copy()
    assert seen == {0.0, 1.0, 2.0}
    s.close()


def test_training_and_warm_start_cut_newton_iterations():
    say("[gpu] network training and warm start")
    A, B = sq.overlapping_pairs(24, 60000)
    s = pydem.SqSolver()
    s.set_pairs(A, B)
    plain = s.solve(pydem.DEM_SQ_FRESH)
    s.net_init(seed=1)
    curve, ms = s.net_train(steps=1500, batch=4096, lr=2e-3, seed=2)
    say(f"  loss {curve[0]:.3e} -> {curve[-50:].mean():.3e} in {ms:.0f} ms")
    assert curve[-50:].mean() < 0.1 * curve[0]
    # unseen pairs
    A2, B2 = sq.overlapping_pairs(25, 4000)
    s.set_pairs(A2, B2)
    p0 = s.solve(pydem.DEM_SQ_FRESH)
    w = s.solve(pydem.DEM_SQ_NETWORK)
    ratio = p0["iters"].mean() / w["iters"].mean()
    fb = ((w["flags"] & 2) != 0).mean()
    conv = (_merit(A2, B2, p0["point"]) < 1e-10) & (_merit(A2, B2, w["point"]) < 1e-10)
    d = np.linalg.norm(w["point"] - p0["point"], axis=1)[conv]
    say(f"  iterations plain {p0['iters'].mean():.2f} warm {w['iters'].mean():.2f} ({ratio:.2f}x), "
        f"fallbacks {fb:.3f}, |dp| median {np.median(d):.2e} q99 {np.quantile(d, 0.99):.2e}")
    assert ratio >= 2.0
    assert np.median(d) < 1e-9
    # past the reference's merit tolerance: both polished by the same Newton
    # refinement, the warm-started points equal the plain ones to 1e-9 m
    p1 = s.solve(pydem.DEM_SQ_FRESH, polish=4)
    w1 = s.solve(pydem.DEM_SQ_NETWORK, polish=4)
    ok = ((p1["flags"] & 1) == 0) & ((w1["flags"] & 1) == 0)
    d1 = np.linalg.norm(w1["point"] - p1["point"], axis=1)[ok]
    say(f"  polished: iterations plain {p1['iters'].mean():.2f} warm {w1['iters'].mean():.2f}, "
        f"|dp| q99 {np.quantile(d1, 0.99):.2e} max {d1.max():.2e}, within 1e-9 m {(d1 <= 1e-9).mean():.4f}")
    assert (d1 <= 1e-9).mean() >= 0.99
    assert p1["iters"].mean() / w1["iters"].mean() >= 2.0
    # contact detection of new pairs from the network: the same contacts and
    # overlaps as from the homotopy, with fewer Newton iterations
    cf = s.check(None, start=pydem.DEM_SQ_FRESH)
    cw = s.check(None, start=pydem.DEM_SQ_NETWORK)
    both = (cf["contact"] == 1) & (cw["contact"] == 1)
    ddn = np.abs(cw["deltan"] - cf["deltan"])[both]
    say(f"  contact detection: contacts {int(cf['contact'].sum())} / {int(cw['contact'].sum())}, "
        f"iterations {cf['iters'].mean():.2f} -> {cw['iters'].mean():.2f}, |d deltan| median {np.median(ddn):.2e}")
    assert (cf["contact"] == cw["contact"]).mean() >= 0.995
    assert np.median(ddn) < 1e-9
    assert cf["iters"].mean() / cw["iters"].mean() >= 2.0
    s.close()
