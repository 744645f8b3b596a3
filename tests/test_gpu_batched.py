"""GPU checks of independent simulations batched in one system
(dem_set_instances): each instance of a heap-formation sweep, advanced
together with the others in one step kernel, against the same instance run
alone.

  same build schedule (rebuild every step)  bitwise equal to the instance
                                            alone; 1e-8 against the oracle
  default schedule (skin/2 check)           kinetic energy and contact
                                            counts within 1 % of the
                                            instance alone in FP64
"""
import sys

import numpy as np
import pytest

import pydem
import pyoracle as orc
import workloads as wl

pytestmark = pytest.mark.gpu

COUNTS = [300, 800, 1500, 500, 1200]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sweep(every_step):
    params, base, P, walls, meshes, counts, imats = wl.heap_sweep(seed=7, counts=COUNTS)
    if every_step:
        params = dict(params, neigh_every=1, neigh_check=False)
    # random velocities (collisions inside the lattice from the first steps)
    # and a downward speed so the columns reach the floor within the test
    rng = np.random.default_rng(11)
    P = dict(P)
    P["v"] = rng.normal(0.0, 0.2, P["v"].shape)
    P["v"][:, 2] -= 0.4
    return params, base, P, walls, meshes, counts, imats


def _ke(P, v):
    return 0.5 * float(P["rmass"] @ (v ** 2).sum(1))


def test_batched_instances_equal_alone_bitwise():
    """With the same build schedule (every step), batching changes nothing:
    every instance's state is bitwise the state of the instance run alone,
    and the per-instance contact counts are equal."""
    say("[gpu] batched == alone (bitwise, rebuild every step)")
    params, base, P, walls, meshes, counts, imats = _sweep(True)
    nsteps = 200
    g = pydem.GpuSystem(params, base, P, walls, meshes=meshes, instances=counts,
                        instance_materials=imats)
    g.run(nsteps)
    gs = g.state()
    gc, gw = g.instance_stats()
    off = 0
    for q, n in enumerate(counts):
        Pq = wl.instance_slice(P, counts, q)
        a = pydem.GpuSystem(params, imats[q], Pq, walls, meshes=meshes)
        a.run(nsteps)
        st = a.state()
        st_stats = a.stats()
        for k in ("x", "v", "omega"):
            assert np.array_equal(gs[k][off:off + n], st[k]), (q, k)
        say(f"  instance {q}: n={n} contacts {gc[q]} / {st_stats['ncontacts']} "
            f"wall {gw[q]} / {st_stats['nwall_contacts']}")
        assert gc[q] == st_stats["ncontacts"] and gw[q] == st_stats["nwall_contacts"]
        a.close()
        off += n
    assert gw.sum() > 50 and gc.sum() > 50
    g.close()


def test_batched_instance_matches_oracle_fp64():
    """One instance of the batch against the oracle run alone with that
    instance's materials (same build schedule): state within 1e-8."""
    say("[gpu] batched instance vs oracle")
    params, base, P, walls, meshes, counts, imats = _sweep(True)
    nsteps = 100
    g = pydem.GpuSystem(params, base, P, walls, meshes=meshes, instances=counts,
                        instance_materials=imats)
    g.run(nsteps)
    gs = g.state()
    q = 3
    off = int(np.sum(counts[:q]))
    n = int(counts[q])
    o = orc.OracleSystem(params, imats[q], wl.instance_slice(P, counts, q), walls, meshes=meshes)
    o.run(nsteps)
    os_ = o.state()
    for k in ("x", "v", "omega"):
        a, b = gs[k][off:off + n], os_[k]
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        say(f"  {k}: {err:.2e}")
        assert err <= 1e-8, k
    g.close()


@pytest.mark.parametrize("precision", [pydem.DEM_FP64, pydem.DEM_MIXED])
def test_batched_statistics_vs_alone_fp64(precision):
    """Default build schedule (a rebuild when any instance moved skin/2):
    per-instance kinetic energy and contact counts within 1 % of the instance
    run alone in FP64."""
    say(f"[gpu] batched statistics precision={precision}")
    params, base, P, walls, meshes, counts, imats = _sweep(False)
    nsteps = 500
    g = pydem.GpuSystem(params, base, P, walls, meshes=meshes, instances=counts,
                        instance_materials=imats, precision=precision)
    g.run(nsteps)
    gs = g.state()
    gc, gw = g.instance_stats()
    off = 0
    tot_c = 0
    for q, n in enumerate(counts):
        Pq = wl.instance_slice(P, counts, q)
        a = pydem.GpuSystem(params, imats[q], Pq, walls, meshes=meshes)
        a.run(nsteps)
        st = a.state()
        s2 = a.stats()
        ke_b = _ke(Pq, gs["v"][off:off + n])
        ke_a = _ke(Pq, st["v"])
        say(f"  instance {q}: KE {ke_b:.6e} / {ke_a:.6e}  contacts {gc[q]} / {s2['ncontacts']}  "
            f"wall {gw[q]} / {s2['nwall_contacts']}")
        assert ke_b == pytest.approx(ke_a, rel=0.01)
        assert gc[q] == pytest.approx(s2["ncontacts"], rel=0.01, abs=3)
        assert gw[q] == pytest.approx(s2["nwall_contacts"], rel=0.01, abs=3)
        tot_c += gc[q]
        a.close()
        off += n
    assert tot_c > 200
    g.close()
