"""Particle insertion: the oracle's fix insert/pack restatement pinned by the
generator's published check value, closed forms of the distribution and of
the region volume, the positions the documented random streams fix, and the
invariants of a packing (no overlaps, all_in margins, uniform positions);
then the library's host implementation (dem_inserter_*) equal to the oracle
bit for bit. No GPU is used."""
import numpy as np
import pytest
from scipy import stats

import pydem
import pyoracle

PI = 3.14159265358979323846
BOX = ((-0.05, -0.05, 0.0), (0.05, 0.05, 0.15))
# examples/LIGGGHTS/Tutorials_public/packing/in.packing
PACKING = dict(templates=[(15485863, 1, 0.005, 2500.0), (15485867, 1, 0.008, 2500.0)], weights=[0.3, 0.7],
               dist_seed=32452843, seed=32452867, maxattempt=200, all_in=1, volumefraction_region=0.25,
               region=BOX, box=BOX, box_change=1)


def park(seed):
    """Park-Miller minimal standard generator written from Park & Miller (1988)
    with 64-bit integers (no Schrage decomposition)."""
    s = seed
    while True:
        s = (16807 * s) % 2147483647
        yield s / 2147483647.0


def draws(seed, skip, n):
    g = park(seed)
    for _ in range(skip):
        next(g)
    return np.array([next(g) for _ in range(n)])


def test_park_miller_check_value():
    # Park & Miller's check: from seed 1, the 10,000th value is 1043618065
    u, seed = pyoracle.park_uniforms(1, 10000)
    assert seed == 1043618065
    assert np.allclose(u[:100], draws(1, 0, 100), rtol=0, atol=1e-15)


def test_gaussian_pairs_share_their_radius():
    r = pyoracle.OriPark()
    lib = pyoracle.load()
    lib.ori_park_init(pyoracle.ctypes.byref(r), 12345)
    g = np.array([lib.ori_park_gaussian(pyoracle.ctypes.byref(r)) for _ in range(40000)])
    assert stats.kstest(g, "norm").pvalue > 1e-3
    # polar Box-Muller: both values of a pair come from one (v1, v2), so their
    # squares add up to -2 ln(rsq), exponentially distributed with mean 2
    s2 = g[0::2] ** 2 + g[1::2] ** 2
    assert stats.kstest(s2, "expon", args=(0, 2)).pvalue > 1e-3


def test_distribution_closed_forms():
    o = pyoracle.OracleInserter(**PACKING)
    ve, me, w, order = o.distribution()
    m = [2500.0 * r ** 3 * 4 * PI / 3 for r in (0.005, 0.008)]
    n = np.array([0.3 / m[0], 0.7 / m[1]])
    n /= n.sum()
    assert np.allclose(w, n, rtol=1e-14)
    assert np.isclose(ve, n[0] * 0.005 ** 3 * 4 * PI / 3 + n[1] * 0.008 ** 3 * 4 * PI / 3, rtol=1e-14)
    assert np.isclose(me, 2500.0 * ve, rtol=1e-14)
    assert list(order) == [1, 0]  # the larger template is inserted first
    # number-based weights are taken as given
    o2 = pyoracle.OracleInserter(**dict(PACKING, dist_numberbased=1))
    assert np.allclose(o2.distribution()[2], [0.3, 0.7], rtol=1e-15)


def test_region_volume_monte_carlo():
    o = pyoracle.OracleInserter(**PACKING)
    assert o.region_volume() == 0.1 * 0.1 * 0.15  # every test point falls inside
    # half of the region outside the box: about half the volume
    kw = dict(PACKING, region=((0.0, -0.05, 0.0), (0.1, 0.05, 0.15)), all_in=0)
    v = pyoracle.OracleInserter(**kw).region_volume()
    full = 0.1 * 0.1 * 0.15
    assert abs(v / full - 0.5) < 4 * np.sqrt(0.25 / 100000)


@pytest.mark.parametrize("box_change", [0, 1])
def test_first_positions_follow_the_region_stream(box_change):
    """No overlap check, all_in no, region = box: the count is
    int(vf V / volexpect + u) with u the first draw of the fix's stream (seed),
    and the k-th particle sits at lo + u * L with the region stream (seed + 12)
    after the Monte Carlo volume's 3 x 100,000 draws (twice when the box can
    change); large particles come first."""
    kw = dict(PACKING, overlapcheck=0, all_in=0, box_change=box_change, volumefraction_region=0.05)
    o = pyoracle.OracleInserter(**kw)
    out, req = o.step()
    ve = o.distribution()[0]
    u_fix = draws(kw["seed"], 0, 1)[0]
    assert req == int((0.05 * 0.1 * 0.1 * 0.15) / ve + u_fix)
    assert len(out["x"]) == req
    skip = 3 * 100000 * (1 + box_change)
    u = draws(kw["seed"] + 12, skip, 3 * 5).reshape(5, 3)
    lo, hi = np.array(BOX[0]), np.array(BOX[1])
    assert np.array_equal(out["x"][:5], lo + u * (hi - lo))
    nl = int(np.sum(out["radius"] == 0.008))
    assert np.all(out["radius"][:nl] == 0.008) and np.all(out["radius"][nl:] == 0.005)
    r = out["radius"]
    assert np.array_equal(out["rmass"], 2500.0 * (r * r * r * 4 * PI / 3))


def _check_packing(out, lo, hi, all_in):
    x, r = out["x"], out["radius"]
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    np.fill_diagonal(d, np.inf)
    assert np.all(d > r[:, None] + r[None, :])  # no two inserted particles touch
    if all_in:
        assert np.all(x - r[:, None] >= lo) and np.all(x + r[:, None] <= hi)
    else:
        assert np.all(x >= lo) and np.all(x <= hi)


def test_packing_invariants_and_volume_fraction():
    o = pyoracle.OracleInserter(**PACKING)
    out, req = o.step()
    assert len(out["x"]) == req  # room enough at 25 %
    _check_packing(out, np.array(BOX[0]), np.array(BOX[1]), True)
    vol = (out["radius"] ** 3).sum() * 4 * PI / 3
    # the requested volume fraction up to the rounding of one particle
    assert abs(vol / (0.1 * 0.1 * 0.15) - 0.25) < 0.008 ** 3 * 4 * PI / 3 / (0.1 * 0.1 * 0.15)
    # the centres are uniform over the shrunken region along each axis
    for d, (a, b) in enumerate(zip(BOX[0], BOX[1])):
        small = out["x"][out["radius"] == 0.005, d]
        assert stats.kstest((small - (a + 0.005)) / (b - a - 0.01), "uniform").pvalue > 1e-3


def test_maxattempt_limits_a_crowded_region():
    kw = dict(PACKING, volumefraction_region=0.6, maxattempt=5)
    out, req = pyoracle.OracleInserter(**kw).step()
    assert len(out["x"]) < req
    out2, _ = pyoracle.OracleInserter(**dict(kw, maxattempt=50)).step()
    assert len(out2["x"]) > len(out["x"])
    _check_packing(out2, np.array(BOX[0]), np.array(BOX[1]), True)


def _existing(n=300, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.uniform((-0.045, -0.045, 0.005), (0.045, 0.045, 0.145), (n, 3))
    keep = [0]
    for i in range(1, n):  # non-overlapping 3 mm spheres
        if np.all(np.linalg.norm(x[keep] - x[i], axis=1) > 0.0061):
            keep.append(i)
    x = x[keep]
    r = np.full(len(x), 0.003)
    return x, r, 2500.0 * r ** 3 * 4 * PI / 3


def test_particles_in_region_counts_the_present_ones():
    x, r, m = _existing()
    reg = ((-0.05, -0.05, 0.0), (0.05, 0.05, 0.075))
    inside = np.all((x >= reg[0]) & (x <= reg[1]), axis=1).sum()
    kw = dict(templates=[(15485863, 1, 0.004, 2500.0)], weights=[1.0], dist_seed=15485867, seed=32452843,
              particles_in_region=inside + 40, region=reg, box=BOX, all_in=1, vel_style=0,
              v_insert=(0.0, 0.0, -0.8))
    out, req = pyoracle.OracleInserter(**kw).step(x, r, m)
    assert req == 40 and len(out["x"]) == 40
    assert np.all(out["v"] == (0.0, 0.0, -0.8))
    # nothing overlaps the present particles either
    d = np.sqrt(((out["x"][:, None, :] - x[None, :, :]) ** 2).sum(-1))
    assert np.all(d > 0.004 + 0.003)


def test_random_velocities():
    kw = dict(PACKING, vel_style=1, v_insert=(0.1, 0.0, -1.0), v_fluct=(0.05, 0.2, 0.3))
    out, _ = pyoracle.OracleInserter(**kw).step()
    v = out["v"]
    lo, hi = np.array([0.05, -0.2, -1.3]), np.array([0.15, 0.2, -0.7])
    assert np.all(v >= lo) and np.all(v <= hi)
    kw = dict(kw, vel_style=2)
    g = pyoracle.OracleInserter(**kw).step()[0]["v"]
    z = (g - np.array([0.1, 0.0, -1.0])) / np.array([0.05, 0.2, 0.3])
    assert stats.kstest(z.reshape(-1), "norm").pvalue > 1e-3


CASES = {
    "packing": PACKING,
    "no_overlapcheck": dict(PACKING, overlapcheck=0, all_in=0),
    "uniform_velocity_numberbased": dict(PACKING, vel_style=1, v_fluct=(0.1, 0.1, 0.1), dist_numberbased=1,
                                         box_change=0),
    "gaussian_velocity_mass_target": dict(PACKING, vel_style=2, v_fluct=(0.1, 0.1, 0.1),
                                          volumefraction_region=0.0, mass_in_region=0.5),
    "uncorrelated": dict(PACKING, exact_number=0),
    "region_past_the_box": dict(PACKING, region=((0.0, -0.05, 0.0), (0.1, 0.05, 0.15)), all_in=0),
    "three_templates": dict(PACKING, templates=[(15485863, 1, 0.003, 2500.0), (15485867, 2, 0.004, 1500.0),
                                                (49979687, 1, 0.004, 3000.0)],
                            weights=[0.2, 0.5, 0.3], volumefraction_region=0.3, maxattempt=100),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_inserter_equals_oracle(name):
    kw = CASES[name]
    o = pyoracle.OracleInserter(**kw)
    g = pydem.Inserter(**kw)
    assert g.info()["region_volume"] == o.region_volume()
    x = r = m = None
    for _ in range(2):  # a second insertion on top of the first
        a, ra = o.step(x, r, m)
        b, rb = g.step(x, r, m)
        assert ra == rb
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        x = a["x"] if x is None else np.vstack([x, a["x"]])
        r = a["radius"] if r is None else np.concatenate([r, a["radius"]])
        m = a["rmass"] if m is None else np.concatenate([m, a["rmass"]])


def test_library_inserter_with_present_particles_equals_oracle():
    x, r, m = _existing()
    kw = dict(templates=[(15485863, 1, 0.004, 2500.0)], weights=[1.0], dist_seed=15485867, seed=32452843,
              particles_in_region=500, region=((-0.05, -0.05, 0.0), (0.05, 0.05, 0.1)), box=BOX, all_in=1,
              v_insert=(0.0, 0.0, -0.8))
    a, _ = pyoracle.OracleInserter(**kw).step(x, r, m)
    b, _ = pydem.Inserter(**kw).step(x, r, m)
    assert len(a["x"]) > 0
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_library_inserter_errors():
    with pytest.raises(pydem.DemError, match="exactly one"):
        pydem.Inserter(**dict(PACKING, particles_in_region=10))
    g = pydem.Inserter(**dict(PACKING, region=((0.0, -0.05, 0.0), (0.1, 0.05, 0.15)), fixed_boundary=1))
    with pytest.raises(pydem.DemError, match="extends outside"):
        g.step()
