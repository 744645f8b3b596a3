"""CPU-side checks of the superquadric commands of the input-script front end
(atom_style superquadric, its read_data format, surface superquadric,
fix nve/superquadric, set ... quat / quat/random), and of the random
orientations set ... quat/random draws.

quat/random follows the reference: a Park-Miller minimal standard generator
(random_park.cpp:54-80) seeded per atom by the Jenkins one-at-a-time hash of
the seed and the atom's coordinates (random_park.cpp:125-157). The
restatement below is pinned by Park and Miller's published check value (the
10,000th draw from seed 1 is 1,043,618,065) and by one-at-a-time hash values
of ASCII strings; the GPU test compares the script's run with an API run
whose orientations come from this restatement."""
import math
import struct

import numpy as np
import pytest

import pydem
import workloads as wl

HEADER = """atom_style superquadric
atom_modify map array
boundary p p f
newton off
communicate single vel yes
units si
neighbor 0.001 bin
neigh_modify delay 0
fix m1 all property/global youngsModulus peratomtype 5.e6
fix m2 all property/global poissonsRatio peratomtype 0.45
fix m3 all property/global coefficientRestitution peratomtypepair 1 0.3
fix m4 all property/global coefficientFriction peratomtypepair 1 0.5
pair_style gran model hertz tangential history surface superquadric
pair_coeff * *
timestep 0.00001
fix gravi all gravity 9.81 vector 0.0 0.0 -1.0
fix zwall all wall/gran model hertz tangential history surface superquadric primitive type 1 zplane 0.0
"""


class RanPark:
    IA, IM, IQ, IR = 16807, 2147483647, 127773, 2836

    def __init__(self, seed=1):
        self.seed = seed

    def uniform(self):
        k = int(self.seed / self.IQ)
        self.seed = self.IA * (self.seed - k * self.IQ) - self.IR * k
        if self.seed < 0:
            self.seed += self.IM
        return (1.0 / self.IM) * self.seed

    def reset(self, ibase, coord):
        data = struct.pack("<i", ibase) + struct.pack("<3d", *coord)
        self.seed = one_at_a_time(data, signed=True, final=True) & 0x7ffffff
        if not self.seed:
            self.seed = 1
        for _ in range(5):
            self.uniform()


def one_at_a_time(data, signed=False, final=True):
    h = 0
    for b in data:
        if signed and b >= 128:
            b -= 256  # char is signed on x86-64
        h = (h + b) & 0xFFFFFFFF
        h = (h + (h << 10)) & 0xFFFFFFFF
        h ^= h >> 6
    if final:
        h = (h + (h << 3)) & 0xFFFFFFFF
        h ^= h >> 11
        h = (h + (h << 15)) & 0xFFFFFFFF
    return h


def quat_random(seed, x):
    """set ... quat/random for atoms at positions x (set.cpp:1195-1225)."""
    r = RanPark()
    out = np.zeros((len(x), 4))
    for i, xi in enumerate(x):
        r.reset(seed, xi)
        s = r.uniform()
        t1, t2 = math.sqrt(1.0 - s), math.sqrt(s)
        th1 = 2.0 * 3.14159265358979323846 * r.uniform()
        th2 = 2.0 * 3.14159265358979323846 * r.uniform()
        out[i] = [math.cos(th2) * t2, math.sin(th1) * t1, math.cos(th1) * t1, math.sin(th2) * t2]
    return out


def test_park_miller_check_value():
    r = RanPark(1)
    for _ in range(10000):
        r.uniform()
    assert r.seed == 1043618065


def test_one_at_a_time_hash_values():
    assert one_at_a_time(b"a") == 0xCA2E9442
    assert one_at_a_time(b"The quick brown fox jumps over the lazy dog") == 0x519E91F5


def test_quat_random_uniform_on_the_sphere():
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 0.1, (4000, 3))
    q = quat_random(12345, x)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-14)
    # uniform orientations: each component squared has mean 1/4
    assert np.allclose((q ** 2).mean(0), 0.25, atol=0.02)


def _gas(n=200):
    params, mats, P, walls = wl.superquadric_settling(5, n, base=(0.04, 0.04), axes=(0.001, 0.002), gap=0.0003)
    P["v"] = np.random.default_rng(1).normal(0, 0.3, (n, 3))
    return params, mats, P, walls


def run_script(tmp_path, text, name="in.sq"):
    p = tmp_path / name
    p.write_text(text)
    log = tmp_path / (name + ".log")
    rc = pydem.run_script(str(p), str(log))
    return rc, log.read_text() if log.exists() else ""


def test_superquadric_commands_parse(tmp_path):
    params, mats, P, walls = _gas()
    data = tmp_path / "sq.data"
    wl.write_superquadric_data(str(data), params, P)
    rc, log = run_script(tmp_path, HEADER + f"read_data {data}\nset group all quat/random 4711\n"
                         "set atom 3 quat 0 0 1 45\nfix integr all nve/superquadric integration_scheme 1\n")
    assert rc == 0, log
    assert "200 atoms" in log and "200 settings made for quat/random" in log


@pytest.mark.parametrize("extra,msg", [
    ("fix integr all nve/sphere\n", "requires atom style sphere"),
    ("fix integr all nve/superquadric integration_scheme 2\n", "integration_scheme 1"),
    ("set group all dipole 1 0 0\n", "not supported"),
])
def test_superquadric_errors(tmp_path, extra, msg):
    params, mats, P, walls = _gas(20)
    data = tmp_path / "sq.data"
    wl.write_superquadric_data(str(data), params, P)
    rc, log = run_script(tmp_path, HEADER + f"read_data {data}\n" + extra)
    assert rc != 0 and msg in log, log


def test_surface_superquadric_needs_the_atom_style(tmp_path):
    rc, log = run_script(tmp_path, "atom_style granular\n"
                         "pair_style gran model hertz tangential history surface superquadric\n")
    assert rc != 0 and "requires atom_style superquadric" in log
