"""Superquadric pose pairs and the contact-point network's encoding.

A pose is 12 doubles: centre (3), unit quaternion (w, i, j, k) (4), half-axes
a, b, c (3) and blockiness n1, n2 (2) -- the per-atom shape, blockiness and
quaternion of atom_style superquadric.

overlapping_pairs() draws seeded random pairs that overlap along their line of
centres: A at the origin, B along a random direction at a distance a little
below the sum of the two centre-to-surface distances along that line (closed
form for a superquadric ray from its centre).

features() / to_point() / from_point() define what the warm-start network
sees and predicts: 18 inputs (B's centre in A's body frame, the relative
quaternion, both shapes relative to their volume-equivalent radii, the
blockiness scaled by 1/8, and the radius ratio), and 3 outputs (the contact
point in A's body frame, in units of the sum of the equivalent radii).
"""
import numpy as np

N_IN = 18
N_OUT = 3


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def quat_to_mat(q):
    """Rotation matrices of quaternions (w, i, j, k), body -> global."""
    w, i, j, k = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = w * w + i * i - j * j - k * k
    R[:, 0, 1] = 2 * i * j - 2 * k * w
    R[:, 0, 2] = 2 * j * w + 2 * i * k
    R[:, 1, 0] = 2 * i * j + 2 * k * w
    R[:, 1, 1] = w * w - i * i + j * j - k * k
    R[:, 1, 2] = 2 * j * k - 2 * i * w
    R[:, 2, 0] = 2 * i * k - 2 * j * w
    R[:, 2, 1] = 2 * j * k + 2 * i * w
    R[:, 2, 2] = w * w - i * i - j * j + k * k
    return R


def quat_mul(a, b):
    w1, x1, y1, z1 = a.T
    w2, x2, y2, z2 = b.T
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=1)


def quat_conj(q):
    return q * np.array([1.0, -1.0, -1.0, -1.0])


def ray_extent(shape, blk, u_local):
    """Distance from the centre to the surface along unit body-frame
    directions u_local: F(t u) = 0 solved in closed form."""
    a, b, c = shape[:, 0], shape[:, 1], shape[:, 2]
    n1, n2 = blk[:, 0], blk[:, 1]
    s = (np.abs(u_local[:, 0] / a) ** n2 + np.abs(u_local[:, 1] / b) ** n2) ** (n1 / n2) \
        + np.abs(u_local[:, 2] / c) ** n1
    return s ** (-1.0 / n1)


def overlapping_pairs(seed, n, axes=(0.001, 0.003), blockiness=(2.0, 8.0), overlap=(0.0, 0.15)):
    """n seeded pose pairs (A, B), each (n, 12); B's centre sits at
    (1 - s) * (extent_A + extent_B) along a random direction, s ~ U(overlap)."""
    rng = np.random.default_rng(seed)
    shA = rng.uniform(*axes, size=(n, 3))
    shB = rng.uniform(*axes, size=(n, 3))
    bkA = rng.uniform(*blockiness, size=(n, 2))
    bkB = rng.uniform(*blockiness, size=(n, 2))
    qA = random_quats(rng, n)
    qB = random_quats(rng, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    RA, RB = quat_to_mat(qA), quat_to_mat(qB)
    uA = np.einsum("nji,nj->ni", RA, u)    # R^T u
    uB = np.einsum("nji,nj->ni", RB, -u)
    d = (ray_extent(shA, bkA, uA) + ray_extent(shB, bkB, uB)) * (1.0 - rng.uniform(*overlap, size=n))
    A = np.concatenate([np.zeros((n, 3)), qA, shA, bkA], axis=1)
    B = np.concatenate([u * d[:, None], qB, shB, bkB], axis=1)
    return A, B


def equiv_radius(shape):
    return np.cbrt(shape[:, 0] * shape[:, 1] * shape[:, 2])


def features(A, B):
    """(n, 18) float32 network inputs of pose pairs."""
    rA, rB = equiv_radius(A[:, 7:10]), equiv_radius(B[:, 7:10])
    L = (rA + rB)[:, None]
    RA = quat_to_mat(A[:, 3:7])
    rel = np.einsum("nji,nj->ni", RA, B[:, 0:3] - A[:, 0:3]) / L
    qr = quat_mul(quat_conj(A[:, 3:7]), B[:, 3:7])
    qr *= np.where(qr[:, :1] < 0.0, -1.0, 1.0)
    X = np.concatenate([rel, qr, A[:, 7:10] / rA[:, None], A[:, 10:12] / 8.0,
                        B[:, 7:10] / rB[:, None], B[:, 10:12] / 8.0,
                        (rA / (rA + rB))[:, None]], axis=1)
    return X.astype(np.float32)


def to_point(A, B, y):
    """Global contact points from network outputs (A's body frame, scaled)."""
    rA, rB = equiv_radius(A[:, 7:10]), equiv_radius(B[:, 7:10])
    RA = quat_to_mat(A[:, 3:7])
    return A[:, 0:3] + np.einsum("nij,nj->ni", RA, np.asarray(y, dtype=np.float64) * (rA + rB)[:, None])


def from_point(A, B, p):
    """Network targets of global contact points p."""
    rA, rB = equiv_radius(A[:, 7:10]), equiv_radius(B[:, 7:10])
    RA = quat_to_mat(A[:, 3:7])
    return (np.einsum("nji,nj->ni", RA, p - A[:, 0:3]) / (rA + rB)[:, None]).astype(np.float32)
