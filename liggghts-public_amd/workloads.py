"""Seeded generators of the granular workloads, shared by the GPU library's
tests, the CPU oracle and bench.py, plus a writer of the reference's
granular data-file format (read_data Atoms ``id type diameter density x y z``
and Velocities ``id vx vy vz wx wy wz``, atom_vec_sphere.cpp:1055-1140).

Every generator returns ``(params, materials, particles, walls)``:
  params     run parameters (dt, skin, box, periodicity, model switches, gravity)
  materials  per-type Young's modulus / Poisson ratio and per-type-pair
             restitution / friction / rolling friction
  particles  x, v, omega (n,3), radius, rmass (n,), type, tag (n,) int32
  walls      [(axis, position, wall_type)] plane walls
"""
import math

import numpy as np

MY_PI = 3.14159265358979323846


def sphere_mass(radius, density):
    """rmass as read_data computes it (atom_vec_sphere.cpp:1081-1082)."""
    return 4.0 * MY_PI / 3.0 * radius * radius * radius * density


def materials_single(youngs=5e6, poisson=0.45, restitution=0.3, friction=0.5,
                     rolling_friction=None):
    m = {"ntypes": 1, "youngs": [youngs], "poisson": [poisson],
         "restitution": [[restitution]], "friction": [[friction]]}
    if rolling_friction is not None:
        m["rolling_friction"] = [[rolling_friction]]
    return m


def lattice(n, box_lo, box_hi, spacing, jitter, rng, z0=None):
    """First n sites of a simple cubic lattice filling the box bottom-up,
    each moved by U(-jitter, +jitter) per component."""
    nx = int((box_hi[0] - box_lo[0]) / spacing + 1e-9)
    ny = int((box_hi[1] - box_lo[1]) / spacing + 1e-9)
    per_layer = nx * ny
    nz = (n + per_layer - 1) // per_layer
    k = np.arange(n)
    ix, iy, iz = k % nx, (k // nx) % ny, k // per_layer
    if z0 is None:
        z0 = box_lo[2] + 0.5 * spacing
    x = np.stack([box_lo[0] + (ix + 0.5) * spacing,
                  box_lo[1] + (iy + 0.5) * spacing,
                  z0 + iz * spacing], axis=1)
    if z0 + (nz - 1) * spacing > box_hi[2]:
        raise ValueError("lattice does not fit in the box")
    x += rng.uniform(-jitter, jitter, size=x.shape)
    return x


def settling_4k(seed=0, n=4000):
    """Plain FP64 Hertz/history sphere settling, CPU-runnable: 4,000 monodisperse
    spheres r=2 mm, density 2500, Y=5e6, nu=0.45, e=0.3, mu=0.5, box 0.1x0.1x0.2
    periodic in x and y, zplane wall at 0, gravity -9.81 z, cubic lattice 4.4 mm
    + U(+-0.1 mm), v ~ N(0, 0.01 m/s), dt=1e-5, skin 1 mm."""
    rng = np.random.default_rng(seed)
    lo, hi = (0.0, 0.0, 0.0), (0.1, 0.1, 0.2)
    r = np.full(n, 0.002)
    x = lattice(n, lo, hi, 0.0044, 0.0001, rng)
    v = rng.normal(0.0, 0.01, size=(n, 3))
    params = {"dt": 1e-5, "skin": 0.001, "cdf": 1.0, "boxlo": lo, "boxhi": hi,
              "periodic": (1, 1, 0), "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1,
              "rolling_cdt": 0}
    particles = {"x": x, "v": v, "omega": np.zeros((n, 3)), "radius": r,
                 "rmass": sphere_mass(r, 2500.0), "type": np.ones(n, dtype=np.int32),
                 "tag": np.arange(1, n + 1, dtype=np.int32)}
    return params, materials_single(), particles, [(2, 0.0, 1)]


def settling_1m(seed=0, n=1_000_000):
    """1M polydisperse spheres settling: radii U(1.0, 1.5) mm, materials as
    settling_4k plus CDT rolling friction 0.1, box 0.3x0.3x0.6 periodic in x and y,
    zplane floor, jittered 3.3 mm lattice, dt=5e-6, skin 1 mm."""
    rng = np.random.default_rng(seed)
    lo, hi = (0.0, 0.0, 0.0), (0.3, 0.3, 0.6)
    r = rng.uniform(0.001, 0.0015, size=n)
    x = lattice(n, lo, hi, 0.0033, 0.0001, rng)
    v = rng.normal(0.0, 0.01, size=(n, 3))
    params = {"dt": 5e-6, "skin": 0.001, "cdf": 1.0, "boxlo": lo, "boxhi": hi,
              "periodic": (1, 1, 0), "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1,
              "rolling_cdt": 1}
    particles = {"x": x, "v": v, "omega": np.zeros((n, 3)), "radius": r,
                 "rmass": sphere_mass(r, 2500.0), "type": np.ones(n, dtype=np.int32),
                 "tag": np.arange(1, n + 1, dtype=np.int32)}
    return params, materials_single(rolling_friction=0.1), particles, [(2, 0.0, 1)]


def dense_pack(seed=0, n=2000, rolling=True, overlap=0.02, box=None, periodic=(1, 1, 0),
               spin=20.0, speed=0.05, radius=(0.001, 0.0015)):
    """A compressed polydisperse pack for single-step parity checks: lattice
    spacing shorter than the mean diameter, so most neighbors overlap by a few
    percent of a radius; random translational and angular velocities so every
    term of the contact law (damping, sliding, rolling) is exercised."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(radius[0], radius[1], size=n)
    spacing = (radius[0] + radius[1]) * (1.0 - overlap)
    side = int(math.ceil(n ** (1.0 / 3.0)))
    if box is None:
        L = side * spacing
        box = ((0.0, 0.0, 0.0), (L, L, L * 1.5))
    lo, hi = box
    x = lattice(n, lo, hi, spacing, 0.05 * spacing, rng, z0=lo[2] + 0.5 * spacing)
    v = rng.normal(0.0, speed, size=(n, 3))
    w = rng.normal(0.0, spin, size=(n, 3))
    # boundary m on the non-periodic faces (as the packing tutorial): the
    # compressed pack spreads, and an f face would lose the spheres it pushes out
    faces = sum(3 << (2 * d) for d in range(3) if not periodic[d])
    params = {"dt": 5e-6, "skin": 0.0005, "cdf": 1.0, "boxlo": lo, "boxhi": hi,
              "periodic": periodic, "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1,
              "rolling_cdt": int(rolling), "shrink_faces": faces, "shrink_min": faces}
    particles = {"x": x, "v": v, "omega": w, "radius": r,
                 "rmass": sphere_mass(r, 2500.0), "type": np.ones(n, dtype=np.int32),
                 "tag": np.arange(1, n + 1, dtype=np.int32)}
    mats = materials_single(rolling_friction=0.1 if rolling else None)
    return params, mats, particles, [(2, lo[2], 1)]


def packed_1m(seed=0, n=1_000_000, overlap=0.01):
    """The 1M polydisperse settling spheres (radii U(1.0, 1.5) mm, CDT rolling
    friction 0.1, dt 5e-6 s, skin 1 mm, periodic in x and y, zplane floor)
    started as a compressed lattice instead of a dilute one: the lattice
    spacing is 1 % below the mean diameter, so about three neighbors of every
    sphere touch and the step carries tangential history for each of them."""
    params, mats, P, walls = dense_pack(seed=seed, n=n, rolling=True, overlap=overlap,
                                        radius=(0.001, 0.0015), speed=0.01, spin=1.0)
    params = dict(params, skin=0.001)
    return params, mats, P, walls


def materials_two(mu_pp=0.5, mu_pw=0.4, restitution=0.3, youngs=5e6, poisson=0.45,
                  rolling_friction=None):
    """Type 1 = particles, type 2 = wall (fix mesh/surface ... type 2)."""
    m = {"ntypes": 2, "youngs": [youngs, youngs], "poisson": [poisson, poisson],
         "restitution": [[restitution] * 2] * 2,
         "friction": [[mu_pp, mu_pw], [mu_pw, mu_pw]]}
    if rolling_friction is not None:
        m["rolling_friction"] = [[rolling_friction] * 2] * 2
    return m


def drum(seed=0, radius=0.15, length=0.1, n_az=64, n_ax=16, r=0.001, spread=0.1, fill=0.5,
         spacing=None, period=2.0, dt=5e-6, skin=0.0005, max_particles=None):
    """Rotating drum with mesh walls: closed cylinder (axis along y) of
    n_az x n_ax mantle quads split into triangles plus fan end caps, turning at
    2*pi/period about its axis (fix move/mesh rotate); spheres of radius
    r*(1 +- spread) on a jittered cubic lattice filling the lower `fill`
    fraction of the drum volume; mu_pp 0.5, mu_pw 0.4, e 0.3, gravity -z."""
    import meshes as msh
    rng = np.random.default_rng(seed)
    rmax = r * (1.0 + spread)
    a = spacing or 2.0 * rmax * 1.02
    margin = rmax + 0.25 * a
    xs = np.arange(-radius + margin, radius - margin + 1e-12, a)
    ys = np.arange(-0.5 * length + margin, 0.5 * length - margin + 1e-12, a)
    X, Y, Z = np.meshgrid(xs, ys, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    inside = pts[:, 0] ** 2 + pts[:, 2] ** 2 < (radius - margin) ** 2
    pts = pts[inside]
    # lower fill fraction of the disc area: z below the chord height h with
    # segment area fraction = fill
    lo_, hi_ = -radius, radius
    for _ in range(60):
        h = 0.5 * (lo_ + hi_)
        th = 2.0 * math.acos(max(-1.0, min(1.0, -h / radius)))
        frac = (th - math.sin(th)) / (2.0 * math.pi)
        lo_, hi_ = (h, hi_) if frac < fill else (lo_, h)
    pts = pts[pts[:, 2] < h]
    order = np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))
    pts = pts[order]
    if max_particles is not None:
        pts = pts[:max_particles]
    n = len(pts)
    x = pts + rng.uniform(-0.05 * (a - 2 * rmax) - 1e-9, 0.05 * (a - 2 * rmax) + 1e-9, pts.shape)
    rad = r * (1.0 + rng.uniform(-spread, spread, n))
    box_hi = (radius * 1.05, 0.5 * length * 1.2, radius * 1.05)
    params = {"dt": dt, "skin": skin, "cdf": 1.0,
              "boxlo": tuple(-b for b in box_hi), "boxhi": box_hi, "periodic": (0, 0, 0),
              "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1, "rolling_cdt": 0}
    particles = {"x": x, "v": np.zeros((n, 3)), "omega": np.zeros((n, 3)), "radius": rad,
                 "rmass": sphere_mass(rad, 2500.0), "type": np.ones(n, dtype=np.int32),
                 "tag": np.arange(1, n + 1, dtype=np.int32)}
    tris = msh.drum(radius, length, n_az, n_ax)
    meshes = [(tris, 2, None, ((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), period))]
    return params, materials_two(), particles, [], meshes


def write_data_file(path, params, particles, ntypes=1, density=2500.0):
    """Write a LIGGGHTS granular data file for read_data."""
    x = np.asarray(particles["x"], dtype=np.float64).tolist()
    v = np.asarray(particles["v"], dtype=np.float64).tolist()
    w = np.asarray(particles["omega"], dtype=np.float64).tolist()
    r = np.asarray(particles["radius"], dtype=np.float64).tolist()
    t = np.asarray(particles["type"]).astype(int).tolist()
    tag = np.asarray(particles["tag"]).astype(int).tolist()
    lo = [float(a) for a in params["boxlo"]]
    hi = [float(a) for a in params["boxhi"]]
    density = float(density)
    with open(path, "w") as f:
        f.write("LIGGGHTS data file (granular), seeded synthetic workload\n\n")
        f.write(f"{len(r)} atoms\n{ntypes} atom types\n\n")
        for d, nm in enumerate("xyz"):
            f.write(f"{lo[d]!r} {hi[d]!r} {nm}lo {nm}hi\n")
        f.write("\nAtoms\n\n")
        for i in range(len(r)):
            f.write(f"{tag[i]} {t[i]} {2.0 * r[i]!r} {density!r} "
                    f"{x[i][0]!r} {x[i][1]!r} {x[i][2]!r}\n")
        f.write("\nVelocities\n\n")
        for i in range(len(r)):
            f.write(f"{tag[i]} {v[i][0]!r} {v[i][1]!r} {v[i][2]!r} "
                    f"{w[i][0]!r} {w[i][1]!r} {w[i][2]!r}\n")


def heap_sweep(seed=0, ninst=512, nmin=2000, nmax=50000, r=0.0015, floor=0.1, dt=1e-5,
               skin=0.001, footprint=0.09, counts=None):
    """Batched calibration sweep: `ninst` independent heap-formation runs.
    Instance q has N_q spheres (log-uniform in [nmin, nmax], or `counts`) of
    radius r on a jittered cubic lattice (spacing 2.2 r) over a centred
    footprint x footprint patch, starting just above a finite floor x floor
    square at z = 0 (two triangles, fix mesh/surface, wall type 2) with open
    sides, falling under gravity. Materials per instance: Y 5e6, nu 0.45,
    friction U(0.2, 0.8) (particle-particle and particle-wall), restitution
    U(0.1, 0.9), CDT rolling friction U(0, 0.3). All instances share one box
    (x, y in +-2 floor, z from -floor/2 up to the tallest column) and local
    coordinates.
    Returns (params, base_materials, particles, walls, meshes, counts,
    instance_materials)."""
    import meshes as msh
    rng = np.random.default_rng(seed)
    if counts is None:
        counts = np.exp(rng.uniform(math.log(nmin), math.log(nmax), ninst)).round().astype(np.int64)
    counts = np.asarray(counts, dtype=np.int64)
    a = 2.2 * r
    nside = max(1, int(footprint / a))
    per_layer = nside * nside
    z0 = r + 0.25 * a
    tallest = int(math.ceil(counts.max() / per_layer))
    box_hi = (floor, floor, z0 + tallest * a + a)
    box_lo = (-floor, -floor, -0.5 * floor)
    xs, vs, rads, tags, inst_mats = [], [], [], [], []
    c0 = -0.5 * (nside - 1) * a
    for q, n in enumerate(counts):
        k = np.arange(n)
        layer, rem = k // per_layer, k % per_layer
        ix, iy = rem % nside, rem // nside
        x = np.stack([c0 + ix * a, c0 + iy * a, z0 + layer * a], axis=1).astype(np.float64)
        x += rng.uniform(-0.05 * (a - 2 * r), 0.05 * (a - 2 * r), x.shape)
        xs.append(x)
        vs.append(np.zeros((n, 3)))
        rads.append(np.full(n, r))
        tags.append(np.arange(1, n + 1, dtype=np.int32))
        mu = float(rng.uniform(0.2, 0.8))
        e = float(rng.uniform(0.1, 0.9))
        mur = float(rng.uniform(0.0, 0.3))
        inst_mats.append({"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
                          "restitution": [[e, e], [e, e]], "friction": [[mu, mu], [mu, mu]],
                          "rolling_friction": [[mur, mur], [mur, mur]]})
    rad = np.concatenate(rads)
    N = len(rad)
    particles = {"x": np.concatenate(xs), "v": np.concatenate(vs), "omega": np.zeros((N, 3)),
                 "radius": rad, "rmass": sphere_mass(rad, 2500.0),
                 "type": np.ones(N, dtype=np.int32), "tag": np.concatenate(tags)}
    params = {"dt": dt, "skin": skin, "cdf": 1.0, "boxlo": box_lo, "boxhi": box_hi,
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -9.81), "gravity_first": 1,
              "rolling_cdt": 1}
    tris = msh.grid(-0.5 * floor, 0.5 * floor, -0.5 * floor, 0.5 * floor, 0.0, 1, 1, diag="same")
    meshes = [(tris, 2, {"rolling_cdt": 1}, None)]
    base = inst_mats[0]
    return params, base, particles, [], meshes, counts, inst_mats


def instance_slice(particles, counts, q):
    """Particles of instance q of a batched system, as a stand-alone set."""
    off = int(np.sum(counts[:q]))
    n = int(counts[q])
    return {k: v[off:off + n] for k, v in particles.items()}


def superquadric_settling(seed=0, n=200_000, base=(0.2, 0.2), axes=(0.001, 0.003), blockiness=(2.0, 8.0),
                          density=2500.0, dt=1e-5, skin=0.001, gap=0.0005, youngs=5e6, poisson=0.45,
                          restitution=0.3, friction=0.5):
    """Superquadrics settling onto a floor: half-axes uniform in ``axes``,
    blockiness n1, n2 uniform in ``blockiness``, uniformly random unit
    quaternions, atom_style superquadric data (shape, blockiness, density).
    The base is periodic in x and y with a primitive zplane floor at z = 0.
    Particles start at rest on a jittered layer lattice whose spacing in each
    layer is the largest circumscribing box diagonal there plus ``gap`` (the
    largest shapes in the lowest layers), so no two start in contact.
    Returns (params, materials, particles, walls) with particles a dict of x,
    v, omega (n,3), quat (n,4), shape (n,3), blockiness (n,2), density (n,),
    type (n,) int32."""
    rng = np.random.default_rng(seed)
    shape = rng.uniform(axes[0], axes[1], (n, 3))
    blk = rng.uniform(blockiness[0], blockiness[1], (n, 2))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    bound = np.sqrt((shape * shape).sum(axis=1)) * 1.0001  # a superquadric with n >= 2 fits in its box
    order = np.argsort(-bound, kind="stable")
    x = np.zeros((n, 3))
    z = 0.0
    k = 0
    while k < n:
        s = 2.0 * bound[order[k]] + gap
        nx, ny = max(1, int(base[0] / s)), max(1, int(base[1] / s))
        sx, sy = base[0] / nx, base[1] / ny
        m = min(nx * ny, n - k)
        ids = order[k:k + m]
        cell = np.arange(m)
        jit = 0.25 * gap
        x[ids, 0] = (cell % nx + 0.5) * sx + rng.uniform(-jit, jit, m)
        x[ids, 1] = (cell // nx + 0.5) * sy + rng.uniform(-jit, jit, m)
        x[ids, 2] = z + 0.5 * s
        z += s
        k += m
    params = {"dt": dt, "skin": skin, "boxlo": [0.0, 0.0, 0.0], "boxhi": [base[0], base[1], max(z, 0.4) + 0.01],
              "periodic": [1, 1, 0], "grav_acc": [0.0, 0.0, -9.81], "tangential_damping": 1}
    mats = materials_single(youngs, poisson, restitution, friction)
    P = {"x": x, "v": np.zeros((n, 3)), "omega": np.zeros((n, 3)), "quat": q, "shape": shape,
         "blockiness": blk, "density": np.full(n, density), "type": np.ones(n, dtype=np.int32)}
    return params, mats, P, [(2, 0.0, 1)]


def write_superquadric_data(path, params, particles, ntypes=1):
    """Writes particles in the reference's data-file format for
    atom_style superquadric (AtomVecSuperquadric::data_atom / data_vel,
    atom_vec_superquadric.cpp:1150-1239): Atoms ``id type density shapex
    shapey shapez blockiness1 blockiness2 x y z``, Velocities ``id vx vy vz
    wx wy wz``. Orientations are not part of the format (set ... quat)."""
    P = particles
    n = len(P["x"])
    lo, hi = params["boxlo"], params["boxhi"]
    with open(path, "w") as f:
        f.write("LIGGGHTS data file (superquadric)\n\n")
        f.write(f"{n} atoms\n{ntypes} atom types\n\n")
        for d, ax in enumerate("xyz"):
            f.write(f"{float(lo[d])!r} {float(hi[d])!r} {ax}lo {ax}hi\n")
        f.write("\nAtoms\n\n")
        r = lambda a: " ".join(repr(float(t)) for t in a)
        for i in range(n):
            f.write(f"{i + 1} {int(P['type'][i])} {float(P['density'][i])!r} {r(P['shape'][i])} "
                    f"{r(P['blockiness'][i])} {r(P['x'][i])}\n")
        f.write("\nVelocities\n\n")
        for i in range(n):
            f.write(f"{i + 1} {r(P['v'][i])} {r(P['omega'][i])}\n")
