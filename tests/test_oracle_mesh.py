"""Pins the oracle's triangle-mesh walls (fix mesh/surface + fix wall/gran
mesh + fix move/mesh rotate, linear, wiggle) against geometry and physics
fixed independently:
closest points on triangles by an independent projection/clamping routine,
mesh topology invariants on meshes whose answer is known by construction,
a sphere on a tessellated floor behaving exactly like the same sphere on a
primitive plane (history carried across coplanar triangles), rigid rotation
against the rotation matrix, translation against x0 + n v dt, the wiggle
against the closed-form discrete sine sum, and the skin/2 rebuild trigger of
moving meshes.
"""
import math

import numpy as np
import pytest

import meshes
import pyoracle as orc
import workloads as wl


def point_triangle_distance(p, a, b, c):
    """Independent closest distance: inside-projection test, else the best of
    the three clamped segment projections."""
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # barycentric by areas
    def area(u, v, w):
        return np.dot(np.cross(v - u, w - u), n)
    A = area(a, b, c)
    l0, l1, l2 = area(q, b, c) / A, area(a, q, c) / A, area(a, b, q) / A
    if l0 >= 0 and l1 >= 0 and l2 >= 0:
        return np.linalg.norm(p - q)
    best = np.inf
    for u, v in ((a, b), (b, c), (c, a)):
        t = np.clip(np.dot(p - u, v - u) / np.dot(v - u, v - u), 0.0, 1.0)
        best = min(best, np.linalg.norm(p - (u + t * (v - u))))
    return best


def angles(t):
    out = []
    for k in range(3):
        u, v = t[(k + 1) % 3] - t[k], t[(k + 2) % 3] - t[k]
        out.append(math.degrees(math.acos(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))))
    return out


def test_tri_contact_is_exact_unless_obtuse_at_node_0_or_1():
    """With all edges and corners active the contact distance equals the true
    point-triangle distance for triangles with no obtuse angle or an obtuse
    angle at node 2. The reference keeps only the last node's obtuse test
    (surface_mesh_I.h:438-447), so an obtuse angle at node 0 or 1 is treated
    as acute: then the distance is exact or an overestimate, never less."""
    rng = np.random.default_rng(0)
    n_exact = n_over = 0
    for _ in range(3000):
        t = rng.uniform(-1, 1, (3, 3))
        if min(angles(t)) < 5:
            continue
        p = t.mean(0) + rng.normal(0, 1.2, 3)
        d, delta, bary, bs = orc.tri_contact(t, p, 0.0)
        ref = point_triangle_distance(p, *t)
        ang = angles(t)
        obtuse01 = ang[0] > 90 or ang[1] > 90
        if obtuse01:
            assert d >= ref * (1 - 1e-12)
            n_over += d > ref * (1 + 1e-9)
        else:
            assert d == pytest.approx(ref, rel=1e-12, abs=1e-14)
            n_exact += 1
        # delta points from the centre to the closest point it reports
        assert np.linalg.norm(delta) == pytest.approx(d, rel=1e-12, abs=1e-14)
        assert bary.sum() == pytest.approx(1.0, abs=1e-12)
    assert n_exact > 1000 and n_over > 0


def test_inactive_edges_and_corners_are_skipped():
    t = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    p_edge = np.array([0.5, -0.3, 0.1])    # beyond edge 0 (nodes 0-1)
    p_corner = np.array([-0.3, -0.3, 0.1])  # beyond corner 0
    d, *_ = orc.tri_contact(t, p_edge, 0.0)
    assert d == pytest.approx(math.hypot(0.3, 0.1), rel=1e-12)
    d, *_ = orc.tri_contact(t, p_edge, 0.0, edge_active=(0, 1, 1))
    assert d == pytest.approx(1e6)
    d, *_ = orc.tri_contact(t, p_corner, 0.0, corner_active=(0, 1, 1))
    assert d == pytest.approx(1e6)
    d, delta, bary, bs = orc.tri_contact(t, np.array([0.2, 0.2, 0.5]), 0.1)
    assert bs == 7 and d == pytest.approx(0.4, rel=1e-12)


def _oracle_with_mesh(tris, n=1, box=((-1, -1, -1), (1, 1, 1))):
    params = {"dt": 1e-5, "skin": 0.001, "boxlo": box[0], "boxhi": box[1], "periodic": (0, 0, 0)}
    P = {"x": np.zeros((n, 3)) + 0.5, "v": np.zeros((n, 3)), "omega": np.zeros((n, 3)),
         "radius": np.full(n, 0.001), "rmass": np.full(n, 1e-5), "type": np.ones(n, dtype=np.int32),
         "tag": np.arange(1, n + 1, dtype=np.int32)}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    return orc.OracleSystem(params, mats, P, meshes=[(tris, 2, None, None)])


def test_topology_flat_square_and_grid():
    sq = meshes.grid(0, 1, 0, 1, 0.0, 1, 1, diag="same")
    o = _oracle_with_mesh(sq)
    ea, ca, ob = o.mesh_topology()
    # the shared diagonal is coplanar: inactive on both; the border edges are active
    assert ea.sum() == 4
    # each of the 4 square corners is active on exactly one triangle
    assert ca.sum() == 4
    g = meshes.grid(0, 1, 0, 1, 0.0, 2, 2, diag="same")
    o = _oracle_with_mesh(g)
    ea, ca, ob = o.mesh_topology()
    assert ea.sum() == 8  # only the outer border (8 unit-half edges)
    # 9 nodes: the centre node has colinear edges through it -> inactive;
    # the 4 mid-side nodes have colinear border edges -> inactive; 4 corners active
    assert ca.sum() == 4


def test_topology_cube_and_drum_invariants():
    for tris in (meshes.cube((0, 0, 0), (0.1, 0.1, 0.1)), meshes.drum(0.15, 0.1, 16, 4)):
        o = _oracle_with_mesh(tris)
        ea, ca, ob = o.mesh_topology()
        T = len(tris)
        key = lambda p: tuple(np.round(p, 12))
        edges = {}
        for t in range(T):
            for k in range(3):
                e = tuple(sorted((key(tris[t][k]), key(tris[t][(k + 1) % 3]))))
                edges.setdefault(e, []).append((t, k))
        normal = lambda t: np.cross(tris[t][1] - tris[t][0], tris[t][2] - tris[t][0]) / \
            np.linalg.norm(np.cross(tris[t][1] - tris[t][0], tris[t][2] - tris[t][0]))
        for e, owners in edges.items():
            assert len(owners) == 2  # closed manifold
            (t1, k1), (t2, k2) = owners
            coplanar = abs(np.dot(normal(t1), normal(t2))) >= 1 - 1e-5
            if coplanar:
                assert ea[t1, k1] == 0 and ea[t2, k2] == 0
            else:
                # exactly one side active: the element with the higher index
                assert ea[t1, k1] + ea[t2, k2] == 1
                hi = max((t1, k1), (t2, k2))
                assert ea[hi[0], hi[1]] == 1
        nodes = {}
        for t in range(T):
            for k in range(3):
                nodes.setdefault(key(tris[t][k]), []).append((t, k))
        for p, owners in nodes.items():
            act = [ca[t, k] for t, k in owners]
            assert sum(act) <= 1
            if sum(act) == 1:  # only on the highest-index element
                assert ca[max(owners)[0], max(owners)[1]] == 1


def _floor_vs_plane(v0, steps, nx=6):
    r, rho, g = 0.002, 2500.0, 9.81
    m = wl.sphere_mass(np.array([r]), rho)
    Ys = 5e6 / (2.0 * (1.0 - 0.45 ** 2))
    dz = (m[0] * g / (4.0 / 3.0 * Ys * math.sqrt(r))) ** (2.0 / 3.0)
    params = {"dt": 2e-6, "skin": 0.001, "boxlo": (-0.05, -0.05, -0.01), "boxhi": (0.05, 0.05, 0.05),
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -g)}
    P = {"x": np.array([[-0.03, -0.0283, r - dz]]), "v": np.array([[v0, 0.6 * v0, 0.0]]),
         "omega": np.zeros((1, 3)), "radius": np.array([r]), "rmass": m, "type": np.array([1]),
         "tag": np.array([1])}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    floor = meshes.grid(-0.04, 0.04, -0.04, 0.04, 0.0, nx, nx)
    a = orc.OracleSystem(params, mats, P, meshes=[(floor, 2, None, None)])
    b = orc.OracleSystem(params, mats, P, walls=[(2, 0.0, 2)])
    a.run(steps)
    b.run(steps)
    return a, b


def test_sphere_on_tessellated_floor_equals_plane_wall():
    """A sphere sliding across many coplanar triangles feels the same contact
    as on a primitive plane: history is carried from triangle to triangle
    (face contacts, coplanar history copy) and the sliding-to-rolling
    transition lands at 5/7 of the initial speed."""
    v0 = 0.5
    t_roll = 2.0 * v0 * math.hypot(1, 0.6) / (7.0 * 0.4 * 9.81)
    a, b = _floor_vs_plane(v0, int(1.3 * t_roll / 2e-6), nx=12)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "omega"):
        assert np.allclose(sa[k], sb[k], rtol=1e-9, atol=1e-15), k
    # it crossed several triangles and ended rolling at 5/7 v0
    assert abs(sa["x"][0, 0] - (-0.03)) > 0.015  # > 2 grid cells of 6.7 mm
    speed = np.linalg.norm(sa["v"][0, :2])
    assert speed == pytest.approx(5.0 / 7.0 * v0 * math.hypot(1, 0.6), rel=3e-3)
    tri, hist = a.mesh_history()
    assert (tri[0] >= 0).sum() == 1


def test_rigid_rotation_matches_rotation_matrix_and_triggers_rebuilds():
    tris = meshes.drum(0.15, 0.1, 16, 4)
    params = {"dt": 5e-6, "skin": 0.0005, "boxlo": (-0.2, -0.2, -0.2), "boxhi": (0.2, 0.2, 0.2),
              "periodic": (0, 0, 0)}
    P = {"x": np.array([[0.0, 0.0, 0.0]]), "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)),
         "radius": np.array([0.001]), "rmass": np.array([1e-5]), "type": np.array([1]),
         "tag": np.array([1])}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    period = 2.0
    o = orc.OracleSystem(params, mats, P, meshes=[(tris, 2, None, ((0, 0, 0), (0, 1, 0), period))])
    nsteps = 1000
    o.run(nsteps)
    nodes, vn = o.mesh_nodes()
    w = 2 * math.pi / period
    th = w * nsteps * params["dt"]
    # rotation about +y by th
    R = np.array([[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]])
    expect = tris @ R.T
    assert np.allclose(nodes, expect, rtol=0, atol=1e-13)
    assert np.allclose(vn, np.cross(np.array([0, w, 0]), nodes), rtol=0, atol=1e-12)
    # first forced rebuild: the step after the first node moved > skin/2
    o2 = orc.OracleSystem(params, mats, P, meshes=[(tris, 2, None, ((0, 0, 0), (0, 1, 0), period))])
    Rmax = np.linalg.norm(tris[:, :, [0, 2]], axis=2).max()
    k = 1
    while 2 * Rmax * math.sin(k * w * params["dt"] / 2) <= 0.5 * params["skin"]:
        k += 1
    o2.run(k)
    assert o2.mesh_forced_rebuilds() == 0
    o2.run_continue(1)
    assert o2.mesh_forced_rebuilds() == 1


def test_rotating_floor_drags_a_resting_sphere():
    """A sphere at rest on a floor spinning about z: after one step the
    tangential force on it points along omega x r_contact (the wall velocity
    interpolated from the node velocities is the rigid-body field)."""
    r = 0.002
    m = wl.sphere_mass(np.array([r]), 2500.0)
    params = {"dt": 1e-6, "skin": 0.001, "boxlo": (-0.1, -0.1, -0.01), "boxhi": (0.1, 0.1, 0.05),
              "periodic": (0, 0, 0)}
    xc = np.array([0.031, -0.017, r - 1e-6])
    P = {"x": xc[None], "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)), "radius": np.array([r]),
         "rmass": m, "type": np.array([1]), "tag": np.array([1])}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    floor = meshes.grid(-0.08, 0.08, -0.08, 0.08, 0.0, 4, 4)
    o = orc.OracleSystem(params, mats, P, meshes=[(floor, 2, None, ((0, 0, 0), (0, 0, 1), 0.5))])
    o.run(1)
    st = o.state()
    ft = st["f"][0, :2]
    w = np.array([0, 0, 2 * math.pi / 0.5])
    vdir = np.cross(w, np.array([xc[0], xc[1], 0.0]))[:2]
    cosang = np.dot(ft, vdir) / np.linalg.norm(ft) / np.linalg.norm(vdir)
    assert cosang > 1 - 1e-6


def _mover_case(motion, dt=5e-6, skin=0.0005):
    tris = meshes.grid(-0.05, 0.05, -0.05, 0.05, 0.0, 4, 4)
    params = {"dt": dt, "skin": skin, "boxlo": (-0.2, -0.2, -0.2), "boxhi": (0.2, 0.2, 0.2),
              "periodic": (0, 0, 0)}
    P = {"x": np.array([[0.0, 0.0, 0.1]]), "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)),
         "radius": np.array([0.001]), "rmass": np.array([1e-5]), "type": np.array([1]),
         "tag": np.array([1])}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    return tris, orc.OracleSystem(params, mats, P, meshes=[(tris, 2, None, motion)]), params


def test_linear_mover_translates_and_triggers_rebuilds():
    """fix move/mesh linear: after n steps every node sits at x0 + n v dt, node
    velocities are v, and the first mesh-forced rebuild comes on the step after
    the displacement first exceeds skin/2."""
    v = np.array([0.3, -0.2, 0.1])
    tris, o, params = _mover_case({"linear": v})
    n = 700
    o.run(n)
    nodes, vn = o.mesh_nodes()
    assert np.allclose(nodes, tris + n * v * params["dt"], rtol=0, atol=1e-14)
    assert np.array_equal(vn, np.broadcast_to(v, vn.shape))
    k = 1
    while (k * np.linalg.norm(v) * params["dt"]) ** 2 <= 0.25 * params["skin"] ** 2:
        k += 1
    _, o2, _ = _mover_case({"linear": v})
    o2.run(k)
    assert o2.mesh_forced_rebuilds() == 0
    o2.run_continue(1)
    assert o2.mesh_forced_rebuilds() == 1


def test_wiggle_mover_is_the_discrete_sine_sum():
    """fix move/mesh wiggle: the node velocity at step k is A w cos(w k dt)
    (the fix's time is advanced before the move) and the displacement after
    n steps is A w dt sum_k cos(w k dt) = A w dt sin(n th/2) cos((n+1) th/2) /
    sin(th/2), th = w dt; over a whole number of periods it returns to the
    start to O(dt). The time runs on across runs (fix lifetime)."""
    A = np.array([0.002, 0.0, 0.001])
    period = 0.01
    tris, o, params = _mover_case({"wiggle": (A, period)})
    dt = params["dt"]
    w = 2 * math.pi / period
    th = w * dt
    n1, n2 = 900, 1100
    o.run(n1)
    o.run(n2)  # a second run continues the fix's time
    n = n1 + n2
    nodes, vn = o.mesh_nodes()
    disp = A * w * dt * math.sin(n * th / 2) * math.cos((n + 1) * th / 2) / math.sin(th / 2)
    assert np.allclose(nodes, tris + disp, rtol=0, atol=1e-13)
    assert np.allclose(vn, np.broadcast_to(A * w * math.cos(w * n * dt), vn.shape), rtol=0, atol=1e-12)
    assert n * dt / period == pytest.approx(1.0)
    assert np.abs(nodes - tris).max() < 2 * np.abs(A).max() * th


def test_linear_floor_drags_a_resting_sphere():
    """A sphere at rest on a floor translating along +x: after one step the
    tangential force on it points along +x (the wall velocity is the
    translation velocity)."""
    r = 0.002
    m = wl.sphere_mass(np.array([r]), 2500.0)
    params = {"dt": 1e-6, "skin": 0.001, "boxlo": (-0.1, -0.1, -0.01), "boxhi": (0.1, 0.1, 0.05),
              "periodic": (0, 0, 0)}
    P = {"x": np.array([[0.01, 0.02, r - 1e-6]]), "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)),
         "radius": np.array([r]), "rmass": m, "type": np.array([1]), "tag": np.array([1])}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    floor = meshes.grid(-0.08, 0.08, -0.08, 0.08, 0.0, 4, 4)
    o = orc.OracleSystem(params, mats, P, meshes=[(floor, 2, None, {"linear": (0.2, 0.0, 0.0)})])
    o.run(1)
    f = o.state()["f"][0]
    assert f[0] > 0 and abs(f[1]) < 1e-12 * abs(f[0])


def test_mesh_stress_of_a_resting_sphere():
    """fix mesh/surface/stress: a sphere resting on a floor pushes it with its
    weight (force on the mesh (0, 0, -m g)); the torque about a reference
    point p is (c - p) x F with the contact point c under the sphere's centre,
    zero about that point itself."""
    r, g = 0.002, 9.81
    m = wl.sphere_mass(np.array([r]), 2500.0)
    Ys = 5e6 / (2.0 * (1.0 - 0.45 ** 2))
    d_eq = (m[0] * g / (4.0 / 3.0 * Ys * math.sqrt(r))) ** (2.0 / 3.0)
    x0, y0 = 0.013, -0.021
    params = {"dt": 2e-6, "skin": 0.001, "boxlo": (-0.1, -0.1, -0.01), "boxhi": (0.1, 0.1, 0.05),
              "periodic": (0, 0, 0), "grav_acc": (0.0, 0.0, -g)}
    P = {"x": np.array([[x0, y0, r - d_eq]]), "v": np.zeros((1, 3)), "omega": np.zeros((1, 3)),
         "radius": np.array([r]), "rmass": m, "type": np.array([1]), "tag": np.array([1])}
    mats = {"ntypes": 2, "youngs": [5e6, 5e6], "poisson": [0.45, 0.45],
            "restitution": [[0.3, 0.3], [0.3, 0.3]], "friction": [[0.5, 0.4], [0.4, 0.4]]}
    floor = meshes.grid(-0.08, 0.08, -0.08, 0.08, 0.0, 4, 4)
    W = m[0] * g
    for pref in ((0.0, 0.0, 0.0), (x0, y0, 0.0), (0.05, 0.01, 0.02)):
        o = orc.OracleSystem(params, mats, P, meshes=[(floor, 2, None, None, pref)])
        o.run(3000)
        F, T = o.mesh_force()
        assert F[2] == pytest.approx(-W, rel=1e-4)
        assert abs(F[0]) < 1e-9 * W and abs(F[1]) < 1e-9 * W
        c = o.state()["x"][0] - np.array([0.0, 0.0, o.state()["x"][0, 2]])  # contact point on the floor
        assert np.allclose(T, np.cross(c - np.array(pref), F), rtol=1e-6, atol=1e-9 * W * 0.01)
        o.close()


def _sin_sum(w, phi, dt, n):
    """sum_{k=1..n} sin(w k dt + phi) in closed form"""
    th = w * dt
    return math.sin(phi + (n + 1) * th / 2) * math.sin(n * th / 2) / math.sin(th / 2)


def test_viblin_mover_is_the_discrete_sum_of_its_modes():
    """fix move/mesh viblin (two modes): displacement after n steps =
    -axis dt sum_j a_j w_j sum_k sin(w_j k dt + phi_j), velocity at the last
    step axis * (-sum_j a_j w_j sin(w_j n dt + phi_j)); the axis is normalized."""
    axis = np.array([1.0, 2.0, 2.0])
    am, ph, pe = np.array([0.002, 0.0005]), np.array([0.3, -1.1]), np.array([0.01, 0.003])
    tris, o, params = _mover_case({"viblin": (axis, am, ph, pe)})
    dt, n = params["dt"], 1500
    o.run(n)
    u = axis / 3.0
    w = 2 * math.pi / pe
    disp = -u * dt * sum(am[j] * w[j] * _sin_sum(w[j], ph[j], dt, n) for j in range(2))
    nodes, vn = o.mesh_nodes()
    assert np.allclose(nodes, tris + disp, rtol=0, atol=1e-13)
    vA = -sum(am[j] * w[j] * math.sin(w[j] * n * dt + ph[j]) for j in range(2))
    assert np.allclose(vn, np.broadcast_to(u * vA, vn.shape), rtol=0, atol=1e-12)


def test_vibrot_mover_turns_by_the_integrated_angle():
    """fix move/mesh vibrot: after n steps the mesh is turned about the axis
    through the origin by dt sum_k vR(k dt), vR = -sum_j a_j w_j sin(w_j t + phi_j);
    node velocities are (axis vR) x (node - origin)."""
    origin = np.array([0.01, -0.02, 0.0])
    am, ph, pe = np.array([0.3]), np.array([0.5]), np.array([0.02])
    tris, o, params = _mover_case({"vibrot": (origin, (0, 0, 2.0), am, ph, pe)})
    dt, n = params["dt"], 1700
    o.run(n)
    w = 2 * math.pi / pe[0]
    ang = -dt * am[0] * w * _sin_sum(w, ph[0], dt, n)
    R = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    nodes, vn = o.mesh_nodes()
    assert np.allclose(nodes, (tris - origin) @ R.T + origin, rtol=0, atol=1e-13)
    vR = -am[0] * w * math.sin(w * n * dt + ph[0])
    assert np.allclose(vn, np.cross(np.array([0, 0, vR]), nodes - origin), rtol=0, atol=1e-12)
    assert abs(ang) > 0.05
