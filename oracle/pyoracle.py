"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the shipped library.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAXT = 8
MAXPLANE = 6

_D = ctypes.c_double
_Mat = (_D * (MAXT + 1)) * (MAXT + 1)


class OrcMaterials(ctypes.Structure):
    _fields_ = [("ntypes", ctypes.c_int), ("Yeff", _Mat), ("Geff", _Mat), ("betaeff", _Mat),
                ("mu", _Mat), ("mur", _Mat), ("kn", _Mat), ("kt", _Mat), ("gamman", _Mat),
                ("gammat", _Mat), ("ced", _Mat), ("loge", _Mat), ("charvel", _D)]


class OrcModel(ctypes.Structure):
    _fields_ = [("rolling_cdt", ctypes.c_int), ("tangential_damping", ctypes.c_int),
                ("limit_force", ctypes.c_int), ("torsion_torque", ctypes.c_int),
                ("normal_model", ctypes.c_int), ("tangential_model", ctypes.c_int),
                ("cohesion_model", ctypes.c_int), ("kt_to_kn", ctypes.c_int),
                ("absolute_damping", ctypes.c_int), ("tangential_reduce", ctypes.c_int),
                ("rolling_epsd2", ctypes.c_int)]


# model <normal> of pair_style gran / fix wall/gran
NORMAL_MODELS = {"hertz": 0, "hooke": 1, "hertz/stiffness": 2, "hooke/stiffness": 3}
TANGENTIAL_MODELS = {"history": 0, "no_history": 1}
COHESION_MODELS = {None: 0, "off": 0, "sjkr": 1}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dt", _D), ("skin", _D), ("cdf", _D),
        ("boxlo", _D * 3), ("boxhi", _D * 3), ("periodic", ctypes.c_int * 3),
        ("model", OrcModel),
        ("has_gravity", ctypes.c_int), ("grav_acc", _D * 3),
        ("nplanes", ctypes.c_int), ("plane_axis", ctypes.c_int * MAXPLANE),
        ("plane_pos", _D * MAXPLANE), ("plane_type", ctypes.c_int * MAXPLANE),
        ("gravity_first", ctypes.c_int),
        ("plane_model", OrcModel * MAXPLANE),
        ("neigh_every", ctypes.c_int), ("neigh_no_check", ctypes.c_int),
        ("shrink_faces", ctypes.c_int), ("shrink_min", ctypes.c_int),
    ]


ORI_MAXTMPL = 8


class OriParams(ctypes.Structure):
    _fields_ = [
        ("ntemplates", ctypes.c_int), ("tmpl_seed", ctypes.c_int * ORI_MAXTMPL),
        ("tmpl_type", ctypes.c_int * ORI_MAXTMPL), ("tmpl_radius", _D * ORI_MAXTMPL),
        ("tmpl_density", _D * ORI_MAXTMPL), ("dist_seed", ctypes.c_int),
        ("dist_numberbased", ctypes.c_int), ("dist_weight", _D * ORI_MAXTMPL),
        ("seed", ctypes.c_int), ("maxattempt", ctypes.c_int), ("overlapcheck", ctypes.c_int),
        ("all_in", ctypes.c_int), ("exact_number", ctypes.c_int), ("vel_style", ctypes.c_int),
        ("v_insert", _D * 3), ("v_fluct", _D * 3), ("omega", _D * 3),
        ("check_dist_from_subdomain_border", ctypes.c_int), ("ntry_mc", ctypes.c_int),
        ("volumefraction_region", _D), ("particles_in_region", ctypes.c_int),
        ("mass_in_region", _D), ("region_lo", _D * 3), ("region_hi", _D * 3),
        ("boxlo", _D * 3), ("boxhi", _D * 3), ("box_change", ctypes.c_int),
    ]


class OriPark(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int), ("save", ctypes.c_int), ("second", _D)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-C", HERE, "liboracle.so"])
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    dp, ip = P(_D), P(ctypes.c_int)
    lib.orc_mix_materials.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, P(OrcMaterials)]
    lib.orc_set_model_coeffs.argtypes = [P(OrcMaterials), ctypes.c_int, dp, dp, dp, dp, dp, _D]
    lib.orc_contact_pair.argtypes = [P(OrcModel), P(OrcMaterials), ctypes.c_int, ctypes.c_int,
                                     dp, _D, _D, dp, dp, dp, dp, _D, _D, _D, ctypes.c_int,
                                     dp, dp, dp, dp]
    lib.orc_contact_pair.restype = ctypes.c_int
    lib.orc_contact_wall.argtypes = [P(OrcModel), P(OrcMaterials), ctypes.c_int, ctypes.c_int,
                                     _D, dp, _D, dp, dp, dp, _D, _D, ctypes.c_int, dp, dp, dp]
    lib.orc_contact_pair_r.argtypes = [P(OrcModel), P(OrcMaterials), ctypes.c_int, ctypes.c_int,
                                       dp, _D, _D, dp, dp, dp, dp, _D, _D, _D, ctypes.c_int,
                                       dp, dp, dp, dp, dp]
    lib.orc_contact_pair_r.restype = ctypes.c_int
    lib.orc_contact_wall_r.argtypes = [P(OrcModel), P(OrcMaterials), ctypes.c_int, ctypes.c_int,
                                       _D, dp, _D, dp, dp, dp, _D, _D, ctypes.c_int, dp, dp, dp,
                                       dp]
    lib.orc_get_pair_rolling.argtypes = [ctypes.c_void_p, ctypes.c_int64, dp]
    lib.orc_get_pair_rolling.restype = ctypes.c_int64
    lib.orc_get_wall_rolling.argtypes = [ctypes.c_void_p, dp]
    lib.orc_pairs_bruteforce.argtypes = [ctypes.c_int, dp, dp, P(OrcParams), ctypes.c_int64, ip]
    lib.orc_pairs_bruteforce.restype = ctypes.c_int64
    lib.orc_create.argtypes = [P(OrcParams), P(OrcMaterials), ctypes.c_int, dp, dp, dp, dp, dp,
                               ip, ip]
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_setup.argtypes = [ctypes.c_void_p]
    lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_compute_forces.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_get_state.argtypes = [ctypes.c_void_p, dp, dp, dp, dp, dp]
    lib.orc_get_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int64, ip, ip, dp]
    lib.orc_get_pairs.restype = ctypes.c_int64
    lib.orc_pair_local.argtypes = [ctypes.c_void_p, ctypes.c_int64, ip, dp]
    lib.orc_pair_local.restype = ctypes.c_int64
    lib.orc_wall_local.argtypes = [ctypes.c_void_p, ctypes.c_int64, ip, dp]
    lib.orc_wall_local.restype = ctypes.c_int64
    lib.orc_get_wall_hist.argtypes = [ctypes.c_void_p, dp]
    lib.orc_lost_atoms.argtypes = [ctypes.c_void_p, P(ctypes.c_int64)]
    lib.orc_lost_atoms.restype = ctypes.c_int64
    lib.orc_get_box.argtypes = [ctypes.c_void_p, dp, dp]
    lib.orc_nrebuilds.argtypes = [ctypes.c_void_p]
    lib.orc_nrebuilds.restype = ctypes.c_int64
    lib.orc_ntimestep.argtypes = [ctypes.c_void_p]
    lib.orc_ntimestep.restype = ctypes.c_int64
    lib.orc_contact_counts.argtypes = [ctypes.c_void_p, P(ctypes.c_int64), P(ctypes.c_int64)]
    lib.orc_add_mesh.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_int, P(OrcModel)]
    lib.orc_add_mesh.restype = ctypes.c_int
    lib.orc_mesh_rotate.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, _D]
    lib.orc_mesh_linear.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_mesh_stress.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
    lib.orc_mesh_force.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
    lib.orc_mesh_wiggle.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, _D]
    lib.orc_mesh_viblin.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_int, dp, dp, dp]
    lib.orc_mesh_vibrot.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, dp]
    lib.orc_mesh_topology.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ip, ip]
    lib.orc_mesh_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
    lib.orc_mesh_history.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ip, dp]
    lib.orc_mesh_history.restype = ctypes.c_int
    lib.orc_tri_contact.argtypes = [dp, ip, ip, dp, _D, dp, dp, ip]
    lib.orc_tri_contact.restype = _D
    lib.orc_mesh_forced_rebuilds.argtypes = [ctypes.c_void_p]
    lib.orc_mesh_forced_rebuilds.restype = ctypes.c_int64
    lib.orc_insert.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp, dp, dp, ip, ip]
    lib.orc_set_max_radius.argtypes = [ctypes.c_void_p, _D]
    lib.orc_natoms.argtypes = [ctypes.c_void_p]
    lib.orc_natoms.restype = ctypes.c_int
    lib.ori_park_init.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.ori_park_uniform.argtypes = [ctypes.c_void_p]
    lib.ori_park_uniform.restype = _D
    lib.ori_park_gaussian.argtypes = [ctypes.c_void_p]
    lib.ori_park_gaussian.restype = _D
    lib.ori_create.argtypes = [P(OriParams)]
    lib.ori_create.restype = ctypes.c_void_p
    lib.ori_destroy.argtypes = [ctypes.c_void_p]
    lib.ori_region_volume.argtypes = [ctypes.c_void_p]
    lib.ori_region_volume.restype = _D
    lib.ori_distribution.argtypes = [ctypes.c_void_p, dp, dp, dp, ip]
    lib.ori_insert_step.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_int,
                                    dp, dp, dp, dp, dp, dp, ip, ip]
    lib.ori_insert_step.restype = ctypes.c_int
    lib.ori_rng_states.argtypes = [ctypes.c_void_p, ip]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(_D))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


MODEL_COEFFS = ("kn", "kt", "gamman", "gammat", "cohesion_energy_density")


def materials(mat):
    """Mix a material dict (youngs, poisson, restitution, friction[, rolling_friction]
    [, kn, kt, gamman, gammat, cohesion_energy_density (ntypes x ntypes),
    characteristic_velocity])."""
    lib = load()
    nt = mat["ntypes"]
    Y = np.ascontiguousarray(mat["youngs"], dtype=np.float64)
    nu = np.ascontiguousarray(mat["poisson"], dtype=np.float64)
    e = np.ascontiguousarray(mat["restitution"], dtype=np.float64).reshape(-1)
    mu = np.ascontiguousarray(mat["friction"], dtype=np.float64).reshape(-1)
    mur = mat.get("rolling_friction")
    mur = None if mur is None else np.ascontiguousarray(mur, dtype=np.float64).reshape(-1)
    out = OrcMaterials()
    lib.orc_mix_materials(nt, _dp(Y), _dp(nu), _dp(e), _dp(mu),
                          _dp(mur) if mur is not None else None, ctypes.byref(out))
    if any(k in mat for k in MODEL_COEFFS + ("characteristic_velocity",)):
        arrs = [None if mat.get(k) is None else np.ascontiguousarray(mat[k], dtype=np.float64).reshape(-1)
                for k in MODEL_COEFFS]
        lib.orc_set_model_coeffs(ctypes.byref(out), nt, *[_dp(a) if a is not None else None for a in arrs],
                                 float(mat.get("characteristic_velocity", 0.0)))
    return out


def model(params):
    m = OrcModel()
    m.rolling_cdt = int(params.get("rolling_cdt", 0))
    m.tangential_damping = int(params.get("tangential_damping", 1))
    m.limit_force = int(params.get("limit_force", 0))
    m.torsion_torque = int(params.get("torsion_torque", 0))
    m.normal_model = NORMAL_MODELS[params.get("normal", "hertz")]
    m.tangential_model = TANGENTIAL_MODELS[params.get("tangential", "history")]
    m.cohesion_model = COHESION_MODELS[params.get("cohesion")]
    m.kt_to_kn = int(params.get("kt_to_kn", 0))
    m.absolute_damping = int(params.get("absolute_damping", 0))
    m.tangential_reduce = int(params.get("tangential_reduce", 0))
    m.rolling_epsd2 = int(params.get("rolling_epsd2", 0))
    return m


def make_params(params, walls=()):
    p = OrcParams()
    p.dt = params["dt"]
    p.skin = params["skin"]
    p.cdf = params.get("cdf", 1.0)
    for d in range(3):
        p.boxlo[d] = params["boxlo"][d]
        p.boxhi[d] = params["boxhi"][d]
        p.periodic[d] = int(params["periodic"][d])
    p.model = model(params)
    g = params.get("grav_acc")
    p.has_gravity = int(g is not None)
    if g is not None:
        for d in range(3):
            p.grav_acc[d] = g[d]
    p.nplanes = len(walls)
    for k, wall in enumerate(walls):
        axis, pos, wtype = wall[:3]
        p.plane_axis[k] = int(axis)
        p.plane_pos[k] = float(pos)
        p.plane_type[k] = int(wtype)
        p.plane_model[k] = model(wall[3] if len(wall) > 3 else params)
    p.gravity_first = int(params.get("gravity_first", 1))
    p.neigh_every = int(params.get("neigh_every", 1))
    p.neigh_no_check = int(not params.get("neigh_check", True))
    p.shrink_faces = int(params.get("shrink_faces", 0))
    p.shrink_min = int(params.get("shrink_min", 0))
    return p


def contact_pair(params, mats, itype, jtype, del_, radi, radj, vi, vj, wi, wj, mi, mj, dt,
                 shearupdate, shear):
    """One pair contact; returns (touch, dFi, dTi, dTj, shear_after)."""
    lib = load()
    mdl = model(params)
    arr = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    d, vi, vj, wi, wj = arr(del_), arr(vi), arr(vj), arr(wi), arr(wj)
    sh = arr(shear).copy()
    dF, dTi, dTj = np.zeros(3), np.zeros(3), np.zeros(3)
    touch = lib.orc_contact_pair(ctypes.byref(mdl), ctypes.byref(mats), itype, jtype, _dp(d),
                                 radi, radj, _dp(vi), _dp(vj), _dp(wi), _dp(wj), mi, mj, dt,
                                 int(shearupdate), _dp(sh), _dp(dF), _dp(dTi), _dp(dTj))
    return bool(touch), dF, dTi, dTj, sh


def contact_wall(params, mats, itype, wtype, overlap, sdelta, radi, vi, wi, vwall, mi, dt,
                 shearupdate, shear):
    lib = load()
    mdl = model(params)
    arr = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    sd, vi, wi, vw = arr(sdelta), arr(vi), arr(wi), arr(vwall)
    sh = arr(shear).copy()
    dF, dT = np.zeros(3), np.zeros(3)
    lib.orc_contact_wall(ctypes.byref(mdl), ctypes.byref(mats), itype, wtype, overlap, _dp(sd),
                         radi, _dp(vi), _dp(wi), _dp(vw), mi, dt, int(shearupdate), _dp(sh),
                         _dp(dF), _dp(dT))
    return dF, dT, sh


def contact_pair_rolling(params, mats, itype, jtype, del_, radi, radj, vi, vj, wi, wj, mi, mj,
                         dt, shearupdate, shear, rhist):
    """One pair contact with the epsd2 rolling history; returns
    (touch, dFi, dTi, dTj, shear_after, rhist_after)."""
    lib = load()
    mdl = model(params)
    arr = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    d, vi, vj, wi, wj = arr(del_), arr(vi), arr(vj), arr(wi), arr(wj)
    sh, rh = arr(shear).copy(), arr(rhist).copy()
    dF, dTi, dTj = np.zeros(3), np.zeros(3), np.zeros(3)
    touch = lib.orc_contact_pair_r(ctypes.byref(mdl), ctypes.byref(mats), itype, jtype, _dp(d),
                                   radi, radj, _dp(vi), _dp(vj), _dp(wi), _dp(wj), mi, mj, dt,
                                   int(shearupdate), _dp(sh), _dp(rh), _dp(dF), _dp(dTi),
                                   _dp(dTj))
    return bool(touch), dF, dTi, dTj, sh, rh


def contact_wall_rolling(params, mats, itype, wtype, overlap, sdelta, radi, vi, wi, vwall, mi,
                         dt, shearupdate, shear, rhist):
    lib = load()
    mdl = model(params)
    arr = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    sd, vi, wi, vw = arr(sdelta), arr(vi), arr(wi), arr(vwall)
    sh, rh = arr(shear).copy(), arr(rhist).copy()
    dF, dT = np.zeros(3), np.zeros(3)
    lib.orc_contact_wall_r(ctypes.byref(mdl), ctypes.byref(mats), itype, wtype, overlap, _dp(sd),
                           radi, _dp(vi), _dp(wi), _dp(vw), mi, dt, int(shearupdate), _dp(sh),
                           _dp(rh), _dp(dF), _dp(dT))
    return dF, dT, sh, rh


def tri_contact(nodes, p, radius, edge_active=(1, 1, 1), corner_active=(1, 1, 1)):
    """TriMesh::resolveTriSphereContactBary on one triangle:
    returns (distance - radius, delta, bary, barysign)."""
    lib = load()
    n = np.ascontiguousarray(nodes, dtype=np.float64).reshape(9)
    ea = np.ascontiguousarray(edge_active, dtype=np.int32)
    ca = np.ascontiguousarray(corner_active, dtype=np.int32)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    delta, bary = np.zeros(3), np.zeros(3)
    bs = ctypes.c_int(0)
    d = lib.orc_tri_contact(_dp(n), _ip(ea), _ip(ca), _dp(pp), float(radius), _dp(delta),
                            _dp(bary), ctypes.byref(bs))
    return d, delta, bary, bs.value


def pairs_bruteforce(params, x, radius):
    lib = load()
    p = make_params(params)
    x = np.ascontiguousarray(x, dtype=np.float64)
    r = np.ascontiguousarray(radius, dtype=np.float64)
    cap = 64 * len(r) + 16
    while True:
        out = np.zeros((cap, 2), dtype=np.int32)
        n = lib.orc_pairs_bruteforce(len(r), _dp(x), _dp(r), ctypes.byref(p), cap, _ip(out))
        if n >= 0:
            return out[:n]
        cap = int(-n) + 16


class OracleSystem:
    """The CPU restatement of Verlet::setup / Verlet::run."""

    def __init__(self, params, mat, particles, walls=(), meshes=()):
        """meshes: [(tris (T,3,3), wall_type, model dict or None, motion or None
        [, stress reference point])], motion = (origin, axis, period) for fix
        move/mesh rotate, or {"linear": (vx, vy, vz)}, {"wiggle": ((ax, ay, az),
        period)}, {"viblin": (axis, amplitudes, phases, periods)} or
        {"vibrot": (origin, axis, amplitudes, phases, periods)}; a fifth entry
        makes the mesh a mesh/surface/stress."""
        self.lib = load()
        self.p = make_params(params, walls)
        self.m = materials(mat)
        P = {k: np.ascontiguousarray(particles[k], dtype=np.float64)
             for k in ("x", "v", "omega", "radius", "rmass")}
        ty = np.ascontiguousarray(particles["type"], dtype=np.int32)
        tg = np.ascontiguousarray(particles["tag"], dtype=np.int32)
        self.n = len(ty)
        self.nwalls = len(walls)
        self.s = self.lib.orc_create(ctypes.byref(self.p), ctypes.byref(self.m), self.n,
                                     _dp(P["x"]), _dp(P["v"]), _dp(P["omega"]),
                                     _dp(P["radius"]), _dp(P["rmass"]), _ip(ty), _ip(tg))
        self.ntri = []
        for mesh in meshes:
            tris, wtype, mdl, rot = mesh[:4]
            t = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
            m = model(mdl if mdl is not None else params)
            k = self.lib.orc_add_mesh(self.s, len(t), _dp(t), int(wtype), ctypes.byref(m))
            if k < 0:
                raise RuntimeError("orc_add_mesh failed")
            if isinstance(rot, dict) and "linear" in rot:
                v = np.ascontiguousarray(rot["linear"], dtype=np.float64)
                self.lib.orc_mesh_linear(self.s, k, _dp(v))
            elif isinstance(rot, dict) and "wiggle" in rot:
                a = np.ascontiguousarray(rot["wiggle"][0], dtype=np.float64)
                self.lib.orc_mesh_wiggle(self.s, k, _dp(a), float(rot["wiggle"][1]))
            elif isinstance(rot, dict) and ("viblin" in rot or "vibrot" in rot):
                args = [np.ascontiguousarray(v, dtype=np.float64) for v in rot.get("viblin", rot.get("vibrot"))]
                if "viblin" in rot:
                    ax, am, ph, pe = args
                    self.lib.orc_mesh_viblin(self.s, k, _dp(ax), len(am), _dp(am), _dp(ph), _dp(pe))
                else:
                    o, ax, am, ph, pe = args
                    self.lib.orc_mesh_vibrot(self.s, k, _dp(o), _dp(ax), len(am), _dp(am), _dp(ph), _dp(pe))
            elif rot is not None:
                o = np.ascontiguousarray(rot[0], dtype=np.float64)
                a = np.ascontiguousarray(rot[1], dtype=np.float64)
                self.lib.orc_mesh_rotate(self.s, k, _dp(o), _dp(a), float(rot[2]))
            if len(mesh) > 4 and mesh[4] is not None:
                pr = np.ascontiguousarray(mesh[4], dtype=np.float64)
                self.lib.orc_mesh_stress(self.s, k, _dp(pr))
            self.ntri.append(len(t))

    def mesh_topology(self, k=0):
        T = self.ntri[k]
        ea = np.zeros((T, 3), dtype=np.int32)
        ca = np.zeros((T, 3), dtype=np.int32)
        ob = np.zeros(T, dtype=np.int32)
        self.lib.orc_mesh_topology(self.s, k, _ip(ea), _ip(ca), _ip(ob))
        return ea, ca, ob

    def mesh_nodes(self, k=0):
        T = self.ntri[k]
        nodes, vn = np.zeros((T, 3, 3)), np.zeros((T, 3, 3))
        self.lib.orc_mesh_nodes(self.s, k, _dp(nodes), _dp(vn))
        return nodes, vn

    def mesh_history(self, k=0, maxslot=16):
        tri = np.zeros((self.n, maxslot), dtype=np.int32)
        hist = np.zeros((self.n, maxslot, 3))
        self.lib.orc_mesh_history(self.s, k, maxslot, _ip(tri), _dp(hist))
        return tri, hist

    def mesh_force(self, k=0):
        """(total force, torque about the reference point) on mesh/surface/stress k
        at the last force pass"""
        f, t = np.zeros(3), np.zeros(3)
        self.lib.orc_mesh_force(self.s, k, _dp(f), _dp(t))
        return f, t

    def mesh_forced_rebuilds(self):
        return self.lib.orc_mesh_forced_rebuilds(self.s)

    def insert(self, P):
        """Queues particles inserted at the next step of a run (after its
        initial integrate, before its forced neighbor build)."""
        arr = {k: np.ascontiguousarray(P[k], dtype=np.float64) for k in ("x", "v", "omega", "radius", "rmass")}
        ty = np.ascontiguousarray(P["type"], dtype=np.int32)
        tg = np.ascontiguousarray(P["tag"], dtype=np.int32)
        self.lib.orc_insert(self.s, len(ty), _dp(arr["x"]), _dp(arr["v"]), _dp(arr["omega"]),
                            _dp(arr["radius"]), _dp(arr["rmass"]), _ip(ty), _ip(tg))

    def set_max_radius(self, r):
        self.lib.orc_set_max_radius(self.s, float(r))

    def setup(self):
        self.lib.orc_setup(self.s)

    def run(self, nsteps):
        """Verlet::setup followed by nsteps steps (one `run` command)."""
        self.lib.orc_setup(self.s)
        self.lib.orc_run(self.s, int(nsteps))

    def run_continue(self, nsteps):
        self.lib.orc_run(self.s, int(nsteps))

    def compute_forces(self, shearupdate):
        self.lib.orc_compute_forces(self.s, int(shearupdate))

    def state(self):
        n = self.n = self.lib.orc_natoms(self.s)
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "omega", "f", "torque")}
        self.lib.orc_get_state(self.s, _dp(out["x"]), _dp(out["v"]), _dp(out["omega"]),
                               _dp(out["f"]), _dp(out["torque"]))
        return out

    def pair_local(self):
        """compute pair/gran/local's default columns per half-list pair, in
        the order of pairs() (orc_pair_local): tags (P,2), data (P,ncol)."""
        ncol = 21 + (6 if self.p.model.rolling_epsd2 else 3)
        cap = max(-int(self.lib.orc_pair_local(self.s, 0, None, None)), 1)
        tags = np.zeros((cap, 2), dtype=np.int32)
        data = np.zeros((cap, ncol))
        n = self.lib.orc_pair_local(self.s, cap, _ip(tags), _dp(data))
        if n < 0:
            raise RuntimeError("orc_pair_local: pair count changed")
        return tags[:n], data[:n]

    def wall_local(self):
        """compute wall/gran/local's rows of the mesh contacts
        (orc_wall_local): ids (R,3) = mesh, triangle, tag; data (R,26)."""
        cap = max(-int(self.lib.orc_wall_local(self.s, 0, None, None)), 1)
        ids = np.zeros((cap, 3), dtype=np.int32)
        data = np.zeros((cap, 26))
        n = self.lib.orc_wall_local(self.s, cap, _ip(ids), _dp(data))
        if n < 0:
            raise RuntimeError("orc_wall_local: row count changed")
        return ids[:n], data[:n]

    def pairs(self):
        # ask for the count first (a negative return is minus the pairs held)
        cap = max(-int(self.lib.orc_get_pairs(self.s, 0, None, None, None)), 1)
        while True:
            tags = np.zeros((cap, 2), dtype=np.int32)
            flags = np.zeros(cap, dtype=np.int32)
            hist = np.zeros((cap, 3))
            n = self.lib.orc_get_pairs(self.s, cap, _ip(tags), _ip(flags), _dp(hist))
            if n >= 0:
                return tags[:n], flags[:n], hist[:n]
            cap = int(-n) + 16

    def wall_history(self):
        out = np.zeros((self.n, max(self.nwalls, 1), 3))
        self.lib.orc_get_wall_hist(self.s, _dp(out))
        return out[:, :self.nwalls, :]

    def pair_rolling(self):
        """epsd2 rolling-torque histories, in the order and orientation of pairs()."""
        cap = 64 * self.n + 16
        while True:
            out = np.zeros((cap, 3))
            n = self.lib.orc_get_pair_rolling(self.s, cap, _dp(out))
            if n >= 0:
                return out[:n]
            cap = int(-n) + 16

    def wall_rolling(self):
        out = np.zeros((self.n, max(self.nwalls, 1), 3))
        self.lib.orc_get_wall_rolling(self.s, _dp(out))
        return out[:, :self.nwalls, :]

    def nrebuilds(self):
        return self.lib.orc_nrebuilds(self.s)

    def lost_atoms(self):
        """(atoms beyond an f face at the last build, first timestep whose
        build lost atoms or -1)"""
        first = ctypes.c_int64()
        n = self.lib.orc_lost_atoms(self.s, ctypes.byref(first))
        return int(n), int(first.value)

    def box(self):
        lo, hi = np.zeros(3), np.zeros(3)
        self.lib.orc_get_box(self.s, _dp(lo), _dp(hi))
        return lo, hi

    def contact_counts(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self.lib.orc_contact_counts(self.s, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def close(self):
        if self.s:
            self.lib.orc_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# superquadric contact point (sq_oracle.c)
# ---------------------------------------------------------------------------
SQ_POSE_DOUBLES = 12  # center[3], quat[4] (w,i,j,k), shape[3], blockiness[2]


def _sq_lib():
    lib = load()
    if not getattr(lib, "_sq_ready", False):
        P = ctypes.POINTER
        dp, ip = P(_D), P(ctypes.c_int)
        lib.orc_sq_props.argtypes = [dp, dp, dp, dp, dp]
        lib.orc_sq_contact.argtypes = [dp, dp, dp, dp, dp, dp, ip]
        lib.orc_sq_contact.restype = ctypes.c_int
        lib.orc_sq_contact_batch.argtypes = [ctypes.c_int64, dp, dp, dp, dp, dp, dp, ip, ip]
        lib.orc_sq_check_batch.argtypes = [ctypes.c_int64, dp, dp, dp, dp, dp, _D, ip, dp, dp, dp, dp, ip]
        lib.orc_sq_obb.argtypes = [dp, dp, ip]
        lib.orc_sq_obb.restype = ctypes.c_int
        lib.orc_sq_line.argtypes = [dp, dp, dp, _D, dp]
        lib.orc_sq_line.restype = _D
        lib._sq_ready = True
    return lib


def sq_props(pose, x, hessian=True):
    """Shape function value, gradient and Hessian of one superquadric pose
    (12 doubles) at global point x."""
    lib = _sq_lib()
    pose = np.ascontiguousarray(pose, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.zeros(1)
    g = np.zeros(3)
    h = np.zeros(9) if hessian else None
    lib.orc_sq_props(_dp(pose), _dp(x), _dp(f), _dp(g), _dp(h) if h is not None else None)
    return f[0], g, (h.reshape(3, 3) if h is not None else None)


def sq_contact_batch(A, B, init=None):
    """Contact points of n pose pairs (A, B: (n, 12)); init None = fresh
    contacts (homotopy), else (n, 3) starting points. Returns
    (point (n,3), fi, fj, newton_iterations, fail)."""
    lib = _sq_lib()
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    n = A.shape[0]
    pt = np.zeros((n, 3))
    fi = np.zeros(n)
    fj = np.zeros(n)
    it = np.zeros(n, dtype=np.int32)
    fl = np.zeros(n, dtype=np.int32)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
    lib.orc_sq_contact_batch(n, _dp(A), _dp(B), _dp(ini) if ini is not None else None, _dp(pt),
                             _dp(fi), _dp(fj), it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             fl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return pt, fi, fj, it, fl


SQ_HIST_DOUBLES = 7  # state, first OBB inequality, contact point[3], alpha_i, alpha_j
SQ_FAR, SQ_CLOSE, SQ_INTERSECT = 0, 1, 2


def sq_check_batch(A, B, hist=None, vel=None, mass=None, dt=0.0):
    """checkSurfaceIntersect over n pose pairs. hist (n, 7) is updated in
    place (None = all pairs far apart); vel (n, 12) = v_i, v_j, omega_i,
    omega_j and mass (n, 2) feed the previous-step estimate (None = still).
    Returns (hist, contact, en (n,3), deltan, cp_i (n,3), cp_j (n,3),
    newton_iterations)."""
    lib = _sq_lib()
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    n = A.shape[0]
    if hist is None:
        hist = np.zeros((n, SQ_HIST_DOUBLES))
    assert hist.shape == (n, SQ_HIST_DOUBLES) and hist.dtype == np.float64 and hist.flags.c_contiguous
    v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
    m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
    contact = np.zeros(n, dtype=np.int32)
    en = np.zeros((n, 3))
    dn = np.zeros(n)
    cpa = np.zeros((n, 3))
    cpb = np.zeros((n, 3))
    it = np.zeros(n, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.orc_sq_check_batch(n, _dp(A), _dp(B), _dp(hist), _dp(v) if v is not None else None,
                           _dp(m) if m is not None else None, dt, contact.ctypes.data_as(ip), _dp(en),
                           _dp(dn), _dp(cpa), _dp(cpb), it.ctypes.data_as(ip))
    return hist, contact, en, dn, cpa, cpb, it


def sq_obb(a, b, start=0):
    """obb_intersect of two poses: (intersect, first inequality for next time)."""
    lib = _sq_lib()
    st = ctypes.c_int(start)
    r = lib.orc_sq_obb(_dp(np.ascontiguousarray(a, dtype=np.float64)),
                       _dp(np.ascontiguousarray(b, dtype=np.float64)), ctypes.byref(st))
    return bool(r), st.value


def sq_line(pose, start, direction, alpha_prev=0.0):
    """surface_line_intersection: (alpha, point) with point = start + alpha * direction."""
    lib = _sq_lib()
    res = np.zeros(3)
    a = lib.orc_sq_line(_dp(np.ascontiguousarray(pose, dtype=np.float64)),
                        _dp(np.ascontiguousarray(start, dtype=np.float64)),
                        _dp(np.ascontiguousarray(direction, dtype=np.float64)), alpha_prev, _dp(res))
    return a, res


# ---------------------------------------------------------------------------
# superquadric DEM step (sqdem_oracle.c)
# ---------------------------------------------------------------------------
SQS_HIST = 10  # state, first OBB inequality, contact point[3], alpha_i, alpha_j, shear[3]


class OrcSqsParams(ctypes.Structure):
    _fields_ = [
        ("dt", _D), ("skin", _D),
        ("boxlo", _D * 3), ("boxhi", _D * 3), ("periodic", ctypes.c_int * 3),
        ("tangential_damping", ctypes.c_int), ("limit_force", ctypes.c_int),
        ("has_gravity", ctypes.c_int), ("grav_acc", _D * 3),
        ("nplanes", ctypes.c_int), ("plane_axis", ctypes.c_int * MAXPLANE),
        ("plane_pos", _D * MAXPLANE), ("plane_type", ctypes.c_int * MAXPLANE),
        ("gravity_first", ctypes.c_int),
    ]


def _sqs_lib():
    lib = _sq_lib()
    if not getattr(lib, "_sqs_ready", False):
        P = ctypes.POINTER
        dp, ip, vp = P(_D), P(ctypes.c_int), ctypes.c_void_p
        i64p = P(ctypes.c_int64)
        lib.orc_sqs_mass_props.argtypes = [dp, dp, _D, dp, dp]
        lib.orc_sq_bounding_radius.argtypes = [dp, dp]
        lib.orc_sq_bounding_radius.restype = _D
        lib.orc_sq_plane.argtypes = [dp, dp, dp, dp, dp]
        lib.orc_sq_plane.restype = ctypes.c_int
        lib.orc_sqs_contact_pair.argtypes = [P(OrcMaterials), ctypes.c_int, ctypes.c_int, _D, ctypes.c_int,
                                             ctypes.c_int, dp, dp, dp, dp, dp, dp, _D, _D, _D, _D, ctypes.c_int,
                                             dp, dp, dp, dp]
        lib.orc_sqs_contact_pair.restype = ctypes.c_int
        lib.orc_sqs_create.argtypes = [P(OrcSqsParams), P(OrcMaterials), ctypes.c_int, dp, dp, dp, dp, dp, dp,
                                       dp, ip]
        lib.orc_sqs_create.restype = vp
        lib.orc_sqs_destroy.argtypes = [vp]
        lib.orc_sqs_setup.argtypes = [vp]
        lib.orc_sqs_run.argtypes = [vp, ctypes.c_int]
        lib.orc_sqs_get_state.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp]
        lib.orc_sqs_get_props.argtypes = [vp, dp, dp, dp, dp]
        lib.orc_sqs_get_pairs.argtypes = [vp, ctypes.c_int64, ip, ip, dp]
        lib.orc_sqs_get_pairs.restype = ctypes.c_int64
        lib.orc_sqs_get_wall_hist.argtypes = [vp, dp]
        lib.orc_sqs_stats.argtypes = [vp, i64p, i64p, i64p, i64p, i64p]
        lib._sqs_ready = True
    return lib


def sq_mass_props(shape, blk, density):
    """(volume, principal inertia (3,)) of one superquadric."""
    lib = _sqs_lib()
    vol = np.zeros(1)
    I = np.zeros(3)
    lib.orc_sqs_mass_props(_dp(np.ascontiguousarray(shape, dtype=np.float64)),
                           _dp(np.ascontiguousarray(blk, dtype=np.float64)), float(density), _dp(vol), _dp(I))
    return vol[0], I


def sq_bounding_radius(shape, blk):
    return _sqs_lib().orc_sq_bounding_radius(_dp(np.ascontiguousarray(shape, dtype=np.float64)),
                                             _dp(np.ascontiguousarray(blk, dtype=np.float64)))


def sq_plane(pose, normal, x0):
    """Superquadric::plane_intersection: (intersect, point of maximal penetration, point of lowest potential)."""
    lib = _sqs_lib()
    pm, pl = np.zeros(3), np.zeros(3)
    hit = lib.orc_sq_plane(_dp(np.ascontiguousarray(pose, dtype=np.float64)),
                           _dp(np.ascontiguousarray(normal, dtype=np.float64)),
                           _dp(np.ascontiguousarray(x0, dtype=np.float64)), _dp(pm), _dp(pl))
    return bool(hit), pm, pl


def sqs_contact_pair(mats, a, b, vi, vj, wi, wj, mi, mj, rvi, rvj, dt, shearupdate, hist=None, itype=1, jtype=1,
                     tangential_damping=1, limit_force=0):
    """One superquadric pair contact: (touch, dFi, dTi, dTj, hist_after)."""
    lib = _sqs_lib()
    arr = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    h = np.zeros(SQS_HIST) if hist is None else arr(hist).copy()
    dF, dTi, dTj = np.zeros(3), np.zeros(3), np.zeros(3)
    touch = lib.orc_sqs_contact_pair(ctypes.byref(mats), int(tangential_damping), int(limit_force), float(dt),
                                     int(itype), int(jtype), _dp(arr(a)), _dp(arr(b)), _dp(arr(vi)), _dp(arr(vj)),
                                     _dp(arr(wi)), _dp(arr(wj)), float(mi), float(mj), float(rvi), float(rvj),
                                     int(shearupdate), _dp(h), _dp(dF), _dp(dTi), _dp(dTj))
    return bool(touch), dF, dTi, dTj, h


def make_sqs_params(params, walls=()):
    p = OrcSqsParams()
    p.dt = params["dt"]
    p.skin = params["skin"]
    for d in range(3):
        p.boxlo[d] = params["boxlo"][d]
        p.boxhi[d] = params["boxhi"][d]
        p.periodic[d] = int(params["periodic"][d])
    p.tangential_damping = int(params.get("tangential_damping", 1))
    p.limit_force = int(params.get("limit_force", 0))
    g = params.get("grav_acc")
    p.has_gravity = int(g is not None)
    if g is not None:
        for d in range(3):
            p.grav_acc[d] = g[d]
    p.nplanes = len(walls)
    for k, wall in enumerate(walls):
        axis, pos, wtype = wall[:3]
        p.plane_axis[k] = int(axis)
        p.plane_pos[k] = float(pos)
        p.plane_type[k] = int(wtype)
    p.gravity_first = int(params.get("gravity_first", 1))
    return p


class SqsSystem:
    """The oracle's superquadric DEM system. P: dict of arrays x, v, omega (n,3),
    quat (n,4), shape (n,3), blockiness (n,2), density (n,), type (n,)."""

    def __init__(self, params, mats, P, walls=()):
        lib = _sqs_lib()
        self.lib = lib
        self.n = len(P["x"])
        self.nplanes = len(walls)
        self.p = make_sqs_params(params, walls)
        self.mats = mats
        arr = lambda k, shape: np.ascontiguousarray(P[k], dtype=np.float64).reshape(shape)
        self._keep = [arr("x", (self.n, 3)), arr("v", (self.n, 3)), arr("omega", (self.n, 3)),
                      arr("quat", (self.n, 4)), arr("shape", (self.n, 3)), arr("blockiness", (self.n, 2)),
                      arr("density", (self.n,)), np.ascontiguousarray(P["type"], dtype=np.int32)]
        k = self._keep
        self.h = lib.orc_sqs_create(ctypes.byref(self.p), ctypes.byref(mats), self.n, _dp(k[0]), _dp(k[1]),
                                    _dp(k[2]), _dp(k[3]), _dp(k[4]), _dp(k[5]), _dp(k[6]), _ip(k[7]))

    def setup(self):
        self.lib.orc_sqs_setup(self.h)

    def run(self, nsteps):
        self.lib.orc_sqs_run(self.h, int(nsteps))

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "omega", "angmom", "f", "torque")}
        out["quat"] = np.zeros((n, 4))
        self.lib.orc_sqs_get_state(self.h, _dp(out["x"]), _dp(out["v"]), _dp(out["omega"]), _dp(out["angmom"]),
                                   _dp(out["quat"]), _dp(out["f"]), _dp(out["torque"]))
        return out

    def props(self):
        n = self.n
        m, rv, rb, I = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros((n, 3))
        self.lib.orc_sqs_get_props(self.h, _dp(m), _dp(rv), _dp(rb), _dp(I))
        return {"mass": m, "rvol": rv, "rbound": rb, "inertia": I}

    def pairs(self):
        need = -self.lib.orc_sqs_get_pairs(self.h, 0, None, None, None)
        cap = max(int(need), 1)
        ij = np.zeros((cap, 2), dtype=np.int32)
        fl = np.zeros(cap, dtype=np.int32)
        h = np.zeros((cap, SQS_HIST))
        m = self.lib.orc_sqs_get_pairs(self.h, cap, _ip(ij), _ip(fl), _dp(h))
        return ij[:m], fl[:m], h[:m]

    def wall_hist(self):
        h = np.zeros((self.n, max(self.nplanes, 1), 3))
        if self.nplanes:
            self.lib.orc_sqs_get_wall_hist(self.h, _dp(h))
        return h[:, :self.nplanes]

    def stats(self):
        v = [ctypes.c_int64() for _ in range(5)]
        self.lib.orc_sqs_stats(self.h, *[ctypes.byref(x) for x in v])
        return dict(zip(("pair_touch", "wall_touch", "newton", "rebuilds", "steps"), [x.value for x in v]))

    def close(self):
        if self.h:
            self.lib.orc_sqs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# particle insertion: fix insert/pack + particletemplate/sphere +
# particledistribution/discrete (insert_oracle.c)
# ---------------------------------------------------------------------------
INSERT_DEFAULTS = dict(maxattempt=50, overlapcheck=1, all_in=0, exact_number=1, vel_style=0,
                       v_insert=(0.0, 0.0, 0.0), v_fluct=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0),
                       check_dist_from_subdomain_border=1, ntry_mc=100000, volumefraction_region=0.0,
                       particles_in_region=0, mass_in_region=0.0, box_change=0, dist_numberbased=0)


def insert_params(cls, **kw):
    """Fills an OriParams-like ctypes structure (the oracle's or pydem's, whose
    field names are given by `names`) from keyword arguments:
    templates=[(seed, type, radius, density)], weights=[...], dist_seed, seed,
    region=(lo, hi), box=(lo, hi) and the insert/pack options."""
    d = dict(INSERT_DEFAULTS, **kw)
    p = cls()
    t = d["templates"]
    p.ntemplates = len(t)
    for i, (sd, ty, r, rho) in enumerate(t):
        p.tmpl_seed[i], p.tmpl_type[i], p.tmpl_radius[i], p.tmpl_density[i] = sd, ty, r, rho
        p.dist_weight[i] = d["weights"][i]
    p.dist_seed = d["dist_seed"]
    p.dist_numberbased = int(d["dist_numberbased"])
    for k in ("seed", "maxattempt", "overlapcheck", "all_in", "exact_number", "vel_style",
              "check_dist_from_subdomain_border", "ntry_mc", "particles_in_region", "box_change"):
        setattr(p, k, int(d[k]))
    p.volumefraction_region = float(d["volumefraction_region"])
    p.mass_in_region = float(d["mass_in_region"])
    for dim in range(3):
        p.v_insert[dim], p.v_fluct[dim], p.omega[dim] = d["v_insert"][dim], d["v_fluct"][dim], d["omega"][dim]
        p.region_lo[dim], p.region_hi[dim] = d["region"][0][dim], d["region"][1][dim]
        p.boxlo[dim], p.boxhi[dim] = d["box"][0][dim], d["box"][1][dim]
    return p


class OracleInserter:
    """fix insert/pack (oracle restatement)."""

    def __init__(self, **kw):
        self.lib = load()
        self.p = insert_params(OriParams, **kw)
        self.s = self.lib.ori_create(ctypes.byref(self.p))
        if not self.s:
            raise ValueError("insert/pack: invalid setting")

    def region_volume(self):
        return self.lib.ori_region_volume(self.s)

    def distribution(self):
        nt = self.p.ntemplates
        ve, me = ctypes.c_double(), ctypes.c_double()
        w = np.zeros(nt)
        order = np.zeros(nt, dtype=np.int32)
        self.lib.ori_distribution(self.s, ctypes.byref(ve), ctypes.byref(me), _dp(w), _ip(order))
        return ve.value, me.value, w, order

    def rng_states(self):
        st = np.zeros(4, dtype=np.int32)
        self.lib.ori_rng_states(self.s, _ip(st))
        return st

    def step(self, x=None, radius=None, rmass=None, box=None, cap=None):
        """One insertion; returns (dict of the inserted particles, requested)."""
        x = np.zeros((0, 3)) if x is None else np.ascontiguousarray(x, dtype=np.float64)
        n = len(x)
        radius = np.zeros(0) if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
        rmass = np.zeros(0) if rmass is None else np.ascontiguousarray(rmass, dtype=np.float64)
        blo = np.ascontiguousarray(box[0] if box else self.p.boxlo, dtype=np.float64)
        bhi = np.ascontiguousarray(box[1] if box else self.p.boxhi, dtype=np.float64)
        cap = cap or 200000
        out = {k: np.zeros((cap, 3)) for k in ("x", "v", "omega")}
        out.update({k: np.zeros(cap) for k in ("radius", "density", "rmass")})
        out["type"] = np.zeros(cap, dtype=np.int32)
        req = ctypes.c_int()
        k = self.lib.ori_insert_step(self.s, n, _dp(x), _dp(radius), _dp(rmass), _dp(blo), _dp(bhi), cap,
                                     _dp(out["x"]), _dp(out["v"]), _dp(out["omega"]), _dp(out["radius"]),
                                     _dp(out["density"]), _dp(out["rmass"]), _ip(out["type"]),
                                     ctypes.byref(req))
        if k < 0:
            raise ValueError("insert/pack: insertion failed")
        return {kk: vv[:k] for kk, vv in out.items()}, req.value

    def close(self):
        if self.s:
            self.lib.ori_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def park_uniforms(seed, n):
    """n draws of the oracle's Park-Miller stream from `seed`."""
    lib = load()
    r = OriPark()
    lib.ori_park_init(ctypes.byref(r), int(seed))
    return np.array([lib.ori_park_uniform(ctypes.byref(r)) for _ in range(n)]), r.seed
