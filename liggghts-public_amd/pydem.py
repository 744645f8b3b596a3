"""ctypes binding of libdem_amd.so (the C ABI in include/liggghts_amd.h).

Used by the tests and bench.py to drive the GPU library from Python with
numpy arrays. The library must have been built in-tree (``make -C
liggghts-public_amd/csrc``); loading fails loudly otherwise, there is no
fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdem_amd.so")

DEM_FP64 = 0
DEM_MIXED = 1

# every symbol declared in include/liggghts_amd.h
EXPORTED = [
    "dem_create", "dem_destroy", "dem_last_error", "dem_set_materials",
    "dem_set_particles", "dem_add_plane_wall", "dem_setup", "dem_run",
    "dem_run_continue", "dem_compute_forces", "dem_get_state", "dem_get_pairs",
    "dem_get_wall_history", "dem_get_pair_rolling", "dem_get_wall_rolling", "dem_get_stats",
    "dem_set_kernel_timing",
    "dem_run_script", "dem_add_mesh", "dem_mesh_rotate", "dem_get_mesh_nodes", "dem_mesh_linear", "dem_mesh_wiggle", "dem_mesh_stress", "dem_get_mesh_force", "dem_mesh_viblin", "dem_mesh_vibrot",
    "dem_write_restart", "dem_read_restart", "dem_restart_info", "dem_restart_particles",
    "dem_get_mesh_topology", "dem_get_mesh_history", "dem_set_instances", "dem_get_box", "dem_compute_pair_local", "dem_compute_wall_local",
    "dem_mesh_generate",
    "dem_set_instance_materials", "dem_get_instance_stats", "dem_set_step_shape",
    "dem_get_step_shape", "dem_set_graph_capture", "dem_set_pair_model", "dem_set_model_coefficients",
    "dem_sq_create", "dem_sq_destroy",
    "dem_sq_last_error", "dem_sq_set_pairs", "dem_sq_solve", "dem_sq_check", "dem_sq_get_results",
    "dem_sq_get_features", "dem_sq_net_init", "dem_sq_net_set_params", "dem_sq_net_get_params",
    "dem_sq_net_forward", "dem_sq_net_gradient", "dem_sq_net_gradient_noalloc", "dem_sq_net_train",
    "dem_sqdem_create", "dem_sqdem_destroy", "dem_sqdem_last_error", "dem_sqdem_set_materials",
    "dem_sqdem_set_particles", "dem_sqdem_set_warm_start", "dem_sqdem_setup", "dem_sqdem_run",
    "dem_sqdem_get_state",
    "dem_sqdem_get_props", "dem_sqdem_get_pairs", "dem_sqdem_get_wall_history", "dem_sqdem_get_stats",
    "dem_insert_particles", "dem_set_max_radius", "dem_inserter_create", "dem_inserter_destroy",
    "dem_inserter_last_error", "dem_inserter_step", "dem_inserter_info",
]

DEM_SQ_FRESH, DEM_SQ_GIVEN, DEM_SQ_NETWORK = 0, 1, 2
DEM_INSERT_MAXTEMPLATES = 8


class DemInsertParams(ctypes.Structure):
    _fields_ = [
        ("ntemplates", ctypes.c_int),
        ("template_seed", ctypes.c_int * DEM_INSERT_MAXTEMPLATES),
        ("template_type", ctypes.c_int * DEM_INSERT_MAXTEMPLATES),
        ("template_radius", ctypes.c_double * DEM_INSERT_MAXTEMPLATES),
        ("template_density", ctypes.c_double * DEM_INSERT_MAXTEMPLATES),
        ("distribution_seed", ctypes.c_int),
        ("distribution_numberbased", ctypes.c_int),
        ("distribution_weight", ctypes.c_double * DEM_INSERT_MAXTEMPLATES),
        ("seed", ctypes.c_int), ("maxattempt", ctypes.c_int), ("overlapcheck", ctypes.c_int),
        ("all_in", ctypes.c_int), ("exact_number", ctypes.c_int), ("vel_style", ctypes.c_int),
        ("v_insert", ctypes.c_double * 3), ("v_fluct", ctypes.c_double * 3), ("omega", ctypes.c_double * 3),
        ("check_dist_from_subdomain_border", ctypes.c_int), ("ntry_mc", ctypes.c_int),
        ("volumefraction_region", ctypes.c_double), ("particles_in_region", ctypes.c_int),
        ("mass_in_region", ctypes.c_double), ("region_lo", ctypes.c_double * 3),
        ("region_hi", ctypes.c_double * 3), ("boxlo", ctypes.c_double * 3), ("boxhi", ctypes.c_double * 3),
        ("box_change", ctypes.c_int), ("warn_region", ctypes.c_int), ("fixed_boundary", ctypes.c_int),
    ]


INSERT_DEFAULTS = dict(maxattempt=50, overlapcheck=1, all_in=0, exact_number=1, vel_style=0,
                       v_insert=(0.0, 0.0, 0.0), v_fluct=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0),
                       check_dist_from_subdomain_border=1, ntry_mc=100000, volumefraction_region=0.0,
                       particles_in_region=0, mass_in_region=0.0, box_change=0, dist_numberbased=0,
                       warn_region=1, fixed_boundary=0)


def insert_params(**kw):
    """DemInsertParams from keyword arguments: templates=[(seed, type,
    radius, density)], weights, dist_seed, seed, region=(lo, hi), box=(lo, hi)
    and the insert/pack options (INSERT_DEFAULTS)."""
    d = dict(INSERT_DEFAULTS, **kw)
    p = DemInsertParams()
    p.ntemplates = len(d["templates"])
    for i, (sd, ty, r, rho) in enumerate(d["templates"]):
        p.template_seed[i], p.template_type[i] = sd, ty
        p.template_radius[i], p.template_density[i] = r, rho
        p.distribution_weight[i] = d["weights"][i]
    p.distribution_seed = d["dist_seed"]
    p.distribution_numberbased = int(d["dist_numberbased"])
    for k in ("seed", "maxattempt", "overlapcheck", "all_in", "exact_number", "vel_style",
              "check_dist_from_subdomain_border", "ntry_mc", "particles_in_region", "box_change",
              "warn_region", "fixed_boundary"):
        setattr(p, k, int(d[k]))
    p.volumefraction_region = float(d["volumefraction_region"])
    p.mass_in_region = float(d["mass_in_region"])
    for dim in range(3):
        p.v_insert[dim], p.v_fluct[dim], p.omega[dim] = d["v_insert"][dim], d["v_fluct"][dim], d["omega"][dim]
        p.region_lo[dim], p.region_hi[dim] = d["region"][0][dim], d["region"][1][dim]
        p.boxlo[dim], p.boxhi[dim] = d["box"][0][dim], d["box"][1][dim]
    return p
SQ_HIST_DOUBLES = 7  # state, first OBB inequality, contact point[3], alpha_i, alpha_j
SQ_NUM_PARAMS = 5571
DEM_SQ_GRAD_NOALLOC_MAX = 16384  # include/liggghts_amd.h
DEM_ERR_ARG, DEM_ERR_HIP, DEM_ERR_OVERFLOW, DEM_ERR_STATE = -1, -2, -3, -4


class DemParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("skin", ctypes.c_double),
        ("cdf", ctypes.c_double),
        ("boxlo", ctypes.c_double * 3),
        ("boxhi", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("rolling_cdt", ctypes.c_int),
        ("tangential_damping", ctypes.c_int),
        ("limit_force", ctypes.c_int),
        ("torsion_torque", ctypes.c_int),
        ("has_gravity", ctypes.c_int),
        ("grav_acc", ctypes.c_double * 3),
        ("gravity_first", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("max_neighbors", ctypes.c_int),
        ("device", ctypes.c_int),
        ("neigh_every", ctypes.c_int),
        ("neigh_no_check", ctypes.c_int),
        ("shrink_faces", ctypes.c_int),
        ("shrink_min", ctypes.c_int),
    ]


class DemContactModel(ctypes.Structure):
    _fields_ = [("rolling_cdt", ctypes.c_int), ("tangential_damping", ctypes.c_int),
                ("limit_force", ctypes.c_int), ("torsion_torque", ctypes.c_int),
                ("normal_model", ctypes.c_int), ("tangential_model", ctypes.c_int),
                ("cohesion_model", ctypes.c_int), ("kt_to_kn", ctypes.c_int),
                ("absolute_damping", ctypes.c_int), ("tangential_reduce", ctypes.c_int),
                ("rolling_epsd2", ctypes.c_int)]


# enum dem_normal_model and the tangential / cohesion model codes
NORMAL_MODELS = {"hertz": 0, "hooke": 1, "hertz/stiffness": 2, "hooke/stiffness": 3}
TANGENTIAL_MODELS = {"history": 0, "no_history": 1}
COHESION_MODELS = {None: 0, "off": 0, "sjkr": 1}


def contact_model(d):
    """dem_contact_model from a dict: rolling_cdt, tangential_damping,
    limit_force, torsion_torque, normal ("hertz" | "hooke" | "hertz/stiffness" |
    "hooke/stiffness"), tangential ("history" | "no_history"), cohesion
    (None | "sjkr"), kt_to_kn, absolute_damping, tangential_reduce, rolling_epsd2."""
    m = DemContactModel()
    m.rolling_cdt = int(d.get("rolling_cdt", 0))
    m.tangential_damping = int(d.get("tangential_damping", 1))
    m.limit_force = int(d.get("limit_force", 0))
    m.torsion_torque = int(d.get("torsion_torque", 0))
    m.normal_model = NORMAL_MODELS[d.get("normal", "hertz")]
    m.tangential_model = TANGENTIAL_MODELS[d.get("tangential", "history")]
    m.cohesion_model = COHESION_MODELS[d.get("cohesion")]
    m.kt_to_kn = int(d.get("kt_to_kn", 0))
    m.absolute_damping = int(d.get("absolute_damping", 0))
    m.tangential_reduce = int(d.get("tangential_reduce", 0))
    m.rolling_epsd2 = int(d.get("rolling_epsd2", 0))
    return m


MODEL_COEFFS = ("kn", "kt", "gamman", "gammat", "cohesion_energy_density")


class DemStats(ctypes.Structure):
    _fields_ = [
        ("ntimestep", ctypes.c_int64),
        ("nrebuilds", ctypes.c_int64),
        ("npairs", ctypes.c_int64),
        ("ncontacts", ctypes.c_int64),
        ("nwall_contacts", ctypes.c_int64),
        ("max_neighbors_seen", ctypes.c_int),
        ("step_kernel_ms", ctypes.c_double),
        ("step_kernel_launches", ctypes.c_int64),
        ("rebuild_ms_total", ctypes.c_double),
        ("nmesh_candidates", ctypes.c_int64),
    ]


_lib = None


class SqDemParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double), ("skin", ctypes.c_double),
        ("boxlo", ctypes.c_double * 3), ("boxhi", ctypes.c_double * 3), ("periodic", ctypes.c_int * 3),
        ("tangential_damping", ctypes.c_int), ("limit_force", ctypes.c_int),
        ("has_gravity", ctypes.c_int), ("gravity", ctypes.c_double * 3), ("gravity_first", ctypes.c_int),
        ("nplanes", ctypes.c_int), ("plane_axis", ctypes.c_int * 6), ("plane_pos", ctypes.c_double * 6),
        ("plane_type", ctypes.c_int * 6),
    ]


class SqDemStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("steps", "rebuilds", "pair_touch", "wall_touch", "newton_last",
                                              "newton", "checks", "fresh", "fresh_iters", "fallbacks")] + [
        ("run_ms", ctypes.c_double)]


def load_library(path=LIB_PATH):
    """Load libdem_amd.so; raise if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `make -C liggghts-public_amd/csrc` "
            "(or __graft_entry__.build()) before using the GPU path")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    ip = P(ctypes.c_int)
    h = ctypes.c_void_p
    lib.dem_create.argtypes = [P(DemParams), P(ctypes.c_void_p)]
    lib.dem_destroy.argtypes = [h]
    lib.dem_last_error.argtypes = [h]
    lib.dem_last_error.restype = ctypes.c_char_p
    lib.dem_set_materials.argtypes = [h, ctypes.c_int, dp, dp, dp, dp, dp]
    lib.dem_set_pair_model.argtypes = [h, P(DemContactModel)]
    lib.dem_set_model_coefficients.argtypes = [h, ctypes.c_int, dp, dp, dp, dp, dp, ctypes.c_double]
    lib.dem_set_particles.argtypes = [h, ctypes.c_int64, dp, dp, dp, dp, dp, ip, ip]
    lib.dem_add_plane_wall.argtypes = [h, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       P(DemContactModel)]
    lib.dem_setup.argtypes = [h]
    lib.dem_run.argtypes = [h, ctypes.c_int64]
    lib.dem_run_continue.argtypes = [h, ctypes.c_int64]
    lib.dem_compute_forces.argtypes = [h, ctypes.c_int]
    lib.dem_get_state.argtypes = [h, dp, dp, dp, dp, dp]
    lib.dem_get_pairs.argtypes = [h, ctypes.c_int64, ip, ip, dp, P(ctypes.c_int64)]
    lib.dem_get_wall_history.argtypes = [h, dp]
    lib.dem_get_pair_rolling.argtypes = [h, ctypes.c_int64, dp, P(ctypes.c_int64)]
    lib.dem_get_wall_rolling.argtypes = [h, dp]
    lib.dem_get_stats.argtypes = [h, P(DemStats)]
    lib.dem_set_kernel_timing.argtypes = [h, ctypes.c_int]
    lib.dem_run_script.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    lib.dem_add_mesh.argtypes = [h, ctypes.c_int64, dp, ctypes.c_int, P(DemContactModel),
                                 P(ctypes.c_int)]
    lib.dem_mesh_rotate.argtypes = [h, ctypes.c_int, dp, dp, ctypes.c_double]
    lib.dem_mesh_linear.argtypes = [h, ctypes.c_int, dp]
    lib.dem_mesh_stress.argtypes = [h, ctypes.c_int, dp]
    lib.dem_write_restart.argtypes = [h, ctypes.c_char_p]
    lib.dem_read_restart.argtypes = [h, ctypes.c_char_p]
    lib.dem_restart_info.argtypes = [ctypes.c_char_p, P(ctypes.c_int64), ip, dp, dp, ip, P(ctypes.c_int64)]
    lib.dem_restart_particles.argtypes = [ctypes.c_char_p, ctypes.c_int64, ip, ip, dp, dp, dp, dp, dp]
    lib.dem_mesh_viblin.argtypes = [h, ctypes.c_int, dp, ctypes.c_int, dp, dp, dp]
    lib.dem_mesh_vibrot.argtypes = [h, ctypes.c_int, dp, dp, ctypes.c_int, dp, dp, dp]
    lib.dem_get_mesh_force.argtypes = [h, ctypes.c_int, dp, dp]
    lib.dem_mesh_wiggle.argtypes = [h, ctypes.c_int, dp, ctypes.c_double]
    lib.dem_get_mesh_nodes.argtypes = [h, ctypes.c_int, dp, dp]
    lib.dem_get_mesh_topology.argtypes = [h, ctypes.c_int, ip, ip, ip]
    lib.dem_get_mesh_history.argtypes = [h, ctypes.c_int, ctypes.c_int, ip, dp]
    lib.dem_set_instances.argtypes = [h, ctypes.c_int, P(ctypes.c_int64)]
    lib.dem_set_instance_materials.argtypes = [h, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp]
    lib.dem_get_instance_stats.argtypes = [h, ctypes.c_int, P(ctypes.c_int64), P(ctypes.c_int64)]
    lib.dem_set_step_shape.argtypes = [h, ctypes.c_int, ctypes.c_int]
    lib.dem_set_graph_capture.argtypes = [h, ctypes.c_int]
    fp = P(ctypes.c_float)
    lib.dem_sq_create.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
    lib.dem_sq_destroy.argtypes = [h]
    lib.dem_sq_last_error.argtypes = [h]
    lib.dem_sq_last_error.restype = ctypes.c_char_p
    lib.dem_sq_set_pairs.argtypes = [h, ctypes.c_int64, dp, dp]
    lib.dem_mesh_generate.argtypes = [ctypes.c_char_p, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int64, dp,
                                      P(ctypes.c_int64)]
    lib.dem_get_box.argtypes = [h, dp, dp]
    lib.dem_compute_pair_local.argtypes = [h, ctypes.c_int64, ip, dp, P(ctypes.c_int64), ip]
    lib.dem_compute_wall_local.argtypes = [h, ctypes.c_int64, ip, dp, P(ctypes.c_int64), ip]
    lib.dem_sq_solve.argtypes = [h, ctypes.c_int, dp, ctypes.c_int]
    lib.dem_sq_get_results.argtypes = [h, dp, dp, dp, ip, ip, dp, dp]
    lib.dem_sq_check.argtypes = [h, ctypes.c_int, ctypes.c_double, dp, dp, dp, ip, dp, dp, dp, dp]
    lib.dem_sq_get_features.argtypes = [h, fp]
    lib.dem_sq_net_init.argtypes = [h, ctypes.c_uint64]
    lib.dem_sq_net_set_params.argtypes = [h, fp]
    lib.dem_sq_net_get_params.argtypes = [h, fp]
    lib.dem_sq_net_forward.argtypes = [h, ctypes.c_int64, fp, fp]
    lib.dem_sq_net_gradient.argtypes = [h, ctypes.c_int64, fp, fp, fp, fp]
    vp = ctypes.c_void_p
    lib.dem_sq_net_gradient_noalloc.argtypes = [h, ctypes.c_int64, vp, vp, vp, vp]
    lib.dem_sq_net_train.argtypes = [h, ctypes.c_int64, ctypes.c_int64, ctypes.c_float, ctypes.c_uint64,
                                     fp, dp]
    lib.dem_get_step_shape.argtypes = [h, ip, ip, ctypes.c_int, ip, ip, ip, dp]
    lib.dem_sqdem_create.argtypes = [ctypes.c_int, P(SqDemParams), P(ctypes.c_void_p)]
    lib.dem_sqdem_destroy.argtypes = [h]
    lib.dem_sqdem_last_error.argtypes = [h]
    lib.dem_sqdem_last_error.restype = ctypes.c_char_p
    lib.dem_sqdem_set_materials.argtypes = [h, ctypes.c_int, dp, dp, dp, dp]
    lib.dem_sqdem_set_particles.argtypes = [h, ctypes.c_int64, dp, dp, dp, dp, dp, dp, dp, ip]
    lib.dem_sqdem_set_warm_start.argtypes = [h, h]
    lib.dem_sqdem_run.argtypes = [h, ctypes.c_int]
    lib.dem_sqdem_setup.argtypes = [h]
    lib.dem_sqdem_get_state.argtypes = [h, dp, dp, dp, dp, dp, dp, dp]
    lib.dem_sqdem_get_props.argtypes = [h, dp, dp, dp, dp]
    lib.dem_sqdem_get_pairs.argtypes = [h, ctypes.c_int64, ip, ip, dp]
    lib.dem_sqdem_get_pairs.restype = ctypes.c_int64
    lib.dem_sqdem_get_wall_history.argtypes = [h, dp]
    lib.dem_sqdem_get_stats.argtypes = [h, P(SqDemStats)]
    lib.dem_insert_particles.argtypes = [h, ctypes.c_int64, dp, dp, dp, dp, dp, ip, ip]
    lib.dem_set_max_radius.argtypes = [h, ctypes.c_double]
    lib.dem_inserter_create.argtypes = [P(DemInsertParams), P(ctypes.c_void_p)]
    lib.dem_inserter_destroy.argtypes = [h]
    lib.dem_inserter_last_error.argtypes = [h]
    lib.dem_inserter_last_error.restype = ctypes.c_char_p
    i64p = P(ctypes.c_int64)
    lib.dem_inserter_step.argtypes = [h, ctypes.c_int64, dp, dp, dp, dp, dp, ctypes.c_int64,
                                      dp, dp, dp, dp, dp, dp, ip, i64p, i64p]
    lib.dem_inserter_info.argtypes = [h, dp, dp, dp, dp]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class DemError(RuntimeError):
    pass


def _opt(a, flat=True):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1) if flat else a


def _material_args(materials):
    """(ntypes, Y, nu, e, mu, mur) pointers for dem_set_materials (youngs,
    poisson and restitution may be absent together); the arrays are kept
    alive on this function until the next call."""
    nt = materials["ntypes"]
    Y, nu, e = (_opt(materials.get(k)) for k in ("youngs", "poisson", "restitution"))
    mu = _opt(materials["friction"])
    mur = _opt(materials.get("rolling_friction"))
    _material_args.keep = (Y, nu, e, mu, mur)
    ptr = lambda a: _dp(a) if a is not None else None
    return (nt, ptr(Y), ptr(nu), ptr(e), ptr(mu), ptr(mur))


def _coefficient_args(materials):
    """dem_set_model_coefficients arguments, or None when the material dict
    holds none of kn / kt / gamman / gammat / cohesion_energy_density /
    characteristic_velocity."""
    if not any(k in materials for k in MODEL_COEFFS + ("characteristic_velocity",)):
        return None
    arrs = [_opt(materials.get(k)) for k in MODEL_COEFFS]
    _coefficient_args.keep = arrs
    return (materials["ntypes"], *[_dp(a) if a is not None else None for a in arrs],
            float(materials.get("characteristic_velocity", 0.0)))


class GpuSystem:
    """One granular system on the GPU (a dem_handle)."""

    def __init__(self, params, materials, particles, walls=(), precision=DEM_FP64,
                 max_neighbors=0, device=0, meshes=(), instances=None, instance_materials=None):
        """meshes: [(tris (T,3,3), wall_type, model dict or None, rotate or None)],
        rotate = (origin, axis, period) for fix move/mesh rotate, or
        {"linear": (vx, vy, vz)}, {"wiggle": ((ax, ay, az), period)},
        {"viblin": (axis, amplitudes, phases, periods)} or {"vibrot": (origin,
        axis, amplitudes, phases, periods)}; an optional
        fifth entry (a reference point) makes the mesh a mesh/surface/stress.
        instances: particle counts of independent simulations batched in this
        system (particles concatenated in that order); instance_materials: one
        material dict per instance."""
        self.lib = load_library()
        p = DemParams()
        p.dt = params["dt"]
        p.skin = params["skin"]
        p.cdf = params.get("cdf", 1.0)
        for d in range(3):
            p.boxlo[d] = params["boxlo"][d]
            p.boxhi[d] = params["boxhi"][d]
            p.periodic[d] = int(params["periodic"][d])
        p.rolling_cdt = int(params.get("rolling_cdt", 0))
        p.tangential_damping = int(params.get("tangential_damping", 1))
        p.limit_force = int(params.get("limit_force", 0))
        p.torsion_torque = int(params.get("torsion_torque", 0))
        g = params.get("grav_acc")
        p.has_gravity = int(g is not None)
        if g is not None:
            for d in range(3):
                p.grav_acc[d] = g[d]
        p.gravity_first = int(params.get("gravity_first", 1))
        p.precision = precision
        p.max_neighbors = max_neighbors
        p.device = device
        p.neigh_every = int(params.get("neigh_every", 1))
        p.neigh_no_check = int(not params.get("neigh_check", True))
        # boundary s / m faces (bit 2d lower, 2d+1 upper face of dimension d)
        p.shrink_faces = int(params.get("shrink_faces", 0))
        p.shrink_min = int(params.get("shrink_min", 0))
        self.h = ctypes.c_void_p()
        rc = self.lib.dem_create(ctypes.byref(p), ctypes.byref(self.h))
        if rc != 0:
            raise DemError(f"dem_create failed ({rc}): {self.lib.dem_last_error(None).decode()}")
        self.precision = precision
        self._check(self.lib.dem_set_materials(self.h, *_material_args(materials)))
        coeffs = _coefficient_args(materials)
        if coeffs is not None:
            self._check(self.lib.dem_set_model_coefficients(self.h, *coeffs))
        pm = contact_model(params)
        self._check(self.lib.dem_set_pair_model(self.h, ctypes.byref(pm)))
        self.ninst = 1
        if instances is not None:
            counts = np.ascontiguousarray(instances, dtype=np.int64)
            self.ninst = len(counts)
            self._check(self.lib.dem_set_instances(
                self.h, self.ninst, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
            for q, m in enumerate(instance_materials or ()):
                self._check(self.lib.dem_set_instance_materials(self.h, q, *_material_args(m)))
        self.nwalls = 0
        for wall in walls:
            axis, pos, wtype = wall[:3]
            m = contact_model(wall[3] if len(wall) > 3 else params)
            self._check(self.lib.dem_add_plane_wall(self.h, int(axis), float(pos), int(wtype),
                                                    ctypes.byref(m)))
            self.nwalls += 1
        self.ntri = []
        for mesh in meshes:
            tris, wtype, mdl, rot = mesh[:4]
            t = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
            m = contact_model(mdl if mdl is not None else params)
            mid = ctypes.c_int(-1)
            self._check(self.lib.dem_add_mesh(self.h, len(t), _dp(t), int(wtype), ctypes.byref(m),
                                              ctypes.byref(mid)))
            if isinstance(rot, dict) and "linear" in rot:
                v = np.ascontiguousarray(rot["linear"], dtype=np.float64)
                self._check(self.lib.dem_mesh_linear(self.h, mid.value, _dp(v)))
            elif isinstance(rot, dict) and "wiggle" in rot:
                a = np.ascontiguousarray(rot["wiggle"][0], dtype=np.float64)
                self._check(self.lib.dem_mesh_wiggle(self.h, mid.value, _dp(a), float(rot["wiggle"][1])))
            elif isinstance(rot, dict) and ("viblin" in rot or "vibrot" in rot):
                args = [np.ascontiguousarray(v, dtype=np.float64) for v in rot.get("viblin", rot.get("vibrot"))]
                if "viblin" in rot:
                    ax, am, ph, pe = args
                    self._check(self.lib.dem_mesh_viblin(self.h, mid.value, _dp(ax), len(am), _dp(am), _dp(ph),
                                                         _dp(pe)))
                else:
                    o, ax, am, ph, pe = args
                    self._check(self.lib.dem_mesh_vibrot(self.h, mid.value, _dp(o), _dp(ax), len(am), _dp(am),
                                                         _dp(ph), _dp(pe)))
            elif rot is not None:
                o = np.ascontiguousarray(rot[0], dtype=np.float64)
                a = np.ascontiguousarray(rot[1], dtype=np.float64)
                self._check(self.lib.dem_mesh_rotate(self.h, mid.value, _dp(o), _dp(a),
                                                     float(rot[2])))
            if len(mesh) > 4 and mesh[4] is not None:
                pr = np.ascontiguousarray(mesh[4], dtype=np.float64)
                self._check(self.lib.dem_mesh_stress(self.h, mid.value, _dp(pr)))
            self.ntri.append(len(t))
        self.set_particles(particles)

    def _check(self, rc):
        if rc != 0:
            raise DemError(f"libdem_amd error {rc}: {self.lib.dem_last_error(self.h).decode()}")

    def insert(self, P):
        """dem_insert_particles: the particles join at the next run's first step."""
        m = int(P["x"].shape[0])
        arr = {k: np.ascontiguousarray(P[k], dtype=np.float64)
               for k in ("x", "v", "omega", "radius", "rmass")}
        ty = np.ascontiguousarray(P["type"], dtype=np.int32)
        tg = np.ascontiguousarray(P["tag"], dtype=np.int32)
        self._check(self.lib.dem_insert_particles(
            self.h, m, _dp(arr["x"]), _dp(arr["v"]), _dp(arr["omega"]),
            _dp(arr["radius"]), _dp(arr["rmass"]), _ip(ty), _ip(tg)))
        self._pending += m

    def set_max_radius(self, r):
        self._check(self.lib.dem_set_max_radius(self.h, float(r)))

    def _joined(self):
        self.n += self._pending
        self._pending = 0

    def set_particles(self, P):
        self._pending = 0
        if P is None:  # an empty system, filled by insert()
            self.n = 0
            return
        self.n = int(P["x"].shape[0])
        arr = {k: np.ascontiguousarray(P[k], dtype=np.float64)
               for k in ("x", "v", "omega", "radius", "rmass")}
        ty = np.ascontiguousarray(P["type"], dtype=np.int32)
        tg = np.ascontiguousarray(P["tag"], dtype=np.int32)
        self._check(self.lib.dem_set_particles(
            self.h, self.n, _dp(arr["x"]), _dp(arr["v"]), _dp(arr["omega"]),
            _dp(arr["radius"]), _dp(arr["rmass"]), _ip(ty), _ip(tg)))

    def setup(self):
        self._check(self.lib.dem_setup(self.h))

    def run(self, nsteps):
        self._check(self.lib.dem_run(self.h, int(nsteps)))
        if nsteps > 0:
            self._joined()

    def run_continue(self, nsteps):
        self._check(self.lib.dem_run_continue(self.h, int(nsteps)))
        if nsteps > 0:
            self._joined()

    def compute_forces(self, shearupdate):
        self._check(self.lib.dem_compute_forces(self.h, int(shearupdate)))

    def set_kernel_timing(self, on):
        self._check(self.lib.dem_set_kernel_timing(self.h, int(on)))

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "omega", "f", "torque")}
        self._check(self.lib.dem_get_state(self.h, _dp(out["x"]), _dp(out["v"]),
                                           _dp(out["omega"]), _dp(out["f"]),
                                           _dp(out["torque"])))
        return out

    def pairs(self):
        """Half-list pairs: (tags (P,2) lower tag first, flags (P,), hist (P,3))."""
        total = ctypes.c_int64(0)
        self._check(self.lib.dem_get_pairs(self.h, 0, None, None, None, ctypes.byref(total)))
        cap = max(int(total.value), 1)
        tags = np.zeros((cap, 2), dtype=np.int32)
        flags = np.zeros(cap, dtype=np.int32)
        hist = np.zeros((cap, 3))
        self._check(self.lib.dem_get_pairs(self.h, cap, _ip(tags), _ip(flags), _dp(hist),
                                           ctypes.byref(total)))
        npairs = int(total.value)
        return tags[:npairs], flags[:npairs], hist[:npairs]

    def pair_local(self):
        """compute pair/gran/local on the current state: (tags (P,2), data (P,ncol)),
        data columns x_i, x_j, v_i, v_j, force on i, torque on i, periodic flag,
        r_i, r_j, history (see dem_compute_pair_local)."""
        total = ctypes.c_int64(0)
        nc = ctypes.c_int(0)
        self._check(self.lib.dem_compute_pair_local(self.h, 0, None, None, ctypes.byref(total), ctypes.byref(nc)))
        cap = max(int(total.value), 1)
        tags = np.zeros((cap, 2), dtype=np.int32)
        data = np.zeros((cap, nc.value))
        self._check(self.lib.dem_compute_pair_local(self.h, cap, _ip(tags), _dp(data), ctypes.byref(total),
                                                    ctypes.byref(nc)))
        npairs = int(total.value)
        return tags[:npairs], data[:npairs]

    def wall_local(self):
        """compute wall/gran/local on the current state: (ids (R,3) = mesh,
        triangle, tag; data (R,26)), data columns contact point, x, wall
        velocity, v, force, torque, unit normal, r^2, radius, history (see
        dem_compute_wall_local)."""
        total = ctypes.c_int64(0)
        nc = ctypes.c_int(0)
        self._check(self.lib.dem_compute_wall_local(self.h, 0, None, None, ctypes.byref(total), ctypes.byref(nc)))
        cap = max(int(total.value), 1)
        ids = np.zeros((cap, 3), dtype=np.int32)
        data = np.zeros((cap, nc.value))
        self._check(self.lib.dem_compute_wall_local(self.h, cap, _ip(ids), _dp(data), ctypes.byref(total),
                                                    ctypes.byref(nc)))
        nrows = int(total.value)
        return ids[:nrows], data[:nrows]

    def box(self):
        """(boxlo, boxhi) at the last neighbor build (shrink-wrapped faces
        follow the atoms)."""
        lo, hi = np.zeros(3), np.zeros(3)
        self._check(self.lib.dem_get_box(self.h, _dp(lo), _dp(hi)))
        return lo, hi

    def wall_history(self):
        out = np.zeros((self.n, max(self.nwalls, 1), 3))
        if self.nwalls:
            self._check(self.lib.dem_get_wall_history(self.h, _dp(out)))
        return out[:, :self.nwalls, :]

    def pair_rolling(self):
        """epsd2 rolling-torque histories (P,3), in the order and orientation of pairs()."""
        total = ctypes.c_int64(0)
        self._check(self.lib.dem_get_pair_rolling(self.h, 0, None, ctypes.byref(total)))
        cap = max(int(total.value), 1)
        out = np.zeros((cap, 3))
        self._check(self.lib.dem_get_pair_rolling(self.h, cap, _dp(out), ctypes.byref(total)))
        return out[:int(total.value)]

    def wall_rolling(self):
        out = np.zeros((self.n, max(self.nwalls, 1), 3))
        if self.nwalls:
            self._check(self.lib.dem_get_wall_rolling(self.h, _dp(out)))
        return out[:, :self.nwalls, :]

    def mesh_nodes(self, k=0):
        T = self.ntri[k]
        nodes, vn = np.zeros((T, 3, 3)), np.zeros((T, 3, 3))
        self._check(self.lib.dem_get_mesh_nodes(self.h, k, _dp(nodes), _dp(vn)))
        return nodes, vn

    def mesh_topology(self, k=0):
        T = self.ntri[k]
        ea = np.zeros((T, 3), dtype=np.int32)
        ca = np.zeros((T, 3), dtype=np.int32)
        ob = np.zeros(T, dtype=np.int32)
        self._check(self.lib.dem_get_mesh_topology(self.h, k, _ip(ea), _ip(ca), _ip(ob)))
        return ea, ca, ob

    def mesh_history(self, k=0, maxslot=16):
        tri = np.zeros((self.n, maxslot), dtype=np.int32)
        hist = np.zeros((self.n, maxslot, 3))
        self._check(self.lib.dem_get_mesh_history(self.h, k, maxslot, _ip(tri), _dp(hist)))
        return tri, hist

    def set_step_shape(self, variant, persist=0):
        """Tile shape of the step kernel; variant -1 autotunes during the next run."""
        self._check(self.lib.dem_set_step_shape(self.h, int(variant), int(persist)))

    def set_graph_capture(self, on):
        """Replay step batches as captured HIP graphs."""
        self._check(self.lib.dem_set_graph_capture(self.h, int(bool(on))))

    def step_shape(self):
        """(variant, persist, [(variant, persist, us_per_step) per tuned candidate])."""
        v, f, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        cap = 64
        cv = np.zeros(cap, dtype=np.int32)
        cf = np.zeros(cap, dtype=np.int32)
        cu = np.zeros(cap)
        self._check(self.lib.dem_get_step_shape(self.h, ctypes.byref(v), ctypes.byref(f), cap,
                                                ctypes.byref(nc), _ip(cv), _ip(cf), _dp(cu)))
        n = min(nc.value, cap)
        return v.value, f.value, [(int(cv[k]), int(cf[k]), float(cu[k])) for k in range(n)]

    def write_restart(self, path):
        self._check(self.lib.dem_write_restart(self.h, str(path).encode()))

    def read_restart(self, path):
        self._check(self.lib.dem_read_restart(self.h, str(path).encode()))

    def mesh_force(self, k=0):
        """(total force, torque about the reference point) on mesh/surface/stress k
        at the last counted force pass"""
        f, t = np.zeros(3), np.zeros(3)
        self._check(self.lib.dem_get_mesh_force(self.h, k, _dp(f), _dp(t)))
        return f, t

    def instance_stats(self):
        """(touching pairs, wall contacts) per instance at the last step."""
        c = np.zeros(self.ninst, dtype=np.int64)
        w = np.zeros(self.ninst, dtype=np.int64)
        P = ctypes.POINTER(ctypes.c_int64)
        self._check(self.lib.dem_get_instance_stats(self.h, self.ninst, c.ctypes.data_as(P),
                                                    w.ctypes.data_as(P)))
        return c, w

    def stats(self):
        s = DemStats()
        self._check(self.lib.dem_get_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in DemStats._fields_}

    def close(self):
        if self.h:
            self.lib.dem_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mesh_generate(kind, dims, nsegments=0, mask=0):
    """Triangles (T, 3, 3) of fix mesh/surface ... generate <kind> (host only)."""
    lib = load_library()
    d = np.zeros(3)
    d[:len(dims)] = dims
    nt = ctypes.c_int64()
    rc = lib.dem_mesh_generate(kind.encode(), _dp(d), int(nsegments), int(mask), 0, None, ctypes.byref(nt))
    if rc:
        raise DemError(lib.dem_last_error(None).decode())
    out = np.zeros((max(nt.value, 1), 9))
    lib.dem_mesh_generate(kind.encode(), _dp(d), int(nsegments), int(mask), nt.value, _dp(out), ctypes.byref(nt))
    return out[:nt.value].reshape(-1, 3, 3)


def run_script(script, log=None):
    lib = load_library()
    return lib.dem_run_script(script.encode(), None if log is None else log.encode())


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class SqSolver:
    """Superquadric contact points on the GPU (dem_sq_* of the C ABI) and the
    warm-start network."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.h = ctypes.c_void_p()
        rc = self.lib.dem_sq_create(device, ctypes.byref(self.h))
        if rc != 0:
            raise DemError(f"dem_sq_create failed ({rc}): {self.lib.dem_sq_last_error(None).decode()}")
        self.n = 0

    def _check(self, rc):
        if rc != 0:
            raise DemError(f"libdem_amd error {rc}: {self.lib.dem_sq_last_error(self.h).decode()}")

    def set_pairs(self, A, B):
        A = np.ascontiguousarray(A, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        self.n = A.shape[0]
        self._check(self.lib.dem_sq_set_pairs(self.h, self.n, _dp(A), _dp(B)))

    def solve(self, start=DEM_SQ_FRESH, init=None, polish=0):
        ini = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
        self._check(self.lib.dem_sq_solve(self.h, int(start), _dp(ini) if ini is not None else None,
                                          int(polish)))
        return self.results()

    def check(self, hist=None, vel=None, mass=None, dt=0.0, start=DEM_SQ_FRESH):
        """checkSurfaceIntersect over the pairs. hist (n, 7) is the contact
        history, updated in place (None: all pairs new); vel (n, 12) and mass
        (n, 2) feed the previous-step estimate. Returns a dict with hist,
        contact, en, deltan, cp_i, cp_j and the results() of the solve."""
        n = self.n
        if hist is None:
            hist = np.zeros((n, SQ_HIST_DOUBLES))
        if hist.shape != (n, SQ_HIST_DOUBLES) or hist.dtype != np.float64 or not hist.flags.c_contiguous:
            raise ValueError("hist must be a C-contiguous float64 array of shape (n, 7)")
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64)
        m = None if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        contact = np.zeros(n, dtype=np.int32)
        en = np.zeros((n, 3))
        dn = np.zeros(n)
        cpi = np.zeros((n, 3))
        cpj = np.zeros((n, 3))
        self._check(self.lib.dem_sq_check(self.h, int(start), float(dt), _dp(v) if v is not None else None,
                                          _dp(m) if m is not None else None, _dp(hist), _ip(contact), _dp(en),
                                          _dp(dn), _dp(cpi), _dp(cpj)))
        out = self.results()
        out.update(hist=hist, contact=contact, en=en, deltan=dn, cp_i=cpi, cp_j=cpj)
        return out

    def results(self):
        n = self.n
        pt = np.zeros((n, 3))
        fi = np.zeros(n)
        fj = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        fl = np.zeros(n, dtype=np.int32)
        sm = ctypes.c_double()
        nm = ctypes.c_double()
        self._check(self.lib.dem_sq_get_results(self.h, _dp(pt), _dp(fi), _dp(fj), _ip(it), _ip(fl),
                                                ctypes.byref(sm), ctypes.byref(nm)))
        return {"point": pt, "fi": fi, "fj": fj, "iters": it, "flags": fl, "solve_ms": sm.value,
                "net_ms": nm.value}

    def features(self):
        X = np.zeros((self.n, 18), dtype=np.float32)
        self._check(self.lib.dem_sq_get_features(self.h, _fp(X)))
        return X

    def net_init(self, seed=0):
        self._check(self.lib.dem_sq_net_init(self.h, int(seed)))

    def net_params(self):
        p = np.zeros(SQ_NUM_PARAMS, dtype=np.float32)
        self._check(self.lib.dem_sq_net_get_params(self.h, _fp(p)))
        return p

    def set_net_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.lib.dem_sq_net_set_params(self.h, _fp(p)))

    def net_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 3), dtype=np.float32)
        self._check(self.lib.dem_sq_net_forward(self.h, X.shape[0], _fp(X), _fp(Y)))
        return Y

    def net_gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(SQ_NUM_PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.lib.dem_sq_net_gradient(self.h, X.shape[0], _fp(X), _fp(T), _fp(g), _fp(loss)))
        return g, float(loss[0])

    def net_gradient_noalloc(self, n, X_ptr, T_ptr, grad_ptr, loss_ptr):
        """dem_sq_net_gradient_noalloc on device pointers (integers): X [n][18],
        T [n][3] in, grad [5571] and loss [1] (or 0) out. Returns the status
        code (raises only for a missing library)."""
        return self.lib.dem_sq_net_gradient_noalloc(self.h, int(n), X_ptr, T_ptr, grad_ptr, loss_ptr or None)

    def net_train(self, steps, batch, lr, seed=0):
        curve = np.zeros(steps, dtype=np.float32)
        ms = ctypes.c_double()
        self._check(self.lib.dem_sq_net_train(self.h, int(steps), int(batch), float(lr), int(seed),
                                              _fp(curve), ctypes.byref(ms)))
        return curve, ms.value

    def close(self):
        if self.h:
            self.lib.dem_sq_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SqDemSystem:
    """Superquadric DEM step on the GPU (dem_sqdem_* of the C ABI).
    params / materials / particles / walls as workloads.superquadric_settling
    returns them."""

    def __init__(self, params, materials, P, walls=(), device=0):
        self.lib = load_library()
        p = SqDemParams()
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
                p.gravity[d] = g[d]
        p.gravity_first = int(params.get("gravity_first", 1))
        p.nplanes = len(walls)
        for k, (axis, pos, wtype) in enumerate(walls):
            p.plane_axis[k] = int(axis)
            p.plane_pos[k] = float(pos)
            p.plane_type[k] = int(wtype)
        self.h = ctypes.c_void_p()
        rc = self.lib.dem_sqdem_create(device, ctypes.byref(p), ctypes.byref(self.h))
        if rc != 0:
            raise DemError(f"dem_sqdem_create failed ({rc}): {self.lib.dem_sqdem_last_error(None).decode()}")
        self.nplanes = len(walls)
        nt, Yp, nup, ep, mup, _ = _material_args(materials)
        self._check(self.lib.dem_sqdem_set_materials(self.h, int(nt), Yp, nup, ep, mup))
        arr = lambda k, shape: np.ascontiguousarray(P[k], dtype=np.float64).reshape(shape)
        self.n = len(P["x"])
        n = self.n
        ty = np.ascontiguousarray(P["type"], dtype=np.int32)
        self._check(self.lib.dem_sqdem_set_particles(
            self.h, n, _dp(arr("x", (n, 3))), _dp(arr("v", (n, 3))), _dp(arr("omega", (n, 3))),
            _dp(arr("quat", (n, 4))), _dp(arr("shape", (n, 3))), _dp(arr("blockiness", (n, 2))),
            _dp(arr("density", (n,))), _ip(ty)))

    def _check(self, rc):
        if rc != 0:
            raise DemError(f"libdem_amd error {rc}: {self.lib.dem_sqdem_last_error(self.h).decode()}")

    def set_warm_start(self, solver):
        self._check(self.lib.dem_sqdem_set_warm_start(self.h, solver.h if solver is not None else None))

    def setup(self):
        self._check(self.lib.dem_sqdem_setup(self.h))

    def run(self, nsteps):
        self._check(self.lib.dem_sqdem_run(self.h, int(nsteps)))

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "omega", "angmom", "f", "torque")}
        out["quat"] = np.zeros((n, 4))
        self._check(self.lib.dem_sqdem_get_state(self.h, _dp(out["x"]), _dp(out["v"]), _dp(out["omega"]),
                                                 _dp(out["angmom"]), _dp(out["quat"]), _dp(out["f"]),
                                                 _dp(out["torque"])))
        return out

    def props(self):
        n = self.n
        m, rv, rb, I = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros((n, 3))
        self._check(self.lib.dem_sqdem_get_props(self.h, _dp(m), _dp(rv), _dp(rb), _dp(I)))
        return {"mass": m, "rvol": rv, "rbound": rb, "inertia": I}

    def pairs(self):
        need = -self.lib.dem_sqdem_get_pairs(self.h, 0, None, None, None)
        if need < 0:
            raise DemError(self.lib.dem_sqdem_last_error(self.h).decode())
        cap = max(int(need), 1)
        ij = np.zeros((cap, 2), dtype=np.int32)
        fl = np.zeros(cap, dtype=np.int32)
        hist = np.zeros((cap, 10))
        m = self.lib.dem_sqdem_get_pairs(self.h, cap, _ip(ij), _ip(fl), _dp(hist))
        if m < 0:
            raise DemError(self.lib.dem_sqdem_last_error(self.h).decode())
        return ij[:m], fl[:m], hist[:m]

    def wall_hist(self):
        h = np.zeros((self.n, max(self.nplanes, 1), 3))
        self._check(self.lib.dem_sqdem_get_wall_history(self.h, _dp(h)))
        return h[:, :self.nplanes]

    def stats(self):
        st = SqDemStats()
        self._check(self.lib.dem_sqdem_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in SqDemStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.dem_sqdem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def restart_particles(path):
    """(header dict, particle dict in the writer's caller order) of a restart file"""
    lib = load_library()
    n, nt, ts = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int64()
    lo, hi = np.zeros(3), np.zeros(3)
    per = np.zeros(3, dtype=np.int32)
    b = str(path).encode()
    if lib.dem_restart_info(b, ctypes.byref(n), ctypes.byref(nt), _dp(lo), _dp(hi), _ip(per), ctypes.byref(ts)):
        raise DemError(lib.dem_last_error(None).decode())
    N = n.value
    P = {"tag": np.zeros(N, dtype=np.int32), "type": np.zeros(N, dtype=np.int32), "x": np.zeros((N, 3)),
         "v": np.zeros((N, 3)), "omega": np.zeros((N, 3)), "radius": np.zeros(N), "rmass": np.zeros(N)}
    if lib.dem_restart_particles(b, N, _ip(P["tag"]), _ip(P["type"]), _dp(P["x"]), _dp(P["v"]),
                                 _dp(P["omega"]), _dp(P["radius"]), _dp(P["rmass"])):
        raise DemError(lib.dem_last_error(None).decode())
    return {"n": N, "ntypes": nt.value, "boxlo": lo, "boxhi": hi, "periodic": per, "ntimestep": ts.value}, P


class Inserter:
    """fix insert/pack + particletemplate/sphere + particledistribution/discrete
    (dem_inserter_*)."""

    def __init__(self, **kw):
        self.lib = load_library()
        self.p = insert_params(**kw)
        self.h = ctypes.c_void_p()
        rc = self.lib.dem_inserter_create(ctypes.byref(self.p), ctypes.byref(self.h))
        if rc != 0:
            raise DemError(f"dem_inserter_create failed ({rc}): {self.lib.dem_inserter_last_error(None).decode()}")

    def info(self):
        v = [ctypes.c_double() for _ in range(4)]
        self.lib.dem_inserter_info(self.h, *[ctypes.byref(a) for a in v])
        return dict(zip(("region_volume", "volume_expect", "mass_expect", "max_radius"), (a.value for a in v)))

    def step(self, x=None, radius=None, rmass=None, box=None, cap=200000):
        """One insertion; returns (dict of the inserted particles, requested)."""
        x = np.zeros((0, 3)) if x is None else np.ascontiguousarray(x, dtype=np.float64)
        radius = np.zeros(0) if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
        rmass = np.zeros(0) if rmass is None else np.ascontiguousarray(rmass, dtype=np.float64)
        blo = np.ascontiguousarray(box[0] if box else self.p.boxlo, dtype=np.float64)
        bhi = np.ascontiguousarray(box[1] if box else self.p.boxhi, dtype=np.float64)
        out = {k: np.zeros((cap, 3)) for k in ("x", "v", "omega")}
        out.update({k: np.zeros(cap) for k in ("radius", "density", "rmass")})
        out["type"] = np.zeros(cap, dtype=np.int32)
        nins, nreq = ctypes.c_int64(), ctypes.c_int64()
        rc = self.lib.dem_inserter_step(self.h, len(x), _dp(x), _dp(radius), _dp(rmass), _dp(blo), _dp(bhi),
                                        cap, _dp(out["x"]), _dp(out["v"]), _dp(out["omega"]),
                                        _dp(out["radius"]), _dp(out["density"]), _dp(out["rmass"]),
                                        _ip(out["type"]), ctypes.byref(nins), ctypes.byref(nreq))
        if rc != 0:
            raise DemError(f"dem_inserter_step failed ({rc}): {self.lib.dem_inserter_last_error(self.h).decode()}")
        k = nins.value
        return {kk: vv[:k] for kk, vv in out.items()}, nreq.value

    def close(self):
        if self.h:
            self.lib.dem_inserter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
