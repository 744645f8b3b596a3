/*
 * liggghts_amd.h -- C ABI of the MI355X-native DEM step library (libdem_amd.so).
 *
 * One handle owns one granular system resident in GPU memory (SoA arrays,
 * neighbor list, contact histories) and one HIP stream. Calls on one handle
 * must come from one host thread at a time. Every entry point returns a
 * dem_status (0 = success, negative = error) instead of exiting the process;
 * dem_last_error() gives the message. Arrays passed in or out are host
 * pointers in the caller's particle order (the order of dem_set_particles).
 *
 * Each entry point replaces a piece of LIGGGHTS-PUBLIC 3.8.0's plugin seam
 * for the per-timestep path; the reference interface is cited per function.
 */
#ifndef LIGGGHTS_AMD_H
#define LIGGGHTS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dem_handle dem_handle;

enum dem_status {
  DEM_OK = 0,
  DEM_ERR_ARG = -1,      /* invalid argument (error->all in the reference) */
  DEM_ERR_HIP = -2,      /* a HIP runtime call failed */
  DEM_ERR_OVERFLOW = -3, /* neighbor list overflow ("boost neigh_modify one") */
  DEM_ERR_STATE = -4,    /* call out of order (e.g. run before particles) */
  DEM_ERR_PHYSICS = -5,  /* non-finite state or negative overlap */
  DEM_ERR_NODEV = -6     /* no usable gfx950 device */
};

enum dem_precision {
  DEM_FP64 = 0,  /* every quantity in FP64 */
  DEM_MIXED = 1  /* FP64 x, v, x_i-x_j and overlap; FP32 contact law, history, omega */
};

/* Run parameters gathered from the input script:
 *   timestep                          -> dt            (update.cpp)
 *   neighbor <skin> bin               -> skin          (neighbor.cpp:1903-1930)
 *   neigh_modify contact_distance_factor -> cdf
 *   neigh_modify one <n>              -> max_neighbors (per-particle list width)
 *   region/create_box + boundary p p f -> boxlo, boxhi, periodic (domain.cpp)
 *   pair_style gran model hertz tangential history [rolling_friction cdt]
 *     [tangential_damping on|off] [limitForce on|off] [torsionTorque on|off]
 *                                     -> rolling_cdt, tangential_damping,
 *                                        limit_force, torsion_torque
 *                                        (pair_gran_proxy.cpp:63-79, settings.h:237-242)
 *   fix gravity <g> vector ax ay az   -> grav_acc = g * a/|a| (fix_gravity.cpp:361-398)
 */
typedef struct dem_params {
  double dt;
  double skin;
  double cdf;
  double boxlo[3];
  double boxhi[3];
  int periodic[3];
  int rolling_cdt;
  int tangential_damping;
  int limit_force;
  int torsion_torque;
  int has_gravity;
  double grav_acc[3];
  int gravity_first;  /* 1: fix gravity defined before fix wall/gran (post_force order) */
  int precision;      /* enum dem_precision */
  int max_neighbors;  /* 0 selects 32 */
  int device;         /* HIP device ordinal */
  int neigh_every;    /* neigh_modify every N (neighbor.cpp:1907-1911); 0 selects 1 */
  int neigh_no_check; /* neigh_modify check no (neighbor.cpp:1922-1927): build every
                         `neigh_every` steps without the skin/2 displacement test */
  /* boundary s / m faces of the non-periodic dimensions: bit 2*d the lower,
   * bit 2*d+1 the upper face of dimension d (shrink_min: those that are m).
   * A shrink-wrapped face follows the atoms at every neighbor build
   * (Domain::reset_box, domain.cpp:381-460) and loses none. A non-periodic
   * face without its bit is f: an atom beyond it at a neighbor build is lost
   * (Comm::exchange, comm.cpp:945-1003), and the step returns DEM_ERR_STATE
   * "Lost atoms: original N current M" (Thermo::lost_check, thermo.cpp:
   * 410-452, with thermo_modify lost error, the default). */
  int shrink_faces;
  int shrink_min;
} dem_params;

enum dem_normal_model {
  DEM_NORMAL_HERTZ = 0,           /* model hertz (normal_model_hertz.h) */
  DEM_NORMAL_HOOKE = 1,           /* model hooke (normal_model_hooke.h), viscous off */
  DEM_NORMAL_HERTZ_STIFFNESS = 2, /* model hertz/stiffness (normal_model_hertz_stiffness.h) */
  DEM_NORMAL_HOOKE_STIFFNESS = 3  /* model hooke/stiffness (normal_model_hooke_stiffness.h) */
};

/* Switches of one contact model (pair_style gran / fix wall/gran keywords):
 * model <normal> tangential history|no_history [cohesion sjkr]
 * [rolling_friction cdt|epsd2] and the on/off settings the models register
 * (Settings::registerOnOff in each model's registerSettings). Zero-filled
 * fields beyond the first four select hertz, history, no cohesion. */
typedef struct dem_contact_model {
  int rolling_cdt;        /* rolling_friction cdt */
  int tangential_damping; /* tangential_damping on (default) */
  int limit_force;        /* limitForce on */
  int torsion_torque;     /* torsionTorque on */
  int normal_model;       /* enum dem_normal_model */
  int tangential_model;   /* 0 history (tangential_model_history.h), 1 no_history
                             (tangential_model_no_history.h) */
  int cohesion_model;     /* 0 off, 1 sjkr (cohesion_model_sjkr.h) */
  int kt_to_kn;           /* hooke: ktToKnUser on (normal_model_hooke.h:100) */
  int absolute_damping;   /* hooke/stiffness: absolute_damping on (normal_model_hooke_stiffness.h:81) */
  int tangential_reduce;  /* sjkr: tangential_reduce on (cohesion_model_sjkr.h:72) */
  int rolling_epsd2;      /* rolling_friction epsd2 (rolling_model_epsd2.h): an elastic
                             rolling spring whose torque is kept as contact history;
                             pair and plane-wall contacts with tangential history */
} dem_contact_model;

typedef struct dem_stats {
  int64_t ntimestep;       /* steps run since creation (update->ntimestep) */
  int64_t nrebuilds;       /* neighbor list builds, setup builds included */
  int64_t npairs;          /* half-list pairs at the last build */
  int64_t ncontacts;       /* touching half-list pairs at the last force pass */
  int64_t nwall_contacts;  /* touching particle-wall contacts at the last force pass */
  int max_neighbors_seen;  /* widest per-particle list at the last build */
  double step_kernel_ms;   /* average time per fused step-kernel launch (HIP events) */
  int64_t step_kernel_launches;
  double rebuild_ms_total; /* wall time inside neighbor rebuilds */
  int64_t nmesh_candidates; /* (particle, triangle) candidate entries at the last build */
} dem_stats;

/* Creates a system. Replaces the construction done by input.cpp:560-650 for
 * atom_style granular + pair_style gran + neighbor + fix nve/sphere. */
int dem_create(const dem_params *params, dem_handle **out);

/* Frees every device allocation of the handle. */
int dem_destroy(dem_handle *h);

/* Message of the last failed call on this handle (or of dem_create when h is NULL). */
const char *dem_last_error(const dem_handle *h);

/* Material properties of fix property/global youngsModulus / poissonsRatio
 * (peratomtype, ntypes values) and coefficientRestitution / coefficientFriction /
 * coefficientRollingFriction (peratomtypepair, ntypes*ntypes row-major, mur may
 * be NULL; Y, nu and e may be NULL together when no contact model reads them,
 * as with the stiffness models). Mixed on the host into Y*, G*, beta per type pair exactly as
 * MODEL_PARAMS::createYeff/createGeff/createBetaEff (global_properties.cpp:411-558),
 * with the reference's range checks (5e6 <= Y <= 1e9, 0 <= nu <= 0.5,
 * 0.05 < e <= 1). Replaces PropertyRegistry::connect of Yeff/Geff/betaeff/
 * coeffFrict/coeffRollFrict (normal_model_hertz.h:146-160). */
int dem_set_materials(dem_handle *h, int ntypes, const double *youngs, const double *poisson,
                      const double *restitution, const double *friction,
                      const double *rolling_friction);

/* The pair style's contact model (pair_style gran model ...: the model
 * selection of PairGranProxy::settings, pair_gran_proxy.cpp:63-79, and the
 * ContactModel<> chain of contact_models.h:137-260), replacing the four
 * switches of dem_params.
 * Walls and meshes added afterwards without their own model take this one.
 * Call before dem_set_particles. With a model other than hertz + history
 * the step kernel runs its general contact law in tile shape 0. */
int dem_set_pair_model(dem_handle *h, const dem_contact_model *model);

/* Coefficients of the sibling models, after dem_set_materials (which clears
 * them): fix property/global kn / kt / gamman (or gamman_abs) / gammat (or
 * gammat_abs) peratomtypepair (hertz/stiffness, hooke/stiffness:
 * MODEL_PARAMS::createKn/Kt/Gamman/Gammat/GammanAbs/GammatAbs,
 * global_properties.cpp:735-812), cohesionEnergyDensity peratomtypepair (sjkr,
 * :291-294) and characteristicVelocity scalar (hooke, :227-240; 0 if unused).
 * Arrays are ntypes*ntypes row-major; kn, kt, gamman, gammat come together or
 * not at all; cohesion_energy_density may be NULL. With dem_set_instances,
 * call before dem_set_instance_materials (every instance gets these values). */
int dem_set_model_coefficients(dem_handle *h, int ntypes, const double *kn, const double *kt,
                               const double *gamman, const double *gammat,
                               const double *cohesion_energy_density, double characteristic_velocity);

/* Uploads n particles (read_data Atoms + Velocities, atom_vec_sphere.cpp:1055-1140).
 * x, v, omega are n*3 row-major; rmass is the mass of each sphere; type is 1-based;
 * tag is the unique positive atom ID. Replaces Atom::x/v/omega/radius/rmass/type/tag. */
int dem_set_particles(dem_handle *h, int64_t n, const double *x, const double *v,
                      const double *omega, const double *radius, const double *rmass,
                      const int *type, const int *tag);

/* Adds a planar primitive wall: fix wall/gran model hertz tangential history
 * [rolling_friction cdt] ... primitive type <t> {x,y,z}plane <pos>
 * (fix_wall_gran.cpp:168-227, primitive_wall_definitions.h:130-150). axis 0/1/2.
 * model holds the fix's own contact-model switches (each fix wall/gran selects
 * its model independently of pair_style); NULL means the pair style's. */
int dem_add_plane_wall(dem_handle *h, int axis, double position, int wall_type,
                       const dem_contact_model *model);

/* Adds a triangle mesh wall: fix mesh/surface file <stl> type <wall_type>
 * (fix_mesh.cpp:75-271, input_mesh_tri.cpp:318) used by fix wall/gran model hertz
 * tangential history ... mesh n_meshes k meshes <ids> (fix_wall_gran.cpp:228-310).
 * nodes holds ntri*9 coordinates (three nodes per triangle, as listed in the STL).
 * Element geometry and topology (edge/corner activity, coplanar neighbors) are
 * computed here as SurfaceMesh::buildNeighbours does (surface_mesh_I.h:454-560,
 * 1035-1190). model: the wall fix's contact-model switches (NULL: the pair
 * style's). *mesh_id receives the mesh index. Call before dem_set_particles. */
int dem_add_mesh(dem_handle *h, int64_t ntri, const double *nodes, int wall_type,
                 const dem_contact_model *model, int *mesh_id);

/* The triangles of fix mesh/surface ... generate <type> (InputMeshTri::
 * meshgenerator, input_mesh_tri.cpp:608-1107), centred on the origin, in the
 * generator's order: type "cube" (dims[0] = size), "box" (x, y, z sizes),
 * "cylinder" (radius, height), "pipe" (inner radius, outer radius, height),
 * "disk" (radius), "plane" (size); nsegments for cylinder / pipe / disk; mask
 * the faces (cube / box: top 1, bottom 2, front 4, back 8, left 16, right 32;
 * cylinder: top 1, bottom 2, side 4; pipe: inner top 1, outer top 2, inner
 * bottom 4, outer bottom 8, inner side 16, outer side 32). Writes up to cap
 * triangles (9 doubles each) to nodes (may be NULL) and their number to *ntri.
 * Host only; errors through dem_last_error(NULL). */
int dem_mesh_generate(const char *type, const double *dims, int nsegments, int mask, int64_t cap, double *nodes,
                      int64_t *ntri);

/* fix move/mesh mesh <id> rotate origin ox oy oz axis ax ay az period T
 * (mesh_mover_rotation.cpp:66-139): the mesh turns by 2*pi*dt/T about the axis
 * each step and its nodes carry the rigid-body velocity. */
int dem_mesh_rotate(dem_handle *h, int mesh_id, const double *origin, const double *axis,
                    double period);

/* fix mesh/surface/stress ... [reference_point x y z] (MeshModuleStress,
 * mesh_module_stress.cpp:51-120): the mesh accumulates, at every counted
 * force pass (the last step of each dem_run / dem_run_continue, dem_setup,
 * dem_compute_forces), the total force the particles exert on it and its
 * torque about reference_point (NULL: the origin), as
 * add_particle_contribution does (:276-300): minus each contact's force on
 * the particle, lever arm from the reference point to the contact point. */
int dem_mesh_stress(dem_handle *h, int mesh_id, const double *reference_point);

/* write_restart <file> (write_restart.cpp with FixContactHistory::
 * write_restart / pack_restart, fix_contact_history.cpp:588-659): the
 * particles, their contact histories (pair, wall and mesh), the mesh
 * geometry and the step counters, in this library's own binary layout. */
int dem_write_restart(dem_handle *h, const char *path);

/* read_restart <file> (read_restart.cpp): the box, particle count, atom
 * types and timestep of a restart file, and its particles in the caller order
 * they were given in (arrays n, n, 3n, 3n, 3n, n, n; any may be NULL), so a
 * caller can rebuild the system with dem_set_particles before
 * dem_read_restart. These two need no handle (errors: dem_last_error(NULL)). */
int dem_restart_info(const char *path, int64_t *n, int *ntypes, double *boxlo, double *boxhi,
                     int *periodic, int64_t *ntimestep);
int dem_restart_particles(const char *path, int64_t n, int *tag, int *type, double *x, double *v,
                          double *omega, double *radius, double *rmass);

/* Restores the device state of a restart file into a handle configured as
 * the one that wrote it (same particles set in the same order, materials,
 * walls, meshes, precision, list width, timestep). The next dem_run sets up
 * and continues bitwise as the writing system's next run would. */
int dem_read_restart(dem_handle *h, const char *path);

/* The total force and torque on a mesh/surface/stress mesh at the last
 * counted pass: what FixMeshSurface::compute_vector (fix_mesh_surface.cpp:560)
 * returns as f_ID[1..6] for thermo output. Either pointer may be NULL. */
int dem_get_mesh_force(dem_handle *h, int mesh_id, double *force, double *torque);

/* fix move/mesh mesh <id> linear vx vy vz (MeshMoverLinear,
 * mesh_mover_linear.cpp:62-101): each step the mesh is shifted by v dt
 * (MultiNodeMesh::move, multi_node_mesh_I.h:502-527) and its nodes carry the
 * velocity v. One fix move/mesh per mesh. Call before dem_set_particles. */
int dem_mesh_linear(dem_handle *h, int mesh_id, const double *velocity);

/* fix move/mesh mesh <id> wiggle amplitude ax ay az period T (MeshMoverWiggle,
 * mesh_mover_linear.cpp:160-218): with omega = 2 pi / T and the fix's time t
 * (t += dt at every step over the fix's lifetime, fix_move_mesh.cpp:226) the
 * nodes move with velocity A omega cos(omega t) and are shifted by that
 * velocity times dt. One fix move/mesh per mesh. Call before dem_set_particles. */
int dem_mesh_wiggle(dem_handle *h, int mesh_id, const double *amplitude, double period);

/* fix move/mesh mesh <id> viblin axis ax ay az order n amplitude a1..an
 * phase p1..pn period T1..Tn (MeshMoverVibLin, mesh_mover_linear.cpp:230-305):
 * velocity axis * sum_j -a_j w_j sin(w_j t + p_j), w_j = 2 pi / T_j, applied
 * as for wiggle. Arrays hold `order` (1..30) values. */
int dem_mesh_viblin(dem_handle *h, int mesh_id, const double *axis, int order, const double *amplitude,
                    const double *phase, const double *period);

/* fix move/mesh mesh <id> vibrot origin ox oy oz axis ax ay az order n
 * amplitude a1..an phase p1..pn period T1..Tn (MeshMoverVibRot,
 * mesh_mover_rotation.cpp:333-430): rotation about the axis through origin
 * with angular velocity sum_j -a_j w_j sin(w_j t + p_j). */
int dem_mesh_vibrot(dem_handle *h, int mesh_id, const double *origin, const double *axis, int order,
                    const double *amplitude, const double *phase, const double *period);

/* Current mesh nodes and node velocities (ntri*9 each; either may be NULL). */
int dem_get_mesh_nodes(dem_handle *h, int mesh_id, double *nodes, double *vnodes);

/* Mesh topology: edge_active, corner_active (ntri*3, 0/1) and the obtuse-angle
 * node per triangle (-1 for none), as SurfaceMesh computes them. */
int dem_get_mesh_topology(dem_handle *h, int mesh_id, int *edge_active, int *corner_active,
                          int *obtuse);

/* Mesh contact history (FixContactHistoryMesh, fix_contact_history_mesh.cpp):
 * per particle in caller order, maxslot slots of (triangle index within the
 * mesh, or -1 for an empty slot) and the shear history (3 values per slot). */
int dem_get_mesh_history(dem_handle *h, int mesh_id, int maxslot, int *tri, double *hist);

/* Verlet::setup (verlet.cpp:134-205): carries contact history out, wraps periodic
 * coordinates, reorders particles by Morton cell key, rebuilds the neighbor list
 * with history carry-in, and computes forces with shearupdate = 0. */
int dem_setup(dem_handle *h);

/* Verlet::run (verlet.cpp:264-392) for nsteps timesteps: nve/sphere initial
 * integrate, Neighbor::decide (every step, delay 0), rebuild if any particle
 * moved more than skin/2, pair force (PairGran::compute, pair_gran.cpp:641-670),
 * post_force fixes (gravity, wall/gran) and nve/sphere final integrate.
 * Runs dem_setup first when the state changed since the last setup. */
int dem_run(dem_handle *h, int64_t nsteps);

/* `run N pre no` (run.cpp:116-119, :209): continues integration from the last
 * force pass without a new setup (no forced rebuild, no shearupdate-0 pass), so
 * dem_run(a) followed by dem_run_continue(b) equals one run of a+b steps.
 * Needs a preceding dem_run, dem_run_continue or dem_setup on this state. */
int dem_run_continue(dem_handle *h, int64_t nsteps);

/* One force evaluation on the current state with the given shearupdate flag,
 * no integration: Pair::compute + post_force of the wall and gravity fixes.
 * f and torque become readable with dem_get_state. Histories and contact flags
 * are updated exactly as in a step. */
int dem_compute_forces(dem_handle *h, int shearupdate);

/* Copies the state back, in caller order; any pointer may be NULL. f and torque
 * are those of the last force pass (setup, step or dem_compute_forces). */
/* The simulation box at the last neighbor build (boxlo, boxhi, 3 each): the
 * created box in periodic and f dimensions; a shrink-wrapped face at the
 * extent of the atoms there plus 1e-4 of the created box's length, an m face
 * never inside its created position (Domain::reset_box, domain.cpp:432-449). */
int dem_get_box(dem_handle *h, double *boxlo, double *boxhi);

/* Per-particle state in the caller's order ([n][3] each; f and torque of the
 * last force pass). A system without particles yet (dem_insert_particles
 * pending) has nothing to copy and returns DEM_OK. */
int dem_get_state(dem_handle *h, double *x, double *v, double *omega, double *f,
                  double *torque);

/* Pairs of the current half neighbor list (NeighList + listgranhistory,
 * neigh_gran.cpp:485-639): tags[2k] < tags[2k+1] are atom tags, flags[k] the
 * contact flag, hist[3k..3k+2] the tangential history oriented from the lower
 * tag to the higher. *npairs receives the number of pairs; min(cap, *npairs)
 * are written. */
int dem_get_pairs(dem_handle *h, int64_t cap, int *tags, int *flags, double *hist,
                  int64_t *npairs);

/* Primitive-wall shear history, n*nwalls*3 values, wall-major inside each particle
 * (FixWallGran's property/atom history, fix_wall_gran.cpp:498-522). */
int dem_get_wall_history(dem_handle *h, double *hist);

/* compute pair/gran/local (compute_pair_gran_local.cpp:81-110, 312-470) as
 * evaluated through PairGran::compute_pgl (pair_gran.cpp:676-708): one pass
 * over the current half neighbor list with shearupdate off, on the GPU,
 * changing nothing. Every half-list pair gets a row (cpl_add_pair is called
 * for each pair of the list, pair_gran_base.h:446-447): tags[2k] < tags[2k+1],
 * and data[k * ncol ...] = x_i[3], x_j[3] (remapped into the periodic box),
 * v_i[3], v_j[3], the force and the torque the contact exerts on particle i
 * (the lower tag; zero when the pair does not touch), 1 if the pair crosses a
 * periodic boundary else 0, r_i, r_j, then the pair's stored history oriented
 * from i to j (3 shear values, and 3 rolling-torque values with epsd2). *ncol
 * receives 21 + the history width. Rows of one particle are consecutive; *npairs
 * receives the number of pairs, min(cap, *npairs) rows are written (cap 0:
 * count only). Deviation: the reference orients a pair by its local atom
 * indices, this path by tags. */
#define DEM_PAIR_LOCAL_FIXED_COLS 21
int dem_compute_pair_local(dem_handle *h, int64_t cap, int *tags, double *data, int64_t *npairs,
                           int *ncol);

/* compute wall/gran/local (compute_pair_gran_local.cpp:143-144, 247-270,
 * 648-870) as evaluated through FixWallGran::post_force_pgl
 * (fix_wall_gran.cpp:752-765): one pass over the mesh contacts with
 * shearupdate off, on the GPU, changing nothing. A row per contact that the
 * mesh walls' contact pass evaluates (within contact, or within the close band
 * of contact_distance_factor, and accepted by the contact-history bookkeeping;
 * fix_wall_gran.cpp:955-1006), mesh by mesh, the rows of one particle
 * consecutive: ids[3k..3k+2] = the mesh (in dem_add_mesh order), the
 * triangle's index within its mesh (the reference's element id), the
 * particle's tag; data[k * ncol ...] = the contact point x + delta[3], x[3],
 * the wall's velocity at the contact[3] (zero for a resting mesh), v[3], the
 * force and the torque on the particle[3 + 3] (zero in the close band), the
 * unit normal from the particle's centre to the contact point[3] (minus the
 * contact's en, fix_wall_gran_base.h:332-334), the squared distance from the
 * centre to the contact point (sidata.rsq), the particle's radius, then the
 * slot's shear history[3]. *ncol receives 26; *nrows the number of rows,
 * min(cap, *nrows) are written (cap 0: count only). Primitive walls give no
 * rows, as in the reference. */
#define DEM_WALL_LOCAL_FIXED_COLS 23
int dem_compute_wall_local(dem_handle *h, int64_t cap, int *ids, double *data, int64_t *nrows, int *ncol);

/* The rolling-torque history of rolling_friction epsd2 (the r_torque{x,y,z}_old
 * values RollingModel<EPSD2> registers, rolling_model_epsd2.h:75-77), for the
 * pairs of dem_get_pairs in the same order and orientation (rhist[3k..3k+2]),
 * and for the primitive walls as dem_get_wall_history. DEM_ERR_STATE when no
 * contact model uses epsd2. */
int dem_get_pair_rolling(dem_handle *h, int64_t cap, double *rhist, int64_t *npairs);
int dem_get_wall_rolling(dem_handle *h, double *rhist);

/* Counters and timing of the handle. */
int dem_get_stats(dem_handle *h, dem_stats *out);

/* Brackets each batch of step-kernel launches of dem_run / dem_run_continue
 * with a pair of HIP events on the handle's stream and accumulates the
 * per-launch average into dem_stats (on != 0), or stops doing so; batches
 * that end in a rebuild request are not counted. Resets the accumulated
 * timing either way. No reference counterpart:
 * the reference's Timer buckets (timer.h:51-52) play this role on the CPU. */
int dem_set_kernel_timing(dem_handle *h, int on);

/* ---- Independent simulations batched in one handle ----
 * No counterpart in the reference, which runs one simulation per process (a
 * calibration sweep there is many lmp_serial runs, each through the entry
 * points above). Instance q owns particles [c_0 + ... + c_{q-1},
 * c_0 + ... + c_q) of the arrays dem_set_particles takes, in the same box and
 * local coordinates as every other instance, with its own copy of the cell
 * grid and of every wall and mesh, and its own material block (initially the
 * dem_set_materials one). Instances never interact; one step of the handle
 * advances all of them, and a neighbor rebuild requested by any instance
 * rebuilds all. Call before dem_set_particles; ninst <= 65536. */
int dem_set_instances(dem_handle *h, int ninst, const int64_t *counts);

/* Materials of one instance, as dem_set_materials (same ntypes). */
int dem_set_instance_materials(dem_handle *h, int inst, int ntypes, const double *Y,
                               const double *nu, const double *e, const double *mu,
                               const double *mur);

/* Per instance (first ninst of them): touching particle pairs and touching
 * wall (plane or mesh) contacts at the last step of the last run. */
int dem_get_instance_stats(dem_handle *h, int ninst, int64_t *ncontacts, int64_t *nwall);

/* Tile shape of the fused step kernel: variant 0..10 (threads per block /
 * neighbor entries loaded together: 0-4 per-lane pair phase 256/4, 256/2,
 * 256/8, 128/4, 64/4; 5-6 compacted group by group over the wavefront 256/4,
 * 256/8; 7-8 queued over whole list rows 256/4, 256/8; 9-10 per lane with the
 * histories loaded in the broad phase 256/4, 128/4) and a persistent grid of
 * `persist` blocks per compute unit (0: one block per tile). variant = -1
 * asks for autotuning: the next runs spend their first batches timing every
 * (variant, persist in {0, 4, 8}) candidate, 8 steps each, in three passes,
 * and keep the fastest. Every shape gives bitwise the same results.
 * No reference counterpart (the reference has one CPU loop). */
int dem_set_step_shape(dem_handle *h, int variant, int persist);

/* on != 0: dem_run / dem_run_continue replay batches of 32 steps as HIP
 * graphs (captured once per buffer parity and kernel shape, rebuilt after
 * any change of the particle set, materials or shape) instead of launching
 * each step; results are bitwise the same. Event timing (dem_set_kernel_timing)
 * and autotuning batches launch step by step. No reference counterpart. */
int dem_set_graph_capture(dem_handle *h, int on);

/* The shape in use and, after autotuning, each candidate's measured
 * microseconds per step (0 if it was not measured); cap bounds the arrays. */
int dem_get_step_shape(dem_handle *h, int *variant, int *persist, int cap, int *ncand,
                       int *cand_variant, int *cand_persist, double *cand_us);

/* ---- Superquadric contact points ----
 * The contact point of two superquadrics (atom_style superquadric: centre,
 * quaternion, half-axes, blockiness) is the point where their shape functions
 * are equal with antiparallel gradients. Poses are 12 doubles: centre[3],
 * quaternion (w, i, j, k)[4], half-axes a, b, c [3], blockiness n1, n2 [2]. */
typedef struct dem_sq_handle dem_sq_handle;

enum dem_sq_start {
  DEM_SQ_FRESH = 0,   /* a fresh contact: the homotopy from volume-equivalent spheres,
                         calc_contact_point_if_no_previous_point_avaialable
                         (math_extra_liggghts_superquadric.cpp:1390-1430) */
  DEM_SQ_GIVEN = 1,   /* one calc_contact_point solve from caller-given points
                         (math_extra_liggghts_superquadric.cpp:778-1022), the fresh
                         homotopy where it fails or misses the merit tolerance */
  DEM_SQ_NETWORK = 2  /* as GIVEN, from the warm-start network's predicted points */
};

int dem_sq_create(int device, dem_sq_handle **out);
int dem_sq_destroy(dem_sq_handle *h);
const char *dem_sq_last_error(const dem_sq_handle *h);

/* n pose pairs (A, B), n*12 doubles each; copied to the device. */
int dem_sq_set_pairs(dem_sq_handle *h, int64_t n, const double *pose_a, const double *pose_b);

/* Solves every pair on the GPU (FP64, one thread per pair). init: n*3 host
 * points for DEM_SQ_GIVEN, else NULL. Replaces calc_contact_point as called
 * from SurfaceModel<SURFACE_SUPERQUADRIC>::checkSurfaceIntersect
 * (surface_model_superquadric.h:108-205). polish > 0 adds up to that many
 * full Newton steps after the reference's stopping rule (each kept only if it
 * lowers the residual): no reference counterpart, for callers that need the
 * point below the reference's merit tolerance; 0 is the reference's solve. */
int dem_sq_solve(dem_sq_handle *h, int start, const double *init, int polish);

/* Contact detection for every pair: SurfaceModel<SURFACE_SUPERQUADRIC>::
 * checkSurfaceIntersect (surface_model_superquadric.h:108-205) -- the OBB
 * pre-check (obb_intersect, math_extra_liggghts_superquadric.cpp:166-203)
 * unless the pair intersected last step, the contact point (the homotopy for
 * a pair far apart last step, else calc_contact_point_using_prev_step,
 * :1434-1466), and for intersecting surfaces the normal en (from j to i) and
 * the overlap (extended_overlap_algorithm, :1489-1504).
 *   start   DEM_SQ_FRESH, or DEM_SQ_NETWORK: pairs far apart last step start
 *           from the warm-start network's point (the homotopy as fallback)
 *   dt      time step of the previous-step estimate
 *   vel     n*12 host doubles v_i, v_j, omega_i, omega_j, with mass (n*2 m_i,
 *           m_j); both NULL for pairs at rest (the estimate is the last point)
 *   hist    n*7 host doubles, in/out, the pair's contact history: state
 *           (0 far, 1 close, 2 intersecting), first OBB inequality to try,
 *           contact point[3], alpha_i, alpha_j; all zero for a new pair
 * Outputs (each may be NULL): contact (n ints), en (n*3), deltan (n), the
 * surface points cp_i, cp_j (n*3 each); contact points, Newton iterations and
 * flags (bit 2: the OBBs are apart, no solve) through dem_sq_get_results. */
int dem_sq_check(dem_sq_handle *h, int start, double dt, const double *vel, const double *mass, double *hist,
                 int *contact, double *en, double *deltan, double *cp_i, double *cp_j);

/* Results of the last solve (any pointer may be NULL): contact points (n*3),
 * shape-function values fi, fj there, Newton iterations per pair, flags
 * (bit 0: the solve failed, bit 1: the fresh fallback ran), and the time of
 * the solve and of the network inference before it (ms, HIP events). */
int dem_sq_get_results(dem_sq_handle *h, double *point, double *fi, double *fj, int *iters, int *flags,
                       double *solve_ms, double *net_ms);

/* The warm-start network: 18 inputs (B's centre in A's body frame and the
 * relative quaternion, both shapes relative to their volume-equivalent radii,
 * blockiness / 8, radius ratio) -> 64 -> 64 (tanh) -> 3 (the contact point in
 * A's body frame over the sum of the equivalent radii); BF16 matrix-core
 * products with FP32 accumulation, FP32 master weights, Adam. No counterpart
 * in the reference. Parameters: W1[18][64], b1[64], W2[64][64], b2[64],
 * W3[64][3], b3[3] (5571 floats). */
int dem_sq_get_features(dem_sq_handle *h, float *X);
int dem_sq_net_init(dem_sq_handle *h, uint64_t seed);
int dem_sq_net_set_params(dem_sq_handle *h, const float *params);
int dem_sq_net_get_params(dem_sq_handle *h, float *params);
int dem_sq_net_forward(dem_sq_handle *h, int64_t n, const float *X, float *Y);
/* Gradient of the mean squared error over n samples (no update). */
int dem_sq_net_gradient(dem_sq_handle *h, int64_t n, const float *X, const float *T, float *grad,
                        float *loss);
/* dem_sq_net_gradient on device memory, allocating nothing. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (tight
 * training or serving loops, harnesses that forbid hipMalloc). X [n][18] and
 * T [n][3] are device pointers on the handle's device; grad (5571 floats) and
 * loss (1 float, or NULL) receive the results in device memory; the call
 * returns once they are written. The per-block partial sums go to a fixed
 * workspace, one per device, so n must be 1 .. DEM_SQ_GRAD_NOALLOC_MAX
 * (DEM_ERR_ARG otherwise) and the network must be initialised (DEM_ERR_STATE
 * otherwise). Every call on a device shares that device's workspace: calls
 * that overlap on one device (several host threads, or handles on different
 * streams of the same device) race and must be serialised by the caller. */
#define DEM_SQ_GRAD_NOALLOC_MAX 16384
int dem_sq_net_gradient_noalloc(dem_sq_handle *h, int64_t n, const float *X, const float *T, float *grad,
                                float *loss);
/* `steps` Adam steps on batches drawn from the handle's pairs, the targets
 * being the points of the last DEM_SQ_FRESH solve; loss_curve (steps floats)
 * and train_ms may be NULL. */
int dem_sq_net_train(dem_sq_handle *h, int64_t steps, int64_t batch, float lr, uint64_t seed,
                     float *loss_curve, double *train_ms);

/* ------------------------------------------------------------------------
 * Superquadric DEM step (atom_style superquadric): FP64 on the GPU, one
 * handle per simulation. Replaces, for superquadric particles, the per-step
 * path of Verlet::setup / Verlet::run (verlet.cpp:134-205, 264-392) with
 *   pair_style gran model hertz tangential history surface superquadric
 *       (PairGran::compute, pair_gran_base.h:187-500, with
 *       SurfaceModel<SURFACE_SUPERQUADRIC>, surface_model_superquadric.h:108-220)
 *   fix wall/gran ... primitive type t {x,y,z}plane pos
 *       (FixWallGran::post_force_primitive, fix_wall_gran.cpp:1024-1153)
 *   fix gravity (fix_gravity.cpp:301-350)
 *   fix nve/superquadric (fix_nve_superquadric.cpp:45-75 over
 *       FixNVEAsphereBase, fix_nve_asphere_base.cpp:355-610, scheme 1)
 * and the granular neighbor list on the bounding spheres with contact
 * history (neigh_gran.cpp:485-639, fix_contact_history.cpp:305-425).
 * ------------------------------------------------------------------------ */
typedef struct dem_sqdem_handle dem_sqdem_handle;

typedef struct {
  double dt;                 /* timestep */
  double skin;               /* neighbor <skin> bin */
  double boxlo[3], boxhi[3]; /* region of the simulation box */
  int periodic[3];           /* boundary p (1) or f (0) per dimension */
  int tangential_damping;    /* pair_style gran ... tangential_damping on|off (on) */
  int limit_force;           /* limitForce on|off (off) */
  int has_gravity;           /* fix gravity present */
  double gravity[3];         /* magnitude times the unit direction */
  int gravity_first;         /* fix gravity defined before the walls */
  int nplanes;               /* fix wall/gran ... primitive planes (<= 6) */
  int plane_axis[6];         /* 0 xplane, 1 yplane, 2 zplane */
  double plane_pos[6];
  int plane_type[6];         /* wall atom type */
} dem_sqdem_params;

typedef struct {
  int64_t steps, rebuilds;
  int64_t pair_touch, wall_touch; /* touching pairs / particle-plane contacts, last step */
  int64_t newton_last;            /* Newton iterations, last step */
  int64_t newton, checks;         /* totals: Newton iterations, pair contact detections */
  int64_t fresh, fresh_iters;     /* totals: solves of pairs far apart the step before (OBBs
                                     overlapping), their Newton iterations */
  int64_t fallbacks;              /* totals: network starts replaced by the homotopy */
  double run_ms;                  /* GPU time of the last dem_sqdem_run (HIP events) */
} dem_sqdem_stats;

int dem_sqdem_create(int device, const dem_sqdem_params *p, dem_sqdem_handle **out);
int dem_sqdem_destroy(dem_sqdem_handle *h);
const char *dem_sqdem_last_error(const dem_sqdem_handle *h);

/* fix property/global youngsModulus / poissonsRatio (per type),
 * coefficientRestitution / coefficientFriction (ntypes*ntypes), mixed as
 * global_properties.cpp:411-558 */
int dem_sqdem_set_materials(dem_sqdem_handle *h, int ntypes, const double *Y, const double *nu, const double *e,
                            const double *mu);

/* The Atoms of read_data for atom_style superquadric
 * (AtomVecSuperquadric::data_atom, atom_vec_superquadric.cpp:1150-1212): x, v,
 * omega (n*3), quat (n*4, w first, unit), shape (half-axes, n*3), blockiness
 * (n*2, >= 2), density (n), type (n, 1-based). The library derives the
 * bounding radius, volume, mass and principal inertia as data_atom does. */
int dem_sqdem_set_particles(dem_sqdem_handle *h, int64_t n, const double *x, const double *v, const double *omega,
                            const double *quat, const double *shape, const double *blockiness,
                            const double *density, const int *type);

/* Warm start for pairs far apart the step before: start their Newton solve
 * from the contact point predicted by the trained network of `net` (same
 * device; it must outlive its use here and not train meanwhile), with the
 * homotopy of calc_contact_point_if_no_previous_point_avaialable
 * (math_extra_liggghts_superquadric.cpp:1390-1430) as fallback. NULL: the
 * reference's homotopy only. No reference counterpart. */
int dem_sqdem_set_warm_start(dem_sqdem_handle *h, dem_sq_handle *net);

/* Verlet::setup (verlet.cpp:134-205): neighbor build with history carried
 * over, forces without history update; what `run N` (pre yes) does first. */
int dem_sqdem_setup(dem_sqdem_handle *h);

/* Runs nsteps timesteps; the first call also runs the setup (forces without
 * history update), as `run` does; later calls continue (`run N pre no`). */
int dem_sqdem_run(dem_sqdem_handle *h, int nsteps);

/* Current state (each pointer may be NULL): x, v, omega, angmom, f, torque
 * (n*3), quat (n*4). */
int dem_sqdem_get_state(dem_sqdem_handle *h, double *x, double *v, double *omega, double *angmom, double *quat,
                        double *f, double *torque);
/* mass, volume-equivalent radius, bounding radius (n each), inertia (n*3) */
int dem_sqdem_get_props(dem_sqdem_handle *h, double *mass, double *rvol, double *rbound, double *inertia);
/* The pair list (i < j by index): ij (2 per pair), touch flags, histories
 * (10 doubles: state, first OBB inequality, contact point[3], alpha_i,
 * alpha_j, shear[3]). Returns the count, or -count when cap is too small. */
int64_t dem_sqdem_get_pairs(dem_sqdem_handle *h, int64_t cap, int *ij, int *flags, double *hist);
/* wall shear history, n*nplanes*3 */
int dem_sqdem_get_wall_history(dem_sqdem_handle *h, double *hist);
int dem_sqdem_get_stats(dem_sqdem_handle *h, dem_sqdem_stats *out);

/* ---- Particle insertion ----
 * Appends n particles at the next step boundary: they join at the first step
 * of the next dem_run / dem_run_continue, after that step's nve/sphere initial
 * integrate of the particles already there and before its neighbor build,
 * which is forced (FixInsert::pre_exchange, fix_insert.cpp:652-826, on a step
 * whose next_reneighbor the fix set; ParticleToInsert::insert,
 * particleToInsert.cpp:99-147, creates the atoms with no contact history).
 * Arrays as dem_set_particles; tags must be new. Works on a handle that has no
 * particles yet (dem_set_particles not called): it then starts with these.
 * Calls before one run accumulate. */
int dem_insert_particles(dem_handle *h, int64_t n, const double *x, const double *v,
                         const double *omega, const double *radius, const double *rmass,
                         const int *type, const int *tag);

/* The largest particle radius the cell grid must cover, beyond the particles
 * present: PairGran::init_one's "future particles as dynamic" (pair_gran.cpp:
 * 538-551) takes max_rad of every insertion fix into the neighbor cutoff. */
int dem_set_max_radius(dem_handle *h, double rmax);

/* fix insert/pack with fix particletemplate/sphere and
 * fix particledistribution/discrete on a block region: where the particles of
 * one insertion go (FixInsertPack, fix_insert_pack.cpp:74-500; FixInsert,
 * fix_insert.cpp:78-972; FixParticledistributionDiscrete::randomize_list,
 * fix_particledistribution_discrete.cpp:395-469; FixTemplateSphere,
 * fix_template_sphere.cpp:72-321; Region::generate_random*, volume_mc,
 * region.cpp:499-701; RanPark, random_park.cpp:54-116), as host code of this
 * library; the particles then enter a system through dem_insert_particles.
 * Radius and density are constant per template, as the reference allows
 * (fix_template_sphere.cpp:180-215). */
#define DEM_INSERT_MAXTEMPLATES 8
typedef struct dem_insert_params {
  /* particletemplate/sphere <seed> atom_type <t> density constant <rho>
     radius constant <r>, in the order the distribution lists them */
  int ntemplates;
  int template_seed[DEM_INSERT_MAXTEMPLATES];
  int template_type[DEM_INSERT_MAXTEMPLATES];
  double template_radius[DEM_INSERT_MAXTEMPLATES];
  double template_density[DEM_INSERT_MAXTEMPLATES];
  /* particledistribution/discrete[/numberbased] <seed> <n> <id> <weight> ... */
  int distribution_seed;
  int distribution_numberbased;
  double distribution_weight[DEM_INSERT_MAXTEMPLATES];
  /* insert/pack seed <s> maxattempt <m> overlapcheck yes|no all_in yes|no
     random_distribute exact|uncorrelated vel constant|uniform|gaussian ...
     omega constant ... check_dist_from_subdomain_border yes|no ntry_mc <n>,
     exactly one of volumefraction_region / particles_in_region /
     mass_in_region, region <block> */
  int seed;
  int maxattempt;
  int overlapcheck;
  int all_in;
  int exact_number;
  int vel_style; /* 0 constant, 1 uniform (v +- dv), 2 gaussian (v, sigma dv) */
  double v_insert[3], v_fluct[3], omega[3];
  int check_dist_from_subdomain_border;
  int ntry_mc;
  double volumefraction_region;
  int particles_in_region;
  double mass_in_region;
  double region_lo[3], region_hi[3];
  /* the simulation box when the fix is set up (its first run) */
  double boxlo[3], boxhi[3];
  /* boundary s or m in some dimension (Domain::box_change, domain.cpp:192-
     208): FixInsertPack::insertion_fraction re-measures the region volume
     with the current box at every insertion (fix_insert_pack.cpp:356-362) */
  int box_change;
  /* warn_region yes (default) and some face of the box is f: a region whose
     bounding box reaches outside the box is an error
     (fix_insert_pack.cpp:252-261) */
  int warn_region;
  int fixed_boundary;
} dem_insert_params;

typedef struct dem_inserter dem_inserter;

/* Builds the fix (constructors + FixInsertPack::calc_insertion_properties:
 * region random stream reset to seed + 12, region volume by Monte Carlo).
 * Errors are the reference's error->all cases of these steps. */
int dem_inserter_create(const dem_insert_params *params, dem_inserter **out);
int dem_inserter_destroy(dem_inserter *ins);
const char *dem_inserter_last_error(const dem_inserter *ins);
/* One insertion with the current particles (x n*3, radius, rmass, any order)
 * and the current box: writes the placed particles in insertion order
 * (position, velocity, angular velocity n*3 each; radius, density, mass;
 * atom type). *ninserted receives how many the fix placed, *nrequested how
 * many it asked for. When cap is below the number asked for, the call
 * returns DEM_ERR_OVERFLOW having changed nothing (its random streams
 * included), with that number in *nrequested, so it can be repeated with
 * room enough. */
int dem_inserter_step(dem_inserter *ins, int64_t n, const double *x, const double *radius,
                      const double *rmass, const double *boxlo, const double *boxhi, int64_t cap,
                      double *x_out, double *v_out, double *omega_out, double *radius_out,
                      double *density_out, double *rmass_out, int *type_out, int64_t *ninserted,
                      int64_t *nrequested);
/* region volume the fix measured and the distribution's expectations */
int dem_inserter_info(const dem_inserter *ins, double *region_volume, double *volume_expect,
                      double *mass_expect, double *max_radius);

/* Runs an input script with the reference's command subset for this path
 * (Input::file, input.cpp:183, and Input::execute_command, input.cpp:560-650).
 * log_path may be NULL (log to stdout). Returns a dem_status. */
int dem_run_script(const char *script_path, const char *log_path);

#ifdef __cplusplus
}
#endif
#endif
