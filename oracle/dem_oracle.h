/*
 * CPU oracle for the DEM step: a plain-C, FP64, single-threaded restatement of
 * LIGGGHTS-PUBLIC 3.8.0's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, as the checker.
 *
 * Every function cites the reference file:line it restates. Ghost atoms of the
 * reference are replaced by the minimum-image convention on x_i - x_j: each
 * periodic pair is stored once instead of once per side (the reference's two
 * copies are mirror images up to rounding of the ghost coordinate).
 */
#ifndef DEM_ORACLE_H
#define DEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXTYPE 8
#define ORC_MAXPLANE 6
#define ORC_MAXMESH 4

/* per-type-pair mixed material matrices, indexed [itype][jtype], types 1..ntypes */
typedef struct {
  int ntypes;
  double Yeff[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double Geff[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double betaeff[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double mu[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double mur[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  /* coefficients of the sibling models: fix property/global kn, kt, gamman
   * (or gamman_abs), gammat (or gammat_abs), cohesionEnergyDensity
   * (peratomtypepair), characteristicVelocity (scalar) and log(e)
   * (global_properties.cpp:227-240, 291-294, 502-520, 735-812) */
  double kn[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double kt[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double gamman[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double gammat[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double ced[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double loge[ORC_MAXTYPE + 1][ORC_MAXTYPE + 1];
  double charvel;
} orc_materials;

/* normal models (model <name>) */
#define ORC_NORMAL_HERTZ 0
#define ORC_NORMAL_HOOKE 1
#define ORC_NORMAL_HERTZ_STIFFNESS 2
#define ORC_NORMAL_HOOKE_STIFFNESS 3

/* contact-model switches: pair_style gran model <normal> tangential <t>
 * [cohesion sjkr] [rolling_friction cdt] [...] */
typedef struct {
  int rolling_cdt;        /* rolling_friction cdt */
  int tangential_damping; /* default on  (normal_model_hertz.h:90) */
  int limit_force;        /* default off (normal_model_hertz.h:91) */
  int torsion_torque;     /* default off (rolling_model_cdt.h:73) */
  int normal_model;       /* ORC_NORMAL_* */
  int tangential_model;   /* 0 history, 1 no_history */
  int cohesion_model;     /* 0 off, 1 sjkr */
  int kt_to_kn;           /* hooke: ktToKnUser on (normal_model_hooke.h:101) */
  int absolute_damping;   /* hooke/stiffness: absolute_damping on (normal_model_hooke_stiffness.h:81) */
  int tangential_reduce;  /* sjkr: tangential_reduce on (cohesion_model_sjkr.h:72) */
  int rolling_epsd2;      /* rolling_friction epsd2 (rolling_model_epsd2.h), instead of cdt */
} orc_model;

typedef struct {
  double dt;
  double skin;
  double cdf; /* neigh_modify contact_distance_factor (1.0 default) */
  double boxlo[3], boxhi[3];
  int periodic[3];
  orc_model model;
  int has_gravity;
  double grav_acc[3]; /* magnitude * unit direction */
  int nplanes;
  int plane_axis[ORC_MAXPLANE]; /* 0=x 1=y 2=z */
  double plane_pos[ORC_MAXPLANE];
  int plane_type[ORC_MAXPLANE];
  int gravity_first; /* order of post_force fixes: 1 = gravity before walls */
  orc_model plane_model[ORC_MAXPLANE]; /* each fix wall/gran's own model switches */
  int neigh_every;    /* neigh_modify every N (0 is taken as 1) */
  int neigh_no_check; /* neigh_modify check no */
  /* boundary s / m faces of non-periodic dimensions (bit 2d lower, 2d+1 upper
   * face of dimension d; shrink_min: the m ones); other non-periodic faces are f */
  int shrink_faces;
  int shrink_min;
} orc_params;

typedef struct orc_sys orc_sys;

/* --- material mixing (global_properties.cpp:411-558) ------------------- */
void orc_mix_materials(int ntypes, const double *Y, const double *nu,
                       const double *e /* nt*nt */, const double *mu /* nt*nt */,
                       const double *mur /* nt*nt or NULL */, orc_materials *out);

/* Fills the sibling models' per-type-pair tables (nt*nt row-major each, any
 * may be NULL) and the characteristic velocity of model hooke. */
void orc_set_model_coeffs(orc_materials *m, int nt, const double *kn, const double *kt,
                          const double *gamman, const double *gammat, const double *ced,
                          double charvel);

/* --- one sphere-sphere contact, i side and j side ----------------------- */
/* del = x_i - x_j. shear (3) is read and, if shearupdate, written.
 * Outputs are the contributions to i (dFi, dTi) and to j (dTj; force on j is -dFi).
 * Returns 1 if the pair touches (rsq < radsum^2) and a force was produced. */
int orc_contact_pair(const orc_model *mdl, const orc_materials *mat, int itype, int jtype,
                     const double *del, double radi, double radj, const double *vi,
                     const double *vj, const double *wi, const double *wj, double mi,
                     double mj, double dt, int shearupdate, double *shear, double *dFi,
                     double *dTi, double *dTj);
/* the same with the rolling-torque history of rolling_friction epsd2 (3
 * values, read and, if shearupdate, written; may be NULL without epsd2) */
int orc_contact_pair_r(const orc_model *mdl, const orc_materials *mat, int itype, int jtype,
                       const double *del, double radi, double radj, const double *vi,
                       const double *vj, const double *wi, const double *wj, double mi,
                       double mj, double dt, int shearupdate, double *shear, double *rhist,
                       double *dFi, double *dTi, double *dTj);

/* --- one sphere-wall contact (fix_wall_gran_base.h:148-361) ------------- */
/* sidata_deltan = overlap (>0 when touching), delta = vector from centre to
 * wall contact point (sign as passed to compute_force after negation),
 * vwall = wall velocity at the contact. */
void orc_contact_wall(const orc_model *mdl, const orc_materials *mat, int itype, int wtype,
                      double sidata_deltan, const double *sidata_delta, double radi,
                      const double *vi, const double *wi, const double *vwall, double mi,
                      double dt, int shearupdate, double *shear, double *dFi, double *dTi);
void orc_contact_wall_r(const orc_model *mdl, const orc_materials *mat, int itype, int wtype,
                        double sidata_deltan, const double *sidata_delta, double radi,
                        const double *vi, const double *wi, const double *vwall, double mi,
                        double dt, int shearupdate, double *shear, double *rhist, double *dFi,
                        double *dTi);

/* --- brute-force O(N^2) half pair set (neigh_gran.cpp:571-582 predicate) - */
/* writes unordered pairs (i<j by index) into pairs[2*k]; returns the count or
 * -(needed) if cap is too small */
int64_t orc_pairs_bruteforce(int n, const double *x, const double *radius,
                             const orc_params *p, int64_t cap, int *pairs);

/* --- system life cycle (Verlet::setup / Verlet::run) --------------------- */
orc_sys *orc_create(const orc_params *p, const orc_materials *m, int n, const double *x,
                    const double *v, const double *omega, const double *radius,
                    const double *rmass, const int *type, const int *tag);
void orc_destroy(orc_sys *s);
/* Queues n particles that are inserted at the next step of orc_run, after its
 * initial integrate and before its (forced) neighbor build (FixInsert::
 * pre_exchange, fix_insert.cpp:652-826). Arrays as orc_create. */
void orc_insert(orc_sys *s, int n, const double *x, const double *v, const double *omega,
                const double *radius, const double *rmass, const int *type, const int *tag);
/* the largest radius of particles still to come (PairGran::init_one,
 * pair_gran.cpp:538-551); sets the cell size of the binned build */
void orc_set_max_radius(orc_sys *s, double rmax_future);
void orc_setup(orc_sys *s);
void orc_run(orc_sys *s, int nsteps);
/* one force evaluation on the current state (no integration), with
 * shearupdate as given; leaves f and torque in the system */
void orc_compute_forces(orc_sys *s, int shearupdate);
void orc_get_state(const orc_sys *s, double *x, double *v, double *omega, double *f,
                   double *torque);
/* pairs of the current half list, as (tag_a, tag_b) with tag_a < tag_b,
 * contact flag, and the shear history oriented from tag_a towards tag_b */
int64_t orc_get_pairs(const orc_sys *s, int64_t cap, int *tags, int *flags, double *hist);
/* compute pair/gran/local's default columns per half-list pair, oriented
 * from the lower tag (see dem_oracle.c); data [pairs][21 + 3 or 6] */
int64_t orc_pair_local(const orc_sys *s, int64_t cap, int *tags, double *data);
/* compute wall/gran/local rows of the mesh contacts (26 columns), changing
 * nothing; negative: -(rows) when above cap */
int64_t orc_wall_local(const orc_sys *s, int64_t cap, int *ids, double *data);
void orc_get_wall_hist(const orc_sys *s, double *hist /* n*nplanes*3 */);
/* rolling-torque histories (epsd2), in the orders of orc_get_pairs / orc_get_wall_hist */
int64_t orc_get_pair_rolling(const orc_sys *s, int64_t cap, double *rhist);
void orc_get_wall_rolling(const orc_sys *s, double *rhist);
int64_t orc_nrebuilds(const orc_sys *s);
/* atoms beyond an f face at the last neighbor build (lost by Comm::exchange),
 * and the first timestep whose build lost atoms (-1: none so far) */
int64_t orc_lost_atoms(const orc_sys *s, int64_t *first_step);
/* the box after the last build's Domain::reset_box (shrink-wrapped faces) */
void orc_get_box(const orc_sys *s, double *boxlo, double *boxhi);
int64_t orc_ntimestep(const orc_sys *s);
int orc_natoms(const orc_sys *s);
/* count of touching half-list pairs and of touching particle-wall contacts */
void orc_contact_counts(const orc_sys *s, int64_t *npair_touch, int64_t *nwall_touch);

/* --- triangle-mesh walls (fix mesh/surface + fix wall/gran ... mesh) --- */
/* Adds a mesh of ntri triangles (nodes: ntri*9, node-major). Computes the
 * element geometry (surface_mesh_I.h:361-447) and the topology: edge and
 * corner activity, neighbor faces (surface_mesh_I.h:454-560, 1035-1155).
 * Returns the mesh index or -1. Call after orc_create, before orc_setup. */
int orc_add_mesh(orc_sys *s, int ntri, const double *nodes, int wall_type,
                 const orc_model *model);
/* fix move/mesh mesh <id> rotate origin ... axis ... period T
 * (mesh_mover_rotation.cpp:66-139) */
void orc_mesh_stress(orc_sys *s, int imesh, const double *p_ref);
void orc_mesh_force(const orc_sys *s, int imesh, double *force, double *torque);
void orc_mesh_linear(orc_sys *s, int imesh, const double *vel);
void orc_mesh_wiggle(orc_sys *s, int imesh, const double *amplitude, double period);
void orc_mesh_viblin(orc_sys *s, int imesh, const double *axis, int order, const double *amplitude,
                     const double *phase, const double *period);
void orc_mesh_vibrot(orc_sys *s, int imesh, const double *origin, const double *axis, int order,
                     const double *amplitude, const double *phase, const double *period);
void orc_mesh_rotate(orc_sys *s, int imesh, const double *origin, const double *axis,
                     double period);
/* topology outputs: edge_active/corner_active ntri*3 (0/1), obtuse ntri (-1 = none) */
void orc_mesh_topology(const orc_sys *s, int imesh, int *edge_active, int *corner_active,
                       int *obtuse);
/* current nodes (ntri*9) and node velocities (ntri*9) */
void orc_mesh_nodes(const orc_sys *s, int imesh, double *nodes, double *vnodes);
/* mesh contact history: per particle up to maxslot (triangle index, shear[3]);
 * triangle -1 marks an empty slot. Returns the slot capacity used. */
int orc_mesh_history(const orc_sys *s, int imesh, int maxslot, int *tri, double *hist);
/* TriMesh::resolveTriSphereContactBary on one triangle given by its 3 nodes,
 * with the given activity flags; returns distance - radius and fills delta
 * (closest point - centre), bary and barysign. Geometry as orc_add_mesh. */
double orc_tri_contact(const double *nodes9, const int *edge_active3, const int *corner_active3,
                       const double *p, double radius, double *delta, double *bary,
                       int *barysign);
/* number of neighbor-list rebuilds forced by mesh motion */
int64_t orc_mesh_forced_rebuilds(const orc_sys *s);

#ifdef __cplusplus
}
#endif
#endif
