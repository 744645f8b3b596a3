// Input-script front end: runs a LIGGGHTS input script whose commands are
// the reference's subset for the per-timestep granular path, on the GPU
// library (Input::file, input.cpp:183; Input::execute_command, input.cpp:560-650).
//
// Supported physics commands keep the reference's syntax and meaning:
//   units si | atom_style granular|sphere | atom_modify ... | boundary
//   newton off | communicate single vel yes | processors ...
//   region <id> block xlo xhi ylo yhi zlo zhi [units box] | create_box <n> <region>
//   read_data <file>  (granular Atoms: id type diameter density x y z; Velocities)
//   neighbor <skin> bin | neigh_modify delay 0 [every N] [check yes|no]
//                                       [contact_distance_factor f] [one n]
//   fix <id> <grp> property/global youngsModulus|poissonsRatio peratomtype ...
//   fix <id> <grp> property/global coefficientRestitution|coefficientFriction|
//                  coefficientRollingFriction|kn|kt|gamman|gammat|gamman_abs|
//                  gammat_abs|cohesionEnergyDensity peratomtypepair <n> ...
//   fix <id> <grp> property/global characteristicVelocity scalar <v>
//   pair_style gran model hertz|hooke|hertz/stiffness|hooke/stiffness
//              tangential history|no_history [cohesion sjkr] [rolling_friction cdt|epsd2]
//              [tangential_damping on|off] [limitForce on|off] [torsionTorque on|off]
//              [ktToKnUser on|off] [absolute_damping on|off] [tangential_reduce on|off]
//   pair_coeff * * | timestep <dt>
//   fix <id> <grp> gravity <g> vector <x> <y> <z>
//   fix <id> <grp> wall/gran model <normal> tangential <tangential> [...]
//                  primitive type <t> xplane|yplane|zplane <pos>
//   fix <id> <grp> mesh/surface file <stl> | generate <cube|box|cylinder|pipe|disk|plane> <values>
//       type <t> [scale s] [move dx dy dz]
//                  [rotate axis ax ay az angle deg] [precision p] [verbose/heal ...]
//   fix <id> <grp> mesh/surface/stress ... [reference_point x y z] [stress on|off]
//                  (thermo f_<id>[1..6]: total force and torque on the mesh)
//   compute <id> <grp> erotate/sphere | ke  (thermo c_<id>)
//   fix <id> <grp> wall/gran model <normal> tangential <tangential> [...] mesh n_meshes <k>
//                  meshes <id1> ... <idk>
//   fix <id> <grp> move/mesh mesh <meshid> rotate origin ox oy oz axis ax ay az period T
//                  | linear vx vy vz | wiggle amplitude ax ay az period T
//                  | viblin axis x y z order n amplitude .. phase .. period ..
//                  | vibrot origin x y z axis x y z order n amplitude .. phase .. period ..
//   fix <id> <grp> nve/sphere
//   atom_style superquadric  (read_data Atoms: id type density shapex shapey shapez
//                             blockiness1 blockiness2 x y z), with
//   pair_style gran model hertz tangential history surface superquadric [...],
//   fix <id> <grp> wall/gran model hertz tangential history surface superquadric
//                  primitive type <t> xplane|yplane|zplane <pos>,
//   fix <id> <grp> nve/superquadric [integration_scheme 1],
//   set atom|type|group <id> quat <a> <b> <c> <theta> | quat/random <seed>
//   thermo <N> | thermo_style custom step atoms ke ... | thermo_modify ...
//   variable <name> equal|index <value> | print | echo | log
//   run <N> [pre no|yes] [post no|yes]
//   package amd precision fp64|mixed   (selects the GPU arithmetic mode)
//   dump <id> all mesh/stl <N> <file*.stl> [binary] [region <id>] [mesh ids]
//   dump <id> all custom <N> <file> <columns> | dump_modify <id> ... | undump <id>
//                  (columns id type x y z vx vy vz fx fy fz omegax omegay omegaz
//                   tqx tqy tqz radius diameter mass; '*' in the file name: one
//                   file per snapshot; atoms in ID order)
//   write_restart <file> | read_restart <file>  (this library's restart layout:
//                  atoms, contact histories, meshes, timestep)
// Other output-only commands (compute, restart, ...) are accepted and ignored
// with a warning; anything that would change the physics and is not listed
// is an error, as the reference reports for an unknown style.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/liggghts_amd.h"
#include "mesh_host.h"

namespace {

constexpr double kMyPi = 3.14159265358979323846;

struct ScriptError {
  std::string msg;
};

[[noreturn]] void die(const std::string &m) { throw ScriptError{m}; }

double num(const std::string &s) {
  char *end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end != '\0') die("Expected floating point parameter instead of '" + s + "'");
  return v;
}

int inum(const std::string &s) {
  char *end = nullptr;
  const long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || *end != '\0') die("Expected integer parameter instead of '" + s + "'");
  return (int)v;
}

// RanPark (random_park.cpp:54-80, 125-157): Park-Miller minimal standard
// generator, seeded per atom by the Jenkins one-at-a-time hash of the caller's
// seed and the atom's coordinates, as set ... quat/random does it
struct RanPark {
  int seed = 1;
  double uniform() {
    const int k = seed / 127773;
    seed = 16807 * (seed - k * 127773) - 2836 * k;
    if (seed < 0) seed += 2147483647;
    return (1.0 / 2147483647) * seed;
  }
  void reset(int ibase, const double *coord) {
    unsigned int hash = 0;
    auto mix = [&hash](const void *p, int n) {
      const signed char *c = (const signed char *)p;  // char is signed on x86-64
      for (int i = 0; i < n; i++) {
        hash += (unsigned int)(int)c[i];
        hash += (hash << 10);
        hash ^= (hash >> 6);
      }
    };
    mix(&ibase, (int)sizeof(int));
    mix(coord, 3 * (int)sizeof(double));
    hash += (hash << 3);
    hash ^= (hash >> 11);
    hash += (hash << 15);
    seed = (int)(hash & 0x7ffffff);
    if (!seed) seed = 1;
    for (int i = 0; i < 5; i++) uniform();
  }
};

struct Interp {
  FILE *out = stdout;
  // box / domain
  bool have_box = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int periodic[3] = {1, 1, 1};
  // boundary s / m faces (bit 2d lower, 2d+1 upper face of dimension d) and
  // the face letters for dump box headers (Domain::boundary_string)
  int shrink_faces = 0, shrink_min = 0;
  char face[3][2] = {{'p', 'p'}, {'p', 'p'}, {'p', 'p'}};
  // thermo_modify lost error | warn | ignore (Thermo::lostflag, thermo.cpp:410-452)
  std::string lost_policy = "error";
  int ntypes = 0;
  std::map<std::string, std::vector<double>> regions;
  // particles
  std::vector<double> x, v, w, r, m;
  std::vector<int> type, tag;
  // neighbor / run
  double skin = -1.0, cdf = 1.0, dt = -1.0;
  int every = 1, no_check = 0;
  int one = 0;
  // materials
  std::vector<double> Y, nu, e, mu, mur;
  // sibling-model properties: kn, kt, gamman, gammat, gamman_abs, gammat_abs,
  // cohesionEnergyDensity (peratomtypepair) and characteristicVelocity
  std::map<std::string, std::vector<double>> model_props;
  // superquadrics (atom_style superquadric)
  bool sq = false;
  std::vector<double> sq_shape, sq_blk, sq_dens, sq_quat, sq_mass;
  bool pair_sq_surface = false, nve_sq = false, omega_warned = false;
  dem_sqdem_handle *hs = nullptr;
  // pair
  bool pair_set = false;
  // pair_style gran model switches (tangential damping on by default)
  dem_contact_model pair_m{0, 1, 0, 0, DEM_NORMAL_HERTZ, 0, 0, 0, 0, 0};
  // fixes
  bool nve = false;
  bool gravity = false, gravity_first = true, wall_seen = false;
  double gacc[3] = {0, 0, 0};
  struct PlaneWall {
    int axis;
    double pos;
    int type;
    dem_contact_model model;
  };
  std::vector<PlaneWall> walls;
  struct MeshDef {
    std::vector<double> nodes;
    int type = 1;
    bool used = false;
    dem_contact_model model{};
    double origin[3] = {0, 0, 0}, axis[3] = {0, 0, 1}, period = 1.0;
    int mover = 0;  // 0 none, 1 rotate, 2 linear, 3 wiggle, 4 viblin, 5 vibrot
    double vec[3] = {0, 0, 0};  // linear: velocity; wiggle: amplitude
    std::vector<double> vampl, vphase, vperiod;  // viblin / vibrot modes
    bool stress = false;        // fix mesh/surface/stress (stress on)
    double pref[3] = {0, 0, 0};  // its reference_point
  };
  std::map<std::string, int> mesh_ids;  // library mesh id of each fix mesh/surface
  std::map<std::string, std::string> computes;  // compute id -> style (erotate/sphere, ke)
  // compute pair/gran/local (compute_pair_gran_local.cpp:71-124): the value
  // groups of a row, in the reference's column order
  struct LocalCompute {
    std::vector<std::string> keys;
    int ncols = 0;
    bool wall = false;      // wall/gran/local
    std::string wall_fix;   // its fix wall/gran (the last one with meshes)
    int dnum = 0;           // history values per row
  };
  std::map<std::string, LocalCompute> local_computes;
  std::map<std::string, MeshDef> mesh_defs;
  std::vector<std::string> mesh_order;  // meshes in fix wall/gran order
  // per mesh: its fix wall/gran and its index among that fix's meshes
  std::map<std::string, std::pair<std::string, int>> mesh_fix;
  std::string last_mesh_wall_fix;
  int precision = DEM_FP64;
  // thermo
  int thermo_every = 0;
  std::vector<std::string> thermo_cols = {"step", "atoms", "ke"};
  std::map<std::string, std::string> vars;
  std::map<std::string, std::string> var_style;  // equal | index | string
  // system
  dem_handle *h = nullptr;
  bool dirty = true;  // particles/params changed since the handle was built
  int64_t step = 0;

  // dump <id> all custom <N> <file> <columns...> (DumpCustom, dump_custom.cpp)
  struct Dump {
    std::string file;  // '*' is replaced by the timestep: one file per snapshot
    long long every = 0;
    std::vector<std::string> cols;
    FILE *fp = nullptr;
    long long last = -1;  // timestep of the last snapshot written
    bool stl = false;     // dump mesh/stl
    bool binary = false, ghost = false;
    std::vector<std::string> meshes;  // mesh/stl: the meshes dumped, in order
    std::string region;               // mesh/stl: only triangles whose centre is inside
    bool local = false;               // dump local: rows of a local compute
    bool vtk = false;                 // dump custom/vtk: legacy VTK polydata
  };
  std::vector<std::string> mesh_def_order;  // fix mesh/surface in definition order
  // thermo `time` (Update::atime / atimestep, update.cpp:434-446) and `cpu`
  double atime = 0.0;
  long long atimestep = 0;
  std::chrono::steady_clock::time_point run_t0 = std::chrono::steady_clock::now();
  std::map<std::string, Dump> dumps;
  std::string restart_file;  // read_restart: device state applied when the system is built

  // fix particletemplate/sphere (fix_template_sphere.cpp:72-245): constant
  // radius and density
  struct Template {
    int seed = 0, type = 1;
    double radius = 0, density = 0;
  };
  std::map<std::string, Template> templates;
  // fix particledistribution/discrete (fix_particledistribution_discrete.cpp:69-244)
  struct Distribution {
    int seed = 0;
    bool numberbased = false;
    std::vector<std::string> tmpl;
    std::vector<double> weight;
  };
  std::map<std::string, Distribution> distributions;
  // fix insert/pack (fix_insert.cpp:78-340, fix_insert_pack.cpp:74-160)
  struct InsertFix {
    dem_insert_params p{};
    std::string group, region;
    long long next = 0;           // next_reneighbor: the step of the next insertion, 0 none
    long long every = 0;          // insert_every (0: once)
    dem_inserter *ins = nullptr;  // built at its first run's setup (FixInsert::setup)
  };
  std::map<std::string, InsertFix> inserts;
  std::vector<std::string> insert_order;  // pre_exchange runs the fixes in definition order
  // groups (group <id> region <rid>): whether every atom belongs to it; the
  // GPU path integrates and walls every atom, so a fix on a group that does
  // not hold them all is refused at run time
  std::map<std::string, bool> group_all;
  std::vector<std::pair<std::string, std::string>> physics_groups;  // (fix id, group) of the GPU fixes

  ~Interp() {
    for (auto &d : dumps)
      if (d.second.fp) fclose(d.second.fp);
    for (auto &kv : inserts)
      if (kv.second.ins) dem_inserter_destroy(kv.second.ins);
    if (h) dem_destroy(h);
    if (hs) dem_sqdem_destroy(hs);
  }

  // Park-Miller seeds must be primes above 10000 (Random::Random,
  // random.cpp:71-92, with seed_check_error on by default, input.cpp:157)
  static int prime_seed(const std::string &s) {
    const int v = inum(s);
    bool prime = v > 1;
    for (long long d = 2; prime && d * d <= v; d++) prime = v % d != 0;
    if (v < 10000 || !prime)
      die("Random number generation: It is required that the random seed value is > 10000 and a prime "
          "number. The random seed used was " + s);
    return v;
  }

  // fix <id> <group> particletemplate/sphere <seed> atom_type <t>
  //     density constant <rho> radius constant <r> [volume_limit <v>]
  void fix_template(const std::vector<std::string> &t) {
    if (t.size() < 5) die("fix particletemplate/sphere: not enough arguments");
    Template tp;
    tp.seed = prime_seed(t[4]);
    double vol_limit = 1e-12;
    for (size_t k = 5; k < t.size();) {
      if (t[k] == "atom_type" && k + 1 < t.size()) {
        tp.type = inum(t[k + 1]);
        if (tp.type < 1) die("fix particletemplate/sphere: invalid atom type (must be >=1)");
        k += 2;
      } else if ((t[k] == "density" || t[k] == "radius") && k + 2 < t.size()) {
        if (t[k + 1] != "constant")
          die("fix particletemplate/sphere: invalid " + t[k] + " random style");  // :180-215
        const double v = num(t[k + 2]);
        if (v <= 0.) die("fix particletemplate/sphere: " + t[k] + " must be >= 0");
        (t[k] == "density" ? tp.density : tp.radius) = v;
        k += 3;
      } else if (t[k] == "volume_limit" && k + 1 < t.size()) {
        vol_limit = num(t[k + 1]);
        if (vol_limit <= 0) die("fix particletemplate/sphere: volume_limit > 0 required");
        k += 2;
      } else {
        die("fix particletemplate/sphere: unrecognized keyword " + t[k]);
      }
    }
    if (tp.density <= 0) die("fix particletemplate/sphere: have to define 'density'");
    if (tp.radius <= 0) die("fix particletemplate/sphere: have to define 'radius'");
    if (tp.radius * tp.radius * tp.radius * 4. * kMyPi / 3. < vol_limit)
      die("fix particletemplate/sphere: Volume expectancy too small");
    if (tp.type > ntypes) die("fix particletemplate/sphere: atom type beyond create_box's types");
    templates[t[1]] = tp;
  }

  // fix <id> <group> particledistribution/discrete[/numberbased|/massbased]
  //     <seed> <n> <template1> <weight1> ...
  void fix_distribution(const std::vector<std::string> &t) {
    if (t.size() < 7) die("fix particledistribution/discrete: not enough arguments");
    Distribution d;
    d.numberbased = t[3].find("numberbased") != std::string::npos;
    d.seed = prime_seed(t[4]);
    const int nt = inum(t[5]);
    if (nt < 1) die("fix particledistribution/discrete: illegal number of templates");
    if (nt > DEM_INSERT_MAXTEMPLATES) die("fix particledistribution/discrete: at most 8 templates on this path");
    if ((int)t.size() != 6 + 2 * nt)
      die("fix particledistribution/discrete: # of templates does not match # of arguments");
    for (int i = 0; i < nt; i++) {
      const std::string &id = t[6 + 2 * i];
      if (!templates.count(id)) die("fix particledistribution/discrete: invalid ID for fix particletemplate provided");
      for (const auto &prev : d.tmpl)
        if (prev == id) die("fix particledistribution/discrete: cannot use the same template twice");
      d.tmpl.push_back(id);
      const double w = num(t[7 + 2 * i]);
      if (w < 0) die("fix particledistribution/discrete: invalid weight");
      d.weight.push_back(w);
    }
    double ws = 0;
    for (double w : d.weight) ws += w;
    if (std::fabs(ws - 1.) > 0.00001)
      fprintf(out, "WARNING: particledistribution/discrete: sum of distribution weights != 1, normalizing "
                   "distribution\n");
    distributions[t[1]] = d;
  }

  // fix <id> <group> insert/pack seed <s> distributiontemplate <d> [...] region <r>
  //     volumefraction_region <f> | particles_in_region <n> | mass_in_region <m>
  void fix_insert_pack(const std::vector<std::string> &t) {
    if (t.size() < 6 || t[4] != "seed") die("fix insert/pack: expecting keyword 'seed'");
    InsertFix f;
    dem_insert_params &p = f.p;
    p.seed = prime_seed(t[5]);
    p.maxattempt = 50;
    p.overlapcheck = 1;
    p.exact_number = 1;
    p.check_dist_from_subdomain_border = 1;
    p.ntry_mc = 100000;
    p.warn_region = 1;
    f.group = t[2];
    f.next = step + 1;  // first insertion on the next step (fix_insert.cpp:394-405)
    f.every = -1;
    std::string dist;
    for (size_t k = 6; k < t.size();) {
      const std::string &kw = t[k];
      auto need = [&](size_t m) {
        if (k + m >= t.size()) die("fix insert/pack: not enough arguments for " + kw);
      };
      if (kw == "distributiontemplate") {
        need(1);
        dist = t[k + 1];
        k += 2;
      } else if (kw == "maxattempt") {
        need(1);
        p.maxattempt = inum(t[k + 1]);
        k += 2;
      } else if (kw == "insert_every" || kw == "every") {
        need(1);
        f.every = t[k + 1] == "once" ? 0 : inum(t[k + 1]);
        if (f.every < 0) die("fix insert/pack: insert_every must be >= 0");
        k += 2;
      } else if (kw == "insert_every_time") {
        need(1);
        if (dt <= 0) die("fix insert/pack: need so set 'timestep' before");
        f.every = t[k + 1] == "once" ? 0 : (long long)(num(t[k + 1]) / dt);
        k += 2;
      } else if (kw == "start") {
        need(1);
        f.next = inum(t[k + 1]);
        if (f.next < step + 1) die("fix insert/pack: 'start' step can not be before current step");
        k += 2;
      } else if (kw == "overlapcheck" || kw == "all_in" || kw == "warn_region" ||
                 kw == "check_dist_from_subdomain_border") {
        need(1);
        if (t[k + 1] != "yes" && t[k + 1] != "no") die("fix insert/pack: expecting 'yes' or 'no' after " + kw);
        const int v = t[k + 1] == "yes";
        if (kw == "overlapcheck") p.overlapcheck = v;
        else if (kw == "all_in") p.all_in = v;
        else if (kw == "warn_region") p.warn_region = v;
        else p.check_dist_from_subdomain_border = v;
        k += 2;
      } else if (kw == "random_distribute") {
        need(1);
        if (t[k + 1] == "uncorrelated") p.exact_number = 0;
        else if (t[k + 1] == "exact") p.exact_number = 1;
        else die("fix insert/pack: expecting 'uncorrelated' or 'exact' after 'random_distribute'");
        k += 2;
      } else if (kw == "vel") {
        need(4);
        const std::string &st = t[k + 1];
        const int nv = st == "constant" ? 3 : (st == "uniform" || st == "gaussian") ? 6 : -1;
        if (nv < 0) die("fix insert/pack: expecting 'constant', 'uniform' or 'gaussian' after 'vel'");
        need(nv);
        p.vel_style = st == "constant" ? 0 : st == "uniform" ? 1 : 2;
        for (int d = 0; d < 3; d++) p.v_insert[d] = num(t[k + 2 + d]);
        if (nv == 6)
          for (int d = 0; d < 3; d++) p.v_fluct[d] = num(t[k + 5 + d]);
        k += 2 + nv;
      } else if (kw == "omega") {
        need(4);
        if (t[k + 1] != "constant") die("fix insert/pack: expecting 'constant' after 'omega'");
        for (int d = 0; d < 3; d++) p.omega[d] = num(t[k + 2 + d]);
        k += 5;
      } else if (kw == "region") {
        need(1);
        if (!regions.count(t[k + 1])) die("fix insert/pack: region ID does not exist");
        f.region = t[k + 1];
        k += 2;
      } else if (kw == "volumefraction_region") {
        need(1);
        p.volumefraction_region = num(t[k + 1]);
        if (p.volumefraction_region < 0. || p.volumefraction_region > 1.)
          die("fix insert/pack: Invalid volumefraction");
        k += 2;
      } else if (kw == "particles_in_region") {
        need(1);
        p.particles_in_region = inum(t[k + 1]);
        if (p.particles_in_region <= 0) die("fix insert/pack: 'ntotal_region' > 0 required");
        k += 2;
      } else if (kw == "mass_in_region") {
        need(1);
        p.mass_in_region = num(t[k + 1]);
        if (p.mass_in_region <= 0.) die("fix insert/pack: 'masstotal_region' > 0 required");
        k += 2;
      } else if (kw == "ntry_mc") {
        need(1);
        p.ntry_mc = inum(t[k + 1]);
        if (p.ntry_mc < 1000) die("fix insert/pack: ntry_mc must be > 1000");
        k += 2;
      } else {
        die("fix insert/pack: unknown keyword " + kw);
      }
    }
    if (dist.empty() || !distributions.count(dist))
      die("Fix insert requires you to define a valid ID for a fix of type particledistribution/discrete");
    if (f.region.empty()) die("fix insert/pack: must define an insertion region");
    if (f.every < 0) die("fix insert/pack: must define 'insert_every'");
    const Distribution &d = distributions[dist];
    p.ntemplates = (int)d.tmpl.size();
    p.distribution_seed = d.seed;
    p.distribution_numberbased = d.numberbased ? 1 : 0;
    std::vector<int> seeds{p.seed, d.seed};
    for (int i = 0; i < p.ntemplates; i++) {
      const Template &tp = templates[d.tmpl[i]];
      p.template_seed[i] = tp.seed;
      p.template_type[i] = tp.type;
      p.template_radius[i] = tp.radius;
      p.template_density[i] = tp.density;
      p.distribution_weight[i] = d.weight[i];
      seeds.push_back(tp.seed);
    }
    // FixInsert::sanity_check (fix_insert.cpp:434-482): every stream its own seed
    std::sort(seeds.begin(), seeds.end());
    if (std::unique(seeds.begin(), seeds.end()) != seeds.end())
      die("Fix insert/pack: Random number generation: It is required that all the random seeds of this fix "
          "insert/*, the random seed of particle distribution fix and all random seeds of the templates are "
          "different");
    const std::vector<double> &b = regions[f.region];
    for (int k = 0; k < 3; k++) {
      p.region_lo[k] = b[2 * k];
      p.region_hi[k] = b[2 * k + 1];
    }
    if (inserts.count(t[1])) die("Replacing a fix insert is not available on this path");
    inserts[t[1]] = f;
    insert_order.push_back(t[1]);
  }

  // the largest radius insertion fixes may bring (FixInsert::max_rad for
  // PairGran::init_one, pair_gran.cpp:538-551)
  double insert_max_radius() {
    double r = 0.0;
    for (auto &kv : inserts)
      for (int i = 0; i < kv.second.p.ntemplates; i++) r = std::max(r, kv.second.p.template_radius[i]);
    return r;
  }

  // One FixInsert::pre_exchange (fix_insert.cpp:652-826) at the current step:
  // place the fix's particles among the present ones, append them to the
  // script's atoms with new tags (Atom::tag_extend) and hand them to the
  // library, where they join at the next step's force pass.
  void do_insertion(const std::string &id, InsertFix &f) {
    double blo[3], bhi[3];
    current_box(blo, bhi);
    if (!f.ins) {  // FixInsert::setup: the region volume with the box of the first run
      dem_insert_params p = f.p;
      for (int d = 0; d < 3; d++) {
        p.boxlo[d] = blo[d];
        p.boxhi[d] = bhi[d];
      }
      p.box_change = shrink_faces ? 1 : 0;
      bool fixed = false;
      for (int d = 0; d < 3; d++)
        for (int s = 0; s < 2; s++) fixed = fixed || face[d][s] == 'f';
      p.fixed_boundary = fixed ? 1 : 0;
      if (dem_inserter_create(&p, &f.ins)) die("fix " + id + ": " + dem_inserter_last_error(nullptr));
    }
    if (h && !tag.empty()) pull_state();
    const int64_t n = (int64_t)tag.size();
    int64_t cap = 4096, nins = 0, nreq = 0;
    std::vector<double> xo, vo, wo, ro, rho, mo;
    std::vector<int> to;
    for (int attempt = 0;; attempt++) {
      xo.resize(3 * cap);
      vo.resize(3 * cap);
      wo.resize(3 * cap);
      ro.resize(cap);
      rho.resize(cap);
      mo.resize(cap);
      to.resize(cap);
      const int rc = dem_inserter_step(f.ins, n, x.data(), r.data(), m.data(), blo, bhi, cap, xo.data(), vo.data(),
                                       wo.data(), ro.data(), rho.data(), mo.data(), to.data(), &nins, &nreq);
      if (rc == DEM_ERR_OVERFLOW && attempt == 0) {  // nothing changed: again with room enough
        cap = nreq;
        continue;
      }
      if (rc) die("fix " + id + ": " + dem_inserter_last_error(f.ins));
      break;
    }
    if (nins < nreq)
      fprintf(out, "WARNING: Particle insertion: Less insertions than requested\n");
    fprintf(out, "INFO: Particle insertion %s: inserted %lld particle templates (mass %g) at step %lld\n",
            id.c_str(), (long long)nins, [&] {
              double s = 0;
              for (int64_t k = 0; k < nins; k++) s += mo[k];
              return s;
            }(), (long long)step + 1);
    // the next insertion (fix_insert.cpp:700-708, 821-825)
    f.next = f.every > 0 ? f.next + f.every : 0;
    if (nins == 0) return;
    int maxtag = 0;
    for (int tg : tag) maxtag = std::max(maxtag, tg);
    std::vector<int> newtag(nins);
    for (int64_t k = 0; k < nins; k++) newtag[k] = maxtag + 1 + (int)k;
    if (h && dem_insert_particles(h, nins, xo.data(), vo.data(), wo.data(), ro.data(), mo.data(), to.data(),
                                  newtag.data()))
      die(dem_last_error(h));
    x.insert(x.end(), xo.begin(), xo.begin() + 3 * nins);
    v.insert(v.end(), vo.begin(), vo.begin() + 3 * nins);
    w.insert(w.end(), wo.begin(), wo.begin() + 3 * nins);
    r.insert(r.end(), ro.begin(), ro.begin() + nins);
    m.insert(m.end(), mo.begin(), mo.begin() + nins);
    type.insert(type.end(), to.begin(), to.begin() + nins);
    tag.insert(tag.end(), newtag.begin(), newtag.end());
    // the new atoms join "all" and the fix's group only
    for (auto &g : group_all)
      if (g.first != f.group) g.second = false;
  }

  // the earliest pending insertion step, 0 if none
  long long next_insertion() const {
    long long s = 0;
    for (const auto &kv : inserts)
      if (kv.second.next > 0 && (s == 0 || kv.second.next < s)) s = kv.second.next;
    return s;
  }

  static bool dump_column_known(const std::string &c) {
    static const char *known[] = {"id", "type", "x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz",
                                  "omegax", "omegay", "omegaz", "tqx", "tqy", "tqz", "radius",
                                  "diameter", "mass"};
    for (const char *k : known)
      if (c == k) return true;
    return false;
  }

  // dump id group mesh/stl N file*.stl [binary] [all|local|ghost] [region id] [mesh ...]
  // (DumpMeshSTL's constructor, dump_mesh_stl.cpp:71-158): no mesh named
  // means every fix mesh/surface, in the order they were defined
  void dump_stl_command(const std::vector<std::string> &t) {
    if (t.size() < 6) die("Illegal dump mesh/stl command");
    if (dumps.count(t[1])) die("Reuse of dump ID " + t[1]);
    Dump d;
    d.stl = true;
    d.every = std::stoll(t[4]);
    if (d.every <= 0) die("Illegal dump command");
    d.file = t[5];
    if (d.file.find('*') == std::string::npos)
      die("You should use a filename like 'dump*.stl' for the 'mesh/stl' command to produce one file "
          "per time-step");
    for (size_t k = 6; k < t.size(); k++) {
      if (t[k].compare(0, 6, "binary") == 0) d.binary = true;
      else if (t[k] == "all" || t[k] == "local") d.ghost = false;
      else if (t[k] == "ghost") d.ghost = true;
      else if (t[k] == "region") {
        if (k + 1 >= t.size()) die("Illegal dump mesh/stl command, not enough arguments");
        if (!regions.count(t[k + 1])) die("Illegal dump mesh/stl command, region ID does not exist");
        d.region = t[++k];
      } else if (mesh_defs.count(t[k])) {
        d.meshes.push_back(t[k]);
      } else {
        die("Illegal dump mesh/stl command, unknown keyword or mesh");
      }
    }
    if (d.meshes.empty()) d.meshes = mesh_def_order;
    if (d.meshes.empty()) fprintf(out, "WARNING: Dump mesh/stl cannot find any fix of type 'mesh/surface' to dump\n");
    dumps[t[1]] = d;
  }

  // one STL snapshot (DumpMeshSTL::pack / write_data_ascii / write_data_binary,
  // dump_mesh_stl.cpp:306-384): per triangle the unit normal of its current
  // nodes and the three nodes; one process owns every triangle, so `ghost`
  // selects none
  void write_stl(const Dump &d) {
    std::string name = d.file;
    name.replace(name.find('*'), 1, std::to_string((long long)step));
    std::vector<double> rec;  // 12 values per triangle
    for (const std::string &mn : d.meshes) {
      auto it = mesh_ids.find(mn);
      if (it == mesh_ids.end())
        die("dump mesh/stl: mesh " + mn + " is not used by a fix wall/gran on this path");
      if (d.ghost) continue;
      const int T = (int)(mesh_defs[mn].nodes.size() / 9);
      std::vector<double> nodes(9 * (size_t)T), vn(9 * (size_t)T);
      if (dem_get_mesh_nodes(h, it->second, nodes.data(), vn.data())) die(dem_last_error(h));
      for (int tri = 0; tri < T; tri++) {
        const double *p = &nodes[9 * (size_t)tri];
        if (!d.region.empty()) {
          const std::vector<double> &b = regions[d.region];
          bool in = true;
          for (int k = 0; k < 3; k++) {
            const double c = (p[k] + p[3 + k] + p[6 + k]) / 3.0;
            in = in && c >= b[2 * k] && c <= b[2 * k + 1];
          }
          if (!in) continue;
        }
        const double e0[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
        const double e1[3] = {p[6] - p[3], p[7] - p[4], p[8] - p[5]};
        double nrm[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                         e0[0] * e1[1] - e0[1] * e1[0]};
        const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int k = 0; k < 3; k++) rec.push_back(nrm[k] / len);
        for (int k = 0; k < 9; k++) rec.push_back(p[k]);
      }
    }
    const size_t nt = rec.size() / 12;
    FILE *fp = fopen(name.c_str(), d.binary ? "wb" : "w");
    if (!fp) die("Cannot open dump file " + name);
    if (d.binary) {
      char header[81];
      std::snprintf(header, sizeof header, "%-80s", "STL File written by LIGGGHTS");
      fwrite(header, 1, 80, fp);
      const uint32_t n32 = (uint32_t)nt;
      fwrite(&n32, 4, 1, fp);
      for (size_t i = 0; i < nt; i++) {
        for (int k = 0; k < 12; k++) {
          const float f = (float)rec[12 * i + k];
          fwrite(&f, 4, 1, fp);
        }
        const uint16_t attr = 0;
        fwrite(&attr, 2, 1, fp);
      }
    } else {
      fprintf(fp, "solid LIGGGHTS_STL_EXPORT\n");
      for (size_t i = 0; i < nt; i++) {
        const double *q = &rec[12 * i];
        fprintf(fp,
                "  facet normal %g %g %g\n    outer loop\n      vertex %g %g %g\n"
                "      vertex %g %g %g\n      vertex %g %g %g\n    endloop\n  endfacet\n",
                q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11]);
      }
      fprintf(fp, "endsolid LIGGGHTS_STL_EXPORT\n");
    }
    fclose(fp);
  }

  // history values per pair the pair style registers (dnum): the tangential
  // shear (3, tangential history) and the epsd2 rolling torque (3)
  int pair_dnum() const {
    return (pair_m.tangential_model == 0 ? 3 : 0) + (pair_m.rolling_epsd2 ? 3 : 0);
  }

  // compute ID group pair/gran/local|wall/gran/local [keywords]
  // (compute_pair_gran_local.cpp:71-124, the column count of :281-283). Rows
  // come from dem_compute_pair_local / dem_compute_wall_local. wall/gran/local
  // takes the mesh contacts of the last fix wall/gran with meshes defined
  // before it (:247-270).
  void local_compute_command(const std::vector<std::string> &t) {
    const bool wall = t[3] == "wall/gran/local";
    if (step > 0) die("Compute pair/gran/local: Need to define this compute before first run");
    if (t[2] != "all") die("compute pair/gran/local: only group all is available on this path");
    static const char *order[] = {"pos", "vel", "id", "force", "force_normal", "force_tangential", "torque",
                                  "torque_normal", "torque_tangential", "history", "contactArea", "delta",
                                  "contactPoint"};
    std::set<std::string> want;
    if (t.size() == 4) {
      want = {"pos", "vel", "id", "force", "torque", "history", "contactArea"};
    } else {
      for (size_t k = 4; k < t.size(); k++) {
        const std::string &w = t[k];
        if (w == "verbose") continue;
        if (w == "extraSurfDistance")
          die("this keyword is deprecated; neighbor->contactDistanceFactor is now used directly");
        if (w == "heatFlux")
          die("Compute pair/gran/local can not calculate heat flux values since no fix heat/gran/conduction "
              "not compute them");
        if (w == "ms_id") die("compute pair/gran/local ms_id needs fix multisphere, not available on this path");
        bool known = false;
        for (const char *o : order) known = known || w == o;
        if (!known) die("Compute pair/gran/local: illegal/unrecognized keyword " + w);
        want.insert(w);
      }
    }
    LocalCompute lc;
    for (const char *o : order)
      if (want.count(o)) lc.keys.push_back(o);
    int dnum = pair_dnum();
    if (wall) {
      if (last_mesh_wall_fix.empty())
        die("Compute wall/gran/local requires a fix of type wall/gran using one or more mesh walls. This fix "
            "has to come before the compute in the script");
      lc.wall = true;
      lc.wall_fix = last_mesh_wall_fix;
      // the shear of tangential history; no rolling history on meshes here
      dnum = 0;
      for (const auto &mf : mesh_fix)
        if (mf.second.first == last_mesh_wall_fix)
          dnum = mesh_defs[mf.first].model.tangential_model == 0 ? 3 : 0;
    }
    lc.dnum = dnum;
    if (want.count("history") && dnum == 0)
      die("Compute pair/gran/local or wall/gran/local can not calculate history values since pair or wall "
          "style does not compute them");
    for (const std::string &k : lc.keys) {
      if (k == "pos" || k == "vel") lc.ncols += 6;
      else if (k == "history") lc.ncols += dnum;
      else if (k == "contactArea" || k == "delta") lc.ncols += 1;
      else lc.ncols += 3;
    }
    local_computes[t[1]] = lc;
  }

  // dump ID group local N file c_ID[k] ... | index (dump_local.cpp)
  void dump_local_command(const std::vector<std::string> &t) {
    if (dumps.count(t[1])) die("Reuse of dump ID " + t[1]);
    Dump d;
    d.local = true;
    d.every = std::stoll(t[4]);
    if (d.every <= 0) die("Illegal dump command");
    d.file = t[5];
    for (size_t k = 6; k < t.size(); k++) {
      const std::string &c = t[k];
      if (c != "index") {
        const size_t lb = c.find('[');
        if (c.compare(0, 2, "c_") != 0 || lb == std::string::npos || c.back() != ']')
          die("Invalid attribute " + c + " in dump local command");
        const std::string id = c.substr(2, lb - 2);
        auto it = local_computes.find(id);
        if (it == local_computes.end()) die("Could not find dump local compute ID " + id);
        const int k1 = std::atoi(c.substr(lb + 1, c.size() - lb - 2).c_str());
        if (k1 < 1 || k1 > it->second.ncols) die("Dump local compute does not calculate local array column " + c);
      }
      d.cols.push_back(c);
    }
    if (d.cols.empty()) die("Illegal dump command");
    if (d.file.find('*') == std::string::npos) {
      d.fp = fopen(d.file.c_str(), "w");
      if (!d.fp) die("Cannot open dump file " + d.file);
    }
    dumps[t[1]] = d;
  }

  // The rows of one local compute at this step: ComputePairGranLocal::add_pair
  // (compute_pair_gran_local.cpp:420-633) from the library's per-pair record.
  // Normal parts use the unit vector from i to j (x_j - x_i with the minimum
  // image), the contact point (x_i r_j + x_j r_i) / (r_i + r_j), the area the
  // sphere-sphere intersection circle's.
  std::vector<std::vector<double>> local_rows(const LocalCompute &lc) {
    if (lc.wall) return wall_local_rows(lc);
    int64_t np = 0;
    int nc = 0;
    if (dem_compute_pair_local(h, 0, nullptr, nullptr, &np, &nc)) die(dem_last_error(h));
    std::vector<int> tg(2 * std::max<int64_t>(np, 1));
    std::vector<double> rec((size_t)nc * std::max<int64_t>(np, 1));
    if (np > 0 && dem_compute_pair_local(h, np, tg.data(), rec.data(), &np, &nc)) die(dem_last_error(h));
    const int dnum = pair_dnum();
    double L[3];
    for (int d = 0; d < 3; d++) L[d] = hi[d] - lo[d];
    std::vector<std::vector<double>> rows((size_t)np);
    for (int64_t q = 0; q < np; q++) {
      const double *rc = &rec[(size_t)q * nc];
      const double *xi = rc, *xj = rc + 3, *vi = rc + 6, *vj = rc + 9, *fc = rc + 12, *tc = rc + 15;
      const double ri = rc[19], rj = rc[20];
      double del[3];  // x_i - x_j, minimum image
      for (int d = 0; d < 3; d++) {
        del[d] = xi[d] - xj[d];
        if (face[d][0] == 'p') {
          if (del[d] > 0.5 * L[d]) del[d] -= L[d];
          else if (del[d] < -0.5 * L[d]) del[d] += L[d];
        }
      }
      const double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
      const double dist = std::sqrt(rsq);
      const double nrm[3] = {-del[0] / dist, -del[1] / dist, -del[2] / dist};
      std::vector<double> &row = rows[(size_t)q];
      for (const std::string &k : lc.keys) {
        if (k == "pos") row.insert(row.end(), {xi[0], xi[1], xi[2], xj[0], xj[1], xj[2]});
        else if (k == "vel") row.insert(row.end(), {vi[0], vi[1], vi[2], vj[0], vj[1], vj[2]});
        else if (k == "id") row.insert(row.end(), {(double)tg[2 * q], (double)tg[2 * q + 1], rc[18]});
        else if (k == "force") row.insert(row.end(), {fc[0], fc[1], fc[2]});
        else if (k == "torque") row.insert(row.end(), {tc[0], tc[1], tc[2]});
        else if (k == "force_normal" || k == "force_tangential" || k == "torque_normal" ||
                 k == "torque_tangential") {
          const double *v = k[0] == 'f' ? fc : tc;
          const double mag = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
          const bool normal = k.find("normal") != std::string::npos;
          for (int d = 0; d < 3; d++) row.push_back(normal ? nrm[d] * mag : v[d] - nrm[d] * mag);
        } else if (k == "history") {
          // the record holds the shear (3) and, with epsd2, the rolling torque (3)
          const int off = pair_m.tangential_model == 0 ? 0 : 3;
          for (int d = 0; d < dnum; d++) row.push_back(rc[DEM_PAIR_LOCAL_FIXED_COLS + off + d]);
        } else if (k == "contactArea") {
          row.push_back(-3.14159265358979323846 / 4 * ((dist - ri - rj) * (dist + ri - rj) * (dist - ri + rj) * (dist + ri + rj)) / rsq);
        } else if (k == "delta") {
          row.push_back(ri + rj - dist);
        } else if (k == "contactPoint") {
          for (int d = 0; d < 3; d++) row.push_back((xi[d] * rj + (xi[d] - del[d]) * ri) / (ri + rj));
        }
      }
    }
    return rows;
  }

  // The rows of a wall/gran/local compute: ComputePairGranLocal::add_wall_1 /
  // add_wall_2 (compute_pair_gran_local.cpp:648-870) from the library's
  // per-contact record, for the meshes of the compute's fix. Normal parts use
  // the unit vector from the particle's centre to the contact point; the
  // reference writes no contactPoint columns for walls (zeros here).
  // Deviation: rows come particle by particle, the reference's triangle by
  // triangle.
  std::vector<std::vector<double>> wall_local_rows(const LocalCompute &lc) {
    int64_t nr = 0;
    int nc = 0;
    if (dem_compute_wall_local(h, 0, nullptr, nullptr, &nr, &nc)) die(dem_last_error(h));
    std::vector<int> ids(3 * std::max<int64_t>(nr, 1));
    std::vector<double> rec((size_t)nc * std::max<int64_t>(nr, 1));
    if (nr > 0 && dem_compute_wall_local(h, nr, ids.data(), rec.data(), &nr, &nc)) die(dem_last_error(h));
    // library mesh index -> index among the compute's fix's meshes, or -1
    std::vector<int> fix_index(mesh_order.size(), -1);
    for (size_t m = 0; m < mesh_order.size(); m++) {
      const auto &mf = mesh_fix[mesh_order[m]];
      if (mf.first == lc.wall_fix) fix_index[m] = mf.second;
    }
    std::vector<std::vector<double>> rows;
    for (int64_t q = 0; q < nr; q++) {
      const int m = ids[3 * q];
      if (m < 0 || m >= (int)fix_index.size() || fix_index[m] < 0) continue;
      const double *rc = &rec[(size_t)q * nc];
      const double *cp = rc, *x = rc + 3, *vw = rc + 6, *v = rc + 9, *fc = rc + 12, *tc = rc + 15, *nrm = rc + 18;
      const double rsq = rc[21], rad = rc[22];
      rows.emplace_back();
      std::vector<double> &row = rows.back();
      for (const std::string &k : lc.keys) {
        if (k == "pos") row.insert(row.end(), {cp[0], cp[1], cp[2], x[0], x[1], x[2]});
        else if (k == "vel") row.insert(row.end(), {vw[0], vw[1], vw[2], v[0], v[1], v[2]});
        else if (k == "id") row.insert(row.end(), {(double)fix_index[m], (double)ids[3 * q + 1], (double)ids[3 * q + 2]});
        else if (k == "force") row.insert(row.end(), {fc[0], fc[1], fc[2]});
        else if (k == "torque") row.insert(row.end(), {tc[0], tc[1], tc[2]});
        else if (k == "force_normal" || k == "force_tangential" || k == "torque_normal" ||
                 k == "torque_tangential") {
          const double *w = k[0] == 'f' ? fc : tc;
          const double mag = w[0] * nrm[0] + w[1] * nrm[1] + w[2] * nrm[2];
          const bool normal = k.find("normal") != std::string::npos;
          for (int d = 0; d < 3; d++) row.push_back(normal ? nrm[d] * mag : w[d] - nrm[d] * mag);
        } else if (k == "history") {
          for (int d = 0; d < lc.dnum; d++) row.push_back(rc[DEM_WALL_LOCAL_FIXED_COLS + d]);
        } else if (k == "contactArea") {
          row.push_back((rad * rad - rsq) * 3.14159265358979323846);
        } else if (k == "delta") {
          row.push_back(rad - std::sqrt(rsq));
        } else if (k == "contactPoint") {
          row.insert(row.end(), {0.0, 0.0, 0.0});
        }
      }
    }
    return rows;
  }

  // one snapshot of a dump local (DumpLocal::header_item / write_lines,
  // dump_local.cpp:239-254, 350-363: "%g " per value)
  void write_local_dump(Dump &d, const char *bstr, const double *blo, const double *bhi) {
    std::map<std::string, std::vector<std::vector<double>>> data;
    size_t nrows = 0;
    bool first = true;
    for (const std::string &c : d.cols) {
      if (c == "index") continue;
      const std::string id = c.substr(2, c.find('[') - 2);
      if (!data.count(id)) {
        if (h && !tag.empty()) data[id] = local_rows(local_computes[id]);
        else data[id] = {};
        if (!first && data[id].size() != nrows) die("Dump local count is not consistent across input fields");
        nrows = data[id].size();
        first = false;
      }
    }
    FILE *fp = d.fp;
    if (!fp) {
      std::string name = d.file;
      name.replace(name.find('*'), 1, std::to_string((long long)step));
      fp = fopen(name.c_str(), "w");
      if (!fp) die("Cannot open dump file " + name);
    }
    std::string cols;
    for (const std::string &c : d.cols) cols += c + " ";
    fprintf(fp, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ENTRIES\n%zu\nITEM: BOX BOUNDS %s\n", (long long)step,
            nrows, bstr);
    for (int k = 0; k < 3; k++) fprintf(fp, "%g %g\n", blo[k], bhi[k]);
    fprintf(fp, "ITEM: ENTRIES %s\n", cols.c_str());
    for (size_t q = 0; q < nrows; q++) {
      for (const std::string &c : d.cols) {
        if (c == "index") {
          fprintf(fp, "%d ", (int)q + 1);
          continue;
        }
        const size_t lb = c.find('[');
        const std::string id = c.substr(2, lb - 2);
        const int k1 = std::atoi(c.substr(lb + 1, c.size() - lb - 2).c_str());
        fprintf(fp, "%g ", data[id][q][(size_t)k1 - 1]);
      }
      fprintf(fp, "\n");
    }
    if (d.fp) fflush(d.fp);
    else fclose(fp);
  }

  void dump_command(const std::vector<std::string> &t) {
    // dump id group custom N file col1 col2 ...
    if (t.size() < 6) die("Illegal dump command");
    if (t[3] == "mesh/stl") {
      if (sq) die("dump mesh/stl for superquadrics is not available on this path");
      dump_stl_command(t);
      return;
    }
    if (t[3] == "local") {
      dump_local_command(t);
      return;
    }
    if (t[3] != "custom" && t[3] != "custom/vtk")
      die("dump style " + t[3] + " is not available on this path (custom, custom/vtk, local, mesh/stl)");
    if (t[2] != "all") die("dump: only group all is available on this path");
    if (sq) die("dump " + t[3] + " for superquadrics is not available on this path");
    if (dumps.count(t[1])) die("Reuse of dump ID " + t[1]);
    Dump d;
    d.vtk = t[3] == "custom/vtk";
    d.every = std::stoll(t[4]);
    if (d.every <= 0) die("Illegal dump command");
    d.file = t[5];
    for (size_t k = 6; k < t.size(); k++) {
      const bool image = t[k] == "ix" || t[k] == "iy" || t[k] == "iz";
      if (!dump_column_known(t[k]) && !(d.vtk && image))
        die("Invalid attribute " + t[k] + " in dump " + t[3] + " command");
      d.cols.push_back(t[k]);
    }
    if (d.cols.empty()) die("Illegal dump command");
    if (d.vtk) {
      // DumpVTK's file formats: the legacy .vtk is written here; .vtp / .vtu
      // and their parallel forms go through the VTK library's XML writers
      if (d.file.size() < 4 || d.file.compare(d.file.size() - 4, 4, ".vtk") != 0)
        die("dump custom/vtk: only the legacy .vtk file format is available on this path");
      dumps[t[1]] = d;
      return;
    }
    if (d.file.find('*') == std::string::npos) {
      d.fp = fopen(d.file.c_str(), "w");
      if (!d.fp) die("Cannot open dump file " + d.file);
    }
    dumps[t[1]] = d;
  }

  // dump custom/vtk (DumpCustomVTK, dump_custom_vtk.cpp:228-322, with
  // DumpParticle's arrays, dump_particle.cpp:1372-1620, 1735-1800): a legacy VTK
  // polydata file per snapshot, the atoms as points with one vertex cell each
  // and the dump's columns as point-data arrays, in the order of DumpParticle's
  // attribute list (not the order of the command). x, y, z are the points; three
  // components of one vector (ix iy iz, vx vy vz, fx fy fz, omegax omegay omegaz,
  // tqx tqy tqz) become one 3-component array named by the text before its 'x'
  // ("i", "v", "f", "omega", "tq"); ids, types and image counts are int arrays,
  // the rest double. The text follows the legacy VTK format as
  // vtkPolyDataWriter writes it in ASCII (header "Generated by LIGGGHTS",
  // values "%.11lg " nine to a line). Beside each snapshot the simulation box is
  // written as a 2x2x2 rectilinear grid (DumpCustomVTK::setFileCurrent:
  // <prefix>boundingBox_<step>.vtk for a '*' file name, <name>_boundingBox.vtk
  // otherwise). Atoms are in ID order, as dump_modify sort id would give.
  void write_vtk_dump(Dump &d, const std::vector<size_t> &order, const std::vector<double> &f,
                      const std::vector<double> &tq, const double *blo, const double *bhi) {
    static const char *attrs[] = {"id", "type", "mass", "ix", "iy", "iz", "vx", "vy", "vz", "fx", "fy", "fz",
                                  "radius", "diameter", "omegax", "omegay", "omegaz", "tqx", "tqy", "tqz"};
    std::set<std::string> have(d.cols.begin(), d.cols.end());
    if (have.count("ix") || have.count("iy") || have.count("iz"))
      for (int k = 0; k < 3; k++)
        if (periodic[k]) die("dump custom/vtk: image flags of periodic boundaries are not available on this path");
    const size_t n = order.size();
    auto value = [&](const std::string &c, size_t i) -> double {
      if (c == "vx" || c == "vy" || c == "vz") return v[3 * i + (c[1] - 'x')];
      if (c == "fx" || c == "fy" || c == "fz") return f[3 * i + (c[1] - 'x')];
      if (c == "omegax" || c == "omegay" || c == "omegaz") return w[3 * i + (c[5] - 'x')];
      if (c == "tqx" || c == "tqy" || c == "tqz") return tq[3 * i + (c[2] - 'x')];
      if (c == "radius") return r[i];
      if (c == "diameter") return 2.0 * r[i];
      if (c == "mass") return m[i];
      return 0.0;  // image counts: no periodic boundary, so always 0
    };
    struct Arr {
      std::string name;
      std::vector<std::string> comps;
      bool integer;
    };
    std::vector<Arr> arrs;
    const int na = (int)(sizeof attrs / sizeof attrs[0]);
    for (int a = 0; a < na; a++) {
      const std::string c = attrs[a];
      if (!have.count(c)) continue;
      const bool vec_start = c.back() == 'x' && c != "id" && a + 2 < na && have.count(attrs[a + 1]) &&
                             have.count(attrs[a + 2]) && std::string(attrs[a + 1]).back() == 'y';
      const bool integer = c == "id" || c == "type" || c[0] == 'i';
      if (vec_start) {
        arrs.push_back({c.substr(0, c.find('x')), {c, attrs[a + 1], attrs[a + 2]}, integer});
        a += 2;
      } else {
        arrs.push_back({c, {c}, integer});
      }
    }
    std::string name = d.file, bbox;
    const size_t star = name.find('*');
    if (star != std::string::npos) {
      const std::string st = std::to_string((long long)step);
      bbox = name.substr(0, star) + "boundingBox_" + st + name.substr(star + 1);
      name.replace(star, 1, st);
    } else {
      bbox = name.substr(0, name.size() - 4) + "_boundingBox.vtk";
    }
    FILE *fp = fopen(name.c_str(), "w");
    if (!fp) die("Cannot open dump file " + name);
    fprintf(fp, "# vtk DataFile Version 3.0\nGenerated by LIGGGHTS\nASCII\nDATASET POLYDATA\nPOINTS %zu double\n", n);
    size_t col = 0;
    auto put = [&](double val, bool integer) {
      if (integer) fprintf(fp, "%d ", (int)val);
      else fprintf(fp, "%.11lg ", val);
      if (++col % 9 == 0) fprintf(fp, "\n");
    };
    auto end_block = [&]() {
      if (col % 9 != 0) fprintf(fp, "\n");
      col = 0;
    };
    for (size_t oi = 0; oi < n; oi++)
      for (int k = 0; k < 3; k++) put(x[3 * order[oi] + k], false);
    end_block();
    fprintf(fp, "VERTICES %zu %zu\n", n, 2 * n);
    for (size_t oi = 0; oi < n; oi++) fprintf(fp, "1 %zu\n", oi);
    fprintf(fp, "\nPOINT_DATA %zu\nFIELD FieldData %zu\n", n, arrs.size());
    for (const Arr &A : arrs) {
      fprintf(fp, "%s %zu %zu %s\n", A.name.c_str(), A.comps.size(), n, A.integer ? "int" : "double");
      for (size_t oi = 0; oi < n; oi++) {
        const size_t i = order[oi];
        for (const std::string &c : A.comps) {
          if (c == "id") put(tag[i], true);
          else if (c == "type") put(type[i], true);
          else put(value(c, i), A.integer);
        }
      }
      end_block();
    }
    fclose(fp);
    FILE *fb = fopen(bbox.c_str(), "w");
    if (!fb) die("Cannot open dump file " + bbox);
    fprintf(fb, "# vtk DataFile Version 3.0\nGenerated by LIGGGHTS\nASCII\nDATASET RECTILINEAR_GRID\n"
                "DIMENSIONS 2 2 2\n");
    for (int k = 0; k < 3; k++)
      fprintf(fb, "%c_COORDINATES 2 double\n%.11lg %.11lg \n", 'X' + k, blo[k], bhi[k]);
    fclose(fb);
  }

  // one snapshot in the reference's text layout (DumpCustom::header_item,
  // dump_custom.cpp:378-393, write_lines: "%d " / "%g " per column), atoms in
  // ID order (dump_modify sort id)
  void write_dumps(bool force_all) {
    bool any = false;
    for (auto &kv : dumps)
      if (kv.second.last != step && (force_all || step % kv.second.every == 0)) any = true;
    if (!any) return;
    const size_t n = tag.size();
    std::vector<double> f(3 * n), tq(3 * n);
    // an empty system (its particles still to be inserted) dumps no atoms
    if (n > 0 && dem_get_state(h, x.data(), v.data(), w.data(), f.data(), tq.data())) die(dem_last_error(h));
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return tag[a] < tag[b]; });
    char bstr[9];
    for (int d = 0; d < 3; d++) {
      bstr[3 * d] = face[d][0];
      bstr[3 * d + 1] = face[d][1];
    }
    bstr[2] = bstr[5] = ' ';
    bstr[8] = '\0';
    double blo[3], bhi[3];
    current_box(blo, bhi);
    for (auto &kv : dumps) {
      Dump &d = kv.second;
      if (d.last == step || !(force_all || step % d.every == 0)) continue;
      if (d.stl) {
        write_stl(d);
        d.last = step;
        continue;
      }
      if (d.local) {
        write_local_dump(d, bstr, blo, bhi);
        d.last = step;
        continue;
      }
      if (d.vtk) {
        write_vtk_dump(d, order, f, tq, blo, bhi);
        d.last = step;
        continue;
      }
      FILE *fp = d.fp;
      if (!fp) {
        std::string name = d.file;
        name.replace(name.find('*'), 1, std::to_string((long long)step));
        fp = fopen(name.c_str(), "w");
        if (!fp) die("Cannot open dump file " + name);
      }
      std::string cols;
      for (size_t c = 0; c < d.cols.size(); c++) cols += (c ? " " : "") + d.cols[c];
      fprintf(fp, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ATOMS\n%zu\nITEM: BOX BOUNDS %s\n", (long long)step, n,
              bstr);
      for (int k = 0; k < 3; k++) fprintf(fp, "%g %g\n", blo[k], bhi[k]);
      fprintf(fp, "ITEM: ATOMS %s\n", cols.c_str());
      for (size_t oi = 0; oi < n; oi++) {
        const size_t i = order[oi];
        for (const std::string &c : d.cols) {
          if (c == "id") fprintf(fp, "%d ", tag[i]);
          else if (c == "type") fprintf(fp, "%d ", type[i]);
          else {
            double val = 0.0;
            if (c == "x" || c == "y" || c == "z") val = x[3 * i + (c[0] - 'x')];
            else if (c == "vx" || c == "vy" || c == "vz") val = v[3 * i + (c[1] - 'x')];
            else if (c == "fx" || c == "fy" || c == "fz") val = f[3 * i + (c[1] - 'x')];
            else if (c == "omegax" || c == "omegay" || c == "omegaz") val = w[3 * i + (c[5] - 'x')];
            else if (c == "tqx" || c == "tqy" || c == "tqz") val = tq[3 * i + (c[2] - 'x')];
            else if (c == "radius") val = r[i];
            else if (c == "diameter") val = 2.0 * r[i];
            else if (c == "mass") val = m[i];
            fprintf(fp, "%g ", val);
          }
        }
        fprintf(fp, "\n");
      }
      if (d.fp) fflush(d.fp);
      else fclose(fp);
      d.last = step;
    }
  }

  // the next step at which a thermo line or a dump is due, at most `limit`
  long long next_output(long long limit) const {
    long long k = limit;
    if (thermo_every > 0) k = std::min<long long>(k, thermo_every - step % thermo_every);
    for (const auto &kv : dumps) k = std::min<long long>(k, kv.second.every - step % kv.second.every);
    return k;
  }

  std::string substitute(const std::string &line) {
    std::string outl;
    for (size_t k = 0; k < line.size(); k++) {
      if (line[k] == '$' && k + 1 < line.size()) {
        std::string name;
        size_t adv;
        if (line[k + 1] == '{') {
          const size_t close = line.find('}', k + 2);
          if (close == std::string::npos) die("Invalid variable name in " + line);
          name = line.substr(k + 2, close - k - 2);
          adv = close;
        } else {
          name = line.substr(k + 1, 1);
          adv = k + 1;
        }
        auto it = vars.find(name);
        if (it == vars.end()) die("Substitution for illegal variable " + name);
        if (var_style[name] == "equal") {  // Variable::retrieve: the formula's value, %.15g
          char buf[64];
          snprintf(buf, sizeof buf, "%.15g", eval_variable(name));
          outl += buf;
        } else {
          outl += it->second;
        }
        k = adv;
      } else {
        outl += line[k];
      }
    }
    return outl;
  }

  // words of a command; text in double quotes is one word, without the quotes
  // (Input::parse, input.cpp:440-520)
  static std::vector<std::string> split(const std::string &s) {
    std::vector<std::string> t;
    size_t k = 0;
    while (k < s.size()) {
      while (k < s.size() && std::isspace((unsigned char)s[k])) k++;
      if (k >= s.size()) break;
      if (s[k] == '"') {
        const size_t e = s.find('"', k + 1);
        if (e == std::string::npos) die("Unbalanced quotes in input line");
        t.push_back(s.substr(k + 1, e - k - 1));
        k = e + 1;
      } else {
        const size_t b = k;
        while (k < s.size() && !std::isspace((unsigned char)s[k])) k++;
        t.push_back(s.substr(b, k - b));
      }
    }
    return t;
  }

  // ---- equal-style variables (Variable::evaluate, variable.cpp:800-1620) ----
  // Numbers, + - * / % ^, unary - and !, comparisons, && ||, parentheses,
  // the math functions of Variable::math_function (variable.cpp:2481-2700),
  // v_name, thermo keywords and c_ID of the computes thermo knows; operator
  // precedence as variable.cpp:132-140 (unary above ^ above * / % above + -),
  // every binary operator left-associative.
  struct Expr {
    Interp &in;
    const std::string &s;
    size_t k = 0;
    int depth;
    Expr(Interp &i, const std::string &str, int d) : in(i), s(str), depth(d) {}
    void ws() {
      while (k < s.size() && std::isspace((unsigned char)s[k])) k++;
    }
    [[noreturn]] void bad(const std::string &m) { die("Invalid syntax in variable formula '" + s + "': " + m); }
    static int prec(const std::string &op) {
      if (op == "||") return 1;
      if (op == "&&") return 2;
      if (op == "==" || op == "!=") return 3;
      if (op == "<" || op == "<=" || op == ">" || op == ">=") return 4;
      if (op == "+" || op == "-") return 5;
      if (op == "*" || op == "/" || op == "%") return 6;
      if (op == "^") return 7;
      return -1;
    }
    std::string peek_op() {
      ws();
      if (k >= s.size()) return "";
      static const char *two[] = {"||", "&&", "==", "!=", "<=", ">="};
      for (const char *o : two)
        if (s.compare(k, 2, o) == 0) return o;
      const char c = s[k];
      if (std::strchr("+-*/%^<>", c)) return std::string(1, c);
      return "";
    }
    double binary(int minp) {
      double a = unary();
      for (;;) {
        const std::string op = peek_op();
        const int p = op.empty() ? -1 : prec(op);
        if (p < minp) return a;
        k += op.size();
        const double b = binary(p + 1);
        if (op == "+") a = a + b;
        else if (op == "-") a = a - b;
        else if (op == "*") a = a * b;
        else if (op == "/") {
          if (b == 0.0) die("Divide by 0 in variable formula");
          a = a / b;
        } else if (op == "%") {
          if (b == 0.0) die("Modulo 0 in variable formula");
          a = std::fmod(a, b);
        } else if (op == "^") {
          if (b == 0.0 && a == 0.0) die("Power by 0 in variable formula");
          a = std::pow(a, b);
        } else if (op == "<") a = a < b;
        else if (op == "<=") a = a <= b;
        else if (op == ">") a = a > b;
        else if (op == ">=") a = a >= b;
        else if (op == "==") a = a == b;
        else if (op == "!=") a = a != b;
        else if (op == "&&") a = (a != 0.0 && b != 0.0);
        else a = (a != 0.0 || b != 0.0);
      }
    }
    double unary() {
      ws();
      if (k < s.size() && s[k] == '-') {
        k++;
        return -unary();
      }
      if (k < s.size() && s[k] == '!' && (k + 1 >= s.size() || s[k + 1] != '=')) {
        k++;
        return unary() == 0.0 ? 1.0 : 0.0;
      }
      return primary();
    }
    double primary() {
      ws();
      if (k >= s.size()) bad("unexpected end");
      if (s[k] == '(') {
        k++;
        const double v = binary(0);
        ws();
        if (k >= s.size() || s[k] != ')') bad("missing )");
        k++;
        return v;
      }
      if (std::isdigit((unsigned char)s[k]) || s[k] == '.') {
        char *end = nullptr;
        const double v = std::strtod(s.c_str() + k, &end);
        k = end - s.c_str();
        return v;
      }
      size_t b = k;
      while (k < s.size() && (std::isalnum((unsigned char)s[k]) || s[k] == '_' || s[k] == '[' || s[k] == ']'))
        k++;
      const std::string word = s.substr(b, k - b);
      if (word.empty()) bad("unexpected '" + s.substr(k, 1) + "'");
      ws();
      if (k < s.size() && s[k] == '(') {  // math function
        k++;
        std::vector<double> args;
        ws();
        if (k < s.size() && s[k] == ')') k++;
        else
          for (;;) {
            args.push_back(binary(0));
            ws();
            if (k < s.size() && s[k] == ',') {
              k++;
              continue;
            }
            if (k < s.size() && s[k] == ')') {
              k++;
              break;
            }
            bad("missing ) after arguments of " + word);
          }
        return func(word, args);
      }
      if (word.compare(0, 2, "v_") == 0) return in.eval_variable(word.substr(2), depth + 1);
      double val = 0.0;
      if (in.formula_keyword(word, val)) return val;
      bad("unknown word '" + word + "'");
    }
    double func(const std::string &f, const std::vector<double> &a) {
      auto one = [&]() {
        if (a.size() != 1) bad("function " + f + " takes one argument");
        return a[0];
      };
      if (f == "sqrt") {
        if (one() < 0.0) die("Sqrt of negative value in variable formula");
        return std::sqrt(a[0]);
      }
      if (f == "exp") return std::exp(one());
      if (f == "ln") {
        if (one() <= 0.0) die("Log of zero/negative value in variable formula");
        return std::log(a[0]);
      }
      if (f == "log") {
        if (one() <= 0.0) die("Log of zero/negative value in variable formula");
        return std::log10(a[0]);
      }
      if (f == "abs") return std::fabs(one());
      if (f == "sin") return std::sin(one());
      if (f == "cos") return std::cos(one());
      if (f == "tan") return std::tan(one());
      if (f == "asin") {
        if (one() < -1.0 || a[0] > 1.0) die("Arcsin of invalid value in variable formula");
        return std::asin(a[0]);
      }
      if (f == "acos") {
        if (one() < -1.0 || a[0] > 1.0) die("Arccos of invalid value in variable formula");
        return std::acos(a[0]);
      }
      if (f == "atan") return std::atan(one());
      if (f == "atan2") {
        if (a.size() != 2) bad("atan2 takes two arguments");
        return std::atan2(a[0], a[1]);
      }
      if (f == "ceil") return std::ceil(one());
      if (f == "floor") return std::floor(one());
      if (f == "round") return std::round(one());
      die("function " + f + " in a variable formula is not available on this path");
    }
  };

  // the value of an equal-style variable (or of an index/string one holding a number)
  double eval_variable(const std::string &name, int depth = 0) {
    if (depth > 50) die("Variable formula refers to itself: " + name);
    auto it = vars.find(name);
    if (it == vars.end()) die("Invalid variable name in variable formula: " + name);
    Expr e(*this, it->second, depth);
    const double v = e.binary(0);
    e.ws();
    if (e.k != it->second.size()) e.bad("trailing text");
    return v;
  }

  // thermo keywords and computes a formula may read (Variable::evaluate's
  // thermo-keyword branch, variable.cpp:1380-1410)
  bool formula_keyword(const std::string &word, double &val) {
    if (word == "step") {
      val = (double)step;
      return true;
    }
    if (word == "atoms") {
      val = (double)tag.size();
      return true;
    }
    if (word == "dt") {
      val = dt;
      return true;
    }
    if (word == "PI") {
      val = kMyPi;
      return true;
    }
    double ke = 0.0;
    if (word == "ke" || word.compare(0, 2, "c_") == 0) {
      if (h && !tag.empty()) pull_state();
      for (size_t i = 0; i < tag.size(); i++)
        ke += 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
      if (word == "ke") {
        val = ke;
        return true;
      }
    }
    return thermo_value(word, ke, val);
  }

  void need_box(const char *cmd) {
    if (!have_box) die(std::string(cmd) + " command before simulation box is defined");
  }

  void read_data(const std::string &path) {
    std::ifstream f(path);
    if (!f) die("Cannot open file " + path);
    std::string line;
    std::getline(f, line);  // title line
    int64_t natoms = -1;
    bool box_given[3] = {false, false, false};
    double blo[3], bhi[3];
    std::string section;
    // header
    while (std::getline(f, line)) {
      const size_t hash = line.find('#');
      if (hash != std::string::npos) line = line.substr(0, hash);
      auto t = split(line);
      if (t.empty()) continue;
      if (t.size() >= 2 && t[1] == "atoms") {
        natoms = std::stoll(t[0]);
      } else if (t.size() >= 3 && t[1] == "atom" && t[2] == "types") {
        const int nt = inum(t[0]);
        if (!have_box) ntypes = nt;
        else if (nt > ntypes) die("Data file atom types exceed create_box types");
      } else if (t.size() >= 4 && (t[2] == "xlo" || t[2] == "ylo" || t[2] == "zlo")) {
        const int d = t[2][0] - 'x';
        blo[d] = num(t[0]);
        bhi[d] = num(t[1]);
        box_given[d] = true;
      } else if (t[0] == "Atoms" || t[0] == "Velocities") {
        section = t[0];
        break;
      }
    }
    if (natoms < 0) die("read_data: no atom count in header");
    if (!have_box) {
      for (int d = 0; d < 3; d++) {
        if (!box_given[d]) die("read_data: box bounds missing");
        lo[d] = blo[d];
        hi[d] = bhi[d];
      }
      have_box = true;
    }
    const size_t base = tag.size();
    std::map<int, size_t> index;
    for (size_t k = 0; k < base; k++) index[tag[k]] = k;
    while (!section.empty()) {
      std::string next;
      int64_t got = 0;
      while (got < natoms && std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        auto t = split(line);
        if (t.empty()) continue;
        if (section == "Atoms" && sq) {
          // AtomVecSuperquadric::data_atom (atom_vec_superquadric.cpp:1150-1212):
          // id type density shapex shapey shapez blockiness1 blockiness2 x y z
          if (t.size() < 11) die("Incorrect format of Atoms section");
          const int id = inum(t[0]);
          const int ty = inum(t[1]);
          if (id <= 0) die("Invalid atom ID in Atoms section of data file");
          if (ty <= 0 || ty > ntypes) die("Invalid atom type in Atoms section of data file");
          const double dens = num(t[2]);
          if (dens <= 0.0) die("Invalid density in Atoms section of data file");
          for (int d = 0; d < 3; d++)
            if (num(t[3 + d]) <= 0.0) die("Invalid shape in Atoms section of data file");
          for (int d = 0; d < 2; d++)
            if (num(t[6 + d]) < 2.0) die("Invalid blockiness in Atoms section of data file");
          index[id] = tag.size();
          tag.push_back(id);
          type.push_back(ty);
          r.push_back(0.0);
          m.push_back(0.0);
          sq_dens.push_back(dens);
          for (int d = 0; d < 3; d++) sq_shape.push_back(num(t[3 + d]));
          for (int d = 0; d < 2; d++) sq_blk.push_back(num(t[6 + d]));
          const double qid[4] = {1.0, 0.0, 0.0, 0.0};  // quatIdentity4D
          for (int d = 0; d < 4; d++) sq_quat.push_back(qid[d]);
          for (int d = 0; d < 3; d++) x.push_back(num(t[8 + d]));
          for (int d = 0; d < 3; d++) {
            v.push_back(0.0);
            w.push_back(0.0);
          }
        } else if (section == "Atoms") {
          // atom_vec_sphere.cpp:1055-1100: id type diameter density x y z
          if (t.size() < 7) die("Incorrect format of Atoms section");
          const int id = inum(t[0]);
          const int ty = inum(t[1]);
          if (id <= 0) die("Invalid atom ID in Atoms section of data file");
          if (ty <= 0 || ty > ntypes) die("Invalid atom type in Atoms section of data file");
          const double rad = 0.5 * num(t[2]);
          const double dens = num(t[3]);
          if (rad < 0.0) die("Invalid radius in Atoms section of data file");
          if (dens <= 0.0) die("Invalid density in Atoms section of data file");
          const double mass = rad == 0.0 ? dens : 4.0 * kMyPi / 3.0 * rad * rad * rad * dens;
          index[id] = tag.size();
          tag.push_back(id);
          type.push_back(ty);
          r.push_back(rad);
          m.push_back(mass);
          for (int d = 0; d < 3; d++) x.push_back(num(t[4 + d]));
          for (int d = 0; d < 3; d++) {
            v.push_back(0.0);
            w.push_back(0.0);
          }
        } else {
          // atom_vec_sphere.cpp:1129-1140: id vx vy vz wx wy wz
          if (t.size() < 7) die("Incorrect format of Velocities section");
          const int id = inum(t[0]);
          auto it = index.find(id);
          if (it == index.end()) die("Invalid atom ID in Velocities section of data file");
          for (int d = 0; d < 3; d++) {
            v[3 * it->second + d] = num(t[1 + d]);
            w[3 * it->second + d] = num(t[4 + d]);
          }
          if (sq) {
            // AtomVecSuperquadric::data_vel sets omega but not the angular
            // momentum fix nve/superquadric integrates (atom_vec_superquadric.cpp:
            // 1231-1239): the particles start from zero angular momentum
            if (!omega_warned && (w[3 * it->second] != 0.0 || w[3 * it->second + 1] != 0.0 ||
                                  w[3 * it->second + 2] != 0.0)) {
              fprintf(out, "WARNING: Velocities omega of superquadrics does not set their angular "
                           "momentum; they start without rotation\n");
              omega_warned = true;
            }
            for (int d = 0; d < 3; d++) w[3 * it->second + d] = 0.0;
          }
        }
        got++;
      }
      if (got < natoms) die("Unexpected end of data file");
      section.clear();
      while (std::getline(f, line)) {
        auto t = split(line);
        if (t.empty()) continue;
        if (t[0] == "Atoms" || t[0] == "Velocities") {
          section = t[0];
          break;
        }
      }
    }
    fprintf(out, "  %zu atoms\n", tag.size() - base);
    dirty = true;
  }

  void fix_property_global(const std::vector<std::string> &t) {
    // fix id group property/global <name> peratomtype v1..vn
    // fix id group property/global <name> peratomtypepair n v11 v12 ... vnn
    // (fix_property_global.cpp:85-102)
    if (t.size() < 7) die("Illegal fix property/global command");
    const std::string &name = t[4];
    const std::string &style = t[5];
    std::vector<double> vals;
    size_t k = 6;
    if (style == "peratomtypepair") {
      const int n = inum(t[6]);
      if (n < 1) die("Illegal fix property/global command");
      k = 7;
      if (t.size() - k != (size_t)n * n) die("fix property/global " + name + ": expected n*n values");
    } else if (style != "peratomtype" && style != "scalar" && style != "vector") {
      die("Unknown fix property/global style " + style);
    }
    for (; k < t.size(); k++) vals.push_back(num(t[k]));
    if (name == "youngsModulus") Y = vals;
    else if (name == "poissonsRatio") nu = vals;
    else if (name == "coefficientRestitution") e = vals;
    else if (name == "coefficientFriction") mu = vals;
    else if (name == "coefficientRollingFriction") mur = vals;
    else if (name == "kn" || name == "kt" || name == "gamman" || name == "gammat" ||
             name == "gamman_abs" || name == "gammat_abs" || name == "cohesionEnergyDensity" ||
             name == "characteristicVelocity")
      model_props[name] = vals;
    else fprintf(out, "WARNING: property/global %s is not used by this path\n", name.c_str());
    dirty = true;
  }

  void fix_mesh_surface(const std::vector<std::string> &t) {
    // fix id group mesh/surface file <stl> type <t> [scale s] [move x y z]
    //     [rotate axis x y z angle a] [precision p] [verbose yes|no] [heal ...]
    // mesh/surface/stress also takes [reference_point x y z] [stress on|off]
    // (MeshModuleStress, mesh_module_stress.cpp:75-120)
    if (t.size() < 6 || (t[4] != "file" && t[4] != "generate"))
      die("fix mesh/surface: expecting keyword 'file', 'generate' or 'fix'");
    MeshDef md;
    md.stress = t[3] == "mesh/surface/stress";
    std::string err;
    size_t k0 = 6;
    if (t[4] == "file") {
      if (!dem::read_ascii_stl(t[5], md.nodes, err)) die(err);
    } else {
      // FixMesh::parse_and_consume_generator_args (fix_mesh.cpp:767-874): the
      // values in order, each > 0, and at least one more word after each
      const std::string &gt = t[5];
      std::vector<std::string> names;
      if (gt == "cube") names = {"size", "mask"};
      else if (gt == "box") names = {"xsize", "ysize", "zsize", "mask"};
      else if (gt == "cylinder") names = {"radius", "zsize", "nsegments", "mask"};
      else if (gt == "pipe") names = {"inner_radius", "outer_radius", "zsize", "nsegments", "mask"};
      else if (gt == "disk") names = {"radius", "nsegments"};
      else if (gt == "plane") names = {"size"};
      else die("fix mesh/surface: unknown generator type '" + gt + "'");
      double dv[3] = {0, 0, 0};
      int nd = 0, nseg = 0, mask = 0;
      size_t a = 6;
      for (const std::string &nm : names) {
        if (a + 1 >= t.size()) die("fix mesh/surface: missing generator argument: " + nm);
        const bool integer = nm == "mask" || nm == "nsegments";
        char *end = nullptr;
        const double val = integer ? (double)std::strtol(t[a].c_str(), &end, 10) : std::strtod(t[a].c_str(), &end);
        if (end == t[a].c_str() || *end != '\0')
          die("fix mesh/surface: generator argument '" + nm + "' is not a valid value");
        if (val <= 0) die("fix mesh/surface: generator argument '" + nm + "' must be > 0");
        if (nm == "mask") mask = (int)val;
        else if (nm == "nsegments") nseg = (int)val;
        else dv[nd++] = val;
        a++;
      }
      if (!dem::generate_mesh(gt, dv, nseg, mask, md.nodes, err)) die(err);
      fprintf(out, "%s mesh generated: %zu triangels\n", gt == "cube" ? "box" : gt.c_str(), md.nodes.size() / 9);
      k0 = a;
    }
    for (size_t k = k0; k < t.size();) {
      if (t[k] == "type" && k + 1 < t.size()) {
        md.type = inum(t[k + 1]);
        k += 2;
      } else if (t[k] == "scale" && k + 1 < t.size()) {
        const double f = num(t[k + 1]);
        for (double &v : md.nodes) v *= f;
        k += 2;
      } else if (t[k] == "move" && k + 3 < t.size()) {
        const double dv[3] = {num(t[k + 1]), num(t[k + 2]), num(t[k + 3])};
        for (size_t q = 0; q < md.nodes.size(); q++) md.nodes[q] += dv[q % 3];
        k += 4;
      } else if (t[k] == "rotate" && k + 6 < t.size() && t[k + 1] == "axis" && t[k + 5] == "angle") {
        double ax[3] = {num(t[k + 2]), num(t[k + 3]), num(t[k + 4])};
        const double ang = num(t[k + 6]) * kMyPi / 180.0;
        const double nrm = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
        if (nrm == 0.) die("fix mesh/surface: rotation axis is zero");
        for (double &a : ax) a /= nrm;
        // Rodrigues rotation about the axis through the origin
        const double c = std::cos(ang), sn = std::sin(ang);
        for (size_t q = 0; q + 2 < md.nodes.size(); q += 3) {
          const double v[3] = {md.nodes[q], md.nodes[q + 1], md.nodes[q + 2]};
          const double kv = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
          const double cr[3] = {ax[1] * v[2] - ax[2] * v[1], ax[2] * v[0] - ax[0] * v[2],
                                ax[0] * v[1] - ax[1] * v[0]};
          for (int d = 0; d < 3; d++) md.nodes[q + d] = v[d] * c + cr[d] * sn + ax[d] * kv * (1 - c);
        }
        k += 7;
      } else if (t[3] == "mesh/surface/stress" && t[k] == "reference_point" && k + 3 < t.size()) {
        for (int d = 0; d < 3; d++) md.pref[d] = num(t[k + 1 + d]);
        k += 4;
      } else if (t[3] == "mesh/surface/stress" && t[k] == "stress" && k + 1 < t.size()) {
        if (t[k + 1] != "on" && t[k + 1] != "off") die("expecting 'on' or 'off' as stress argument");
        md.stress = t[k + 1] == "on";
        k += 2;
      } else if (t[3] == "mesh/surface/stress" && t[k] == "wear" && k + 1 < t.size()) {
        if (t[k + 1] != "off") die("fix mesh/surface/stress: wear is not available on this path");
        k += 2;
      } else if ((t[k] == "precision" || t[k] == "verbose" || t[k] == "heal" ||
                  t[k] == "curvature" || t[k] == "curvature_tolerant") && k + 1 < t.size()) {
        if (t[k] == "precision" || t[k] == "curvature")
          fprintf(out, "WARNING: fix mesh/surface %s uses the default value on this path\n",
                  t[k].c_str());
        k += 2;
      } else {
        die("fix mesh/surface: unsupported keyword " + t[k]);
      }
    }
    if (!mesh_defs.count(t[1])) mesh_def_order.push_back(t[1]);
    mesh_defs[t[1]] = md;
    dirty = true;
  }

  void parse_model(const std::vector<std::string> &t, size_t k, size_t &end, bool pair,
                   dem_contact_model *wall_model = nullptr) {
    // model <normal> tangential <tangential> [cohesion sjkr] [rolling_friction cdt]
    // [settings on|off ...] (the keyword loop of PairGranProxy::settings and
    // Settings::parseArguments)
    bool saw_model = false, saw_tang = false;
    dem_contact_model mm{0, 1, 0, 0, DEM_NORMAL_HERTZ, 0, 0, 0, 0, 0};
    while (k < t.size()) {
      if (t[k] == "model" && k + 1 < t.size()) {
        const std::string &nm = t[k + 1];
        if (nm == "hertz") mm.normal_model = DEM_NORMAL_HERTZ;
        else if (nm == "hooke") mm.normal_model = DEM_NORMAL_HOOKE;
        else if (nm == "hertz/stiffness") mm.normal_model = DEM_NORMAL_HERTZ_STIFFNESS;
        else if (nm == "hooke/stiffness") mm.normal_model = DEM_NORMAL_HOOKE_STIFFNESS;
        else die("model " + nm + " is not available on this path (hertz, hooke, hertz/stiffness, hooke/stiffness)");
        saw_model = true;
        k += 2;
      } else if (t[k] == "tangential" && k + 1 < t.size()) {
        if (t[k + 1] == "history") mm.tangential_model = 0;
        else if (t[k + 1] == "no_history") mm.tangential_model = 1;
        else die("tangential " + t[k + 1] + " is not available on this path (history, no_history)");
        saw_tang = true;
        k += 2;
      } else if (t[k] == "rolling_friction" && k + 1 < t.size()) {
        if (t[k + 1] == "cdt") mm.rolling_cdt = 1;
        else if (t[k + 1] == "epsd2") mm.rolling_epsd2 = 1;
        else if (t[k + 1] != "off")
          die("Only 'rolling_friction cdt' and 'rolling_friction epsd2' are available, got " + t[k + 1]);
        k += 2;
      } else if (t[k] == "cohesion" && k + 1 < t.size()) {
        if (t[k + 1] == "sjkr") mm.cohesion_model = 1;
        else if (t[k + 1] != "off") die("cohesion " + t[k + 1] + " is not available on this path (sjkr)");
        k += 2;
      } else if (t[k] == "viscous" && k + 1 < t.size()) {
        if (t[k + 1] != "off") die("model hooke viscous on is not available on this path");
        k += 2;
      } else if ((t[k] == "ktToKnUser" || t[k] == "absolute_damping" || t[k] == "tangential_reduce") &&
                 k + 1 < t.size()) {
        const int on = t[k + 1] == "on" ? 1 : (t[k + 1] == "off" ? 0 : -1);
        if (on < 0) die("Expected on/off for " + t[k]);
        if (t[k] == "ktToKnUser") mm.kt_to_kn = on;
        else if (t[k] == "absolute_damping") mm.absolute_damping = on;
        else mm.tangential_reduce = on;
        k += 2;
      } else if (t[k] == "surface" && k + 1 < t.size()) {
        if (t[k + 1] == "superquadric") {
          if (!sq) die("surface superquadric requires atom_style superquadric");
          if (pair) pair_sq_surface = true;
        } else if (t[k + 1] != "default") {
          die("Only the default and superquadric surface models are available");
        } else if (sq) {
          die("atom_style superquadric needs 'surface superquadric'");
        }
        k += 2;
      } else if ((t[k] == "tangential_damping" || t[k] == "limitForce" || t[k] == "torsionTorque") &&
                 k + 1 < t.size()) {
        const int on = t[k + 1] == "on" ? 1 : (t[k + 1] == "off" ? 0 : -1);
        if (on < 0) die("Expected on/off for " + t[k]);
        if (t[k] == "tangential_damping") mm.tangential_damping = on;
        else if (t[k] == "limitForce") mm.limit_force = on;
        else mm.torsion_torque = on;
        k += 2;
      } else {
        break;
      }
    }
    if (!saw_model || !saw_tang) die("gran model needs 'model <normal> tangential <tangential>'");
    end = k;
    if (pair) pair_m = mm;
    else if (wall_model) *wall_model = mm;
  }

  // Set::command for quat and quat/random (set.cpp:345-365, 1041-1065,
  // 1195-1225); atom, type and group all selectors
  void set_command(const std::vector<std::string> &t) {
    if (t.size() < 4) die("Illegal set command");
    if (!sq) die("set: only quaternions of superquadrics are supported on this path");
    std::vector<size_t> sel;
    const std::string &style = t[1], &id = t[2];
    for (size_t i = 0; i < tag.size(); i++) {
      bool on = false;
      if (style == "atom") on = id == "*" || tag[i] == inum(id);
      else if (style == "type") on = id == "*" || type[i] == inum(id);
      else if (style == "group") {
        if (id != "all") die("set group: only group all is available on this path");
        on = true;
      } else die("set: style " + style + " is not supported on this path");
      if (on) sel.push_back(i);
    }
    const std::string &kw = t[3];
    if (kw == "quat") {
      if (t.size() < 8) die("Illegal set command");
      const double a = num(t[4]), b = num(t[5]), cc = num(t[6]), wv = num(t[7]);
      const double theta2 = 1.57079632679489661923 * wv / 180.0;  // MY_PI2
      const double st = std::sin(theta2);
      for (size_t i : sel) {
        double *q = &sq_quat[4 * i];
        q[0] = std::cos(theta2);
        q[1] = a * st;
        q[2] = b * st;
        q[3] = cc * st;
        const double nrm = 1.0 / std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        for (int d = 0; d < 4; d++) q[d] *= nrm;
      }
    } else if (kw == "quat/random") {
      if (t.size() < 5) die("Illegal set command");
      const int seed = inum(t[4]);
      if (seed <= 0) die("Invalid random number seed in set command");
      RanPark rng;
      for (size_t i : sel) {
        rng.reset(seed, &x[3 * i]);
        const double sv = rng.uniform();
        const double t1 = std::sqrt(1.0 - sv), t2 = std::sqrt(sv);
        const double theta1 = 2.0 * kMyPi * rng.uniform();
        const double th2 = 2.0 * kMyPi * rng.uniform();
        double *q = &sq_quat[4 * i];
        q[0] = std::cos(th2) * t2;
        q[1] = std::sin(theta1) * t1;
        q[2] = std::cos(theta1) * t1;
        q[3] = std::sin(th2) * t2;
      }
    } else {
      die("set keyword " + kw + " is not supported on this path");
    }
    fprintf(out, "Setting atom values ...\n  %zu settings made for %s\n", sel.size(), kw.c_str());
    dirty = true;
  }

  void build_handle_sq() {
    need_box("run");
    if (!pair_set || !pair_sq_surface) die("run: pair_style gran ... surface superquadric not defined");
    if (!nve_sq) die("run: no fix nve/superquadric, particles would not move");
    if (dt <= 0) die("run: timestep not set");
    if (skin < 0) die("run: neighbor skin not set");
    if (every != 1 || no_check || cdf != 1.0)
      die("superquadrics: neigh_modify every/check/contact_distance_factor must keep their defaults");
    if (pair_m.rolling_cdt || pair_m.rolling_epsd2 || pair_m.torsion_torque)
      die("superquadrics: rolling friction is not available on this path");
    for (auto &wl : walls)
      if (wl.model.rolling_cdt || wl.model.rolling_epsd2 || wl.model.torsion_torque)
        die("superquadrics: rolling friction is not available on this path");
    std::vector<const dem_contact_model *> sq_models{&pair_m};
    for (auto &wl : walls) sq_models.push_back(&wl.model);
    for (const dem_contact_model *cm : sq_models)
      if (cm->normal_model != DEM_NORMAL_HERTZ || cm->tangential_model || cm->cohesion_model)
        die("superquadrics: only model hertz tangential history is available on this path");
    if (!mesh_order.empty()) die("superquadrics: mesh walls are not available on this path");
    if ((int)Y.size() != ntypes || (int)nu.size() != ntypes)
      die("youngsModulus/poissonsRatio need one value per atom type");
    if ((int)e.size() != ntypes * ntypes || (int)mu.size() != ntypes * ntypes)
      die("coefficientRestitution/coefficientFriction need ntypes^2 values");
    if (walls.size() > 6) die("at most 6 primitive walls");
    if (hs) {
      dem_sqdem_destroy(hs);
      hs = nullptr;
    }
    dem_sqdem_params p{};
    p.dt = dt;
    p.skin = skin;
    for (int d = 0; d < 3; d++) {
      p.boxlo[d] = lo[d];
      p.boxhi[d] = hi[d];
      p.periodic[d] = periodic[d];
    }
    p.tangential_damping = pair_m.tangential_damping;
    p.limit_force = pair_m.limit_force;
    p.has_gravity = gravity;
    for (int d = 0; d < 3; d++) p.gravity[d] = gacc[d];
    p.gravity_first = gravity_first;
    p.nplanes = (int)walls.size();
    for (size_t k = 0; k < walls.size(); k++) {
      p.plane_axis[k] = walls[k].axis;
      p.plane_pos[k] = walls[k].pos;
      p.plane_type[k] = walls[k].type;
    }
    if (dem_sqdem_create(0, &p, &hs)) die(std::string("GPU library: ") + dem_sqdem_last_error(nullptr));
    if (dem_sqdem_set_materials(hs, ntypes, Y.data(), nu.data(), e.data(), mu.data())) die(dem_sqdem_last_error(hs));
    if (tag.empty()) die("run: no atoms");
    if (dem_sqdem_set_particles(hs, (int64_t)tag.size(), x.data(), v.data(), w.data(), sq_quat.data(),
                                sq_shape.data(), sq_blk.data(), sq_dens.data(), type.data()))
      die(dem_sqdem_last_error(hs));
    sq_mass.assign(tag.size(), 0.0);
    if (dem_sqdem_get_props(hs, sq_mass.data(), nullptr, nullptr, nullptr)) die(dem_sqdem_last_error(hs));
    m = sq_mass;
    dirty = false;
  }

  void build_handle() {
    if (sq) {
      build_handle_sq();
      return;
    }
    need_box("run");
    if (!pair_set) die("run: pair_style gran not defined");
    if (!nve) die("run: no fix nve/sphere, particles would not move");
    if (dt <= 0) die("run: timestep not set");
    if (skin < 0) die("run: neighbor skin not set");
    std::vector<const dem_contact_model *> models{&pair_m};
    for (auto &wl : walls) models.push_back(&wl.model);
    for (auto &name : mesh_order) models.push_back(&mesh_defs[name].model);
    bool any_rolling = false, elastic = false, stiff = false, abs_damp = false, rel_damp = false;
    bool hooke = false, sjkr = false;
    for (const dem_contact_model *cm : models) {
      any_rolling = any_rolling || cm->rolling_cdt || cm->rolling_epsd2;
      const bool st = cm->normal_model == DEM_NORMAL_HERTZ_STIFFNESS ||
                      cm->normal_model == DEM_NORMAL_HOOKE_STIFFNESS;
      elastic = elastic || !st;
      stiff = stiff || st;
      const bool ab = cm->normal_model == DEM_NORMAL_HOOKE_STIFFNESS && cm->absolute_damping;
      abs_damp = abs_damp || ab;
      rel_damp = rel_damp || (st && !ab);
      hooke = hooke || cm->normal_model == DEM_NORMAL_HOOKE;
      sjkr = sjkr || cm->cohesion_model;
    }
    // Yeff / Geff / restitution are connected by hertz and hooke only
    // (normal_model_*.h connectToProperties)
    const bool have_el = !Y.empty() || !nu.empty() || !e.empty();
    if (elastic || have_el) {
      if ((int)Y.size() != ntypes || (int)nu.size() != ntypes)
        die("youngsModulus/poissonsRatio need one value per atom type");
      if ((int)e.size() != ntypes * ntypes) die("coefficientRestitution needs ntypes^2 values");
    }
    if ((int)mu.size() != ntypes * ntypes) die("coefficientFriction needs ntypes^2 values");
    if (any_rolling && (int)mur.size() != ntypes * ntypes)
      die("rolling_friction needs coefficientRollingFriction (ntypes^2 values)");
    if (abs_damp && rel_damp)
      die("hooke/stiffness with and without absolute_damping in one run is not available on this path");
    auto pair_prop = [&](const char *name, bool need) -> const double * {
      auto it = model_props.find(name);
      if (it == model_props.end()) {
        if (need) die(std::string("property/global ") + name + " (peratomtypepair) is needed by the contact model");
        return nullptr;
      }
      if ((int)it->second.size() != ntypes * ntypes)
        die(std::string("property/global ") + name + " needs ntypes^2 values");
      return it->second.data();
    };
    const double *kn = pair_prop("kn", stiff), *kt = pair_prop("kt", stiff);
    const double *gn = pair_prop(abs_damp ? "gamman_abs" : "gamman", stiff);
    const double *gt = pair_prop(abs_damp ? "gammat_abs" : "gammat", stiff);
    const double *ced = pair_prop("cohesionEnergyDensity", sjkr);
    double cv = 0.0;
    if (model_props.count("characteristicVelocity")) {
      if (model_props["characteristicVelocity"].size() != 1) die("characteristicVelocity is one scalar");
      cv = model_props["characteristicVelocity"][0];
    } else if (hooke) {
      die("model hooke needs fix property/global characteristicVelocity scalar");
    }
    if (h) {
      dem_destroy(h);
      h = nullptr;
    }
    dem_params p{};
    p.dt = dt;
    p.skin = skin;
    p.cdf = cdf;
    p.neigh_every = every;
    p.neigh_no_check = no_check;
    for (int d = 0; d < 3; d++) {
      p.boxlo[d] = lo[d];
      p.boxhi[d] = hi[d];
      p.periodic[d] = periodic[d];
    }
    p.shrink_faces = shrink_faces;
    p.shrink_min = shrink_min;
    p.rolling_cdt = pair_m.rolling_cdt;
    p.tangential_damping = pair_m.tangential_damping;
    p.limit_force = pair_m.limit_force;
    p.torsion_torque = pair_m.torsion_torque;
    p.has_gravity = gravity;
    for (int d = 0; d < 3; d++) p.grav_acc[d] = gacc[d];
    p.gravity_first = gravity_first;
    p.precision = precision;
    p.max_neighbors = one;
    p.device = 0;
    int rc = dem_create(&p, &h);
    if (rc) die(std::string("GPU library: ") + dem_last_error(nullptr));
    if (dem_set_materials(h, ntypes, have_el ? Y.data() : nullptr, have_el ? nu.data() : nullptr,
                          have_el ? e.data() : nullptr, mu.data(), any_rolling ? mur.data() : nullptr))
      die(dem_last_error(h));
    if ((stiff || sjkr || hooke) && dem_set_model_coefficients(h, ntypes, stiff ? kn : nullptr,
                                                               stiff ? kt : nullptr, stiff ? gn : nullptr,
                                                               stiff ? gt : nullptr, ced, cv))
      die(dem_last_error(h));
    if (dem_set_pair_model(h, &pair_m)) die(dem_last_error(h));
    for (auto &wl : walls)
      if (dem_add_plane_wall(h, wl.axis, wl.pos, wl.type, &wl.model)) die(dem_last_error(h));
    for (auto &name : mesh_order) {
      MeshDef &md = mesh_defs[name];
      int id = -1;
      if (dem_add_mesh(h, (int64_t)(md.nodes.size() / 9), md.nodes.data(), md.type, &md.model, &id))
        die(dem_last_error(h));
      if (md.mover == 1 && dem_mesh_rotate(h, id, md.origin, md.axis, md.period)) die(dem_last_error(h));
      if (md.mover == 2 && dem_mesh_linear(h, id, md.vec)) die(dem_last_error(h));
      if (md.mover == 3 && dem_mesh_wiggle(h, id, md.vec, md.period)) die(dem_last_error(h));
      if (md.mover == 4 && dem_mesh_viblin(h, id, md.axis, (int)md.vampl.size(), md.vampl.data(),
                                           md.vphase.data(), md.vperiod.data()))
        die(dem_last_error(h));
      if (md.mover == 5 && dem_mesh_vibrot(h, id, md.origin, md.axis, (int)md.vampl.size(), md.vampl.data(),
                                           md.vphase.data(), md.vperiod.data()))
        die(dem_last_error(h));
      if (md.stress && dem_mesh_stress(h, id, md.pref)) die(dem_last_error(h));
      mesh_ids[name] = id;
    }
    if (dem_set_max_radius(h, insert_max_radius())) die(dem_last_error(h));
    if (tag.empty()) {
      if (next_insertion() == 0) die("run: no atoms");
    } else if (dem_set_particles(h, (int64_t)tag.size(), x.data(), v.data(), w.data(), r.data(), m.data(),
                                 type.data(), tag.data())) {
      die(dem_last_error(h));
    }
    if (!restart_file.empty()) {  // contact histories, list and mesh state of read_restart
      if (dem_read_restart(h, restart_file.c_str())) die(dem_last_error(h));
      restart_file.clear();
    }
    dirty = false;
  }

  void pull_state() {
    if (tag.empty()) return;
    if (sq) {
      if (dem_sqdem_get_state(hs, x.data(), v.data(), w.data(), nullptr, sq_quat.data(), nullptr, nullptr))
        die(dem_sqdem_last_error(hs));
      return;
    }
    if (dem_get_state(h, x.data(), v.data(), w.data(), nullptr, nullptr)) die(dem_last_error(h));
  }

  void thermo_header() {
    std::string hdr;
    for (auto &c : thermo_cols) {
      std::string name = c == "step" ? "Step" : c == "atoms" ? "Atoms" : c == "ke" ? "KinEng" : c;
      char buf[64];
      snprintf(buf, sizeof buf, "%14s ", name.c_str());
      hdr += buf;
    }
    fprintf(out, "%s\n", hdr.c_str());
  }

  // the box at the last neighbor build: shrink-wrapped faces follow the atoms
  // (Domain::reset_box); the created box before any run
  void current_box(double *blo, double *bhi) {
    for (int d = 0; d < 3; d++) {
      blo[d] = lo[d];
      bhi[d] = hi[d];
    }
    if (!sq && h && shrink_faces && dem_get_box(h, blo, bhi)) die(dem_last_error(h));
  }

  // f_<mesh>[1..6]: force and torque on a mesh/surface/stress
  // (FixMeshSurface::compute_vector, fix_mesh_surface.cpp:560-570);
  // c_<id> of compute erotate/sphere (0.5 I w^2 with I = 0.4 m r^2,
  // compute_erotate_sphere.cpp) or compute ke
  bool thermo_value(const std::string &c, double ke, double &val) {
    // Thermo::compute_vol / compute_time / compute_cpu (thermo.cpp:931-1003)
    if (c == "vol") {
      double blo[3], bhi[3];
      current_box(blo, bhi);
      val = (bhi[0] - blo[0]) * (bhi[1] - blo[1]) * (bhi[2] - blo[2]);
      return true;
    }
    if (c == "time") {
      val = atime + (double)(step - atimestep) * dt;
      return true;
    }
    if (c == "cpu") {
      val = std::chrono::duration<double>(std::chrono::steady_clock::now() - run_t0).count();
      return true;
    }
    if (c.size() > 5 && c.compare(0, 2, "f_") == 0 && c.back() == ']') {
      const size_t lb = c.find('[');
      if (lb == std::string::npos) return false;
      const std::string id = c.substr(2, lb - 2);
      const int k = std::atoi(c.substr(lb + 1, c.size() - lb - 2).c_str());
      auto it = mesh_ids.find(id);
      if (sq || !h || it == mesh_ids.end() || !mesh_defs[id].stress || k < 1 || k > 6) return false;
      double f[3], tq[3];
      if (dem_get_mesh_force(h, it->second, f, tq)) die(dem_last_error(h));
      val = k <= 3 ? f[k - 1] : tq[k - 4];
      return true;
    }
    if (c.size() > 2 && c.compare(0, 2, "c_") == 0) {
      auto it = computes.find(c.substr(2));
      if (it == computes.end()) return false;
      if (it->second == "ke") {
        val = ke;
        return true;
      }
      if (it->second == "erotate/sphere" && !sq) {
        double er = 0.0;
        for (size_t i = 0; i < tag.size(); i++)
          er += 0.4 * m[i] * r[i] * r[i] *
                (w[3 * i] * w[3 * i] + w[3 * i + 1] * w[3 * i + 1] + w[3 * i + 2] * w[3 * i + 2]);
        val = 0.5 * er;
        return true;
      }
    }
    return false;
  }

  void thermo_line() {
    pull_state();
    double ke = 0.0;
    for (size_t i = 0; i < tag.size(); i++)
      ke += 0.5 * m[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    std::string row;
    for (auto &c : thermo_cols) {
      char buf[64];
      double val = 0.0;
      if (c == "step") snprintf(buf, sizeof buf, "%14lld ", (long long)step);
      else if (c == "atoms") snprintf(buf, sizeof buf, "%14zu ", tag.size());
      else if (c == "ke") snprintf(buf, sizeof buf, "%14.8g ", ke);
      else if (thermo_value(c, ke, val)) snprintf(buf, sizeof buf, "%14.8g ", val);
      else snprintf(buf, sizeof buf, "%14s ", "n/a");
      row += buf;
    }
    fprintf(out, "%s\n", row.c_str());
  }

  // fixes of the GPU path act on every atom: their groups must hold them all
  void check_groups() {
    for (const auto &fg : physics_groups)
      if (fg.second != "all" && !group_all[fg.second])
        die("fix " + fg.first + ": group " + fg.second +
            " does not hold every atom; fixes on part of the atoms are not available on this path");
  }

  // run N [upto] [pre yes|no] [post yes|no] (Run::command, run.cpp:40-230)
  void run(const std::vector<std::string> &t) {
    if (t.size() < 2) die("Illegal run command");
    long long n = std::stoll(t[1]);
    bool pre = true, upto = false;
    for (size_t k = 2; k < t.size();) {
      if (t[k] == "upto") {
        upto = true;
        k++;
        continue;
      }
      if (k + 1 >= t.size()) die("Illegal run command");
      if (t[k] == "pre") pre = t[k + 1] != "no";
      else if (t[k] == "every" || t[k] == "start" || t[k] == "stop")
        die("run option " + t[k] + " is not supported on this path");
      else if (t[k] != "post") die("Illegal run command");
      k += 2;
    }
    if (upto) {  // N is the last timestep of the run (run.cpp:118-121)
      if (n < step) die("Invalid run command upto value");
      n -= step;
    }
    if (n < 0) die("Illegal run command");
    check_groups();
    if (dirty || (sq ? !hs : !h)) {
      build_handle();
      pre = true;
    } else if (!sq && dem_set_max_radius(h, insert_max_radius())) {
      die(dem_last_error(h));
    }
    const auto t0 = std::chrono::steady_clock::now();
    run_t0 = t0;
    long long done = 0;
    const long long chunk = thermo_every > 0 ? thermo_every : n;
    bool first = true;
    if (pre) {
      if (sq ? dem_sqdem_setup(hs) : dem_setup(h)) die(sq ? dem_sqdem_last_error(hs) : dem_last_error(h));
    }
    thermo_header();
    thermo_line();
    if (!sq) write_dumps(false);  // Output::setup writes the dumps due at this step
    while (done < n) {
      // insertions due at the next step: FixInsert::pre_exchange on the step
      // whose next_reneighbor it set (fix_insert.cpp:652-660), fixes in order
      for (const std::string &id : insert_order) {
        InsertFix &f = inserts[id];
        if (f.next == step + 1) do_insertion(id, f);
      }
      check_groups();
      long long k = sq ? chunk - (step % (chunk > 0 ? chunk : 1)) : next_output(n - done);
      if (k <= 0 || k > n - done) k = n - done;
      const long long ni = next_insertion();  // the next piece ends just before it
      if (ni > step + 1 && ni - 1 - step < k) k = ni - 1 - step;
      int rc = sq ? dem_sqdem_run(hs, (int)k) : dem_run_continue(h, k);
      (void)first;
      first = false;
      if (rc) {
        std::string msg = sq ? dem_sqdem_last_error(hs) : dem_last_error(h);
        // thermo_modify lost ignore / warn would delete the atom and go on; this
        // path keeps every atom, so a loss ends the run whatever the setting
        if (lost_policy != "error" && msg.rfind("Lost atoms", 0) == 0)
          msg += " (thermo_modify lost " + lost_policy + ": atoms are not deleted on this path)";
        die(msg);
      }
      done += k;
      step += k;
      if ((thermo_every > 0 && step % thermo_every == 0) || done == n) thermo_line();
      if (!sq) write_dumps(false);
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(out, "Loop time of %g on 1 procs (1 GPU) for %lld steps with %zu atoms\n", secs, n,
            tag.size());
    if (sq) {
      dem_sqdem_stats st{};
      dem_sqdem_get_stats(hs, &st);
      fprintf(out, "Neighbor list builds = %lld\n", (long long)st.rebuilds);
      fprintf(out, "Newton iterations = %lld (%lld contact detections)\n", (long long)st.newton,
              (long long)st.checks);
    } else {
      dem_stats st{};
      dem_get_stats(h, &st);
      fprintf(out, "Neighbor list builds = %lld\n", (long long)st.nrebuilds);
    }
    if (secs > 0)
      fprintf(out, "Performance: %.4g particle-timesteps/s\n", (double)n * tag.size() / secs);
  }

  void execute(const std::string &raw) {
    std::string line = raw;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = substitute(line);
    auto t = split(line);
    if (t.empty()) return;
    const std::string &c = t[0];
    if (c == "units") {
      if (t.size() < 2 || t[1] != "si") die("Only 'units si' is supported on this path");
    } else if (c == "atom_style") {
      if (t.size() >= 2 && t[1] == "superquadric") {
        if (!tag.empty()) die("atom_style command after simulation box is defined");
        sq = true;
      } else if (t.size() < 2 || (t[1] != "granular" && t[1] != "sphere")) {
        die("Only atom_style granular/sphere/superquadric is supported on this path");
      }
    } else if (c == "atom_modify" || c == "processors" || c == "echo" || c == "log" ||
               (c == "compute" && !(t.size() >= 4 && (t[3] == "erotate/sphere" || t[3] == "ke" ||
                                                       t[3] == "pair/gran/local" || t[3] == "wall/gran/local"))) ||
               c == "uncompute" || c == "restart" ||
               c == "write_dump" || c == "soft_particles" ||
               c == "hard_particles") {
      if (c != "atom_modify" && c != "processors" && c != "echo" && c != "log")
        fprintf(out, "WARNING: '%s' is output/bookkeeping only and is ignored on this path\n", c.c_str());
    } else if (c == "group") {
      // group <id> region <rid> (Group::assign, group.cpp:120-200) holds the
      // atoms inside the region now; atoms inserted later join only "all" and
      // their fix's group. Other group styles are kept as groups that do not
      // hold every atom.
      if (t.size() < 3) die("Illegal group command");
      if (t[1] == "all") die("Cannot redefine group all");
      bool all = false;
      if (t[2] == "region" && t.size() >= 4) {
        auto it = regions.find(t[3]);
        if (it == regions.end()) die("Group region ID does not exist");
        const std::vector<double> &b = it->second;
        all = true;
        for (size_t i = 0; i < tag.size() && all; i++)
          for (int d = 0; d < 3; d++)
            if (x[3 * i + d] < b[2 * d] || x[3 * i + d] > b[2 * d + 1]) all = false;
      } else {
        fprintf(out, "WARNING: group style %s is kept as a group that does not hold every atom\n", t[2].c_str());
      }
      group_all[t[1]] = all;
    } else if (c == "thermo_modify") {
      // Thermo::modify_params (thermo.cpp:476-602): lost error|warn|ignore is
      // kept; the formatting keywords change output only
      for (size_t k = 1; k < t.size(); k++)
        if (t[k] == "lost") {
          if (k + 1 >= t.size() || (t[k + 1] != "error" && t[k + 1] != "warn" && t[k + 1] != "ignore"))
            die("Illegal thermo_modify command");
          lost_policy = t[++k];
        }
    } else if (c == "dump") {
      dump_command(t);
    } else if (c == "dump_modify") {
      if (t.size() < 2 || !dumps.count(t[1])) die("Could not find dump_modify ID");
      for (size_t k = 2; k < t.size(); k++)
        if (t[k] == "sort" && k + 1 < t.size() && t[k + 1] != "id" && t[k + 1] != "off")
          die("dump_modify sort: atoms are always written in ID order on this path");
      fprintf(out, "WARNING: dump_modify: atoms are written in ID order with the default formats\n");
    } else if (c == "undump") {
      auto it = dumps.find(t.size() > 1 ? t[1] : "");
      if (it == dumps.end()) die("Could not find undump ID");
      if (it->second.fp) fclose(it->second.fp);
      dumps.erase(it);
    } else if (c == "compute" && t.size() >= 4 && (t[3] == "pair/gran/local" || t[3] == "wall/gran/local")) {
      local_compute_command(t);
    } else if (c == "compute") {
      // compute id group erotate/sphere | ke: read by thermo as c_id
      computes[t[1]] = t[3];
    } else if (c == "boundary") {
      // Domain::set_boundary (domain.cpp:194-238): p, or one or two of f s m
      if (t.size() < 4) die("Illegal boundary command");
      shrink_faces = shrink_min = 0;
      for (int d = 0; d < 3; d++) {
        const std::string &b = t[1 + d];
        if (b.empty() || b.size() > 2) die("Illegal boundary command");
        const char lo_c = b[0], hi_c = b.size() == 2 ? b[1] : b[0];
        for (char ch : {lo_c, hi_c})
          if (ch != 'p' && ch != 'f' && ch != 's' && ch != 'm') die("Illegal boundary command");
        if ((lo_c == 'p') != (hi_c == 'p')) die("Both sides of boundary must be periodic");
        periodic[d] = lo_c == 'p' ? 1 : 0;
        face[d][0] = lo_c;
        face[d][1] = hi_c;
        for (int side = 0; side < 2; side++) {
          const char ch = face[d][side];
          if (ch == 's' || ch == 'm') shrink_faces |= 1 << (2 * d + side);
          if (ch == 'm') shrink_min |= 1 << (2 * d + side);
        }
      }
    } else if (c == "newton") {
      if (t.size() < 2 || t[1] != "off")
        die("Pair granular with shear history requires newton pair off");  // pair_gran.cpp:260
    } else if (c == "communicate" || c == "comm_modify") {
      // ghost velocities are always available on the GPU path
    } else if (c == "region") {
      if (t.size() < 9 || t[2] != "block") die("Only 'region <id> block ...' is supported");
      std::vector<double> b;
      for (int k = 3; k < 9; k++) b.push_back(num(t[k]));
      regions[t[1]] = b;
    } else if (c == "create_box") {
      if (t.size() < 3) die("Illegal create_box command");
      auto it = regions.find(t[2]);
      if (it == regions.end()) die("Create_box region ID does not exist");
      ntypes = inum(t[1]);
      for (int d = 0; d < 3; d++) {
        lo[d] = it->second[2 * d];
        hi[d] = it->second[2 * d + 1];
      }
      have_box = true;
    } else if (c == "read_data") {
      if (t.size() < 2) die("Illegal read_data command");
      read_data(t[1]);
    } else if (c == "write_restart") {
      // write_restart <file> (write_restart.cpp): after a run, the system and
      // its contact histories in this library's restart layout
      if (t.size() < 2) die("Illegal write_restart command");
      if (sq) die("write_restart for superquadrics is not available on this path");
      if (!h || dirty) die("write_restart needs a preceding run on this path");
      if (dem_write_restart(h, t[1].c_str())) die(dem_last_error(h));
    } else if (c == "read_restart") {
      // read_restart <file> (read_restart.cpp): box, types, atoms and timestep
      // now; the device state (histories, list, meshes) when the system is built
      if (t.size() < 2) die("Illegal read_restart command");
      if (have_box) die("Cannot read_restart after simulation box is defined");
      int64_t nr = 0;
      int nt = 0, per[3];
      long long ts = 0;
      int64_t ts64 = 0;
      if (dem_restart_info(t[1].c_str(), &nr, &nt, lo, hi, per, &ts64)) die(dem_last_error(nullptr));
      ts = ts64;
      tag.assign(nr, 0);
      type.assign(nr, 0);
      x.assign(3 * nr, 0.0);
      v.assign(3 * nr, 0.0);
      w.assign(3 * nr, 0.0);
      r.assign(nr, 0.0);
      m.assign(nr, 0.0);
      if (dem_restart_particles(t[1].c_str(), nr, tag.data(), type.data(), x.data(), v.data(), w.data(),
                                r.data(), m.data()))
        die(dem_last_error(nullptr));
      for (int d = 0; d < 3; d++) {
        periodic[d] = per[d];
        // the file keeps periodicity only: non-periodic faces come back as f
        for (int side = 0; side < 2; side++) face[d][side] = per[d] ? 'p' : 'f';
      }
      shrink_faces = shrink_min = 0;
      ntypes = nt;
      have_box = true;
      step = ts;
      restart_file = t[1];
      fprintf(out, "  %lld atoms\n", (long long)nr);
      dirty = true;
    } else if (c == "neighbor") {
      if (t.size() < 3 || t[2] != "bin") die("neighbor style must be 'bin' on this path");
      skin = num(t[1]);
    } else if (c == "neigh_modify") {
      for (size_t k = 1; k + 1 < t.size(); k += 2) {
        if (t[k] == "delay") {
          if (inum(t[k + 1]) != 0) die("neigh_modify delay must be 0 on this path");
        } else if (t[k] == "every") {
          every = inum(t[k + 1]);
          if (every <= 0) die("Illegal neigh_modify command");
        } else if (t[k] == "check") {
          if (t[k + 1] == "yes") no_check = 0;
          else if (t[k + 1] == "no") no_check = 1;
          else die("Illegal neigh_modify command");
        } else if (t[k] == "contact_distance_factor") {
          cdf = num(t[k + 1]);
        } else if (t[k] == "one") {
          one = inum(t[k + 1]);
        } else if (t[k] == "page" || t[k] == "binsize") {
        } else die("Illegal neigh_modify command");
      }
      dirty = true;
    } else if (c == "pair_style") {
      if (t.size() < 2 || t[1] != "gran") die("Only pair_style gran is available on this path");
      size_t end;
      parse_model(t, 2, end, true);
      if (end != t.size()) die("Illegal pair_style gran argument " + t[end]);
      pair_set = true;
      dirty = true;
    } else if (c == "pair_coeff") {
      if (t.size() < 3 || t[1] != "*" || t[2] != "*") die("pair_coeff must be '* *' for granular");
    } else if (c == "timestep") {
      if (t.size() < 2) die("Illegal timestep command");
      if (dt > 0) atime += (double)(step - atimestep) * dt;  // Update::reset_timestep bookkeeping
      atimestep = step;
      dt = num(t[1]);
      dirty = true;
    } else if (c == "fix") {
      if (t.size() < 4) die("Illegal fix command");
      const std::string &style = t[3];
      if (style == "nve/sphere" || style == "gravity" || style == "wall/gran" || style == "nve/superquadric") {
        if (t[2] != "all" && !group_all.count(t[2])) die("Could not find fix group ID " + t[2]);
        physics_groups.push_back({t[1], t[2]});
      }
      if (style == "property/global") {
        fix_property_global(t);
      } else if (style == "nve/sphere") {
        if (sq) die("Fix nve/sphere requires atom style sphere");
        nve = true;
      } else if (style == "nve/superquadric") {
        // fix_nve_asphere_base.cpp:60-75: integration_scheme N (default 1)
        if (!sq) die("Fix nve/superquadric requires atom style superquadric");
        for (size_t k = 4; k < t.size(); k += 2) {
          if (t[k] != "integration_scheme" || k + 1 >= t.size()) die("fix nve/superquadric: unknown keyword");
          if (inum(t[k + 1]) != 1)
            die("fix nve/superquadric: only integration_scheme 1 (the default) is available on this path");
        }
        nve_sq = true;
        dirty = true;
      } else if (style == "gravity") {
        if (t.size() < 9 || t[5] != "vector") die("fix gravity: only '<g> vector x y z' is supported");
        const double g = num(t[4]);
        const double ax = num(t[6]), ay = num(t[7]), az = num(t[8]);
        const double len = std::sqrt(ax * ax + ay * ay + az * az);  // fix_gravity.cpp:379
        if (len == 0.) die("Gravity direction vector = 0");
        gacc[0] = g * (ax / len);
        gacc[1] = g * (ay / len);
        gacc[2] = g * (az / len);
        gravity = true;
        gravity_first = !wall_seen;
        dirty = true;
      } else if (style == "mesh/surface" || style == "mesh/surface/planar" ||
                 style == "mesh/surface/stress") {
        fix_mesh_surface(t);
      } else if (style == "move/mesh") {
        // fix id group move/mesh mesh <meshid> rotate origin x y z axis x y z period T
        //                                      | linear vx vy vz
        //                                      | wiggle amplitude ax ay az period T
        // (FixMoveMesh, fix_move_mesh.cpp:70-120; the movers' argument lists)
        if (t.size() < 7 || t[4] != "mesh") die("fix move/mesh: expected 'mesh <id> <movement>'");
        auto it = mesh_defs.find(t[5]);
        if (it == mesh_defs.end()) die("fix move/mesh: unknown mesh " + t[5]);
        MeshDef &md = it->second;
        if (md.mover) die("fix move/mesh: one fix move/mesh per mesh on this path");
        if (t[6] == "rotate") {
          if (t.size() < 17 || t[7] != "origin" || t[11] != "axis" || t[15] != "period")
            die("fix move/mesh rotate: expected 'origin x y z axis x y z period T'");
          for (int d = 0; d < 3; d++) {
            md.origin[d] = num(t[8 + d]);
            md.axis[d] = num(t[12 + d]);
          }
          md.period = num(t[16]);
          md.mover = 1;
        } else if (t[6] == "linear") {
          if (t.size() < 10) die("Not enough arguments for movement type linear");
          for (int d = 0; d < 3; d++) md.vec[d] = num(t[7 + d]);
          md.mover = 2;
        } else if (t[6] == "wiggle") {
          if (t.size() < 13) die("Not enough arguments for movement type wiggle");
          if (t[7] != "amplitude") die("Expected keyword 'amplitude'");
          if (t[11] != "period") die("Expected keyword 'period'");
          for (int d = 0; d < 3; d++) md.vec[d] = num(t[8 + d]);
          md.period = num(t[12]);
          md.mover = 3;
        } else if (t[6] == "viblin" || t[6] == "vibrot") {
          // viblin axis x y z order n amplitude a.. phase p.. period T..
          // vibrot origin x y z axis x y z order n amplitude ... (same tail)
          const bool rot = t[6] == "vibrot";
          size_t k = 7;
          if (rot) {
            if (t.size() < k + 4 || t[k] != "origin") die("Expected keyword 'origin'");
            for (int d = 0; d < 3; d++) md.origin[d] = num(t[k + 1 + d]);
            k += 4;
          }
          if (t.size() < k + 6 || t[k] != "axis") die("Expected keyword 'axis'");
          for (int d = 0; d < 3; d++) md.axis[d] = num(t[k + 1 + d]);
          k += 4;
          if (t[k] != "order") die("Expected keyword 'order'");
          const int ord = inum(t[k + 1]);
          if (ord > 30 || ord < 1) die("order can be at most 30 and must be greater 0");
          k += 2;
          if (t.size() < k + 3 + 3 * (size_t)ord) die("Not enough arguments for movement type " + t[6]);
          if (t[k] != "amplitude") die("Expected keyword 'amplitude'");
          if (t[k + 1 + ord] != "phase") die("Expected keyword 'phase'");
          if (t[k + 2 + 2 * ord] != "period") die("Expected keyword 'period'");
          md.vampl.clear();
          md.vphase.clear();
          md.vperiod.clear();
          for (int j = 0; j < ord; j++) {
            md.vampl.push_back(num(t[k + 1 + j]));
            md.vphase.push_back(num(t[k + 2 + ord + j]));
            md.vperiod.push_back(num(t[k + 3 + 2 * ord + j]));
          }
          md.mover = rot ? 5 : 4;
        } else {
          die("fix move/mesh: movement " + t[6] +
              " is not available on this path (rotate, linear, wiggle, viblin, vibrot)");
        }
        dirty = true;
      } else if (style == "wall/gran" && t.size() > 4 &&
                 std::find(t.begin(), t.end(), std::string("mesh")) != t.end() &&
                 std::find(t.begin(), t.end(), std::string("primitive")) == t.end()) {
        if (sq) die("fix wall/gran: mesh walls for superquadrics are not available on this path");
        size_t k;
        dem_contact_model wm{};
        parse_model(t, 4, k, false, &wm);
        // mesh n_meshes <k> meshes <ids...>
        if (k + 3 >= t.size() || t[k] != "mesh" || t[k + 1] != "n_meshes" || t[k + 3] != "meshes")
          die("fix wall/gran: expected 'mesh n_meshes <k> meshes <ids>'");
        const int nm = inum(t[k + 2]);
        if (nm < 1 || k + 4 + nm > t.size()) die("Illegal fix wall/gran command");
        for (int q = 0; q < nm; q++) {
          auto it = mesh_defs.find(t[k + 4 + q]);
          if (it == mesh_defs.end()) die("fix wall/gran: unknown mesh " + t[k + 4 + q]);
          if (it->second.used) die("fix wall/gran: mesh " + t[k + 4 + q] + " used twice");
          it->second.used = true;
          it->second.model = wm;
          mesh_order.push_back(t[k + 4 + q]);
          mesh_fix[t[k + 4 + q]] = {t[1], q};
        }
        last_mesh_wall_fix = t[1];
        if (!gravity) gravity_first = false;
        wall_seen = true;
        dirty = true;
      } else if (style == "wall/gran") {
        size_t k;
        dem_contact_model wm{};
        parse_model(t, 4, k, false, &wm);
        if (k >= t.size() || t[k] != "primitive")
          die("fix wall/gran: only 'primitive ... xplane|yplane|zplane' and mesh walls are available");
        if (k + 4 >= t.size() || t[k + 1] != "type") die("Illegal fix wall/gran command");
        const int wt = inum(t[k + 2]);
        const std::string &shape = t[k + 3];
        int ax = -1;
        if (shape == "xplane") ax = 0;
        else if (shape == "yplane") ax = 1;
        else if (shape == "zplane") ax = 2;
        else die("fix wall/gran: primitive wall style " + shape + " is not available");
        walls.push_back({ax, num(t[k + 4]), wt, wm});
        if (!gravity) gravity_first = false;
        wall_seen = true;
        dirty = true;
      } else if (style == "particletemplate/sphere") {
        need_box("fix particletemplate/sphere");
        fix_template(t);
      } else if (style.compare(0, 29, "particledistribution/discrete") == 0) {
        fix_distribution(t);
      } else if (style == "insert/pack") {
        need_box("fix insert/pack");
        if (sq) die("fix insert/pack for superquadrics is not available on this path");
        if (t[2] != "all" && !group_all.count(t[2])) die("Could not find fix group ID " + t[2]);
        fix_insert_pack(t);
      } else {
        die("fix style " + style + " is not available on this path");
      }
    } else if (c == "set") {
      set_command(t);
    } else if (c == "unfix") {
      // Modify::delete_fix: insertion fixes and their templates leave; the
      // other fixes of this path shape the GPU system and stay
      if (t.size() < 2) die("Illegal unfix command");
      if (inserts.count(t[1])) {
        if (inserts[t[1]].ins) dem_inserter_destroy(inserts[t[1]].ins);
        inserts.erase(t[1]);
        insert_order.erase(std::find(insert_order.begin(), insert_order.end(), t[1]));
      } else if (templates.count(t[1])) {
        templates.erase(t[1]);
      } else if (distributions.count(t[1])) {
        distributions.erase(t[1]);
      } else {
        die("unfix " + t[1] + ": only insertion fixes (insert/pack, particletemplate/sphere, "
            "particledistribution/discrete) can be removed on this path");
      }
    } else if (c == "thermo") {
      if (t.size() < 2) die("Illegal thermo command");
      thermo_every = inum(t[1]);
    } else if (c == "thermo_style") {
      if (t.size() < 2) die("Illegal thermo_style command");
      thermo_cols.clear();
      if (t[1] == "custom") {
        for (size_t k = 2; k < t.size(); k++) thermo_cols.push_back(t[k]);
      } else {
        thermo_cols = {"step", "atoms", "ke"};
      }
    } else if (c == "variable") {
      if (t.size() < 4) die("Illegal variable command");
      if (t[2] != "equal" && t[2] != "index" && t[2] != "string")
        die("variable style " + t[2] + " is not supported on this path");
      if ((t[2] == "equal" || t[2] == "string") && t.size() != 4) die("Illegal variable command");
      // an index or string variable keeps its first definition (variable.cpp:160-200)
      if (vars.find(t[1]) == vars.end() || t[2] == "equal") {
        vars[t[1]] = t[3];
        var_style[t[1]] = t[2];
      }
    } else if (c == "print") {
      // Input::print (input.cpp:1010-1060): the first argument, quotes removed
      if (t.size() < 2) die("Illegal print command");
      fprintf(out, "%s\n", t[1].c_str());
    } else if (c == "package") {
      if (t.size() >= 4 && t[1] == "amd" && t[2] == "precision") {
        if (t[3] == "fp64") precision = DEM_FP64;
        else if (t[3] == "mixed") precision = DEM_MIXED;
        else die("package amd precision must be fp64 or mixed");
        dirty = true;
      } else die("Illegal package command");
    } else if (c == "run") {
      run(t);
    } else {
      die("Unknown command: " + line);
    }
  }
};

}  // namespace

extern "C" int dem_run_script(const char *script_path, const char *log_path) {
  if (!script_path) return DEM_ERR_ARG;
  std::ifstream f(script_path);
  if (!f) {
    fprintf(stderr, "ERROR: Cannot open input script %s\n", script_path);
    return DEM_ERR_ARG;
  }
  Interp in;
  FILE *logf = nullptr;
  if (log_path) {
    logf = fopen(log_path, "w");
    if (!logf) {
      fprintf(stderr, "ERROR: Cannot open logfile %s\n", log_path);
      return DEM_ERR_ARG;
    }
    in.out = logf;
  }
  std::string line, acc;
  int lineno = 0;
  int rc = DEM_OK;
  try {
    while (std::getline(f, line)) {
      lineno++;
      // continuation lines end with '&' (input.cpp:211-220)
      size_t endpos = line.find_last_not_of(" \t\r");
      if (endpos != std::string::npos && line[endpos] == '&') {
        acc += line.substr(0, endpos) + " ";
        continue;
      }
      acc += line;
      in.execute(acc);
      acc.clear();
    }
  } catch (const ScriptError &err) {
    fprintf(in.out, "ERROR: %s (line %d)\n", err.msg.c_str(), lineno);
    if (in.out != stderr) fprintf(stderr, "ERROR: %s (line %d)\n", err.msg.c_str(), lineno);
    rc = DEM_ERR_ARG;
  } catch (const std::exception &err) {
    fprintf(stderr, "ERROR: %s (line %d)\n", err.what(), lineno);
    rc = DEM_ERR_ARG;
  }
  if (logf) fclose(logf);
  return rc;
}
