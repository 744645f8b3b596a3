// Particle insertion: fix insert/pack with fix particletemplate/sphere and
// fix particledistribution/discrete on a block region (C ABI of
// include/liggghts_amd.h, dem_inserter_*).
//
// Insertion is sequential by nature: each candidate position is accepted
// only if it overlaps none of the particles already placed, so the order of
// the random streams defines the result. This host code keeps every stream
// the reference keeps (the fix's, the region's, the distribution's), draws
// from them in the reference's order and places the same particles at the
// same positions. Overlap candidates come from a uniform cell grid over the
// fix's bounding box whose cells are no smaller than the largest possible
// contact distance, so each test visits the 27 cells around the candidate.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/liggghts_amd.h"

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI
constexpr int kSeedOffset = 12;                  // fix_insert_pack.cpp:66
std::string g_ins_create_error;

// RanPark (random_park.cpp:54-116)
struct Park {
  int seed = 1;
  bool save = false;
  double second = 0.0;
  explicit Park(int s = 1) : seed(s) {}
  double uniform() {
    const int k = seed / 127773;
    seed = 16807 * (seed - k * 127773) - 2836 * k;
    if (seed < 0) seed += 2147483647;
    return (1.0 / 2147483647) * seed;
  }
  double gaussian() {
    if (save) {
      save = false;
      return second;
    }
    double v1, v2, rsq;
    do {
      v1 = 2.0 * uniform() - 1.0;
      v2 = 2.0 * uniform() - 1.0;
      rsq = v1 * v1 + v2 * v2;
    } while (!(rsq < 1.0 && rsq != 0.0));
    const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
    second = v1 * fac;
    save = true;
    return v2 * fac;
  }
};

// region <id> block, side in (RegBlock, region_block.cpp:60-259)
struct Block {
  double lo[3], hi[3];
  bool inside(const double *x) const {
    return x[0] >= lo[0] && x[0] <= hi[0] && x[1] >= lo[1] && x[1] <= hi[1] && x[2] >= lo[2] &&
           x[2] <= hi[2];
  }
  // RegBlock::surface_interior: a face closer than cutoff (the point inside)
  bool near_face(const double *x, double cutoff) const {
    if (!inside(x)) return false;
    for (int d = 0; d < 3; d++)
      if (x[d] - lo[d] < cutoff || hi[d] - x[d] < cutoff) return true;
    return false;
  }
  // Region::match_expandby_cut (region.cpp:608-619): inside, or outside
  // within cutoff of the block (RegBlock::surface_exterior)
  bool within(const double *x, double cutoff) const {
    if (inside(x)) return true;
    for (int d = 0; d < 3; d++)
      if (x[d] <= lo[d] - cutoff || x[d] >= hi[d] + cutoff) return false;
    double dd[3];
    for (int d = 0; d < 3; d++) dd[d] = x[d] - std::min(std::max(x[d], lo[d]), hi[d]);
    return std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]) < cutoff;
  }
};

// Particles a candidate must not overlap, binned on a uniform grid.
class OverlapGrid {
 public:
  void reset(const double *lo, const double *hi, double cell) {
    pts_.clear();
    for (int d = 0; d < 3; d++) {
      lo_[d] = lo[d];
      const double ext = std::max(hi[d] - lo[d], 0.0);
      n_[d] = std::max(1, (int)std::floor(ext / cell));
      // cap the grid at ~8M cells (the reference's own bin limit)
      while ((double)n_[0] * n_[1] * n_[2] > 8.0e6 && n_[d] > 1) n_[d] = (n_[d] + 1) / 2;
      inv_[d] = ext > 0 ? n_[d] / ext : 0.0;
    }
    head_.assign((size_t)n_[0] * n_[1] * n_[2], -1);
    next_.clear();
  }
  void add(const double *x, double r) {
    const int c = cell_of(x);
    pts_.push_back({x[0], x[1], x[2], r});
    next_.push_back(head_[c]);
    head_[c] = (int)pts_.size() - 1;
  }
  // RegionNeighborList::hasOverlap (region_neighbor_list_I.h:72-98): the
  // contact test rsq <= (r + r_k)^2 against every listed particle near x
  bool overlaps(const double *x, double r) const {
    int c[3];
    coords(x, c);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int ix = c[0] + dx, iy = c[1] + dy, iz = c[2] + dz;
          if (ix < 0 || iy < 0 || iz < 0 || ix >= n_[0] || iy >= n_[1] || iz >= n_[2]) continue;
          for (int k = head_[((size_t)iz * n_[1] + iy) * n_[0] + ix]; k >= 0; k = next_[k]) {
            const P &p = pts_[k];
            const double d0 = x[0] - p.x, d1 = x[1] - p.y, d2 = x[2] - p.z;
            const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
            const double radsum = r + p.r;
            if (rsq <= radsum * radsum) return true;
          }
        }
    return false;
  }

 private:
  struct P {
    double x, y, z, r;
  };
  void coords(const double *x, int *c) const {
    for (int d = 0; d < 3; d++) {
      int k = (int)std::floor((x[d] - lo_[d]) * inv_[d]);
      c[d] = std::min(std::max(k, 0), n_[d] - 1);
    }
  }
  int cell_of(const double *x) const {
    int c[3];
    coords(x, c);
    return ((c[2] * n_[1]) + c[1]) * n_[0] + c[0];
  }
  double lo_[3] = {0, 0, 0}, inv_[3] = {0, 0, 0};
  int n_[3] = {1, 1, 1};
  std::vector<int> head_, next_;
  std::vector<P> pts_;
};

}  // namespace

struct dem_inserter {
  dem_insert_params p{};
  std::string err;
  Park rnd, region_rnd, dist_rnd;
  Block region{};
  int nt = 0;
  double vol[DEM_INSERT_MAXTEMPLATES]{}, mass[DEM_INSERT_MAXTEMPLATES]{};
  double weight[DEM_INSERT_MAXTEMPLATES]{};  // number-based
  int order[DEM_INSERT_MAXTEMPLATES]{};      // largest expected volume first
  double volexpect = 0, massexpect = 0, maxrad = 0;
  double region_volume = 0, region_volume_local = 0;
  OverlapGrid grid;
};

namespace {

int ins_fail(dem_inserter *s, int code, const std::string &m) {
  if (s) s->err = m;
  else g_ins_create_error = m;
  return code;
}

// Region::volume_mc (region.cpp:634-701) on one process: every test point
// costs three draws of the region's stream.
bool volume_mc(dem_inserter *s, const double *boxlo, const double *boxhi) {
  const dem_insert_params &p = s->p;
  const double *lo = p.region_lo, *hi = p.region_hi;
  const double cut = s->maxrad;
  int nin = 0, nloc = 0;
  for (int i = 0; i < p.ntry_mc; i++) {
    double x[3];
    for (int d = 0; d < 3; d++) x[d] = lo[d] + s->region_rnd.uniform() * (hi[d] - lo[d]);
    bool in_box = true;
    for (int d = 0; d < 3; d++) in_box = in_box && x[d] >= boxlo[d] && x[d] <= boxhi[d];
    if (!in_box || !s->region.inside(x)) continue;
    nin++;
    if (!p.all_in || !s->region.near_face(x, cut)) nloc++;
  }
  if (nin == 0) return false;
  const double vbox = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  const double vg = (double)nin / (double)p.ntry_mc * vbox;
  double vl = (double)nloc / (double)p.ntry_mc * vbox;
  if (vl < 1.e-10) return false;  // Region volume_limit_ default (region.cpp:93)
  const double vl_all = vl;
  vl *= vg / vl_all;
  s->region_volume = vg;
  s->region_volume_local = vl;
  return true;
}

double dist_to_box(const double *lo, const double *hi, const double *x) {
  // Domain::dist_subbox_borders (domain_I.h:126-143) with vectorMin3D's
  // tie-breaking (vector_liggghts.h:176-185)
  auto min3 = [](const double *v) {
    if (v[0] < v[1] && v[0] < v[2]) return v[0];
    return v[1] < v[2] ? v[1] : v[2];
  };
  double a[3], b[3];
  for (int d = 0; d < 3; d++) {
    a[d] = std::fabs(lo[d] - x[d]);
    b[d] = std::fabs(hi[d] - x[d]);
  }
  const double dl = min3(a), dh = min3(b);
  return dl < dh ? dl : dh;
}

}  // namespace

extern "C" {

int dem_inserter_create(const dem_insert_params *params, dem_inserter **out) {
  if (!params || !out) return ins_fail(nullptr, DEM_ERR_ARG, "null argument");
  *out = nullptr;
  const dem_insert_params &p = *params;
  const int nt = p.ntemplates;
  if (nt < 1 || nt > DEM_INSERT_MAXTEMPLATES)
    return ins_fail(nullptr, DEM_ERR_ARG, "particledistribution/discrete: illegal number of templates");
  if (p.seed <= 0 || p.distribution_seed <= 0) return ins_fail(nullptr, DEM_ERR_ARG, "illegal seed");
  for (int i = 0; i < nt; i++) {
    if (!(p.template_radius[i] > 0)) return ins_fail(nullptr, DEM_ERR_ARG, "radius must be >= 0");
    if (!(p.template_density[i] > 0)) return ins_fail(nullptr, DEM_ERR_ARG, "density must be >= 0");
    if (p.template_type[i] < 1) return ins_fail(nullptr, DEM_ERR_ARG, "invalid atom type (must be >=1)");
    if (p.distribution_weight[i] < 0) return ins_fail(nullptr, DEM_ERR_ARG, "invalid weight");
  }
  for (int d = 0; d < 3; d++)
    if (p.region_lo[d] > p.region_hi[d]) return ins_fail(nullptr, DEM_ERR_ARG, "Illegal region block command");
  if (p.maxattempt <= 0) return ins_fail(nullptr, DEM_ERR_ARG, "maxattempt must be > 0");
  if (p.ntry_mc < 1000) return ins_fail(nullptr, DEM_ERR_ARG, "ntry_mc must be > 1000");
  const int ndef = (p.volumefraction_region > 0.) + (p.particles_in_region > 0) + (p.mass_in_region > 0.);
  if (ndef != 1)
    return ins_fail(nullptr, DEM_ERR_ARG,
                    "must define exactly one keyword out of 'volumefraction_region', "
                    "'particles_in_region', and 'mass_in_region'");
  if (p.volumefraction_region > 1.) return ins_fail(nullptr, DEM_ERR_ARG, "Invalid volumefraction");

  dem_inserter *s = new dem_inserter;
  s->p = p;
  s->nt = nt;
  for (int d = 0; d < 3; d++) {
    s->region.lo[d] = p.region_lo[d];
    s->region.hi[d] = p.region_hi[d];
  }
  // templates: constant radius and density (fix_template_sphere.cpp:236-244)
  for (int i = 0; i < nt; i++) {
    const double r = p.template_radius[i];
    s->vol[i] = r * r * r * 4. * kPi / 3.;
    s->mass[i] = p.template_density[i] * s->vol[i];
    s->maxrad = std::max(s->maxrad, r);
  }
  // distribution weights, mass-based unless numberbased
  // (fix_particledistribution_discrete.cpp:126-175)
  double wsum = 0;
  for (int i = 0; i < nt; i++) wsum += p.distribution_weight[i];
  if (!(wsum > 0)) {
    delete s;
    return ins_fail(nullptr, DEM_ERR_ARG, "invalid weight");
  }
  for (int i = 0; i < nt; i++) s->weight[i] = p.distribution_weight[i] / wsum;
  if (!p.distribution_numberbased) {
    for (int i = 0; i < nt; i++) s->weight[i] = s->weight[i] / s->mass[i];
    wsum = 0;
    for (int i = 0; i < nt; i++) wsum += s->weight[i];
    for (int i = 0; i < nt; i++) s->weight[i] /= wsum;
  }
  for (int i = 0; i < nt; i++) {
    s->volexpect += s->vol[i] * s->weight[i];
    s->massexpect += s->mass[i] * s->weight[i];
  }
  // insertion order: descending expected volume, ties keep the listed order
  // (the reference's bubble sort, :199-221, is stable)
  for (int i = 0; i < nt; i++) s->order[i] = i;
  std::stable_sort(s->order, s->order + nt, [s](int a, int b) { return s->vol[a] > s->vol[b]; });
  s->rnd = Park(p.seed);
  s->dist_rnd = Park(p.distribution_seed);
  s->region_rnd = Park(p.seed + kSeedOffset);
  if (!volume_mc(s, p.boxlo, p.boxhi) || s->region_volume <= 0. || s->region_volume_local < 0. ||
      (s->region_volume_local - s->region_volume) / s->region_volume > 1e-3) {
    delete s;
    return ins_fail(nullptr, DEM_ERR_ARG, "Fix insert: Region volume calculation with MC failed");
  }
  *out = s;
  return DEM_OK;
}

int dem_inserter_destroy(dem_inserter *ins) {
  delete ins;
  return DEM_OK;
}

const char *dem_inserter_last_error(const dem_inserter *ins) {
  return ins ? ins->err.c_str() : g_ins_create_error.c_str();
}

int dem_inserter_info(const dem_inserter *s, double *region_volume, double *volume_expect,
                      double *mass_expect, double *max_radius) {
  if (!s) return DEM_ERR_ARG;
  if (region_volume) *region_volume = s->region_volume;
  if (volume_expect) *volume_expect = s->volexpect;
  if (mass_expect) *mass_expect = s->massexpect;
  if (max_radius) *max_radius = s->maxrad;
  return DEM_OK;
}

int dem_inserter_step(dem_inserter *s, int64_t n, const double *x, const double *radius,
                      const double *rmass, const double *boxlo, const double *boxhi, int64_t cap,
                      double *x_out, double *v_out, double *omega_out, double *radius_out,
                      double *density_out, double *rmass_out, int *type_out, int64_t *ninserted,
                      int64_t *nrequested) {
  if (!s || n < 0 || (n > 0 && (!x || !radius || !rmass)) || !boxlo || !boxhi || cap < 0 ||
      (cap > 0 && (!x_out || !v_out || !omega_out || !radius_out || !density_out || !rmass_out ||
                   !type_out)))
    return s ? ins_fail(s, DEM_ERR_ARG, "bad arguments") : DEM_ERR_ARG;
  const dem_insert_params &p = s->p;
  if (ninserted) *ninserted = 0;
  if (nrequested) *nrequested = 0;
  // what a call that finds too little room in the output gives back
  const Park keep_rnd = s->rnd, keep_region = s->region_rnd, keep_dist = s->dist_rnd;
  const double keep_vol = s->region_volume, keep_vol_local = s->region_volume_local;

  // a region reaching outside a box with f faces (fix_insert_pack.cpp:252-261,
  // Region::bbox_extends_outside_box, region.cpp:704-711)
  if (p.warn_region && p.fixed_boundary) {
    bool out = false;
    for (int d = 0; d < 3; d++) {
      const double a = p.region_lo[d] + 1e-8, b = p.region_hi[d] - 1e-8;
      out = out || !(a >= boxlo[d] && a <= boxhi[d]) || !(b >= boxlo[d] && b <= boxhi[d]);
    }
    if (out)
      return ins_fail(s, DEM_ERR_ARG,
                      "Insertion region extends outside simulation box and a fixed boundary is used."
                      "Please use non-fixed boundaries in this case only");
  }
  // how many are missing in the region (FixInsertPack::calc_ninsert_this,
  // fix_insert_pack.cpp:236-330); all particles count, not only the group's
  int64_t np_region = 0;
  double vol_region = 0., mass_region = 0.;
  const double c43 = 4. * kPi / 3.;
  for (int64_t i = 0; i < n; i++)
    if (s->region.inside(x + 3 * i)) {
      np_region++;
      vol_region += c43 * radius[i] * radius[i] * radius[i];
      mass_region += rmass[i];
    }
  long long want;
  double ratio;
  if (p.volumefraction_region > 0.) {
    want = (int)((p.volumefraction_region * s->region_volume - vol_region) / s->volexpect + s->rnd.uniform());
    ratio = vol_region / (p.volumefraction_region * s->region_volume);
  } else if (p.particles_in_region > 0) {
    want = (long long)p.particles_in_region - np_region;
    ratio = (double)np_region / (double)p.particles_in_region;
  } else {
    want = (int)((p.mass_in_region - mass_region) / s->massexpect + s->rnd.uniform());
    ratio = mass_region / p.mass_in_region;
  }
  if (want < -200000)
    return ins_fail(s, DEM_ERR_ARG,
                    "overflow in particle number calculation: inserting too many particles in one step");
  want = std::max(want, 0LL);
  ratio = std::min(std::max(ratio, 0.), 1.);

  // one process takes the whole share (FixInsert::distribute_ninsert_this,
  // fix_insert.cpp:829-913: the exact split draws nothing); its
  // insertion_fraction re-measures the region when the box can change
  if (p.box_change && !volume_mc(s, boxlo, boxhi))
    return ins_fail(s, DEM_ERR_ARG, "Unable to calculate region volume");
  int nlocal = (int)want;
  if (!p.exact_number)
    nlocal = (int)(s->region_volume_local / s->region_volume * (double)want + s->rnd.uniform());

  // how many of each template (randomize_list,
  // fix_particledistribution_discrete.cpp:395-469)
  int count[DEM_INSERT_MAXTEMPLATES];
  if (!p.exact_number) {
    for (int i = 0; i < s->nt; i++) count[i] = (int)((double)nlocal * s->weight[i] + s->dist_rnd.uniform());
  } else {
    double rem[DEM_INSERT_MAXTEMPLATES];
    int placed = 0;
    for (int i = 0; i < s->nt; i++) {
      count[i] = (int)((double)nlocal * s->weight[i]);
      placed += count[i];
      rem[i] = (double)nlocal * s->weight[i] - (double)count[i];
    }
    const int gap = nlocal - placed;
    for (int g = 0; g < gap; g++) {
      const double r = s->dist_rnd.uniform() * (double)gap;
      int j = 0;
      double acc = rem[0];
      while (acc < r && j < s->nt - 1) acc += rem[++j];
      count[j]++;
    }
  }
  std::vector<int> plan;  // template of each particle, largest templates first
  for (int i = 0; i < s->nt; i++) plan.insert(plan.end(), count[s->order[i]], s->order[i]);
  const int64_t nplan = (int64_t)plan.size();
  if (nrequested) *nrequested = nplan;
  if (nplan == 0) return DEM_OK;
  if (nplan > cap) {  // nothing happened: the caller can retry with room for nplan
    s->rnd = keep_rnd;
    s->region_rnd = keep_region;
    s->dist_rnd = keep_dist;
    s->region_volume = keep_vol;
    s->region_volume_local = keep_vol_local;
    return ins_fail(s, DEM_ERR_OVERFLOW, "the output holds fewer particles than the insertion asks for");
  }

  // particles to check against: those near the region inside the fix's
  // bounding box (FixInsert::load_xnear, fix_insert.cpp:916-948;
  // FixInsertPack::is_nearby / getBoundingBox, fix_insert_pack.cpp:345-374)
  double bblo[3], bbhi[3];
  for (int d = 0; d < 3; d++) {
    bblo[d] = std::max(p.region_lo[d], boxlo[d]) - 2 * s->maxrad;
    bbhi[d] = std::min(p.region_hi[d], boxhi[d]) + 2 * s->maxrad;
  }
  std::vector<int64_t> listed;
  double rlist = s->maxrad;
  if (p.overlapcheck) {
    for (int64_t i = 0; i < n; i++) {
      const double *xi = x + 3 * i;
      if (!s->region.within(xi, p.all_in ? s->maxrad : radius[i] + s->maxrad)) continue;
      bool in = true;
      for (int d = 0; d < 3; d++) in = in && xi[d] >= bblo[d] && xi[d] <= bbhi[d];
      if (!in) continue;
      listed.push_back(i);
      rlist = std::max(rlist, radius[i]);
    }
  }
  s->grid.reset(bblo, bbhi, s->maxrad + rlist);
  for (int64_t i : listed) s->grid.add(x + 3 * i, radius[i]);

  // candidate positions (FixInsertPack::x_v_omega, fix_insert_pack.cpp:390-500;
  // Region::generate_random / generate_random_shrinkby_cut, region.cpp:531-571)
  double lo[3], hi[3], span[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = std::max(p.region_lo[d], boxlo[d]);
    hi[d] = std::min(p.region_hi[d], boxhi[d]);
    if (lo[d] >= hi[d]) return ins_fail(s, DEM_ERR_ARG, "Impossible to generate random points on wrong sub-domain");
    span[d] = hi[d] - lo[d];
  }
  if (p.all_in)
    for (int d = 0; d < 3; d++)
      if (p.region_hi[d] - p.region_lo[d] < 2. * s->maxrad)
        return ins_fail(s, DEM_ERR_ARG,
                        "Impossible to generate random points within region - region too small "
                        "(smaller than twice the particle cutoff)");
  auto candidate = [&](double rb, double *pos) {
    do {
      for (int d = 0; d < 3; d++) pos[d] = lo[d] + s->region_rnd.uniform() * span[d];
    } while (!s->region.inside(pos) || (p.all_in && s->region.near_face(pos, rb)));
  };
  auto velocity = [&](double *v) {
    for (int d = 0; d < 3; d++) v[d] = p.v_insert[d];
    if (p.vel_style == 1)
      for (int d = 0; d < 3; d++) v[d] = p.v_insert[d] + p.v_fluct[d] * 2.0 * (s->rnd.uniform() - 0.50);
    else if (p.vel_style == 2)
      for (int d = 0; d < 3; d++) v[d] = p.v_insert[d] + p.v_fluct[d] * s->rnd.gaussian();
  };
  const long long maxtry = ratio >= 1. ? (long long)(nplan * p.maxattempt)
                                       : (long long)(int)((double)(int)(nplan * p.maxattempt) / (1. - ratio));
  long long ntry = 0;
  int64_t done = 0;
  while (done < nplan && (!p.overlapcheck || ntry < maxtry)) {
    const int t = plan[done];
    const double rb = p.template_radius[t];
    double pos[3], v[3];
    if (p.overlapcheck) {
      bool ok = false;
      while (!ok && ntry < maxtry) {
        do {
          candidate(rb, pos);
          ntry++;
        } while (p.check_dist_from_subdomain_border && ntry < maxtry && dist_to_box(boxlo, boxhi, pos) < rb);
        if (ntry == maxtry) break;
        velocity(v);
        if (s->grid.overlaps(pos, rb)) continue;
        s->grid.add(pos, rb);
        ok = true;
      }
      if (!ok) break;
    } else {
      candidate(rb, pos);
      velocity(v);
    }
    {
      for (int d = 0; d < 3; d++) {
        x_out[3 * done + d] = pos[d];
        v_out[3 * done + d] = v[d];
        omega_out[3 * done + d] = p.omega[d];
      }
      radius_out[done] = rb;
      density_out[done] = p.template_density[t];
      rmass_out[done] = p.template_density[t] * (rb * rb * rb * 4. * kPi / 3.);
      type_out[done] = p.template_type[t];
    }
    done++;
  }
  if (ninserted) *ninserted = done;
  return DEM_OK;
}

}  // extern "C"
