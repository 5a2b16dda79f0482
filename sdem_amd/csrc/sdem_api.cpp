// sdem_api.cpp -- the extern "C" entry points declared in include/sdem.h.
//
// Host responsibilities: option checks, the pair-coefficient tables
// (Prtcl_Property.f90:62-247), plane-wall construction (Prtcl_Geometry.f90:
// 457-537), the search-level layout (Init_NBSM / Init_Munjiza_Hrchl), device
// allocation, the step loop and the readers.  All arithmetic that must agree
// with the reference bit for bit is compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/sdem.h"
#include "sdem_internal.h"

namespace sdem {
hipError_t launch_pre(const Params& P, const Buf& A, Scalars* S, Slot* slots, int n_upper,
                      hipStream_t st);
hipError_t launch_finish(Scalars* S, const Slot* slots, hipStream_t st);
hipError_t launch_ctl(const Params& P, Scalars* S, Slot* slots, hipStream_t st);
hipError_t launch_sort(const Params& P, const Buf& A, Scalars* S, int n_upper, int64_t dead_key,
                       int* keys, int* rank, int* counts, int* start, int* tile_sum, int* perm_tmp,
                       int* perm, double4* pk, hipStream_t st);
hipError_t launch_vlist(const Params& P, Scalars* S, int n_upper, const double4* pk,
                        const int* cstart, int* vcnt, int* vlist, hipStream_t st);
hipError_t launch_force(const Params& P, const Buf& A, const Buf& B, Scalars* S, int n_upper,
                        const int* perm, const int* vcnt, const int* vlist, unsigned* clist,
                        int* ccnt, const double4* fpk, const int* fid,
                        const int* ftype, const int* fstart, Slot* slots, hipStream_t st,
                        hipEvent_t mid);
int64_t scan_tiles(int64_t n);
}  // namespace sdem

using namespace sdem;

namespace {

struct PureProp {
  double radius, density, poisson, youngs, mass, volume, inertia;
};

struct V3 {
  double x, y, z;
};
inline V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 vmul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 vdiv(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 vcross(V3 u, V3 v) {
  return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
inline double vnorm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline double fmn(double a, double b) { return a < b ? a : b; }
inline double fmx(double a, double b) { return a > b ? a : b; }
inline V3 vmin(V3 a, V3 b) { return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }

// ---- plane-wall geometry used at construction (Prtcl_Geometry.f90) --------
bool in_plane(V3 p1, V3 p2, V3 p3, V3 p4, V3 cp) {  // PW_IsInPlane :185-221
  V3 a = vsub(p1, cp), b = vsub(p2, cp), c = vsub(p3, cp), d = vsub(p4, cp);
  double p1p3 = vdot(a, c), p2p4 = vdot(b, d), p1p4 = vdot(a, d), p2p3 = vdot(b, c);
  if (p1p3 * p2p4 - p1p4 * p2p3 < 0.0) return false;
  double p1p2 = vdot(a, b), p3p4 = vdot(c, d);
  if (p1p3 * p2p4 - p1p2 * p3p4 < 0.0) return false;
  double p2p2 = vdot(b, b);
  if (p1p2 * p2p3 - p1p3 * p2p2 < 0.0) return false;
  return true;
}

bool line_insert_inner(V3 a1, V3 a2, V3 b1, V3 b2) {  // IsLineInsertInner :331-421
  V3 v1 = vsub(a2, a1), v2 = vsub(b2, b1), l21 = vsub(b1, a1);
  V3 v1xv2 = vcross(v1, v2), l21xv1 = vcross(l21, v1), l21xv2 = vcross(l21, v2);
  double c = vnorm(l21xv1), d = vnorm(l21xv2), e = vnorm(v1xv2);
  if (e < 1.0e-10) {
    double t1 = vdot(l21, v1) / vdot(v1, v1);
    V3 cp = vadd(vmul(v1, t1), a1);
    double dist = vnorm(vsub(cp, b1));
    if (dist > 1.0e-6) return false;
    double t2 = vdot(vsub(b2, a1), v1) / vdot(v1, v1);
    return !((t1 < 0.0 && t2 < 0.0) || (t1 > 1.0 && t2 > 1.0));
  }
  double a = vdot(l21xv2, v1xv2) > 0.0 ? d / e : -d / e;
  double b = vdot(l21xv1, v1xv2) > 0.0 ? c / e : -c / e;
  const double t = 1.0e-10;
  return a >= 0.0 - t && a <= 1.0 + t && b >= 0.0 - t && b <= 1.0 + t;
}

bool reflection_ovlp(const V3 f[4], const V3 s[4]) {  // IsReflectionOvlp :285-326
  for (int k = 0; k < 4; k++)
    if (in_plane(f[0], f[1], f[2], f[3], s[k])) return true;
  for (int k = 0; k < 4; k++)
    if (in_plane(s[0], s[1], s[2], s[3], f[k])) return true;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++)
      if (line_insert_inner(f[a], f[(a + 1) % 4], s[b], s[(b + 1) % 4])) return true;
  return false;
}

bool point_in_domain(V3 p, V3 mn, V3 mx) {  // PointIsInDomain :266-280
  const double e = 1.0e-10;
  if (p.x < mn.x - e || p.x > mx.x + e) return false;
  if (p.y < mn.y - e || p.y > mx.y + e) return false;
  if (p.z < mn.z - e || p.z > mx.z + e) return false;
  return true;
}

bool wall_in_domain(const Wall& w, V3 mn, V3 mx) {  // PW_IsInDomain :114-180
  V3 p[4] = {v3(w.p1), v3(w.p2), v3(w.p3), v3(w.p4)};
  for (int k = 0; k < 4; k++)
    if (point_in_domain(p[k], mn, mx)) return true;
  {
    V3 f[4], s[4] = {{mn.x, mn.y, 0}, {mx.x, mn.y, 0}, {mx.x, mx.y, 0}, {mn.x, mx.y, 0}};
    for (int k = 0; k < 4; k++) f[k] = {p[k].x, p[k].y, 0};
    if (!reflection_ovlp(f, s)) return false;
  }
  {
    V3 f[4], s[4] = {{mn.x, 0, mn.z}, {mn.x, 0, mx.z}, {mx.x, 0, mx.z}, {mx.x, 0, mn.z}};
    for (int k = 0; k < 4; k++) f[k] = {p[k].x, 0, p[k].z};
    if (!reflection_ovlp(f, s)) return false;
  }
  {
    V3 f[4], s[4] = {{0, mn.y, mn.z}, {0, mx.y, mn.z}, {0, mx.y, mx.z}, {0, mn.y, mx.z}};
    for (int k = 0; k < 4; k++) f[k] = {0, p[k].y, p[k].z};
    if (!reflection_ovlp(f, s)) return false;
  }
  V3 dom[8] = {{mn.x, mn.y, mn.z}, {mx.x, mn.y, mn.z}, {mx.x, mx.y, mn.z}, {mn.x, mx.y, mn.z},
               {mn.x, mn.y, mx.z}, {mx.x, mn.y, mx.z}, {mx.x, mx.y, mx.z}, {mn.x, mx.y, mx.z}};
  int sum = 0;
  V3 n = v3(w.n);
  for (int k = 0; k < 8; k++) {
    double t = vdot(n, dom[k]) + w.d;
    sum += t > 1.0e-8 ? 1 : (t < -1.0e-8 ? 0 : -99);
  }
  return !(sum == 0 || sum == 8);
}

// clc_BnryPrtcl_Prop (Prtcl_Property.f90:200-247)
Bnry bnry_prop(int cf_type, const PureProp& pi, const PureProp& pj, double fs, double fk,
               double froll, double en, bool iswall, double lin_kn, double lin_kt_ratio) {
  Bnry b;
  if (!iswall) {
    b.rad_eff = (pi.radius * pj.radius) / (pi.radius + pj.radius);
    b.mass_eff = (pi.mass * pj.mass) / (pi.mass + pj.mass);
  } else {
    b.rad_eff = pi.radius;
    b.mass_eff = pi.mass;
  }
  b.fric_s = fs;
  b.fric_k = fk;
  b.fric_roll = froll;
  const double pri = pi.poisson, prj = pj.poisson;
  const double gi = pi.youngs / (2.0 * (1.0 + pri));
  const double gj = pj.youngs / (2.0 * (1.0 + prj));
  const double e_eff = 1.0 / ((1.0 - pri * pri) / pi.youngs + (1.0 - prj * prj) / pj.youngs);
  const double g_eff = 1.0 / ((2.0 - pri) / gi + (2.0 - prj) / gj);
  double kappa = ((1.0 - pri) / gi + (1.0 - prj) / gj) / ((1.0 - 0.5 * pri) / gi + (1.0 - 0.5 * prj) / gj);
  kappa = std::fabs(kappa);
  double eta = std::log(en);
  eta = eta * eta;
  const double PI = 3.141592653589793238462643383279502884;
  if (cf_type == SDEM_CFT_LSD) {
    b.kn = 1.2024 * std::pow(std::sqrt(b.mass_eff) * (e_eff * e_eff) * b.rad_eff, 0.4);
    b.kt = kappa * b.kn;
    b.dn = -2.0 * std::log(en) * std::sqrt(b.mass_eff * b.kn) / std::sqrt(PI * PI + eta);
  } else if (cf_type == SDEM_CFT_LINEAR) {
    // no counterpart in the reference: constant kn, kt = ratio * kn, and the
    // damping of the LSD formula at unit velocity factor
    b.kn = lin_kn;
    b.kt = lin_kt_ratio * lin_kn;
    b.dn = -2.0 * std::log(en) * std::sqrt(b.mass_eff * b.kn) / std::sqrt(PI * PI + eta);
  } else {
    const double kh = 4.0 / 3.0 * e_eff * std::sqrt(b.rad_eff);
    b.kn = kh;
    b.kt = 16.0 / 3.0 * g_eff * std::sqrt(b.rad_eff);
    b.dn = -2.2664 * std::log(en) * std::sqrt(b.mass_eff * kh) / std::sqrt(eta + 10.1354);
  }
  return b;
}

int level_size(int m) { return m == SDEM_PIM_FE ? 1 : (m == SDEM_PIM_AB2 ? 2 : 3); }

}  // namespace

struct sdem_ctx {
  sdem_options o{};
  std::string err;
  int tsize = 1, rsize = 1, ta = 1, tr = 1;
  hipStream_t stream = nullptr;

  int ntype = 0, nwtype = 0;
  std::vector<PureProp> pprop, wprop;
  std::vector<Bnry> pp_tab, pw_tab;
  bool have_pprop = false, have_wprop = false;
  std::vector<Wall> walls;
  int num_pwall_total = 0;

  std::vector<int> fix_id, fix_type;
  std::vector<double> fix_pos;

  // device
  int cap = 0, hcap = 0;
  Buf buf[2]{};
  int cur = 0;
  Scalars* S = nullptr;
  int *keys = nullptr, *rank = nullptr, *counts = nullptr, *start = nullptr;
  int *perm_tmp = nullptr, *perm = nullptr;
  double4* pk = nullptr;
  int *vcnt = nullptr, *vlist = nullptr;
  unsigned* clist = nullptr;
  int* ccnt = nullptr;
  Slot* slots = nullptr;
  int lcap = 0;
  double cell_factor[3] = {1.0, 1.0, 1.0};
  double skin_factor = 0.1;  // candidate-list skin as a fraction of d_max
  int keep_forces = 1;
  bool first_step_pending = false;
  double lin_kn = 0.0, lin_kt_ratio = 0.0;  // constant-stiffness linear model
  bool need_pre = false;  // the state in buffer A has no near-wall pass yet
  bool use_graph = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // two steps, by starting buffer
  int hist_slots_hint = 0;
  int list_slots_hint = 0;
  int* tile_sum = nullptr;
  Bnry *d_pp = nullptr, *d_pw = nullptr;
  double *d_mass = nullptr, *d_inertia = nullptr;
  Wall* d_walls = nullptr;
  int* d_type_level = nullptr;
  double4* d_fpk = nullptr;
  int *d_fid = nullptr, *d_ftype = nullptr, *d_fstart = nullptr;
  Params P{};
  int64_t dead_key = 0;
  int n_upper = 0;
  bool have_state = false, stepped = false, prepared = false;
  std::vector<double> host_lacc1, host_racc1;  // level-1 accelerations before the first step
  std::vector<double> level_rmax;              // per level, from the state
  std::vector<int> type_level;

  bool prof = false;
  double prof_ms[SDEM_NUM_PHASES] = {};
  int64_t prof_n[SDEM_NUM_PHASES] = {};
  std::vector<void*> allocs_state, allocs_grid, allocs_tab;
  int64_t dev_bytes = 0;

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

namespace {

int hip_fail(sdem_ctx* c, hipError_t e, const char* where) {
  return c->fail(SDEM_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

void invalidate_graphs(sdem_ctx* c) {
  for (auto& g : c->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

template <class T>
hipError_t dalloc(sdem_ctx* c, std::vector<void*>& list, T** p, size_t n) {
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e == hipSuccess) {
    list.push_back(*p);
    c->dev_bytes += (int64_t)bytes;
  }
  return e;
}

void free_list(sdem_ctx* c, std::vector<void*>& list) {
  for (void* p : list) (void)hipFree(p);
  list.clear();
  (void)c;
}

double max_type_radius(const sdem_ctx* c) {
  double m = c->pprop[0].radius;
  for (int i = 1; i < c->ntype; i++) m = fmx(m, c->pprop[i].radius);
  return m;
}
double min_type_radius(const sdem_ctx* c) {
  double m = c->pprop[0].radius;
  for (int i = 1; i < c->ntype; i++) m = fmn(m, c->pprop[i].radius);
  return m;
}

// Level of each type: one level for CS_Method=1, the reference's hierarchy
// (Init_Munjiza_Hrchl, Prtcl_Hrchl_Munjiza.f90:84-237) for CS_Method=2.
// Returns per-level cell lengths.
std::vector<double> plan_levels(sdem_ctx* c, const std::vector<int>& count_per_type) {
  std::vector<double> cells;
  c->type_level.assign(c->ntype, 0);
  const double maxd = 2.0 * max_type_radius(c), mind = 2.0 * min_type_radius(c);
  double cell_len = c->o.prtcl_cs_ratio * maxd;
  if (c->o.cs_method != SDEM_CSM_NBS_MUNJIZA_HRCHL) {
    cells.push_back(cell_len);
    return cells;
  }
  int dmax_min = (int)(maxd / mind);
  int numlevels;
  if (c->o.cs_numlvls <= 0) {
    if (dmax_min <= 1) numlevels = 1;
    else if (dmax_min <= 3) numlevels = 2;
    else if (dmax_min <= 7) numlevels = 3;
    else if (dmax_min <= 15) numlevels = 4;
    else if (dmax_min <= 31) numlevels = 5;
    else numlevels = 6;
  } else {
    numlevels = std::min(c->o.cs_numlvls, MAX_LEVELS);
  }
  double maxd_lvl = maxd;
  for (int idh = 1; idh <= numlevels; idh++) {
    double mind_lvl = maxd_lvl / 2.0;
    if (idh == numlevels) mind_lvl = 0.0;
    int num = 0;
    for (int i = 0; i < c->ntype; i++) {
      double diam = 2.0 * c->pprop[i].radius;
      if (diam > mind_lvl && (idh == 1 || diam <= maxd_lvl)) num += count_per_type[i];
    }
    if (num > 0) {
      int lvl = (int)cells.size();
      for (int i = 0; i < c->ntype; i++) {
        double diam = 2.0 * c->pprop[i].radius;
        if (diam > mind_lvl && (idh == 1 || diam <= maxd_lvl)) c->type_level[i] = lvl;
      }
      cells.push_back(cell_len);
    }
    maxd_lvl = mind_lvl;
    cell_len = cell_len / 2.0;
  }
  if (cells.empty()) cells.push_back(c->o.prtcl_cs_ratio * maxd);
  return cells;
}

int make_grid(sdem_ctx* c, double cell_len0, double rmax, int64_t key_base, Grid* G) {
  G->key_base = key_base;
  G->rmax = rmax;
  G->ncell = 1;
  for (int d = 0; d < 3; d++) {
    double L = c->o.domain_max[d] - c->o.domain_min[d];
    double cell_len = cell_len0 * c->cell_factor[d];
    G->origin[d] = c->o.domain_min[d];
    if (c->o.is_periodic[d]) {
      int n = (int)std::floor(L / cell_len);
      // the search range of one particle must not wrap onto itself
      double reach = 2.0 * max_type_radius(c) + rmax;
      int span = (int)std::floor(2.0 * reach * (1.0 + 1e-6) / (L / n)) + 2;
      if (n < 3 || span > n)
        return c->fail(SDEM_E_ARG, "periodic dimension too short for the search range");
      G->n[d] = n;
      G->w[d] = L / n;
      G->inv_w[d] = (double)n / L;
    } else {
      int n = (int)std::ceil(L / cell_len);
      if (n < 1) n = 1;
      G->n[d] = n;
      G->w[d] = cell_len;
      G->inv_w[d] = 1.0 / cell_len;
    }
    G->ncell *= G->n[d];
  }
  if (G->ncell > (int64_t)1 << 30) return c->fail(SDEM_E_ARG, "search grid too large");
  return SDEM_OK;
}

}  // namespace

// ===========================================================================
// The entry points keep C linkage from their declarations in sdem.h.

int sdem_options_default(sdem_options* o) {
  if (!o) return SDEM_E_ARG;
  std::memset(o, 0, sizeof(*o));
  // DEMS_Options defaults, Prtcl_Parameters.f90:45-95
  o->dt = 1.0e-5;
  o->gravity[1] = -9.81;
  o->prtcl_cs_ratio = 1.0;
  o->wall_neighbor_ratio = 2.0;
  o->cs_method = SDEM_CSM_NBS_MUNJIZA;
  o->cf_type = SDEM_CFT_LSD;
  o->ct_model = SDEM_CTM_CONSTANT_TORQUE;
  o->pi_method = SDEM_PIM_FE;
  o->pri_method = SDEM_PIM_FE;
  o->cntct_list_size = 6;
  o->base_wall_id = 10000000;
  o->wall_max_update_iter = 100;
  o->num_prtcl = 8000;
  o->ifirst = 1;
  return SDEM_OK;
}

int sdem_create(const sdem_options* o, sdem_ctx** out) {
  if (!o || !out) return SDEM_E_ARG;
  *out = nullptr;
  if (o->pi_method < 1 || o->pi_method > 3 || o->pri_method < 1 || o->pri_method > 3 ||
      o->cf_type < 1 || o->cf_type > 3 || o->ct_model < 1 || o->ct_model > 2 ||
      o->cs_method < 1 || o->cs_method > 2 || o->num_prtcl < 1 || !(o->dt > 0.0) ||
      !(o->prtcl_cs_ratio >= 1.0) || o->wall_max_update_iter < 1)
    return SDEM_E_ARG;
  for (int d = 0; d < 3; d++)
    if (!(o->domain_max[d] > o->domain_min[d])) return SDEM_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SDEM_E_NODEV;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SDEM_E_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return SDEM_E_NODEV;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SDEM_E_NODEV;
  sdem_ctx* c = new sdem_ctx();
  c->o = *o;
  c->tsize = level_size(o->pi_method);
  c->rsize = level_size(o->pri_method);
  c->ta = c->tsize == 1 ? 1 : c->tsize - 1;
  c->tr = c->rsize == 1 ? 1 : c->rsize - 1;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return SDEM_E_HIP;
  }
  *out = c;
  return SDEM_OK;
}

int sdem_set_linear_stiffness(sdem_ctx* c, double kn, double kt_over_kn) {
  if (!c) return SDEM_E_ARG;
  if (c->o.cf_type != SDEM_CFT_LINEAR) return c->fail(SDEM_E_STATE, "cf_type is not the linear model");
  if (!(kn > 0.0) || !(kt_over_kn >= 0.0)) return c->fail(SDEM_E_ARG, "bad linear stiffness");
  if (c->have_pprop) return c->fail(SDEM_E_STATE, "stiffness must be set before the properties");
  c->lin_kn = kn;
  c->lin_kt_ratio = kt_over_kn;
  return SDEM_OK;
}

int sdem_set_particle_property(sdem_ctx* c, int32_t ntype, const double* diameter,
                               const double* density, const double* youngs, const double* poisson,
                               double fs, double fk, double froll, double en, double et) {
  (void)et;
  if (!c) return SDEM_E_ARG;
  if (ntype < 1 || ntype > MAX_TYPES || !diameter || !density || !youngs || !poisson)
    return c->fail(SDEM_E_ARG, "bad particle property arguments");
  if (c->have_state) return c->fail(SDEM_E_STATE, "properties must be set before the state");
  if (c->o.cf_type == SDEM_CFT_LINEAR && !(c->lin_kn > 0.0))
    return c->fail(SDEM_E_STATE, "linear model: set the stiffness first");
  const double PI = 3.141592653589793238462643383279502884;
  c->ntype = ntype;
  c->pprop.assign(ntype, PureProp{});
  for (int i = 0; i < ntype; i++) {
    PureProp& p = c->pprop[i];
    p.density = density[i];
    p.radius = 0.5 * diameter[i];
    p.youngs = youngs[i];
    p.poisson = poisson[i];
    p.volume = 4.0 / 3.0 * PI * (p.radius * p.radius * p.radius);
    p.mass = density[i] * p.volume;
    p.inertia = 2.0 / 5.0 * p.mass * (p.radius * p.radius);
    if (!(p.radius > 0.0) || !(p.mass > 0.0)) return c->fail(SDEM_E_ARG, "non-positive diameter or density");
  }
  c->pp_tab.assign((size_t)ntype * ntype, Bnry{});
  for (int i = 0; i < ntype; i++)
    for (int j = 0; j < ntype; j++)  // pari = parj = type i (Prtcl_Property.f90:137-138)
      // the reference's table takes the row type only (Prtcl_Property.f90:137-138);
      // the linear model, which it does not have, takes the true pair
      c->pp_tab[(size_t)i * ntype + j] =
          bnry_prop(c->o.cf_type, c->pprop[i], c->o.cf_type == SDEM_CFT_LINEAR ? c->pprop[j] : c->pprop[i],
                    fs, fk, froll, en, false, c->lin_kn, c->lin_kt_ratio);
  c->have_pprop = true;
  c->have_wprop = false;
  return SDEM_OK;
}

int sdem_set_wall_property(sdem_ctx* c, int32_t nwtype, const double* youngs, const double* poisson,
                           double fs, double fk, double froll, double en, double et) {
  (void)et;
  if (!c) return SDEM_E_ARG;
  if (!c->have_pprop) return c->fail(SDEM_E_STATE, "particle property first");
  if (nwtype < 1 || !youngs || !poisson) return c->fail(SDEM_E_ARG, "bad wall property arguments");
  if (c->have_state) return c->fail(SDEM_E_STATE, "properties must be set before the state");
  c->nwtype = nwtype;
  c->wprop.assign(nwtype, PureProp{});
  for (int i = 0; i < nwtype; i++) {
    PureProp& w = c->wprop[i];
    w.density = 1.0e50; w.radius = 1.0e50; w.youngs = youngs[i]; w.poisson = poisson[i];
    w.mass = 1.0e50; w.volume = 1.0e50; w.inertia = 1.0e50;
  }
  c->pw_tab.assign((size_t)c->ntype * nwtype, Bnry{});
  for (int i = 0; i < c->ntype; i++)
    for (int j = 0; j < nwtype; j++)
      c->pw_tab[(size_t)i * nwtype + j] =
          bnry_prop(c->o.cf_type, c->pprop[i], c->wprop[j], fs, fk, froll, en, true, c->lin_kn, c->lin_kt_ratio);
  c->have_wprop = true;
  return SDEM_OK;
}

int sdem_add_plane_wall(sdem_ctx* c, const double p1[3], const double p2[3], const double p3[3],
                        const double p4[3], int32_t user_id, int32_t wall_type, int32_t both_side,
                        int32_t is_infinite, const double trans_vel[3], int32_t* wall_id) {
  if (!c || !p1 || !p2 || !p3 || !p4) return SDEM_E_ARG;
  if (!c->have_pprop) return c->fail(SDEM_E_STATE, "particle property first");
  if (c->have_state) return c->fail(SDEM_E_STATE, "walls must be added before the state");
  (void)user_id;
  Wall w;
  std::memset(&w, 0, sizeof(w));
  for (int k = 0; k < 3; k++) { w.p1[k] = p1[k]; w.p2[k] = p2[k]; w.p3[k] = p3[k]; w.p4[k] = p4[k]; }
  V3 P1 = v3(p1), P2 = v3(p2), P3 = v3(p3), P4 = v3(p4);
  V3 ln = vcross(vsub(P2, P1), vsub(P3, P1));
  V3 mn = vmin(P4, vmin(P3, vmin(P2, P1))), mx = vmax(P4, vmax(P3, vmax(P2, P1)));
  V3 n = vdiv(ln, vnorm(ln));
  w.n[0] = n.x; w.n[1] = n.y; w.n[2] = n.z;
  w.d = -(vdot(n, P1));
  w.min_point[0] = mn.x; w.min_point[1] = mn.y; w.min_point[2] = mn.z;
  w.max_point[0] = mx.x; w.max_point[1] = mx.y; w.max_point[2] = mx.z;
  if (trans_vel) { w.trans_vel[0] = trans_vel[0]; w.trans_vel[1] = trans_vel[1]; w.trans_vel[2] = trans_vel[2]; }
  w.both_side = both_side ? 1 : 0;
  w.is_infinite = is_infinite ? 1 : 0;
  w.wall_type = wall_type;
  if (wall_type < 1) return c->fail(SDEM_E_ARG, "wall_type must be >= 1");
  if (std::fabs(vdot(n, P4) + w.d) >= 0.00001) return c->fail(SDEM_E_ARG, "Cannot create a plane wall");
  if (wall_id) *wall_id = 0;
  double lenexp = 1.2 * max_type_radius(c);
  V3 dmin = v3(c->o.domain_min), dmax = v3(c->o.domain_max);
  V3 one = {1.0, 1.0, 1.0};
  V3 pmin = vsub(dmin, vmul(one, lenexp)), pmax = vadd(dmax, vmul(one, lenexp));
  if (!wall_in_domain(w, pmin, pmax)) return SDEM_OK;  // skipped, as G_add_PlaneWall does
  if ((int)c->walls.size() >= MAX_WALLS) return c->fail(SDEM_E_CAPACITY, "more than 64 walls");
  c->num_pwall_total++;
  w.wall_id = c->num_pwall_total + c->o.base_wall_id;
  if (w.wall_id > KEY_MASK) return c->fail(SDEM_E_ARG, "wall id exceeds 2^29-1");
  c->walls.push_back(w);
  if (wall_id) *wall_id = w.wall_id;
  return SDEM_OK;
}

int sdem_set_fixed_particles(sdem_ctx* c, int32_t nfix, const int32_t* id, const int32_t* ptype,
                             const double* posr) {
  if (!c || nfix < 0 || (nfix > 0 && (!id || !ptype || !posr))) return SDEM_E_ARG;
  if (!c->have_pprop) return c->fail(SDEM_E_STATE, "particle property first");
  if (c->have_state) return c->fail(SDEM_E_STATE, "fixed particles must be set before the state");
  c->fix_id.assign(id, id + nfix);
  c->fix_type.assign(ptype, ptype + nfix);
  c->fix_pos.assign(posr, posr + 4 * (size_t)nfix);
  for (int i = 0; i < nfix; i++) {
    if (ptype[i] < 1 || ptype[i] > c->ntype) return c->fail(SDEM_E_ARG, "fixed ptype out of range");
    if (id[i] < 0 || id[i] > KEY_MASK) return c->fail(SDEM_E_ARG, "fixed id out of range");
  }
  return SDEM_OK;
}

static int alloc_buf(sdem_ctx* c, Buf& b) {
  size_t cap = (size_t)c->cap, hc = (size_t)c->hcap;
  hipError_t e = hipSuccess;
  auto A = [&](auto** p, size_t n) { if (e == hipSuccess) e = dalloc(c, c->allocs_state, p, n); };
  A(&b.pos, cap);
  A(&b.lv, cap * 3 * c->tsize); A(&b.la, cap * 3 * c->ta);
  A(&b.th, cap * 3); A(&b.rv, cap * 3 * c->rsize); A(&b.ra, cap * 3 * c->tr);
  A(&b.id, cap); A(&b.type, cap); A(&b.mark, cap); A(&b.wmask, cap);
  A(&b.hcnt, cap); A(&b.hkey, cap * hc); A(&b.hval, cap * hc);
  A(&b.force, cap * 6);
  return e == hipSuccess ? SDEM_OK : hip_fail(c, e, "hipMalloc(state)");
}

// history slots per particle: a kissing-number bound for the largest type
// touched by the smallest, with room for soft overlaps and walls, at most 48
// to start with (an overflow doubles it and replays the step)
static int history_capacity(const sdem_ctx* c) {
  double rmax = max_type_radius(c), rmin = min_type_radius(c);
  double s = rmin / (rmax + rmin);
  double cosv = std::sqrt(std::max(0.0, 1.0 - s * s));
  double bound = 2.0 / (1.0 - cosv);
  int h = (int)std::ceil(1.25 * bound) + 4 + (int)c->walls.size();
  h = std::max(h, 2 * c->o.cntct_list_size);
  h = std::max(h, 16);
  return std::min(h, 48);
}

int sdem_set_state(sdem_ctx* c, int32_t n, const int32_t* id, const int32_t* ptype,
                   const int32_t* usr_mark, const double* posr, const double* lin_vel,
                   const double* lin_acc, const double* theta, const double* rot_vel,
                   const double* rot_acc) {
  if (!c) return SDEM_E_ARG;
  if (!c->have_pprop || !c->have_wprop) return c->fail(SDEM_E_STATE, "particle and wall properties first");
  if (n < 0 || (n > 0 && (!id || !ptype || !posr || !lin_vel || !lin_acc || !theta || !rot_vel || !rot_acc)))
    return c->fail(SDEM_E_ARG, "bad state arguments");
  if (n > c->o.num_prtcl) return c->fail(SDEM_E_CAPACITY, "n exceeds numPrtcl");
  for (int i = 0; i < n; i++) {
    if (ptype[i] < 1 || ptype[i] > c->ntype) return c->fail(SDEM_E_ARG, "ptype out of range");
    if (id[i] < 0 || id[i] > KEY_MASK) return c->fail(SDEM_E_ARG, "particle id out of range");
    if (!(posr[4 * i + 3] > 0.0)) return c->fail(SDEM_E_ARG, "non-positive radius");
    for (int d = 0; d < 3; d++) {
      double p = posr[4 * i + d];
      if (!(p >= c->o.domain_min[d] && p < c->o.domain_max[d]))
        return c->fail(SDEM_E_ARG, "particle outside the domain");
    }
  }
  // (re)allocate state buffers on first use
  if (c->allocs_state.empty()) {
    c->cap = c->o.num_prtcl;
    c->hcap = c->hist_slots_hint > 0 ? c->hist_slots_hint : history_capacity(c);
    for (int b = 0; b < 2; b++)
      if (int r = alloc_buf(c, c->buf[b])) return r;
    hipError_t e = dalloc(c, c->allocs_state, &c->S, 1);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->keys, c->cap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->rank, c->cap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->perm_tmp, c->cap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->perm, c->cap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->pk, c->cap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->vcnt, (size_t)c->cap + 2);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->clist, (size_t)c->cap * c->hcap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->ccnt, c->cap);
    c->lcap = c->list_slots_hint > 0 ? c->list_slots_hint : std::max(48, 2 * c->hcap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->vlist, (size_t)c->cap * c->lcap);
    if (e == hipSuccess) e = dalloc(c, c->allocs_state, &c->slots, NSLOT);
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(sort)");
  }
  const int cap = c->cap, ts = c->tsize, rs = c->rsize;
  // SoA transposition on the host
  std::vector<double> x(n), y(n), z(n), r(n), lv((size_t)3 * ts * cap, 0.0), la((size_t)3 * c->ta * cap, 0.0);
  std::vector<double> th((size_t)3 * cap, 0.0), rv((size_t)3 * rs * cap, 0.0), ra((size_t)3 * c->tr * cap, 0.0);
  std::vector<int> mark(n, 1), hc(n, 0);
  std::vector<uint64_t> wm(n, 0);
  c->host_lacc1.assign((size_t)3 * n, 0.0);
  c->host_racc1.assign((size_t)3 * n, 0.0);
  double host_max_a = 0.0;
  for (int i = 0; i < n; i++) {
    x[i] = posr[4 * i]; y[i] = posr[4 * i + 1]; z[i] = posr[4 * i + 2]; r[i] = posr[4 * i + 3];
    if (usr_mark) mark[i] = usr_mark[i];
    for (int l = 0; l < ts; l++)
      for (int d = 0; d < 3; d++) lv[((size_t)l * 3 + d) * cap + i] = lin_vel[((size_t)i * ts + l) * 3 + d];
    // stored acceleration levels: FE keeps level 1, AB2/AB3 keep levels 2..tsize
    for (int l = 0; l < c->ta; l++) {
      int src = ts == 1 ? 0 : l + 1;
      for (int d = 0; d < 3; d++) la[((size_t)l * 3 + d) * cap + i] = lin_acc[((size_t)i * ts + src) * 3 + d];
    }
    for (int d = 0; d < 3; d++) th[(size_t)d * cap + i] = theta[3 * i + d];
    for (int l = 0; l < rs; l++)
      for (int d = 0; d < 3; d++) rv[((size_t)l * 3 + d) * cap + i] = rot_vel[((size_t)i * rs + l) * 3 + d];
    for (int l = 0; l < c->tr; l++) {
      int src = rs == 1 ? 0 : l + 1;
      for (int d = 0; d < 3; d++) ra[((size_t)l * 3 + d) * cap + i] = rot_acc[((size_t)i * rs + src) * 3 + d];
    }
    for (int d = 0; d < 3; d++) {
      c->host_lacc1[3 * i + d] = lin_acc[((size_t)i * ts) * 3 + d];
      c->host_racc1[3 * i + d] = rot_acc[((size_t)i * rs) * 3 + d];
    }
    const double* a = lin_acc + (size_t)i * ts * 3;
    host_max_a = fmx(host_max_a, std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]));
  }
  Buf& A = c->buf[0];
  c->cur = 0;
  hipError_t e = hipSuccess;
  auto H = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  std::vector<double4> p4(n);
  for (int i = 0; i < n; i++) p4[i] = make_double4(x[i], y[i], z[i], r[i]);
  H(A.pos, p4.data(), (size_t)n * sizeof(double4));
  H(A.lv, lv.data(), lv.size() * 8); H(A.la, la.data(), la.size() * 8); H(A.th, th.data(), th.size() * 8);
  H(A.rv, rv.data(), rv.size() * 8); H(A.ra, ra.data(), ra.size() * 8);
  H(A.id, id, n * 4); H(A.type, ptype, n * 4); H(A.mark, mark.data(), n * 4);
  H(A.wmask, wm.data(), n * 8); H(A.hcnt, hc.data(), n * 4);
  Scalars s;
  std::memset(&s, 0, sizeof(s));
  s.n_stored = n;
  s.n_alive = n;
  s.iter = c->o.ifirst - 1;
  s.next_wall_update = 0;
  s.use_host_max_a = 1;
  s.host_max_a = host_max_a;
  s.need_rebuild = 1;
  H(c->S, &s, sizeof(s));
  if (e == hipSuccess) e = hipMemset(c->slots, 0, sizeof(Slot) * NSLOT);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(state)");
  // search levels: radii never change, so the per-level largest radius is
  // fixed by the state
  std::vector<int> cnt(c->ntype, 0);
  for (int i = 0; i < n; i++) cnt[ptype[i] - 1]++;
  std::vector<double> cells = plan_levels(c, cnt);
  c->level_rmax.assign(cells.size(), 0.0);
  for (int i = 0; i < n; i++) {
    int L = c->type_level[ptype[i] - 1];
    c->level_rmax[L] = fmx(c->level_rmax[L], r[i]);
  }
  c->n_upper = n;
  c->have_state = true;
  c->stepped = false;
  c->prepared = false;
  c->first_step_pending = true;
  c->need_pre = true;
  // plan the grids now so configuration errors surface here
  free_list(c, c->allocs_grid);
  int64_t base = 0;
  c->P.nlvl = (int)cells.size();
  for (int L = 0; L < c->P.nlvl; L++) {
    double rm = c->level_rmax[L] > 0.0 ? c->level_rmax[L] : 0.0;
    if (int rc = make_grid(c, cells[L], rm, base, &c->P.lvl[L])) return rc;
    base += c->P.lvl[L].ncell;
  }
  if (base + 2 > INT32_MAX) return c->fail(SDEM_E_ARG, "search grid too large");
  c->dead_key = base;
  e = dalloc(c, c->allocs_grid, &c->counts, (size_t)base + 2);
  if (e == hipSuccess) e = hipMemset(c->counts, 0, ((size_t)base + 2) * sizeof(int));
  if (e == hipSuccess) e = dalloc(c, c->allocs_grid, &c->start, (size_t)base + 2);
  int64_t nt = scan_tiles(base + 2);
  if (e == hipSuccess) e = dalloc(c, c->allocs_grid, &c->tile_sum, (size_t)nt + 1);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(grid)");
  return SDEM_OK;
}

// history records (Add_RestartCntctlink-style install; replaces all history)
int sdem_set_contacts(sdem_ctx* c, int64_t m, const int32_t* kind, const int32_t* owner_id,
                      const int32_t* partner_id, const double* tan_delta, const double* vrel_init) {
  if (!c) return SDEM_E_ARG;
  if (!c->have_state) return c->fail(SDEM_E_STATE, "state first");
  if (m < 0 || (m > 0 && (!kind || !owner_id || !partner_id || !tan_delta || !vrel_init)))
    return c->fail(SDEM_E_ARG, "bad contact arguments");
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "sync");
  Buf& A = c->buf[c->cur];
  int n = c->n_upper;
  std::vector<int> ids(n), types(n);
  e = hipMemcpy(ids.data(), A.id, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(types.data(), A.type, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(ids)");
  std::vector<std::pair<int, int>> idx;
  idx.reserve(n);
  for (int s = 0; s < n; s++)
    if (types[s] >= 1) idx.push_back({ids[s], s});
  std::sort(idx.begin(), idx.end());
  std::vector<int> hcnt(n, 0), hkey((size_t)n * c->hcap, 0);
  std::vector<double4> hval((size_t)n * c->hcap);  // record q of particle s at q*n+s here
  for (int64_t k = 0; k < m; k++) {
    auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(owner_id[k], -1));
    if (it == idx.end() || it->first != owner_id[k]) return c->fail(SDEM_E_ARG, "owner id not present");
    if (kind[k] < 1 || kind[k] > 4) return c->fail(SDEM_E_ARG, "bad contact kind");
    if (partner_id[k] < 0 || partner_id[k] > KEY_MASK) return c->fail(SDEM_E_ARG, "partner id out of range");
    int s = it->second;
    if (hcnt[s] >= c->hcap) return c->fail(SDEM_E_CAPACITY, "too many history records for one particle");
    size_t slot = (size_t)hcnt[s]++ * n + s;
    hkey[slot] = partner_id[k] | (kind[k] << KEY_BITS);
    hval[slot] = make_double4(tan_delta[3 * k], tan_delta[3 * k + 1], tan_delta[3 * k + 2], vrel_init[k]);
  }
  e = hipMemcpy(A.hcnt, hcnt.data(), (size_t)n * 4, hipMemcpyHostToDevice);
  for (int q = 0; q < c->hcap && e == hipSuccess; q++) {
    e = hipMemcpy(A.hkey + (size_t)q * c->cap, hkey.data() + (size_t)q * n, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(A.hval + (size_t)q * c->cap, hval.data() + (size_t)q * n, (size_t)n * sizeof(double4),
                    hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(contacts)");
  return SDEM_OK;
}

static int prepare(sdem_ctx* c) {
  if (c->prepared) return SDEM_OK;
  invalidate_graphs(c);
  free_list(c, c->allocs_tab);
  Params& P = c->P;
  P.dt = c->o.dt;
  for (int d = 0; d < 3; d++) {
    P.g[d] = c->o.gravity[d];
    P.dmin[d] = c->o.domain_min[d];
    P.dmax[d] = c->o.domain_max[d];
    P.simlen[d] = c->o.domain_max[d] - c->o.domain_min[d];
    P.periodic[d] = c->o.is_periodic[d] ? 1 : 0;
  }
  double maxd = 2.0 * max_type_radius(c);
  double lenforcs = c->o.prtcl_cs_ratio * maxd;
  for (int d = 0; d < 3; d++)
    if (1.02 * lenforcs > P.simlen[d]) return c->fail(SDEM_E_ARG, "so big Diameter");
  P.cf_type = c->o.cf_type;
  P.ct_model = c->o.ct_model;
  P.pi_method = c->o.pi_method;
  P.pri_method = c->o.pri_method;
  P.tsize = c->tsize; P.rsize = c->rsize; P.ta = c->ta; P.tr = c->tr;
  P.cap = c->cap;
  P.hcap = c->hcap;
  P.ntype = c->ntype;
  P.nwtype = c->nwtype;
  P.nwall = (int)c->walls.size();
  P.search_fixed = c->o.cs_method == SDEM_CSM_NBS_MUNJIZA ? 1 : 0;
  P.keep_forces = c->keep_forces;
  double rmax = max_type_radius(c);
  P.dxmax1 = rmax * c->o.wall_neighbor_ratio;
  P.dxmax2 = rmax * (c->o.wall_neighbor_ratio + 1.0);
  double mwv = 0.0;
  for (auto& w : c->walls) mwv = fmx(mwv, vnorm(v3(w.trans_vel)));
  P.max_wall_vel = mwv;
  P.wall_max_update_iter = c->o.wall_max_update_iter;
  P.skin = c->skin_factor * maxd;
  P.rebuild_disp = 0.5 * P.skin * (1.0 - 1.0e-6);
  P.lcap = c->lcap;
  for (auto& w : c->walls)
    if (w.wall_type > c->nwtype) return c->fail(SDEM_E_ARG, "wall_type exceeds numWall_type");
  std::vector<double> mass(c->ntype), inertia(c->ntype);
  for (int i = 0; i < c->ntype; i++) { mass[i] = c->pprop[i].mass; inertia[i] = c->pprop[i].inertia; }
  hipError_t e = hipSuccess;
  auto U = [&](auto** dst, const auto& vec) {
    using T = typename std::remove_reference<decltype(vec)>::type::value_type;
    if (e == hipSuccess) e = dalloc(c, c->allocs_tab, dst, vec.size());
    if (e == hipSuccess && !vec.empty()) e = hipMemcpy(*dst, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice);
  };
  U(&c->d_pp, c->pp_tab);
  U(&c->d_pw, c->pw_tab);
  U(&c->d_mass, mass);
  U(&c->d_inertia, inertia);
  U(&c->d_walls, c->walls);
  U(&c->d_type_level, c->type_level);
  // fixed particles, sorted by the cells of a single grid
  int nfix = (int)c->fix_id.size();
  P.nfix = nfix;
  double fr = 0.0;
  for (int i = 0; i < nfix; i++) fr = fmx(fr, c->fix_pos[4 * i + 3]);
  if (int rc = make_grid(c, c->o.prtcl_cs_ratio * maxd, fr, 0, &P.fgrid)) return rc;
  {
    Grid& G = P.fgrid;
    std::vector<int64_t> key(nfix);
    auto cl = [&](int d, double x) {
      int cc = (int)std::floor((x - G.origin[d]) * G.inv_w[d]);
      return cc < 0 ? 0 : (cc >= G.n[d] ? G.n[d] - 1 : cc);
    };
    for (int i = 0; i < nfix; i++) {
      const double* p = &c->fix_pos[4 * i];
      key[i] = ((int64_t)cl(2, p[2]) * G.n[1] + cl(1, p[1])) * G.n[0] + cl(0, p[0]);
    }
    std::vector<int> ord(nfix);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<double4> fpk(nfix);
    std::vector<int> fid(nfix), ftype(nfix), fstart(nfix > 0 ? G.ncell + 1 : 1, 0);
    for (int k = 0; k < nfix; k++) {
      int i = ord[k];
      const double* p = &c->fix_pos[4 * i];
      fpk[k] = make_double4(p[0], p[1], p[2], p[3]);
      fid[k] = c->fix_id[i];
      ftype[k] = c->fix_type[i];
    }
    if (nfix > 0) {
      for (int i = 0; i < nfix; i++) fstart[key[i] + 1]++;
      for (int64_t q = 0; q < G.ncell; q++) fstart[q + 1] += fstart[q];
    }
    U(&c->d_fpk, fpk);
    U(&c->d_fid, fid);
    U(&c->d_ftype, ftype);
    U(&c->d_fstart, fstart);
  }
  if (e != hipSuccess) return hip_fail(c, e, "upload tables");
  P.pp = c->d_pp;
  P.pw = c->d_pw;
  P.mass = c->d_mass;
  P.inertia = c->d_inertia;
  P.walls = c->d_walls;
  P.type_level = c->d_type_level;
  c->prepared = true;
  return SDEM_OK;
}

template <class T>
static hipError_t dl(std::vector<T>& v, const T* src, size_t n) {
  v.resize(n);
  if (!n) return hipSuccess;
  return hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost);
}

// Replace a tracked device allocation by a larger one (contents not kept).
template <class T>
static hipError_t regrow(sdem_ctx* c, T** p, size_t n) {
  auto& L = c->allocs_state;
  auto it = std::find(L.begin(), L.end(), (void*)*p);
  if (it != L.end()) L.erase(it);
  (void)hipFree(*p);
  *p = nullptr;
  return dalloc(c, L, p, n);
}

// History slots per particle doubled.  Record q of particle s sits at
// q*cap+s, so the records of buffer A keep their places in the larger arrays.
static int grow_history(sdem_ctx* c) {
  const int old = c->hcap, neu = std::min(old * 2, 4096);
  if (neu <= old) return c->fail(SDEM_E_CAPACITY, "contact history slots exhausted for a particle");
  const size_t cap = (size_t)c->cap;
  Buf& A = c->buf[c->cur];
  int* hk = nullptr;
  double4* hv = nullptr;
  hipError_t e = dalloc(c, c->allocs_state, &hk, cap * neu);
  if (e == hipSuccess) e = dalloc(c, c->allocs_state, &hv, cap * neu);
  if (e == hipSuccess) e = hipMemcpy(hk, A.hkey, cap * old * sizeof(int), hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = hipMemcpy(hv, A.hval, cap * old * sizeof(double4), hipMemcpyDeviceToDevice);
  if (e != hipSuccess) return hip_fail(c, e, "grow history");
  auto drop = [&](void* p) {
    auto it = std::find(c->allocs_state.begin(), c->allocs_state.end(), p);
    if (it != c->allocs_state.end()) c->allocs_state.erase(it);
    (void)hipFree(p);
  };
  drop(A.hkey);
  drop(A.hval);
  A.hkey = hk;
  A.hval = hv;
  Buf& B = c->buf[1 - c->cur];
  e = regrow(c, &B.hkey, cap * neu);
  if (e == hipSuccess) e = regrow(c, &B.hval, cap * neu);
  if (e == hipSuccess) e = regrow(c, &c->clist, cap * neu);
  if (e != hipSuccess) return hip_fail(c, e, "grow history");
  c->hcap = neu;
  c->P.hcap = neu;
  invalidate_graphs(c);
  return SDEM_OK;
}

// Candidate-list slots per particle sized to what the failed rebuild counted.
static int grow_vlist(sdem_ctx* c, int list_need) {
  int neu = std::max(list_need + list_need / 4 + 8, 2 * c->lcap);
  if ((int64_t)neu * c->cap > ((int64_t)1 << 34)) return c->fail(SDEM_E_CAPACITY, "candidate lists too large");
  hipError_t e = regrow(c, &c->vlist, (size_t)c->cap * neu);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(candidate lists)");
  c->lcap = neu;
  c->P.lcap = neu;
  invalidate_graphs(c);
  return SDEM_OK;
}

// Queue one step on the context's stream; ev (SDEM_NUM_PHASES+1 events) may
// be null.  The buffers swap roles afterwards.
static hipError_t enqueue_step(sdem_ctx* c, hipEvent_t* ev) {
  Buf& A = c->buf[c->cur];
  Buf& B = c->buf[1 - c->cur];
  if (ev) (void)hipEventRecord(ev[0], c->stream);
  hipError_t e = launch_ctl(c->P, c->S, c->slots, c->stream);
  if (ev) (void)hipEventRecord(ev[1], c->stream);
  if (e == hipSuccess)
    e = launch_sort(c->P, A, c->S, c->n_upper, c->dead_key, c->keys, c->rank, c->counts, c->start,
                    c->tile_sum, c->perm_tmp, c->perm, c->pk, c->stream);
  if (ev) (void)hipEventRecord(ev[2], c->stream);
  if (e == hipSuccess) e = launch_vlist(c->P, c->S, c->n_upper, c->pk, c->start, c->vcnt, c->vlist, c->stream);
  if (ev) (void)hipEventRecord(ev[3], c->stream);
  if (e == hipSuccess)
    e = launch_force(c->P, A, B, c->S, c->n_upper, c->perm, c->vcnt, c->vlist, c->clist, c->ccnt,
                     c->d_fpk, c->d_fid, c->d_ftype, c->d_fstart, c->slots, c->stream,
                     ev ? ev[4] : nullptr);
  if (ev) (void)hipEventRecord(ev[5], c->stream);
  c->cur = 1 - c->cur;
  return e;
}

// A graph of two consecutive steps starting from buffer c->cur (the kernels
// read their control flags from device memory, so the graph is fixed).
static hipError_t pair_graph(sdem_ctx* c, hipGraphExec_t* out) {
  if (c->gexec[c->cur]) {
    *out = c->gexec[c->cur];
    return hipSuccess;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t e1 = enqueue_step(c, nullptr);
  if (e1 == hipSuccess) e1 = enqueue_step(c, nullptr);
  e = hipStreamEndCapture(c->stream, &g);
  if (e1 != hipSuccess) e = e1;
  if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec[c->cur], g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess) *out = c->gexec[c->cur];
  return e;
}

int sdem_step(sdem_ctx* c, int32_t nsteps) {
  if (!c) return SDEM_E_ARG;
  if (!c->have_state) return c->fail(SDEM_E_STATE, "state must be set before stepping");
  if (nsteps < 0) return c->fail(SDEM_E_ARG, "nsteps < 0");
  if (int rc = prepare(c)) return rc;
  const int NE = SDEM_NUM_PHASES + 1;
  int remaining = nsteps, retries = 0;
  while (remaining > 0) {
    std::vector<hipEvent_t> evs;
    if (c->prof) evs.resize((size_t)remaining * NE);
    for (auto& x : evs)
      if (hipEventCreate(&x) != hipSuccess) return c->fail(SDEM_E_HIP, "hipEventCreate");
    const int c0 = c->cur;
    hipError_t e = hipSuccess;
    if (c->need_pre) {
      // near-wall pass of a state that no step has produced
      e = launch_pre(c->P, c->buf[c->cur], c->S, c->slots, c->n_upper, c->stream);
      c->need_pre = false;
    }
    int it = 0;
    if (!c->prof && c->use_graph) {
      for (; it + 2 <= remaining && e == hipSuccess; it += 2) {
        hipGraphExec_t g = nullptr;
        e = pair_graph(c, &g);
        if (e == hipSuccess) e = hipGraphLaunch(g, c->stream);
      }
    }
    for (; it < remaining && e == hipSuccess; it++)
      e = enqueue_step(c, c->prof ? &evs[(size_t)it * NE] : nullptr);
    if (e == hipSuccess) e = launch_finish(c->S, c->slots, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      for (auto& x : evs) (void)hipEventDestroy(x);
      return hip_fail(c, e, "sdem_step");
    }
    Scalars s;
    e = hipMemcpy(&s, c->S, sizeof(s), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      for (auto& x : evs) (void)hipEventDestroy(x);
      return hip_fail(c, e, "hipMemcpy(scalars)");
    }
    const int ran = remaining - s.halted_steps;  // steps that did work, the last one failed if overflow
    const int ok_steps = s.overflow ? ran - 1 : ran;
    for (int it = 0; it < ok_steps && c->prof; it++)
      for (int q = 0; q < SDEM_NUM_PHASES; q++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, evs[(size_t)it * NE + q], evs[(size_t)it * NE + q + 1]) == hipSuccess) {
          c->prof_ms[q] += ms;
          c->prof_n[q]++;
        }
      }
    for (auto& x : evs) (void)hipEventDestroy(x);
    if (ok_steps > 0) {
      c->stepped = true;
      c->first_step_pending = false;
    }
    if (!s.overflow) break;
    // a capacity ran out in step `ok_steps` of this batch: its input buffer
    // is intact; grow the capacity and replay from there
    if (++retries > 12) return c->fail(SDEM_E_CAPACITY, "capacity still exhausted after regrowing");
    c->cur = (c0 + ok_steps) % 2;
    const int kind = s.overflow, list_need = s.list_need;
    s.iter = s.iter_at_begin;
    s.next_wall_update = s.next_wall_at_begin;
    s.n_alive = s.n_stored;
    s.need_rebuild = 1;
    s.have_parts = 0;
    s.n_deleted = 0;
    s.ticket_scan = 0;
    s.list_need = 0;
    s.overflow = 0;
    s.halt = 0;
    s.halted_steps = 0;
    if (c->first_step_pending) s.use_host_max_a = 1;
    c->need_pre = true;
    e = hipMemcpy(c->S, &s, sizeof(s), hipMemcpyHostToDevice);
    // the slots hold what the failing and the halted steps folded in; the
    // replayed step rebuilds its lists, so only its own maxima are needed
    if (e == hipSuccess) e = hipMemset(c->slots, 0, NSLOT * sizeof(Slot));
    if (e != hipSuccess) return hip_fail(c, e, "reset scalars");
    int rc = kind == 2 ? grow_vlist(c, list_need) : grow_history(c);
    if (rc) return rc;
    remaining -= ok_steps;
  }
  return SDEM_OK;
}

// live particles of buffer A as (id, storage index), ascending id
static int live_index(sdem_ctx* c, std::vector<std::pair<int, int>>& idx) {
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "sync");
  Scalars s;
  e = hipMemcpy(&s, c->S, sizeof(s), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(scalars)");
  int n = c->stepped ? s.n_alive : s.n_stored;
  Buf& A = c->buf[c->cur];
  std::vector<int> ids(n), types(n);
  e = hipMemcpy(ids.data(), A.id, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(types.data(), A.type, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(ids)");
  idx.clear();
  for (int k = 0; k < n; k++)
    if (types[k] >= 1) idx.push_back({ids[k], k});
  std::sort(idx.begin(), idx.end());
  return SDEM_OK;
}

int sdem_get_num_particles(sdem_ctx* c, int32_t* n) {
  if (!c || !n) return SDEM_E_ARG;
  if (!c->have_state) { *n = 0; return SDEM_OK; }
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  *n = (int32_t)idx.size();
  return SDEM_OK;
}

int sdem_get_state(sdem_ctx* c, int32_t capn, int32_t* nout, int32_t* id, int32_t* ptype,
                   int32_t* usr_mark, double* posr, double* lin_vel, double* lin_acc, double* theta,
                   double* rot_vel, double* rot_acc) {
  if (!c || !nout) return SDEM_E_ARG;
  if (!c->have_state) return c->fail(SDEM_E_STATE, "no state");
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  int n = (int)idx.size();
  *nout = n;
  if (capn < n) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
  Buf& A = c->buf[c->cur];
  const size_t cap = c->cap;
  const int ts = c->tsize, rs = c->rsize;
  std::vector<double> lv, la, th, rv, ra;
  std::vector<double4> p4;
  std::vector<int> ty, mk, ids;
  hipError_t e = dl(p4, A.pos, cap);
  if (e == hipSuccess) e = dl(lv, A.lv, cap * 3 * ts);
  if (e == hipSuccess) e = dl(la, A.la, cap * 3 * c->ta);
  if (e == hipSuccess) e = dl(th, A.th, cap * 3);
  if (e == hipSuccess) e = dl(rv, A.rv, cap * 3 * rs);
  if (e == hipSuccess) e = dl(ra, A.ra, cap * 3 * c->tr);
  if (e == hipSuccess) e = dl(ty, A.type, cap);
  if (e == hipSuccess) e = dl(mk, A.mark, cap);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(state)");
  // before the first step the storage order is the caller's order
  for (int q = 0; q < n; q++) {
    const int k = idx[q].second;
    if (id) id[q] = idx[q].first;
    if (ptype) ptype[q] = ty[k];
    if (usr_mark) usr_mark[q] = mk[k];
    if (posr) { posr[4 * q] = p4[k].x; posr[4 * q + 1] = p4[k].y; posr[4 * q + 2] = p4[k].z; posr[4 * q + 3] = p4[k].w; }
    for (int l = 0; l < ts; l++)
      for (int d = 0; d < 3; d++) {
        if (lin_vel) lin_vel[((size_t)q * ts + l) * 3 + d] = lv[((size_t)l * 3 + d) * cap + k];
        if (lin_acc) {
          double v;
          if (ts == 1) v = la[(size_t)d * cap + k];
          else if (l == 0) v = c->stepped ? la[(size_t)d * cap + k] : c->host_lacc1[3 * (size_t)k + d];
          else v = la[((size_t)(l - 1) * 3 + d) * cap + k];
          lin_acc[((size_t)q * ts + l) * 3 + d] = v;
        }
      }
    if (theta) for (int d = 0; d < 3; d++) theta[3 * q + d] = th[(size_t)d * cap + k];
    for (int l = 0; l < rs; l++)
      for (int d = 0; d < 3; d++) {
        if (rot_vel) rot_vel[((size_t)q * rs + l) * 3 + d] = rv[((size_t)l * 3 + d) * cap + k];
        if (rot_acc) {
          double v;
          if (rs == 1) v = ra[(size_t)d * cap + k];
          else if (l == 0) v = c->stepped ? ra[(size_t)d * cap + k] : c->host_racc1[3 * (size_t)k + d];
          else v = ra[((size_t)(l - 1) * 3 + d) * cap + k];
          rot_acc[((size_t)q * rs + l) * 3 + d] = v;
        }
      }
  }
  return SDEM_OK;
}

int sdem_get_forces(sdem_ctx* c, int32_t capn, int32_t* id, double* f, double* t) {
  if (!c) return SDEM_E_ARG;
  if (!c->have_state) return c->fail(SDEM_E_STATE, "no state");
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  int n = (int)idx.size();
  if (capn < n) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
  std::vector<double> fb;
  hipError_t e = dl(fb, c->buf[c->cur].force, (size_t)c->cap * 6);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(forces)");
  const size_t cap = c->cap;
  for (int q = 0; q < n; q++) {
    int k = idx[q].second;
    if (id) id[q] = idx[q].first;
    for (int d = 0; d < 3; d++) {
      double fv = c->stepped ? fb[(size_t)d * cap + k] : 0.0;
      double tv = c->stepped ? fb[(size_t)(3 + d) * cap + k] : 0.0;
      if (f) f[3 * q + d] = fv;
      if (t) t[3 * q + d] = tv;
    }
  }
  return SDEM_OK;
}

struct Rec {
  int owner, partner, kind;
  size_t slot;
};

static int collect(sdem_ctx* c, std::vector<Rec>& recs, std::vector<double4>& hv) {
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  Buf& A = c->buf[c->cur];
  std::vector<int> hc, hk;
  size_t n = (size_t)c->n_upper;
  hipError_t e = dl(hc, A.hcnt, n);
  if (e == hipSuccess) e = dl(hk, A.hkey, (size_t)c->cap * c->hcap);
  if (e == hipSuccess) e = dl(hv, A.hval, (size_t)c->cap * c->hcap);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(history)");
  recs.clear();
  for (auto& pr : idx) {
    int k = pr.second;
    for (int q = 0; q < hc[k]; q++) {
      size_t slot = (size_t)q * c->cap + k;
      recs.push_back({pr.first, hk[slot] & KEY_MASK, (int)((uint32_t)hk[slot] >> KEY_BITS), slot});
    }
  }
  std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
    return a.owner != b.owner ? a.owner < b.owner : a.partner < b.partner;
  });
  return SDEM_OK;
}

int sdem_get_num_contacts(sdem_ctx* c, int64_t* m) {
  if (!c || !m) return SDEM_E_ARG;
  if (!c->have_state) { *m = 0; return SDEM_OK; }
  std::vector<Rec> recs;
  std::vector<double4> hv;
  if (int rc = collect(c, recs, hv)) return rc;
  *m = (int64_t)recs.size();
  return SDEM_OK;
}

int sdem_get_contacts(sdem_ctx* c, int64_t capm, int64_t* m, int32_t* kind, int32_t* owner,
                      int32_t* partner, double* td, double* vr) {
  if (!c || !m) return SDEM_E_ARG;
  if (!c->have_state) { *m = 0; return SDEM_OK; }
  std::vector<Rec> recs;
  std::vector<double4> hv;
  if (int rc = collect(c, recs, hv)) return rc;
  *m = (int64_t)recs.size();
  if (capm < *m) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
  for (size_t q = 0; q < recs.size(); q++) {
    const Rec& r = recs[q];
    if (kind) kind[q] = r.kind;
    if (owner) owner[q] = r.owner;
    if (partner) partner[q] = r.partner;
    const double4 v = hv[r.slot];
    if (td) { td[3 * q] = v.x; td[3 * q + 1] = v.y; td[3 * q + 2] = v.z; }
    if (vr) vr[q] = v.w;
  }
  return SDEM_OK;
}

int sdem_get_counts(sdem_ctx* c, int32_t num[4]) {
  if (!c || !num) return SDEM_E_ARG;
  if (!c->have_state) { for (int k = 0; k < 4; k++) num[k] = 0; return SDEM_OK; }
  hipError_t e = hipStreamSynchronize(c->stream);
  Scalars s;
  if (e == hipSuccess) e = hipMemcpy(&s, c->S, sizeof(s), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(scalars)");
  for (int k = 0; k < 4; k++) num[k] = c->stepped ? s.num_cntcts[k] : 0;
  return SDEM_OK;
}

int sdem_get_iter(sdem_ctx* c, int32_t* iter, int32_t* next_update) {
  if (!c) return SDEM_E_ARG;
  if (!c->have_state) {
    if (iter) *iter = c->o.ifirst - 1;
    if (next_update) *next_update = 0;
    return SDEM_OK;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  Scalars s;
  if (e == hipSuccess) e = hipMemcpy(&s, c->S, sizeof(s), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(scalars)");
  if (iter) *iter = s.iter;
  if (next_update) *next_update = s.next_wall_update;
  return SDEM_OK;
}

int sdem_set_profiling(sdem_ctx* c, int32_t enable) {
  if (!c) return SDEM_E_ARG;
  c->prof = enable != 0;
  for (int q = 0; q < SDEM_NUM_PHASES; q++) { c->prof_ms[q] = 0.0; c->prof_n[q] = 0; }
  return SDEM_OK;
}

int sdem_get_profile(sdem_ctx* c, double ms[SDEM_NUM_PHASES], int64_t launches[SDEM_NUM_PHASES]) {
  if (!c) return SDEM_E_ARG;
  for (int q = 0; q < SDEM_NUM_PHASES; q++) {
    if (ms) ms[q] = c->prof_ms[q];
    if (launches) launches[q] = c->prof_n[q];
  }
  return SDEM_OK;
}

int sdem_set_tuning(sdem_ctx* c, const double cell_factor[3], double skin_factor, int32_t keep_forces) {
  if (!c) return SDEM_E_ARG;
  if (skin_factor >= 0.0) {
    if (skin_factor > 0.45) return c->fail(SDEM_E_ARG, "skin factor must be below 0.45");
    c->skin_factor = skin_factor;
    c->prepared = false;
  }
  if (cell_factor) {
    for (int d = 0; d < 3; d++)
      if (cell_factor[d] < 0.0 || cell_factor[d] > 64.0) return c->fail(SDEM_E_ARG, "cell factor out of range");
    for (int d = 0; d < 3; d++)
      if (cell_factor[d] > 0.0) c->cell_factor[d] = cell_factor[d];
  }
  if (keep_forces >= 0) {
    c->keep_forces = keep_forces ? 1 : 0;
    c->P.keep_forces = c->keep_forces;
    invalidate_graphs(c);
  }
  return SDEM_OK;
}

int sdem_set_capacity(sdem_ctx* c, int32_t hist_slots, int32_t list_slots) {
  if (!c) return SDEM_E_ARG;
  if (!c->allocs_state.empty()) return c->fail(SDEM_E_STATE, "capacities must be set before the state");
  if (hist_slots < 0 || hist_slots > 4096 || list_slots < 0 || list_slots > 65536)
    return c->fail(SDEM_E_ARG, "capacity out of range");
  c->hist_slots_hint = hist_slots;
  c->list_slots_hint = list_slots;
  return SDEM_OK;
}

int sdem_get_device_bytes(sdem_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return SDEM_E_ARG;
  *bytes = c->dev_bytes;
  return SDEM_OK;
}

const char* sdem_last_error(const sdem_ctx* c) { return c ? c->err.c_str() : "null context"; }

void sdem_destroy(sdem_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  invalidate_graphs(c);
  free_list(c, c->allocs_state);
  free_list(c, c->allocs_grid);
  free_list(c, c->allocs_tab);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}
