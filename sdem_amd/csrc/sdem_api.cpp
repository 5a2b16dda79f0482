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
#include <unordered_map>
#include <vector>

#include "../../include/sdem.h"
#include "sdem_host.h"
#include "sdem_internal.h"

namespace sdem {
hipError_t launch_pre(const Dev& D, const Buf& A, hipStream_t st);
hipError_t launch_ctl(const Dev& D, const Buf& A, hipStream_t st);
hipError_t launch_finish(const Dev& D, const Buf& A, hipStream_t st);
hipError_t launch_sort(const Dev& D, const Buf& A, hipStream_t st);
hipError_t launch_vlist(const Dev& D, hipStream_t st);
hipError_t launch_force(const Dev& D, const Buf& A, const Buf& B, hipStream_t st, hipEvent_t mid, bool fast);
hipError_t launch_phase(const Dev& D, const Buf& A, const Buf& B, hipStream_t st, int phase);
int64_t scan_tiles(int64_t n);
int scan_tile_keys();
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


// ---------------------------------------------------------------------------
// One box: an independent system (domain, materials, walls, particles).  A
// plain context has one; an ensemble has many, stepped together.
// ---------------------------------------------------------------------------
struct Box {
  sdem_options o{};
  int ntype = 0, nwtype = 0;
  std::vector<PureProp> pprop, wprop;
  std::vector<Bnry> pp_tab, pw_tab;
  bool have_pprop = false, have_wprop = false;
  double lin_kn = 0.0, lin_kt_ratio = 0.0;  // constant-stiffness linear model
  std::vector<Wall> walls;
  int num_pwall_total = 0;
  std::vector<int> fix_id, fix_type;
  std::vector<double> fix_pos;
  // the state as set (host layouts of sdem_set_state), and its contacts
  bool have_state = false;
  int n = 0;
  std::vector<int> id, ptype, mark;
  std::vector<double> posr, lin_vel, lin_acc, theta, rot_vel, rot_acc;
  std::vector<int> ckind, cowner, cpartner;
  std::vector<double> ctan, cvri;
  // placement and search layout
  int off = 0, bcap = 0;
  std::vector<int> type_level;
  std::vector<double> level_rmax, cells;
  double host_max_a = 0.0;
};

struct sdem_ctx {
  std::vector<Box> boxes;
  int sel = 0;  // box that the setters and getters address
  std::string err;
  int tsize = 1, rsize = 1, ta = 1, tr = 1;
  hipStream_t stream = nullptr;

  // search tuning (results do not depend on it)
  double cell_factor[3] = {1.0, 1.0, 1.0};
  double skin_factor = 0.1;  // candidate-list skin as a fraction of a box's d_max
  int keep_forces = 1;
  int bs_contact = 256, bs_force = 128;
  int hist_slots_hint = 0, list_slots_hint = 0;

  // device
  bool dev_ready = false;     // buffers allocated and the state uploaded
  bool params_ready = false;  // box parameters, tables and grids uploaded
  bool stepped = false, first_step_pending = false;
  bool need_pre = false;      // the state in buffer A has no near-wall pass yet
  int64_t nslots = 0;
  int hcap = 0, lcap = 0;
  Buf buf[2]{};
  int cur = 0;
  Dev D{};
  Scalars* S = nullptr;
  BoxState* BS = nullptr;
  BoxSlot* bslots = nullptr;
  Params* d_params = nullptr;
  int* d_blk_box = nullptr;
  int *keys = nullptr, *rank = nullptr, *counts = nullptr, *start = nullptr, *tile_sum = nullptr;
  int *perm_tmp = nullptr, *perm = nullptr;
  double4* pk = nullptr;
  int* pid = nullptr;
  int *vlist = nullptr, *vkey = nullptr, *cand = nullptr;
  int4* act = nullptr;
  uint2* clist = nullptr;
  int* ccnt = nullptr;
  double* ft = nullptr;  // phase entry points: F, T, then a, alpha (6 x cap)
  int64_t nkeys = 0;
  bool use_graph = true;
  // graphs of one and of two steps, without / with the speculative fast
  // path (single box: no sort and list kernels), by starting buffer
  hipGraphExec_t gexec[2][2][4][2] = {};  // [force occupancy 2/3][fast][1, 2, 8 or 32 steps][buffer]
  int occ3 = 0;                            // the force kernel at three waves per SIMD (Dev::occ3)
  int force_build = 0;                     // 0: chosen per batch; 2 or 3: fixed (sdem_set_force_build)
  bool spec_ready = false;  // the next step may try the fast path
  int spec_chunk = 8;       // steps queued per fast batch (adapted)
  int spec_full = 1;        // steps queued per full batch between fast ones: grows
                            // while fast batches fail at their first step

  bool prof = false;
  double prof_ms[SDEM_NUM_PHASES] = {};
  int64_t prof_n[SDEM_NUM_PHASES] = {};
  // the reference's step timers (Prtcl_DEMSystem.f90:29-37), kept while
  // profiling is on
  double timer_ms[SDEM_NUM_TIMERS] = {};
  int64_t timer_n[SDEM_NUM_TIMERS] = {};
  // a step run phase by phase: the last phase entry point that ran
  // (PH_NONE: no step open)
  int phase = 0;
  std::vector<void*> allocs_state, allocs_grid, allocs_tab;
  int64_t dev_bytes = 0;

  Box& box() { return boxes[sel]; }
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
  for (auto& o : c->gexec)
    for (auto& a : o)
      for (auto& b : a)
        for (auto& g : b)
          if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
          }
  c->spec_ready = false;
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

double max_type_radius(const Box& b) {
  double m = b.pprop[0].radius;
  for (int i = 1; i < b.ntype; i++) m = fmx(m, b.pprop[i].radius);
  return m;
}
double min_type_radius(const Box& b) {
  double m = b.pprop[0].radius;
  for (int i = 1; i < b.ntype; i++) m = fmn(m, b.pprop[i].radius);
  return m;
}

// Level of each type: one level for CS_Method=1, the reference's hierarchy
// (Init_Munjiza_Hrchl, Prtcl_Hrchl_Munjiza.f90:84-237) for CS_Method=2.
// Returns per-level cell lengths.
std::vector<double> plan_levels(Box& b, const std::vector<int>& count_per_type) {
  std::vector<double> cells;
  b.type_level.assign(b.ntype, 0);
  const double maxd = 2.0 * max_type_radius(b), mind = 2.0 * min_type_radius(b);
  double cell_len = b.o.prtcl_cs_ratio * maxd;
  if (b.o.cs_method != SDEM_CSM_NBS_MUNJIZA_HRCHL) {
    cells.push_back(cell_len);
    return cells;
  }
  int dmax_min = (int)(maxd / mind);
  int numlevels;
  if (b.o.cs_numlvls <= 0) {
    if (dmax_min <= 1) numlevels = 1;
    else if (dmax_min <= 3) numlevels = 2;
    else if (dmax_min <= 7) numlevels = 3;
    else if (dmax_min <= 15) numlevels = 4;
    else if (dmax_min <= 31) numlevels = 5;
    else numlevels = 6;
  } else {
    numlevels = std::min(b.o.cs_numlvls, MAX_LEVELS);
  }
  double maxd_lvl = maxd;
  for (int idh = 1; idh <= numlevels; idh++) {
    double mind_lvl = maxd_lvl / 2.0;
    if (idh == numlevels) mind_lvl = 0.0;
    int num = 0;
    for (int i = 0; i < b.ntype; i++) {
      double diam = 2.0 * b.pprop[i].radius;
      if (diam > mind_lvl && (idh == 1 || diam <= maxd_lvl)) num += count_per_type[i];
    }
    if (num > 0) {
      int lvl = (int)cells.size();
      for (int i = 0; i < b.ntype; i++) {
        double diam = 2.0 * b.pprop[i].radius;
        if (diam > mind_lvl && (idh == 1 || diam <= maxd_lvl)) b.type_level[i] = lvl;
      }
      cells.push_back(cell_len);
    }
    maxd_lvl = mind_lvl;
    cell_len = cell_len / 2.0;
  }
  if (cells.empty()) cells.push_back(b.o.prtcl_cs_ratio * maxd);
  return cells;
}

int make_grid(sdem_ctx* c, const Box& b, double cell_len0, double rmax, int64_t key_base, Grid* G) {
  G->key_base = key_base;
  G->rmax = rmax;
  G->ncell = 1;
  for (int d = 0; d < 3; d++) {
    double L = b.o.domain_max[d] - b.o.domain_min[d];
    double cell_len = cell_len0 * c->cell_factor[d];
    G->origin[d] = b.o.domain_min[d];
    // the cell count is checked before it is converted to an int (a tiny
    // cell factor would overflow it)
    if (!(L / cell_len <= (double)(1 << 30))) return c->fail(SDEM_E_ARG, "search grid too large");
    if (b.o.is_periodic[d]) {
      int n = (int)std::floor(L / cell_len);
      // the search range of one particle must not wrap onto itself
      double reach = 2.0 * max_type_radius(b) + rmax;
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
  G->nkey = (int64_t)BRICK * BRICK * BRICK;
  for (int d = 0; d < 3; d++) {
    G->nb[d] = (G->n[d] + BRICK - 1) / BRICK;
    G->nkey *= G->nb[d];
  }
  if (G->nkey > (int64_t)1 << 30) return c->fail(SDEM_E_ARG, "search grid too large");
  return SDEM_OK;
}

// The most history records one particle of the box holds when its contact
// records are installed: its own, and a copy of each PP record of a lower id
// partner (KIND_COPY).
static int installed_records(const Box& b) {
  std::unordered_map<int, int> n;
  for (size_t k = 0; k < b.ckind.size(); k++) {
    n[b.cowner[k]]++;
    if (b.ckind[k] == SDEM_KIND_PP) n[b.cpartner[k]]++;
  }
  int m = 0;
  for (auto& kv : n) m = std::max(m, kv.second);
  return m;
}

// history slots per particle: a kissing-number bound for the largest type
// touched by the smallest, with room for soft overlaps and walls, at most 48
// to start with (an overflow doubles it and replays the step)
int history_capacity(const Box& b) {
  double rmax = max_type_radius(b), rmin = min_type_radius(b);
  double s = rmin / (rmax + rmin);
  double cosv = std::sqrt(std::max(0.0, 1.0 - s * s));
  double bound = 2.0 / (1.0 - cosv);
  int h = (int)std::ceil(1.25 * bound) + 4 + (int)b.walls.size();
  h = std::max(h, 2 * b.o.cntct_list_size);
  h = std::max(h, 16);
  return std::min(h, 48);
}

bool check_options(const sdem_options* o) {
  if (o->pi_method < 1 || o->pi_method > 3 || o->pri_method < 1 || o->pri_method > 3 ||
      o->cf_type < 1 || o->cf_type > 3 || o->ct_model < 1 || o->ct_model > 2 ||
      o->cs_method < 1 || o->cs_method > 2 || o->num_prtcl < 1 || !(o->dt > 0.0) ||
      !(o->prtcl_cs_ratio >= 1.0) || o->wall_max_update_iter < 1)
    return false;
  for (int d = 0; d < 3; d++)
    if (!(o->domain_max[d] > o->domain_min[d])) return false;
  return true;
}

int setup_device(sdem_ctx* c) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return SDEM_E_NODEV;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SDEM_E_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return SDEM_E_NODEV;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SDEM_E_NODEV;
  if (hipStreamCreate(&c->stream) != hipSuccess) return SDEM_E_HIP;
  return SDEM_OK;
}

int round_up(int n, int a) { return (n + a - 1) / a * a; }

template <class T>
hipError_t dl(std::vector<T>& v, const T* src, size_t n) {
  v.resize(n);
  if (!n) return hipSuccess;
  return hipMemcpy(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost);
}

// rows x len elements of a row-major device array whose rows are `cap` long,
// from column `off` on (one box's slots of a per-particle array):
// v[r * len + i] = src[r * cap + off + i]
template <class T>
hipError_t dl_box(std::vector<T>& v, const T* src, size_t cap, size_t rows, size_t off, size_t len) {
  v.resize(rows * len);
  if (!rows || !len) return hipSuccess;
  return hipMemcpy2D(v.data(), len * sizeof(T), src + off, cap * sizeof(T), len * sizeof(T), rows,
                     hipMemcpyDeviceToHost);
}

template <class T>
hipError_t put_field(T* dst, const T& v) {
  return hipMemcpy(dst, &v, sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

// ===========================================================================
// Device state: allocation and upload of every box's staged state
// ===========================================================================
static int alloc_buf(sdem_ctx* c, Buf& b) {
  size_t cap = (size_t)c->nslots, hc = (size_t)c->hcap;
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

static int upload_state(sdem_ctx* c) {
  invalidate_graphs(c);
  free_list(c, c->allocs_state);
  // placement: every box owns a BOX_ALIGN-multiple of slots
  int64_t off = 0;
  int hc = 0, hc_inst = 0;
  for (auto& b : c->boxes) {
    if (!b.have_state) return c->fail(SDEM_E_STATE, "every box needs a state before stepping");
    b.off = (int)off;
    b.bcap = round_up(std::max(b.n, 1), BOX_ALIGN);
    off += b.bcap;
    hc = std::max(hc, history_capacity(b));
    hc_inst = std::max(hc_inst, installed_records(b));
  }
  if (off > INT32_MAX / 2) return c->fail(SDEM_E_CAPACITY, "too many particles for one context");
  c->nslots = off;
  const size_t cap = (size_t)off;
  c->hcap = std::max({KEY_HEAD, c->hist_slots_hint > 0 ? c->hist_slots_hint : hc, hc_inst});
  c->lcap = c->list_slots_hint > 0 ? c->list_slots_hint : std::max(48, 2 * c->hcap);
  const int nbox = (int)c->boxes.size();
  for (int k = 0; k < 2; k++) {
    if (int r = alloc_buf(c, c->buf[k])) return r;
    c->buf[k].idx = k;
  }
  hipError_t e = dalloc(c, c->allocs_state, &c->S, 1);
  auto M = [&](auto** p, size_t n) { if (e == hipSuccess) e = dalloc(c, c->allocs_state, p, n); };
  M(&c->BS, nbox); M(&c->bslots, (size_t)nbox * NSET * NSLOT);
  M(&c->keys, cap); M(&c->rank, cap); M(&c->perm_tmp, cap); M(&c->perm, cap); M(&c->pk, cap); M(&c->pid, cap);
  M(&c->vlist, cap * c->lcap); M(&c->vkey, cap * c->lcap);
  M(&c->act, cap); M(&c->cand, cap * c->lcap);
  M(&c->ccnt, cap); M(&c->clist, cap * c->hcap); M(&c->ft, cap * 6);
  M(&c->d_blk_box, cap / BOX_ALIGN);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");

  const int ts = c->tsize, rs = c->rsize;
  std::vector<double4> p4(cap, make_double4(0, 0, 0, 0));
  std::vector<double> lv(cap * 3 * ts, 0.0), la(cap * 3 * c->ta, 0.0), th(cap * 3, 0.0);
  std::vector<double> rv(cap * 3 * rs, 0.0), ra(cap * 3 * c->tr, 0.0);
  std::vector<int> id(cap, 0), ty(cap, 0), mk(cap, 0), hcn(cap, 0), hk(cap * c->hcap, 0), blk(cap / BOX_ALIGN);
  std::vector<double4> hv(cap * c->hcap, make_double4(0, 0, 0, 0));
  std::vector<uint64_t> wm(cap, 0);
  std::vector<BoxState> bs(nbox);
  for (int bi = 0; bi < nbox; bi++) {
    Box& b = c->boxes[bi];
    for (int g = b.off / BOX_ALIGN; g < (b.off + b.bcap) / BOX_ALIGN; g++) blk[g] = bi;
    double host_max_a = 0.0;
    for (int i = 0; i < b.n; i++) {
      const size_t s = (size_t)b.off + i;
      const double* pr = &b.posr[4 * (size_t)i];
      p4[s] = make_double4(pr[0], pr[1], pr[2], pr[3]);
      id[s] = b.id[i]; ty[s] = b.ptype[i]; mk[s] = b.mark[i];
      for (int l = 0; l < ts; l++)
        for (int d = 0; d < 3; d++) lv[((size_t)l * 3 + d) * cap + s] = b.lin_vel[((size_t)i * ts + l) * 3 + d];
      // stored acceleration levels: FE keeps level 1, AB2/AB3 keep levels 2..tsize
      for (int l = 0; l < c->ta; l++) {
        int src = ts == 1 ? 0 : l + 1;
        for (int d = 0; d < 3; d++) la[((size_t)l * 3 + d) * cap + s] = b.lin_acc[((size_t)i * ts + src) * 3 + d];
      }
      for (int d = 0; d < 3; d++) th[(size_t)d * cap + s] = b.theta[3 * (size_t)i + d];
      for (int l = 0; l < rs; l++)
        for (int d = 0; d < 3; d++) rv[((size_t)l * 3 + d) * cap + s] = b.rot_vel[((size_t)i * rs + l) * 3 + d];
      for (int l = 0; l < c->tr; l++) {
        int src = rs == 1 ? 0 : l + 1;
        for (int d = 0; d < 3; d++) ra[((size_t)l * 3 + d) * cap + s] = b.rot_acc[((size_t)i * rs + src) * 3 + d];
      }
      const double* a = &b.lin_acc[(size_t)i * ts * 3];
      host_max_a = fmx(host_max_a, std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]));
    }
    b.host_max_a = host_max_a;
    // history records (Add_RestartCntctlink-style install)
    std::vector<std::pair<int, int>> idx;
    idx.reserve(b.n);
    for (int i = 0; i < b.n; i++) idx.push_back({b.id[i], i});
    std::sort(idx.begin(), idx.end());
    auto put = [&](size_t s, int partner, int kind, size_t k) -> bool {
      if (hcn[s] >= c->hcap) return false;
      const int r = hcn[s]++;
      hk[key_slot(r, s, cap)] = partner | (kind << KEY_BITS);
      hv[(size_t)r * cap + s] = make_double4(b.ctan[3 * k], b.ctan[3 * k + 1], b.ctan[3 * k + 2], b.cvri[k]);
      return true;
    };
    for (size_t k = 0; k < b.ckind.size(); k++) {
      auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(b.cowner[k], -1));
      if (it == idx.end() || it->first != b.cowner[k]) return c->fail(SDEM_E_ARG, "owner id not present");
      if (!put((size_t)b.off + it->second, b.cpartner[k], b.ckind[k], k))
        return c->fail(SDEM_E_CAPACITY, "too many history records for one particle");
    }
    // the higher id's copy of each PP record (KIND_COPY), after the
    // records proper
    for (size_t k = 0; k < b.ckind.size(); k++) {
      if (b.ckind[k] != SDEM_KIND_PP) continue;
      auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(b.cpartner[k], -1));
      if (it == idx.end() || it->first != b.cpartner[k]) continue;
      if (!put((size_t)b.off + it->second, b.cowner[k], KIND_COPY, k))
        return c->fail(SDEM_E_CAPACITY, "too many history records for one particle");
    }
    BoxState& B = bs[bi];
    std::memset(&B, 0, sizeof(B));
    B.n_stored = b.n;
    B.n_alive = b.n;
    B.next_wall_update = 0;
    B.use_host_max_a = 1;
    B.host_max_a = host_max_a;
  }
  Buf& A = c->buf[0];
  auto H = [&](void* dst, const void* src, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  H(A.pos, p4.data(), cap * sizeof(double4));
  H(A.lv, lv.data(), lv.size() * 8); H(A.la, la.data(), la.size() * 8); H(A.th, th.data(), th.size() * 8);
  H(A.rv, rv.data(), rv.size() * 8); H(A.ra, ra.data(), ra.size() * 8);
  H(A.id, id.data(), cap * 4); H(A.type, ty.data(), cap * 4); H(A.mark, mk.data(), cap * 4);
  H(A.wmask, wm.data(), cap * 8); H(A.hcnt, hcn.data(), cap * 4);
  H(A.hkey, hk.data(), hk.size() * 4); H(A.hval, hv.data(), hv.size() * sizeof(double4));
  H(c->d_blk_box, blk.data(), blk.size() * 4);
  H(c->BS, bs.data(), bs.size() * sizeof(BoxState));
  Scalars s;
  std::memset(&s, 0, sizeof(s));
  s.iter = c->boxes[0].o.ifirst - 1;
  s.need_rebuild = 1;
  s.rv2_stored[0] = s.rv2_stored[1] = 1;
  H(c->S, &s, sizeof(s));
  if (e == hipSuccess) e = hipMemset(c->bslots, 0, sizeof(BoxSlot) * NSET * NSLOT * nbox);
  if (e != hipSuccess) return hip_fail(c, e, "upload state");
  c->cur = 0;
  c->dev_ready = true;
  c->params_ready = false;
  c->stepped = false;
  c->first_step_pending = true;
  c->need_pre = true;
  return SDEM_OK;
}

// Box parameters, pair tables, walls and search grids for the current tuning.
static int upload_params(sdem_ctx* c) {
  invalidate_graphs(c);
  free_list(c, c->allocs_grid);
  free_list(c, c->allocs_tab);
  const int nbox = (int)c->boxes.size();
  std::vector<Params> P(nbox);
  std::vector<Bnry> pp, pw;
  std::vector<double> mass, inertia;
  std::vector<Wall> walls;
  std::vector<int> tlev;
  std::vector<size_t> o_pp(nbox), o_pw(nbox), o_m(nbox), o_w(nbox), o_t(nbox);
  int64_t key = 0;
  Grid fgrid{};
  int nfix = 0;
  for (int bi = 0; bi < nbox; bi++) {
    Box& b = c->boxes[bi];
    Params& p = P[bi];
    std::memset(&p, 0, sizeof(p));
    p.dt = b.o.dt;
    for (int d = 0; d < 3; d++) {
      p.g[d] = b.o.gravity[d];
      p.dmin[d] = b.o.domain_min[d];
      p.dmax[d] = b.o.domain_max[d];
      p.simlen[d] = b.o.domain_max[d] - b.o.domain_min[d];
      p.periodic[d] = b.o.is_periodic[d] ? 1 : 0;
    }
    const double maxd = 2.0 * max_type_radius(b);
    for (int d = 0; d < 3; d++)
      if (1.02 * b.o.prtcl_cs_ratio * maxd > p.simlen[d]) return c->fail(SDEM_E_ARG, "so big Diameter");
    p.cf_type = b.o.cf_type;
    p.ct_model = b.o.ct_model;
    p.pi_method = b.o.pi_method;
    p.pri_method = b.o.pri_method;
    p.tsize = c->tsize; p.rsize = c->rsize; p.ta = c->ta; p.tr = c->tr;
    p.cap = (int)c->nslots;
    p.hcap = c->hcap;
    p.lcap = c->lcap;
    p.ntype = b.ntype;
    p.nwtype = b.nwtype;
    p.nwall = (int)b.walls.size();
    p.search_fixed = b.o.cs_method == SDEM_CSM_NBS_MUNJIZA ? 1 : 0;
    p.keep_forces = c->keep_forces;
    const double rmax = max_type_radius(b);
    p.dxmax1 = rmax * b.o.wall_neighbor_ratio;
    p.dxmax2 = rmax * (b.o.wall_neighbor_ratio + 1.0);
    double mwv = 0.0;
    for (auto& w : b.walls) mwv = fmx(mwv, vnorm(v3(w.trans_vel)));
    p.max_wall_vel = mwv;
    p.wall_max_update_iter = b.o.wall_max_update_iter;
    p.skin = c->skin_factor * maxd;
    p.rebuild_disp = 0.5 * p.skin * (1.0 - 1.0e-6);
    for (auto& w : b.walls)
      if (w.wall_type > b.nwtype) return c->fail(SDEM_E_ARG, "wall_type exceeds numWall_type");
    // search levels: radii never change, so the per-level largest radius is
    // fixed by the state
    std::vector<int> cnt(b.ntype, 0);
    for (int i = 0; i < b.n; i++) cnt[b.ptype[i] - 1]++;
    b.cells = plan_levels(b, cnt);
    b.level_rmax.assign(b.cells.size(), 0.0);
    for (int i = 0; i < b.n; i++) {
      int L = b.type_level[b.ptype[i] - 1];
      b.level_rmax[L] = fmx(b.level_rmax[L], b.posr[4 * (size_t)i + 3]);
    }
    p.nlvl = (int)b.cells.size();
    p.key_first = (int)key;
    for (int L = 0; L < p.nlvl; L++) {
      if (int rc = make_grid(c, b, b.cells[L], b.level_rmax[L], key, &p.lvl[L])) return rc;
      key += p.lvl[L].nkey;
    }
    p.dead_key = (int)key;
    key += 1;
    if (key > INT32_MAX / 2) return c->fail(SDEM_E_ARG, "search grids too large");
    p.box = bi;
    p.off = b.off;
    p.bcap = b.bcap;
    // tables, concatenated over the boxes
    o_pp[bi] = pp.size(); pp.insert(pp.end(), b.pp_tab.begin(), b.pp_tab.end());
    o_pw[bi] = pw.size(); pw.insert(pw.end(), b.pw_tab.begin(), b.pw_tab.end());
    o_m[bi] = mass.size();
    for (int i = 0; i < b.ntype; i++) { mass.push_back(b.pprop[i].mass); inertia.push_back(b.pprop[i].inertia); }
    o_w[bi] = walls.size(); walls.insert(walls.end(), b.walls.begin(), b.walls.end());
    o_t[bi] = tlev.size(); tlev.insert(tlev.end(), b.type_level.begin(), b.type_level.end());
    // fixed particles: single-box contexts only (checked at set time)
    if (!b.fix_id.empty()) {
      nfix = (int)b.fix_id.size();
      double fr = 0.0;
      for (int i = 0; i < nfix; i++) fr = fmx(fr, b.fix_pos[4 * i + 3]);
      if (int rc = make_grid(c, b, b.o.prtcl_cs_ratio * maxd, fr, 0, &fgrid)) return rc;
    }
    p.nfix = b.fix_id.empty() ? 0 : nfix;
  }
  c->nkeys = key;
  hipError_t e = hipSuccess;
  auto U = [&](auto** dst, const auto& vec) {
    using T = typename std::remove_reference<decltype(vec)>::type::value_type;
    if (e == hipSuccess) e = dalloc(c, c->allocs_tab, dst, vec.size());
    if (e == hipSuccess && !vec.empty()) e = hipMemcpy(*dst, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice);
  };
  Bnry *d_pp = nullptr, *d_pw = nullptr;
  double *d_mass = nullptr, *d_inertia = nullptr;
  Wall* d_walls = nullptr;
  int* d_tlev = nullptr;
  U(&d_pp, pp); U(&d_pw, pw); U(&d_mass, mass); U(&d_inertia, inertia); U(&d_walls, walls); U(&d_tlev, tlev);
  // fixed particles sorted by the cells of their grid
  double4* d_fpk = nullptr;
  int *d_fid = nullptr, *d_ftype = nullptr, *d_fstart = nullptr;
  {
    const Box& b = c->boxes[0];
    const Grid& G = fgrid;
    std::vector<int64_t> fkey(nfix);
    auto cl = [&](int d, double x) {
      int cc = (int)std::floor((x - G.origin[d]) * G.inv_w[d]);
      return cc < 0 ? 0 : (cc >= G.n[d] ? G.n[d] - 1 : cc);
    };
    for (int i = 0; i < nfix; i++) {
      const double* p = &b.fix_pos[4 * i];
      fkey[i] = ((int64_t)cl(2, p[2]) * G.n[1] + cl(1, p[1])) * G.n[0] + cl(0, p[0]);
    }
    std::vector<int> ord(nfix);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return fkey[x] < fkey[y]; });
    std::vector<double4> fpk(nfix);
    std::vector<int> fid(nfix), ftype(nfix), fstart(nfix > 0 ? G.ncell + 1 : 1, 0);
    for (int k = 0; k < nfix; k++) {
      int i = ord[k];
      const double* p = &b.fix_pos[4 * i];
      fpk[k] = make_double4(p[0], p[1], p[2], p[3]);
      fid[k] = b.fix_id[i];
      ftype[k] = b.fix_type[i];
    }
    if (nfix > 0) {
      for (int i = 0; i < nfix; i++) fstart[fkey[i] + 1]++;
      for (int64_t q = 0; q < G.ncell; q++) fstart[q + 1] += fstart[q];
    }
    U(&d_fpk, fpk); U(&d_fid, fid); U(&d_ftype, ftype); U(&d_fstart, fstart);
  }
  for (int bi = 0; bi < nbox; bi++) {
    Params& p = P[bi];
    p.pp = d_pp + o_pp[bi];
    p.pw = d_pw + o_pw[bi];
    p.mass = d_mass + o_m[bi];
    p.inertia = d_inertia + o_m[bi];
    p.walls = d_walls + o_w[bi];
    p.type_level = d_tlev + o_t[bi];
    p.fgrid = fgrid;
  }
  U(&c->d_params, P);
  if (e == hipSuccess) e = dalloc(c, c->allocs_grid, &c->counts, (size_t)key + 1);
  if (e == hipSuccess) e = hipMemset(c->counts, 0, ((size_t)key + 1) * sizeof(int));
  if (e == hipSuccess) e = dalloc(c, c->allocs_grid, &c->start, (size_t)key + 1);
  if (e == hipSuccess) e = dalloc(c, c->allocs_grid, &c->tile_sum, (size_t)scan_tiles(key + 1) + 1);
  // boxes covered by each scan tile (keys are box-contiguous; the end
  // sentinel belongs to the last box)
  int2* d_tile_box = nullptr;
  {
    const int64_t nt = scan_tiles(key + 1), tk = scan_tile_keys();
    std::vector<int2> tb(nt);
    int bx = 0;
    auto box_of = [&](int64_t k) {
      while (bx + 1 < nbox && k > P[bx].dead_key) bx++;
      return bx;
    };
    for (int64_t t = 0; t < nt; t++) {
      tb[t].x = box_of(t * tk);
      tb[t].y = box_of(std::min<int64_t>((t + 1) * tk - 1, key));
    }
    if (e == hipSuccess) e = dalloc(c, c->allocs_grid, &d_tile_box, tb.size());
    if (e == hipSuccess) e = hipMemcpy(d_tile_box, tb.data(), tb.size() * sizeof(int2), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = put_field(&c->S->need_rebuild, 1);  // grids or skin may have changed
  if (e != hipSuccess) return hip_fail(c, e, "upload parameters");
  Dev& D = c->D;
  D.PB = c->d_params; D.blk_box = c->d_blk_box; D.nbox = nbox; D.nslots = c->nslots; D.nkeys = key;
  D.S = c->S; D.BS = c->BS; D.bslots = c->bslots;
  D.keys = c->keys; D.rank = c->rank; D.counts = c->counts; D.start = c->start; D.tile_sum = c->tile_sum;
  D.tile_box = d_tile_box;
  D.perm_tmp = c->perm_tmp; D.perm = c->perm; D.pk = c->pk; D.pid = c->pid;
  D.vlist = c->vlist; D.vkey = c->vkey; D.clist = c->clist; D.ccnt = c->ccnt; D.ft = c->ft;
  D.act = c->act; D.cand = c->cand;
  D.fpk = d_fpk; D.fid = d_fid; D.ftype = d_ftype; D.fstart = d_fstart;
  D.bs_contact = c->bs_contact; D.bs_force = c->bs_force;
  D.pim = c->boxes[0].o.pi_method; D.prim = c->boxes[0].o.pri_method;
  D.occ3 = c->occ3;
  c->params_ready = true;
  return SDEM_OK;
}

static int prepare(sdem_ctx* c) {
  if (!c->dev_ready)
    if (int rc = upload_state(c)) return rc;
  if (!c->params_ready)
    if (int rc = upload_params(c)) return rc;
  return SDEM_OK;
}

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

int sdem_create_ensemble(int32_t nbox, const sdem_options* opts, sdem_ctx** out) {
  if (!out) return SDEM_E_ARG;
  *out = nullptr;
  if (nbox < 1 || !opts) return SDEM_E_ARG;
  for (int b = 0; b < nbox; b++) {
    if (!check_options(&opts[b])) return SDEM_E_ARG;
    if (opts[b].pi_method != opts[0].pi_method || opts[b].pri_method != opts[0].pri_method ||
        opts[b].ifirst != opts[0].ifirst)
      return SDEM_E_ARG;  // one state layout and one iterNumber for all boxes
  }
  sdem_ctx* c = new sdem_ctx();
  if (int rc = setup_device(c)) {
    delete c;
    return rc;
  }
  c->boxes.resize(nbox);
  for (int b = 0; b < nbox; b++) c->boxes[b].o = opts[b];
  c->tsize = level_size(opts[0].pi_method);
  c->rsize = level_size(opts[0].pri_method);
  c->ta = c->tsize == 1 ? 1 : c->tsize - 1;
  c->tr = c->rsize == 1 ? 1 : c->rsize - 1;
  *out = c;
  return SDEM_OK;
}

int sdem_create(const sdem_options* o, sdem_ctx** out) { return sdem_create_ensemble(1, o, out); }

int sdem_get_num_boxes(sdem_ctx* c, int32_t* nbox) {
  if (!c || !nbox) return SDEM_E_ARG;
  *nbox = (int32_t)c->boxes.size();
  return SDEM_OK;
}

int sdem_select_box(sdem_ctx* c, int32_t box) {
  if (!c) return SDEM_E_ARG;
  if (box < 0 || box >= (int)c->boxes.size()) return c->fail(SDEM_E_ARG, "box index out of range");
  c->sel = box;
  return SDEM_OK;
}

int sdem_set_linear_stiffness(sdem_ctx* c, double kn, double kt_over_kn) {
  if (!c) return SDEM_E_ARG;
  Box& b = c->box();
  if (b.o.cf_type != SDEM_CFT_LINEAR) return c->fail(SDEM_E_STATE, "cf_type is not the linear model");
  if (!(kn > 0.0) || !(kt_over_kn >= 0.0)) return c->fail(SDEM_E_ARG, "bad linear stiffness");
  if (b.have_pprop) return c->fail(SDEM_E_STATE, "stiffness must be set before the properties");
  b.lin_kn = kn;
  b.lin_kt_ratio = kt_over_kn;
  return SDEM_OK;
}

int sdem_set_particle_property(sdem_ctx* c, int32_t ntype, const double* diameter,
                               const double* density, const double* youngs, const double* poisson,
                               double fs, double fk, double froll, double en, double et) {
  (void)et;
  if (!c) return SDEM_E_ARG;
  Box& b = c->box();
  if (ntype < 1 || ntype > MAX_TYPES || !diameter || !density || !youngs || !poisson)
    return c->fail(SDEM_E_ARG, "bad particle property arguments");
  if (b.have_state) return c->fail(SDEM_E_STATE, "properties must be set before the state");
  if (b.o.cf_type == SDEM_CFT_LINEAR && !(b.lin_kn > 0.0))
    return c->fail(SDEM_E_STATE, "linear model: set the stiffness first");
  const double PI = 3.141592653589793238462643383279502884;
  b.ntype = ntype;
  b.pprop.assign(ntype, PureProp{});
  for (int i = 0; i < ntype; i++) {
    PureProp& p = b.pprop[i];
    p.density = density[i];
    p.radius = 0.5 * diameter[i];
    p.youngs = youngs[i];
    p.poisson = poisson[i];
    p.volume = 4.0 / 3.0 * PI * (p.radius * p.radius * p.radius);
    p.mass = density[i] * p.volume;
    p.inertia = 2.0 / 5.0 * p.mass * (p.radius * p.radius);
    if (!(p.radius > 0.0) || !(p.mass > 0.0)) return c->fail(SDEM_E_ARG, "non-positive diameter or density");
  }
  b.pp_tab.assign((size_t)ntype * ntype, Bnry{});
  for (int i = 0; i < ntype; i++)
    for (int j = 0; j < ntype; j++)
      // the reference's table takes the row type only (Prtcl_Property.f90:137-138);
      // the linear model, which it does not have, takes the true pair
      b.pp_tab[(size_t)i * ntype + j] =
          bnry_prop(b.o.cf_type, b.pprop[i], b.o.cf_type == SDEM_CFT_LINEAR ? b.pprop[j] : b.pprop[i],
                    fs, fk, froll, en, false, b.lin_kn, b.lin_kt_ratio);
  b.have_pprop = true;
  b.have_wprop = false;
  return SDEM_OK;
}

int sdem_set_wall_property(sdem_ctx* c, int32_t nwtype, const double* youngs, const double* poisson,
                           double fs, double fk, double froll, double en, double et) {
  (void)et;
  if (!c) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_pprop) return c->fail(SDEM_E_STATE, "particle property first");
  if (nwtype < 1 || !youngs || !poisson) return c->fail(SDEM_E_ARG, "bad wall property arguments");
  if (b.have_state) return c->fail(SDEM_E_STATE, "properties must be set before the state");
  b.nwtype = nwtype;
  b.wprop.assign(nwtype, PureProp{});
  for (int i = 0; i < nwtype; i++) {
    PureProp& w = b.wprop[i];
    w.density = 1.0e50; w.radius = 1.0e50; w.youngs = youngs[i]; w.poisson = poisson[i];
    w.mass = 1.0e50; w.volume = 1.0e50; w.inertia = 1.0e50;
  }
  b.pw_tab.assign((size_t)b.ntype * nwtype, Bnry{});
  for (int i = 0; i < b.ntype; i++)
    for (int j = 0; j < nwtype; j++)
      b.pw_tab[(size_t)i * nwtype + j] =
          bnry_prop(b.o.cf_type, b.pprop[i], b.wprop[j], fs, fk, froll, en, true, b.lin_kn, b.lin_kt_ratio);
  b.have_wprop = true;
  return SDEM_OK;
}

int sdem_add_plane_wall(sdem_ctx* c, const double p1[3], const double p2[3], const double p3[3],
                        const double p4[3], int32_t user_id, int32_t wall_type, int32_t both_side,
                        int32_t is_infinite, const double trans_vel[3], int32_t* wall_id) {
  if (!c || !p1 || !p2 || !p3 || !p4) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_pprop) return c->fail(SDEM_E_STATE, "particle property first");
  if (b.have_state) return c->fail(SDEM_E_STATE, "walls must be added before the state");
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
  double lenexp = 1.2 * max_type_radius(b);
  V3 dmin = v3(b.o.domain_min), dmax = v3(b.o.domain_max);
  V3 one = {1.0, 1.0, 1.0};
  V3 pmin = vsub(dmin, vmul(one, lenexp)), pmax = vadd(dmax, vmul(one, lenexp));
  if (!wall_in_domain(w, pmin, pmax)) return SDEM_OK;  // skipped, as G_add_PlaneWall does
  if ((int)b.walls.size() >= MAX_WALLS) return c->fail(SDEM_E_CAPACITY, "more than 64 walls");
  b.num_pwall_total++;
  w.wall_id = b.num_pwall_total + b.o.base_wall_id;
  if (w.wall_id > KEY_MASK) return c->fail(SDEM_E_ARG, "wall id exceeds 2^29-1");
  b.walls.push_back(w);
  if (wall_id) *wall_id = w.wall_id;
  return SDEM_OK;
}

int sdem_set_fixed_particles(sdem_ctx* c, int32_t nfix, const int32_t* id, const int32_t* ptype,
                             const double* posr) {
  if (!c || nfix < 0 || (nfix > 0 && (!id || !ptype || !posr))) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_pprop) return c->fail(SDEM_E_STATE, "particle property first");
  if (b.have_state) return c->fail(SDEM_E_STATE, "fixed particles must be set before the state");
  if (nfix > 0 && c->boxes.size() > 1) return c->fail(SDEM_E_ARG, "fixed particles need a single-box context");
  for (int i = 0; i < nfix; i++) {
    if (ptype[i] < 1 || ptype[i] > b.ntype) return c->fail(SDEM_E_ARG, "fixed ptype out of range");
    if (id[i] < 0 || id[i] > KEY_MASK) return c->fail(SDEM_E_ARG, "fixed id out of range");
  }
  b.fix_id.assign(id, id + nfix);
  b.fix_type.assign(ptype, ptype + nfix);
  b.fix_pos.assign(posr, posr + 4 * (size_t)nfix);
  return SDEM_OK;
}

// a new state or new contacts of one box
static int restage(sdem_ctx* c) {
  if (c->stepped && c->boxes.size() > 1)
    return c->fail(SDEM_E_STATE, "an ensemble's states and contacts are set before its first step");
  invalidate_graphs(c);
  c->dev_ready = false;
  c->params_ready = false;
  c->stepped = false;
  return SDEM_OK;
}

int sdem_set_state(sdem_ctx* c, int32_t n, const int32_t* id, const int32_t* ptype,
                   const int32_t* usr_mark, const double* posr, const double* lin_vel,
                   const double* lin_acc, const double* theta, const double* rot_vel,
                   const double* rot_acc) {
  if (!c) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_pprop || !b.have_wprop) return c->fail(SDEM_E_STATE, "particle and wall properties first");
  if (n < 0 || (n > 0 && (!id || !ptype || !posr || !lin_vel || !lin_acc || !theta || !rot_vel || !rot_acc)))
    return c->fail(SDEM_E_ARG, "bad state arguments");
  if (n > b.o.num_prtcl) return c->fail(SDEM_E_CAPACITY, "n exceeds numPrtcl");
  for (int i = 0; i < n; i++) {
    if (ptype[i] < 1 || ptype[i] > b.ntype) return c->fail(SDEM_E_ARG, "ptype out of range");
    if (id[i] < 0 || id[i] > KEY_MASK) return c->fail(SDEM_E_ARG, "particle id out of range");
    if (!(posr[4 * i + 3] > 0.0)) return c->fail(SDEM_E_ARG, "non-positive radius");
    for (int d = 0; d < 3; d++) {
      double p = posr[4 * i + d];
      if (!(p >= b.o.domain_min[d] && p < b.o.domain_max[d]))
        return c->fail(SDEM_E_ARG, "particle outside the domain");
    }
  }
  if (int rc = restage(c)) return rc;
  const int ts = c->tsize, rs = c->rsize;
  b.n = n;
  b.id.assign(id, id + n);
  b.ptype.assign(ptype, ptype + n);
  if (usr_mark) b.mark.assign(usr_mark, usr_mark + n);
  else b.mark.assign(n, 1);
  b.posr.assign(posr, posr + 4 * (size_t)n);
  b.lin_vel.assign(lin_vel, lin_vel + 3 * (size_t)ts * n);
  b.lin_acc.assign(lin_acc, lin_acc + 3 * (size_t)ts * n);
  b.theta.assign(theta, theta + 3 * (size_t)n);
  b.rot_vel.assign(rot_vel, rot_vel + 3 * (size_t)rs * n);
  b.rot_acc.assign(rot_acc, rot_acc + 3 * (size_t)rs * n);
  b.ckind.clear(); b.cowner.clear(); b.cpartner.clear(); b.ctan.clear(); b.cvri.clear();
  b.have_state = true;
  return SDEM_OK;
}

// history records (Add_RestartCntctlink-style install; replaces all history)
int sdem_set_contacts(sdem_ctx* c, int64_t m, const int32_t* kind, const int32_t* owner_id,
                      const int32_t* partner_id, const double* tan_delta, const double* vrel_init) {
  if (!c) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_state) return c->fail(SDEM_E_STATE, "state first");
  if (m < 0 || (m > 0 && (!kind || !owner_id || !partner_id || !tan_delta || !vrel_init)))
    return c->fail(SDEM_E_ARG, "bad contact arguments");
  for (int64_t k = 0; k < m; k++) {
    if (kind[k] < 1 || kind[k] > 4) return c->fail(SDEM_E_ARG, "bad contact kind");
    if (partner_id[k] < 0 || partner_id[k] > KEY_MASK) return c->fail(SDEM_E_ARG, "partner id out of range");
  }
  if (c->stepped) return c->fail(SDEM_E_STATE, "contacts are set with the state, before stepping");
  if (int rc = restage(c)) return rc;
  std::vector<int> ids(b.id);
  std::sort(ids.begin(), ids.end());
  for (int64_t k = 0; k < m; k++) {
    if (!std::binary_search(ids.begin(), ids.end(), owner_id[k])) return c->fail(SDEM_E_ARG, "owner id not present");
    // a PP record sits in the lower id's bucket (AddContactPP,
    // Prtcl_CL_and_CF.f90:142-160), with its partner among the particles
    if (kind[k] == SDEM_KIND_PP &&
        (owner_id[k] >= partner_id[k] || !std::binary_search(ids.begin(), ids.end(), partner_id[k])))
      return c->fail(SDEM_E_ARG, "a PP record belongs to the lower id of two particles of the state");
  }
  b.ckind.assign(kind, kind + m);
  b.cowner.assign(owner_id, owner_id + m);
  b.cpartner.assign(partner_id, partner_id + m);
  b.ctan.assign(tan_delta, tan_delta + 3 * (size_t)m);
  b.cvri.assign(vrel_init, vrel_init + (size_t)m);
  return SDEM_OK;
}

// ===========================================================================
// Stepping
// ===========================================================================

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
  const size_t cap = (size_t)c->nslots;
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
  c->params_ready = false;
  return SDEM_OK;
}

// Candidate-list slots per particle sized to what the failed rebuild counted.
static int grow_vlist(sdem_ctx* c, int list_need) {
  int neu = std::max(list_need + list_need / 4 + 8, 2 * c->lcap);
  if ((int64_t)neu * c->nslots > ((int64_t)1 << 34)) return c->fail(SDEM_E_CAPACITY, "candidate lists too large");
  hipError_t e = regrow(c, &c->vlist, (size_t)c->nslots * neu);
  if (e == hipSuccess) e = regrow(c, &c->vkey, (size_t)c->nslots * neu);
  if (e == hipSuccess) e = regrow(c, &c->cand, (size_t)c->nslots * neu);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(candidate lists)");
  c->lcap = neu;
  c->params_ready = false;
  return SDEM_OK;
}

// After a step that ran out of a capacity (s.overflow 1: history slots, 2:
// candidate-list entries), read from the device into s: the step's input
// buffer is intact, so the step counters and box states go back to the start
// of that step and the capacity grows; the step then runs again.  A state
// that has blown up (3) is an error.
static int rewind_and_grow(sdem_ctx* c, Scalars& s) {
  if (s.overflow == 3)
    return c->fail(SDEM_E_STATE, "a particle is more than one period outside a periodic dimension, or a "
                                 "coordinate is not finite: the state has blown up");
  const int nbox = (int)c->boxes.size();
  const int kind = s.overflow, list_need = s.list_need;
  s.halt = 0;
  s.halted_steps = 0;
  s.cnt_over = 0;
  s.iter = s.iter_at_begin;
  s.need_rebuild = 1;
  s.have_parts = 0;
  s.list_need = 0;
  s.overflow = 0;
  std::vector<BoxState> bs;
  hipError_t e = dl(bs, c->BS, (size_t)nbox);
  for (auto& B : bs) {
    B.n_alive = B.n_stored;
    B.next_wall_update = B.next_wall_at_begin;
    if (c->first_step_pending) B.use_host_max_a = 1;
  }
  if (e == hipSuccess) e = hipMemcpy(c->BS, bs.data(), bs.size() * sizeof(BoxState), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->S, &s, sizeof(s), hipMemcpyHostToDevice);
  // the slots hold what the failing and the halted steps folded in; the
  // replayed step rebuilds its lists, so only its own maxima are needed
  if (e == hipSuccess) e = hipMemset(c->bslots, 0, (size_t)nbox * NSET * NSLOT * sizeof(BoxSlot));
  if (e != hipSuccess) return hip_fail(c, e, "reset scalars");
  c->need_pre = true;
  c->spec_ready = false;
  return kind == 2 ? grow_vlist(c, list_need) : grow_history(c);
}

// Queue one step on the context's stream; ev (SDEM_NUM_PHASES+1 events) may
// be null.  The buffers swap roles afterwards.  A fast step (single box)
// leaves out the sort and list kernels, and its control runs inside the
// contact-test kernel; it halts the call if the box must rebuild.
static hipError_t enqueue_step(sdem_ctx* c, hipEvent_t* ev, bool fast = false) {
  Buf& A = c->buf[c->cur];
  Buf& B = c->buf[1 - c->cur];
  fast = fast && c->D.nbox == 1;
  if (ev) (void)hipEventRecord(ev[0], c->stream);
  hipError_t e = fast ? hipSuccess : launch_ctl(c->D, A, c->stream);
  if (ev) (void)hipEventRecord(ev[1], c->stream);
  if (e == hipSuccess && !fast) e = launch_sort(c->D, A, c->stream);
  if (ev) (void)hipEventRecord(ev[2], c->stream);
  if (e == hipSuccess && !fast) e = launch_vlist(c->D, c->stream);
  if (ev) (void)hipEventRecord(ev[3], c->stream);
  if (e == hipSuccess) e = launch_force(c->D, A, B, c->stream, ev ? ev[4] : nullptr, fast);
  if (ev) (void)hipEventRecord(ev[5], c->stream);
  c->cur = 1 - c->cur;
  return e;
}

// A graph of `n` (1, 2, 8 or 32) consecutive steps starting from buffer
// c->cur (the kernels read their control flags from device memory, so the
// graph is fixed); c->cur is left as it was.  Long batches go in the larger
// graphs: each graph launch costs several microseconds more than a kernel
// inside a graph.
static hipError_t step_graph(sdem_ctx* c, int n, bool fast, hipGraphExec_t* out) {
  hipGraphExec_t& slot = c->gexec[c->occ3][fast ? 1 : 0][n == 1 ? 0 : (n == 2 ? 1 : (n == 8 ? 2 : 3))][c->cur];
  if (slot) {
    *out = slot;
    return hipSuccess;
  }
  const int cur0 = c->cur;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t e1 = hipSuccess;
  for (int k = 0; k < n && e1 == hipSuccess; k++) e1 = enqueue_step(c, nullptr, fast);
  e = hipStreamEndCapture(c->stream, &g);
  c->cur = cur0;
  if (e1 != hipSuccess) e = e1;
  if (e == hipSuccess) e = hipGraphInstantiate(&slot, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess) *out = slot;
  return e;
}

// The force kernel's build for the batch about to run: three waves per
// SIMD where the spheres touch (0.25 contact records per live sphere or
// more at the last step, or among the records installed with the state),
// two otherwise.  Both give the same bits; in contact the third wave's
// gathers in flight outweigh its spills (k_force 458 -> 409 us on a 1M bed
// at 2.3 records per sphere, 979 -> 909 us on 256 settled boxes at 1.0; the
// step 2 % faster at 0.42), without contacts they do not (63.2 -> 63.9 us on
// the falling 1M bed).  A single box's totals come with the scalars read
// after each batch; otherwise the box states are read (once per call: only
// a single box runs its steps in several batches).
static int choose_force_build(sdem_ctx* c, const Scalars* last) {
  int64_t rec = 0, alive = 0;
  if (last) {
    // a single box between the batches of a call: the totals k_finish left
    rec = last->fin_records;
    alive = last->fin_alive;
  } else if (!c->stepped) {
    for (const Box& b : c->boxes) {
      rec += (int64_t)b.ckind.size();
      alive += b.n;
    }
  } else {
    std::vector<BoxState> bs(c->boxes.size());
    hipError_t e = hipMemcpy(bs.data(), c->BS, bs.size() * sizeof(BoxState), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(box states)");
    for (const BoxState& B : bs) {
      for (int k = 0; k < 4; k++) rec += B.num_cntcts[k];
      alive += B.n_alive;
    }
  }
  c->occ3 = (alive > 0 && 4 * rec >= alive) ? 1 : 0;
  if (c->force_build) c->occ3 = c->force_build == 3 ? 1 : 0;
  c->D.occ3 = c->occ3;
  return SDEM_OK;
}

// Steps in batches.  With graphs and a single box, steps are queued on the
// fast path (no sort and list kernels) in batches of spec_chunk, growing
// while they run clean; when a step finds that its box must rebuild, it
// halts the batch before its force kernel (the state buffers and the step
// counters are untouched; contact lists its contact test wrote are written
// again by the replay), the host replays it on the full path and
// speculation resumes after it.  Results do not depend on the
// path: the same criterion decides every rebuild.  A capacity overflow
// halts a batch too: the capacity grows and the failing step is replayed.
int sdem_step(sdem_ctx* c, int32_t nsteps) {
  if (!c) return SDEM_E_ARG;
  if (nsteps < 0) return c->fail(SDEM_E_ARG, "nsteps < 0");
  if (c->phase != 0) return c->fail(SDEM_E_STATE, "a step run phase by phase is open: finish it with sdem_remv_released");
  if (int rc = prepare(c)) return rc;
  const int nbox = (int)c->boxes.size();
  const int NE = SDEM_NUM_PHASES + 1;
  const bool spec_mode = !c->prof && c->use_graph && nbox == 1;
  int remaining = nsteps, retries = 0;
  Scalars last;
  bool have_last = false;  // the scalars after a batch of this call that did work
  while (remaining > 0) {
    if (int rc = prepare(c)) return rc;
    if (int rc = choose_force_build(c, have_last && nbox == 1 ? &last : nullptr)) return rc;
    const bool fast = spec_mode && c->spec_ready;
    const int batch = !spec_mode ? remaining : std::min(remaining, fast ? c->spec_chunk : c->spec_full);
    std::vector<hipEvent_t> evs;
    if (c->prof) evs.resize((size_t)batch * NE);
    for (auto& x : evs)
      if (hipEventCreate(&x) != hipSuccess) return c->fail(SDEM_E_HIP, "hipEventCreate");
    const int c0 = c->cur;
    hipError_t e = hipSuccess;
    if (c->need_pre) {
      // near-wall pass of a state that no step has produced
      e = launch_pre(c->D, c->buf[c->cur], c->stream);
      c->need_pre = false;
    }
    int it = 0;
    if (!c->prof && c->use_graph) {
      for (int big : {32, 8})
        for (; it + big <= batch && e == hipSuccess; it += big) {
          hipGraphExec_t g = nullptr;
          e = step_graph(c, big, fast, &g);
          if (e == hipSuccess) e = hipGraphLaunch(g, c->stream);
        }
      for (; it + 2 <= batch && e == hipSuccess; it += 2) {
        hipGraphExec_t g = nullptr;
        e = step_graph(c, 2, fast, &g);
        if (e == hipSuccess) e = hipGraphLaunch(g, c->stream);
      }
      if (it < batch && e == hipSuccess) {
        hipGraphExec_t g = nullptr;
        e = step_graph(c, 1, fast, &g);
        if (e == hipSuccess) e = hipGraphLaunch(g, c->stream);
        c->cur = 1 - c->cur;
        it++;
      }
    }
    for (; it < batch && e == hipSuccess; it++)
      e = enqueue_step(c, c->prof ? &evs[(size_t)it * NE] : nullptr);
    if (e == hipSuccess) e = launch_finish(c->D, c->buf[c->cur], c->stream);
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
    const int ran = batch - s.halted_steps;  // steps that did work, the last one failed if overflow
    const int ok_steps = s.overflow ? ran - 1 : ran;
    if (ok_steps > 0 && !s.halt) {
      last = s;
      have_last = true;
    }
    for (int k = 0; k < ok_steps && c->prof; k++) {
      for (int q = 0; q < SDEM_NUM_PHASES; q++) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, evs[(size_t)k * NE + q], evs[(size_t)k * NE + q + 1]) == hipSuccess) {
          c->prof_ms[q] += ms;
          c->prof_n[q]++;
          // the reference's timers: the control is the pre-iteration; the
          // fused kernels do the contact search and every force, the
          // accelerations, integration and exchange in one pass, all of it
          // under "CS and CF P-P"
          const int t = q == SDEM_PHASE_CONTROL ? SDEM_TIMER_PRE_ITER : SDEM_TIMER_CSCF_PP;
          c->timer_ms[t] += ms;
          c->timer_ms[SDEM_TIMER_TOTAL] += ms;
        }
      }
      c->timer_n[SDEM_TIMER_TOTAL]++;
      c->timer_n[SDEM_TIMER_PRE_ITER]++;
      c->timer_n[SDEM_TIMER_CSCF_PP]++;
    }
    for (auto& x : evs) (void)hipEventDestroy(x);
    if (ok_steps > 0) {
      c->stepped = true;
      c->first_step_pending = false;
    }
    if (!s.overflow && !s.spec_fail) {
      remaining -= batch;
      if (fast) c->spec_chunk = std::min(2 * c->spec_chunk, 1024);
      c->spec_ready = spec_mode;
      continue;
    }
    c->cur = (c0 + ok_steps) % 2;
    remaining -= ok_steps;
    s.halt = 0;
    s.halted_steps = 0;
    s.cnt_over = 0;
    if (s.spec_fail) {
      // the step after the clean ones must rebuild: replay it on the full
      // path (its force kernel did not run), then speculate again in
      // batches of about the run just seen
      s.spec_fail = 0;
      e = hipMemcpy(c->S, &s, sizeof(s), hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(c, e, "reset scalars");
      c->spec_ready = false;
      c->spec_chunk = std::max(4, ok_steps);
      // boxes that rebuild every step gain nothing from speculation: back off
      c->spec_full = ok_steps == 0 ? std::min(2 * c->spec_full + 1, 256) : 1;
      continue;
    }
    if (s.overflow != 3 && ++retries > 12) return c->fail(SDEM_E_CAPACITY, "capacity still exhausted after regrowing");
    if (int rc = rewind_and_grow(c, s)) return rc;
  }
  return SDEM_OK;
}

// ===========================================================================
// One step phase by phase, in the order of DEMS_iterate
// (Prtcl_DEMSystem.f90:213-248); together the calls are one sdem_step(c, 1)
// and give its bits (the phase kernels are the fused ones' parts)
// ===========================================================================
namespace {
enum { PH_NONE = 0, PH_PRE, PH_COMM, PH_PP, PH_PW, PH_ACC, PH_INT, PH_EXCH, PH_REMV };
const char* const PHASE_CALL[] = {"", "sdem_pre_iteration", "sdem_comm_for_cntct", "sdem_find_contacts",
                                  "sdem_find_contacts_pw", "sdem_clc_acceleration", "sdem_integrate",
                                  "sdem_comm_for_exchange", "sdem_remv_released"};
const int PHASE_TIMER[] = {0, SDEM_TIMER_PRE_ITER, SDEM_TIMER_COMM_CS, SDEM_TIMER_CSCF_PP, SDEM_TIMER_CSCF_PW,
                           SDEM_TIMER_ACCELERATION, SDEM_TIMER_INTEGRATION, SDEM_TIMER_COMM_EXCHANGE,
                           SDEM_TIMER_COMM_EXCHANGE};
}  // namespace

static int run_phase(sdem_ctx* c, int ph) {
  if (!c) return SDEM_E_ARG;
  if (c->phase != ph - 1)
    return c->fail(SDEM_E_STATE, std::string(PHASE_CALL[ph]) + " out of order: the calls follow DEMS_iterate (" +
                                     (c->phase == PH_NONE ? "the step starts with sdem_pre_iteration"
                                                          : std::string("next: ") + PHASE_CALL[c->phase + 1]) + ")");
  if (ph == PH_PRE)
    if (int rc = prepare(c)) return rc;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (c->prof)
    for (auto& x : ev)
      if (hipEventCreate(&x) != hipSuccess) return c->fail(SDEM_E_HIP, "hipEventCreate");
  if (ev[0]) (void)hipEventRecord(ev[0], c->stream);
  Buf& A = c->buf[c->cur];
  Buf& B = c->buf[1 - c->cur];
  hipError_t e = hipSuccess;
  switch (ph) {
    case PH_PRE:  // UpdateNearPrtclsPW's cadence, the step's control; F/T and stale marks need nothing here
      if (c->need_pre) {
        e = launch_pre(c->D, A, c->stream);
        c->need_pre = false;
      }
      if (e == hipSuccess) e = launch_ctl(c->D, A, c->stream);
      break;
    case PH_COMM:  // ghost images: the kernels take minimum images
      break;
    case PH_PP:  // the cell sort and candidate lists (rebuild steps), the contact test, PP/PPG/PPFix forces
      e = launch_sort(c->D, A, c->stream);
      if (e == hipSuccess) e = launch_vlist(c->D, c->stream);
      if (e == hipSuccess) e = launch_phase(c->D, A, B, c->stream, 1);
      break;
    case PH_PW: e = launch_phase(c->D, A, B, c->stream, 2); break;
    case PH_ACC: e = launch_phase(c->D, A, B, c->stream, 3); break;
    case PH_INT: e = launch_phase(c->D, A, B, c->stream, 4); break;
    case PH_EXCH: e = launch_phase(c->D, A, B, c->stream, 5); break;
    case PH_REMV:  // the new history holds this step's records only; the buffers swap
      c->cur = 1 - c->cur;
      e = launch_finish(c->D, c->buf[c->cur], c->stream);
      break;
    default: break;
  }
  if (ev[1]) (void)hipEventRecord(ev[1], c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && ev[0]) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) {
      c->timer_ms[PHASE_TIMER[ph]] += ms;
      c->timer_ms[SDEM_TIMER_TOTAL] += ms;
      if (ph != PH_REMV) c->timer_n[PHASE_TIMER[ph]]++;
      if (ph == PH_REMV) c->timer_n[SDEM_TIMER_TOTAL]++;
    }
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  if (e != hipSuccess) return hip_fail(c, e, PHASE_CALL[ph]);
  if (ph == PH_PP || ph == PH_PW) {
    Scalars s;
    e = hipMemcpy(&s, c->S, sizeof(s), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(scalars)");
    if (s.overflow) {
      // the step's input is intact: back to its start with more capacity
      c->phase = PH_NONE;
      const bool blown = s.overflow == 3;
      if (int rc = rewind_and_grow(c, s)) return rc;
      if (blown) return SDEM_E_STATE;
      return c->fail(SDEM_E_CAPACITY, "a capacity ran out and has been grown: run the step again from "
                                      "sdem_pre_iteration");
    }
  }
  if (ph == PH_REMV) {
    c->phase = PH_NONE;
    c->stepped = true;
    c->first_step_pending = false;
    c->spec_ready = false;
  } else {
    c->phase = ph;
  }
  return SDEM_OK;
}

int sdem_pre_iteration(sdem_ctx* c) { return run_phase(c, PH_PRE); }
int sdem_comm_for_cntct(sdem_ctx* c) { return run_phase(c, PH_COMM); }
int sdem_find_contacts(sdem_ctx* c) { return run_phase(c, PH_PP); }
int sdem_find_contacts_pw(sdem_ctx* c) { return run_phase(c, PH_PW); }
int sdem_clc_acceleration(sdem_ctx* c) { return run_phase(c, PH_ACC); }
int sdem_integrate(sdem_ctx* c) { return run_phase(c, PH_INT); }
int sdem_comm_for_exchange(sdem_ctx* c) { return run_phase(c, PH_EXCH); }
int sdem_remv_released(sdem_ctx* c) { return run_phase(c, PH_REMV); }

// ===========================================================================
// Readers (the selected box; ascending particle id)
// ===========================================================================

// live particles of the selected box in buffer A as (id, storage slot)
static int live_index(sdem_ctx* c, std::vector<std::pair<int, int>>& idx) {
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "sync");
  const Box& b = c->box();
  BoxState B;
  e = hipMemcpy(&B, c->BS + c->sel, sizeof(B), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(box state)");
  const int n = B.n_alive;
  Buf& A = c->buf[c->cur];
  std::vector<int> ids(n), types(n);
  e = hipMemcpy(ids.data(), A.id + b.off, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(types.data(), A.type + b.off, (size_t)n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(ids)");
  idx.clear();
  for (int k = 0; k < n; k++)
    if (types[k] >= 1) idx.push_back({ids[k], b.off + k});
  std::sort(idx.begin(), idx.end());
  return SDEM_OK;
}

// the staged state of the selected box, ascending id
static std::vector<std::pair<int, int>> staged_index(const Box& b) {
  std::vector<std::pair<int, int>> idx;
  for (int i = 0; i < b.n; i++) idx.push_back({b.id[i], i});
  std::sort(idx.begin(), idx.end());
  return idx;
}

int sdem_get_num_particles(sdem_ctx* c, int32_t* n) {
  if (!c || !n) return SDEM_E_ARG;
  const Box& b = c->box();
  if (!c->stepped) { *n = b.have_state ? b.n : 0; return SDEM_OK; }
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  *n = (int32_t)idx.size();
  return SDEM_OK;
}

// With AB2 rotation the force kernel leaves rotVel(2) out of the buffers
// it writes (it equals their linVel(1)); *stored tells whether the current
// buffer's rv array holds it.
static hipError_t rv2_stored(sdem_ctx* c, bool* stored) {
  *stored = true;
  if (c->rsize != 2) return hipSuccess;
  int v = 1;
  const char* at = (const char*)c->S + offsetof(Scalars, rv2_stored) + sizeof(int) * c->cur;
  hipError_t e = hipMemcpy(&v, at, sizeof(int), hipMemcpyDeviceToHost);
  *stored = v != 0;
  return e;
}

// Write rotVel(2) into the current buffer's rv array, so that code outside
// the step kernels may change linVel(1).
static hipError_t store_rv2(sdem_ctx* c) {
  bool st = true;
  hipError_t e = rv2_stored(c, &st);
  if (e != hipSuccess || st) return e;
  const Buf& A = c->buf[c->cur];
  const size_t cap = (size_t)c->nslots;
  e = hipMemcpy(A.rv + 3 * cap, A.lv, 3 * cap * sizeof(double), hipMemcpyDeviceToDevice);
  const int one = 1;
  char* at = (char*)c->S + offsetof(Scalars, rv2_stored) + sizeof(int) * c->cur;
  if (e == hipSuccess) e = hipMemcpy(at, &one, sizeof(int), hipMemcpyHostToDevice);
  return e;
}

int sdem_get_state(sdem_ctx* c, int32_t capn, int32_t* nout, int32_t* id, int32_t* ptype,
                   int32_t* usr_mark, double* posr, double* lin_vel, double* lin_acc, double* theta,
                   double* rot_vel, double* rot_acc) {
  if (!c || !nout) return SDEM_E_ARG;
  const Box& b = c->box();
  if (!b.have_state) return c->fail(SDEM_E_STATE, "no state");
  const int ts = c->tsize, rs = c->rsize;
  if (!c->stepped) {
    auto idx = staged_index(b);
    const int n = (int)idx.size();
    *nout = n;
    if (capn < n) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
    for (int q = 0; q < n; q++) {
      const size_t i = (size_t)idx[q].second;
      if (id) id[q] = b.id[i];
      if (ptype) ptype[q] = b.ptype[i];
      if (usr_mark) usr_mark[q] = b.mark[i];
      if (posr) std::copy(&b.posr[4 * i], &b.posr[4 * i] + 4, posr + 4 * (size_t)q);
      if (lin_vel) std::copy(&b.lin_vel[3 * ts * i], &b.lin_vel[3 * ts * i] + 3 * ts, lin_vel + 3 * ts * (size_t)q);
      if (lin_acc) std::copy(&b.lin_acc[3 * ts * i], &b.lin_acc[3 * ts * i] + 3 * ts, lin_acc + 3 * ts * (size_t)q);
      if (theta) std::copy(&b.theta[3 * i], &b.theta[3 * i] + 3, theta + 3 * (size_t)q);
      if (rot_vel) std::copy(&b.rot_vel[3 * rs * i], &b.rot_vel[3 * rs * i] + 3 * rs, rot_vel + 3 * rs * (size_t)q);
      if (rot_acc) std::copy(&b.rot_acc[3 * rs * i], &b.rot_acc[3 * rs * i] + 3 * rs, rot_acc + 3 * rs * (size_t)q);
    }
    return SDEM_OK;
  }
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  int n = (int)idx.size();
  *nout = n;
  if (capn < n) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
  Buf& A = c->buf[c->cur];
  // the selected box's slots only: [off, off + bcap) of every array
  const size_t cap = (size_t)c->nslots, off = (size_t)b.off, len = (size_t)b.bcap;
  std::vector<double> lv, la, th, rv, ra;
  std::vector<double4> p4;
  std::vector<int> ty, mk;
  hipError_t e = dl_box(p4, A.pos, cap, 1, off, len);
  if (e == hipSuccess) e = dl_box(lv, A.lv, cap, 3 * ts, off, len);
  if (e == hipSuccess) e = dl_box(la, A.la, cap, 3 * c->ta, off, len);
  if (e == hipSuccess) e = dl_box(th, A.th, cap, 3, off, len);
  if (e == hipSuccess) e = dl_box(rv, A.rv, cap, 3 * rs, off, len);
  if (e == hipSuccess) e = dl_box(ra, A.ra, cap, 3 * c->tr, off, len);
  if (e == hipSuccess) e = dl_box(ty, A.type, cap, 1, off, len);
  if (e == hipSuccess) e = dl_box(mk, A.mark, cap, 1, off, len);
  bool rv2 = true;
  if (e == hipSuccess) e = rv2_stored(c, &rv2);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(state)");
  for (int q = 0; q < n; q++) {
    const size_t k = (size_t)idx[q].second - off;
    if (id) id[q] = idx[q].first;
    if (ptype) ptype[q] = ty[k];
    if (usr_mark) usr_mark[q] = mk[k];
    if (posr) { posr[4 * q] = p4[k].x; posr[4 * q + 1] = p4[k].y; posr[4 * q + 2] = p4[k].z; posr[4 * q + 3] = p4[k].w; }
    for (int l = 0; l < ts; l++)
      for (int d = 0; d < 3; d++) {
        if (lin_vel) lin_vel[((size_t)q * ts + l) * 3 + d] = lv[((size_t)l * 3 + d) * len + k];
        // stored level l-1 holds linAcc(l); linAcc(1) is the stored level 0 (FE: the only one)
        if (lin_acc) lin_acc[((size_t)q * ts + l) * 3 + d] = la[((size_t)(l == 0 ? 0 : l - 1) * 3 + d) * len + k];
      }
    if (theta) for (int d = 0; d < 3; d++) theta[3 * q + d] = th[(size_t)d * len + k];
    for (int l = 0; l < rs; l++)
      for (int d = 0; d < 3; d++) {
        if (rot_vel)
          rot_vel[((size_t)q * rs + l) * 3 + d] =
              (l == 1 && !rv2) ? lv[(size_t)d * len + k] : rv[((size_t)l * 3 + d) * len + k];
        if (rot_acc) rot_acc[((size_t)q * rs + l) * 3 + d] = ra[((size_t)(l == 0 ? 0 : l - 1) * 3 + d) * len + k];
      }
  }
  return SDEM_OK;
}

int sdem_get_forces(sdem_ctx* c, int32_t capn, int32_t* id, double* f, double* t) {
  if (!c) return SDEM_E_ARG;
  const Box& b = c->box();
  if (!b.have_state) return c->fail(SDEM_E_STATE, "no state");
  if (!c->stepped) {
    auto idx = staged_index(b);
    if (capn < (int)idx.size()) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
    for (size_t q = 0; q < idx.size(); q++) {
      if (id) id[q] = idx[q].first;
      for (int d = 0; d < 3; d++) {
        if (f) f[3 * q + d] = 0.0;
        if (t) t[3 * q + d] = 0.0;
      }
    }
    return SDEM_OK;
  }
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  int n = (int)idx.size();
  if (capn < n) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
  std::vector<double> fb;
  const size_t cap = (size_t)c->nslots, off = (size_t)b.off, len = (size_t)b.bcap;
  hipError_t e = dl_box(fb, c->buf[c->cur].force, cap, 6, off, len);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(forces)");
  for (int q = 0; q < n; q++) {
    const size_t k = (size_t)idx[q].second - off;
    if (id) id[q] = idx[q].first;
    for (int d = 0; d < 3; d++) {
      if (f) f[3 * q + d] = fb[(size_t)d * len + k];
      if (t) t[3 * q + d] = fb[(size_t)(3 + d) * len + k];
    }
  }
  return SDEM_OK;
}

struct Rec {
  int owner, partner, kind;
  double4 v;
};

static int collect(sdem_ctx* c, std::vector<Rec>& recs) {
  recs.clear();
  const Box& b = c->box();
  if (!c->stepped) {
    for (size_t k = 0; k < b.ckind.size(); k++)
      recs.push_back({b.cowner[k], b.cpartner[k], b.ckind[k],
                      make_double4(b.ctan[3 * k], b.ctan[3 * k + 1], b.ctan[3 * k + 2], b.cvri[k])});
  } else {
    std::vector<std::pair<int, int>> idx;
    if (int rc = live_index(c, idx)) return rc;
    Buf& A = c->buf[c->cur];
    const size_t cap = (size_t)c->nslots, off = (size_t)b.off, len = (size_t)b.bcap;
    std::vector<int> hc, hh, hk;
    std::vector<double4> hv;
    // keys: the box's heads (KEY_HEAD per particle, side by side), then
    // rows KEY_HEAD.. of the box's slots
    hipError_t e = dl_box(hc, A.hcnt, cap, 1, off, len);
    if (e == hipSuccess) e = dl_box(hh, A.hkey, KEY_HEAD * cap, 1, KEY_HEAD * off, KEY_HEAD * len);
    if (e == hipSuccess)
      e = dl_box(hk, A.hkey + KEY_HEAD * cap, cap, (size_t)(c->hcap - KEY_HEAD), off, len);
    if (e == hipSuccess) e = dl_box(hv, A.hval, cap, (size_t)c->hcap, off, len);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(history)");
    for (auto& pr : idx) {
      const size_t k = (size_t)pr.second - off;
      for (int q = 0; q < hc[k]; q++) {
        const int key = q < KEY_HEAD ? hh[KEY_HEAD * k + q] : hk[(size_t)(q - KEY_HEAD) * len + k];
        if ((int)((uint32_t)key >> KEY_BITS) == KIND_COPY) continue;
        recs.push_back({pr.first, key & KEY_MASK, (int)((uint32_t)key >> KEY_BITS), hv[(size_t)q * len + k]});
      }
    }
  }
  std::sort(recs.begin(), recs.end(), [](const Rec& x, const Rec& y) {
    return x.owner != y.owner ? x.owner < y.owner : x.partner < y.partner;
  });
  return SDEM_OK;
}

int sdem_get_num_contacts(sdem_ctx* c, int64_t* m) {
  if (!c || !m) return SDEM_E_ARG;
  std::vector<Rec> recs;
  if (int rc = collect(c, recs)) return rc;
  *m = (int64_t)recs.size();
  return SDEM_OK;
}

int sdem_get_contacts(sdem_ctx* c, int64_t capm, int64_t* m, int32_t* kind, int32_t* owner,
                      int32_t* partner, double* td, double* vr) {
  if (!c || !m) return SDEM_E_ARG;
  std::vector<Rec> recs;
  if (int rc = collect(c, recs)) return rc;
  *m = (int64_t)recs.size();
  if (capm < *m) return c->fail(SDEM_E_CAPACITY, "output arrays too small");
  for (size_t q = 0; q < recs.size(); q++) {
    const Rec& r = recs[q];
    if (kind) kind[q] = r.kind;
    if (owner) owner[q] = r.owner;
    if (partner) partner[q] = r.partner;
    if (td) { td[3 * q] = r.v.x; td[3 * q + 1] = r.v.y; td[3 * q + 2] = r.v.z; }
    if (vr) vr[q] = r.v.w;
  }
  return SDEM_OK;
}

static int read_box_state(sdem_ctx* c, BoxState* B, Scalars* s) {
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && B) e = hipMemcpy(B, c->BS + c->sel, sizeof(*B), hipMemcpyDeviceToHost);
  if (e == hipSuccess && s) e = hipMemcpy(s, c->S, sizeof(*s), hipMemcpyDeviceToHost);
  return e == hipSuccess ? SDEM_OK : hip_fail(c, e, "hipMemcpy(scalars)");
}

int sdem_get_counts(sdem_ctx* c, int32_t num[4]) {
  if (!c || !num) return SDEM_E_ARG;
  for (int k = 0; k < 4; k++) num[k] = 0;
  if (!c->stepped) return SDEM_OK;
  BoxState B;
  if (int rc = read_box_state(c, &B, nullptr)) return rc;
  for (int k = 0; k < 4; k++) num[k] = B.num_cntcts[k];
  return SDEM_OK;
}

int sdem_get_rebuilds(sdem_ctx* c, int64_t* rebuilds) {
  if (!c || !rebuilds) return SDEM_E_ARG;
  *rebuilds = 0;
  if (!c->stepped) return SDEM_OK;
  BoxState B;
  Scalars s;
  if (int rc = read_box_state(c, &B, &s)) return rc;
  *rebuilds = s.rebuilds;
  return SDEM_OK;
}

int sdem_get_iter(sdem_ctx* c, int32_t* iter, int32_t* next_update) {
  if (!c) return SDEM_E_ARG;
  if (!c->stepped) {
    if (iter) *iter = c->box().o.ifirst - 1;
    if (next_update) *next_update = 0;
    return SDEM_OK;
  }
  BoxState B;
  Scalars s;
  if (int rc = read_box_state(c, &B, &s)) return rc;
  if (iter) *iter = s.iter;
  if (next_update) *next_update = B.next_wall_update;
  return SDEM_OK;
}

int sdem_set_profiling(sdem_ctx* c, int32_t enable) {
  if (!c) return SDEM_E_ARG;
  c->prof = enable != 0;
  for (int q = 0; q < SDEM_NUM_PHASES; q++) { c->prof_ms[q] = 0.0; c->prof_n[q] = 0; }
  for (int q = 0; q < SDEM_NUM_TIMERS; q++) { c->timer_ms[q] = 0.0; c->timer_n[q] = 0; }
  return SDEM_OK;
}

int sdem_get_timers(sdem_ctx* c, double tot_ms[SDEM_NUM_TIMERS], int64_t calls[SDEM_NUM_TIMERS]) {
  if (!c || !tot_ms || !calls) return SDEM_E_ARG;
  for (int q = 0; q < SDEM_NUM_TIMERS; q++) {
    tot_ms[q] = c->timer_ms[q];
    calls[q] = c->timer_n[q];
  }
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
  if (skin_factor >= 0.0 && skin_factor > 0.45) return c->fail(SDEM_E_ARG, "skin factor must be below 0.45");
  if (cell_factor)
    for (int d = 0; d < 3; d++)
      if (cell_factor[d] < 0.0 || cell_factor[d] > 64.0) return c->fail(SDEM_E_ARG, "cell factor out of range");
  if (skin_factor >= 0.0) c->skin_factor = skin_factor;
  if (cell_factor)
    for (int d = 0; d < 3; d++)
      if (cell_factor[d] > 0.0) c->cell_factor[d] = cell_factor[d];
  if (keep_forces >= 0) c->keep_forces = keep_forces ? 1 : 0;
  // new grids and parameters at the next step, which rebuilds the lists
  c->params_ready = false;
  return SDEM_OK;
}

int sdem_set_force_build(sdem_ctx* c, int32_t waves_per_simd) {
  if (!c) return SDEM_E_ARG;
  if (waves_per_simd != 0 && waves_per_simd != 2 && waves_per_simd != 3)
    return c->fail(SDEM_E_ARG, "force build: 0 (chosen per batch), 2 or 3 waves per SIMD");
  c->force_build = waves_per_simd;
  return SDEM_OK;
}

int sdem_set_block_sizes(sdem_ctx* c, int32_t contact_block, int32_t force_block) {
  if (!c) return SDEM_E_ARG;
  auto ok = [](int b) { return b == 64 || b == 128 || b == 256; };
  if (!ok(contact_block) || !ok(force_block)) return c->fail(SDEM_E_ARG, "block sizes must be 64, 128 or 256");
  c->bs_contact = contact_block;
  c->bs_force = force_block;
  c->params_ready = false;
  return SDEM_OK;
}

int sdem_get_tuning(sdem_ctx* c, double cell_factor[3], double* skin_factor, int32_t* contact_block,
                    int32_t* force_block) {
  if (!c) return SDEM_E_ARG;
  if (cell_factor) for (int d = 0; d < 3; d++) cell_factor[d] = c->cell_factor[d];
  if (skin_factor) *skin_factor = c->skin_factor;
  if (contact_block) *contact_block = c->bs_contact;
  if (force_block) *force_block = c->bs_force;
  return SDEM_OK;
}

// Autotuner: every candidate (cell shape, skin, block sizes) runs
// `steps_per_trial` real steps after one settling step, timed with HIP
// events on the library's stream; the fastest is kept.  The results do not
// depend on the tuning, so the trial steps are ordinary steps of the run.
int sdem_autotune(sdem_ctx* c, int32_t steps_per_trial, int32_t* steps_run, double* best_ms_per_step) {
  if (!c) return SDEM_E_ARG;
  if (steps_per_trial < 2) return c->fail(SDEM_E_ARG, "steps_per_trial must be at least 2");
  struct Cand { double cf, skin; int bc, bf; double ms; };
  std::vector<Cand> cands;
  for (double cf : {0.5, 0.75, 1.0})
    for (double sk : {0.03, 0.06, 0.12, 0.2})
      for (int bf : {128, 256}) cands.push_back({cf, sk, 256, bf, 0.0});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return c->fail(SDEM_E_HIP, "hipEventCreate");
  int total = 0, rc = SDEM_OK;
  const bool prof = c->prof;
  c->prof = false;
  // The run moves on while the trials run (a settling pack speeds up, so
  // does its rebuild rate): every setting is timed once in a forward and
  // once in a backward pass and scored by the mean, which cancels a drift
  // that is linear in time.
  const int nc = (int)cands.size();
  for (int pass = 0; pass < 2 && rc == SDEM_OK; pass++)
    for (int q = 0; q < nc; q++) {
      Cand& k = cands[pass == 0 ? q : nc - 1 - q];
      const double cf[3] = {k.cf, k.cf, k.cf};
      if ((rc = sdem_set_tuning(c, cf, k.skin, -1))) break;
      if ((rc = sdem_set_block_sizes(c, k.bc, k.bf))) break;
      if ((rc = sdem_step(c, 2))) break;  // new grids: the first steps rebuild
      total += 2;
      (void)hipEventRecord(e0, c->stream);
      if ((rc = sdem_step(c, steps_per_trial))) break;
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      total += steps_per_trial;
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      k.ms += 0.5 * ms / steps_per_trial;
    }
  c->prof = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  const Cand* bestc = &cands[0];
  for (const Cand& k : cands)
    if (k.ms < bestc->ms) bestc = &k;
  const double cf[3] = {bestc->cf, bestc->cf, bestc->cf};
  if ((rc = sdem_set_tuning(c, cf, bestc->skin, -1))) return rc;
  if ((rc = sdem_set_block_sizes(c, bestc->bc, bestc->bf))) return rc;
  if (steps_run) *steps_run = total;
  if (best_ms_per_step) *best_ms_per_step = bestc->ms;
  return SDEM_OK;
}

int sdem_set_capacity(sdem_ctx* c, int32_t hist_slots, int32_t list_slots) {
  if (!c) return SDEM_E_ARG;
  if (c->dev_ready) return c->fail(SDEM_E_STATE, "capacities must be set before the first step");
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

// ===========================================================================
// Access for the learned surrogate (gnn_api.cpp), single-box contexts only.
// ===========================================================================
extern "C" int sdem_internal_view(sdem_ctx* c, sdem_dem_view* v) {
  if (!c || !v) return SDEM_E_ARG;
  if (c->boxes.size() != 1) return c->fail(SDEM_E_STATE, "the surrogate drives single-box contexts");
  if (int rc = prepare(c)) return rc;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = store_rv2(c);  // the caller may change linVel(1)
  BoxState B;
  if (e == hipSuccess) e = hipMemcpy(&B, c->BS, sizeof(B), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "surrogate view");
  const Box& b = c->boxes[0];
  Buf& A = c->buf[c->cur];
  v->pos = A.pos;
  v->lv = A.lv;
  v->la = A.la;
  v->id = A.id;
  v->cap = (int)c->nslots;
  v->n = B.n_alive;
  v->dt = b.o.dt;
  for (int d = 0; d < 3; d++) {
    v->g[d] = b.o.gravity[d];
    v->lo[d] = b.o.domain_min[d];
    v->hi[d] = b.o.domain_max[d];
    v->per[d] = b.o.is_periodic[d] ? 1 : 0;
  }
  v->stream = c->stream;
  return SDEM_OK;
}

// The state in buffer A was advanced by `steps` steps outside the DEM
// kernels: the velocity history restarts from the new velocity, contact
// history is cleared, the near-wall list and the candidate lists are
// rebuilt at the next step, and iterNumber moves on.
extern "C" int sdem_internal_after_update(sdem_ctx* c, int32_t steps) {
  if (!c || steps < 0) return SDEM_E_ARG;
  Buf& A = c->buf[c->cur];
  const size_t cap = (size_t)c->nslots;
  hipError_t e = hipSuccess;
  for (int l = 1; l < c->tsize && e == hipSuccess; l++)
    e = hipMemcpy(A.lv + (size_t)l * 3 * cap, A.lv, 3 * cap * sizeof(double), hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = hipMemset(A.hcnt, 0, cap * sizeof(int));
  Scalars s;
  BoxState B;
  if (e == hipSuccess) e = hipMemcpy(&s, c->S, sizeof(s), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&B, c->BS, sizeof(B), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(c, e, "surrogate update");
  s.iter += steps;
  s.need_rebuild = 1;
  s.have_parts = 0;
  B.next_wall_update = s.iter;
  B.n_stored = B.n_alive;
  for (int k = 0; k < 4; k++) B.num_cntcts[k] = 0;
  e = hipMemcpy(c->S, &s, sizeof(s), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->BS, &B, sizeof(B), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(c->bslots, 0, c->boxes.size() * NSET * NSLOT * sizeof(BoxSlot));
  if (e != hipSuccess) return hip_fail(c, e, "surrogate update");
  c->need_pre = true;
  c->spec_ready = false;  // the next step rebuilds
  c->stepped = true;
  c->first_step_pending = false;
  return SDEM_OK;
}

extern "C" int sdem_internal_box_info(sdem_ctx* c, sdem_box_info* out) {
  if (!c || !out) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_pprop) return c->fail(SDEM_E_STATE, "particle properties first");
  static thread_local std::vector<double> rad;
  static thread_local std::vector<int32_t> wid;
  rad.resize(b.ntype);
  for (int t = 0; t < b.ntype; t++) rad[t] = b.pprop[t].radius;
  wid.resize(b.walls.size());
  for (size_t w = 0; w < b.walls.size(); w++) wid[w] = b.walls[w].wall_id;
  out->tsize = c->tsize;
  out->rsize = c->rsize;
  out->num_prtcl = b.o.num_prtcl;
  out->ntype = b.ntype;
  out->nwall = (int32_t)b.walls.size();
  out->nfix = (int32_t)b.fix_id.size();
  out->keep_forces = c->keep_forces;
  for (int d = 0; d < 3; d++) {
    out->lo[d] = b.o.domain_min[d];
    out->hi[d] = b.o.domain_max[d];
    out->periodic[d] = b.o.is_periodic[d] ? 1 : 0;
  }
  out->type_radius = rad.data();
  out->wall_id = wid.data();
  out->fix_id = b.fix_id.data();
  return SDEM_OK;
}

extern "C" int sdem_internal_set_marks(sdem_ctx* c, int32_t n, const int32_t* id, const int32_t* mark) {
  if (!c || n < 0 || (n > 0 && (!id || !mark))) return SDEM_E_ARG;
  Box& b = c->box();
  if (!b.have_state) return c->fail(SDEM_E_STATE, "no state");
  std::vector<std::pair<int, int>> want(n);
  for (int i = 0; i < n; i++) want[i] = {id[i], mark[i]};
  std::sort(want.begin(), want.end());
  auto find = [&](int pid) -> const std::pair<int, int>* {
    auto it = std::lower_bound(want.begin(), want.end(), std::make_pair(pid, INT32_MIN));
    return (it != want.end() && it->first == pid) ? &*it : nullptr;
  };
  if (!c->stepped) {
    for (int i = 0; i < b.n; i++)
      if (auto p = find(b.id[i])) b.mark[i] = p->second;
    return SDEM_OK;
  }
  std::vector<std::pair<int, int>> idx;
  if (int rc = live_index(c, idx)) return rc;
  Buf& A = c->buf[c->cur];
  std::vector<int> mk;
  hipError_t e = dl(mk, A.mark, (size_t)c->nslots);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(marks)");
  for (const auto& q : idx)
    if (auto p = find(q.first)) mk[q.second] = p->second;
  // both buffers: the force kernel carries marks over only after re-sorts
  // (after one, the other buffer is rewritten from this one anyway)
  e = hipMemcpy(A.mark, mk.data(), mk.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->buf[1 - c->cur].mark, mk.data(), mk.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(c, e, "hipMemcpy(marks)");
  return SDEM_OK;
}

extern "C" int sdem_internal_fail(sdem_ctx* c, int32_t code, const char* msg) {
  if (!c) return code;
  return c->fail(code, msg ? msg : "");
}
