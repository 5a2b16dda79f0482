/*
 * sdem_oracle.c -- CPU restatement of sDEM's per-timestep DEM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libsdem_hip.so) never does.
 *
 * It restates, single-threaded and in the reference's own evaluation order,
 * the Fortran code of GongZheng-Justin/sDEM for one MPI rank owning the whole
 * domain (all citations are paths under src/DEM/):
 *   - vector operators and their operation order     Prtcl_TypeDef.f90:105-288
 *   - pair-coefficient tables                         Prtcl_Property.f90:62-247
 *   - plane walls (construction, in-domain test,
 *     contact test with quad/edge/vertex fallbacks)   Prtcl_Geometry.f90:55-537
 *   - near-wall lists and their update cadence        Prtcl_ContactSearchPW.f90:49-319
 *   - periodic ghost images (single rank)             Prtcl_Comm.f90:217-1205
 *   - NBS-Munjiza broad phase with linked lists       Prtcl_NBS_Munjiza.f90:57-557
 *   - hierarchical (multi-level) NBS-Munjiza          Prtcl_Hrchl_Munjiza.f90:84-863
 *   - contact list buckets, Find_insert, PreIteration,
 *     RemvReleased                                    Prtcl_CL_and_CF.f90:66-572
 *   - contact force include                           Prtcl_ContactForce_inc.f90:1-86
 *     and its four call sites                         Prtcl_CL_and_CF.f90:732-856
 *   - accelerations                                   Prtcl_DEMSystem.f90:360-375
 *   - FE / AB2 / AB3 integration (incl. AB2 rotation
 *     reading the linear velocity)                    Prtcl_Integration.f90:15-88
 *   - periodic wrap / open-side deletion              Prtcl_Comm.f90:1287-2111
 *   - step order                                      Prtcl_DEMSystem.f90:201-331
 *
 * Build with -ffp-contract=off: every multiply and add is rounded separately,
 * as the reference's operator functions are.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/sdem.h"

/* ------------------------------------------------------------------------ */
/* real3 / real4 and the operators of Prtcl_TypeDef.f90                      */
/* ------------------------------------------------------------------------ */
typedef struct { double x, y, z; } r3;
typedef struct { double x, y, z, w; } r4;

static inline r3 R3(double x, double y, double z) { r3 r = {x, y, z}; return r; }
static const r3 ZERO3 = {0.0, 0.0, 0.0};
/* real3_add1 :105 */
static inline r3 add3(r3 a, r3 b) { return R3(a.x + b.x, a.y + b.y, a.z + b.z); }
/* real3_sub :144 */
static inline r3 sub3(r3 a, r3 b) { return R3(a.x - b.x, a.y - b.y, a.z - b.z); }
/* real3_prod2 / real3_prod2_2 :158-172 (p%x*v) */
static inline r3 mul3(r3 a, double v) { return R3(a.x * v, a.y * v, a.z * v); }
/* real3_division2 :174 */
static inline r3 div3(r3 a, double v) { return R3(a.x / v, a.y / v, a.z / v); }
/* real3_dotProd :203  (x*x' + y*y') + z*z' */
static inline double dot3(r3 a, r3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* real3_CrossProd :210 */
static inline r3 cross3(r3 u, r3 v) {
  return R3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
/* real3_norm :240 */
static inline double norm3(r3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* real3_nv / real4_nv :224-281: unit(p1-p2), zero when |p1-p2| <= 1e-10 */
static inline r3 nv3(r3 p1, r3 p2) {
  r3 s = sub3(p1, p2);
  double n = norm3(s);
  if (n > 1.0e-10) return div3(s, n);
  return ZERO3;
}
static inline r3 xyz(r4 p) { return R3(p.x, p.y, p.z); }
/* min_v / max_v :297-309 (Fortran min/max of two reals) */
static inline double fmin_f(double a, double b) { return a < b ? a : b; }
static inline double fmax_f(double a, double b) { return a > b ? a : b; }
static inline r3 min3(r3 a, r3 b) { return R3(fmin_f(a.x, b.x), fmin_f(a.y, b.y), fmin_f(a.z, b.z)); }
static inline r3 max3(r3 a, r3 b) { return R3(fmax_f(a.x, b.x), fmax_f(a.y, b.y), fmax_f(a.z, b.z)); }

/* ------------------------------------------------------------------------ */
/* properties, Prtcl_Property.f90                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  double radius, density, poisson, youngs, mass, volume, inertia;
} pure_prop;

typedef struct {
  double rad_eff, mass_eff, kn, kt, dn, dt, rest_n, rest_t, fric_s, fric_k, fric_roll;
} bnry_prop;

/* ------------------------------------------------------------------------ */
/* plane wall, Prtcl_Geometry.f90:17-37                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  r3 p1, p2, p3, p4, trans_vel;
  int user_id, wall_type, both_side, is_infinite, wall_id;
  double d;
  r3 n, min_point, max_point;
} plane_wall;

/* ------------------------------------------------------------------------ */
/* hierarchical level, Prtcl_Hrchl_Munjiza.f90                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nx, ny, nz, multiple;
  double cell_len;
} hlevel;

typedef struct oracle_ctx {
  sdem_options o;
  int tsize, rsize;
  char err[256];

  /* properties */
  int ntype, nwtype;
  pure_prop *pprop;   /* 1..ntype */
  pure_prop *wprop;   /* 1..nwtype */
  bnry_prop *pp_tab;  /* (i,j) -> [(i-1)*ntype + (j-1)] */
  bnry_prop *pw_tab;  /* (i,w) -> [(i-1)*nwtype + (w-1)] */
  int have_pprop, have_wprop;
  double lin_kn, lin_kt_ratio; /* constant-stiffness linear model */

  /* geometry */
  int nwall, mwall, num_pwall_total;
  plane_wall *walls;  /* 1..nwall */

  /* moving particles, 1-based, capacity mlocal */
  int mlocal, nlocal;
  int *id, *ptype, *mark;
  r4 *pos;
  r3 *lvel, *lacc, *rvel, *racc; /* [(i-1)*size + (l-1)] */
  r3 *theta, *force, *torque;
  int have_state;

  /* fixed particles (local + their periodic images) */
  int nfix_local, nfix; /* nfix = local + ghost images */
  int *fix_id, *fix_type;
  r4 *fix_pos;

  /* ghosts (periodic images), 1-based */
  int nghost, mghost;
  int *g_id, *g_type;
  r4 *g_pos;
  r3 *g_lvel, *g_rvel;

  /* contact list, Prtcl_CL_and_CF.f90:18-27 */
  int max_cntcts, next_insert;
  int *bucket;     /* 1..mlocal */
  int *id_j, *next, *status;
  r3 *tandel;
  double *vrel;
  int num_cntcts[4];
  int *kind;       /* kind of each record (for output only) */

  /* near-wall lists, Prtcl_ContactSearchPW.f90:19-31 */
  int *bucket_pw, *id_wall, *next_pw;
  int max_near, next_insert_pw, next_iter_update;
  double max_wall_vel, dxmax1, dxmax2;

  /* search grid (NBS) */
  double xst_cs, yst_cs, zst_cs, cell_len;
  int nx, ny, nz;
  int mbox;
  int *bx, *by, *bz, *nextx, *nexty, *nextz;
  int *heady, *headx, *headx0;
  int *headz, *headz0; /* (0:1,0:nz+1) and (0:2,0:nz+1) */
  int64_t num_cnsv[2];

  /* hierarchical */
  int nlvl;
  hlevel lvl[8];
  int *type_level; /* 1..ntype */

  int pw_initialized;
  int iter_number;
  double simlen[3];
  double lenforcs;
} oracle_ctx;

static int fail(oracle_ctx *c, int code, const char *msg) {
  if (c) snprintf(c->err, sizeof(c->err), "%s", msg);
  return code;
}

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ------------------------------------------------------------------------ */
/* options                                                                   */
/* ------------------------------------------------------------------------ */
int oracle_options_default(sdem_options *o) {
  /* DEMS_Options defaults, Prtcl_Parameters.f90:45-95 */
  memset(o, 0, sizeof(*o));
  o->dt = 1.0e-5;
  o->gravity[0] = 0.0; o->gravity[1] = -9.81; o->gravity[2] = 0.0;
  o->prtcl_cs_ratio = 1.0;
  o->wall_neighbor_ratio = 2.0;
  o->cs_method = SDEM_CSM_NBS_MUNJIZA;
  o->cf_type = SDEM_CFT_LSD;
  o->ct_model = SDEM_CTM_CONSTANT_TORQUE;
  o->pi_method = SDEM_PIM_FE;
  o->pri_method = SDEM_PIM_FE;
  o->cs_numlvls = 0;
  o->cntct_list_size = 6;
  o->base_wall_id = 10000000;
  o->wall_max_update_iter = 100;
  o->num_prtcl = 8000;
  o->ifirst = 1;
  return SDEM_OK;
}

static int level_size(int m) { return m == SDEM_PIM_FE ? 1 : (m == SDEM_PIM_AB2 ? 2 : 3); }

int oracle_create(const sdem_options *o, oracle_ctx **out) {
  if (!o || !out) return SDEM_E_ARG;
  oracle_ctx *c = (oracle_ctx *)xcalloc(1, sizeof(oracle_ctx));
  c->o = *o;
  if (o->pi_method < 1 || o->pi_method > 3 || o->pri_method < 1 || o->pri_method > 3 ||
      o->cf_type < 1 || o->cf_type > 3) {
    free(c);
    return SDEM_E_ARG;
  }
  c->tsize = level_size(o->pi_method);
  c->rsize = level_size(o->pri_method);
  /* single rank: mlocal = min(int(1.5*numPrtcl), numPrtcl), >= 10 (Prtcl_Variables.f90:118-122) */
  c->mlocal = o->num_prtcl < 10 ? 10 : o->num_prtcl;
  int m = c->mlocal;
  c->id = xcalloc(m + 1, sizeof(int));
  c->ptype = xcalloc(m + 1, sizeof(int));
  c->mark = xcalloc(m + 1, sizeof(int));
  c->pos = xcalloc(m + 1, sizeof(r4));
  c->lvel = xcalloc((size_t)(m + 1) * c->tsize, sizeof(r3));
  c->lacc = xcalloc((size_t)(m + 1) * c->tsize, sizeof(r3));
  c->rvel = xcalloc((size_t)(m + 1) * c->rsize, sizeof(r3));
  c->racc = xcalloc((size_t)(m + 1) * c->rsize, sizeof(r3));
  c->theta = xcalloc(m + 1, sizeof(r3));
  c->force = xcalloc(m + 1, sizeof(r3));
  c->torque = xcalloc(m + 1, sizeof(r3));
  /* contact list, CL_InitContactList :66-103 */
  c->max_cntcts = o->cntct_list_size * m;
  int mc = c->max_cntcts;
  c->bucket = xcalloc(m + 1, sizeof(int));
  c->id_j = xcalloc(mc + 1, sizeof(int));
  c->next = xcalloc(mc + 1, sizeof(int));
  c->status = xcalloc(mc + 1, sizeof(int));
  c->tandel = xcalloc(mc + 1, sizeof(r3));
  c->vrel = xcalloc(mc + 1, sizeof(double));
  c->kind = xcalloc(mc + 1, sizeof(int));
  for (int i = 1; i <= mc; i++) { c->next[i] = -i - 1; c->status[i] = -1; }
  c->next[mc] = 0;
  c->next_insert = 1;
  /* near-wall list, InitContactSearchPW :56-63 */
  c->max_near = (int)(1.5 * m) + 1;
  c->bucket_pw = xcalloc(m + 1, sizeof(int));
  c->id_wall = xcalloc(c->max_near + 1, sizeof(int));
  c->next_pw = xcalloc(c->max_near + 1, sizeof(int));
  c->next_iter_update = 0;
  c->iter_number = o->ifirst - 1;
  for (int d = 0; d < 3; d++) c->simlen[d] = o->domain_max[d] - o->domain_min[d];
  c->mwall = 6;
  c->walls = xcalloc(c->mwall + 1, sizeof(plane_wall));
  *out = c;
  return SDEM_OK;
}

void oracle_destroy(oracle_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->pprop, c->wprop, c->pp_tab, c->pw_tab, c->walls, c->id, c->ptype, c->mark,
                  c->pos, c->lvel, c->lacc, c->rvel, c->racc, c->theta, c->force, c->torque,
                  c->fix_id, c->fix_type, c->fix_pos, c->g_id, c->g_type, c->g_pos, c->g_lvel,
                  c->g_rvel, c->bucket, c->id_j, c->next, c->status, c->tandel, c->vrel, c->kind,
                  c->bucket_pw, c->id_wall, c->next_pw, c->bx, c->by, c->bz, c->nextx, c->nexty,
                  c->nextz, c->heady, c->headx, c->headx0, c->headz, c->headz0, c->type_level};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
  free(c);
}

const char *oracle_last_error(const oracle_ctx *c) { return c ? c->err : "null context"; }

/* ------------------------------------------------------------------------ */
/* clc_BnryPrtcl_Prop, Prtcl_Property.f90:200-247                            */
/* ------------------------------------------------------------------------ */
static bnry_prop clc_bnry(const oracle_ctx *c, const pure_prop *pi, const pure_prop *pj,
                          double fs, double fk, double froll, double en, double et, int iswall) {
  bnry_prop b;
  memset(&b, 0, sizeof(b));
  if (!iswall) {
    b.rad_eff = (pi->radius * pj->radius) / (pi->radius + pj->radius);
    b.mass_eff = (pi->mass * pj->mass) / (pi->mass + pj->mass);
  } else {
    b.rad_eff = pi->radius;
    b.mass_eff = pi->mass;
  }
  b.rest_n = en; b.rest_t = et; b.fric_s = fs; b.fric_k = fk; b.fric_roll = froll;
  double pri = pi->poisson, prj = pj->poisson;
  double gi = pi->youngs / (2.0 * (1.0 + pri));
  double gj = pj->youngs / (2.0 * (1.0 + prj));
  double e_eff = 1.0 / ((1.0 - pri * pri) / pi->youngs + (1.0 - prj * prj) / pj->youngs);
  double g_eff = 1.0 / ((2.0 - pri) / gi + (2.0 - prj) / gj);
  double kappa = ((1.0 - pri) / gi + (1.0 - prj) / gj) / ((1.0 - 0.5 * pri) / gi + (1.0 - 0.5 * prj) / gj);
  kappa = fabs(kappa);
  double eta = log(en);
  eta = eta * eta;
  const double PI = 3.141592653589793238462643383279502884;
  if (c->o.cf_type == SDEM_CFT_LSD) {
    b.kn = 1.2024 * pow(sqrt(b.mass_eff) * (e_eff * e_eff) * b.rad_eff, 0.4);
    b.kt = kappa * b.kn;
    b.dn = -2.0 * log(en) * sqrt(b.mass_eff * b.kn) / sqrt(PI * PI + eta);
    b.dt = 0.0;
  } else if (c->o.cf_type == SDEM_CFT_LINEAR) {
    /* no counterpart in the reference: constant kn, kt = ratio*kn, and the
     * damping of the LSD formula at unit velocity factor */
    b.kn = c->lin_kn;
    b.kt = c->lin_kt_ratio * b.kn;
    b.dn = -2.0 * log(en) * sqrt(b.mass_eff * b.kn) / sqrt(PI * PI + eta);
    b.dt = 0.0;
  } else {
    double kh = 4.0 / 3.0 * e_eff * sqrt(b.rad_eff);
    b.kn = kh;
    b.kt = 16.0 / 3.0 * g_eff * sqrt(b.rad_eff);
    b.dn = -2.2664 * log(en) * sqrt(b.mass_eff * kh) / sqrt(eta + 10.1354);
    b.dt = 0.0;
  }
  return b;
}

/* Constant-stiffness linear model: no counterpart in the reference. */
int oracle_set_linear_stiffness(oracle_ctx *c, double kn, double kt_over_kn) {
  if (!c) return SDEM_E_ARG;
  if (c->o.cf_type != SDEM_CFT_LINEAR) return fail(c, SDEM_E_STATE, "cf_type is not the linear model");
  if (!(kn > 0.0) || !(kt_over_kn >= 0.0)) return fail(c, SDEM_E_ARG, "bad linear stiffness");
  if (c->have_pprop) return fail(c, SDEM_E_STATE, "stiffness must be set before the properties");
  c->lin_kn = kn;
  c->lin_kt_ratio = kt_over_kn;
  return SDEM_OK;
}

/* InitPrtclProperty :62-143 (property part; the bin counts are not needed) */
int oracle_set_particle_property(oracle_ctx *c, int ntype, const double *diam, const double *dens,
                                 const double *E, const double *nu, double fs, double fk,
                                 double froll, double en, double et) {
  if (ntype < 1) return fail(c, SDEM_E_ARG, "ntype < 1");
  if (c->o.cf_type == SDEM_CFT_LINEAR && !(c->lin_kn > 0.0))
    return fail(c, SDEM_E_STATE, "linear model: set the stiffness first");
  const double PI = 3.141592653589793238462643383279502884;
  c->ntype = ntype;
  free(c->pprop); free(c->pp_tab); free(c->type_level);
  c->pprop = xcalloc(ntype + 1, sizeof(pure_prop));
  c->pp_tab = xcalloc((size_t)ntype * ntype, sizeof(bnry_prop));
  c->type_level = xcalloc(ntype + 1, sizeof(int));
  for (int i = 1; i <= ntype; i++) {
    pure_prop *p = &c->pprop[i];
    p->density = dens[i - 1];
    p->radius = 0.5 * diam[i - 1];
    p->youngs = E[i - 1];
    p->poisson = nu[i - 1];
    /* four/three*Pi*(R)**3 ; R**3 is R*R*R */
    p->volume = 4.0 / 3.0 * PI * (p->radius * p->radius * p->radius);
    p->mass = dens[i - 1] * p->volume;
    p->inertia = 2.0 / 5.0 * p->mass * (p->radius * p->radius);
  }
  for (int i = 1; i <= ntype; i++)
    for (int j = 1; j <= ntype; j++) {
      /* pari = parj = Prtcl_PureProp(i)  (:137-138); the linear model, which
       * the reference does not have, takes the true pair */
      const pure_prop *pj = c->o.cf_type == SDEM_CFT_LINEAR ? &c->pprop[j] : &c->pprop[i];
      c->pp_tab[(i - 1) * ntype + (j - 1)] = clc_bnry(c, &c->pprop[i], pj, fs, fk, froll, en, et, 0);
    }
  c->have_pprop = 1;
  return SDEM_OK;
}

/* InitWallProperty :148-195 */
int oracle_set_wall_property(oracle_ctx *c, int nwtype, const double *E, const double *nu,
                             double fs, double fk, double froll, double en, double et) {
  if (!c->have_pprop) return fail(c, SDEM_E_STATE, "particle property first");
  if (nwtype < 1) return fail(c, SDEM_E_ARG, "nwtype < 1");
  c->nwtype = nwtype;
  free(c->wprop); free(c->pw_tab);
  c->wprop = xcalloc(nwtype + 1, sizeof(pure_prop));
  c->pw_tab = xcalloc((size_t)c->ntype * nwtype, sizeof(bnry_prop));
  for (int i = 1; i <= nwtype; i++) {
    pure_prop *w = &c->wprop[i];
    w->density = 1.0e50; w->radius = 1.0e50;
    w->youngs = E[i - 1]; w->poisson = nu[i - 1];
    w->mass = 1.0e50; w->volume = 1.0e50; w->inertia = 1.0e50;
  }
  for (int i = 1; i <= c->ntype; i++)
    for (int j = 1; j <= nwtype; j++)
      c->pw_tab[(i - 1) * nwtype + (j - 1)] =
          clc_bnry(c, &c->pprop[i], &c->wprop[j], fs, fk, froll, en, et, 1);
  c->have_wprop = 1;
  return SDEM_OK;
}

/* ------------------------------------------------------------------------ */
/* geometry, Prtcl_Geometry.f90                                              */
/* ------------------------------------------------------------------------ */
/* PW_IsInPlane :185-221 */
static int pw_is_in_plane(r3 p1, r3 p2, r3 p3, r3 p4, r3 cp) {
  r3 p1p = sub3(p1, cp), p2p = sub3(p2, cp), p3p = sub3(p3, cp), p4p = sub3(p4, cp);
  double p1p3 = dot3(p1p, p3p), p2p4 = dot3(p2p, p4p), p1p4 = dot3(p1p, p4p), p2p3 = dot3(p2p, p3p);
  if (p1p3 * p2p4 - p1p4 * p2p3 < 0.0) return 0;
  double p1p2 = dot3(p1p, p2p), p3p4 = dot3(p3p, p4p);
  if (p1p3 * p2p4 - p1p2 * p3p4 < 0.0) return 0;
  double p2p2 = dot3(p2p, p2p);
  if (p1p2 * p2p3 - p1p3 * p2p2 < 0.0) return 0;
  return 1;
}

/* Line_point_check :226-261 (the lp2 branch keeps the reference's
 * condition t>one .and. t>=(one+r/length)) */
static int line_point_check(r3 lp1, r3 lp2, r4 dpos, r3 *nv, double *ovrlp) {
  r3 pos = xyz(dpos);
  r3 w = sub3(pos, lp1);
  r3 v = sub3(lp2, lp1);
  double length = norm3(sub3(lp1, lp2));
  double r = dpos.w;
  double t = dot3(w, v) / dot3(v, v);
  r3 cp;
  if (t >= 0.0 && t <= 1.0) {
    cp = add3(mul3(v, t), lp1);
  } else if (t >= (-r / length) && t < 0.0) {
    cp = lp1;
  } else if (t > 1.0 && t >= (1.0 + r / length)) {
    cp = lp2;
  } else {
    return 0;
  }
  double ov = r - norm3(sub3(cp, pos));
  if (ov >= 0.0) {
    *ovrlp = ov;
    *nv = nv3(pos, cp);
    return 1;
  }
  return 0;
}

/* PW_isInContact :55-109 */
static int pw_is_in_contact(const plane_wall *w, r4 box, double *ovrlp, r3 *nv) {
  double radius = box.w;
  r3 p = xyz(box);
  double dist = dot3(w->n, p) + w->d;
  double ov;
  if (w->both_side) {
    ov = radius - fabs(dist);
  } else {
    if (dist < 0.0) return 0;
    ov = radius - dist;
  }
  if (ov < 0.0) return 0;
  if (w->is_infinite) {
    *ovrlp = ov;
    *nv = mul3(w->n, copysign(1.0, dist));
    return 1;
  }
  double t = -dist;
  r3 cp = add3(mul3(w->n, t), p);
  if (pw_is_in_plane(w->p1, w->p2, w->p3, w->p4, cp)) {
    *ovrlp = ov;
    *nv = mul3(w->n, copysign(1.0, dist));
    return 1;
  }
  if (line_point_check(w->p1, w->p2, box, nv, ovrlp)) return 1;
  if (line_point_check(w->p2, w->p3, box, nv, ovrlp)) return 1;
  if (line_point_check(w->p3, w->p4, box, nv, ovrlp)) return 1;
  if (line_point_check(w->p4, w->p1, box, nv, ovrlp)) return 1;
  return 0;
}

/* PointIsInDomain :266-280 */
static int point_in_domain(r3 p, r3 mn, r3 mx) {
  const double e = 1.0e-10;
  if (p.x < mn.x - e || p.x > mx.x + e) return 0;
  if (p.y < mn.y - e || p.y > mx.y + e) return 0;
  if (p.z < mn.z - e || p.z > mx.z + e) return 0;
  return 1;
}

/* IsLineInsertInner :331-421 */
static int line_insert_inner(r3 l1p1, r3 l1p2, r3 l2p1, r3 l2p2) {
  r3 v1 = sub3(l1p2, l1p1), v2 = sub3(l2p2, l2p1), l21 = sub3(l2p1, l1p1);
  r3 v1xv2 = cross3(v1, v2), l21xv1 = cross3(l21, v1), l21xv2 = cross3(l21, v2);
  double cc = norm3(l21xv1), dd = norm3(l21xv2), ee = norm3(v1xv2);
  if (ee < 1.0e-10) {
    double t1 = dot3(l21, v1) / dot3(v1, v1);
    r3 cp = add3(mul3(v1, t1), l1p1);
    double dist = norm3(sub3(cp, l2p1));
    if (dist > 1.0e-6) return 0;
    double t2 = dot3(sub3(l2p2, l1p1), v1) / dot3(v1, v1);
    if ((t1 < 0.0 && t2 < 0.0) || (t1 > 1.0 && t2 > 1.0)) return 0;
    return 1;
  }
  double a = (dot3(l21xv2, v1xv2) > 0.0) ? dd / ee : -dd / ee;
  double b = (dot3(l21xv1, v1xv2) > 0.0) ? cc / ee : -cc / ee;
  const double e = 1.0e-10;
  return (a >= 0.0 - e && a <= 1.0 + e && b >= 0.0 - e && b <= 1.0 + e);
}

/* IsReflectionOvlp :285-326 */
static int reflection_ovlp(r3 f1, r3 f2, r3 f3, r3 f4, r3 s1, r3 s2, r3 s3, r3 s4) {
  if (pw_is_in_plane(f1, f2, f3, f4, s1) || pw_is_in_plane(f1, f2, f3, f4, s2) ||
      pw_is_in_plane(f1, f2, f3, f4, s3) || pw_is_in_plane(f1, f2, f3, f4, s4) ||
      pw_is_in_plane(s1, s2, s3, s4, f1) || pw_is_in_plane(s1, s2, s3, s4, f2) ||
      pw_is_in_plane(s1, s2, s3, s4, f3) || pw_is_in_plane(s1, s2, s3, s4, f4))
    return 1;
  r3 F[5] = {ZERO3, f1, f2, f3, f4}, S[5] = {ZERO3, s1, s2, s3, s4};
  for (int a = 1; a <= 4; a++) {
    int a2 = a % 4 + 1;
    for (int b = 1; b <= 4; b++) {
      int b2 = b % 4 + 1;
      if (line_insert_inner(F[a], F[a2], S[b], S[b2])) return 1;
    }
  }
  return 0;
}

/* PW_IsInDomain :114-180 */
static int pw_is_in_domain(const plane_wall *w, r3 mn, r3 mx) {
  r3 dom[9];
  dom[1] = R3(mn.x, mn.y, mn.z); dom[2] = R3(mx.x, mn.y, mn.z);
  dom[3] = R3(mx.x, mx.y, mn.z); dom[4] = R3(mn.x, mx.y, mn.z);
  dom[5] = R3(mn.x, mn.y, mx.z); dom[6] = R3(mx.x, mn.y, mx.z);
  dom[7] = R3(mx.x, mx.y, mx.z); dom[8] = R3(mn.x, mx.y, mx.z);
  if (point_in_domain(w->p1, mn, mx) || point_in_domain(w->p2, mn, mx) ||
      point_in_domain(w->p3, mn, mx) || point_in_domain(w->p4, mn, mx))
    return 1;
  /* x-y reflection */
  if (!reflection_ovlp(R3(w->p1.x, w->p1.y, 0), R3(w->p2.x, w->p2.y, 0), R3(w->p3.x, w->p3.y, 0),
                       R3(w->p4.x, w->p4.y, 0), R3(mn.x, mn.y, 0), R3(mx.x, mn.y, 0),
                       R3(mx.x, mx.y, 0), R3(mn.x, mx.y, 0)))
    return 0;
  /* x-z reflection */
  if (!reflection_ovlp(R3(w->p1.x, 0, w->p1.z), R3(w->p2.x, 0, w->p2.z), R3(w->p3.x, 0, w->p3.z),
                       R3(w->p4.x, 0, w->p4.z), R3(mn.x, 0, mn.z), R3(mn.x, 0, mx.z),
                       R3(mx.x, 0, mx.z), R3(mx.x, 0, mn.z)))
    return 0;
  /* y-z reflection */
  if (!reflection_ovlp(R3(0, w->p1.y, w->p1.z), R3(0, w->p2.y, w->p2.z), R3(0, w->p3.y, w->p3.z),
                       R3(0, w->p4.y, w->p4.z), R3(0, mn.y, mn.z), R3(0, mx.y, mn.z),
                       R3(0, mx.y, mx.z), R3(0, mn.y, mx.z)))
    return 0;
  int sum = 0;
  for (int i = 1; i <= 8; i++) {
    double t = dot3(w->n, dom[i]) + w->d;
    int v = -99;
    if (t > 1.0e-8) v = 1;
    else if (t < -1.0e-8) v = 0;
    sum += v;
  }
  if (sum == 0 || sum == 8) return 0;
  return 1;
}

static double max_radius(const oracle_ctx *c) {
  double m = c->pprop[1].radius;
  for (int i = 2; i <= c->ntype; i++) m = fmax_f(m, c->pprop[i].radius);
  return m;
}
static double min_radius(const oracle_ctx *c) {
  double m = c->pprop[1].radius;
  for (int i = 2; i <= c->ntype; i++) m = fmin_f(m, c->pprop[i].radius);
  return m;
}

/* G_add_PlaneWall :457-537 with G_InitAllocate's expanded domain :435-439 */
int oracle_add_plane_wall(oracle_ctx *c, const double *p1, const double *p2, const double *p3,
                          const double *p4, int user_id, int wall_type, int both, int infinite,
                          const double *tvel, int *wall_id) {
  if (!c->have_pprop) return fail(c, SDEM_E_STATE, "particle property first");
  plane_wall w;
  memset(&w, 0, sizeof(w));
  w.p1 = R3(p1[0], p1[1], p1[2]); w.p2 = R3(p2[0], p2[1], p2[2]);
  w.p3 = R3(p3[0], p3[1], p3[2]); w.p4 = R3(p4[0], p4[1], p4[2]);
  r3 ln = cross3(sub3(w.p2, w.p1), sub3(w.p3, w.p1));
  w.min_point = min3(w.p4, min3(w.p3, min3(w.p2, w.p1)));
  w.max_point = max3(w.p4, max3(w.p3, max3(w.p2, w.p1)));
  w.user_id = user_id; w.wall_type = wall_type; w.both_side = both ? 1 : 0;
  w.is_infinite = infinite ? 1 : 0;
  w.n = div3(ln, norm3(ln));
  w.d = -(dot3(w.n, w.p1));
  w.trans_vel = tvel ? R3(tvel[0], tvel[1], tvel[2]) : ZERO3;
  if (fabs(dot3(w.n, w.p4) + w.d) >= 0.00001) return fail(c, SDEM_E_ARG, "Cannot create a plane wall");
  double lenexp = 1.2 * max_radius(c);
  r3 dmin = R3(c->o.domain_min[0], c->o.domain_min[1], c->o.domain_min[2]);
  r3 dmax = R3(c->o.domain_max[0], c->o.domain_max[1], c->o.domain_max[2]);
  r3 pmin = sub3(dmin, mul3(R3(1, 1, 1), lenexp));
  r3 pmax = add3(dmax, mul3(R3(1, 1, 1), lenexp));
  if (wall_id) *wall_id = 0;
  if (!pw_is_in_domain(&w, pmin, pmax)) return SDEM_OK;
  c->num_pwall_total++;
  w.wall_id = c->num_pwall_total + c->o.base_wall_id;
  if (c->nwall == c->mwall) {
    c->mwall = (int)(c->mwall * 1.2) + 1;
    c->walls = realloc(c->walls, sizeof(plane_wall) * (c->mwall + 1));
  }
  c->walls[++c->nwall] = w;
  if (wall_id) *wall_id = w.wall_id;
  return SDEM_OK;
}

/* ------------------------------------------------------------------------ */
/* fixed particles and their periodic images (Comm_For_Cntct_fixed, built   */
/* once, with the same image rule as the moving ghosts)                     */
/* ------------------------------------------------------------------------ */
static void fix_push(oracle_ctx *c, int *cap, int id, int type, r4 p) {
  if (c->nfix + 1 > *cap) {
    *cap = 2 * (*cap) + 8;
    c->fix_id = realloc(c->fix_id, sizeof(int) * (*cap + 1));
    c->fix_type = realloc(c->fix_type, sizeof(int) * (*cap + 1));
    c->fix_pos = realloc(c->fix_pos, sizeof(r4) * (*cap + 1));
  }
  int k = ++c->nfix;
  c->fix_id[k] = id; c->fix_type[k] = type; c->fix_pos[k] = p;
}

static inline double comp(r4 p, int d);
static inline void setcomp(r4 *p, int d, double v);

int oracle_set_fixed_particles(oracle_ctx *c, int nfix, const int *id, const int *ptype,
                               const double *posr) {
  if (!c->have_pprop) return fail(c, SDEM_E_STATE, "particle property first");
  free(c->fix_id); free(c->fix_type); free(c->fix_pos);
  int cap = nfix + 8;
  c->fix_id = xcalloc(cap + 1, sizeof(int));
  c->fix_type = xcalloc(cap + 1, sizeof(int));
  c->fix_pos = xcalloc(cap + 1, sizeof(r4));
  c->nfix = 0;
  for (int i = 1; i <= nfix; i++) {
    r4 p = {posr[4 * (i - 1)], posr[4 * (i - 1) + 1], posr[4 * (i - 1) + 2], posr[4 * (i - 1) + 3]};
    fix_push(c, &cap, id[i - 1], ptype[i - 1], p);
  }
  c->nfix_local = nfix;
  const double L = c->o.prtcl_cs_ratio * 2.0 * max_radius(c);
  if (c->o.is_periodic[2]) {
    double st2 = c->o.domain_min[2] + L, ed0 = c->o.domain_max[2] - L;
    for (int i = 1; i <= nfix; i++) {
      double pz = c->fix_pos[i].z;
      if (pz <= st2) { r4 p = c->fix_pos[i]; p.z = pz + c->simlen[2]; fix_push(c, &cap, c->fix_id[i], c->fix_type[i], p); }
      if (pz >= ed0) { r4 p = c->fix_pos[i]; p.z = pz - c->simlen[2]; fix_push(c, &cap, c->fix_id[i], c->fix_type[i], p); }
    }
  }
  for (int k = 0; k < 2; k++) {
    int d = k; /* x then y */
    if (!c->o.is_periodic[d]) continue;
    double st1 = c->o.domain_min[d], st2 = st1 + L, ed0 = c->o.domain_max[d] - L;
    int ng0 = c->nfix;
    /* "+" side: previous images, then locals */
    for (int i = nfix + 1; i <= ng0; i++) {
      double pd = comp(c->fix_pos[i], d);
      if (pd >= ed0) { r4 p = c->fix_pos[i]; setcomp(&p, d, pd - c->simlen[d]); fix_push(c, &cap, c->fix_id[i], c->fix_type[i], p); }
    }
    for (int i = 1; i <= nfix; i++) {
      double pd = comp(c->fix_pos[i], d);
      if (pd >= ed0) { r4 p = c->fix_pos[i]; setcomp(&p, d, pd - c->simlen[d]); fix_push(c, &cap, c->fix_id[i], c->fix_type[i], p); }
    }
    /* "-" side */
    for (int i = nfix + 1; i <= ng0; i++) {
      double pd = comp(c->fix_pos[i], d);
      if (pd <= st2 && pd >= st1) { r4 p = c->fix_pos[i]; setcomp(&p, d, pd + c->simlen[d]); fix_push(c, &cap, c->fix_id[i], c->fix_type[i], p); }
    }
    for (int i = 1; i <= nfix; i++) {
      double pd = comp(c->fix_pos[i], d);
      if (pd <= st2 && pd >= st1) { r4 p = c->fix_pos[i]; setcomp(&p, d, pd + c->simlen[d]); fix_push(c, &cap, c->fix_id[i], c->fix_type[i], p); }
    }
  }
  return SDEM_OK;
}

/* ------------------------------------------------------------------------ */
/* state                                                                     */
/* ------------------------------------------------------------------------ */
int oracle_set_state(oracle_ctx *c, int n, const int *id, const int *ptype, const int *mark,
                     const double *posr, const double *lvel, const double *lacc, const double *theta,
                     const double *rvel, const double *racc) {
  if (n < 0 || n > c->mlocal) return fail(c, SDEM_E_CAPACITY, "n exceeds numPrtcl");
  if (!c->have_pprop) return fail(c, SDEM_E_STATE, "particle property first");
  int ts = c->tsize, rs = c->rsize;
  for (int i = 1; i <= n; i++) {
    int k = i - 1;
    c->id[i] = id[k];
    c->ptype[i] = ptype[k];
    if (ptype[k] < 1 || ptype[k] > c->ntype) return fail(c, SDEM_E_ARG, "ptype out of range");
    c->mark[i] = mark ? mark[k] : 1;
    r4 p = {posr[4 * k], posr[4 * k + 1], posr[4 * k + 2], posr[4 * k + 3]};
    c->pos[i] = p;
    for (int l = 0; l < ts; l++) {
      const double *v = lvel + ((size_t)k * ts + l) * 3;
      const double *a = lacc + ((size_t)k * ts + l) * 3;
      c->lvel[(size_t)k * ts + l] = R3(v[0], v[1], v[2]);
      c->lacc[(size_t)k * ts + l] = R3(a[0], a[1], a[2]);
    }
    c->theta[i] = R3(theta[3 * k], theta[3 * k + 1], theta[3 * k + 2]);
    for (int l = 0; l < rs; l++) {
      const double *v = rvel + ((size_t)k * rs + l) * 3;
      const double *a = racc + ((size_t)k * rs + l) * 3;
      c->rvel[(size_t)k * rs + l] = R3(v[0], v[1], v[2]);
      c->racc[(size_t)k * rs + l] = R3(a[0], a[1], a[2]);
    }
    c->force[i] = ZERO3;
    c->torque[i] = ZERO3;
  }
  c->nlocal = n;
  c->have_state = 1;
  return SDEM_OK;
}

#define LV(c, i, l) ((c)->lvel[(size_t)((i)-1) * (c)->tsize + ((l)-1)])
#define LA(c, i, l) ((c)->lacc[(size_t)((i)-1) * (c)->tsize + ((l)-1)])
#define RV(c, i, l) ((c)->rvel[(size_t)((i)-1) * (c)->rsize + ((l)-1)])
#define RA(c, i, l) ((c)->racc[(size_t)((i)-1) * (c)->rsize + ((l)-1)])

/* (id, local index) pairs sorted by id, so that a record's owner is found
   by bisection: the ids are unique, and a scan per record is quadratic in
   the size of a settled pack (millions of records) */
typedef struct {
  int id, idx;
} id_index;

static int cmp_id_index(const void *a, const void *b) {
  const int x = ((const id_index *)a)->id, y = ((const id_index *)b)->id;
  return (x > y) - (x < y);
}

static int local_index_of(const id_index *t, int n, int id) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int mid = lo + (hi - lo) / 2;
    if (t[mid].id == id) return t[mid].idx;
    if (t[mid].id < id) lo = mid + 1;
    else hi = mid - 1;
  }
  return 0;
}

/* Add_RestartCntctlink-style install of history records (status 2) */
int oracle_set_contacts(oracle_ctx *c, int64_t m, const int *kind, const int *owner,
                        const int *partner, const double *td, const double *vr) {
  id_index *t = (id_index *)malloc(sizeof(id_index) * (size_t)(c->nlocal > 0 ? c->nlocal : 1));
  if (!t) return fail(c, SDEM_E_CAPACITY, "out of host memory");
  for (int i = 1; i <= c->nlocal; i++) {
    t[i - 1].id = c->id[i];
    t[i - 1].idx = i;
  }
  qsort(t, (size_t)c->nlocal, sizeof(id_index), cmp_id_index);
  for (int64_t k = 0; k < m; k++) {
    int pid = local_index_of(t, c->nlocal, owner[k]);
    if (pid == 0) {
      free(t);
      return fail(c, SDEM_E_ARG, "owner id not present");
    }
    int ni = c->next_insert;
    if (ni == 0) {
      free(t);
      return fail(c, SDEM_E_CAPACITY, "contact container is full");
    }
    c->next_insert = -c->next[ni];
    c->id_j[ni] = partner[k];
    c->next[ni] = c->bucket[pid];
    c->bucket[pid] = ni;
    c->status[ni] = 2;
    c->kind[ni] = kind[k];
    c->tandel[ni] = R3(td[3 * k], td[3 * k + 1], td[3 * k + 2]);
    c->vrel[ni] = vr[k];
  }
  free(t);
  return SDEM_OK;
}

/* ------------------------------------------------------------------------ */
/* contact list, Prtcl_CL_and_CF.f90                                         */
/* ------------------------------------------------------------------------ */
/* Find_insert :256-296 */
static int find_insert(oracle_ctx *c, int bkt, int value, int *slot) {
  if (bkt > c->mlocal) { *slot = -1; return -1; }
  int n = c->bucket[bkt];
  while (n > 0) {
    if (c->id_j[n] == value) { *slot = n; return 1; }
    n = c->next[n];
  }
  if (c->next_insert == 0) { *slot = -2; return -2; }
  int ni = c->next_insert;
  c->next_insert = -c->next[ni];
  c->id_j[ni] = value;
  c->next[ni] = c->bucket[bkt];
  c->bucket[bkt] = ni;
  c->tandel[ni] = ZERO3;
  c->vrel[ni] = -2.0;
  *slot = ni;
  return 2;
}

/* CL_PreIteration :560-572 */
static void cl_pre_iteration(oracle_ctx *c) {
  memset(c->num_cntcts, 0, sizeof(c->num_cntcts));
  for (int i = 1; i <= c->max_cntcts; i++)
    if (c->status[i] > 0) c->status[i] = -2;
}

/* CL_RemvReleased :500-555 (single rank: no status-3 counterpart links) */
static void cl_remv_released(oracle_ctx *c) {
  for (int pid = 1; pid <= c->nlocal; pid++) {
    int prev = 0, n = c->bucket[pid];
    while (n > 0) {
      int nexti = c->next[n];
      if (c->status[n] == -2) {
        if (prev == 0) c->bucket[pid] = nexti;
        else c->next[prev] = nexti;
        c->status[n] = -1;
        c->next[n] = -c->next_insert;
        c->next_insert = n;
      } else {
        prev = n;
      }
      n = nexti;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Prtcl_ContactForce_inc.f90:1-66 -- the common part of all four call sites */
/* ------------------------------------------------------------------------ */
typedef struct { r3 fsum, moment, mrij; } cf_out;

static cf_out contact_force_inc(oracle_ctx *c, const bnry_prop *P, r3 norm_v, double ri, double rj,
                                r3 veli, r3 velj, r3 rvei, r3 rvej, double ovrlp, int ind) {
  const double dt = c->o.dt;
  r3 vel_w = cross3(add3(mul3(rvei, ri), mul3(rvej, rj)), norm_v);
  r3 vrij = add3(sub3(veli, velj), vel_w);
  double vrn = dot3(vrij, norm_v);
  if (c->vrel[ind] < -1.0) c->vrel[ind] = fabs(vrn);
  r3 vij_n = mul3(norm_v, vrn);
  r3 vij_t = sub3(vrij, vij_n);

  r3 ovlp_t = c->tandel[ind];
  double n1 = norm3(ovlp_t);
  ovlp_t = sub3(ovlp_t, mul3(norm_v, dot3(ovlp_t, norm_v)));
  double n2 = norm3(ovlp_t);
  if (n2 > 1.0e-10) ovlp_t = mul3(ovlp_t, n1 / n2);
  else ovlp_t = ZERO3;
  ovlp_t = add3(mul3(vij_t, dt), ovlp_t);

  double fn, k_t, d_t;
  r3 ftij;
  if (c->o.cf_type == SDEM_CFT_LINEAR) {
    fn = -P->kn * ovrlp + -P->dn * vrn;
    k_t = -P->kt;
    d_t = 0.0;
    ftij = add3(mul3(ovlp_t, k_t), mul3(vij_t, d_t));
  } else if (c->o.cf_type == SDEM_CFT_LSD) {
    double vel_in = fmax_f(c->vrel[ind], 0.01);
    double k_n = -(P->kn * pow(vel_in, 0.4));
    double d_n = -(P->dn * pow(vel_in, 0.2));
    fn = k_n * ovrlp + d_n * vrn;
    k_t = -(P->kt * pow(vel_in, 0.4));
    d_t = 0.0;
    ftij = add3(mul3(ovlp_t, k_t), mul3(vij_t, d_t));
  } else {
    double k_n = -P->kn;
    double d_n = -P->dn;
    fn = k_n * pow(ovrlp, 1.5) + d_n * pow(ovrlp, 0.25) * vrn;
    k_t = -(P->kt * sqrt(ovrlp));
    d_t = 0.0;
    ftij = add3(mul3(ovlp_t, k_t), mul3(vij_t, d_t));
  }
  r3 fnij = mul3(norm_v, fn);

  double ft = norm3(ftij);
  double ft_fric = P->fric_s * fabs(fn);
  if (ft > ft_fric) {
    ft_fric = P->fric_k * fabs(fn);
    ftij = mul3(ftij, ft_fric / ft);
    ovlp_t = div3(ftij, k_t);
  }

  r3 w_hat = sub3(rvei, rvej);
  double wm = norm3(w_hat);
  if (wm > 1.0e-10) w_hat = div3(w_hat, wm);
  else w_hat = ZERO3;
  r3 mrij;
  if (c->o.ct_model == SDEM_CTM_CONSTANT_TORQUE)
    mrij = mul3(w_hat, -(P->fric_roll * fabs(fn) * P->rad_eff));
  else
    mrij = mul3(w_hat, -(P->fric_roll * fabs(fn) * P->rad_eff * norm3(vel_w)));

  c->tandel[ind] = ovlp_t;
  cf_out o;
  o.moment = cross3(norm_v, ftij);
  o.fsum = add3(fnij, ftij);
  o.mrij = mrij;
  return o;
}

/* ContactForce_PP :732-756 */
static void contact_force_pp(oracle_ctx *c, int pid, int pjd, int ind, double ovrlp) {
  const bnry_prop *P = &c->pp_tab[(c->ptype[pid] - 1) * c->ntype + (c->ptype[pjd] - 1)];
  r4 posi = c->pos[pid], posj = c->pos[pjd];
  double ri = posi.w, rj = posj.w;
  r3 norm_v = nv3(xyz(posj), xyz(posi));
  cf_out o = contact_force_inc(c, P, norm_v, ri, rj, LV(c, pid, 1), LV(c, pjd, 1), RV(c, pid, 1),
                               RV(c, pjd, 1), ovrlp, ind);
  c->force[pid] = add3(c->force[pid], o.fsum);
  c->torque[pid] = add3(add3(c->torque[pid], mul3(o.moment, ri)), o.mrij);
  c->force[pjd] = sub3(c->force[pjd], o.fsum);
  c->torque[pjd] = sub3(add3(c->torque[pjd], mul3(o.moment, rj)), o.mrij);
}

/* ContactForce_PPG :761-798 */
static void contact_force_ppg(oracle_ctx *c, int pid, int g, int ind, double ovrlp) {
  const bnry_prop *P = &c->pp_tab[(c->ptype[pid] - 1) * c->ntype + (c->g_type[g] - 1)];
  if (c->id[pid] < c->g_id[g]) {
    r4 posi = c->pos[pid], posj = c->g_pos[g];
    double ri = posi.w, rj = posj.w;
    r3 norm_v = nv3(xyz(posj), xyz(posi));
    cf_out o = contact_force_inc(c, P, norm_v, ri, rj, LV(c, pid, 1), c->g_lvel[g], RV(c, pid, 1),
                                 c->g_rvel[g], ovrlp, ind);
    c->force[pid] = add3(c->force[pid], o.fsum);
    c->torque[pid] = add3(add3(c->torque[pid], mul3(o.moment, ri)), o.mrij);
  } else {
    r4 posi = c->g_pos[g], posj = c->pos[pid];
    double ri = posi.w, rj = posj.w;
    r3 norm_v = nv3(xyz(posj), xyz(posi));
    cf_out o = contact_force_inc(c, P, norm_v, ri, rj, c->g_lvel[g], LV(c, pid, 1), c->g_rvel[g],
                                 RV(c, pid, 1), ovrlp, ind);
    c->force[pid] = sub3(c->force[pid], o.fsum);
    c->torque[pid] = sub3(add3(c->torque[pid], mul3(o.moment, rj)), o.mrij);
  }
}

/* ContactForce_PPFix :803-827 */
static void contact_force_ppfix(oracle_ctx *c, int pid, int f, int ind, double ovrlp) {
  const bnry_prop *P = &c->pp_tab[(c->ptype[pid] - 1) * c->ntype + (c->fix_type[f] - 1)];
  r4 posi = c->pos[pid], posj = c->fix_pos[f];
  double ri = posi.w, rj = posj.w;
  r3 norm_v = nv3(xyz(posj), xyz(posi));
  cf_out o = contact_force_inc(c, P, norm_v, ri, rj, LV(c, pid, 1), ZERO3, RV(c, pid, 1), ZERO3,
                               ovrlp, ind);
  c->force[pid] = add3(c->force[pid], o.fsum);
  c->torque[pid] = add3(add3(c->torque[pid], mul3(o.moment, ri)), o.mrij);
}

/* ContactForce_PW :832-856 */
static void contact_force_pw(oracle_ctx *c, int pid, int mwi, int ind, double ovrlp, r3 nv) {
  const plane_wall *w = &c->walls[mwi];
  const bnry_prop *P = &c->pw_tab[(c->ptype[pid] - 1) * c->nwtype + (w->wall_type - 1)];
  double ri = c->pos[pid].w;
  double rj = 1.00e20 * ri;
  r3 norm_v = mul3(nv, -1.0);
  cf_out o = contact_force_inc(c, P, norm_v, ri, rj, LV(c, pid, 1), w->trans_vel, RV(c, pid, 1),
                               ZERO3, ovrlp, ind);
  c->force[pid] = add3(c->force[pid], o.fsum);
  c->torque[pid] = add3(add3(c->torque[pid], mul3(o.moment, ri)), o.mrij);
}

static int add_contact(oracle_ctx *c, int kind, int bkt, int value, int *slot) {
  int st = find_insert(c, bkt, value, slot);
  if (st > 0) {
    c->status[*slot] = st;
    c->kind[*slot] = kind;
    c->num_cntcts[kind - 1]++;
    return 1;
  }
  if (st == -1) fail(c, SDEM_E_ARG, "The inserted item's id is greater than allowed value");
  else fail(c, SDEM_E_CAPACITY, "The container is full and there is no space for new item");
  return 0;
}

/* AddContactPP :142-168 */
static int add_contact_pp(oracle_ctx *c, int id1, int id2, double ov) {
  int slot;
  if (!add_contact(c, SDEM_KIND_PP, id1, c->id[id2], &slot)) return 0;
  contact_force_pp(c, id1, id2, slot, ov);
  return 1;
}
/* AddContactPPG :173-195 */
static int add_contact_ppg(oracle_ctx *c, int id1, int g, double ov) {
  int slot;
  if (!add_contact(c, SDEM_KIND_PPG, id1, c->g_id[g], &slot)) return 0;
  contact_force_ppg(c, id1, g, slot, ov);
  return 1;
}
/* AddContactPPFix :200-222 */
static int add_contact_ppfix(oracle_ctx *c, int id1, int f, double ov) {
  int slot;
  if (!add_contact(c, SDEM_KIND_PPFIX, id1, c->fix_id[f], &slot)) return 0;
  contact_force_ppfix(c, id1, f, slot, ov);
  return 1;
}
/* AddContactPW :227-249 */
static int add_contact_pw(oracle_ctx *c, int id1, int wi, double ov, r3 nv) {
  int slot;
  if (!add_contact(c, SDEM_KIND_PW, id1, c->walls[wi].wall_id, &slot)) return 0;
  contact_force_pw(c, id1, wi, slot, ov, nv);
  return 1;
}

/* ------------------------------------------------------------------------ */
/* ghosts: PC_Comm_For_Cntct for one rank owning the whole domain            */
/* (z-pencil order: z first, then x, then y; Prtcl_Comm.f90:878-1204)        */
/* ------------------------------------------------------------------------ */
static void ghost_push(oracle_ctx *c, int id, int type, r4 p, r3 lv, r3 rv) {
  if (c->nghost + 1 > c->mghost) {
    int m = c->mghost < 16 ? 64 : (int)(1.2 * c->mghost) + 1;
    c->g_id = realloc(c->g_id, sizeof(int) * (m + 1));
    c->g_type = realloc(c->g_type, sizeof(int) * (m + 1));
    c->g_pos = realloc(c->g_pos, sizeof(r4) * (m + 1));
    c->g_lvel = realloc(c->g_lvel, sizeof(r3) * (m + 1));
    c->g_rvel = realloc(c->g_rvel, sizeof(r3) * (m + 1));
    c->mghost = m;
  }
  int g = ++c->nghost;
  c->g_id[g] = id; c->g_type[g] = type; c->g_pos[g] = p; c->g_lvel[g] = lv; c->g_rvel[g] = rv;
}

static inline double comp(r4 p, int d) { return d == 0 ? p.x : (d == 1 ? p.y : p.z); }
static inline void setcomp(r4 *p, int d, double v) {
  if (d == 0) p->x = v; else if (d == 1) p->y = v; else p->z = v;
}

static void comm_for_cntct(oracle_ctx *c) {
  c->nghost = 0;
  double L = c->lenforcs;
  /* pencil direction (z): both sides from the locals, :880-905 */
  if (c->o.is_periodic[2]) {
    double st2 = c->o.domain_min[2] + L, ed0 = c->o.domain_max[2] - L;
    for (int i = 1; i <= c->nlocal; i++) {
      double pz = c->pos[i].z;
      if (pz <= st2) { r4 p = c->pos[i]; p.z = pz + c->simlen[2]; ghost_push(c, c->id[i], c->ptype[i], p, LV(c, i, 1), RV(c, i, 1)); }
      if (pz >= ed0) { r4 p = c->pos[i]; p.z = pz - c->simlen[2]; ghost_push(c, c->id[i], c->ptype[i], p, LV(c, i, 1), RV(c, i, 1)); }
    }
  }
  /* x then y: "+" side (p>=ed0 -> p-L) from previous ghosts then locals,
   * then "-" side (st1<=p<=st2 -> p+L) from previous ghosts then locals */
  int dd[2] = {0, 1};
  for (int k = 0; k < 2; k++) {
    int d = dd[k];
    if (!c->o.is_periodic[d]) continue;
    double st1 = c->o.domain_min[d], st2 = st1 + L, ed0 = c->o.domain_max[d] - L;
    int ngp = c->nghost;
    for (int i = 1; i <= ngp; i++) {
      double pd = comp(c->g_pos[i], d);
      if (pd >= ed0) { r4 p = c->g_pos[i]; setcomp(&p, d, pd - c->simlen[d]); ghost_push(c, c->g_id[i], c->g_type[i], p, c->g_lvel[i], c->g_rvel[i]); }
    }
    for (int i = 1; i <= c->nlocal; i++) {
      double pd = comp(c->pos[i], d);
      if (pd >= ed0) { r4 p = c->pos[i]; setcomp(&p, d, pd - c->simlen[d]); ghost_push(c, c->id[i], c->ptype[i], p, LV(c, i, 1), RV(c, i, 1)); }
    }
    for (int i = 1; i <= ngp; i++) {
      double pd = comp(c->g_pos[i], d);
      if (pd <= st2 && pd >= st1) { r4 p = c->g_pos[i]; setcomp(&p, d, pd + c->simlen[d]); ghost_push(c, c->g_id[i], c->g_type[i], p, c->g_lvel[i], c->g_rvel[i]); }
    }
    for (int i = 1; i <= c->nlocal; i++) {
      double pd = comp(c->pos[i], d);
      if (pd <= st2 && pd >= st1) { r4 p = c->pos[i]; setcomp(&p, d, pd + c->simlen[d]); ghost_push(c, c->id[i], c->ptype[i], p, LV(c, i, 1), RV(c, i, 1)); }
    }
  }
}

/* GL_copy :351-373 + CL_copy :449-458 + CSPW_copy :240-259 */
static void delete_particle(oracle_ctx *c, int i1, int i2) {
  /* CSPW_copy: release i1's near-wall nodes, move i2's list */
  int n = c->bucket_pw[i1];
  while (n > 0) {
    int nx = c->next_pw[n];
    c->next_pw[n] = -c->next_insert_pw;
    c->next_insert_pw = n;
    n = nx;
  }
  c->bucket_pw[i1] = c->bucket_pw[i2];
  c->bucket_pw[i2] = 0;
  /* CL_copy: the deleted particle's records are not released (they leak) */
  c->bucket[i1] = c->bucket[i2];
  c->bucket[i2] = 0;
  if (i1 == i2) return;
  c->id[i1] = c->id[i2]; c->ptype[i1] = c->ptype[i2]; c->mark[i1] = c->mark[i2];
  c->pos[i1] = c->pos[i2];
  for (int l = 1; l <= c->tsize; l++) { LV(c, i1, l) = LV(c, i2, l); LA(c, i1, l) = LA(c, i2, l); }
  c->theta[i1] = c->theta[i2];
  for (int l = 1; l <= c->rsize; l++) { RV(c, i1, l) = RV(c, i2, l); RA(c, i1, l) = RA(c, i2, l); }
  c->force[i1] = c->force[i2];
  c->torque[i1] = c->torque[i2];
}

/* PC_Comm_For_Exchange, one rank (z-pencil order: z, x, y) */
static void comm_for_exchange(oracle_ctx *c) {
  int order[3] = {2, 0, 1};
  for (int k = 0; k < 3; k++) {
    int d = order[k];
    double st = c->o.domain_min[d], ed = c->o.domain_max[d];
    if (c->o.is_periodic[d]) {
      for (int i = 1; i <= c->nlocal; i++) {
        double p = comp(c->pos[i], d);
        if (p >= ed) setcomp(&c->pos[i], d, p - c->simlen[d]);
        else if (p < st) setcomp(&c->pos[i], d, p + c->simlen[d]);
      }
    } else {
      int i = 1;
      while (i <= c->nlocal) {
        double p = comp(c->pos[i], d);
        if (p >= ed || p < st) {
          delete_particle(c, i, c->nlocal);
          c->nlocal--;
        } else {
          i++;
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* NBS-Munjiza, Prtcl_NBS_Munjiza.f90                                        */
/* ------------------------------------------------------------------------ */
static int nfixed_cs(const oracle_ctx *c) { return c->nfix; }

static void ensure_box(oracle_ctx *c, int n) {
  if (n + 2 <= c->mbox) return;
  int m = n + 16;
  c->bx = realloc(c->bx, sizeof(int) * m);
  c->by = realloc(c->by, sizeof(int) * m);
  c->bz = realloc(c->bz, sizeof(int) * m);
  c->nextx = realloc(c->nextx, sizeof(int) * m);
  c->nexty = realloc(c->nexty, sizeof(int) * m);
  c->nextz = realloc(c->nextz, sizeof(int) * m);
  c->mbox = m;
}

/* Init_NBSM :57-113 */
static void init_nbsm(oracle_ctx *c) {
  double maxd = 2.0 * max_radius(c);
  c->cell_len = c->o.prtcl_cs_ratio * maxd;
  c->xst_cs = c->o.domain_min[0] - c->cell_len * 1.05;
  c->yst_cs = c->o.domain_min[1] - c->cell_len * 1.05;
  c->zst_cs = c->o.domain_min[2] - c->cell_len * 1.05;
  double xed = c->o.domain_max[0] + c->cell_len * 1.05;
  double yed = c->o.domain_max[1] + c->cell_len * 1.05;
  double zed = c->o.domain_max[2] + c->cell_len * 1.05;
  c->nx = (int)((xed - c->xst_cs) / c->cell_len) + 1;
  c->ny = (int)((yed - c->yst_cs) / c->cell_len) + 1;
  c->nz = (int)((zed - c->zst_cs) / c->cell_len) + 1;
  free(c->heady); free(c->headx); free(c->headx0); free(c->headz); free(c->headz0);
  c->heady = xcalloc(c->ny + 2, sizeof(int));
  c->headx = xcalloc(c->nx + 2, sizeof(int));
  c->headx0 = xcalloc(c->nx + 2, sizeof(int));
  c->headz = xcalloc(2 * (c->nz + 2), sizeof(int));
  c->headz0 = xcalloc(3 * (c->nz + 2), sizeof(int));
}

#define HZ(c, m, iz) ((c)->headz[(m) * ((c)->nz + 2) + (iz)])
#define HZ0(c, m, iz) ((c)->headz0[(m) * ((c)->nz + 2) + (iz)])

static inline int floor_i(double v) { return (int)floor(v); }

/* FineSearch1 :478-498 (moving m vs fixed f) */
static int fine1(oracle_ctx *c, int pid1, int pid2) {
  int p = pid1 - nfixed_cs(c);
  r4 a = c->pos[p], b = c->fix_pos[pid2];
  double dr = a.w + b.w, dr2 = dr * dr;
  double dx = a.x - b.x, d2 = dx * dx;
  if (d2 > dr2) return 1;
  double dy = a.y - b.y; d2 = dy * dy + d2;
  if (d2 > dr2) return 1;
  double dz = a.z - b.z; d2 = dz * dz + d2;
  if (d2 > dr2) return 1;
  double ov = dr - sqrt(d2);
  return add_contact_ppfix(c, p, pid2, ov);
}
/* FineSearch2 :503-530 */
static int fine2(oracle_ctx *c, int pid1, int pid2) {
  int p1 = pid1 - nfixed_cs(c), p2 = pid2 - nfixed_cs(c);
  r4 a = c->pos[p1], b = c->pos[p2];
  double dr = a.w + b.w, dr2 = dr * dr;
  double dx = a.x - b.x, d2 = dx * dx;
  if (d2 > dr2) return 1;
  double dy = a.y - b.y; d2 = dy * dy + d2;
  if (d2 > dr2) return 1;
  double dz = a.z - b.z; d2 = dz * dz + d2;
  if (d2 > dr2) return 1;
  double ov = dr - sqrt(d2);
  if (c->id[p1] < c->id[p2]) return add_contact_pp(c, p1, p2, ov);
  return add_contact_pp(c, p2, p1, ov);
}
/* FineSearch3 :535-557 */
static int fine3(oracle_ctx *c, int pid1, int pid2) {
  int p = pid1 - nfixed_cs(c);
  int g = pid2 - (nfixed_cs(c) + c->nlocal);
  r4 a = c->pos[p], b = c->g_pos[g];
  double dr = a.w + b.w, dr2 = dr * dr;
  double dx = a.x - b.x, d2 = dx * dx;
  if (d2 > dr2) return 1;
  double dy = a.y - b.y; d2 = dy * dy + d2;
  if (d2 > dr2) return 1;
  double dz = a.z - b.z; d2 = dz * dz + d2;
  if (d2 > dr2) return 1;
  double ov = dr - sqrt(d2);
  return add_contact_ppg(c, p, g, ov);
}

/* route one candidate pair as LoopNBSMask does for a target m */
static int nbs_pair(oracle_ctx *c, int m, int n) {
  int nf = nfixed_cs(c), nfl = nf + c->nlocal;
  if (m <= nf) {
    if (n <= nfl && n > nf) { c->num_cnsv[0]++; return fine1(c, n, m); }
    return 1;
  } else if (m > nfl) {
    if (n <= nfl && n > nf) { c->num_cnsv[0]++; return fine3(c, n, m); }
    return 1;
  } else {
    c->num_cnsv[0]++;
    if (n <= nf) return fine1(c, m, n);
    if (n > nfl) return fine3(c, m, n);
    return fine2(c, m, n);
  }
}

/* NBSM_LoopNBSMask :302-473 */
static int loop_nbs_mask(oracle_ctx *c, int iz) {
  int m = HZ(c, 1, iz);
  while (m != -1) {
    int n = c->nextz[m];
    while (n != -1) { if (!nbs_pair(c, m, n)) return 0; n = c->nextz[n]; }
    n = HZ(c, 1, iz - 1);
    while (n != -1) { if (!nbs_pair(c, m, n)) return 0; n = c->nextz[n]; }
    for (int i = -1; i <= 1; i++) {
      n = HZ(c, 0, iz + i);
      while (n != -1) { if (!nbs_pair(c, m, n)) return 0; n = c->nextz[n]; }
    }
    for (int lx = 0; lx <= 2; lx++)
      for (int i = -1; i <= 1; i++) {
        n = HZ0(c, lx, iz + i);
        while (n != -1) { if (!nbs_pair(c, m, n)) return 0; n = c->nextz[n]; }
      }
    m = c->nextz[m];
  }
  return 1;
}

static void build_zlist(oracle_ctx *c, int ix, int m) {
  for (int iz = 0; iz <= c->nz + 1; iz++) HZ(c, m, iz) = -1;
  int n = c->headx[ix + m - 1];
  while (n != -1) {
    int iz = c->bz[n];
    c->nextz[n] = HZ(c, m, iz);
    HZ(c, m, iz) = n;
    n = c->nextx[n];
  }
}
static void build_zlist0(oracle_ctx *c, int ix, int m) {
  for (int iz = 0; iz <= c->nz + 1; iz++) HZ0(c, m, iz) = -1;
  int ixm = ix + m - 1;
  if (ixm > c->nx) return;
  int n = c->headx0[ixm];
  while (n != -1) {
    int iz = c->bz[n];
    c->nextz[n] = HZ0(c, m, iz);
    HZ0(c, m, iz) = n;
    n = c->nextx[n];
  }
}

/* NBSM_ContactSearch :118-158 */
static int nbsm_contact_search(oracle_ctx *c) {
  c->num_cnsv[0] = c->num_cnsv[1] = 0;
  if (c->nlocal == 0) return 1;
  int nf = nfixed_cs(c);
  int ntot = nf + c->nlocal + c->nghost;
  ensure_box(c, ntot);
  /* fixed box indices (Init_NBSM :107-111), then locals and ghosts (clcBoxIndex :191-211) */
  for (int i = 1; i <= nf; i++) {
    c->bx[i] = floor_i((c->fix_pos[i].x - c->xst_cs) / c->cell_len) + 1;
    c->by[i] = floor_i((c->fix_pos[i].y - c->yst_cs) / c->cell_len) + 1;
    c->bz[i] = floor_i((c->fix_pos[i].z - c->zst_cs) / c->cell_len) + 1;
  }
  for (int m = 1; m <= c->nlocal; m++) {
    int i = nf + m;
    c->bx[i] = floor_i((c->pos[m].x - c->xst_cs) / c->cell_len) + 1;
    c->by[i] = floor_i((c->pos[m].y - c->yst_cs) / c->cell_len) + 1;
    c->bz[i] = floor_i((c->pos[m].z - c->zst_cs) / c->cell_len) + 1;
  }
  for (int m = 1; m <= c->nghost; m++) {
    int i = nf + c->nlocal + m;
    c->bx[i] = floor_i((c->g_pos[m].x - c->xst_cs) / c->cell_len) + 1;
    c->by[i] = floor_i((c->g_pos[m].y - c->yst_cs) / c->cell_len) + 1;
    c->bz[i] = floor_i((c->g_pos[m].z - c->zst_cs) / c->cell_len) + 1;
  }
  for (int i = 1; i <= ntot; i++) {
    if (c->bx[i] < 1 || c->bx[i] > c->nx || c->by[i] < 1 || c->by[i] > c->ny || c->bz[i] < 1 ||
        c->bz[i] > c->nz) {
      fail(c, SDEM_E_ARG, "particle outside the search grid");
      return 0;
    }
  }
  /* BuildYList :217-229 */
  for (int iy = 0; iy <= c->ny + 1; iy++) c->heady[iy] = -1;
  for (int i = 1; i <= ntot; i++) {
    int iy = c->by[i];
    c->nexty[i] = c->heady[iy];
    c->heady[iy] = i;
  }
  for (int ix = 0; ix <= c->nx + 1; ix++) c->headx0[ix] = -1;
  for (int iy = 1; iy <= c->ny; iy++) {
    /* BuildXList :234-250 */
    for (int ix = 0; ix <= c->nx + 1; ix++) c->headx[ix] = -1;
    int n = c->heady[iy];
    while (n != -1) {
      int ix = c->bx[n];
      c->nextx[n] = c->headx[ix];
      c->headx[ix] = n;
      n = c->nexty[n];
    }
    if (c->heady[iy] != -1) {
      for (int iz = 0; iz <= c->nz + 1; iz++) { HZ(c, 0, iz) = -1; HZ0(c, 0, iz) = -1; }
      build_zlist0(c, 1, 1);
      for (int ix = 1; ix <= c->nx; ix++) {
        build_zlist(c, ix, 1);
        build_zlist0(c, ix, 2);
        if (c->headx[ix] != -1)
          for (int iz = 1; iz <= c->nz; iz++)
            if (!loop_nbs_mask(c, iz)) return 0;
        for (int iz = 0; iz <= c->nz + 1; iz++) {
          HZ(c, 0, iz) = HZ(c, 1, iz);
          HZ0(c, 0, iz) = HZ0(c, 1, iz);
          HZ0(c, 1, iz) = HZ0(c, 2, iz);
        }
      }
    }
    for (int ix = 0; ix <= c->nx + 1; ix++) c->headx0[ix] = c->headx[ix];
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* hierarchical NBS-Munjiza, Prtcl_Hrchl_Munjiza.f90                          */
/* Levels and level membership follow Init_Munjiza_Hrchl :84-237; every      */
/* level is searched with its own half stencil, and every particle of a     */
/* finer level is checked against the 27 cells of each coarser level around */
/* its coarse cell (NBSMH_Loop_CrossMask :763-863).  Fixed particles are    */
/* not searched (NBSMH_FineSearch :705-755 has no fixed branch).            */
/* ------------------------------------------------------------------------ */
static void init_hrchl(oracle_ctx *c) {
  double maxd = 2.0 * max_radius(c), mind = 2.0 * min_radius(c);
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
    numlevels = c->o.cs_numlvls;
  }
  /* particle counts per type from the current state (nPrtcl_in_Bin) */
  int *cnt = xcalloc(c->ntype + 1, sizeof(int));
  for (int i = 1; i <= c->nlocal; i++) cnt[c->ptype[i]]++;
  double cell_len = c->o.prtcl_cs_ratio * maxd;
  c->xst_cs = c->o.domain_min[0] - cell_len * 1.05;
  c->yst_cs = c->o.domain_min[1] - cell_len * 1.05;
  c->zst_cs = c->o.domain_min[2] - cell_len * 1.05;
  double xed = c->o.domain_max[0] + cell_len * 1.05;
  double yed = c->o.domain_max[1] + cell_len * 1.05;
  double zed = c->o.domain_max[2] + cell_len * 1.05;
  int nx = (int)((xed - c->xst_cs) / cell_len) + 1;
  int ny = (int)((yed - c->yst_cs) / cell_len) + 1;
  int nz = (int)((zed - c->zst_cs) / cell_len) + 1;
  int id_level = 0;
  double maxd_lvl = maxd;
  for (int idh = 1; idh <= numlevels; idh++) {
    double mind_lvl = maxd_lvl / 2.0;
    if (idh == numlevels) mind_lvl = 0.0;
    int num = 0;
    for (int i = 1; i <= c->ntype; i++) {
      double diam = 2.0 * c->pprop[i].radius;
      if (diam > mind_lvl && (idh == 1 || diam <= maxd_lvl)) num += cnt[i];
    }
    if (num > 0) {
      id_level++;
      for (int i = 1; i <= c->ntype; i++) {
        double diam = 2.0 * c->pprop[i].radius;
        if (diam > mind_lvl && (idh == 1 || diam <= maxd_lvl)) c->type_level[i] = id_level;
      }
      c->lvl[id_level].nx = nx; c->lvl[id_level].ny = ny; c->lvl[id_level].nz = nz;
      c->lvl[id_level].multiple = idh;
      c->lvl[id_level].cell_len = cell_len;
    }
    nx *= 2; ny *= 2; nz *= 2;
    maxd_lvl = mind_lvl;
    cell_len = cell_len / 2.0;
  }
  c->nlvl = id_level;
  free(cnt);
}

/* NBSMH_FineSearch :705-755; indices: 1..nlocal locals, nlocal+g ghosts */
static int hfine(oracle_ctx *c, int pid1, int pid2) {
  int nl = c->nlocal;
  if (pid1 > nl && pid2 > nl) return 1;
  if (pid1 > nl) {
    int g = pid1 - nl;
    r4 a = c->g_pos[g], b = c->pos[pid2];
    double dr = a.w + b.w, dr2 = dr * dr;
    double dx = a.x - b.x, d2 = dx * dx; if (d2 > dr2) return 1;
    double dy = a.y - b.y; d2 = dy * dy + d2; if (d2 > dr2) return 1;
    double dz = a.z - b.z; d2 = dz * dz + d2; if (d2 > dr2) return 1;
    return add_contact_ppg(c, pid2, g, dr - sqrt(d2));
  } else if (pid2 > nl) {
    int g = pid2 - nl;
    r4 a = c->pos[pid1], b = c->g_pos[g];
    double dr = a.w + b.w, dr2 = dr * dr;
    double dx = a.x - b.x, d2 = dx * dx; if (d2 > dr2) return 1;
    double dy = a.y - b.y; d2 = dy * dy + d2; if (d2 > dr2) return 1;
    double dz = a.z - b.z; d2 = dz * dz + d2; if (d2 > dr2) return 1;
    return add_contact_ppg(c, pid1, g, dr - sqrt(d2));
  } else {
    r4 a = c->pos[pid1], b = c->pos[pid2];
    double dr = a.w + b.w, dr2 = dr * dr;
    double dx = a.x - b.x, d2 = dx * dx; if (d2 > dr2) return 1;
    double dy = a.y - b.y; d2 = dy * dy + d2; if (d2 > dr2) return 1;
    double dz = a.z - b.z; d2 = dz * dz + d2; if (d2 > dr2) return 1;
    double ov = dr - sqrt(d2);
    if (c->id[pid1] < c->id[pid2]) return add_contact_pp(c, pid1, pid2, ov);
    return add_contact_pp(c, pid2, pid1, ov);
  }
}

typedef struct { int n; int *start, *list; } cellmap; /* CSR of a level's cells */

static int hrchl_contact_search(oracle_ctx *c) {
  c->num_cnsv[0] = c->num_cnsv[1] = 0;
  if (c->nlocal == 0) return 1;
  init_hrchl(c);
  int nl = c->nlocal, ntot = nl + c->nghost;
  int *lv = xcalloc(ntot + 1, sizeof(int));
  int *ix = xcalloc(ntot + 1, sizeof(int)), *iy = xcalloc(ntot + 1, sizeof(int)), *iz = xcalloc(ntot + 1, sizeof(int));
  for (int i = 1; i <= ntot; i++) {
    r4 p = i <= nl ? c->pos[i] : c->g_pos[i - nl];
    int t = i <= nl ? c->ptype[i] : c->g_type[i - nl];
    int l = c->type_level[t];
    double cl = c->lvl[l].cell_len;
    lv[i] = l;
    ix[i] = floor_i((p.x - c->xst_cs) / cl) + 1;
    iy[i] = floor_i((p.y - c->yst_cs) / cl) + 1;
    iz[i] = floor_i((p.z - c->zst_cs) / cl) + 1;
    if (ix[i] < 1 || ix[i] > c->lvl[l].nx || iy[i] < 1 || iy[i] > c->lvl[l].ny || iz[i] < 1 || iz[i] > c->lvl[l].nz) {
      free(lv); free(ix); free(iy); free(iz);
      fail(c, SDEM_E_ARG, "particle outside the search grid");
      return 0;
    }
  }
  /* per-level cell lists in (iy, ix, iz) order with push-front inside a cell */
  cellmap cm[8];
  for (int l = 1; l <= c->nlvl; l++) {
    int nxl = c->lvl[l].nx, nyl = c->lvl[l].ny, nzl = c->lvl[l].nz;
    int64_t ncell = (int64_t)nxl * nyl * nzl;
    cm[l].n = (int)ncell;
    cm[l].start = xcalloc(ncell + 2, sizeof(int));
    cm[l].list = xcalloc(ntot + 1, sizeof(int));
    for (int i = 1; i <= ntot; i++)
      if (lv[i] == l) cm[l].start[((int64_t)(iy[i] - 1) * nxl + (ix[i] - 1)) * nzl + (iz[i] - 1) + 1]++;
    for (int64_t k = 0; k < ncell; k++) cm[l].start[k + 1] += cm[l].start[k];
    int *fill = xcalloc(ncell + 1, sizeof(int));
    for (int i = ntot; i >= 1; i--)
      if (lv[i] == l) {
        int64_t cc = ((int64_t)(iy[i] - 1) * nxl + (ix[i] - 1)) * nzl + (iz[i] - 1);
        cm[l].list[cm[l].start[cc] + fill[cc]++] = i;
      }
    free(fill);
  }
#define CELL(l, X, Y, Z) ((((int64_t)((Y)-1) * c->lvl[l].nx + ((X)-1)) * c->lvl[l].nz) + ((Z)-1))
  int ok = 1;
  for (int base = c->nlvl; base >= 1 && ok; base--) {
    int nxl = c->lvl[base].nx, nyl = c->lvl[base].ny, nzl = c->lvl[base].nz;
    for (int y = 1; y <= nyl && ok; y++)
      for (int x = 1; x <= nxl && ok; x++)
        for (int z = 1; z <= nzl && ok; z++) {
          int64_t cc = CELL(base, x, y, z);
          for (int a = cm[base].start[cc]; a < cm[base].start[cc + 1] && ok; a++) {
            int m = cm[base].list[a];
            /* same cell successors, then the 13-cell half stencil */
            for (int b = a + 1; b < cm[base].start[cc + 1] && ok; b++) {
              int n = cm[base].list[b];
              if (m <= nl || n <= nl) { c->num_cnsv[0]++; ok = hfine(c, m, n); }
            }
            static const int hs[13][3] = {{0, 0, -1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1},
                                          {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {0, -1, -1},
                                          {0, -1, 0}, {0, -1, 1}, {1, -1, -1}, {1, -1, 0}, {1, -1, 1}};
            for (int s = 0; s < 13 && ok; s++) {
              int X = x + hs[s][0], Y = y + hs[s][1], Z = z + hs[s][2];
              if (X < 1 || X > nxl || Y < 1 || Y > nyl || Z < 1 || Z > nzl) continue;
              int64_t c2 = CELL(base, X, Y, Z);
              for (int b = cm[base].start[c2]; b < cm[base].start[c2 + 1] && ok; b++) {
                int n = cm[base].list[b];
                if (m <= nl || n <= nl) { c->num_cnsv[0]++; ok = hfine(c, m, n); }
              }
            }
            /* cross-level: 27 coarse cells of every coarser level */
            for (int idh = base - 1; idh >= 1 && ok; idh--) {
              int coe = 1 << (c->lvl[base].multiple - c->lvl[idh].multiple);
              int cx = (x - 1) / coe + 1, cy = (y - 1) / coe + 1, cz = (z - 1) / coe + 1;
              for (int dy = -1; dy <= 1 && ok; dy++)
                for (int dx = -1; dx <= 1 && ok; dx++)
                  for (int dz = -1; dz <= 1 && ok; dz++) {
                    int X = cx + dx, Y = cy + dy, Z = cz + dz;
                    if (X < 1 || X > c->lvl[idh].nx || Y < 1 || Y > c->lvl[idh].ny || Z < 1 || Z > c->lvl[idh].nz) continue;
                    int64_t c2 = CELL(idh, X, Y, Z);
                    for (int b = cm[idh].start[c2]; b < cm[idh].start[c2 + 1] && ok; b++) {
                      int n = cm[idh].list[b];
                      if (m <= nl || n <= nl) { c->num_cnsv[1]++; ok = hfine(c, m, n); }
                    }
                  }
            }
          }
        }
  }
#undef CELL
  for (int l = 1; l <= c->nlvl; l++) { free(cm[l].start); free(cm[l].list); }
  free(lv); free(ix); free(iy); free(iz);
  return ok;
}

/* ------------------------------------------------------------------------ */
/* near-wall lists, Prtcl_ContactSearchPW.f90                                */
/* ------------------------------------------------------------------------ */
static void init_pwcs(oracle_ctx *c) {
  /* Init_PWCS_List :81-92 */
  for (int i = 0; i <= c->mlocal; i++) c->bucket_pw[i] = 0;
  for (int i = 1; i <= c->max_near; i++) c->next_pw[i] = -i - 1;
  c->next_pw[c->max_near] = 0;
  c->next_insert_pw = 1;
}

static void realloc_pwcs(oracle_ctx *c) {
  /* Reallocate_PWCS_List :122-154 */
  int sizep = c->max_near, sizen = (int)(1.2 * sizep);
  c->id_wall = realloc(c->id_wall, sizeof(int) * (sizen + 1));
  c->next_pw = realloc(c->next_pw, sizeof(int) * (sizen + 1));
  for (int i = sizep + 1; i <= sizen; i++) c->next_pw[i] = -i - 1;
  c->next_pw[sizen] = 0;
  c->next_insert_pw = sizep + 1;
  c->max_near = sizen;
}

/* InitContactSearchPW :49-76 */
static void init_contact_search_pw(oracle_ctx *c) {
  init_pwcs(c);
  double mv = 0.0;
  for (int i = 1; i <= c->nwall; i++) mv = fmax_f(mv, norm3(c->walls[i].trans_vel));
  c->max_wall_vel = mv;
  double rmax = max_radius(c);
  c->dxmax1 = rmax * c->o.wall_neighbor_ratio;
  c->dxmax2 = rmax * (c->o.wall_neighbor_ratio + 1.0);
}

/* UpdateNearPrtclsPW :159-235 */
static void update_near_prtcls_pw(oracle_ctx *c, int iter) {
  if (iter < c->next_iter_update) return;
  init_pwcs(c);
  for (int wid = 1; wid <= c->nwall; wid++) {
    r3 mn = c->walls[wid].min_point, mx = c->walls[wid].max_point;
    r3 pmn = sub3(mn, mul3(R3(1, 1, 1), c->dxmax2));
    r3 pmx = add3(mx, mul3(R3(1, 1, 1), c->dxmax2));
    for (int i = 1; i <= c->nlocal; i++) {
      r4 p = c->pos[i];
      if (p.x >= pmn.x && p.x <= pmx.x && p.y >= pmn.y && p.y <= pmx.y && p.z >= pmn.z && p.z <= pmx.z) {
        if (c->next_insert_pw == 0) realloc_pwcs(c);
        int ni = c->next_insert_pw;
        c->next_insert_pw = -c->next_pw[ni];
        c->id_wall[ni] = wid;
        c->next_pw[ni] = c->bucket_pw[i];
        c->bucket_pw[i] = ni;
      }
    }
  }
  double maxreal = 0.0;
  for (int i = 1; i <= c->nlocal; i++) maxreal = fmax_f(maxreal, norm3(LV(c, i, 1)));
  double max_v = maxreal + c->max_wall_vel;
  max_v = fmax_f(max_v, 0.3);
  max_v = max_v * 1.5;
  maxreal = 0.0;
  for (int i = 1; i <= c->nlocal; i++) maxreal = fmax_f(maxreal, norm3(LA(c, i, 1)));
  double max_a = fmax_f(maxreal, norm3(R3(c->o.gravity[0], c->o.gravity[1], c->o.gravity[2])));
  max_a = max_a * 1.5;
  double t = (sqrt(max_v * max_v + 2.0 * max_a * c->dxmax1) - max_v) / max_a;
  /* int(t/dt) (:225): t is NaN (0/0) when nothing accelerates and there is
     no gravity; int() of NaN or of 2^31 or more gives the most negative
     integer on x86 (cvttsd2si), which max(., 1) makes 1.  Spelled out here,
     since the C cast of such values is undefined. */
  double q = t / c->o.dt;
  int num_iter = (q >= 1.0 && q < 2147483648.0) ? (int)q : 1;
  if (num_iter > c->o.wall_max_update_iter) num_iter = c->o.wall_max_update_iter;
  c->next_iter_update = iter + num_iter;
}

/* FindContactsPW :298-319 */
static int find_contacts_pw(oracle_ctx *c) {
  for (int i = 1; i <= c->nlocal; i++) {
    int nid = c->bucket_pw[i];
    while (nid > 0) {
      int wid = c->id_wall[nid];
      double ov;
      r3 nv;
      if (pw_is_in_contact(&c->walls[wid], c->pos[i], &ov, &nv))
        if (!add_contact_pw(c, i, wid, ov, nv)) return 0;
      nid = c->next_pw[nid];
    }
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* DEMS_clc_Acceleration :360-375 and Prtcl_Integrate :15-88                 */
/* ------------------------------------------------------------------------ */
static void clc_acceleration(oracle_ctx *c) {
  r3 g = R3(c->o.gravity[0], c->o.gravity[1], c->o.gravity[2]);
  for (int i = 1; i <= c->nlocal; i++) {
    const pure_prop *p = &c->pprop[c->ptype[i]];
    LA(c, i, 1) = add3(div3(c->force[i], p->mass), g);
    RA(c, i, 1) = div3(c->torque[i], p->inertia);
  }
}

static void integrate(oracle_ctx *c) {
  const double dt = c->o.dt;
  const double ab2[2] = {3.0 / 2.0, -1.0 / 2.0};
  const double ab3[3] = {23.0 / 12.0, -16.0 / 12.0, 5.0 / 12.0};
  int n = c->nlocal;
  if (c->o.pi_method == SDEM_PIM_FE) {
    for (int i = 1; i <= n; i++) {
      r3 s = mul3(LV(c, i, 1), dt);
      c->pos[i].x = c->pos[i].x + s.x; c->pos[i].y = c->pos[i].y + s.y; c->pos[i].z = c->pos[i].z + s.z;
      LV(c, i, 1) = add3(LV(c, i, 1), mul3(LA(c, i, 1), dt));
    }
  } else if (c->o.pi_method == SDEM_PIM_AB2) {
    for (int i = 1; i <= n; i++) {
      r3 v1 = LV(c, i, 1);
      r3 s = mul3(add3(mul3(v1, ab2[0]), mul3(LV(c, i, 2), ab2[1])), dt);
      c->pos[i].x = c->pos[i].x + s.x; c->pos[i].y = c->pos[i].y + s.y; c->pos[i].z = c->pos[i].z + s.z;
      LV(c, i, 1) = add3(v1, mul3(add3(mul3(LA(c, i, 1), ab2[0]), mul3(LA(c, i, 2), ab2[1])), dt));
      LV(c, i, 2) = v1;
      LA(c, i, 2) = LA(c, i, 1);
    }
  } else {
    for (int i = 1; i <= n; i++) {
      r3 v1 = LV(c, i, 1), v2 = LV(c, i, 2);
      r3 s = mul3(add3(add3(mul3(v1, ab3[0]), mul3(v2, ab3[1])), mul3(LV(c, i, 3), ab3[2])), dt);
      c->pos[i].x = c->pos[i].x + s.x; c->pos[i].y = c->pos[i].y + s.y; c->pos[i].z = c->pos[i].z + s.z;
      LV(c, i, 1) = add3(v1, mul3(add3(add3(mul3(LA(c, i, 1), ab3[0]), mul3(LA(c, i, 2), ab3[1])),
                                       mul3(LA(c, i, 3), ab3[2])), dt));
      LV(c, i, 3) = v2;
      LV(c, i, 2) = v1;
      LA(c, i, 3) = LA(c, i, 2);
      LA(c, i, 2) = LA(c, i, 1);
    }
  }
  if (c->o.pri_method == SDEM_PIM_FE) {
    for (int i = 1; i <= n; i++) {
      c->theta[i] = add3(c->theta[i], mul3(RV(c, i, 1), dt));
      RV(c, i, 1) = add3(RV(c, i, 1), mul3(RA(c, i, 1), dt));
    }
  } else if (c->o.pri_method == SDEM_PIM_AB2) {
    for (int i = 1; i <= n; i++) {
      r3 w1 = LV(c, i, 1); /* Prtcl_Integration.f90:64 reads GPrtcl_linVel(1,i) */
      c->theta[i] = add3(c->theta[i], mul3(add3(mul3(w1, ab2[0]), mul3(RV(c, i, 2), ab2[1])), dt));
      RV(c, i, 1) = add3(w1, mul3(add3(mul3(RA(c, i, 1), ab2[0]), mul3(RA(c, i, 2), ab2[1])), dt));
      RV(c, i, 2) = w1;
      RA(c, i, 2) = RA(c, i, 1);
    }
  } else {
    for (int i = 1; i <= n; i++) {
      r3 w1 = RV(c, i, 1), w2 = RV(c, i, 2);
      c->theta[i] = add3(c->theta[i], mul3(add3(add3(mul3(w1, ab3[0]), mul3(w2, ab3[1])),
                                                mul3(RV(c, i, 3), ab3[2])), dt));
      RV(c, i, 1) = add3(w1, mul3(add3(add3(mul3(RA(c, i, 1), ab3[0]), mul3(RA(c, i, 2), ab3[1])),
                                       mul3(RA(c, i, 3), ab3[2])), dt));
      RV(c, i, 3) = w2;
      RV(c, i, 2) = w1;
      RA(c, i, 3) = RA(c, i, 2);
      RA(c, i, 2) = RA(c, i, 1);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* one step: DEMS_iterate :201-248 (no I/O, no logging)                       */
/* ------------------------------------------------------------------------ */
int oracle_step(oracle_ctx *c, int nsteps) {
  if (!c->have_state || !c->have_pprop || !c->have_wprop)
    return fail(c, SDEM_E_STATE, "properties and state must be set before stepping");
  double maxd = 2.0 * max_radius(c);
  c->lenforcs = c->o.prtcl_cs_ratio * maxd;
  for (int d = 0; d < 3; d++)
    if (1.02 * c->lenforcs > c->simlen[d]) return fail(c, SDEM_E_ARG, "so big Diameter");
  if (!c->pw_initialized) { init_contact_search_pw(c); c->pw_initialized = 1; }
  init_nbsm(c);
  for (int s = 0; s < nsteps; s++) {
    /* preIteration :322-331 */
    update_near_prtcls_pw(c, c->iter_number);
    for (int i = 1; i <= c->nlocal; i++) { c->force[i] = ZERO3; c->torque[i] = ZERO3; }
    cl_pre_iteration(c);
    comm_for_cntct(c);
    int ok = (c->o.cs_method == SDEM_CSM_NBS_MUNJIZA_HRCHL) ? hrchl_contact_search(c)
                                                           : nbsm_contact_search(c);
    if (!ok) return c->err[0] && strstr(c->err, "full") ? SDEM_E_CAPACITY : SDEM_E_ARG;
    if (!find_contacts_pw(c)) return SDEM_E_CAPACITY;
    clc_acceleration(c);
    integrate(c);
    comm_for_exchange(c);
    cl_remv_released(c);
    c->iter_number++;
  }
  return SDEM_OK;
}

/* ------------------------------------------------------------------------ */
/* readers (outputs in ascending id order)                                  */
/* ------------------------------------------------------------------------ */
static int *order_by_id(const oracle_ctx *c) {
  int n = c->nlocal;
  int *ord = xcalloc(n + 1, sizeof(int));
  for (int i = 0; i < n; i++) ord[i] = i + 1;
  /* insertion sort is fine for test sizes; use a simple merge for larger */
  for (int i = 1; i < n; i++) {
    int v = ord[i], j = i - 1;
    while (j >= 0 && c->id[ord[j]] > c->id[v]) { ord[j + 1] = ord[j]; j--; }
    ord[j + 1] = v;
  }
  return ord;
}

int oracle_get_num_particles(oracle_ctx *c, int *n) { *n = c->nlocal; return SDEM_OK; }

int oracle_get_state(oracle_ctx *c, int cap, int *nout, int *id, int *ptype, int *mark,
                     double *posr, double *lvel, double *lacc, double *theta, double *rvel,
                     double *racc) {
  int n = c->nlocal;
  *nout = n;
  if (cap < n) return fail(c, SDEM_E_CAPACITY, "output arrays too small");
  int *ord = order_by_id(c);
  int ts = c->tsize, rs = c->rsize;
  for (int k = 0; k < n; k++) {
    int i = ord[k];
    if (id) id[k] = c->id[i];
    if (ptype) ptype[k] = c->ptype[i];
    if (mark) mark[k] = c->mark[i];
    if (posr) { posr[4 * k] = c->pos[i].x; posr[4 * k + 1] = c->pos[i].y; posr[4 * k + 2] = c->pos[i].z; posr[4 * k + 3] = c->pos[i].w; }
    for (int l = 1; l <= ts; l++) {
      size_t o = ((size_t)k * ts + (l - 1)) * 3;
      if (lvel) { lvel[o] = LV(c, i, l).x; lvel[o + 1] = LV(c, i, l).y; lvel[o + 2] = LV(c, i, l).z; }
      if (lacc) { lacc[o] = LA(c, i, l).x; lacc[o + 1] = LA(c, i, l).y; lacc[o + 2] = LA(c, i, l).z; }
    }
    if (theta) { theta[3 * k] = c->theta[i].x; theta[3 * k + 1] = c->theta[i].y; theta[3 * k + 2] = c->theta[i].z; }
    for (int l = 1; l <= rs; l++) {
      size_t o = ((size_t)k * rs + (l - 1)) * 3;
      if (rvel) { rvel[o] = RV(c, i, l).x; rvel[o + 1] = RV(c, i, l).y; rvel[o + 2] = RV(c, i, l).z; }
      if (racc) { racc[o] = RA(c, i, l).x; racc[o + 1] = RA(c, i, l).y; racc[o + 2] = RA(c, i, l).z; }
    }
  }
  free(ord);
  return SDEM_OK;
}

int oracle_get_forces(oracle_ctx *c, int cap, int *id, double *f, double *t) {
  int n = c->nlocal;
  if (cap < n) return fail(c, SDEM_E_CAPACITY, "output arrays too small");
  int *ord = order_by_id(c);
  for (int k = 0; k < n; k++) {
    int i = ord[k];
    if (id) id[k] = c->id[i];
    if (f) { f[3 * k] = c->force[i].x; f[3 * k + 1] = c->force[i].y; f[3 * k + 2] = c->force[i].z; }
    if (t) { t[3 * k] = c->torque[i].x; t[3 * k + 1] = c->torque[i].y; t[3 * k + 2] = c->torque[i].z; }
  }
  free(ord);
  return SDEM_OK;
}

typedef struct { int owner, partner, kind, slot; } rec_t;
static int rec_cmp(const void *a, const void *b) {
  const rec_t *x = a, *y = b;
  if (x->owner != y->owner) return x->owner < y->owner ? -1 : 1;
  if (x->partner != y->partner) return x->partner < y->partner ? -1 : 1;
  return 0;
}

static int64_t collect(const oracle_ctx *c, rec_t **out) {
  int64_t m = 0;
  for (int pid = 1; pid <= c->nlocal; pid++)
    for (int n = c->bucket[pid]; n > 0; n = c->next[n])
      if (c->status[n] > 0) m++;
  rec_t *r = xcalloc(m + 1, sizeof(rec_t));
  int64_t k = 0;
  for (int pid = 1; pid <= c->nlocal; pid++)
    for (int n = c->bucket[pid]; n > 0; n = c->next[n])
      if (c->status[n] > 0) { r[k].owner = c->id[pid]; r[k].partner = c->id_j[n]; r[k].kind = c->kind[n]; r[k].slot = n; k++; }
  qsort(r, m, sizeof(rec_t), rec_cmp);
  *out = r;
  return m;
}

int oracle_get_num_contacts(oracle_ctx *c, int64_t *m) {
  rec_t *r;
  *m = collect(c, &r);
  free(r);
  return SDEM_OK;
}

int oracle_get_contacts(oracle_ctx *c, int64_t cap, int64_t *mout, int *kind, int *owner,
                        int *partner, double *td, double *vr) {
  rec_t *r;
  int64_t m = collect(c, &r);
  *mout = m;
  if (cap < m) { free(r); return fail(c, SDEM_E_CAPACITY, "output arrays too small"); }
  for (int64_t k = 0; k < m; k++) {
    if (kind) kind[k] = r[k].kind;
    if (owner) owner[k] = r[k].owner;
    if (partner) partner[k] = r[k].partner;
    if (td) { td[3 * k] = c->tandel[r[k].slot].x; td[3 * k + 1] = c->tandel[r[k].slot].y; td[3 * k + 2] = c->tandel[r[k].slot].z; }
    if (vr) vr[k] = c->vrel[r[k].slot];
  }
  free(r);
  return SDEM_OK;
}

int oracle_get_counts(oracle_ctx *c, int *num) {
  for (int k = 0; k < 4; k++) num[k] = c->num_cntcts[k];
  return SDEM_OK;
}

int oracle_get_iter(oracle_ctx *c, int *iter, int *next_update) {
  *iter = c->iter_number;
  *next_update = c->next_iter_update;
  return SDEM_OK;
}

int oracle_get_num_cnsv(oracle_ctx *c, int64_t *same_level, int64_t *cross_level) {
  *same_level = c->num_cnsv[0];
  *cross_level = c->num_cnsv[1];
  return SDEM_OK;
}

/* pair-table and wall accessors for the tests */
int oracle_get_pair_prop(oracle_ctx *c, int ti, int tj, double *out11) {
  if (ti < 1 || ti > c->ntype || tj < 1 || tj > c->ntype) return SDEM_E_ARG;
  const bnry_prop *b = &c->pp_tab[(ti - 1) * c->ntype + (tj - 1)];
  double v[11] = {b->rad_eff, b->mass_eff, b->kn, b->kt, b->dn, b->dt, b->rest_n, b->rest_t, b->fric_s, b->fric_k, b->fric_roll};
  memcpy(out11, v, sizeof(v));
  return SDEM_OK;
}

int oracle_get_wall(oracle_ctx *c, int k, double *n3, double *d, int *wall_id) {
  if (k < 1 || k > c->nwall) return SDEM_E_ARG;
  n3[0] = c->walls[k].n.x; n3[1] = c->walls[k].n.y; n3[2] = c->walls[k].n.z;
  *d = c->walls[k].d;
  *wall_id = c->walls[k].wall_id;
  return SDEM_OK;
}

int oracle_get_num_walls(oracle_ctx *c, int *n) { *n = c->nwall; return SDEM_OK; }

/* one wall-contact query, for geometry unit tests */
int oracle_wall_contact(oracle_ctx *c, int k, const double *posr4, double *ovrlp, double *nv3_out) {
  if (k < 1 || k > c->nwall) return -1;
  r4 p = {posr4[0], posr4[1], posr4[2], posr4[3]};
  double ov;
  r3 nv;
  int hit = pw_is_in_contact(&c->walls[k], p, &ov, &nv);
  if (hit) { *ovrlp = ov; nv3_out[0] = nv.x; nv3_out[1] = nv.y; nv3_out[2] = nv.z; }
  return hit;
}
