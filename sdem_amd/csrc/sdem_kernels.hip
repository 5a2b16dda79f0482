// sdem_kernels.hip -- gfx950 kernels of one DEM time step.
//
// A step is a short pipeline on one stream.  Candidate (Verlet) lists are
// rebuilt, together with the cell sort, only when the particles may have
// moved half the skin since the last rebuild (the sum over steps of the
// largest move of a particle passes half the skin); the contact set itself
// is decided every step with the reference's exact predicate, so it does
// not depend on when lists are rebuilt.
//
//   k_pre                   : near-wall bit masks, max|v|, max|a| of the
//                             state as set or replayed (first step of a run
//                             only; later steps get them from k_force)
//   k_ctl                   : one wave: decides the rebuild, runs the
//                             near-wall cadence and advances iterNumber
//                             (on steps queued without the rebuild
//                             kernels, the first wave of k_contact does it)
//   -- rebuild steps only (the kernels return at once otherwise) --
//   k_bin                   : (level, cell) key per particle, per-key counts
//   k_scan_tiles/_sums/_add : exclusive scan of the counts
//   k_scatter / k_rank      : deterministic counting sort (in-cell order =
//                             storage order) + packed positions, ids in
//                             cell order
//   k_vlist                 : candidate lists, pairs within ri+rj+skin
//                             (fixed particles included), sorted by
//                             partner id; the particles that have any are
//                             packed per box with their smallest gap
//   -- every step --
//   k_contact               : exact contact test over the packed lists
//                             (a particle whose smallest gap exceeds twice
//                             the box's summed largest moves since the
//                             build touches nothing and is skipped)
//   k_force                 : history match, contact forces, accelerations,
//                             integration, wrap/delete, the largest move of
//                             the step, and the next step's near-wall masks
//
// Floating-point kernels are built with -ffp-contract=off so that every
// product and sum is rounded on its own, in the order the reference's
// operator functions impose (Prtcl_TypeDef.f90:105-288).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "sdem_internal.h"

namespace sdem {

// ---------------------------------------------------------------------------
// real3 arithmetic in the reference's evaluation order (Prtcl_TypeDef.f90)
// ---------------------------------------------------------------------------
struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 D3(double x, double y, double z) {
  d3 r;
  r.x = x; r.y = y; r.z = z;
  return r;
}
__device__ __forceinline__ d3 add3(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 sub3(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 mul3(d3 a, double v) { return D3(a.x * v, a.y * v, a.z * v); }
__device__ __forceinline__ d3 div3(d3 a, double v) { return D3(a.x / v, a.y / v, a.z / v); }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross3(d3 u, d3 v) {
  return D3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
__device__ __forceinline__ double norm3(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// .nv. : unit(p1-p2), zero when |p1-p2| <= 1e-10 (Prtcl_TypeDef.f90:224-281)
__device__ __forceinline__ d3 nv3(d3 p1, d3 p2) {
  d3 s = sub3(p1, p2);
  double n = norm3(s);
  if (n > 1.0e-10) return div3(s, n);
  return D3(0.0, 0.0, 0.0);
}
__device__ __forceinline__ d3 ld3(const double* a, int cap, int i) {
  return D3(a[i], a[cap + i], a[2 * cap + i]);
}
__device__ __forceinline__ void st3(double* a, int cap, int i, d3 v) {
  a[i] = v.x;
  a[cap + i] = v.y;
  a[2 * cap + i] = v.z;
}
// The same with nontemporal hints, for the per-particle streams the force
// kernel reads or writes once per step and no other thread gathers (the
// history levels and the new state): they pass through the caches without
// displacing the positions and velocities that neighbours gather (the plain
// copy of this part reaches 6.6 TB/s so, 6.2 TB/s without the hints).
__device__ __forceinline__ d3 ld3n(const double* a, int cap, int i) {
  return D3(__builtin_nontemporal_load(a + i), __builtin_nontemporal_load(a + cap + i),
            __builtin_nontemporal_load(a + 2 * cap + i));
}
__device__ __forceinline__ void st3n(double* a, int cap, int i, d3 v) {
  __builtin_nontemporal_store(v.x, a + i);
  __builtin_nontemporal_store(v.y, a + cap + i);
  __builtin_nontemporal_store(v.z, a + 2 * cap + i);
}

// ---------------------------------------------------------------------------
// Prtcl_ContactForce_inc.f90:1-68 for one contact record
// ---------------------------------------------------------------------------
struct CF {
  d3 fsum, moment, mrij, tandel;
  double vri;
};

__device__ __forceinline__ CF contact_force(const Params& P, const Bnry& B, d3 norm_v, double ri,
                                            double rj, d3 veli, d3 velj, d3 rvei, d3 rvej,
                                            double ovrlp, d3 tandel, double vri) {
  const double dt = P.dt;
  d3 vel_w = cross3(add3(mul3(rvei, ri), mul3(rvej, rj)), norm_v);
  d3 vrij = add3(sub3(veli, velj), vel_w);
  double vrn = dot3(vrij, norm_v);
  if (vri < -1.0) vri = fabs(vrn);
  d3 vij_n = mul3(norm_v, vrn);
  d3 vij_t = sub3(vrij, vij_n);

  d3 ot = tandel;
  double n1 = norm3(ot);
  ot = sub3(ot, mul3(norm_v, dot3(ot, norm_v)));
  double n2 = norm3(ot);
  if (n2 > 1.0e-10) ot = mul3(ot, n1 / n2);
  else ot = D3(0.0, 0.0, 0.0);
  ot = add3(mul3(vij_t, dt), ot);

  double fn, k_t;
  const double d_t = 0.0;
  d3 ftij;
  if (P.cf_type == 3) {  // constant-stiffness linear model (not in the reference)
    fn = -B.kn * ovrlp + -B.dn * vrn;
    k_t = -B.kt;
  } else if (P.cf_type == 1) {  // CFT_LSD
    // Vel_in**0.4 as (Vel_in**0.2)**2: one pow instead of three, within
    // an ulp or two of the reference's pow results
    double vel_in = vri > 0.01 ? vri : 0.01;
    const double v02 = pow(vel_in, 0.2);
    const double v04 = v02 * v02;
    double k_n = -(B.kn * v04);
    double d_n = -(B.dn * v02);
    fn = k_n * ovrlp + d_n * vrn;
    k_t = -(B.kt * v04);
  } else {               // CFT_nLin
    // ovrlp**1.5 and ovrlp**0.25 from square roots (a few instructions
    // each, where FP64 pow takes well over a hundred), within an ulp or
    // two of the reference's pow results
    double k_n = -B.kn;
    double d_n = -B.dn;
    const double sq = sqrt(ovrlp);
    fn = k_n * (ovrlp * sq) + d_n * sqrt(sq) * vrn;
    k_t = -(B.kt * sq);
  }
  ftij = add3(mul3(ot, k_t), mul3(vij_t, d_t));
  d3 fnij = mul3(norm_v, fn);

  double ft = norm3(ftij);
  double ft_fric = B.fric_s * fabs(fn);
  if (ft > ft_fric) {
    ft_fric = B.fric_k * fabs(fn);
    ftij = mul3(ftij, ft_fric / ft);
    ot = div3(ftij, k_t);
  }

  d3 w_hat = sub3(rvei, rvej);
  double wm = norm3(w_hat);
  const double mr = P.ct_model == 1 ? -(B.fric_roll * fabs(fn) * B.rad_eff)
                                    : -(B.fric_roll * fabs(fn) * B.rad_eff * norm3(vel_w));
  d3 mrij;
  if (mr == 0.0 && wm <= 1.0e308) {
    // no rolling resistance (a zero coefficient): the products are zeros,
    // signed as w_hat times mr, and w_hat = w / wm has the signs of w
    // (+0 below 1e-10), so the three divisions are not needed
    if (wm > 1.0e-10) mrij = D3(copysign(0.0, w_hat.x) * mr, copysign(0.0, w_hat.y) * mr, copysign(0.0, w_hat.z) * mr);
    else mrij = D3(0.0 * mr, 0.0 * mr, 0.0 * mr);
  } else {
    if (wm > 1.0e-10) w_hat = div3(w_hat, wm);
    else w_hat = D3(0.0, 0.0, 0.0);
    mrij = mul3(w_hat, mr);
  }

  CF o;
  o.tandel = ot;
  o.vri = vri;
  o.moment = cross3(norm_v, ftij);
  o.fsum = add3(fnij, ftij);
  o.mrij = mrij;
  return o;
}

// ---------------------------------------------------------------------------
// PW_isInContact and helpers (Prtcl_Geometry.f90:55-109, 185-261)
// ---------------------------------------------------------------------------
__device__ __forceinline__ d3 wp(const double* p) { return D3(p[0], p[1], p[2]); }

__device__ bool pw_in_plane(d3 p1, d3 p2, d3 p3, d3 p4, d3 cp) {
  d3 p1p = sub3(p1, cp), p2p = sub3(p2, cp), p3p = sub3(p3, cp), p4p = sub3(p4, cp);
  double p1p3 = dot3(p1p, p3p), p2p4 = dot3(p2p, p4p), p1p4 = dot3(p1p, p4p), p2p3 = dot3(p2p, p3p);
  if (p1p3 * p2p4 - p1p4 * p2p3 < 0.0) return false;
  double p1p2 = dot3(p1p, p2p), p3p4 = dot3(p3p, p4p);
  if (p1p3 * p2p4 - p1p2 * p3p4 < 0.0) return false;
  double p2p2 = dot3(p2p, p2p);
  if (p1p2 * p2p3 - p1p3 * p2p2 < 0.0) return false;
  return true;
}

__device__ bool line_point_check(d3 lp1, d3 lp2, d3 pos, double r, d3* nv, double* ovrlp) {
  d3 w = sub3(pos, lp1);
  d3 v = sub3(lp2, lp1);
  double length = norm3(sub3(lp1, lp2));
  double t = dot3(w, v) / dot3(v, v);
  d3 cp;
  if (t >= 0.0 && t <= 1.0) {
    cp = add3(mul3(v, t), lp1);
  } else if (t >= (-r / length) && t < 0.0) {
    cp = lp1;
  } else if (t > 1.0 && t >= (1.0 + r / length)) {  // as written in the reference
    cp = lp2;
  } else {
    return false;
  }
  double ov = r - norm3(sub3(cp, pos));
  if (ov >= 0.0) {
    *ovrlp = ov;
    *nv = nv3(pos, cp);
    return true;
  }
  return false;
}

__device__ bool pw_in_contact(const Wall& W, d3 p, double radius, double* ovrlp, d3* nv) {
  d3 n = wp(W.n);
  double dist = dot3(n, p) + W.d;
  double ov;
  if (W.both_side) {
    ov = radius - fabs(dist);
  } else {
    if (dist < 0.0) return false;
    ov = radius - dist;
  }
  if (ov < 0.0) return false;
  if (W.is_infinite) {
    *ovrlp = ov;
    *nv = mul3(n, copysign(1.0, dist));
    return true;
  }
  double t = -dist;
  d3 cp = add3(mul3(n, t), p);
  d3 p1 = wp(W.p1), p2 = wp(W.p2), p3 = wp(W.p3), p4 = wp(W.p4);
  if (pw_in_plane(p1, p2, p3, p4, cp)) {
    *ovrlp = ov;
    *nv = mul3(n, copysign(1.0, dist));
    return true;
  }
  if (line_point_check(p1, p2, p, radius, nv, ovrlp)) return true;
  if (line_point_check(p2, p3, p, radius, nv, ovrlp)) return true;
  if (line_point_check(p3, p4, p, radius, nv, ovrlp)) return true;
  if (line_point_check(p4, p1, p, radius, nv, ovrlp)) return true;
  return false;
}

// ---------------------------------------------------------------------------
// grid helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_raw(const Grid& G, int d, double x) {
  return (int)floor((x - G.origin[d]) * G.inv_w[d]);
}
__device__ __forceinline__ int cell_clamped(const Grid& G, int d, double x) {
  int c = cell_raw(G, d, x);
  return c < 0 ? 0 : (c >= G.n[d] ? G.n[d] - 1 : c);
}
// key of cell (cx, cy, cz) of a level: its brick, then its place in the
// brick (x fastest), so that the cells of one brick row are consecutive keys
__device__ __forceinline__ int64_t brick_key(const Grid& G, int cx, int cy, int cz) {
  constexpr int M = BRICK - 1;
  const int64_t brick = ((int64_t)(cz >> BRICK_BITS) * G.nb[1] + (cy >> BRICK_BITS)) * G.nb[0] + (cx >> BRICK_BITS);
  return G.key_base + (brick << (3 * BRICK_BITS)) +
         ((((cz & M) << BRICK_BITS) | (cy & M)) << BRICK_BITS | (cx & M));
}
__device__ __forceinline__ int64_t cell_key(const Grid& G, double x, double y, double z) {
  return brick_key(G, cell_clamped(G, 0, x), cell_clamped(G, 1, y), cell_clamped(G, 2, z));
}

// ---------------------------------------------------------------------------
// step control + near-wall list: k_pre
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = fmax(r, sh[w]);
  return r;
}
__device__ __forceinline__ int block_sum(int v, int* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) r += sh[w];
  return r;
}
__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

// Near-wall list (UpdateNearPrtclsPW, Prtcl_ContactSearchPW.f90:159-235)
// of a state that no step has produced yet (after sdem_set_state or a
// replay): on an update iteration the near-wall bit mask of each particle is
// rebuilt and max|v|, max|a| are folded into the slots (one atomic per
// block, spread over NSLOT addresses).  Every later step gets the same from
// the k_force of the step before, which holds the new state in registers.
__device__ __forceinline__ uint64_t near_wall_mask(const Params& P, double x, double y, double z) {
  uint64_t mask = 0;
  for (int w = 0; w < P.nwall; w++) {
    const Wall& W = P.walls[w];
    const double mnx = W.min_point[0] - P.dxmax2, mny = W.min_point[1] - P.dxmax2;
    const double mnz = W.min_point[2] - P.dxmax2;
    const double mxx = W.max_point[0] + P.dxmax2, mxy = W.max_point[1] + P.dxmax2;
    const double mxz = W.max_point[2] + P.dxmax2;
    if (x >= mnx && x <= mxx && y >= mny && y <= mxy && z >= mnz && z <= mxz) mask |= (1ull << w);
  }
  return mask;
}

// The box of a block: boxes own whole BOX_ALIGN-slot groups, so the first
// slot of a block (of at most BOX_ALIGN threads) names it.
__device__ __forceinline__ int block_box(const int* __restrict__ blk_box) {
  return blk_box[(blockIdx.x * blockDim.x) / BOX_ALIGN];
}

// The workgroups of a launch go to the chip's NXCD compute dies in turn
// (block b to die b % NXCD), and each die has an L2 of its own.  The force
// kernel, the contact test and the list kernel number their tiles so that
// each die takes runs of XCD_RUN consecutive tiles: the partners a tile
// gathers (close in cell order) are then more often in the L2 of the die
// that gathers them, while all dies still work in one region of memory at
// a time (one contiguous range per die was slower: 8 streams far apart).
// A bijection of [0, gridDim.x); the last incomplete round of runs keeps
// the plain numbering.
constexpr int NXCD = 8, XCD_RUN = 16;
__device__ __forceinline__ int xcd_tile() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int full = (nb / (NXCD * XCD_RUN)) * (NXCD * XCD_RUN);
  if (b >= full) return b;
  const int x = b % NXCD, i = b / NXCD;
  return ((i / XCD_RUN) * NXCD + x) * XCD_RUN + i % XCD_RUN;
}

__global__ __launch_bounds__(256) void k_pre(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                             Buf A, Scalars* S, BoxState* BS, BoxSlot* bslots) {
  __shared__ double shd[4];
  const int box = block_box(blk_box);
  const Params& P = PB[box];
  if (S->halt || S->iter < BS[box].next_wall_update) return;  // max|v|, max|a| are taken even without walls
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double mv = 0.0, ma = 0.0;
  if (s - P.off < BS[box].n_alive && A.type[s] >= 1) {
    const double4 p = A.pos[s];
    A.wmask[s] = near_wall_mask(P, p.x, p.y, p.z);
    mv = norm3(ld3(A.lv, P.cap, s));
    ma = norm3(ld3(A.la, P.cap, s));  // stored level 0 == linAcc(1) at step start
  }
  mv = block_max(mv, shd);
  ma = block_max(ma, shd);
  if (threadIdx.x == 0) {
    BoxSlot& sl = bslots[box * NSLOT + blockIdx.x % NSLOT];
    atomicMax(&sl.max_v_bits, dbits(mv));
    atomicMax(&sl.max_a_bits, dbits(ma));
  }
}

// Per-box part of the step control, one wave: reduces the box's slots `rd`
// (the last step's largest move, deletions, max|v|, max|a|), clears the
// other set `wr` for this step's force kernel, decides whether the box is re-sorted and its candidate lists rebuilt
// (forced, a particle was removed, or the summed largest moves since the
// last rebuild may exceed half the skin: the exact contact test runs every
// step, so the decision never changes results), and runs the near-wall
// cadence (Prtcl_ContactSearchPW.f90:201-229) at iterNumber `iter`.
// Returns the decision on lane 0.
__device__ bool box_control(const Params& P, int iter, BoxState& B, const BoxSlot* rd, BoxSlot* wr,
                            bool forced) {
  const int lane = threadIdx.x & 63;
  const BoxSlot& m = rd[lane];
  double mv = bitsd(m.max_v_bits), ma = bitsd(m.max_a_bits), step = bitsd(m.step_bits);
  int ndel = m.ndel;
  for (int off = 32; off > 0; off >>= 1) {
    mv = fmax(mv, __shfl_down(mv, off, 64));
    ma = fmax(ma, __shfl_down(ma, off, 64));
    step = fmax(step, __shfl_down(step, off, 64));
    ndel += __shfl_down(ndel, off, 64);
  }
  BoxSlot& z = wr[lane];
  z.max_v_bits = 0ull;
  z.max_a_bits = 0ull;
  z.step_bits = 0ull;
  z.ndel = 0;
  for (int q = 0; q < 4; q++) z.cnt[q] = 0;
  if (lane != 0) return false;
  // |x_now - x_rebuild| <= sum of the steps' largest moves (triangle inequality)
  const double disp = B.disp_acc + step;
  const bool rebuild = forced || ndel > 0 || !(disp <= P.rebuild_disp);
  B.prev_rebuild = B.rebuild;
  B.rebuild = rebuild ? 1 : 0;
  if (rebuild) {
    // the list kernel lays the lists out anew: in cell order when nearly
    // every particle had candidates at the last rebuild, else packed
    const int listed = B.dense ? B.n_alive - B.n_unlisted : B.n_act + B.n_far;
    B.dense = (B.n_alive > 0 && listed >= B.n_alive - B.n_alive / 10) ? 1 : 0;
    B.n_act = B.n_far = B.n_unlisted = 0;
  }
  B.disp_acc = rebuild ? 0.0 : disp;
  const int next = B.next_wall_update;
  B.n_stored = B.n_alive;
  B.next_wall_at_begin = next;
  B.wall_prev = iter >= next ? 1 : 0;  // the last force kernel wrote fresh near-wall lists
  if (iter >= next) {
    double max_v = mv + P.max_wall_vel;
    max_v = max_v > 0.3 ? max_v : 0.3;
    max_v = max_v * 1.5;
    double max_a = B.use_host_max_a ? B.host_max_a : ma;
    const double gn = sqrt(P.g[0] * P.g[0] + P.g[1] * P.g[1] + P.g[2] * P.g[2]);
    max_a = max_a > gn ? max_a : gn;
    max_a = max_a * 1.5;
    // t is 0/0 (NaN) when nothing accelerates and there is no gravity.
    // int() of NaN, or of 2^31 or more, is the most negative integer with the
    // reference's compilers on x86, which max(., 1) turns into one step.
    const double t = (sqrt(max_v * max_v + 2.0 * max_a * P.dxmax1) - max_v) / max_a;
    const double q = t / P.dt;
    int num_iter = (q >= 1.0 && q < 2147483648.0) ? (int)q : 1;
    if (num_iter > P.wall_max_update_iter) num_iter = P.wall_max_update_iter;
    B.next_wall_update = iter + num_iter;
  }
  B.use_host_max_a = 0;
  return rebuild;
}

// Step control, one wave (lanes 0-63 of the calling block): advances
// iterNumber (Prtcl_DEMSystem.f90:248) and decides whether this step is
// forced to rebuild every box; with a single box it also runs that box's
// control.  After a capacity overflow it only marks the call halted, so
// every later kernel of the call does nothing and the failing step's input
// buffer stays intact for the replay; every halted step is counted so that
// the host knows how many steps completed.  `fast` (single box): the step
// was queued without the sort and list kernels; if the box must rebuild,
// the step halts the call before its force kernel (spec_fail), and the host
// replays it with them.  Returns on every lane whether the step goes ahead.
// The flags are read by lane 0 alone, which also writes them.
__device__ bool step_control(const Params* __restrict__ PB, Scalars* S, BoxState* BS, const BoxSlot* rd,
                             BoxSlot* wr, int nbox, int fast) {
  const int lane = threadIdx.x & 63;
  int flags = 0, iter = 0;
  if (lane == 0) {
    flags = (S->halt || S->overflow) ? 1 : 0;
    flags |= (S->need_rebuild || !S->have_parts) ? 2 : 0;
    iter = S->iter;
  }
  flags = __shfl(flags, 0, 64);
  if (flags & 1) {
    if (lane == 0) {
      S->halt = 1;
      S->halted_steps++;
    }
    return false;
  }
  const bool forced = (flags & 2) != 0;
  if (fast) {
    // box_control's rebuild decision, read-only
    const BoxSlot& m = rd[lane];
    double step = bitsd(m.step_bits);
    int ndel = m.ndel;
    for (int off = 32; off > 0; off >>= 1) {
      step = fmax(step, __shfl_down(step, off, 64));
      ndel += __shfl_down(ndel, off, 64);
    }
    int need = 0;
    if (lane == 0) need = (forced || ndel > 0 || !(BS[0].disp_acc + step <= PB[0].rebuild_disp)) ? 1 : 0;
    if (__shfl(need, 0, 64)) {
      if (lane == 0) {
        S->halt = 1;
        S->spec_fail = 1;
        S->halted_steps++;
      }
      return false;
    }
  }
  if (lane == 0) {
    S->forced = forced ? 1 : 0;
    S->need_rebuild = 0;
    S->have_parts = 0;
    S->rebuild = 0;  // the boxes' decisions are folded in below / by k_box
    S->iter_at_begin = iter;
    S->iter = iter + 1;  // iterNumber advances once per step
  }
  if (nbox == 1) {
    const bool r = box_control(PB[0], iter, BS[0], rd, wr, forced);
    if (lane == 0 && r) {
      S->rebuild = 1;
      S->rebuilds++;
    }
  }
  return true;
}

// The control of a step queued with every kernel (the fast steps run it in
// k_contact).
__global__ __launch_bounds__(64) void k_ctl(const Params* __restrict__ PB, Scalars* S, BoxState* BS,
                                            const BoxSlot* rd, BoxSlot* wr, int nbox) {
  (void)step_control(PB, S, BS, rd, wr, nbox, 0);
}

// Per-box step control of an ensemble, one wave per box, after k_ctl.
__global__ __launch_bounds__(64) void k_box(const Params* __restrict__ PB, Scalars* S, BoxState* BS,
                                            const BoxSlot* rd, BoxSlot* wr) {
  if (S->halt) return;
  const int box = blockIdx.x;
  const bool r = box_control(PB[box], S->iter_at_begin, BS[box], rd + (size_t)box * NSLOT,
                             wr + (size_t)box * NSLOT, S->forced != 0);
  if (threadIdx.x == 0 && r) {
    atomicOr(&S->rebuild, 1);
    atomicAdd(&S->rebuilds, 1);
  }
}

// counters of the last step for the host (once per call), one wave per box,
// from the slot set of the buffer the last step wrote; a single box also
// leaves its totals beside the scalars the host reads after every batch
__global__ __launch_bounds__(64) void k_finish(Scalars* S, BoxState* BS, const BoxSlot* bslots) {
  if (S->halt || !S->have_parts) return;
  const int box = blockIdx.x;
  const int lane = threadIdx.x;
  const BoxSlot& m = bslots[(size_t)box * NSLOT + lane];
  int c[4], nd = m.ndel;
  for (int q = 0; q < 4; q++) c[q] = m.cnt[q];
  for (int off = 32; off > 0; off >>= 1) {
    for (int q = 0; q < 4; q++) c[q] += __shfl_down(c[q], off, 64);
    nd += __shfl_down(nd, off, 64);
  }
  if (lane == 0) {
    for (int q = 0; q < 4; q++) BS[box].num_cntcts[q] = c[q];
    BS[box].n_deleted = nd;
    if (gridDim.x == 1) {
      S->fin_records = c[0] + c[1] + c[2] + c[3];
      S->fin_alive = BS[box].n_alive;
    }
  }
}

// ---------------------------------------------------------------------------
// cell sort (rebuild steps only)
// ---------------------------------------------------------------------------
// (level, cell) key and per-key counts of the boxes that rebuild; dead
// particles go to the box's last key.  Lanes of a wave holding the same key
// (neighbours in storage order usually share a cell) take their slots with
// one atomic per run.  The counts array is all zero on entry (zeroed at
// allocation and by k_scan_add).  The scan then runs over every key; the
// boxes that do not rebuild contribute no counts, so a box's offsets are
// taken relative to its first key (box_start).
__device__ __forceinline__ int box_start_base(const Params& P, const int* __restrict__ start) {
  return start[P.key_first] - P.off;  // start[key] - base = storage slot of the key's first particle
}

__global__ __launch_bounds__(256) void k_bin(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                             Buf A, const Scalars* S, const BoxState* BS,
                                             int* __restrict__ keys, int* __restrict__ rank,
                                             int* __restrict__ counts) {
  if (!S->rebuild || S->halt) return;
  const int box = block_box(blk_box);
  if (!BS[box].rebuild) return;
  const Params& P = PB[box];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  // every slot of the box takes part: removed and unused ones go to the
  // box's last key, so the box keeps its slot range
  int key = P.dead_key;
  if (s - P.off < BS[box].n_stored) {
    const int t = A.type[s];
    if (t >= 1) {
      const Grid& G = P.lvl[P.type_level[t - 1]];
      const double4 p = A.pos[s];
      key = (int)cell_key(G, p.x, p.y, p.z);
    }
  }
  keys[s] = key;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = (lane == 0) || (prev != key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  const int my_head = 63 - __clzll((long long)(heads & le));
  const unsigned long long after = heads & ~le;
  const int next_head = after ? (__ffsll((long long)after) - 1) : 64;
  int base = 0;
  if (head) base = atomicAdd(&counts[key], next_head - lane);
  base = __shfl(base, my_head, 64);
  rank[s] = base + (lane - my_head);
}

// Exclusive scan of the counts in three kernels: tile scans, the scan of the
// tile sums, and the tile offsets added (while the counts are zeroed for the
// next rebuild).  A tile whose keys all belong to boxes that do not rebuild
// holds only zero counts and nobody reads its offsets: it is skipped (its
// sum is zero).
constexpr int SCAN_T = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_T * SCAN_ITEMS;

__device__ __forceinline__ bool tile_active(const int2* __restrict__ tile_box, const BoxState* BS, int t) {
  const int2 r = tile_box[t];
  for (int b = r.x; b <= r.y; b++)
    if (BS[b].rebuild) return true;
  return false;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                        int* __restrict__ tile_sum, int64_t n,
                                                        const Scalars* S, const BoxState* BS,
                                                        const int2* __restrict__ tile_box) {
  if (!S->rebuild || S->halt) return;
  if (!tile_active(tile_box, BS, blockIdx.x)) {
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = 0;
    return;
  }
  using Scan = hipcub::BlockScan<int, SCAN_T>;
  __shared__ typename Scan::TempStorage tmp;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  for (int i = 0; i < SCAN_ITEMS; i++) v[i] = (base + i < n) ? in[base + i] : 0;
  int total;
  Scan(tmp).ExclusiveSum(v, v, total);
  for (int i = 0; i < SCAN_ITEMS; i++)
    if (base + i < n) out[base + i] = v[i];
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// exclusive scan of the tile sums, one block (a kernel of its own: a
// last-block election would need a device-wide fence in every tile block)
__global__ __launch_bounds__(SCAN_T) void k_scan_sums(int* __restrict__ tile_sum, int ntiles, const Scalars* S) {
  if (!S->rebuild || S->halt) return;
  using BS = hipcub::BlockScan<int, SCAN_T>;
  __shared__ typename BS::TempStorage tmp;
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int start = 0; start < ntiles; start += SCAN_TILE) {
    const int b0 = start + threadIdx.x * SCAN_ITEMS;
    int w[SCAN_ITEMS];
    for (int i = 0; i < SCAN_ITEMS; i++) w[i] = (b0 + i < ntiles) ? tile_sum[b0 + i] : 0;
    int tot;
    BS(tmp).ExclusiveSum(w, w, tot);
    const int c0 = carry;
    for (int i = 0; i < SCAN_ITEMS; i++)
      if (b0 + i < ntiles) tile_sum[b0 + i] = w[i] + c0;
    __syncthreads();
    if (threadIdx.x == 0) carry = c0 + tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_add(int* __restrict__ in, int* __restrict__ out,
                                                      const int* __restrict__ tile_sum, int64_t n,
                                                      const Scalars* S, const BoxState* BS,
                                                      const int2* __restrict__ tile_box) {
  if (!S->rebuild || S->halt) return;
  if (!tile_active(tile_box, BS, blockIdx.x)) return;
  const int add = tile_sum[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_T)
    if (base + i < n) {
      out[base + i] += add;
      in[base + i] = 0;
    }
}

__global__ __launch_bounds__(256) void k_scatter(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                                 const Scalars* S, const BoxState* BS,
                                                 const int* __restrict__ keys, const int* __restrict__ rank,
                                                 const int* __restrict__ start, int* __restrict__ perm_tmp) {
  if (!S->rebuild || S->halt) return;
  const int box = block_box(blk_box);
  if (!BS[box].rebuild) return;
  const int base = box_start_base(PB[box], start);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  perm_tmp[start[keys[s]] - base + rank[s]] = s;
}

// Make the in-cell order the storage order (deterministic) and pack the
// positions and ids in cell order.  A box's live particles are the slots before its
// last key: n_alive = start[dead_key] - off.  Removed and unused slots keep
// the order the counting sort gave them; nothing reads them.
__global__ __launch_bounds__(256) void k_rank(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                              Buf A, const Scalars* S, BoxState* BS,
                                              const int* __restrict__ keys,
                                              const int* __restrict__ start,
                                              const int* __restrict__ perm_tmp,
                                              int* __restrict__ perm, double4* __restrict__ pk,
                                              int* __restrict__ pid) {
  if (!S->rebuild || S->halt) return;
  const int box = block_box(blk_box);
  if (!BS[box].rebuild) return;
  const Params& P = PB[box];
  const int base = box_start_base(P, start);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == P.off) BS[box].n_alive = start[P.dead_key] - base - P.off;
  const int s = perm_tmp[k];
  const int key = keys[s];
  if (key == P.dead_key) {
    perm[k] = s;
    return;
  }
  const int a = start[key] - base, b = start[key + 1] - base;
  int r = 0;
  for (int j = a; j < b; j++) r += (perm_tmp[j] < s);
  perm[a + r] = s;
  pk[a + r] = A.pos[s];
  pid[a + r] = A.id[s];
}

// ---------------------------------------------------------------------------
// candidate (Verlet) lists, rebuild steps only
// ---------------------------------------------------------------------------
// Contact-list entries: index of the partner in cell order (moving) or in the
// fixed-particle arrays, with a flag for fixed partners.
constexpr unsigned CL_INDEX = 0x3fffffffu, CL_FIXED = 0x40000000u;

// Every particle lists the particles j (any level, periodic images included)
// with ((dx*dx + dy*dy) + dz*dz) <= (ri+rj+skin)^2 over the cells of each
// level overlapping [x - reach, x + reach], reach = ri + (largest radius of
// the level) + skin, and, with the plain search (CS_Method 1), the fixed
// particles of the fixed grid's cells within the same distance (the
// hierarchical search does not look for fixed particles,
// Prtcl_Hrchl_Munjiza.f90:705-755).  The list is sorted by partner id (2 id,
// +1 for a fixed partner), so that a particle's contact forces are summed in
// an order that depends neither on the grid, the skin nor the storage order.
// Entry q of particle k sits at vlist[q*cap + k] (lanes of a wave read
// consecutive words); vkey holds the sort keys while a list longer than
// the LDS part is sorted.

// A particle's list is appended to while the cells are walked (entries
// only) and sorted once at the end, after its keys are loaded.  Its first
// VL_LDS entries sit in LDS (entry q of thread tid at q*256 + tid: a wave's
// lanes hit consecutive banks), the rest in vlist/vkey.  One 256-thread
// block holds 24 KB of them.
constexpr int VL_LDS = 12, VL_T = 256;
struct VlStage {
  int key[VL_LDS * VL_T];
  int ent[VL_LDS * VL_T];
};

__device__ __forceinline__ void vlist_append(VlStage& L, int tid, int* __restrict__ vlist, size_t cap, int k,
                                             int lcap, int& c, unsigned entry) {
  if (c < lcap) {
    if (c < VL_LDS) L.ent[c * VL_T + tid] = (int)entry;
    else vlist[(size_t)c * cap + k] = (int)entry;
  }
  c++;
}

// sort key of a list entry: 2 id of a moving partner, 2 id + 1 of a fixed one
__device__ __forceinline__ int entry_key(unsigned e, const int* __restrict__ pid, const int* __restrict__ fid) {
  return (e & CL_FIXED) ? 2 * fid[e & CL_INDEX] + 1 : 2 * pid[e & CL_INDEX];
}

// The keys of a particle's nq <= lcap entries (loaded four at a time), then
// an insertion sort by key (stable; equal keys only come from one partner
// seen through two periodic images, whose entries are equal).
__device__ __forceinline__ void vlist_sort(VlStage& L, int tid, int* __restrict__ vlist, int* __restrict__ vkey,
                                           size_t cap, int k, int nq, const int* __restrict__ pid,
                                           const int* __restrict__ fid) {
  if (nq <= VL_LDS) {
    for (int q = 0; q < nq; q += 4) {
      int kk[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (q + u < nq) kk[u] = entry_key((unsigned)L.ent[(q + u) * VL_T + tid], pid, fid);
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (q + u < nq) L.key[(q + u) * VL_T + tid] = kk[u];
    }
    for (int q = 1; q < nq; q++) {
      const int key = L.key[q * VL_T + tid], ent = L.ent[q * VL_T + tid];
      int b = q - 1;
      while (b >= 0) {
        const int kb = L.key[b * VL_T + tid];
        if (kb <= key) break;
        L.key[(b + 1) * VL_T + tid] = kb;
        L.ent[(b + 1) * VL_T + tid] = L.ent[b * VL_T + tid];
        b--;
      }
      L.key[(b + 1) * VL_T + tid] = key;
      L.ent[(b + 1) * VL_T + tid] = ent;
    }
    return;
  }
  for (int q = 0; q < VL_LDS; q++) vlist[(size_t)q * cap + k] = L.ent[q * VL_T + tid];
  for (int q = 0; q < nq; q += 4) {
    int kk[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (q + u < nq) kk[u] = entry_key((unsigned)vlist[(size_t)(q + u) * cap + k], pid, fid);
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (q + u < nq) vkey[(size_t)(q + u) * cap + k] = kk[u];
  }
  for (int q = 1; q < nq; q++) {
    const int key = vkey[(size_t)q * cap + k], ent = vlist[(size_t)q * cap + k];
    int b = q - 1;
    while (b >= 0) {
      const int kb = vkey[(size_t)b * cap + k];
      if (kb <= key) break;
      vkey[(size_t)(b + 1) * cap + k] = kb;
      vlist[(size_t)(b + 1) * cap + k] = vlist[(size_t)b * cap + k];
      b--;
    }
    vkey[(size_t)(b + 1) * cap + k] = key;
    vlist[(size_t)(b + 1) * cap + k] = ent;
  }
}

// The cells a particle searches on a grid, [lo, hi] per axis: those
// overlapping [x - reach, x + reach], clamped to the grid (periodic axes to
// one period either side).
__device__ __forceinline__ void search_range(const Params& P, const Grid& G, const double xv[3], double reach,
                                             int lo[3], int hi[3]) {
  for (int d = 0; d < 3; d++) {
    const double pad = 1.0e-6 * G.w[d];
    lo[d] = cell_raw(G, d, xv[d] - reach - pad);
    hi[d] = cell_raw(G, d, xv[d] + reach + pad);
    if (!P.periodic[d]) {
      lo[d] = lo[d] < 0 ? 0 : lo[d];
      hi[d] = hi[d] >= G.n[d] ? G.n[d] - 1 : hi[d];
    } else {
      lo[d] = lo[d] < -G.n[d] ? -G.n[d] : lo[d];
      hi[d] = hi[d] > 2 * G.n[d] - 1 ? 2 * G.n[d] - 1 : hi[d];
    }
  }
}

// One candidate: the reference's predicate with the skin, on the image
// coordinates; the particle itself (no image) is left out.
__device__ __forceinline__ void vl_test(const Params& P, const double4& pi, int k, const double4& pj, double shx,
                                        double shy, double shz, unsigned entry, bool self_ok, VlStage& L, int tid,
                                        int* __restrict__ vlist, size_t cap, int lcap, int& c, double& gmin) {
  const double ri = pi.w;
  const double ddx = pi.x - (pj.x + shx);
  const double ddy = pi.y - (pj.y + shy);
  const double ddz = pi.z - (pj.z + shz);
  const double dr = ri + pj.w + P.skin;
  const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
  if (d2 > dr * dr) return;
  const bool shifted = (shx != 0.0) || (shy != 0.0) || (shz != 0.0);
  if (!self_ok && entry == (unsigned)k && !shifted) return;
  gmin = fmin(gmin, sqrt(d2) - (ri + pj.w));
  vlist_append(L, tid, vlist, cap, k, lcap, c, entry);
}

// A particle's walk over one level's cells.  The walk is a chain of
// dependent loads (a run's bounds, then its particles), so it is laid out
// to keep several in flight: the x extent of the range splits into the
// same runs of consecutive keys on every row (a row's cells within one
// brick and one period), the bounds of a row's first three runs are loaded
// together, one row ahead, and the row's particles four at a time.  Runs
// beyond the third (ranges wider than three bricks) are walked one by one.
__device__ __forceinline__ void vl_level(const Params& P, const Grid& G, const double4& pi, int k,
                                         const int* __restrict__ cstart, int base, const double4* __restrict__ pk,
                                         VlStage& L, int tid, int* __restrict__ vlist, size_t cap, int lcap, int& c,
                                         double& gmin) {
  const double xv[3] = {pi.x, pi.y, pi.z};
  int lo[3], hi[3];
  search_range(P, G, xv, pi.w + G.rmax + P.skin, lo, hi);
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
  const int n0 = G.n[0];
  // x runs: key offsets of their first and last cells, the image shift
  int a0 = 0, b0 = -1, a1 = 0, b1 = -1, a2 = 0, b2 = -1;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  int x = lo[0];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (x <= hi[0]) {
      int wx = x;
      double s = 0.0;
      if (wx < 0) { wx += n0; s = -P.simlen[0]; }
      else if (wx >= n0) { wx -= n0; s = P.simlen[0]; }
      const int wend = min(wx | (BRICK - 1), n0 - 1);
      const int x1 = min(hi[0], x + (wend - wx));
      const int wx1 = wx + (x1 - x);
      const int ka = ((wx >> BRICK_BITS) << (3 * BRICK_BITS)) | (wx & (BRICK - 1));
      const int kb = ((wx1 >> BRICK_BITS) << (3 * BRICK_BITS)) | (wx1 & (BRICK - 1));
      if (i == 0) { a0 = ka; b0 = kb; s0 = s; }
      else if (i == 1) { a1 = ka; b1 = kb; s1 = s; }
      else { a2 = ka; b2 = kb; s2 = s; }
      x = x1 + 1;
    }
  }
  const int xrest = x;  // first cell of the runs walked one by one
  const int nrow = (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  // bounds of row (cy, cz): first index and length of each of the three
  // runs; the rows are walked y fastest
  int ja0, ja1, ja2, n0r, n1r, n2r, wy, wz;
  double shy, shz;
  auto row_at = [&](int cy, int cz, int& j0, int& j1, int& j2, int& m0, int& m1, int& m2, int& ry, int& rz,
                    double& sy, double& sz) {
    sy = 0.0;
    sz = 0.0;
    if (cy < 0) { cy += G.n[1]; sy = -P.simlen[1]; }
    else if (cy >= G.n[1]) { cy -= G.n[1]; sy = P.simlen[1]; }
    if (cz < 0) { cz += G.n[2]; sz = -P.simlen[2]; }
    else if (cz >= G.n[2]) { cz -= G.n[2]; sz = P.simlen[2]; }
    ry = cy;
    rz = cz;
    const int64_t row = G.key_base +
                        ((((int64_t)(cz >> BRICK_BITS) * G.nb[1] + (cy >> BRICK_BITS)) * G.nb[0]) << (3 * BRICK_BITS)) +
                        ((((cz & (BRICK - 1)) << BRICK_BITS) | (cy & (BRICK - 1))) << BRICK_BITS);
    const int* r = cstart + row;
    j0 = r[a0] - base;
    m0 = r[b0 + 1] - base - j0;
    j1 = 0; m1 = 0; j2 = 0; m2 = 0;
    if (b1 >= 0) {
      j1 = r[a1] - base;
      m1 = r[b1 + 1] - base - j1;
    }
    if (b2 >= 0) {
      j2 = r[a2] - base;
      m2 = r[b2 + 1] - base - j2;
    }
  };
  int cy = lo[1], cz = lo[2];
  row_at(cy, cz, ja0, ja1, ja2, n0r, n1r, n2r, wy, wz, shy, shz);
  for (int rho = 0; rho < nrow; rho++) {
    if (++cy > hi[1]) { cy = lo[1]; cz++; }
    int na0 = 0, na1 = 0, na2 = 0, nm0 = 0, nm1 = 0, nm2 = 0, nwy = 0, nwz = 0;
    double nsy = 0.0, nsz = 0.0;
    if (rho + 1 < nrow) row_at(cy, cz, na0, na1, na2, nm0, nm1, nm2, nwy, nwz, nsy, nsz);
    const int n01 = n0r + n1r, ntot = n01 + n2r;
    for (int t = 0; t < ntot; t += 4) {
      double4 q[4];
      int kq[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int tt = t + u;
        kq[u] = tt < n0r ? ja0 + tt : (tt < n01 ? ja1 + (tt - n0r) : ja2 + (tt - n01));
        if (tt < ntot) q[u] = pk[kq[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int tt = t + u;
        if (tt < ntot) {
          const double sx = tt < n0r ? s0 : (tt < n01 ? s1 : s2);
          vl_test(P, pi, k, q[u], sx, shy, shz, (unsigned)kq[u], false, L, tid, vlist, cap, lcap, c, gmin);
        }
      }
    }
    // runs past the third, one by one
    for (int x0 = xrest; x0 <= hi[0];) {
      int wx = x0;
      double sx = 0.0;
      if (wx < 0) { wx += n0; sx = -P.simlen[0]; }
      else if (wx >= n0) { wx -= n0; sx = P.simlen[0]; }
      const int wend = min(wx | (BRICK - 1), n0 - 1);
      const int x1 = min(hi[0], x0 + (wend - wx));
      const int ja = cstart[brick_key(G, wx, wy, wz)] - base;
      const int jb = cstart[brick_key(G, wx + (x1 - x0), wy, wz) + 1] - base;
      for (int kj = ja; kj < jb; kj++)
        vl_test(P, pi, k, pk[kj], sx, shy, shz, (unsigned)kj, false, L, tid, vlist, cap, lcap, c, gmin);
      x0 = x1 + 1;
    }
    ja0 = na0; ja1 = na1; ja2 = na2;
    n0r = nm0; n1r = nm1; n2r = nm2;
    wy = nwy; wz = nwz;
    shy = nsy; shz = nsz;
  }
}

__global__ __launch_bounds__(256) void k_vlist(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                               Scalars* S, BoxState* BS, const double4* __restrict__ pk,
                                               const int* __restrict__ pid, const int* __restrict__ cstart,
                                               int* __restrict__ vlist,
                                               int* __restrict__ vkey, int4* __restrict__ act,
                                               int* __restrict__ cand, int* __restrict__ ccnt,
                                               const double4* __restrict__ fpk,
                                               const int* __restrict__ fid, const int* __restrict__ fstart) {
  if (!S->rebuild || S->halt) return;
  const int tile = xcd_tile();
  const int box = blk_box[(tile * blockDim.x) / BOX_ALIGN];
  if (!BS[box].rebuild) return;
  const Params& P = PB[box];
  __shared__ VlStage stage;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int k = tile * blockDim.x + threadIdx.x;
  const bool live = k - P.off < BS[box].n_alive;
  const bool dense = BS[box].dense != 0;  // read before this kernel's atomics on the box's state
  const size_t cap = P.cap;
  const int lcap = P.lcap;
  int c = 0;
  double gmin = 1.0e300;  // smallest gap to a candidate, |xi - xj| - (ri + rj)
  double ri = 0.0;
  if (live) {
  const int base = box_start_base(P, cstart);
  const double4 pi = pk[k];
  const double xv[3] = {pi.x, pi.y, pi.z};
  ri = pi.w;
  // The cell walk wraps a periodic dimension by one period at most: a
  // coordinate that is not finite, or that lies more than a period outside
  // the domain (a particle that crossed it in one step), is refused here
  // with an error instead of being read past the grid.
  bool bad = false;
  for (int d = 0; d < 3; d++)
    if (!isfinite(xv[d]) || (P.periodic[d] && !(xv[d] >= P.dmin[d] - P.simlen[d] &&
                                                xv[d] <= P.dmax[d] + P.simlen[d])))
      bad = true;
  if (bad) atomicMax(&S->overflow, 3);
  const int nlvl = bad ? 0 : P.nlvl;
  for (int lv = 0; lv < nlvl; lv++)
    vl_level(P, P.lvl[lv], pi, k, cstart, base, pk, stage, tid, vlist, cap, lcap, c, gmin);
  if (!bad && P.search_fixed && P.nfix > 0) {
    const Grid& G = P.fgrid;
    int lo[3], hi[3];
    search_range(P, G, xv, ri + G.rmax + P.skin, lo, hi);
    for (int cz = lo[2]; cz <= hi[2]; cz++) {
      int wz = cz;
      double shz = 0.0;
      if (wz < 0) { wz += G.n[2]; shz = -P.simlen[2]; }
      else if (wz >= G.n[2]) { wz -= G.n[2]; shz = P.simlen[2]; }
      for (int cy = lo[1]; cy <= hi[1]; cy++) {
        int wy = cy;
        double shy = 0.0;
        if (wy < 0) { wy += G.n[1]; shy = -P.simlen[1]; }
        else if (wy >= G.n[1]) { wy -= G.n[1]; shy = P.simlen[1]; }
        for (int cx = lo[0]; cx <= hi[0]; cx++) {
          int wx = cx;
          double shx = 0.0;
          if (wx < 0) { wx += G.n[0]; shx = -P.simlen[0]; }
          else if (wx >= G.n[0]) { wx -= G.n[0]; shx = P.simlen[0]; }
          const int64_t key = ((int64_t)wz * G.n[1] + wy) * G.n[0] + wx;
          for (int kj = fstart[key]; kj < fstart[key + 1]; kj++)
            vl_test(P, pi, k, fpk[kj], shx, shy, shz, (unsigned)kj | CL_FIXED, true, stage, tid, vlist, cap,
                    lcap, c, gmin);
        }
      }
    }
  }
  if (c > lcap) {
    atomicMax(&S->overflow, 2);
    atomicMax(&S->list_need, c);
  }
  vlist_sort(stage, tid, vlist, vkey, cap, k, c < lcap ? c : lcap, pid, fid);
  }  // live
  const double ri_live = ri;
  // The contact test runs over slots t with (particle, count, smallest gap)
  // in act[t] and the entries in cand[q*cap + t]; the layout does not enter
  // any result.  Packed (most boxes): the particles with candidates only,
  // one atomic per wave, those whose smallest gap is under half the skin
  // from slot P.off up, the others from the box's last slot down, so that
  // the waves of the latter skip as a whole until the box has moved a
  // quarter of the skin.  Dense (nearly every particle had candidates at
  // the last rebuild): slot t = k for every live particle, so that the
  // contact test walks the particles in cell order (a particle without
  // candidates has count 0 and an infinite gap).  The particles without
  // candidates have no contact until the next rebuild.
  const bool has = live && c > 0;
  // the gap less a margin for the roundings of this and the contact test
  // (both are within a few units in the last place of ri + rj)
  const double gstore = gmin - 1.0e-9 * ri_live;
  if (live && !has) ccnt[k] = 0;
  if (dense) {
    // the live particles without candidates are counted (few in a dense box)
    const unsigned long long bu = __ballot(live && !has);
    if (bu && lane == __ffsll((long long)bu) - 1) atomicAdd(&BS[box].n_unlisted, __popcll(bu));
    if (k == P.off) BS[box].n_act = BS[box].n_alive;
    if (!live) return;
    act[k] = make_int4(k, c, __double2loint(gstore), __double2hiint(gstore));
    const int nq = c < lcap ? c : lcap;
    if (c <= VL_LDS || lcap <= VL_LDS)  // the list never left LDS
      for (int q = 0; q < nq; q++) cand[(size_t)q * cap + k] = stage.ent[q * VL_T + tid];
    else
      for (int q = 0; q < nq; q++) cand[(size_t)q * cap + k] = vlist[(size_t)q * cap + k];
    return;
  }
  const bool far = has && gstore > 0.5 * P.skin, near = has && !far;
  const unsigned long long bn = __ballot(near), bfar = __ballot(far);
  if (!(bn | bfar)) return;
  const int ln = bn ? __ffsll((long long)bn) - 1 : 0, lf = bfar ? __ffsll((long long)bfar) - 1 : 0;
  int tn = 0, tf = 0;
  if (bn && lane == ln) tn = atomicAdd(&BS[box].n_act, __popcll(bn));
  if (bfar && lane == lf) tf = atomicAdd(&BS[box].n_far, __popcll(bfar));
  tn = __shfl(tn, ln, 64);
  tf = __shfl(tf, lf, 64);
  if (!has) return;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int t = near ? P.off + tn + __popcll(bn & below) : P.off + P.bcap - 1 - (tf + __popcll(bfar & below));
  act[t] = make_int4(k, c, __double2loint(gstore), __double2hiint(gstore));
  const int nq = c < lcap ? c : lcap;
  if (c <= VL_LDS || lcap <= VL_LDS)  // the list never left LDS
    for (int q = 0; q < nq; q++) cand[(size_t)q * cap + t] = stage.ent[q * VL_T + tid];
  else
    for (int q = 0; q < nq; q++) cand[(size_t)q * cap + t] = vlist[(size_t)q * cap + k];
}

// ---------------------------------------------------------------------------
// the per-step force kernel
// ---------------------------------------------------------------------------
struct Hist {
  d3 td;
  double vri;
};

// The row of particle b's history (values at b, b + cap, ...) that holds
// the record against `key`, the first match in record order, or NO_ROW:
// the count and the first KEY_HEAD keys (one int4) are loaded by the
// caller, in one batch (their addresses depend on b only), the rest one by
// one.
constexpr unsigned NO_ROW = 0xffffffffu;
struct HeadKeys {
  int n;
  int4 key;
};

__device__ __forceinline__ HeadKeys head_keys(const Buf& A, int b) {
  HeadKeys h;
  h.n = A.hcnt[b];
  h.key = reinterpret_cast<const int4*>(A.hkey)[b];
  return h;
}

__device__ __forceinline__ unsigned hist_row(const Buf& A, size_t cap, int b, const HeadKeys& h, int key) {
  if (h.n > 0 && (h.key.x & KEY_MASK) == key) return 0u;
  if (h.n > 1 && (h.key.y & KEY_MASK) == key) return 1u;
  if (h.n > 2 && (h.key.z & KEY_MASK) == key) return 2u;
  if (h.n > 3 && (h.key.w & KEY_MASK) == key) return 3u;
  for (int r = KEY_HEAD; r < h.n; r++)
    if ((A.hkey[key_slot(r, b, cap)] & KEY_MASK) == key) return (unsigned)r;
  return NO_ROW;
}

// Lookups in the particle's own list as the reference's bucket holds it
// (PP copies, KIND_COPY, aside).  A particle's records are written in the
// order of its contacts, so the record of its contact number `hint` is
// usually in row `hint`: that row is tried first, then the whole list.
__device__ __forceinline__ bool key_is(int k, int key) { return (k & KEY_MASK) == key; }
__device__ __forceinline__ bool is_copy(int k) { return ((unsigned)k >> KEY_BITS) == (unsigned)KIND_COPY; }
__device__ __forceinline__ int key_at(const Buf& A, size_t cap, int b, const HeadKeys& h, int r) {
  if (r < KEY_HEAD) return r == 0 ? h.key.x : (r == 1 ? h.key.y : (r == 2 ? h.key.z : h.key.w));
  return A.hkey[key_slot(r, b, cap)];
}
// The first record against `key` that is not a copy, or NO_ROW (with
// `copy`: the first record against `key`, if it is a copy, else NO_ROW).
__device__ __forceinline__ unsigned own_row(const Buf& A, size_t cap, int b, const HeadKeys& h, int key, int hint,
                                            bool copy) {
  if (hint < h.n) {
    const int k = key_at(A, cap, b, h, hint);
    if (key_is(k, key) && is_copy(k) == copy) return (unsigned)hint;
  }
#pragma unroll
  for (int r = 0; r < KEY_HEAD; r++) {
    const int k = key_at(A, cap, b, h, r);
    if (r < h.n && key_is(k, key) && (copy || !is_copy(k))) return is_copy(k) == copy ? (unsigned)r : NO_ROW;
  }
  for (int r = KEY_HEAD; r < h.n; r++) {
    const int k = A.hkey[key_slot(r, b, cap)];
    if (key_is(k, key) && (copy || !is_copy(k))) return is_copy(k) == copy ? (unsigned)r : NO_ROW;
  }
  return NO_ROW;
}

__device__ __forceinline__ Hist hist_at(const Buf& A, size_t cap, int b, unsigned row) {
  Hist h;
  h.td = D3(0.0, 0.0, 0.0);
  h.vri = -2.0;
  if (row != NO_ROW) {
    const double4 v = A.hval[(size_t)row * cap + b];
    h.td = D3(v.x, v.y, v.z);
    h.vri = v.w;
  }
  return h;
}

// A contact-list row names its bucket as well: ROW_PARTNER set = the record
// sits in the partner's bucket (a PP pair whose partner has the lower id),
// so the force kernel loads it together with the partner's state instead
// of after the partner's id.  NO_ROW (all bits) stays NO_ROW.
constexpr unsigned ROW_PARTNER = 0x80000000u;
__device__ __forceinline__ Hist hist_of_entry(const Buf& A, size_t cap, int s, int j, unsigned row) {
  if (row == NO_ROW) return hist_at(A, cap, s, NO_ROW);
  return hist_at(A, cap, (row & ROW_PARTNER) ? j : s, row & ~ROW_PARTNER);
}

struct Acc {
  d3 F, T;
  int hc;
  unsigned long long nkind;  // contacts by kind, 16 bits each (kind 1 lowest)
  int overflow;
};

__device__ __forceinline__ void push_hist(const Params& P, const Buf& B, int k, Acc& acc, int kind,
                                          int partner, const CF& cf) {
  if (acc.hc < P.hcap) {
    const size_t slot = (size_t)acc.hc * P.cap + k;
    B.hkey[key_slot(acc.hc, k, P.cap)] = partner | (kind << KEY_BITS);
    B.hval[slot] = make_double4(cf.tandel.x, cf.tandel.y, cf.tandel.z, cf.vri);
    acc.hc++;
  } else {
    acc.overflow = 1;
  }
  if (kind != KIND_COPY) acc.nkind += 1ull << (16 * (kind - 1));
}

// image of a partner coordinate: the reference's ghost copies hold
// fl(x + L) or fl(x - L) (Prtcl_Comm.f90:242,252); the image within half a
// period of the particle is the only one that can touch it
__device__ __forceinline__ double image(double xi, double xj, double L, int periodic, int* shifted) {
  if (periodic) {
    const double dx = xi - xj;
    if (dx > 0.5 * L) { *shifted = 1; return xj + L; }
    if (dx < -0.5 * L) { *shifted = 1; return xj + (-L); }
  }
  return xj;
}

// Contact test of one step: every particle walks its candidate list (sorted
// by partner id at the rebuild) and keeps the partners that touch it under
// the reference's predicate ((dx*dx + dy*dy) + dz*dz <= (ri+rj)^2 with
// early-outs, NBSM_FineSearch1/2/3, Prtcl_NBS_Munjiza.f90:478-557) on the
// image coordinates, so the contact list comes out sorted as well.  Entries:
// the partner's index in buffer A (moving) or in the fixed arrays, with the
// fixed flag, and a flag for a periodic image, each with the row of last
// step's history that holds the pair's record (the lower id's bucket for
// two moving particles side by side, the particle's own otherwise: the
// buckets the force routine reads, ContactForce_PP :732-756, _PPG
// :761-798, _PPFix :803-827), so that the force kernel loads the record
// directly.  A light kernel (few registers), so that many waves hide the
// gather latency, the history search's chain of loads included.
//
// FAST (single box, a step queued without the sort and list kernels; an
// instantiation of its own, so that the other one compiles as before):
// the kernel also runs the step's control in the first wave of block 0,
// which saves a dependent launch per step.  No other wave reads what that
// wave writes: the lists are those of the last rebuild (a fast step never
// rebuilds: if the box must, the control halts the call and the host
// replays the step; what the other waves wrote meanwhile is rewritten by
// the replay), and each wave adds the last step's largest move (slot set
// rd, which the control leaves alone) to the summed moves itself.  It
// may read the summed moves before or after the control wave has added
// the same move to them; after, it adds it twice, which only makes the gap
// test below skip less.
constexpr unsigned CL_SHIFT = 0x80000000u;

template <int FAST>
__global__ __launch_bounds__(256) void k_contact(const Params* __restrict__ PB,
                                                 const int* __restrict__ blk_box, Buf A, Scalars* S,
                                                 BoxState* BS,
                                                 const int* __restrict__ perm,
                                                 const int4* __restrict__ act,
                                                 const int* __restrict__ cand,
                                                 uint2* __restrict__ clist,
                                                 int* __restrict__ ccnt,
                                                 const double4* __restrict__ fpk,
                                                 const int* __restrict__ fid,
                                                 const BoxSlot* rd, BoxSlot* wr) {
  const int tile = xcd_tile();
  const int box = blk_box[(tile * blockDim.x) / BOX_ALIGN];
  // FAST: summed largest moves since the last rebuild, before the control
  // adds the last step's
  double disp = FAST ? BS[box].disp_acc : 0.0;
  if (FAST && tile == 0 && threadIdx.x < 64)
    if (!step_control(PB, S, BS, rd, wr, 1, 1)) return;
  int stop = (S->halt || S->overflow >= 2) ? 1 : 0;
  if (FAST) stop = __shfl(stop, 0, 64);  // one decision per wave: the whole wave reduces below
  if (stop) return;
  const Params& P = PB[box];
  // packed slot t: a particle of the box with candidates, from either end of
  // the box's slots
  const int t = tile * blockDim.x + threadIdx.x;
  const bool listed = !(t - P.off >= BS[box].n_act && t - P.off < P.bcap - BS[box].n_far);
  // a fast step's waves stay whole until they have reduced the move below
  if (FAST ? !__ballot(listed) : !listed) return;
  const size_t cap = P.cap;
  // the packed entry and the first CPF list entries, all in flight at once
  // (their addresses depend on t only), then every partner position of
  // those: two latencies for most particles instead of a chain
  constexpr int CPF = 4;
  unsigned ent[CPF] = {0u, 0u, 0u, 0u};
  int4 a = make_int4(0, 0, 0, 0);
  if (listed) {
    for (int q = 0; q < CPF; q++)
      if (q < P.lcap) ent[q] = (unsigned)cand[(size_t)q * cap + t];
    a = act[t];
  }
  if (FAST) {
    // the last step's largest move, reduced by the whole wave while the
    // loads above are in flight
    double step = bitsd(rd[(size_t)box * NSLOT + (threadIdx.x & 63)].step_bits);
    for (int off = 32; off > 0; off >>= 1) step = fmax(step, __shfl_xor(step, off, 64));
    disp = disp + step;
    if (!listed) return;
  }
  const int k = a.x, nq = min(a.y, P.lcap);
  const bool rebuilt = !FAST && BS[box].rebuild != 0;
  if (!FAST) disp = BS[box].disp_acc;  // the control kernel has added the last step's move
  // No particle of the box has moved more than disp (the summed moves, or
  // more) since the lists were built, so no listed pair has closed its gap
  // by more than twice that: while the smallest gap of the list (less a
  // rounding margin) is larger, no candidate touches and nothing more is
  // loaded.  The summed moves only grow until the next rebuild, so a
  // particle skipped now touched nothing at any step since the rebuild at
  // which it was tested: its count is the 0 written then.
  const double gap = __hiloint2double(a.w, a.z);
  if (gap > 2.0 * disp * (1.0 + 1.0e-9)) {
    if (rebuilt) ccnt[k] = 0;
    return;
  }
  const int s = rebuilt ? perm[k] : k;
  double4 pjs[CPF];
  for (int q = 0; q < CPF; q++)
    if (q < nq) {
      const unsigned e = ent[q];
      const bool fixed = (e & CL_FIXED) != 0;
      const int j = fixed ? (int)(e & CL_INDEX) : (rebuilt ? perm[e & CL_INDEX] : (int)(e & CL_INDEX));
      pjs[q] = fixed ? fpk[j] : A.pos[j];
      ent[q] = fixed ? e : (unsigned)j;
    }
  const double4 pi = A.pos[s];
  const int idi = A.id[s];
  const HeadKeys own = head_keys(A, s);
  int c = 0;
  // one candidate: the reference's predicate, then the entry and its
  // history row
  auto consider = [&](unsigned e, double4 pj, bool fixed, int j) {
    int shifted = 0;
    const double xj = image(pi.x, pj.x, P.simlen[0], P.periodic[0], &shifted);
    const double yj = image(pi.y, pj.y, P.simlen[1], P.periodic[1], &shifted);
    const double zj = image(pi.z, pj.z, P.simlen[2], P.periodic[2], &shifted);
    const double dr = pi.w + pj.w;
    const double dr2 = dr * dr;
    const double ddx = pi.x - xj;
    double d2 = ddx * ddx;
    if (d2 > dr2) return;
    const double ddy = pi.y - yj;
    d2 = ddy * ddy + d2;
    if (d2 > dr2) return;
    const double ddz = pi.z - zj;
    d2 = ddz * ddz + d2;
    if (d2 > dr2) return;
    if (c < P.hcap) {
      // the record in the particle's own list; for PP from the higher id
      // that is its copy of the lower id's record, and failing one, the
      // record in the lower id's list (a pair that was periodic, or a copy
      // the host did not install)
      const int idj = fixed ? fid[j] : A.id[j];
      const bool copy = !fixed && !shifted && idj < idi;
      unsigned row = own_row(A, cap, s, own, idj, c, copy);
      if (copy && row == NO_ROW) {
        const unsigned r_other = hist_row(A, cap, j, head_keys(A, j), idi);
        row = r_other == NO_ROW ? NO_ROW : (r_other | ROW_PARTNER);
      }
      clist[(size_t)c * cap + k] = make_uint2(fixed ? e : ((unsigned)j | (shifted ? CL_SHIFT : 0u)), row);
    }
    c++;
  };
  // the prefetched candidates (unrolled: the staged values stay in
  // registers), then the rest of the list
#pragma unroll
  for (int q = 0; q < CPF; q++)
    if (q < nq) {
      const bool fixed = (ent[q] & CL_FIXED) != 0;
      consider(ent[q], pjs[q], fixed, (int)(ent[q] & CL_INDEX));
    }
  for (int q = CPF; q < nq; q++) {
    const unsigned e = (unsigned)cand[(size_t)q * cap + t];
    const bool fixed = (e & CL_FIXED) != 0;
    const int j = fixed ? (int)(e & CL_INDEX) : (rebuilt ? perm[e & CL_INDEX] : (int)(e & CL_INDEX));
    consider(e, fixed ? fpk[j] : A.pos[j], fixed, j);
  }
  ccnt[k] = c;
  // More contacts than history slots: the force kernel of the step turns
  // this into an overflow (the host grows the slots and replays the step).
  // A fast step that its control halts runs no force kernel, so what its
  // other waves found here never halts a call.
  if (c > P.hcap) atomicMax(&S->cnt_over, 1);
}

// ---------------------------------------------------------------------------
// The force step, in the parts the kernel k_force is made of:
//   contact_loop  : every contact of a particle -> F, T, the new history
//   Integ / integrate : accelerations, integration, wrap / delete, the
//                   new state and next step's near-wall mask
//   wave_flush    : per-wave reductions, one atomic per wave and value
// ---------------------------------------------------------------------------

// the start-of-step state of one particle (buffer A, slot s) the contact
// loop reads (its spin rotVel(1) is loaded where it is used: by the contact
// loop of a particle that touches something, and by the integration when it
// reads rotVel(1), which AB2 rotation does not: a particle without contacts
// does not load it)
struct Own {
  double4 pi;
  d3 vi;
  int idi, ti;
  uint64_t mask0;  // near-wall list
};

__device__ __forceinline__ Own load_own(const Params& P, const Buf& A, int s) {
  Own o;
  o.pi = A.pos[s];
  o.idi = A.id[s];
  o.ti = A.type[s];
  o.vi = ld3(A.lv, P.cap, s);
  o.mask0 = A.wmask[s];
  return o;
}

// Every contact of the particle in slot s of A (new slot k): the contacts
// k_contact found, in ascending partner id, then the walls of the near-wall
// list, newest first (FindContactsPW :298-319 walks the list from the
// highest index).  Each contact is reduced to the arguments of the force
// routine from the perspective the reference evaluates it from
// (ContactForce_PP :732-756, _PPG :761-798, _PPFix :803-827, _PW :832-856).
// The walls run in a loop of their own after the particle contacts.  The
// records go to slot k of B in the same order (a PP pair's in both lists:
// KIND_COPY).
// `parts` selects the particle contacts (1), the walls (2) or both (3), in
// that order: the phase entry points run them in two kernels, which continue
// the same sums.
__device__ __forceinline__ void contact_loop(const Params& P, const Buf& A, const Buf& B, int k, int s,
                                             const Own& o, const uint2* __restrict__ clist,
                                             const int* __restrict__ ccnt, const double4* __restrict__ fpk,
                                             const int* __restrict__ fid, const int* __restrict__ ftype,
                                             Acc& acc, int parts = 3) {
  const int cap = P.cap;
  const d3 xi = D3(o.pi.x, o.pi.y, o.pi.z);
  const double ri = o.pi.w;
  const int idi = o.idi, ti = o.ti;
  const int nc = (parts & 1) ? min(ccnt[k], P.hcap) : 0;  // more: cnt_over is set
  d3 wi = D3(0.0, 0.0, 0.0);
  if (nc > 0 || ((parts & 2) && o.mask0 != 0ull)) wi = ld3(A.rv, cap, s);
  auto apply = [&](const CF& cf, bool flip, int kind, int partner) {
    if (!flip) {
      acc.F = add3(acc.F, cf.fsum);
      acc.T = add3(add3(acc.T, mul3(cf.moment, ri)), cf.mrij);
    } else {
      acc.F = sub3(acc.F, cf.fsum);
      acc.T = sub3(add3(acc.T, mul3(cf.moment, ri)), cf.mrij);
    }
    push_hist(P, B, k, acc, kind, partner, cf);
  };
  // each entry is loaded one contact ahead of its use, and names the bucket
  // of its record, so that a contact costs one round of gathers (partner
  // state and record together)
  uint2 ce_next = make_uint2(0u, NO_ROW);
  if (nc > 0) ce_next = clist[k];
  for (int q = 0; q < nc; q++) {
    const uint2 ce = ce_next;
    if (q + 1 < nc) ce_next = clist[(size_t)(q + 1) * cap + k];
    const unsigned e = ce.x;
    const int j = (int)(e & CL_INDEX);
    const Hist h = hist_of_entry(A, cap, s, j, ce.y);
    double4 pj;
    d3 vj, wj;
    int idj, tj;
    if (e & CL_FIXED) {
      pj = fpk[j];
      vj = D3(0.0, 0.0, 0.0);
      wj = vj;
      idj = fid[j];
      tj = ftype[j];
    } else {
      pj = A.pos[j];
      vj = ld3(A.lv, cap, j);
      wj = ld3(A.rv, cap, j);
      idj = A.id[j];
      tj = A.type[j];
    }
    int shifted = 0;
    const double xj = image(o.pi.x, pj.x, P.simlen[0], P.periodic[0], &shifted);
    const double yj = image(o.pi.y, pj.y, P.simlen[1], P.periodic[1], &shifted);
    const double zj = image(o.pi.z, pj.z, P.simlen[2], P.periodic[2], &shifted);
    const double dr = ri + pj.w;
    const double ddx = xi.x - xj, ddy = xi.y - yj, ddz = xi.z - zj;
    double d2 = ddx * ddx;
    d2 = ddy * ddy + d2;
    d2 = ddz * ddz + d2;
    // |xi - xj| as .nv. takes it too: the components are the same up to
    // sign, and the sums the same up to order of two exact operands
    const double dist = sqrt(d2);
    const double ovrlp = dr - dist;
    const d3 xjv = D3(xj, yj, zj);
    const Bnry* Bp;
    int kind;
    bool flip = false;
    if (e & CL_FIXED) {
      // PPFix: the moving particle's record, its own perspective
      kind = 3;
      Bp = &P.pp[(ti - 1) * P.ntype + (tj - 1)];
    } else if (!shifted) {
      // PP: the record sits in the lower id's bucket and the force is
      // taken from that side (FineSearch2 :525-529)
      const bool ilo = idi < idj;
      kind = ilo ? 1 : KIND_COPY;
      const int tlo = ilo ? ti : tj, thi = ilo ? tj : ti;
      Bp = &P.pp[(tlo - 1) * P.ntype + (thi - 1)];
      flip = !ilo;
    } else {
      // PPG: each side keeps its own record against the other's image,
      // the force from the lower id's side; row type = the local particle
      kind = 2;
      Bp = &P.pp[(ti - 1) * P.ntype + (tj - 1)];
      flip = idi > idj;
    }
    // .nv. of the two centres (zero within 1e-10 of each other)
    const d3 sep = flip ? sub3(xi, xjv) : sub3(xjv, xi);
    const d3 nvec = dist > 1.0e-10 ? div3(sep, dist) : D3(0.0, 0.0, 0.0);
    const CF cf = contact_force(P, *Bp, nvec, flip ? pj.w : ri, flip ? ri : pj.w, flip ? vj : o.vi,
                                flip ? o.vi : vj, flip ? wj : wi, flip ? wi : wj, ovrlp, h.td, h.vri);
    apply(cf, flip, kind, idj);
  }
  if (!(parts & 2)) return;
  // the walls of the particle's list, highest index first
  uint64_t mask = o.mask0;
  while (mask) {
    const int w = 63 - __clzll((long long)mask);
    mask &= ~(1ull << w);
    const Wall& W = P.walls[w];
    d3 nv;
    double ovrlp;
    if (!pw_in_contact(W, xi, ri, &ovrlp, &nv)) continue;
    const Hist h = hist_at(A, cap, s, own_row(A, cap, s, head_keys(A, s), W.wall_id, acc.hc, false));
    const Bnry& Bw = P.pw[(ti - 1) * P.nwtype + (W.wall_type - 1)];
    const d3 nvec = mul3(nv, -1.0);  // Norm_v = -nv
    const CF cf = contact_force(P, Bw, nvec, ri, 1.00e20 * ri, o.vi, wp(W.trans_vel), wi, D3(0.0, 0.0, 0.0),
                                ovrlp, h.td, h.vri);
    apply(cf, false, 4, W.wall_id);
  }
}

// Accelerations (DEMS_clc_Acceleration, Prtcl_DEMSystem.f90:360-375):
// a = F / m + g, alpha = T / I of the particle's type.
__device__ __forceinline__ void accel(const Params& P, int ti, d3 F, d3 T, d3& a1, d3& al1) {
  const double mass = P.mass[ti - 1], inertia = P.inertia[ti - 1];
  a1 = add3(div3(F, mass), D3(P.g[0], P.g[1], P.g[2]));
  al1 = div3(T, inertia);
}

// The integration inputs beyond the contact loop's (history levels, theta,
// mark).  PIM / PRIM: translational / rotational integration method fixed
// at compile time (0: read from Params); with fixed methods they are loaded
// by `preload` in one batch ahead of their use, otherwise where they are
// used.
template <int PIM, int PRIM>
struct Integ {
  d3 v2, v3, a2, a3, th, w2, w3, al2, al3;
  int mark;

  __device__ __forceinline__ void preload(const Params& P, const Scalars* S, const Buf& A, int s, d3 vi,
                                          bool put_ids) {
    const int cap = P.cap;
    const int pim = PIM ? PIM : P.pi_method, prim = PRIM ? PRIM : P.pri_method;
    const d3 z3 = D3(0.0, 0.0, 0.0);
    v2 = v3 = a2 = a3 = th = w2 = w3 = al2 = al3 = z3;
    mark = 0;
    if (!PIM) return;
    // rotVel(2) of a state the force step wrote is its linVel(1) (below)
    const bool rv2 = S->rv2_stored[A.idx] != 0;
    if (pim >= 2) {
      v2 = ld3n(A.lv + 3 * cap, cap, s);
      a2 = ld3n(A.la, cap, s);
    }
    if (pim == 3) {
      v3 = ld3n(A.lv + 6 * cap, cap, s);
      a3 = ld3n(A.la + 3 * cap, cap, s);
    }
    th = ld3n(A.th, cap, s);
    if (prim == 2) {
      if (rv2) w2 = ld3n(A.rv + 3 * cap, cap, s);
      else w2 = vi;
      al2 = ld3n(A.ra, cap, s);
    } else if (prim == 3) {
      w2 = ld3n(A.rv + 3 * cap, cap, s);
      w3 = ld3n(A.rv + 6 * cap, cap, s);
      al2 = ld3n(A.ra, cap, s);
      al3 = ld3n(A.ra + 3 * cap, cap, s);
    }
    if (put_ids) mark = A.mark[s];
  }

  // Integration (Prtcl_Integrate :15-88) of the particle in slot s of A
  // with accelerations a1, al1: the new velocity levels, accelerations and
  // rotation into slot k of B; returns the new position, before any wrap,
  // and linVel(1) in lin1.
  __device__ __forceinline__ d3 advance(const Params& P, const Scalars* S, const Buf& A, const Buf& B, int k, int s,
                                        const Own& o, d3 a1, d3 al1, d3& lin1) {
    const int cap = P.cap;
    const int pim = PIM ? PIM : P.pi_method, prim = PRIM ? PRIM : P.pri_method;
    const d3 xi = D3(o.pi.x, o.pi.y, o.pi.z), vi = o.vi;
    const bool rv2 = PIM ? false : S->rv2_stored[A.idx] != 0;
    const double dt = P.dt;
    d3 pos = xi;
    if (pim == 1) {
      d3 sft = mul3(vi, dt);
      pos = D3(pos.x + sft.x, pos.y + sft.y, pos.z + sft.z);
      lin1 = add3(vi, mul3(a1, dt));
      st3n(B.lv, cap, k, lin1);
      st3n(B.la, cap, k, a1);
    } else if (pim == 2) {
      const double c1 = 3.0 / 2.0, c2 = -1.0 / 2.0;
      if (!PIM) {
        v2 = ld3n(A.lv + 3 * cap, cap, s);
        a2 = ld3n(A.la, cap, s);
      }
      d3 sft = mul3(add3(mul3(vi, c1), mul3(v2, c2)), dt);
      pos = D3(pos.x + sft.x, pos.y + sft.y, pos.z + sft.z);
      lin1 = add3(vi, mul3(add3(mul3(a1, c1), mul3(a2, c2)), dt));
      st3n(B.lv, cap, k, lin1);
      st3n(B.lv + 3 * cap, cap, k, vi);
      st3n(B.la, cap, k, a1);  // linAcc(2) <- linAcc(1)
    } else {
      const double c1 = 23.0 / 12.0, c2 = -16.0 / 12.0, c3 = 5.0 / 12.0;
      if (!PIM) {
        v2 = ld3n(A.lv + 3 * cap, cap, s);
        v3 = ld3n(A.lv + 6 * cap, cap, s);
        a2 = ld3n(A.la, cap, s);
        a3 = ld3n(A.la + 3 * cap, cap, s);
      }
      d3 sft = mul3(add3(add3(mul3(vi, c1), mul3(v2, c2)), mul3(v3, c3)), dt);
      pos = D3(pos.x + sft.x, pos.y + sft.y, pos.z + sft.z);
      lin1 = add3(vi, mul3(add3(add3(mul3(a1, c1), mul3(a2, c2)), mul3(a3, c3)), dt));
      st3n(B.lv, cap, k, lin1);
      st3n(B.lv + 3 * cap, cap, k, vi);
      st3n(B.lv + 6 * cap, cap, k, v2);
      st3n(B.la, cap, k, a1);
      st3n(B.la + 3 * cap, cap, k, a2);
    }
    if (!PIM) th = ld3n(A.th, cap, s);
    const d3 wi = prim == 2 ? D3(0.0, 0.0, 0.0) : ld3(A.rv, cap, s);  // rotVel(1), unused by AB2
    if (prim == 1) {
      th = add3(th, mul3(wi, dt));
      st3n(B.rv, cap, k, add3(wi, mul3(al1, dt)));
      st3n(B.ra, cap, k, al1);
    } else if (prim == 2) {
      // Prtcl_Integration.f90:64: rotVel1 = GPrtcl_linVel(1,i), and
      // rotVel(2) <- rotVel1 (:69), so rotVel(2) of the new state is its
      // linVel(1): it is not stored, and read back as linVel(1) (vi) from a
      // buffer the force step wrote
      const double c1 = 3.0 / 2.0, c2 = -1.0 / 2.0;
      d3 w1 = lin1;
      if (!PIM) {
        if (rv2) w2 = ld3n(A.rv + 3 * cap, cap, s);
        else w2 = vi;
        al2 = ld3n(A.ra, cap, s);
      }
      th = add3(th, mul3(add3(mul3(w1, c1), mul3(w2, c2)), dt));
      st3n(B.rv, cap, k, add3(w1, mul3(add3(mul3(al1, c1), mul3(al2, c2)), dt)));
      st3n(B.ra, cap, k, al1);
    } else {
      const double c1 = 23.0 / 12.0, c2 = -16.0 / 12.0, c3 = 5.0 / 12.0;
      if (!PIM) {
        w2 = ld3n(A.rv + 3 * cap, cap, s);
        w3 = ld3n(A.rv + 6 * cap, cap, s);
        al2 = ld3n(A.ra, cap, s);
        al3 = ld3n(A.ra + 3 * cap, cap, s);
      }
      th = add3(th, mul3(add3(add3(mul3(wi, c1), mul3(w2, c2)), mul3(w3, c3)), dt));
      st3n(B.rv, cap, k, add3(wi, mul3(add3(add3(mul3(al1, c1), mul3(al2, c2)), mul3(al3, c3)), dt)));
      st3n(B.rv + 3 * cap, cap, k, wi);
      st3n(B.rv + 6 * cap, cap, k, w2);
      st3n(B.ra, cap, k, al1);
      st3n(B.ra + 3 * cap, cap, k, al2);
    }
    st3n(B.th, cap, k, th);
    return pos;
  }

  // Periodic wrap / deletion (Comm_For_Exchange) of the new position `pos`
  // and the stores of the new state's position, id, mark, type and next
  // step's near-wall mask into slot k of B.  Returns whether the particle
  // left through an open side.
  __device__ __forceinline__ bool exchange(const Params& P, const Buf& A, const Buf& B, int k, int s, const Own& o,
                                           d3 pos, bool put_ids, bool put_wall, bool upd_next) {
    bool dead = false;
    double pv[3] = {pos.x, pos.y, pos.z};
    for (int d = 0; d < 3; d++) {
      if (P.periodic[d]) {
        if (pv[d] >= P.dmax[d]) pv[d] = pv[d] - P.simlen[d];
        else if (pv[d] < P.dmin[d]) pv[d] = pv[d] + P.simlen[d];
      } else if (pv[d] >= P.dmax[d] || pv[d] < P.dmin[d]) {
        dead = true;
      }
    }
    const int ti = o.ti;
    B.pos[k] = make_double4(pv[0], pv[1], pv[2], o.pi.w);
    if (put_ids) {
      B.id[k] = o.idi;
      B.mark[k] = PIM ? mark : A.mark[s];
    }
    if (put_ids || dead) B.type[k] = dead ? -ti : ti;
    // the next step's near-wall list is built here from the new state
    // (UpdateNearPrtclsPW runs at the start of that step, after deletion)
    if (put_wall) B.wmask[k] = (upd_next && !dead) ? near_wall_mask(P, pv[0], pv[1], pv[2]) : o.mask0;
    return dead;
  }

  // Accelerations, integration, wrap / deletion and the new state in slot
  // k of B, for the contact forces F, T.  Returns whether the particle left
  // through an open side; move2 = its squared move, vnext / anext = |v|,
  // |a| of the new state when the next step refreshes the near-wall lists.
  __device__ __forceinline__ bool integrate(const Params& P, const Scalars* S, const Buf& A, const Buf& B, int k,
                                            int s, const Own& o, d3 F, d3 T, bool put_ids, bool put_wall,
                                            bool upd_next, double& move2, double& vnext, double& anext) {
    d3 a1, al1, lin1;
    accel(P, o.ti, F, T, a1, al1);
    const d3 pos = advance(P, S, A, B, k, s, o, a1, al1, lin1);
    // the move of this step, before any periodic wrap
    {
      const double mx = pos.x - o.pi.x, my = pos.y - o.pi.y, mz = pos.z - o.pi.z;
      move2 = mx * mx + my * my + mz * mz;
    }
    const bool dead = exchange(P, A, B, k, s, o, pos, put_ids, put_wall, upd_next);
    if (upd_next && !dead) {
      vnext = norm3(lin1);
      anext = norm3(a1);  // linAcc(1) at the start of the next step
    }
    if (P.keep_forces) {
      st3n(B.force, P.cap, k, F);
      st3n(B.force + 3 * P.cap, P.cap, k, T);
    }
    return dead;
  }
};

// What the force step decides from the box's control for one block: A is
// read through the sort permutation on a step that re-sorted the box; id,
// type, mark and the near-wall list change only where the box is re-sorted,
// a particle is removed or the lists are refreshed, so they are written when
// the buffer written here may hold other values (this step or the one
// before re-sorted the box, or refreshed the lists); otherwise both buffers
// hold the same values in the same order and the stores are left out.
struct StepFlags {
  bool rebuilt, upd_next, put_ids, put_wall;
};
__device__ __forceinline__ StepFlags step_flags(const Scalars* S, const BoxState& bs) {
  StepFlags f;
  f.rebuilt = bs.rebuild != 0;
  // the control has advanced iterNumber: is the next step a near-wall update?
  f.upd_next = S->iter >= bs.next_wall_update;
  f.put_ids = f.rebuilt || bs.prev_rebuild != 0;
  f.put_wall = f.put_ids || f.upd_next || bs.wall_prev != 0;
  return f;
}

// Wave reductions, then one atomic per wave and value on slot
// (wave % NSLOT) of the box: no barrier.  The counts are exact integer sums
// and the maxima exact, whatever the order.  `counts`: the contact counts
// (the kernel ran the contact loop); `motion`: the move, max|v|, max|a| and
// deletions (the kernel integrated).
__device__ __forceinline__ void wave_flush(BoxSlot* __restrict__ bslots, int box, bool counts, bool motion,
                                           unsigned long long nkind, double move2, bool upd_next, double vnext,
                                           double anext, bool dead) {
  unsigned long long c01 = (nkind & 0xffffull) | (((nkind >> 16) & 0xffffull) << 32);
  unsigned long long c23 = ((nkind >> 32) & 0xffffull) | (((nkind >> 48) & 0xffffull) << 32);
  double wmove = move2;
  if (counts)
    for (int off = 32; off > 0; off >>= 1) {
      c01 += __shfl_xor(c01, off, 64);
      c23 += __shfl_xor(c23, off, 64);
    }
  int wdel = 0;
  double wv = 0.0, wa = 0.0;
  if (motion) {
    for (int off = 32; off > 0; off >>= 1) wmove = fmax(wmove, __shfl_xor(wmove, off, 64));
    wdel = __popcll(__ballot(dead));
    if (upd_next) {
      wv = vnext;
      wa = anext;
      for (int off = 32; off > 0; off >>= 1) {
        wv = fmax(wv, __shfl_xor(wv, off, 64));
        wa = fmax(wa, __shfl_xor(wa, off, 64));
      }
    }
  }
  if ((threadIdx.x & 63) == 0) {
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    BoxSlot& bl = bslots[box * NSLOT + wave % NSLOT];
    if (motion) {
      if (wmove > 0.0) atomicMax(&bl.step_bits, dbits(sqrt(wmove)));
      if (upd_next) {
        atomicMax(&bl.max_v_bits, dbits(wv));
        atomicMax(&bl.max_a_bits, dbits(wa));
      }
      if (wdel) atomicAdd(&bl.ndel, wdel);
    }
    if (counts) {
      const int cnt[4] = {(int)(c01 & 0xffffffffull), (int)(c01 >> 32), (int)(c23 & 0xffffffffull),
                          (int)(c23 >> 32)};
      for (int q = 0; q < 4; q++)
        if (cnt[q]) atomicAdd(&bl.cnt[q], cnt[q]);
    }
  }
}

__device__ __forceinline__ Acc acc_zero() {
  Acc acc;
  acc.F = D3(0.0, 0.0, 0.0);
  acc.T = D3(0.0, 0.0, 0.0);
  acc.hc = 0;
  acc.nkind = 0ull;
  acc.overflow = 0;
  return acc;
}

// The fused force step: contact loop, accelerations and integration in one
// kernel, reading and writing each particle's stream once (the choice for
// beds with few contacts, where the step is a streaming pass).  With fixed
// integration methods (PIM, PRIM) the integration inputs are loaded in one
// batch after the contact loop: loaded before it, they stay live across
// the loop and spill at the cap of two waves per SIMD (256 VGPRs), and the
// spill traffic of every particle costs a streaming step more than the
// latency they hide.  OCC, the waves per SIMD the kernel is built for: 2
// (no spills) for beds with few contacts; 3 (168 VGPRs, ~30 spilled) where
// spheres touch, whose contact loop is bound by the partner gathers in
// flight, which a third wave adds more of than its spills cost (the same
// arithmetic: the same bits).
template <int PIM, int PRIM, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void k_force(const Params* __restrict__ PB,
                                               const int* __restrict__ blk_box, Buf A, Buf B, Scalars* S,
                                               BoxState* BS,
                                               const int* __restrict__ perm,
                                               const uint2* __restrict__ clist,
                                               const int* __restrict__ ccnt,
                                               const double4* __restrict__ fpk,
                                               const int* __restrict__ fid,
                                               const int* __restrict__ ftype,
                                               BoxSlot* __restrict__ bslots) {
  // a candidate-list overflow in this step's rebuild leaves the lists
  // incomplete: do nothing, the host regrows them and replays the step
  if (S->halt || S->overflow >= 2) return;
  const int tile = xcd_tile();
  const int box = blk_box[(tile * blockDim.x) / BOX_ALIGN];
  const Params& P = PB[box];
  const int k = tile * blockDim.x + threadIdx.x;
  const StepFlags f = step_flags(S, BS[box]);
  Acc acc = acc_zero();
  double move2 = 0.0, vnext = 0.0, anext = 0.0;
  bool dead = false;
  if (k - P.off < BS[box].n_alive) {
    const int s = f.rebuilt ? perm[k] : k;
    const Own o = load_own(P, A, s);
    Integ<PIM, PRIM> in;
    contact_loop(P, A, B, k, s, o, clist, ccnt, fpk, fid, ftype, acc);
    in.preload(P, S, A, s, o.vi, f.put_ids);
    dead = in.integrate(P, S, A, B, k, s, o, acc.F, acc.T, f.put_ids, f.put_wall, f.upd_next, move2, vnext, anext);
    B.hcnt[k] = dead ? 0 : acc.hc;
  }
  wave_flush(bslots, box, true, true, acc.nkind, move2, f.upd_next, vnext, anext, dead);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->have_parts = 1;
    S->rv2_stored[B.idx] = 0;
    // this step's contact test found more contacts than history slots
    if (S->cnt_over) atomicMax(&S->overflow, 1);
  }
  if (acc.overflow) atomicMax(&S->overflow, 1);
}

// ---------------------------------------------------------------------------
// The force step in the reference's phases (the entry points
// sdem_find_contacts ... sdem_comm_for_exchange): the parts of k_force in
// kernels of their own, the same operations in the same order, so that a
// step run phase by phase gives k_force's bits.  F and T, then the
// accelerations, cross between the phases in `ft` (6 x cap doubles, slot k),
// the new position before the wrap in B.pos.
// ---------------------------------------------------------------------------
// FindContacts (Prtcl_DEMSystem.f90:225): the particle contacts' forces and
// records
__global__ __launch_bounds__(256) void k_ph_pp(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                               Buf A, Buf B, Scalars* S, BoxState* BS, const int* __restrict__ perm,
                                               const uint2* __restrict__ clist, const int* __restrict__ ccnt,
                                               const double4* __restrict__ fpk, const int* __restrict__ fid,
                                               const int* __restrict__ ftype, double* __restrict__ ft,
                                               BoxSlot* __restrict__ bslots) {
  if (S->halt || S->overflow >= 2) return;
  const int box = block_box(blk_box);
  const Params& P = PB[box];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  Acc acc = acc_zero();
  if (k - P.off < BS[box].n_alive) {
    const int s = BS[box].rebuild ? perm[k] : k;
    const Own o = load_own(P, A, s);
    contact_loop(P, A, B, k, s, o, clist, ccnt, fpk, fid, ftype, acc, 1);
    st3(ft, P.cap, k, acc.F);
    st3(ft + 3 * (size_t)P.cap, P.cap, k, acc.T);
    B.hcnt[k] = acc.hc;
  }
  wave_flush(bslots, box, true, false, acc.nkind, 0.0, false, 0.0, 0.0, false);
  if (blockIdx.x == 0 && threadIdx.x == 0 && S->cnt_over) atomicMax(&S->overflow, 1);
  if (acc.overflow) atomicMax(&S->overflow, 1);
}

// FindContactsPW (:230): the walls, continuing the sums and the records
__global__ __launch_bounds__(256) void k_ph_pw(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                               Buf A, Buf B, Scalars* S, BoxState* BS, const int* __restrict__ perm,
                                               const uint2* __restrict__ clist, const int* __restrict__ ccnt,
                                               const double4* __restrict__ fpk, const int* __restrict__ fid,
                                               const int* __restrict__ ftype, double* __restrict__ ft,
                                               BoxSlot* __restrict__ bslots) {
  if (S->halt || S->overflow >= 2) return;
  const int box = block_box(blk_box);
  const Params& P = PB[box];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  Acc acc = acc_zero();
  if (k - P.off < BS[box].n_alive) {
    const int s = BS[box].rebuild ? perm[k] : k;
    const Own o = load_own(P, A, s);
    acc.F = ld3(ft, P.cap, k);
    acc.T = ld3(ft + 3 * (size_t)P.cap, P.cap, k);
    acc.hc = B.hcnt[k];
    contact_loop(P, A, B, k, s, o, clist, ccnt, fpk, fid, ftype, acc, 2);
    st3(ft, P.cap, k, acc.F);
    st3(ft + 3 * (size_t)P.cap, P.cap, k, acc.T);
    B.hcnt[k] = acc.hc;
  }
  wave_flush(bslots, box, true, false, acc.nkind, 0.0, false, 0.0, 0.0, false);
  if (acc.overflow) atomicMax(&S->overflow, 1);
}

// clc_Acceleration (:235): ft holds a and alpha afterwards
__global__ __launch_bounds__(256) void k_ph_accel(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                                  Buf A, Buf B, const Scalars* S, const BoxState* BS,
                                                  const int* __restrict__ perm, double* __restrict__ ft) {
  if (S->halt || S->overflow >= 2) return;
  const int box = block_box(blk_box);
  const Params& P = PB[box];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k - P.off >= BS[box].n_alive) return;
  const int s = BS[box].rebuild ? perm[k] : k;
  const d3 F = ld3(ft, P.cap, k), T = ld3(ft + 3 * (size_t)P.cap, P.cap, k);
  if (P.keep_forces) {
    st3n(B.force, P.cap, k, F);
    st3n(B.force + 3 * P.cap, P.cap, k, T);
  }
  d3 a1, al1;
  accel(P, A.type[s], F, T, a1, al1);
  st3(ft, P.cap, k, a1);
  st3(ft + 3 * (size_t)P.cap, P.cap, k, al1);
}

// Prtcl_Integrate (:240): the new velocities, accelerations and rotation,
// the new position before any wrap in B.pos, the step's largest move
template <int PIM, int PRIM>
__global__ __launch_bounds__(256) void k_ph_integ(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                                  Buf A, Buf B, const Scalars* S, const BoxState* BS,
                                                  const int* __restrict__ perm, const double* __restrict__ ft,
                                                  BoxSlot* __restrict__ bslots) {
  if (S->halt || S->overflow >= 2) return;
  const int box = block_box(blk_box);
  const Params& P = PB[box];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const StepFlags f = step_flags(S, BS[box]);
  double move2 = 0.0;
  if (k - P.off < BS[box].n_alive) {
    const int s = f.rebuilt ? perm[k] : k;
    Own o;
    o.pi = A.pos[s];
    o.ti = A.type[s];
    o.idi = 0;
    o.vi = ld3(A.lv, P.cap, s);
    o.mask0 = 0ull;
    Integ<PIM, PRIM> in;
    in.preload(P, S, A, s, o.vi, false);
    const d3 a1 = ld3(ft, P.cap, k), al1 = ld3(ft + 3 * (size_t)P.cap, P.cap, k);
    d3 lin1;
    const d3 pos = in.advance(P, S, A, B, k, s, o, a1, al1, lin1);
    const double mx = pos.x - o.pi.x, my = pos.y - o.pi.y, mz = pos.z - o.pi.z;
    move2 = mx * mx + my * my + mz * mz;
    B.pos[k] = make_double4(pos.x, pos.y, pos.z, o.pi.w);
  }
  wave_flush(bslots, box, false, true, 0ull, move2, false, 0.0, 0.0, false);
}

// Comm_For_Exchange (:245): wrap or delete the new positions; ids, marks,
// types and the next step's near-wall masks of the new state
__global__ __launch_bounds__(256) void k_ph_exch(const Params* __restrict__ PB, const int* __restrict__ blk_box,
                                                 Buf A, Buf B, Scalars* S, const BoxState* BS,
                                                 const int* __restrict__ perm, BoxSlot* __restrict__ bslots) {
  if (S->halt || S->overflow >= 2) return;
  const int box = block_box(blk_box);
  const Params& P = PB[box];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const StepFlags f = step_flags(S, BS[box]);
  double vnext = 0.0, anext = 0.0;
  bool dead = false;
  if (k - P.off < BS[box].n_alive) {
    const int s = f.rebuilt ? perm[k] : k;
    Own o;
    o.pi = A.pos[s];
    o.ti = A.type[s];
    o.idi = f.put_ids ? A.id[s] : 0;
    o.mask0 = (f.put_wall && !f.upd_next) ? A.wmask[s] : 0ull;
    const double4 q = B.pos[k];
    Integ<0, 0> in;
    dead = in.exchange(P, A, B, k, s, o, D3(q.x, q.y, q.z), f.put_ids, f.put_wall, f.upd_next);
    if (f.upd_next && !dead) {
      vnext = norm3(ld3(B.lv, P.cap, k));
      anext = norm3(ld3(B.la, P.cap, k));  // linAcc(1) at the start of the next step
    }
    if (dead) B.hcnt[k] = 0;
  }
  wave_flush(bslots, box, false, true, 0ull, 0.0, f.upd_next, vnext, anext, dead);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->have_parts = 1;
    S->rv2_stored[B.idx] = 0;
  }
}

// ---------------------------------------------------------------------------
// host-side launchers (called from sdem_api.cpp).  Particle kernels cover all
// `nslots` storage slots (a multiple of BOX_ALIGN) with blocks of `bs`
// threads (a divisor of BOX_ALIGN).
// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

int64_t scan_tiles(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }
int scan_tile_keys() { return SCAN_TILE; }

// slot set `set` (numbered like the state buffers) of every box
static inline BoxSlot* slot_set(const Dev& D, int set) { return D.bslots + (size_t)set * D.nbox * NSLOT; }

hipError_t launch_pre(const Dev& D, const Buf& A, hipStream_t st) {
  hipLaunchKernelGGL(k_pre, dim3(nblk(D.nslots, 256)), dim3(256), 0, st, D.PB, D.blk_box, A, D.S, D.BS,
                     slot_set(D, A.idx));
  return hipGetLastError();
}

hipError_t launch_ctl(const Dev& D, const Buf& A, hipStream_t st) {
  hipLaunchKernelGGL(k_ctl, dim3(1), dim3(64), 0, st, D.PB, D.S, D.BS, slot_set(D, A.idx),
                     slot_set(D, 1 - A.idx), D.nbox);
  if (D.nbox > 1)
    hipLaunchKernelGGL(k_box, dim3(D.nbox), dim3(64), 0, st, D.PB, D.S, D.BS, slot_set(D, A.idx),
                       slot_set(D, 1 - A.idx));
  return hipGetLastError();
}

hipError_t launch_finish(const Dev& D, const Buf& A, hipStream_t st) {
  hipLaunchKernelGGL(k_finish, dim3(D.nbox), dim3(64), 0, st, D.S, D.BS, slot_set(D, A.idx));
  return hipGetLastError();
}

hipError_t launch_sort(const Dev& D, const Buf& A, hipStream_t st) {
  const int64_t n = D.nkeys + 1;
  const unsigned nt = nblk(n, SCAN_TILE);
  const unsigned nb = nblk(D.nslots, 256);
  hipLaunchKernelGGL(k_bin, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, D.S, D.BS, D.keys, D.rank,
                     D.counts);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SCAN_T), 0, st, D.counts, D.start, D.tile_sum, n, D.S, D.BS,
                     D.tile_box);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, st, D.tile_sum, (int)nt, D.S);
  hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(SCAN_T), 0, st, D.counts, D.start, D.tile_sum, n, D.S, D.BS,
                     D.tile_box);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, D.S, D.BS, D.keys, D.rank, D.start,
                     D.perm_tmp);
  hipLaunchKernelGGL(k_rank, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, D.S, D.BS, D.keys, D.start,
                     D.perm_tmp, D.perm, D.pk, D.pid);
  return hipGetLastError();
}

hipError_t launch_vlist(const Dev& D, hipStream_t st) {
  hipLaunchKernelGGL(k_vlist, dim3(nblk(D.nslots, 256)), dim3(256), 0, st, D.PB, D.blk_box,
                     D.S, D.BS, D.pk, D.pid, D.start, D.vlist, D.vkey, D.act, D.cand, D.ccnt, D.fpk,
                     D.fid, D.fstart);
  return hipGetLastError();
}

// one phase of the step's force part (sdem_find_contacts ... _exchange):
// 1 contact test + particle contacts, 2 walls, 3 accelerations,
// 4 integration, 5 exchange
hipError_t launch_phase(const Dev& D, const Buf& A, const Buf& B, hipStream_t st, int phase) {
  const unsigned nb = nblk(D.nslots, 256);
  switch (phase) {
    case 1:
      hipLaunchKernelGGL(k_contact<0>, dim3(nblk(D.nslots, D.bs_contact)), dim3(D.bs_contact), 0, st, D.PB,
                         D.blk_box, A, D.S, D.BS, D.perm, D.act, D.cand, D.clist, D.ccnt, D.fpk, D.fid,
                         slot_set(D, A.idx), slot_set(D, B.idx));
      hipLaunchKernelGGL(k_ph_pp, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, D.perm, D.clist,
                         D.ccnt, D.fpk, D.fid, D.ftype, D.ft, slot_set(D, B.idx));
      break;
    case 2:
      hipLaunchKernelGGL(k_ph_pw, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, D.perm, D.clist,
                         D.ccnt, D.fpk, D.fid, D.ftype, D.ft, slot_set(D, B.idx));
      break;
    case 3:
      hipLaunchKernelGGL(k_ph_accel, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, D.perm, D.ft);
      break;
    case 4:
      if (D.pim == 2 && D.prim == 2)
        hipLaunchKernelGGL((k_ph_integ<2, 2>), dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, D.perm,
                           D.ft, slot_set(D, B.idx));
      else
        hipLaunchKernelGGL((k_ph_integ<0, 0>), dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, D.perm,
                           D.ft, slot_set(D, B.idx));
      break;
    case 5:
      hipLaunchKernelGGL(k_ph_exch, dim3(nb), dim3(256), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, D.perm,
                         slot_set(D, B.idx));
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_force(const Dev& D, const Buf& A, const Buf& B, hipStream_t st, hipEvent_t mid, bool fast) {
  if (fast && D.nbox == 1)
    hipLaunchKernelGGL(k_contact<1>, dim3(nblk(D.nslots, D.bs_contact)), dim3(D.bs_contact), 0, st, D.PB,
                       D.blk_box, A, D.S, D.BS, D.perm, D.act, D.cand, D.clist, D.ccnt, D.fpk, D.fid,
                       slot_set(D, A.idx), slot_set(D, B.idx));
  else
    hipLaunchKernelGGL(k_contact<0>, dim3(nblk(D.nslots, D.bs_contact)), dim3(D.bs_contact), 0, st, D.PB,
                       D.blk_box, A, D.S, D.BS, D.perm, D.act, D.cand, D.clist, D.ccnt, D.fpk, D.fid,
                       slot_set(D, A.idx), slot_set(D, B.idx));
  if (mid) (void)hipEventRecord(mid, st);
  const unsigned nf = nblk(D.nslots, D.bs_force);
#define SDEM_FORCE(P, R, O)                                                                                  \
  hipLaunchKernelGGL((k_force<P, R, O>), dim3(nf), dim3(D.bs_force), 0, st, D.PB, D.blk_box, A, B, D.S, D.BS, \
                     D.perm, D.clist, D.ccnt, D.fpk, D.fid, D.ftype, slot_set(D, B.idx))
  if (D.pim == 2 && D.prim == 2) {
    if (D.occ3) SDEM_FORCE(2, 2, 3);
    else SDEM_FORCE(2, 2, 2);
  } else {
    if (D.occ3) SDEM_FORCE(0, 0, 3);
    else SDEM_FORCE(0, 0, 2);
  }
#undef SDEM_FORCE
  return hipGetLastError();
}

}  // namespace sdem
