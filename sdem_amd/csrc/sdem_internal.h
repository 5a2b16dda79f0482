// sdem_internal.h -- device-side data layout shared by the kernels and the
// C-ABI implementation of libsdem_hip.so.
//
// Particle state lives in two SoA buffers (A = state at the start of a step,
// B = state at its end).  The force kernel reads A (through the cell-sort
// permutation on the steps that rebuild the candidate lists) and writes B,
// so the particles stay stored in the cell order of the last rebuild.
// Contact history is held per owner particle in hcap slots interleaved with
// the other particles' (owner = the particle whose bucket holds the record
// in the reference: the lower id for PP, the local particle for PPG / PPFix
// / PW; Prtcl_CL_and_CF.f90:142-249).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdem {

// partner key: id in the low 29 bits, contact kind (1..4) above it
constexpr int KEY_BITS = 29;
constexpr int KEY_MASK = (1 << KEY_BITS) - 1;
// A PP record lives in the lower id's list (as the reference's buckets);
// the higher id's list holds a copy of it under this kind, written by the
// same step with the same bits, so that each side of a pair reads the
// pair's history from its own rows (which line up with its contact list
// from step to step).  Copies are not contact records: the readers leave
// them out, and the reference's lookups in a particle's own bucket (PPG,
// PPFix, PW) pass over them.
constexpr int KIND_COPY = 5;

// Where record r of particle s keeps its partner key: the first KEY_HEAD
// records of a particle side by side (one 16-byte load gives all four, as
// an int4 at s), the others in rows of cap, as the record values (hval at
// r*cap + s).  The hcap >= KEY_HEAD rows of an hcap x cap array hold both.
constexpr int KEY_HEAD = 4;
__host__ __device__ inline size_t key_slot(int r, size_t s, size_t cap) {
  return r < KEY_HEAD ? (size_t)KEY_HEAD * s + (size_t)r : (size_t)r * cap + s;
}
constexpr int MAX_WALLS = 64;     // walls seen by one context (near-wall bit mask)
constexpr int MAX_TYPES = 64;

// binary (pair) coefficients, Prtcl_Property.f90:28-41
struct Bnry {
  double rad_eff, mass_eff, kn, kt, dn, fric_s, fric_k, fric_roll;
};

// plane wall, Prtcl_Geometry.f90:17-37 (+ derived quantities)
struct Wall {
  double p1[3], p2[3], p3[3], p4[3];
  double n[3], d;
  double min_point[3], max_point[3];
  double trans_vel[3];
  int both_side, is_infinite, wall_type, wall_id;
};

// uniform cell grid of one search level (x fastest).  A periodic dimension
// tiles its period with n cells of width L/n; a closed one starts at the
// domain minimum.  Cell indices are clamped to [0,n) for stored particles.
struct Grid {
  double origin[3];
  double inv_w[3];
  double w[3];
  int n[3];
  int64_t ncell;
  int64_t key_base;   // first key of this level in the combined key space
  // Keys of the moving particles' cells: bricks of BRICK^3 cells in
  // row-major brick order, the cells of a brick row-major inside it, so
  // that particles close in space sit close in slot order along all three
  // axes (a wave's neighbours share its cache lines).  nb: bricks per axis,
  // nkey = nb[0] nb[1] nb[2] BRICK^3 keys.
  int nb[3];
  int64_t nkey;
  double rmax;        // largest radius among this level's particles
};
constexpr int MAX_LEVELS = 6;
constexpr int BRICK_BITS = 2, BRICK = 1 << BRICK_BITS;

// one SoA particle buffer
struct Buf {
  double4* pos;   // x, y, z, radius (GPrtcl_PosR)
  double *lv;     // (tsize*3) x cap  : lv[(l*3+c)*cap + i]
  double *la;     // (ta*3) x cap     : stored acceleration levels
  double *th;     // 3 x cap
  double *rv;     // (rsize*3) x cap
  double *ra;     // (tr*3) x cap
  int *id, *type, *mark;
  uint64_t *wmask;  // near-wall list as a bit mask over wall indices
  int *hcnt;        // history records owned by the particle
  int *hkey;        // hcap x cap partner keys, record c of particle s at key_slot(c, s, cap)
  double4 *hval;    // hcap x cap (TanDelta.x,.y,.z, VelRel_Init), same layout
  double *force;    // 6 x cap (F, T) when forces are kept
  int idx;          // which of the context's two buffers this is
};

// everything the kernels read that does not change inside a step
struct Params {
  double dt;
  double g[3];
  double dmin[3], dmax[3], simlen[3];
  int periodic[3];
  int cf_type, ct_model, pi_method, pri_method;
  int tsize, rsize, ta, tr;   // velocity / stored acceleration levels
  int cap;                    // particle capacity (stride of SoA arrays)
  int hcap;                   // history slots per particle
  int ntype, nwtype, nwall;
  int search_fixed;           // 0 when CS_Method=2 (hierarchical search has no fixed branch)
  int keep_forces;
  int nlvl;                   // number of search levels (1 = single grid)
  Grid lvl[MAX_LEVELS];
  const int* type_level;      // ntype: search level of each particle type
  Grid fgrid;                 // grid of the fixed particles
  int nfix;
  double dxmax1, dxmax2, max_wall_vel;
  double skin;                // candidate-list skin (search tuning)
  double rebuild_disp;        // half the skin, with a margin: the box's lists are
                              // rebuilt once its particles may have moved that far
  int lcap;                   // candidate-list entries per particle
  int wall_max_update_iter;
  const Bnry* pp;             // ntype*ntype, row = type of the record's lower-id particle
  const Bnry* pw;             // ntype*nwtype
  const double* mass;         // ntype (1-based types -> index t-1)
  const double* inertia;
  const Wall* walls;          // nwall
  // the box's place in the context
  int box;                    // index of the box
  int off, bcap;              // storage slots [off, off + bcap)
  int key_first;              // first sort key of the box (its keys are
                              // [key_first, dead_key])
  int dead_key;               // sort key of the box's removed / unused slots
};

// Partial results of the per-particle kernels, accumulated with atomics
// spread over NSLOT addresses per box (wave w uses slot w % NSLOT) so that
// no single address takes thousands of atomics; reduced by the next step's
// control and by k_finish.  There are two sets of NSLOT slots per box,
// numbered like the two state buffers: a step's force kernel adds into the
// set of the buffer it writes, and the next step's control reduces that
// set (the one of its input buffer) and clears the other for its own force
// kernel, so that the reads and the clearing never meet in one kernel.
// Set s of every box: bslots[s * nbox * NSLOT + box * NSLOT + slot].
constexpr int NSLOT = 64, NSET = 2;
struct BoxSlot {                 // one slot (NSLOT per box and set)
  unsigned long long max_v_bits; // near-wall list: max |v|, max |a| (as bits of
  unsigned long long max_a_bits; // non-negative doubles)
  unsigned long long step_bits;  // largest distance a particle moved in the step
  int cnt[4];                    // contacts by kind
  int ndel;                      // particles removed
  int pad;
};

// Every box (an independent system: one for a plain context, many for an
// ensemble) owns the storage slots [off, off + bcap) of every per-particle
// array; bcap is a multiple of BOX_ALIGN, so a block of up to BOX_ALIGN
// threads never spans two boxes and reads its box once, as a scalar.
constexpr int BOX_ALIGN = 256;

// device-side state of one box
struct BoxState {
  int n_stored;            // entries of the box in buffer A at the start of the step
  int n_alive;             // live particles after this step's sort
  int n_deleted;           // particles removed at the end of the last step
  int next_wall_update;    // Next_iter_update
  int next_wall_at_begin;  // ... at the start of the last step that ran
  int use_host_max_a;      // first step after the state was set: max|a| from the host
  double host_max_a;
  int num_cntcts[4];       // numCntcts of the last step
  int rebuild;             // this step re-sorts the box and rebuilds its lists
  int n_act;               // particles with candidates and a gap under half the skin (packed)
  int prev_rebuild;        // the step before this one re-sorted the box
  int wall_prev;           // the step before refreshed the near-wall lists
  int n_far;               // ... of them with all gaps over half the skin (packed
                           // from the end of the box's slots down)
  int n_unlisted;          // dense layout: live particles without candidates
  int dense;               // the candidate lists are indexed by the particle's
                           // own slot (cell order; most particles have
                           // candidates), not packed
  double disp_acc;         // sum over steps since the box's last rebuild of the
                           // largest move of one of its particles: bounds every
                           // displacement since then
};

// device-side scalars of one context
struct Scalars {
  int iter;                // iterNumber (all boxes step together)
  int iter_at_begin;       // iterNumber at the start of the last step that ran
  int overflow;            // 1: history slots, 2: candidate-list entries, 3: a
                           // coordinate outside the periodic domain (error)
  int halt;                // set by the step after an overflow: no more work
  int halted_steps;        // steps of the call skipped after an overflow
  int rebuild;             // some box rebuilds this step (the sort kernels run)
  int need_rebuild;        // every box rebuilds at the next step
  int forced;              // ... this step (need_rebuild, or no partial results)
  int rebuilds;            // box rebuilds so far (statistics)
  int have_parts;          // the force kernel's slots are current
  int list_need;           // largest candidate count seen when a list overflowed
  int spec_fail;           // a step queued without the rebuild kernels needed them:
                           // it halted the call before doing anything
  int cnt_over;            // a contact test found more contacts than history
                           // slots (the force kernel raises overflow)
  int rv2_stored[2];       // with AB2 rotation, rotVel(2) sits in buffer b's rv
                           // array; 0 once the force kernel wrote the buffer:
                           // rotVel(2) is then linVel(1) of the same buffer
                           // (Prtcl_Integration.f90:64 stores one as the other)
  int fin_records;         // a single box: contact records and live particles at
  int fin_alive;           // the last step of the last call's batch (k_finish)
};

// Device pointers and launch shapes the host hands to the launchers.
struct Dev {
  const Params* PB;      // one per box
  const int* blk_box;    // box of each BOX_ALIGN-slot group
  int nbox;
  int64_t nslots;        // storage slots, sum of the boxes' bcap
  int64_t nkeys;         // sort keys of all boxes (cells + one last key each)
  Scalars* S;
  BoxState* BS;
  BoxSlot* bslots;       // NSET sets of NSLOT per box, set-major
  int *keys, *rank, *counts, *start, *tile_sum, *perm_tmp, *perm;
  const int2* tile_box;  // first and last box whose keys a scan tile covers
  double4* pk;           // positions in cell order (rebuild steps)
  int* pid;              // ids in cell order (rebuild steps)
  int *vlist, *vkey;     // candidate lists by particle (lists longer than the LDS stage)
  int4* act;             // (particle, candidate count, smallest gap as a double) per packed slot
  int* cand;             // candidate entries by packed slot: entry q of slot t at q*cap+t
  uint2* clist;          // contacts of the step by particle: (entry, history row), entry q of k at q*cap+k
  int* ccnt;
  double* ft;            // phase entry points: F, T, then a, alpha (6 x cap)
  const double4* fpk;
  const int *fid, *ftype, *fstart;
  int bs_contact, bs_force;  // block sizes (divisors of BOX_ALIGN)
  int pim, prim;             // integration methods (the same in every box)
  int occ3;                  // the force kernel built for three waves per SIMD (spheres in contact)
};

}  // namespace sdem
