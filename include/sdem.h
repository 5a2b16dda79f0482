/*
 * sdem.h -- C ABI of libsdem_hip.so, the MI355X (gfx950) implementation of
 * sDEM's per-timestep DEM path.
 *
 * The library replaces the body of DEMS_iterate (src/DEM/Prtcl_DEMSystem.f90:201-317)
 * on one GPU: near-wall list upkeep, contact search (NBS-Munjiza grid or the
 * hierarchical multi-level variant), the persistent contact list with its
 * tangential history, linear (LSD) and Hertz-Mindlin (nLin) particle-particle,
 * particle-ghost, particle-fixed and particle-wall forces, accelerations,
 * FE/AB2/AB3 integration, and the periodic wrap / open-boundary deletion.
 *
 * Conventions (all entry points):
 *   - return 0 on success, a negative SDEM_E_* code on failure; the message
 *     is then available from sdem_last_error();
 *   - host arrays are owned by the caller and copied before return; no
 *     pointer is retained;
 *   - a context is used by one host thread at a time;
 *   - array layouts follow the Fortran arrays of the reference so a Fortran
 *     host can pass them through ISO_C_BINDING unchanged:
 *       posr      : n x 4 doubles   (GPrtcl_PosR, real4 x,y,z,w=radius)
 *       lin_vel   : n x tsize x 3   (GPrtcl_linVel(tsize,n), real3)
 *       lin_acc   : n x tsize x 3
 *       theta     : n x 3
 *       rot_vel   : n x rsize x 3,  rot_acc : n x rsize x 3
 *     where tsize/rsize = 1,2,3 for FE/AB2/AB3 (Prtcl_Variables.f90:100-113).
 */
#ifndef SDEM_H
#define SDEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes -------------------------------------------------------- */
#define SDEM_OK           0
#define SDEM_E_ARG       -1  /* invalid argument or configuration            */
#define SDEM_E_HIP       -2  /* a HIP runtime call failed                      */
#define SDEM_E_CAPACITY  -3  /* particle / contact-history capacity exceeded   */
#define SDEM_E_STATE     -4  /* call order violated (e.g. step before state)   */
#define SDEM_E_NODEV     -5  /* no usable gfx950 device                        */

/* ---- enumerations, same values as Prtcl_Parameters.f90:13-24 ------------- */
#define SDEM_CSM_NBS_MUNJIZA        1
#define SDEM_CSM_NBS_MUNJIZA_HRCHL  2
#define SDEM_CFT_LSD   1
#define SDEM_CFT_NLIN  2
/* constant-stiffness linear spring-dashpot: no counterpart in the reference
 * (its CF_Type has only the two models above); set with
 * sdem_set_linear_stiffness */
#define SDEM_CFT_LINEAR 3
#define SDEM_CTM_CONSTANT_TORQUE  1
#define SDEM_CTM_VARIABLE_TORQUE  2
#define SDEM_PIM_FE   1
#define SDEM_PIM_AB2  2
#define SDEM_PIM_AB3  3

/* contact kinds, the four counters of ContactList%numCntcts
 * (Prtcl_CL_and_CF.f90:156,187,214,242) */
#define SDEM_KIND_PP     1
#define SDEM_KIND_PPG    2
#define SDEM_KIND_PPFIX  3
#define SDEM_KIND_PW     4

/* profiling phases reported by sdem_get_profile */
#define SDEM_PHASE_CONTROL    0   /* rebuild decision, near-wall cadence        */
#define SDEM_PHASE_BIN        1   /* cell keys + counting sort (rebuild steps)  */
#define SDEM_PHASE_SEARCH     2   /* candidate lists (rebuild steps)            */
#define SDEM_PHASE_CONTACT    3   /* contact test over the candidate lists      */
#define SDEM_PHASE_FORCE      4   /* history/force/accel/integrate kernel       */
#define SDEM_NUM_PHASES       5

/* the reference's step timers (Prtcl_DEMSystem.f90:29-37, logged at
 * :280-305), reported by sdem_get_timers while profiling is on */
#define SDEM_TIMER_TOTAL          0  /* m_total_timer                             */
#define SDEM_TIMER_PRE_ITER       1  /* m_pre_iter_timer      "PreItertion time"     */
#define SDEM_TIMER_COMM_CS        2  /* m_comm_cs_timer       "Comm_For_Contact"     */
#define SDEM_TIMER_CSCF_PP        3  /* m_CSCF_PP_timer       "CS and CF P-P time"   */
#define SDEM_TIMER_CSCF_PW        4  /* m_CSCF_PW_timer       "CS and CF P-W time"   */
#define SDEM_TIMER_ACCELERATION   5  /* m_Acceleration_timer  "Acceleration time"    */
#define SDEM_TIMER_INTEGRATION    6  /* m_integration_timer   "Integration time"     */
#define SDEM_TIMER_COMM_EXCHANGE  7  /* m_comm_exchange_timer "Comm_For_Exchange"    */
#define SDEM_TIMER_WRITE          8  /* m_write_prtcl_timer   "Write to file time"   */
#define SDEM_NUM_TIMERS           9

/*
 * The DEMS_Options fields the step reads (Prtcl_Parameters.f90:45-95,
 * namelist &DEMOptions read at :120-124).  Doubles first, then 32-bit
 * integers, so the layout is the same as a BIND(C) Fortran derived type.
 */
typedef struct sdem_options {
  double  dt;                    /* dtDEM                                      */
  double  gravity[3];            /* gravity                                    */
  double  domain_min[3];         /* minpoint  (SimDomain_min)                  */
  double  domain_max[3];         /* maxpoint  (SimDomain_max)                  */
  double  prtcl_cs_ratio;        /* Prtcl_cs_ratio (cell = ratio * d_max)      */
  double  wall_neighbor_ratio;   /* Wall_neighbor_ratio                        */
  int32_t is_periodic[3];        /* isPeriodic (0/1)                           */
  int32_t cs_method;             /* CS_Method   (SDEM_CSM_*)                   */
  int32_t cf_type;               /* CF_Type     (SDEM_CFT_*)                   */
  int32_t ct_model;              /* CT_Model    (SDEM_CTM_*)                   */
  int32_t pi_method;             /* PI_Method   (SDEM_PIM_*)                   */
  int32_t pri_method;            /* PRI_Method  (SDEM_PIM_*)                   */
  int32_t cs_numlvls;            /* CS_numlvls  (0 = levels from d_max/d_min)  */
  int32_t cntct_list_size;       /* CntctList_Size: history slots / particle   */
  int32_t base_wall_id;          /* Base_wall_id                               */
  int32_t wall_max_update_iter;  /* Wall_max_update_iter                       */
  int32_t num_prtcl;             /* numPrtcl: particle capacity                */
  int32_t ifirst;                /* ifirstDEM: iterNumber starts at ifirst-1   */
} sdem_options;

typedef struct sdem_ctx sdem_ctx;

/* Fill *o with the defaults of DEMS_Options (Prtcl_Parameters.f90:45-95). */
int sdem_options_default(sdem_options* o);

/* Create a context on the current HIP device. Replaces the DEM_opt setup of
 * DO_ReadDEMOption (Prtcl_Parameters.f90:104-205); everything the MPI
 * decomposition computed is replaced by the whole domain on one GPU. */
int sdem_create(const sdem_options* o, sdem_ctx** out);

/* Ensemble of independent boxes (no counterpart in the reference): nbox
 * systems, each with its own options (domain, walls, materials, contact
 * model, time step), stepped together by one sdem_step -- every kernel of a
 * step covers all boxes at once.  All boxes share PI_Method, PRI_Method and
 * ifirst; fixed particles need a single-box context.  sdem_create is the
 * one-box case.  The setters and readers below address the box chosen with
 * sdem_select_box (box 0 at creation); an ensemble's states and contacts are
 * set before its first step.  Each box gives the same bits as a context of
 * its own. */
int sdem_create_ensemble(int32_t nbox, const sdem_options* opts /* nbox */, sdem_ctx** out);
int sdem_get_num_boxes(sdem_ctx* c, int32_t* nbox);
int sdem_select_box(sdem_ctx* c, int32_t box);

/* InitPrtclProperty (Prtcl_Property.f90:62-143): per-type pure properties
 * and the particle-particle binary table, built with the reference's
 * row-type rule (pair (i,j) takes both partners' properties from type i,
 * Prtcl_Property.f90:137-138). Arrays have ntype entries (type ids 1..ntype). */
int sdem_set_particle_property(sdem_ctx* c, int32_t ntype,
                               const double* diameter, const double* density,
                               const double* youngs_modulus, const double* poisson_ratio,
                               double fric_s, double fric_k, double fric_roll,
                               double rest_n, double rest_t);

/* Constant-stiffness linear model (cf_type SDEM_CFT_LINEAR; no counterpart in
 * the reference): normal stiffness kn [N/m] for every particle-particle and
 * particle-wall pair, tangential stiffness kt = kt_over_kn * kn, normal
 * damping from the pair's restitution e and effective mass m* as the
 * reference's linear model has it at unit velocity factor,
 * dn = -2 ln(e) sqrt(m* kn) / sqrt(pi^2 + ln(e)^2), which gives the
 * restitution e exactly for a head-on collision.  Here m* is the true pair
 * value (both types), and walls count as infinitely heavy.  Call before
 * sdem_set_particle_property. */
int sdem_set_linear_stiffness(sdem_ctx* c, double kn, double kt_over_kn);

/* InitWallProperty (Prtcl_Property.f90:148-195): wall pure properties and the
 * particle-wall binary table. Call after sdem_set_particle_property. */
int sdem_set_wall_property(sdem_ctx* c, int32_t nwtype,
                           const double* youngs_modulus, const double* poisson_ratio,
                           double fric_s, double fric_k, double fric_roll,
                           double rest_n, double rest_t);

/* G_add_PlaneWall (Prtcl_Geometry.f90:457-537): adds one plane quad; normal,
 * offset and bounding box are derived as the reference derives them.
 * wall_id = Base_wall_id + running count. Walls outside the (expanded)
 * domain are skipped as in the reference (returns 0, *wall_id = 0). */
int sdem_add_plane_wall(sdem_ctx* c, const double p1[3], const double p2[3],
                        const double p3[3], const double p4[3],
                        int32_t user_id, int32_t wall_type,
                        int32_t both_side, int32_t is_infinite,
                        const double trans_vel[3], int32_t* wall_id);

/* Fixed particles (PIO_ReadFixedCoord, Prtcl_IOAndVisu.f90:164-291): ids
 * must not collide with moving ids (the reference uses numPrtcl+k). */
int sdem_set_fixed_particles(sdem_ctx* c, int32_t nfix, const int32_t* id,
                             const int32_t* ptype, const double* posr);

/* Particle state (what Read_Restart / MakingAllPrtcl produce,
 * Prtcl_IOAndVisu.f90:731-844).  usr_mark may be NULL. */
int sdem_set_state(sdem_ctx* c, int32_t n, const int32_t* id, const int32_t* ptype,
                   const int32_t* usr_mark, const double* posr,
                   const double* lin_vel, const double* lin_acc, const double* theta,
                   const double* rot_vel, const double* rot_acc);

/* Contact-history records (what RestartCL / Add_RestartCntctlink install,
 * Prtcl_CL_and_CF.f90:577-622).  A record belongs to owner_id's bucket:
 * the lower-id particle for PP, the local particle for PPG/PPFix/PW.
 * partner_id is a particle id, fixed-particle id or wall id. */
int sdem_set_contacts(sdem_ctx* c, int64_t m, const int32_t* kind,
                      const int32_t* owner_id, const int32_t* partner_id,
                      const double* tan_delta, const double* vrel_init);

/* nsteps x DEMS_iterate (Prtcl_DEMSystem.f90:201-248) without I/O or logs.
 * Asynchronous w.r.t. the host only in the sense that it returns once the
 * work is queued and checked; it synchronises before returning. */
int sdem_step(sdem_ctx* c, int32_t nsteps);

/* ---- one step phase by phase ---------------------------------------------
 * The calls of DEMS_iterate (Prtcl_DEMSystem.f90:213-248) one at a time, so
 * that each reference call can be stood in for on its own.  They run on the
 * context's device state in the reference's order -- pre-iteration, ghost
 * communication, particle contacts, wall contacts, accelerations,
 * integration, exchange, release -- and together are one sdem_step(c, 1),
 * with its bits (the same operations in the same order).  A call out of that
 * order returns SDEM_E_STATE, and so does sdem_step while a phased step is
 * open.  If sdem_find_contacts or sdem_find_contacts_pw runs out of
 * history or candidate-list capacity it grows it, puts the step back to its
 * start and returns SDEM_E_CAPACITY: run the step again from
 * sdem_pre_iteration.  iterNumber reads as advanced from sdem_pre_iteration
 * on (the step's control advances it with the near-wall cadence). */
int sdem_pre_iteration(sdem_ctx* c);     /* DEMS_preIteration :322-331: UpdateNearPrtclsPW, F/T zeroing, PreIteration */
int sdem_comm_for_cntct(sdem_ctx* c);    /* Comm_For_Cntct :220 (ghost images: minimum image here, nothing to do) */
int sdem_find_contacts(sdem_ctx* c);     /* FindContacts :225: search, PP / PPG / PPFix forces and history */
int sdem_find_contacts_pw(sdem_ctx* c);  /* FindContactsPW :230: wall forces and history */
int sdem_clc_acceleration(sdem_ctx* c);  /* clc_Acceleration :235 */
int sdem_integrate(sdem_ctx* c);         /* Prtcl_Integrate :240 */
int sdem_comm_for_exchange(sdem_ctx* c); /* Comm_For_Exchange :245: periodic wrap, deletion */
int sdem_remv_released(sdem_ctx* c);     /* RemvReleased :246 and the end of the step (iterNumber + 1 :248) */

/* Number of particles in the domain (np_InDomain, Prtcl_DEMSystem.f90:252). */
int sdem_get_num_particles(sdem_ctx* c, int32_t* n);

/* State in ascending id order; cap = capacity of the arrays in particles.
 * Any output pointer may be NULL. */
int sdem_get_state(sdem_ctx* c, int32_t cap, int32_t* n,
                   int32_t* id, int32_t* ptype, int32_t* usr_mark, double* posr,
                   double* lin_vel, double* lin_acc, double* theta,
                   double* rot_vel, double* rot_acc);

/* GPrtcl_cntctForce / GPrtcl_torque of the last step, ascending id order. */
int sdem_get_forces(sdem_ctx* c, int32_t cap, int32_t* id,
                    double* cntct_force, double* torque);

/* Contact records after the last step, sorted by (owner_id, partner_id). */
int sdem_get_num_contacts(sdem_ctx* c, int64_t* m);
int sdem_get_contacts(sdem_ctx* c, int64_t cap, int64_t* m, int32_t* kind,
                      int32_t* owner_id, int32_t* partner_id,
                      double* tan_delta, double* vrel_init);

/* numCntcts(1:4) of the last step: P-P | P-GP | P-FP | P-W
 * (Prtcl_DEMSystem.f90:313). */
int sdem_get_counts(sdem_ctx* c, int32_t num_cntcts[4]);

/* iterNumber and the near-wall list's Next_iter_update
 * (Prtcl_ContactSearchPW.f90:29,229). */
int sdem_get_iter(sdem_ctx* c, int32_t* iter_number, int32_t* next_wall_update);

/* Re-sorts and candidate-list builds so far, summed over the boxes (a
 * statistic of this implementation, which keeps candidate lists between
 * rebuilds; the reference searches the cells every step and has no
 * counterpart).  A box rebuilds when the sum over steps of its largest
 * particle move since its last rebuild may exceed half the skin. */
int sdem_get_rebuilds(sdem_ctx* c, int64_t* rebuilds);

/* Per-phase kernel timing with HIP events on the library's stream.
 * enable=1 starts collecting (resets totals); totals are in milliseconds. */
int sdem_set_profiling(sdem_ctx* c, int32_t enable);
int sdem_get_profile(sdem_ctx* c, double ms[SDEM_NUM_PHASES],
                     int64_t launches[SDEM_NUM_PHASES]);
/* The reference's nine step timers (SDEM_TIMER_*): total milliseconds and
 * number of timed steps or calls since profiling was enabled.  Steps run
 * phase by phase fill each timer from its phase entry point; sdem_step runs
 * the contact search, every force, the accelerations, integration and
 * exchange in fused kernels and books all of them under SDEM_TIMER_CSCF_PP
 * (the step's control under SDEM_TIMER_PRE_ITER).  The library writes no
 * files during a step: SDEM_TIMER_WRITE is the caller's. */
int sdem_get_timers(sdem_ctx* c, double tot_ms[SDEM_NUM_TIMERS], int64_t calls[SDEM_NUM_TIMERS]);

/* Search and output tuning (no counterpart in the reference; results do not
 * depend on it).  cell_factor[d] scales the search-cell edge (the reference's
 * Prtcl_cs_ratio * d_max) in dimension d, 0 keeps the current value.
 * skin_factor (fraction of d_max, below 0.45; negative keeps the current
 * value, default 0.1) is the margin of the candidate lists, which are rebuilt
 * with the cell sort only once the sum over steps of the largest move of a
 * particle passes half the skin (or a particle was removed); the
 * contact set is decided every step with the exact predicate, so it is the
 * same for any factors.  keep_forces = 0 stops writing the per-particle
 * contact force / torque arrays that sdem_get_forces reads (the forces are
 * still computed and integrated); -1 keeps the current setting.  Call before
 * sdem_set_state for the cell factors to apply. */
int sdem_set_tuning(sdem_ctx* c, const double cell_factor[3], double skin_factor,
                    int32_t keep_forces);

/* Initial capacities (no counterpart in the reference, whose contact list
 * holds CntctList_Size * mlocal records in one pool): hist_slots history
 * records per particle (0 = a kissing-number bound from the type radii) and
 * list_slots candidate-list entries per particle (0 = twice hist_slots, at
 * least 48).  Both grow on demand: a step that runs out is replayed from its
 * intact input after the storage has been enlarged.  Call before
 * sdem_set_state. */
int sdem_set_capacity(sdem_ctx* c, int32_t hist_slots, int32_t list_slots);

/* Block sizes of the contact-test and force kernels (64, 128 or 256; no
 * counterpart in the reference, results do not depend on them). */
int sdem_set_block_sizes(sdem_ctx* c, int32_t contact_block, int32_t force_block);

/* Which build of the force kernel the steps launch: 0 (default) chooses per
 * batch of steps, three waves per SIMD where the last step had 0.25 contact
 * records per live particle or more, two otherwise; 2 or 3 fixes it.  No
 * counterpart in the reference; results do not depend on it. */
int sdem_set_force_build(sdem_ctx* c, int32_t waves_per_simd);
int sdem_get_tuning(sdem_ctx* c, double cell_factor[3], double* skin_factor, int32_t* contact_block,
                    int32_t* force_block);

/* Autotuner (no counterpart in the reference): tries cell factors, skins
 * and block sizes, each for steps_per_trial timed steps (plus two settling
 * steps) in a forward and a backward pass over the settings, and keeps the
 * one with the smallest mean (a run whose cost drifts while the trials run
 * favours no setting for its place in the order).  The trials are real
 * steps of the run (the results do not depend on the tuning); *steps_run
 * says how many. */
int sdem_autotune(sdem_ctx* c, int32_t steps_per_trial, int32_t* steps_run, double* best_ms_per_step);

/* Memory the context holds on the device, in bytes. */
int sdem_get_device_bytes(sdem_ctx* c, int64_t* bytes);

/* ---- utilities ------------------------------------------------------------
 * Seeded random sequential addition of n non-overlapping spheres in the box
 * [lo,hi) (largest first; each sphere retries up to max_attempts uniform
 * centres).  periodic[d] lets spheres straddle that face.  Writes pos (n x 3).
 * Replaces the clock-seeded GL_MakingAllPrtcl (Prtcl_Variables.f90:191-307)
 * for reproducible initial states; host only, no device needed.
 * Returns SDEM_E_CAPACITY when a sphere could not be placed. */
int sdem_util_random_pack(int64_t n, const double* radius, const double lo[3], const double hi[3],
                          const int32_t periodic[3], uint64_t seed, int32_t max_attempts,
                          double* pos);

const char* sdem_last_error(const sdem_ctx* c);
void sdem_destroy(sdem_ctx* c);

#ifdef __cplusplus
}
#endif
#endif /* SDEM_H */
