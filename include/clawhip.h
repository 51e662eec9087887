/* clawhip.h -- C ABI of libclawhip.so, the MI355X (gfx950) implementation of
 * PyClaw's time-step hot path.
 *
 * Plain pointers and sizes only.  Every array argument is DEVICE memory; every
 * entry point enqueues work on `stream` (a hipStream_t, NULL = default stream)
 * and returns without synchronising.  Return value: 0 on success, a negative
 * CLAWHIP_E* code for an invalid argument, or a positive hipError_t.
 *
 * Layout of a state array (struct clawhip_geom): field f of cell (i, j, k)
 * (1-based interior indices) is at
 *     base[f*field_stride + (i-1+ghosts) + (j-1+ghosts)*row_stride
 *          + (k-1+ghosts)*slab_stride]
 * `ghosts` is the number of ghost layers present in storage: 0 for the compact
 * state the kernels keep, num_ghost for a reference-style qbc whose ghosts the
 * caller filled (boundary code CLAWHIP_BC_GIVEN).  With any other boundary code
 * the kernels produce the ghost values themselves, exactly as
 * Solver._bc_lower/_bc_upper (pyclaw/solver.py:439-519) would.
 *
 * Numerics: dtype CLAWHIP_F64 reproduces the reference's FP64 Fortran step bit
 * for bit.  CLAWHIP_F32K keeps q in FP32 with a Kahan compensation array of the
 * same layout (`comp`); the CFL number is always reported in FP64.
 */
#ifndef CLAWHIP_H
#define CLAWHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLAWHIP_E_ARG (-1)     /* invalid size, stride or pointer */
#define CLAWHIP_E_KIND (-2)    /* unknown Riemann solver / unsupported combination */
#define CLAWHIP_E_METHOD (-3)  /* unsupported method flags */

/* dtype: the arithmetic of a step.
 *   CLAWHIP_F64       FP64, every expression rounded as the reference's Fortran
 *                     (no contraction): bit for bit the reference's step;
 *   CLAWHIP_F32K      FP32 state and arithmetic, Kahan-compensated update, FP64
 *                     CFL; a*b + c within one expression is one FMA;
 *   CLAWHIP_F64_FAST  FP64 state and arithmetic with a*b + c within one
 *                     expression as one FMA: not bitwise, within about 1e-13
 *                     relative of CLAWHIP_F64 per step.  Arrays as CLAWHIP_F64.
 * The step controllers' FP64 arithmetic (dt, accept/reject, t) is the same in
 * all three. */
enum { CLAWHIP_F64 = 0, CLAWHIP_F32K = 1, CLAWHIP_F64_FAST = 2 };

/* Riemann solver identifiers: the clawpack.riemann modules the reference's
 * examples pass to ClawSolver2D(...) (e.g. examples/euler_2d/quadrants.py:80). */
enum {
  CLAWHIP_RP_ACOUSTICS = 1, /* acoustics_2D             */
  CLAWHIP_RP_EULER = 2,     /* euler_4wave_2D           */
  CLAWHIP_RP_SHALLOW = 3,   /* shallow_roe_with_efix_2D */
  CLAWHIP_RP_ADVECTION = 4, /* advection_2D             */
  CLAWHIP_RP_EULER3 = 5,    /* euler_3D                 */
  CLAWHIP_RP_VC_ADVECTION = 6, /* variable-coefficient advection of rp->meqn (1..4) fields,
                                * edge velocities in 2 aux fields
                                * (examples/advection_reaction_2d/rpn2_vc_advection_vector.f90);
                                * its rpt2 is a dummy: dimension-split steps, or unsplit
                                * with transverse_waves = 0 */
  CLAWHIP_RP_VC_ADVECTION_2D = 7, /* vc_advection_2D (clawpack/riemann): scalar variable-
                                   * coefficient advection, the same rpn2 with a transverse
                                   * solver (rpt2_vc_advection.f90); 2D steps */
  CLAWHIP_RP_VC_ACOUSTICS = 8,     /* vc_acoustics_2D (clawpack/riemann): acoustics with
                                   * density and sound speed per cell in 2 aux fields;
                                   * 2D steps */
  CLAWHIP_RP_SHALLOW_HLLE = 9,     /* shallow_hlle_2D (clawpack/riemann): HLLE shallow
                                   * water; no transverse solver (split or unsplit with
                                   * transverse_waves = 0; SharpClaw) */
  CLAWHIP_RP_EULER5 = 10,          /* euler_5wave_2D (clawpack/riemann; in-tree copy
                                   * development/rp_approaches/rpn2_euler_5wave.f,
                                   * rpt2_euler_5wave.f): euler_4wave_2D plus a passive
                                   * tracer q[4]; classic 2D steps, SharpClaw with
                                   * char_decomp 0 */
  CLAWHIP_RP_SHALLOW_FWAVE = 11    /* shallow_bathymetry_fwave_2D (clawpack/riemann): shallow
                                   * water over bathymetry b = aux[0], f-waves (set
                                   * rp.fwave = 1); no transverse solver (split, or unsplit
                                   * with transverse_waves = 0); classic 2D steps */
};

/* Boundary codes: pyclaw.BC (pyclaw/solver.py:13-19) plus GIVEN. */
enum { CLAWHIP_BC_EXTRAP = 1, CLAWHIP_BC_PERIODIC = 2, CLAWHIP_BC_WALL = 3, CLAWHIP_BC_GIVEN = 4 };

/* Problem parameters: the values state.problem_data puts in the Riemann
 * solver's /cparam/ common block (pyclaw/state.py:206-231). */
typedef struct {
  int kind;
  int efix;
  double rho, bulk, cc, zz; /* acoustics */
  double gamma;             /* Euler */
  double grav;              /* shallow water */
  double u, v;              /* advection */
  int meqn;                 /* CLAWHIP_RP_VC_ADVECTION: number of fields */
  int fwave;                /* the solver returns f-waves: step2 / step2ds's use_fwave
                             * (classic/flux2.f90:134-160: dsign(1, s) in place of |s| in
                             * the second-order correction); classic 2D steps */
  double dry_tolerance;     /* shallow water over bathymetry */
  double sea_level;
} clawhip_rp;

typedef struct {
  int mx, my, mz;
  int ghosts;
  long long field_stride, row_stride, slab_stride;
} clawhip_geom;

/* Device-resident time-step control: the state Solver.evolve_to_time keeps
 * between steps (pyclaw/solver.py:554-666).  The caller fills it (t, dt of the
 * first attempt after its own get_dt, tstart, tend, CFL targets, max_steps,
 * dt_variable, cur = buffer holding the current state; zeros elsewhere) and
 * reads it back. */
typedef struct {
  double t, dt, tstart, tend;
  double cfl, cfl_desired, cfl_max, dt_max;
  double cflmax, dtmin, dtmax;
  int dt_variable, done, cur, attempts;
  int max_steps, accepted, nrejected, pad0;
  unsigned int pad2, pad1;
  unsigned long long cfl_bits;
} clawhip_stepctl;

/* Device-resident control of SharpClaw's SSP33 step (clawhip_sharp2_ssp33_run):
 * what SharpClawSolver.step and Solver.evolve_to_time keep between attempts
 * (pyclaw/sharpclaw/solver.py:283-323, 622-685; pyclaw/solver.py:535-666).
 *   t, dt, dt_old, tstart, tend; cfl = the CFL object's cached maximum;
 *   cfl_desired, cfl_max, dt_max; cflmax, dtmin, dtmax = solver.status;
 *   have_dt_old: dt_old is set; accept_prev = solver.accept_step;
 *   done: 0 running, 1 reached tend, 3 max_steps attempts made; cur: the buffer
 *   (q0/q1, comp0/comp1) holding the state; attempts, max_steps, accepted,
 *   nrejected; judge: the learned indicator's path (flags + positivity judge);
 *   age, every: its flag-refresh cadence; judge_steps, fallbacks: its judge's
 *   counts; run_flags, run1, run23, run_fb: the next attempt's gates (flags,
 *   stage 1, stages 2-3, the full-WENO5 recomputation); the rest is the
 *   launches' scratch (zero on entry). */
typedef struct clawhip_sharpctl {
  double t, dt, dt_old, tstart, tend;
  double cfl, cfl_desired, cfl_max, dt_max;
  double cflmax, dtmin, dtmax;
  int have_dt_old, accept_prev, done, cur;
  int attempts, max_steps, accepted, nrejected;
  int judge, age, every, judge_steps;
  int fallbacks, run_flags, run1, run23;
  int run_fb, pad0;
  unsigned int pad1, nbad;
  unsigned long long cfl1_bits, cfl23_bits, cflfb_bits;
} clawhip_sharpctl;

/* Library / device information: writes the HIP device count; 0 on success. */
int clawhip_device_count(int *count);

/* *cfl_dev = 0.0 (the reference's cfl.set_global_max(0.), pyclaw/cfl.py:18-19),
 * as a kernel so that it may be captured in a HIP graph. */
int clawhip_cfl_reset(double *cfl_dev, void *stream);

/* One unsplit 2D classic step: replaces fmod.step2(...) as called by
 * ClawSolver2D.step_hyperbolic (pyclaw/classic/solver.py:539-541; Fortran
 * classic/step2.f90:2-250 with flux2.f90, limiter.f90, philim.f90) together with
 * the ghost fill that precedes it (_apply_bcs, pyclaw/solver.py:351-437).
 *   method[7] as ClawSolver._set_method builds it (classic/solver.py:173-193);
 *   mthlim[num_waves] as _set_mthlim (classic/solver.py:162-171);
 *   q_in (geometry gin) -> q_out (geometry gout, ghosts = 0); distinct buffers;
 *   comp_in -> comp_out: Kahan compensation (CLAWHIP_F32K only; may alias);
 *   *cfl_dev = max(*cfl_dev, Courant number of this step) (the `cfl` output).
 * reflect[d] is solver.reflect_index[d] (classic/solver.py:445).
 * aux (geometry gaux, method[6] fields, codes aux_bc_lower/upper as for
 * clawhip_step2_split) for the solvers that read aux (kinds 6 and 7, at least
 * their 2 velocity fields, at most 3 fields), else NULL.  method[5] =
 * index_capa + 1 > 0 names a capacity field among them: the capa branches of
 * step2.f90:101-105,152-163,183-187,232-244 with flux2.f90's per-cell dtdx1d. */
int clawhip_step2_unsplit(const clawhip_rp *rp, int dtype, const void *q_in,
                          const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                          const void *comp_in, void *comp_out,
                          const int *bc_lower, const int *bc_upper,
                          const int *reflect, int num_ghost, const int *method,
                          const int *mthlim, double dx, double dy, double dt,
                          const void *aux, const clawhip_geom *gaux,
                          const int *aux_bc_lower, const int *aux_bc_upper,
                          double *cfl_dev, void *stream);

/* Up to nsteps attempted unsplit steps with the control on the device: each
 * launch reads dt and the current buffer (q0/q1, comp0/comp1 by ctl->cur) from
 * ctl_dev, steps, and its last wavefront applies accept_reject_step and the next
 * get_dt (pyclaw/solver.py:524-552, 621-657) -- the arithmetic of
 * evolve_to_time(solution, tend).  Launches after ctl->done != 0 exit at once:
 * 1 reached tend, 2 CFL too large with fixed dt, 3 max_steps attempts made.
 * The launches contain no host synchronisation and may be captured in a graph.
 * Boundary codes 1..3. */
int clawhip_step2_unsplit_run(const clawhip_rp *rp, int dtype, const clawhip_geom *g,
                              const int *bc_lower, const int *bc_upper, const int *reflect,
                              int num_ghost, const int *method, const int *mthlim, double dx,
                              double dy, void *q0, void *q1, void *comp0, void *comp1,
                              clawhip_stepctl *ctl_dev, int nsteps, void *stream);

/* One dimension-split 2D step: both sweeps of ClawSolver2D.step_hyperbolic with
 * dimensional_split=True (pyclaw/classic/solver.py:523-535), i.e. fmod.step2ds
 * ids=1 then ids=2 (classic/step2ds.f90:2-264; the x-sweep covers the ghost rows,
 * step2ds.f90:96).  method[2] must be -1.
 *   work:      device buffer of num_eqn*(mx+2*num_ghost)*(my+2*num_ghost) elements
 *              of the state's type (the intermediate qbc between the sweeps);
 *   comp_work: the same for the Kahan compensation (CLAWHIP_F32K only);
 *   aux:       the solver's aux fields (state.aux; required exactly when the
 *              solver has them, method[6] = their number), planar like q in the
 *              state's type, geometry gaux; aux_bc_lower/upper: the
 *              solver.aux_bc codes (a wall mirrors without negating).
 * q_out either compact (gout->ghosts = 0: the new interior) or the whole qbc
 * (gout->ghosts = num_ghost, with q_in storing its ghosts too): then every cell
 * the reference's two sweeps write is written as they write it -- the x-sweep's
 * ghost rows, the y-sweep's ghost columns -- and the corners keep q_in's values,
 * so a caller that keeps qbc between steps (pyclaw/solver.py:341-343; custom BC
 * functions that set only some ghost cells) sees the reference's ghost values;
 * work / comp_work are then unused (the sweeps run in q_out / comp_out). */
int clawhip_step2_split(const clawhip_rp *rp, int dtype, const void *q_in,
                        const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                        const void *comp_in, void *comp_out, void *work, void *comp_work,
                        const int *bc_lower, const int *bc_upper, const int *reflect,
                        int num_ghost, const int *method, const int *mthlim, double dx,
                        double dy, double dt, const void *aux, const clawhip_geom *gaux,
                        const int *aux_bc_lower, const int *aux_bc_upper, double *cfl_dev,
                        void *stream);

/* One unsplit 3D classic step: replaces fmod.step3(...) as called by
 * ClawSolver3D.step_hyperbolic with dimensional_split=False
 * (pyclaw/classic/solver.py:694-699; Fortran classic/step3.f90:2-597 with
 * flux3.f90:2-594: transverse and double-transverse propagation), FP64
 * (dtype CLAWHIP_F64), rp->kind = CLAWHIP_RP_EULER3 (euler_3D's rpn3, rpt3,
 * rptt3).  method[2] = 0, 10, 11, 20, 21 or 22 (ClawSolver3D.trans_inc = 11,
 * trans_cor = 22; 20-22 need method[1] = 2); no aux, no capacity.  q_in with
 * geometry gin (compact, or with num_ghost stored ghosts for BC code 4), q_out
 * compact; work: clawhip_step3_unsplit_work_elems(meqn, mx, my, mz) doubles
 * (the slices' flux3 results).  Bitwise equal to the reference's step3. */
long long clawhip_step3_unsplit_work_elems(int meqn, int mx, int my, int mz);
int clawhip_step3_unsplit(const clawhip_rp *rp, int dtype, const void *q_in,
                          const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                          void *work, const int *bc_lower, const int *bc_upper,
                          const int *reflect, int num_ghost, const int *method,
                          const int *mthlim, double dx, double dy, double dz, double dt,
                          double *cfl_dev, void *stream);

/* One dimension-split 3D step: ClawSolver3D.step_hyperbolic with
 * dimensional_split=True (pyclaw/classic/solver.py:676-692), i.e. fmod.step3ds
 * idir=1,2,3 (classic/step3ds.f90:112-378, flux3.f90:168-263).  method[2] = -1;
 * rp->kind = CLAWHIP_RP_EULER3; q_in compact (ghosts = 0).
 *   work / comp_work: num_eqn*(mx+2ng)*(my+2ng)*(mz+2ng) elements. */
int clawhip_step3_split(const clawhip_rp *rp, int dtype, const void *q_in,
                        const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                        const void *comp_in, void *comp_out, void *work, void *comp_work,
                        const int *bc_lower, const int *bc_upper, const int *reflect,
                        int num_ghost, const int *method, const int *mthlim, double dx,
                        double dy, double dz, double dt, double *cfl_dev, void *stream);

/* SharpClaw 2D operator with a fused Runge-Kutta stage.  Computes
 * D = dt*L(Y) -- SharpClawSolver2D.dq_hyperbolic (sharpclaw/solver.py:848-898),
 * i.e. the ghost fill plus fmod.flux2 (sharpclaw/flux2.f90, flux1.f90; WENO5 via
 * weno.f90:5-83 for char_decomp 0, reconstruct.f90:201-299 for char_decomp 2) --
 * and writes one stage of SharpClawSolver.step (sharpclaw/solver.py:283-430):
 *   y_mode   0: Y = P;  1: Y = P + s*Q;  2: Y = P + (s*Q)/6.
 *   out_mode 0: out = D/dt                      (dq_dt, solver.py:293)
 *            1: out = 0.75*A + 0.25*(Y + D)     (SSP33 stage 2, solver.py:318)
 *            2: out = 1/3*A + 2/3*(Y + D)       (SSP33 stage 3, solver.py:323)
 *            3: out = Y + D/6.                  (SSP104 stages, solver.py:415,425)
 *            4: out = A + (0.6*Y + 0.1*D)       (SSP104 last stage, solver.py:430)
 *            5: FP32 SSP33 stage 2: inc = s*B + D; out = A + 0.25*inc; aux_out = inc/6
 *            6: FP32 SSP33 stage 3: Kahan update of A by aux_in + 2/3*D (comp_in->comp_out)
 *            7: out = D                         (fmod.flux2's dq on the interior)
 *            8: out = 0.5*((A + Y) + D)         (SSP22, solver.py:388-393)
 * All arrays compact (ghosts = 0) with geometry g, except P when gin (may be
 * NULL) gives it >= 3 stored ghosts: then the caller has filled P's ghosts
 * (custom boundary functions run on the host) and boundary code 4 (GIVEN)
 * reads them; that form takes y_mode 0 and no flags.  num_ghost is 3 (WENO5).
 * Boundary codes 1..3, or 4 with a ghosted P.
 * src (may be NULL, geometry g): a source increment added to D before the
 * stage combination, dq = dq_hyperbolic + dq_src (sharpclaw/solver.py:537-540).
 * *cfl_dev = max(*cfl_dev, CFL of this evaluation).
 * flags (may be NULL): per-interface troubled flags from clawhip_mlp_flags.
 * An interface whose flag and whose two neighbours' flags (along its normal) are
 * 0 uses WENO5's linear 5th-order weights; the others use the WENO5 selected by
 * char_decomp.  NULL flags = WENO5 everywhere (the reference's operator).
 * With flags: flag_list / flag_count from clawhip_mlp_flag_list of the same
 * flags (flag_list sized as there, (mx+1)*my + mx*(my+1) entries: the full-WENO5
 * pass reads its first entries before it reads the count), and recon, a
 * device buffer of 2*num_eqn*((mx+1)*my + mx*(my+1))
 * elements of the state's type that takes the full-WENO5 edge values of the
 * listed interfaces (work space; its contents on entry do not matter).
 * aux (geometry gaux, num_aux fields, codes aux_bc_lower/upper 1..4 as for
 * clawhip_step2_split; 4 needs >= 3 stored ghosts): the solver's aux fields for
 * the solvers that read them (kinds 6, 7, 8; char_decomp 0, no flags), else
 * NULL.  The solve at an interface takes the aux of its two cells, the solve
 * inside a cell that cell's aux (sharpclaw/flux1.f90:131, 174-181). */
int clawhip_sharp2_stage(const clawhip_rp *rp, int dtype, int char_decomp,
                         const clawhip_geom *g, const int *bc_lower, const int *bc_upper,
                         const int *reflect, int y_mode, int out_mode, double dt, double s,
                         double dx, double dy, const void *P, const clawhip_geom *gin,
                         const void *Q, const void *A, const void *B, const void *aux_in,
                         const void *comp_in, void *out, void *aux_out, void *comp_out,
                         const void *src, const unsigned char *flags, const int *flag_list,
                         const int *flag_count, void *recon, const void *aux,
                         const clawhip_geom *gaux, const int *aux_bc_lower,
                         const int *aux_bc_upper, int num_aux, double *cfl_dev, void *stream);

/* nsteps attempts of SharpClawSolver's SSP33 step (sharpclaw/solver.py:283-323)
 * within Solver.evolve_to_time (solver.py:554-666), controlled on the device:
 * each stage kernel reads dt and the current buffer from ctl_dev and runs only
 * while its gate is set; a one-wavefront kernel after stage 1 applies get_dt
 * and the CFL rescale, one after stage 3 accept_reject_step, t and the next
 * attempt's gates.  Launches after ctl->done != 0 do nothing.  No host
 * synchronisation: the launches may be captured in a HIP graph.
 *   q0/q1 (comp0/comp1 for CLAWHIP_F32K): the state, compact planar, current
 *   one by ctl->cur; dq_dt, r2: work of the state's size; aux23: FP32 work of
 *   the state's size (CLAWHIP_F32K; may be NULL for F64).
 *   flags (NULL: full WENO5, no judge): the learned indicator's path -- flags
 *   of the state refreshed by ctl's cadence (mlp_params, flat, logit_threshold
 *   as clawhip_mlp_flags), their list (flag_list, flag_count as
 *   clawhip_mlp_flag_list), recon (as clawhip_sharp2_stage), and the judge:
 *   a step that leaves a field of positive_mask not > 0 (or not finite) is
 *   recomputed with full WENO5 into the same buffers (dq_dt_full: work of the
 *   state's size).  The recomputation is launched once, after the nsteps
 *   attempts: a failed judgement closes the gates of the attempts after it, so
 *   one call completes between 1 and nsteps attempts -- the caller reads
 *   ctl->attempts and calls again for the rest.  Boundary codes 1..3.
 *   stage_events (may be NULL; not while the stream is captured): 6*nsteps
 *   hipEvent_t handles, recorded before and after the march launch of stages
 *   1, 2, 3 of each attempt (timing). */
int clawhip_sharp2_ssp33_run(const clawhip_rp *rp, int dtype, int char_decomp,
                             const clawhip_geom *g, const int *bc_lower, const int *bc_upper,
                             const int *reflect, double dx, double dy, void *q0, void *q1,
                             void *comp0, void *comp1, void *dq_dt, void *r2, void *aux23,
                             unsigned char *flags, int *flag_list, int *flag_count, void *recon,
                             void *dq_dt_full, const float *mlp_params, float flat,
                             float logit_threshold, unsigned positive_mask,
                             clawhip_sharpctl *ctl_dev, int nsteps, void *const *stage_events,
                             void *stream);

/* The elementwise middle of SSP104 (sharpclaw/solver.py:418-419):
 *   q = q/25. + 0.36*s1;  s1 = 15.*q - 5.*s1   over n elements. */
int clawhip_ssp104_mid(int dtype, void *q, void *s1, long long n, void *stream);

/* The time-integrator updates that take no operator evaluation, over n
 * elements (SharpClawSolver.step, sharpclaw/solver.py:306-307 Euler, 328-348
 * RK, 351-374 SSPLMMk2/k3 and 376-384 LMM), with the reference's grouping;
 * c[5] are the FP64 scalars the solver computed:
 *   mode 0: out = q + c1*u                                 (Euler: c1 = dt, u = dq_dt;
 *                                                           RK: one a[i,j] or b[j] term)
 *   mode 1: out = c0*(c1*q + c2*u) + c3*R                  (SSPLMMk2)
 *   mode 2: out = (c0*(c1*q + c2*u) + c3*R) + c4*v         (SSPLMMk3)
 *   mode 3: out = c1*u                                     (RK: dt*dq_dt)
 *   mode 4: out = c1*u + c2*v                              (LMM: the newest term)
 *   mode 5: out = q + (c1*u + c2*v)                        (LMM: each older term)
 * R and v may be NULL where the mode does not read them; out may alias q. */
int clawhip_lmm_update(int dtype, int mode, long long n, const void *q, const void *u,
                       const void *R, const void *v, const double *c, void *out, void *stream);

/* ---- ensembles: many independent unsplit 2D problems of uneven sizes (one
 * Riemann solver and method), each with its own dt and step controller,
 * advanced by the same launches.  No counterpart in the reference (which steps
 * one grid per Solver); each member's arithmetic is that of
 * clawhip_step2_unsplit_run on its own. */
typedef struct {
  int mx, my;                  /* compact planar (meqn, my, mx) arrays            */
  void *q0, *q1;               /* ping-pong state (current one: ctl->cur)         */
  void *comp0, *comp1;         /* Kahan compensation (CLAWHIP_F32K), else NULL    */
  double dx, dy;
  clawhip_stepctl *ctl;        /* this member's device controller                 */
} clawhip_ens_member;

/* Bytes of the device table clawhip_ensemble_prepare writes for n members. */
long long clawhip_ensemble_table_bytes(int n);

/* Validates the members and writes the launch table (member order, wavefront
 * offsets, method) into table_dev (device memory of clawhip_ensemble_table_bytes
 * bytes; a synchronous host-to-device copy).  rows: rows per wavefront (0 =
 * chosen from the occupancy and the total work).  waves[3] receives the
 * wavefronts of the step launch, the member count and the wavefronts per
 * workgroup (4; the caller may set 1 or 2 instead), to pass to
 * clawhip_ensemble_run.  Boundary codes 1..3, num_ghost 2, method/mthlim as
 * clawhip_step2_unsplit. */
int clawhip_ensemble_prepare(const clawhip_rp *rp, int dtype, const clawhip_ens_member *members,
                             int n, const int *bc_lower, const int *bc_upper, const int *reflect,
                             int num_ghost, const int *method, const int *mthlim, int rows,
                             void *table_dev, long long *waves);

/* nsteps attempted steps of every member (one launch per step; members whose
 * controller is done skip).  Capturable in a graph. */
int clawhip_ensemble_run(const clawhip_rp *rp, int dtype, const void *table_dev,
                         const long long *waves, int nsteps, void *stream);

/* ---- entry points on ghost-storing arrays: one-to-one with the reference's
 * f2py kernels, for a `solver.fmod` drop-in (pyclaw_amd/fmod).  The caller has
 * filled the ghost cells (the reference's _apply_bcs); every side is BC_GIVEN.
 * Arrays are planar with `num_ghost` stored ghost layers (gin/gout->ghosts ==
 * num_ghost), FP64 only.
 *
 * One dimension-split sweep, replacing step2ds (classic/step2ds.f90:2-264;
 * called at classic/solver.py:527-533).  ids 1: x-sweep q_in -> q_out (distinct);
 * every cell of q_out is written: interior columns of all rows 1-ng..my+ng
 * updated, ghost columns copied.  ids 2: y-sweep in place (q_in == q_out),
 * columns 1-ng..mx+ng swept, interior cells updated.  *cfl_dev = max(*cfl_dev,
 * CFL of the sweep).  aux: the caller-filled auxbc (ghosts stored, as q), or NULL
 * for solvers without aux fields. */
int clawhip_step2_sweep(const clawhip_rp *rp, int dtype, int ids, const void *q_in,
                        const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                        int num_ghost, const int *method, const int *mthlim, double dx, double dy,
                        double dt, const void *aux, const clawhip_geom *gaux, double *cfl_dev,
                        void *stream);

/* One sweep of step3ds (classic/step3ds.f90:2-378 with flux3.f90; called at
 * classic/solver.py:680-690).  idir 1 out of place (every cell written),
 * idir 2 and 3 in place; slices 0..m+1 as the reference. */
int clawhip_step3_sweep(const clawhip_rp *rp, int dtype, int idir, const void *q_in,
                        const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                        int num_ghost, const int *method, const int *mthlim, double dx, double dy,
                        double dz, double dt, double *cfl_dev, void *stream);

/* SharpClaw's flux2 (sharpclaw/flux2.f90:2-74; called at sharpclaw/solver.py:893):
 * dq = dt*L(q) on the interior (gout compact) from q with >= 3 stored, filled
 * ghost layers; char_decomp 0 (component WENO5) or 2 (characteristic).  aux
 * (geometry gaux with >= 3 filled ghost layers): the solver's aux fields for the
 * solvers that read them, else NULL. */
int clawhip_sharp2_flux(const clawhip_rp *rp, int dtype, int char_decomp, const void *q,
                        const clawhip_geom *gin, void *dq, const clawhip_geom *gout, double dt,
                        double dx, double dy, const void *aux, const clawhip_geom *gaux,
                        double *cfl_dev, void *stream);

/* ---- learned troubled-interface indicator (no counterpart in the reference;
 * the reference always runs the full WENO5 of reconstruct.f90:201-299).
 * A 6->32->32->1 MLP in BF16 MFMA with FP32 accumulation.  Parameters: one FP32
 * array of clawhip_mlp_param_count() floats: W1[16][32] (row 6 = bias, rows 7..15
 * zero), W2[32][32], b2[32], w3[32], b3.
 * Samples are (interface, component): x-interfaces (i = 1..mx+1, j = 1..my)
 * then y-interfaces (i = 1..mx, j = 1..my+1), row-major, times num_eqn. */
int clawhip_mlp_param_count(void);

/* Features (8 floats per sample: the min-max normalised 6-cell stencil, 1, 0)
 * and labels (1 = the component's WENO5 weights depart from the linear weights by
 * more than wtol while the stencil range exceeds flat) of every sample of q. */
int clawhip_mlp_features(int dtype, const void *q, const clawhip_geom *g, const int *bc_lower,
                         const int *bc_upper, const int *reflect, int meqn, float flat,
                         float wtol, float *feat, unsigned char *label, void *stream);

/* One training step on the samples batch_idx[0..nbatch) of feat/label (nsamples
 * rows; indices outside [0, nsamples) are skipped): grad += d(loss)/d(params)
 * (weighted binary cross-entropy, positive weight pos_weight, mean over nbatch),
 * *loss += loss.  step > 0 then applies Adam step `step` (lr, 0.9, 0.999, 1e-8) to
 * params and clears grad; step == 0 only accumulates (gradient checks). */
int clawhip_mlp_train_step(const float *feat, long long nsamples, const unsigned char *label,
                           const int *batch_idx, int nbatch, float *params, float *grad,
                           float *adam_m, float *adam_v, int step, float lr, float pos_weight,
                           float *loss, void *stream);

/* Logits of n samples (feat as from clawhip_mlp_features). */
int clawhip_mlp_logits(const float *feat, long long n, const float *params, float *logits,
                       void *stream);

/* The interfaces whose flag, or a neighbour's along the normal, is set -- the
 * ones clawhip_sharp2_stage reconstructs with the full WENO5 -- as interface ids
 * in the flags' layout, in no particular order; *count = their number.  list
 * holds up to (mx+1)*my + mx*(my+1) entries; flags is 4-byte aligned (read as
 * dwords). */
int clawhip_mlp_flag_list(const unsigned char *flags, int mx, int my, int *list, int *count,
                          void *stream);

/* Per-interface flags of a state: 1 when any component whose stencil range exceeds
 * flat has a logit > logit_threshold. */
int clawhip_mlp_flags(int dtype, const void *q, const clawhip_geom *g, const int *bc_lower,
                      const int *bc_upper, const int *reflect, int meqn, float flat,
                      const float *params, float logit_threshold, unsigned char *flags,
                      void *stream);

/* ---- FP64 reductions over a state (no counterpart in the reference's kernels:
 * they replace the numpy sums of q PyClaw scripts take for conservation checks,
 * e.g. sum(q)*dx*dy, and give the indicator path its positivity and L1 checks).
 * Deterministic (fixed reduction order).  work: device scratch of
 * clawhip_sums_work_doubles(nfields) doubles, nfields = meqn + 1 for
 * clawhip_state_sums and 2 for clawhip_l1_diff. */
long long clawhip_sums_work_doubles(int nfields);

/* out[f] = sum over the interior cells of field f (f < meqn <= 7), in FP64;
 * out[meqn] = number of cells where a field whose bit is set in positive_mask is
 * <= 0 or not finite.  2D (mz = 1) or 3D geometry, any stored ghost width. */
int clawhip_state_sums(int dtype, const void *q, const clawhip_geom *g, int meqn,
                       unsigned positive_mask, double *out, double *work, void *stream);

/* clawhip_state_sums without the work argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It has no counterpart in the reference.
 * The per-block partials go to a fixed array in the library's device memory
 * (2048 x 8 doubles = 128 KiB per device, held until the process ends); the
 * call itself allocates nothing.  meqn above CLAWHIP_SUMS_NOALLOC_MAX_MEQN is
 * rejected with CLAWHIP_E_ARG.  Results equal clawhip_state_sums bit for bit.
 * Every call on one device shares that array: calls that overlap on one device
 * (two streams, or two host threads) race and must be serialised by the caller. */
#define CLAWHIP_SUMS_NOALLOC_MAX_MEQN 7
int clawhip_state_sums_noalloc(int dtype, const void *q, const clawhip_geom *g, int meqn,
                               unsigned positive_mask, double *out, void *stream);

/* out[0] = sum |a[i] - b[i]|, out[1] = sum |b[i]| over n elements, in FP64. */
int clawhip_l1_diff(int dtype, const void *a, const void *b, long long n, double *out,
                    double *work, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* CLAWHIP_H */
