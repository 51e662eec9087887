/* oracle.h -- CPU restatement (FP64, serial, strict IEEE evaluation order) of the
 * PyClaw time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pyclaw_amd/, libclawhip.so)
 * links, loads or calls this code.  It is the checker that tests/, the smoke test
 * and bench.py's cpu_baseline leg compare the HIP path against.
 *
 * Array layout is the reference's: Fortran column-major with the equation index
 * fastest, ghost cells included, e.g. q(meqn, 1-ng:mx+ng, 1-ng:my+ng)
 * (classic/step2.f90:21-24).  Indices in the comments below are the reference's
 * 1-based Fortran indices.
 */
#ifndef CLAW_ORACLE_H
#define CLAW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Riemann solver identifiers (the clawpack.riemann modules the examples use). */
enum {
  ORC_RP_ACOUSTICS = 1,  /* acoustics_2D:            meqn 3, mwaves 2 */
  ORC_RP_EULER = 2,      /* euler_4wave_2D:          meqn 4, mwaves 4 */
  ORC_RP_SHALLOW = 3,    /* shallow_roe_with_efix_2D: meqn 3, mwaves 3 */
  ORC_RP_ADVECTION = 4,  /* advection_2D:            meqn 1, mwaves 1 */
  ORC_RP_EULER3 = 5,     /* euler_3D:                meqn 5, mwaves 3 */
  ORC_RP_VC_ADVECTION = 6, /* examples/advection_reaction_2d/rpn2_vc_advection_vector.f90:
                            * meqn = orc_rp.meqn, mwaves 1, 2 aux fields (edge u, v) */
  ORC_RP_VC_ADVECTION_2D = 7, /* clawpack.riemann vc_advection_2D: the same rpn2 with
                               * meqn 1, and its transverse solver rpt2 */
  ORC_RP_VC_ACOUSTICS = 8,     /* clawpack.riemann vc_acoustics_2D: meqn 3, mwaves 2,
                               * aux = (density, sound speed) per cell */
  ORC_RP_SHALLOW_HLLE = 9,     /* clawpack.riemann shallow_hlle_2D: meqn 3, mwaves 2 */
  ORC_RP_EULER5 = 10,          /* euler_5wave_2D (development/rp_approaches/
                               * rpn2_euler_5wave.f, rpt2_euler_5wave.f): meqn 5, mwaves 5 */
  ORC_RP_SHALLOW_FWAVE = 11    /* clawpack.riemann shallow_bathymetry_fwave_2D: meqn 3,
                               * mwaves 3, f-waves, aux = bathymetry b per cell */
};

/* Problem parameters: the reference passes these through the Fortran /cparam/
 * common block filled from state.problem_data (state.py:206-231). */
typedef struct {
  int kind;
  int efix;                     /* Harten-Hyman entropy fix on/off (Euler, SWE) */
  double rho, bulk, cc, zz;     /* acoustics */
  double gamma;                 /* Euler */
  double grav;                  /* shallow water */
  double u, v;                  /* advection velocities */
  int meqn;                     /* ORC_RP_VC_ADVECTION: number of fields */
  double dry_tolerance;         /* shallow water with bathymetry */
  double sea_level;
} orc_rp;

/* Boundary condition codes: pyclaw.BC (solver.py:13-19). */
enum { ORC_BC_CUSTOM = 0, ORC_BC_EXTRAP = 1, ORC_BC_PERIODIC = 2, ORC_BC_WALL = 3 };

/* --- pointwise Riemann solvers --------------------------------------------
 * qL = state left of the interface (qr(:,i-1) in the reference ABI),
 * qR = state right of it (ql(:,i)).  wave[mw*meqn + m] = wave(m,mw,i). */
void orc_rpn(const orc_rp *p, int ixy, const double *qL, const double *qR,
             double *wave, double *s, double *amdq, double *apdq);
void orc_rpt(const orc_rp *p, int ixy, int imp, const double *qL, const double *qR,
             const double *asdq, double *bmasdq, double *bpasdq);
/* The normal solve with the aux values of the left and right cells (auxL =
 * auxr(:,i-1), auxR = auxl(:,i) in the reference ABI); solvers without aux
 * fields ignore them. */
void orc_rpn_aux(const orc_rp *p, int ixy, const double *qL, const double *qR,
                 const double *auxL, const double *auxR, double *wave, double *s,
                 double *amdq, double *apdq);
/* The transverse split with aux (the rpt2 ABI's aux1/aux2/aux3 slices):
 * a1, a2, a3 point at aux(:, i-1) of the previous, this and the next slice
 * (maux values per cell); the solver reads cell i1 = i-2+imp of them. */
void orc_rpt_aux(const orc_rp *p, int ixy, int imp, const double *qL, const double *qR,
                 int maux, const double *a1, const double *a2, const double *a3,
                 const double *asdq, double *bmasdq, double *bpasdq);
/* Test hook: plant mistake k in a restated transverse solver (1: vc_advection
 * normal velocity; 2: vc_acoustics cells below/above swapped; 3: vc_acoustics
 * cell i for both imp; 4, 5: euler_5wave_2D tracer; 6: a NaN Courant-number term
 * restarts the running maximum); 0 removes it. */
void orc_set_plant(int k);
int orc_rp_meqn(int kind);
int orc_rp_mwaves(int kind);

/* --- boundary conditions (solver.py:351-519) ------------------------------
 * Fills the ghost cells of qbc in place: dimension by dimension (x, then y,
 * then z), lower side before upper side.  reflect[d] is the component negated
 * by a wall in dimension d (solver.reflect_index). */
void orc_bc2(double *qbc, int meqn, int mx, int my, int ng,
             const int *bc_lower, const int *bc_upper, const int *reflect);
void orc_bc3(double *qbc, int meqn, int mx, int my, int mz, int ng,
             const int *bc_lower, const int *bc_upper, const int *reflect);

/* --- classic Clawpack steps -----------------------------------------------
 * method[7] and mthlim[mwaves] as built by ClawSolver._set_method /
 * _set_mthlim (classic/solver.py:162-193).  qold and qnew may alias for
 * step2ds(ids=2) and step3ds(idir=2,3), exactly as the reference calls them. */
/* aux: NULL (maux = 0) or aux(maux, 1-ng:mx+ng, 1-ng:my+ng), ghosts filled;
 * method[5] = index_capa (1-based, 0: none) names the capacity field. */
double orc_step2(int meqn, int mwaves, int ng, int mx, int my,
                 const double *qold, double *qnew, double dx, double dy, double dt,
                 const int *method, const int *mthlim, int use_fwave, const orc_rp *p,
                 int maux, const double *aux);
/* aux: NULL (maux = 0) or aux(maux, 1-ng:mx+ng, 1-ng:my+ng), ghosts filled. */
double orc_step2ds(int meqn, int mwaves, int ng, int mx, int my,
                   const double *qold, double *qnew, double dx, double dy, double dt,
                   const int *method, const int *mthlim, int ids, int use_fwave,
                   const orc_rp *p, int maux, const double *aux);
/* classic/step3.f90: the unsplit 3D step (transverse method(3) = 10, 11, 20,
 * 21, 22), qnew starting as a copy of qold; rpt3 / rptt3 of euler_3D. */
double orc_step3(int meqn, int mwaves, int ng, int mx, int my, int mz, const double *qold,
                 double *qnew, double dx, double dy, double dz, double dt,
                 const int *method, const int *mthlim, int use_fwave, const orc_rp *p);
void orc_rpt3(const orc_rp *p, int ixyz, int icoor, const double *qL, const double *qR,
              const double *asdq, double *bmasdq, double *bpasdq);
double orc_step3ds(int meqn, int mwaves, int ng, int mx, int my, int mz,
                   const double *qold, double *qnew, double dx, double dy, double dz,
                   double dt, const int *method, const int *mthlim, int idir,
                   int use_fwave, const orc_rp *p);

/* --- SharpClaw semi-discrete operator (sharpclaw/flux2.f90) ----------------
 * dq (same shape as q, ghosts included) is overwritten: dq = dt * L(q) on the
 * interior, 0 elsewhere.  char_decomp 0: component WENO5 (weno.f90:5-83);
 * 2: characteristic WENO5 (reconstruct.f90:201-299) with the eigenvectors of
 * orc_evec.  Returns the CFL number. */
/* The WENO epsilon of weno5 / weno5_char (default the reference's 1.e-36 REAL(4)). */
void orc_set_weno_eps(double eps);
double orc_sharp_flux2(int meqn, int mwaves, int ng, int mx, int my,
                       const double *q, double *dq, double dx, double dy, double dt,
                       int char_decomp, const orc_rp *p, int maux, const double *aux);

/* One slice of the SharpClaw operator (sharpclaw/flux1.f90) for a slice of n
 * cells plus ng ghosts each side: q1d(meqn, 1-ng:n+ng) -> dq1d(meqn, 1-ng:n+ng).
 * Returns the slice CFL. */
double orc_sharp_flux1(int meqn, int mwaves, int ng, int n, int ixy,
                       const double *q1d, double *dq1d, double dtdx,
                       int char_decomp, const orc_rp *p, int maux, const double *aux1d);

/* Eigenvectors at each interface of a slice (evec ABI of
 * examples/euler_1d/evec.f90:2-21, with a direction argument added):
 * q(meqn, n2) 1-based; evl/evr(meqn, meqn, n2) filled for i = 2..n2 with the
 * Roe average of cells i-1 and i.  evl(ip,m,i) = row ip of R^-1. */
void orc_evec(const orc_rp *p, int ixy, int meqn, int n2, const double *q,
              double *evl, double *evr);

/* Characteristic WENO5 of one slice (reconstruct.f90:201-299).  All arrays are
 * 1-based of length n2 = size(q,2): q(meqn,n2), ql/qr(meqn,n2),
 * evl/evr(meqn,meqn,n2).  Interfaces i = ng+1 .. n2-ng+1 are reconstructed
 * (interface ng, whose projection the reference reads before it is computed,
 * is left alone; flux1 never uses it). */
void orc_weno5_char(int meqn, int ng, int n2, const double *q, double *ql, double *qr,
                    const double *evl, const double *evr);
/* Component WENO5 (weno.f90:5-83) over cells i = ng .. n2-ng+1 (1-based). */
void orc_weno5(int meqn, int ng, int n2, const double *q, double *ql, double *qr);

/* Sums over the interior cells (conservation checks), FP64, row order. */
void orc_interior_sums2(const double *q, int meqn, int mx, int my, int ng, double *sums);

#ifdef __cplusplus
}
#endif
#endif
