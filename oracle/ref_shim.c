/* ref_shim.c -- TEST INFRASTRUCTURE: the callbacks the reference's own Fortran
 * step routines take as arguments, and C entry points that call those routines.
 *
 * The reference's classic step (classic/step2.f90, step2ds.f90, step3ds.f90) and
 * SharpClaw operator (sharpclaw/flux2.f90, flux1.f90) receive the Riemann solvers
 * as procedure arguments (f2py intent(callback), classic/flux2.f90:53-55); the
 * user supplies `evec` for characteristic WENO (examples/euler_1d/evec.f90 is the
 * reference's own example of one).  This file supplies those from the oracle's
 * Riemann solvers (oracle.c) so the reference's compiled step can be run on the
 * same inputs as the oracle and the HIP path.  Built only into oracle/_ref/ by
 * oracle/ref.mk; the product never loads it.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static orc_rp g_rp;
static int g_evec_mode = 0;   /* 0: fixed direction; 1: count flux1 calls of a 2D flux2 */
static int g_evec_dir = 1;
static int g_evec_calls = 0;
static int g_evec_xcalls = 0; /* number of x-slices flux2 makes: my + 2 */

static int g_fwave = 0;        /* use_fwave passed to the step routines */

void shim_set_rp(const orc_rp *p) { g_rp = *p; }
void shim_set_fwave(int on) { g_fwave = on != 0; }
void shim_set_evec_dir(int dir) { g_evec_mode = 0; g_evec_dir = dir; }
void shim_set_evec_count(int xcalls) {
  g_evec_mode = 1; g_evec_calls = 0; g_evec_xcalls = xcalls;
}

/* rpn2(ixy,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq):
 * problem i (2-mbc..mx+mbc) has left state qr(:,i-1), right state ql(:,i). */
void shim_rpn2_(const int *ixy, const int *maxm, const int *meqn, const int *mwaves,
                const int *maux, const int *mbc, const int *mx, const double *ql,
                const double *qr, const double *auxl, const double *auxr, double *wave,
                double *s, double *amdq, double *apdq) {
  (void)maxm; (void)maux; (void)auxl; (void)auxr;
  const int me = *meqn, mw = *mwaves, ng = *mbc;
  for (int i = 2 - ng; i <= *mx + ng; i++) {
    long ci = i + ng - 1, cm = ci - 1;
    orc_rpn(&g_rp, *ixy, qr + (long)me * cm, ql + (long)me * ci, wave + (long)me * mw * ci,
            s + (long)mw * ci, amdq + (long)me * ci, apdq + (long)me * ci);
  }
}

/* rpt2(ixy,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3,asdq,bmasdq,bpasdq) */
void shim_rpt2_(const int *ixy, const int *imp, const int *maxm, const int *meqn,
                const int *mwaves, const int *maux, const int *mbc, const int *mx,
                const double *ql, const double *qr, const double *aux1, const double *aux2,
                const double *aux3, const double *asdq, double *bmasdq, double *bpasdq) {
  (void)maxm; (void)mwaves; (void)maux; (void)aux1; (void)aux2; (void)aux3;
  const int me = *meqn, ng = *mbc;
  for (int i = 2 - ng; i <= *mx + ng; i++) {
    long ci = i + ng - 1, cm = ci - 1;
    orc_rpt(&g_rp, *ixy, *imp, qr + (long)me * cm, ql + (long)me * ci, asdq + (long)me * ci,
            bmasdq + (long)me * ci, bpasdq + (long)me * ci);
  }
}

/* rpn2 for solvers with aux, from the oracle: problem i takes the aux of
 * cell i-1 from auxr and of cell i from auxl (both are flux2's aux2). */
void shim_rpn2_aux_(const int *ixy, const int *maxm, const int *meqn, const int *mwaves,
                    const int *maux, const int *mbc, const int *mx, const double *ql,
                    const double *qr, const double *auxl, const double *auxr, double *wave,
                    double *s, double *amdq, double *apdq) {
  (void)maxm;
  const int me = *meqn, mw = *mwaves, ng = *mbc, ma = *maux;
  for (int i = 2 - ng; i <= *mx + ng; i++) {
    long ci = i + ng - 1, cm = ci - 1;
    orc_rpn_aux(&g_rp, *ixy, qr + (long)me * cm, ql + (long)me * ci, auxr + (long)ma * cm,
                auxl + (long)ma * ci, wave + (long)me * mw * ci, s + (long)mw * ci,
                amdq + (long)me * ci, apdq + (long)me * ci);
  }
}

/* rpt2 for solvers with aux: aux1/aux2/aux3 are the three slices
 * (maux values per cell); problem i's solver reads cell i1 = i-2+imp. */
void shim_rpt2_aux_(const int *ixy, const int *imp, const int *maxm, const int *meqn,
                    const int *mwaves, const int *maux, const int *mbc, const int *mx,
                    const double *ql, const double *qr, const double *aux1, const double *aux2,
                    const double *aux3, const double *asdq, double *bmasdq, double *bpasdq) {
  (void)maxm; (void)mwaves;
  const int me = *meqn, ng = *mbc, ma = *maux;
  for (int i = 2 - ng; i <= *mx + ng; i++) {
    long ci = i + ng - 1, cm = ci - 1;
    orc_rpt_aux(&g_rp, *ixy, *imp, qr + (long)me * cm, ql + (long)me * ci, ma,
                aux1 + (long)ma * cm, aux2 + (long)ma * cm, aux3 + (long)ma * cm,
                asdq + (long)me * ci, bmasdq + (long)me * ci, bpasdq + (long)me * ci);
  }
}

/* rpn3 has the rpn2 argument list (flux3.f90:201-202). */
void shim_rpn3_(const int *ixyz, const int *maxm, const int *meqn, const int *mwaves,
                const int *maux, const int *mbc, const int *mx, const double *ql,
                const double *qr, const double *auxl, const double *auxr, double *wave,
                double *s, double *amdq, double *apdq) {
  shim_rpn2_(ixyz, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq);
}

/* rpt3(ixyz,icoor,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3,asdq,
 *      bmasdq,bpasdq) (flux3.f90:272-289): problem i splits asdq(:,i) with the
 * states qr(:,i-1) | ql(:,i). */
void shim_rpt3_(const int *ixyz, const int *icoor, const int *imp, const int *maxm,
                const int *meqn, const int *mwaves, const int *maux, const int *mbc,
                const int *mx, const double *ql, const double *qr, const double *aux1,
                const double *aux2, const double *aux3, const double *asdq, double *bmasdq,
                double *bpasdq) {
  (void)imp; (void)maxm; (void)mwaves; (void)maux; (void)aux1; (void)aux2; (void)aux3;
  const int me = *meqn, ng = *mbc;
  for (int i = 2 - ng; i <= *mx + ng; i++) {
    long ci = i + ng - 1, cm = ci - 1;
    orc_rpt3(&g_rp, *ixyz, *icoor, qr + (long)me * cm, ql + (long)me * ci,
             asdq + (long)me * ci, bmasdq + (long)me * ci, bpasdq + (long)me * ci);
  }
}

/* rptt3(ixyz,icoor,imp,impt,maxm,...): the same split of its input. */
void shim_rptt3_(const int *ixyz, const int *icoor, const int *imp, const int *impt,
                 const int *maxm, const int *meqn, const int *mwaves, const int *maux,
                 const int *mbc, const int *mx, const double *ql, const double *qr,
                 const double *aux1, const double *aux2, const double *aux3,
                 const double *bsasdq, double *cmbsasdq, double *cpbsasdq) {
  (void)impt;
  shim_rpt3_(ixyz, icoor, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3,
             bsasdq, cmbsasdq, cpbsasdq);
}

/* never called on the split / tfluct_solver=False paths */
void shim_unused_(void) { abort(); }

/* evec(maxnx,num_eqn,num_ghost,mx,q,auxl,auxr,evl,evr) as flux1.f90:180 calls it:
 * q(num_eqn, maxnx+2*num_ghost) with maxnx = mx. */
void evec_(const int *maxnx, const int *num_eqn, const int *num_ghost, const int *mx,
           const double *q, const double *auxl, const double *auxr, double *evl, double *evr) {
  (void)mx; (void)auxl; (void)auxr;
  int dir = g_evec_dir;
  if (g_evec_mode == 1) dir = (g_evec_calls++ < g_evec_xcalls) ? 1 : 2;
  orc_evec(&g_rp, dir, *num_eqn, *maxnx + 2 * *num_ghost, q, evl, evr);
}

/* ---- entry points into the reference's compiled routines --------------- */
#ifdef REF_CLASSIC
typedef void (*rp_fn)(void);
extern void step2_(const int *, const int *, const int *, const int *, const int *,
                   const int *, const int *, const double *, double *, const double *,
                   const double *, const double *, const double *, const int *, const int *,
                   double *, double *, double *, double *, double *, double *, double *,
                   double *, double *, double *, double *, const int *, const int *, rp_fn,
                   rp_fn);
extern void step2ds_(const int *, const int *, const int *, const int *, const int *,
                     const int *, const int *, const double *, double *, const double *,
                     const double *, const double *, const double *, const int *,
                     const int *, double *, double *, double *, double *, double *,
                     double *, double *, double *, double *, double *, double *,
                     const int *, const int *, const int *, rp_fn, rp_fn);
extern void step3ds_(const int *, const int *, const int *, const int *, const int *,
                     const int *, const int *, const double *, double *, const double *,
                     const double *, const double *, const double *, const double *,
                     const int *, const int *, double *, double *, double *, double *,
                     double *, double *, double *, double *, double *, double *, double *,
                     double *, const int *, double *, const int *, const int *, const int *,
                     rp_fn, rp_fn, rp_fn);

/* ids = 0: unsplit step2; 1,2: step2ds sweep, with the given rpn2/rpt2 (the
 * reference argument lists) -- NULL: the oracle's solver p.  Work arrays sized
 * as ClawSolver2D._allocate_workspace does (classic/solver.py:466-494). */
double ref_step2_fns(const orc_rp *p, rp_fn rpn2, rp_fn rpt2, int ids, int meqn, int mwaves,
                     int ng, int mx, int my, const double *qold, double *qnew, double dx,
                     double dy, double dt, const int *method, const int *mthlim) {
  shim_set_rp(p);
  if (!rpn2) rpn2 = (rp_fn)shim_rpn2_;
  if (!rpt2) rpt2 = (rp_fn)shim_rpt2_;
  int maxm = mx > my ? mx : my, maux = 0, use_fwave = g_fwave;
  long L = maxm + 2 * ng;
  int mwork = (int)(L * (5 * meqn + mwaves + meqn * mwaves));
  double *buf = (double *)calloc(L * (8 * meqn + 4) + mwork + 16, sizeof(double));
  double *qadd = buf, *fadd = qadd + L * meqn, *gadd = fadd + L * meqn,
         *q1d = gadd + 2 * L * meqn, *dtdx1d = q1d + L * meqn, *dtdy1d = dtdx1d + L,
         *aux1 = dtdy1d + L, *aux2 = aux1 + L, *aux3 = aux2 + L, *work = aux3 + L;
  double aux[1] = {0.0};
  double cfl = 0.0;
  if (ids == 0)
    step2_(&maxm, &meqn, &mwaves, &maux, &ng, &mx, &my, qold, qnew, aux, &dx, &dy, &dt, method,
           mthlim, &cfl, qadd, fadd, gadd, q1d, dtdx1d, dtdy1d, aux1, aux2, aux3, work, &mwork,
           &use_fwave, rpn2, rpt2);
  else
    step2ds_(&maxm, &meqn, &mwaves, &maux, &ng, &mx, &my, qold, qnew, aux, &dx, &dy, &dt,
             method, mthlim, &cfl, qadd, fadd, gadd, q1d, dtdx1d, dtdy1d, aux1, aux2, aux3,
             work, &mwork, &ids, &use_fwave, rpn2, rpt2);
  free(buf);
  return cfl;
}

double ref_step2(const orc_rp *p, int ids, int meqn, int mwaves, int ng, int mx, int my,
                 const double *qold, double *qnew, double dx, double dy, double dt,
                 const int *method, const int *mthlim) {
  return ref_step2_fns(p, 0, 0, ids, meqn, mwaves, ng, mx, my, qold, qnew, dx, dy, dt, method,
                       mthlim);
}

/* The unsplit step2 with aux fields (method[6] = maux; method[5] = index_capa
 * for a capacity field) and a caller-supplied rpn2 (the advection_reaction_2d
 * example's, compiled from its source; NULL: the oracle's solver p); rpt2 from
 * the oracle's solver p. */
double ref_step2_aux(const orc_rp *p, rp_fn rpn2, int meqn, int mwaves, int maux, int ng,
                     int mx, int my, const double *qold, double *qnew, const double *aux,
                     double dx, double dy, double dt, const int *method, const int *mthlim) {
  shim_set_rp(p);
  int maxm = mx > my ? mx : my, use_fwave = g_fwave;
  long L = maxm + 2 * ng;
  int mwork = (int)(L * (5 * meqn + mwaves + meqn * mwaves));
  double *buf = (double *)calloc(L * (6 * meqn + 2 + 3 * maux) + mwork + 16, sizeof(double));
  double *qadd = buf, *fadd = qadd + L * meqn, *gadd = fadd + L * meqn,
         *q1d = gadd + 2 * L * meqn, *dtdx1d = q1d + L * meqn, *dtdy1d = dtdx1d + L,
         *aux1 = dtdy1d + L, *aux2 = aux1 + L * maux, *aux3 = aux2 + L * maux,
         *work = aux3 + L * maux;
  double cfl = 0.0;
  step2_(&maxm, &meqn, &mwaves, &maux, &ng, &mx, &my, qold, qnew, aux, &dx, &dy, &dt, method,
         mthlim, &cfl, qadd, fadd, gadd, q1d, dtdx1d, dtdy1d, aux1, aux2, aux3, work, &mwork,
         &use_fwave, rpn2 ? rpn2 : (rp_fn)shim_rpn2_aux_, (rp_fn)shim_rpt2_aux_);
  free(buf);
  return cfl;
}

/* step2ds (ids 1, 2) with aux fields and a caller-supplied rpn2 -- the
 * reference example's own Riemann solver, compiled from its source by ref.mk
 * (examples/advection_reaction_2d/rpn2_vc_advection_vector.f90), or NULL for
 * the oracle's solver set by shim_set_rp.
 * aux(maux, 1-ng:mx+ng, 1-ng:my+ng), ghosts filled. */
double ref_step2ds_aux(rp_fn rpn2, int ids, int meqn, int mwaves, int maux, int ng, int mx,
                       int my, const double *qold, double *qnew, const double *aux, double dx,
                       double dy, double dt, const int *method, const int *mthlim) {
  int maxm = mx > my ? mx : my, use_fwave = g_fwave;
  long L = maxm + 2 * ng;
  int mwork = (int)(L * (5 * meqn + mwaves + meqn * mwaves));
  double *buf = (double *)calloc(L * (6 * meqn + 2 + 3 * maux) + mwork + 16, sizeof(double));
  double *qadd = buf, *fadd = qadd + L * meqn, *gadd = fadd + L * meqn,
         *q1d = gadd + 2 * L * meqn, *dtdx1d = q1d + L * meqn, *dtdy1d = dtdx1d + L,
         *aux1 = dtdy1d + L, *aux2 = aux1 + L * maux, *aux3 = aux2 + L * maux,
         *work = aux3 + L * maux;
  double cfl = 0.0;
  step2ds_(&maxm, &meqn, &mwaves, &maux, &ng, &mx, &my, qold, qnew, aux, &dx, &dy, &dt,
           method, mthlim, &cfl, qadd, fadd, gadd, q1d, dtdx1d, dtdy1d, aux1, aux2, aux3,
           work, &mwork, &ids, &use_fwave, rpn2 ? rpn2 : (rp_fn)shim_rpn2_aux_,
           (rp_fn)shim_rpt2_);
  free(buf);
  return cfl;
}

extern void step3_(const int *, const int *, const int *, const int *, const int *,
                   const int *, const int *, const double *, double *, const double *,
                   const double *, const double *, const double *, const double *,
                   const int *, const int *, double *, double *, double *, double *,
                   double *, double *, double *, double *, double *, double *, double *,
                   double *, const int *, double *, const int *, const int *,
                   rp_fn, rp_fn, rp_fn);

/* The reference's unsplit step3 (classic/step3.f90) with the oracle's euler_3D
 * rpn3 / rpt3 / rptt3; qnew starts as the caller's array. */
double ref_step3(const orc_rp *p, int meqn, int mwaves, int ng, int mx, int my, int mz,
                 const double *qold, double *qnew, double dx, double dy, double dz, double dt,
                 const int *method, const int *mthlim) {
  shim_set_rp(p);
  int maxm = mx > my ? mx : my, maux = 0, use_fwave = g_fwave;
  if (mz > maxm) maxm = mz;
  long L = maxm + 2 * ng;
  int mwork = (int)(L * (31 * meqn + mwaves + meqn * mwaves));
  double *buf = (double *)calloc(L * (16 * meqn + 16) + mwork + 16, sizeof(double));
  double *qadd = buf, *fadd = qadd + L * meqn, *gadd = fadd + L * meqn,
         *hadd = gadd + 6 * L * meqn, *q1d = hadd + 6 * L * meqn, *dtdx1d = q1d + L * meqn,
         *dtdy1d = dtdx1d + L, *dtdz1d = dtdy1d + L, *aux1 = dtdz1d + L, *aux2 = aux1 + 3 * L,
         *aux3 = aux2 + 3 * L, *work = aux3 + 3 * L;
  double aux[1] = {0.0};
  double cfl = 0.0;
  step3_(&maxm, &meqn, &mwaves, &ng, &mx, &my, &mz, qold, qnew, aux, &dx, &dy, &dz, &dt,
         method, mthlim, &cfl, qadd, fadd, gadd, hadd, q1d, dtdx1d, dtdy1d, dtdz1d, aux1,
         aux2, aux3, &maux, work, &mwork, &use_fwave, (rp_fn)shim_rpn3_, (rp_fn)shim_rpt3_,
         (rp_fn)shim_rptt3_);
  free(buf);
  return cfl;
}

double ref_step3ds(const orc_rp *p, int idir, int meqn, int mwaves, int ng, int mx, int my,
                   int mz, const double *qold, double *qnew, double dx, double dy, double dz,
                   double dt, const int *method, const int *mthlim) {
  shim_set_rp(p);
  int maxm = mx > my ? mx : my, maux = 0, use_fwave = g_fwave;
  if (mz > maxm) maxm = mz;
  long L = maxm + 2 * ng;
  int mwork = (int)(L * (31 * meqn + mwaves + meqn * mwaves));
  double *buf = (double *)calloc(L * (16 * meqn + 16) + mwork + 16, sizeof(double));
  double *qadd = buf, *fadd = qadd + L * meqn, *gadd = fadd + L * meqn,
         *hadd = gadd + 6 * L * meqn, *q1d = hadd + 6 * L * meqn, *dtdx1d = q1d + L * meqn,
         *dtdy1d = dtdx1d + L, *dtdz1d = dtdy1d + L, *aux1 = dtdz1d + L, *aux2 = aux1 + 3 * L,
         *aux3 = aux2 + 3 * L, *work = aux3 + 3 * L;
  double aux[1] = {0.0};
  double cfl = 0.0;
  step3ds_(&maxm, &meqn, &mwaves, &ng, &mx, &my, &mz, qold, qnew, aux, &dx, &dy, &dz, &dt,
           method, mthlim, &cfl, qadd, fadd, gadd, hadd, q1d, dtdx1d, dtdy1d, dtdz1d, aux1,
           aux2, aux3, &maux, work, &mwork, &idir, &use_fwave, (rp_fn)shim_rpn3_,
           (rp_fn)shim_unused_, (rp_fn)shim_unused_);
  free(buf);
  return cfl;
}
#endif /* REF_CLASSIC */

#ifdef REF_SHARP
typedef void (*rp_fn)(void);
extern void ref_sharp_init(int num_dim, int lim_type, int weno_order, int char_decomp,
                           int num_waves, const double *dxs, const int *mthlim, int maxnx,
                           int num_ghost, int num_eqn);
extern void flux2_(const double *q, double *dq, const double *aux, const double *dt,
                   double *cfl, const double *t, const int *num_aux, const int *num_eqn,
                   const int *num_ghost, const int *maxnx, const int *mx, const int *my,
                   rp_fn rpn2, rp_fn tfluct2);
extern void flux1_(const double *q1d, double *dq1d, const double *aux, const double *dt,
                   double *cfl, const double *t, const int *ixyz, const int *num_aux,
                   const int *num_eqn, const int *mx, const int *num_ghost, const int *maxnx,
                   rp_fn rp, rp_fn tfluct);

/* An rpn2 (the reference argument list) to use instead of the oracle's solver
 * in the next ref_sharp_flux2 calls without aux; NULL restores the oracle's. */
static rp_fn g_sharp_rpn2 = 0;
void shim_set_sharp_rpn2(rp_fn f) { g_sharp_rpn2 = f; }

/* SharpClawSolver2D.dq_hyperbolic's call (sharpclaw/solver.py:893) with the
 * module set-up of _set_fortran_parameters (sharpclaw/solver.py:563-588). */
double ref_sharp_flux2(const orc_rp *p, int char_decomp, int meqn, int mwaves, int ng, int mx,
                       int my, const double *q, double *dq, double dx, double dy, double dt,
                       int num_aux, const double *auxq) {
  shim_set_rp(p);
  int maxm = mx > my ? mx : my;
  double dxs[2] = {dx, dy};
  int mthlim[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  ref_sharp_init(2, 2, 5, char_decomp, mwaves, dxs, mthlim, maxm + 2 * ng, ng, meqn);
  if (char_decomp == 2) shim_set_evec_count(my + 2);
  long nq = (long)meqn * (mx + 2 * ng) * (my + 2 * ng);
  memset(dq, 0, sizeof(double) * nq);
  double aux0[1] = {0.0}, t = 0.0, cfl = 0.0;
  flux2_(q, dq, num_aux > 0 ? auxq : aux0, &dt, &cfl, &t, &num_aux, &meqn, &ng, &maxm, &mx, &my,
         num_aux > 0 ? (rp_fn)shim_rpn2_aux_ : g_sharp_rpn2 ? g_sharp_rpn2 : (rp_fn)shim_rpn2_,
         (rp_fn)shim_unused_);
  return cfl;
}

/* one slice through flux1 (sharpclaw/flux1.f90), direction given */
double ref_sharp_flux1(const orc_rp *p, int char_decomp, int meqn, int mwaves, int ng, int n,
                       int ixy, const double *q1d, double *dq1d, double dxd, double dt) {
  shim_set_rp(p);
  int num_aux = 0;
  double dxs[2] = {dxd, dxd};
  int mthlim[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  ref_sharp_init(2, 2, 5, char_decomp, mwaves, dxs, mthlim, n + 2 * ng, ng, meqn);
  shim_set_evec_dir(ixy);
  double aux[1] = {0.0}, t = 0.0, cfl = 0.0;
  flux1_(q1d, dq1d, aux, &dt, &cfl, &t, &ixy, &num_aux, &meqn, &n, &ng, &n, (rp_fn)shim_rpn2_,
         (rp_fn)shim_unused_);
  return cfl;
}
#endif /* REF_SHARP */
