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

enum { CLAWHIP_F64 = 0, CLAWHIP_F32K = 1 };

/* Riemann solver identifiers: the clawpack.riemann modules the reference's
 * examples pass to ClawSolver2D(...) (e.g. examples/euler_2d/quadrants.py:80). */
enum {
  CLAWHIP_RP_ACOUSTICS = 1, /* acoustics_2D             */
  CLAWHIP_RP_EULER = 2,     /* euler_4wave_2D           */
  CLAWHIP_RP_SHALLOW = 3,   /* shallow_roe_with_efix_2D */
  CLAWHIP_RP_ADVECTION = 4, /* advection_2D             */
  CLAWHIP_RP_EULER3 = 5     /* euler_3D                 */
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
} clawhip_rp;

typedef struct {
  int mx, my, mz;
  int ghosts;
  long long field_stride, row_stride, slab_stride;
} clawhip_geom;

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
 * reflect[d] is solver.reflect_index[d] (classic/solver.py:445). */
int clawhip_step2_unsplit(const clawhip_rp *rp, int dtype, const void *q_in,
                          const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                          const void *comp_in, void *comp_out, const int *bc_lower, const int *bc_upper,
                          const int *reflect, int num_ghost, const int *method,
                          const int *mthlim, double dx, double dy, double dt,
                          double *cfl_dev, void *stream);

/* One dimension-split 2D step: both sweeps of ClawSolver2D.step_hyperbolic with
 * dimensional_split=True (pyclaw/classic/solver.py:523-535), i.e. fmod.step2ds
 * ids=1 then ids=2 (classic/step2ds.f90:2-264; the x-sweep covers the ghost rows,
 * step2ds.f90:96).  method[2] must be -1.
 *   work:      device buffer of num_eqn*(mx+2*num_ghost)*(my+2*num_ghost) elements
 *              of the state's type (the intermediate qbc between the sweeps);
 *   comp_work: the same for the Kahan compensation (CLAWHIP_F32K only). */
int clawhip_step2_split(const clawhip_rp *rp, int dtype, const void *q_in,
                        const clawhip_geom *gin, void *q_out, const clawhip_geom *gout,
                        const void *comp_in, void *comp_out, void *work, void *comp_work,
                        const int *bc_lower, const int *bc_upper, const int *reflect,
                        int num_ghost, const int *method, const int *mthlim, double dx,
                        double dy, double dt, double *cfl_dev, void *stream);

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

#ifdef __cplusplus
}
#endif
#endif /* CLAWHIP_H */
