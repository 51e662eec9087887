// capi.hip -- the extern "C" entry points of libclawhip.so (see include/clawhip.h).
#include <hip/hip_runtime.h>

#include "../../include/clawhip.h"
#include "launch.hpp"

using namespace claw;

namespace {

__global__ void k_set_zero_u64(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

bool geom_ok(const clawhip_geom* g, int ndim) {
  if (!g || g->mx < 1 || g->my < 1 || g->ghosts < 0) return false;
  if (ndim == 3 && g->mz < 1) return false;
  if (g->row_stride < g->mx + 2LL * g->ghosts) return false;
  if (g->field_stride < 1) return false;
  return true;
}

bool valid_dtype(int d) { return d == CLAWHIP_F64 || d == CLAWHIP_F32K || d == CLAWHIP_F64_FAST; }

// the element type of the arrays: CLAWHIP_F64_FAST stores FP64 like CLAWHIP_F64
int storage_dtype(int d) { return d == CLAWHIP_F64_FAST ? CLAWHIP_F64 : d; }

Geom to_geom(const clawhip_geom* g) {
  Geom o;
  o.mx = g->mx; o.my = g->my; o.mz = g->mz; o.g = g->ghosts;
  o.fs = g->field_stride; o.rs = g->row_stride; o.ss = g->slab_stride;
  return o;
}

}  // namespace

extern "C" {

int clawhip_device_count(int* count) {
  if (!count) return CLAWHIP_E_ARG;
  hipError_t e = hipGetDeviceCount(count);
  return e == hipSuccess ? 0 : (int)e;
}

int clawhip_cfl_reset(double* cfl_dev, void* stream) {
  if (!cfl_dev) return CLAWHIP_E_ARG;
  hipLaunchKernelGGL(k_set_zero_u64, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     (unsigned long long*)cfl_dev);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

// aux: NULL, or the solver's aux fields (planar like q, geometry gaux with the
// interior of q; stored ghosts only with CLAWHIP_BC_GIVEN aux codes).
static int make_aux_in(const void* aux, const clawhip_geom* gaux, const int* alo, const int* ahi,
                  const clawhip_geom* gq, int num_ghost, AuxIn& out) {
  if (!aux) return 0;
  if (!gaux || !alo || !ahi || !geom_ok(gaux, 2)) return CLAWHIP_E_ARG;
  if (gaux->mx != gq->mx || gaux->my != gq->my) return CLAWHIP_E_ARG;
  for (int d = 0; d < 2; d++) {
    if (alo[d] < 1 || alo[d] > 4 || ahi[d] < 1 || ahi[d] > 4) return CLAWHIP_E_ARG;
    if ((alo[d] == CLAWHIP_BC_GIVEN || ahi[d] == CLAWHIP_BC_GIVEN) && gaux->ghosts < num_ghost)
      return CLAWHIP_E_ARG;
    out.lo[d] = alo[d];
    out.hi[d] = ahi[d];
  }
  out.p = aux;
  out.g = to_geom(gaux);
  out.g.ss = 0;
  return 0;
}

static int rp_num_aux(const clawhip_rp* rp) {
  if (rp->kind == CLAWHIP_RP_SHALLOW_FWAVE) return 1;
  return rp->kind == CLAWHIP_RP_VC_ADVECTION || rp->kind == CLAWHIP_RP_VC_ADVECTION_2D ||
                 rp->kind == CLAWHIP_RP_VC_ACOUSTICS
             ? 2
             : 0;
}

// solvers without a transverse solver (rpt2): split, or transverse_waves = 0
static bool rp_no_rpt(const clawhip_rp* rp) {
  return rp->kind == CLAWHIP_RP_VC_ADVECTION || rp->kind == CLAWHIP_RP_SHALLOW_HLLE ||
         rp->kind == CLAWHIP_RP_SHALLOW_FWAVE;
}

int clawhip_step2_unsplit(const clawhip_rp* rp, int dtype, const void* q_in,
                          const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                          const void* comp_in, void* comp_out,
                          const int* bc_lower, const int* bc_upper,
                          const int* reflect, int num_ghost, const int* method,
                          const int* mthlim, double dx, double dy, double dt,
                          const void* aux, const clawhip_geom* gaux, const int* aux_bc_lower,
                          const int* aux_bc_upper, double* cfl_dev, void* stream) {
  if (!rp || !q_in || !q_out || !cfl_dev || !bc_lower || !bc_upper || !reflect || !method ||
      !mthlim)
    return CLAWHIP_E_ARG;
  if (!geom_ok(gin, 2) || !geom_ok(gout, 2) || gout->ghosts != 0) return CLAWHIP_E_ARG;
  if (gin->mx != gout->mx || gin->my != gout->my) return CLAWHIP_E_ARG;
  if (num_ghost != 2 || gout->mx < num_ghost || gout->my < num_ghost) return CLAWHIP_E_ARG;
  if (q_in == q_out) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F32K && (!comp_in || !comp_out)) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  if (method[2] < 0 || method[2] > 2) return CLAWHIP_E_METHOD;   // unsplit only
  // aux fields: the solver's own (velocities first) and at most one more; a
  // capacity field (method[5] = index_capa + 1) is one of them
  const int na = rp_num_aux(rp);
  if (na == 0 ? method[6] != 0 : (method[6] < na || method[6] > 3)) return CLAWHIP_E_METHOD;
  if (method[5] < 0 || method[5] > method[6]) return CLAWHIP_E_METHOD;
  // solvers without a transverse solver
  if (rp_no_rpt(rp) && method[2] != 0) return CLAWHIP_E_METHOD;
  for (int d = 0; d < 2; d++) {
    int lo = bc_lower[d], hi = bc_upper[d];
    if (lo < 1 || lo > 4 || hi < 1 || hi > 4) return CLAWHIP_E_ARG;
    if ((lo == CLAWHIP_BC_GIVEN || hi == CLAWHIP_BC_GIVEN) && gin->ghosts < num_ghost)
      return CLAWHIP_E_ARG;
  }
  AuxIn ax;
  int rc = make_aux_in(aux, gaux, aux_bc_lower, aux_bc_upper, gout, num_ghost, ax);
  if (rc) return rc;
  if ((na > 0) != (ax.p != nullptr)) return CLAWHIP_E_ARG;
  return launch_unsplit(*rp, dtype, q_in, to_geom(gin), q_out, to_geom(gout), comp_in, comp_out,
                        bc_lower, bc_upper, reflect, num_ghost, method, mthlim, dx, dy, dt,
                        (unsigned long long*)cfl_dev, (hipStream_t)stream, ax);
}

static int check_common(const clawhip_rp* rp, int dtype, const void* q_in, const void* q_out,
                        const void* comp_in, const void* comp_out, double* cfl_dev,
                        const int* bc_lower, const int* bc_upper, const int* reflect,
                        const int* method, const int* mthlim, int ndim, int num_ghost,
                        const clawhip_geom* gin, const clawhip_geom* gout, int num_aux = 0,
                        bool ghost_out = false) {
  if (!rp || !q_in || !q_out || !cfl_dev || !bc_lower || !bc_upper || !reflect || !method ||
      !mthlim)
    return CLAWHIP_E_ARG;
  if (!geom_ok(gin, ndim) || !geom_ok(gout, ndim)) return CLAWHIP_E_ARG;
  if (gout->ghosts != 0 && !(ghost_out && gout->ghosts == num_ghost && gin->ghosts == num_ghost))
    return CLAWHIP_E_ARG;
  if (rp->fwave && ndim != 2) return CLAWHIP_E_METHOD;
  if (gin->mx != gout->mx || gin->my != gout->my) return CLAWHIP_E_ARG;
  if (ndim == 3 && gin->mz != gout->mz) return CLAWHIP_E_ARG;
  if (num_ghost != 2 || gout->mx < num_ghost || gout->my < num_ghost) return CLAWHIP_E_ARG;
  if (ndim == 3 && gout->mz < num_ghost) return CLAWHIP_E_ARG;
  if (q_in == q_out) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F32K && (!comp_in || !comp_out)) return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  // no capacity; a solver with aux fields may be given more than it reads
  if (method[5] != 0) return CLAWHIP_E_METHOD;
  if (num_aux == 0 ? method[6] != 0 : (method[6] < num_aux || method[6] > 3))
    return CLAWHIP_E_METHOD;
  for (int d = 0; d < ndim; d++) {
    int lo = bc_lower[d], hi = bc_upper[d];
    if (lo < 1 || lo > 4 || hi < 1 || hi > 4) return CLAWHIP_E_ARG;
    if ((lo == CLAWHIP_BC_GIVEN || hi == CLAWHIP_BC_GIVEN) && gin->ghosts < num_ghost)
      return CLAWHIP_E_ARG;
  }
  return 0;
}

int clawhip_step2_split(const clawhip_rp* rp, int dtype, const void* q_in,
                        const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                        const void* comp_in, void* comp_out, void* work, void* comp_work,
                        const int* bc_lower, const int* bc_upper, const int* reflect,
                        int num_ghost, const int* method, const int* mthlim, double dx,
                        double dy, double dt, const void* aux, const clawhip_geom* gaux,
                        const int* aux_bc_lower, const int* aux_bc_upper, double* cfl_dev,
                        void* stream) {
  if (!rp) return CLAWHIP_E_ARG;
  int rc = check_common(rp, dtype, q_in, q_out, comp_in, comp_out, cfl_dev, bc_lower, bc_upper,
                        reflect, method, mthlim, 2, num_ghost, gin, gout, rp_num_aux(rp), true);
  if (rc) return rc;
  if (gout->ghosts == 0 && (!work || (dtype == CLAWHIP_F32K && !comp_work))) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  if (method[2] != -1) return CLAWHIP_E_METHOD;
  AuxIn ax;
  rc = make_aux_in(aux, gaux, aux_bc_lower, aux_bc_upper, gout, num_ghost, ax);
  if (rc) return rc;
  if ((rp_num_aux(rp) > 0) != (ax.p != nullptr)) return CLAWHIP_E_ARG;
  return launch_split2(*rp, dtype, q_in, to_geom(gin), q_out, to_geom(gout), comp_in, comp_out,
                       work, comp_work, bc_lower, bc_upper, reflect, num_ghost, method, mthlim,
                       dx, dy, dt, ax, (unsigned long long*)cfl_dev, (hipStream_t)stream);
}

int clawhip_step3_split(const clawhip_rp* rp, int dtype, const void* q_in,
                        const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                        const void* comp_in, void* comp_out, void* work, void* comp_work,
                        const int* bc_lower, const int* bc_upper, const int* reflect,
                        int num_ghost, const int* method, const int* mthlim, double dx,
                        double dy, double dz, double dt, double* cfl_dev, void* stream) {
  int rc = check_common(rp, dtype, q_in, q_out, comp_in, comp_out, cfl_dev, bc_lower, bc_upper,
                        reflect, method, mthlim, 3, num_ghost, gin, gout);
  if (rc) return rc;
  if (!work || (dtype == CLAWHIP_F32K && !comp_work)) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  if (method[2] != -1) return CLAWHIP_E_METHOD;
  if (gin->ghosts != 0) return CLAWHIP_E_ARG;
  return launch_split3(*rp, dtype, q_in, to_geom(gin), q_out, to_geom(gout), comp_in, comp_out,
                       work, comp_work, bc_lower, bc_upper, reflect, num_ghost, method, mthlim,
                       dx, dy, dz, dt, (unsigned long long*)cfl_dev, (hipStream_t)stream);
}

long long clawhip_step3_unsplit_work_elems(int meqn, int mx, int my, int mz) {
  if (meqn < 1 || mx < 1 || my < 1 || mz < 1) return 0;
  return step3_work_elems(meqn, mx, my, mz);
}

int clawhip_step3_unsplit(const clawhip_rp* rp, int dtype, const void* q_in,
                          const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                          void* work, const int* bc_lower, const int* bc_upper,
                          const int* reflect, int num_ghost, const int* method,
                          const int* mthlim, double dx, double dy, double dz, double dt,
                          double* cfl_dev, void* stream) {
  int rc = check_common(rp, dtype, q_in, q_out, nullptr, nullptr, cfl_dev, bc_lower, bc_upper,
                        reflect, method, mthlim, 3, num_ghost, gin, gout);
  if (rc) return rc;
  if (!work) return CLAWHIP_E_ARG;
  if (dtype != CLAWHIP_F64) return CLAWHIP_E_ARG;
  if (rp->kind != CLAWHIP_RP_EULER3) return CLAWHIP_E_KIND;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  // transverse method(3) = 10 m3 + m4: 0 (none), 10, 11, 20, 21, 22; the
  // correction's transverse split (m3 = 2) needs the second-order correction
  const int t = method[2], m3 = t / 10, m4 = t - 10 * (t / 10);
  if (!(t == 0 || ((m3 == 1 || m3 == 2) && m4 >= 0 && m4 <= m3))) return CLAWHIP_E_METHOD;
  if (m3 == 2 && method[1] != 2) return CLAWHIP_E_METHOD;
  Geom g = to_geom(gin);
  return launch_step3(*rp, q_in, g, q_out, work, bc_lower, bc_upper, reflect, num_ghost, method,
                      mthlim, dx, dy, dz, dt, (unsigned long long*)cfl_dev, (hipStream_t)stream);
}

int clawhip_sharp2_stage(const clawhip_rp* rp, int dtype, int char_decomp,
                         const clawhip_geom* g, const int* bc_lower, const int* bc_upper,
                         const int* reflect, int y_mode, int out_mode, double dt, double s,
                         double dx, double dy, const void* P, const clawhip_geom* gin,
                         const void* Q, const void* A, const void* B, const void* aux_in,
                         const void* comp_in, void* out, void* aux_out, void* comp_out,
                         const void* src, const unsigned char* flags, const int* flag_list,
                         const int* flag_count, void* recon, const void* aux,
                         const clawhip_geom* gaux, const int* aux_bc_lower,
                         const int* aux_bc_upper, int num_aux, double* cfl_dev, void* stream) {
  if (!rp || !g || !bc_lower || !bc_upper || !reflect || !P || !out || !cfl_dev)
    return CLAWHIP_E_ARG;
  if (rp->fwave) return CLAWHIP_E_METHOD;   // SharpClaw's f-wave form is not built
  // P with stored ghosts (gin, BC code 4 = given): a stage input whose ghosts
  // the caller filled, read as it is (y_mode 0, no indicator flags)
  if (gin && (!geom_ok(gin, 2) || gin->mx != g->mx || gin->my != g->my || gin->ghosts < 0))
    return CLAWHIP_E_ARG;
  const bool ghosted = gin && gin->ghosts > 0;
  if (ghosted && (y_mode != 0 || flags)) return CLAWHIP_E_ARG;
  if (flags && (!flag_list || !flag_count || !recon)) return CLAWHIP_E_ARG;
  if (!geom_ok(g, 2) || g->ghosts != 0 || g->mx < 3 || g->my < 3) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (char_decomp != 0 && char_decomp != 2) return CLAWHIP_E_METHOD;
  if (y_mode < 0 || y_mode > 2 || out_mode < 0 || out_mode > 8) return CLAWHIP_E_ARG;
  if (y_mode != 0 && !Q) return CLAWHIP_E_ARG;
  if ((out_mode == 1 || out_mode == 2 || (out_mode >= 4 && out_mode <= 6) || out_mode == 8) && !A)
    return CLAWHIP_E_ARG;
  if (out_mode == 5 && (!B || !aux_out || dtype != CLAWHIP_F32K)) return CLAWHIP_E_ARG;
  if (out_mode == 6 && (!aux_in || !comp_in || !comp_out || dtype != CLAWHIP_F32K))
    return CLAWHIP_E_ARG;
  if (out == P || (Q && out == Q) || (A && out == A)) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt > 0.0)) return CLAWHIP_E_ARG;
  for (int d = 0; d < 2; d++) {
    for (int bc : {bc_lower[d], bc_upper[d]}) {
      if (bc == CLAWHIP_BC_GIVEN ? !ghosted || gin->ghosts < 3 : bc < 1 || bc > 3)
        return CLAWHIP_E_ARG;
    }
  }
  // aux: the solver's fields (at least those it reads, at most 3)
  const int na = rp_num_aux(rp);
  if (na > 0 && (num_aux < na || num_aux > 3)) return CLAWHIP_E_ARG;
  AuxIn ax;
  int rc = make_aux_in(aux, gaux, aux_bc_lower, aux_bc_upper, g, 3, ax);
  if (rc) return rc;
  if ((na > 0) != (ax.p != nullptr)) return CLAWHIP_E_ARG;
  FlagSet fs;
  fs.flags = flags;
  fs.list = flag_list;
  fs.count = flag_count;
  fs.recon = recon;
  return launch_sharp_stage(*rp, dtype, char_decomp, to_geom(g), to_geom(gin ? gin : g), bc_lower,
                            bc_upper, reflect, y_mode, out_mode, dt, s, dx, dy, P, Q, A, B,
                            aux_in, comp_in, out, aux_out, comp_out, (unsigned long long*)cfl_dev,
                            fs, (hipStream_t)stream, ax, src);
}

int clawhip_sharp2_ssp33_run(const clawhip_rp* rp, int dtype, int char_decomp,
                             const clawhip_geom* g, const int* bc_lower, const int* bc_upper,
                             const int* reflect, double dx, double dy, void* q0, void* q1,
                             void* comp0, void* comp1, void* dq_dt, void* r2, void* aux23,
                             unsigned char* flags, int* flag_list, int* flag_count, void* recon,
                             void* dq_dt_full, const float* mlp_params, float flat,
                             float logit_threshold, unsigned positive_mask,
                             clawhip_sharpctl* ctl_dev, int nsteps, void* const* stage_events,
                             void* stream) {
  static_assert(sizeof(clawhip_sharpctl) == sizeof(SharpCtl), "clawhip_sharpctl layout");
  if (!rp || !g || !bc_lower || !bc_upper || !reflect || !q0 || !q1 || !dq_dt || !r2 ||
      !ctl_dev || nsteps < 0)
    return CLAWHIP_E_ARG;
  if (rp->fwave) return CLAWHIP_E_METHOD;
  if (!geom_ok(g, 2) || g->ghosts != 0 || g->mx < 3 || g->my < 3) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (char_decomp != 0 && char_decomp != 2) return CLAWHIP_E_METHOD;
  if (q0 == q1) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F32K && (!comp0 || !comp1 || comp0 == comp1 || !aux23))
    return CLAWHIP_E_ARG;
  if (flags && (!flag_list || !flag_count || !recon || !dq_dt_full || !mlp_params))
    return CLAWHIP_E_ARG;
  if (rp_num_aux(rp) > 0) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0)) return CLAWHIP_E_ARG;
  for (int d = 0; d < 2; d++)
    if (bc_lower[d] < 1 || bc_lower[d] > 3 || bc_upper[d] < 1 || bc_upper[d] > 3)
      return CLAWHIP_E_ARG;
  SharpRunBufs b;
  b.q[0] = q0; b.q[1] = q1; b.comp[0] = comp0; b.comp[1] = comp1;
  b.dqdt = dq_dt; b.dqdt_fb = dq_dt_full; b.r2 = r2; b.aux23 = aux23;
  b.flags = flags; b.flag_list = flag_list; b.flag_count = flag_count; b.recon = recon;
  b.params = mlp_params; b.flat = flat; b.threshold = logit_threshold;
  b.pos_mask = positive_mask;
  b.stage_events = stage_events;
  return launch_sharp_ssp33_run(*rp, dtype, char_decomp, to_geom(g), bc_lower, bc_upper, reflect,
                               dx, dy, b, (SharpCtl*)ctl_dev, nsteps, (hipStream_t)stream);
}

int clawhip_lmm_update(int dtype, int mode, long long n, const void* q, const void* u,
                       const void* R, const void* v, const double* c, void* out,
                       void* stream) {
  if (!q || !u || !c || !out || n < 0 || mode < 0 || mode > 5) return CLAWHIP_E_ARG;
  if ((mode == 1 || mode == 2) && !R) return CLAWHIP_E_ARG;
  if ((mode == 2 || mode == 4 || mode == 5) && !v) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  return launch_lmm_update(dtype, mode, n, q, u, R, v, c, out, (hipStream_t)stream);
}

int clawhip_ssp104_mid(int dtype, void* q, void* s1, long long n, void* stream) {
  if (!q || !s1 || n < 0 || q == s1) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  return launch_ssp104_mid(dtype, q, s1, n, (hipStream_t)stream);
}

int clawhip_step2_unsplit_run(const clawhip_rp* rp, int dtype, const clawhip_geom* g,
                              const int* bc_lower, const int* bc_upper, const int* reflect,
                              int num_ghost, const int* method, const int* mthlim, double dx,
                              double dy, void* q0, void* q1, void* comp0, void* comp1,
                              clawhip_stepctl* ctl_dev, int nsteps, void* stream) {
  if (!rp || !g || !bc_lower || !bc_upper || !reflect || !method || !mthlim || !q0 || !q1 ||
      !ctl_dev || nsteps < 0)
    return CLAWHIP_E_ARG;
  if (!geom_ok(g, 2) || g->ghosts != 0 || num_ghost != 2 || g->mx < 2 || g->my < 2)
    return CLAWHIP_E_ARG;
  if (q0 == q1) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F32K && (!comp0 || !comp1 || comp0 == comp1)) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0)) return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  if (method[2] < 0 || method[2] > 2) return CLAWHIP_E_METHOD;
  if (method[5] != 0 || method[6] != 0) return CLAWHIP_E_METHOD;
  if (rp_no_rpt(rp) && method[2] != 0) return CLAWHIP_E_METHOD;
  for (int d = 0; d < 2; d++)
    if (bc_lower[d] < 1 || bc_lower[d] > 3 || bc_upper[d] < 1 || bc_upper[d] > 3)
      return CLAWHIP_E_ARG;
  static_assert(sizeof(clawhip_stepctl) == sizeof(StepCtl), "clawhip_stepctl layout");
  CtlBuffers b;
  b.q[0] = q0; b.q[1] = q1; b.comp[0] = comp0; b.comp[1] = comp1;
  b.dx = dx; b.dy = dy;
  return launch_unsplit_ctl(*rp, dtype, to_geom(g), bc_lower, bc_upper, reflect, num_ghost,
                            method, mthlim, (StepCtl*)ctl_dev, b, nsteps, (hipStream_t)stream);
}

// ---- ensembles
long long clawhip_ensemble_table_bytes(int n) { return n < 1 ? 0 : ensemble_table_bytes(n); }

int clawhip_ensemble_prepare(const clawhip_rp* rp, int dtype, const clawhip_ens_member* members,
                             int n, const int* bc_lower, const int* bc_upper, const int* reflect,
                             int num_ghost, const int* method, const int* mthlim, int rows,
                             void* table_dev, long long* waves) {
  if (!rp || !members || n < 1 || !bc_lower || !bc_upper || !reflect || !method || !mthlim ||
      !table_dev || !waves || rows < 0)
    return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (num_ghost != 2) return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  if (method[2] < 0 || method[2] > 2) return CLAWHIP_E_METHOD;
  if (method[5] != 0 || method[6] != 0) return CLAWHIP_E_METHOD;
  if (rp->kind < CLAWHIP_RP_ACOUSTICS || rp->kind > CLAWHIP_RP_ADVECTION) return CLAWHIP_E_KIND;
  for (int d = 0; d < 2; d++)
    if (bc_lower[d] < 1 || bc_lower[d] > 3 || bc_upper[d] < 1 || bc_upper[d] > 3)
      return CLAWHIP_E_ARG;
  for (int k = 0; k < n; k++) {
    const clawhip_ens_member& m = members[k];
    if (m.mx < 2 || m.my < 2 || !m.q0 || !m.q1 || !m.ctl) return CLAWHIP_E_ARG;
    if (m.q0 == m.q1) return CLAWHIP_E_ARG;
    if (dtype == CLAWHIP_F32K && (!m.comp0 || !m.comp1 || m.comp0 == m.comp1))
      return CLAWHIP_E_ARG;
    if (!(m.dx > 0.0) || !(m.dy > 0.0)) return CLAWHIP_E_ARG;
  }
  return ensemble_prepare(*rp, members, n, bc_lower, bc_upper, reflect, num_ghost, method, mthlim,
                          rows, table_dev, waves);
}

int clawhip_ensemble_run(const clawhip_rp* rp, int dtype, const void* table_dev,
                         const long long* waves, int nsteps, void* stream) {
  if (!rp || !table_dev || !waves || nsteps < 0 || waves[0] < 0 || waves[1] < 0)
    return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  return ensemble_run(*rp, dtype, table_dev, waves, nsteps, (hipStream_t)stream);
}

// ---- sweeps and operator on ghost-storing arrays (the f2py entry points) ----
static int check_sweep(const clawhip_rp* rp, int dtype, int ndim, int dir, const void* q_in,
                       const clawhip_geom* gin, const void* q_out, const clawhip_geom* gout,
                       int num_ghost, const int* method, const int* mthlim, double* cfl_dev,
                       int num_aux = 0) {
  if (!rp || !q_in || !q_out || !gin || !gout || !method || !mthlim || !cfl_dev)
    return CLAWHIP_E_ARG;
  if (dtype != CLAWHIP_F64) return CLAWHIP_E_ARG;
  if (dir < 1 || dir > ndim) return CLAWHIP_E_ARG;
  if (!geom_ok(gin, ndim) || !geom_ok(gout, ndim)) return CLAWHIP_E_ARG;
  if (num_ghost != 2 || gin->ghosts != num_ghost || gout->ghosts != num_ghost) return CLAWHIP_E_ARG;
  if (gin->mx != gout->mx || gin->my != gout->my || (ndim == 3 && gin->mz != gout->mz))
    return CLAWHIP_E_ARG;
  if (gout->mx < num_ghost || gout->my < num_ghost || (ndim == 3 && gout->mz < num_ghost))
    return CLAWHIP_E_ARG;
  if ((dir == 1) == (q_in == q_out)) return CLAWHIP_E_ARG;   // sweep 1 out of place, 2/3 in place
  if (dir > 1 && (gin->field_stride != gout->field_stride || gin->row_stride != gout->row_stride ||
                  gin->slab_stride != gout->slab_stride))
    return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  if (method[2] != -1 || method[5] != 0) return CLAWHIP_E_METHOD;
  if (num_aux == 0 ? method[6] != 0 : (method[6] < num_aux || method[6] > 3))
    return CLAWHIP_E_METHOD;
  if (rp->fwave && ndim != 2) return CLAWHIP_E_METHOD;
  return 0;
}

int clawhip_step2_sweep(const clawhip_rp* rp, int dtype, int ids, const void* q_in,
                        const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                        int num_ghost, const int* method, const int* mthlim, double dx, double dy,
                        double dt, const void* aux, const clawhip_geom* gaux, double* cfl_dev,
                        void* stream) {
  if (!rp) return CLAWHIP_E_ARG;
  int rc = check_sweep(rp, dtype, 2, ids, q_in, gin, q_out, gout, num_ghost, method, mthlim,
                       cfl_dev, rp_num_aux(rp));
  if (rc) return rc;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  const int given[2] = {CLAWHIP_BC_GIVEN, CLAWHIP_BC_GIVEN};
  AuxIn ax;
  rc = make_aux_in(aux, gaux, given, given, gout, num_ghost, ax);
  if (rc) return rc;
  if ((rp_num_aux(rp) > 0) != (ax.p != nullptr)) return CLAWHIP_E_ARG;
  Geom a = to_geom(gin), b = to_geom(gout);
  a.ss = 0;   // 2D: the sweep kernels add (k-1+g)*ss with k = 1
  b.ss = 0;
  return launch_sweep2(*rp, ids, q_in, a, q_out, b, num_ghost, method, mthlim,
                       ids == 1 ? dt / dx : dt / dy, ax, (unsigned long long*)cfl_dev,
                       (hipStream_t)stream);
}

int clawhip_step3_sweep(const clawhip_rp* rp, int dtype, int idir, const void* q_in,
                        const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                        int num_ghost, const int* method, const int* mthlim, double dx, double dy,
                        double dz, double dt, double* cfl_dev, void* stream) {
  int rc = check_sweep(rp, dtype, 3, idir, q_in, gin, q_out, gout, num_ghost, method, mthlim,
                       cfl_dev);
  if (rc) return rc;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  const double dtdx = idir == 1 ? dt / dx : (idir == 2 ? dt / dy : dt / dz);
  return launch_sweep3(*rp, idir, q_in, to_geom(gin), q_out, to_geom(gout), num_ghost, method,
                       mthlim, dtdx, (unsigned long long*)cfl_dev, (hipStream_t)stream);
}

int clawhip_sharp2_flux(const clawhip_rp* rp, int dtype, int char_decomp, const void* q,
                        const clawhip_geom* gin, void* dq, const clawhip_geom* gout, double dt,
                        double dx, double dy, const void* aux, const clawhip_geom* gaux,
                        double* cfl_dev, void* stream) {
  if (!rp || !q || !gin || !dq || !gout || !cfl_dev) return CLAWHIP_E_ARG;
  if (rp->fwave) return CLAWHIP_E_METHOD;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  if (char_decomp != 0 && char_decomp != 2) return CLAWHIP_E_ARG;
  if (!geom_ok(gin, 2) || !geom_ok(gout, 2) || gin->ghosts < 3 || gout->ghosts != 0)
    return CLAWHIP_E_ARG;
  if (gin->mx != gout->mx || gin->my != gout->my || gout->mx < 3 || gout->my < 3)
    return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  const int given[2] = {CLAWHIP_BC_GIVEN, CLAWHIP_BC_GIVEN};
  const int refl[2] = {1, 2};
  AuxIn ax;
  int rc = make_aux_in(aux, gaux, given, given, gout, 3, ax);
  if (rc) return rc;
  if ((rp_num_aux(rp) > 0) != (ax.p != nullptr)) return CLAWHIP_E_ARG;
  return launch_sharp_stage(*rp, dtype, char_decomp, to_geom(gout), to_geom(gin), given, given,
                            refl, 0, 7, dt, 0.0, dx, dy, q, nullptr, nullptr, nullptr, nullptr,
                            nullptr, dq, nullptr, nullptr, (unsigned long long*)cfl_dev, FlagSet(),
                            (hipStream_t)stream, ax);
}

int clawhip_mlp_param_count(void) { return 16 * 32 + 32 * 32 + 32 + 32 + 1; }

int clawhip_mlp_features(int dtype, const void* q, const clawhip_geom* g, const int* bc_lower,
                         const int* bc_upper, const int* reflect, int meqn, float flat,
                         float wtol, float* feat, unsigned char* label, void* stream) {
  if (!q || !g || !bc_lower || !bc_upper || !reflect || !feat || !label) return CLAWHIP_E_ARG;
  if (!geom_ok(g, 2) || g->ghosts != 0 || g->mx < 3 || g->my < 3 || meqn < 1) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  dtype = storage_dtype(dtype);
  for (int d = 0; d < 2; d++)
    if (bc_lower[d] < 1 || bc_lower[d] > 3 || bc_upper[d] < 1 || bc_upper[d] > 3)
      return CLAWHIP_E_ARG;
  return launch_features_labels(dtype, q, to_geom(g), bc_lower, bc_upper, reflect, meqn, flat,
                                wtol, feat, label, (hipStream_t)stream);
}

int clawhip_mlp_train_step(const float* feat, long long nsamples, const unsigned char* label,
                           const int* batch_idx, int nbatch, float* params, float* grad,
                           float* adam_m, float* adam_v, int step, float lr, float pos_weight,
                           float* loss, void* stream) {
  if (!feat || !label || !batch_idx || nbatch < 1 || nsamples < 1 || !params || !grad || !loss)
    return CLAWHIP_E_ARG;
  if (step < 0 || (step > 0 && (!adam_m || !adam_v))) return CLAWHIP_E_ARG;
  return launch_mlp_train(feat, nsamples, label, batch_idx, nbatch, params, grad, adam_m, adam_v,
                          step, lr, pos_weight, loss, (hipStream_t)stream);
}

int clawhip_mlp_logits(const float* feat, long long n, const float* params, float* logits,
                       void* stream) {
  if (!feat || !params || !logits || n < 0) return CLAWHIP_E_ARG;
  if (n == 0) return 0;
  return launch_mlp_logits(feat, n, params, logits, (hipStream_t)stream);
}

int clawhip_mlp_flag_list(const unsigned char* flags, int mx, int my, int* list, int* count,
                          void* stream) {
  if (!flags || !list || !count || mx < 3 || my < 3) return CLAWHIP_E_ARG;
  return launch_flag_list(flags, mx, my, list, count, (hipStream_t)stream);
}

int clawhip_mlp_flags(int dtype, const void* q, const clawhip_geom* g, const int* bc_lower,
                      const int* bc_upper, const int* reflect, int meqn, float flat,
                      const float* params, float logit_threshold, unsigned char* flags,
                      void* stream) {
  if (!q || !g || !bc_lower || !bc_upper || !reflect || !params || !flags) return CLAWHIP_E_ARG;
  if (!geom_ok(g, 2) || g->ghosts != 0 || g->mx < 3 || g->my < 3 || meqn < 1) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  dtype = storage_dtype(dtype);
  for (int d = 0; d < 2; d++)
    if (bc_lower[d] < 1 || bc_lower[d] > 3 || bc_upper[d] < 1 || bc_upper[d] > 3)
      return CLAWHIP_E_ARG;
  return launch_mlp_flags(dtype, q, to_geom(g), bc_lower, bc_upper, reflect, meqn, flat, params,
                          logit_threshold, flags, (hipStream_t)stream);
}

long long clawhip_sums_work_doubles(int nfields) {
  return nfields < 1 ? 0 : 2048LL * nfields;
}

int clawhip_state_sums(int dtype, const void* q, const clawhip_geom* g, int meqn,
                       unsigned positive_mask, double* out, double* work, void* stream) {
  if (!q || !g || !out || !work || meqn < 1 || meqn > 7) return CLAWHIP_E_ARG;
  if (!geom_ok(g, g->mz > 1 ? 3 : 2) || g->mz < 1) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  dtype = storage_dtype(dtype);
  return launch_state_sums(dtype, q, to_geom(g), meqn, positive_mask, out, work,
                           (hipStream_t)stream);
}

int clawhip_state_sums_noalloc(int dtype, const void* q, const clawhip_geom* g, int meqn,
                               unsigned positive_mask, double* out, void* stream) {
  if (!q || !g || !out) return CLAWHIP_E_ARG;
  if (meqn < 1 || meqn > CLAWHIP_SUMS_NOALLOC_MAX_MEQN) return CLAWHIP_E_ARG;
  if (!geom_ok(g, g->mz > 1 ? 3 : 2) || g->mz < 1) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  dtype = storage_dtype(dtype);
  return launch_state_sums_noalloc(dtype, q, to_geom(g), meqn, positive_mask, out,
                                   (hipStream_t)stream);
}

int clawhip_l1_diff(int dtype, const void* a, const void* b, long long n, double* out,
                    double* work, void* stream) {
  if (!a || !b || !out || !work || n < 0) return CLAWHIP_E_ARG;
  if (!valid_dtype(dtype)) return CLAWHIP_E_ARG;
  dtype = storage_dtype(dtype);
  return launch_l1_diff(dtype, a, b, n, out, work, (hipStream_t)stream);
}

}  // extern "C"
