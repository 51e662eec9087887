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

int clawhip_step2_unsplit(const clawhip_rp* rp, int dtype, const void* q_in,
                          const clawhip_geom* gin, void* q_out, const clawhip_geom* gout,
                          const void* comp_in, void* comp_out, const int* bc_lower, const int* bc_upper,
                          const int* reflect, int num_ghost, const int* method,
                          const int* mthlim, double dx, double dy, double dt,
                          double* cfl_dev, void* stream) {
  if (!rp || !q_in || !q_out || !cfl_dev || !bc_lower || !bc_upper || !reflect || !method ||
      !mthlim)
    return CLAWHIP_E_ARG;
  if (!geom_ok(gin, 2) || !geom_ok(gout, 2) || gout->ghosts != 0) return CLAWHIP_E_ARG;
  if (gin->mx != gout->mx || gin->my != gout->my) return CLAWHIP_E_ARG;
  if (num_ghost != 2 || gout->mx < num_ghost || gout->my < num_ghost) return CLAWHIP_E_ARG;
  if (q_in == q_out) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F32K && (!comp_in || !comp_out)) return CLAWHIP_E_ARG;
  if (dtype != CLAWHIP_F64 && dtype != CLAWHIP_F32K) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  if (method[2] < 0 || method[2] > 2) return CLAWHIP_E_METHOD;   // unsplit only
  if (method[5] != 0 || method[6] != 0) return CLAWHIP_E_METHOD;  // no capa / aux
  for (int d = 0; d < 2; d++) {
    int lo = bc_lower[d], hi = bc_upper[d];
    if (lo < 1 || lo > 4 || hi < 1 || hi > 4) return CLAWHIP_E_ARG;
    if ((lo == CLAWHIP_BC_GIVEN || hi == CLAWHIP_BC_GIVEN) && gin->ghosts < num_ghost)
      return CLAWHIP_E_ARG;
  }
  return launch_unsplit(*rp, dtype, q_in, to_geom(gin), q_out, to_geom(gout), comp_in, comp_out,
                        bc_lower,
                        bc_upper, reflect, num_ghost, method, mthlim, dx, dy, dt,
                        (unsigned long long*)cfl_dev, (hipStream_t)stream);
}

static int check_common(const clawhip_rp* rp, int dtype, const void* q_in, const void* q_out,
                        const void* comp_in, const void* comp_out, double* cfl_dev,
                        const int* bc_lower, const int* bc_upper, const int* reflect,
                        const int* method, const int* mthlim, int ndim, int num_ghost,
                        const clawhip_geom* gin, const clawhip_geom* gout) {
  if (!rp || !q_in || !q_out || !cfl_dev || !bc_lower || !bc_upper || !reflect || !method ||
      !mthlim)
    return CLAWHIP_E_ARG;
  if (!geom_ok(gin, ndim) || !geom_ok(gout, ndim) || gout->ghosts != 0) return CLAWHIP_E_ARG;
  if (gin->mx != gout->mx || gin->my != gout->my) return CLAWHIP_E_ARG;
  if (ndim == 3 && gin->mz != gout->mz) return CLAWHIP_E_ARG;
  if (num_ghost != 2 || gout->mx < num_ghost || gout->my < num_ghost) return CLAWHIP_E_ARG;
  if (ndim == 3 && gout->mz < num_ghost) return CLAWHIP_E_ARG;
  if (q_in == q_out) return CLAWHIP_E_ARG;
  if (dtype != CLAWHIP_F64 && dtype != CLAWHIP_F32K) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F32K && (!comp_in || !comp_out)) return CLAWHIP_E_ARG;
  if (method[1] != 1 && method[1] != 2) return CLAWHIP_E_METHOD;
  if (method[5] != 0 || method[6] != 0) return CLAWHIP_E_METHOD;
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
                        double dy, double dt, double* cfl_dev, void* stream) {
  int rc = check_common(rp, dtype, q_in, q_out, comp_in, comp_out, cfl_dev, bc_lower, bc_upper,
                        reflect, method, mthlim, 2, num_ghost, gin, gout);
  if (rc) return rc;
  if (!work || (dtype == CLAWHIP_F32K && !comp_work)) return CLAWHIP_E_ARG;
  if (!(dx > 0.0) || !(dy > 0.0) || !(dt >= 0.0)) return CLAWHIP_E_ARG;
  if (method[2] != -1) return CLAWHIP_E_METHOD;
  return launch_split2(*rp, dtype, q_in, to_geom(gin), q_out, to_geom(gout), comp_in, comp_out,
                       work, comp_work, bc_lower, bc_upper, reflect, num_ghost, method, mthlim,
                       dx, dy, dt, (unsigned long long*)cfl_dev, (hipStream_t)stream);
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

}  // extern "C"
