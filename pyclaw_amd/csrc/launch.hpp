// launch.hpp -- host-side launchers behind the C ABI (capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clawhip.h"
#include "common.hpp"
#include "stepctl.hpp"

namespace claw {

template <typename T>
inline RPParams<T> to_params(const clawhip_rp& r) {
  RPParams<T> p;
  p.rho = T(r.rho); p.bulk = T(r.bulk); p.cc = T(r.cc); p.zz = T(r.zz);
  p.gamma = T(r.gamma); p.gamma1 = T(r.gamma - 1.0);
  p.grav = T(r.grav); p.u = T(r.u); p.v = T(r.v); p.efix = r.efix;
  return p;
}

inline BCSpec to_bc(const int* lo, const int* hi, const int* refl, int ndim) {
  BCSpec b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = d < ndim ? lo[d] : BC_EXTRAP;
    b.hi[d] = d < ndim ? hi[d] : BC_EXTRAP;
    b.refl[d] = d < ndim ? refl[d] : -1;
  }
  return b;
}

// Aux fields of a 2D step (p == nullptr: none): planar like q, their
// boundary codes per side (never negating).
struct AuxIn {
  const void* p = nullptr;
  Geom g{};
  int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
};

int launch_unsplit(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                   void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                   const int* lo, const int* hi, const int* refl, int ng, const int* method,
                   const int* mthlim, double dx, double dy, double dt,
                   unsigned long long* cfl_bits, hipStream_t stream, const AuxIn& aux);

int launch_split2(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dt, const AuxIn& aux, unsigned long long* cfl, hipStream_t s);

int launch_split3(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dz, double dt, unsigned long long* cfl, hipStream_t s);

int launch_sharp_stage(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                       const Geom& gin, const int* lo, const int* hi, const int* refl, int ymode, int omode,
                       double dt, double s, double dx, double dy, const void* P, const void* Q,
                       const void* A, const void* B, const void* aux_in, const void* comp_in,
                       void* out, void* aux_out, void* comp_out, unsigned long long* cfl,
                       const unsigned char* flags, const int* cells, const int* ncells,
                       hipStream_t stream, const AuxIn& aux, const void* src = nullptr);
int launch_ssp104_mid(int dtype, void* q, void* s1, long long n, hipStream_t stream);
int launch_lmm_update(int dtype, int mode, long long n, const void* q, const void* u,
                      const void* R, const void* v, const double* c, void* out,
                      hipStream_t stream);

int launch_unsplit_ctl(const clawhip_rp& rp, int dtype, const Geom& g, const int* lo,
                       const int* hi, const int* refl, int ng, const int* method,
                       const int* mthlim, StepCtl* ctl, const CtlBuffers& bufs, int nsteps,
                       hipStream_t stream);

long long ensemble_table_bytes(int n);
int ensemble_prepare(const clawhip_rp& rp, const clawhip_ens_member* members, int n,
                     const int* lo, const int* hi, const int* refl, int ng, const int* method,
                     const int* mthlim, int rows, void* table_dev, long long* waves);
int ensemble_run(const clawhip_rp& rp, int dtype, const void* table_dev, const long long* waves,
                 int nsteps, hipStream_t s);

int launch_sweep2(const clawhip_rp& rp, int ids, const void* q_in, const Geom& gin, void* q_out,
                  const Geom& gout, int ng, const int* method, const int* mthlim, double dtdx,
                  const AuxIn& aux, unsigned long long* cfl, hipStream_t s);
int launch_sweep3(const clawhip_rp& rp, int idir, const void* q_in, const Geom& gin, void* q_out,
                  const Geom& gout, int ng, const int* method, const int* mthlim, double dtdx,
                  unsigned long long* cfl, hipStream_t s);

int launch_features_labels(int dtype, const void* q, const Geom& g, const int* lo, const int* hi,
                           const int* refl, int meqn, float flat, float wtol, float* feat,
                           unsigned char* label, hipStream_t s);
int launch_mlp_train(const float* feat, long long nsamples, const unsigned char* label,
                     const int* batch_idx, int nbatch, float* master, float* grad, float* m1,
                     float* m2, int step, float lr, float pos_weight, float* loss, hipStream_t s);
int launch_mlp_logits(const float* feat, long long n, const float* master, float* logits,
                      hipStream_t s);
int launch_troubled_cells(const unsigned char* flags, int mx, int my, int* cells, int* count,
                          hipStream_t s);
int launch_mlp_flags(int dtype, const void* q, const Geom& g, const int* lo, const int* hi,
                     const int* refl, int meqn, float flat, const float* master, float zthr,
                     unsigned char* flags, hipStream_t s);
int launch_state_sums(int dtype, const void* q, const Geom& g, int meqn, unsigned positive_mask,
                      double* out, double* work, hipStream_t s);
int launch_state_sums_noalloc(int dtype, const void* q, const Geom& g, int meqn,
                              unsigned positive_mask, double* out, hipStream_t s);
int launch_l1_diff(int dtype, const void* a, const void* b, long long n, double* out,
                   double* work, hipStream_t s);

}  // namespace claw
