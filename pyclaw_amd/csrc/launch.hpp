// launch.hpp -- host-side launchers behind the C ABI (capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/clawhip.h"
#include "common.hpp"
#include "stepctl.hpp"
#include "sharpctl.hpp"

#include <map>
#include <mutex>
#include <utility>

namespace claw {
// Wavefronts of `kernel` (launched in 256-thread blocks) resident on the current
// device at once, from its occupancy, counting at most per_simd_cap per SIMD
// (0: no cap); cached per (device, kernel address), so every instantiation gets
// its own figure.  0 when the runtime cannot tell.
inline long long resident_waves(const void* kernel, int per_simd_cap) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, long long> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lock(mu);
  const auto key = std::make_pair(dev, kernel);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int ncu = 0, nb = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess ||
      ncu < 1 || nb < 1)
    return 0;
  // nb blocks of 4 wavefronts per CU = nb wavefronts per SIMD
  const int per_simd = per_simd_cap > 0 && nb > per_simd_cap ? per_simd_cap : nb;
  const long long r = (long long)ncu * per_simd * 4;
  cache[key] = r;
  return r;
}
}  // namespace claw

namespace claw {

template <typename T>
inline RPParams<T> to_params(const clawhip_rp& r) {
  RPParams<T> p;
  p.rho = T(r.rho); p.bulk = T(r.bulk); p.cc = T(r.cc); p.zz = T(r.zz);
  p.gamma = T(r.gamma); p.gamma1 = T(r.gamma - 1.0);
  p.grav = T(r.grav); p.u = T(r.u); p.v = T(r.v); p.efix = r.efix;
  p.dry_tolerance = T(r.dry_tolerance); p.fwave = r.fwave;
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

long long step3_work_elems(int meqn, int mx, int my, int mz);
int launch_step3(const clawhip_rp& rp, const void* q_in, const Geom& gin, void* q_out,
                 void* work, const int* lo, const int* hi, const int* refl, int ng,
                 const int* method, const int* mthlim, double dx, double dy, double dz,
                 double dt, unsigned long long* cfl, hipStream_t stream);
int launch_split3(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dz, double dt, unsigned long long* cfl, hipStream_t s);

// A kernel of a device-controlled run: it runs only while *gate != 0 and reads
// the state buffer q[*cur] (gate NULL: an eager launch on the given pointer).
// zero_count (flags kernels; may be NULL): a counter the kernel sets to 0 when
// it runs, for the flag list launched after it.
struct BufSel {
  const int* gate = nullptr;
  const int* cur = nullptr;
  const void* q[2] = {nullptr, nullptr};
  int* zero_count = nullptr;
};

// The learned indicator's selection for a SharpClaw stage: per-interface flags
// (NULL: full WENO5 everywhere), the ids of the interfaces whose dilated flag is
// set and their number (device), and the buffer their full-WENO5 edge values go
// to (2*meqn values of the state's type per interface id).
struct FlagSet {
  const unsigned char* flags = nullptr;
  const int* list = nullptr;
  const int* count = nullptr;
  void* recon = nullptr;
};

// The buffers of a device-controlled SSP33 run: the state and its Kahan
// compensation (ping-pong by SharpCtl::cur), dq_dt, the second stage's state
// and FP32 increment, and -- on the indicator path -- the flags, their list, the
// full-WENO5 edge values, the network and the fields the judge keeps positive.
struct SharpRunBufs {
  void* q[2] = {nullptr, nullptr};
  void* comp[2] = {nullptr, nullptr};
  void* dqdt = nullptr;
  void* dqdt_fb = nullptr;
  void* r2 = nullptr;
  void* aux23 = nullptr;
  unsigned char* flags = nullptr;
  int* flag_list = nullptr;
  int* flag_count = nullptr;
  void* recon = nullptr;
  const float* params = nullptr;
  float flat = 0.0f, threshold = 0.0f;
  unsigned pos_mask = 0u;
  // optional timing: hipEvent_t handles, two per main stage per attempt, recorded
  // around that stage's march launch (eager runs only)
  void* const* stage_events = nullptr;
};

int launch_sharp_ssp33_run(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                          const int* lo, const int* hi, const int* refl, double dx, double dy,
                          const SharpRunBufs& b, SharpCtl* ctl, int nsteps,
                          hipStream_t stream);
int launch_flag_list(const unsigned char* flags, int mx, int my, int* list, int* count,
                     hipStream_t s, const int* gate = nullptr, bool zeroed = false);
int launch_sharp_stage(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                       const Geom& gin, const int* lo, const int* hi, const int* refl, int ymode, int omode,
                       double dt, double s, double dx, double dy, const void* P, const void* Q,
                       const void* A, const void* B, const void* aux_in, const void* comp_in,
                       void* out, void* aux_out, void* comp_out, unsigned long long* cfl,
                       const FlagSet& flags, hipStream_t stream, const AuxIn& aux,
                       const void* src = nullptr);
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
int launch_mlp_flags(int dtype, const void* q, const Geom& g, const int* lo, const int* hi,
                     const int* refl, int meqn, float flat, const float* master, float zthr,
                     unsigned char* flags, hipStream_t s, const BufSel& sel = BufSel());
int launch_state_sums(int dtype, const void* q, const Geom& g, int meqn, unsigned positive_mask,
                      double* out, double* work, hipStream_t s);
int launch_state_sums_noalloc(int dtype, const void* q, const Geom& g, int meqn,
                              unsigned positive_mask, double* out, hipStream_t s);
int launch_l1_diff(int dtype, const void* a, const void* b, long long n, double* out,
                   double* work, hipStream_t s);

}  // namespace claw

// The launchers of the two arithmetic variants (variant.hpp).
namespace claw {
namespace exact {
#include "variant_decls.inc"
}  // namespace exact
namespace fast {
#include "variant_decls.inc"
}  // namespace fast
}  // namespace claw
