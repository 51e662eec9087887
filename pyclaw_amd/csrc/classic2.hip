// classic2.hip -- 2D classic Clawpack unsplit wave-propagation step for gfx950:
// one fused kernel per step (unsplit.hpp: x- and y-slices of classic/step2.f90
// with flux2.f90, limiter.f90, philim.f90, the ghost fill as an index map and
// the CFL max-reduction), plus the device-controlled variant that applies
// Solver.evolve_to_time's accept/reject and next dt in its last wavefront.
//
// A fused kernel was measured against a two-launch form (an x-sweep into a work
// array, then a y-sweep): on 1024^2 Euler FP64 the fused kernel takes 162 us,
// the two launches 78 + 132 us, because the slices' FP64 instruction count is
// the bound and the split pays the halo recomputation and the work round trip
// twice.
#include "variant.hpp"  // first: selects the arithmetic variant
#include <stdlib.h>

#include "launch.hpp"
#include "unsplit.hpp"

namespace claw {
namespace CLAW_VNS {

template <class RS, typename T, bool KAHAN, bool AUX = false, bool CAPA = false>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(unsplit_waves<RS, T, AUX, CAPA>())))
    k_unsplit(const UnsplitArgs<T> a) {
  // one Courant-number atomic per workgroup: one per wavefront serialises
  // (same address) over the thousands of wavefronts of a large grid
  block_atomic_max(
      unsplit_body<RS, T, KAHAN, AUX, CAPA>(a, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)),
      a.cfl_bits);
}

// Device-resident stepping: one launch = one attempted step of
// Solver.evolve_to_time.  The state ping-pongs between two buffers; `ctl->cur`
// names the current one.  The last workgroup to finish applies the CFL check,
// accepts or rejects the step and computes the next dt (k_step_ctl after it).
template <class RS, typename T, bool KAHAN>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(unsplit_waves<RS, T>())))
    k_unsplit_ctl(UnsplitArgs<T> a, StepCtl* ctl, CtlBuffers bufs) {
  if (ctl->done) return;   // uniform across the grid
  const int cur = ctl->cur;
  const double dt = ctl->dt;
  a.qin = (const T*)bufs.q[cur];
  a.qout = (T*)bufs.q[1 - cur];
  a.comp_in = (const T*)bufs.comp[cur];
  a.comp_out = (T*)bufs.comp[1 - cur];
  a.dtdx = T(dt / bufs.dx);
  a.dtdy = T(dt / bufs.dy);
  a.cfl_bits = &ctl->cfl_bits;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  block_atomic_max(unsplit_body<RS, T, KAHAN>(a, wid), a.cfl_bits);
}

// The step's control after it (step_ctl_apply).
__global__ void __launch_bounds__(64) k_step_ctl(StepCtl* ctl) {
  if (threadIdx.x == 0) step_ctl_apply(ctl);
}

}  // namespace CLAW_VNS
}  // namespace claw

// ---------------------------------------------------------------- launchers
namespace claw {
namespace CLAW_VNS {

namespace {
// Rows per wavefront (march length H).  Each wavefront recomputes the x-slices of
// 2 halo rows, so long marches waste less; but the grid must also fill the
// device: with R wavefronts resident at once (from the kernel's occupancy), a
// launch of W wavefronts runs in ceil(W / R) rounds of H + 2 row-steps each.
// H minimises that product (ties: the longer march).  CLAWHIP_UNSPLIT_ROWS
// overrides the choice (tuning sweeps).
template <class Kernel>
int choose_rows(Kernel kernel, int nstrips, int my, int waves_hint) {
  // beyond two wavefronts per SIMD the measured gain of a third is smaller than
  // the halo rows the shorter marches add, so count at most two, unless the
  // kernel is compiled for more (unsplit_waves)
  const long long resident = resident_waves((const void*)kernel, waves_hint > 2 ? waves_hint : 2);
  if (resident == 0) return 16;
  if (const char* env = getenv("CLAWHIP_UNSPLIT_ROWS")) {
    int h = atoi(env);
    if (h >= 1 && h <= 4096) return h;
  }
  // the shortest march considered: 3 rows.  A grid that fills less than one
  // resident round takes it (256^2 FP64 acoustics step: H = 1, 2, 3, 4, 6 ran
  // 26.8, 21.3, 21.0, 22.0, 25.0 us); shorter marches add more wavefronts than
  // they save in each one's dependent row chain
  int best = 16;
  long long best_cost = -1;
  for (int h = 3; h <= 64; h++) {
    const long long waves = (long long)nstrips * ((my + h - 1) / h);
    const long long cost = ((waves + resident - 1) / resident) * (long long)(h + 2);
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && h > best)) {
      best = h;
      best_cost = cost;
    }
    if ((my + h - 1) / h == 1) break;
  }
  return best;
}

template <class RS, typename T, bool KAHAN, bool AUX = false, bool CAPA = false>
int run_unsplit(const clawhip_rp& rp, const void* q_in, const Geom& gin, void* q_out,
                const Geom& gout, const void* comp_in, void* comp_out, const int* lo, const int* hi, const int* refl,
                int ng, const int* method, const int* mthlim, double dx, double dy, double dt,
                unsigned long long* cfl_bits, hipStream_t stream, const AuxIn* aux = nullptr) {
  UnsplitArgs<T> a;
  a.aux = nullptr;
  a.naux = 0;
  a.capa = -1;
  if constexpr (AUX) {
    a.aux = (const T*)aux->p;
    a.gaux = aux->g;
    const int never[3] = {-1, -1, -1};
    a.bcaux = to_bc(aux->lo, aux->hi, never, 2);
    a.naux = method[6];
    a.capa = method[5] - 1;
  }
  a.rp = to_params<T>(rp);
  a.gin = gin; a.gout = gout;
  a.bc = to_bc(lo, hi, refl, 2);
  a.ng = ng; a.order = method[1]; a.trans = method[2];
  for (int w = 0; w < kMaxWaves; w++) a.mthlim[w] = w < RS::mwaves ? mthlim[w] : 0;
  a.dtdx = T(dt / dx); a.dtdy = T(dt / dy);
  a.nstrips = (gout.mx + 59) / 60;
  a.H = choose_rows(k_unsplit<RS, T, KAHAN, AUX, CAPA>, a.nstrips, gout.my,
                    unsplit_waves<RS, T, AUX, CAPA>());
  a.nrb = (gout.my + a.H - 1) / a.H;
  a.qin = (const T*)q_in; a.qout = (T*)q_out;
  a.comp_in = (const T*)comp_in; a.comp_out = (T*)comp_out;
  a.cfl_bits = cfl_bits;
  long long waves = (long long)a.nstrips * a.nrb;
  int blocks = (int)((waves + 3) / 4);
  hipLaunchKernelGGL((k_unsplit<RS, T, KAHAN, AUX, CAPA>), dim3(blocks), dim3(256), 0, stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

template <class RS, typename T, bool KAHAN>
int run_unsplit_ctl(const clawhip_rp& rp, const Geom& g, const int* lo, const int* hi,
                    const int* refl, int ng, const int* method, const int* mthlim,
                    StepCtl* ctl, const CtlBuffers& bufs, int nsteps, hipStream_t stream) {
  UnsplitArgs<T> a;
  a.rp = to_params<T>(rp);
  a.gin = g; a.gout = g;
  a.bc = to_bc(lo, hi, refl, 2);
  a.ng = ng; a.order = method[1]; a.trans = method[2];
  for (int w = 0; w < kMaxWaves; w++) a.mthlim[w] = w < RS::mwaves ? mthlim[w] : 0;
  a.dtdx = T(0); a.dtdy = T(0);
  a.nstrips = (g.mx + 59) / 60;
  a.H = choose_rows(k_unsplit_ctl<RS, T, KAHAN>, a.nstrips, g.my, unsplit_waves<RS, T>());
  a.nrb = (g.my + a.H - 1) / a.H;
  a.qin = nullptr; a.qout = nullptr; a.comp_in = nullptr; a.comp_out = nullptr;
  a.cfl_bits = nullptr;
  long long waves = (long long)a.nstrips * a.nrb;
  int blocks = (int)((waves + 3) / 4);
  for (int n = 0; n < nsteps; n++) {
    hipLaunchKernelGGL((k_unsplit_ctl<RS, T, KAHAN>), dim3(blocks), dim3(256), 0, stream, a,
                       ctl, bufs);
    hipLaunchKernelGGL(k_step_ctl, dim3(1), dim3(64), 0, stream, ctl);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

template <template <typename> class RSF>
int dispatch_unsplit_ctl(int dtype, const clawhip_rp& rp, const Geom& g, const int* lo,
                         const int* hi, const int* refl, int ng, const int* method,
                         const int* mthlim, StepCtl* ctl, const CtlBuffers& bufs, int nsteps,
                         hipStream_t stream) {
  if (dtype != CLAWHIP_F32K)
    return run_unsplit_ctl<RSF<double>, double, false>(rp, g, lo, hi, refl, ng, method, mthlim,
                                                       ctl, bufs, nsteps, stream);
#if CLAW_FAST
  return run_unsplit_ctl<RSF<float>, float, true>(rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                                  bufs, nsteps, stream);
#else
  // FP32 + Kahan is dispatched to the fast variant only
  return CLAWHIP_E_ARG;
#endif
}

template <template <typename> class RSF>
int dispatch_unsplit(int dtype, const clawhip_rp& rp, const void* q_in, const Geom& gin,
                     void* q_out, const Geom& gout, const void* comp_in, void* comp_out, const int* lo, const int* hi,
                     const int* refl, int ng, const int* method, const int* mthlim, double dx,
                     double dy, double dt, unsigned long long* cfl_bits, hipStream_t stream) {
  if (dtype != CLAWHIP_F32K)
    return run_unsplit<RSF<double>, double, false>(rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi,
                                                   refl, ng, method, mthlim, dx, dy, dt,
                                                   cfl_bits, stream);
#if CLAW_FAST
  return run_unsplit<RSF<float>, float, true>(rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi, refl,
                                              ng, method, mthlim, dx, dy, dt, cfl_bits, stream);
#else
  // FP32 + Kahan is dispatched to the fast variant only
  return CLAWHIP_E_ARG;
#endif
}
// Solvers with aux fields: the aux window, with or without a capacity field.
template <class RS>
int dispatch_unsplit_aux(int dtype, const clawhip_rp& rp, const void* q_in, const Geom& gin,
                         void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                         const int* lo, const int* hi, const int* refl, int ng,
                         const int* method, const int* mthlim, double dx, double dy, double dt,
                         unsigned long long* cfl_bits, hipStream_t stream, const AuxIn& aux) {
  (void)rp;
  const bool capa = method[5] > 0;
  if (dtype != CLAWHIP_F32K) {
    using R = typename RS::template with<double>;
    return capa ? run_unsplit<R, double, false, true, true>(rp, q_in, gin, q_out, gout, comp_in,
                                                            comp_out, lo, hi, refl, ng, method,
                                                            mthlim, dx, dy, dt, cfl_bits, stream,
                                                            &aux)
                : run_unsplit<R, double, false, true, false>(rp, q_in, gin, q_out, gout, comp_in,
                                                             comp_out, lo, hi, refl, ng, method,
                                                             mthlim, dx, dy, dt, cfl_bits, stream,
                                                             &aux);
  }
#if CLAW_FAST
  using R = typename RS::template with<float>;
  return capa ? run_unsplit<R, float, true, true, true>(rp, q_in, gin, q_out, gout, comp_in,
                                                        comp_out, lo, hi, refl, ng, method, mthlim,
                                                        dx, dy, dt, cfl_bits, stream, &aux)
              : run_unsplit<R, float, true, true, false>(rp, q_in, gin, q_out, gout, comp_in,
                                                         comp_out, lo, hi, refl, ng, method,
                                                         mthlim, dx, dy, dt, cfl_bits, stream,
                                                         &aux);
#else
  // FP32 + Kahan is dispatched to the fast variant only
  return CLAWHIP_E_ARG;
#endif
}

template <int ME>
struct VCAdvectionOf {
  template <typename T>
  using with = VCAdvection<T, ME>;
};
template <template <typename> class RSF>
struct SolverOf {
  template <typename T>
  using with = RSF<T>;
};
}  // namespace

int launch_unsplit(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                   void* q_out, const Geom& gout, const void* comp_in, void* comp_out, const int* lo, const int* hi,
                   const int* refl, int ng, const int* method, const int* mthlim, double dx,
                   double dy, double dt, unsigned long long* cfl_bits, hipStream_t stream,
                   const AuxIn& aux) {
  if (rp.kind == CLAWHIP_RP_VC_ADVECTION || rp.kind == CLAWHIP_RP_VC_ADVECTION_2D) {
    if (!aux.p) return CLAWHIP_E_ARG;
    const int me = rp.kind == CLAWHIP_RP_VC_ADVECTION_2D ? 1 : rp.meqn;
#define CLAWHIP_VCA(ME)                                                                         \
  return dispatch_unsplit_aux<VCAdvectionOf<ME>>(dtype, rp, q_in, gin, q_out, gout, comp_in,   \
                                                 comp_out, lo, hi, refl, ng, method, mthlim,   \
                                                 dx, dy, dt, cfl_bits, stream, aux)
    switch (me) {
      case 1: CLAWHIP_VCA(1);
      case 2: CLAWHIP_VCA(2);
      case 3: CLAWHIP_VCA(3);
      case 4: CLAWHIP_VCA(4);
    }
#undef CLAWHIP_VCA
    return CLAWHIP_E_ARG;
  }
  if (rp.kind == CLAWHIP_RP_VC_ACOUSTICS) {
    if (!aux.p) return CLAWHIP_E_ARG;
    return dispatch_unsplit_aux<SolverOf<VCAcoustics>>(dtype, rp, q_in, gin, q_out, gout, comp_in,
                                                      comp_out, lo, hi, refl, ng, method, mthlim,
                                                      dx, dy, dt, cfl_bits, stream, aux);
  }
  if (rp.kind == CLAWHIP_RP_SHALLOW_FWAVE) {
    if (!aux.p) return CLAWHIP_E_ARG;
    return dispatch_unsplit_aux<SolverOf<ShallowFwave>>(dtype, rp, q_in, gin, q_out, gout,
                                                       comp_in, comp_out, lo, hi, refl, ng,
                                                       method, mthlim, dx, dy, dt, cfl_bits,
                                                       stream, aux);
  }
  if (aux.p) return CLAWHIP_E_ARG;
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS:
      return dispatch_unsplit<Acoustics>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi, refl,
                                         ng, method, mthlim, dx, dy, dt, cfl_bits, stream);
    case CLAWHIP_RP_EULER:
      return dispatch_unsplit<Euler>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi, refl, ng,
                                     method, mthlim, dx, dy, dt, cfl_bits, stream);
    case CLAWHIP_RP_SHALLOW:
      return dispatch_unsplit<Shallow>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi, refl,
                                       ng, method, mthlim, dx, dy, dt, cfl_bits, stream);
    case CLAWHIP_RP_ADVECTION:
      return dispatch_unsplit<Advection>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi, refl,
                                         ng, method, mthlim, dx, dy, dt, cfl_bits, stream);
    case CLAWHIP_RP_SHALLOW_HLLE:
      return dispatch_unsplit<ShallowHLLE>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                           lo, hi, refl, ng, method, mthlim, dx, dy, dt,
                                           cfl_bits, stream);
    case CLAWHIP_RP_EULER5:
      return dispatch_unsplit<Euler5>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, lo, hi,
                                      refl, ng, method, mthlim, dx, dy, dt, cfl_bits, stream);
  }
  return CLAWHIP_E_KIND;
}

int launch_unsplit_ctl(const clawhip_rp& rp, int dtype, const Geom& g, const int* lo,
                       const int* hi, const int* refl, int ng, const int* method,
                       const int* mthlim, StepCtl* ctl, const CtlBuffers& bufs, int nsteps,
                       hipStream_t stream) {
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS:
      return dispatch_unsplit_ctl<Acoustics>(dtype, rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                             bufs, nsteps, stream);
    case CLAWHIP_RP_EULER:
      return dispatch_unsplit_ctl<Euler>(dtype, rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                         bufs, nsteps, stream);
    case CLAWHIP_RP_SHALLOW:
      return dispatch_unsplit_ctl<Shallow>(dtype, rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                           bufs, nsteps, stream);
    case CLAWHIP_RP_ADVECTION:
      return dispatch_unsplit_ctl<Advection>(dtype, rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                             bufs, nsteps, stream);
    case CLAWHIP_RP_SHALLOW_HLLE:
      return dispatch_unsplit_ctl<ShallowHLLE>(dtype, rp, g, lo, hi, refl, ng, method, mthlim,
                                               ctl, bufs, nsteps, stream);
    case CLAWHIP_RP_EULER5:
      return dispatch_unsplit_ctl<Euler5>(dtype, rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                          bufs, nsteps, stream);
  }
  return CLAWHIP_E_KIND;
}

}  // namespace CLAW_VNS
}  // namespace claw
