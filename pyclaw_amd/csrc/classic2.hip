// classic2.hip -- 2D classic Clawpack unsplit wave-propagation step for gfx950.
//
// The unsplit step (classic/step2.f90 + flux2.f90 + limiter.f90 + philim.f90,
// with the ghost-cell fill of solver.py:351-519 applied as an index map on load
// and the CFL max-reduction) runs as two launches of one sweep kernel:
//
//   k_usweep<AX = 1>: the x-slices j = 0..my+1 (step2.f90:96-163).  Their
//       updates of each cell are summed, in the reference's order, into `work`:
//       exactly the reference's qnew after its x-sweep.
//   k_usweep<AX = 2>: the y-slices i = 0..mx+1 (step2.f90:169-247), continuing
//       each cell's sum from `work` and writing q_new.
//
// Both sweeps are the same code with the roles of x and y exchanged.  A
// wavefront owns 64 consecutive cells of a slice (lanes 0,1,62,63 only supply
// neighbour data, so 60 cells are written) and marches H slices across.  All
// interfaces of a slice are solved lane-parallel: neighbour states, waves and
// fluctuations arrive by cross-lane moves, and no state is carried between
// slices except the three running sums of the cells the slice touches
// (slices s-1, s, s+1 update cell s).  With only the slice's temporaries live,
// the kernel needs far fewer registers than a fused x+y march, so several
// wavefronts share each SIMD and hide the FP64 latency of the Riemann solvers.
//
// Every cell sums its six contributions in the reference's order and
// parenthesisation:
//   after x:  ((((q + dtdy*g2[j-1]) + qadd) - dtdx*dF) - dtdy*dG) - dtdy*g1[j+1]
//   after y:  ((w + dtdx*g2[i-1]) + ((qadd - dtdy*dF) - dtdx*dG)) - dtdx*g1[i+1]
// FP64: bitwise equal to the reference step (built with -ffp-contract=off).
// FP32 (KAHAN): the same sums start from 0 instead of q, and the total increment
// is applied with a Kahan-compensated update (compensation array alongside q).
#include <stdlib.h>

#include "launch.hpp"
#include "usweep.hpp"

namespace claw {

// 3 wavefronts per SIMD: the FP64 Euler sweep fits 168 VGPRs (smaller solvers
// need fewer and run more wavefronts).
template <class RS, typename T, bool KAHAN, int AX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_usweep(const USweepArgs<T> a) {
  usweep_body<RS, T, KAHAN, AX>(a, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
}

// Device-resident stepping: the two sweeps of one attempted step of
// Solver.evolve_to_time.  The state ping-pongs between two buffers; `ctl->cur`
// names the current one.  The last wavefront of the y-sweep applies the CFL
// check, accepts or rejects the step and computes the next dt (step_ctl_finish).
template <class RS, typename T, bool KAHAN, int AX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_usweep_ctl(USweepArgs<T> a, StepCtl* ctl, CtlBuffers bufs) {
  if (ctl->done) return;   // uniform across the grid
  const int cur = ctl->cur;
  const double dt = ctl->dt;
  a.qin = (const T*)bufs.q[cur];
  if (AX == 1) {
    a.wout = (T*)bufs.work;
    a.dn = T(dt / bufs.dx);
    a.dtr = T(dt / bufs.dy);
  } else {
    a.win = (const T*)bufs.work;
    a.wout = (T*)bufs.q[1 - cur];
    a.comp_in = (const T*)bufs.comp[cur];
    a.comp_out = (T*)bufs.comp[1 - cur];
    a.dn = T(dt / bufs.dy);
    a.dtr = T(dt / bufs.dx);
  }
  a.cfl_bits = &ctl->cfl_bits;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  usweep_body<RS, T, KAHAN, AX>(a, wid);
  if (AX == 2) {
    const int total = a.nstrips * a.nrb;
    if (wid < total) step_ctl_finish(ctl, (unsigned)total);
  }
}

template <class RS, typename T, bool KAHAN>
__global__ void __launch_bounds__(256) k_ywin(const USweepArgs<T> a) {
  ywin_body<RS, T, KAHAN>(a, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
}

template <class RS, typename T, bool KAHAN>
__global__ void __launch_bounds__(256) k_ywin_ctl(USweepArgs<T> a, StepCtl* ctl,
                                                  CtlBuffers bufs) {
  if (ctl->done) return;   // uniform across the grid
  const int cur = ctl->cur;
  const double dt = ctl->dt;
  a.qin = (const T*)bufs.q[cur];
  a.win = (const T*)bufs.work;
  a.wout = (T*)bufs.q[1 - cur];
  a.comp_in = (const T*)bufs.comp[cur];
  a.comp_out = (T*)bufs.comp[1 - cur];
  a.dn = T(dt / bufs.dy);
  a.dtr = T(dt / bufs.dx);
  a.cfl_bits = &ctl->cfl_bits;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  ywin_body<RS, T, KAHAN>(a, wid);
  const int total = a.nstrips * a.nrb;
  if (wid < total) step_ctl_finish(ctl, (unsigned)total);
}

}  // namespace claw

// ---------------------------------------------------------------- launchers
namespace claw {

namespace {
// Slices per wavefront (march length H).  Each wavefront recomputes 2 halo
// slices, so long marches waste less; but the grid must also fill the device:
// with R wavefronts resident at once (from the kernel's occupancy), a launch of
// W wavefronts runs in ceil(W / R) rounds of H + 2 slice steps each.  H
// minimises that product (ties: the longer march).  CLAWHIP_UNSPLIT_ROWS
// overrides the choice (tuning sweeps).
template <class Kernel>
int choose_rows(Kernel kernel, int nstrips, int mm) {
  static long long resident = 0;   // per kernel instantiation
  if (resident == 0) {
    int dev = 0, ncu = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess ||
        ncu < 1 || nb < 1)
      return 16;
    resident = (long long)ncu * nb * 4;
  }
  if (const char* env = getenv("CLAWHIP_UNSPLIT_ROWS")) {
    int h = atoi(env);
    if (h >= 1 && h <= 4096) return h;
  }
  int best = 16;
  long long best_cost = -1;
  for (int h = 2; h <= 64; h++) {
    const long long waves = (long long)nstrips * ((mm + h - 1) / h);
    const long long cost = ((waves + resident - 1) / resident) * (long long)(h + 2);
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && h > best)) {
      best = h;
      best_cost = cost;
    }
    if ((mm + h - 1) / h == 1) break;
  }
  return best;
}

// y-sweep variant: lanes across the columns marching down (default) or lanes
// along the slices (CLAWHIP_YSWEEP=lane).
bool ysweep_lane() {
  const char* e = getenv("CLAWHIP_YSWEEP");
  return e && e[0] == 'l';
}

template <class RS, typename T, bool KAHAN>
USweepArgs<T> ywin_args(const clawhip_rp& rp, const Geom& gin, const Geom& gout, const int* lo,
                        const int* hi, const int* refl, int ng, const int* method,
                        const int* mthlim, bool ctl) {
  USweepArgs<T> a;
  a.rp = to_params<T>(rp);
  a.gin = gin; a.gout = gout;
  a.bc = to_bc(lo, hi, refl, 2);
  a.ng = ng; a.order = method[1]; a.trans = method[2];
  for (int w = 0; w < 4; w++) a.mthlim[w] = w < RS::mwaves ? mthlim[w] : 0;
  a.dn = T(0); a.dtr = T(0);
  a.nstrips = (gout.mx + 59) / 60;
  a.H = ctl ? choose_rows(k_ywin_ctl<RS, T, KAHAN>, a.nstrips, gout.my)
            : choose_rows(k_ywin<RS, T, KAHAN>, a.nstrips, gout.my);
  a.nrb = (gout.my + a.H - 1) / a.H;
  a.qin = nullptr; a.win = nullptr; a.wout = nullptr;
  a.comp_in = nullptr; a.comp_out = nullptr; a.cfl_bits = nullptr;
  return a;
}

template <class RS, typename T, bool KAHAN, int AX>
USweepArgs<T> sweep_args(const clawhip_rp& rp, const Geom& gin, const Geom& gout,
                         const int* lo, const int* hi, const int* refl, int ng,
                         const int* method, const int* mthlim, bool ctl) {
  USweepArgs<T> a;
  a.rp = to_params<T>(rp);
  a.gin = gin; a.gout = gout;
  a.bc = to_bc(lo, hi, refl, 2);
  a.ng = ng; a.order = method[1]; a.trans = method[2];
  for (int w = 0; w < 4; w++) a.mthlim[w] = w < RS::mwaves ? mthlim[w] : 0;
  a.dn = T(0); a.dtr = T(0);
  const int ms = AX == 1 ? gout.mx : gout.my, mm = AX == 1 ? gout.my : gout.mx;
  a.nstrips = (ms + 59) / 60;
  a.H = ctl ? choose_rows(k_usweep_ctl<RS, T, KAHAN, AX>, a.nstrips, mm)
            : choose_rows(k_usweep<RS, T, KAHAN, AX>, a.nstrips, mm);
  a.nrb = (mm + a.H - 1) / a.H;
  a.qin = nullptr; a.win = nullptr; a.wout = nullptr;
  a.comp_in = nullptr; a.comp_out = nullptr; a.cfl_bits = nullptr;
  return a;
}

template <typename T>
unsigned sweep_blocks(const USweepArgs<T>& a) {
  return (unsigned)(((long long)a.nstrips * a.nrb + 3) / 4);
}

template <class RS, typename T, bool KAHAN>
int run_unsplit(const clawhip_rp& rp, const void* q_in, const Geom& gin, void* q_out,
                const Geom& gout, const void* comp_in, void* comp_out, void* work,
                const int* lo, const int* hi, const int* refl, int ng, const int* method,
                const int* mthlim, double dx, double dy, double dt,
                unsigned long long* cfl_bits, hipStream_t stream) {
  auto ax = sweep_args<RS, T, KAHAN, 1>(rp, gin, gout, lo, hi, refl, ng, method, mthlim, false);
  ax.dn = T(dt / dx); ax.dtr = T(dt / dy);
  ax.qin = (const T*)q_in; ax.wout = (T*)work; ax.cfl_bits = cfl_bits;
  hipLaunchKernelGGL((k_usweep<RS, T, KAHAN, 1>), dim3(sweep_blocks(ax)), dim3(256), 0, stream,
                     ax);
  const bool lane = ysweep_lane();
  auto ay = lane ? sweep_args<RS, T, KAHAN, 2>(rp, gin, gout, lo, hi, refl, ng, method, mthlim,
                                               false)
                 : ywin_args<RS, T, KAHAN>(rp, gin, gout, lo, hi, refl, ng, method, mthlim,
                                           false);
  ay.dn = T(dt / dy); ay.dtr = T(dt / dx);
  ay.qin = (const T*)q_in; ay.win = (const T*)work; ay.wout = (T*)q_out;
  ay.comp_in = (const T*)comp_in; ay.comp_out = (T*)comp_out; ay.cfl_bits = cfl_bits;
  if (lane)
    hipLaunchKernelGGL((k_usweep<RS, T, KAHAN, 2>), dim3(sweep_blocks(ay)), dim3(256), 0,
                       stream, ay);
  else
    hipLaunchKernelGGL((k_ywin<RS, T, KAHAN>), dim3(sweep_blocks(ay)), dim3(256), 0, stream,
                       ay);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

template <class RS, typename T, bool KAHAN>
int run_unsplit_ctl(const clawhip_rp& rp, const Geom& g, const int* lo, const int* hi,
                    const int* refl, int ng, const int* method, const int* mthlim,
                    StepCtl* ctl, const CtlBuffers& bufs, int nsteps, hipStream_t stream) {
  auto ax = sweep_args<RS, T, KAHAN, 1>(rp, g, g, lo, hi, refl, ng, method, mthlim, true);
  const bool lane = ysweep_lane();
  auto ay = lane ? sweep_args<RS, T, KAHAN, 2>(rp, g, g, lo, hi, refl, ng, method, mthlim, true)
                 : ywin_args<RS, T, KAHAN>(rp, g, g, lo, hi, refl, ng, method, mthlim, true);
  for (int n = 0; n < nsteps; n++) {
    hipLaunchKernelGGL((k_usweep_ctl<RS, T, KAHAN, 1>), dim3(sweep_blocks(ax)), dim3(256), 0,
                       stream, ax, ctl, bufs);
    if (lane)
      hipLaunchKernelGGL((k_usweep_ctl<RS, T, KAHAN, 2>), dim3(sweep_blocks(ay)), dim3(256), 0,
                         stream, ay, ctl, bufs);
    else
      hipLaunchKernelGGL((k_ywin_ctl<RS, T, KAHAN>), dim3(sweep_blocks(ay)), dim3(256), 0,
                         stream, ay, ctl, bufs);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

template <template <typename> class RSF>
int dispatch_unsplit_ctl(int dtype, const clawhip_rp& rp, const Geom& g, const int* lo,
                         const int* hi, const int* refl, int ng, const int* method,
                         const int* mthlim, StepCtl* ctl, const CtlBuffers& bufs, int nsteps,
                         hipStream_t stream) {
  if (dtype == CLAWHIP_F64)
    return run_unsplit_ctl<RSF<double>, double, false>(rp, g, lo, hi, refl, ng, method, mthlim,
                                                       ctl, bufs, nsteps, stream);
  return run_unsplit_ctl<RSF<float>, float, true>(rp, g, lo, hi, refl, ng, method, mthlim, ctl,
                                                  bufs, nsteps, stream);
}

template <template <typename> class RSF>
int dispatch_unsplit(int dtype, const clawhip_rp& rp, const void* q_in, const Geom& gin,
                     void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                     void* work, const int* lo, const int* hi, const int* refl, int ng,
                     const int* method, const int* mthlim, double dx, double dy, double dt,
                     unsigned long long* cfl_bits, hipStream_t stream) {
  if (dtype == CLAWHIP_F64)
    return run_unsplit<RSF<double>, double, false>(rp, q_in, gin, q_out, gout, comp_in,
                                                   comp_out, work, lo, hi, refl, ng, method,
                                                   mthlim, dx, dy, dt, cfl_bits, stream);
  return run_unsplit<RSF<float>, float, true>(rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                              work, lo, hi, refl, ng, method, mthlim, dx, dy,
                                              dt, cfl_bits, stream);
}
}  // namespace

int launch_unsplit(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                   void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                   void* work, const int* lo, const int* hi, const int* refl, int ng,
                   const int* method, const int* mthlim, double dx, double dy, double dt,
                   unsigned long long* cfl_bits, hipStream_t stream) {
#define ARGS dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, work, lo, hi, refl, ng, \
             method, mthlim, dx, dy, dt, cfl_bits, stream
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS: return dispatch_unsplit<Acoustics>(ARGS);
    case CLAWHIP_RP_EULER: return dispatch_unsplit<Euler>(ARGS);
    case CLAWHIP_RP_SHALLOW: return dispatch_unsplit<Shallow>(ARGS);
    case CLAWHIP_RP_ADVECTION: return dispatch_unsplit<Advection>(ARGS);
  }
#undef ARGS
  return CLAWHIP_E_KIND;
}

int launch_unsplit_ctl(const clawhip_rp& rp, int dtype, const Geom& g, const int* lo,
                       const int* hi, const int* refl, int ng, const int* method,
                       const int* mthlim, StepCtl* ctl, const CtlBuffers& bufs, int nsteps,
                       hipStream_t stream) {
#define ARGS dtype, rp, g, lo, hi, refl, ng, method, mthlim, ctl, bufs, nsteps, stream
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS: return dispatch_unsplit_ctl<Acoustics>(ARGS);
    case CLAWHIP_RP_EULER: return dispatch_unsplit_ctl<Euler>(ARGS);
    case CLAWHIP_RP_SHALLOW: return dispatch_unsplit_ctl<Shallow>(ARGS);
    case CLAWHIP_RP_ADVECTION: return dispatch_unsplit_ctl<Advection>(ARGS);
  }
#undef ARGS
  return CLAWHIP_E_KIND;
}

}  // namespace claw
