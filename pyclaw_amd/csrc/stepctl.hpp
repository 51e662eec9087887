// stepctl.hpp -- device-resident time-step control.
//
// StepCtl holds what Solver.evolve_to_time keeps between steps (pyclaw/solver.py:
// 554-666): t, dt, the cached CFL, the method's CFL targets and the status
// counters.  A step kernel reads dt and the current buffer from it; a
// one-wavefront kernel after it calls step_ctl_apply, which performs -- in FP64,
// with the reference's arithmetic -- the accept/reject decision
// (solver.py:621-651, accept_reject_step 524-529) and the next get_dt
// (solver.py:535-552, ClawSolver uses get_dt_new 531-533).  The layout equals
// clawhip_stepctl in include/clawhip.h.
#pragma once
#include <hip/hip_runtime.h>

namespace claw {

struct StepCtl {
  double t, dt, tstart, tend;
  double cfl, cfl_desired, cfl_max, dt_max;
  double cflmax, dtmin, dtmax;
  int dt_variable, done, cur, attempts;
  int max_steps, accepted, nrejected, pad0;
  unsigned int pad2, pad1;
  unsigned long long cfl_bits;
};

struct CtlBuffers {
  const void* q[2];
  void* comp[2];
  double dx, dy;
};

// done codes
enum { CTL_RUNNING = 0, CTL_REACHED_TEND = 1, CTL_CFL_FIXED_DT = 2, CTL_MAX_STEPS = 3 };

// The end of a step: applied by one thread of a kernel launched after the step
// kernel on the same stream (the kernel boundary orders the step's CFL atomics
// before this read, so the step kernel needs no fences).  Nothing happens once
// ctl->done is set (the step kernel did not run either).
__device__ __forceinline__ void step_ctl_apply(StepCtl* ctl) {
#pragma clang fp contract(off)   // the host loop's FP64 rounding, in either variant
  if (ctl->done) return;
  const unsigned long long bits = ctl->cfl_bits;
  const double cfl = __longlong_as_double((long long)bits);
  ctl->cfl = cfl;
  const double dt = ctl->dt;
  double t = ctl->t;
  const bool accept = !(cfl > ctl->cfl_max);
  int done = CTL_RUNNING;
  if (accept) {
    ctl->cflmax = cfl > ctl->cflmax ? cfl : ctl->cflmax;
    if (ctl->dt_variable) t = t + dt;
    else t = ctl->tstart + (double)(ctl->accepted + 1) * dt;
    ctl->t = t;
    ctl->accepted += 1;
    ctl->cur = 1 - ctl->cur;
  } else {
    if (ctl->dt_variable) ctl->nrejected += 1;
    else done = CTL_CFL_FIXED_DT;
  }
  ctl->attempts += 1;
  if (done == CTL_RUNNING) {
    if (accept && t >= ctl->tend) done = CTL_REACHED_TEND;
    else if (ctl->attempts >= ctl->max_steps) done = CTL_MAX_STEPS;
  }
  if (done == CTL_RUNNING) {
    // get_dt for the next attempt (dt_old is set after the first step)
    double ndt = dt;
    if (ctl->dt_variable) {
      if (cfl > 0.0) {
        double cand = dt * ctl->cfl_desired / cfl;
        ndt = ctl->dt_max < cand ? ctl->dt_max : cand;   // min(dt_max, cand)
        ctl->dtmin = ndt < ctl->dtmin ? ndt : ctl->dtmin;
        ctl->dtmax = ndt > ctl->dtmax ? ndt : ctl->dtmax;
      } else {
        ndt = ctl->dt_max;
      }
    }
    const double tend = ctl->tend;
    if (t + ndt > tend && ctl->tstart < tend) ndt = tend - t;
    if (tend - t - ndt < 1.e-14 * t) ndt = tend - t;
    ctl->dt = ndt;
  }
  ctl->done = done;
  ctl->cfl_bits = 0ull;
}

}  // namespace claw
