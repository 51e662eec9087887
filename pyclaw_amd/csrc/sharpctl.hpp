// sharpctl.hpp -- device-resident step control of SharpClaw's SSP33 step.
//
// SharpCtl holds what SharpClawSolver.step and Solver.evolve_to_time keep
// between attempts (pyclaw/sharpclaw/solver.py:283-323, 622-685;
// pyclaw/solver.py:535-666): t, dt, dt_old, the CFL object's cached maximum,
// the status counters, the learned indicator's flag-refresh cadence and its
// positivity judge.  The stage kernels read dt and the current buffer from it
// and run only when their gate is set; a one-wavefront kernel after a stage
// applies that stage's part of the control, in FP64 with the reference's
// arithmetic:
//   after stage 1 (PH_MID): the cached CFL = stage 1's (set_global_max(0) then
//     update_global_max), get_dt, the rescale by dt/dt_old;
//   after stage 3 (PH_END), or after the full-WENO5 recomputation (PH_END_FB):
//     update_global_max with the later stages', accept_reject_step, t, the
//     buffer swap, and the next attempt's gates (its flag refresh, and its get_dt
//     when this attempt was rejected, since that attempt skips stage 1);
//   after stage 3 on the indicator path (PH_END_JUDGE): the judge's step count
//     and positivity check -- a non-positive or non-finite depth / density /
//     energy opens the gate of the full-WENO5 recomputation instead of ending,
//     and closes every other gate: the attempts still queued in the launch do
//     nothing, the recomputation (launched once, after them) ends the attempt
//     (PH_END_FB) and reopens the gates for the next launch.
// The layout equals clawhip_sharpctl in include/clawhip.h.
#pragma once
#include <hip/hip_runtime.h>

namespace claw {

struct SharpCtl {
  double t, dt, dt_old, tstart, tend;
  double cfl, cfl_desired, cfl_max, dt_max;
  double cflmax, dtmin, dtmax;
  int have_dt_old, accept_prev, done, cur;
  int attempts, max_steps, accepted, nrejected;
  int judge, age, every, judge_steps;
  int fallbacks, run_flags, run1, run23;
  int run_fb, pad0;
  unsigned int pad1, nbad;
  unsigned long long cfl1_bits, cfl23_bits, cflfb_bits;
};

// done codes (as StepCtl's)
enum { SCTL_RUNNING = 0, SCTL_REACHED_TEND = 1, SCTL_MAX_STEPS = 3 };
// stage control phases
enum { PH_NONE = 0, PH_MID = 1, PH_END = 2, PH_END_JUDGE = 3, PH_END_FB = 4 };

__device__ __forceinline__ double bits_to_double(unsigned long long b) {
  return __longlong_as_double((long long)b);
}

// Python's max(a, b): b when b > a, else a (pyclaw/cfl.py update_global_max)
__device__ __forceinline__ double py_max(double a, double b) { return b > a ? b : a; }

// Solver.get_dt (solver.py:535-552) with SharpClawSolver.get_dt_new for the
// Runge-Kutta methods (sharpclaw/solver.py:653-685), tend clipping for
// evolve_to_time(solution, tend), then the rescale of the cached CFL by
// dt/dt_old and dt_old = dt (sharpclaw/solver.py:299-303).
__device__ __forceinline__ void sharp_get_dt(SharpCtl* c) {
#pragma clang fp contract(off)   // the host loop's FP64 rounding, in either variant
  double cfl = c->cfl;
  if (c->have_dt_old) {
    if (cfl > 0.0) {
      c->dt_old = c->dt;
      c->dt = c->dt * c->cfl_desired / cfl;
      c->dtmin = c->dt < c->dtmin ? c->dt : c->dtmin;   // min(dt, dtmin)
      c->dtmax = c->dt > c->dtmax ? c->dt : c->dtmax;   // max(dt, dtmax)
    } else {
      c->dt = c->dt_max;
    }
  } else {
    c->dt_old = c->dt;
    c->have_dt_old = 1;
  }
  const double t = c->t, tend = c->tend;
  if (t + c->dt > tend && c->tstart < tend) c->dt = tend - t;
  if (tend - t - c->dt < 1.e-14 * t) c->dt = tend - t;
  cfl = c->cfl;
  c->cfl = c->dt / c->dt_old * cfl;
  c->dt_old = c->dt;
}

// The end of an attempt: evolve_to_time's accept/reject (solver.py:616-651,
// accept_reject_step sharpclaw/solver.py:622-650) and the next attempt's gates.
__device__ __forceinline__ void sharp_end(SharpCtl* c, bool fallback_ran) {
#pragma clang fp contract(off)   // the host loop's FP64 rounding, in either variant
  double cfl = py_max(c->cfl, bits_to_double(c->cfl23_bits));
  if (fallback_ran) cfl = py_max(cfl, bits_to_double(c->cflfb_bits));
  c->cfl = cfl;
  const bool accept = !(cfl > c->cfl_max);
  int done = SCTL_RUNNING;
  if (accept) {
    c->cflmax = c->cflmax > cfl ? c->cflmax : cfl;   // max(cfl, cflmax)
    c->t = c->t + c->dt;
    c->accepted += 1;
    c->cur = 1 - c->cur;
  } else {
    c->nrejected += 1;
  }
  c->attempts += 1;
  if (accept && c->t >= c->tend) done = SCTL_REACHED_TEND;
  else if (c->attempts >= c->max_steps) done = SCTL_MAX_STEPS;
  c->done = done;
  c->accept_prev = accept ? 1 : 0;
  c->cfl1_bits = 0ull;
  c->cfl23_bits = 0ull;
  c->cflfb_bits = 0ull;
  c->nbad = 0u;
  c->run_fb = 0;
  c->run1 = (done == SCTL_RUNNING && accept) ? 1 : 0;
  c->run23 = done == SCTL_RUNNING ? 1 : 0;
  c->run_flags = 0;
  if (done != SCTL_RUNNING) return;
  // the rejected attempt's successor skips stage 1: its get_dt happens here
  if (!accept) sharp_get_dt(c);
  // the indicator refreshes its flags every `every` attempts (step_flags)
  if (c->judge) {
    if (c->age >= c->every) {
      c->run_flags = 1;
      c->age = 0;
    }
    c->age += 1;
  }
}

// The control phase after a stage, applied by k_sharp_ctl (one wavefront,
// launched after the stage on the same stream: the kernel boundary orders the
// stage's atomics before these reads, so the stage kernels need no fences).
__device__ __forceinline__ void sharp_ctl_phase(SharpCtl* c, int phase) {
#pragma clang fp contract(off)   // the host loop's FP64 rounding, in either variant
  switch (phase) {
    case PH_MID: {
      // set_global_max(0.) then update_global_max(stage 1's CFL)
      c->cfl = py_max(0.0, bits_to_double(c->cfl1_bits));
      sharp_get_dt(c);
      break;
    }
    case PH_END_JUDGE: {
      c->judge_steps += 1;
      if (c->nbad > 0u) {
        // the judge's positivity fallback: the same step with full WENO5
        c->fallbacks += 1;
        c->run_fb = 1;
        c->run_flags = 0;
        c->run1 = 0;
        c->run23 = 0;
        break;
      }
      sharp_end(c, false);
      break;
    }
    case PH_END: sharp_end(c, false); break;
    case PH_END_FB: sharp_end(c, true); break;
    default: break;
  }
}

}  // namespace claw
