// common.hpp -- grid descriptors, boundary-mapped loads, wavefront helpers and the
// wave limiter shared by the classic and SharpClaw kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "riemann.hpp"

namespace claw {

// Most waves of any solver here (euler_5wave_2D): the size of the per-wave
// limiter table the classic kernels take.
constexpr int kMaxWaves = 5;

// Boundary-condition codes: pyclaw.BC (solver.py:13-19) plus GIVEN, which means
// the ghost cells are stored in the array (a user-filled qbc) and read as they are.
enum BCCode { BC_CUSTOM = 0, BC_EXTRAP = 1, BC_PERIODIC = 2, BC_WALL = 3, BC_GIVEN = 4 };

// Storage of one state array on the device.  Cell (i, j, k) (1-based interior
// indices, ghosts at <= 0 and > m) of field f lives at
//   base + f*fs + (i-1+g)*1 + (j-1+g)*rs + (k-1+g)*ss
// where g is the number of ghost layers present in storage (0: compact).
struct Geom {
  int mx, my, mz;
  int g;
  long long fs, rs, ss;
};

struct BCSpec {
  int lo[3], hi[3];      // per dimension, BCCode
  int refl[3];           // component negated by a wall in that dimension
};

// Map one index into the stored range following Solver._bc_lower/_bc_upper
// (solver.py:439-519).  Indices are first clamped to the reference's ghost range
// [1-ng, m+ng]; returns the storage index and whether a wall flipped it.
__device__ __forceinline__ int bc_map(int i, int m, int ng, int lo, int hi, bool& flip) {
  i = i < 1 - ng ? 1 - ng : (i > m + ng ? m + ng : i);
  if (i < 1) {
    if (lo == BC_EXTRAP) i = 1;
    else if (lo == BC_PERIODIC) i += m;
    else if (lo == BC_WALL) { i = 1 - i; flip = !flip; }
  } else if (i > m) {
    if (hi == BC_EXTRAP) i = m;
    else if (hi == BC_PERIODIC) i -= m;
    else if (hi == BC_WALL) { i = 2 * m + 1 - i; flip = !flip; }
  }
  return i;
}

// Load cell (c, r[, k]) with the ghost-cell value the reference's BC fill would
// have placed there (x, then y, then z, so corners compose the per-dimension maps).
template <int ME, typename T>
__device__ __forceinline__ void load_cell2(const T* __restrict__ base, const Geom& g,
                                           const BCSpec& bc, int ng, int c, int r, T (&out)[ME]) {
  bool fx = false, fy = false;
  c = bc_map(c, g.mx, ng, bc.lo[0], bc.hi[0], fx);
  r = bc_map(r, g.my, ng, bc.lo[1], bc.hi[1], fy);
  const T* p = base + (long long)(c - 1 + g.g) + (long long)(r - 1 + g.g) * g.rs;
#pragma unroll
  for (int m = 0; m < ME; m++) out[m] = p[(long long)m * g.fs];
#pragma unroll
  for (int m = 0; m < ME; m++) {
    if (fx && m == bc.refl[0]) out[m] = -out[m];
    if (fy && m == bc.refl[1]) out[m] = -out[m];
  }
}

template <int ME, typename T>
__device__ __forceinline__ void load_cell3(const T* __restrict__ base, const Geom& g,
                                           const BCSpec& bc, int ng, int c, int r, int k,
                                           T (&out)[ME]) {
  bool fx = false, fy = false, fz = false;
  c = bc_map(c, g.mx, ng, bc.lo[0], bc.hi[0], fx);
  r = bc_map(r, g.my, ng, bc.lo[1], bc.hi[1], fy);
  k = bc_map(k, g.mz, ng, bc.lo[2], bc.hi[2], fz);
  const T* p = base + (long long)(c - 1 + g.g) + (long long)(r - 1 + g.g) * g.rs +
               (long long)(k - 1 + g.g) * g.ss;
#pragma unroll
  for (int m = 0; m < ME; m++) out[m] = p[(long long)m * g.fs];
#pragma unroll
  for (int m = 0; m < ME; m++) {
    if (fx && m == bc.refl[0]) out[m] = -out[m];
    if (fy && m == bc.refl[1]) out[m] = -out[m];
    if (fz && m == bc.refl[2]) out[m] = -out[m];
  }
}

template <typename T>
__device__ __forceinline__ T* cell_ptr2(T* base, const Geom& g, int c, int r) {
  return base + (long long)(c - 1 + g.g) + (long long)(r - 1 + g.g) * g.rs;
}
template <typename T>
__device__ __forceinline__ T* cell_ptr3(T* base, const Geom& g, int c, int r, int k) {
  return base + (long long)(c - 1 + g.g) + (long long)(r - 1 + g.g) * g.rs +
         (long long)(k - 1 + g.g) * g.ss;
}

// ---- cross-lane moves: value of lane-1 / lane+1 ------------------------------
// DPP whole-wavefront shifts (wave_shr:1 = 0x138: lane l reads lane l-1;
// wave_shl:1 = 0x130: lane l reads lane l+1), one VALU move per dword; the lane
// at the end (0, resp. 63) keeps its own value.
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_wave_shift(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 4) {
    const int x = __builtin_bit_cast(int, v);
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(x, x, CTRL, 0xf, 0xf, false));
  } else {
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = (int)x, hi = (int)(x >> 32);
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(T, ((long long)h2 << 32) | (long long)(unsigned)l2);
  }
}
template <typename T>
__device__ __forceinline__ T from_left(T v) { return dpp_wave_shift<0x138>(v); }
template <typename T>
__device__ __forceinline__ T from_right(T v) { return dpp_wave_shift<0x130>(v); }

template <int N, typename T>
__device__ __forceinline__ void from_left(const T (&v)[N], T (&o)[N]) {
#pragma unroll
  for (int m = 0; m < N; m++) o[m] = from_left(v[m]);
}
template <int N, typename T>
__device__ __forceinline__ void from_right(const T (&v)[N], T (&o)[N]) {
#pragma unroll
  for (int m = 0; m < N; m++) o[m] = from_right(v[m]);
}

// Running maximum of Courant-number terms (wave speeds, or dt/dx times them).  A
// NaN term -- a wave speed of a state that overflowed -- is left out, whatever
// the lane or the order: v > acc is false for a NaN v.  The reference leaves MAX
// with a NaN argument to the Fortran processor; the oracle applies this rule.
template <typename T>
__device__ __forceinline__ T cfl_term(T acc, T v) { return v > acc ? v : acc; }

// Wave-wide max of a lane's Courant number, then one atomic per wave.  Values
// that are not >= 0 (a NaN) count as 0, so the result does not depend on which
// lane held them.  Ordering of non-negative IEEE doubles equals the ordering of
// their bit patterns.
__device__ __forceinline__ void wave_atomic_max(double v, unsigned long long* dst) {
  if (!(v >= 0.0)) v = 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(v));
}

// The same for a workgroup of up to 256 threads: one atomic per workgroup.
// Every thread of the workgroup calls it (it holds a barrier).
__device__ __forceinline__ void block_atomic_max(double v, unsigned long long* dst) {
  __shared__ double part[256 / 64];
  if (!(v >= 0.0)) v = 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = part[w] > m ? part[w] : m;
    if (m > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// philim.f90:2-36
template <typename T>
__device__ __forceinline__ T philim(T a, T b, int meth) {
  T r = b / a;
  switch (meth) {
    case 1: return dmax(T(0), dmin(T(1), r));
    case 2: return dmax(dmax(T(0), dmin(T(1), T(2) * r)), dmin(T(2), r));
    case 3: return (r + fabs(r)) / (T(1) + fabs(r));
    case 4: {
      T c = (T(1) + r) / T(2);
      return dmax(T(0), dmin(dmin(c, T(2)), T(2) * r));
    }
    case 5: return r;
  }
  return T(0);
}

// One interface of limiter.f90:30-54: wnorm2 and the two dot products are of the
// unlimited waves; the upwind one (by the sign of s) enters phi.
template <int ME, int MW, int NL, typename T>
__device__ __forceinline__ void limit_waves(T (&W)[MW][ME], const T (&s)[MW], const T (&dotl)[MW],
                                            const T (&dotr)[MW], const int (&mthlim)[NL]) {
  static_assert(MW <= NL, "one limiter per wave");
#pragma unroll
  for (int w = 0; w < MW; w++) {
    if (mthlim[w] == 0) continue;
    T wnorm2 = T(0);
#pragma unroll
    for (int m = 0; m < ME; m++) wnorm2 = wnorm2 + W[w][m] * W[w][m];
    if (wnorm2 == T(0)) continue;
    T phi = s[w] > T(0) ? philim(wnorm2, dotl[w], mthlim[w]) : philim(wnorm2, dotr[w], mthlim[w]);
#pragma unroll
    for (int m = 0; m < ME; m++) W[w][m] = phi * W[w][m];
  }
}

template <int ME, int MW, typename T>
__device__ __forceinline__ void wave_dots(const T (&A)[MW][ME], const T (&B)[MW][ME], T (&d)[MW]) {
#pragma unroll
  for (int w = 0; w < MW; w++) {
    d[w] = T(0);
#pragma unroll
    for (int m = 0; m < ME; m++) d[w] = d[w] + A[w][m] * B[w][m];
  }
}

// flux2.f90:134-160: cqxx from limited waves, fadd = 0.5*cqxx.  dtdxave is
// 0.5*(dtdx1d(i-1) + dtdx1d(i)), the mean over the interface's two cells.
// With f-waves the factor |s| becomes dsign(1, s) (the f-wave already carries
// the speed); copysign keeps dsign's sign of a negative zero.
template <int ME, int MW, typename T>
__device__ __forceinline__ void second_order_ave(const T (&W)[MW][ME], const T (&s)[MW],
                                                 T dtdxave, T (&cq)[ME], T (&F)[ME],
                                                 int fwave = 0) {
#pragma unroll
  for (int m = 0; m < ME; m++) {
    cq[m] = T(0);
#pragma unroll
    for (int w = 0; w < MW; w++) {
      T sa = fabs(s[w]);
      T f = fwave ? copysign(T(1), s[w]) : sa;
      cq[m] = cq[m] + f * (T(1) - sa * dtdxave) * W[w][m];
    }
    F[m] = T(0) + T(0.5) * cq[m];
  }
}

// The same with one dtdx for every cell (no capacity function).
template <int ME, int MW, typename T>
__device__ __forceinline__ void second_order(const T (&W)[MW][ME], const T (&s)[MW], T dtdx,
                                             T (&cq)[ME], T (&F)[ME], int fwave = 0) {
  second_order_ave<ME, MW>(W, s, T(0.5) * (dtdx + dtdx), cq, F, fwave);
}

}  // namespace claw
