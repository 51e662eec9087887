// riemann.hpp -- device Riemann solvers for the classic and SharpClaw kernels.
//
// Each solver is a struct with compile-time meqn/mwaves and two pointwise
// device functions:
//   rpn<IXY>(p, qL, qR, wave, s, amdq, apdq, roe)   normal solve at one interface
//   rpt<IXY>(p, roe, asdq, bmasdq, bpasdq)           transverse split of asdq
// Solvers with aux fields (maux > 0) take rpn_aux / rpt_aux instead, with the
// aux values of the cells involved; rpn_call / rpt_call pick the right form.
// qL is the state left of the interface (the reference ABI's qr(:,i-1)), qR the
// state right of it (ql(:,i)).  IXY selects the normal momentum component.  The
// Roe averages rpn computes are handed to rpt in `roe`; the reference's rpt2
// recomputes them from the same two states with the same operations, so the
// values are identical.
//
// Algorithms (clawpack.riemann, restated; the in-tree form is
// development/rp_approaches/rpn2_euler_5wave.f and rpt2_euler_5wave.f):
//   acoustics_2D              2 waves (-c, +c)
//   euler_4wave_2D            Roe, 4 waves (u-a, shear, entropy, u+a), Harten-Hyman fix
//   shallow_roe_with_efix_2D  Roe, 3 waves (u-c, u, u+c), Harten-Hyman fix
//   advection_2D              1 wave
//   euler_3D                  Roe, 3 waves (u-a, lumped entropy+shear, u+a), fix
//   vc_advection(_vector)_2D  1 wave, edge velocities in aux
//   vc_acoustics_2D           2 waves (-c left, +c right), rho and c in aux
//   shallow_hlle_2D           HLLE, 2 waves at the Einfeldt speeds
//   shallow_bathymetry_fwave_2D  f-waves, 3 waves at the Einfeldt speeds, bathymetry in aux
// Every expression keeps the reference's evaluation order; the library is
// compiled with -ffp-contract=off so FP64 results match the CPU oracle bitwise.
// In the contracted (fast) build of the kernel sources (variant.hpp) the Euler
// solvers (2D and 3D) also share their reciprocals: one division per density, per Roe
// denominator and per sound speed squared, multiplications elsewhere
// (kSharedReciprocals).
#pragma once
#include <hip/hip_runtime.h>

#ifndef CLAW_FAST
#define CLAW_FAST 0
#endif

namespace claw {

// The fast variant's solvers divide once per shared denominator (internal
// linkage: each translation unit sees its own variant's value).
constexpr bool kSharedReciprocals = CLAW_FAST != 0;

template <typename T>
struct RPParams {
  T rho, bulk, cc, zz;   // acoustics
  T gamma, gamma1;       // Euler
  T grav;                // shallow water
  T u, v;                // advection
  T dry_tolerance;       // shallow water with bathymetry
  int efix;
  int fwave;             // the solver returns f-waves (step2's use_fwave)
};

template <typename T>
__device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }

// ---------------------------------------------------------------- acoustics
template <typename T>
struct Acoustics {
  static constexpr int meqn = 3, mwaves = 2, maux = 0;
  struct Roe {};
  template <int IXY>
  __device__ static __forceinline__ void rpn(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                             T (&amdq)[meqn], T (&apdq)[meqn], Roe&) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T zz = p.zz, cc = p.cc;
    T delta1 = qR[0] - qL[0];
    T delta2 = qR[mu] - qL[mu];
    T a1 = (-delta1 + zz * delta2) / (T(2) * zz);
    T a2 = (delta1 + zz * delta2) / (T(2) * zz);
    wave[0][0] = -a1 * zz; wave[0][mu] = a1; wave[0][mv] = T(0); s[0] = -cc;
    wave[1][0] = a2 * zz; wave[1][mu] = a2; wave[1][mv] = T(0); s[1] = cc;
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      amdq[m] = s[0] * wave[0][m];
      apdq[m] = s[1] * wave[1][m];
    }
  }
  template <int IXY>
  __device__ static __forceinline__ void rpt(const RPParams<T>& p, const Roe&, const T (&asdq)[meqn],
                                             T (&bm)[meqn], T (&bp)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T zz = p.zz, cc = p.cc;
    T a1 = (-asdq[0] + zz * asdq[mv]) / (T(2) * zz);
    T a2 = (asdq[0] + zz * asdq[mv]) / (T(2) * zz);
    bm[0] = cc * a1 * zz; bm[mu] = T(0); bm[mv] = -cc * a1;
    bp[0] = cc * a2 * zz; bp[mu] = T(0); bp[mv] = cc * a2;
  }
  // normal flux A q (bulk = rho c^2): amdq + apdq of rpn equals flux(qR) - flux(qL)
  template <int IXY>
  __device__ static __forceinline__ void flux(const RPParams<T>& p, const T* q, T (&f)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    f[0] = p.bulk * q[mu];
    f[mu] = q[0] / p.rho;
    f[mv] = T(0);
  }
};

// ---------------------------------------------------------------- 2D Euler
// TR = 0: euler_4wave_2D.  TR = 1: euler_5wave_2D, the same with a passive tracer
// q[4] carried by a fifth wave (its jump, speed u) in the normal solve and a
// fourth transverse wave (asdq[4], speed v); its components of the other waves
// are zero (development/rp_approaches/rpn2_euler_5wave.f:96-103,
// rpt2_euler_5wave.f:62-83).
template <typename T, int TR>
struct EulerT {
  static constexpr int meqn = 4 + TR, mwaves = 4 + TR, maux = 0;
  struct Roe { T u, v, enth, u2v2, a, g1a2, euv; };
  // normal flux; the Roe linearisation gives amdq + apdq = flux(qR) - flux(qL)
  template <int IXY>
  __device__ static __forceinline__ void flux(const RPParams<T>& p, const T* q, T (&f)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    T un, pr;
    if constexpr (kSharedReciprocals) {
      const T ir = T(1) / q[0];
      un = q[mu] * ir;
      pr = p.gamma1 * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) * ir);
    } else {
      un = q[mu] / q[0];
      pr = p.gamma1 * (q[3] - T(0.5) * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    }
    f[0] = q[mu];
    f[mu] = q[mu] * un + pr;
    f[mv] = q[mv] * un;
    f[3] = (q[3] + pr) * un;
    if constexpr (TR) f[4] = q[4] * un;
  }

  // The fast variant's Roe average: 1/rho per side (also 1/sqrt(rho) =
  // sqrt(rho)/rho), 1/(sqrt(rhoL) + sqrt(rhoR)) and 1/a^2 once each.
  template <int IXY>
  __device__ static __forceinline__ Roe roe_avg_shared(T gamma1, const T* qL, const T* qR,
                                                       T& irl, T& irr, T& ia2) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    Roe r;
    irl = T(1) / qL[0];
    irr = T(1) / qR[0];
    const T rhsqrtl = sqrt(qL[0]), rhsqrtr = sqrt(qR[0]);
    const T isl = rhsqrtl * irl, isr = rhsqrtr * irr;
    const T pl = gamma1 * (qL[3] - T(0.5) * (qL[1] * qL[1] + qL[2] * qL[2]) * irl);
    const T pr = gamma1 * (qR[3] - T(0.5) * (qR[1] * qR[1] + qR[2] * qR[2]) * irr);
    const T irq = T(1) / (rhsqrtl + rhsqrtr);
    r.u = (qL[mu] * isl + qR[mu] * isr) * irq;
    r.v = (qL[mv] * isl + qR[mv] * isr) * irq;
    r.enth = ((qL[3] + pl) * isl + (qR[3] + pr) * isr) * irq;
    r.u2v2 = r.u * r.u + r.v * r.v;
    const T a2 = gamma1 * (r.enth - T(.5) * r.u2v2);
    r.a = sqrt(a2);
    ia2 = T(1) / a2;
    r.g1a2 = gamma1 * ia2;
    r.euv = r.enth - r.u2v2;
    return r;
  }

  template <int IXY>
  __device__ static __forceinline__ Roe roe_avg(T gamma1, const T* qL, const T* qR) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    Roe r;
    T rhsqrtl = sqrt(qL[0]);
    T rhsqrtr = sqrt(qR[0]);
    T pl = gamma1 * (qL[3] - T(0.5) * (qL[1] * qL[1] + qL[2] * qL[2]) / qL[0]);
    T pr = gamma1 * (qR[3] - T(0.5) * (qR[1] * qR[1] + qR[2] * qR[2]) / qR[0]);
    T rhsq2 = rhsqrtl + rhsqrtr;
    r.u = (qL[mu] / rhsqrtl + qR[mu] / rhsqrtr) / rhsq2;
    r.v = (qL[mv] / rhsqrtl + qR[mv] / rhsqrtr) / rhsq2;
    r.enth = (((qL[3] + pl) / rhsqrtl + (qR[3] + pr) / rhsqrtr)) / rhsq2;
    r.u2v2 = r.u * r.u + r.v * r.v;
    T a2 = gamma1 * (r.enth - T(.5) * r.u2v2);
    r.a = sqrt(a2);
    r.g1a2 = gamma1 / a2;
    r.euv = r.enth - r.u2v2;
    return r;
  }

  template <int IXY>
  __device__ static __forceinline__ void rpn(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                             T (&amdq)[meqn], T (&apdq)[meqn], Roe& r) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T gamma = p.gamma, gamma1 = p.gamma1;
    T irl = T(0), irr = T(0), ia2 = T(0);   // shared reciprocals (fast variant)
    if constexpr (kSharedReciprocals) r = roe_avg_shared<IXY>(gamma1, qL, qR, irl, irr, ia2);
    else r = roe_avg<IXY>(gamma1, qL, qR);
    T d0 = qR[0] - qL[0], d1 = qR[mu] - qL[mu], d2 = qR[mv] - qL[mv], d3 = qR[3] - qL[3];
    T a3 = r.g1a2 * (r.euv * d0 + r.u * d1 + r.v * d2 - d3);
    T a2 = d2 - r.v * d0;
    T a4;
    if constexpr (kSharedReciprocals)   // 1/(2a) = a/(2a^2)
      a4 = (d1 + (r.a - r.u) * d0 - r.a * a3) * (T(0.5) * r.a * ia2);
    else
      a4 = (d1 + (r.a - r.u) * d0 - r.a * a3) / (T(2) * r.a);
    T a1 = d0 - a3 - a4;
    wave[0][0] = a1; wave[0][mu] = a1 * (r.u - r.a); wave[0][mv] = a1 * r.v;
    wave[0][3] = a1 * (r.enth - r.u * r.a); s[0] = r.u - r.a;
    wave[1][0] = T(0); wave[1][mu] = T(0); wave[1][mv] = a2; wave[1][3] = a2 * r.v; s[1] = r.u;
    wave[2][0] = a3; wave[2][mu] = a3 * r.u; wave[2][mv] = a3 * r.v;
    wave[2][3] = a3 * T(0.5) * r.u2v2; s[2] = r.u;
    wave[3][0] = a4; wave[3][mu] = a4 * (r.u + r.a); wave[3][mv] = a4 * r.v;
    wave[3][3] = a4 * (r.enth + r.u * r.a); s[3] = r.u + r.a;
    if constexpr (TR) {
      wave[0][4] = T(0); wave[1][4] = T(0); wave[2][4] = T(0); wave[3][4] = T(0);
      wave[4][0] = T(0); wave[4][1] = T(0); wave[4][2] = T(0); wave[4][3] = T(0);
      wave[4][4] = qR[4] - qL[4]; s[4] = r.u;
    }

    if (!p.efix) {
#pragma unroll
      for (int m = 0; m < meqn; m++) {
        amdq[m] = T(0); apdq[m] = T(0);
#pragma unroll
        for (int w = 0; w < mwaves; w++) {
          if (s[w] < T(0)) amdq[m] = amdq[m] + s[w] * wave[w][m];
          else apdq[m] = apdq[m] + s[w] * wave[w][m];
        }
      }
      return;
    }
    do {
      T rhoim1 = qL[0];
      T pim1, cim1, s0;
      if constexpr (kSharedReciprocals) {
        pim1 = gamma1 * (qL[3] - T(0.5) * (qL[mu] * qL[mu] + qL[mv] * qL[mv]) * irl);
        cim1 = sqrt(gamma * pim1 * irl);
        s0 = qL[mu] * irl - cim1;
      } else {
        pim1 = gamma1 * (qL[3] - T(0.5) * (qL[mu] * qL[mu] + qL[mv] * qL[mv]) / rhoim1);
        cim1 = sqrt(gamma * pim1 / rhoim1);
        s0 = qL[mu] / rhoim1 - cim1;
      }
      if (s0 >= T(0) && s[0] > T(0)) {
#pragma unroll
        for (int m = 0; m < meqn; m++) amdq[m] = T(0);
        break;
      }
      T rho1 = qL[0] + wave[0][0];
      T rhou1 = qL[mu] + wave[0][mu];
      T rhov1 = qL[mv] + wave[0][mv];
      T en1 = qL[3] + wave[0][3];
      T p1, c1, s1;
      if constexpr (kSharedReciprocals) {
        const T ir1 = T(1) / rho1;
        p1 = gamma1 * (en1 - T(0.5) * (rhou1 * rhou1 + rhov1 * rhov1) * ir1);
        c1 = sqrt(gamma * p1 * ir1);
        s1 = rhou1 * ir1 - c1;
      } else {
        p1 = gamma1 * (en1 - T(0.5) * (rhou1 * rhou1 + rhov1 * rhov1) / rho1);
        c1 = sqrt(gamma * p1 / rho1);
        s1 = rhou1 / rho1 - c1;
      }
      T sfract;
      if (s0 < T(0) && s1 > T(0)) sfract = s0 * (s1 - s[0]) / (s1 - s0);
      else if (s[0] < T(0)) sfract = s[0];
      else sfract = T(0);
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = sfract * wave[0][m];
      if (s[1] >= T(0)) break;
#pragma unroll
      for (int m = 0; m < meqn; m++) {
        amdq[m] = amdq[m] + s[1] * wave[1][m];
        amdq[m] = amdq[m] + s[2] * wave[2][m];
        if constexpr (TR) amdq[m] = amdq[m] + s[4] * wave[4][m];
      }
      T rhoi = qR[0];
      T pi, ci, s3;
      if constexpr (kSharedReciprocals) {
        pi = gamma1 * (qR[3] - T(0.5) * (qR[mu] * qR[mu] + qR[mv] * qR[mv]) * irr);
        ci = sqrt(gamma * pi * irr);
        s3 = qR[mu] * irr + ci;
      } else {
        pi = gamma1 * (qR[3] - T(0.5) * (qR[mu] * qR[mu] + qR[mv] * qR[mv]) / rhoi);
        ci = sqrt(gamma * pi / rhoi);
        s3 = qR[mu] / rhoi + ci;
      }
      T rho2 = qR[0] - wave[3][0];
      T rhou2 = qR[mu] - wave[3][mu];
      T rhov2 = qR[mv] - wave[3][mv];
      T en2 = qR[3] - wave[3][3];
      T p2, c2, s2;
      if constexpr (kSharedReciprocals) {
        const T ir2 = T(1) / rho2;
        p2 = gamma1 * (en2 - T(0.5) * (rhou2 * rhou2 + rhov2 * rhov2) * ir2);
        c2 = sqrt(gamma * p2 * ir2);
        s2 = rhou2 * ir2 + c2;
      } else {
        p2 = gamma1 * (en2 - T(0.5) * (rhou2 * rhou2 + rhov2 * rhov2) / rho2);
        c2 = sqrt(gamma * p2 / rho2);
        s2 = rhou2 / rho2 + c2;
      }
      if (s2 < T(0) && s3 > T(0)) sfract = s2 * (s3 - s[3]) / (s3 - s2);
      else if (s[3] < T(0)) sfract = s[3];
      else break;
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = amdq[m] + sfract * wave[3][m];
    } while (0);
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      T df = T(0);
#pragma unroll
      for (int w = 0; w < mwaves; w++) df = df + s[w] * wave[w][m];
      apdq[m] = df - amdq[m];
    }
  }

  template <int IXY>
  __device__ static __forceinline__ void rpt(const RPParams<T>&, const Roe& r, const T (&asdq)[meqn],
                                             T (&bm)[meqn], T (&bp)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    T a3 = r.g1a2 * (r.euv * asdq[0] + r.u * asdq[mu] + r.v * asdq[mv] - asdq[3]);
    T a2 = asdq[mu] - r.u * asdq[0];
    T a4 = (asdq[mv] + (r.a - r.v) * asdq[0] - r.a * a3) / (T(2) * r.a);
    T a1 = asdq[0] - a3 - a4;
    T wb[3 + TR][meqn], sb[3 + TR];
    wb[0][0] = a1; wb[0][mu] = a1 * r.u; wb[0][mv] = a1 * (r.v - r.a);
    wb[0][3] = a1 * (r.enth - r.v * r.a); sb[0] = r.v - r.a;
    wb[1][0] = a3; wb[1][mu] = a3 * r.u + a2; wb[1][mv] = a3 * r.v;
    wb[1][3] = a3 * T(0.5) * r.u2v2 + a2 * r.u; sb[1] = r.v;
    wb[2][0] = a4; wb[2][mu] = a4 * r.u; wb[2][mv] = a4 * (r.v + r.a);
    wb[2][3] = a4 * (r.enth + r.v * r.a); sb[2] = r.v + r.a;
    if constexpr (TR) {
      wb[0][4] = T(0); wb[1][4] = T(0); wb[2][4] = T(0);
      wb[3][0] = T(0); wb[3][1] = T(0); wb[3][2] = T(0); wb[3][3] = T(0);
      wb[3][4] = asdq[4]; sb[3] = r.v;
    }
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      bm[m] = T(0); bp[m] = T(0);
#pragma unroll
      for (int w = 0; w < 3 + TR; w++) {
        bm[m] = bm[m] + dmin(sb[w], T(0)) * wb[w][m];
        bp[m] = bp[m] + dmax(sb[w], T(0)) * wb[w][m];
      }
    }
  }
};

template <typename T> using Euler = EulerT<T, 0>;
template <typename T> using Euler5 = EulerT<T, 1>;

// ---------------------------------------------------------------- shallow water
template <typename T>
struct Shallow {
  static constexpr int meqn = 3, mwaves = 3, maux = 0;
  struct Roe { T u, v, a; };
  // normal flux; the Roe linearisation gives amdq + apdq = flux(qR) - flux(qL)
  template <int IXY>
  __device__ static __forceinline__ void flux(const RPParams<T>& p, const T* q, T (&f)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T un = q[mu] / q[0];
    f[0] = q[mu];
    f[mu] = q[mu] * un + T(0.5) * p.grav * q[0] * q[0];
    f[mv] = q[mv] * un;
  }

  template <int IXY>
  __device__ static __forceinline__ void rpn(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                             T (&amdq)[meqn], T (&apdq)[meqn], Roe& r) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T grav = p.grav;
    T h = (qL[0] + qR[0]) * T(0.50);
    T hsqrtl = sqrt(qL[0]);
    T hsqrtr = sqrt(qR[0]);
    T hsq2 = hsqrtl + hsqrtr;
    T u = (qL[mu] / hsqrtl + qR[mu] / hsqrtr) / hsq2;
    T v = (qL[mv] / hsqrtl + qR[mv] / hsqrtr) / hsq2;
    T a = sqrt(grav * h);
    r.u = u; r.v = v; r.a = a;
    T delta0 = qR[0] - qL[0];
    T delta1 = qR[mu] - qL[mu];
    T delta2 = qR[mv] - qL[mv];
    T a1 = ((u + a) * delta0 - delta1) * (T(0.50) / a);
    T a2 = -v * delta0 + delta2;
    T a3 = (-(u - a) * delta0 + delta1) * (T(0.50) / a);
    wave[0][0] = a1; wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v; s[0] = u - a;
    wave[1][0] = T(0); wave[1][mu] = T(0); wave[1][mv] = a2; s[1] = u;
    wave[2][0] = a3; wave[2][mu] = a3 * (u + a); wave[2][mv] = a3 * v; s[2] = u + a;
    if (!p.efix) {
#pragma unroll
      for (int m = 0; m < meqn; m++) {
        amdq[m] = T(0); apdq[m] = T(0);
#pragma unroll
        for (int w = 0; w < mwaves; w++) {
          if (s[w] < T(0)) amdq[m] = amdq[m] + s[w] * wave[w][m];
          else apdq[m] = apdq[m] + s[w] * wave[w][m];
        }
      }
      return;
    }
    do {
      T him1 = qL[0];
      T s0 = qL[mu] / him1 - sqrt(grav * him1);
      if (s0 >= T(0) && s[0] > T(0)) {
#pragma unroll
        for (int m = 0; m < meqn; m++) amdq[m] = T(0);
        break;
      }
      T h1 = qL[0] + wave[0][0];
      T hu1 = qL[mu] + wave[0][mu];
      T s1 = hu1 / h1 - sqrt(grav * h1);
      T sfract;
      if (s0 < T(0) && s1 > T(0)) sfract = s0 * (s1 - s[0]) / (s1 - s0);
      else if (s[0] < T(0)) sfract = s[0];
      else sfract = T(0);
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = sfract * wave[0][m];
      if (s[1] >= T(0)) break;
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = amdq[m] + s[1] * wave[1][m];
      T hi = qR[0];
      T s03 = qR[mu] / hi + sqrt(grav * hi);
      T h3 = qR[0] - wave[2][0];
      T hu3 = qR[mu] - wave[2][mu];
      T s3 = hu3 / h3 + sqrt(grav * h3);
      if (s3 < T(0) && s03 > T(0)) sfract = s3 * (s03 - s[2]) / (s03 - s3);
      else if (s[2] < T(0)) sfract = s[2];
      else break;
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = amdq[m] + sfract * wave[2][m];
    } while (0);
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      T df = T(0);
#pragma unroll
      for (int w = 0; w < mwaves; w++) df = df + s[w] * wave[w][m];
      apdq[m] = df - amdq[m];
    }
  }

  // The transverse split recomputes the Roe average with the same operations as
  // rpn except that the wave speed uses the mean depth; it is carried in `r`.
  template <int IXY>
  __device__ static __forceinline__ void rpt(const RPParams<T>&, const Roe& r, const T (&asdq)[meqn],
                                             T (&bm)[meqn], T (&bp)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T u = r.u, v = r.v, a = r.a;
    T a1 = (T(0.50) / a) * ((v + a) * asdq[0] - asdq[mv]);
    T a2 = asdq[mu] - u * asdq[0];
    T a3 = (T(0.50) / a) * (-(v - a) * asdq[0] + asdq[mv]);
    T wb[3][meqn], sb[3];
    wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); sb[0] = v - a;
    wb[1][0] = T(0); wb[1][mu] = a2; wb[1][mv] = T(0); sb[1] = v;
    wb[2][0] = a3; wb[2][mu] = a3 * u; wb[2][mv] = a3 * (v + a); sb[2] = v + a;
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      bm[m] = T(0); bp[m] = T(0);
#pragma unroll
      for (int w = 0; w < 3; w++) {
        bm[m] = bm[m] + dmin(sb[w], T(0)) * wb[w][m];
        bp[m] = bp[m] + dmax(sb[w], T(0)) * wb[w][m];
      }
    }
  }
};

// ---------------------------------------------------------------- advection
template <typename T>
struct Advection {
  static constexpr int meqn = 1, mwaves = 1, maux = 0;
  struct Roe {};
  template <int IXY>
  __device__ static __forceinline__ void rpn(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                             T (&amdq)[meqn], T (&apdq)[meqn], Roe&) {
    wave[0][0] = qR[0] - qL[0];
    s[0] = IXY == 1 ? p.u : p.v;
    amdq[0] = dmin(s[0], T(0)) * wave[0][0];
    apdq[0] = dmax(s[0], T(0)) * wave[0][0];
  }
  template <int IXY>
  __device__ static __forceinline__ void rpt(const RPParams<T>& p, const Roe&, const T (&asdq)[meqn],
                                             T (&bm)[meqn], T (&bp)[meqn]) {
    T stran = IXY == 1 ? p.v : p.u;
    bm[0] = dmin(stran, T(0)) * asdq[0];
    bp[0] = dmax(stran, T(0)) * asdq[0];
  }
  template <int IXY>
  __device__ static __forceinline__ void flux(const RPParams<T>& p, const T* q, T (&f)[meqn]) {
    f[0] = (IXY == 1 ? p.u : p.v) * q[0];
  }
};

// ---------------------------------------------------------------- 3D Euler
template <typename T>
struct Euler3 {
  static constexpr int meqn = 5, mwaves = 3, maux = 0;
  // the interface's Roe averages, kept for the transverse splits (u is the
  // normal velocity, v and w the next two in the cycle x -> y -> z)
  struct Roe { T u, v, w, enth, a, g1a2, euv, u2v2w2; };
  // IXY = 1, 2, 3 selects x, y, z as the normal direction.
  template <int IXY>
  __device__ static __forceinline__ void rpn(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                             T (&amdq)[meqn], T (&apdq)[meqn], Roe& roe) {
    constexpr int mu = IXY == 1 ? 1 : (IXY == 2 ? 2 : 3);
    constexpr int mv = IXY == 1 ? 2 : (IXY == 2 ? 3 : 1);
    constexpr int mw = IXY == 1 ? 3 : (IXY == 2 ? 1 : 2);
    const T gamma = p.gamma, gamma1 = p.gamma1;
    T rhsqrtl = sqrt(qL[0]);
    T rhsqrtr = sqrt(qR[0]);
    // the fast variant's shared reciprocals (kSharedReciprocals): 1/rho per side
    // (and 1/sqrt(rho) = sqrt(rho)/rho), the Roe denominator, 1/a^2
    T irl = T(0), irr = T(0), ia2 = T(0);
    T pl, pr, u, v, w, enth;
    if constexpr (kSharedReciprocals) {
      irl = T(1) / qL[0];
      irr = T(1) / qR[0];
      const T isl = rhsqrtl * irl, isr = rhsqrtr * irr;
      pl = gamma1 * (qL[4] - T(0.5) * (qL[1] * qL[1] + qL[2] * qL[2] + qL[3] * qL[3]) * irl);
      pr = gamma1 * (qR[4] - T(0.5) * (qR[1] * qR[1] + qR[2] * qR[2] + qR[3] * qR[3]) * irr);
      const T irq = T(1) / (rhsqrtl + rhsqrtr);
      u = (qL[mu] * isl + qR[mu] * isr) * irq;
      v = (qL[mv] * isl + qR[mv] * isr) * irq;
      w = (qL[mw] * isl + qR[mw] * isr) * irq;
      enth = ((qL[4] + pl) * isl + (qR[4] + pr) * isr) * irq;
    } else {
      pl = gamma1 * (qL[4] - T(0.5) * (qL[1] * qL[1] + qL[2] * qL[2] + qL[3] * qL[3]) / qL[0]);
      pr = gamma1 * (qR[4] - T(0.5) * (qR[1] * qR[1] + qR[2] * qR[2] + qR[3] * qR[3]) / qR[0]);
      T rhsq2 = rhsqrtl + rhsqrtr;
      u = (qL[mu] / rhsqrtl + qR[mu] / rhsqrtr) / rhsq2;
      v = (qL[mv] / rhsqrtl + qR[mv] / rhsqrtr) / rhsq2;
      w = (qL[mw] / rhsqrtl + qR[mw] / rhsqrtr) / rhsq2;
      enth = (((qL[4] + pl) / rhsqrtl + (qR[4] + pr) / rhsqrtr)) / rhsq2;
    }
    T u2v2w2 = u * u + v * v + w * w;
    T a2 = gamma1 * (enth - T(.5) * u2v2w2);
    T a = sqrt(a2);
    T g1a2;
    if constexpr (kSharedReciprocals) {
      ia2 = T(1) / a2;
      g1a2 = gamma1 * ia2;
    } else {
      g1a2 = gamma1 / a2;
    }
    T euv = enth - u2v2w2;
    roe.u = u; roe.v = v; roe.w = w; roe.enth = enth; roe.a = a; roe.g1a2 = g1a2;
    roe.euv = euv; roe.u2v2w2 = u2v2w2;
    T d0 = qR[0] - qL[0], d1 = qR[mu] - qL[mu], d2 = qR[mv] - qL[mv], d3 = qR[mw] - qL[mw],
      d4 = qR[4] - qL[4];
    T a4 = g1a2 * (euv * d0 + u * d1 + v * d2 + w * d3 - d4);
    T b2 = d2 - v * d0;
    T b3 = d3 - w * d0;
    T a5;
    if constexpr (kSharedReciprocals)   // 1/(2a) = a/(2a^2)
      a5 = (d1 + (a - u) * d0 - a * a4) * (T(0.5) * a * ia2);
    else
      a5 = (d1 + (a - u) * d0 - a * a4) / (T(2) * a);
    T a1 = d0 - a4 - a5;
    wave[0][0] = a1; wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v; wave[0][mw] = a1 * w;
    wave[0][4] = a1 * (enth - u * a); s[0] = u - a;
    wave[1][0] = a4; wave[1][mu] = a4 * u; wave[1][mv] = a4 * v + b2; wave[1][mw] = a4 * w + b3;
    wave[1][4] = a4 * T(0.5) * u2v2w2 + b2 * v + b3 * w; s[1] = u;
    wave[2][0] = a5; wave[2][mu] = a5 * (u + a); wave[2][mv] = a5 * v; wave[2][mw] = a5 * w;
    wave[2][4] = a5 * (enth + u * a); s[2] = u + a;
    if (!p.efix) {
#pragma unroll
      for (int m = 0; m < meqn; m++) {
        amdq[m] = T(0); apdq[m] = T(0);
#pragma unroll
        for (int k = 0; k < mwaves; k++) {
          if (s[k] < T(0)) amdq[m] = amdq[m] + s[k] * wave[k][m];
          else apdq[m] = apdq[m] + s[k] * wave[k][m];
        }
      }
      return;
    }
    do {
      T rhoim1 = qL[0];
      T pim1, cim1, s0;
      if constexpr (kSharedReciprocals) {
        pim1 = gamma1 * (qL[4] - T(0.5) * (qL[mu] * qL[mu] + qL[mv] * qL[mv] + qL[mw] * qL[mw]) * irl);
        cim1 = sqrt(gamma * pim1 * irl);
        s0 = qL[mu] * irl - cim1;
      } else {
        pim1 = gamma1 * (qL[4] - T(0.5) * (qL[mu] * qL[mu] + qL[mv] * qL[mv] + qL[mw] * qL[mw]) / rhoim1);
        cim1 = sqrt(gamma * pim1 / rhoim1);
        s0 = qL[mu] / rhoim1 - cim1;
      }
      if (s0 >= T(0) && s[0] > T(0)) {
#pragma unroll
        for (int m = 0; m < meqn; m++) amdq[m] = T(0);
        break;
      }
      T rho1 = qL[0] + wave[0][0];
      T rhou1 = qL[mu] + wave[0][mu];
      T rhov1 = qL[mv] + wave[0][mv];
      T rhow1 = qL[mw] + wave[0][mw];
      T en1 = qL[4] + wave[0][4];
      T p1, c1, s1;
      if constexpr (kSharedReciprocals) {
        const T ir1 = T(1) / rho1;
        p1 = gamma1 * (en1 - T(0.5) * (rhou1 * rhou1 + rhov1 * rhov1 + rhow1 * rhow1) * ir1);
        c1 = sqrt(gamma * p1 * ir1);
        s1 = rhou1 * ir1 - c1;
      } else {
        p1 = gamma1 * (en1 - T(0.5) * (rhou1 * rhou1 + rhov1 * rhov1 + rhow1 * rhow1) / rho1);
        c1 = sqrt(gamma * p1 / rho1);
        s1 = rhou1 / rho1 - c1;
      }
      T sfract;
      if (s0 < T(0) && s1 > T(0)) sfract = s0 * (s1 - s[0]) / (s1 - s0);
      else if (s[0] < T(0)) sfract = s[0];
      else sfract = T(0);
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = sfract * wave[0][m];
      if (s[1] >= T(0)) break;
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = amdq[m] + s[1] * wave[1][m];
      T rhoi = qR[0];
      T pi, ci, s3;
      if constexpr (kSharedReciprocals) {
        pi = gamma1 * (qR[4] - T(0.5) * (qR[mu] * qR[mu] + qR[mv] * qR[mv] + qR[mw] * qR[mw]) * irr);
        ci = sqrt(gamma * pi * irr);
        s3 = qR[mu] * irr + ci;
      } else {
        pi = gamma1 * (qR[4] - T(0.5) * (qR[mu] * qR[mu] + qR[mv] * qR[mv] + qR[mw] * qR[mw]) / rhoi);
        ci = sqrt(gamma * pi / rhoi);
        s3 = qR[mu] / rhoi + ci;
      }
      T rho2 = qR[0] - wave[2][0];
      T rhou2 = qR[mu] - wave[2][mu];
      T rhov2 = qR[mv] - wave[2][mv];
      T rhow2 = qR[mw] - wave[2][mw];
      T en2 = qR[4] - wave[2][4];
      T p2, c2, s2;
      if constexpr (kSharedReciprocals) {
        const T ir2 = T(1) / rho2;
        p2 = gamma1 * (en2 - T(0.5) * (rhou2 * rhou2 + rhov2 * rhov2 + rhow2 * rhow2) * ir2);
        c2 = sqrt(gamma * p2 * ir2);
        s2 = rhou2 * ir2 + c2;
      } else {
        p2 = gamma1 * (en2 - T(0.5) * (rhou2 * rhou2 + rhov2 * rhov2 + rhow2 * rhow2) / rho2);
        c2 = sqrt(gamma * p2 / rho2);
        s2 = rhou2 / rho2 + c2;
      }
      if (s2 < T(0) && s3 > T(0)) sfract = s2 * (s3 - s[2]) / (s3 - s2);
      else if (s[2] < T(0)) sfract = s[2];
      else break;
#pragma unroll
      for (int m = 0; m < meqn; m++) amdq[m] = amdq[m] + sfract * wave[2][m];
    } while (0);
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      T df = T(0);
#pragma unroll
      for (int k = 0; k < mwaves; k++) df = df + s[k] * wave[k][m];
      apdq[m] = df - amdq[m];
    }
  }
  template <int IXY>
  __device__ static __forceinline__ void rpt(const RPParams<T>&, const Roe&, const T (&)[meqn],
                                             T (&bm)[meqn], T (&bp)[meqn]) {
#pragma unroll
    for (int m = 0; m < meqn; m++) { bm[m] = T(0); bp[m] = T(0); }
  }
  // euler_3D rpt3 / rptt3 (clawpack.riemann rpt3_euler.f90, rptt3_euler.f90,
  // published form): asdq split in the transverse direction ICOOR (2: the
  // direction after IXY in the cycle, momentum mv; 3: the next, mw) with the
  // interface's Roe averages; three waves at vt - a, vt, vt + a with the shear
  // parts lumped into the middle one; bm / bp take the negative / positive speeds.
  template <int IXY, int ICOOR>
  __device__ static __forceinline__ void rpt3(const Roe& r, const T (&asdq)[meqn],
                                              T (&bm)[meqn], T (&bp)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : (IXY == 2 ? 2 : 3);
    constexpr int mv = IXY == 1 ? 2 : (IXY == 2 ? 3 : 1);
    constexpr int mw = IXY == 1 ? 3 : (IXY == 2 ? 1 : 2);
    constexpr int mt = ICOOR == 2 ? mv : mw, m1 = mu, m2 = ICOOR == 2 ? mw : mv;
    const T u = r.u, v = r.v, w = r.w, a = r.a, enth = r.enth;
    const T vt = ICOOR == 2 ? v : w, c1 = u, c2 = ICOOR == 2 ? w : v;
    T a4 = r.g1a2 * (r.euv * asdq[0] + u * asdq[mu] + v * asdq[mv] + w * asdq[mw] - asdq[4]);
    T b1 = asdq[m1] - c1 * asdq[0];
    T b2 = asdq[m2] - c2 * asdq[0];
    T a5 = (asdq[mt] + (a - vt) * asdq[0] - a * a4) / (T(2) * a);
    T a1 = asdq[0] - a4 - a5;
    T wv[3][meqn], sb[3];
    wv[0][0] = a1; wv[0][m1] = a1 * c1; wv[0][mt] = a1 * (vt - a); wv[0][m2] = a1 * c2;
    wv[0][4] = a1 * (enth - vt * a); sb[0] = vt - a;
    wv[1][0] = a4; wv[1][m1] = a4 * c1 + b1; wv[1][mt] = a4 * vt; wv[1][m2] = a4 * c2 + b2;
    wv[1][4] = a4 * T(0.5) * r.u2v2w2 + b1 * c1 + b2 * c2; sb[1] = vt;
    wv[2][0] = a5; wv[2][m1] = a5 * c1; wv[2][mt] = a5 * (vt + a); wv[2][m2] = a5 * c2;
    wv[2][4] = a5 * (enth + vt * a); sb[2] = vt + a;
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      bm[m] = T(0);
      bp[m] = T(0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        bm[m] = bm[m] + dmin(sb[k], T(0)) * wv[k][m];
        bp[m] = bp[m] + dmax(sb[k], T(0)) * wv[k][m];
      }
    }
  }
};

// ---------------------------------------------------------------- shallow water, HLLE
// shallow_hlle_2D of the clawpack/riemann package (rpn2_shallow_hlle.f90, not
// part of the reference tree: restated from its published form).  Two waves
// at the Einfeldt speeds s1 = min(u_l - c_l, ubar - cbar) and
// s2 = max(u_r + c_r, ubar + cbar) (ubar the Roe average, cbar =
// sqrt(g (h_l + h_r) / 2)) around the HLL middle state
// q* = (f(q_r) - f(q_l) - s2 q_r + s1 q_l) / (s1 - s2); each wave's flux
// difference goes by the sign of its speed.  No transverse solver is paired
// with it here: unsplit steps take it with transverse_waves = 0.
template <typename T>
struct ShallowHLLE {
  static constexpr int meqn = 3, mwaves = 2, maux = 0;
  struct Roe {};
  // normal flux; the HLL waves give amdq + apdq = flux(qR) - flux(qL)
  template <int IXY>
  __device__ static __forceinline__ void flux(const RPParams<T>& p, const T* q, T (&f)[meqn]) {
    Shallow<T>::template flux<IXY>(p, q, f);
  }
  template <int IXY>
  __device__ static __forceinline__ void rpn(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                             T (&amdq)[meqn], T (&apdq)[meqn], Roe&) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T grav = p.grav;
    const T h_l = qL[0], h_r = qR[0];
    const T u_l = qL[mu] / qL[0], u_r = qR[mu] / qR[0];
    const T hsqrt_l = sqrt(qL[0]), hsqrt_r = sqrt(qR[0]);
    const T hsq2 = hsqrt_l + hsqrt_r;
    const T ubar = (u_l * hsqrt_l + u_r * hsqrt_r) / hsq2;
    const T cbar = sqrt(T(0.5) * grav * (h_l + h_r));
    const T s1 = dmin(u_l - sqrt(grav * h_l), ubar - cbar);
    const T s2 = dmax(u_r + sqrt(grav * h_r), ubar + cbar);
    const T hm = (qR[mu] - qL[mu] - s2 * qR[0] + s1 * qL[0]) / (s1 - s2);
    const T hum = (qR[mu] * u_r - qL[mu] * u_l + T(0.5) * grav * (h_r * h_r - h_l * h_l) -
                   s2 * qR[mu] + s1 * qL[mu]) / (s1 - s2);
    const T hvm = (qR[mv] * u_r - qL[mv] * u_l - s2 * qR[mv] + s1 * qL[mv]) / (s1 - s2);
    wave[0][0] = hm - h_l;
    wave[0][mu] = hum - qL[mu];
    wave[0][mv] = hvm - qL[mv];
    s[0] = s1;
    wave[1][0] = h_r - hm;
    wave[1][mu] = qR[mu] - hum;
    wave[1][mv] = qR[mv] - hvm;
    s[1] = s2;
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      amdq[m] = T(0);
      apdq[m] = T(0);
#pragma unroll
      for (int w = 0; w < mwaves; w++) {
        if (s[w] < T(0)) amdq[m] = amdq[m] + s[w] * wave[w][m];
        else apdq[m] = apdq[m] + s[w] * wave[w][m];
      }
    }
  }
  // Not reached: the entry points refuse transverse_waves > 0 for this solver.
  template <int IXY>
  __device__ static __forceinline__ void rpt(const RPParams<T>&, const Roe&, const T (&)[meqn],
                                             T (&bm)[meqn], T (&bp)[meqn]) {
#pragma unroll
    for (int m = 0; m < meqn; m++) { bm[m] = T(0); bp[m] = T(0); }
  }
};

// ---------------------------------------------------------------- variable-coefficient advection
// examples/advection_reaction_2d/rpn2_vc_advection_vector.f90:53-60: the
// non-conservative transport q_t + u q_x + v q_y = 0 of ME fields, velocities
// given at cell edges in aux (aux[0] = u at the cell's left edge, aux[1] = v
// at its lower edge).  Interface i (left cell i-1, right cell i): one wave
// q_i - q_{i-1} with speed the right cell's edge velocity; its flux
// difference goes downwind.
//
// rpt_aux is the transverse split of the scalar solver vc_advection_2D of the
// clawpack/riemann package (rpt2_vc_advection.f90; that package is not part of
// the reference tree, so this restates its published form): the transverse
// velocity kv (v for x-slices, u for y-slices) of the cell i1 that receives
// asdq, taken at i1's lower edge (aux2, in the slice) for the down-going part
// and at its upper edge (aux3, the next slice) for the up-going part:
//   bmasdq = min(aux2(kv, i1), 0) * asdq,  bpasdq = max(aux3(kv, i1), 0) * asdq.
// Applied to each of the ME fields.  The vector solver of
// examples/advection_reaction_2d pairs rpn2 with a dummy rpt2, so kind 6 is
// used with dimension splitting or with transverse_waves = 0.
template <typename T, int ME>
struct VCAdvection {
  static constexpr int meqn = ME, mwaves = 1, maux = 2;
  struct Roe {};
  template <int IXY>
  __device__ static __forceinline__ void rpn_aux(const RPParams<T>&, const T* qL, const T* qR,
                                                 const T* auxL, const T* auxR,
                                                 T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                                 T (&amdq)[meqn], T (&apdq)[meqn], Roe&) {
    (void)auxL;
    const T u = auxR[IXY - 1];
    const T um = u < T(0) ? u : T(0);   // dmin1(u, 0)
    const T up = u > T(0) ? u : T(0);   // dmax1(u, 0)
    s[0] = u;
#pragma unroll
    for (int m = 0; m < ME; m++) {
      wave[0][m] = qR[m] - qL[m];
      amdq[m] = um * wave[0][m];
      apdq[m] = up * wave[0][m];
    }
  }
  // a1, a2, a3: aux of cell i1 in the previous, this and the next slice.
  template <int IXY>
  __device__ static __forceinline__ void rpt_aux(const RPParams<T>&, const Roe&, const T* a1,
                                                 const T* a2, const T* a3, const T (&asdq)[ME],
                                                 T (&bm)[ME], T (&bp)[ME]) {
    (void)a1;
    constexpr int kv = 2 - IXY;   // x-slices: v (aux[1]); y-slices: u (aux[0])
    const T vm = a2[kv] < T(0) ? a2[kv] : T(0);
    const T vp = a3[kv] > T(0) ? a3[kv] : T(0);
#pragma unroll
    for (int m = 0; m < ME; m++) {
      bm[m] = vm * asdq[m];
      bp[m] = vp * asdq[m];
    }
  }
};
template <typename T> using VCAdvection1 = VCAdvection<T, 1>;
template <typename T> using VCAdvection2 = VCAdvection<T, 2>;
template <typename T> using VCAdvection3 = VCAdvection<T, 3>;
template <typename T> using VCAdvection4 = VCAdvection<T, 4>;

// ---------------------------------------------------------------- variable-coefficient acoustics
// vc_acoustics_2D of the clawpack/riemann package (rpn2_vc_acoustics.f90 and
// rpt2_vc_acoustics.f90, not part of the reference tree: restated from their
// published form), as examples/acoustics_2d_variable uses it.  q = (p, u, v);
// aux[0] = density rho, aux[1] = sound speed c of each cell; impedance Z = rho*c.
// Normal solve at interface i (cells i-1 | i): the two waves move at -c(i-1)
// and +c(i) with the transmission coefficients of the impedance jump.
// Transverse split of asdq into cell i1's neighbours across the slice: the
// down-going part moves at -c of the cell below (aux1), the up-going part at +c
// of the cell above (aux3), each weighted by the impedance jump it crosses.
template <typename T>
struct VCAcoustics {
  static constexpr int meqn = 3, mwaves = 2, maux = 2;
  struct Roe {};
  template <int IXY>
  __device__ static __forceinline__ void rpn_aux(const RPParams<T>&, const T* qL, const T* qR,
                                                 const T* auxL, const T* auxR,
                                                 T (&wave)[mwaves][meqn], T (&s)[mwaves],
                                                 T (&amdq)[meqn], T (&apdq)[meqn], Roe&) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T zi = auxR[0] * auxR[1];     // impedance of cell i
    const T zim = auxL[0] * auxL[1];    // impedance of cell i-1
    const T d0 = qR[0] - qL[0], d1 = qR[mu] - qL[mu];
    const T a1 = (-d0 + zi * d1) / (zim + zi);
    const T a2 = (d0 + zim * d1) / (zim + zi);
    wave[0][0] = -(a1 * zim);
    wave[0][mu] = a1;
    wave[0][mv] = T(0);
    s[0] = -auxL[1];
    wave[1][0] = a2 * zi;
    wave[1][mu] = a2;
    wave[1][mv] = T(0);
    s[1] = auxR[1];
#pragma unroll
    for (int m = 0; m < meqn; m++) {
      amdq[m] = s[0] * wave[0][m];
      apdq[m] = s[1] * wave[1][m];
    }
  }
  // a1, a2, a3: aux of cell i1 in the previous, this and the next slice.
  template <int IXY>
  __device__ static __forceinline__ void rpt_aux(const RPParams<T>&, const Roe&, const T* a1,
                                                 const T* a2, const T* a3, const T (&asdq)[meqn],
                                                 T (&bm)[meqn], T (&bp)[meqn]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T cm = a1[1], cp = a3[1];
    const T zm = a1[0] * a1[1], zz = a2[0] * a2[1], zp = a3[0] * a3[1];
    const T b1 = (-asdq[0] + asdq[mv] * zz) / (zm + zz);   // transmitted down-going
    const T b2 = (asdq[0] + asdq[mv] * zz) / (zz + zp);    // transmitted up-going
    bm[0] = cm * b1 * zm;
    bm[mu] = T(0);
    bm[mv] = -(cm * b1);
    bp[0] = cp * b2 * zp;
    bp[mu] = T(0);
    bp[mv] = cp * b2;
  }
};

// ---------------------------------------------------------------- shallow water over bathymetry, f-waves
// shallow_bathymetry_fwave_2D (clawpack.riemann rpn2_shallow_bathymetry_fwave.f90,
// published form; pinned by the reference's sill fixture through the oracle):
// the Roe f-wave solver with the source -g h b_x folded into the flux
// difference, so a flat surface over any bathymetry stays at rest.  aux[0] = b.
//   u, v = hu/h, hv/h where h > dry_tolerance, else 0; both sides dry: no waves;
//   u^ = (sqrt(hL) uL + sqrt(hR) uR)/(sqrt(hL) + sqrt(hR)), hbar = (hL + hR)/2,
//   c^ = sqrt(g hbar); s1 = min(uL - sqrt(g hL), u^ - c^),
//   s3 = max(uR + sqrt(g hR), u^ + c^), s2 = (s1 + s3)/2;
//   d1 = (hu)R - (hu)L, d2 = (hu^2 + g h^2/2)R - (...)L + g hbar (bR - bL),
//   d3 = (huv)R - (huv)L; b1 = (s3 d1 - d2)/(s3 - s1), b3 = (d2 - s1 d1)/(s3 - s1),
//   b2 = d3 - b1 vL - b3 vR; f-waves b1 (1, s1, vL), b2 (0, 0, 1), b3 (1, s3, vR);
//   each f-wave goes to amdq (s < 0) or apdq (s > 0), half to each at s = 0.
// No transverse solver: dimension-split steps, or unsplit with transverse_waves = 0.
template <typename T>
struct ShallowFwave {
  static constexpr int meqn = 3, mwaves = 3, maux = 1;
  struct Roe {};
  template <int IXY>
  __device__ static __forceinline__ void rpn_aux(const RPParams<T>& p, const T* qL, const T* qR,
                                                 const T* auxL, const T* auxR,
                                                 T (&fw)[mwaves][meqn], T (&s)[mwaves],
                                                 T (&amdq)[meqn], T (&apdq)[meqn], Roe&) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T g = p.grav, dry = p.dry_tolerance;
#pragma unroll
    for (int w = 0; w < mwaves; w++) {
      s[w] = T(0);
#pragma unroll
      for (int m = 0; m < meqn; m++) fw[w][m] = T(0);
    }
#pragma unroll
    for (int m = 0; m < meqn; m++) { amdq[m] = T(0); apdq[m] = T(0); }
    const T hl = qL[0], hr = qR[0];
    if (hl <= dry && hr <= dry) return;
    T ul = T(0), vl = T(0), ur = T(0), vr = T(0);
    if (hl > dry) { ul = qL[mu] / hl; vl = qL[mv] / hl; }
    if (hr > dry) { ur = qR[mu] / hr; vr = qR[mv] / hr; }
    const T sl = sqrt(hl), sr = sqrt(hr);
    const T hbar = T(0.5) * (hl + hr);
    const T uhat = (sl * ul + sr * ur) / (sl + sr);
    const T chat = sqrt(g * hbar);
    s[0] = dmin(ul - sqrt(g * hl), uhat - chat);
    s[2] = dmax(ur + sqrt(g * hr), uhat + chat);
    s[1] = T(0.5) * (s[0] + s[2]);
    const T d1 = hr * ur - hl * ul;
    const T d2 = (hr * ur * ur + T(0.5) * g * hr * hr) - (hl * ul * ul + T(0.5) * g * hl * hl) +
                 g * hbar * (auxR[0] - auxL[0]);
    const T d3 = hr * ur * vr - hl * ul * vl;
    const T b1 = (s[2] * d1 - d2) / (s[2] - s[0]);
    const T b3 = (d2 - s[0] * d1) / (s[2] - s[0]);
    const T b2 = d3 - b1 * vl - b3 * vr;
    fw[0][0] = b1; fw[0][mu] = b1 * s[0]; fw[0][mv] = b1 * vl;
    fw[1][mv] = b2;
    fw[2][0] = b3; fw[2][mu] = b3 * s[2]; fw[2][mv] = b3 * vr;
#pragma unroll
    for (int w = 0; w < mwaves; w++) {
#pragma unroll
      for (int m = 0; m < meqn; m++) {
        if (s[w] < T(0)) amdq[m] = amdq[m] + fw[w][m];
        else if (s[w] > T(0)) apdq[m] = apdq[m] + fw[w][m];
        else { amdq[m] = amdq[m] + T(0.5) * fw[w][m]; apdq[m] = apdq[m] + T(0.5) * fw[w][m]; }
      }
    }
  }
  // Not reached: the entry points refuse transverse_waves > 0 for this solver.
  template <int IXY>
  __device__ static __forceinline__ void rpt_aux(const RPParams<T>&, const Roe&, const T*,
                                                 const T*, const T*, const T (&)[meqn],
                                                 T (&bm)[meqn], T (&bp)[meqn]) {
#pragma unroll
    for (int m = 0; m < meqn; m++) { bm[m] = T(0); bp[m] = T(0); }
  }
};

// The normal solve with the aux values of the left and right cells; solvers
// without aux fields take the plain call.
template <class RS, int IXY, typename T>
__device__ __forceinline__ void rpn_call(const RPParams<T>& p, const T* qL, const T* qR,
                                         const T* auxL, const T* auxR,
                                         T (&wave)[RS::mwaves][RS::meqn], T (&s)[RS::mwaves],
                                         T (&amdq)[RS::meqn], T (&apdq)[RS::meqn],
                                         typename RS::Roe& roe) {
  if constexpr (RS::maux > 0) {
    RS::template rpn_aux<IXY>(p, qL, qR, auxL, auxR, wave, s, amdq, apdq, roe);
  } else {
    (void)auxL; (void)auxR;
    RS::template rpn<IXY>(p, qL, qR, wave, s, amdq, apdq, roe);
  }
}

// The transverse split with the aux values of cell i1 in the three slices
// (previous, this, next); solvers without aux fields take the plain call.
template <class RS, int IXY, typename T>
__device__ __forceinline__ void rpt_call(const RPParams<T>& p, const typename RS::Roe& roe,
                                         const T* a1, const T* a2, const T* a3,
                                         const T (&asdq)[RS::meqn], T (&bm)[RS::meqn],
                                         T (&bp)[RS::meqn]) {
  if constexpr (RS::maux > 0) {
    RS::template rpt_aux<IXY>(p, roe, a1, a2, a3, asdq, bm, bp);
  } else {
    (void)a1; (void)a2; (void)a3;
    RS::template rpt<IXY>(p, roe, asdq, bm, bp);
  }
}

}  // namespace claw
