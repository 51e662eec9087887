// sharp2.hip -- SharpClaw 2D semi-discrete operator with fused SSP-RK stage
// combination, gfx950.
//
// Follows sharpclaw/flux2.f90:2-74 and flux1.f90:76-268 (lim_type = 2,
// tfluct_solver = False): per slice, WENO5 reconstruction -- component-wise
// (weno.f90:5-83) or characteristic-wise (reconstruct.f90:201-299) --, the
// Riemann solve at every interface, the CFL, the Riemann solve inside every
// cell between its two reconstructed edge values, and
//   dq1d(i) = -dtdx*(amdq(i+1) + apdq(i) + amdq2(i) + apdq2(i)).
// The x-slices' and y-slices' contributions are summed as flux2.f90 does
// (dq = 0 + dqx + dqy).  The x-slices cover rows 0..my+1 and the y-slices
// columns 0..mx+1; their ghost-row / ghost-column parts only enter the CFL.
//
// k_sharp_march computes a stage in one pass over HBM: each wavefront marches a
// 64-column strip down its rows (details below), and the stage combination of
// SharpClawSolver.step (sharpclaw/solver.py:312-323, 400-430) is applied as each
// cell's dq is complete.  With the learned indicator's flags, each interface
// takes WENO5's linear weights unless its (dilated) flag is set; a flagged
// interface takes the full WENO5, in the lanes that hold one.
#include "variant.hpp"  // first: selects the arithmetic variant
#include <stdlib.h>

#include "launch.hpp"
#include "sharpctl.hpp"

namespace claw {
namespace CLAW_VNS {

// The WENO epsilon.  FP64: the reference's REAL(4) literal 1.e-36 widened to
// double (reconstruct.f90:11, weno.f90:52-57).  FP32 keeps flat stencils away
// from 0/0: both forms divide by (eps + beta)^2 (the FP32 characteristic weights
// are formed as d_k / (eps + beta_k)^2, not through the reference's products of
// two such squares), so eps = 1e-18 keeps that square a normal float.  On the
// shallow-water humps and dam breaks at 256^2 it moves dq by 1e-11 of its
// largest value, against FP32's rounding of 6e-8.
template <typename T>
__device__ __forceinline__ T weno_eps() { return sizeof(T) == 8 ? T((double)1.0e-36f) : T(1.0e-18f); }
template <typename T>
__device__ __forceinline__ T weno_eps_char() {
  return sizeof(T) == 8 ? T((double)1.0e-36f) : T(1.0e-18f);
}

// ---------------------------------------------------------------- stage modes
enum StageOut {
  OUT_DQDT = 0,      // out = D / dt               (dq_dt = dq(state)/dt, solver.py:293)
  OUT_SSP33_2 = 1,   // out = 0.75*A + 0.25*(Y + D)          (solver.py:318)
  OUT_SSP33_3 = 2,   // out = 1/3*A + 2/3*(Y + D)            (solver.py:323)
  OUT_ADD6 = 3,      // out = Y + D/6.                       (solver.py:415, 425)
  OUT_SSP104_FIN = 4,// out = A + (0.6*Y + 0.1*D)            (solver.py:430)
  OUT_F32_S2 = 5,    // FP32: out = A + 0.25*(e1 + D), aux = (e1 + D)/6, e1 = s*B
  OUT_F32_S3 = 6,    // FP32: Kahan(A, aux + 2/3*D, comp)
  OUT_D = 7,         // out = D (the operator itself, fmod.flux2's dq)
  OUT_SSP22 = 8      // out = 0.5*((A + Y) + D)              (solver.py:393)
};
enum StageY {
  Y_ARRAY = 0,       // Y = P
  Y_AXPY = 1,        // Y = P + s*Q          (r = q + dt*dq_dt, solver.py:313)
  Y_AXPY6 = 2        // Y = P + (s*Q)/6.     (s1 = q + dt*dq_dt/6., solver.py:409)
};

template <typename T>
struct SharpArgs {
  RPParams<T> rp;
  Geom g;               // A, B, aux, comp and the outputs: compact
  Geom gin;             // P and Q (the stage input); may store ghost cells (BC_GIVEN)
  BCSpec bc;
  int ng;
  int char_decomp;
  int ymode, omode;
  T dt;                 // dt passed to the operator (dq = dt*L)
  T s;                  // scalar of the Y / e1 axpy
  T dtdx, dtdy;
  T rdt;                // T(1/dt), dt in FP64 (the FP32 stages' dq_dt = D * rdt)
  const T* P;           // Y source
  const T* Q;           // axpy partner of P
  const T* A;           // stage anchor (q at the start of the step)
  const T* B;           // OUT_F32_S2: dq_dt
  const T* aux_in;      // OUT_F32_S3: accumulated increment
  const T* comp_in;     // OUT_F32_S3: Kahan compensation
  T* out;
  T* aux_out;
  T* comp_out;
  const T* src;         // optional source increment added to D (SharpClaw dq_src)
  unsigned long long* cfl_bits;
  const unsigned char* flags;   // troubled-interface flags (learned indicator) or null
  const int* flag_list;         // with flags: ids of the interfaces whose dilated flag is set
  const int* flag_count;        //   and their number (device)
  T* recon;                     //   their full-WENO5 edge values, 2*meqn per interface id
  // device-controlled stage (ctl non-null): it runs while *gate != 0, takes dt
  // from ctl and the pointers named in `res` from the state's two buffers by
  // ctl->cur; its last wavefront applies phase ctl_phase; with count_bad it
  // counts the stored cells whose fields in pos_mask are not > 0 (or not finite)
  SharpCtl* ctl;
  const int* gate;
  int ctl_phase, count_bad;
  unsigned res, pos_mask;
  void* qb[2];
  void* cb[2];
  double dx, dy;
  const T* qaux;                // the solver's aux fields (solvers with maux > 0), geometry
  Geom gaux;                    //   gaux, boundary codes bcaux (never negated)
  BCSpec bcaux;
};


// ---------------------------------------------------------------- WENO5
// weno.f90:32-79 for one cell: ql = left edge (fs0), qr = right edge (fs1).
// `v` holds cells i-2..i+2.
template <typename T>
__device__ __forceinline__ void weno5_cell(const T (&v)[5], T& fs0, T& fs1) {
  const T qm2 = v[0], qm1 = v[1], q0 = v[2], qp1 = v[3], qp2 = v[4];
  const T eps = weno_eps<T>();
  T sigma0 = T(3.3333333333333333333333333333333333) * (q0 * q0) -
             T(10.333333333333333333333333333333333) * q0 * qp1 +
             T(3.6666666666666666666666666666666667) * q0 * qp2 +
             T(8.3333333333333333333333333333333333) * (qp1 * qp1) -
             T(6.3333333333333333333333333333333333) * qp1 * qp2 +
             T(1.3333333333333333333333333333333333) * (qp2 * qp2);
  T sigma1 = T(4.3333333333333333333333333333333333) * (q0 * q0) -
             T(4.3333333333333333333333333333333333) * q0 * qp1 -
             T(4.3333333333333333333333333333333333) * q0 * qm1 +
             T(1.3333333333333333333333333333333333) * (qp1 * qp1) +
             T(1.6666666666666666666666666666666667) * qp1 * qm1 +
             T(1.3333333333333333333333333333333333) * (qm1 * qm1);
  T sigma2 = T(3.3333333333333333333333333333333333) * (q0 * q0) -
             T(10.333333333333333333333333333333333) * q0 * qm1 +
             T(3.6666666666666666666666666666666667) * q0 * qm2 +
             T(8.3333333333333333333333333333333333) * (qm1 * qm1) -
             T(6.3333333333333333333333333333333333) * qm1 * qm2 +
             T(1.3333333333333333333333333333333333) * (qm2 * qm2);
  T d0 = eps + sigma0, d1 = eps + sigma1, d2 = eps + sigma2;
  T omega0 = T(0.1) / (d0 * d0);
  T omega1 = T(0.6) / (d1 * d1);
  T omega2 = T(0.3) / (d2 * d2);
  T omega3 = T(0.3) / (d0 * d0);
  T omega4 = T(0.6) / (d1 * d1);
  T omega5 = T(0.1) / (d2 * d2);
  T fr0 = T(1.8333333333333333333333333333333333) * q0 -
          T(1.1666666666666666666666666666666667) * qp1 +
          T(0.33333333333333333333333333333333333) * qp2;
  T fr1 = T(0.83333333333333333333333333333333333) * q0 -
          T(0.16666666666666666666666666666666667) * qp1 +
          T(0.33333333333333333333333333333333333) * qm1;
  T fr2 = T(0.33333333333333333333333333333333333) * q0 +
          T(0.83333333333333333333333333333333333) * qm1 -
          T(0.16666666666666666666666666666666667) * qm2;
  T fr3 = T(0.33333333333333333333333333333333333) * q0 +
          T(0.83333333333333333333333333333333333) * qp1 -
          T(0.16666666666666666666666666666666667) * qp2;
  T fr4 = T(0.83333333333333333333333333333333333) * q0 +
          T(0.33333333333333333333333333333333333) * qp1 -
          T(0.16666666666666666666666666666666667) * qm1;
  T fr5 = T(1.8333333333333333333333333333333333) * q0 -
          T(1.1666666666666666666666666666666667) * qm1 +
          T(0.33333333333333333333333333333333333) * qm2;
  fs0 = (fr0 * omega0 + fr1 * omega1 + fr2 * omega2) / (omega0 + omega1 + omega2);
  fs1 = (fr3 * omega3 + fr4 * omega4 + fr5 * omega5) / (omega3 + omega4 + omega5);
}

// reconstruct.f90:264-284 for one side: hh(-2..2) -> uu.
template <typename T>
__device__ __forceinline__ T weno_char_side(T hm2, T hm1, T h0, T hp1, T hp2, int im) {
  // im = +1: (intwo, inone, ione) = (-2, -1, +1); im = -1: (2, 1, -1)
  T a2 = im == 1 ? hm2 : hp2;   // hh(intwo)
  T a1 = im == 1 ? hm1 : hp1;   // hh(inone)
  T b1 = im == 1 ? hp1 : hm1;   // hh(ione)
  const T eps = weno_eps_char<T>();
  T fim = T(im);
  T t1 = fim * (a2 - a1);
  T t2 = fim * (a1 - h0);
  T t3 = fim * (h0 - b1);
  T e1 = a2 - T(3) * a1;
  T e2 = a1 + h0;
  T e3 = T(3) * h0 - b1;
  T tt1 = T(13) * (t1 * t1) + T(3) * (e1 * e1);
  T tt2 = T(13) * (t2 * t2) + T(3) * (e2 * e2);
  T tt3 = T(13) * (t3 * t3) + T(3) * (e3 * e3);
  tt1 = (eps + tt1) * (eps + tt1);
  tt2 = (eps + tt2) * (eps + tt2);
  tt3 = (eps + tt3) * (eps + tt3);
  T s1, s3;
  if constexpr (sizeof(T) == 4) {
    // the same weights as d_k / (eps + beta_k)^2: no fourth powers to underflow
    const T r1 = T(1) / tt1, r2 = T(6) / tt2, r3 = T(3) / tt3;
    const T t0 = T(1) / (r1 + r2 + r3);
    s1 = r1 * t0;
    s3 = r3 * t0;
  } else {   // the reference's products of two squares, bit for bit
    s1 = tt2 * tt3;
    T s2 = T(6) * tt1 * tt3;
    s3 = T(3) * tt1 * tt2;
    T t0 = T(1) / (s1 + s2 + s3);
    s1 = s1 * t0;
    s3 = s3 * t0;
  }
  return (s1 * (t2 - t1) + (T(0.5) * s3 - T(0.25)) * (t3 - t2)) / T(3);
}

// Eigenvectors at the Roe average of (qL, qR) in direction IXY (the same
// formulas as the oracle's evec).  evl[ip][m] = row ip of R^-1, evr[m][ip].
template <class RS, int IXY, typename T>
struct Evec;

template <int IXY, typename T>
struct Evec<Shallow<T>, IXY, T> {
  __device__ static __forceinline__ void get(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&evl)[3][3], T (&evr)[3][3]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    T h = (qL[0] + qR[0]) * T(0.50);
    T hsqrtl = sqrt(qL[0]), hsqrtr = sqrt(qR[0]);
    T hsq2 = hsqrtl + hsqrtr;
    T u = (qL[mu] / hsqrtl + qR[mu] / hsqrtr) / hsq2;
    T v = (qL[mv] / hsqrtl + qR[mv] / hsqrtr) / hsq2;
    T c = sqrt(p.grav * h);
    evr[0][0] = T(1); evr[mu][0] = u - c; evr[mv][0] = v;
    evr[0][1] = T(0); evr[mu][1] = T(0); evr[mv][1] = T(1);
    evr[0][2] = T(1); evr[mu][2] = u + c; evr[mv][2] = v;
    T r2c = T(0.5) / c;
    evl[0][0] = (u + c) * r2c; evl[0][mu] = -r2c; evl[0][mv] = T(0);
    evl[1][0] = -v; evl[1][mu] = T(0); evl[1][mv] = T(1);
    evl[2][0] = -(u - c) * r2c; evl[2][mu] = r2c; evl[2][mv] = T(0);
  }
};

// The shallow-water eigenvectors do not depend on the Riemann solver.
template <int IXY, typename T>
struct Evec<ShallowHLLE<T>, IXY, T> : Evec<Shallow<T>, IXY, T> {};

template <int IXY, typename T>
struct Evec<Euler<T>, IXY, T> {
  __device__ static __forceinline__ void get(const RPParams<T>& p, const T* qL, const T* qR,
                                             T (&evl)[4][4], T (&evr)[4][4]) {
    constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
    const T gamma1 = p.gamma1;
    auto r = Euler<T>::template roe_avg<IXY>(gamma1, qL, qR);
    T u = r.u, v = r.v, c = r.a, H = r.enth;
    T q2 = T(0.5) * r.u2v2;
    evr[0][0] = T(1); evr[mu][0] = u - c; evr[mv][0] = v; evr[3][0] = H - u * c;
    evr[0][1] = T(0); evr[mu][1] = T(0); evr[mv][1] = T(1); evr[3][1] = v;
    evr[0][2] = T(1); evr[mu][2] = u; evr[mv][2] = v; evr[3][2] = q2;
    evr[0][3] = T(1); evr[mu][3] = u + c; evr[mv][3] = v; evr[3][3] = H + u * c;
    T b1 = gamma1 / (c * c), b2 = q2 * b1;
    T r2c = T(0.5) / c;
    evl[0][0] = T(0.5) * b2 + u * r2c; evl[0][mu] = -(T(0.5) * b1 * u + r2c);
    evl[0][mv] = -T(0.5) * b1 * v; evl[0][3] = T(0.5) * b1;
    evl[1][0] = -v; evl[1][mu] = T(0); evl[1][mv] = T(1); evl[1][3] = T(0);
    evl[2][0] = T(1) - b2; evl[2][mu] = b1 * u; evl[2][mv] = b1 * v; evl[2][3] = -b1;
    evl[3][0] = T(0.5) * b2 - u * r2c; evl[3][mu] = -(T(0.5) * b1 * u - r2c);
    evl[3][mv] = -T(0.5) * b1 * v; evl[3][3] = T(0.5) * b1;
  }
};

// WENO5 with its linear weights (0.1, 0.6, 0.3): the 5th-order upwind-biased
// reconstructions of both edges, component-wise (with linear weights the
// characteristic projection cancels).  st[0..5] = cells c-3..c+2.
template <int ME, typename T>
__device__ __forceinline__ void recon_linear(const T (&st)[6][ME], T (&qrL)[ME], T (&qlR)[ME]) {
#pragma unroll
  for (int m = 0; m < ME; m++) {
    qrL[m] = (T(2) * st[0][m] - T(13) * st[1][m] + T(47) * st[2][m] + T(27) * st[3][m] -
              T(3) * st[4][m]) * T(1. / 60.);
    qlR[m] = (-T(3) * st[1][m] + T(27) * st[2][m] + T(47) * st[3][m] - T(13) * st[4][m] +
              T(2) * st[5][m]) * T(1. / 60.);
  }
}

// Reconstruct both sides of the interface between stencil cells 2 and 3 of
// st[0..5] (cells c-3..c+2): qrL = right edge of cell c-1, qlR = left edge of c.
template <class RS, int IXY, typename T, bool CHAR>
__device__ __forceinline__ void recon_interface(const RPParams<T>& p,
                                                const T (&st)[6][RS::meqn],
                                                T (&qrL)[RS::meqn], T (&qlR)[RS::meqn]) {
  constexpr int ME = RS::meqn;
  if constexpr (!CHAR) {
#pragma unroll
    for (int m = 0; m < ME; m++) {
      T a[5] = {st[0][m], st[1][m], st[2][m], st[3][m], st[4][m]};
      T b[5] = {st[1][m], st[2][m], st[3][m], st[4][m], st[5][m]};
      T f0, f1;
      weno5_cell(a, f0, f1);
      qrL[m] = f1;
      weno5_cell(b, f0, f1);
      qlR[m] = f0;
    }
  } else {
    T evl[ME][ME], evr[ME][ME];
    Evec<RS, IXY, T>::get(p, st[2], st[3], evl, evr);
    T dq[5][ME];   // dq(m, c+m2) = q(c+m2) - q(c+m2-1), m2 = -2..2
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int m = 0; m < ME; m++) dq[k][m] = st[k + 1][m] - st[k][m];
#pragma unroll
    for (int m = 0; m < ME; m++) {
      T lin = (-st[1][m] + T(7) * (st[2][m] + st[3][m]) - st[4][m]) / T(12);
      qrL[m] = lin;
      qlR[m] = lin;
    }
#pragma unroll
    for (int ip = 0; ip < ME; ip++) {
      T hh[5];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        hh[k] = T(0);
#pragma unroll
        for (int m = 0; m < ME; m++) hh[k] = hh[k] + evl[ip][m] * dq[k][m];
      }
      T uu1 = weno_char_side(hh[0], hh[1], hh[2], hh[3], hh[4], 1);
      T uu2 = weno_char_side(hh[0], hh[1], hh[2], hh[3], hh[4], -1);
#pragma unroll
      for (int m = 0; m < ME; m++) {
        qrL[m] = qrL[m] + evr[m][ip] * uu1;
        qlR[m] = qlR[m] + evr[m][ip] * uu2;
      }
    }
  }
}

// Pointers a controlled stage takes from the state's buffers (SharpArgs::res):
// q[cur] as P, q[cur] as A, q[1-cur] as out, comp[cur] as comp_in,
// comp[1-cur] as comp_out.
enum { RES_P = 1, RES_A = 2, RES_OUT = 4, RES_CIN = 8, RES_COUT = 16 };

// What a stage takes from its SharpCtl when it is device-controlled: dt (and s,
// dtdx, dtdy as the eager launch forms them on the host: T(dt / dx)) and the
// pointers named in SharpArgs::res.  A few uniform values, kept apart from the
// (large, constant) argument block.
template <typename T>
struct StageView {
  const T* P;
  const T* A;
  T* out;
  const T* comp_in;
  T* comp_out;
  T dt, s, dtdx, dtdy;
  T rdt;   // FP32: T(1/dt) (dq_dt = D * rdt; FP64 divides, as the reference)
};

template <typename T>
__device__ __forceinline__ StageView<T> stage_view(const SharpArgs<T>& a) {
  StageView<T> v{a.P, a.A, a.out, a.comp_in, a.comp_out, a.dt, a.s, a.dtdx, a.dtdy, T(0)};
  v.rdt = a.rdt;
  if (a.ctl) {
    const int cur = a.ctl->cur;
    const double dt = a.ctl->dt;
    v.dt = T(dt);
    if (sizeof(T) == 4) v.rdt = T(1.0 / dt);
    v.s = T(dt);
    v.dtdx = T(dt / a.dx);
    v.dtdy = T(dt / a.dy);
    if (a.res & RES_P) v.P = (const T*)a.qb[cur];
    if (a.res & RES_A) v.A = (const T*)a.qb[cur];
    if (a.res & RES_OUT) v.out = (T*)a.qb[1 - cur];
    if (a.res & RES_CIN) v.comp_in = (const T*)a.cb[cur];
    if (a.res & RES_COUT) v.comp_out = (T*)a.cb[1 - cur];
  }
  return v;
}

// The inputs a cell's stage combination reads besides D and Y, loaded at the top
// of the march iteration that stores the cell (so that their latency overlaps
// the row's reconstruction and Riemann solves): pa = A; pb = B (OUT_F32_S2) or
// aux_in (OUT_F32_S3); pc = comp_in (OUT_F32_S3).
template <int ME, typename T>
struct StoreIn {
  T pa[ME], pb[ME], pc[ME];
};

template <int ME, typename T>
__device__ __forceinline__ void load_store_in(const SharpArgs<T>& a, const StageView<T>& v,
                                              long long off, StoreIn<ME, T>& si) {
  const long long fs = a.g.fs;
  const int om = a.omode;
  // uniform choices; each array is read whole or not at all
  const T* pa = (om == OUT_DQDT || om == OUT_ADD6 || om == OUT_D) ? nullptr : v.A;
  const T* pb = om == OUT_F32_S2 ? a.B : om == OUT_F32_S3 ? a.aux_in : nullptr;
  const T* pc = om == OUT_F32_S3 ? v.comp_in : nullptr;
#pragma unroll
  for (int m = 0; m < ME; m++) {
    si.pa[m] = pa ? pa[off + m * fs] : T(0);
    si.pb[m] = pb ? pb[off + m * fs] : T(0);
    si.pc[m] = pc ? pc[off + m * fs] : T(0);
  }
}

// Stage combination of cell offset `off`, field m (SharpClawSolver.step,
// sharpclaw/solver.py:312-323, 400-430; the FP32 stages keep a compensated sum).
template <int ME, typename T>
__device__ __forceinline__ T stage_store(const SharpArgs<T>& a, const StageView<T>& v,
                                         const StoreIn<ME, T>& si, long long off, int m, T D,
                                         T yv) {
  const long long fs = a.g.fs;
  // dq = dq_hyperbolic + dq_src (sharpclaw/solver.py:537-540)
  if (a.src) D = D + a.src[off + m * fs];
  const int om = a.omode;
  T o;
  if (om == OUT_DQDT) {
    if constexpr (sizeof(T) == 4) o = D * v.rdt;
    else o = D / v.dt;
  } else if (om == OUT_SSP33_2) {
    o = T(0.75) * si.pa[m] + T(0.25) * (yv + D);
  } else if (om == OUT_SSP33_3) {
    o = T(1. / 3.) * si.pa[m] + T(2. / 3.) * (yv + D);
  } else if (om == OUT_ADD6) {
    o = yv + D / T(6);
  } else if (om == OUT_SSP104_FIN) {
    o = si.pa[m] + (T(0.6) * yv + T(0.1) * D);
  } else if (om == OUT_D) {
    o = D;
  } else if (om == OUT_SSP22) {
    o = T(0.5) * ((si.pa[m] + yv) + D);
  } else if (om == OUT_F32_S2) {
    T e1 = v.s * si.pb[m];
    T inc = e1 + D;
    o = si.pa[m] + T(0.25) * inc;
    a.aux_out[off + m * fs] = inc * T(1. / 6.);
  } else {   // OUT_F32_S3
    T delta = si.pb[m] + T(2. / 3.) * D;
    T qa = si.pa[m];
    T yk = delta - si.pc[m];
    T tq = qa + yk;
    v.comp_out[off + m * fs] = (tq - qa) - yk;
    o = tq;
  }
  v.out[off + m * fs] = o;
  return o;
}

// The stage input Y of cell (c, r): P, P + s*Q or P + (s*Q)/6 (y_mode), with the
// boundary map applied to both.
template <int ME, typename T>
__device__ __forceinline__ void stage_input(const SharpArgs<T>& a, const StageView<T>& v, int c,
                                            int r, T (&y)[ME]) {
  load_cell2<ME>(v.P, a.gin, a.bc, a.ng, c, r, y);
  if (a.ymode != Y_ARRAY) {
    T qv[ME];
    load_cell2<ME>(a.Q, a.gin, a.bc, a.ng, c, r, qv);
#pragma unroll
    for (int m = 0; m < ME; m++)
      y[m] = a.ymode == Y_AXPY ? y[m] + v.s * qv[m] : y[m] + (v.s * qv[m]) / T(6);
  }
}

// Interface ids (the layout of the indicator's flags): x-interface c (left face
// of cell c, c = 1..mx+1) of row r (1..my) is (r-1)*(mx+1) + c-1; y-interface t
// (lower face of row t, t = 1..my+1) of column c (1..mx) follows them at
// (mx+1)*my + (t-1)*mx + c-1.
__device__ __forceinline__ long long xface_id(int mx, int c, int r) {
  return (long long)(r - 1) * (mx + 1) + (c - 1);
}
__device__ __forceinline__ long long yface_id(int mx, int my, int c, int t) {
  return (long long)(mx + 1) * my + (long long)(t - 1) * mx + (c - 1);
}

// An interface is listed (reconstructed with the full WENO5) when its flag or
// the flag of a neighbour along its normal is set: x-interface c of row r when
// one of x-interfaces c-1, c, c+1 of that row is (clamped to 1..mx+1), y-interface
// t of column c when one of y-interfaces t-1, t, t+1 of that column is (clamped
// to 1..my+1).  k_flag_list builds the list, the flagged march tests the same
// rule on flag bytes it loads a row ahead.

template <int ME, typename T>
__device__ __forceinline__ void load_recon(const T* p, T (&qrL)[ME], T (&qlR)[ME]) {
#pragma unroll
  for (int m = 0; m < ME; m++) {
    qrL[m] = p[m];
    qlR[m] = p[ME + m];
  }
}

// ---------------------------------------------------------------- flagged interfaces
// The ids of the listed interfaces (the dilation rule above), appended
// to `list`; the order is free, each entry is computed on its own.  A thread
// takes 4 consecutive ids (one aligned dword of flags, its neighbours as
// byte-aligned dwords: the x-neighbours at +-1, the y-neighbours at +-mx), a
// block of 1024 threads FLAG_LIST_GROUPS such groups of 4096 ids; the hits are
// scanned over the block group by group, so a block's ids land in the list in
// increasing order (the full-WENO5 pass over the list then reads neighbouring
// stencils together), and the block takes its range with one atomic.  Same-
// address atomics serialise, so a block covers 16384 ids: about 500 atomics at
// 2048^2.  *count must be 0 on entry (k_zero_flag_count, or the flags kernel of
// a device-controlled run).
__global__ void k_zero_flag_count(int* count, const int* gate) {
  if (gate && !*gate) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0;
}

constexpr int FLAG_LIST_THREADS = 1024;
constexpr int FLAG_LIST_GROUPS = 4;

// dword i of the flags (bytes past the end read as 0; only the last partial
// dword is assembled bytewise)
__device__ __forceinline__ unsigned flag_dword(const unsigned char* f, int i, int nfull, int n) {
  if (i < nfull) return reinterpret_cast<const unsigned*>(f)[i];
  unsigned v = 0u;
  for (int k = 0; k < 4; k++)
    if (4 * i + k < n) v |= (unsigned)f[4 * i + k] << (8 * k);
  return v;
}

// the 4 flag bytes from byte offset o >= 0
__device__ __forceinline__ unsigned flag_bytes4(const unsigned char* f, int o, int nfull, int n) {
  const int i = o >> 2;
  return __builtin_amdgcn_alignbyte(flag_dword(f, i + 1, nfull, n), flag_dword(f, i, nfull, n),
                                    (unsigned)(o & 3));
}

// bit k: interface 4q+k is listed (its flag or a neighbour's along its normal)
__device__ __forceinline__ unsigned dilated_mask4(const unsigned char* f, int mx, int my, int q) {
  const int nx = (mx + 1) * my, n = nx + mx * (my + 1), nfull = n >> 2;
  const int p0 = 4 * q;
  if (p0 >= n) return 0u;
  const unsigned c = flag_dword(f, q, nfull, n);
  unsigned m = 0u;
  if (p0 < nx) {   // x-interfaces: neighbours in the row, clamped at its ends
    const unsigned left = q > 0 ? flag_bytes4(f, p0 - 1, nfull, n) : (c << 8);
    const unsigned right = flag_bytes4(f, p0 + 1, nfull, n);
    int col = p0 % (mx + 1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const unsigned b = (c >> (8 * k)) & 0xffu;
      const unsigned a = col > 0 ? (left >> (8 * k)) & 0xffu : b;
      const unsigned d = col < mx ? (right >> (8 * k)) & 0xffu : b;
      if (p0 + k < nx) m |= (unsigned)((a | b | d) != 0u) << k;
      col = col == mx ? 0 : col + 1;
    }
  }
  if (p0 + 3 >= nx) {   // y-interfaces: neighbours one row of mx away, clamped
    const unsigned up = flag_bytes4(f, max(p0 - mx, 0), nfull, n);
    const unsigned dn = flag_bytes4(f, p0 + mx, nfull, n);
    int t = p0 >= nx ? (p0 - nx) / mx : -1;
    int col = p0 >= nx ? (p0 - nx) - t * mx : 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int p = p0 + k;
      if (p >= nx && p < n) {
        if (p == nx) { t = 0; col = 0; }
        const unsigned b = (c >> (8 * k)) & 0xffu;
        const unsigned a = t > 0 ? (up >> (8 * k)) & 0xffu : b;
        const unsigned d = t < my ? (dn >> (8 * k)) & 0xffu : b;
        m |= (unsigned)((a | b | d) != 0u) << k;
        if (++col == mx) { col = 0; t++; }
      }
    }
  }
  return m;
}

__global__ void __launch_bounds__(FLAG_LIST_THREADS)
    k_flag_list(const unsigned char* f, int mx, int my, int* list, int* count, const int* gate) {
  constexpr int U = FLAG_LIST_GROUPS, NW = FLAG_LIST_THREADS / 64;
  if (gate && !*gate) return;
  __shared__ int wave_base[U][NW];
  __shared__ int block_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned m[U];
  int incl[U], cnt[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    m[u] = dilated_mask4(f, mx, my, (blockIdx.x * U + u) * FLAG_LIST_THREADS + threadIdx.x);
    cnt[u] = __popc(m[u]);
    incl[u] = cnt[u];
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int v = __shfl_up(incl[u], d, 64);
      if (lane >= d) incl[u] += v;
    }
  }
  if (lane == 63) {
#pragma unroll
    for (int u = 0; u < U; u++) wave_base[u][wave] = incl[u];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int u = 0; u < U; u++)
      for (int w = 0; w < NW; w++) {
        const int v = wave_base[u][w];
        wave_base[u][w] = tot;
        tot += v;
      }
    block_base = tot > 0 ? atomicAdd(count, tot) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; u++) {
    int pos = block_base + wave_base[u][wave] + incl[u] - cnt[u];
    const int p0 = 4 * ((blockIdx.x * U + u) * FLAG_LIST_THREADS + threadIdx.x);
    unsigned mm = m[u];
    while (mm) {
      const int k = __ffs(mm) - 1;
      mm &= mm - 1u;
      list[pos++] = p0 + k;
    }
  }
}

// The full WENO5 (characteristic or component, as configured) of every listed
// interface of the stage input, stored at its id: the right edge of the cell
// before it (qrL), then the left edge of the cell after it (qlR) -- the values
// recon_interface gives k_sharp_march for the same stencil.  One thread per
// listed interface, so the cost follows the flagged fraction; the list length
// is read on the device.
template <class RS, typename T, bool CHAR>
__global__ void __launch_bounds__(256) k_weno_flagged(const SharpArgs<T> a) {
  constexpr int ME = RS::meqn;
  // the gate, the controller's state, the list length and this thread's first
  // list entry are independent loads: issued together, they cost one memory
  // round trip instead of four (the list holds (mx+1)*my + mx*(my+1) entries,
  // so the first entry is readable before the length says it is valid)
  const int open = a.ctl ? *a.gate : 1;
  const StageView<T> v = stage_view(a);
  const int n = *a.flag_count;
  const int mx = a.g.mx, my = a.g.my;
  const long long nx = (long long)(mx + 1) * my;
  const long long nint = nx + (long long)mx * (my + 1);
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int first = i0 < nint ? a.flag_list[i0] : 0;
  if (!open) return;
  for (int i = i0; i < n; i += gridDim.x * blockDim.x) {
    const long long id = i == i0 ? first : a.flag_list[i];
    T st[6][ME], qrL[ME], qlR[ME];
    if (id < nx) {
      const int c = 1 + (int)(id % (mx + 1)), r = 1 + (int)(id / (mx + 1));
#pragma unroll
      for (int k = 0; k < 6; k++) stage_input<ME>(a, v, c - 3 + k, r, st[k]);
      recon_interface<RS, 1, T, CHAR>(a.rp, st, qrL, qlR);
    } else {
      const long long j = id - nx;
      const int c = 1 + (int)(j % mx), t = 1 + (int)(j / mx);
#pragma unroll
      for (int k = 0; k < 6; k++) stage_input<ME>(a, v, c, t - 3 + k, st[k]);
      recon_interface<RS, 2, T, CHAR>(a.rp, st, qrL, qlR);
    }
    T* out = a.recon + id * (2 * ME);
#pragma unroll
    for (int m = 0; m < ME; m++) {
      out[m] = qrL[m];
      out[ME + m] = qlR[m];
    }
  }
}

int launch_flag_list(const unsigned char* flags, int mx, int my, int* list, int* count,
                     hipStream_t s, const int* gate, bool zeroed) {
  const long long n = (long long)(mx + 1) * my + (long long)mx * (my + 1);
  if (n > 0x7fffffffLL - 4LL * FLAG_LIST_THREADS * FLAG_LIST_GROUPS) return CLAWHIP_E_ARG;
  if ((reinterpret_cast<unsigned long long>(flags) & 3ull) != 0ull) return CLAWHIP_E_ARG;
  if (!zeroed) hipLaunchKernelGGL(k_zero_flag_count, dim3(1), dim3(64), 0, s, count, gate);
  const long long per_block = 4LL * FLAG_LIST_THREADS * FLAG_LIST_GROUPS;
  hipLaunchKernelGGL(k_flag_list, dim3((unsigned)((n + per_block - 1) / per_block)),
                     dim3(FLAG_LIST_THREADS), 0, s, flags, mx, my, list, count, gate);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

// ---------------------------------------------------------------- march form
// The stage kernel: a wavefront owns a strip of 64 columns c0-3 .. c0+60
// (lanes 3..60 write, the others supply stencil neighbours) and marches down H
// rows, with no LDS and no barriers.  Per row, the x-interfaces are
// reconstructed lane-parallel (the 6-cell stencil arrives by cross-lane moves)
// and each lane's y-interface is reconstructed from a register window of the
// 6 rows around it; the interface results of the previous row are carried, so
// every interface is solved once.  An LDS-tiled form (a 32x8 tile with a 3-cell
// halo staged in LDS, interfaces then cells in barrier-separated passes) was
// measured at 0.91 ms (FP64) / 0.80 ms (FP32) per 2048^2 characteristic stage
// against 0.48 / 0.32 ms for this march: its partial second passes and barriers
// left the SIMDs idle.
template <typename T>
struct MarchArgs {
  SharpArgs<T> s;
  int H, nstrips, nrb;
};

// FP32 (fp32k) path: the fluctuation inside a cell, amdq + apdq of the Riemann
// problem between its two reconstructed edges (flux1.f90:174-186), is the flux
// difference f(qr) - f(ql), which the Roe solvers reproduce exactly in exact
// arithmetic (their linearisation satisfies A_hat (qr - ql) = f(qr) - f(ql), and
// the entropy fixes only move fluctuation between amdq and apdq).  FP64 keeps the
// reference's Riemann solve, bit for bit.
//
// Solvers with aux fields (flux1.f90:131, 174-181): the solve at an interface
// takes the aux of its two cells, the solve inside a cell that cell's aux on
// both sides; they keep that solve in FP32 too (no flux-difference shortcut).
template <class RS, typename T, bool CHAR, bool FLAGS>
__global__ void __launch_bounds__(256) k_sharp_march(const MarchArgs<T> ma) {
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  constexpr bool AUX = RS::maux > 0;
  constexpr int MA = AUX ? RS::maux : 1;
  constexpr bool TFLUCT = sizeof(T) == 4 && !AUX;
  const SharpArgs<T>& a = ma.s;
  if (a.ctl && !*a.gate) return;   // uniform over the launch
  const StageView<T> v = stage_view(a);
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  // wavefronts past the last strip skip the march but still reach the
  // workgroup's Courant-number reduction (a barrier)
  const bool live = wid < ma.nstrips * ma.nrb;   // uniform per wavefront
  T smax_x = T(0), smax_y = T(0);
  int nbad = 0;   // stored cells that lost positivity (count_bad)
  if (live) {
    const int strip = wid % ma.nstrips, rb = wid / ma.nstrips;
    const int mx = a.g.mx, my = a.g.my;
    const int c0 = 1 + strip * 58;
    const int c = c0 - 3 + lane;                 // this lane's column
    const int r0 = 1 + rb * ma.H;
    const int rend = min(r0 + ma.H - 1, my);     // last output row
    const bool out_lane = lane >= 3 && lane <= 60 && c >= 1 && c <= mx;
    const bool xint_lane = lane >= 3 && lane <= 61 && c >= 1 && c <= mx + 1;
    const bool ycfl_lane = c >= 0 && c <= mx + 1;
    const RPParams<T>& rp = a.rp;

    auto load_y = [&](int r, T (&y)[ME]) { stage_input<ME>(a, v, c, r, y); };
    // a window row in two parts: its loads (issued one iteration ahead) and the
    // stage-input combination stage_input applies (the same expression)
    auto load_raw = [&](int r, T (&p)[ME], T (&q)[ME]) {
      load_cell2<ME>(v.P, a.gin, a.bc, a.ng, c, r, p);
      if (a.ymode != Y_ARRAY) load_cell2<ME>(a.Q, a.gin, a.bc, a.ng, c, r, q);
    };
    auto combine = [&](const T (&p)[ME], const T (&q)[ME], T (&y)[ME]) {
  #pragma unroll
      for (int m = 0; m < ME; m++) {
        y[m] = p[m];
        if (a.ymode != Y_ARRAY)
          y[m] = a.ymode == Y_AXPY ? y[m] + v.s * q[m] : y[m] + (v.s * q[m]) / T(6);
      }
    };
    auto load_a = [&](int r, T (&A)[MA]) {
      if constexpr (AUX) load_cell2<MA>(a.qaux, a.gaux, a.bcaux, a.ng, c, r, A);
      else A[0] = T(0);
    };
    // x-slice of row r (aux Ar of this lane's cell): interface c (left of this
    // lane's cell) -> qlR, qrL, am, ap
    auto x_interface = [&](const T (&row)[ME], const T (&Ar)[MA], int r, bool flagged,
                           T (&qrL)[ME], T (&qlR)[ME], T (&am)[ME], T (&ap)[ME]) {
      // st[k] = the row value of lane l + k - 3 (chained one-lane moves)
      T st[6][ME];
  #pragma unroll
      for (int m = 0; m < ME; m++) st[3][m] = row[m];
      from_left<ME>(st[3], st[2]);
      from_left<ME>(st[2], st[1]);
      from_left<ME>(st[1], st[0]);
      from_right<ME>(st[3], st[4]);
      from_right<ME>(st[4], st[5]);
      // flagged path: the linear weights, or, where the interface's dilated flag
      // is set (`flagged`, from the flag bytes loaded a row ahead), the full WENO5
      // edge values k_weno_flagged stored (interfaces of ghost rows enter only the
      // CFL and take the linear weights)
      if (FLAGS) {
        recon_linear<ME>(st, qrL, qlR);
        if (flagged) load_recon<ME>(a.recon + xface_id(mx, c, r) * (2 * ME), qrL, qlR);
      } else {
        recon_interface<RS, 1, T, CHAR>(rp, st, qrL, qlR);
      }
      T W[MW][ME], sp[MW];
      typename RS::Roe roe;
      T AL[MA];
      if constexpr (AUX) from_left<MA>(Ar, AL);
      rpn_call<RS, 1>(rp, qrL, qlR, AL, Ar, W, sp, am, ap, roe);
      if (xint_lane) {
  #pragma unroll
        for (int w = 0; w < MW; w++) smax_x = cfl_term(smax_x, (T)fabs(sp[w]));
      }
    };
    // y-interface t (rows t-1 | t, aux Alo | Ahi) of this lane's column from the
    // window win[0..5] = rows t-3..t+2
    auto y_interface = [&](const T (&win)[6][ME], const T (&Alo)[MA], const T (&Ahi)[MA], int t,
                           bool flagged, T (&qrL)[ME], T (&qlR)[ME], T (&am)[ME], T (&ap)[ME]) {
      if (FLAGS) {
        recon_linear<ME>(win, qrL, qlR);
        if (flagged) load_recon<ME>(a.recon + yface_id(mx, my, c, t) * (2 * ME), qrL, qlR);
      } else {
        recon_interface<RS, 2, T, CHAR>(rp, win, qrL, qlR);
      }
      T W[MW][ME], sp[MW];
      typename RS::Roe roe;
      rpn_call<RS, 2>(rp, qrL, qlR, Alo, Ahi, W, sp, am, ap, roe);
      if (ycfl_lane && t <= my + 1) {
  #pragma unroll
        for (int w = 0; w < MW; w++) smax_y = cfl_term(smax_y, (T)fabs(sp[w]));
      }
    };
    // x contribution of this lane's cell in row r (a whole x-slice pass)
    auto x_cell = [&](const T (&row)[ME], const T (&Ar)[MA], int r, bool flagged, T (&dqx)[ME]) {
      T qrL[ME], qlR[ME], am[ME], ap[ME];
      x_interface(row, Ar, r, flagged, qrL, qlR, am, ap);
      T qr[ME], amn[ME];
      from_right<ME>(qrL, qr);     // right edge of this cell: interface c+1
      from_right<ME>(am, amn);     // amdq at interface c+1
      if constexpr (TFLUCT) {
        T fl[ME], fr[ME];
        RS::template flux<1>(rp, qlR, fl);
        RS::template flux<1>(rp, qr, fr);
  #pragma unroll
        for (int m = 0; m < ME; m++) dqx[m] = T(0) - v.dtdx * ((amn[m] + ap[m]) + (fr[m] - fl[m]));
      } else {
        T W[MW][ME], sp[MW], am2[ME], ap2[ME];
        typename RS::Roe roe;
        rpn_call<RS, 1>(rp, qlR, qr, Ar, Ar, W, sp, am2, ap2, roe);
  #pragma unroll
        for (int m = 0; m < ME; m++)
          dqx[m] = T(0) - v.dtdx * (((amn[m] + ap[m]) + am2[m]) + ap2[m]);
      }
    };

    // ---- window rows r0-3 .. r0+2; y-interface r0; aux of rows t-1 (Alo), t (Ahi)
    T win[6][ME];
  #pragma unroll
    for (int k = 0; k < 6; k++) load_y(r0 - 3 + k, win[k]);
    T Alo[MA], Ahi[MA];
    load_a(r0 - 1, Alo);
    load_a(r0, Ahi);
    if (r0 == 1) {   // x-slice of row 0 (CFL only)
      T d[ME];
      x_cell(win[2], Alo, 0, false, d);
    }
    // FLAGS: this lane's flag bytes, loaded one row iteration before their use so
    // that their latency hides behind a row's arithmetic.  xbyte(r): the flag of
    // x-interface c of row r (the column clamped to 1..mx+1, so the clamped row
    // neighbours come from the lanes beside this one); ybyte(k): the flag of
    // y-interface c in y-row k (0-based, clamped to 0..my).
    auto xbyte = [&](int r) -> int {
      if (!FLAGS || r < 1 || r > my) return 0;
      return a.flags[(long long)(r - 1) * (mx + 1) + min(max(c - 1, 0), mx)];
    };
    auto ybyte = [&](int k) -> int {
      if (!FLAGS || c < 1 || c > mx) return 0;
      return a.flags[(long long)(mx + 1) * my + (long long)min(max(k, 0), my) * mx + (c - 1)];
    };
    // x-interface c of a row is listed when its byte or a row neighbour's is set
    // (the cross-lane moves run on every lane, before any lane-dependent test)
    auto xlisted = [&](int b) {
      const int nb = b | from_left(b) | from_right(b);
      return FLAGS && c >= 1 && c <= mx + 1 && nb != 0;
    };
    // y-interface t: 0-based y-rows t-2, t-1, t (y-interfaces t-1, t, t+1)
    int ym2 = ybyte(r0 - 2), ym1 = ybyte(r0 - 1), y0 = ybyte(r0);
    T qlY[ME], apY[ME];   // interface t-1 carried: lower edge of row t-1, apdq
    {
      T qrL[ME], am[ME];
      y_interface(win, Alo, Ahi, r0, (ym2 | ym1 | y0) != 0, qrL, qlY, am, apY);
    }
    // the bytes of the first loop iteration (y-interface r0+1, x-row r0)
    ym2 = ym1;
    ym1 = y0;
    y0 = ybyte(r0 + 1);
    int xb = xbyte(r0);
    T pf_p[ME], pf_q[ME];   // the loads of the next iteration's window row
  #pragma unroll
    for (int m = 0; m < ME; m++) pf_q[m] = T(0);
    load_raw(r0 + 3, pf_p, pf_q);
    for (int t = r0 + 1; t <= rend + 1; t++) {
      // row t-1 = win[3]; shift the window to rows t-3..t+2
  #pragma unroll
      for (int k = 0; k < 5; k++)
  #pragma unroll
        for (int m = 0; m < ME; m++) win[k][m] = win[k + 1][m];
      combine(pf_p, pf_q, win[5]);
      if (t <= rend) load_raw(t + 3, pf_p, pf_q);
      // the store-time inputs of row t-1
      StoreIn<ME, T> si;
      const long long off = (long long)(c - 1) + (long long)(t - 2) * a.g.rs;
      if (out_lane) load_store_in<ME>(a, v, off, si);
  #pragma unroll
      for (int k = 0; k < MA; k++) Alo[k] = Ahi[k];
      load_a(t, Ahi);
      // the flag bytes of the next iteration (y-row t+1, x-row t)
      const int y_next = t <= rend ? ybyte(t + 1) : 0;
      const int x_next = t <= rend ? xbyte(t) : 0;
      // y-interface t closes row t-1
      T qrL[ME], qlR[ME], am[ME], ap[ME];
      y_interface(win, Alo, Ahi, t, (ym2 | ym1 | y0) != 0, qrL, qlR, am, ap);
      T dqy[ME];
      if constexpr (TFLUCT) {
        T fl[ME], fr[ME];
        RS::template flux<2>(rp, qlY, fl);
        RS::template flux<2>(rp, qrL, fr);
  #pragma unroll
        for (int m = 0; m < ME; m++) dqy[m] = T(0) - v.dtdy * ((am[m] + apY[m]) + (fr[m] - fl[m]));
      } else {
        T W[MW][ME], sp[MW], am2[ME], ap2[ME];
        typename RS::Roe roe;
        rpn_call<RS, 2>(rp, qlY, qrL, Alo, Alo, W, sp, am2, ap2, roe);
  #pragma unroll
        for (int m = 0; m < ME; m++)
          dqy[m] = T(0) - v.dtdy * (((am[m] + apY[m]) + am2[m]) + ap2[m]);
      }
      // x-slice of row t-1
      T dqx[ME];
      x_cell(win[2], Alo, t - 1, xlisted(xb), dqx);
      if (out_lane) {
        bool bad = false;
  #pragma unroll
        for (int m = 0; m < ME; m++) {
          const T o = stage_store<ME>(a, v, si, off, m, (T(0) + dqx[m]) + dqy[m], win[2][m]);
          bad = bad || (((a.pos_mask >> m) & 1u) && (!(o > T(0)) || !isfinite(o)));
        }
        nbad += bad ? 1 : 0;
      }
  #pragma unroll
      for (int m = 0; m < ME; m++) { qlY[m] = qlR[m]; apY[m] = ap[m]; }
      ym2 = ym1;
      ym1 = y0;
      y0 = y_next;
      xb = x_next;
    }
    if (rend == my) {   // x-slice of row my+1 (CFL only) = window row index 3 now
      T d[ME];
      x_cell(win[3], Ahi, my + 1, false, d);
    }
  }
  double cx = (double)v.dtdx * (double)smax_x;
  double cy = (double)v.dtdy * (double)smax_y;
  block_atomic_max(cx > cy ? cx : cy, a.cfl_bits);
  if (a.ctl && a.count_bad) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nbad += __shfl_xor(nbad, off, 64);
    if (lane == 0 && nbad > 0) atomicAdd(&a.ctl->nbad, (unsigned)nbad);
  }
}

// A controlled stage's control phase (sharp_ctl_phase), when its gate is open.
__global__ void __launch_bounds__(64) k_sharp_ctl(SharpCtl* ctl, const int* gate, int phase) {
  if (threadIdx.x != 0 || !*gate) return;
  sharp_ctl_phase(ctl, phase);
}

// elementwise SSP104 middle update (solver.py:418-419):
//   q = q/25. + 0.36*s1 ;  s1 = 15.*q - 5.*s1
template <typename T>
__global__ void k_ssp104_mid(T* q, T* s1, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T qn = q[i] / T(25) + T(0.36) * s1[i];
  q[i] = qn;
  s1[i] = T(15) * qn - T(5) * s1[i];
}

// The multistep updates that need no operator evaluation (SharpClawSolver.step,
// sharpclaw/solver.py:306-307, 351-374), elementwise with the reference's
// grouping; c = the scalar coefficients the solver computed in FP64:
//   mode 0 (Euler):    out = q + c1*u                       (c1 = dt)
//   mode 1 (SSPLMMk2): out = c0*(c1*q + c2*u) + c3*R        (beta, r, dt, delta)
//   mode 2 (SSPLMMk3): out = (c0*(c1*q + c2*u) + c3*R) + c4*v
//                      (beta_{k-1}, r, dt, r*beta0 + delta0, beta0*dt)
template <typename T>
__global__ void k_lmm_update(int mode, long long n, const T* q, const T* u, const T* R,
                             const T* v, T c0, T c1, T c2, T c3, T c4, T* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T o;
  switch (mode) {
    case 0: o = q[i] + c1 * u[i]; break;
    case 3: o = c1 * u[i]; break;
    case 4: o = c1 * u[i] + c2 * v[i]; break;
    case 5: o = q[i] + (c1 * u[i] + c2 * v[i]); break;
    default:
      o = c0 * (c1 * q[i] + c2 * u[i]) + c3 * R[i];
      if (mode == 2) o = o + c4 * v[i];
  }
  out[i] = o;
}

// ---------------------------------------------------------------- launchers
namespace {
// Rows per wavefront of the march form: with R wavefronts resident, W in the
// launch, minimise ceil(W / R) * (H + 6) (6 = the window rows loaded and the
// interface reconstructed before the first output row).  CLAWHIP_SHARP_ROWS
// overrides it.
template <class Kernel>
int march_rows(Kernel kernel, int nstrips, int my) {
  const long long resident = resident_waves((const void*)kernel, 0);
  if (resident == 0) return 16;
  if (const char* env = getenv("CLAWHIP_SHARP_ROWS")) {
    int h = atoi(env);
    if (h >= 1 && h <= 65536) return h;
  }
  int best = 16;
  long long best_cost = -1;
  for (int h = 2; h <= 256; h++) {
    const long long waves = (long long)nstrips * ((my + h - 1) / h);
    const long long cost = ((waves + resident - 1) / resident) * (long long)(h + 6);
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && h > best)) {
      best = h;
      best_cost = cost;
    }
    if ((my + h - 1) / h == 1) break;
  }
  return best;
}

// The march launch of a stage (and, with flags, the full-WENO5 pass over the
// flagged interfaces before it).
template <class RS, typename T, bool CHAR>
int launch_args(const SharpArgs<T>& a, hipStream_t stream, void* const* ev = nullptr) {
  const bool flags = a.flags != nullptr;
  if constexpr (RS::maux > 0) {
    if (flags) return CLAWHIP_E_ARG;   // the indicator path takes solvers without aux
  }
  MarchArgs<T> ma;
  ma.s = a;
  ma.nstrips = (a.g.mx + 57) / 58;
  ma.H = (flags && RS::maux == 0)
             ? march_rows(k_sharp_march<RS, T, CHAR, RS::maux == 0>, ma.nstrips, a.g.my)
             : march_rows(k_sharp_march<RS, T, CHAR, false>, ma.nstrips, a.g.my);
  ma.nrb = (a.g.my + ma.H - 1) / ma.H;
  const unsigned blocks = (unsigned)(((long long)ma.nstrips * ma.nrb + 3) / 4);
  if (flags) {
    if constexpr (RS::maux == 0) {
      hipLaunchKernelGGL((k_weno_flagged<RS, T, CHAR>), dim3(2048), dim3(256), 0, stream, a);
      if (ev) hipEventRecord((hipEvent_t)ev[0], stream);
      hipLaunchKernelGGL((k_sharp_march<RS, T, CHAR, true>), dim3(blocks), dim3(256), 0, stream,
                         ma);
      if (ev) hipEventRecord((hipEvent_t)ev[1], stream);
    }
  } else {
    if (ev) hipEventRecord((hipEvent_t)ev[0], stream);
    hipLaunchKernelGGL((k_sharp_march<RS, T, CHAR, false>), dim3(blocks), dim3(256), 0, stream,
                       ma);
    if (ev) hipEventRecord((hipEvent_t)ev[1], stream);
  }
  if (a.ctl && a.ctl_phase != PH_NONE)
    hipLaunchKernelGGL(k_sharp_ctl, dim3(1), dim3(64), 0, stream, a.ctl, a.gate, a.ctl_phase);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

template <class RS, typename T, bool CHAR>
SharpArgs<T> make_args(const clawhip_rp& rp, const Geom& g, const Geom& gin, const int* lo,
                       const int* hi, const int* refl, int ymode, int omode, double dt, double s,
                       double dx, double dy, const void* P, const void* Q, const void* A,
                       const void* B, const void* aux_in, const void* comp_in, void* out,
                       void* aux_out, void* comp_out, unsigned long long* cfl, const FlagSet& fs,
                       const AuxIn& ax, const void* src) {
  SharpArgs<T> a;
  a.rp = to_params<T>(rp);
  a.g = g;
  a.gin = gin;
  a.bc = to_bc(lo, hi, refl, 2);
  a.ng = 3;
  a.char_decomp = CHAR ? 2 : 0;
  a.ymode = ymode; a.omode = omode;
  a.dt = T(dt); a.s = T(s);
  a.rdt = T(1.0 / dt);
  a.dtdx = T(dt / dx); a.dtdy = T(dt / dy);
  a.P = (const T*)P; a.Q = (const T*)Q; a.A = (const T*)A; a.B = (const T*)B;
  a.aux_in = (const T*)aux_in; a.comp_in = (const T*)comp_in;
  a.out = (T*)out; a.aux_out = (T*)aux_out; a.comp_out = (T*)comp_out;
  a.src = (const T*)src;
  a.cfl_bits = cfl;
  a.flags = fs.flags;
  a.flag_list = fs.list;
  a.flag_count = fs.count;
  a.recon = (T*)fs.recon;
  a.qaux = (const T*)ax.p;
  a.gaux = ax.g;
  const int never[3] = {-1, -1, -1};
  a.bcaux = to_bc(ax.lo, ax.hi, never, 2);
  a.ctl = nullptr;
  a.gate = nullptr;
  a.ctl_phase = PH_NONE;
  a.count_bad = 0;
  a.res = 0u;
  a.pos_mask = 0u;
  a.qb[0] = a.qb[1] = nullptr;
  a.cb[0] = a.cb[1] = nullptr;
  a.dx = dx;
  a.dy = dy;
  return a;
}

template <class RS, typename T, bool CHAR>
int run_stage(const clawhip_rp& rp, const Geom& g, const Geom& gin, const int* lo, const int* hi,
              const int* refl, int ymode, int omode, double dt, double s, double dx, double dy,
              const void* P, const void* Q, const void* A, const void* B, const void* aux_in,
              const void* comp_in, void* out, void* aux_out, void* comp_out,
              unsigned long long* cfl, const FlagSet& fs, hipStream_t stream,
              const AuxIn& ax, const void* src) {
  if ((RS::maux > 0) != (ax.p != nullptr)) return CLAWHIP_E_ARG;
  const SharpArgs<T> a = make_args<RS, T, CHAR>(rp, g, gin, lo, hi, refl, ymode, omode, dt, s,
                                                dx, dy, P, Q, A, B, aux_in, comp_in, out,
                                                aux_out, comp_out, cfl, fs, ax, src);
  return launch_args<RS, T, CHAR>(a, stream);
}

// ---------------------------------------------------------------- controlled SSP33
// nsteps attempts of SharpClawSolver's SSP33 step under the device controller:
//   [indicator: flags of q[cur] and their list]   gate run_flags
//   stage 1: dq_dt = D(q)/dt                         gate run1, PH_MID
//   stage 2: r2 (SSP33 or FP32 stage 2)              gate run23
//   stage 3: q[1-cur] (SSP33 or Kahan stage 3)       gate run23, PH_END / PH_END_JUDGE
// and, with the indicator, once after the nsteps attempts:
//   the judged step again with full WENO5            gate run_fb, PH_END_FB
// A judged step that fails positivity closes the attempt gates, so the attempts
// after it in this launch do nothing and the recomputation ends it: a launch
// completes at least one attempt and at most nsteps (ctl->attempts counts them).
// No host synchronisation: the launches may be captured in a HIP graph.
template <class RS, typename T, bool CHAR>
int run_ssp33(const clawhip_rp& rp, const Geom& g, const int* lo, const int* hi, const int* refl,
              double dx, double dy, const SharpRunBufs& b, SharpCtl* ctl, int nsteps,
              hipStream_t stream) {
  constexpr bool F32 = sizeof(T) == 4;
  const AuxIn noaux;
  const bool judge = b.flags != nullptr;
  FlagSet fs;
  if (judge) {
    fs.flags = b.flags;
    fs.list = b.flag_list;
    fs.count = b.flag_count;
    fs.recon = b.recon;
  }
  auto base = [&](int ymode, int omode, const void* Q, const void* B, void* out, void* aux_out,
                  unsigned long long* cfl, const FlagSet& f, const int* gate, int phase,
                  unsigned res) {
    // dt, s and the buffers named in res come from ctl when the kernel runs
    SharpArgs<T> a = make_args<RS, T, CHAR>(rp, g, g, lo, hi, refl, ymode, omode, 1.0, 1.0, dx,
                                            dy, nullptr, Q, nullptr, B, nullptr, nullptr, out,
                                            aux_out, nullptr, cfl, f, noaux, nullptr);
    a.ctl = ctl;
    a.gate = gate;
    a.ctl_phase = phase;
    a.res = res;
    a.qb[0] = b.q[0]; a.qb[1] = b.q[1];
    a.cb[0] = b.comp[0]; a.cb[1] = b.comp[1];
    return a;
  };
  const int end_phase = judge ? PH_END_JUDGE : PH_END;
  // stage 1 -> dq_dt
  SharpArgs<T> s1 = base(Y_ARRAY, OUT_DQDT, nullptr, nullptr, b.dqdt, nullptr, &ctl->cfl1_bits,
                         fs, &ctl->run1, PH_MID, RES_P);
  // stage 2 -> r2 (Y = q + dt*dq_dt)
  SharpArgs<T> s2 = base(Y_AXPY, F32 ? OUT_F32_S2 : OUT_SSP33_2, b.dqdt, b.dqdt, b.r2, b.aux23,
                         &ctl->cfl23_bits, fs, &ctl->run23, PH_NONE, RES_P | RES_A);
  // stage 3 -> q[1-cur]
  SharpArgs<T> s3 = base(Y_ARRAY, F32 ? OUT_F32_S3 : OUT_SSP33_3, nullptr, nullptr, nullptr,
                         nullptr, &ctl->cfl23_bits, fs, &ctl->run23, end_phase,
                         RES_A | RES_OUT | (F32 ? RES_CIN | RES_COUT : 0u));
  s3.P = (const T*)b.r2;
  s3.aux_in = (const T*)b.aux23;
  s3.count_bad = judge ? 1 : 0;
  s3.pos_mask = b.pos_mask;
  // the judge's recomputation with full WENO5 (no flags)
  const FlagSet none;
  SharpArgs<T> f1 = base(Y_ARRAY, OUT_DQDT, nullptr, nullptr, b.dqdt_fb, nullptr,
                         &ctl->cflfb_bits, none, &ctl->run_fb, PH_NONE, RES_P);
  SharpArgs<T> f2 = base(Y_AXPY, F32 ? OUT_F32_S2 : OUT_SSP33_2, b.dqdt_fb, b.dqdt_fb, b.r2,
                         b.aux23, &ctl->cflfb_bits, none, &ctl->run_fb, PH_NONE, RES_P | RES_A);
  SharpArgs<T> f3 = base(Y_ARRAY, F32 ? OUT_F32_S3 : OUT_SSP33_3, nullptr, nullptr, nullptr,
                         nullptr, &ctl->cflfb_bits, none, &ctl->run_fb, PH_END_FB,
                         RES_A | RES_OUT | (F32 ? RES_CIN | RES_COUT : 0u));
  f3.P = (const T*)b.r2;
  f3.aux_in = (const T*)b.aux23;
  BufSel sel;
  sel.gate = &ctl->run_flags;
  sel.cur = &ctl->cur;
  sel.q[0] = b.q[0];
  sel.q[1] = b.q[1];
  sel.zero_count = b.flag_count;   // the flags kernel clears the list length
  for (int n = 0; n < nsteps; n++) {
    int rc = 0;
    if (judge) {
      rc = launch_mlp_flags(F32 ? CLAWHIP_F32K : CLAWHIP_F64, b.q[0], g, lo, hi, refl, RS::meqn,
                            b.flat, b.params, b.threshold, b.flags, stream, sel);
      if (!rc) rc = launch_flag_list(b.flags, g.mx, g.my, b.flag_list, b.flag_count, stream,
                                     &ctl->run_flags, true);
    }
    void* const* ev = b.stage_events ? b.stage_events + 6LL * n : nullptr;
    if (!rc) rc = launch_args<RS, T, CHAR>(s1, stream, ev);
    if (!rc) rc = launch_args<RS, T, CHAR>(s2, stream, ev ? ev + 2 : nullptr);
    if (!rc) rc = launch_args<RS, T, CHAR>(s3, stream, ev ? ev + 4 : nullptr);
    if (rc) return rc;
  }
  if (judge) {
    int rc = launch_args<RS, T, CHAR>(f1, stream);
    if (!rc) rc = launch_args<RS, T, CHAR>(f2, stream);
    if (!rc) rc = launch_args<RS, T, CHAR>(f3, stream);
    return rc;
  }
  return 0;
}

template <template <typename> class RSF, bool CHAR>
int dispatch_stage(int dtype, const clawhip_rp& rp, const Geom& g, const Geom& gin, const int* lo,
                   const int* hi,
                   const int* refl, int ymode, int omode, double dt, double s, double dx,
                   double dy, const void* P, const void* Q, const void* A, const void* B,
                   const void* aux_in, const void* comp_in, void* out, void* aux_out,
                   void* comp_out, unsigned long long* cfl, const FlagSet& flags,
                   hipStream_t stream, const AuxIn& ax, const void* src) {
  if (dtype != CLAWHIP_F32K)
    return run_stage<RSF<double>, double, CHAR>(rp, g, gin, lo, hi, refl, ymode, omode, dt, s, dx,
                                                dy, P, Q, A, B, aux_in, comp_in, out, aux_out,
                                                comp_out, cfl, flags, stream, ax, src);
#if CLAW_FAST
  return run_stage<RSF<float>, float, CHAR>(rp, g, gin, lo, hi, refl, ymode, omode, dt, s, dx, dy, P,
                                            Q, A, B, aux_in, comp_in, out, aux_out, comp_out,
                                            cfl, flags, stream, ax, src);
#else
  // FP32 + Kahan is dispatched to the fast variant only
  return CLAWHIP_E_ARG;
#endif
}
}  // namespace

int launch_sharp_stage(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                       const Geom& gin, const int* lo, const int* hi, const int* refl, int ymode, int omode,
                       double dt, double s, double dx, double dy, const void* P, const void* Q,
                       const void* A, const void* B, const void* aux_in, const void* comp_in,
                       void* out, void* aux_out, void* comp_out, unsigned long long* cfl,
                       const FlagSet& flags, hipStream_t stream, const AuxIn& ax,
                       const void* src) {
#define ARGS dtype, rp, g, gin, lo, hi, refl, ymode, omode, dt, s, dx, dy, P, Q, A, B, aux_in, \
             comp_in, out, aux_out, comp_out, cfl, flags, stream, ax, src
  if (rp.kind == CLAWHIP_RP_VC_ACOUSTICS || rp.kind == CLAWHIP_RP_VC_ADVECTION ||
      rp.kind == CLAWHIP_RP_VC_ADVECTION_2D) {
    if (char_decomp != 0) return CLAWHIP_E_KIND;   // component WENO5 only
    switch (rp.kind == CLAWHIP_RP_VC_ACOUSTICS ? 0
            : rp.kind == CLAWHIP_RP_VC_ADVECTION_2D ? 1 : rp.meqn) {
      case 0: return dispatch_stage<VCAcoustics, false>(ARGS);
      case 1: return dispatch_stage<VCAdvection1, false>(ARGS);
      case 2: return dispatch_stage<VCAdvection2, false>(ARGS);
      case 3: return dispatch_stage<VCAdvection3, false>(ARGS);
      case 4: return dispatch_stage<VCAdvection4, false>(ARGS);
    }
    return CLAWHIP_E_ARG;
  }
  if (ax.p) return CLAWHIP_E_ARG;
  if (char_decomp == 2) {
    switch (rp.kind) {
      case CLAWHIP_RP_SHALLOW: return dispatch_stage<Shallow, true>(ARGS);
      case CLAWHIP_RP_SHALLOW_HLLE: return dispatch_stage<ShallowHLLE, true>(ARGS);
      case CLAWHIP_RP_EULER: return dispatch_stage<Euler, true>(ARGS);
    }
    return CLAWHIP_E_KIND;
  }
  switch (rp.kind) {
    case CLAWHIP_RP_SHALLOW: return dispatch_stage<Shallow, false>(ARGS);
    case CLAWHIP_RP_SHALLOW_HLLE: return dispatch_stage<ShallowHLLE, false>(ARGS);
    case CLAWHIP_RP_EULER: return dispatch_stage<Euler, false>(ARGS);
    case CLAWHIP_RP_EULER5: return dispatch_stage<Euler5, false>(ARGS);
    case CLAWHIP_RP_ACOUSTICS: return dispatch_stage<Acoustics, false>(ARGS);
    case CLAWHIP_RP_ADVECTION: return dispatch_stage<Advection, false>(ARGS);
  }
#undef ARGS
  return CLAWHIP_E_KIND;
}

int launch_sharp_ssp33_run(const clawhip_rp& rp, int dtype, int char_decomp, const Geom& g,
                          const int* lo, const int* hi, const int* refl, double dx, double dy,
                          const SharpRunBufs& b, SharpCtl* ctl, int nsteps, hipStream_t stream) {
  const bool f64 = dtype != CLAWHIP_F32K;
#define RUN(RSF, CH)                                                                          \
  (f64 ? run_ssp33<RSF<double>, double, CH>(rp, g, lo, hi, refl, dx, dy, b, ctl, nsteps, stream) \
       : RUN_F32(RSF, CH))
#if CLAW_FAST
#define RUN_F32(RSF, CH) run_ssp33<RSF<float>, float, CH>(rp, g, lo, hi, refl, dx, dy, b, ctl, nsteps, stream)
#else
#define RUN_F32(RSF, CH) CLAWHIP_E_ARG   // FP32 + Kahan: the fast variant only
#endif
  if (char_decomp == 2) {
    switch (rp.kind) {
      case CLAWHIP_RP_SHALLOW: return RUN(Shallow, true);
      case CLAWHIP_RP_SHALLOW_HLLE: return RUN(ShallowHLLE, true);
      case CLAWHIP_RP_EULER: return RUN(Euler, true);
    }
    return CLAWHIP_E_KIND;
  }
  switch (rp.kind) {
    case CLAWHIP_RP_SHALLOW: return RUN(Shallow, false);
    case CLAWHIP_RP_SHALLOW_HLLE: return RUN(ShallowHLLE, false);
    case CLAWHIP_RP_EULER: return RUN(Euler, false);
    case CLAWHIP_RP_EULER5: return RUN(Euler5, false);
    case CLAWHIP_RP_ACOUSTICS: return RUN(Acoustics, false);
    case CLAWHIP_RP_ADVECTION: return RUN(Advection, false);
  }
#undef RUN
#undef RUN_F32
  return CLAWHIP_E_KIND;
}

int launch_lmm_update(int dtype, int mode, long long n, const void* q, const void* u,
                      const void* R, const void* v, const double* c, void* out,
                      hipStream_t stream) {
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (dtype != CLAWHIP_F32K)
    hipLaunchKernelGGL(k_lmm_update<double>, dim3(blocks), dim3(256), 0, stream, mode, n,
                       (const double*)q, (const double*)u, (const double*)R, (const double*)v,
                       c[0], c[1], c[2], c[3], c[4], (double*)out);
  else
    hipLaunchKernelGGL(k_lmm_update<float>, dim3(blocks), dim3(256), 0, stream, mode, n,
                       (const float*)q, (const float*)u, (const float*)R, (const float*)v,
                       (float)c[0], (float)c[1], (float)c[2], (float)c[3], (float)c[4],
                       (float*)out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

int launch_ssp104_mid(int dtype, void* q, void* s1, long long n, hipStream_t stream) {
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (dtype != CLAWHIP_F32K)
    hipLaunchKernelGGL(k_ssp104_mid<double>, dim3(blocks), dim3(256), 0, stream, (double*)q,
                       (double*)s1, n);
  else
    hipLaunchKernelGGL(k_ssp104_mid<float>, dim3(blocks), dim3(256), 0, stream, (float*)q,
                       (float*)s1, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace CLAW_VNS
}  // namespace claw
