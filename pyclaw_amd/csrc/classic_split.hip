// classic_split.hip -- dimension-split classic sweeps for 2D (classic/step2ds.f90)
// and 3D (classic/step3ds.f90 + flux3.f90:168-263) on gfx950.
//
// k_xsweep: x-sweeps, lane-parallel along a row (like the unsplit kernel's x-pass,
//   no transverse terms).  It writes the reference's intermediate qbc: swept rows
//   get the updated interior plus the BC-filled ghost columns; unswept rows are
//   copied from the BC-filled input.  2D sweeps every row 1-ng..my+ng
//   (step2ds.f90:96); 3D sweeps rows j,k = 0..m+1 (step3ds.f90:134-135).
// k_march: one sweep along y or z: lanes take consecutive x columns (coalesced),
//   each lane marches along the sweep axis with a three-interface register window
//   for the limiter.  Used for the 2D y-sweep (columns 1-ng..mx+ng,
//   step2ds.f90:180), and the 3D y- and z-sweeps (pencils 0..m+1,
//   step3ds.f90:215-216, 298-299).
// The 2D split uses flux2's 0.5*cqxx; the 3D split flux3's cqxx with the 0.5
// inside the sum (flux3.f90:244) -- the HALF_IN_SUM template flag.
#include "variant.hpp"  // first: selects the arithmetic variant
#include <stdlib.h>

#include "launch.hpp"

namespace claw {
namespace CLAW_VNS {

// flux3.f90:231-259: the 0.5 is applied inside the wave sum, fadd += cqxx.
template <int ME, int MW, typename T>
__device__ __forceinline__ void second_order3(const T (&W)[MW][ME], const T (&s)[MW], T dtdx,
                                              T (&F)[ME]) {
  T dtdxave = T(0.5) * (dtdx + dtdx);
  T cq[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) cq[m] = T(0);
#pragma unroll
  for (int w = 0; w < MW; w++) {
#pragma unroll
    for (int m = 0; m < ME; m++) {
      T sa = fabs(s[w]);
      cq[m] = cq[m] + T(0.5) * sa * (T(1) - sa * dtdxave) * W[w][m];
    }
  }
#pragma unroll
  for (int m = 0; m < ME; m++) F[m] = T(0) + cq[m];
}

template <int ME, int MW, bool HALF_IN_SUM, typename T>
__device__ __forceinline__ void corr_flux(const T (&W)[MW][ME], const T (&s)[MW], T dtdx,
                                          T (&F)[ME], int fwave) {
  if (HALF_IN_SUM) {
    second_order3<ME, MW>(W, s, dtdx, F);   // 3D: the entry points refuse f-waves
  } else {
    T cq[ME];
    second_order<ME, MW>(W, s, dtdx, cq, F, fwave);
  }
}

template <typename T>
struct SweepArgs {
  RPParams<T> rp;
  Geom gin, gout;       // input / output geometry
  BCSpec bc;            // used when reading a compact input (x-sweeps)
  int ng, order;
  int mthlim[kMaxWaves];
  T dtdx;               // dt / delta of the sweep direction
  // k_xsweep
  int nstrips;
  int xrows;            // rows per wavefront
  // k_march
  int axis;             // 2: sweep along j, 3: sweep along k
  int i_lo, i_hi;       // columns swept
  int p_lo, p_hi;       // pencils swept (the third index)
  int w_ilo, w_ihi, w_plo, w_phi;  // cells written to the output (plus s in 1..m)
  int H, nrb, npencil, nistrips;
  const T* qin;
  T* qout;
  const T* comp_in;
  T* comp_out;
  unsigned long long* cfl_bits;
  // aux fields (solvers with RS::maux > 0; 2D): compact or with stored ghosts,
  // boundary values by bcaux (aux BCs never negate)
  const T* aux;
  Geom gaux;
  BCSpec bcaux;
};

// ---------------------------------------------------------------- x-sweeps
// A wavefront sweeps a run of a.xrows rows of one 60-column strip (the next
// row's cells are loaded before the current row's arithmetic), and a workgroup
// adds its Courant number with one atomic: one row per wavefront with an atomic
// each made the 384^3 x-sweep 1.05 M same-address atomics, which serialise
// (11.9 ms of an 18 ms step).  Every wavefront reaches the workgroup barrier.
template <class RS, typename T, bool KAHAN, int DIM, bool HALF_IN_SUM>
__global__ void __launch_bounds__(256) k_xsweep(const SweepArgs<T> a) {
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int ng = a.ng;
  const int mx = a.gin.mx, my = a.gin.my, mz = DIM == 3 ? a.gin.mz : 1;
  const int nrows_j = my + 2 * ng;
  const int nrows = DIM == 3 ? nrows_j * (mz + 2 * ng) : nrows_j;
  const int ngroups = (nrows + a.xrows - 1) / a.xrows;
  const bool active = wid < a.nstrips * ngroups;   // uniform per wavefront
  const int strip = active ? wid % a.nstrips : 0, grp = active ? wid / a.nstrips : 0;
  const int row0 = grp * a.xrows;
  const int row_end = active ? min(row0 + a.xrows, nrows) : row0;
  const int c = 1 + strip * 60 - 2 + lane;
  constexpr int MA = RS::maux > 0 ? RS::maux : 1;
  auto row_jk = [&](int row, int& j, int& k) {
    j = 1 - ng + row % nrows_j;
    k = DIM == 3 ? 1 - ng + row / nrows_j : 1;
  };
  auto load_row = [&](int row, T (&Q)[ME]) {
    int j, k;
    row_jk(row, j, k);
    if (DIM == 2) load_cell2<ME>(a.qin, a.gin, a.bc, ng, c, j, Q);
    else load_cell3<ME>(a.qin, a.gin, a.bc, ng, c, j, k, Q);
  };
  double cfl = 0.0;
  T Qn[ME];
  if (row0 < row_end) load_row(row0, Qn);
  for (int row = row0; row < row_end; row++) {
    T Q[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) Q[m] = Qn[m];
    if (row + 1 < row_end) load_row(row + 1, Qn);
    int j, k;
    row_jk(row, j, k);
    const bool swept = DIM == 2 ? true : (j >= 0 && j <= my + 1 && k >= 0 && k <= mz + 1);
    const bool interior_row = j >= 1 && j <= my && (DIM == 2 || (k >= 1 && k <= mz));
    T A[MA], AL[MA];
    if constexpr (RS::maux > 0) {
      static_assert(DIM == 2, "aux fields: 2D sweeps");
      load_cell2<MA>(a.aux, a.gaux, a.bcaux, ng, c, j, A);
      from_left(A, AL);
    }
    T upd[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) upd[m] = Q[m];
    if (swept) {   // uniform across the wavefront
      T QL[ME];
      from_left(Q, QL);
      T W[MW][ME], s[MW], am[ME], ap[ME];
      typename RS::Roe roe;
      rpn_call<RS, 1>(a.rp, QL, Q, AL, A, W, s, am, ap, roe);
      T smax = T(0);
      if (lane >= 1 && c >= 1 && c <= mx + 1) {
#pragma unroll
        for (int w = 0; w < MW; w++) smax = cfl_term(smax, fabs(s[w]));
      }
      const double row_cfl = (double)a.dtdx * (double)smax;
      cfl = row_cfl > cfl ? row_cfl : cfl;
      T F[ME];
#pragma unroll
      for (int m = 0; m < ME; m++) F[m] = T(0);
      if (a.order != 1) {
        T Wn[MW][ME];
#pragma unroll
        for (int w = 0; w < MW; w++) from_right(W[w], Wn[w]);
        T dotr[MW], dotl[MW];
        wave_dots<ME, MW>(W, Wn, dotr);
        from_left(dotr, dotl);
        limit_waves<ME, MW>(W, s, dotl, dotr, a.mthlim);
        corr_flux<ME, MW, HALF_IN_SUM>(W, s, a.dtdx, F, a.rp.fwave);
      }
      T amn[ME], Fr[ME];
      from_right(am, amn);
      from_right(F, Fr);
#pragma unroll
      for (int m = 0; m < ME; m++) {
        T qadd = (T(0) - a.dtdx * ap[m]) - a.dtdx * amn[m];
        upd[m] = KAHAN ? (qadd - a.dtdx * (Fr[m] - F[m]))
                       : ((Q[m] + qadd) - a.dtdx * (Fr[m] - F[m]));
      }
    }
    const bool own = lane >= 2 && lane < 62 && c >= 1 && c <= mx;
    const bool ghost_col = (c < 1 || c > mx) && c >= 1 - ng && c <= mx + ng;
    if (own || ghost_col) {
      const bool update = own && swept;
      T* po = DIM == 2 ? cell_ptr2(a.qout, a.gout, c, j) : cell_ptr3(a.qout, a.gout, c, j, k);
      if (KAHAN) {
        T* pco = DIM == 2 ? cell_ptr2(a.comp_out, a.gout, c, j) : cell_ptr3(a.comp_out, a.gout, c, j, k);
        if (update && interior_row) {
          // compensation of the compact input state
          const T* pci = DIM == 2 ? cell_ptr2(a.comp_in, a.gin, c, j) : cell_ptr3(a.comp_in, a.gin, c, j, k);
#pragma unroll
          for (int m = 0; m < ME; m++) {
            T y = upd[m] - pci[(long long)m * a.gin.fs];
            T tq = Q[m] + y;
            pco[(long long)m * a.gout.fs] = (tq - Q[m]) - y;
            po[(long long)m * a.gout.fs] = tq;
          }
        } else {
#pragma unroll
          for (int m = 0; m < ME; m++) {
            po[(long long)m * a.gout.fs] = update ? Q[m] + upd[m] : Q[m];
            pco[(long long)m * a.gout.fs] = T(0);
          }
        }
      } else {
#pragma unroll
        for (int m = 0; m < ME; m++) po[(long long)m * a.gout.fs] = update ? upd[m] : Q[m];
      }
    }
  }
  block_atomic_max(cfl, a.cfl_bits);
}

// ---------------------------------------------------------------- y / z march
template <int AXIS>
__device__ __forceinline__ long long cell_off(const Geom& g, int i, int s, int p) {
  // AXIS 2: s = j, p = k ; AXIS 3: s = k, p = j
  int j = AXIS == 2 ? s : p;
  int k = AXIS == 2 ? p : s;
  return (long long)(i - 1 + g.g) + (long long)(j - 1 + g.g) * g.rs + (long long)(k - 1 + g.g) * g.ss;
}

template <class RS, typename T, bool KAHAN, int AXIS, bool HALF_IN_SUM>
__global__ void __launch_bounds__(256) k_march(const SweepArgs<T> a) {
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  constexpr int IXY = AXIS;   // Riemann direction: 2 = y, 3 = z
  using Roe = typename RS::Roe;
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wid >= a.nistrips * a.npencil * a.nrb) return;
  const int istrip = wid % a.nistrips;
  const int rest = wid / a.nistrips;
  const int pidx = rest % a.npencil, rb = rest / a.npencil;
  const int i = a.i_lo + istrip * 64 + lane;
  const int p = a.p_lo + pidx;
  const bool live = i <= a.i_hi;
  const int ic = live ? i : a.i_hi;     // dead lanes read a valid column
  const int ng = a.ng;
  const int m_ax = AXIS == 2 ? a.gin.my : a.gin.mz;
  const int j0 = 1 + rb * a.H;
  const int jend = min(j0 + a.H, m_ax + 1);
  const T dtdx = a.dtdx;
  const bool wcol = live && ic >= a.w_ilo && ic <= a.w_ihi && p >= a.w_plo && p <= a.w_phi;
  T smax = T(0);

  auto load = [&](int s, T (&out)[ME]) {
    s = s < 1 - ng ? 1 - ng : (s > m_ax + ng ? m_ax + ng : s);
    const T* ptr = a.qin + cell_off<AXIS>(a.gin, ic, s, p);
#pragma unroll
    for (int m = 0; m < ME; m++) out[m] = ptr[(long long)m * a.gin.fs];
  };
  constexpr int MA = RS::maux > 0 ? RS::maux : 1;
  auto load_aux = [&](int s, T (&out)[MA]) {
    if constexpr (RS::maux > 0) {
      static_assert(AXIS == 2, "aux fields: 2D sweeps");
      load_cell2<MA>(a.aux, a.gaux, a.bcaux, ng, ic, s, out);
    } else {
      (void)s; out[0] = T(0);
    }
  };

  T Qm[ME], Q0[ME], Q1[ME], Am[MA], A0[MA], A1[MA];
  load(j0 - 2, Qm);
  load(j0 - 1, Q0);
  load(j0, Q1);
  load_aux(j0 - 2, Am);
  load_aux(j0 - 1, A0);
  load_aux(j0, A1);
  T Wt[MW][ME], st[MW], amt[ME], apt[ME], dotl[MW];
  Roe roet;
  {
    T Wp[MW][ME], sp[MW], amp[ME], app[ME];
    Roe roep;
    rpn_call<RS, IXY>(a.rp, Qm, Q0, Am, A0, Wp, sp, amp, app, roep);
    if (j0 - 1 >= 1) {
#pragma unroll
      for (int w = 0; w < MW; w++) smax = cfl_term(smax, fabs(sp[w]));
    }
    rpn_call<RS, IXY>(a.rp, Q0, Q1, A0, A1, Wt, st, amt, apt, roet);
#pragma unroll
    for (int w = 0; w < MW; w++) smax = cfl_term(smax, fabs(st[w]));
    wave_dots<ME, MW>(Wp, Wt, dotl);
  }
  T qold_prev[ME], qadd_p[ME], F_p[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) { Q0[m] = Q1[m]; qold_prev[m] = T(0); qadd_p[m] = T(0); F_p[m] = T(0); }
#pragma unroll
  for (int m = 0; m < MA; m++) A0[m] = A1[m];

  for (int t = j0; t <= jend; t++) {
    load(t + 1, Q1);
    load_aux(t + 1, A1);
    T Wn[MW][ME], sn[MW], amn[ME], apn[ME], dotn[MW];
    Roe roen;
    rpn_call<RS, IXY>(a.rp, Q0, Q1, A0, A1, Wn, sn, amn, apn, roen);
    if (t + 1 <= m_ax + 1) {
#pragma unroll
      for (int w = 0; w < MW; w++) smax = cfl_term(smax, fabs(sn[w]));
    }
    T F[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) F[m] = T(0);
#pragma unroll
    for (int w = 0; w < MW; w++) dotn[w] = T(0);
    if (a.order != 1) {
      wave_dots<ME, MW>(Wt, Wn, dotn);
      limit_waves<ME, MW>(Wt, st, dotl, dotn, a.mthlim);
      corr_flux<ME, MW, HALF_IN_SUM>(Wt, st, dtdx, F, a.rp.fwave);
    }
    if (t > j0 && wcol) {
      // cell t-1: apdq of interface t-1 (carried), amdq of interface t
      const int s = t - 1;
      T* po = a.qout + cell_off<AXIS>(a.gout, ic, s, p);
      if (KAHAN) {
        const T* pci = a.comp_in + cell_off<AXIS>(a.gin, ic, s, p);
        T* pco = a.comp_out + cell_off<AXIS>(a.gout, ic, s, p);
#pragma unroll
        for (int m = 0; m < ME; m++) {
          T qadd = qadd_p[m] - dtdx * amt[m];
          T d = qadd - dtdx * (F[m] - F_p[m]);
          T y = d - pci[(long long)m * a.gin.fs];
          T tq = qold_prev[m] + y;
          pco[(long long)m * a.gout.fs] = (tq - qold_prev[m]) - y;
          po[(long long)m * a.gout.fs] = tq;
        }
      } else {
#pragma unroll
        for (int m = 0; m < ME; m++) {
          T qadd = qadd_p[m] - dtdx * amt[m];
          po[(long long)m * a.gout.fs] = (qold_prev[m] + qadd) - dtdx * (F[m] - F_p[m]);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < ME; m++) {
      qadd_p[m] = T(0) - dtdx * apt[m];
      F_p[m] = F[m];
      amt[m] = amn[m];
      apt[m] = apn[m];
      qold_prev[m] = Q0[m];
      Q0[m] = Q1[m];
    }
#pragma unroll
    for (int m = 0; m < MA; m++) A0[m] = A1[m];
#pragma unroll
    for (int w = 0; w < MW; w++) {
      st[w] = sn[w];
      dotl[w] = dotn[w];
#pragma unroll
      for (int m = 0; m < ME; m++) Wt[w][m] = Wn[w][m];
    }
    roet = roen;
  }
  double cfl = live ? (double)a.dtdx * (double)smax : 0.0;
  wave_atomic_max(cfl, a.cfl_bits);
}

// ---------------------------------------------------------------- launchers
namespace {
constexpr int kMarchRows = 32;

// Rows per wavefront of an out-of-place y/z sweep: kMarchRows, or
// CLAWHIP_SPLIT_ROWS (tuning sweeps).  In-place sweeps march whole pencils
// (launch_march).
int march_rows_split() {
  if (const char* env = getenv("CLAWHIP_SPLIT_ROWS")) {
    const int h = atoi(env);
    if (h >= 1 && h <= 65536) return h;
  }
  return kMarchRows;
}

template <typename T>
SweepArgs<T> base_args(const clawhip_rp& rp, const Geom& gin, const Geom& gout, const int* lo,
                       const int* hi, const int* refl, int ndim, int ng, const int* method,
                       const int* mthlim, int mwaves, double dtdx, const void* qin, void* qout,
                       const void* cin, void* cout, unsigned long long* cfl) {
  SweepArgs<T> a;
  a.rp = to_params<T>(rp);
  a.gin = gin; a.gout = gout;
  a.bc = to_bc(lo, hi, refl, ndim);
  a.ng = ng; a.order = method[1];
  for (int w = 0; w < kMaxWaves; w++) a.mthlim[w] = w < mwaves ? mthlim[w] : 0;
  a.dtdx = T(dtdx);
  a.qin = (const T*)qin; a.qout = (T*)qout;
  a.comp_in = (const T*)cin; a.comp_out = (T*)cout;
  a.cfl_bits = cfl;
  a.nstrips = 0; a.xrows = 1; a.axis = 0; a.i_lo = a.i_hi = a.p_lo = a.p_hi = 0;
  a.w_ilo = a.w_ihi = a.w_plo = a.w_phi = 0;
  a.H = march_rows_split(); a.nrb = a.npencil = a.nistrips = 0;
  a.aux = nullptr;
  a.gaux = gout;
  a.bcaux = a.bc;
  return a;
}

// The aux fields of a 2D split step: bcaux from the aux BC codes, never negated.
template <typename T>
void set_aux(SweepArgs<T>& a, const void* aux, const Geom* gaux, const int* alo,
             const int* ahi) {
  if (!aux) return;
  const int norefl[3] = {-1, -1, -1};
  a.aux = (const T*)aux;
  a.gaux = *gaux;
  a.bcaux = to_bc(alo, ahi, norefl, 2);
}

template <class RS, typename T, bool KAHAN, int DIM, bool HALF>
int launch_x(SweepArgs<T> a, hipStream_t stream) {
  const int mx = a.gin.mx, my = a.gin.my, mz = DIM == 3 ? a.gin.mz : 1;
  a.nstrips = (mx + 59) / 60;
  long long rows = (long long)(my + 2 * a.ng) * (DIM == 3 ? (mz + 2 * a.ng) : 1);
  // rows per wavefront: one resident round of wavefronts over the (strip, row)
  // tasks, or CLAWHIP_XSWEEP_ROWS (tuning sweeps)
  const long long tasks = (long long)a.nstrips * rows;
  const long long resident = resident_waves((const void*)k_xsweep<RS, T, KAHAN, DIM, HALF>, 0);
  long long xr = resident > 0 ? (tasks + resident - 1) / resident : 1;
  if (const char* env = getenv("CLAWHIP_XSWEEP_ROWS")) {
    const long long h = atoll(env);
    if (h >= 1 && h <= 65536) xr = h;
  }
  a.xrows = (int)(xr < 1 ? 1 : xr > rows ? rows : xr);
  const long long waves = (long long)a.nstrips * ((rows + a.xrows - 1) / a.xrows);
  hipLaunchKernelGGL((k_xsweep<RS, T, KAHAN, DIM, HALF>), dim3((unsigned)((waves + 3) / 4)),
                     dim3(256), 0, stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

template <class RS, typename T, bool KAHAN, int AXIS, bool HALF>
int launch_march(SweepArgs<T> a, hipStream_t stream) {
  const int m_ax = AXIS == 2 ? a.gin.my : a.gin.mz;
  a.axis = AXIS;
  // An in-place sweep must march whole pencils: a block's halo rows would
  // otherwise be rewritten by the next block of the same pencil.
  if ((const void*)a.qin == (const void*)a.qout) a.H = m_ax;
  a.nistrips = (a.i_hi - a.i_lo + 1 + 63) / 64;
  a.npencil = a.p_hi - a.p_lo + 1;
  a.nrb = (m_ax + a.H - 1) / a.H;
  long long waves = (long long)a.nistrips * a.npencil * a.nrb;
  hipLaunchKernelGGL((k_march<RS, T, KAHAN, AXIS, HALF>), dim3((unsigned)((waves + 3) / 4)),
                     dim3(256), 0, stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

// 2D: x-sweep into the padded intermediate, y-sweep into the compact output.
template <class RS, typename T, bool KAHAN>
int run_split2(const clawhip_rp& rp, const void* q_in, const Geom& gin, void* q_out,
               const Geom& gout, const void* cin, void* cout, void* work, void* cwork,
               const int* lo, const int* hi, const int* refl, int ng, const int* method,
               const int* mthlim, double dx, double dy, double dt, const AuxIn& aux,
               unsigned long long* cfl, hipStream_t stream) {
  Geom gw;
  gw.mx = gout.mx; gw.my = gout.my; gw.mz = 1; gw.g = ng;
  gw.rs = gout.mx + 2 * ng; gw.fs = (long long)gw.rs * (gout.my + 2 * ng);
  gw.ss = 0;   // 2D: no k index (cell_off adds (k-1+g)*ss with k = 1)
  // The whole qbc as output: the x-sweep writes it (ghost rows swept, ghost
  // columns copied), then the y-sweep runs in place on it and also writes the
  // ghost columns' interior rows, as step2ds.f90:180-255 does.
  const bool whole = gout.g == ng;
  if (whole) {
    gw = gout;
    gw.ss = 0;
    work = q_out;
    cwork = cout;
  }
  auto ax = base_args<T>(rp, gin, gw, lo, hi, refl, 2, ng, method, mthlim, RS::mwaves, dt / dx,
                         q_in, work, cin, cwork, cfl);
  set_aux(ax, aux.p, &aux.g, aux.lo, aux.hi);
  int rc = launch_x<RS, T, KAHAN, 2, false>(ax, stream);
  if (rc) return rc;
  auto ay = base_args<T>(rp, gw, whole ? gw : gout, lo, hi, refl, 2, ng, method, mthlim,
                         RS::mwaves, dt / dy, work, q_out, cwork, cout, cfl);
  set_aux(ay, aux.p, &aux.g, aux.lo, aux.hi);
  ay.i_lo = 1 - ng; ay.i_hi = gout.mx + ng; ay.p_lo = ay.p_hi = 1;
  ay.w_ilo = whole ? 1 - ng : 1; ay.w_ihi = whole ? gout.mx + ng : gout.mx;
  ay.w_plo = 1; ay.w_phi = 1;
  return launch_march<RS, T, KAHAN, 2, false>(ay, stream);
}

// 3D: x into w (padded), y in place on w, z into the compact output.
template <class RS, typename T, bool KAHAN>
int run_split3(const clawhip_rp& rp, const void* q_in, const Geom& gin, void* q_out,
               const Geom& gout, const void* cin, void* cout, void* work, void* cwork,
               const int* lo, const int* hi, const int* refl, int ng, const int* method,
               const int* mthlim, double dx, double dy, double dz, double dt,
               unsigned long long* cfl, hipStream_t stream) {
  Geom gw;
  gw.mx = gout.mx; gw.my = gout.my; gw.mz = gout.mz; gw.g = ng;
  gw.rs = gout.mx + 2 * ng;
  gw.ss = gw.rs * (long long)(gout.my + 2 * ng);
  gw.fs = gw.ss * (long long)(gout.mz + 2 * ng);
  auto ax = base_args<T>(rp, gin, gw, lo, hi, refl, 3, ng, method, mthlim, RS::mwaves, dt / dx,
                         q_in, work, cin, cwork, cfl);
  int rc = launch_x<RS, T, KAHAN, 3, true>(ax, stream);
  if (rc) return rc;
  auto ay = base_args<T>(rp, gw, gw, lo, hi, refl, 3, ng, method, mthlim, RS::mwaves, dt / dy,
                         work, work, cwork, cwork, cfl);
  ay.i_lo = 0; ay.i_hi = gout.mx + 1; ay.p_lo = 0; ay.p_hi = gout.mz + 1;
  ay.w_ilo = 0; ay.w_ihi = gout.mx + 1; ay.w_plo = 0; ay.w_phi = gout.mz + 1;
  rc = launch_march<RS, T, KAHAN, 2, true>(ay, stream);
  if (rc) return rc;
  auto az = base_args<T>(rp, gw, gout, lo, hi, refl, 3, ng, method, mthlim, RS::mwaves, dt / dz,
                         work, q_out, cwork, cout, cfl);
  az.i_lo = 0; az.i_hi = gout.mx + 1; az.p_lo = 0; az.p_hi = gout.my + 1;
  az.w_ilo = 1; az.w_ihi = gout.mx; az.w_plo = 1; az.w_phi = gout.my;
  return launch_march<RS, T, KAHAN, 3, true>(az, stream);
}

// One sweep of step2ds / step3ds between arrays that store their ghost cells
// (the reference's qbc, filled by the caller: BC_GIVEN on every side).  Sweep 1
// reads q_in and writes every cell of q_out (interior columns updated on all
// rows, ghost columns copied), as step2ds/step3ds with ids/idir = 1 do; sweeps
// 2 and 3 update q_out in place (q_in == q_out), marching whole pencils.
template <class RS>
int run_sweep(const clawhip_rp& rp, int ndim, int dir, const void* q_in, const Geom& gin,
              void* q_out, const Geom& gout, int ng, const int* method, const int* mthlim,
              double dtdx, const AuxIn& aux, unsigned long long* cfl, hipStream_t stream) {
  const int given[3] = {BC_GIVEN, BC_GIVEN, BC_GIVEN};
  const int refl[3] = {1, 2, 3};
  auto a = base_args<double>(rp, gin, gout, given, given, refl, ndim, ng, method, mthlim,
                             RS::mwaves, dtdx, q_in, q_out, nullptr, nullptr, cfl);
  set_aux(a, aux.p, &aux.g, given, given);
  if (ndim == 2) {
    if (dir == 1) return launch_x<RS, double, false, 2, false>(a, stream);
    // step2ds ids = 2 (step2ds.f90:180-255): columns 1-ng..mx+ng are swept and
    // written (rows 1..my), ghost columns included
    a.i_lo = 1 - ng; a.i_hi = gout.mx + ng; a.p_lo = a.p_hi = 1;
    a.w_ilo = 1 - ng; a.w_ihi = gout.mx + ng; a.w_plo = 1; a.w_phi = 1;
    return launch_march<RS, double, false, 2, false>(a, stream);
  }
  if constexpr (RS::maux > 0) {
    return CLAWHIP_E_KIND;   // aux fields: 2D only
  } else {
    if (dir == 1) return launch_x<RS, double, false, 3, true>(a, stream);
    if (dir == 2) {    // step3ds idir = 2: i, k in 0..m+1
      a.i_lo = 0; a.i_hi = gout.mx + 1; a.p_lo = 0; a.p_hi = gout.mz + 1;
      a.w_ilo = 0; a.w_ihi = gout.mx + 1; a.w_plo = 0; a.w_phi = gout.mz + 1;
      return launch_march<RS, double, false, 2, true>(a, stream);
    }
    a.i_lo = 0; a.i_hi = gout.mx + 1; a.p_lo = 0; a.p_hi = gout.my + 1;   // idir = 3
    a.w_ilo = 0; a.w_ihi = gout.mx + 1; a.w_plo = 0; a.w_phi = gout.my + 1;
    return launch_march<RS, double, false, 3, true>(a, stream);
  }
}

template <template <typename> class RSF>
int dispatch_split2(int dtype, const clawhip_rp& rp, const void* q_in, const Geom& gin,
                    void* q_out, const Geom& gout, const void* cin, void* cout, void* work,
                    void* cwork, const int* lo, const int* hi, const int* refl, int ng,
                    const int* method, const int* mthlim, double dx, double dy, double dt,
                    const AuxIn& aux, unsigned long long* cfl, hipStream_t s) {
  if (dtype != CLAWHIP_F32K)
    return run_split2<RSF<double>, double, false>(rp, q_in, gin, q_out, gout, cin, cout, work,
                                                  cwork, lo, hi, refl, ng, method, mthlim, dx,
                                                  dy, dt, aux, cfl, s);
#if CLAW_FAST
  return run_split2<RSF<float>, float, true>(rp, q_in, gin, q_out, gout, cin, cout, work, cwork,
                                             lo, hi, refl, ng, method, mthlim, dx, dy, dt, aux,
                                             cfl, s);
#else
  // FP32 + Kahan is dispatched to the fast variant only
  return CLAWHIP_E_ARG;
#endif
}
}  // namespace

int launch_split2(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dt, const AuxIn& aux, unsigned long long* cfl, hipStream_t s) {
  if (rp.kind == CLAWHIP_RP_VC_ADVECTION || rp.kind == CLAWHIP_RP_VC_ADVECTION_2D) {
    if (!aux.p) return CLAWHIP_E_ARG;
    switch (rp.kind == CLAWHIP_RP_VC_ADVECTION_2D ? 1 : rp.meqn) {
      case 1: return dispatch_split2<VCAdvection1>(dtype, rp, q_in, gin, q_out, gout, comp_in,
                                                   comp_out, work, cwork, lo, hi, refl, ng,
                                                   method, mthlim, dx, dy, dt, aux, cfl, s);
      case 2: return dispatch_split2<VCAdvection2>(dtype, rp, q_in, gin, q_out, gout, comp_in,
                                                   comp_out, work, cwork, lo, hi, refl, ng,
                                                   method, mthlim, dx, dy, dt, aux, cfl, s);
      case 3: return dispatch_split2<VCAdvection3>(dtype, rp, q_in, gin, q_out, gout, comp_in,
                                                   comp_out, work, cwork, lo, hi, refl, ng,
                                                   method, mthlim, dx, dy, dt, aux, cfl, s);
      case 4: return dispatch_split2<VCAdvection4>(dtype, rp, q_in, gin, q_out, gout, comp_in,
                                                   comp_out, work, cwork, lo, hi, refl, ng,
                                                   method, mthlim, dx, dy, dt, aux, cfl, s);
    }
    return CLAWHIP_E_ARG;
  }
  if (rp.kind == CLAWHIP_RP_VC_ACOUSTICS) {
    if (!aux.p) return CLAWHIP_E_ARG;
    return dispatch_split2<VCAcoustics>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                        work, cwork, lo, hi, refl, ng, method, mthlim, dx, dy,
                                        dt, aux, cfl, s);
  }
  if (rp.kind == CLAWHIP_RP_SHALLOW_FWAVE) {
    if (!aux.p) return CLAWHIP_E_ARG;
    return dispatch_split2<ShallowFwave>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                         work, cwork, lo, hi, refl, ng, method, mthlim, dx, dy,
                                         dt, aux, cfl, s);
  }
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS:
      return dispatch_split2<Acoustics>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                        work, cwork, lo, hi, refl, ng, method, mthlim, dx, dy,
                                        dt, aux, cfl, s);
    case CLAWHIP_RP_EULER:
      return dispatch_split2<Euler>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, work,
                                    cwork, lo, hi, refl, ng, method, mthlim, dx, dy, dt, aux, cfl,
                                    s);
    case CLAWHIP_RP_SHALLOW:
      return dispatch_split2<Shallow>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, work,
                                      cwork, lo, hi, refl, ng, method, mthlim, dx, dy, dt, aux,
                                      cfl, s);
    case CLAWHIP_RP_ADVECTION:
      return dispatch_split2<Advection>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                        work, cwork, lo, hi, refl, ng, method, mthlim, dx, dy,
                                        dt, aux, cfl, s);
    case CLAWHIP_RP_SHALLOW_HLLE:
      return dispatch_split2<ShallowHLLE>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                          work, cwork, lo, hi, refl, ng, method, mthlim, dx, dy,
                                          dt, aux, cfl, s);
    case CLAWHIP_RP_EULER5:
      return dispatch_split2<Euler5>(dtype, rp, q_in, gin, q_out, gout, comp_in, comp_out, work,
                                     cwork, lo, hi, refl, ng, method, mthlim, dx, dy, dt, aux,
                                     cfl, s);
  }
  return CLAWHIP_E_KIND;
}

int launch_split3(const clawhip_rp& rp, int dtype, const void* q_in, const Geom& gin,
                  void* q_out, const Geom& gout, const void* comp_in, void* comp_out,
                  void* work, void* cwork, const int* lo, const int* hi, const int* refl,
                  int ng, const int* method, const int* mthlim, double dx, double dy,
                  double dz, double dt, unsigned long long* cfl, hipStream_t s) {
  if (rp.kind != CLAWHIP_RP_EULER3) return CLAWHIP_E_KIND;
  if (dtype != CLAWHIP_F32K)
    return run_split3<Euler3<double>, double, false>(rp, q_in, gin, q_out, gout, comp_in,
                                                     comp_out, work, cwork, lo, hi, refl, ng,
                                                     method, mthlim, dx, dy, dz, dt, cfl, s);
#if CLAW_FAST
  return run_split3<Euler3<float>, float, true>(rp, q_in, gin, q_out, gout, comp_in, comp_out,
                                                work, cwork, lo, hi, refl, ng, method, mthlim,
                                                dx, dy, dz, dt, cfl, s);
#else
  // FP32 + Kahan is dispatched to the fast variant only
  return CLAWHIP_E_ARG;
#endif
}

int launch_sweep2(const clawhip_rp& rp, int ids, const void* q_in, const Geom& gin, void* q_out,
                  const Geom& gout, int ng, const int* method, const int* mthlim, double dtdx,
                  const AuxIn& aux, unsigned long long* cfl, hipStream_t s) {
  if (rp.kind == CLAWHIP_RP_VC_ADVECTION || rp.kind == CLAWHIP_RP_VC_ADVECTION_2D) {
    if (!aux.p) return CLAWHIP_E_ARG;
    switch (rp.kind == CLAWHIP_RP_VC_ADVECTION_2D ? 1 : rp.meqn) {
      case 1: return run_sweep<VCAdvection1<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng,
                                                     method, mthlim, dtdx, aux, cfl, s);
      case 2: return run_sweep<VCAdvection2<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng,
                                                     method, mthlim, dtdx, aux, cfl, s);
      case 3: return run_sweep<VCAdvection3<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng,
                                                     method, mthlim, dtdx, aux, cfl, s);
      case 4: return run_sweep<VCAdvection4<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng,
                                                     method, mthlim, dtdx, aux, cfl, s);
    }
    return CLAWHIP_E_ARG;
  }
  if (rp.kind == CLAWHIP_RP_VC_ACOUSTICS) {
    if (!aux.p) return CLAWHIP_E_ARG;
    return run_sweep<VCAcoustics<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method, mthlim,
                                          dtdx, aux, cfl, s);
  }
  if (rp.kind == CLAWHIP_RP_SHALLOW_FWAVE) {
    if (!aux.p) return CLAWHIP_E_ARG;
    return run_sweep<ShallowFwave<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method,
                                           mthlim, dtdx, aux, cfl, s);
  }
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS:
      return run_sweep<Acoustics<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method, mthlim,
                                          dtdx, aux, cfl, s);
    case CLAWHIP_RP_EULER:
      return run_sweep<Euler<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method, mthlim,
                                      dtdx, aux, cfl, s);
    case CLAWHIP_RP_SHALLOW:
      return run_sweep<Shallow<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method, mthlim,
                                        dtdx, aux, cfl, s);
    case CLAWHIP_RP_ADVECTION:
      return run_sweep<Advection<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method, mthlim,
                                          dtdx, aux, cfl, s);
    case CLAWHIP_RP_SHALLOW_HLLE:
      return run_sweep<ShallowHLLE<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method,
                                            mthlim, dtdx, aux, cfl, s);
    case CLAWHIP_RP_EULER5:
      return run_sweep<Euler5<double>>(rp, 2, ids, q_in, gin, q_out, gout, ng, method, mthlim,
                                       dtdx, aux, cfl, s);
  }
  return CLAWHIP_E_KIND;
}

int launch_sweep3(const clawhip_rp& rp, int idir, const void* q_in, const Geom& gin, void* q_out,
                  const Geom& gout, int ng, const int* method, const int* mthlim, double dtdx,
                  unsigned long long* cfl, hipStream_t s) {
  if (rp.kind != CLAWHIP_RP_EULER3) return CLAWHIP_E_KIND;
  const AuxIn none;
  return run_sweep<Euler3<double>>(rp, 3, idir, q_in, gin, q_out, gout, ng, method, mthlim, dtdx,
                                   none, cfl, s);
}

}  // namespace CLAW_VNS
}  // namespace claw
