// unsplit.hpp -- the 2D classic unsplit step for one wavefront's share of the
// grid (classic/step2.f90 + flux2.f90 + limiter.f90 + philim.f90 with the ghost
// fill of solver.py:351-519 applied as an index map on load), shared by the
// single-grid kernels (classic2.hip) and the ensemble kernels (ensemble.hip).
//
// Work decomposition (wave64):
//   * a wavefront owns a strip of 64 columns c0-2 .. c0+61 and writes the 60
//     columns c0 .. c0+59; lanes 0,1,62,63 only supply neighbour data.
//   * it marches down rows j0 .. j0+H-1.  x-interfaces of a row are computed
//     lane-parallel (neighbour states, waves and fluctuations arrive by
//     cross-lane moves); y-interfaces are lane-local, kept in a register
//     window so each one is solved once.
//   * every cell update applies the six contributions in the reference's order:
//     x-slice j-1 (+dtdy*g2), x-slice j (main), x-slice j+1 (-dtdy*g1), then
//     y-slice i-1, y-slice i, y-slice i+1 (step2.f90:136-145, 218-228).
// FP64: bitwise equal to the reference step (built with -ffp-contract=off).
// FP32 (KAHAN): the same increment summed in FP32 and applied with a
// Kahan-compensated update (compensation array alongside q).
//
// AUX: the solver's aux fields (up to 3) ride along in a row window like q's;
// the Riemann solvers get the aux of the cells they involve (step2.f90:107-115,
// 189-197: aux1/aux2/aux3 = the previous, this and the next slice).
// CAPA: one aux field is a capacity kappa; every per-cell dtdx becomes
// dtdx/kappa of that cell (step2.f90:101-105, 183-187; flux2.f90:105-119,
// 134-146, 181-203) and each update divides by the target cell's kappa in the
// reference's order (step2.f90:152-163, 232-244).
#pragma once
#include "common.hpp"

namespace claw {

template <typename T>
struct UnsplitArgs {
  RPParams<T> rp;
  Geom gin, gout;        // input (any ghosts) / output (compact) geometry
  BCSpec bc;
  int ng, order, trans;
  int mthlim[kMaxWaves];
  T dtdx, dtdy;
  int H, nstrips, nrb;
  const T* qin;
  T* qout;
  const T* comp_in;              // Kahan compensation (FP32 path), output geometry
  T* comp_out;
  unsigned long long* cfl_bits;  // FP64 max CFL (bit pattern)
  const T* aux;                  // AUX kernels: naux fields, geometry gaux, codes bcaux
  Geom gaux;
  BCSpec bcaux;
  int naux, capa;                // capa: the capacity field (CAPA kernels)
};

constexpr int UNSPLIT_MAX_AUX = 3;

// aux fields of cell (c, r) through the aux boundary codes (never negated).
template <bool AUX, typename T>
__device__ __forceinline__ void load_aux(const UnsplitArgs<T>& a, int c, int r,
                                         T (&A)[UNSPLIT_MAX_AUX]) {
  if constexpr (AUX) {
    bool fx = false, fy = false;
    c = bc_map(c, a.gaux.mx, a.ng, a.bcaux.lo[0], a.bcaux.hi[0], fx);
    r = bc_map(r, a.gaux.my, a.ng, a.bcaux.lo[1], a.bcaux.hi[1], fy);
    const T* p = a.aux + (long long)(c - 1 + a.gaux.g) + (long long)(r - 1 + a.gaux.g) * a.gaux.rs;
#pragma unroll
    for (int k = 0; k < UNSPLIT_MAX_AUX; k++) A[k] = k < a.naux ? p[(long long)k * a.gaux.fs] : T(1);
  } else {
    (void)a; (void)c; (void)r;
#pragma unroll
    for (int k = 0; k < UNSPLIT_MAX_AUX; k++) A[k] = T(0);
  }
}

// The capacity of a cell from its aux values.
template <typename T>
__device__ __forceinline__ T capa_of(const UnsplitArgs<T>& a, const T (&A)[UNSPLIT_MAX_AUX]) {
  T r = A[0];
#pragma unroll
  for (int k = 1; k < UNSPLIT_MAX_AUX; k++) r = k == a.capa ? A[k] : r;
  return r;
}

// Per-row result of the x-pass for the cell owned by this lane.
template <int ME, typename T>
struct XRow {
  T qadd[ME], Fl[ME], Fr[ME], g1[ME], g2[ME];
};

// x-interfaces of one row (flux2.f90 with ixy = 1) for the cell owned by the lane.
// Qc = state of this lane's cell; the left neighbour comes from lane-1.
// Am, A0, A1: aux of this lane's cell in rows t-1, t, t+1 (AUX kernels).
// smax collects max |s| (dtdx applied at the end), or with CAPA the Courant
// numbers themselves.
template <class RS, bool AUX, bool CAPA, typename T>
__device__ __forceinline__ void xpass(const UnsplitArgs<T>& a, const T (&Qc)[RS::meqn],
                                      const T (&Am)[UNSPLIT_MAX_AUX],
                                      const T (&A0)[UNSPLIT_MAX_AUX],
                                      const T (&A1)[UNSPLIT_MAX_AUX], bool cfl_ok, T& smax,
                                      XRow<RS::meqn, T>& o) {
  constexpr int ME = RS::meqn, MW = RS::mwaves, NA = UNSPLIT_MAX_AUX;
  T QL[ME];
  from_left(Qc, QL);
  T AL[NA];
  if constexpr (AUX) from_left(A0, AL);
  // dtdx1d of this cell and of its left neighbour
  T dtdx_i = a.dtdx, dtdx_l = a.dtdx;
  if constexpr (CAPA) {
    dtdx_i = a.dtdx / capa_of(a, A0);
    dtdx_l = from_left(dtdx_i);
  }
  T W[MW][ME], s[MW], am[ME], ap[ME];
  typename RS::Roe roe;
  rpn_call<RS, 1>(a.rp, QL, Qc, AL, A0, W, s, am, ap, roe);
  if (cfl_ok) {
#pragma unroll
    for (int w = 0; w < MW; w++) {
      if constexpr (CAPA) smax = cfl_term(cfl_term(smax, dtdx_i * s[w]), -(dtdx_l * s[w]));
      else smax = cfl_term(smax, fabs(s[w]));
    }
  }
  T F[ME], cq[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) { F[m] = T(0); cq[m] = T(0); }
  if (a.order != 1) {
    T Wn[MW][ME];
#pragma unroll
    for (int w = 0; w < MW; w++) from_right(W[w], Wn[w]);
    T dotr[MW], dotl[MW];
    wave_dots<ME, MW>(W, Wn, dotr);
    from_left(dotr, dotl);
    limit_waves<ME, MW>(W, s, dotl, dotr, a.mthlim);
    if constexpr (CAPA) second_order_ave<ME, MW>(W, s, T(0.5) * (dtdx_l + dtdx_i), cq, F, a.rp.fwave);
    else second_order<ME, MW>(W, s, a.dtdx, cq, F, a.rp.fwave);
  }
  T amn[ME];
  from_right(am, amn);
  from_right(F, o.Fr);
#pragma unroll
  for (int m = 0; m < ME; m++) {
    o.Fl[m] = F[m];
    o.qadd[m] = (T(0) - dtdx_i * ap[m]) - dtdx_i * amn[m];
    o.g1[m] = T(0);
    o.g2[m] = T(0);
  }
  if (a.trans > 0) {
    T amm[ME], apm[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) {
      bool cor = a.order > 1 && a.trans == 2;
      amm[m] = cor ? am[m] + cq[m] : am[m];
      apm[m] = cor ? ap[m] - cq[m] : ap[m];
    }
    T bmA[ME], bpA[ME], bmP[ME], bpP[ME];
    // imp = 1 goes to cell i-1 (its aux from the left lane), imp = 2 to cell i
    T AmL[NA], A1L[NA];
    if constexpr (AUX) {
      from_left(Am, AmL);
      from_left(A1, A1L);
    }
    rpt_call<RS, 1>(a.rp, roe, AmL, AL, A1L, amm, bmA, bpA);
    rpt_call<RS, 1>(a.rp, roe, Am, A0, A1, apm, bmP, bpP);
    T bmAn[ME], bpAn[ME];
    from_right(bmA, bmAn);
    from_right(bpA, bpAn);
    const T h = T(0.5) * dtdx_i;
#pragma unroll
    for (int m = 0; m < ME; m++) {
      o.g1[m] = (T(0) - h * bmAn[m]) - h * bmP[m];
      o.g2[m] = (T(0) - h * bpAn[m]) - h * bpP[m];
    }
  }
}

// dtdy1d of a cell with aux A (with CAPA: dtdy / kappa).
template <bool CAPA, typename T>
__device__ __forceinline__ T dy1d(const UnsplitArgs<T>& a, const T (&A)[UNSPLIT_MAX_AUX]) {
  if constexpr (CAPA) return a.dtdy / capa_of(a, A);
  else return a.dtdy;
}

// Courant number of a y-interface between cells with aux Alo (below) and Ahi
// (above): flux2.f90:111-119, s > 0 against the upper cell, s < 0 the lower
// (CAPA); without a capacity acc collects max |s| (dtdy applied at the end).
template <bool CAPA, int MW, typename T>
__device__ __forceinline__ void cfl_y(const UnsplitArgs<T>& a, T& acc, const T (&s)[MW],
                                      const T (&Alo)[UNSPLIT_MAX_AUX],
                                      const T (&Ahi)[UNSPLIT_MAX_AUX]) {
  if constexpr (CAPA) {
    const T dlo = dy1d<CAPA>(a, Alo), dhi = dy1d<CAPA>(a, Ahi);
#pragma unroll
    for (int w = 0; w < MW; w++) acc = dmax(dmax(acc, dhi * s[w]), -(dlo * s[w]));
  } else {
#pragma unroll
    for (int w = 0; w < MW; w++) acc = dmax(acc, fabs(s[w]));
  }
}

// Wavefronts per SIMD a step kernel is compiled for (the register budget is
// 512 / that number), where a solver gains from more than the compiler's own
// choice.  FP64 acoustics (180 VGPRs, 2 per SIMD) and FP32 + Kahan Euler (172
// VGPRs, 2 per SIMD) fit 168 VGPRs with a few spilled dwords:
//   - 4096^2 FP64 acoustics step: 476 -> 431 us;
//   - 1024^2 FP32 + Kahan Euler step: 76.7 -> 71.0 us.
// Forcing FP64 Euler (314 VGPRs) to 3 per SIMD doubles its time (heavy
// spills); FP64 acoustics at 4 per SIMD (37 spilled dwords) takes 907 us.
// 1 = no constraint.  The aux and capacity forms keep the compiler's choice.
// The step kernels (k_unsplit, k_unsplit_ctl, k_ens_step) take it as their
// amdgpu_waves_per_eu attribute, and their march-length choice counts that
// many resident wavefronts per SIMD.
template <class RS, typename T>
struct UnsplitWaves {
  static constexpr int value = 1;
};
template <>
struct UnsplitWaves<Acoustics<double>, double> {
  static constexpr int value = 3;
};
template <>
struct UnsplitWaves<Euler<float>, float> {
  static constexpr int value = 3;
};
template <class RS, typename T, bool AUX = false, bool CAPA = false>
constexpr int unsplit_waves() {
  return (AUX || CAPA) ? 1 : UnsplitWaves<RS, T>::value;
}

// `wid`: this wavefront's index among the a.nstrips * a.nrb of the launch.
// Returns this lane's Courant number (0 past the last strip); the kernel adds it
// to *a.cfl_bits (one atomic per workgroup where every wavefront of the
// workgroup reaches that point, else one per wavefront).
template <class RS, typename T, bool KAHAN, bool AUX = false, bool CAPA = false>
__device__ __forceinline__ double unsplit_body(const UnsplitArgs<T>& a, int wid) {
  constexpr int ME = RS::meqn, MW = RS::mwaves, NA = UNSPLIT_MAX_AUX;
  static_assert(AUX || (RS::maux == 0 && !CAPA), "aux solvers and capacity need the aux window");
  using Roe = typename RS::Roe;
  const int lane = threadIdx.x & 63;
  if (wid >= a.nstrips * a.nrb) return 0.0;   // uniform per wavefront
  const int strip = wid % a.nstrips, rb = wid / a.nstrips;
  const int mx = a.gout.mx, my = a.gout.my;
  const int c = 1 + strip * 60 - 2 + lane;
  const int j0 = 1 + rb * a.H;
  const int jend = min(j0 + a.H, my + 1);
  const bool out_col = lane >= 2 && lane < 62 && c >= 1 && c <= mx;
  const bool cflx_col = lane >= 1 && c >= 1 && c <= mx + 1;   // lane 0 has no left state
  const bool cfly_col = c >= 0 && c <= mx + 1;
  const T dtdx = a.dtdx, dtdy = a.dtdy, hy = T(0.5) * a.dtdy;
  T smx = T(0), smy = T(0);

  T Qm[ME], Q0[ME], Q1[ME];
  T Am[NA], A0[NA], A1[NA];   // aux rows t-1, t, t+1 (AUX)
  XRow<ME, T> xr;
  // ---- warm-up: rows j0-2, j0-1, j0; x-pass of row j0-1; y-interfaces j0-1, j0
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0 - 2, Qm);
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0 - 1, Q0);
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0, Q1);
  load_aux<AUX>(a, c, j0 - 2, Am);
  load_aux<AUX>(a, c, j0 - 1, A0);
  load_aux<AUX>(a, c, j0, A1);
  xpass<RS, AUX, CAPA>(a, Q0, Am, A0, A1, cflx_col && j0 - 1 >= 0, smx, xr);
  T acc_prev[ME], acc_cur[ME], qold_prev[ME], qold_cur[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) {
    if constexpr (CAPA) acc_cur[m] = (KAHAN ? T(0) : Q1[m]) + (dtdy * xr.g2[m]) / capa_of(a, A1);
    else acc_cur[m] = (KAHAN ? T(0) : Q1[m]) + dtdy * xr.g2[m];
    qold_cur[m] = Q1[m];
    acc_prev[m] = T(0);
    qold_prev[m] = T(0);
  }
  T Wt[MW][ME], st[MW], amt[ME], apt[ME], dotl[MW];
  Roe roet;
  {
    T Wp[MW][ME], sp[MW], amp[ME], app[ME];
    Roe roep;
    rpn_call<RS, 2>(a.rp, Qm, Q0, Am, A0, Wp, sp, amp, app, roep);
    if (cfly_col && j0 - 1 >= 1) cfl_y<CAPA>(a, smy, sp, Am, A0);
    rpn_call<RS, 2>(a.rp, Q0, Q1, A0, A1, Wt, st, amt, apt, roet);
    if (cfly_col && j0 <= my + 1) cfl_y<CAPA>(a, smy, st, A0, A1);
    wave_dots<ME, MW>(Wp, Wt, dotl);
  }
#pragma unroll
  for (int m = 0; m < ME; m++) Q0[m] = Q1[m];
#pragma unroll
  for (int k = 0; k < NA; k++) { Am[k] = A0[k]; A0[k] = A1[k]; }
  T qaddy_p[ME], hbm_p[ME], hbp_p[ME], Fy_p[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) { qaddy_p[m] = T(0); hbm_p[m] = T(0); hbp_p[m] = T(0); Fy_p[m] = T(0); }

  for (int t = j0; t <= jend; t++) {
    load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, t + 1, Q1);
    load_aux<AUX>(a, c, t + 1, A1);
    // ---- x-slice t
    xpass<RS, AUX, CAPA>(a, Q0, Am, A0, A1, cflx_col && t <= my + 1, smx, xr);
    T acc_next[ME];
    if constexpr (CAPA) {
      const T km = capa_of(a, Am), k0 = capa_of(a, A0), k1 = capa_of(a, A1);
#pragma unroll
      for (int m = 0; m < ME; m++) {
        acc_prev[m] = acc_prev[m] - (dtdy * xr.g1[m]) / km;
        acc_cur[m] = (acc_cur[m] + xr.qadd[m]) -
                     (dtdx * (xr.Fr[m] - xr.Fl[m]) + dtdy * (xr.g2[m] - xr.g1[m])) / k0;
        acc_next[m] = (KAHAN ? T(0) : Q1[m]) + (dtdy * xr.g2[m]) / k1;
      }
    } else {
#pragma unroll
      for (int m = 0; m < ME; m++) {
        acc_prev[m] = acc_prev[m] - dtdy * xr.g1[m];
        acc_cur[m] = ((acc_cur[m] + xr.qadd[m]) - dtdx * (xr.Fr[m] - xr.Fl[m])) - dtdy * (xr.g2[m] - xr.g1[m]);
        acc_next[m] = (KAHAN ? T(0) : Q1[m]) + dtdy * xr.g2[m];
      }
    }
    // ---- y-interface t+1 (rows t, t+1); limit interface t
    T Wn[MW][ME], sn[MW], amn[ME], apn[ME], dotn[MW];
    Roe roen;
    rpn_call<RS, 2>(a.rp, Q0, Q1, A0, A1, Wn, sn, amn, apn, roen);
    if (cfly_col && t + 1 <= my + 1) cfl_y<CAPA>(a, smy, sn, A0, A1);
    // dtdy1d of rows t-1 (imp = 1 side of interface t) and t (imp = 2 side)
    const T dyA = dy1d<CAPA>(a, Am), dyP = dy1d<CAPA>(a, A0);
    T Fy[ME], cq[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) { Fy[m] = T(0); cq[m] = T(0); }
    if (a.order != 1) {
      wave_dots<ME, MW>(Wt, Wn, dotn);
      limit_waves<ME, MW>(Wt, st, dotl, dotn, a.mthlim);
      if constexpr (CAPA) second_order_ave<ME, MW>(Wt, st, T(0.5) * (dyA + dyP), cq, Fy, a.rp.fwave);
      else second_order<ME, MW>(Wt, st, a.dtdy, cq, Fy, a.rp.fwave);
    }
    T bmA[ME], bpA[ME], bmP[ME], bpP[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) { bmA[m] = T(0); bpA[m] = T(0); bmP[m] = T(0); bpP[m] = T(0); }
    if (a.trans > 0) {
      T amm[ME], apm[ME];
#pragma unroll
      for (int m = 0; m < ME; m++) {
        bool cor = a.order > 1 && a.trans == 2;
        amm[m] = cor ? amt[m] + cq[m] : amt[m];
        apm[m] = cor ? apt[m] - cq[m] : apt[m];
      }
      // cell i1 = row t-1 (imp = 1) or t (imp = 2); its neighbours across the
      // slice are the lanes either side
      T AmL[NA], AmR[NA], A0L[NA], A0R[NA];
      if constexpr (AUX) {
        from_left(Am, AmL);
        from_right(Am, AmR);
        from_left(A0, A0L);
        from_right(A0, A0R);
      }
      rpt_call<RS, 2>(a.rp, roet, AmL, Am, AmR, amm, bmA, bpA);
      rpt_call<RS, 2>(a.rp, roet, A0L, A0, A0R, apm, bmP, bpP);
    }
    const T hyA = CAPA ? T(0.5) * dyA : hy, hyP = CAPA ? T(0.5) * dyP : hy;
    if (t > j0) {
      // row t-1 is complete: interface t-1 (apdq side, carried) and t (amdq side)
      T g1y[ME], g2y[ME], g2l[ME], g1r[ME];
#pragma unroll
      for (int m = 0; m < ME; m++) {
        g1y[m] = (T(0) - hyA * bmA[m]) - hbm_p[m];
        g2y[m] = (T(0) - hyA * bpA[m]) - hbp_p[m];
      }
      from_left(g2y, g2l);
      from_right(g1y, g1r);
      T upd[ME];
      if constexpr (CAPA) {
        const T km = capa_of(a, Am);
#pragma unroll
        for (int m = 0; m < ME; m++) {
          T qaddy = qaddy_p[m] - dyA * amt[m];
          T v = acc_prev[m] + (dtdx * g2l[m]) / km;
          v = (v + qaddy) - (dtdy * (Fy[m] - Fy_p[m]) + dtdx * (g2y[m] - g1y[m])) / km;
          upd[m] = v - (dtdx * g1r[m]) / km;
        }
      } else {
#pragma unroll
        for (int m = 0; m < ME; m++) {
          T qaddy = qaddy_p[m] - dtdy * amt[m];
          T v = acc_prev[m] + dtdx * g2l[m];
          v = v + ((qaddy - dtdy * (Fy[m] - Fy_p[m])) - dtdx * (g2y[m] - g1y[m]));
          upd[m] = v - dtdx * g1r[m];
        }
      }
      if (out_col) {
        T* po = cell_ptr2(a.qout, a.gout, c, t - 1);
        if (KAHAN) {
          const T* pci = cell_ptr2(a.comp_in, a.gout, c, t - 1);
          T* pco = cell_ptr2(a.comp_out, a.gout, c, t - 1);
#pragma unroll
          for (int m = 0; m < ME; m++) {
            T y = upd[m] - pci[(long long)m * a.gout.fs];
            T tq = qold_prev[m] + y;
            pco[(long long)m * a.gout.fs] = (tq - qold_prev[m]) - y;
            po[(long long)m * a.gout.fs] = tq;
          }
        } else {
#pragma unroll
          for (int m = 0; m < ME; m++) po[(long long)m * a.gout.fs] = upd[m];
        }
      }
    }
    // ---- carry: row t partials from interface t (apdq side); shift windows
#pragma unroll
    for (int m = 0; m < ME; m++) {
      qaddy_p[m] = T(0) - dyP * apt[m];
      hbm_p[m] = hyP * bmP[m];
      hbp_p[m] = hyP * bpP[m];
      Fy_p[m] = Fy[m];
      amt[m] = amn[m];
      apt[m] = apn[m];
      acc_prev[m] = acc_cur[m];
      acc_cur[m] = acc_next[m];
      qold_prev[m] = qold_cur[m];
      qold_cur[m] = Q1[m];
      Q0[m] = Q1[m];
    }
#pragma unroll
    for (int k = 0; k < NA; k++) { Am[k] = A0[k]; A0[k] = A1[k]; }
#pragma unroll
    for (int w = 0; w < MW; w++) {
      st[w] = sn[w];
      dotl[w] = dotn[w];
#pragma unroll
      for (int m = 0; m < ME; m++) Wt[w][m] = Wn[w][m];
    }
    roet = roen;
  }
  double cfl, cy;
  if constexpr (CAPA) {   // smx, smy hold Courant numbers
    cfl = (double)smx;
    cy = (double)smy;
  } else {
    cfl = (double)a.dtdx * (double)smx;
    cy = (double)a.dtdy * (double)smy;
  }
  if (cy > cfl) cfl = cy;
  return cfl;
}

}  // namespace claw
