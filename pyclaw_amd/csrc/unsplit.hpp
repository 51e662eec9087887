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
#pragma once
#include "common.hpp"

namespace claw {

template <typename T>
struct UnsplitArgs {
  RPParams<T> rp;
  Geom gin, gout;        // input (any ghosts) / output (compact) geometry
  BCSpec bc;
  int ng, order, trans;
  int mthlim[4];
  T dtdx, dtdy;
  int H, nstrips, nrb;
  const T* qin;
  T* qout;
  const T* comp_in;              // Kahan compensation (FP32 path), output geometry
  T* comp_out;
  unsigned long long* cfl_bits;  // FP64 max CFL (bit pattern)
};

// Per-row result of the x-pass for the cell owned by this lane.
template <int ME, typename T>
struct XRow {
  T qadd[ME], Fl[ME], Fr[ME], g1[ME], g2[ME];
};

// x-interfaces of one row (flux2.f90 with ixy = 1) for the cell owned by the lane.
// Qc = state of this lane's cell; the left neighbour comes from lane-1.
template <class RS, typename T>
__device__ __forceinline__ void xpass(const UnsplitArgs<T>& a, const T (&Qc)[RS::meqn],
                                      bool cfl_ok, T& smax, XRow<RS::meqn, T>& o) {
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  T QL[ME];
  from_left(Qc, QL);
  T W[MW][ME], s[MW], am[ME], ap[ME];
  typename RS::Roe roe;
  RS::template rpn<1>(a.rp, QL, Qc, W, s, am, ap, roe);
  if (cfl_ok) {
#pragma unroll
    for (int w = 0; w < MW; w++) smax = dmax(smax, fabs(s[w]));
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
    second_order<ME, MW>(W, s, a.dtdx, cq, F);
  }
  T amn[ME];
  from_right(am, amn);
  from_right(F, o.Fr);
#pragma unroll
  for (int m = 0; m < ME; m++) {
    o.Fl[m] = F[m];
    o.qadd[m] = (T(0) - a.dtdx * ap[m]) - a.dtdx * amn[m];
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
    RS::template rpt<1>(a.rp, roe, amm, bmA, bpA);
    RS::template rpt<1>(a.rp, roe, apm, bmP, bpP);
    T bmAn[ME], bpAn[ME];
    from_right(bmA, bmAn);
    from_right(bpA, bpAn);
    const T h = T(0.5) * a.dtdx;
#pragma unroll
    for (int m = 0; m < ME; m++) {
      o.g1[m] = (T(0) - h * bmAn[m]) - h * bmP[m];
      o.g2[m] = (T(0) - h * bpAn[m]) - h * bpP[m];
    }
  }
}

// `wid`: this wavefront's index among the a.nstrips * a.nrb of the launch.
template <class RS, typename T, bool KAHAN>
__device__ __forceinline__ void unsplit_body(const UnsplitArgs<T>& a, int wid) {
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  using Roe = typename RS::Roe;
  const int lane = threadIdx.x & 63;
  if (wid >= a.nstrips * a.nrb) return;   // uniform per wavefront
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
  XRow<ME, T> xr;
  // ---- warm-up: rows j0-2, j0-1, j0; x-pass of row j0-1; y-interfaces j0-1, j0
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0 - 2, Qm);
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0 - 1, Q0);
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0, Q1);
  xpass<RS>(a, Q0, cflx_col && j0 - 1 >= 0, smx, xr);
  T acc_prev[ME], acc_cur[ME], qold_prev[ME], qold_cur[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) {
    acc_cur[m] = (KAHAN ? T(0) : Q1[m]) + dtdy * xr.g2[m];
    qold_cur[m] = Q1[m];
    acc_prev[m] = T(0);
    qold_prev[m] = T(0);
  }
  T Wt[MW][ME], st[MW], amt[ME], apt[ME], dotl[MW];
  Roe roet;
  {
    T Wp[MW][ME], sp[MW], amp[ME], app[ME];
    Roe roep;
    RS::template rpn<2>(a.rp, Qm, Q0, Wp, sp, amp, app, roep);
    if (cfly_col && j0 - 1 >= 1) {
#pragma unroll
      for (int w = 0; w < MW; w++) smy = dmax(smy, fabs(sp[w]));
    }
    RS::template rpn<2>(a.rp, Q0, Q1, Wt, st, amt, apt, roet);
    if (cfly_col && j0 <= my + 1) {
#pragma unroll
      for (int w = 0; w < MW; w++) smy = dmax(smy, fabs(st[w]));
    }
    wave_dots<ME, MW>(Wp, Wt, dotl);
  }
#pragma unroll
  for (int m = 0; m < ME; m++) Q0[m] = Q1[m];
  T qaddy_p[ME], hbm_p[ME], hbp_p[ME], Fy_p[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) { qaddy_p[m] = T(0); hbm_p[m] = T(0); hbp_p[m] = T(0); Fy_p[m] = T(0); }

  for (int t = j0; t <= jend; t++) {
    load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, t + 1, Q1);
    // ---- x-slice t
    xpass<RS>(a, Q0, cflx_col && t <= my + 1, smx, xr);
    T acc_next[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) {
      acc_prev[m] = acc_prev[m] - dtdy * xr.g1[m];
      acc_cur[m] = ((acc_cur[m] + xr.qadd[m]) - dtdx * (xr.Fr[m] - xr.Fl[m])) - dtdy * (xr.g2[m] - xr.g1[m]);
      acc_next[m] = (KAHAN ? T(0) : Q1[m]) + dtdy * xr.g2[m];
    }
    // ---- y-interface t+1 (rows t, t+1); limit interface t
    T Wn[MW][ME], sn[MW], amn[ME], apn[ME], dotn[MW];
    Roe roen;
    RS::template rpn<2>(a.rp, Q0, Q1, Wn, sn, amn, apn, roen);
    if (cfly_col && t + 1 <= my + 1) {
#pragma unroll
      for (int w = 0; w < MW; w++) smy = dmax(smy, fabs(sn[w]));
    }
    T Fy[ME], cq[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) { Fy[m] = T(0); cq[m] = T(0); }
    if (a.order != 1) {
      wave_dots<ME, MW>(Wt, Wn, dotn);
      limit_waves<ME, MW>(Wt, st, dotl, dotn, a.mthlim);
      second_order<ME, MW>(Wt, st, a.dtdy, cq, Fy);
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
      RS::template rpt<2>(a.rp, roet, amm, bmA, bpA);
      RS::template rpt<2>(a.rp, roet, apm, bmP, bpP);
    }
    if (t > j0) {
      // row t-1 is complete: interface t-1 (apdq side, carried) and t (amdq side)
      T g1y[ME], g2y[ME], g2l[ME], g1r[ME];
#pragma unroll
      for (int m = 0; m < ME; m++) {
        g1y[m] = (T(0) - hy * bmA[m]) - hbm_p[m];
        g2y[m] = (T(0) - hy * bpA[m]) - hbp_p[m];
      }
      from_left(g2y, g2l);
      from_right(g1y, g1r);
      T upd[ME];
#pragma unroll
      for (int m = 0; m < ME; m++) {
        T qaddy = qaddy_p[m] - dtdy * amt[m];
        T v = acc_prev[m] + dtdx * g2l[m];
        v = v + ((qaddy - dtdy * (Fy[m] - Fy_p[m])) - dtdx * (g2y[m] - g1y[m]));
        upd[m] = v - dtdx * g1r[m];
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
      qaddy_p[m] = T(0) - dtdy * apt[m];
      hbm_p[m] = hy * bmP[m];
      hbp_p[m] = hy * bpP[m];
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
    for (int w = 0; w < MW; w++) {
      st[w] = sn[w];
      dotl[w] = dotn[w];
#pragma unroll
      for (int m = 0; m < ME; m++) Wt[w][m] = Wn[w][m];
    }
    roet = roen;
  }
  double cfl = (double)a.dtdx * (double)smx;
  double cy = (double)a.dtdy * (double)smy;
  if (cy > cfl) cfl = cy;
  wave_atomic_max(cfl, a.cfl_bits);
}

}  // namespace claw
