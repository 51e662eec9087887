// usweep.hpp -- one sweep of the 2D classic unsplit step (the x-slices or the
// y-slices of classic/step2.f90 with flux2.f90, limiter.f90, philim.f90), shared
// by the single-grid kernels (classic2.hip) and the ensemble kernels
// (ensemble.hip).  See classic2.hip for the decomposition and the summation order.
#pragma once
#include "common.hpp"

namespace claw {

template <typename T>
struct USweepArgs {
  RPParams<T> rp;
  Geom gin, gout;        // q_in geometry (any stored ghosts) / compact geometry
  BCSpec bc;
  int ng, order, trans;
  int mthlim[4];
  T dn;                  // dt / (spacing along the slice): fluxes, fluctuations
  T dtr;                 // dt / (spacing across the slice): transverse terms
  int H, nstrips, nrb;
  const T* qin;          // state the Riemann problems see
  const T* win;          // AX = 2: sums after the x-sweep (compact)
  T* wout;               // AX = 1: sums after the x-sweep; AX = 2: q_new
  const T* comp_in;      // AX = 2, KAHAN: compensation of q (compact)
  T* comp_out;
  unsigned long long* cfl_bits;  // FP64 max CFL (bit pattern)
};

// Per-slice result of the slice pass for the cell owned by this lane.
template <int ME, typename T>
struct SliceRow {
  T qadd[ME], Fl[ME], Fr[ME], g1[ME], g2[ME];
};

// One slice of flux2.f90 (ixy = IXY) for the cell owned by the lane: the
// Riemann problem at the cell's left interface is solved by this lane, its
// right interface by lane+1.  Qc = state of this lane's cell.
template <class RS, int IXY, typename T>
__device__ __forceinline__ void slice_pass(const USweepArgs<T>& a, const T (&Qc)[RS::meqn],
                                           bool cfl_ok, T& smax, SliceRow<RS::meqn, T>& o) {
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  T QL[ME];
  from_left(Qc, QL);
  T W[MW][ME], s[MW], am[ME], ap[ME];
  typename RS::Roe roe;
  RS::template rpn<IXY>(a.rp, QL, Qc, W, s, am, ap, roe);
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
    second_order<ME, MW>(W, s, a.dn, cq, F);
  }
  T amn[ME];
  from_right(am, amn);
  from_right(F, o.Fr);
#pragma unroll
  for (int m = 0; m < ME; m++) {
    o.Fl[m] = F[m];
    o.qadd[m] = (T(0) - a.dn * ap[m]) - a.dn * amn[m];
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
    RS::template rpt<IXY>(a.rp, roe, amm, bmA, bpA);
    RS::template rpt<IXY>(a.rp, roe, apm, bmP, bpP);
    T bmAn[ME], bpAn[ME];
    from_right(bmA, bmAn);
    from_right(bpA, bpAn);
    const T h = T(0.5) * a.dn;
#pragma unroll
    for (int m = 0; m < ME; m++) {
      o.g1[m] = (T(0) - h * bmAn[m]) - h * bmP[m];
      o.g2[m] = (T(0) - h * bpAn[m]) - h * bpP[m];
    }
  }
}

// `wid`: this wavefront's index among the a.nstrips * a.nrb of the sweep.
template <class RS, typename T, bool KAHAN, int AX>
__device__ __forceinline__ void usweep_body(const USweepArgs<T>& a, int wid) {
  constexpr int ME = RS::meqn;
  const int lane = threadIdx.x & 63;
  if (wid >= a.nstrips * a.nrb) return;   // uniform per wavefront
  const int strip = wid % a.nstrips, rb = wid / a.nstrips;
  const int mx = a.gout.mx, my = a.gout.my;
  const int ms = AX == 1 ? mx : my;       // cells along a slice
  const int mm = AX == 1 ? my : mx;       // slices (march direction)
  const int p = 1 + strip * 60 - 2 + lane;
  const int s0 = 1 + rb * a.H;
  const int send = min(s0 + a.H, mm + 1);
  const bool out_lane = lane >= 2 && lane < 62 && p >= 1 && p <= ms;
  const bool cfl_lane = lane >= 1 && p >= 1 && p <= ms + 1;   // lane 0 has no left state
  const int pc = p < 1 ? 1 : (p > ms ? ms : p);               // clamped, for compact loads
  const T dn = a.dn, dtr = a.dtr;
  T smax = T(0);

  auto load_q = [&](int s, T (&o)[ME]) {
    if (AX == 1) load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, p, s, o);
    else load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, s, p, o);
  };
  // value each cell's running sum starts from: q (FP64 x-sweep), 0 (FP32 x-sweep),
  // the x-sweep result (y-sweep)
  auto start = [&](int s, const T (&q)[ME], T (&o)[ME]) {
    if (AX == 1) {
#pragma unroll
      for (int m = 0; m < ME; m++) o[m] = KAHAN ? T(0) : q[m];
    } else {
      const int sc = s < 1 ? 1 : (s > mm ? mm : s);
      const T* pw = cell_ptr2(a.win, a.gout, sc, pc);
#pragma unroll
      for (int m = 0; m < ME; m++) o[m] = pw[(long long)m * a.gout.fs];
    }
  };

  SliceRow<ME, T> sr;
  T Qc[ME], Qn[ME], Qp[ME], st[ME];
  // ---- slice s0-1: its update of cell s0 opens that cell's sum
  load_q(s0 - 1, Qc);
  slice_pass<RS, AX>(a, Qc, cfl_lane && s0 - 1 >= 0, smax, sr);
  load_q(s0, Qn);
  T acc_prev[ME], acc_cur[ME];
  start(s0, Qn, st);
#pragma unroll
  for (int m = 0; m < ME; m++) {
    acc_cur[m] = st[m] + dtr * sr.g2[m];
    acc_prev[m] = T(0);
    Qc[m] = Qn[m];
    Qp[m] = Qn[m];
  }

  for (int t = s0; t <= send; t++) {
    load_q(t + 1, Qn);
    // ---- slice t: closes cell t-1, updates cell t, opens cell t+1
    slice_pass<RS, AX>(a, Qc, cfl_lane && t <= mm + 1, smax, sr);
    start(t + 1, Qn, st);
    T acc_next[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) {
      acc_prev[m] = acc_prev[m] - dtr * sr.g1[m];
      if (AX == 1)
        acc_cur[m] = ((acc_cur[m] + sr.qadd[m]) - dn * (sr.Fr[m] - sr.Fl[m])) -
                     dtr * (sr.g2[m] - sr.g1[m]);
      else
        acc_cur[m] = acc_cur[m] + ((sr.qadd[m] - dn * (sr.Fr[m] - sr.Fl[m])) -
                                   dtr * (sr.g2[m] - sr.g1[m]));
      acc_next[m] = st[m] + dtr * sr.g2[m];
    }
    if (t > s0 && out_lane) {
      // cell t-1 is complete
      const int c = AX == 1 ? p : t - 1, r = AX == 1 ? t - 1 : p;
      T* po = cell_ptr2(a.wout, a.gout, c, r);
      if (AX == 2 && KAHAN) {
        const T* pci = cell_ptr2(a.comp_in, a.gout, c, r);
        T* pco = cell_ptr2(a.comp_out, a.gout, c, r);
#pragma unroll
        for (int m = 0; m < ME; m++) {
          T y = acc_prev[m] - pci[(long long)m * a.gout.fs];
          T tq = Qp[m] + y;
          pco[(long long)m * a.gout.fs] = (tq - Qp[m]) - y;
          po[(long long)m * a.gout.fs] = tq;
        }
      } else {
#pragma unroll
        for (int m = 0; m < ME; m++) po[(long long)m * a.gout.fs] = acc_prev[m];
      }
    }
#pragma unroll
    for (int m = 0; m < ME; m++) {
      acc_prev[m] = acc_cur[m];
      acc_cur[m] = acc_next[m];
      Qp[m] = Qc[m];
      Qc[m] = Qn[m];
    }
  }
  wave_atomic_max((double)dn * (double)smax, a.cfl_bits);
}

// The y-sweep with lanes across the slices: lane l holds column c of a strip of
// 64 columns and marches down the rows, keeping the y-interface just solved in
// registers (its waves for the limiter of the next one, its fluctuations for
// the cell above).  Loads and stores stay coalesced along x; the transverse
// terms of columns c-1 and c+1 arrive by cross-lane moves.  Each cell's sum
// continues from the x-sweep result in `win` (step2.f90:221-228).
template <class RS, typename T, bool KAHAN>
__device__ __forceinline__ void ywin_body(const USweepArgs<T>& a, int wid) {
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
  const bool cfly_col = c >= 0 && c <= mx + 1;
  const int cc = c < 1 ? 1 : (c > mx ? mx : c);   // clamped, for compact loads
  const T dtdy = a.dn, dtdx = a.dtr, hy = T(0.5) * a.dn;
  T smy = T(0);

  T Qm[ME], Q0[ME], Q1[ME];
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0 - 2, Qm);
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0 - 1, Q0);
  load_cell2<ME>(a.qin, a.gin, a.bc, a.ng, c, j0, Q1);
  T qold_prev[ME], qold_cur[ME];
#pragma unroll
  for (int m = 0; m < ME; m++) { qold_cur[m] = Q1[m]; qold_prev[m] = T(0); }
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
      second_order<ME, MW>(Wt, st, dtdy, cq, Fy);
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
      const T* pw = cell_ptr2(a.win, a.gout, cc, t - 1);
      T upd[ME];
#pragma unroll
      for (int m = 0; m < ME; m++) {
        T qaddy = qaddy_p[m] - dtdy * amt[m];
        T v = pw[(long long)m * a.gout.fs] + dtdx * g2l[m];
        v = v + ((qaddy - dtdy * (Fy[m] - Fy_p[m])) - dtdx * (g2y[m] - g1y[m]));
        upd[m] = v - dtdx * g1r[m];
      }
      if (out_col) {
        T* po = cell_ptr2(a.wout, a.gout, c, t - 1);
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
  wave_atomic_max((double)a.dn * (double)smy, a.cfl_bits);
}

}  // namespace claw
