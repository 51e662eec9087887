// unsplit3.hip -- the unsplit 3D classic step (classic/step3.f90 with
// flux3.f90's transverse and double-transverse terms), FP64.
//
// step3 adds to each cell, in a fixed order, the updates of the 27 slices
// around it: 9 x-slices (k outer, j inner), 9 y-slices (k outer, i inner), 9
// z-slices (j outer, i inner) (step3.f90:120-580).  Every slice reads the old
// state, so the slices are independent; the order only matters for the sums.
// Two kernels per step:
//   k_flux3<DIR> (one per direction): for every slice of that direction, the
//     flux3 results of each cell along it -- qadd, fadd at its left interface,
//     and the transverse modifications gadd(m, 1:2, -1:1) and hadd -- into a
//     work array; one wavefront per 64-cell strip of a slice, lane l at cell
//     c = base + l, neighbours' values by DPP lane shifts as in the split sweeps.
//   k_apply3: one thread per cell, qnew = qold plus the 27 slices' terms in the
//     reference's order and parenthesisation (bitwise equal to the oracle, whose
//     step3 equals the reference's own Fortran on every qbc cell).
// Work: 14 * meqn doubles per cell and direction.  Bound: HBM, by the work
// array's write and its 27-fold gather (about 3.4 KB per cell-step for
// euler_3D against 80 B algorithmic); the path exists for parity with
// ClawSolver3D's unsplit default, the split step is the fast 3D path.
#include <type_traits>

#include "launch.hpp"

namespace claw {

template <typename T>
struct Step3Args {
  RPParams<T> rp;
  Geom g;            // the state's geometry (compact, or with stored ghosts)
  BCSpec bc;
  int ng, order, m3, m4;
  int mthlim[kMaxWaves];
  T dtdx, dtdy, dtdz;
  const T* qin;
  T* qout;           // compact output (k_apply3)
  T* work[3];        // per direction: component-major slice results
  long long wstride[3];   // elements per component
  int nstrips;       // k_flux3
  unsigned long long* cfl_bits;
};

// Extents of direction DIR's slices: normal length n, first transverse
// (gadd's, the next axis in the cycle) a, second (hadd's) b.
template <int DIR>
__device__ __host__ __forceinline__ void dir_extents(const Geom& g, int& n, int& na, int& nb) {
  if (DIR == 1) { n = g.mx; na = g.my; nb = g.mz; }
  else if (DIR == 2) { n = g.my; na = g.mz; nb = g.mx; }
  else { n = g.mz; na = g.mx; nb = g.my; }
}

// Work index of (component, slice (a, b), position p = 1..n+1).
__device__ __forceinline__ long long widx(int comp, long long stride, int a, int b, int p,
                                          int n, int na) {
  return comp * stride + ((long long)b * (na + 2) + a) * (n + 1) + (p - 1);
}

// component slots: 0..ME-1 qadd, ME..2ME-1 fadd, then gadd / hadd as
// 2ME + m + ME * (side - 1 + 2 * (off + 1)), H after the 6 ME of G.
template <int ME>
struct W3 {
  static constexpr int QA = 0, FA = ME, G = 2 * ME, H = 8 * ME, N = 14 * ME;
  __device__ static constexpr int gh(int base, int m, int side, int off) {
    return base + m + ME * ((side - 1) + 2 * (off + 1));
  }
};

template <class RS, int DIR>
__global__ void __launch_bounds__(256) k_flux3(const Step3Args<double> a) {
  using T = double;
  constexpr int ME = RS::meqn, MW = RS::mwaves;
  using WL = W3<ME>;
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int n, na, nb;
  dir_extents<DIR>(a.g, n, na, nb);
  const long long slices = (long long)(na + 2) * (nb + 2);
  // wavefronts past the last slice skip the work but reach the workgroup's
  // Courant-number reduction (a barrier): one atomic per workgroup
  const bool live = wid < a.nstrips * slices;   // uniform per wavefront
  double cfl_wave = 0.0;
  if (live) {
    const int strip = wid % a.nstrips;
    const long long sl = wid / a.nstrips;
    const int sa = (int)(sl % (na + 2)), sb = (int)(sl / (na + 2));   // 0..na+1, 0..nb+1
    const int c = 1 + strip * 60 - 2 + lane;
    const T dtn = DIR == 1 ? a.dtdx : (DIR == 2 ? a.dtdy : a.dtdz);
    // flux3's dtdy / dtdz: the first / second transverse direction's dt/d
    const T dt1 = DIR == 1 ? a.dtdy : (DIR == 2 ? a.dtdz : a.dtdx);
    const T dt2 = DIR == 1 ? a.dtdz : (DIR == 2 ? a.dtdx : a.dtdy);
    int i, j, k;
    if (DIR == 1) { i = c; j = sa; k = sb; }
    else if (DIR == 2) { j = c; k = sa; i = sb; }
    else { k = c; i = sa; j = sb; }
    T Q[ME], QL[ME];
    load_cell3<ME>(a.qin, a.g, a.bc, a.ng, i, j, k, Q);
    from_left(Q, QL);
    T W[MW][ME], s[MW], am[ME], ap[ME];
    typename RS::Roe roe;
    RS::template rpn<DIR>(a.rp, QL, Q, W, s, am, ap, roe);
    // Courant number over interfaces 1..n+1 (flux3.f90:214-221)
    T smax = T(0);
    if (lane >= 1 && c >= 1 && c <= n + 1) {
  #pragma unroll
      for (int w = 0; w < MW; w++) smax = cfl_term(smax, fabs(s[w]));
    }
    cfl_wave = (double)dtn * (double)smax;
    // qadd(c) = (0 - dtn apdq(c)) - dtn amdq(c+1)  (flux3.f90:205-208)
    T amn[ME], qadd[ME];
    from_right(am, amn);
  #pragma unroll
    for (int m = 0; m < ME; m++) qadd[m] = (T(0) - dtn * ap[m]) - dtn * amn[m];
    // second-order correction (flux3.f90:224-259); cqxx carries the 0.5
    T cq[ME];
  #pragma unroll
    for (int m = 0; m < ME; m++) cq[m] = T(0);
    if (a.order != 1) {
      T Wn[MW][ME];
  #pragma unroll
      for (int w = 0; w < MW; w++) from_right(W[w], Wn[w]);
      T dotr[MW], dotl[MW];
      wave_dots<ME, MW>(W, Wn, dotr);
      from_left(dotr, dotl);
      limit_waves<ME, MW>(W, s, dotl, dotr, a.mthlim);
      const T dtdxave = T(0.5) * (dtn + dtn);
  #pragma unroll
      for (int w = 0; w < MW; w++) {
  #pragma unroll
        for (int m = 0; m < ME; m++) {
          const T sa2 = fabs(s[w]);
          const T f = a.rp.fwave ? copysign(T(1), s[w]) : sa2;
          cq[m] = cq[m] + T(0.5) * f * (T(1) - sa2 * dtdxave) * W[w][m];
        }
      }
    }
    T fadd[ME];
  #pragma unroll
    for (int m = 0; m < ME; m++) fadd[m] = T(0) + cq[m];

    // stores of this lane's cell c (cells 1..n; fadd at interfaces 1..n+1)
    const bool own = lane >= 2 && lane < 62 && c >= 1;
    const bool own_cell = own && c <= n;
    T* wk = a.work[DIR - 1];
    const long long st = a.wstride[DIR - 1];
    if (own && c <= n + 1) {
  #pragma unroll
      for (int m = 0; m < ME; m++) wk[widx(WL::FA + m, st, sa, sb, c, n, na)] = fadd[m];
    }
    if (own_cell) {
  #pragma unroll
      for (int m = 0; m < ME; m++) wk[widx(WL::QA + m, st, sa, sb, c, n, na)] = qadd[m];
    }
    const int m3 = a.m3, m4 = a.m4;
    T bmam[ME], bpam[ME], bmap[ME], bpap[ME], cmam[ME], cpam[ME], cmap[ME], cpap[ME];
    T bmcq[ME], bpcq[ME], cmcq[ME], cpcq[ME];
  #pragma unroll
    for (int m = 0; m < ME; m++) {
      bmam[m] = bpam[m] = bmap[m] = bpap[m] = cmam[m] = cpam[m] = cmap[m] = cpap[m] = T(0);
      bmcq[m] = bpcq[m] = cmcq[m] = cpcq[m] = T(0);
    }
    if (m3 > 0) {
      RS::template rpt3<DIR, 2>(roe, am, bmam, bpam);
      RS::template rpt3<DIR, 2>(roe, ap, bmap, bpap);
      RS::template rpt3<DIR, 3>(roe, am, cmam, cpam);
      RS::template rpt3<DIR, 3>(roe, ap, cmap, cpap);
      if (m3 == 2) {
        RS::template rpt3<DIR, 2>(roe, cq, bmcq, bpcq);
        RS::template rpt3<DIR, 3>(roe, cq, cmcq, cpcq);
      }
    }
    // One family of transverse modifications (gadd: family 0, split direction
    // 2, double-transverse factor dt2; hadd: family 1, direction 3, factor dt1)
    // for cell c: interface c's own terms as a running sum in statement order,
    // then interface c+1's terms for its left cell, shifted from the right lane
    // (flux3.f90:374-456, 488-580).  x*: the 2D splits (bm/bp of am and ap for
    // gadd, cm/cp for hadd); y*: the correction's split in that direction;
    // z*: the other direction's splits, which the double-transverse solve
    // splits again.
    auto family = [&](auto fam, const T (&xmam)[ME], const T (&xpam)[ME], const T (&xmap)[ME],
                      const T (&xpap)[ME], const T (&ymcq)[ME], const T (&ypcq)[ME],
                      const T (&zmam)[ME], const T (&zpam)[ME], const T (&zmap)[ME],
                      const T (&zpap)[ME], const T (&zmcq)[ME], const T (&zpcq)[ME], T dtf) {
      constexpr int FAM = decltype(fam)::value;
      constexpr int ICO = FAM == 0 ? 2 : 3;
      const T sixth = T(1) / T(6);
      T cmcpap[ME], cpcpap[ME], cmcpam[ME], cpcpam[ME], cmcmap[ME], cpcmap[ME], cmcmam[ME],
          cpcmam[ME];
      if (m3 > 0 && m4 > 0) {
        T x2[ME];
  #pragma unroll
        for (int m = 0; m < ME; m++) x2[m] = m4 == 2 ? zpap[m] - T(3) * zpcq[m] : zpap[m];
        RS::template rpt3<DIR, ICO>(roe, x2, cmcpap, cpcpap);
  #pragma unroll
        for (int m = 0; m < ME; m++) x2[m] = m4 == 2 ? zpam[m] + T(3) * zpcq[m] : zpam[m];
        RS::template rpt3<DIR, ICO>(roe, x2, cmcpam, cpcpam);
  #pragma unroll
        for (int m = 0; m < ME; m++) x2[m] = m4 == 2 ? zmap[m] - T(3) * zmcq[m] : zmap[m];
        RS::template rpt3<DIR, ICO>(roe, x2, cmcmap, cpcmap);
  #pragma unroll
        for (int m = 0; m < ME; m++) x2[m] = m4 == 2 ? zmam[m] + T(3) * zmcq[m] : zmam[m];
        RS::template rpt3<DIR, ICO>(roe, x2, cmcmam, cpcmam);
      }
  #pragma unroll
      for (int m = 0; m < ME; m++) {
        // [side][off + 1]: the running sums of cell c
        T v[2][3];
        v[0][0] = v[0][1] = v[0][2] = v[1][0] = v[1][1] = v[1][2] = T(0);
        // interface c+1's terms for this cell, shifted from the right lane
        T l0[2], l4[2], l3[2], lp[2], lm[2];
        if (m3 > 0) {
          v[0][1] = v[0][1] - T(0.5) * dtn * xmap[m];
          v[1][1] = v[1][1] - T(0.5) * dtn * xpap[m];
          l0[0] = T(0.5) * dtn * xmam[m];
          l0[1] = T(0.5) * dtn * xpam[m];
          if (m4 > 0) {
            v[1][1] = v[1][1] + sixth * dtn * dtf * (cpcpap[m] - cpcmap[m]);
            v[0][1] = v[0][1] + sixth * dtn * dtf * (cmcpap[m] - cmcmap[m]);
            v[1][2] = v[1][2] - sixth * dtn * dtf * cpcpap[m];
            v[0][2] = v[0][2] - sixth * dtn * dtf * cmcpap[m];
            v[1][0] = v[1][0] + sixth * dtn * dtf * cpcmap[m];
            v[0][0] = v[0][0] + sixth * dtn * dtf * cmcmap[m];
            l4[1] = sixth * dtn * dtf * (cpcpam[m] - cpcmam[m]);
            l4[0] = sixth * dtn * dtf * (cmcpam[m] - cmcmam[m]);
            lp[1] = sixth * dtn * dtf * cpcpam[m];
            lp[0] = sixth * dtn * dtf * cmcpam[m];
            lm[1] = sixth * dtn * dtf * cpcmam[m];
            lm[0] = sixth * dtn * dtf * cmcmam[m];
          }
          if (m3 >= 2) {
            v[1][1] = v[1][1] + dtn * ypcq[m];
            v[0][1] = v[0][1] + dtn * ymcq[m];
            l3[1] = dtn * ypcq[m];
            l3[0] = dtn * ymcq[m];
          }
  #pragma unroll
          for (int sd = 0; sd < 2; sd++) {
            v[sd][1] = v[sd][1] - from_right(l0[sd]);
            if (m4 > 0) {
              v[sd][1] = v[sd][1] + from_right(l4[sd]);
              v[sd][2] = v[sd][2] - from_right(lp[sd]);
              v[sd][0] = v[sd][0] + from_right(lm[sd]);
            }
            if (m3 >= 2) v[sd][1] = v[sd][1] - from_right(l3[sd]);
          }
        }
        if (own_cell) {
          constexpr int base = FAM == 0 ? WL::G : WL::H;
  #pragma unroll
          for (int sd = 0; sd < 2; sd++)
  #pragma unroll
            for (int o = 0; o < 3; o++)
              wk[widx(WL::gh(base, m, sd + 1, o - 1), st, sa, sb, c, n, na)] = v[sd][o];
        }
      }
    };
    family(std::integral_constant<int, 0>(), bmam, bpam, bmap, bpap, bmcq, bpcq, cmam, cpam,
           cmap, cpap, cmcq, cpcq, dt2);
    // hadd's double-transverse input: the y-like splits with the correction
    family(std::integral_constant<int, 1>(), cmam, cpam, cmap, cpap, cmcq, cpcq, bmam, bpam,
           bmap, bpap, bmcq, bpcq, dt1);
  }
  block_atomic_max(cfl_wave, a.cfl_bits);
}

// One slice's update of a cell at transverse offsets (da, db) from the slice
// (step3.f90:189-225 and the y / z equivalents): v is the cell's running value.
template <int ME>
__device__ __forceinline__ double term3(double v, const double* wk, long long st, int sa, int sb,
                                        int p, int n, int na, int m, int da, int db, double dtn,
                                        double dt1, double dt2) {
  using WL = W3<ME>;
  auto G = [&](int side, int off) { return wk[widx(WL::gh(WL::G, m, side, off), st, sa, sb, p, n, na)]; };
  auto H = [&](int side, int off) { return wk[widx(WL::gh(WL::H, m, side, off), st, sa, sb, p, n, na)]; };
  if (da == 0 && db == 0) {
    const double qa = wk[widx(WL::QA + m, st, sa, sb, p, n, na)];
    const double f0 = wk[widx(WL::FA + m, st, sa, sb, p, n, na)];
    const double f1 = wk[widx(WL::FA + m, st, sa, sb, p + 1, n, na)];
    return v + qa - dtn * (f1 - f0) - dt1 * (G(2, 0) - G(1, 0)) - dt2 * (H(2, 0) - H(1, 0));
  }
  if (da == -1 && db == 0) return v - dt1 * G(1, 0) - dt2 * (H(2, -1) - H(1, -1));
  if (da == -1 && db == -1) return v - dt1 * G(1, -1) - dt2 * H(1, -1);
  if (da == 0 && db == -1) return v - dt1 * (G(2, -1) - G(1, -1)) - dt2 * H(1, 0);
  if (da == 1 && db == -1) return v + dt1 * G(2, -1) - dt2 * H(1, 1);
  if (da == 1 && db == 0) return v + dt1 * G(2, 0) - dt2 * (H(2, 1) - H(1, 1));
  if (da == 1 && db == 1) return v + dt1 * G(2, 1) + dt2 * H(2, 1);
  if (da == 0 && db == 1) return v - dt1 * (G(2, 1) - G(1, 1)) + dt2 * H(2, 0);
  return v - dt1 * G(1, 1) + dt2 * H(2, -1);   // (-1, 1)
}

template <class RS>
__global__ void __launch_bounds__(256) k_apply3(const Step3Args<double> a) {
  constexpr int ME = RS::meqn;
  const Geom& g = a.g;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ncell = (long long)g.mx * g.my * g.mz;
  if (cell >= ncell) return;
  const int i = 1 + (int)(cell % g.mx);
  const int j = 1 + (int)((cell / g.mx) % g.my);
  const int k = 1 + (int)(cell / ((long long)g.mx * g.my));
  const double* pin = cell_ptr3(a.qin, g, i, j, k);
  Geom go = g;
  go.g = 0; go.rs = g.mx; go.ss = (long long)g.mx * g.my; go.fs = go.ss * g.mz;
  double* pout = cell_ptr3(a.qout, go, i, j, k);
#pragma unroll 1
  for (int m = 0; m < ME; m++) {
    double v = pin[(long long)m * g.fs];
    // x-slices (a = j', b = k'): k' outer, j' inner
    for (int bb = k - 1; bb <= k + 1; bb++)
      for (int aa = j - 1; aa <= j + 1; aa++)
        v = term3<ME>(v, a.work[0], a.wstride[0], aa, bb, i, g.mx, g.my, m, j - aa, k - bb,
                      a.dtdx, a.dtdy, a.dtdz);
    // y-slices (a = k', b = i'): k' outer, i' inner
    for (int aa = k - 1; aa <= k + 1; aa++)
      for (int bb = i - 1; bb <= i + 1; bb++)
        v = term3<ME>(v, a.work[1], a.wstride[1], aa, bb, j, g.my, g.mz, m, k - aa, i - bb,
                      a.dtdy, a.dtdz, a.dtdx);
    // z-slices (a = i', b = j'): j' outer, i' inner
    for (int bb = j - 1; bb <= j + 1; bb++)
      for (int aa = i - 1; aa <= i + 1; aa++)
        v = term3<ME>(v, a.work[2], a.wstride[2], aa, bb, k, g.mz, g.mx, m, i - aa, j - bb,
                      a.dtdz, a.dtdx, a.dtdy);
    pout[(long long)m * go.fs] = v;
  }
}

namespace {
template <class RS, int DIR>
int launch_flux3(Step3Args<double> a, hipStream_t stream) {
  int n, na, nb;
  dir_extents<DIR>(a.g, n, na, nb);
  a.nstrips = (n + 1 + 59) / 60;
  const long long waves = (long long)a.nstrips * (na + 2) * (nb + 2);
  hipLaunchKernelGGL((k_flux3<RS, DIR>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                     stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}
}  // namespace

long long step3_work_elems(int meqn, int mx, int my, int mz) {
  const long long per = 14LL * meqn;
  return per * ((long long)(mx + 1) * (my + 2) * (mz + 2) + (long long)(my + 1) * (mz + 2) *
                (mx + 2) + (long long)(mz + 1) * (mx + 2) * (my + 2));
}

int launch_step3(const clawhip_rp& rp, const void* q_in, const Geom& gin, void* q_out,
                 void* work, const int* lo, const int* hi, const int* refl, int ng,
                 const int* method, const int* mthlim, double dx, double dy, double dz,
                 double dt, unsigned long long* cfl, hipStream_t stream) {
  if (rp.kind != CLAWHIP_RP_EULER3) return CLAWHIP_E_KIND;
  using RS = Euler3<double>;
  Step3Args<double> a;
  a.rp = to_params<double>(rp);
  a.g = gin;
  a.bc = to_bc(lo, hi, refl, 3);
  a.ng = ng;
  a.order = method[1];
  a.m3 = method[2] / 10;
  a.m4 = method[2] - 10 * a.m3;
  for (int w = 0; w < kMaxWaves; w++) a.mthlim[w] = w < RS::mwaves ? mthlim[w] : 0;
  a.dtdx = dt / dx; a.dtdy = dt / dy; a.dtdz = dt / dz;
  a.qin = (const double*)q_in;
  a.qout = (double*)q_out;
  a.cfl_bits = cfl;
  const int mx = gin.mx, my = gin.my, mz = gin.mz, ME = RS::meqn;
  const long long s1 = (long long)(mx + 1) * (my + 2) * (mz + 2);
  const long long s2 = (long long)(my + 1) * (mz + 2) * (mx + 2);
  const long long s3 = (long long)(mz + 1) * (mx + 2) * (my + 2);
  double* w = (double*)work;
  a.work[0] = w; a.wstride[0] = s1;
  a.work[1] = w + 14LL * ME * s1; a.wstride[1] = s2;
  a.work[2] = a.work[1] + 14LL * ME * s2; a.wstride[2] = s3;
  int rc = launch_flux3<RS, 1>(a, stream);
  if (!rc) rc = launch_flux3<RS, 2>(a, stream);
  if (!rc) rc = launch_flux3<RS, 3>(a, stream);
  if (rc) return rc;
  const long long ncell = (long long)mx * my * mz;
  hipLaunchKernelGGL((k_apply3<RS>), dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0,
                     stream, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace claw
