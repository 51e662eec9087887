// ensemble.hip -- many independent 2D unsplit problems (uneven grid sizes, one
// Riemann solver and method) advanced together: every launch sweeps all
// members, each with its own dt and device step controller.
//
// The members' wavefronts are concatenated into one grid (largest members
// first, so the small ones fill the tail); a wavefront finds its member by
// binary search over the members' first-wavefront offsets and then runs the
// same step body as the single-grid kernel (unsplit.hpp).  The last wavefront of
// a member applies that member's accept/reject and next dt (stepctl.hpp).  A
// step of the whole ensemble is one launch with no host synchronisation, so
// runs of steps are captured in a HIP graph.
#include "variant.hpp"  // first: selects the arithmetic variant
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "launch.hpp"
#include "unsplit.hpp"

namespace claw {
namespace CLAW_VNS {

struct EnsMember {
  int mx, my;
  int nstrips, nrb;
  long long base;             // first wavefront of this member
  const void* q[2];
  void* comp[2];
  double dx, dy;
  StepCtl* ctl;
};

struct EnsHeader {
  int n, H, ng, order, trans;
  int mthlim[4];
  BCSpec bc;
  RPParams<double> rp64;
  RPParams<float> rp32;
  long long waves;
};

constexpr long long kEnsAlign = 256;
inline long long ens_members_offset() {
  return (sizeof(EnsHeader) + kEnsAlign - 1) / kEnsAlign * kEnsAlign;
}

template <typename T>
__device__ __forceinline__ const RPParams<T>& ens_rp(const EnsHeader& h);
template <>
__device__ __forceinline__ const RPParams<double>& ens_rp<double>(const EnsHeader& h) {
  return h.rp64;
}
template <>
__device__ __forceinline__ const RPParams<float>& ens_rp<float>(const EnsHeader& h) {
  return h.rp32;
}

template <class RS, typename T, bool KAHAN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(unsplit_waves<RS, T>())))
k_ens_step(const EnsHeader* __restrict__ hdr, const EnsMember* __restrict__ mem) {
  const long long gw = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gw >= hdr->waves) return;
  // member k: the last one whose first wavefront is <= gw
  int lo = 0, hi = hdr->n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].base <= gw) lo = mid;
    else hi = mid - 1;
  }
  const EnsMember& m = mem[lo];
  StepCtl* ctl = m.ctl;
  if (ctl->done) return;   // uniform across the member
  const int cur = ctl->cur;
  const double dt = ctl->dt;
  UnsplitArgs<T> a;
  a.rp = ens_rp<T>(*hdr);
  a.gin.mx = m.mx; a.gin.my = m.my; a.gin.mz = 1; a.gin.g = 0;
  a.gin.fs = (long long)m.mx * m.my; a.gin.rs = m.mx; a.gin.ss = 0;
  a.gout = a.gin;
  a.bc = hdr->bc;
  a.ng = hdr->ng; a.order = hdr->order; a.trans = hdr->trans;
  for (int w = 0; w < 4; w++) a.mthlim[w] = hdr->mthlim[w];
  a.H = hdr->H; a.nstrips = m.nstrips; a.nrb = m.nrb;
  a.qin = (const T*)m.q[cur];
  a.qout = (T*)m.q[1 - cur];
  a.comp_in = (const T*)m.comp[cur];
  a.comp_out = (T*)m.comp[1 - cur];
  a.dtdx = T(dt / m.dx);
  a.dtdy = T(dt / m.dy);
  a.cfl_bits = &ctl->cfl_bits;
  const int local = (int)(gw - m.base);
  // the members' Courant numbers go to their own controllers (many
  // addresses), and a workgroup may span members that return above
  wave_atomic_max(unsplit_body<RS, T, KAHAN>(a, local), a.cfl_bits);
}

// Every member's step control after the step launch (step_ctl_apply), one
// thread per member.
__global__ void __launch_bounds__(256) k_ens_ctl(const EnsHeader* __restrict__ hdr,
                                                 const EnsMember* __restrict__ mem) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < hdr->n) step_ctl_apply(mem[k].ctl);
}

namespace {

// The table does not know the state's type, so the default rows follow the
// occupancy of the FP32 + Kahan Euler step (the ensemble workload's, counted up
// to the wavefronts per SIMD it is compiled for); the autotuner
// (Ensemble.autotune) replaces them by measured rows.
long long ens_resident_waves() {
  constexpr int hint = unsplit_waves<Euler<float>, float>();
  const long long r =
      resident_waves((const void*)k_ens_step<Euler<float>, float, true>, hint > 2 ? hint : 2);
  return r > 0 ? r : 1024;
}

long long ens_waves(const std::vector<EnsMember>& v, int H) {
  long long w = 0;
  for (const auto& m : v) w += (long long)((m.mx + 59) / 60) * ((m.my + H - 1) / H);
  return w;
}

// Rows per wavefront for the whole ensemble: the single-grid rule
// (rounds x (H + 2)) over the total wavefront count.
int ens_choose_rows(const std::vector<EnsMember>& v) {
  const long long R = ens_resident_waves();
  int best = 16;
  long long best_cost = -1;
  for (int h = 2; h <= 64; h++) {
    const long long cost = ((ens_waves(v, h) + R - 1) / R) * (h + 2);
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && h > best)) {
      best = h;
      best_cost = cost;
    }
  }
  return best;
}

template <class RS, typename T, bool KAHAN>
void ens_launch(const EnsHeader* h, const EnsMember* m, const long long* waves, int nsteps,
                hipStream_t s) {
  // wavefronts per workgroup: 1, 2 or 4 (the autotuner's second parameter)
  const long long wpb = (waves[2] == 1 || waves[2] == 2) ? waves[2] : 4;
  const unsigned blocks = (unsigned)((waves[0] + wpb - 1) / wpb);
  const unsigned ctl_blocks = (unsigned)((waves[1] + 255) / 256);
  for (int n = 0; n < nsteps; n++) {
    hipLaunchKernelGGL((k_ens_step<RS, T, KAHAN>), dim3(blocks), dim3((unsigned)(64 * wpb)), 0,
                       s, h, m);
    hipLaunchKernelGGL(k_ens_ctl, dim3(ctl_blocks), dim3(256), 0, s, h, m);
  }
}

template <template <typename> class RSF>
void ens_dispatch(int dtype, const EnsHeader* h, const EnsMember* m, const long long* waves,
                  int nsteps, hipStream_t s) {
  if (dtype != CLAWHIP_F32K) ens_launch<RSF<double>, double, false>(h, m, waves, nsteps, s);
#if CLAW_FAST
  else ens_launch<RSF<float>, float, true>(h, m, waves, nsteps, s);
#endif
}

}  // namespace

long long ensemble_table_bytes(int n) {
  return ens_members_offset() + (long long)n * (long long)sizeof(EnsMember);
}

int ensemble_prepare(const clawhip_rp& rp, const clawhip_ens_member* members, int n,
                     const int* lo, const int* hi, const int* refl, int ng, const int* method,
                     const int* mthlim, int rows, void* table_dev, long long* waves) {
  std::vector<EnsMember> v(n);
  for (int k = 0; k < n; k++) {
    const clawhip_ens_member& s = members[k];
    EnsMember& m = v[k];
    m.mx = s.mx; m.my = s.my;
    m.q[0] = s.q0; m.q[1] = s.q1; m.comp[0] = s.comp0; m.comp[1] = s.comp1;
    m.dx = s.dx; m.dy = s.dy;
    m.ctl = (StepCtl*)s.ctl;
  }
  // largest members first: their long wavefront lists start early, the small
  // members fill the tail of each launch
  std::stable_sort(v.begin(), v.end(), [](const EnsMember& a, const EnsMember& b) {
    return (long long)a.mx * a.my > (long long)b.mx * b.my;
  });
  const int H = rows > 0 ? rows : ens_choose_rows(v);
  long long base = 0;
  for (auto& m : v) {
    m.nstrips = (m.mx + 59) / 60;
    m.nrb = (m.my + H - 1) / H;
    m.base = base;
    base += (long long)m.nstrips * m.nrb;
  }
  EnsHeader h;
  h.n = n; h.H = H; h.ng = ng; h.order = method[1]; h.trans = method[2];
  for (int w = 0; w < 4; w++) h.mthlim[w] = mthlim[w];
  h.bc = to_bc(lo, hi, refl, 2);
  h.rp64 = to_params<double>(rp);
  h.rp32 = to_params<float>(rp);
  h.waves = base;
  std::vector<char> buf((size_t)ensemble_table_bytes(n), 0);
  ::memcpy(buf.data(), &h, sizeof(h));
  ::memcpy(buf.data() + ens_members_offset(), v.data(), (size_t)n * sizeof(EnsMember));
  hipError_t e = hipMemcpy(table_dev, buf.data(), buf.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return (int)e;
  waves[0] = base; waves[1] = n;   // wavefronts of a step launch; members
  waves[2] = 4;                    // wavefronts per workgroup (the caller may change it)
  return 0;
}

int ensemble_run(const clawhip_rp& rp, int dtype, const void* table_dev, const long long* waves,
                 int nsteps, hipStream_t s) {
  const EnsHeader* h = (const EnsHeader*)table_dev;
  const EnsMember* m = (const EnsMember*)((const char*)table_dev + ens_members_offset());
  switch (rp.kind) {
    case CLAWHIP_RP_ACOUSTICS: ens_dispatch<Acoustics>(dtype, h, m, waves, nsteps, s); break;
    case CLAWHIP_RP_EULER: ens_dispatch<Euler>(dtype, h, m, waves, nsteps, s); break;
    case CLAWHIP_RP_SHALLOW: ens_dispatch<Shallow>(dtype, h, m, waves, nsteps, s); break;
    case CLAWHIP_RP_ADVECTION: ens_dispatch<Advection>(dtype, h, m, waves, nsteps, s); break;
    default: return CLAWHIP_E_KIND;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace CLAW_VNS
}  // namespace claw
