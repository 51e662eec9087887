// stats.hip -- FP64 reductions over a state: the conserved totals of each field
// with a count of cells that lost positivity, and the L1 distance between two
// states (the check the SharpClaw indicator path is judged by).
//
// Both are deterministic: a fixed grid of SUMS_BLOCKS blocks, each thread
// walking a fixed set of cells in order, a fixed-order tree in LDS per block,
// then one block summing the per-block partials in block order.  Sums are FP64
// for FP32 states too.
#include <mutex>

#include "launch.hpp"

namespace claw {

constexpr int SUMS_BLOCKS = 2048;
constexpr int SUMS_THREADS = 256;
constexpr int SUMS_MAXF = 8;      // fields reduced at once (meqn + 1 <= 8)

// Fixed-order tree sum of v over the block; returns the total in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = SUMS_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// partial[b * (nf) + f]: block b's sum of field f (f < meqn) and, at f = meqn,
// its count of cells where a field in positive_mask is <= 0 or not finite.
// Block b takes the rows (j, k) = b, b + SUMS_BLOCKS, ...; its threads stride
// along each row.
template <typename T>
__global__ void __launch_bounds__(SUMS_THREADS) k_state_sums(const T* q, Geom g, int meqn,
                                                             unsigned positive_mask,
                                                             double* partial) {
  __shared__ double sh[SUMS_THREADS];
  double acc[SUMS_MAXF - 1], nbad = 0.0;
#pragma unroll
  for (int f = 0; f < SUMS_MAXF - 1; f++) acc[f] = 0.0;
  const int nrow = g.my * g.mz;
  for (int row = blockIdx.x; row < nrow; row += SUMS_BLOCKS) {
    const int r = row % g.my, k = row / g.my;
    const T* base = q + (long long)g.g + (long long)(r + g.g) * g.rs +
                    (g.mz > 1 ? (long long)(k + g.g) * g.ss : 0LL);
#pragma unroll 4
    for (int c = threadIdx.x; c < g.mx; c += SUMS_THREADS) {
      bool bad = false;
#pragma unroll
      for (int f = 0; f < SUMS_MAXF - 1; f++) {
        if (f < meqn) {
          const double x = (double)base[(long long)f * g.fs + c];
          acc[f] += x;
          if ((positive_mask >> f) & 1u) bad = bad || !(x > 0.0) || !isfinite(x);
        }
      }
      if (bad) nbad += 1.0;
    }
  }
#pragma unroll
  for (int f = 0; f < SUMS_MAXF; f++) {
    if (f <= meqn) {   // uniform across the block
      const double s = block_sum(f < meqn ? acc[f < SUMS_MAXF - 1 ? f : 0] : nbad, sh);
      if (threadIdx.x == 0) partial[(long long)blockIdx.x * (meqn + 1) + f] = s;
    }
  }
}

// partial[b * 2 + 0] = sum |a - b|, [b * 2 + 1] = sum |b| over block b's elements.
template <typename T>
__global__ void __launch_bounds__(SUMS_THREADS) k_l1_diff(const T* a, const T* b, long long n,
                                                          double* partial) {
  __shared__ double sh[SUMS_THREADS];
  double d = 0.0, s = 0.0;
  const long long stride = (long long)SUMS_BLOCKS * SUMS_THREADS;
  for (long long i = (long long)blockIdx.x * SUMS_THREADS + threadIdx.x; i < n; i += stride) {
    const double x = (double)a[i], y = (double)b[i];
    d += fabs(x - y);
    s += fabs(y);
  }
  d = block_sum(d, sh);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    partial[2LL * blockIdx.x] = d;
    partial[2LL * blockIdx.x + 1] = s;
  }
}

// out[f] = sum over blocks of partial[b * nf + f], in block order (one block).
__global__ void __launch_bounds__(SUMS_THREADS) k_finish_sums(const double* partial, int nf,
                                                              double* out) {
  __shared__ double sh[SUMS_THREADS];
  for (int f = 0; f < nf; f++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < SUMS_BLOCKS; b += SUMS_THREADS) v += partial[(long long)b * nf + f];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[f] = v;
  }
}

int launch_state_sums(int dtype, const void* q, const Geom& g, int meqn, unsigned positive_mask,
                      double* out, double* work, hipStream_t s) {
  if (meqn < 1 || meqn + 1 > SUMS_MAXF) return CLAWHIP_E_ARG;
  if (dtype == CLAWHIP_F64)
    hipLaunchKernelGGL(k_state_sums<double>, dim3(SUMS_BLOCKS), dim3(SUMS_THREADS), 0, s,
                       (const double*)q, g, meqn, positive_mask, work);
  else
    hipLaunchKernelGGL(k_state_sums<float>, dim3(SUMS_BLOCKS), dim3(SUMS_THREADS), 0, s,
                       (const float*)q, g, meqn, positive_mask, work);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(SUMS_THREADS), 0, s, work, meqn + 1, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

// ---------------------------------------------------------------- no-allocation form
// The per-block partials of launch_state_sums_noalloc live in this file-scope
// device array instead of a caller's work buffer: SUMS_BLOCKS * SUMS_MAXF doubles
// (128 KiB), held by every device the library runs on until the process ends.
// Every block of k_state_sums writes all meqn + 1 of its partials, and
// k_finish_sums reads only those, so nothing in it has to start at zero.  Calls
// on one device share the array: calls that overlap on one device (two streams,
// or two host threads) race.
static __device__ __align__(256) double scratch[SUMS_BLOCKS * SUMS_MAXF];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy at its own address).  Host code uses
// only this address: the array's name is a host-side stand-in.
hipError_t scratch_address(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace

int launch_state_sums_noalloc(int dtype, const void* q, const Geom& g, int meqn,
                              unsigned positive_mask, double* out, hipStream_t s) {
  if (meqn < 1 || meqn + 1 > SUMS_MAXF) return CLAWHIP_E_ARG;
  static_assert(SUMS_BLOCKS * SUMS_MAXF * sizeof(double) == sizeof(scratch),
                "scratch holds the partials of SUMS_MAXF fields");
  double* work = nullptr;
  hipError_t e = scratch_address(&work);
  if (e != hipSuccess) return (int)e;
  return launch_state_sums(dtype, q, g, meqn, positive_mask, out, work, s);
}

int launch_l1_diff(int dtype, const void* a, const void* b, long long n, double* out,
                   double* work, hipStream_t s) {
  if (dtype == CLAWHIP_F64)
    hipLaunchKernelGGL(k_l1_diff<double>, dim3(SUMS_BLOCKS), dim3(SUMS_THREADS), 0, s,
                       (const double*)a, (const double*)b, n, work);
  else
    hipLaunchKernelGGL(k_l1_diff<float>, dim3(SUMS_BLOCKS), dim3(SUMS_THREADS), 0, s,
                       (const float*)a, (const float*)b, n, work);
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(SUMS_THREADS), 0, s, work, 2, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace claw
