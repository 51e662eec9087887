// indicator.hip -- learned troubled-interface indicator for the SharpClaw WENO5
// path: a 6 -> 32 -> 32 -> 1 MLP on BF16 MFMA (v_mfma_f32_16x16x16bf16_1k, FP32
// accumulation) with hand-written training (forward, backward, Adam) and
// inference kernels.
//
// Sample = (interface, conserved component): the six cell values of the WENO5
// stencil of that interface (cells c-3..c+2), min-max normalised to [-1, 1]
// (all zeros when the stencil range is below `flat`).  Label = the component's
// Jiang-Shu WENO5 weights (the weight formulas of reconstruct.f90:264-284
// applied to the component's differences, both sides) depart from the linear
// weights (0.1, 0.6, 0.3) by more than `wtol` while the stencil range exceeds
// `flat`.  An interface is smooth when no component is flagged; the stage kernel
// then uses the linear 5th-order reconstruction (WENO5 with its linear weights)
// and skips eigenvectors, projections and nonlinear weights.
//
// MFMA 16x16x16 bf16 operand layout (wave64): lane l holds A[l%16][4*(l/16)+0..3],
// B[4*(l/16)+0..3][l%16]; the 4 FP32 results are D[4*(l/16)+0..3][l%16].
#include <hip/hip_bf16.h>
#include <stdlib.h>

#include "launch.hpp"

namespace claw {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = 16;      // padded input width (6 features, 1.0 bias at index 6)
constexpr int NH = 32;      // hidden width

// Parameter block layout (FP32 master and BF16 working copies use the same order)
//   W1[NF][NH] (row 6 is b1), W2[NH][NH], b2[NH], w3[NH], b3
constexpr int P_W1 = 0, P_W2 = P_W1 + NF * NH, P_B2 = P_W2 + NH * NH, P_W3 = P_B2 + NH,
              P_B3 = P_W3 + NH, P_N = P_B3 + 1;

// float -> BF16 bits, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((unsigned)(unsigned short)b) << 16); }

// ---------------------------------------------------------------- features
// stencil v[0..5] -> normalised features f[0..5]; returns the range.  Computed in
// the state's precision (features are rounded to BF16 afterwards).
template <typename T>
__device__ __forceinline__ float norm_stencil(const T (&v)[6], float flat, float (&f)[6]) {
  T lo = v[0], hi = v[0];
#pragma unroll
  for (int k = 1; k < 6; k++) { lo = v[k] < lo ? v[k] : lo; hi = v[k] > hi ? v[k] : hi; }
  T rng = hi - lo;
  if (rng <= (T)flat) {
#pragma unroll
    for (int k = 0; k < 6; k++) f[k] = 0.f;
  } else {
    T inv = T(2) / rng;
#pragma unroll
    for (int k = 0; k < 6; k++) f[k] = (float)((v[k] - lo) * inv - T(1));
  }
  return (float)rng;
}

// Departure of the scalar WENO5 weights from the linear weights, both sides.
__device__ __forceinline__ double weno_departure(const double (&v)[6]) {
  double hh[5];
#pragma unroll
  for (int k = 0; k < 5; k++) hh[k] = v[k + 1] - v[k];
  double worst = 0.0;
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int im = side == 0 ? 1 : -1;
    double a2 = im == 1 ? hh[0] : hh[4], a1 = im == 1 ? hh[1] : hh[3];
    double h0 = hh[2], b1 = im == 1 ? hh[3] : hh[1];
    double t1 = im * (a2 - a1), t2 = im * (a1 - h0), t3 = im * (h0 - b1);
    double e1 = a2 - 3.0 * a1, e2 = a1 + h0, e3 = 3.0 * h0 - b1;
    const double eps = (double)1.0e-36f;
    double tt1 = 13.0 * (t1 * t1) + 3.0 * (e1 * e1);
    double tt2 = 13.0 * (t2 * t2) + 3.0 * (e2 * e2);
    double tt3 = 13.0 * (t3 * t3) + 3.0 * (e3 * e3);
    tt1 = (eps + tt1) * (eps + tt1);
    tt2 = (eps + tt2) * (eps + tt2);
    tt3 = (eps + tt3) * (eps + tt3);
    double s1 = tt2 * tt3, s2 = 6.0 * tt1 * tt3, s3 = 3.0 * tt1 * tt2;
    double sum = s1 + s2 + s3;
    double w1 = s1 / sum, w2 = s2 / sum, w3 = s3 / sum;
    double d = fabs(w1 - 0.1);
    d = fmax(d, fabs(w2 - 0.6));
    d = fmax(d, fabs(w3 - 0.3));
    worst = fmax(worst, d);
  }
  return worst;
}

// Gather the stencil of interface `idx` (x-interfaces first, then y) for field m.
// x-interface (c, r): c in 1..mx+1, r in 1..my; y-interface (c, r): r in 1..my+1.
template <typename T, typename V>
__device__ __forceinline__ void gather_stencil(const T* q, const Geom& g, const BCSpec& bc,
                                               long long idx, int m, V (&v)[6]) {
  const long long nx = (long long)(g.mx + 1) * g.my;
  int c, r, dir;
  if (idx < nx) {
    c = 1 + (int)(idx % (g.mx + 1)); r = 1 + (int)(idx / (g.mx + 1)); dir = 0;
  } else {
    long long k = idx - nx;
    c = 1 + (int)(k % g.mx); r = 1 + (int)(k / g.mx); dir = 1;
  }
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int cc = dir == 0 ? c - 3 + s : c;
    const int rr = dir == 0 ? r : r - 3 + s;
    bool fx = false, fy = false;
    int ci = bc_map(cc, g.mx, 3, bc.lo[0], bc.hi[0], fx);
    int ri = bc_map(rr, g.my, 3, bc.lo[1], bc.hi[1], fy);
    V x = (V)q[(long long)m * g.fs + (ci - 1) + (long long)(ri - 1) * g.rs];
    if ((fx && m == bc.refl[0]) != (fy && m == bc.refl[1])) x = -x;
    v[s] = x;
  }
}

// Stencil of the interface (c, r) in direction dir (0: x-interface between cells
// (c-1, r) and (c, r); 1: y-interface between (c, r-1) and (c, r)), field m.
// Interior stencils load directly; others go through the boundary map.
template <typename T, typename V>
__device__ __forceinline__ void gather_stencil_at(const T* q, const Geom& g, const BCSpec& bc,
                                                  int dir, int c, int r, int m, V (&v)[6]) {
  const T* base = q + (long long)m * g.fs;
  const bool interior = dir == 0 ? (c - 3 >= 1 && c + 2 <= g.mx) : (r - 3 >= 1 && r + 2 <= g.my);
  if (interior) {
    const T* p0 = base + (long long)(r - 1) * g.rs + (c - 1);
    const long long step = dir == 0 ? 1 : g.rs;
#pragma unroll
    for (int s = 0; s < 6; s++) v[s] = (V)p0[(s - 3) * step];
    return;
  }
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int cc = dir == 0 ? c - 3 + s : c;
    const int rr = dir == 0 ? r : r - 3 + s;
    bool fx = false, fy = false;
    int ci = bc_map(cc, g.mx, 3, bc.lo[0], bc.hi[0], fx);
    int ri = bc_map(rr, g.my, 3, bc.lo[1], bc.hi[1], fy);
    V x = (V)base[(ci - 1) + (long long)(ri - 1) * g.rs];
    if ((fx && m == bc.refl[0]) != (fy && m == bc.refl[1])) x = -x;
    v[s] = x;
  }
}

// Features and labels of every (interface, component) sample of a state.
template <typename T>
__global__ void k_features_labels(const T* q, Geom g, BCSpec bc, int meqn, float flat,
                                  float wtol, float* feat, unsigned char* label) {
  const long long nint = (long long)(g.mx + 1) * g.my + (long long)g.mx * (g.my + 1);
  const long long n = nint * meqn;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = (int)(i % meqn);
  const long long idx = i / meqn;
  double v[6];
  gather_stencil(q, g, bc, idx, m, v);
  float f[6];
  float rng = norm_stencil(v, flat, f);
#pragma unroll
  for (int k = 0; k < 6; k++) feat[i * 8 + k] = f[k];
  feat[i * 8 + 6] = 1.0f;
  feat[i * 8 + 7] = 0.0f;
  label[i] = (rng > flat && weno_departure(v) > (double)wtol) ? 1 : 0;
}

// ---------------------------------------------------------------- MLP forward
// One wavefront evaluates 16 samples.  X: 16 rows x NF (bf16) in LDS; returns
// the logit for rows 4*(lane/16)+a in z3[a] (identical across the 16 lanes of a
// row group), and leaves Z1, Z2 (pre-activation, FP32, D layout) for backward.
struct Fwd {
  f32x4 z1[2], z2[2];
  float z3[4];
};

__device__ __forceinline__ bf16x4 load_a(const short* base, int row, int k0, int ld) {
  // A operand: row `row`, k0..k0+3 from a row-major [row][k] bf16 matrix
  bf16x4 a;
  a.x = base[row * ld + k0 + 0]; a.y = base[row * ld + k0 + 1];
  a.z = base[row * ld + k0 + 2]; a.w = base[row * ld + k0 + 3];
  return a;
}
__device__ __forceinline__ bf16x4 load_b(const short* base, int k0, int col, int ld) {
  // B operand: k0..k0+3 of column `col` from a row-major [k][col] bf16 matrix
  bf16x4 b;
  b.x = base[(k0 + 0) * ld + col]; b.y = base[(k0 + 1) * ld + col];
  b.z = base[(k0 + 2) * ld + col]; b.w = base[(k0 + 3) * ld + col];
  return b;
}

// P: bf16 parameters in LDS (layout P_*); sX: [16][NF] bf16; sH: [16][NH] bf16 scratch
__device__ __forceinline__ void mlp_forward(const short* P, const short* sX, short* sH, Fwd& o) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, kq = 4 * (lane >> 4);
  // layer 1: Z1 = X W1 (bias row folded in)
#pragma unroll
  for (int nb = 0; nb < 2; nb++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    bf16x4 a = load_a(sX, col, kq, NF);
    bf16x4 b = load_b(P + P_W1, kq, nb * 16 + col, NH);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    o.z1[nb] = acc;
#pragma unroll
    for (int e = 0; e < 4; e++) sH[(kq + e) * NH + nb * 16 + col] = f2bf(fmaxf(acc[e], 0.f));
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  // layer 2: Z2 = H1 W2 + b2
#pragma unroll
  for (int nb = 0; nb < 2; nb++) {
    const float bias = bf2f(P[P_B2 + nb * 16 + col]);
    f32x4 acc = {bias, bias, bias, bias};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      bf16x4 a = load_a(sH, col, kb * 16 + kq, NH);
      bf16x4 b = load_b(P + P_W2, kb * 16 + kq, nb * 16 + col, NH);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    o.z2[nb] = acc;
  }
  // layer 3: z3 = relu(Z2) . w3 + b3  (rows 4*(lane/16)+e, reduce over 16 lanes x 2 blocks)
#pragma unroll
  for (int e = 0; e < 4; e++) {
    float s = fmaxf(o.z2[0][e], 0.f) * bf2f(P[P_W3 + col]) +
              fmaxf(o.z2[1][e], 0.f) * bf2f(P[P_W3 + 16 + col]);
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
    o.z3[e] = s + bf2f(P[P_B3]);
  }
}

// Inference form on 32x32 MFMA tiles (32 samples per product), computed
// transposed: Z1^T = W1^T X^T (K = 8: six features, the bias input 1 and a zero;
// v_mfma_f32_32x32x8bf16_1k), Z2^T = W2^T relu(Z1^T) + b2 (K = 32 as two
// v_mfma_f32_32x32x16_bf16), z = w3 . relu(Z2^T) + b3.  Result item i of lane l
// of a 32x32 product is row d32_row(l/32, i), column l%32.  Those rows are the
// hidden units the lane holds; they go unchanged into the lane's B items of the
// next product, and the weights (the A operand) are loaded in the same unit
// order, so the hardware's K order never matters: A item t of lane (row + 32h)
// and B item t of lane (col + 32h) always meet.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int d32_row(int h, int i) { return 8 * (i >> 2) + 4 * h + (i & 3); }

// two floats -> packed BF16 pair (round to nearest even; a in the low half)
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __bf16 b2 __attribute__((ext_vector_type(2)));
  const f2 v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, b2));
}
// max(x, 0) of both halves of a BF16 pair: BF16 bit patterns order like int16
// for the sign test, so a packed int16 max against 0 is the ReLU
__device__ __forceinline__ unsigned relu_bf16x2(unsigned u) {
  typedef short s2 __attribute__((ext_vector_type(2)));
  const s2 z = {0, 0};
  return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(s2, u), z));
}
__device__ __forceinline__ float bf16_round(float f) { return bf2f(f2bf(f)); }

struct MlpRegs {
  bf16x4 w1;       // W1[4*(l/32)+t][l%32], t = 0..3 (features 7.. are zero)
  bf16x8 w2[2];    // product b: W2[d32_row(l/32, 8b+t)][l%32], t = 0..7
  f32x16 b2;       // b2[d32_row(l/32, i)]
  float w3[16];    // w3[d32_row(l/32, i)]
  float b3;
};

__device__ __forceinline__ void mlp_load_regs(const float* master, MlpRegs& r) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 31, h = lane >> 5;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int k = 4 * h + t;
    r.w1[t] = k < 7 ? f2bf(master[P_W1 + k * NH + col]) : (short)0;
  }
#pragma unroll
  for (int b = 0; b < 2; b++)
#pragma unroll
    for (int t = 0; t < 8; t++)
      r.w2[b][t] = (__bf16)master[P_W2 + d32_row(h, 8 * b + t) * NH + col];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    r.b2[i] = bf16_round(master[P_B2 + d32_row(h, i)]);
    r.w3[i] = bf16_round(master[P_W3 + d32_row(h, i)]);
  }
  r.b3 = bf16_round(master[P_B3]);
}

// 64 samples, one per lane.  lo / hi: this lane's own sample as BF16 features
// 0-3 and 4-7 (4, 5, the bias input 1, 0).  Tile 0 is samples 0..31, tile 1
// samples 32..63; returns the logit of the lane's own sample.
__device__ __forceinline__ float mlp_forward_regs(const MlpRegs& r, u32x2 lo, u32x2 hi) {
  const int lane = threadIdx.x & 63;
  const bool low = lane < 32;
  // B of tile 0: lanes 0-31 features 0-3 of sample l, lanes 32-63 features 4-7
  // of sample l-32; tile 1: lanes 0-31 features 0-3 of sample l+32, lanes 32-63
  // features 4-7 of sample l.  One exchange with lane l^32 supplies the rest.
  const u32x2 send = low ? hi : lo;
  u32x2 recv;
  recv.x = (unsigned)__shfl_xor((int)send.x, 32, 64);
  recv.y = (unsigned)__shfl_xor((int)send.y, 32, 64);
  const u32x2 x[2] = {low ? lo : recv, low ? recv : hi};
  float part[2];
#pragma unroll
  for (int tile = 0; tile < 2; tile++) {
    const f32x16 zero = {};
    const f32x16 z1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(
        r.w1, __builtin_bit_cast(bf16x4, x[tile]), zero, 0, 0, 0);
    u32x4 hb[2];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        hb[b][j] = relu_bf16x2(pack_bf16(z1[8 * b + 2 * j], z1[8 * b + 2 * j + 1]));
    f32x16 z2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        r.w2[0], __builtin_bit_cast(bf16x8, hb[0]), r.b2, 0, 0, 0);
    z2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
        r.w2[1], __builtin_bit_cast(bf16x8, hb[1]), z2, 0, 0, 0);
    // ReLU as a signed-integer max of the bit pattern (negative floats have the
    // sign bit set), then an explicit FMA: the products and sums of the FP32 dot
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++)
      s = fmaf(__int_as_float(max(__float_as_int(z2[i]), 0)), r.w3[i], s);
    part[tile] = s;
  }
  // lane l's own sample is column l%32 of tile l/32; its other half-sum sits in
  // lane l^32
  const float other = __shfl_xor(low ? part[1] : part[0], 32, 64);
  return (low ? part[0] : part[1]) + other + r.b3;
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + __expf(-z)); }

// ---------------------------------------------------------------- training
// Each wavefront takes tiles of 16 samples; gradients (FP32) are summed per
// wavefront and added to `grad` (P_N floats) with atomics.  Loss: weighted
// binary cross-entropy, mean over the batch.
__global__ void __launch_bounds__(256) k_mlp_train(const float* feat, long long nsamples,
                                                   const unsigned char* label,
                                                   const int* batch_idx, int nbatch,
                                                   const float* master, float pos_weight,
                                                   float* grad, float* loss_out) {
  __shared__ short sP[P_N + 1];
  __shared__ short sX[4][16 * NF];
  __shared__ short sH[4][16 * NH];
  __shared__ short sD[4][16 * NH];
  __shared__ float sG[P_N];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  for (int i = tid; i < P_N; i += 256) { sP[i] = f2bf(master[i]); sG[i] = 0.f; }
  __syncthreads();
  const int col = lane & 15, kq = 4 * (lane >> 4);
  const int ntiles = (nbatch + 15) / 16;
  const float invb = 1.f / (float)nbatch;
  float loss_acc = 0.f;
  // per-wave gradient accumulators in D layout
  f32x4 gW2[2][2], gW1[2];
  for (int a = 0; a < 2; a++) { gW1[a] = f32x4{0, 0, 0, 0}; for (int b = 0; b < 2; b++) gW2[a][b] = f32x4{0, 0, 0, 0}; }
  float gb2[2] = {0.f, 0.f}, gw3[2] = {0.f, 0.f}, gb3 = 0.f;
  for (int tile = blockIdx.x * 4 + wv; tile < ntiles; tile += gridDim.x * 4) {
    // stage X (bf16); batch entries outside [0, nsamples) count as padding
    for (int e = lane; e < 16 * NF; e += 64) {
      int r = e / NF, k = e % NF;
      int s = tile * 16 + r;
      float x = 0.f;
      if (s < nbatch && k < 8) {
        const int bi = batch_idx[s];
        if (bi >= 0 && bi < nsamples) x = feat[(long long)bi * 8 + k];
      }
      sX[wv][r * NF + k] = f2bf(x);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    Fwd f;
    mlp_forward(sP, sX[wv], sH[wv], f);
    // dz3 for rows kq+e
    float dz3[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int s = tile * 16 + kq + e;
      const int bi = s < nbatch ? batch_idx[s] : -1;
      const bool valid = bi >= 0 && bi < nsamples;
      float y = valid ? (float)label[bi] : 0.f;
      float p = sigmoidf(f.z3[e]);
      float w = pos_weight * y + (1.f - y);
      dz3[e] = valid ? (p * w - pos_weight * y) * invb : 0.f;
      if (valid && col == 0) {
        float pc = fminf(fmaxf(p, 1e-7f), 1.f - 1e-7f);
        loss_acc += -(pos_weight * y * __logf(pc) + (1.f - y) * __logf(1.f - pc)) * invb;
      }
      if (col == 0) gb3 += dz3[e];
    }
    // layer 3 grads and dZ2 = dz3 * w3 * [Z2 > 0]  (D layout), staged to sD [row][j]
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
      const float w3 = bf2f(sP[P_W3 + nb * 16 + col]);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float a2 = fmaxf(f.z2[nb][e], 0.f);
        gw3[nb] += a2 * dz3[e];
        float dz2 = f.z2[nb][e] > 0.f ? dz3[e] * w3 : 0.f;
        gb2[nb] += dz2;
        sD[wv][(kq + e) * NH + nb * 16 + col] = f2bf(dz2);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // dW2[k][j] += sum_s H1[s][k] dZ2[s][j]  (M = k, K = s, N = j)
#pragma unroll
    for (int mb = 0; mb < 2; mb++)
#pragma unroll
      for (int nb = 0; nb < 2; nb++) {
        bf16x4 a, b;
        // A = H1^T: row k = mb*16+col, samples kq..kq+3 -> H1[s][k]
        a.x = sH[wv][(kq + 0) * NH + mb * 16 + col]; a.y = sH[wv][(kq + 1) * NH + mb * 16 + col];
        a.z = sH[wv][(kq + 2) * NH + mb * 16 + col]; a.w = sH[wv][(kq + 3) * NH + mb * 16 + col];
        b = load_b(sD[wv], kq, nb * 16 + col, NH);
        gW2[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, gW2[mb][nb], 0, 0, 0);
      }
    // dA1[s][k] = sum_j dZ2[s][j] W2[k][j] ; dZ1 = dA1 * [Z1 > 0]
    f32x4 dA1[2];
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {   // nb: block of k
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 2; kb++) {  // kb: block of j
        bf16x4 a = load_a(sD[wv], col, kb * 16 + kq, NH);
        bf16x4 b;   // B[j][k] = W2[k][j]
        const int k = nb * 16 + col;
        b.x = sP[P_W2 + k * NH + kb * 16 + kq + 0]; b.y = sP[P_W2 + k * NH + kb * 16 + kq + 1];
        b.z = sP[P_W2 + k * NH + kb * 16 + kq + 2]; b.w = sP[P_W2 + k * NH + kb * 16 + kq + 3];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
      }
      dA1[nb] = acc;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float dz1 = f.z1[nb][e] > 0.f ? dA1[nb][e] : 0.f;
        sD[wv][(kq + e) * NH + nb * 16 + col] = f2bf(dz1);
      }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // dW1[f][k] += sum_s X[s][f] dZ1[s][k]  (M = f (16), K = s, N = k)
#pragma unroll
    for (int nb = 0; nb < 2; nb++) {
      bf16x4 a;
      a.x = sX[wv][(kq + 0) * NF + col]; a.y = sX[wv][(kq + 1) * NF + col];
      a.z = sX[wv][(kq + 2) * NF + col]; a.w = sX[wv][(kq + 3) * NF + col];
      bf16x4 b = load_b(sD[wv], kq, nb * 16 + col, NH);
      gW1[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, gW1[nb], 0, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
  // reduce the wave's accumulators into LDS, then global
#pragma unroll
  for (int mb = 0; mb < 2; mb++)
#pragma unroll
    for (int nb = 0; nb < 2; nb++)
#pragma unroll
      for (int e = 0; e < 4; e++)
        atomicAdd(&sG[P_W2 + (mb * 16 + kq + e) * NH + nb * 16 + col], gW2[mb][nb][e]);
#pragma unroll
  for (int nb = 0; nb < 2; nb++)
#pragma unroll
    for (int e = 0; e < 4; e++) atomicAdd(&sG[P_W1 + (kq + e) * NH + nb * 16 + col], gW1[nb][e]);
#pragma unroll
  for (int nb = 0; nb < 2; nb++) {
    atomicAdd(&sG[P_B2 + nb * 16 + col], gb2[nb]);
    atomicAdd(&sG[P_W3 + nb * 16 + col], gw3[nb]);
  }
  atomicAdd(&sG[P_B3], gb3);
  __syncthreads();
  for (int i = tid; i < P_N; i += 256) atomicAdd(&grad[i], sG[i]);
  // loss (lanes with col == 0 hold partial sums)
  for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off, 64);
  if (lane == 0) atomicAdd(loss_out, loss_acc);
}

// Adam on the FP32 master weights; clears the gradient buffer.  Rows 7..15 of W1
// (padding) are kept at zero.
__global__ void k_adam(float* master, float* grad, float* m1, float* m2, int step, float lr,
                       float b1, float b2, float eps) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P_N) return;
  float g = grad[i];
  grad[i] = 0.f;
  if (i < P_W2 && (i / NH) >= 7) { master[i] = 0.f; return; }
  float mm = b1 * m1[i] + (1.f - b1) * g;
  float vv = b2 * m2[i] + (1.f - b2) * g * g;
  m1[i] = mm;
  m2[i] = vv;
  float mh = mm / (1.f - powf(b1, (float)step));
  float vh = vv / (1.f - powf(b2, (float)step));
  master[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- inference
// Logits of a list of samples (for tests and evaluation).
__global__ void __launch_bounds__(256) k_mlp_logits(const float* feat, long long n,
                                                    const float* master, float* logits) {
  __shared__ short sP[P_N + 1];
  __shared__ short sX[4][16 * NF];
  __shared__ short sH[4][16 * NH];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  for (int i = tid; i < P_N; i += 256) sP[i] = f2bf(master[i]);
  __syncthreads();
  const long long ntiles = (n + 15) / 16;
  const int col = lane & 15, kq = 4 * (lane >> 4);
  for (long long tile = (long long)blockIdx.x * 4 + wv; tile < ntiles; tile += (long long)gridDim.x * 4) {
    for (int e = lane; e < 16 * NF; e += 64) {
      int r = e / NF, k = e % NF;
      long long s = tile * 16 + r;
      sX[wv][r * NF + k] = f2bf((s < n && k < 8) ? feat[s * 8 + k] : 0.f);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    Fwd f;
    mlp_forward(sP, sX[wv], sH[wv], f);
    if (col == 0) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        long long s = tile * 16 + kq + e;
        if (s < n) logits[s] = f.z3[e];
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
}

// Interface flags of a state: flag = 1 when any component whose stencil range
// exceeds `flat` has a logit above `zthr` (troubled -> full WENO).  Flags:
// x-interfaces (mx+1)*my, then y-interfaces mx*(my+1), row-major as in
// gather_stencil.  A wavefront takes ni = 64/meqn interfaces of one row per
// pass (x-interfaces c = 1..mx+1 of rows 1..my, then y-interfaces c = 1..mx of
// rows 1..my+1): lane l gathers component l%meqn of interface l/meqn (the
// spare lanes repeat the last interface), the 64 samples go through the MLP as
// two 32-sample tiles, and a ballot of the hits gives each interface its
// components.  A wavefront loads its next pass's stencils before it evaluates
// the current one; the pass position is wavefront-uniform, so the row, the
// column block and whether the stencils need the boundary map are scalar.
struct FlagPass {
  int dir, r, cb;
};

template <typename T>
__device__ __forceinline__ void flag_stencils(const T* q, const Geom& g, const BCSpec& bc,
                                              const FlagPass& p, int ni, int li, int m,
                                              T (&v)[6]) {
  const int cmax = p.dir == 0 ? g.mx + 1 : g.mx;
  const int c = min(1 + p.cb * ni + li, cmax);
  const bool interior = p.dir == 0 ? (p.cb * ni >= 3 && p.cb * ni + ni + 2 <= g.mx)
                                   : (p.r >= 4 && p.r + 2 <= g.my);
  if (interior) {
    const T* p0 = q + (long long)m * g.fs + (long long)(p.r - 1) * g.rs + (c - 1);
    const long long step = p.dir == 0 ? 1 : g.rs;
#pragma unroll
    for (int s = 0; s < 6; s++) v[s] = p0[(s - 3) * step];
  } else {
    gather_stencil_at(q, g, bc, p.dir, c, p.r, m, v);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_mlp_flags(const T* q, Geom g, BCSpec bc, int meqn,
                                                   float flat, const float* master, float zthr,
                                                   unsigned char* flags, const BufSel sel) {
  if (sel.gate) {
    if (!*sel.gate) return;
    q = (const T*)sel.q[*sel.cur];
  }
  if (sel.zero_count && blockIdx.x == 0 && threadIdx.x == 0) *sel.zero_count = 0;
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int stride = gridDim.x * 4;
  const int mx = g.mx, my = g.my;
  const int ni = 64 / meqn;
  const int npx = (mx + 1 + ni - 1) / ni, npy = (mx + ni - 1) / ni;
  const int nxp = npx * my;
  const int npass = nxp + npy * (my + 1);
  if (wid >= npass) return;
  MlpRegs reg;
  mlp_load_regs(master, reg);
  const int li = min(lane / meqn, ni - 1), m = lane % meqn;
  const unsigned long long cmask = meqn >= 64 ? ~0ull : (1ull << meqn) - 1ull;
  auto locate = [&](int pass) {
    FlagPass p;
    if (pass < nxp) {
      p.dir = 0; p.r = 1 + pass / npx; p.cb = pass % npx;
    } else {
      p.dir = 1; p.r = 1 + (pass - nxp) / npy; p.cb = (pass - nxp) % npy;
    }
    return p;
  };
  FlagPass cur = locate(wid);
  T v[6];
  flag_stencils(q, g, bc, cur, ni, li, m, v);
  // passes are interleaved across the grid (wid, wid + stride, ...), which keeps
  // the waves balanced where whole smooth passes skip the network
  for (int pass = wid; pass < npass; pass += stride) {
    const bool more = pass + stride < npass;
    const FlagPass nxt = locate(more ? pass + stride : pass);
    T vn[6];
    if (more) flag_stencils(q, g, bc, nxt, ni, li, m, vn);
    float f[6];
    const bool rough = norm_stencil(v, flat, f) > flat;
    unsigned long long hits = 0;
    if (__ballot(rough) != 0ull) {
      const u32x2 lo = {pack_bf16(f[0], f[1]), pack_bf16(f[2], f[3])};
      const u32x2 hi = {pack_bf16(f[4], f[5]), 0x3f80u};   // bias input 1.0, then 0
      const float z = mlp_forward_regs(reg, lo, hi);
      hits = __ballot(rough && z > zthr);
    }
    const int cmax = cur.dir == 0 ? mx + 1 : mx;
    const int ci = 1 + cur.cb * ni + lane;
    if (lane < ni && ci <= cmax) {
      const long long fi = cur.dir == 0
                               ? (long long)(cur.r - 1) * (mx + 1) + (ci - 1)
                               : (long long)(mx + 1) * my + (long long)(cur.r - 1) * mx + (ci - 1);
      flags[fi] = ((hits >> (lane * meqn)) & cmask) != 0ull ? 1 : 0;
    }
    if (more) {
      cur = nxt;
#pragma unroll
      for (int s = 0; s < 6; s++) v[s] = vn[s];
    }
  }
}

// Flags with one interface per lane and its ME components in registers (ME <=
// 4): a pass takes 64 consecutive interfaces of one row, and the ME component
// samples of the pass go through the network as ME independent 64-sample
// evaluations, which the scheduler interleaves.  Same flags as k_mlp_flags.
template <typename T, int ME>
__global__ void __launch_bounds__(256) k_mlp_flags_row(const T* q, Geom g, BCSpec bc, float flat,
                                                       const float* master, float zthr,
                                                       unsigned char* flags, const BufSel sel) {
  if (sel.gate) {
    if (!*sel.gate) return;
    q = (const T*)sel.q[*sel.cur];
  }
  if (sel.zero_count && blockIdx.x == 0 && threadIdx.x == 0) *sel.zero_count = 0;
  const int lane = threadIdx.x & 63;
  const int wid = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int stride = gridDim.x * 4;
  const int mx = g.mx, my = g.my;
  const int npx = (mx + 1 + 63) / 64, npy = (mx + 63) / 64;
  const int nxp = npx * my;
  const int npass = nxp + npy * (my + 1);
  if (wid >= npass) return;
  MlpRegs reg;
  mlp_load_regs(master, reg);
  for (int pass = wid; pass < npass; pass += stride) {
    int dir, r, cb;
    if (pass < nxp) {
      dir = 0; r = 1 + pass / npx; cb = pass % npx;
    } else {
      dir = 1; r = 1 + (pass - nxp) / npy; cb = (pass - nxp) % npy;
    }
    const int cmax = dir == 0 ? mx + 1 : mx;
    const int ci = 1 + cb * 64 + lane;
    const int c = min(ci, cmax);
    const bool interior = dir == 0 ? (cb * 64 >= 3 && cb * 64 + 64 + 2 <= mx)
                                   : (r >= 4 && r + 2 <= my);
    T v[ME][6];
    if (interior) {
      const long long step = dir == 0 ? 1 : g.rs;
#pragma unroll
      for (int m = 0; m < ME; m++) {
        const T* p0 = q + (long long)m * g.fs + (long long)(r - 1) * g.rs + (c - 1);
#pragma unroll
        for (int k = 0; k < 6; k++) v[m][k] = p0[(k - 3) * step];
      }
    } else {
#pragma unroll
      for (int m = 0; m < ME; m++) gather_stencil_at(q, g, bc, dir, c, r, m, v[m]);
    }
    bool rough[ME];
    u32x2 lo[ME], hi[ME];
#pragma unroll
    for (int m = 0; m < ME; m++) {
      float f[6];
      rough[m] = norm_stencil(v[m], flat, f) > flat;
      lo[m] = u32x2{pack_bf16(f[0], f[1]), pack_bf16(f[2], f[3])};
      hi[m] = u32x2{pack_bf16(f[4], f[5]), 0x3f80u};   // bias input 1.0, then 0
    }
    bool hit = false;
    // a component's network runs only when one of the pass's 64 samples of it
    // is rough (the others cannot set a flag): smooth fields, as the momenta
    // away from moving fronts, skip their evaluation
#pragma unroll
    for (int m = 0; m < ME; m++) {
      if (__ballot(rough[m]) != 0ull) {
        const float z = mlp_forward_regs(reg, lo[m], hi[m]);
        hit = hit || (rough[m] && z > zthr);
      }
    }
    if (ci <= cmax) {
      const long long fi = dir == 0 ? (long long)(r - 1) * (mx + 1) + (ci - 1)
                                    : (long long)(mx + 1) * my + (long long)(r - 1) * mx + (ci - 1);
      flags[fi] = hit ? 1 : 0;
    }
  }
}

// ---------------------------------------------------------------- launchers
int launch_features_labels(int dtype, const void* q, const Geom& g, const int* lo, const int* hi,
                           const int* refl, int meqn, float flat, float wtol, float* feat,
                           unsigned char* label, hipStream_t s) {
  BCSpec bc = to_bc(lo, hi, refl, 2);
  long long n = ((long long)(g.mx + 1) * g.my + (long long)g.mx * (g.my + 1)) * meqn;
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (dtype == CLAWHIP_F64)
    hipLaunchKernelGGL(k_features_labels<double>, dim3(blocks), dim3(256), 0, s, (const double*)q,
                       g, bc, meqn, flat, wtol, feat, label);
  else
    hipLaunchKernelGGL(k_features_labels<float>, dim3(blocks), dim3(256), 0, s, (const float*)q,
                       g, bc, meqn, flat, wtol, feat, label);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

int launch_mlp_train(const float* feat, long long nsamples, const unsigned char* label,
                     const int* batch_idx, int nbatch, float* master, float* grad, float* m1,
                     float* m2, int step, float lr, float pos_weight, float* loss, hipStream_t s) {
  int tiles = (nbatch + 15) / 16;
  int blocks = (tiles + 3) / 4;
  if (blocks > 256) blocks = 256;
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, s, feat, nsamples, label,
                     batch_idx, nbatch, master, pos_weight, grad, loss);
  if (step > 0)
    hipLaunchKernelGGL(k_adam, dim3((P_N + 255) / 256), dim3(256), 0, s, master, grad, m1, m2,
                       step, lr, 0.9f, 0.999f, 1e-8f);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

int launch_mlp_logits(const float* feat, long long n, const float* master, float* logits,
                      hipStream_t s) {
  long long tiles = (n + 15) / 16;
  long long blocks = (tiles + 3) / 4;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_mlp_logits, dim3((unsigned)blocks), dim3(256), 0, s, feat, n, master,
                     logits);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

int launch_mlp_flags(int dtype, const void* q, const Geom& g, const int* lo, const int* hi,
                     const int* refl, int meqn, float flat, const float* master, float zthr,
                     unsigned char* flags, hipStream_t s, const BufSel& sel) {
  BCSpec bc = to_bc(lo, hi, refl, 2);
  if (meqn < 1 || meqn > 64) return CLAWHIP_E_ARG;
  const int ni = 64 / meqn;
  long long passes = (long long)((g.mx + 1 + ni - 1) / ni) * g.my +
                     (long long)((g.mx + ni - 1) / ni) * (g.my + 1);
  if (passes + 4LL * 65536 >= (1LL << 31)) return CLAWHIP_E_ARG;   // 32-bit pass index
  // grid-stride: one resident round of wavefronts (measured at 2048^2: one
  // round 85 us, two 103 us, four 93 us, half a round 119 us), each loading the
  // weights once
  long long cap = 2048;
  if (const char* env = getenv("CLAWHIP_FLAGS_BLOCKS")) {
    const long long b = atoll(env);
    if (b >= 1 && b <= 65536) cap = b;
  }
  long long blocks = (passes + 3) / 4;
  if (blocks > cap) blocks = cap;
  if (meqn <= 4) {
    const long long rpass = (long long)((g.mx + 1 + 63) / 64) * g.my +
                            (long long)((g.mx + 63) / 64) * (g.my + 1);
    long long rblocks = (rpass + 3) / 4;
    if (!getenv("CLAWHIP_FLAGS_BLOCKS")) {
      const void* kern = dtype == CLAWHIP_F64
          ? (meqn == 1 ? (const void*)k_mlp_flags_row<double, 1> : meqn == 2 ? (const void*)k_mlp_flags_row<double, 2>
             : meqn == 3 ? (const void*)k_mlp_flags_row<double, 3> : (const void*)k_mlp_flags_row<double, 4>)
          : (meqn == 1 ? (const void*)k_mlp_flags_row<float, 1> : meqn == 2 ? (const void*)k_mlp_flags_row<float, 2>
             : meqn == 3 ? (const void*)k_mlp_flags_row<float, 3> : (const void*)k_mlp_flags_row<float, 4>);
      const long long resident = resident_waves(kern, 0);
      if (resident >= 4) cap = resident / 4;
    }
    if (rblocks > cap) rblocks = cap;
    const dim3 grid((unsigned)rblocks);
    if (dtype == CLAWHIP_F64) {
      const double* qd = (const double*)q;
      switch (meqn) {
        case 1: hipLaunchKernelGGL((k_mlp_flags_row<double, 1>), grid, dim3(256), 0, s, qd, g, bc, flat, master, zthr, flags, sel); break;
        case 2: hipLaunchKernelGGL((k_mlp_flags_row<double, 2>), grid, dim3(256), 0, s, qd, g, bc, flat, master, zthr, flags, sel); break;
        case 3: hipLaunchKernelGGL((k_mlp_flags_row<double, 3>), grid, dim3(256), 0, s, qd, g, bc, flat, master, zthr, flags, sel); break;
        default: hipLaunchKernelGGL((k_mlp_flags_row<double, 4>), grid, dim3(256), 0, s, qd, g, bc, flat, master, zthr, flags, sel); break;
      }
    } else {
      const float* qf = (const float*)q;
      switch (meqn) {
        case 1: hipLaunchKernelGGL((k_mlp_flags_row<float, 1>), grid, dim3(256), 0, s, qf, g, bc, flat, master, zthr, flags, sel); break;
        case 2: hipLaunchKernelGGL((k_mlp_flags_row<float, 2>), grid, dim3(256), 0, s, qf, g, bc, flat, master, zthr, flags, sel); break;
        case 3: hipLaunchKernelGGL((k_mlp_flags_row<float, 3>), grid, dim3(256), 0, s, qf, g, bc, flat, master, zthr, flags, sel); break;
        default: hipLaunchKernelGGL((k_mlp_flags_row<float, 4>), grid, dim3(256), 0, s, qf, g, bc, flat, master, zthr, flags, sel); break;
      }
    }
  } else if (dtype == CLAWHIP_F64) {
    hipLaunchKernelGGL(k_mlp_flags<double>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const double*)q, g, bc, meqn, flat, master, zthr, flags, sel);
  } else {
    hipLaunchKernelGGL(k_mlp_flags<float>, dim3((unsigned)blocks), dim3(256), 0, s,
                       (const float*)q, g, bc, meqn, flat, master, zthr, flags, sel);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace claw
