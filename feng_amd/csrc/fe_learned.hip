// Learned correction of the split IC(0) preconditioner (FE_PC_LEARNED).
//
// The preconditioner applied by the split PCG is M^-1 = G^T D^-1 G with G unit
// lower triangular on the pattern of A (G = I - L_strict from IC(0)) and D the
// IC(0) pivots; it is SPD for any G with a unit diagonal and any D > 0, so a
// correction of G and D can never make PCG break down.  A message-passing
// network on the matrix graph predicts the correction, in the coordinates of
// the unit-diagonal matrix A^ = S A S (s_i = a_ii^-1/2), on the strict lower
// entries (i, j), j < i:
//   w0_ij = (A^ g^0_i)_j                   (row i of G^0 times A^)
//   g^_ij = g^0_ij - (sigma_i + rho_j) w0_ij + a^_ij (p_i . q_j + s_i + r_j)
//   d^_i  = B_ii exp(clamp(tau_i, -4, 4)),  B_ii = g^_i^T A^ g^_i
//   x_i   = [1, d^0_i, sum |a^_ij|, sum a^_ij, deg_i/16, sum_{j<i} g^0_ij^2,
//            sum_{j<i} w0_ij^2, 0..]
//   h0 = relu(x W0 + b0),  m_l = A^_off h_{l-1},
//   h_l = relu([h_{l-1}, m_l] W_l + b_l)  (l = 1..3),
//   p = h3 P, q = h3 Q (8 wide), [tau, s, r, sigma, rho] = h3 Hs + hb.
// w0 is the gradient of the loss below at the IC(0) factor, so the network
// learns step lengths along it plus corrections proportional to a^.
// Activations are stored as bfloat16; the dense layers and both GEMMs of
// their backward run on v_mfma_f32_16x16x16bf16_1k with FP32 accumulation,
// the heads, the factor and the loss in FP32.
//
// Training minimises Kaporin's condition number of the preconditioned
// operator A~ = D^^-1/2 G^ A^ G^^T D^^-1/2, K = (tr A~ / n)^n / det A~, a
// bound on the CG iteration count (Kaporin 1994); up to a constant,
//   L = log K / n = log((1/n) sum_i e^-tau_i) + mean(tau) + mean(log B_ii),
// whose minimiser over the pattern is the FSAI factor (Kolotilina & Yeremin
// 1993).  dL/dg^_ij = 2 w_ij / (n B_ii) with w_ij = (A^ g^_i)_j.  The backward
// is written out by hand; every parameter gradient is a per-block partial
// folded by a second kernel (no atomics, no fences).
#include "fe_hip_internal.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kH = 32, kNin = 16, kNpq = 8, kNs = 5;
constexpr int oW0 = 0, ob0 = oW0 + kNin * kH, oW1 = ob0 + kH, ob1 = oW1 + 2 * kH * kH, oW2 = ob1 + kH,
              ob2 = oW2 + 2 * kH * kH, oW3 = ob2 + kH, ob3 = oW3 + 2 * kH * kH, oP = ob3 + kH,
              oQ = oP + kH * kNpq, oHs = oQ + kH * kNpq, ohb = oHs + kH * kNs, kParams = ohb + kNs;
constexpr int kHeadW = kParams - oP;           // P, Q, Hs, hb: contiguous
constexpr int kW16 = kNin * kH + 6 * kH * kH;  // bf16 copies of W0..W3, packed
constexpr int kW16Off[4] = {0, kNin * kH, kNin * kH + 2 * kH * kH, kNin * kH + 4 * kH * kH};
constexpr int kWOff[4] = {oW0, oW1, oW2, oW3};
constexpr int kBOff[4] = {ob0, ob1, ob2, ob3};
constexpr float kTauClamp = 4.f;
constexpr int kGemmGrid = 128;  // blocks of the parameter-gradient GEMMs
constexpr int kLossGrid = 256;  // blocks of the loss reduction

typedef short feShort4 __attribute__((ext_vector_type(4)));
typedef float feF4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// FP32 -> bfloat16, round to nearest even.
__device__ __forceinline__ bf16_t toBf16(float x)
{
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float fromBf16(bf16_t b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ feF4 mfmaBf16(feShort4 a, feShort4 b, feF4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Graph quantities.

// s_i = a_ii^-1/2, d^0_i = u_ii / a_ii.
__global__ void mpScale(int n, const int *__restrict__ diag, const double *__restrict__ val,
                        const double *__restrict__ lu, double *__restrict__ s, double *__restrict__ d0)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double a = val[diag[i]];
  s[i] = 1. / sqrt(a);
  d0[i] = lu[diag[i]] / a;
}

// a^ on the pattern, g^0 on the strict lower entries (one thread per entry).
__global__ void mpNormalise(int64_t nnz, const int *__restrict__ rowOf, const int *__restrict__ ja,
                            const double *__restrict__ val, const double *__restrict__ lu,
                            const double *__restrict__ s, float *__restrict__ ahat, float *__restrict__ g0)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= nnz) return;
  const int i = rowOf[p], j = ja[p];
  const double si = s[i], sj = s[j];
  ahat[p] = (float)(val[p] * si * sj);
  if(j < i) g0[p] = (float)(-lu[p] * si / sj);
}

// B_ii = g^_i^T A^ g^_i and w_e = (A^ g^_i)_j on the strict lower entries
// e = (i, j) of row i, for G^ = I + g (strict lower values on the pattern).
// Row i's factor pattern P_i = its entries up to the diagonal.  8 lanes per
// row: lane q takes the entries e = ia[i] + q, + 8, ... of P_i and forms
// (A^ g^_i)_j by merging row j of A^ with P_i (both sorted); B_ii is the
// 8-lane sum of g^_ie (A^ g^_i)_j.
__global__ __launch_bounds__(256) void mpRows(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                              const int *__restrict__ diag, const float *__restrict__ ahat,
                                              const float *__restrict__ g, float *__restrict__ B,
                                              float *__restrict__ w)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t >> 3, lane = t & 7;
  if(i >= n) return;  // whole 8-lane groups leave together
  const int p0 = ia[i], pd = diag[i];
  double b = 0.;
  for(int e = p0 + lane; e <= pd; e += 8) {
    const int j = ja[e];
    double tt = 0.;
    int a = ia[j], q = p0;
    const int ae = ia[j + 1];
    while(a < ae && q <= pd) {
      const int ca = ja[a], cq = ja[q];
      if(ca == cq) {
        tt += (double)ahat[a] * (q == pd ? 1. : (double)g[q]);
        ++a;
        ++q;
      } else if(ca < cq) {
        ++a;
      } else {
        ++q;
      }
    }
    b += (e == pd ? 1. : (double)g[e]) * tt;
    if(e < pd) w[e] = (float)tt;
  }
#pragma unroll
  for(int o = 4; o > 0; o >>= 1) b += __shfl_xor(b, o, 8);
  if(lane == 0) B[i] = (float)b;
}

// Input features.
__global__ void mpFeatures(int n, const int *__restrict__ ia, const int *__restrict__ ja, const int *__restrict__ diag,
                           const float *__restrict__ ahat, const float *__restrict__ g0, const float *__restrict__ w0,
                           const double *__restrict__ d0, bf16_t *__restrict__ X)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double sa = 0., ss = 0., sg = 0., sw = 0.;
  int deg = 0;
  for(int p = ia[i]; p < ia[i + 1]; ++p) {
    if(ja[p] == i) continue;
    const double ah = ahat[p];
    sa += fabs(ah);
    ss += ah;
    ++deg;
  }
  for(int p = ia[i]; p < diag[i]; ++p) {
    sg += (double)g0[p] * g0[p];
    sw += (double)w0[p] * w0[p];
  }
  bf16_t *x = X + (size_t)i * kNin;
  x[0] = toBf16(1.f);
  x[1] = toBf16((float)d0[i]);
  x[2] = toBf16((float)sa);
  x[3] = toBf16((float)ss);
  x[4] = toBf16((float)(deg / 16.));
  x[5] = toBf16((float)sg);
  x[6] = toBf16((float)sw);
#pragma unroll
  for(int k = 7; k < kNin; ++k) x[k] = 0;
}

// ---------------------------------------------------------------------------
// Forward.

// H = bf16(relu([A1 | A2] W + b)), W KIN x 32 (bf16), one wavefront per 16
// rows.  KIN = 16: A1 is the n x 16 input; KIN = 64: A1 = h (n x 32),
// A2 = m (n x 32).  Operand layouts of v_mfma_f32_16x16x16bf16_1k: lane l
// holds A[l % 16][4 (l/16) + e] and B[4 (l/16) + e][l % 16], e = 0..3; the
// result D[4 (l/16) + v][l % 16] is in register v.
template <int KIN>
__global__ __launch_bounds__(256) void mpDense(int n, const bf16_t *__restrict__ A1, const bf16_t *__restrict__ A2,
                                               const bf16_t *__restrict__ W, const float *__restrict__ bias,
                                               bf16_t *__restrict__ Hout)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  const int i = lane & 15, kq = lane >> 4;
  const int row = row0 + i;
  feF4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
  for(int s = 0; s < KIN / 16; ++s) {
    feShort4 a = {0, 0, 0, 0};
    if(row < n) {
      const bf16_t *src;
      if(KIN == 16) src = A1 + (size_t)row * 16 + 4 * kq;
      else if(s < 2) src = A1 + (size_t)row * 32 + 16 * s + 4 * kq;
      else src = A2 + (size_t)row * 32 + 16 * (s - 2) + 4 * kq;
      a = *(const feShort4 *)src;
    }
#pragma unroll
    for(int t = 0; t < 2; ++t) {
      feShort4 b;
#pragma unroll
      for(int e = 0; e < 4; ++e) b[e] = (short)W[(16 * s + 4 * kq + e) * 32 + 16 * t + i];
      acc[t] = mfmaBf16(a, b, acc[t]);
    }
  }
#pragma unroll
  for(int t = 0; t < 2; ++t)
#pragma unroll
    for(int v = 0; v < 4; ++v) {
      const int r = row0 + 4 * kq + v, c = 16 * t + i;
      if(r < n) Hout[(size_t)r * 32 + c] = toBf16(fmaxf(acc[t][v] + bias[c], 0.f));
    }
}

// M = bf16(A^_off H), 16 lanes per row, 2 features per lane.
__global__ __launch_bounds__(256) void mpAggregate(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                   const float *__restrict__ ahat, const bf16_t *__restrict__ Hin,
                                                   bf16_t *__restrict__ Mout)
{
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int row = t >> 4, f = (t & 15) * 2;
  if(row >= n) return;
  float s0 = 0.f, s1 = 0.f;
  for(int p = ia[row]; p < ia[row + 1]; ++p) {
    const int j = ja[p];
    if(j == row) continue;
    const float a = ahat[p];
    const unsigned w = *(const unsigned *)(Hin + (size_t)j * 32 + f);
    s0 += a * fromBf16((bf16_t)(w & 0xFFFFu));
    s1 += a * fromBf16((bf16_t)(w >> 16));
  }
  *(unsigned *)(Mout + (size_t)row * 32 + f) = (unsigned)toBf16(s0) | ((unsigned)toBf16(s1) << 16);
}

// p = h3 P, q = h3 Q, hs = h3 Hs + hb (FP32, one thread per row).
__global__ __launch_bounds__(256) void mpHeads(int n, const bf16_t *__restrict__ h3, const float *__restrict__ theta,
                                               float *__restrict__ p, float *__restrict__ q, float *__restrict__ hs)
{
  __shared__ float w[kHeadW];
  for(int k = threadIdx.x; k < kHeadW; k += blockDim.x) w[k] = theta[oP + k];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  float h[kH];
#pragma unroll
  for(int f = 0; f < kH; ++f) h[f] = fromBf16(h3[(size_t)i * kH + f]);
#pragma unroll
  for(int k = 0; k < kNpq; ++k) {
    float a = 0.f, b = 0.f;
#pragma unroll
    for(int f = 0; f < kH; ++f) {
      a += h[f] * w[f * kNpq + k];
      b += h[f] * w[oQ - oP + f * kNpq + k];
    }
    p[(size_t)i * kNpq + k] = a;
    q[(size_t)i * kNpq + k] = b;
  }
#pragma unroll
  for(int c = 0; c < kNs; ++c) {
    float a = w[ohb - oP + c];
#pragma unroll
    for(int f = 0; f < kH; ++f) a += h[f] * w[oHs - oP + f * kNs + c];
    hs[(size_t)i * kNs + c] = a;
  }
}

// g^ = g^0 - (sigma_i + rho_j) w0 + a^ (p_i . q_j + s_i + r_j) on strict lower
// entries (one thread per entry).
__global__ void mpCorrect(int64_t nnz, const int *__restrict__ rowOf, const int *__restrict__ ja,
                          const float *__restrict__ ahat, const float *__restrict__ g0, const float *__restrict__ w0,
                          const float *__restrict__ p, const float *__restrict__ q, const float *__restrict__ hs,
                          float *__restrict__ g)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= nnz) return;
  const int i = rowOf[e], j = ja[e];
  if(j >= i) return;
  float c = hs[(size_t)i * kNs + 1] + hs[(size_t)j * kNs + 2];
#pragma unroll
  for(int k = 0; k < kNpq; ++k) c += p[(size_t)i * kNpq + k] * q[(size_t)j * kNpq + k];
  g[e] = g0[e] - (hs[(size_t)i * kNs + 3] + hs[(size_t)j * kNs + 4]) * w0[e] + ahat[e] * c;
}

// d^_i = B_ii exp(clamp(tau_i)).
__global__ void mpPivots(int n, const float *__restrict__ B, const float *__restrict__ hs, float *__restrict__ d)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float t = fminf(fmaxf(hs[(size_t)i * kNs], -kTauClamp), kTauClamp);
  d[i] = (float)((double)B[i] * exp((double)t));
}

// Corrected split factor in the original coordinates, written over the IC(0)
// one: g_ij = g^_ij s_j / s_i into the lower rows of G and the upper rows of
// G^T, dinv_i = 1 / (d^_i a_ii) (one thread per entry of A).
__global__ void mpWriteFactor(int64_t nnz, const int *__restrict__ rowOf, const int *__restrict__ ia,
                              const int *__restrict__ ja, const int *__restrict__ diag, const int *__restrict__ tpos,
                              const double *__restrict__ val, const double *__restrict__ s, const float *__restrict__ g,
                              const float *__restrict__ d, const int *__restrict__ iaL, const int *__restrict__ iaU,
                              double *__restrict__ gL64, float *__restrict__ gL32, double *__restrict__ gU64,
                              float *__restrict__ gU32, double *__restrict__ dinv64, float *__restrict__ dinv32)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= nnz) return;
  const int i = rowOf[e], j = ja[e];
  if(j < i) {
    const double gij = (double)g[e] * s[j] / s[i];
    const int lpos = iaL[i] + (int)(e - ia[i]);
    gL64[lpos] = gij;
    gL32[lpos] = (float)gij;
    const int t = tpos[e];  // entry (j, i), an upper entry of row j
    const int upos = iaU[j] + (t - diag[j]);
    gU64[upos] = gij;
    gU32[upos] = (float)gij;
  } else if(j == i) {
    const double di = (double)d[i] * val[e];
    dinv64[i] = 1. / di;
    dinv32[i] = (float)(1. / di);
  }
}

// ---------------------------------------------------------------------------
// Loss and its gradient.

// Per-block partials of sum log B_i, sum e^-tau_i, sum tau_i (tau clamped).
__global__ __launch_bounds__(256) void mpLossParts(int n, const float *__restrict__ B, const float *__restrict__ hs,
                                                   double *__restrict__ part)
{
  __shared__ double lds[3][4];
  double lb = 0., et = 0., tt = 0.;
  for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double t = fmin(fmax((double)hs[(size_t)i * kNs], -(double)kTauClamp), (double)kTauClamp);
    lb += log((double)B[i]);
    et += exp(-t);
    tt += t;
  }
  lb = feWaveSum(lb);
  et = feWaveSum(et);
  tt = feWaveSum(tt);
  if((threadIdx.x & 63) == 0) {
    lds[0][threadIdx.x >> 6] = lb;
    lds[1][threadIdx.x >> 6] = et;
    lds[2][threadIdx.x >> 6] = tt;
  }
  __syncthreads();
  if(threadIdx.x < 3)
    part[blockIdx.x * 3 + threadIdx.x] =
        lds[threadIdx.x][0] + lds[threadIdx.x][1] + lds[threadIdx.x][2] + lds[threadIdx.x][3];
}

// out[0] = L, out[1] = sum e^-tau (for the tau gradient).
__global__ void mpLossFold(int n, int G, const double *__restrict__ part, double *__restrict__ out)
{
  if(threadIdx.x != 0 || blockIdx.x != 0) return;
  double lb = 0., et = 0., tt = 0.;
  for(int b = 0; b < G; ++b) {
    lb += part[3 * b];
    et += part[3 * b + 1];
    tt += part[3 * b + 2];
  }
  out[0] = log(et / n) + tt / n + lb / n;
  out[1] = et;
}

// Head gradients into gHead (n x 32: gp at 0..7, gq at 8..15, then gtau, gs,
// gr, gsigma, grho at 16..20), with ge_ij = 2 w_ij / (n B_ii) on the strict
// lower entries, gc = ge a^, gsr = -ge w0:
//   gp_i = sum_{j<i} gc_ij q_j, gs_i = sum_{j<i} gc_ij, gsigma_i = sum_{j<i} gsr_ij,
//   gq_i = sum_{j>i} gc_ji p_j, gr_i = sum_{j>i} gc_ji, grho_i = sum_{j>i} gsr_ji,
//   gtau_i = (1 - n e^-tau_i / sum e^-tau) / n inside the clamp, else 0.
__global__ __launch_bounds__(256) void mpEdgeGrad(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                  const int *__restrict__ diag, const int *__restrict__ tpos,
                                                  const float *__restrict__ ahat, const float *__restrict__ w,
                                                  const float *__restrict__ w0, const float *__restrict__ B,
                                                  const float *__restrict__ p, const float *__restrict__ q,
                                                  const float *__restrict__ hs, const double *__restrict__ lossOut,
                                                  float *__restrict__ gHead)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const float inv = 2.f / (float)n;
  float gp[kNpq], gq[kNpq];
#pragma unroll
  for(int k = 0; k < kNpq; ++k) gp[k] = gq[k] = 0.f;
  float gs = 0.f, gr = 0.f, gsig = 0.f, grho = 0.f;
  const float bi = B[i];
  for(int e = ia[i]; e < diag[i]; ++e) {
    const int j = ja[e];
    const float ge = inv * w[e] / bi;
    const float gc = ge * ahat[e];
#pragma unroll
    for(int k = 0; k < kNpq; ++k) gp[k] += gc * q[(size_t)j * kNpq + k];
    gs += gc;
    gsig -= ge * w0[e];
  }
  for(int e = diag[i] + 1; e < ia[i + 1]; ++e) {
    const int j = ja[e];
    const int t = tpos[e];  // entry (j, i), strict lower in row j
    const float ge = inv * w[t] / B[j];
    const float gc = ge * ahat[t];
#pragma unroll
    for(int k = 0; k < kNpq; ++k) gq[k] += gc * p[(size_t)j * kNpq + k];
    gr += gc;
    grho -= ge * w0[t];
  }
  float *gh = gHead + (size_t)i * 32;
#pragma unroll
  for(int k = 0; k < kNpq; ++k) {
    gh[k] = gp[k];
    gh[8 + k] = gq[k];
  }
  const float tv = hs[(size_t)i * kNs];
  const double S = lossOut[1];
  gh[16] = (tv > -kTauClamp && tv < kTauClamp)
               ? (float)((1. - (double)n * exp(-(double)tv) / S) / (double)n)
               : 0.f;
  gh[17] = gs;
  gh[18] = gr;
  gh[19] = gsig;
  gh[20] = grho;
}

// gz3 = bf16(mask(h3 > 0) (gp P^T + gq Q^T + ghs Hs^T)), 32 lanes per row.
__global__ __launch_bounds__(256) void mpHeadsBwd(int n, const bf16_t *__restrict__ h3, const float *__restrict__ gHead,
                                                  const float *__restrict__ theta, bf16_t *__restrict__ gz)
{
  __shared__ float w[kHeadW];
  for(int k = threadIdx.x; k < kHeadW; k += blockDim.x) w[k] = theta[oP + k];
  __syncthreads();
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i = t >> 5, f = t & 31;
  if(i >= n) return;
  const float *gh = gHead + (size_t)i * 32;
  float s = 0.f;
#pragma unroll
  for(int k = 0; k < kNpq; ++k) s += gh[k] * w[f * kNpq + k] + gh[8 + k] * w[oQ - oP + f * kNpq + k];
#pragma unroll
  for(int c = 0; c < kNs; ++c) s += gh[16 + c] * w[oHs - oP + f * kNs + c];
  gz[(size_t)i * 32 + f] = fromBf16(h3[(size_t)i * 32 + f]) > 0.f ? toBf16(s) : (bf16_t)0;
}

// gX = gz W^T (n x KIN, FP32) on the matrix cores: A = gz rows (K = 32),
// B[k][j] = W[j][k].
template <int KIN>
__global__ __launch_bounds__(256) void mpDenseBwdX(int n, const bf16_t *__restrict__ gz, const bf16_t *__restrict__ W,
                                                   float *__restrict__ gX)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + wave) * 16;
  const int i = lane & 15, kq = lane >> 4;
  const int row = row0 + i;
  feF4 acc[KIN / 16];
#pragma unroll
  for(int t = 0; t < KIN / 16; ++t) acc[t] = (feF4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for(int s = 0; s < 2; ++s) {
    feShort4 a = {0, 0, 0, 0};
    if(row < n) a = *(const feShort4 *)(gz + (size_t)row * 32 + 16 * s + 4 * kq);
#pragma unroll
    for(int t = 0; t < KIN / 16; ++t) {
      const feShort4 b = *(const feShort4 *)(W + (size_t)(16 * t + i) * 32 + 16 * s + 4 * kq);
      acc[t] = mfmaBf16(a, b, acc[t]);
    }
  }
#pragma unroll
  for(int t = 0; t < KIN / 16; ++t)
#pragma unroll
    for(int v = 0; v < 4; ++v) {
      const int r = row0 + 4 * kq + v;
      if(r < n) gX[(size_t)r * KIN + 16 * t + i] = acc[t][v];
    }
}

// gz_prev = bf16(mask(h_prev > 0) (gX[:, 0:32] + A^_off gX[:, 32:64])).
__global__ __launch_bounds__(256) void mpAggBwd(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                const float *__restrict__ ahat, const float *__restrict__ gX,
                                                const bf16_t *__restrict__ hprev, bf16_t *__restrict__ gz)
{
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int row = t >> 5, f = t & 31;
  if(row >= n) return;
  float s = 0.f;
  for(int p = ia[row]; p < ia[row + 1]; ++p) {
    const int j = ja[p];
    if(j == row) continue;
    s += ahat[p] * gX[(size_t)j * 64 + 32 + f];
  }
  s = gX[(size_t)row * 64 + f] + s;
  gz[(size_t)row * 32 + f] = fromBf16(hprev[(size_t)row * 32 + f]) > 0.f ? toBf16(s) : (bf16_t)0;
}

// Per-block partials of X^T Y (F1 x 32) and of the column sums of Y (32):
// X = [X1 | X2] rows (bf16, F1 = 16, 32 or 64 wide), Y (n x 32).  With bf16 Y
// the product runs on v_mfma_f32_16x16x16bf16_1k (16 rows per step, A[i][k] =
// X[row0 + k][f0 + i]); with FP32 Y on v_mfma_f32_16x16x4f32 (4 rows per
// step).  Block b owns rows [b R, (b+1) R).
template <int F1, bool YBF16>
__global__ __launch_bounds__(256) void mpGemmTN(int n, int R, const bf16_t *__restrict__ X1, int ld1,
                                                const bf16_t *__restrict__ X2, const void *__restrict__ Yv,
                                                float *__restrict__ part)
{
  constexpr int NT1 = F1 / 16, NT2 = 2, OUT = F1 * 32 + 32;
  __shared__ float lds[4][F1 * 32];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = lane & 15, kq = lane >> 4;
  const int rb = blockIdx.x * R, re = min(n, rb + R);
  feF4 acc[NT1][NT2];
#pragma unroll
  for(int a = 0; a < NT1; ++a)
#pragma unroll
    for(int b = 0; b < NT2; ++b) acc[a][b] = (feF4){0.f, 0.f, 0.f, 0.f};
  auto xval = [&](int r, int f) -> bf16_t {
    if(F1 == 64 && f >= 32) return X2[(size_t)r * 32 + f - 32];
    return X1[(size_t)r * ld1 + f];
  };
  if(YBF16) {
    const bf16_t *Y = (const bf16_t *)Yv;
    for(int r0 = rb + wave * 16; r0 < re; r0 += 64) {
#pragma unroll
      for(int ta = 0; ta < NT1; ++ta) {
        feShort4 av;
#pragma unroll
        for(int e = 0; e < 4; ++e) {
          const int r = r0 + 4 * kq + e;
          av[e] = r < re ? (short)xval(r, 16 * ta + i) : (short)0;
        }
#pragma unroll
        for(int tb = 0; tb < NT2; ++tb) {
          feShort4 bv;
#pragma unroll
          for(int e = 0; e < 4; ++e) {
            const int r = r0 + 4 * kq + e;
            bv[e] = r < re ? (short)Y[(size_t)r * 32 + 16 * tb + i] : (short)0;
          }
          acc[ta][tb] = mfmaBf16(av, bv, acc[ta][tb]);
        }
      }
    }
  } else {
    const float *Y = (const float *)Yv;
    for(int r0 = rb + wave * 4; r0 < re; r0 += 16) {
      const int r = r0 + kq;
#pragma unroll
      for(int ta = 0; ta < NT1; ++ta) {
        const float av = r < re ? fromBf16(xval(r, 16 * ta + i)) : 0.f;
#pragma unroll
        for(int tb = 0; tb < NT2; ++tb) {
          const float bv = r < re ? Y[(size_t)r * 32 + 16 * tb + i] : 0.f;
          acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[ta][tb], 0, 0, 0);
        }
      }
    }
  }
  // D[4 kq + v][i] of tile (ta, tb) = sum over rows of X[., 16 ta + 4 kq + v] Y[., 16 tb + i]
#pragma unroll
  for(int ta = 0; ta < NT1; ++ta)
#pragma unroll
    for(int tb = 0; tb < NT2; ++tb)
#pragma unroll
      for(int v = 0; v < 4; ++v) lds[wave][(16 * ta + 4 * kq + v) * 32 + 16 * tb + i] = acc[ta][tb][v];
  __syncthreads();
  float *out = part + (size_t)blockIdx.x * OUT;
  for(int e = threadIdx.x; e < F1 * 32; e += 256) out[e] = lds[0][e] + lds[1][e] + lds[2][e] + lds[3][e];
  if(threadIdx.x < 32) {
    float s = 0.f;
    for(int r = rb; r < re; ++r)
      s += YBF16 ? fromBf16(((const bf16_t *)Yv)[(size_t)r * 32 + threadIdx.x]) : ((const float *)Yv)[(size_t)r * 32 + threadIdx.x];
    out[F1 * 32 + threadIdx.x] = s;
  }
}

// Fold the partials of mpGemmTN into the gradient vector.  mode 0: a dense
// layer (weights at wOff, F1 x 32 row-major; biases at bOff); mode 1: the
// heads (columns 0-7 -> P, 8-15 -> Q, 16-20 -> Hs; column sums 16-20 -> hb).
__global__ void mpFoldGrad(int G, int F1, int mode, int wOff, int bOff, const float *__restrict__ part,
                           float *__restrict__ grad)
{
  const int OUT = F1 * 32 + 32;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= OUT) return;
  double s = 0.;
  for(int b = 0; b < G; ++b) s += part[(size_t)b * OUT + e];
  if(mode == 0) {
    if(e < F1 * 32) grad[wOff + e] = (float)s;
    else grad[bOff + (e - F1 * 32)] = (float)s;
  } else if(e < F1 * 32) {
    const int f = e / 32, c = e % 32;
    if(c < 8) grad[oP + f * kNpq + c] = (float)s;
    else if(c < 16) grad[oQ + f * kNpq + c - 8] = (float)s;
    else if(c < 16 + kNs) grad[oHs + f * kNs + c - 16] = (float)s;
  } else {
    const int c = e - F1 * 32;
    if(c >= 16 && c < 16 + kNs) grad[ohb + c - 16] = (float)s;
  }
}

__global__ void mpAdam(int count, float *__restrict__ theta, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= count) return;
  const float g = grad[e];
  const float mm = b1 * m[e] + (1.f - b1) * g;
  const float vv = b2 * v[e] + (1.f - b2) * g * g;
  m[e] = mm;
  v[e] = vv;
  theta[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// bf16 copies of W0..W3.
__global__ void mpRefreshW16(const float *__restrict__ theta, bf16_t *__restrict__ w16)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= kW16) return;
  int l = 3;
  while(l > 0 && e < kW16Off[l]) --l;
  w16[e] = toBf16(theta[kWOff[l] + (e - kW16Off[l])]);
}

inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------
// State

struct FeMpnn {
  bool hasParams = false;
  int step = 0;
  int n = 0;
  int64_t nnz = 0;
  float *theta = nullptr, *grad = nullptr, *adm = nullptr, *adv = nullptr;
  double *s = nullptr, *d0 = nullptr, *lossPart = nullptr, *loss = nullptr;
  float *ahat = nullptr, *g0 = nullptr, *w0 = nullptr, *g = nullptr, *w = nullptr, *B = nullptr, *B0 = nullptr,
        *d = nullptr;
  const int *tpos = nullptr;  // transpose map of the pattern (owned by the preconditioner)
  unsigned short *X = nullptr, *h[4] = {}, *m[4] = {}, *gz = nullptr, *w16 = nullptr;
  float *p = nullptr, *q = nullptr, *hs = nullptr, *gHead = nullptr, *gX = nullptr, *gpart = nullptr;
};

void feMpnnRelease(fe_hip_ctx *ctx)
{
  delete ctx->mpnn;
  ctx->mpnn = nullptr;
}

namespace {

int mpState(fe_hip_ctx *ctx, FeMpnn *&S)
{
  if(!ctx->mpnn) {
    ctx->mpnn = new FeMpnn();
    FeMpnn &M = *ctx->mpnn;
    if(feAlloc(ctx, &M.theta, kParams) || feAlloc(ctx, &M.grad, kParams) || feAlloc(ctx, &M.adm, kParams) ||
       feAlloc(ctx, &M.adv, kParams) || feAlloc(ctx, &M.w16, kW16) || feAlloc(ctx, &M.loss, 8) ||
       feAlloc(ctx, &M.lossPart, 3 * kLossGrid))
      return FE_STATUS_OUT_OF_MEMORY;
  }
  S = ctx->mpnn;
  return FE_STATUS_OK;
}

// Buffers sized by the current system.
int mpBuffers(fe_hip_ctx *ctx, FeMpnn &M)
{
  FE_REQUIRE(ctx, ctx->csrBuilt, "build the CSR pattern first");
  const int n = ctx->n;
  const int64_t nnz = ctx->nnz;
  if(M.n == n && M.nnz == nnz && M.X) return FE_STATUS_OK;
  FE_REQUIRE(ctx, M.n == 0, "the learned preconditioner is bound to the first system it saw");
  M.n = n;
  M.nnz = nnz;
  const int gemmOut = 64 * 32 + 32;
  if(feAlloc(ctx, &M.s, n) || feAlloc(ctx, &M.d0, n) || feAlloc(ctx, &M.ahat, nnz) || feAlloc(ctx, &M.g0, nnz) ||
     feAlloc(ctx, &M.w0, nnz) || feAlloc(ctx, &M.g, nnz) || feAlloc(ctx, &M.w, nnz) || feAlloc(ctx, &M.B, n) ||
     feAlloc(ctx, &M.B0, n) || feAlloc(ctx, &M.d, n) ||
     feAlloc(ctx, &M.X, (size_t)n * kNin) || feAlloc(ctx, &M.gz, (size_t)n * 32) ||
     feAlloc(ctx, &M.p, (size_t)n * kNpq) || feAlloc(ctx, &M.q, (size_t)n * kNpq) ||
     feAlloc(ctx, &M.hs, (size_t)n * kNs) || feAlloc(ctx, &M.gHead, (size_t)n * 32) ||
     feAlloc(ctx, &M.gX, (size_t)n * 64) || feAlloc(ctx, &M.gpart, (size_t)kGemmGrid * gemmOut))
    return FE_STATUS_OUT_OF_MEMORY;
  for(int l = 0; l < 4; ++l)
    if(feAlloc(ctx, &M.h[l], (size_t)n * 32)) return FE_STATUS_OUT_OF_MEMORY;
  for(int l = 1; l < 4; ++l)
    if(feAlloc(ctx, &M.m[l], (size_t)n * 32)) return FE_STATUS_OUT_OF_MEMORY;
  // entries never written by the kernels stay zero
  for(float *a : {M.g0, M.w0, M.g, M.w}) FE_HIP_TRY(ctx, hipMemsetAsync(a, 0, sizeof(float) * nnz, ctx->stream));
  FE_HIP_TRY(ctx, hipMemsetAsync(M.gHead, 0, sizeof(float) * (size_t)n * 32, ctx->stream));
  return FE_STATUS_OK;
}

// Features, forward, corrected factor (g, B, d) of the current matrix and its
// IC(0) factor (ctx->pc.d_lu must hold it).
int mpForward(fe_hip_ctx *ctx, FeMpnn &M)
{
  FE_REQUIRE(ctx, M.hasParams, "no learned preconditioner weights are loaded");
  if(int rc = mpBuffers(ctx, M)) return rc;
  FE_REQUIRE(ctx, ctx->pc.patternReady && ctx->pc.symmetricPattern,
             "the learned preconditioner needs a factored, structurally symmetric matrix");
  M.tpos = ctx->pc.d_tposA;
  const int n = ctx->n;
  hipStream_t st = ctx->stream;
  const int gN = cdiv(n, 256);
  const int *ia = ctx->d_ia, *ja = ctx->d_ja, *dg = ctx->d_diag;
  hipLaunchKernelGGL(mpScale, dim3(gN), dim3(256), 0, st, n, dg, ctx->d_val, ctx->pc.d_lu, M.s, M.d0);
  const int64_t nnz = ctx->nnz;
  const int gE = cdiv(nnz, 256);
  const int *rowOf = ctx->pc.d_rowOf;
  hipLaunchKernelGGL(mpNormalise, dim3(gE), dim3(256), 0, st, nnz, rowOf, ja, ctx->d_val, ctx->pc.d_lu, M.s, M.ahat,
                     M.g0);
  const int g8 = cdiv((int64_t)n * 8, 256);
  hipLaunchKernelGGL(mpRows, dim3(g8), dim3(256), 0, st, n, ia, ja, dg, M.ahat, M.g0, M.B0, M.w0);
  hipLaunchKernelGGL(mpFeatures, dim3(gN), dim3(256), 0, st, n, ia, ja, dg, M.ahat, M.g0, M.w0, M.d0, M.X);
  const int gD = cdiv(n, 64);
  hipLaunchKernelGGL(mpDense<16>, dim3(gD), dim3(256), 0, st, n, M.X, (const bf16_t *)nullptr, M.w16 + kW16Off[0],
                     M.theta + ob0, M.h[0]);
  for(int l = 1; l <= 3; ++l) {
    hipLaunchKernelGGL(mpAggregate, dim3(cdiv((int64_t)n * 16, 256)), dim3(256), 0, st, n, ia, ja, M.ahat,
                       M.h[l - 1], M.m[l]);
    hipLaunchKernelGGL(mpDense<64>, dim3(gD), dim3(256), 0, st, n, M.h[l - 1], M.m[l], M.w16 + kW16Off[l],
                       M.theta + kBOff[l], M.h[l]);
  }
  hipLaunchKernelGGL(mpHeads, dim3(gN), dim3(256), 0, st, n, M.h[3], M.theta, M.p, M.q, M.hs);
  hipLaunchKernelGGL(mpCorrect, dim3(gE), dim3(256), 0, st, nnz, rowOf, ja, M.ahat, M.g0, M.w0, M.p, M.q, M.hs, M.g);
  hipLaunchKernelGGL(mpRows, dim3(g8), dim3(256), 0, st, n, ia, ja, dg, M.ahat, M.g, M.B, M.w);
  hipLaunchKernelGGL(mpPivots, dim3(gN), dim3(256), 0, st, n, M.B, M.hs, M.d);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

int mpLossGrad(fe_hip_ctx *ctx, FeMpnn &M, double *lossOut)
{
  const int n = ctx->n;
  hipStream_t st = ctx->stream;
  const int *ia = ctx->d_ia, *ja = ctx->d_ja, *dg = ctx->d_diag;
  const int gL = std::min(kLossGrid, cdiv(n, 256));
  hipLaunchKernelGGL(mpLossParts, dim3(gL), dim3(256), 0, st, n, M.B, M.hs, M.lossPart);
  hipLaunchKernelGGL(mpLossFold, dim3(1), dim3(64), 0, st, n, gL, M.lossPart, M.loss);
  hipLaunchKernelGGL(mpEdgeGrad, dim3(cdiv(n, 256)), dim3(256), 0, st, n, ia, ja, dg, M.tpos, M.ahat, M.w, M.w0, M.B,
                     M.p, M.q, M.hs, M.loss, M.gHead);
  // heads
  const int R = ((cdiv(n, kGemmGrid) + 63) / 64) * 64;
  const int G = cdiv(n, R);
  hipLaunchKernelGGL((mpGemmTN<32, false>), dim3(G), dim3(256), 0, st, n, R, M.h[3], 32, (const bf16_t *)nullptr,
                     (const void *)M.gHead, M.gpart);
  hipLaunchKernelGGL(mpFoldGrad, dim3(cdiv(32 * 32 + 32, 256)), dim3(256), 0, st, G, 32, 1, 0, 0, M.gpart, M.grad);
  hipLaunchKernelGGL(mpHeadsBwd, dim3(cdiv((int64_t)n * 32, 256)), dim3(256), 0, st, n, M.h[3], M.gHead, M.theta, M.gz);
  // layers 3..1
  for(int l = 3; l >= 1; --l) {
    hipLaunchKernelGGL((mpGemmTN<64, true>), dim3(G), dim3(256), 0, st, n, R, M.h[l - 1], 32, M.m[l],
                       (const void *)M.gz, M.gpart);
    hipLaunchKernelGGL(mpFoldGrad, dim3(cdiv(64 * 32 + 32, 256)), dim3(256), 0, st, G, 64, 0, kWOff[l], kBOff[l],
                       M.gpart, M.grad);
    hipLaunchKernelGGL(mpDenseBwdX<64>, dim3(cdiv(n, 64)), dim3(256), 0, st, n, M.gz, M.w16 + kW16Off[l], M.gX);
    hipLaunchKernelGGL(mpAggBwd, dim3(cdiv((int64_t)n * 32, 256)), dim3(256), 0, st, n, ia, ja, M.ahat, M.gX,
                       M.h[l - 1], M.gz);
  }
  hipLaunchKernelGGL((mpGemmTN<16, true>), dim3(G), dim3(256), 0, st, n, R, M.X, 16, (const bf16_t *)nullptr,
                     (const void *)M.gz, M.gpart);
  hipLaunchKernelGGL(mpFoldGrad, dim3(cdiv(16 * 32 + 32, 256)), dim3(256), 0, st, G, 16, 0, kWOff[0], kBOff[0],
                     M.gpart, M.grad);
  FE_HIP_TRY(ctx, hipGetLastError());
  if(lossOut) {
    FE_HIP_TRY(ctx, hipMemcpyAsync(lossOut, M.loss, sizeof(double), hipMemcpyDeviceToHost, st));
    FE_HIP_TRY(ctx, hipStreamSynchronize(st));
  }
  return FE_STATUS_OK;
}

void hostInit(uint64_t seed, std::vector<float> &th)
{
  th.assign(kParams, 0.f);
  struct Blk {
    int off, rows, cols;
    double scale;
  } blks[] = {{oW0, kNin, kH, 1.}, {oW1, 2 * kH, kH, 1.}, {oW2, 2 * kH, kH, 1.},
              {oW3, 2 * kH, kH, 1.}, {oP, kH, kNpq, 0.05}, {oQ, kH, kNpq, 0.05}};
  for(const Blk &b : blks) {
    const double r = std::sqrt(6.0 / b.rows) * b.scale;
    for(int e = 0; e < b.rows * b.cols; ++e) {
      const uint64_t q = (uint64_t)(b.off + e);
      const double u = (double)(splitmix64(((seed & 0xFFFFFFFFull) << 32) + q) >> 11) * 0x1.0p-53;
      th[b.off + e] = (float)(r * (2.0 * u - 1.0));
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Solver hook and C ABI

int feLearnedCorrect(fe_hip_ctx *ctx)
{
  FE_REQUIRE(ctx, ctx->mpnn && ctx->mpnn->hasParams, "no learned preconditioner weights are loaded");
  FeMpnn &M = *ctx->mpnn;
  if(int rc = mpForward(ctx, M)) return rc;
  FePrecond &P = ctx->pc;
  hipLaunchKernelGGL(mpWriteFactor, dim3(cdiv(ctx->nnz, 256)), dim3(256), 0, ctx->stream, ctx->nnz, P.d_rowOf,
                     ctx->d_ia, ctx->d_ja, ctx->d_diag, M.tpos, ctx->d_val, M.s, M.g, M.d, P.d_iaL, P.d_iaU, P.d_gL64,
                     P.d_gL32, P.d_gU64, P.d_gU32, P.d_dinv64, P.d_dinv32);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

extern "C" {

int fe_hip_mpnn_param_count(void) { return kParams; }

void fe_hip_mpnn_default_opts(fe_hip_mpnn_opts *o)
{
  if(!o) return;
  o->lr = 3e-3;
  o->beta1 = 0.9;
  o->beta2 = 0.999;
  o->eps = 1e-8;
}

int fe_hip_mpnn_set_params(fe_hip_ctx *ctx, const float *theta, int count)
{
  if(!ctx || !theta) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, count == kParams, "parameter vector has the wrong length");
  for(int e = 0; e < count; ++e) FE_REQUIRE(ctx, std::isfinite(theta[e]), "parameters must be finite");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  FeMpnn *M = nullptr;
  if(int rc = mpState(ctx, M)) return rc;
  FE_HIP_TRY(ctx, hipMemcpyAsync(M->theta, theta, sizeof(float) * kParams, hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(ctx, hipMemsetAsync(M->adm, 0, sizeof(float) * kParams, ctx->stream));
  FE_HIP_TRY(ctx, hipMemsetAsync(M->adv, 0, sizeof(float) * kParams, ctx->stream));
  hipLaunchKernelGGL(mpRefreshW16, dim3(cdiv(kW16, 256)), dim3(256), 0, ctx->stream, M->theta, M->w16);
  FE_HIP_TRY(ctx, hipGetLastError());
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  M->hasParams = true;
  M->step = 0;
  return FE_STATUS_OK;
}

int fe_hip_mpnn_init(fe_hip_ctx *ctx, uint64_t seed)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  std::vector<float> th;
  hostInit(seed, th);
  return fe_hip_mpnn_set_params(ctx, th.data(), kParams);
}

int fe_hip_mpnn_get_params(fe_hip_ctx *ctx, float *theta, int count)
{
  if(!ctx || !theta) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, count == kParams, "parameter vector has the wrong length");
  FE_REQUIRE(ctx, ctx->mpnn && ctx->mpnn->hasParams, "no learned preconditioner weights are loaded");
  FE_HIP_TRY(ctx, hipMemcpyAsync(theta, ctx->mpnn->theta, sizeof(float) * kParams, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return FE_STATUS_OK;
}

int fe_hip_mpnn_loss_grad(fe_hip_ctx *ctx, double *loss, float *grad)
{
  if(!ctx) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->mpnn && ctx->mpnn->hasParams, "no learned preconditioner weights are loaded");
  FE_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if(int rc = fePrecondFactor(ctx, kLearnedIcSweeps)) return rc;
  FeMpnn &M = *ctx->mpnn;
  if(int rc = mpForward(ctx, M)) return rc;
  if(int rc = mpLossGrad(ctx, M, loss)) return rc;
  if(grad) {
    FE_HIP_TRY(ctx, hipMemcpyAsync(grad, M.grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  return FE_STATUS_OK;
}

int fe_hip_mpnn_train_step(fe_hip_ctx *ctx, const fe_hip_mpnn_opts *o, double *loss)
{
  if(!ctx || !o) return FE_STATUS_POINTER_ERROR;
  if(int rc = fe_hip_mpnn_loss_grad(ctx, loss, nullptr)) return rc;
  FeMpnn &M = *ctx->mpnn;
  M.step++;
  const double c1 = 1. - std::pow(o->beta1, M.step), c2 = 1. - std::pow(o->beta2, M.step);
  hipLaunchKernelGGL(mpAdam, dim3(cdiv(kParams, 256)), dim3(256), 0, ctx->stream, kParams, M.theta, M.grad, M.adm,
                     M.adv, (float)o->lr, (float)o->beta1, (float)o->beta2, (float)o->eps, (float)c1, (float)c2);
  hipLaunchKernelGGL(mpRefreshW16, dim3(cdiv(kW16, 256)), dim3(256), 0, ctx->stream, M.theta, M.w16);
  FE_HIP_TRY(ctx, hipGetLastError());
  return FE_STATUS_OK;
}

int fe_hip_mpnn_get_tensor(fe_hip_ctx *ctx, const char *name, float *out, int64_t count)
{
  if(!ctx || !name || !out) return FE_STATUS_POINTER_ERROR;
  FE_REQUIRE(ctx, ctx->mpnn && ctx->mpnn->X, "run the network first");
  FeMpnn &M = *ctx->mpnn;
  const std::string k(name);
  const int64_t n = M.n;
  const bf16_t *b16 = nullptr;
  const float *f32 = nullptr;
  int64_t need = 0;
  if(k == "X") { b16 = M.X; need = n * kNin; }
  else if(k.size() == 2 && (k[0] == 'h' || k[0] == 'm') && k[1] >= '0' && k[1] <= '3') {
    const int l = k[1] - '0';
    FE_REQUIRE(ctx, k[0] == 'h' || l >= 1, "no tensor m0");
    b16 = k[0] == 'h' ? M.h[l] : M.m[l];
    need = n * 32;
  }
  else if(k == "p") { f32 = M.p; need = n * kNpq; }
  else if(k == "q") { f32 = M.q; need = n * kNpq; }
  else if(k == "heads") { f32 = M.hs; need = n * kNs; }
  else if(k == "g") { f32 = M.g; need = M.nnz; }
  else if(k == "w0") { f32 = M.w0; need = M.nnz; }
  else if(k == "w") { f32 = M.w; need = M.nnz; }
  else if(k == "B") { f32 = M.B; need = n; }
  else if(k == "d") { f32 = M.d; need = n; }
  else FE_REQUIRE(ctx, false, "unknown tensor " + k);
  FE_REQUIRE(ctx, count == need, "tensor " + k + " has " + std::to_string(need) + " entries");
  FE_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if(f32) {
    FE_HIP_TRY(ctx, hipMemcpy(out, f32, sizeof(float) * need, hipMemcpyDeviceToHost));
  } else {
    std::vector<bf16_t> tmp(need);
    FE_HIP_TRY(ctx, hipMemcpy(tmp.data(), b16, sizeof(bf16_t) * need, hipMemcpyDeviceToHost));
    for(int64_t e = 0; e < need; ++e) {
      const uint32_t u = (uint32_t)tmp[e] << 16;
      std::memcpy(out + e, &u, 4);
    }
  }
  return FE_STATUS_OK;
}

}  // extern "C"
