// Rate surrogate network: 3 -> 64 -> 64 -> 6, SiLU, on gfx950 matrix cores.
//
// Inputs are log10 of the three threshold optical depths (HI, HeI, HeII),
// outputs log10 of the unit-spectrum kernels
//   F_s(tau) = sum_k S_k sigma_s(k) exp(-tau_k),
//   G_s(tau) = sum_k S_k sigma_s(k) (nu_k - nu_s) exp(-tau_k)   (nu_k > nu_s),
// s = HI, HeI, HeII, from which the photon-conserving rates follow.
//
// Weights are bf16 values; the two 64-wide layers run on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation, activations split into a
// bf16 high and low part (two MFMAs) so that each keeps ~16 bits.  The
// 3-wide input layer runs on the VALU in FP32 with bf16-valued weights.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A (16x32, MxK): lane l holds A[l%16][8*(l/16) + i], i = 0..7
//   B (32x16, KxN): lane l holds B[8*(l/16) + i][l%16]
//   D (16x16, MxN): lane l holds D[4*(l/16) + i][l%16], i = 0..3
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace argot {

constexpr int kHidden = 64;
constexpr int kNin = 3;
constexpr int kNout = 6;    // F_HI, F_HeI, F_HeII, G_HI, G_HeI, G_HeII

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 parameter vector (the training master copy; the quantized copy Pq
// holds the same layout with every weight rounded to bf16):
//   W1[3][64], b1[64], W2[64 in][64 out], b2[64], W3[64 in][6 out], b3[6].
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kNin * kHidden;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden * kNout;
constexpr int kNparam = kOffB3 + kNout;
constexpr int kNparamPad = (kNparam + 63) / 64 * 64;

// bf16 weight images in the layouts the MFMA fragments read contiguously.
constexpr int kW2T = 0;                         // [64 out][64 in]   forward
constexpr int kW2N = kW2T + kHidden * kHidden;  // [64 in][64 out]   backward (dH1)
constexpr int kW3T = kW2N + kHidden * kHidden;  // [16 out][64 in]   forward, rows >= 6 zero
constexpr int kW3N = kW3T + 16 * kHidden;       // [64 in][32 out]   backward (dH2), cols >= 6 zero
constexpr int kNbf16 = kW3N + kHidden * 32;

// Polynomial baseline under the network: the monomials of degree <= 3 of
// the normalised input, in the order 1, z0, z1, z2, z0z0, z0z1, z0z2, z1z1,
// z1z2, z2z2, z0z0z0, z0z0z1, z0z0z2, z0z1z1, z0z1z2, z0z2z2, z1z1z1,
// z1z1z2, z1z2z2, z2z2z2; coefficient m of output o at poly[m * kNout + o].
constexpr int kNpoly = 20;

// Device view of a trained network.
struct MlpDev {
  const float* Pq;       // [kNparam] quantized parameters
  const __bf16* Wb;      // [kNbf16] bf16 images
  float xc, xs;          // input normalisation: (log10 tau - xc) * xs
  float ymean[kNout], ystd[kNout];   // output: log10 F = ymean + ystd * y (+ baseline)
  const float* poly;     // [kNpoly * kNout] baseline coefficients, or unused
  int npoly;             // 0 (no baseline) or kNpoly
};

// The baseline at normalised input z: out[o] = sum_m z^m poly[m][o] (FP32,
// the same operations wherever it is evaluated, so that training targets and
// inference agree bit for bit).
__device__ __forceinline__ void poly_eval(const float* __restrict__ c, float z0, float z1,
                                          float z2, float out[kNout]) {
  float m[kNpoly];
  m[0] = 1.0f;
  m[1] = z0;
  m[2] = z1;
  m[3] = z2;
  m[4] = z0 * z0;
  m[5] = z0 * z1;
  m[6] = z0 * z2;
  m[7] = z1 * z1;
  m[8] = z1 * z2;
  m[9] = z2 * z2;
  m[10] = m[4] * z0;
  m[11] = m[4] * z1;
  m[12] = m[4] * z2;
  m[13] = m[7] * z0;
  m[14] = m[5] * z2;
  m[15] = m[9] * z0;
  m[16] = m[7] * z1;
  m[17] = m[7] * z2;
  m[18] = m[9] * z1;
  m[19] = m[9] * z2;
#pragma unroll
  for (int o = 0; o < kNout; o++) {
    float a = 0.0f;
#pragma unroll
    for (int k = 0; k < kNpoly; k++) a = __builtin_fmaf(m[k], c[k * kNout + o], a);
    out[o] = a;
  }
}

__device__ __forceinline__ float sigmoid_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

__device__ __forceinline__ float silu(float z) { return z * sigmoid_f(z); }

// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z)
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid_f(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ void split_bf16(const float v[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    hi[i] = (__bf16)v[i];
    lo[i] = (__bf16)(v[i] - (float)hi[i]);
  }
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// A fragment (rows r0..r0+15, columns c0..c0+31) of a row-major FP32
// matrix in LDS with leading dimension ld.
__device__ __forceinline__ void lds_frag_a(const float* m, int ld, int r0, int c0, float v[8]) {
  const int l = threadIdx.x & 63;
  const float* p = m + (r0 + (l & 15)) * ld + c0 + 8 * (l >> 4);
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = p[i];
}

// A fragment of the transpose: A[r][c] = M[c][r].
__device__ __forceinline__ void lds_frag_at(const float* m, int ld, int r0, int c0, float v[8]) {
  const int l = threadIdx.x & 63;
  const float* p = m + (c0 + 8 * (l >> 4)) * ld + r0 + (l & 15);
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = p[i * ld];
}

// B fragment (rows k0..k0+31, columns n0..n0+15) of a row-major matrix.
__device__ __forceinline__ void lds_frag_b(const float* m, int ld, int k0, int n0, float v[8]) {
  const int l = threadIdx.x & 63;
  const float* p = m + (k0 + 8 * (l >> 4)) * ld + n0 + (l & 15);
#pragma unroll
  for (int i = 0; i < 8; i++) v[i] = p[i * ld];
}

// B fragment of a bf16 matrix stored transposed (Wt[n][k], row length
// ldk): lane l reads 8 consecutive k of column n.
__device__ __forceinline__ bf16x8 w_frag(const __bf16* wt, int ldk, int k0, int n0) {
  const int l = threadIdx.x & 63;
  return *reinterpret_cast<const bf16x8*>(wt + (n0 + (l & 15)) * ldk + k0 + 8 * (l >> 4));
}

// Store a D fragment at (r0, c0) of a row-major matrix.
__device__ __forceinline__ void lds_store_d(float* m, int ld, int r0, int c0, f32x4 d) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 4; i++) m[(r0 + 4 * (l >> 4) + i) * ld + c0 + (l & 15)] = d[i];
}

constexpr int kLdH = 68;                        // padded LDS row (floats)
constexpr int kMlpLds = 16 * kLdH + 64 * 8;     // floats of LDS per wavefront

// Forward pass for the 64 samples of a wavefront, one per lane; every lane
// of the wavefront must be active (the MFMAs span the wavefront).  x: this
// lane's log10 optical depths (already clamped to the input hull); y: this
// lane's log10 kernels.  lds: kMlpLds floats private to the wavefront.
// The 64 samples go through in four blocks of 16 rows, so that only one
// 16x64 activation block is held at a time.
__device__ inline void mlp_forward(const MlpDev& net, float x0, float x1, float x2,
                                   float y[kNout], float* lds) {
  const int l = threadIdx.x & 63;
  float* H = lds;
  // one 64 x 8 table holds each sample's normalised input (columns 0-2)
  // and, once layer 1 of its 16-row block has read them, its outputs
  // (columns 0-5): layer 3 of a block writes only that block's rows
  float* XY = lds + 16 * kLdH;
  const float* P = net.Pq;
  const float z0 = (x0 - net.xc) * net.xs, z1 = (x1 - net.xc) * net.xs,
              z2 = (x2 - net.xc) * net.xs;
  XY[8 * l + 0] = z0;
  XY[8 * l + 1] = z1;
  XY[8 * l + 2] = z2;
  __syncthreads();
  const int r = l & 15, c0 = 16 * (l >> 4);
#pragma unroll 1
  for (int mb = 0; mb < 4; mb++) {
    // layer 1: lane computes row r of the block, hidden units c0..c0+15
    const float a0 = XY[8 * (16 * mb + r) + 0], a1 = XY[8 * (16 * mb + r) + 1],
                a2 = XY[8 * (16 * mb + r) + 2];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int u = c0 + j;
      float z = P[kOffB1 + u];
      z = __builtin_fmaf(a0, P[kOffW1 + u], z);
      z = __builtin_fmaf(a1, P[kOffW1 + kHidden + u], z);
      z = __builtin_fmaf(a2, P[kOffW1 + 2 * kHidden + u], z);
      H[r * kLdH + u] = silu(z);
    }
    __syncthreads();
    // layer 2 on MFMA
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float v[8];
      lds_frag_a(H, kLdH, 0, 32 * kb, v);
      bf16x8 hi, lo;
      split_bf16(v, hi, lo);
#pragma unroll
      for (int nb = 0; nb < 4; nb++) {
        const bf16x8 w = w_frag(net.Wb + kW2T, kHidden, 32 * kb, 16 * nb);
        acc[nb] = mfma(hi, w, acc[nb]);
        acc[nb] = mfma(lo, w, acc[nb]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const float b = P[kOffB2 + 16 * nb + (l & 15)];
#pragma unroll
      for (int i = 0; i < 4; i++) acc[nb][i] = silu(acc[nb][i] + b);
      lds_store_d(H, kLdH, 0, 16 * nb, acc[nb]);
    }
    __syncthreads();
    // layer 3 on MFMA
    f32x4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float v[8];
      lds_frag_a(H, kLdH, 0, 32 * kb, v);
      bf16x8 hi, lo;
      split_bf16(v, hi, lo);
      const bf16x8 w = w_frag(net.Wb + kW3T, kHidden, 32 * kb, 0);
      o = mfma(hi, w, o);
      o = mfma(lo, w, o);
    }
    if ((l & 15) < kNout) {
#pragma unroll
      for (int i = 0; i < 4; i++) XY[(16 * mb + 4 * (l >> 4) + i) * 8 + (l & 15)] = o[i];
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = 0; o < kNout; o++)
    y[o] = net.ymean[o] + net.ystd[o] * (XY[8 * l + o] + P[kOffB3 + o]);
  if (net.npoly) {
    float b[kNout];
    poly_eval(net.poly, z0, z1, z2, b);
#pragma unroll
    for (int o = 0; o < kNout; o++) y[o] += b[o];
  }
  __syncthreads();
}

}  // namespace argot
