// Rate surrogate: sample generator (exact integral in FP64), training
// kernels (bf16 MFMA forward and backward with FP32 accumulation, Adam),
// validation on a trust grid, batch inference, and the host entry points
// declared in include/argot_hip.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <vector>

#include "argot_hip.h"
#include "argot_internal.h"
#include "surrogate.h"

namespace argot {

static_assert(kNparam == ARGOT_SURROGATE_NPARAM, "parameter count of the public header");

// splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014): counter-based
// uniform doubles in [0, 1), reproducible on the host.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__host__ __device__ inline double uniform01(uint64_t seed, uint64_t i) {
  return (double)(mix64(seed ^ mix64(i)) >> 11) * 0x1.0p-53;
}

// Tables of the exact integral (device pointers, FP64).
struct ThinTables {
  int ngrid;
  const double* nu;    // [ngrid] nu / nu_L
  const double* cs;    // [3*ngrid] bin-major sigma (FP32 table values)
  const double* S;     // [ngrid] spectrum shape, sum 1
  double sig_th[3];    // first nonzero sigma of each species
};

// Exact unit-spectrum kernels at columns N[3] (cm^-2): F_s and G_s.
__device__ inline void thin_exact(const ThinTables& T, const double N[3], double f[kNout]) {
#pragma unroll
  for (int o = 0; o < kNout; o++) f[o] = 0.0;
  for (int k = 0; k < T.ngrid; k++) {
    const double c0 = T.cs[3 * k], c1 = T.cs[3 * k + 1], c2 = T.cs[3 * k + 2];
    const double w = T.S[k] * exp(-(N[0] * c0 + N[1] * c1 + N[2] * c2));
    const double nu = T.nu[k];
    f[0] += w * c0;
    f[1] += w * c1;
    f[2] += w * c2;
    if (nu > 1.0) f[3] += w * c0 * (nu - 1.0);
    if (nu > 1.8125) f[4] += w * c1 * (nu - 1.8125);
    if (nu > 4.0) f[5] += w * c2 * (nu - 4.0);
  }
}

// log10 tau of sample i, axis s, drawn uniformly over [lo, hi) and rounded
// to FP32 (the network sees the FP32 value; the target is computed from it).
__device__ inline float sample_log_tau(uint64_t seed, uint64_t i, int s, double lo, double hi) {
  return (float)(lo + (hi - lo) * uniform01(seed, 3 * i + (uint64_t)s));
}

// Samples around seed points (the optical depths a caller's rays occupy):
// seed j = floor(u * nseed), u uniform under seed ^ 0x5eed5eed, plus jitter
// times a Box-Muller normal per axis (Box & Muller, Ann. Math. Stat. 29,
// 1958), clamped to [lo, hi] and rounded to FP32.
struct SeedDist {
  const float* seeds;   // [nseed * 3] log10 tau, or null (uniform samples)
  int64_t nseed;
  float jitter;
};

__device__ inline void sample_point(const SeedDist& D, uint64_t seed, uint64_t i, double lo,
                                    double hi, float x[3]) {
  if (!D.seeds) {
    for (int s = 0; s < 3; s++) x[s] = sample_log_tau(seed, i, s, lo, hi);
    return;
  }
  const int64_t j = min((int64_t)(uniform01(seed ^ 0x5eed5eedull, i) * (double)D.nseed),
                        D.nseed - 1);
  for (int s = 0; s < 3; s++) {
    const double u1 = uniform01(seed, 6 * i + 2 * (uint64_t)s);
    const double u2 = uniform01(seed, 6 * i + 2 * (uint64_t)s + 1);
    const double z = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
    const double v = (double)D.seeds[3 * j + s] + (double)D.jitter * z;
    x[s] = (float)fmin(fmax(v, lo), hi);
  }
}

__global__ void thin_samples_kernel(ThinTables T, SeedDist D, int64_t n, uint64_t seed, double lo,
                                    double hi, float4* __restrict__ X, float* __restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[3];
  double N[3], f[kNout];
  sample_point(D, seed, (uint64_t)i, lo, hi, x);
  for (int s = 0; s < 3; s++) N[s] = pow(10.0, (double)x[s]) / T.sig_th[s];
  thin_exact(T, N, f);
  X[i] = make_float4(x[0], x[1], x[2], 0.0f);
  for (int o = 0; o < kNout; o++) Y[8 * i + o] = (float)log10(f[o]);
  Y[8 * i + 6] = 0.0f;
  Y[8 * i + 7] = 0.0f;
}

template <typename T>
__device__ __forceinline__ T wave_sum64(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------------------------------------
// Training: one wavefront (workgroup of 64) per 64 samples, one sample per
// lane.  Forward keeps the pre-activations Z1, Z2 in LDS; the backward pass
// forms dW3, dH2, dW2, dH1 on MFMA and dW1 and the biases on the VALU.
// Each wavefront writes its gradient to its own slot of `partial`; the
// Adam kernel sums the slots in a fixed order (deterministic).
struct TrainArgs {
  MlpDev net;
  const float4* X;     // [n] log10 tau
  const float* Y;      // [n][8] log10 kernels
  int64_t n;
  int64_t offset;      // first sample of the batch
  float gscale;        // 2 / (batch * kNout)
  float* partial;      // [nwave][kNparamPad]
  double* loss_part;   // [nwave] sum of squared normalised errors
};

__global__ __launch_bounds__(64) void mlp_train_kernel(TrainArgs a) {
  __shared__ float Z1[64 * kLdH], Z2[64 * kLdH], C[64 * kLdH], DY[64 * 32], XS[64 * 4];
  const int l = threadIdx.x;
  const int wv = blockIdx.x;
  const MlpDev& net = a.net;
  const float* P = net.Pq;
  const __bf16* Wb = net.Wb;
  float* G = a.partial + (int64_t)wv * kNparamPad;
  const int64_t s0 = a.offset + 64 * (int64_t)wv;

  const float4 xr = a.X[(s0 + l) % a.n];
  const float x0 = (xr.x - net.xc) * net.xs, x1 = (xr.y - net.xc) * net.xs,
              x2 = (xr.z - net.xc) * net.xs;
  XS[4 * l + 0] = x0;
  XS[4 * l + 1] = x1;
  XS[4 * l + 2] = x2;
  XS[4 * l + 3] = 0.0f;
  for (int j = 0; j < kHidden; j++) {
    float z = P[kOffB1 + j];
    z = __builtin_fmaf(x0, P[kOffW1 + j], z);
    z = __builtin_fmaf(x1, P[kOffW1 + kHidden + j], z);
    z = __builtin_fmaf(x2, P[kOffW1 + 2 * kHidden + j], z);
    Z1[l * kLdH + j] = z;
  }
  for (int j = 0; j < 32; j++) DY[l * 32 + j] = 0.0f;
  __syncthreads();

  // Z2 = silu(Z1) W2 + b2
#pragma unroll 1
  for (int mb = 0; mb < 4; mb++) {
    f32x4 acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float v[8];
      lds_frag_a(Z1, kLdH, 16 * mb, 32 * kb, v);
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = silu(v[i]);
      bf16x8 hi, lo;
      split_bf16(v, hi, lo);
#pragma unroll
      for (int nb = 0; nb < 4; nb++) {
        const bf16x8 w = w_frag(Wb + kW2T, kHidden, 32 * kb, 16 * nb);
        acc[nb] = mfma(hi, w, acc[nb]);
        acc[nb] = mfma(lo, w, acc[nb]);
      }
    }
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const float b = P[kOffB2 + 16 * nb + (l & 15)];
#pragma unroll
      for (int i = 0; i < 4; i++) acc[nb][i] += b;
      lds_store_d(Z2, kLdH, 16 * mb, 16 * nb, acc[nb]);
    }
  }
  __syncthreads();

  // y = silu(Z2) W3 + b3, loss, dY = gscale (y - t)
  double lsum = 0.0;
  const int oc = l & 15;
#pragma unroll 1
  for (int mb = 0; mb < 4; mb++) {
    f32x4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float v[8];
      lds_frag_a(Z2, kLdH, 16 * mb, 32 * kb, v);
#pragma unroll
      for (int i = 0; i < 8; i++) v[i] = silu(v[i]);
      bf16x8 hi, lo;
      split_bf16(v, hi, lo);
      const bf16x8 w = w_frag(Wb + kW3T, kHidden, 32 * kb, 0);
      o = mfma(hi, w, o);
      o = mfma(lo, w, o);
    }
    if (oc < kNout) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 16 * mb + 4 * (l >> 4) + i;
        const float t = (a.Y[8 * ((s0 + row) % a.n) + oc] - net.ymean[oc]) / net.ystd[oc];
        const float e = (o[i] + P[kOffB3 + oc]) - t;
        lsum += (double)e * (double)e;
        DY[row * 32 + oc] = a.gscale * e;
      }
    }
  }
  __syncthreads();

  // dW3[unit][out] = silu(Z2)^T dY  (3 MFMAs: hi*hi + lo*hi + hi*lo)
#pragma unroll 1
  for (int mu = 0; mu < 4; mu++) {
    f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float va[8], vb[8];
      lds_frag_at(Z2, kLdH, 16 * mu, 32 * kb, va);
#pragma unroll
      for (int i = 0; i < 8; i++) va[i] = silu(va[i]);
      lds_frag_b(DY, 32, 32 * kb, 0, vb);
      bf16x8 ah, al, bh, bl;
      split_bf16(va, ah, al);
      split_bf16(vb, bh, bl);
      d = mfma(ah, bh, d);
      d = mfma(al, bh, d);
      d = mfma(ah, bl, d);
    }
    if (oc < kNout) {
#pragma unroll
      for (int i = 0; i < 4; i++) G[kOffW3 + (16 * mu + 4 * (l >> 4) + i) * kNout + oc] = d[i];
    }
  }
  if (l < kNout) {
    float sacc = 0.0f;
    for (int s = 0; s < 64; s++) sacc += DY[s * 32 + l];
    G[kOffB3 + l] = sacc;
  }

  // dZ2 = (dY W3^T) * silu'(Z2)  -> C
#pragma unroll 1
  for (int mb = 0; mb < 4; mb++) {
    float va[8];
    lds_frag_a(DY, 32, 16 * mb, 0, va);
    bf16x8 ah, al;
    split_bf16(va, ah, al);
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const bf16x8 w = w_frag(Wb + kW3N, 32, 0, 16 * nb);
      f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
      d = mfma(ah, w, d);
      d = mfma(al, w, d);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 16 * mb + 4 * (l >> 4) + i;
        d[i] *= silu_grad(Z2[row * kLdH + 16 * nb + (l & 15)]);
      }
      lds_store_d(C, kLdH, 16 * mb, 16 * nb, d);
    }
  }
  __syncthreads();
  {
    float sacc = 0.0f;
    for (int s = 0; s < 64; s++) sacc += C[s * kLdH + l];
    G[kOffB2 + l] = sacc;
  }

  // dW2[in][out] = silu(Z1)^T dZ2
#pragma unroll 1
  for (int mi = 0; mi < 4; mi++) {
    f32x4 d[4];
#pragma unroll
    for (int no = 0; no < 4; no++) d[no] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float va[8];
      lds_frag_at(Z1, kLdH, 16 * mi, 32 * kb, va);
#pragma unroll
      for (int i = 0; i < 8; i++) va[i] = silu(va[i]);
      bf16x8 ah, al;
      split_bf16(va, ah, al);
#pragma unroll
      for (int no = 0; no < 4; no++) {
        float vb[8];
        lds_frag_b(C, kLdH, 32 * kb, 16 * no, vb);
        bf16x8 bh, bl;
        split_bf16(vb, bh, bl);
        d[no] = mfma(ah, bh, d[no]);
        d[no] = mfma(al, bh, d[no]);
        d[no] = mfma(ah, bl, d[no]);
      }
    }
#pragma unroll
    for (int no = 0; no < 4; no++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        G[kOffW2 + (16 * mi + 4 * (l >> 4) + i) * kHidden + 16 * no + (l & 15)] = d[no][i];
  }

  // dZ1 = (dZ2 W2^T) * silu'(Z1)  -> C (after every read of dZ2)
  f32x4 dh[4][4];
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
#pragma unroll
    for (int ni = 0; ni < 4; ni++) dh[mb][ni] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      float va[8];
      lds_frag_a(C, kLdH, 16 * mb, 32 * kb, va);
      bf16x8 ah, al;
      split_bf16(va, ah, al);
#pragma unroll
      for (int ni = 0; ni < 4; ni++) {
        const bf16x8 w = w_frag(Wb + kW2N, kHidden, 32 * kb, 16 * ni);
        dh[mb][ni] = mfma(ah, w, dh[mb][ni]);
        dh[mb][ni] = mfma(al, w, dh[mb][ni]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int mb = 0; mb < 4; mb++)
#pragma unroll
    for (int ni = 0; ni < 4; ni++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 16 * mb + 4 * (l >> 4) + i;
        dh[mb][ni][i] *= silu_grad(Z1[row * kLdH + 16 * ni + (l & 15)]);
      }
      lds_store_d(C, kLdH, 16 * mb, 16 * ni, dh[mb][ni]);
    }
  __syncthreads();

  // dW1[i][j] = sum_s x_i dZ1[s][j], db1[j] = sum_s dZ1[s][j]  (lane j)
  {
    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, gb = 0.0f;
    for (int s = 0; s < 64; s++) {
      const float d = C[s * kLdH + l];
      g0 = __builtin_fmaf(XS[4 * s + 0], d, g0);
      g1 = __builtin_fmaf(XS[4 * s + 1], d, g1);
      g2 = __builtin_fmaf(XS[4 * s + 2], d, g2);
      gb += d;
    }
    G[kOffW1 + l] = g0;
    G[kOffW1 + kHidden + l] = g1;
    G[kOffW1 + 2 * kHidden + l] = g2;
    G[kOffB1 + l] = gb;
  }
  lsum = wave_sum64(lsum);
  if (l == 0) a.loss_part[wv] = lsum;
}

// Quantized copies of parameter p: Pq (weights rounded to bf16) and the
// bf16 images of W2 and W3.
__device__ inline void quantize_param(int p, float val, float* Pq, __bf16* Wb) {
  if (p < kOffB1) {
    Pq[p] = (float)(__bf16)val;
  } else if (p < kOffW2) {
    Pq[p] = val;
  } else if (p < kOffB2) {
    const __bf16 q = (__bf16)val;
    Pq[p] = (float)q;
    const int e = p - kOffW2, in = e / kHidden, out = e % kHidden;
    Wb[kW2T + out * kHidden + in] = q;
    Wb[kW2N + in * kHidden + out] = q;
  } else if (p < kOffW3) {
    Pq[p] = val;
  } else if (p < kOffB3) {
    const __bf16 q = (__bf16)val;
    Pq[p] = (float)q;
    const int e = p - kOffW3, unit = e / kNout, out = e % kNout;
    Wb[kW3T + out * kHidden + unit] = q;
    Wb[kW3N + unit * 32 + out] = q;
  } else {
    Pq[p] = val;
  }
}

__global__ void mlp_quantize_kernel(const float* __restrict__ P, float* Pq, __bf16* Wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNparam) quantize_param(p, P[p], Pq, Wb);
}

struct AdamArgs {
  float* P;
  float* m;
  float* v;
  float* Pq;
  __bf16* Wb;
  const float* partial;
  int nwave;
  const double* loss_part;
  double* loss_out;      // loss_out[0] = mean normalised squared error of the batch
  double inv_count;      // 1 / (batch * kNout)
  float lr, beta1, beta2, eps, bc1, bc2;
  int apply;             // 0: only reduce the gradient into `grad`
  float* grad;
};

// Sum of the per-wavefront gradients in wavefront order (FP64), then one
// Adam step and re-quantization.
__global__ void mlp_adam_kernel(AdamArgs a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && a.loss_out) {
    double s = 0.0;
    for (int w = 0; w < a.nwave; w++) s += a.loss_part[w];
    a.loss_out[0] = s * a.inv_count;
  }
  if (p >= kNparam) return;
  double gs = 0.0;
  for (int w = 0; w < a.nwave; w++) gs += (double)a.partial[(int64_t)w * kNparamPad + p];
  const float g = (float)gs;
  if (a.grad) a.grad[p] = g;
  if (!a.apply) return;
  const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
  const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * (g * g);
  a.m[p] = m;
  a.v[p] = v;
  const float val = a.P[p] - a.lr * (m / a.bc1) / (sqrtf(v / a.bc2) + a.eps);
  a.P[p] = val;
  quantize_param(p, val, a.Pq, a.Wb);
}

__global__ __launch_bounds__(64) void mlp_eval_kernel(MlpDev net, const float* __restrict__ x,
                                                      float* __restrict__ y, int64_t n) {
  __shared__ float lds[kMlpLds];
  const int64_t i = 64 * (int64_t)blockIdx.x + threadIdx.x;
  const int64_t ic = i < n ? i : n - 1;
  float out[kNout];
  mlp_forward(net, x[3 * ic], x[3 * ic + 1], x[3 * ic + 2], out, lds);
  if (i < n)
    for (int o = 0; o < kNout; o++) y[kNout * i + o] = out[o];
}

// Validation: 64 uniform points in each trust cell, exact kernels (FP64)
// against the network; relative error max over the six outputs.
__global__ __launch_bounds__(64) void mlp_trust_kernel(MlpDev net, ThinTables T, int tn, double lo,
                                                       double hi, uint64_t seed,
                                                       float* __restrict__ cell_err,
                                                       float* __restrict__ pt_err) {
  __shared__ float lds[kMlpLds];
  const int64_t c = blockIdx.x;
  const int l = threadIdx.x;
  const int idx[3] = {(int)(c / ((int64_t)tn * tn)), (int)((c / tn) % tn), (int)(c % tn)};
  const double h = (hi - lo) / tn;
  float x[3];
  double N[3], f[kNout];
  for (int s = 0; s < 3; s++) {
    x[s] = (float)(lo + ((double)idx[s] + uniform01(seed, 3 * (64 * (uint64_t)c + l) + s)) * h);
    N[s] = pow(10.0, (double)x[s]) / T.sig_th[s];
  }
  thin_exact(T, N, f);
  float y[kNout];
  mlp_forward(net, x[0], x[1], x[2], y, lds);
  float e = 0.0f;
  for (int o = 0; o < kNout; o++) {
    const double r = fabs(expm1(((double)y[o] - log10(f[o])) * 2.302585092994046));
    e = fmaxf(e, (float)fmin(r, 1.0e30));
  }
  pt_err[64 * c + l] = e;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) e = fmaxf(e, __shfl_xor(e, off, 64));
  if (l == 0) cell_err[c] = e;
}

// Training targets of a network with a baseline: Y -= baseline(z).
__global__ void poly_residual_kernel(const float4* __restrict__ X, float* __restrict__ Y,
                                     int64_t n, float xc, float xs, const float* __restrict__ poly) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = X[i];
  float b[kNout];
  poly_eval(poly, (x.x - xc) * xs, (x.y - xc) * xs, (x.z - xc) * xs, b);
  for (int o = 0; o < kNout; o++) Y[8 * i + o] -= b[o];
}

// Validation with seed points: n points drawn like the training samples
// (independent counter seed), the network against the exact kernels (FP64);
// each point's relative error (max over the outputs) goes to pt_err and,
// as a maximum, to the trust cell it lies in (cell index computed as the
// step kernel computes it), with a count of points per cell.
__global__ __launch_bounds__(64) void mlp_validate_points_kernel(
    MlpDev net, ThinTables T, SeedDist D, int64_t n, uint64_t seed, double lo, double hi, int tn,
    float flo, float inv_cell, unsigned* __restrict__ cell_bits, unsigned* __restrict__ cell_cnt,
    float* __restrict__ pt_err, int* __restrict__ pt_cell) {
  __shared__ float lds[kMlpLds];
  const int64_t i = 64 * (int64_t)blockIdx.x + threadIdx.x;
  const int64_t ic = i < n ? i : n - 1;
  float x[3];
  double N[3], f[kNout];
  sample_point(D, seed, (uint64_t)ic, lo, hi, x);
  for (int s = 0; s < 3; s++) N[s] = pow(10.0, (double)x[s]) / T.sig_th[s];
  thin_exact(T, N, f);
  float y[kNout];
  mlp_forward(net, x[0], x[1], x[2], y, lds);
  float e = 0.0f;
  for (int o = 0; o < kNout; o++) {
    double r = fabs(expm1(((double)y[o] - log10(f[o])) * 2.302585092994046));
    if (!(r < 1.0e30)) r = 1.0e30;
    e = fmaxf(e, (float)r);
  }
  if (i < n) {
    pt_err[i] = e;
    int c[3];
    for (int s = 0; s < 3; s++) c[s] = min(tn - 1, max(0, (int)((x[s] - flo) * inv_cell)));
    const int64_t cell = ((int64_t)c[0] * tn + c[1]) * tn + c[2];
    pt_cell[i] = (int)cell;
    atomicMax(&cell_bits[cell], __float_as_uint(e));
    atomicAdd(&cell_cnt[cell], 1u);
  }
}

// ------------------------------------------------------------------------
// Host side.

namespace {

struct SurrCtx {
  bool have_net = false;     // parameters present (trained or set)
  bool trained = false;      // trust grid valid
  hipStream_t stream = nullptr;
  int ngrid = 0;
  std::vector<double> shape;
  float sig_th[3] = {0, 0, 0};
  float lo = -6.0f, hi = 4.0f, dtau_max = 0.8f, tol = 1e-3f;
  int trust_n = 0;
  float xc = -1.0f, xs = 0.2f, ymean[kNout] = {}, ystd[kNout] = {1, 1, 1, 1, 1, 1};
  float *P = nullptr, *m = nullptr, *v = nullptr, *Pq = nullptr;
  __bf16* Wb = nullptr;
  unsigned char* trust = nullptr;
  float* cell_err = nullptr;
  int64_t ntrust = 0;        // trust cells
  int64_t ntrusted = 0;      // of which within tolerance
  float* poly = nullptr;     // [kNpoly * kNout] baseline coefficients (device)
  int npoly = 0;             // 0: no baseline
};

SurrCtx s_;

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

#define S_OK(expr)                                                                     \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      argot_hip_internal_set_error(_e == hipErrorOutOfMemory ? ARGOT_ERR_NOMEM         \
                                                             : ARGOT_ERR_HIP);         \
      return;                                                                          \
    }                                                                                  \
  } while (0)

#define S_ARG(cond)                                          \
  do {                                                       \
    if (!(cond)) {                                           \
      argot_hip_internal_set_error(ARGOT_ERR_ARG);           \
      return;                                                \
    }                                                        \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
}

// Persistent parameter buffers.
hipError_t ensure_params() {
  hipError_t e = hipSuccess;
  if (!s_.stream && (e = hipStreamCreate(&s_.stream)) != hipSuccess) return e;
  if (s_.P) return hipSuccess;
  if ((e = dalloc(&s_.P, kNparamPad)) != hipSuccess) return e;
  if ((e = dalloc(&s_.m, kNparamPad)) != hipSuccess) return e;
  if ((e = dalloc(&s_.v, kNparamPad)) != hipSuccess) return e;
  if ((e = dalloc(&s_.Pq, kNparamPad)) != hipSuccess) return e;
  if ((e = dalloc(&s_.Wb, kNbf16)) != hipSuccess) return e;
  if ((e = dalloc(&s_.poly, kNpoly * kNout)) != hipSuccess) return e;
  if ((e = hipMemset(s_.poly, 0, sizeof(float) * kNpoly * kNout)) != hipSuccess) return e;
  return hipMemset(s_.Wb, 0, sizeof(__bf16) * kNbf16);
}

MlpDev net_dev() {
  MlpDev d{};
  d.Pq = s_.Pq;
  d.Wb = s_.Wb;
  d.xc = s_.xc;
  d.xs = s_.xs;
  for (int o = 0; o < kNout; o++) {
    d.ymean[o] = s_.ymean[o];
    d.ystd[o] = s_.ystd[o];
  }
  d.poly = s_.poly;
  d.npoly = s_.npoly;
  return d;
}

// Monomials of the baseline at normalised input z, as poly_eval forms them.
void monomials_host(float z0, float z1, float z2, double m[kNpoly]) {
  float f[kNpoly];
  f[0] = 1.0f;
  f[1] = z0;
  f[2] = z1;
  f[3] = z2;
  f[4] = z0 * z0;
  f[5] = z0 * z1;
  f[6] = z0 * z2;
  f[7] = z1 * z1;
  f[8] = z1 * z2;
  f[9] = z2 * z2;
  f[10] = f[4] * z0;
  f[11] = f[4] * z1;
  f[12] = f[4] * z2;
  f[13] = f[7] * z0;
  f[14] = f[5] * z2;
  f[15] = f[9] * z0;
  f[16] = f[7] * z1;
  f[17] = f[7] * z2;
  f[18] = f[9] * z1;
  f[19] = f[9] * z2;
  for (int k = 0; k < kNpoly; k++) m[k] = f[k];
}

// Least-squares cubic baseline of each output over the samples (x, y):
// normal equations in FP64 with a relative ridge of 1e-12, solved by
// Gaussian elimination with partial pivoting.  Returns false if singular.
bool fit_baseline(const std::vector<float4>& x, const std::vector<float>& y, float xc, float xs,
                  float coef[kNpoly * kNout]) {
  const int K = kNpoly;
  std::vector<double> A(K * K, 0.0), B(K * kNout, 0.0);
  double m[kNpoly];
  for (size_t i = 0; i < x.size(); i++) {
    monomials_host((x[i].x - xc) * xs, (x[i].y - xc) * xs, (x[i].z - xc) * xs, m);
    for (int a = 0; a < K; a++) {
      for (int b = 0; b < K; b++) A[a * K + b] += m[a] * m[b];
      for (int o = 0; o < kNout; o++) B[a * kNout + o] += m[a] * y[8 * i + o];
    }
  }
  for (int a = 0; a < K; a++) A[a * K + a] *= 1.0 + 1e-12;
  for (int col = 0; col < K; col++) {
    int piv = col;
    for (int r = col + 1; r < K; r++)
      if (fabs(A[r * K + col]) > fabs(A[piv * K + col])) piv = r;
    if (!(fabs(A[piv * K + col]) > 0.0)) return false;
    if (piv != col) {
      for (int c = 0; c < K; c++) std::swap(A[col * K + c], A[piv * K + c]);
      for (int o = 0; o < kNout; o++) std::swap(B[col * kNout + o], B[piv * kNout + o]);
    }
    for (int r = 0; r < K; r++) {
      if (r == col) continue;
      const double f = A[r * K + col] / A[col * K + col];
      if (f == 0.0) continue;
      for (int c = col; c < K; c++) A[r * K + c] -= f * A[col * K + c];
      for (int o = 0; o < kNout; o++) B[r * kNout + o] -= f * B[col * kNout + o];
    }
  }
  for (int a = 0; a < K; a++)
    for (int o = 0; o < kNout; o++) {
      const double v = B[a * kNout + o] / A[a * K + a];
      if (!std::isfinite(v)) return false;
      coef[a * kNout + o] = (float)v;
    }
  return true;
}

// Device tables of the exact integral for cfg (freed by the caller).
struct ThinHost {
  ThinTables t{};
  double *nu = nullptr, *cs = nullptr, *S = nullptr;
  std::vector<double> shape;
  float sig_th[3] = {0, 0, 0};
  void release() {
    if (nu) (void)hipFree(nu);
    if (cs) (void)hipFree(cs);
    if (S) (void)hipFree(S);
    nu = cs = S = nullptr;
  }
};

int thin_tables(const argot_surrogate_config* c, ThinHost* th) {
  if (!c || c->ngrid <= 0 || c->ngrid > 4096 || !c->nu || !c->cs || !c->spec) return ARGOT_ERR_ARG;
  const int ng = c->ngrid;
  std::vector<double> cs(3 * (size_t)ng), S(ng);
  double tot = 0.0;
  for (int k = 0; k < ng; k++) {
    if (!(c->spec[k] >= 0.0)) return ARGOT_ERR_ARG;
    tot += c->spec[k];
  }
  if (!(tot > 0.0)) return ARGOT_ERR_ARG;
  for (int k = 0; k < ng; k++) S[k] = c->spec[k] / tot;
  for (int s = 0; s < 3; s++) {
    th->sig_th[s] = 0.0f;
    for (int k = 0; k < ng; k++) {
      cs[3 * k + s] = (double)c->cs[3 * k + s];
      if (th->sig_th[s] == 0.0f && c->cs[3 * k + s] > 0.0f) th->sig_th[s] = c->cs[3 * k + s];
    }
    if (!(th->sig_th[s] > 0.0f)) return ARGOT_ERR_ARG;
  }
  if (dalloc(&th->nu, ng) != hipSuccess || dalloc(&th->cs, 3 * (size_t)ng) != hipSuccess ||
      dalloc(&th->S, ng) != hipSuccess)
    return ARGOT_ERR_NOMEM;
  if (hipMemcpy(th->nu, c->nu, sizeof(double) * ng, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(th->cs, cs.data(), sizeof(double) * 3 * ng, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(th->S, S.data(), sizeof(double) * ng, hipMemcpyHostToDevice) != hipSuccess)
    return ARGOT_ERR_HIP;
  th->t.ngrid = ng;
  th->t.nu = th->nu;
  th->t.cs = th->cs;
  th->t.S = th->S;
  for (int s = 0; s < 3; s++) th->t.sig_th[s] = (double)th->sig_th[s];
  th->shape = S;
  return ARGOT_OK;
}

// Xavier-uniform weights, zero biases, from the seed.
void init_params(uint64_t seed, std::vector<float>* P) {
  P->assign(kNparamPad, 0.0f);
  auto fill = [&](int off, int n, double fan_in, double fan_out, uint64_t tag) {
    const double a = sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < n; i++)
      (*P)[off + i] = (float)(a * (2.0 * uniform01(seed ^ tag, (uint64_t)i) - 1.0));
  };
  fill(kOffW1, kNin * kHidden, kNin, kHidden, 0x1111);
  fill(kOffW2, kHidden * kHidden, kHidden, kHidden, 0x2222);
  fill(kOffW3, kHidden * kNout, kHidden, kNout, 0x3333);
}

}  // namespace

bool surrogate_ready() { return s_.have_net && s_.trained; }

bool surrogate_any_trusted() { return surrogate_ready() && s_.ntrusted > 0; }

SurrDev surrogate_dev() {
  SurrDev d{};
  d.net = net_dev();
  d.trust = s_.trust;
  d.trust_n = s_.trust_n;
  d.lo = s_.lo;
  d.hi = s_.hi;
  d.inv_cell = s_.trust_n / (s_.hi - s_.lo);
  for (int s = 0; s < 3; s++) d.sig_th[s] = s_.sig_th[s];
  d.dtau_max = s_.dtau_max;
  d.tau_lo = pow(10.0, (double)s_.lo);
  d.tau_hi = pow(10.0, (double)s_.hi);
  d.node_q = nullptr;
  return d;
}

void surrogate_node_q(const double* spec64, int nnode, int ngrid, std::vector<double>* q) {
  q->assign(std::max(nnode, 1), -1.0);
  if (!surrogate_ready() || ngrid != s_.ngrid) return;
  for (int u = 0; u < nnode; u++) {
    const double* sp = spec64 + (size_t)u * ngrid;
    double Q = 0.0;
    for (int k = 0; k < ngrid; k++) Q += sp[k];
    if (!(Q > 0.0)) continue;
    bool ok = true;
    for (int k = 0; k < ngrid && ok; k++)
      ok = fabs(sp[k] - Q * s_.shape[k]) <= 1.0e-9 * Q * s_.shape[k] + 1.0e-300;
    if (ok) (*q)[u] = Q;
  }
}

// The seed distribution of cfg on the device (*dev freed by the caller):
// ARGOT_OK, or ARGOT_ERR_ARG for an inconsistent one.
int seed_dist(const argot_surrogate_config* cfg, float** dev, SeedDist* D) {
  *dev = nullptr;
  D->seeds = nullptr;
  D->nseed = 0;
  D->jitter = 0.0f;
  if (cfg->nseed == 0) return ARGOT_OK;
  if (cfg->nseed < 0 || !cfg->seed_log_tau || !(cfg->jitter >= 0.0f) ||
      !(cfg->jitter < 10.0f))
    return ARGOT_ERR_ARG;
  for (int64_t i = 0; i < 3 * cfg->nseed; i++)
    if (!std::isfinite(cfg->seed_log_tau[i])) return ARGOT_ERR_ARG;
  if (dalloc(dev, 3 * (size_t)cfg->nseed) != hipSuccess) return ARGOT_ERR_NOMEM;
  if (hipMemcpy(*dev, cfg->seed_log_tau, sizeof(float) * 3 * cfg->nseed,
                hipMemcpyHostToDevice) != hipSuccess)
    return ARGOT_ERR_HIP;
  D->seeds = *dev;
  D->nseed = cfg->nseed;
  D->jitter = cfg->jitter;
  return ARGOT_OK;
}

}  // namespace argot

using namespace argot;

static_assert(kNpoly == ARGOT_SURROGATE_NPOLY, "baseline size of the public header");

extern "C" {

static void argot_hip_surrogate_samples_body(const struct argot_surrogate_config* cfg, int64_t n,
                                             float* log_tau, float* log_kernel) {
  argot_hip_internal_clear_error();
  S_ARG(cfg && n > 0 && log_tau && log_kernel && cfg->log_tau_lo < cfg->log_tau_hi);
  ThinHost th;
  float* dseed = nullptr;
  SeedDist D;
  int e = thin_tables(cfg, &th);
  if (!e) e = seed_dist(cfg, &dseed, &D);
  if (e) {
    th.release();
    if (dseed) (void)hipFree(dseed);
    return argot_hip_internal_set_error(e);
  }
  float4* X = nullptr;
  float* Y = nullptr;
  hipError_t he = dalloc(&X, n);
  if (he == hipSuccess) he = dalloc(&Y, 8 * (size_t)n);
  if (he == hipSuccess) {
    thin_samples_kernel<<<(unsigned)((n + 255) / 256), 256>>>(th.t, D, n, cfg->seed,
                                                              cfg->log_tau_lo, cfg->log_tau_hi, X,
                                                              Y);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipDeviceSynchronize();
  if (dseed) (void)hipFree(dseed);
  std::vector<float4> hx;
  std::vector<float> hy;
  if (he == hipSuccess) {
    hx.resize(n);
    hy.resize(8 * (size_t)n);
    he = hipMemcpy(hx.data(), X, sizeof(float4) * n, hipMemcpyDeviceToHost);
    if (he == hipSuccess) he = hipMemcpy(hy.data(), Y, sizeof(float) * 8 * n, hipMemcpyDeviceToHost);
  }
  if (X) (void)hipFree(X);
  if (Y) (void)hipFree(Y);
  th.release();
  S_OK(he);
  for (int64_t i = 0; i < n; i++) {
    log_tau[3 * i + 0] = hx[i].x;
    log_tau[3 * i + 1] = hx[i].y;
    log_tau[3 * i + 2] = hx[i].z;
    for (int o = 0; o < kNout; o++) log_kernel[kNout * i + o] = hy[8 * i + o];
  }
}

void argot_hip_surrogate_samples(const struct argot_surrogate_config* cfg, int64_t n,
                                 float* log_tau, float* log_kernel) {
  try {
    argot_hip_surrogate_samples_body(cfg, n, log_tau, log_kernel);
  } catch (const std::bad_alloc&) {
    argot_hip_internal_set_error(ARGOT_ERR_NOMEM);
  } catch (...) {
    argot_hip_internal_set_error(ARGOT_ERR_ARG);
  }
}

static void argot_hip_surrogate_train_body(const struct argot_surrogate_config* cfg,
                                           struct argot_surrogate_report* rep) {
  argot_hip_internal_clear_error();
  S_ARG(cfg && rep);
  S_ARG(cfg->batch >= 64 && cfg->batch % 64 == 0 && cfg->nsample >= cfg->batch &&
        cfg->nsample % cfg->batch == 0 && cfg->nstep > 0);
  S_ARG(cfg->log_tau_lo < cfg->log_tau_hi && cfg->trust_n >= 1 && cfg->trust_n <= 128);
  S_ARG(cfg->lr_max > 0.0f && cfg->lr_min >= 0.0f && cfg->trust_tol > 0.0f &&
        cfg->dtau_max > 0.0f);
  S_ARG(cfg->poly_degree == 0 || cfg->poly_degree == 3);
  S_ARG(cfg->nvalidate >= 0 && cfg->nvalidate <= ((int64_t)1 << 30));
  memset(rep, 0, sizeof(*rep));
  S_OK(ensure_params());
  ThinHost th;
  float* dseed = nullptr;
  SeedDist D;
  int e = thin_tables(cfg, &th);
  if (!e) e = seed_dist(cfg, &dseed, &D);
  if (e) {
    th.release();
    if (dseed) (void)hipFree(dseed);
    return argot_hip_internal_set_error(e);
  }
  const bool seeded = D.seeds != nullptr;
  const int64_t n = cfg->nsample;
  const int nwave = cfg->batch / 64;
  const int tn = cfg->trust_n;
  const int64_t ncell = (int64_t)tn * tn * tn;
  // validation points: 64 per trust cell (uniform), or nvalidate drawn
  // like the training samples (seeds; default 2^20)
  const int64_t nval = seeded ? (cfg->nvalidate > 0 ? cfg->nvalidate : ((int64_t)1 << 20))
                              : 64 * ncell;
  float4* X = nullptr;
  float* Y = nullptr;
  float* partial = nullptr;
  double *loss_part = nullptr, *loss_hist = nullptr;
  float *cell_err = nullptr, *pt_err = nullptr;
  unsigned* cell_cnt = nullptr;
  int* pt_cell = nullptr;
  hipError_t he = dalloc(&X, n);
  if (he == hipSuccess) he = dalloc(&Y, 8 * (size_t)n);
  if (he == hipSuccess) he = dalloc(&partial, (size_t)nwave * kNparamPad);
  if (he == hipSuccess) he = dalloc(&loss_part, nwave);
  if (he == hipSuccess) he = dalloc(&loss_hist, cfg->nstep);
  if (he == hipSuccess) he = dalloc(&pt_err, (size_t)nval);
  if (he == hipSuccess && seeded) he = dalloc(&cell_cnt, ncell);
  if (he == hipSuccess && seeded) he = dalloc(&pt_cell, (size_t)nval);
  if (he == hipSuccess && s_.ntrust != ncell) {
    if (s_.trust) (void)hipFree(s_.trust);
    if (s_.cell_err) (void)hipFree(s_.cell_err);
    s_.trust = nullptr;
    s_.cell_err = nullptr;
    s_.ntrust = 0;
    he = dalloc(&s_.trust, ncell);
    if (he == hipSuccess) he = dalloc(&s_.cell_err, ncell);
    if (he == hipSuccess) s_.ntrust = ncell;
  }
  cell_err = s_.cell_err;
  auto cleanup = [&]() {
    if (X) (void)hipFree(X);
    if (Y) (void)hipFree(Y);
    if (partial) (void)hipFree(partial);
    if (loss_part) (void)hipFree(loss_part);
    if (loss_hist) (void)hipFree(loss_hist);
    if (pt_err) (void)hipFree(pt_err);
    if (cell_cnt) (void)hipFree(cell_cnt);
    if (pt_cell) (void)hipFree(pt_cell);
    if (dseed) (void)hipFree(dseed);
    th.release();
  };
  if (he != hipSuccess) {
    cleanup();
    S_OK(he);
  }
  s_.trained = false;
  s_.have_net = false;
  hipStream_t st = s_.stream;

  // 1. samples of the exact integral
  double t0 = now_ms();
  thin_samples_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(th.t, D, n, cfg->seed,
                                                                   cfg->log_tau_lo,
                                                                   cfg->log_tau_hi, X, Y);
  he = hipGetLastError();
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  rep->ms_samples = now_ms() - t0;
  // input normalisation: the hull mapped to [-1, 1]; with seed points the
  // mean and spread of their coordinates instead (one pair for the three
  // axes), so that the network's inputs span the range the rays occupy
  s_.xc = 0.5f * (cfg->log_tau_lo + cfg->log_tau_hi);
  s_.xs = 2.0f / (cfg->log_tau_hi - cfg->log_tau_lo);
  if (seeded) {
    double m1 = 0.0, m2 = 0.0;
    const int64_t m = 3 * cfg->nseed;
    for (int64_t i = 0; i < m; i++) m1 += cfg->seed_log_tau[i];
    m1 /= (double)m;
    for (int64_t i = 0; i < m; i++) {
      const double d = cfg->seed_log_tau[i] - m1;
      m2 += d * d;
    }
    const double sd = sqrt(m2 / (double)m) + (double)cfg->jitter;
    if (sd > 0.0 && std::isfinite(sd)) {
      s_.xc = (float)m1;
      s_.xs = (float)(1.0 / (2.0 * sd));
    }
  }

  // 2a. cubic baseline (least squares on the first samples); the network
  // learns what it leaves
  s_.npoly = 0;
  if (he == hipSuccess && cfg->poly_degree == 3) {
    const int64_t m = std::min<int64_t>(n, 65536);
    std::vector<float4> hx(m);
    std::vector<float> hy(8 * (size_t)m);
    he = hipMemcpy(hx.data(), X, sizeof(float4) * m, hipMemcpyDeviceToHost);
    if (he == hipSuccess) he = hipMemcpy(hy.data(), Y, sizeof(float) * 8 * m, hipMemcpyDeviceToHost);
    float coef[kNpoly * kNout];
    if (he == hipSuccess && !fit_baseline(hx, hy, s_.xc, s_.xs, coef)) {
      cleanup();
      return argot_hip_internal_set_error(ARGOT_ERR_ARG);
    }
    if (he == hipSuccess)
      he = hipMemcpy(s_.poly, coef, sizeof(coef), hipMemcpyHostToDevice);
    if (he == hipSuccess) {
      poly_residual_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(X, Y, n, s_.xc, s_.xs,
                                                                        s_.poly);
      he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he == hipSuccess) s_.npoly = kNpoly;
  }

  // 2b. normalisation from the first samples
  if (he == hipSuccess) {
    const int64_t m = std::min<int64_t>(n, 65536);
    std::vector<float> hy(8 * (size_t)m);
    he = hipMemcpy(hy.data(), Y, sizeof(float) * 8 * m, hipMemcpyDeviceToHost);
    for (int o = 0; o < kNout && he == hipSuccess; o++) {
      double s1 = 0.0, s2 = 0.0;
      for (int64_t i = 0; i < m; i++) {
        s1 += hy[8 * i + o];
        s2 += (double)hy[8 * i + o] * hy[8 * i + o];
      }
      const double mu = s1 / m;
      s_.ymean[o] = (float)mu;
      s_.ystd[o] = (float)std::max(sqrt(std::max(s2 / m - mu * mu, 0.0)), 1e-6);
    }
  }

  // 3. initial parameters, Adam state
  if (he == hipSuccess) {
    std::vector<float> P0;
    init_params(cfg->seed, &P0);
    he = hipMemcpy(s_.P, P0.data(), sizeof(float) * kNparamPad, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemset(s_.m, 0, sizeof(float) * kNparamPad);
    if (he == hipSuccess) he = hipMemset(s_.v, 0, sizeof(float) * kNparamPad);
    if (he == hipSuccess) {
      mlp_quantize_kernel<<<(kNparam + 255) / 256, 256, 0, st>>>(s_.P, s_.Pq, s_.Wb);
      he = hipGetLastError();
    }
  }

  // 4. training
  t0 = now_ms();
  const MlpDev net = net_dev();
  const float b1 = 0.9f, b2 = 0.999f;
  double p1 = 1.0, p2 = 1.0;
  const int warm = std::max(1, cfg->nstep / 50);
  for (int step = 0; step < cfg->nstep && he == hipSuccess; step++) {
    TrainArgs ta;
    ta.net = net;
    ta.X = X;
    ta.Y = Y;
    ta.n = n;
    ta.offset = ((int64_t)step * cfg->batch) % n;
    ta.gscale = 2.0f / (float)(cfg->batch * kNout);
    ta.partial = partial;
    ta.loss_part = loss_part;
    mlp_train_kernel<<<nwave, 64, 0, st>>>(ta);
    p1 *= b1;
    p2 *= b2;
    double lr;
    if (step < warm) {
      lr = cfg->lr_max * (step + 1) / (double)warm;
    } else {
      const double x = (double)(step - warm) / std::max(1, cfg->nstep - warm);
      lr = cfg->lr_min + 0.5 * (cfg->lr_max - cfg->lr_min) * (1.0 + cos(M_PI * x));
    }
    AdamArgs aa;
    aa.P = s_.P;
    aa.m = s_.m;
    aa.v = s_.v;
    aa.Pq = s_.Pq;
    aa.Wb = s_.Wb;
    aa.partial = partial;
    aa.nwave = nwave;
    aa.loss_part = loss_part;
    aa.loss_out = loss_hist + step;
    aa.inv_count = 1.0 / ((double)cfg->batch * kNout);
    aa.lr = (float)lr;
    aa.beta1 = b1;
    aa.beta2 = b2;
    aa.eps = 1e-8f;
    aa.bc1 = (float)(1.0 - p1);
    aa.bc2 = (float)(1.0 - p2);
    aa.apply = 1;
    aa.grad = nullptr;
    mlp_adam_kernel<<<(kNparam + 255) / 256, 256, 0, st>>>(aa);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(st);
  rep->ms_train = now_ms() - t0;
  if (he == hipSuccess) {
    std::vector<double> lh(cfg->nstep);
    he = hipMemcpy(lh.data(), loss_hist, sizeof(double) * cfg->nstep, hipMemcpyDeviceToHost);
    rep->loss_first = lh[0];
    const int tail = std::max(1, cfg->nstep / 100);
    double s = 0.0;
    for (int i = cfg->nstep - tail; i < cfg->nstep; i++) s += lh[i];
    rep->loss_last = s / tail;
  }

  // 5. validation on the trust grid
  t0 = now_ms();
  if (he == hipSuccess && seeded) {
    // points drawn like the samples, binned into the trust cells: a cell is
    // trusted with at least 8 points, all within the tolerance
    he = hipMemsetAsync(cell_err, 0, sizeof(float) * ncell, st);
    if (he == hipSuccess) he = hipMemsetAsync(cell_cnt, 0, sizeof(unsigned) * ncell, st);
    const float flo = cfg->log_tau_lo;
    const float inv_cell = tn / (cfg->log_tau_hi - cfg->log_tau_lo);
    if (he == hipSuccess) {
      mlp_validate_points_kernel<<<(unsigned)((nval + 63) / 64), 64, 0, st>>>(
          net, th.t, D, nval, cfg->seed ^ 0x7a11da7eull, cfg->log_tau_lo, cfg->log_tau_hi, tn,
          flo, inv_cell, reinterpret_cast<unsigned*>(cell_err), cell_cnt, pt_err, pt_cell);
      he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    std::vector<float> ce(ncell), pe(nval);
    std::vector<unsigned> cc(ncell);
    std::vector<int> pc(nval);
    if (he == hipSuccess)
      he = hipMemcpy(ce.data(), cell_err, sizeof(float) * ncell, hipMemcpyDeviceToHost);
    if (he == hipSuccess)
      he = hipMemcpy(cc.data(), cell_cnt, sizeof(unsigned) * ncell, hipMemcpyDeviceToHost);
    if (he == hipSuccess)
      he = hipMemcpy(pe.data(), pt_err, sizeof(float) * nval, hipMemcpyDeviceToHost);
    if (he == hipSuccess)
      he = hipMemcpy(pc.data(), pt_cell, sizeof(int) * nval, hipMemcpyDeviceToHost);
    if (he == hipSuccess) {
      std::vector<unsigned char> tr(ncell);
      int64_t ntr = 0;
      for (int64_t c = 0; c < ncell; c++) {
        tr[c] = cc[c] >= 8 && ce[c] <= cfg->trust_tol ? 1 : 0;
        ntr += tr[c];
        if (cc[c] == 0) ce[c] = INFINITY;
      }
      he = hipMemcpy(s_.trust, tr.data(), ncell, hipMemcpyHostToDevice);
      if (he == hipSuccess)
        he = hipMemcpy(cell_err, ce.data(), sizeof(float) * ncell, hipMemcpyHostToDevice);
      // the fraction of validation points that lie in trusted cells (the
      // share of such rays the network may serve)
      int64_t inside = 0;
      for (int64_t i = 0; i < nval; i++) inside += tr[pc[i]];
      s_.ntrusted = ntr;
      std::sort(pe.begin(), pe.end());
      rep->err_max = pe.back();
      rep->err_p99 = pe[(size_t)(0.99 * (pe.size() - 1))];
      rep->err_median = pe[pe.size() / 2];
      rep->trusted_frac = (double)inside / nval;
    }
  } else if (he == hipSuccess) {
    mlp_trust_kernel<<<(unsigned)ncell, 64, 0, st>>>(net, th.t, tn, cfg->log_tau_lo,
                                                      cfg->log_tau_hi, cfg->seed ^ 0x7a11da7eull,
                                                      cell_err, pt_err);
    he = hipGetLastError();
    if (he == hipSuccess) he = hipStreamSynchronize(st);
  }
  if (he == hipSuccess && !seeded) {
    std::vector<float> ce(ncell), pe(64 * (size_t)ncell);
    he = hipMemcpy(ce.data(), cell_err, sizeof(float) * ncell, hipMemcpyDeviceToHost);
    if (he == hipSuccess)
      he = hipMemcpy(pe.data(), pt_err, sizeof(float) * pe.size(), hipMemcpyDeviceToHost);
    if (he == hipSuccess) {
      std::vector<unsigned char> tr(ncell);
      int64_t ntr = 0;
      for (int64_t c = 0; c < ncell; c++) {
        tr[c] = ce[c] <= cfg->trust_tol ? 1 : 0;
        ntr += tr[c];
      }
      he = hipMemcpy(s_.trust, tr.data(), ncell, hipMemcpyHostToDevice);
      s_.ntrusted = ntr;
      rep->trusted_frac = (double)ntr / ncell;
      std::sort(pe.begin(), pe.end());
      rep->err_max = pe.back();
      rep->err_p99 = pe[(size_t)(0.99 * (pe.size() - 1))];
      rep->err_median = pe[pe.size() / 2];
    }
  }
  rep->ms_validate = now_ms() - t0;
  cleanup();
  S_OK(he);
  s_.ngrid = cfg->ngrid;
  s_.shape = th.shape;
  for (int s = 0; s < 3; s++) s_.sig_th[s] = th.sig_th[s];
  s_.lo = cfg->log_tau_lo;
  s_.hi = cfg->log_tau_hi;
  s_.trust_n = tn;
  s_.tol = cfg->trust_tol;
  s_.dtau_max = cfg->dtau_max;
  s_.have_net = true;
  s_.trained = true;
}

void argot_hip_surrogate_train(const struct argot_surrogate_config* cfg,
                               struct argot_surrogate_report* rep) {
  try {
    argot_hip_surrogate_train_body(cfg, rep);
  } catch (const std::bad_alloc&) {
    argot_hip_internal_set_error(ARGOT_ERR_NOMEM);
  } catch (...) {
    argot_hip_internal_set_error(ARGOT_ERR_ARG);
  }
}

void argot_hip_surrogate_get_params(float* params, float* norm) {
  argot_hip_internal_clear_error();
  S_ARG(params && norm);
  if (!s_.have_net) return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  S_OK(hipMemcpy(params, s_.P, sizeof(float) * kNparam, hipMemcpyDeviceToHost));
  norm[0] = s_.xc;
  norm[1] = s_.xs;
  for (int o = 0; o < kNout; o++) {
    norm[2 + o] = s_.ymean[o];
    norm[2 + kNout + o] = s_.ystd[o];
  }
}

void argot_hip_surrogate_set_params(const float* params, const float* norm) {
  argot_hip_internal_clear_error();
  S_ARG(params && norm && norm[1] != 0.0f);
  for (int o = 0; o < kNout; o++) S_ARG(norm[2 + kNout + o] != 0.0f);
  S_OK(ensure_params());
  S_OK(hipMemcpy(s_.P, params, sizeof(float) * kNparam, hipMemcpyHostToDevice));
  S_OK(hipMemset(s_.m, 0, sizeof(float) * kNparamPad));
  S_OK(hipMemset(s_.v, 0, sizeof(float) * kNparamPad));
  s_.xc = norm[0];
  s_.xs = norm[1];
  for (int o = 0; o < kNout; o++) {
    s_.ymean[o] = norm[2 + o];
    s_.ystd[o] = norm[2 + kNout + o];
  }
  mlp_quantize_kernel<<<(kNparam + 255) / 256, 256, 0, s_.stream>>>(s_.P, s_.Pq, s_.Wb);
  S_OK(hipGetLastError());
  S_OK(hipStreamSynchronize(s_.stream));
  s_.npoly = 0;         // no baseline until argot_hip_surrogate_set_baseline
  s_.have_net = true;
  s_.trained = false;   // the trust grid belongs to the trained network
}

int argot_hip_surrogate_get_baseline(float* coef) {
  argot_hip_internal_clear_error();
  if (!coef) {
    argot_hip_internal_set_error(ARGOT_ERR_ARG);
    return 0;
  }
  if (!s_.have_net) {
    argot_hip_internal_set_error(ARGOT_ERR_STATE);
    return 0;
  }
  if (s_.npoly == 0) {
    memset(coef, 0, sizeof(float) * kNpoly * kNout);
    return 0;
  }
  if (hipMemcpy(coef, s_.poly, sizeof(float) * kNpoly * kNout, hipMemcpyDeviceToHost) !=
      hipSuccess) {
    argot_hip_internal_set_error(ARGOT_ERR_HIP);
    return 0;
  }
  return s_.npoly;
}

void argot_hip_surrogate_set_baseline(const float* coef, int npoly) {
  argot_hip_internal_clear_error();
  S_ARG(npoly == 0 || (npoly == kNpoly && coef));
  if (!s_.have_net) return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  if (npoly) S_OK(hipMemcpy(s_.poly, coef, sizeof(float) * kNpoly * kNout, hipMemcpyHostToDevice));
  s_.npoly = npoly;
  s_.trained = false;   // the trust grid belongs to the trained network
}

static void argot_hip_surrogate_eval_body(const float* log_tau, float* log_kernel, int64_t n) {
  argot_hip_internal_clear_error();
  S_ARG(log_tau && log_kernel && n > 0);
  if (!s_.have_net) return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  float *x = nullptr, *y = nullptr;
  hipError_t he = dalloc(&x, 3 * (size_t)n);
  if (he == hipSuccess) he = dalloc(&y, kNout * (size_t)n);
  if (he == hipSuccess)
    he = hipMemcpy(x, log_tau, sizeof(float) * 3 * n, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    mlp_eval_kernel<<<(unsigned)((n + 63) / 64), 64, 0, s_.stream>>>(net_dev(), x, y, n);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(s_.stream);
  if (he == hipSuccess)
    he = hipMemcpy(log_kernel, y, sizeof(float) * kNout * n, hipMemcpyDeviceToHost);
  if (x) (void)hipFree(x);
  if (y) (void)hipFree(y);
  S_OK(he);
}

void argot_hip_surrogate_eval(const float* log_tau, float* log_kernel, int64_t n) {
  try {
    argot_hip_surrogate_eval_body(log_tau, log_kernel, n);
  } catch (const std::bad_alloc&) {
    argot_hip_internal_set_error(ARGOT_ERR_NOMEM);
  } catch (...) {
    argot_hip_internal_set_error(ARGOT_ERR_ARG);
  }
}

static void argot_hip_surrogate_gradient_body(const float* log_tau, const float* log_kernel,
                                              int64_t n, float* grad, double* loss) {
  argot_hip_internal_clear_error();
  S_ARG(log_tau && log_kernel && grad && loss && n >= 64 && n % 64 == 0);
  if (!s_.have_net) return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  const int nwave = (int)(n / 64);
  std::vector<float4> hx(n);
  std::vector<float> hy(8 * (size_t)n, 0.0f);
  for (int64_t i = 0; i < n; i++) {
    hx[i] = make_float4(log_tau[3 * i], log_tau[3 * i + 1], log_tau[3 * i + 2], 0.0f);
    for (int o = 0; o < kNout; o++) hy[8 * i + o] = log_kernel[kNout * i + o];
  }
  float4* X = nullptr;
  float *Y = nullptr, *partial = nullptr, *g = nullptr;
  double *lp = nullptr, *lo = nullptr;
  hipError_t he = dalloc(&X, n);
  if (he == hipSuccess) he = dalloc(&Y, 8 * (size_t)n);
  if (he == hipSuccess) he = dalloc(&partial, (size_t)nwave * kNparamPad);
  if (he == hipSuccess) he = dalloc(&g, kNparamPad);
  if (he == hipSuccess) he = dalloc(&lp, nwave);
  if (he == hipSuccess) he = dalloc(&lo, 1);
  if (he == hipSuccess) he = hipMemcpy(X, hx.data(), sizeof(float4) * n, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(Y, hy.data(), sizeof(float) * 8 * n, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    TrainArgs ta;
    ta.net = net_dev();
    ta.X = X;
    ta.Y = Y;
    ta.n = n;
    ta.offset = 0;
    ta.gscale = 2.0f / (float)(n * kNout);
    ta.partial = partial;
    ta.loss_part = lp;
    mlp_train_kernel<<<nwave, 64, 0, s_.stream>>>(ta);
    AdamArgs aa{};
    aa.partial = partial;
    aa.nwave = nwave;
    aa.loss_part = lp;
    aa.loss_out = lo;
    aa.inv_count = 1.0 / ((double)n * kNout);
    aa.apply = 0;
    aa.grad = g;
    mlp_adam_kernel<<<(kNparam + 255) / 256, 256, 0, s_.stream>>>(aa);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipStreamSynchronize(s_.stream);
  if (he == hipSuccess) he = hipMemcpy(grad, g, sizeof(float) * kNparam, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(loss, lo, sizeof(double), hipMemcpyDeviceToHost);
  for (void* p : {(void*)X, (void*)Y, (void*)partial, (void*)g, (void*)lp, (void*)lo})
    if (p) (void)hipFree(p);
  S_OK(he);
}

void argot_hip_surrogate_gradient(const float* log_tau, const float* log_kernel, int64_t n,
                                  float* grad, double* loss) {
  try {
    argot_hip_surrogate_gradient_body(log_tau, log_kernel, n, grad, loss);
  } catch (const std::bad_alloc&) {
    argot_hip_internal_set_error(ARGOT_ERR_NOMEM);
  } catch (...) {
    argot_hip_internal_set_error(ARGOT_ERR_ARG);
  }
}

void argot_hip_surrogate_trust_error(float* err, int64_t n) {
  argot_hip_internal_clear_error();
  S_ARG(err);
  if (!s_.trained) return argot_hip_internal_set_error(ARGOT_ERR_STATE);
  S_ARG(n == s_.ntrust);
  S_OK(hipMemcpy(err, s_.cell_err, sizeof(float) * n, hipMemcpyDeviceToHost));
}

}  // extern "C"
