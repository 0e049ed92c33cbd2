// Surrogate of the phase-relaxation Newton solve: a 3-64-64-64-3 tanh MLP
// mapping (log rho, e, alpha_L) of a cell to the equilibrium (T, asinh(p /
// 1e4), alpha_L) that warm-starts an FP64 Newton polish.  No counterpart in
// the reference (its relaxation.cu / thermRlx.cu / chemRlx.cu always start
// from the cell state).
//
// Numerics of the network (the CPU check restates the same; inference takes
// tanh as tanh_fast, within ~1e-7 of it, training tanhf):
//   layer 1: FP32 FMA, tanh, rounded to BF16
//   layers 2, 3: BF16 x BF16 MFMA (v_mfma_f32_16x16x32_bf16), FP32
//                accumulate, + bias, tanh, rounded to BF16
//   layer 4: FP32 FMA on the BF16 activations
// Operand layouts of v_mfma_f32_16x16x32_bf16 (one wave, lane l):
//   A (16x32, row m, col k):  A[l % 16][8 (l / 16) + i], i = 0..7
//   B (32x16, row k, col n):  B[8 (l / 16) + i][l % 16]
//   D (16x16):                D[4 (l / 16) + i][l % 16], i = 0..3
// Training: per-wave forward + backward over 32-sample chunks, weight
// gradients by MFMA with the sample index as K, per-wave partial sums
// reduced in a fixed order (deterministic), then Adam on FP32 master weights.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fooska.h"
#include "fk_device.h"
#include "fk_internal.h"
#include "fk_relax.h"

using namespace fkx;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int HID = 64;
constexpr int kParams = 8771;
constexpr int oW1 = 0, oB1 = 192, oW2 = 256, oB2 = 4352, oW3 = 4416, oB3 = 8512, oW4 = 8576, oB4 = 8768;
// BF16 copies (uint16 offsets): W2, W3 row-major [out][in], W2T, W3T [in][out]
constexpr int bW2 = 0, bW3 = 4096, bW2T = 8192, bW3T = 12288, kBf = 16384;
constexpr float kPScale = 1e4f;  // pressure target is asinh(p / kPScale)

struct Norm {
  float xm[3], xs[3], ym[3], ys[3];
};

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

// tanh(z) = 1 - 2 / (1 + e^(2z)) with the hardware exponential and
// reciprocal: within about 1e-7 of tanh, far below the BF16 rounding
// (2^-9 relative) of every activation.  Inference only (the relaxation's
// warm start); training and its CPU check use tanhf.
__device__ __forceinline__ float tanh_fast(float z) {
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * z));
}

__device__ __forceinline__ bf16x8 ld_bf8(const uint16_t *p) {  // 16-byte aligned
  return *reinterpret_cast<const bf16x8 *>(p);
}

// Per-wave LDS scratch of the forward pass over 16 samples.
struct FwdLds {
  uint16_t h[16 * HID];  // activations of the current layer, [m][k]
  float x[16 * 3];
  float y[16 * 3];
};

// The FP32 parameters of the forward pass, copied into LDS once per block:
// layer 1 as (W1 row k, b1[k]) quads, the biases of layers 2 and 3, W4, b4.
struct SmallW {
  float4 w1b[HID];
  float b2[HID], b3[HID];
  __attribute__((aligned(16))) float w4[3 * HID];
  float b4[4];
};

// Forward of 16 samples by one wave.  x: LDS [16][3] (normalised), result in
// lds.y [16][3] (normalised).  Weights: sw (LDS: W1, b1, b2, b3, W4, b4),
// bf16 wb in LDS (W2, W3).  Keeps the BF16 activations of each layer in hs
// (row-major [16][64]) when hs != nullptr (training).
template <bool FAST = false>
__device__ __forceinline__ void mlp_fwd16(const SmallW &sw, const uint16_t *wb, FwdLds &lds, uint16_t *h1s,
                                         uint16_t *h2s, uint16_t *h3s) {
  const int l = threadIdx.x & 63;
  const int m = l & 15, q = l >> 4;
  // layer 1 (FP32): this lane computes neurons k = 32 ks + 8 q + i of sample m
  const float x0 = lds.x[m * 3 + 0], x1 = lds.x[m * 3 + 1], x2 = lds.x[m * 3 + 2];
  bf16x8 a[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = 32 * ks + 8 * q + i;
      const float4 c = sw.w1b[k];
      const float z = c.w + c.x * x0 + c.y * x1 + c.z * x2;
      const uint16_t hb = f2bf(FAST ? tanh_fast(z) : tanhf(z));
      a[ks][i] = __builtin_bit_cast(__bf16, hb);
      if (h1s) h1s[m * HID + k] = hb;
    }
  // layers 2 and 3 (BF16 MFMA)
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const uint16_t *W = wb + (layer == 0 ? bW2 : bW3);
    const float *bias = layer == 0 ? sw.b2 : sw.b3;
    uint16_t *hs = layer == 0 ? h2s : h3s;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      const int n = 16 * nt + m;  // B column of this lane
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const bf16x8 b = ld_bf8(W + n * HID + 32 * ks + 8 * q);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 4 * q + i;
        const uint16_t hb = f2bf(FAST ? tanh_fast(acc[i] + bias[n]) : tanhf(acc[i] + bias[n]));
        lds.h[row * HID + n] = hb;
        if (hs) hs[row * HID + n] = hb;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) a[ks] = ld_bf8(lds.h + m * HID + 32 * ks + 8 * q);
    __builtin_amdgcn_wave_barrier();
  }
  // layer 4 (FP32) on the last activations: this lane's 16 neurons of sample
  // m (in a[]), then the sum over the sample's 4 lanes q = 0..3,
  // (p_0 + p_1) + (p_2 + p_3) on every one of them
  float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    // the whole vector reinterpreted (a bit_cast of one element of a[ks]
    // reads element 0)
    const u16x8 ab = __builtin_bit_cast(u16x8, a[ks]);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float h = bf2f(ab[i]);
      const int k = 32 * ks + 8 * q + i;
#pragma unroll
      for (int j = 0; j < 3; ++j) p[j] += sw.w4[j * HID + k] * h;
    }
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    p[j] += __shfl_xor(p[j], 16, 64);
    p[j] += __shfl_xor(p[j], 32, 64);
  }
  if (q == 0)
#pragma unroll
    for (int j = 0; j < 3; ++j) lds.y[m * 3 + j] = sw.b4[j] + p[j];
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
}

// Copy the FP32 forward parameters into LDS (block-wide; the caller's
// barrier, that of load_wb, publishes them).
__device__ __forceinline__ void load_small(const float *__restrict__ w, SmallW &s) {
  for (int i = threadIdx.x; i < HID; i += blockDim.x) {
    s.w1b[i] = make_float4(w[oW1 + 3 * i], w[oW1 + 3 * i + 1], w[oW1 + 3 * i + 2], w[oB1 + i]);
    s.b2[i] = w[oB2 + i];
    s.b3[i] = w[oB3 + i];
  }
  for (int i = threadIdx.x; i < 3 * HID; i += blockDim.x) s.w4[i] = w[oW4 + i];
  if (threadIdx.x < 3) s.b4[threadIdx.x] = w[oB4 + threadIdx.x];
}

// Copy the BF16 weights into LDS (block-wide).
__device__ __forceinline__ void load_wb(const uint16_t *__restrict__ g, uint16_t *s, int n) {
  for (int i = threadIdx.x * 8; i < n; i += blockDim.x * 8)
    *reinterpret_cast<uint4 *>(s + i) = *reinterpret_cast<const uint4 *>(g + i);
  __syncthreads();
}

// Inference on host-provided inputs x [n][3] (raw features) -> y [n][3] (raw
// targets).  Grid-stride over 64-sample chunks per wave.
__global__ void __launch_bounds__(256) mlp_infer_k(int64_t n, const float *__restrict__ x, float *__restrict__ y,
                                                   const float *__restrict__ w, const uint16_t *__restrict__ wbg,
                                                   Norm nm) {
  __shared__ __attribute__((aligned(16))) uint16_t wb[8192];
  __shared__ FwdLds lds[4];
  __shared__ SmallW sw;
  load_small(w, sw);
  load_wb(wbg, wb, 8192);
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t nchunk = (n + 63) / 64;
  for (int64_t ch = (int64_t)blockIdx.x * 4 + wv; ch < (nchunk + 3) / 4 * 4; ch += (int64_t)gridDim.x * 4) {
    for (int t = 0; t < 4; ++t) {
      const int64_t base = ch * 64 + t * 16;
      if (l < 48) {
        const int mm = l / 3, j = l % 3;
        const int64_t s = base + mm;
        lds[wv].x[mm * 3 + j] = s < n ? (x[s * 3 + j] - nm.xm[j]) / nm.xs[j] : 0.f;
      }
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      mlp_fwd16<true>(sw, wb, lds[wv], nullptr, nullptr, nullptr);
      if (l < 48) {
        const int mm = l / 3, j = l % 3;
        const int64_t s = base + mm;
        if (s < n) y[s * 3 + j] = lds[wv].y[mm * 3 + j] * nm.ys[j] + nm.ym[j];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// ---------------- training ----------------

struct TrainLds {
  uint16_t h1[32 * HID], h2[32 * HID], h3[32 * HID];  // [m][k]
  uint16_t dz[32 * HID];                              // [m][k] of the current layer
  uint16_t hT[HID * 32];                              // [k][m] of the layer input
  uint16_t dzT[HID * 32];                             // [n][m]
  float x[32 * 3], dy[32 * 3];
  FwdLds f;
};

// One wave per block, 32 samples idx[32 b .. 32 b + 32) of the batch; writes
// this block's share of the gradient of sum_m sum_j (y - t)^2 / (3 B) to
// part[blockIdx.x] (each weight-gradient tile as soon as it is formed).
__global__ void __launch_bounds__(64) mlp_train_k(int64_t B, const int64_t *__restrict__ idx,
                                                  const float *__restrict__ X, const float *__restrict__ Y,
                                                  const float *__restrict__ w, const uint16_t *__restrict__ wbg,
                                                  Norm nm, float *__restrict__ part, float *__restrict__ loss_part) {
  __shared__ __attribute__((aligned(16))) uint16_t wb[kBf];
  __shared__ __attribute__((aligned(16))) TrainLds L;
  __shared__ SmallW sw;
  load_small(w, sw);
  load_wb(wbg, wb, kBf);
  const int l = threadIdx.x, m16 = l & 15, q = l >> 4;
  float *g = part + (int64_t)blockIdx.x * kParams;
  const int64_t s0 = (int64_t)blockIdx.x * 32;
  const int64_t s1 = s0 + 32 < B ? s0 + 32 : B;
  const float inv = 2.0f / (3.0f * (float)B);
  float lsum = 0;
  // inputs (normalised) of 32 samples; padding samples get dy = 0
  if (l < 32) {
    const int64_t s = s0 + l;
    for (int j = 0; j < 3; ++j) L.x[l * 3 + j] = s < s1 ? (X[idx[s] * 3 + j] - nm.xm[j]) / nm.xs[j] : 0.f;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  for (int t = 0; t < 2; ++t) {
    if (l < 48) L.f.x[l] = L.x[t * 48 + l];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    mlp_fwd16(sw, wb, L.f, L.h1 + t * 16 * HID, L.h2 + t * 16 * HID, L.h3 + t * 16 * HID);
    if (l < 48) {
      const int mm = l / 3, j = l % 3;
      const int64_t s = s0 + t * 16 + mm;
      float d = 0.f;
      if (s < s1) {
        const float tgt = (Y[idx[s] * 3 + j] - nm.ym[j]) / nm.ys[j];
        d = L.f.y[mm * 3 + j] - tgt;
        lsum += d * d;
      }
      L.dy[(t * 16 + mm) * 3 + j] = d * inv;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
  // layer 4 (lane k = l): gW4[j][k], gb4[j]; dz3[m][k] = (sum_j dy W4[j][k]) (1 - h3^2)
  {
    const int k = l;
    const float w4[3] = {w[oW4 + k], w[oW4 + HID + k], w[oW4 + 2 * HID + k]};
    float gw4[3] = {0, 0, 0}, gb3 = 0;
    for (int mm = 0; mm < 32; ++mm) {
      const float h = bf2f(L.h3[mm * HID + k]);
      float dh = 0;
      for (int j = 0; j < 3; ++j) {
        const float d = L.dy[mm * 3 + j];
        gw4[j] += d * h;
        dh += d * w4[j];
      }
      const float dz = dh * (1.f - h * h);
      gb3 += dz;
      const uint16_t db = f2bf(dz);
      L.dz[mm * HID + k] = db;
      L.dzT[k * 32 + mm] = db;
      L.hT[k * 32 + mm] = L.h2[mm * HID + k];
    }
    for (int j = 0; j < 3; ++j) g[oW4 + j * HID + k] = gw4[j];
    g[oB3 + k] = gb3;
    if (l < 3) {
      float sb = 0;
      for (int mm = 0; mm < 32; ++mm) sb += L.dy[mm * 3 + l];
      g[oB4 + l] = sb;
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  // layers 3 then 2: gW = dz^T h (K = 32 samples); dh = dz W; dz_below = dh (1 - h^2)
  for (int layer = 3; layer >= 2; --layer) {
    const int oW = layer == 3 ? oW3 : oW2;
    for (int mt = 0; mt < 4; ++mt) {
      const bf16x8 a = ld_bf8(L.dzT + (16 * mt + m16) * 32 + 8 * q);
      for (int nt = 0; nt < 4; ++nt) {
        const bf16x8 b = ld_bf8(L.hT + (16 * nt + m16) * 32 + 8 * q);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        for (int i = 0; i < 4; ++i) g[oW + (16 * mt + 4 * q + i) * HID + 16 * nt + m16] = acc[i];
      }
    }
    const uint16_t *WT = wb + (layer == 3 ? bW3T : bW2T);
    const uint16_t *hin = layer == 3 ? L.h2 : L.h1;  // activations feeding this layer
    float dzn[2][4][4];
    for (int t = 0; t < 2; ++t) {
      bf16x8 a2[2];
      for (int ks = 0; ks < 2; ++ks) a2[ks] = ld_bf8(L.dz + (16 * t + m16) * HID + 32 * ks + 8 * q);
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 2; ++ks) {
          const bf16x8 b = ld_bf8(WT + (16 * nt + m16) * HID + 32 * ks + 8 * q);
          d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[ks], b, d, 0, 0, 0);
        }
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * t + 4 * q + i, col = 16 * nt + m16;
          const float h = bf2f(hin[row * HID + col]);
          dzn[t][nt][i] = d[i] * (1.f - h * h);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int t = 0; t < 2; ++t)
      for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i) {
          const int row = 16 * t + 4 * q + i, col = 16 * nt + m16;
          const uint16_t db = f2bf(dzn[t][nt][i]);
          L.dz[row * HID + col] = db;
          L.dzT[col * 32 + row] = db;
        }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    {
      const int k = l;
      float sb = 0, gw1[3] = {0, 0, 0};
      for (int mm = 0; mm < 32; ++mm) {
        const float dz = bf2f(L.dz[mm * HID + k]);
        sb += dz;
        if (layer == 3) L.hT[k * 32 + mm] = L.h1[mm * HID + k];
        else
          for (int j = 0; j < 3; ++j) gw1[j] += dz * L.x[mm * 3 + j];
      }
      if (layer == 3) g[oB2 + k] = sb;
      else {
        g[oB1 + k] = sb;
        for (int j = 0; j < 3; ++j) g[oW1 + 3 * k + j] = gw1[j];
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
  if (l == 0) loss_part[blockIdx.x] = lsum;
}

// Fixed-order sum of the per-block partials (deterministic).
__global__ void __launch_bounds__(256) mlp_reduce_k(int nb, const float *__restrict__ part, float *__restrict__ g) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kParams) return;
  float s = 0;
  for (int b = 0; b < nb; ++b) s += part[(int64_t)b * kParams + i];
  g[i] = s;
}

// Adam on the FP32 master weights, then refresh of the BF16 copies.
__global__ void __launch_bounds__(256) mlp_adam_k(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ g, float lr, float b1t, float b2t) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kParams) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1 - b1) * gi;
  const float vi = b2 * v[i] + (1 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / (1 - b1t)) / (sqrtf(vi / (1 - b2t)) + eps);
}

__global__ void __launch_bounds__(256) mlp_pack_k(const float *__restrict__ w, uint16_t *__restrict__ wb) {
  const int i = blockIdx.x * 256 + threadIdx.x;  // 0 .. 4095
  if (i >= 4096) return;
  const int r = i / HID, c = i % HID;
  const uint16_t a = f2bf(w[oW2 + i]), b = f2bf(w[oW3 + i]);
  wb[bW2 + i] = a;
  wb[bW3 + i] = b;
  wb[bW2T + c * HID + r] = a;
  wb[bW3T + c * HID + r] = b;
}

// Training target of a state: the chemical (p-T-g) equilibrium of its
// mixture density and internal energy, (T, asinh(p / 1e4), alpha_L), by the
// (ln p, T) Newton from the liquid's (T, p) to 1e-10 (fk::chem_polish), or NaN
// where that solve finds no two-phase equilibrium.  The equilibrium depends
// on (rho, e) alone, so the labels of every state, whatever its class, are
// values of the one map the network learns.  u is not modified.
__device__ inline void chem_label(const fk_params &P, const double *u, float *y) {
  double uc[10];
  for (int v = 0; v < 10; ++v) uc[v] = u[v];
  double T0, p0;
  fk::chem_cold(P, uc, T0, p0);
  if (fk::chem_polish(P, uc, T0, p0, P.polishTol, 30) > 0 && fk::mlp_targets(P, uc, y)) return;
  for (int j = 0; j < 3; ++j) y[j] = NAN;
}

// Labels host-provided states: features of U, and for the chemical-class
// states (those the relaxation hands to the network) the equilibrium target.
__global__ void __launch_bounds__(256) mlp_label_k(fk_params P, int64_t n, const double *__restrict__ U,
                                                   float *__restrict__ x, float *__restrict__ y) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  double u[10];
  for (int v = 0; v < 10; ++v) u[v] = U[c * 10 + v];
  fk::mlp_features(P, u, x + c * 3);
  int ts = 0;
  int chem = 0;
  if (P.chemFlag == 1) {
    const double avf = fk::clampd(P, u[0]);
    const double rho = u[1] / avf, enr = u[6] / (avf * rho);
    const double tL = fk::eos_T(P, 0, rho, enr), pL = fk::eos_p(P, 0, rho, enr);
    chem = tL > fk::tsat_dev(P, pL, tL, P.satTempUrf, &ts) && avf > P.chemLimit;
  }
  if (chem) chem_label(P, u, y + c * 3);
  else  // the network serves chemical cells only
    for (int j = 0; j < 3; ++j) y[c * 3 + j] = NAN;
}

// Surrogate relaxation, stage 1 (every cell, one thread each): exact
// classification by the sign of f(T_L; p_L); thermal cells try at most
// kColdIters analytic-Jacobian Newton iterations from the cell's own state
// (one hydro step away from the previous equilibrium, so most converge in 2-3)
// and are written back when converged (fk::screen_cell).  Chemical cells and
// thermal cells that did not converge go to the work list as
// (cell << 1) | chem for stage 2.  The last RK stage's fused pass F runs the
// same stage 1 on the state it produces (fk_kernels.hip).
// counts: [n_chem, newton iterations, -, fallbacks, warm-started, list
// length from the front (chemical), list length from the back (thermal),
// fallback list length].

__global__ void __launch_bounds__(256) relax_screen_k(fk_params P, int64_t nc, double *__restrict__ U,
                                                      uint32_t *__restrict__ list,
                                                      unsigned long long *__restrict__ counts,
                                                      unsigned long long *__restrict__ part, float *__restrict__ W) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int l = threadIdx.x & 63;
  int chem = 0, it = 0, push = 0;
  if (c < nc) {
    double u[10];
    fk::load_rec<10>(U + c * 10, u);
    push = fk::screen_cell(P, u, &chem, &it);
    if (!push) {
      U[c * 10 + 0] = u[0];
      U[c * 10 + 6] = u[6];
      U[c * 10 + 7] = u[7];
    }
    // primitive records of pass G for the cells solved here (list cells:
    // relax_list_k)
    if (W && !push) fk::store_prims<10>(fk::make_mat<double>(P), u, W + c * 10);
  }
  fk::list_push(push, chem, c, list, nc, counts);
  unsigned long long v[2] = {(unsigned long long)chem, (unsigned long long)it};
  fk::block_partials<2>(v, part);  // [n_chem, newton iterations]
}

// Surrogate relaxation, stage 2a (FK_RELAX_MLP): the network's equilibrium
// (T*, p*) of each chemical entry of the work list (its front),
// by wave-cooperative inference, 64 entries per wave, into start[e].  Apart
// from the polish (stage 2b) so that the registers of neither limit the
// occupancy of the other.
__global__ void __launch_bounds__(256) relax_infer_k(fk_params P, const double *__restrict__ U,
                                                     const uint32_t *__restrict__ list, const float *__restrict__ w,
                                                     const uint16_t *__restrict__ wbg, Norm nm,
                                                     const unsigned long long *__restrict__ counts,
                                                     float2 *__restrict__ start) {
  __shared__ __attribute__((aligned(16))) uint16_t wb[8192];
  __shared__ FwdLds lds[4];
  __shared__ SmallW sw;
  const int64_t nfr = (int64_t)counts[5];
  const int64_t cf = (nfr + 63) / 64;
  if ((int64_t)blockIdx.x * 4 >= cf) return;  // whole block idle: skip the weight load
  load_small(w, sw);
  load_wb(wbg, wb, 8192);
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t per = (cf + 3) / 4 * 4;
  for (int64_t ch = (int64_t)blockIdx.x * 4 + wv; ch < per; ch += (int64_t)gridDim.x * 4) {
    const int64_t e = ch * 64 + l;
    const bool live = e < nfr;
    const int64_t c = live ? (int64_t)(list[e] >> 1) : 0;
    double u[10];
    for (int v = 0; v < 10; ++v) u[v] = live ? U[c * 10 + v] : (v == 1 || v == 2 ? 1.0 : 0.0);
    float xf[3];
    fk::mlp_features(P, u, xf);
    float yw[3];
    for (int t = 0; t < 4; ++t) {  // wave-uniform: every lane joins
      if ((l >> 4) == t)
        for (int j = 0; j < 3; ++j) lds[wv].x[(l & 15) * 3 + j] = (xf[j] - nm.xm[j]) / nm.xs[j];
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      mlp_fwd16<true>(sw, wb, lds[wv], nullptr, nullptr, nullptr);
      if ((l >> 4) == t)
        for (int j = 0; j < 3; ++j) yw[j] = lds[wv].y[(l & 15) * 3 + j] * nm.ys[j] + nm.ym[j];
      __builtin_amdgcn_wave_barrier();
    }
    // (T*, p*); the alpha_L* output is not needed: the polish takes alpha_L
    // from the mass balance at (p, T)
    if (live) start[e] = make_float2(yw[0], (float)((double)kPScale * sinh((double)yw[1])));
  }
}

// Surrogate relaxation, stage 2b (the work list, 64 entries per wave): FP64
// Newton polish from the network's start (WARM, chemical entries) or from the
// cell state (thermal entries, and FK_RELAX_COLD); a cell whose polish fails
// goes to the fallback list.
template <bool WARM>
__global__ void __launch_bounds__(256) relax_list_k(fk_params P, double *__restrict__ U,
                                                    const uint32_t *__restrict__ list,
                                                    const float2 *__restrict__ start,
                                                    unsigned long long *__restrict__ counts, float *__restrict__ W,
                                                    int64_t cap, uint32_t *__restrict__ fall) {
  // chunks of 64 entries: the chemical ones from the list's front, then the
  // thermal ones from its back
  const int64_t nfr = (int64_t)counts[5], nbk = (int64_t)counts[6];
  const int64_t cf = (nfr + 63) / 64, nchunk = cf + (nbk + 63) / 64;
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  unsigned long long nit = 0, nwarm = 0;
  for (int64_t ch = (int64_t)blockIdx.x * 4 + wv; ch < nchunk; ch += (int64_t)gridDim.x * 4) {
    const bool front = ch < cf;  // wave-uniform
    const int64_t e = (front ? ch : ch - cf) * 64 + l;
    const bool live = e < (front ? nfr : nbk);
    const uint32_t ent = live ? list[front ? e : cap - 1 - e] : 0u;
    const int64_t c = (int64_t)(ent >> 1);
    const int chem = (int)(ent & 1u);
    double u[10];
    for (int v = 0; v < 10; ++v) u[v] = live ? U[c * 10 + v] : 0.0;
    double T0 = 0, p0 = 0;  // chemical cells: start (T, p)
    if (WARM && front && live) {
      const float2 y = start[e];
      T0 = (double)y.x;
      p0 = (double)y.y;
    }
    if (!live) continue;
    if (chem && !WARM) fk::chem_cold(P, u, T0, p0);
    nwarm += (WARM && chem) ? 1 : 0;
    int it;
    if (chem) {
      it = fk::chem_polish(P, u, T0, p0, P.polishTol, 30);
    } else {  // thermal cells: the cell state
      fk::ThermProb t;
      fk::therm_setup(P, u, t);
      double rk, ek;
      fk::therm_cold(P, u, t, rk, ek);
      it = fk::therm_polish(P, u, rk, ek, P.polishTol, 30);
    }
    // a failed polish leaves U as it was: the cell goes to the fallback list
    // (relax_fallback_k runs the reference solve from the cell state)
    const bool failed = it < 0;
    const unsigned long long mf = __ballot(failed);
    if (mf) {
      const int leader = __ffsll((long long)mf) - 1;
      unsigned long long b = 0;
      if (l == leader) b = atomicAdd(&counts[7], (unsigned long long)__popcll(mf));
      b = __shfl(b, leader, 64);
      if (failed) fall[b + __popcll(mf & ((1ull << l) - 1ull))] = ent;
    }
    if (failed) continue;
    nit += it;
    U[c * 10 + 0] = u[0];
    U[c * 10 + 1] = u[1];
    U[c * 10 + 2] = u[2];
    U[c * 10 + 6] = u[6];
    U[c * 10 + 7] = u[7];
    U[c * 10 + 9] = u[9];
    if (W) fk::store_prims<10>(fk::make_mat<double>(P), u, W + c * 10);
  }
  for (int off = 32; off > 0; off >>= 1) {
    nit += __shfl_down(nit, off, 64);
    nwarm += __shfl_down(nwarm, off, 64);
  }
  if (l == 0) {
    atomicAdd(&counts[1], nit);
    atomicAdd(&counts[4], nwarm);
  }
}

// Fallback of the surrogate relaxation: the reference solve (thermRlx /
// chemRlx, as the plain relaxation) from the cell state for the cells whose
// polish failed (fall[0 .. counts[7]), entries (cell << 1) | chem).  Apart
// from relax_list_k so that the reference solves' registers do not limit the
// occupancy of the polish.  A fallback counts 30 iterations plus those of the
// reference solve.
__global__ void __launch_bounds__(256) relax_fallback_k(fk_params P, double *__restrict__ U,
                                                        const uint32_t *__restrict__ fall,
                                                        unsigned long long *__restrict__ counts,
                                                        float *__restrict__ W) {
  const int64_t n = (int64_t)counts[7];
  unsigned long long nit = 0, nfb = 0;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const uint32_t ent = fall[e];
    const int64_t c = (int64_t)(ent >> 1);
    double u[10];
    for (int v = 0; v < 10; ++v) u[v] = U[c * 10 + v];
    const int it = (ent & 1u) ? fk::chem_dev(P, u) : fk::therm_dev(P, u);
    nit += 30 + (unsigned long long)it;
    ++nfb;
    for (int v = 0; v < 10; ++v) U[c * 10 + v] = u[v];
    if (W) fk::store_prims<10>(fk::make_mat<double>(P), u, W + c * 10);
  }
  const int l = threadIdx.x & 63;
  for (int off = 32; off > 0; off >>= 1) {
    nit += __shfl_down(nit, off, 64);
    nfb += __shfl_down(nfb, off, 64);
  }
  if (l == 0 && nfb) {
    atomicAdd(&counts[1], nit);
    atomicAdd(&counts[3], nfb);
  }
}

// Emission of solver states: plain relaxation that also writes (features,
// equilibrium target) of the pre-relaxation states of cells c = phase +
// k stride, whatever their class, into slot k of the training set.
__global__ void __launch_bounds__(256) relax_emit_k(fk_params P, int64_t nc, double *__restrict__ U, int64_t stride,
                                                    int64_t phase, float *__restrict__ dx, float *__restrict__ dy,
                                                    int64_t cap) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  double u[10];
  for (int v = 0; v < 10; ++v) u[v] = U[c * 10 + v];
  const bool emit = stride > 0 && (c % stride) == phase;
  float xf[3], yl[3];
  if (emit) {
    fk::mlp_features(P, u, xf);
    chem_label(P, u, yl);  // from the pre-relaxation state, whatever the cell's class
  }
  int chem = 0, ts = 0;
  if (P.chemFlag == 1) {
    const double avf = fk::clampd(P, u[0]);
    const double rho = u[1] / avf, enr = u[6] / (avf * rho);
    const double tL = fk::eos_T(P, 0, rho, enr), pL = fk::eos_p(P, 0, rho, enr);
    chem = tL > fk::tsat_dev(P, pL, tL, P.satTempUrf, &ts) && avf > P.chemLimit;
  }
  if (chem) fk::chem_dev(P, u);
  else fk::therm_dev(P, u);
  for (int v = 0; v < 10; ++v) U[c * 10 + v] = u[v];
  if (emit) {
    const int64_t slot = (c - phase) / stride;
    if (slot < cap) {
      for (int j = 0; j < 3; ++j) {
        dx[slot * 3 + j] = xf[j];
        dy[slot * 3 + j] = yl[j];
      }
    }
  }
}

__global__ void zero_u64_k(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}



// ---------------- host side ----------------

Norm norm_of(const Mlp &m) {
  Norm n;
  for (int j = 0; j < 3; ++j) {
    n.xm[j] = m.norm[j];
    n.xs[j] = m.norm[3 + j];
    n.ym[j] = m.norm[6 + j];
    n.ys[j] = m.norm[9 + j];
  }
  return n;
}

int ensure_mlp(fk_ctx *c) {
  Mlp &m = c->mlp;
  if (m.w) return FK_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMalloc(&m.w, kParams * 4));
  HIPCHK(hipMalloc(&m.m, kParams * 4));
  HIPCHK(hipMalloc(&m.v, kParams * 4));
  HIPCHK(hipMalloc(&m.grad, kParams * 4));
  HIPCHK(hipMalloc(&m.wb, kBf * 2));
  HIPCHK(hipMemset(m.w, 0, kParams * 4));
  HIPCHK(hipMemset(m.m, 0, kParams * 4));
  HIPCHK(hipMemset(m.v, 0, kParams * 4));
  HIPCHK(hipMemset(m.wb, 0, kBf * 2));
  return FK_OK;
}

int pack(fk_ctx *c) {
  hipLaunchKernelGGL(mlp_pack_k, dim3(16), dim3(256), 0, c->stream, c->mlp.w, c->mlp.wb);
  HIPCHK(hipGetLastError());
  return FK_OK;
}

int ensure_data(fk_ctx *c, int64_t cap) {
  Mlp &m = c->mlp;
  if (cap <= m.cap_data) return FK_OK;
  float *nx = nullptr, *ny = nullptr;
  HIPCHK(hipMalloc(&nx, (size_t)cap * 12));
  HIPCHK(hipMalloc(&ny, (size_t)cap * 12));
  if (m.n_data) {
    HIPCHK(hipMemcpyAsync(nx, m.dx, (size_t)m.n_data * 12, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(ny, m.dy, (size_t)m.n_data * 12, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (m.dx) (void)hipFree(m.dx);
  if (m.dy) (void)hipFree(m.dy);
  m.dx = nx;
  m.dy = ny;
  m.cap_data = cap;
  return FK_OK;
}

}  // namespace

// called from the relaxation entry point of fk_kernels.hip; counts layout as
// relax_screen_k / relax_list_k (slots [0, 8) of c->lcount, zeroed here).
// Checks, scratch and zeroed counts of the surrogate relaxation (before stage
// 1, which relax_screen_k or the last stage's fused pass F runs).
int fk_relax_mlp_begin(fk_ctx *c, bool warm) {
  Mlp &m = c->mlp;
  if (warm && !m.ready) return fail(FK_ESTATE, "surrogate relaxation needs trained or loaded MLP weights");
  // the exact classification needs f(T) increasing: B <= 0 and C >= 0
  const fk_params &P = c->p;
  if (P.chemFlag == 1 && !((P.eta[0] - P.eta[1]) <= 0 && (P.Cp[1] - P.Cp[0]) >= 0))
    return fail(FK_EINVAL, "surrogate classification needs eta_L <= eta_G and Cp_G >= Cp_L");
  int rc = ensure_rlist(c);
  if (rc) return rc;
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  LAUNCH_TIMED(c, "zero", hipLaunchKernelGGL(zero_u64_k, dim3(1), dim3(64), 0, c->stream, cnt, 8));
  return FK_OK;
}

int fk_relax_mlp_impl(fk_ctx *c, bool warm) {
  int rc = fk_relax_mlp_begin(c, warm);
  if (rc) return rc;
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  const unsigned gs = (unsigned)((c->nc + 255) / 256);
  // mixed precision: the relaxation also writes the primitive records of pass G
  float *W = (c->p.precision == FK_MIXED && c->prims_w) ? c->wprim : nullptr;
  LAUNCH_TIMED(c, "relax_screen", hipLaunchKernelGGL(relax_screen_k, dim3(gs), dim3(256), 0, c->stream, c->p, c->nc,
                                                     c->U[c->ia], c->rlist, cnt, c->rpart, W));
  LAUNCH_TIMED(c, "relax_count", hipLaunchKernelGGL(fk::reduce_partials<2>, dim3(fk::reduce_grid(gs)), dim3(256), 0,
                                                    c->stream, c->rpart, (int64_t)gs, cnt));
  return fk_relax_mlp_finish(c, warm);
}

// Stage 2 (the work list) after stage 1.
int fk_relax_mlp_finish(fk_ctx *c, bool warm) {
  Mlp &m = c->mlp;
  unsigned long long *cnt = (unsigned long long *)c->lcount;
  float *W = (c->p.precision == FK_MIXED && c->prims_w) ? c->wprim : nullptr;
  const int64_t nchunk = (c->nc + 63) / 64;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nchunk + 3) / 4, 1024));
  if (warm) {
    int rc = ensure_rstart(c);  // fk_run_steps has allocated it before its first stage
    if (rc) return rc;
    float2 *st = static_cast<float2 *>(c->rstart);
    LAUNCH_TIMED(c, "relax_infer", hipLaunchKernelGGL(relax_infer_k, dim3(grid), dim3(256), 0, c->stream, c->p,
                                                      c->U[c->ia], c->rlist, m.w, m.wb, norm_of(m), cnt, st));
    LAUNCH_TIMED(c, "relax_list", hipLaunchKernelGGL(relax_list_k<true>, dim3(grid), dim3(256), 0, c->stream, c->p,
                                                     c->U[c->ia], c->rlist, st, cnt, W, c->nc, c->rfall));
  } else {
    LAUNCH_TIMED(c, "relax_list", hipLaunchKernelGGL(relax_list_k<false>, dim3(grid), dim3(256), 0, c->stream, c->p,
                                                     c->U[c->ia], c->rlist, nullptr, cnt, W, c->nc, c->rfall));
  }
  const unsigned gfb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((c->nc + 255) / 256, 256));
  LAUNCH_TIMED(c, "relax_fallback", hipLaunchKernelGGL(relax_fallback_k, dim3(gfb), dim3(256), 0, c->stream, c->p,
                                                       c->U[c->ia], c->rfall, cnt, W));
  c->w_valid = W != nullptr;
  return FK_OK;
}

// plain relaxation with emission of training pairs (when enabled)
int fk_relax_emit_impl(fk_ctx *c, bool *done) {
  *done = false;
  Mlp &m = c->mlp;
  if (m.emit_stride <= 0) return FK_OK;
  const int64_t room = m.cap_data - m.n_data;
  if (room <= 0) return FK_OK;
  LAUNCH_TIMED(c, "relax_emit", hipLaunchKernelGGL(relax_emit_k, dim3((unsigned)((c->nc + 255) / 256)), dim3(256), 0,
                                                   c->stream, c->p, c->nc, c->U[c->ia], m.emit_stride, m.emit_phase,
                                                   m.dx + m.n_data * 3, m.dy + m.n_data * 3, room));
  const int64_t emitted = m.emit_phase < c->nc ? (c->nc - m.emit_phase + m.emit_stride - 1) / m.emit_stride : 0;
  m.n_data += std::min<int64_t>(emitted, room);
  m.emit_phase = (m.emit_phase + 1) % m.emit_stride;
  *done = true;
  return FK_OK;
}

void fk_mlp_free(fk_ctx *c) {
  Mlp &m = c->mlp;
  void *ptrs[] = {m.w, m.m, m.v, m.wb, m.grad, m.part, m.dx, m.dy};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  m = Mlp{};
}

extern "C" {

int fk_mlp_set_weights(fk_ctx *c, const float *w, const float *norm) {
  if (!c || !w || !norm) return fail(FK_EINVAL, "null argument");
  int rc = ensure_mlp(c);
  if (rc) return rc;
  HIPCHK(hipMemcpy(c->mlp.w, w, kParams * 4, hipMemcpyHostToDevice));
  std::memcpy(c->mlp.norm, norm, sizeof c->mlp.norm);
  for (int j = 0; j < 3; ++j)
    if (!(norm[3 + j] > 0) || !(norm[9 + j] > 0)) return fail(FK_EINVAL, "normalisation scales must be positive");
  rc = pack(c);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->mlp.ready = true;
  c->mlp.adam_t = 0;
  HIPCHK(hipMemset(c->mlp.m, 0, kParams * 4));
  HIPCHK(hipMemset(c->mlp.v, 0, kParams * 4));
  return FK_OK;
}

int fk_mlp_get_weights(fk_ctx *c, float *w, float *norm) {
  if (!c || !w || !norm) return fail(FK_EINVAL, "null argument");
  if (!c->mlp.w) return fail(FK_ESTATE, "no MLP weights");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(w, c->mlp.w, kParams * 4, hipMemcpyDeviceToHost));
  std::memcpy(norm, c->mlp.norm, sizeof c->mlp.norm);
  return FK_OK;
}

int fk_mlp_label(fk_ctx *c, const double *U, int64_t n, float *x, float *y) {
  if (!c || !U || !x || !y || n < 0) return fail(FK_EINVAL, "bad argument");
  if (c->p.model != FK_MODEL_TWO_PHASE) return fail(FK_EINVAL, "labels need the two-phase model");
  if (n == 0) return FK_OK;
  HIPCHK(hipSetDevice(c->device));
  double *dU = nullptr;
  float *dx = nullptr, *dy = nullptr;
  HIPCHK(hipMalloc(&dU, (size_t)n * 80));
  HIPCHK(hipMalloc(&dx, (size_t)n * 12));
  HIPCHK(hipMalloc(&dy, (size_t)n * 12));
  HIPCHK(hipMemcpy(dU, U, (size_t)n * 80, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_label_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->p, n, dU, dx, dy);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(x, dx, (size_t)n * 12, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(y, dy, (size_t)n * 12, hipMemcpyDeviceToHost));
  (void)hipFree(dU);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return FK_OK;
}

int fk_mlp_data(fk_ctx *c, const float *x, const float *y, int64_t n, int append) {
  if (!c || n < 0 || (n > 0 && (!x || !y))) return fail(FK_EINVAL, "bad argument");
  int rc = ensure_mlp(c);
  if (rc) return rc;
  Mlp &m = c->mlp;
  if (!append) m.n_data = 0;
  rc = ensure_data(c, m.n_data + n);
  if (rc) return rc;
  if (n) {
    HIPCHK(hipMemcpy(m.dx + m.n_data * 3, x, (size_t)n * 12, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.dy + m.n_data * 3, y, (size_t)n * 12, hipMemcpyHostToDevice));
  }
  m.n_data += n;
  return FK_OK;
}

int fk_mlp_emit(fk_ctx *c, int64_t stride, int64_t capacity) {
  if (!c || stride < 0 || capacity < 0) return fail(FK_EINVAL, "bad argument");
  int rc = ensure_mlp(c);
  if (rc) return rc;
  c->mlp.emit_stride = stride;
  c->mlp.emit_phase = 0;
  return ensure_data(c, c->mlp.n_data + capacity);
}

int fk_mlp_download_data(fk_ctx *c, float *x, float *y, int64_t max, int64_t *n) {
  if (!c || !n) return fail(FK_EINVAL, "bad argument");
  const int64_t k = std::min<int64_t>(max, c->mlp.n_data);
  *n = c->mlp.n_data;
  if (k > 0) {
    if (!x || !y) return fail(FK_EINVAL, "null buffers");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(x, c->mlp.dx, (size_t)k * 12, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(y, c->mlp.dy, (size_t)k * 12, hipMemcpyDeviceToHost));
  }
  return FK_OK;
}

int fk_mlp_set_norm(fk_ctx *c, const float *norm) {
  if (!c || !norm) return fail(FK_EINVAL, "null argument");
  for (int j = 0; j < 3; ++j)
    if (!(norm[3 + j] > 0) || !(norm[9 + j] > 0)) return fail(FK_EINVAL, "normalisation scales must be positive");
  std::memcpy(c->mlp.norm, norm, sizeof c->mlp.norm);
  return FK_OK;
}

// Gradient of the mean squared error (normalised outputs) over the given
// samples (indices into the device training set), without an update.
static int grad_on(fk_ctx *c, const int64_t *didx, int64_t B, float *loss) {
  Mlp &m = c->mlp;
  const int nb = (int)((B + 31) / 32);
  if (nb > m.part_cap) {
    if (m.part) (void)hipFree(m.part);
    HIPCHK(hipMalloc(&m.part, (size_t)(nb + 1) * kParams * 4));
    m.part_cap = nb;
  }
  float *lossp = m.part + (size_t)nb * kParams;
  hipLaunchKernelGGL(mlp_train_k, dim3(nb), dim3(64), 0, c->stream, B, didx, m.dx, m.dy, m.w, m.wb, norm_of(m), m.part,
                     lossp);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(mlp_reduce_k, dim3((kParams + 255) / 256), dim3(256), 0, c->stream, nb, m.part, m.grad);
  HIPCHK(hipGetLastError());
  if (loss) {
    std::vector<float> lp((size_t)nb);
    HIPCHK(hipMemcpyAsync(lp.data(), lossp, (size_t)nb * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0;
    for (float v : lp) s += v;
    *loss = (float)(s / (3.0 * (double)B));
  }
  return FK_OK;
}

int fk_mlp_grad(fk_ctx *c, const int64_t *idx, int64_t B, float *grad, float *loss) {
  if (!c || !idx || B <= 0 || !grad) return fail(FK_EINVAL, "bad argument");
  Mlp &m = c->mlp;
  if (!m.w || m.n_data <= 0) return fail(FK_ESTATE, "no weights or no training set");
  for (int64_t i = 0; i < B; ++i)
    if (idx[i] < 0 || idx[i] >= m.n_data) return fail(FK_EINVAL, "sample index out of range");
  int64_t *didx = nullptr;
  HIPCHK(hipMalloc(&didx, (size_t)B * 8));
  HIPCHK(hipMemcpy(didx, idx, (size_t)B * 8, hipMemcpyHostToDevice));
  int rc = grad_on(c, didx, B, loss);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(didx);
  if (rc) return rc;
  HIPCHK(hipMemcpy(grad, m.grad, kParams * 4, hipMemcpyDeviceToHost));
  return FK_OK;
}

int fk_mlp_train(fk_ctx *c, int steps, int batch, float lr, uint64_t seed, float *loss_first, float *loss_last) {
  if (!c || steps < 0 || batch <= 0) return fail(FK_EINVAL, "bad argument");
  Mlp &m = c->mlp;
  if (!m.w) return fail(FK_ESTATE, "set initial weights first (fk_mlp_set_weights)");
  if (m.n_data <= 0) return fail(FK_ESTATE, "empty training set");
  HIPCHK(hipSetDevice(c->device));
  // batch indices drawn on the host (splitmix64), uploaded per step
  std::vector<int64_t> idx((size_t)batch);
  int64_t *didx = nullptr;
  HIPCHK(hipMalloc(&didx, (size_t)batch * 8 * 2));
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  };
  int rc = FK_OK;
  for (int st = 0; st < steps && rc == FK_OK; ++st) {
    for (int i = 0; i < batch; ++i) idx[(size_t)i] = (int64_t)(next() % (uint64_t)m.n_data);
    int64_t *cur = didx + (st & 1) * batch;
    HIPCHK(hipMemcpyAsync(cur, idx.data(), (size_t)batch * 8, hipMemcpyHostToDevice, c->stream));
    const bool want = (st == 0 && loss_first) || (st == steps - 1 && loss_last);
    float l = 0;
    rc = grad_on(c, cur, batch, want ? &l : nullptr);
    if (rc) break;
    if (st == 0 && loss_first) *loss_first = l;
    if (st == steps - 1 && loss_last) *loss_last = l;
    ++m.adam_t;
    const float b1t = std::pow(0.9f, (float)m.adam_t), b2t = std::pow(0.999f, (float)m.adam_t);
    hipLaunchKernelGGL(mlp_adam_k, dim3((kParams + 255) / 256), dim3(256), 0, c->stream, m.w, m.m, m.v, m.grad, lr, b1t,
                       b2t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(FK_EDEVICE, hipGetErrorString(e));
    if (!rc) rc = pack(c);
    // the host index buffer is reused: wait for the copy of this step
    if (!rc) {
      e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = fail(FK_EDEVICE, hipGetErrorString(e));
    }
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(didx);
  if (rc == FK_OK) m.ready = true;
  return rc;
}

int fk_mlp_infer(fk_ctx *c, const float *x, int64_t n, float *y) {
  if (!c || n < 0 || (n > 0 && (!x || !y))) return fail(FK_EINVAL, "bad argument");
  if (!c->mlp.w) return fail(FK_ESTATE, "no MLP weights");
  if (n == 0) return FK_OK;
  HIPCHK(hipSetDevice(c->device));
  float *dx = nullptr, *dy = nullptr;
  HIPCHK(hipMalloc(&dx, (size_t)n * 12));
  HIPCHK(hipMalloc(&dy, (size_t)n * 12));
  HIPCHK(hipMemcpy(dx, x, (size_t)n * 12, hipMemcpyHostToDevice));
  const int64_t nchunk = (n + 63) / 64;
  const unsigned grid = (unsigned)std::min<int64_t>((nchunk + 3) / 4, 4096);
  hipLaunchKernelGGL(mlp_infer_k, dim3(grid), dim3(256), 0, c->stream, n, dx, dy, c->mlp.w, c->mlp.wb, norm_of(c->mlp));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(y, dy, (size_t)n * 12, hipMemcpyDeviceToHost));
  (void)hipFree(dx);
  (void)hipFree(dy);
  return FK_OK;
}

}  // extern "C"
