// Small fully-connected ReLU network (h*kappa regressor) trained and run on
// the GPU with hand-written BF16 MFMA kernels (FP32 accumulation).
//
// The reference runs a comparable network for inference only, on the CPU
// with OpenBLAS sgemm (src/my_p4est_curvature_ml.cpp:214-286), and ships no
// training code; this is the MI355X-native counterpart with training.
//
// Layer l: Z_l = A_{l-1} W_l + b_l, A_l = relu(Z_l) (last layer: no ReLU);
// A_0 = bf16((x - mean)/std).  Weights are FP32 masters [in][out] with BF16
// copies for the GEMMs; activations are stored in BF16; every GEMM
// accumulates in FP32.  Loss: mean squared error.
//
// GEMM: one wavefront computes a 32x32 tile with
// v_mfma_f32_32x32x16_bf16 (A lanes hold rows i = lane%32, k = 8*(lane/32)+v;
// B lanes hold columns j = lane%32, same k; the 16 accumulators of a lane are
// rows 8*(r/4) + 4*(lane/32) + r%4 of column lane%32); a 256-thread block
// covers 64x64.  Every operand load is bounds checked, so any M, N, K work.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "curv_device.hpp"
#include "reinit_device.hpp"

using namespace casl;
using casl::dev::curv_hybrid2;
using casl::dev::curv_hybrid3;
using casl::dev::curv_normalize2;
using casl::dev::curv_normalize3;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return 0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// The GEMM epilogue of one 32x32 wave tile at (tm, tn): bias, ReLU or the
// ReLU'-mask, FP32 and/or BF16 stores.
__device__ __forceinline__ void gemm_store(const f32x16 &acc, int M, int N, int tm, int tn, int lane,
                                           const float *__restrict__ bias, const uint16_t *__restrict__ mask, int ldm,
                                           int relu, float *__restrict__ C, uint16_t *__restrict__ C16, int ldc) {
  const int li = lane & 31;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = tm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int col = tn + li;
    if (row < M && col < N) {
      float v = acc[r];
      if (bias) v += bias[col];
      if (relu) v = v > 0.f ? v : 0.f;
      if (mask && !(bf2f(mask[size_t(row) * ldm + col]) > 0.f)) v = 0.f;
      if (C) C[size_t(row) * ldc + col] = v;
      if (C16) C16[size_t(row) * ldc + col] = f2bf(v);
    }
  }
}

// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[N]) then ReLU or ReLU'-mask.
//   op(A)(i, k) = TA ? A[k*lda + i] : A[i*lda + k]
//   op(B)(k, j) = TB ? B[j*ldb + k] : B[k*ldb + j]
//   mask: multiply by (mask[i*ldm + j] > 0), the ReLU derivative of a stored activation
// Split K: block z covers k in [z*kc, (z+1)*kc) and writes its FP32 partial
// to C32 + z*M*ldc (k_sum_splits adds the partials in a fixed order).
template <bool TA, bool TB>
__global__ void __launch_bounds__(256)
    k_gemm(int M, int N, int K, int kc, const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B,
           int ldb, const float *__restrict__ bias, const uint16_t *__restrict__ mask, int ldm, int relu,
           float *__restrict__ C32, uint16_t *__restrict__ C16, int ldc) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tm = blockIdx.y * 64 + (wave >> 1) * 32;
  const int tn = blockIdx.x * 64 + (wave & 1) * 32;
  const int li = lane & 31, lk = (lane >> 5) * 8;
  const int i = tm + li, j = tn + li;
  const int kb = int(blockIdx.z) * kc, ke = min(K, kb + kc);
  // rows of op(A) / columns of op(B) that are contiguous in k can be read 16 bytes at a time
  const bool avec = !TA && (lda & 7) == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
  const bool bvec = TB && (ldb & 7) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int k0 = kb; k0 < ke; k0 += 16) {
    const int kk = k0 + lk;
    union {
      uint4 q;
      uint16_t h[8];
    } ua, ub;
    if (avec && i < M && kk + 8 <= ke) {
      ua.q = *reinterpret_cast<const uint4 *>(A + size_t(i) * lda + kk);
    } else {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int k = kk + v;
        ua.h[v] = (i < M && k < ke) ? (TA ? A[size_t(k) * lda + i] : A[size_t(i) * lda + k]) : uint16_t(0);
      }
    }
    if (bvec && j < N && kk + 8 <= ke) {
      ub.q = *reinterpret_cast<const uint4 *>(B + size_t(j) * ldb + kk);
    } else {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int k = kk + v;
        ub.h[v] = (j < N && k < ke) ? (TB ? B[size_t(j) * ldb + k] : B[size_t(k) * ldb + j]) : uint16_t(0);
      }
    }
    bf16x8 a, b;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      a[v] = __builtin_bit_cast(__bf16, ua.h[v]);
      b[v] = __builtin_bit_cast(__bf16, ub.h[v]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  float *C = C32 ? C32 + size_t(blockIdx.z) * M * ldc : nullptr;
  gemm_store(acc, M, N, tm, tn, lane, bias, mask, ldm, relu, C, C16, ldc);
}

// k_gemm with LDS-staged panels, for M and N multiples of 64 and K ranges
// (K and kc) multiples of 64 with 16-byte aligned operands: per K chunk of
// 64 the block's op(A) panel [64 x 64] and op(B) panel [64 x 64] are loaded
// with 16-byte vectors, all of a thread's loads in flight together, into
// LDS (k contiguous per row of op(A) and per column of op(B); transposed
// layouts are transposed on the way in), double-buffered so the next
// chunk's loads overlap this chunk's MFMAs.  Each wave then reads its
// operands from LDS.  The MFMA sequence (k ascending by 16) and the operand
// layout are k_gemm's, so the results are the same bits.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256)
    k_gemm_lds(int M, int N, int K, int kc, const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ B,
               int ldb, const float *__restrict__ bias, const uint16_t *__restrict__ mask, int ldm, int relu,
               float *__restrict__ C32, uint16_t *__restrict__ C16, int ldc) {
  constexpr int KC = 64, LS = KC + 8;  // chunk, LDS row stride in bf16 (144 bytes: 16-byte aligned rows)
  __shared__ __align__(16) uint16_t sa[2][64 * LS], sb[2][64 * LS];
  const int tid = int(threadIdx.x), wave = tid >> 6, lane = tid & 63;
  const int bm = int(blockIdx.y) * 64, bn = int(blockIdx.x) * 64;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 31, lk = (lane >> 5) * 8;
  const int kb = int(blockIdx.z) * kc, ke = min(K, kb + kc);
  const int nch = (ke - kb) / KC;
  // vector v of a panel (2 per thread): row v / 8 (of op(A), of op(B)^T, or
  // of the stored k-major matrix when transposed), 8 elements at (v % 8) * 8
  uint4 ra[2], rb[2];
  auto gload = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int v = tid + 256 * h, r = v >> 3, c = (v & 7) * 8;
      ra[h] = TA ? *reinterpret_cast<const uint4 *>(A + size_t(k0 + r) * lda + bm + c)
                 : *reinterpret_cast<const uint4 *>(A + size_t(bm + r) * lda + k0 + c);
      rb[h] = TB ? *reinterpret_cast<const uint4 *>(B + size_t(bn + r) * ldb + k0 + c)
                 : *reinterpret_cast<const uint4 *>(B + size_t(k0 + r) * ldb + bn + c);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int v = tid + 256 * h, r = v >> 3, c = (v & 7) * 8;
      if (TA) {
        const uint16_t *e = reinterpret_cast<const uint16_t *>(&ra[h]);
#pragma unroll
        for (int t = 0; t < 8; ++t) sa[buf][(c + t) * LS + r] = e[t];
      } else {
        *reinterpret_cast<uint4 *>(&sa[buf][r * LS + c]) = ra[h];
      }
      if (TB) {
        *reinterpret_cast<uint4 *>(&sb[buf][r * LS + c]) = rb[h];
      } else {
        const uint16_t *e = reinterpret_cast<const uint16_t *>(&rb[h]);
#pragma unroll
        for (int t = 0; t < 8; ++t) sb[buf][(c + t) * LS + r] = e[t];
      }
    }
  };
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  if (nch > 0) {
    gload(kb);
    lstore(0);
  }
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nch) gload(kb + (ch + 1) * KC);
#pragma unroll
    for (int ks = 0; ks < KC / 16; ++ks) {
      const uint4 ua = *reinterpret_cast<const uint4 *>(&sa[buf][(wm + li) * LS + ks * 16 + lk]);
      const uint4 ub = *reinterpret_cast<const uint4 *>(&sb[buf][(wn + li) * LS + ks * 16 + lk]);
      const bf16x8 a = __builtin_bit_cast(bf16x8, ua), b = __builtin_bit_cast(bf16x8, ub);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    if (ch + 1 < nch) lstore(buf ^ 1);
    __syncthreads();
  }
  float *C = C32 ? C32 + size_t(blockIdx.z) * M * ldc : nullptr;
  gemm_store(acc, M, N, bm + wm, bn + wn, lane, bias, mask, ldm, relu, C, C16, ldc);
}

// out[e] = sum_z part[z*n + e], z = 0 .. splits-1 in order; eight loads are
// issued ahead of their additions, so the chain waits on memory once per
// eight splits (64-thread blocks spread the elements over every CU)
__global__ void k_sum_splits(const float *__restrict__ part, int splits, int64_t n, float *__restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = part[e];
  int z = 1;
  for (; z + 8 <= splits; z += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[size_t(z + k) * n + e];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; z < splits; ++z) s += part[size_t(z) * n + e];
  out[e] = s;
}

// A_0 = bf16((x - mean) / std)
__global__ void k_pack_input(const float *__restrict__ x, const float *__restrict__ mean,
                             const float *__restrict__ inv_std, int n, int d, uint16_t *__restrict__ out) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= int64_t(n) * d) return;
  const int c = int(t % d);
  out[t] = f2bf((x[t] - mean[c]) * inv_std[c]);
}

// dY = 2 (y - t) / n (MSE gradient), as FP32 and BF16; loss accumulates sum (y - t)^2
__global__ void k_mse_seed(const float *__restrict__ y, const float *__restrict__ t, int n, float scale,
                           float *__restrict__ dy32, uint16_t *__restrict__ dy16, float *__restrict__ loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (i < n) {
    e = y[i] - t[i];
    const float g = 2.f * e * scale;
    dy32[i] = g;
    dy16[i] = f2bf(g);
  }
  float s = e * e;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, s);
}

// Bias gradients, out[c] = sum over rows of x[r * m + c], in a fixed order.
// k_colsum_part: block b sums rows [b * COLSUM_ROWS, ...) per column with
// coalesced row reads (consecutive threads take consecutive columns; 256 / m
// row lanes when m < 256, added in lane order) into part[b * m + c].
// k_colsum_fin: one block per column; thread t adds partials t, t + 256, ...
// in ascending order, then a fixed tree over the threads in LDS.  Short rows
// per block and a tree at the end keep every dependent chain short.
constexpr int COLSUM_ROWS = 32;
__global__ void __launch_bounds__(256)
    k_colsum_part(const float *__restrict__ x, int n, int m, float *__restrict__ part) {
  __shared__ float red[256];
  const int t = int(threadIdx.x), b = int(blockIdx.x);
  const int r0 = b * COLSUM_ROWS, r1 = min(n, r0 + COLSUM_ROWS);
  for (int c0 = 0; c0 < m; c0 += 256) {
    const int cw = min(m - c0, 256), lanes = 256 / cw, lane = t / cw, c = c0 + t % cw;
    float s = 0.f;
    if (lane < lanes)
      for (int r = r0 + lane; r < r1; r += lanes) s += x[size_t(r) * m + c];
    red[t] = s;
    __syncthreads();
    if (t < cw) {
      float tot = red[t];
      for (int l = 1; l < lanes; ++l) tot += red[l * cw + t];
      part[size_t(b) * m + c0 + t] = tot;
    }
    __syncthreads();
  }
}
__global__ void __launch_bounds__(256) k_colsum_fin(const float *__restrict__ part, int nb, int m,
                                                    float *__restrict__ out) {
  __shared__ float red[256];
  const int c = int(blockIdx.x), t = int(threadIdx.x);
  float s = 0.f;
  for (int b = t; b < nb; b += 256) s += part[size_t(b) * m + c];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[c] = red[0];
}

// Adam on FP32 masters, refreshes the BF16 copy (may be null for biases)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       uint16_t *__restrict__ p16, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  p[i] = pi;
  if (p16) p16[i] = f2bf(pi);
}

__global__ void k_to_bf16(const float *__restrict__ a, uint16_t *__restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

__global__ void k_zero_f(float *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

inline unsigned blocks(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

template <bool TA, bool TB>
void gemm(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, const float *bias,
          const uint16_t *mask, int ldm, int relu, float *C32, uint16_t *C16, int ldc, hipStream_t s,
          int splits = 1) {
  const int kc = splits > 1 ? ((K + splits - 1) / splits + 15) / 16 * 16 : K;
  dim3 grid(unsigned((N + 63) / 64), unsigned((M + 63) / 64), unsigned(splits));
  const bool staged = M % 64 == 0 && N % 64 == 0 && K % 64 == 0 && kc % 64 == 0 && lda % 8 == 0 && ldb % 8 == 0 &&
                      (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0;
  prof_begin("k_gemm", s);
  if (staged)
    hipLaunchKernelGGL((k_gemm_lds<TA, TB>), grid, dim3(256), 0, s, M, N, K, kc, A, lda, B, ldb, bias, mask, ldm,
                       relu, C32, C16, ldc);
  else
    hipLaunchKernelGGL((k_gemm<TA, TB>), grid, dim3(256), 0, s, M, N, K, kc, A, lda, B, ldb, bias, mask, ldm, relu,
                       C32, C16, ldc);
  prof_end("k_gemm", s);
  HIP_CHECK(hipGetLastError());
}

// number of K splits that gives a small output tile grid enough blocks
inline int k_splits(int M, int N, int K) {
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  int s = 1;
  while (s < 64 && tiles * s < 512 && K / (2 * s) >= 256) s *= 2;
  return s;
}

}  // namespace

struct casl_mlp {
  int L = 0;                        // number of layers
  std::vector<int> dim;             // L + 1 widths
  DevBuf<float> W, b, gW, gb, mW, vW, mb, vb;  // concatenated per layer
  DevBuf<uint16_t> W16;
  std::vector<size_t> offW, offb;   // per-layer offsets
  DevBuf<float> mean, inv_std;
  int cap = 0;                      // activation capacity (rows)
  DevBuf<uint16_t> act16;           // L activations [cap x dim[l]] (l = 0 .. L-1) as bf16
  DevBuf<float> z32;                // FP32 scratch [cap x max width]
  DevBuf<uint16_t> d16a, d16b;      // BF16 back-propagated gradients
  DevBuf<float> y32, loss;
  DevBuf<float> part;               // split-K partial sums of weight gradients
  DevBuf<float> colpart;            // per-block column sums of bias gradients
  std::vector<size_t> offA;
  int maxw = 0;
  int64_t step = 0;

  void ensure_cap(int n) {
    if (n <= cap) return;
    cap = n;
    size_t tot = 0;
    offA.assign(L, 0);
    for (int l = 0; l < L; ++l) {
      offA[l] = tot;
      tot += size_t(n) * dim[l];
    }
    act16.ensure(tot);
    z32.ensure(size_t(n) * maxw);
    d16a.ensure(size_t(n) * maxw);
    d16b.ensure(size_t(n) * maxw);
    y32.ensure(size_t(n) * dim[L]);
  }
};

namespace {

// forward pass; leaves activations in act16 and the output in out (FP32 [n x dim[L]])
void forward(casl_mlp *m, const float *x, int n, float *out, hipStream_t s) {
  m->ensure_cap(n);
  hipLaunchKernelGGL(k_pack_input, dim3(blocks(int64_t(n) * m->dim[0])), dim3(256), 0, s, x, m->mean.p,
                     m->inv_std.p, n, m->dim[0], m->act16.p + m->offA[0]);
  HIP_CHECK(hipGetLastError());
  for (int l = 0; l < m->L; ++l) {
    const int in = m->dim[l], o = m->dim[l + 1];
    const uint16_t *A = m->act16.p + m->offA[l];
    const bool last = (l == m->L - 1);
    gemm<false, false>(n, o, in, A, in, m->W16.p + m->offW[l], o, m->b.p + m->offb[l], nullptr, 0, last ? 0 : 1,
                       last ? out : nullptr, last ? nullptr : m->act16.p + m->offA[l + 1], o, s);
  }
}

// ------------------------------------------------ fused network curvature
// The learned h*kappa of samples in one launch: input normalisation, every
// layer of a (10|28)-128-128-128-128-1 network on MFMA with the activations
// in LDS, and the hybrid blend.  Each block holds all weights in LDS (stored
// [out][in], so a column's 8 consecutive k are one 16-byte read) and runs
// tiles of 128 network rows: 2D samples give 2 rows each (as is, reflected),
// 3D samples one.  Wave w owns rows 32w .. 32w + 31 in every layer: it reads
// only its own rows of X and overwrites them with the layer's output, so the
// layers need no barrier.  Every product, bias add, ReLU and BF16 rounding is
// the one k_gemm makes, in the same order (the K loop ascends in steps of
// 16), so the result equals casl_curvature_normalize(_3d) + casl_mlp_forward
// + casl_curvature_hybrid(_3d) bit for bit.
constexpr int NF_H = 128;            // hidden width
constexpr int NF_XS = NF_H + 8;      // LDS row stride (halves) of X and of the hidden weights
constexpr int NF_K0S = 32 + 8;       // LDS row stride of the first layer's weights
constexpr int NF_ROWS = 128;         // network rows per tile

struct FusedNet {
  const uint16_t *W16;  // [in][out] per layer, concatenated
  const float *b;       // per layer, concatenated
  const float *mean, *inv_std;
  size_t offW[5], offb[5];
  int in;               // 10 or 28
};

// LDS footprint (bytes): weights, X, biases, predictions
constexpr size_t NF_LDS = size_t(NF_H) * NF_K0S * 2 + 3 * size_t(NF_H) * NF_XS * 2 + size_t(32) * NF_XS * 2 +
                          size_t(NF_ROWS) * NF_XS * 2 + (4 * NF_H + 1) * 4 + NF_ROWS * 4;
static_assert(NF_LDS <= 160 * 1024, "fused network kernel exceeds the LDS of a compute unit");

template <int D>
__global__ void __launch_bounds__(256)
    k_net_curvature(FusedNet net, const float *__restrict__ feat, int64_t n, float lo, float hi, int residual,
                    float dead, float *__restrict__ hk) {
  constexpr int NFEAT = D == 2 ? 10 : 28;                 // features per sample
  constexpr int RPS = D == 2 ? 2 : 1;                     // network rows per sample
  constexpr int SPT = NF_ROWS / RPS;                      // samples per tile
  extern __shared__ __align__(16) unsigned char lds[];
  uint16_t *W0 = reinterpret_cast<uint16_t *>(lds);                 // [128][NF_K0S]
  uint16_t *Wh = W0 + NF_H * NF_K0S;                                // 3 x [128][NF_XS]
  uint16_t *W4 = Wh + 3 * NF_H * NF_XS;                             // [32][NF_XS], row 0 = last layer
  uint16_t *X = W4 + 32 * NF_XS;                                    // [128][NF_XS]
  float *bias = reinterpret_cast<float *>(X + NF_ROWS * NF_XS);     // 4 x 128 + 1
  float *pred = bias + 4 * NF_H + 1;                                // [128]
  const int tid = int(threadIdx.x), wave = tid >> 6, lane = tid & 63;
  const int in = net.in, k0p = (in + 15) / 16 * 16;
  // weights, transposed into [out][in], zero beyond each layer's width
  for (int e = tid; e < NF_H * 32; e += 256) {
    const int j = e / 32, k = e % 32;
    W0[j * NF_K0S + k] = k < in ? net.W16[net.offW[0] + size_t(k) * NF_H + j] : uint16_t(0);
  }
  for (int l = 1; l <= 3; ++l)
    for (int e = tid; e < NF_H * NF_H; e += 256) {
      const int k = e / NF_H, j = e % NF_H;  // coalesced over j in global memory
      Wh[(l - 1) * NF_H * NF_XS + j * NF_XS + k] = net.W16[net.offW[l] + size_t(k) * NF_H + j];
    }
  for (int e = tid; e < 32 * NF_H; e += 256) {
    const int j = e / NF_H, k = e % NF_H;
    W4[j * NF_XS + k] = j == 0 ? net.W16[net.offW[4] + size_t(k)] : uint16_t(0);
  }
  for (int e = tid; e < 4 * NF_H + 1; e += 256) {
    const int l = e / NF_H, j = e % NF_H;
    bias[e] = net.b[net.offb[l] + size_t(j)];
  }
  const int64_t ntiles = (n + SPT - 1) / SPT;
  const int li = lane & 31, lk = (lane >> 5) * 8;
  const int row0 = wave * 32;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    __syncthreads();  // weights loaded; the previous tile's X and pred consumed
    // inputs: normalised, standardised, BF16 (k_pack_input's rounding)
    for (int t = tid; t < SPT; t += 256) {
      const int64_t smp = tile * SPT + t;
      float o[RPS * NFEAT];
      if (smp < n) {
        if constexpr (D == 2)
          curv_normalize2(feat + NFEAT * smp, o);
        else
          curv_normalize3(feat + NFEAT * smp, o);
      }
#pragma unroll
      for (int r = 0; r < RPS; ++r) {
        uint16_t *xr = X + (t * RPS + r) * NF_XS;
        for (int k = 0; k < 32; ++k)
          xr[k] = (smp < n && k < in) ? f2bf((o[r * NFEAT + k] - net.mean[k]) * net.inv_std[k]) : uint16_t(0);
      }
    }
    __syncthreads();
    // hidden layers: wave-owned rows, in place
    for (int l = 0; l < 4; ++l) {
      const uint16_t *W = l == 0 ? W0 : Wh + (l - 1) * NF_H * NF_XS;
      const int ws = l == 0 ? NF_K0S : NF_XS;
      const int K = l == 0 ? k0p : NF_H;
      f32x16 acc[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
      for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(X + (row0 + li) * NF_XS + k0 + lk);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bf16x8 b = *reinterpret_cast<const bf16x8 *>(W + (32 * c + li) * ws + k0 + lk);
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
        }
      }
      const float *bl = bias + l * NF_H;
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int row = row0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3), col = 32 * c + li;
          float v = acc[c][r];
          v += bl[col];
          v = v > 0.f ? v : 0.f;
          X[row * NF_XS + col] = f2bf(v);
        }
    }
    // last layer (one output column)
    {
      f32x16 acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.f;
      for (int k0 = 0; k0 < NF_H; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(X + (row0 + li) * NF_XS + k0 + lk);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(W4 + li * NF_XS + k0 + lk);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
      if (li == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) pred[row0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] = acc[r] + bias[4 * NF_H];
      }
    }
    __syncthreads();
    for (int t = tid; t < SPT; t += 256) {
      const int64_t smp = tile * SPT + t;
      if (smp >= n) continue;
      const float ihk = feat[NFEAT * smp + NFEAT - 1];
      if constexpr (D == 2)
        hk[smp] = curv_hybrid2(ihk, pred[2 * t], pred[2 * t + 1], lo, hi, residual);
      else
        hk[smp] = curv_hybrid3(ihk, pred[t], lo, hi, residual, dead);
    }
  }
}

}  // namespace

extern "C" {

int casl_mlp_create(const int32_t *dims, int n_layers, uint64_t seed, casl_mlp_t **out) {
  return guarded([&]() {
    if (!out || !dims || n_layers < 1) throw std::invalid_argument("bad arguments");
    auto m = std::make_unique<casl_mlp>();
    m->L = n_layers;
    m->dim.assign(dims, dims + n_layers + 1);
    size_t nw = 0, nb = 0;
    for (int l = 0; l < n_layers; ++l) {
      if (m->dim[l] < 1 || m->dim[l + 1] < 1) throw std::invalid_argument("layer widths must be positive");
      m->offW.push_back(nw);
      m->offb.push_back(nb);
      nw += size_t(m->dim[l]) * m->dim[l + 1];
      nb += size_t(m->dim[l + 1]);
    }
    for (int d : m->dim) m->maxw = std::max(m->maxw, d);
    // He-normal initialisation (deterministic in seed), zero biases
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(nw), bb(nb, 0.f);
    for (int l = 0; l < n_layers; ++l) {
      const double sd = std::sqrt(2.0 / m->dim[l]);
      for (size_t i = 0; i < size_t(m->dim[l]) * m->dim[l + 1]; ++i) w[m->offW[l] + i] = float(sd * nd(rng));
    }
    std::vector<uint16_t> w16(nw);
    for (size_t i = 0; i < nw; ++i) w16[i] = f2bf(w[i]);
    m->W.ensure(nw);
    m->W16.ensure(nw);
    m->gW.ensure(nw);
    m->mW.ensure(nw);
    m->vW.ensure(nw);
    m->b.ensure(nb);
    m->gb.ensure(nb);
    m->mb.ensure(nb);
    m->vb.ensure(nb);
    HIP_CHECK(hipMemcpy(m->W.p, w.data(), nw * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->W16.p, w16.data(), nw * 2, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->b.p, bb.data(), nb * 4, hipMemcpyHostToDevice));
    std::vector<float> zw(nw, 0.f), zb(nb, 0.f);
    HIP_CHECK(hipMemcpy(m->mW.p, zw.data(), nw * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->vW.p, zw.data(), nw * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->mb.p, zb.data(), nb * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->vb.p, zb.data(), nb * 4, hipMemcpyHostToDevice));
    std::vector<float> mean(m->dim[0], 0.f), inv(m->dim[0], 1.f);
    m->mean.ensure(mean.size());
    m->inv_std.ensure(inv.size());
    HIP_CHECK(hipMemcpy(m->mean.p, mean.data(), mean.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->inv_std.p, inv.data(), inv.size() * 4, hipMemcpyHostToDevice));
    m->loss.ensure(1);
    *out = m.release();
  });
}

int casl_curvature_learned(casl_mlp_t *m, const float *feat, int64_t n, float lo, float hi, int residual,
                           float deadband, float *hk, void *stream) {
  return guarded([&]() {
    if (!m || n < 0 || (n > 0 && (!feat || !hk)) || !(hi > lo) || !(lo >= 0.f) || !(deadband >= 0.f))
      throw std::invalid_argument("bad arguments");
    if (m->L != 5 || (m->dim[0] != 10 && m->dim[0] != 28) || m->dim[1] != NF_H || m->dim[2] != NF_H ||
        m->dim[3] != NF_H || m->dim[4] != NF_H || m->dim[5] != 1)
      throw std::invalid_argument("casl_curvature_learned takes a (10|28)-128-128-128-128-1 network");
    if (n == 0) return;
    FusedNet net{};
    net.W16 = m->W16.p;
    net.b = m->b.p;
    net.mean = m->mean.p;
    net.inv_std = m->inv_std.p;
    for (int l = 0; l < 5; ++l) {
      net.offW[l] = m->offW[l];
      net.offb[l] = m->offb[l];
    }
    net.in = m->dim[0];
    hipStream_t s = static_cast<hipStream_t>(stream);
    int dev = 0, cus = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t spt = net.in == 10 ? NF_ROWS / 2 : NF_ROWS;
    const int64_t ntiles = (n + spt - 1) / spt;
    const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>(ntiles, int64_t(cus))));
    prof_begin("k_net_curvature", s);
    if (net.in == 10) {
      HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_net_curvature<2>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(NF_LDS)));
      hipLaunchKernelGGL(k_net_curvature<2>, dim3(grid), dim3(256), NF_LDS, s, net, feat, n, lo, hi, residual,
                         deadband, hk);
    } else {
      HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_net_curvature<3>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(NF_LDS)));
      hipLaunchKernelGGL(k_net_curvature<3>, dim3(grid), dim3(256), NF_LDS, s, net, feat, n, lo, hi, residual,
                         deadband, hk);
    }
    prof_end("k_net_curvature", s);
    HIP_CHECK(hipGetLastError());
  });
}

int casl_mlp_destroy(casl_mlp_t *m) {
  delete m;
  return CASL_OK;
}

int casl_mlp_num_params(const casl_mlp_t *m, int64_t *n_weights, int64_t *n_biases) {
  return guarded([&]() {
    if (!m) throw std::invalid_argument("null mlp");
    if (n_weights) *n_weights = int64_t(m->W.n);
    if (n_biases) *n_biases = int64_t(m->b.n);
  });
}

int casl_mlp_set_scaler(casl_mlp_t *m, const float *mean, const float *stdev) {
  return guarded([&]() {
    if (!m || !mean || !stdev) throw std::invalid_argument("null argument");
    std::vector<float> inv(m->dim[0]);
    for (int i = 0; i < m->dim[0]; ++i) {
      if (!(stdev[i] > 0)) throw std::invalid_argument("scaler deviations must be positive");
      inv[i] = 1.f / stdev[i];
    }
    HIP_CHECK(hipMemcpy(m->mean.p, mean, m->dim[0] * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->inv_std.p, inv.data(), m->dim[0] * 4, hipMemcpyHostToDevice));
  });
}

int casl_mlp_get_params(const casl_mlp_t *m, float *weights, float *biases) {
  return guarded([&]() {
    if (!m) throw std::invalid_argument("null mlp");
    if (weights) HIP_CHECK(hipMemcpy(weights, m->W.p, m->W.n * 4, hipMemcpyDeviceToHost));
    if (biases) HIP_CHECK(hipMemcpy(biases, m->b.p, m->b.n * 4, hipMemcpyDeviceToHost));
  });
}

int casl_mlp_set_params(casl_mlp_t *m, const float *weights, const float *biases) {
  return guarded([&]() {
    if (!m || !weights || !biases) throw std::invalid_argument("null argument");
    std::vector<uint16_t> w16(m->W.n);
    for (size_t i = 0; i < m->W.n; ++i) w16[i] = f2bf(weights[i]);
    HIP_CHECK(hipMemcpy(m->W.p, weights, m->W.n * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->W16.p, w16.data(), m->W.n * 2, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(m->b.p, biases, m->b.n * 4, hipMemcpyHostToDevice));
  });
}

int casl_mlp_get_grads(const casl_mlp_t *m, float *gw, float *gb) {
  return guarded([&]() {
    if (!m) throw std::invalid_argument("null mlp");
    if (gw) HIP_CHECK(hipMemcpy(gw, m->gW.p, m->gW.n * 4, hipMemcpyDeviceToHost));
    if (gb) HIP_CHECK(hipMemcpy(gb, m->gb.p, m->gb.n * 4, hipMemcpyDeviceToHost));
  });
}

int casl_mlp_forward(casl_mlp_t *m, const float *x, int n, float *y, void *stream) {
  return guarded([&]() {
    if (!m || !x || !y || n < 0) throw std::invalid_argument("bad arguments");
    if (n == 0) return;
    forward(m, x, n, y, static_cast<hipStream_t>(stream));
  });
}

// Forward + backward of the mean squared error on (x, t); gradients stay on
// the device (casl_mlp_get_grads).  *loss (host, may be null) receives the MSE.
int casl_mlp_backward(casl_mlp_t *m, const float *x, const float *t, int n, float *loss, void *stream) {
  return guarded([&]() {
    if (!m || !x || !t || n <= 0) throw std::invalid_argument("bad arguments");
    if (m->dim[m->L] != 1) throw std::invalid_argument("training needs a scalar output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    m->ensure_cap(n);  // sizes y32 before its address is taken
    forward(m, x, n, m->y32.p, s);
    hipLaunchKernelGGL(k_zero_f, dim3(1), dim3(64), 0, s, m->loss.p, 1);
    // dZ_L = 2 (y - t) / n  -> z32 (FP32) and d16a (BF16), [n x 1]
    hipLaunchKernelGGL(k_mse_seed, dim3(blocks(n)), dim3(256), 0, s, m->y32.p, t, n, 1.f / float(n), m->z32.p,
                       m->d16a.p, m->loss.p);
    HIP_CHECK(hipGetLastError());
    uint16_t *dcur = m->d16a.p, *dnext = m->d16b.p;
    for (int l = m->L - 1; l >= 0; --l) {
      const int in = m->dim[l], o = m->dim[l + 1];
      // bias gradient from the FP32 dZ_l in z32
      const int nb = (n + COLSUM_ROWS - 1) / COLSUM_ROWS;
      m->colpart.ensure(size_t(std::max(nb, 1)) * o);
      hipLaunchKernelGGL(k_colsum_part, dim3(unsigned(std::max(nb, 1))), dim3(256), 0, s, m->z32.p, n, o,
                         m->colpart.p);
      hipLaunchKernelGGL(k_colsum_fin, dim3(unsigned(o)), dim3(256), 0, s, m->colpart.p, std::max(nb, 1), o,
                         m->gb.p + m->offb[l]);
      HIP_CHECK(hipGetLastError());
      const uint16_t *Aprev = m->act16.p + m->offA[l];  // [n x in]
      // dW_l = A_{l-1}^T dZ_l : (in x n)(n x o)
      const int sp = k_splits(in, o, n);
      if (sp == 1) {
        gemm<true, false>(in, o, n, Aprev, in, dcur, o, nullptr, nullptr, 0, 0, m->gW.p + m->offW[l], nullptr, o, s);
      } else {
        m->part.ensure(size_t(sp) * in * o);
        gemm<true, false>(in, o, n, Aprev, in, dcur, o, nullptr, nullptr, 0, 0, m->part.p, nullptr, o, s, sp);
        const int64_t e = int64_t(in) * o;
        hipLaunchKernelGGL(k_sum_splits, dim3(unsigned((e + 63) / 64)), dim3(64), 0, s, m->part.p, sp, e,
                           m->gW.p + m->offW[l]);
        HIP_CHECK(hipGetLastError());
      }
      if (l > 0) {
        // dZ_{l-1} = (dZ_l W_l^T) * relu'(A_{l-1}) : (n x o)(o x in)
        gemm<false, true>(n, in, o, dcur, o, m->W16.p + m->offW[l], o, nullptr, Aprev, in, 0, m->z32.p, dnext, in, s);
        std::swap(dcur, dnext);
      }
    }
    if (loss) {
      float h = 0.f;
      HIP_CHECK(hipMemcpyAsync(&h, m->loss.p, 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      *loss = h / float(n);
    }
  });
}

int casl_mlp_adam_step(casl_mlp_t *m, float lr, float beta1, float beta2, float eps, void *stream) {
  return guarded([&]() {
    if (!m) throw std::invalid_argument("null mlp");
    hipStream_t s = static_cast<hipStream_t>(stream);
    m->step += 1;
    const float c1 = float(1.0 / (1.0 - std::pow(double(beta1), double(m->step))));
    const float c2 = float(1.0 / (1.0 - std::pow(double(beta2), double(m->step))));
    const int nw = int(m->W.n), nb = int(m->b.n);
    hipLaunchKernelGGL(k_adam, dim3(blocks(nw)), dim3(256), 0, s, m->W.p, m->gW.p, m->mW.p, m->vW.p, m->W16.p, nw,
                       lr, beta1, beta2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3(blocks(nb)), dim3(256), 0, s, m->b.p, m->gb.p, m->mb.p, m->vb.p,
                       (uint16_t *)nullptr, nb, lr, beta1, beta2, eps, c1, c2);
    HIP_CHECK(hipGetLastError());
  });
}

// Plain BF16 MFMA GEMM on device buffers (exposed for testing the kernel):
// C = op(A) op(B) (+ bias) (relu), C32 and/or C16 written with leading dim ldc.
int casl_bf16_gemm(int M, int N, int K, const uint16_t *A, int lda, int trans_a, const uint16_t *B, int ldb,
                   int trans_b, const float *bias, int relu, float *C32, uint16_t *C16, int ldc, void *stream) {
  return guarded([&]() {
    if (M < 0 || N < 0 || K < 0 || !A || !B || (!C32 && !C16)) throw std::invalid_argument("bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!trans_a && !trans_b)
      gemm<false, false>(M, N, K, A, lda, B, ldb, bias, nullptr, 0, relu, C32, C16, ldc, s);
    else if (trans_a && !trans_b)
      gemm<true, false>(M, N, K, A, lda, B, ldb, bias, nullptr, 0, relu, C32, C16, ldc, s);
    else if (!trans_a && trans_b)
      gemm<false, true>(M, N, K, A, lda, B, ldb, bias, nullptr, 0, relu, C32, C16, ldc, s);
    else
      throw std::invalid_argument("both operands transposed is not supported");
  });
}

}  // extern "C"
