// Coordinate MLP of the learned warm start (include/usg_mlp.h) on gfx950:
// every product is v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32
// accumulation).
//
// Operand layout of one 16x16x32 MFMA in a 64-lane wave (lane l):
//   A (16 x 32): row l % 16, k = 8 * (l / 16) .. +7      (8 bf16 in a lane)
//   B (32 x 16): column l % 16, k = 8 * (l / 16) .. +7
//   C (16 x 16): column l % 16, rows 4 * (l / 16) .. +3  (4 floats)
// (checked against the CPU by usg_mlp_mfma_check).  Activations are kept
// twice: row-major [P][K] (the A operand of the forward and data-gradient
// products, 8 consecutive k per lane) and point-blocked [P/8][K][8] (both
// operands of the weight-gradient product, which reduces over points: 8
// consecutive points per lane).  Every operand load is one 16-byte load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/usg_mlp.h"

namespace usgmlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;  // raw bits

constexpr int kWave = 64;
constexpr int kChunk = 256;  // points per partial sum of the weight and bias gradients

thread_local std::string g_err;
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}
#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(-3, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// FP32 -> BF16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ bf16 to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ __forceinline__ bf16x8 load8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// ---------------------------------------------------------------------------
// Forward of one layer: Z = X W + b over 64 points x NT*16 outputs per block
// (4 waves x 16 points).  Wt is W transposed ([N][K]) so that the B operand
// is 8 consecutive k of one output.  ACT: H = bf16(silu(Z)) row-major (and
// point-blocked + Z kept when TRAIN); else Z is the network output.
// ---------------------------------------------------------------------------
template <int NT, bool ACT, bool TRAIN>
__global__ __launch_bounds__(256) void k_fwd(const bf16 *__restrict__ X, int K, const bf16 *__restrict__ Wt,
                                             const float *__restrict__ b, int N, float *__restrict__ Z,
                                             bf16 *__restrict__ H, bf16 *__restrict__ Hb) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int p0 = blockIdx.x * 64 + wave * 16;
  const int n0 = blockIdx.y * NT * 16;
  floatx4 acc[NT];
  for (int t = 0; t < NT; t++) acc[t] = floatx4{0, 0, 0, 0};
  const bf16 *xa = X + (size_t)(p0 + lane % 16) * K + (lane / 16) * 8;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bf16x8 a = load8(xa + k0);
    for (int t = 0; t < NT; t++) {
      const bf16x8 bv = load8(Wt + (size_t)(n0 + t * 16 + lane % 16) * K + k0 + (lane / 16) * 8);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
    }
  }
  for (int t = 0; t < NT; t++) {
    const int n = n0 + t * 16 + lane % 16;
    for (int r = 0; r < 4; r++) {
      const int p = p0 + (lane / 16) * 4 + r;
      const float z = acc[t][r] + b[n];
      if (!ACT) {
        Z[(size_t)p * N + n] = z;
        continue;
      }
      if (TRAIN) Z[(size_t)p * N + n] = z;
      const bf16 h = to_bf16(z * sigmoid(z));
      H[(size_t)p * N + n] = h;
      if (TRAIN) Hb[((size_t)(p / 8) * N + n) * 8 + p % 8] = h;
    }
  }
}

// ---------------------------------------------------------------------------
// Data gradient through one layer: dH = dZ W^T (reduction over its N
// outputs), then dZ_prev = bf16(dH * silu'(Z_prev)) row-major and
// point-blocked.  W is [K][N] so that the B operand is 8 consecutive n of
// one input k.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_bwd_data(const bf16 *__restrict__ dZ, int N, const bf16 *__restrict__ W,
                                                  int K, const float *__restrict__ Zp, bf16 *__restrict__ dZp,
                                                  bf16 *__restrict__ dZpb) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int p0 = blockIdx.x * 64 + wave * 16;
  const int k0 = blockIdx.y * NT * 16;
  floatx4 acc[NT];
  for (int t = 0; t < NT; t++) acc[t] = floatx4{0, 0, 0, 0};
  const bf16 *da = dZ + (size_t)(p0 + lane % 16) * N + (lane / 16) * 8;
  for (int n0 = 0; n0 < N; n0 += 32) {
    const bf16x8 a = load8(da + n0);
    for (int t = 0; t < NT; t++) {
      const bf16x8 bv = load8(W + (size_t)(k0 + t * 16 + lane % 16) * N + n0 + (lane / 16) * 8);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
    }
  }
  for (int t = 0; t < NT; t++) {
    const int k = k0 + t * 16 + lane % 16;
    for (int r = 0; r < 4; r++) {
      const int p = p0 + (lane / 16) * 4 + r;
      const float z = Zp[(size_t)p * K + k];
      const float s = sigmoid(z);
      const bf16 d = to_bf16(acc[t][r] * (s * (1.0f + z * (1.0f - s))));
      dZp[(size_t)p * K + k] = d;
      dZpb[((size_t)(p / 8) * K + k) * 8 + p % 8] = d;
    }
  }
}

// ---------------------------------------------------------------------------
// Row-major product with the B tile in LDS: C[P][N] = A[P][K] x B, B given
// transposed (Bt[N][K]).  A block (4 waves) owns 128 points x 64 outputs;
// each wave 32 points x 64 outputs (2 x 4 MFMA tiles per k step).  The
// block's 64 rows of Bt are staged in LDS once (rows padded by 8 bf16 so
// that the 16 lanes reading one column group hit different banks).
// EPI 0: forward, Z = C + b, H = bf16(silu(Z)) (+ blocked copy and Z when
// TRAIN); EPI 1: data gradient, dZp = bf16(C * silu'(Zp)) row-major and
// blocked.
// ---------------------------------------------------------------------------
constexpr int kBN = 64;    // outputs per block
constexpr int kBP = 128;   // points per block
template <int EPI, bool TRAIN>
__global__ __launch_bounds__(256) void k_gemm(const bf16 *__restrict__ A, int K, const bf16 *__restrict__ Bt,
                                              const float *__restrict__ bias, int N, const float *__restrict__ Zp,
                                              float *__restrict__ Z, bf16 *__restrict__ H, bf16 *__restrict__ Hb) {
  extern __shared__ __align__(16) bf16 sB[];  // [kBN][K + 8]
  const int ldb = K + 8;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int n0 = blockIdx.y * kBN;
  // stage Bt rows n0 .. n0+63 (16-byte chunks)
  const int chunks = kBN * K / 8;
  for (int i = threadIdx.x; i < chunks; i += 256) {
    const int r = i / (K / 8), c8 = i % (K / 8);
    *reinterpret_cast<bf16x8 *>(&sB[r * ldb + c8 * 8]) = load8(Bt + (size_t)(n0 + r) * K + c8 * 8);
  }
  __syncthreads();
  const int p0 = blockIdx.x * kBP + wave * 32;
  floatx4 acc[2][4];
  for (int a = 0; a < 2; a++)
    for (int t = 0; t < 4; t++) acc[a][t] = floatx4{0, 0, 0, 0};
  const bf16 *a0p = A + (size_t)(p0 + lane % 16) * K + (lane / 16) * 8;
  const bf16 *a1p = a0p + (size_t)16 * K;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bf16x8 a0 = load8(a0p + k0), a1 = load8(a1p + k0);
    for (int t = 0; t < 4; t++) {
      const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&sB[(t * 16 + lane % 16) * ldb + k0 + (lane / 16) * 8]);
      acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bv, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bv, acc[1][t], 0, 0, 0);
    }
  }
  for (int a = 0; a < 2; a++)
    for (int t = 0; t < 4; t++) {
      const int n = n0 + t * 16 + lane % 16;
      for (int r = 0; r < 4; r++) {
        const int p = p0 + a * 16 + (lane / 16) * 4 + r;
        const size_t o = (size_t)p * N + n;
        if (EPI == 0) {
          const float z = acc[a][t][r] + bias[n];
          if (TRAIN) Z[o] = z;
          const bf16 h = to_bf16(z * sigmoid(z));
          H[o] = h;
          if (TRAIN) Hb[((size_t)(p / 8) * N + n) * 8 + p % 8] = h;
        } else {
          const float z = Zp[o];
          const float s = sigmoid(z);
          const bf16 d = to_bf16(acc[a][t][r] * (s * (1.0f + z * (1.0f - s))));
          H[o] = d;
          Hb[((size_t)(p / 8) * N + n) * 8 + p % 8] = d;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient partial of one chunk of kChunk points:
// part[chunk][k][n] = sum_{p in chunk} X[p][k] dZ[p][n], 64 k x NT*16 n per
// block (4 waves x 16 k), operands from the point-blocked copies.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_bwd_weight(const bf16 *__restrict__ Xb, int K, const bf16 *__restrict__ dZb,
                                                    int N, int P, float *__restrict__ part) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int k0 = blockIdx.x * 64 + wave * 16;
  const int n0 = blockIdx.y * NT * 16;
  const int pa = blockIdx.z * kChunk, pb = min(P, pa + kChunk);
  floatx4 acc[NT];
  for (int t = 0; t < NT; t++) acc[t] = floatx4{0, 0, 0, 0};
  for (int p = pa; p < pb; p += 32) {
    const size_t blk = (size_t)(p / 8) + lane / 16;
    const bf16x8 a = load8(Xb + (blk * K + k0 + lane % 16) * 8);
    for (int t = 0; t < NT; t++) {
      const bf16x8 bv = load8(dZb + (blk * N + n0 + t * 16 + lane % 16) * 8);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[t], 0, 0, 0);
    }
  }
  float *o = part + (size_t)blockIdx.z * K * N;
  for (int t = 0; t < NT; t++)
    for (int r = 0; r < 4; r++) {
      const int k = k0 + (lane / 16) * 4 + r, n = n0 + t * 16 + lane % 16;
      o[(size_t)k * N + n] = acc[t][r];
    }
}

// bias-gradient partial of one chunk of kChunk points from the point-blocked
// copy: block (chunk, 32 columns); thread (column c, j) sums the 8-point
// blocks j, j + 8, ... of the chunk (16-byte loads), then the 8 partial sums
// of a column are added in j order (reproducible).
__global__ __launch_bounds__(256) void k_colsum(const bf16 *__restrict__ dZb, int N, int P, float *__restrict__ partb) {
  __shared__ float sh[8][32];
  const int c = threadIdx.x % 32, j = threadIdx.x / 32;
  const int n = blockIdx.y * 32 + c;
  const int pa = blockIdx.x * kChunk, pb = min(P, pa + kChunk);
  float s = 0;
  for (int b = pa / 8 + j; b < pb / 8; b += 8) {
    const bf16x8 v = load8(dZb + ((size_t)b * N + n) * 8);
    for (int e = 0; e < 8; e++) s += (float)v[e];
  }
  sh[j][c] = s;
  __syncthreads();
  if (j == 0) {
    float t = 0;
    for (int q = 0; q < 8; q++) t += sh[q][c];
    partb[(size_t)blockIdx.x * N + n] = t;
  }
}

// sum of nchunk partials in chunk order (loads batched ahead of the adds)
__device__ __forceinline__ float chunk_sum(const float *__restrict__ part, int nchunk, long len, long i) {
  float s = 0;
  int c = 0;
  for (; c + 8 <= nchunk; c += 8) {
    float v[8];
    for (int u = 0; u < 8; u++) v[u] = part[(size_t)(c + u) * len + i];
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; c < nchunk; c++) s += part[(size_t)c * len + i];
  return s;
}
__global__ void k_sum_chunks(const float *__restrict__ part, int nchunk, long len, float *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  out[i] = chunk_sum(part, nchunk, len, i);
}

// Loss and output gradient: out[P][No] (fp32), targets y[n][nout] of the
// first nv points; dZ = bf16(2 (out - y) / (nv nout)) (zero on padding),
// row-major and point-blocked; lossp[block] = partial sum of squares.
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ out, int No, const float *__restrict__ y,
                                              int nout, int nv, bf16 *__restrict__ dZ, bf16 *__restrict__ dZb,
                                              float *__restrict__ lossp) {
  __shared__ float sh[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // over P * No
  const int p = i / No, n = i % No;
  float sq = 0, g = 0;
  if (p < nv && n < nout) {
    const float e = out[i] - y[(size_t)p * nout + n];
    sq = e * e;
    g = 2.0f * e / ((float)nv * (float)nout);
  }
  const bf16 d = to_bf16(g);
  dZ[i] = d;
  dZb[((size_t)(p / 8) * No + n) * 8 + p % 8] = d;
  sh[threadIdx.x] = sq;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = sh[0];
}

// out[slot] = scale * (sum of n partials): one block, thread t sums
// t, t + 256, ... then a fixed tree (reproducible)
__global__ __launch_bounds__(256) void k_sum_small(const float *__restrict__ part, int n, float scale,
                                                   float *__restrict__ out, int slot) {
  __shared__ double sh[256];
  double a = 0;
  for (int i = threadIdx.x; i < n; i += 256) a += part[i];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[slot] = (float)(sh[0] * scale);
}

// Batch assembly: rows idx[start + i] of the resident set (bf16 inputs padded
// to Kp, fp32 targets), zero rows past n; also the point-blocked input copy.
__global__ __launch_bounds__(256) void k_gather(const bf16 *__restrict__ xs, int Kp, const float *__restrict__ ys,
                                                int nout, const int *__restrict__ idx, int nv, int P,
                                                bf16 *__restrict__ X, bf16 *__restrict__ Xb, float *__restrict__ Y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over P * Kp
  if (i >= (long)P * Kp) return;
  const int p = (int)(i / Kp), k = (int)(i % Kp);
  bf16 v = 0;
  if (p < nv) v = xs[(size_t)idx[p] * Kp + k];
  X[i] = v;
  Xb[((size_t)(p / 8) * Kp + k) * 8 + p % 8] = v;
  if (k < nout) Y[(size_t)p * nout + k] = p < nv ? ys[(size_t)idx[p] * nout + k] : 0.0f;
}

// Adam on one layer (FP32 master weights, padded K x N) straight from the
// chunk partials of its gradients (summed in chunk order, as k_sum_chunks
// forms them), then the BF16 copies W[k][n] and Wt[n][k].
__global__ void k_adam_w_part(float *w, const float *part, int nchunk, float *m1, float *m2, int K, int N, float lr,
                              float c1, float c2, bf16 *W, bf16 *Wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  const float gi = chunk_sum(part, nchunk, (long)K * N, i);
  const float a = 0.9f * m1[i] + 0.1f * gi;
  const float v = 0.999f * m2[i] + 0.001f * gi * gi;
  m1[i] = a;
  m2[i] = v;
  const float wi = w[i] - lr * (a / c1) / (sqrtf(v / c2) + 1e-8f);
  w[i] = wi;
  const int k = i / N, n = i % N;
  const bf16 h = to_bf16(wi);
  W[i] = h;
  Wt[(size_t)n * K + k] = h;
}
__global__ void k_adam_b_part(float *b, const float *partb, int nchunk, float *m1, float *m2, int n, float lr,
                              float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = chunk_sum(partb, nchunk, n, i);
  const float a = 0.9f * m1[i] + 0.1f * gi;
  const float v = 0.999f * m2[i] + 0.001f * gi * gi;
  m1[i] = a;
  m2[i] = v;
  b[i] = b[i] - lr * (a / c1) / (sqrtf(v / c2) + 1e-8f);
}

__global__ void k_mfma_check(const bf16 *a, const bf16 *b, float *d) {
  const int lane = threadIdx.x;
  bf16x8 av, bv;
  for (int e = 0; e < 8; e++) {
    const int k = (lane / 16) * 8 + e;
    uint16_t ab = a[(lane % 16) * 32 + k], bb = b[k * 16 + lane % 16];
    av[e] = __builtin_bit_cast(__bf16, ab);
    bv[e] = __builtin_bit_cast(__bf16, bb);
  }
  floatx4 c = {0, 0, 0, 0};
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
  for (int r = 0; r < 4; r++) d[((lane / 16) * 4 + r) * 16 + lane % 16] = c[r];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Layer {
  int K = 0, N = 0;     // padded fan-in, fan-out
  int k = 0, n = 0;     // real
  float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;
  float *m1w = nullptr, *m2w = nullptr, *m1b = nullptr, *m2b = nullptr;
  bf16 *W = nullptr, *Wt = nullptr;
};

}  // namespace usgmlp

using namespace usgmlp;

struct usg_mlp {
  usg_mlp_desc d{};
  std::vector<Layer> L;
  hipStream_t s = nullptr;
  std::vector<void *> allocs;
  long step = 0;
  // batch buffers (capacity P points)
  int cap = 0;
  std::vector<void *> bufs;
  bf16 *X = nullptr, *Xb = nullptr;
  std::vector<float *> Z;           // per layer [P][N]
  std::vector<bf16 *> H, Hb;        // per hidden layer output [P][N] row / blocked
  bf16 *dA = nullptr, *dAb = nullptr, *dB = nullptr, *dBb = nullptr;  // dZ ping-pong
  float *Y = nullptr, *lossp = nullptr, *part = nullptr, *partb = nullptr;
  // resident training set
  bf16 *xs = nullptr;
  float *ys = nullptr;
  int *idx = nullptr;
  long nset = 0, idx_cap = 0;
  float *lossbuf = nullptr;  // per-step batch losses of an epoch
  long lossbuf_cap = 0;
};

namespace usgmlp {

template <typename T>
int dalloc(std::vector<void *> &v, T **p, size_t n) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
  v.push_back(q);
  *p = (T *)q;
  return 0;
}

int free_bufs(usg_mlp *m) {
  for (void *p : m->bufs) hipFree(p);
  m->bufs.clear();
  m->cap = 0;
  return 0;
}

int ensure_cap(usg_mlp *m, int P) {
  if (P <= m->cap) return 0;
  HIPCHK(hipStreamSynchronize(m->s));
  free_bufs(m);
  const int nl = (int)m->L.size();
  const int K0 = m->L[0].K, W = m->d.width, No = m->L.back().N;
  auto &v = m->bufs;
  int rc = 0;
  rc |= dalloc(v, &m->X, (size_t)P * K0);
  rc |= dalloc(v, &m->Xb, (size_t)P * K0);
  m->Z.assign(nl, nullptr);
  m->H.assign(nl, nullptr);
  m->Hb.assign(nl, nullptr);
  for (int l = 0; l < nl; l++) {
    rc |= dalloc(v, &m->Z[l], (size_t)P * m->L[l].N);
    if (l < nl - 1) {
      rc |= dalloc(v, &m->H[l], (size_t)P * W);
      rc |= dalloc(v, &m->Hb[l], (size_t)P * W);
    }
  }
  const int mx = std::max(W, No);
  rc |= dalloc(v, &m->dA, (size_t)P * mx);
  rc |= dalloc(v, &m->dAb, (size_t)P * mx);
  rc |= dalloc(v, &m->dB, (size_t)P * mx);
  rc |= dalloc(v, &m->dBb, (size_t)P * mx);
  rc |= dalloc(v, &m->Y, (size_t)P * m->d.nout);
  rc |= dalloc(v, &m->lossp, (size_t)P * No / 256 + 1);
  const int nch = (P + kChunk - 1) / kChunk;
  rc |= dalloc(v, &m->part, (size_t)nch * std::max(K0, W) * std::max(W, No));
  rc |= dalloc(v, &m->partb, (size_t)nch * mx);
  if (rc) return rc;
  m->cap = P;
  return 0;
}

int round_up(int a, int b) { return (a + b - 1) / b * b; }

// forward over P (multiple of 64) points already in m->X
void forward(usg_mlp *m, int P, bool train) {
  const int nl = (int)m->L.size();
  const bf16 *x = m->X;
  for (int l = 0; l < nl; l++) {
    const Layer &y = m->L[l];
    const bool last = l == nl - 1;
    if (last) {
      const int NT = 2;
      hipLaunchKernelGGL((k_fwd<NT, false, false>), dim3(P / 64, y.N / (NT * 16)), dim3(256), 0, m->s, x, y.K,
                         y.Wt, y.b, y.N, m->Z[l], nullptr, nullptr);
    } else if (train) {
      hipLaunchKernelGGL((k_gemm<0, true>), dim3(P / kBP, y.N / kBN), dim3(256), kBN * (y.K + 8) * 2, m->s, x, y.K,
                         y.Wt, y.b, y.N, nullptr, m->Z[l], m->H[l], m->Hb[l]);
    } else {
      hipLaunchKernelGGL((k_gemm<0, false>), dim3(P / kBP, y.N / kBN), dim3(256), kBN * (y.K + 8) * 2, m->s, x, y.K,
                         y.Wt, y.b, y.N, nullptr, nullptr, m->H[l], nullptr);
    }
    x = m->H[l];
  }
}

// Adam step parameters (training); null in the gradient-only path
struct AdamStep {
  float lr, c1, c2;
};

// backward from the output gradient in dA/dAb: gradients into L[l].gw/gb,
// or, with an Adam step, each layer updated right after its last use (the
// data-gradient product reads W_l before W_l changes)
void backward(usg_mlp *m, int P, const AdamStep *adam = nullptr) {
  const int nl = (int)m->L.size();
  const int nch = (P + kChunk - 1) / kChunk;
  bf16 *cur = m->dA, *curb = m->dAb, *nxt = m->dB, *nxtb = m->dBb;
  for (int l = nl - 1; l >= 0; l--) {
    Layer &y = m->L[l];
    const bf16 *inb = l == 0 ? m->Xb : m->Hb[l - 1];
    // N is 32 (output layer: 2 column tiles of 16) or a multiple of 64 (4 tiles)
    if (y.N % 64 == 0)
      hipLaunchKernelGGL((k_bwd_weight<4>), dim3(y.K / 64, y.N / 64, nch), dim3(256), 0, m->s, inb, y.K, curb, y.N,
                         P, m->part);
    else
      hipLaunchKernelGGL((k_bwd_weight<2>), dim3(y.K / 64, y.N / 32, nch), dim3(256), 0, m->s, inb, y.K, curb, y.N,
                         P, m->part);
    hipLaunchKernelGGL(k_colsum, dim3(nch, y.N / 32), dim3(256), 0, m->s, curb, y.N, P, m->partb);
    if (!adam) {
      hipLaunchKernelGGL(k_sum_chunks, dim3((y.K * y.N + 255) / 256), dim3(256), 0, m->s, m->part, nch,
                         (long)y.K * y.N, y.gw);
      hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(256), 0, m->s, m->partb, nch, (long)y.N, y.gb);
    }
    if (l > 0) {
      if (y.N % 32 == 0 && y.N >= 64)
        // dH = dZ W^T: A = dZ [P][N], Bt = W [K][N] (rows = outputs k)
        hipLaunchKernelGGL((k_gemm<1, true>), dim3(P / kBP, y.K / kBN), dim3(256), kBN * (y.N + 8) * 2, m->s, cur,
                           y.N, y.W, nullptr, y.K, m->Z[l - 1], nullptr, nxt, nxtb);
      else
        hipLaunchKernelGGL((k_bwd_data<2>), dim3(P / 64, y.K / 32), dim3(256), 0, m->s, cur, y.N, y.W, y.K,
                           m->Z[l - 1], nxt, nxtb);
    }
    if (adam) {
      hipLaunchKernelGGL(k_adam_w_part, dim3((y.K * y.N + 255) / 256), dim3(256), 0, m->s, y.w, m->part, nch, y.m1w,
                         y.m2w, y.K, y.N, adam->lr, adam->c1, adam->c2, y.W, y.Wt);
      hipLaunchKernelGGL(k_adam_b_part, dim3((y.N + 255) / 256), dim3(256), 0, m->s, y.b, m->partb, nch, y.m1b, y.m2b,
                         y.N, adam->lr, adam->c1, adam->c2);
    }
    if (l == 0) break;
    std::swap(cur, nxt);
    std::swap(curb, nxtb);
  }
}

}  // namespace usgmlp

extern "C" {

const char *usg_mlp_last_error(void) { return usgmlp::g_err.c_str(); }

int usg_mlp_sizes(const usg_mlp_desc *d, long *nw, long *nb) {
  if (!d || !nw || !nb) return fail(-1, "null argument");
  if (d->nin < 1 || d->nin > 128 || d->width < 64 || d->width > 256 || d->width % 64 || d->nlayer < 2 ||
      d->nout < 1 || d->nout > 16)
    return fail(-1, "network shape out of range (nin <= 128, width multiple of 64 <= 256, nout <= 16)");
  *nw = (long)d->nin * d->width + (long)(d->nlayer - 2) * d->width * d->width + (long)d->width * d->nout;
  *nb = (long)(d->nlayer - 1) * d->width + d->nout;
  return 0;
}

int usg_mlp_destroy(usg_mlp *m) {
  if (!m) return 0;
  if (m->s) hipStreamSynchronize(m->s);
  free_bufs(m);
  for (void *p : m->allocs) hipFree(p);
  if (m->xs) hipFree(m->xs);
  if (m->ys) hipFree(m->ys);
  if (m->idx) hipFree(m->idx);
  if (m->lossbuf) hipFree(m->lossbuf);
  if (m->s) hipStreamDestroy(m->s);
  delete m;
  return 0;
}

int usg_mlp_create(const usg_mlp_desc *d, usg_mlp **out) {
  long nw, nb;
  if (!out) return fail(-1, "null argument");
  *out = nullptr;
  int rc = usg_mlp_sizes(d, &nw, &nb);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(-2, "no GPU visible");
  usg_mlp *m = new usg_mlp();
  m->d = *d;
  if (hipStreamCreate(&m->s) != hipSuccess) {
    delete m;
    return fail(-3, "hipStreamCreate");
  }
  m->L.resize(d->nlayer);
  for (int l = 0; l < d->nlayer; l++) {
    Layer &y = m->L[l];
    y.k = l == 0 ? d->nin : d->width;
    y.n = l == d->nlayer - 1 ? d->nout : d->width;
    y.K = l == 0 ? round_up(d->nin, 64) : d->width;
    y.N = l == d->nlayer - 1 ? 32 : d->width;
    const size_t kn = (size_t)y.K * y.N;
    rc = 0;
    rc |= dalloc(m->allocs, &y.w, kn);
    rc |= dalloc(m->allocs, &y.gw, kn);
    rc |= dalloc(m->allocs, &y.m1w, kn);
    rc |= dalloc(m->allocs, &y.m2w, kn);
    rc |= dalloc(m->allocs, &y.b, y.N);
    rc |= dalloc(m->allocs, &y.gb, y.N);
    rc |= dalloc(m->allocs, &y.m1b, y.N);
    rc |= dalloc(m->allocs, &y.m2b, y.N);
    rc |= dalloc(m->allocs, &y.W, kn);
    rc |= dalloc(m->allocs, &y.Wt, kn);
    if (rc) {
      usg_mlp_destroy(m);
      return rc;
    }
  }
  *out = m;
  // zero parameters until set
  std::vector<float> w(nw, 0.0f), b(nb, 0.0f);
  return usg_mlp_set_params(m, w.data(), b.data());
}

int usg_mlp_set_params(usg_mlp *m, const float *w, const float *b) {
  if (!m || !w || !b) return fail(-1, "null argument");
  long ow = 0, ob = 0;
  for (Layer &y : m->L) {
    std::vector<float> pw((size_t)y.K * y.N, 0.0f), pb(y.N, 0.0f);
    std::vector<bf16> W((size_t)y.K * y.N), Wt((size_t)y.K * y.N);
    for (int k = 0; k < y.k; k++)
      for (int n = 0; n < y.n; n++) pw[(size_t)k * y.N + n] = w[ow + (long)k * y.n + n];
    for (int n = 0; n < y.n; n++) pb[n] = b[ob + n];
    ow += (long)y.k * y.n;
    ob += y.n;
    for (int k = 0; k < y.K; k++)
      for (int n = 0; n < y.N; n++) {
        const bf16 h = to_bf16(pw[(size_t)k * y.N + n]);
        W[(size_t)k * y.N + n] = h;
        Wt[(size_t)n * y.K + k] = h;
      }
    const size_t kn = (size_t)y.K * y.N;
    HIPCHK(hipMemcpyAsync(y.w, pw.data(), kn * 4, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(y.b, pb.data(), y.N * 4, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(y.W, W.data(), kn * 2, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemcpyAsync(y.Wt, Wt.data(), kn * 2, hipMemcpyHostToDevice, m->s));
    HIPCHK(hipMemsetAsync(y.m1w, 0, kn * 4, m->s));
    HIPCHK(hipMemsetAsync(y.m2w, 0, kn * 4, m->s));
    HIPCHK(hipMemsetAsync(y.m1b, 0, y.N * 4, m->s));
    HIPCHK(hipMemsetAsync(y.m2b, 0, y.N * 4, m->s));
  }
  m->step = 0;
  HIPCHK(hipStreamSynchronize(m->s));
  return 0;
}

int usg_mlp_get_params(usg_mlp *m, float *w, float *b) {
  if (!m || !w || !b) return fail(-1, "null argument");
  long ow = 0, ob = 0;
  for (Layer &y : m->L) {
    std::vector<float> pw((size_t)y.K * y.N), pb(y.N);
    HIPCHK(hipMemcpyAsync(pw.data(), y.w, pw.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipMemcpyAsync(pb.data(), y.b, pb.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    for (int k = 0; k < y.k; k++)
      for (int n = 0; n < y.n; n++) w[ow + (long)k * y.n + n] = pw[(size_t)k * y.N + n];
    for (int n = 0; n < y.n; n++) b[ob + n] = pb[n];
    ow += (long)y.k * y.n;
    ob += y.n;
  }
  return 0;
}

// Host batch -> device X (bf16, padded) and Y; returns padded P
static int upload_batch(usg_mlp *m, const float *x, const float *y, int n, int *Pout) {
  const int P = round_up(std::max(n, 1), 128);
  int rc = ensure_cap(m, P);
  if (rc) return rc;
  const int K0 = m->L[0].K, nin = m->d.nin, nout = m->d.nout;
  std::vector<bf16> X((size_t)P * K0, 0), Xb((size_t)P * K0, 0);
  for (int p = 0; p < n; p++)
    for (int k = 0; k < nin; k++) {
      const bf16 h = to_bf16(x[(size_t)p * nin + k]);
      X[(size_t)p * K0 + k] = h;
      Xb[((size_t)(p / 8) * K0 + k) * 8 + p % 8] = h;
    }
  HIPCHK(hipMemcpyAsync(m->X, X.data(), X.size() * 2, hipMemcpyHostToDevice, m->s));
  HIPCHK(hipMemcpyAsync(m->Xb, Xb.data(), Xb.size() * 2, hipMemcpyHostToDevice, m->s));
  if (y) {
    std::vector<float> Y((size_t)P * nout, 0.0f);
    std::memcpy(Y.data(), y, (size_t)n * nout * 4);
    HIPCHK(hipMemcpyAsync(m->Y, Y.data(), Y.size() * 4, hipMemcpyHostToDevice, m->s));
  }
  *Pout = P;
  return 0;
}

int usg_mlp_predict(usg_mlp *m, const float *x, long n, float *y) {
  if (!m || !x || !y || n < 0) return fail(-1, "bad argument");
  const int step = 65536;
  const int No = m->L.back().N, nout = m->d.nout;
  std::vector<float> o;
  for (long a = 0; a < n; a += step) {
    const int cnt = (int)std::min<long>(step, n - a);
    int P = 0;
    int rc = upload_batch(m, x + a * m->d.nin, nullptr, cnt, &P);
    if (rc) return rc;
    forward(m, P, false);
    HIPCHK(hipGetLastError());
    o.resize((size_t)P * No);
    HIPCHK(hipMemcpyAsync(o.data(), m->Z.back(), o.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    for (int p = 0; p < cnt; p++)
      for (int j = 0; j < nout; j++) y[(a + p) * nout + j] = o[(size_t)p * No + j];
  }
  return 0;
}

// loss + output gradient + backward of the batch in X/Y (nv valid of P);
// with lossbuf, the batch loss goes to lossbuf[slot] on the device (no sync)
static int loss_and_grads(usg_mlp *m, int P, int nv, float *loss, float *lossbuf = nullptr, int slot = 0,
                          const AdamStep *adam = nullptr) {
  forward(m, P, true);
  const int No = m->L.back().N;
  const int nbl = P * No / 256;
  hipLaunchKernelGGL(k_loss, dim3(nbl), dim3(256), 0, m->s, m->Z.back(), No, m->Y, m->d.nout, nv, m->dA, m->dAb,
                     m->lossp);
  if (lossbuf)
    hipLaunchKernelGGL(k_sum_small, dim3(1), dim3(256), 0, m->s, m->lossp, nbl,
                       (float)(1.0 / ((double)nv * m->d.nout)), lossbuf, slot);
  backward(m, P, adam);
  HIPCHK(hipGetLastError());
  if (loss) {
    std::vector<float> lp(nbl);
    HIPCHK(hipMemcpyAsync(lp.data(), m->lossp, nbl * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    double s = 0;
    for (float v : lp) s += v;
    *loss = (float)(s / ((double)nv * m->d.nout));
  }
  return 0;
}

int usg_mlp_gradients(usg_mlp *m, const float *x, const float *y, int n, float *loss, float *dw, float *db) {
  if (!m || !x || !y || n <= 0) return fail(-1, "bad argument");
  int P = 0;
  int rc = upload_batch(m, x, y, n, &P);
  if (rc) return rc;
  rc = loss_and_grads(m, P, n, loss);
  if (rc) return rc;
  long ow = 0, ob = 0;
  for (Layer &L : m->L) {
    std::vector<float> g((size_t)L.K * L.N), gb(L.N);
    HIPCHK(hipMemcpyAsync(g.data(), L.gw, g.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipMemcpyAsync(gb.data(), L.gb, gb.size() * 4, hipMemcpyDeviceToHost, m->s));
    HIPCHK(hipStreamSynchronize(m->s));
    if (dw)
      for (int k = 0; k < L.k; k++)
        for (int j = 0; j < L.n; j++) dw[ow + (long)k * L.n + j] = g[(size_t)k * L.N + j];
    if (db)
      for (int j = 0; j < L.n; j++) db[ob + j] = gb[j];
    ow += (long)L.k * L.n;
    ob += L.n;
  }
  return 0;
}

int usg_mlp_set_data(usg_mlp *m, const float *x, const float *y, long n) {
  if (!m || !x || !y || n <= 0) return fail(-1, "bad argument");
  // the old set is gone from here on: a failure below leaves no data set
  m->nset = 0;
  if (m->xs || m->ys) {
    HIPCHK(hipStreamSynchronize(m->s));
    if (m->xs) hipFree(m->xs);
    if (m->ys) hipFree(m->ys);
    m->xs = nullptr;
    m->ys = nullptr;
  }
  const int K0 = m->L[0].K, nin = m->d.nin, nout = m->d.nout;
  std::vector<bf16> X((size_t)n * K0, 0);
  for (long p = 0; p < n; p++)
    for (int k = 0; k < nin; k++) X[(size_t)p * K0 + k] = to_bf16(x[(size_t)p * nin + k]);
  HIPCHK(hipMalloc(&m->xs, X.size() * 2));
  HIPCHK(hipMalloc(&m->ys, (size_t)n * nout * 4));
  HIPCHK(hipMemcpyAsync(m->xs, X.data(), X.size() * 2, hipMemcpyHostToDevice, m->s));
  HIPCHK(hipMemcpyAsync(m->ys, y, (size_t)n * nout * 4, hipMemcpyHostToDevice, m->s));
  HIPCHK(hipStreamSynchronize(m->s));
  m->nset = n;
  return 0;
}

int usg_mlp_train_epoch(usg_mlp *m, const int *perm, long n, int batch, float lr, float *loss_out) {
  if (!m || !perm || n <= 0 || batch <= 0) return fail(-1, "bad argument");
  if (n != m->nset) return fail(-1, "perm length differs from the training set");
  for (long i = 0; i < n; i++)
    if (perm[i] < 0 || perm[i] >= n) return fail(-1, "perm entry out of range");
  const int P = round_up(batch, 128);
  int rc = ensure_cap(m, P);
  if (rc) return rc;
  if (m->idx_cap < n) {
    if (m->idx) hipFree(m->idx);
    m->idx = nullptr;
    HIPCHK(hipMalloc(&m->idx, n * sizeof(int)));
    m->idx_cap = n;
  }
  HIPCHK(hipMemcpyAsync(m->idx, perm, n * sizeof(int), hipMemcpyHostToDevice, m->s));
  const int K0 = m->L[0].K;
  const long nsteps = (n + batch - 1) / batch;
  if (m->lossbuf_cap < nsteps) {
    if (m->lossbuf) hipFree(m->lossbuf);
    m->lossbuf = nullptr;
    HIPCHK(hipMalloc(&m->lossbuf, nsteps * sizeof(float)));
    m->lossbuf_cap = nsteps;
  }
  long nb = 0;
  for (long a = 0; a < n; a += batch) {
    const int nv = (int)std::min<long>(batch, n - a);
    hipLaunchKernelGGL(k_gather, dim3((unsigned)(((long)P * K0 + 255) / 256)), dim3(256), 0, m->s, m->xs, K0, m->ys,
                       m->d.nout, m->idx + a, nv, P, m->X, m->Xb, m->Y);
    m->step++;
    const AdamStep st{lr, 1.0f - std::pow(0.9f, (float)m->step), 1.0f - std::pow(0.999f, (float)m->step)};
    rc = loss_and_grads(m, P, nv, nullptr, m->lossbuf, (int)nb, &st);
    if (rc) return rc;
    nb++;
  }
  HIPCHK(hipGetLastError());
  std::vector<float> lb(nb);
  HIPCHK(hipMemcpyAsync(lb.data(), m->lossbuf, nb * sizeof(float), hipMemcpyDeviceToHost, m->s));
  HIPCHK(hipStreamSynchronize(m->s));
  double lsum = 0;
  for (float v : lb) lsum += v;
  if (loss_out) *loss_out = (float)(lsum / std::max<long>(nb, 1));
  return 0;
}

int usg_mlp_mfma_check(const float *a, const float *b, float *d) {
  if (!a || !b || !d) return fail(-1, "null argument");
  // device buffers released on every return
  struct Bufs {
    void *p[3] = {nullptr, nullptr, nullptr};
    ~Bufs() {
      for (void *q : p)
        if (q) (void)hipFree(q);
    }
  } bufs;
  std::vector<bf16> ha(16 * 32), hb(32 * 16);
  for (int i = 0; i < 16 * 32; i++) {
    ha[i] = to_bf16(a[i]);
    hb[i] = to_bf16(b[i]);
  }
  HIPCHK(hipMalloc(&bufs.p[0], ha.size() * 2));
  HIPCHK(hipMalloc(&bufs.p[1], hb.size() * 2));
  HIPCHK(hipMalloc(&bufs.p[2], 256 * 4));
  bf16 *da = (bf16 *)bufs.p[0], *db = (bf16 *)bufs.p[1];
  float *dd = (float *)bufs.p[2];
  HIPCHK(hipMemcpy(da, ha.data(), ha.size() * 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(db, hb.data(), hb.size() * 2, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_mfma_check, dim3(1), dim3(64), 0, 0, da, db, dd);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(d, dd, 256 * 4, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
