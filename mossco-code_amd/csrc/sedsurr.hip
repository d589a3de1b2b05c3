// sedsurr.hip -- the spin-up surrogate (include/sedsurr.h) on gfx950.
//
// A multilayer perceptron trained and evaluated with hand-written kernels:
//   k_gemm     bf16 x bf16 -> fp32 on the MFMA units (v_mfma_f32_32x32x16
//              bf16), 64 x 64 block tiles, four waves of 32 x 32, operands
//              staged through LDS; the three products of training (forward
//              H W, data gradient dZ W^T, weight gradient H^T dZ) are the
//              transpose variants of one kernel, with fused epilogues (bias
//              + GELU storing the pre-activation, bias only, GELU'
//              product, plain fp32)
//   k_loss     squared error, its gradient in bf16, per-block loss sums
//   k_sum_slices  the split-k weight gradients, summed in a fixed order
//              (for the gradient query; in training k_adam sums them);
//              the bias gradients (column sums of dZ) come out of the
//              weight-gradient product as the row of a ones column
//              appended to its left operand
//   k_adam     Adam on the fp32 master parameters, writing the bf16 copy
//   k_gather   minibatch rows of the resident training set, bf16 inputs
// Layer widths are padded on the device (inputs to 32, outputs to a
// multiple of 64) with zero weights, which keeps the products exact and the
// padded parameters at zero through training.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sedcol.h"
#include "sedsurr.h"

namespace ss {

using bf16 = unsigned short;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// fp32 -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16 to_bf16(float f)
{
  unsigned u = __builtin_bit_cast(unsigned, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float from_bf16(bf16 b) { return __builtin_bit_cast(float, (unsigned)b << 16); }

// GELU, tanh form: 0.5 z (1 + tanh(c (z + a z^3))), c = sqrt(2/pi), a = 0.044715
constexpr float kGc = 0.7978845608028654f, kGa = 0.044715f;
__device__ inline float tanh_e(float u) { return 1.0f - 2.0f / (expf(2.0f * u) + 1.0f); }
__device__ inline float gelu(float z)
{
  const float t = tanh_e(kGc * (z + kGa * z * z * z));
  return 0.5f * z * (1.0f + t);
}
__device__ inline float gelu_grad(float z)
{
  const float t = tanh_e(kGc * (z + kGa * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGc * (1.0f + 3.0f * kGa * z * z);
}

constexpr int BM = 64, BN = 64, BK = 64;
enum { EPI_F32 = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_GRAD = 3 };

struct Epi {
  float *out_f;       // fp32 output (EPI_F32, EPI_BIAS; pre-activation for EPI_BIAS_GELU)
  int ldo;
  long long zstride;  // EPI_F32 with split k: distance between the slices' outputs
  bf16 *out_h;        // bf16 output (EPI_BIAS_GELU activation, EPI_GELU_GRAD dZ)
  int ldh;
  const float *bias;  // EPI_BIAS, EPI_BIAS_GELU
  const float *zin;   // EPI_GELU_GRAD: pre-activation of the layer below
  int ldz;
};

// eight consecutive bf16 (16 bytes) of a row-major array, zero outside
// [0, lim) of the fastest index; one 16-byte load when all eight are inside
// and the address is aligned
__device__ inline uint4 load8(const bf16 *__restrict__ p, int i0, int lim, bool row_ok)
{
  if (row_ok && i0 + 7 < lim && ((reinterpret_cast<uintptr_t>(p + i0) & 15u) == 0))
    return *reinterpret_cast<const uint4 *>(p + i0);
  unsigned w[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const unsigned lo = (row_ok && i0 + 2 * j < lim) ? p[i0 + 2 * j] : 0u;
    const unsigned hi = (row_ok && i0 + 2 * j + 1 < lim) ? p[i0 + 2 * j + 1] : 0u;
    w[j] = lo | (hi << 16);
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ inline bf16 elem8(const uint4 &w, int j)
{
  const unsigned u = j < 2 ? w.x : j < 4 ? w.y : j < 6 ? w.z : w.w;
  return (bf16)((j & 1) ? (u >> 16) : (u & 0xffffu));
}

// C (M x N) = op(A) (M x K) op(B) (K x N); op(A)(m,k) = TA ? A[k lda + m] :
// A[m lda + k], op(B)(k,n) = TB ? B[n ldb + k] : B[k ldb + n].  Out-of-range
// rows, columns and k are loaded as zero and not stored.  Block z of the
// grid takes k in [z kchunk, (z+1) kchunk) and (EPI_F32) writes its partial
// product at out_f + z zstride.
//
// Both tiles sit in LDS with k contiguous (A as [m][k], B as [n][k]), so a
// lane's MFMA operand (8 consecutive k of one row or column) is one 16-byte
// LDS read.  v_mfma_f32_32x32x16_bf16: lane l supplies row (A) / column (B)
// l % 32 at k = 8 (l / 32) + 0..7.  The k loop is software pipelined: the
// next 64-wide k tile is loaded from global memory into registers while the
// MFMAs run on the current one, and the LDS tiles are double buffered, so
// one barrier per k tile suffices.  The MFMAs accumulate k in increasing
// order of 16-wide steps, as for any tile width.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, const bf16 *__restrict__ A, int lda,
                                              const bf16 *__restrict__ B, int ldb, Epi e, int ones_row)
{
  constexpr int LDK = BK + 8;  // row pitch of the tiles: 144 bytes, 16-byte aligned rows
  __shared__ __attribute__((aligned(16))) bf16 sA[2][BM][LDK];
  __shared__ __attribute__((aligned(16))) bf16 sB[2][BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bm = blockIdx.y * BM, bn = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int kb0 = blockIdx.z * kchunk;
  const int kend = min(K, kb0 + kchunk);
  f32x16 acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
  const int r = lane & 31, q8 = (lane >> 5) * 8;

  // this thread's share of a 64 x 64 tile of each operand: eight
  // consecutive elements at idx = 8 tid + 2048 h, h = 0, 1
  uint4 ra[2], rb[2];
  // Operands stored k-major (TA for A, !TB for B) are transposed into the
  // k-contiguous LDS tiles: a thread loads the same eight m (or n) at two
  // consecutive k (k even), so every tile element pair (m, k..k+1) is one
  // 4-byte LDS write.
  const int tm = (tid % 8) * 8, tk = 2 * (tid / 8);
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int idx = tid * 8 + 2048 * h;
      if (TA) {  // A stored k-major: m fastest
        const int m = tm, k = tk + h;
        // row ones_row of op(A) is all ones (read nothing there): the
        // product's row ones_row is then the column sums of op(B)
        const int lim = ones_row >= 0 ? ones_row : M;
        ra[h] = load8(A + (size_t)(k0 + k) * lda, bm + m, lim, k0 + k < kend);
        const int j1 = ones_row - (bm + m);
        if (j1 >= 0 && j1 < 8 && k0 + k < kend) {
          const unsigned one = 0x3F80u;  // bf16 1.0
          if (j1 == 0) ra[h].x = (ra[h].x & 0xffff0000u) | one;
          else if (j1 == 1) ra[h].x = (ra[h].x & 0xffffu) | (one << 16);
          else if (j1 == 2) ra[h].y = (ra[h].y & 0xffff0000u) | one;
          else if (j1 == 3) ra[h].y = (ra[h].y & 0xffffu) | (one << 16);
          else if (j1 == 4) ra[h].z = (ra[h].z & 0xffff0000u) | one;
          else if (j1 == 5) ra[h].z = (ra[h].z & 0xffffu) | (one << 16);
          else if (j1 == 6) ra[h].w = (ra[h].w & 0xffff0000u) | one;
          else ra[h].w = (ra[h].w & 0xffffu) | (one << 16);
        }
      } else {
        const int m = idx / BK, k = idx % BK;
        ra[h] = load8(A + (size_t)(bm + m) * lda, k0 + k, kend, bm + m < M);
      }
      if (TB) {  // B stored n-major: k fastest
        const int n = idx / BK, kb = idx % BK;
        rb[h] = load8(B + (size_t)(bn + n) * ldb, k0 + kb, kend, bn + n < N);
      } else {
        const int kb = tk + h, n = tm;
        rb[h] = load8(B + (size_t)(k0 + kb) * ldb, bn + n, N, k0 + kb < kend);
      }
    }
  };
  auto pair = [](const uint4 &lo, const uint4 &hi, int j) {
    return (unsigned)elem8(lo, j) | ((unsigned)elem8(hi, j) << 16);
  };
  auto stage = [&](int buf) {
    if (TA) {
#pragma unroll
      for (int j = 0; j < 8; ++j) *reinterpret_cast<unsigned *>(&sA[buf][tm + j][tk]) = pair(ra[0], ra[1], j);
    } else {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int idx = tid * 8 + 2048 * h;
        const int m = idx / BK, k = idx % BK;
        *reinterpret_cast<uint4 *>(&sA[buf][m][k]) = ra[h];
      }
    }
    if (TB) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int idx = tid * 8 + 2048 * h;
        const int n = idx / BK, kb = idx % BK;
        *reinterpret_cast<uint4 *>(&sB[buf][n][kb]) = rb[h];
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) *reinterpret_cast<unsigned *>(&sB[buf][tm + j][tk]) = pair(rb[0], rb[1], j);
    }
  };

  int buf = 0;
  if (kb0 < kend) fetch(kb0);
  for (int k0 = kb0; k0 < kend; k0 += BK) {
    stage(buf);
    __syncthreads();
    if (k0 + BK < kend) fetch(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sA[buf][wm + r][kk + q8]);
      const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sB[buf][wn + r][kk + q8]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    buf ^= 1;
  }
  // accumulator layout of the 32 x 32 result: lane holds column (lane & 31),
  // register v holds row 8 (v / 4) + 4 (lane / 32) + v % 4
  const int col = bn + wn + r;
  if (col >= N) return;
  const int q = (lane >> 5) * 4;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = bm + wm + 8 * (v >> 2) + q + (v & 3);
    if (row >= M) continue;
    const float x = acc[v];
    if constexpr (EPI == EPI_F32) {
      e.out_f[blockIdx.z * e.zstride + (size_t)row * e.ldo + col] = x;
    } else if constexpr (EPI == EPI_BIAS) {
      e.out_f[(size_t)row * e.ldo + col] = x + e.bias[col];
    } else if constexpr (EPI == EPI_BIAS_GELU) {
      const float z = x + e.bias[col];
      e.out_f[(size_t)row * e.ldo + col] = z;
      e.out_h[(size_t)row * e.ldh + col] = to_bf16(gelu(z));
    } else {
      const float z = e.zin[(size_t)row * e.ldz + col];
      e.out_h[(size_t)row * e.ldh + col] = to_bf16(x * gelu_grad(z));
    }
  }
}

// out[i] = sum over s < S of part[s stride + i], in order of s
__global__ void k_sum_slices(const float *part, int S, long long stride, long long n, float *out)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float a = 0.0f;
    for (int sl = 0; sl < S; ++sl) a += part[sl * stride + i];
    out[i] = a;
  }
}

// rows of the fp32 sets (or of X itself when rows is null) into padded
// batch buffers: Xb bf16 (b x kin, zero beyond n_in), Yb fp32 (b x nout_p)
__global__ void k_gather(const float *X, int n_in, const float *Y, int n_out, const int *rows, int b, int kin,
                         int nout_p, bf16 *Xb, float *Yb)
{
  const int i = blockIdx.x;
  if (i >= b) return;
  const long long src = rows ? rows[i] : i;
  for (int j = threadIdx.x; j < kin; j += blockDim.x)
    Xb[(size_t)i * kin + j] = j < n_in ? to_bf16(X[src * n_in + j]) : (bf16)0;
  if (Yb)
    for (int j = threadIdx.x; j < nout_p; j += blockDim.x)
      Yb[(size_t)i * nout_p + j] = j < n_out ? Y[src * n_out + j] : 0.0f;
}

// squared error over the first n_out columns; dZ = bf16(scale (yh - y));
// one partial sum of squared errors per block
__global__ __launch_bounds__(256) void k_loss(const float *Yh, const float *Y, int b, int n_out, int ld, float scale,
                                              bf16 *dZ, float *part)
{
  __shared__ float red[256];
  float s = 0.0f;
  const long long total = (long long)b * ld;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int col = (int)(i % ld);
    float g = 0.0f;
    if (col < n_out) {
      const float d = Yh[i] - Y[i];
      s += d * d;
      g = scale * d;
    }
    dZ[i] = to_bf16(g);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[0] = sum(part[0..n-1]) * scale (one block)
__global__ __launch_bounds__(256) void k_sum(const float *part, int n, float scale, float *out)
{
  __shared__ float red[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// Adam on the flat parameter vector; wb receives the bf16 copy
// Split-k slices of the weight gradients that Adam sums itself: layer l's
// parameters [woff, woff + welems) take their gradient from nz slices at
// part + pofs (slice z at + z welems) when nz > 1, else from g
constexpr int kMaxLayers = 17;  // n_layers <= 16 hidden layers + the output layer
struct Slices {
  int nl;
  long long woff[kMaxLayers], welems[kMaxLayers], pofs[kMaxLayers];
  int nz[kMaxLayers];
};

__global__ void k_adam(float *w, const float *g, float *m, float *v, bf16 *wb, long long n, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, const float *part, Slices sl)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float gi = g[i];
    for (int l = 0; l < sl.nl; ++l) {
      const long long j = i - sl.woff[l];
      if (j >= 0 && j < sl.welems[l] && sl.nz[l] > 1) {
        // the slices in order, as k_sum_slices adds them
        float a = 0.0f;
        for (int z = 0; z < sl.nz[l]; ++z) a += part[sl.pofs[l] + z * sl.welems[l] + j];
        gi = a;
      }
    }
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    wb[i] = to_bf16(wi);
  }
}

__global__ void k_to_bf16(const float *w, bf16 *wb, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    wb[i] = to_bf16(w[i]);
}

}  // namespace ss

using ss::bf16;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}

#define HIPOK(call)                                                                  \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(SEDCOL_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

}  // namespace

// Makes the handle's device current for the duration of one entry point and
// restores the caller's current device afterwards: a handle belongs to the
// device that was current at sedsurr_create.
struct DeviceScope {
  int prev = -1, dev;
  explicit DeviceScope(int d) : dev(d)
  {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceScope()
  {
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
  }
  DeviceScope(const DeviceScope &) = delete;
  DeviceScope &operator=(const DeviceScope &) = delete;
};

struct sedsurr_handle {
  int device = 0;  // current at sedsurr_create; every call runs on it
  sedsurr_shape s{};
  int kin = 0, nout_p = 0, nl = 0;  // padded input width, padded output width, weight layers
  std::vector<int> fi, fo, fip, fop;    // logical and padded fan-in / fan-out per layer
  std::vector<long long> woff, boff;    // offsets of W_l and b_l in the padded flat vector
  long long np_pad = 0, np = 0;         // padded and logical parameter counts
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  bf16 *wb = nullptr;
  bf16 *xb = nullptr;                   // max_batch x kin
  std::vector<bf16 *> hs;               // hidden activations, max_batch x n_hidden
  std::vector<float *> zs;              // pre-activations
  float *yh = nullptr, *yb = nullptr;   // max_batch x nout_p
  bf16 *dz0 = nullptr, *dz1 = nullptr;  // max_batch x max(n_hidden, nout_p)
  float *part = nullptr;                // loss partial sums
  float *kpart = nullptr;               // split-k partial weight gradients
  float *kpart_all = nullptr;           // the same, one region per layer (training: summed by Adam)
  std::vector<long long> kpofs;         // offset of layer l's region in kpart_all
  ss::Slices slices{};                  // the last training step's deferred slices
  float *loss = nullptr;                // per-step losses of a train call
  int loss_cap = 0;
  float *X = nullptr, *Y = nullptr;     // training set
  long long ndata = 0;
  int *rows = nullptr;
  long long rows_cap = 0;
  float *xin = nullptr, *yin = nullptr; // staging of loss_grad / predict inputs
  float *ydev = nullptr;                // sedsurr_predict_device output (ydev_rows x n_out)
  int64_t ydev_rows = 0;
  long long step = 0;                   // Adam steps taken
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0.0;
  int last_launches = 0;
};

namespace {

constexpr int kLossBlocks = 256;

constexpr int kMaxSplit = 8;     // k slices of a weight-gradient product

template <bool TA, bool TB, int EPI>
void gemm(sedsurr_handle *h, int M, int N, int K, const bf16 *A, int lda, const bf16 *B, int ldb, ss::Epi e,
          int *launches, int split = 1, int ones_row = -1)
{
  const int kchunk = (K + split - 1) / split;
  const int kc = (kchunk + ss::BK - 1) / ss::BK * ss::BK;
  const int nz = (K + kc - 1) / kc;
  dim3 grid((N + ss::BN - 1) / ss::BN, (M + ss::BM - 1) / ss::BM, nz > 0 ? nz : 1);
  hipLaunchKernelGGL((ss::k_gemm<TA, TB, EPI>), grid, dim3(256), 0, h->st, M, N, K, kc, A, lda, B, ldb, e,
                     ones_row);
  ++*launches;
}

// k slices of the weight-gradient product of a batch of b rows: enough to
// give the device about 512 blocks, each slice at least 256 rows
int split_for(int M, int N, int b)
{
  const int tiles = ((M + ss::BM - 1) / ss::BM) * ((N + ss::BN - 1) / ss::BN);
  int s = 1;
  while (s < kMaxSplit && tiles * s < 512 && b / (2 * s) >= 256) s *= 2;
  return s;
}

// width of the activation entering layer l (padded)
inline int in_width(const sedsurr_handle *h, int l) { return h->fip[l]; }

// forward pass of b rows from h->xb; output h->yh (b x nout_p), or, with
// out, the n_out real columns only, compact (b x n_out)
void forward(sedsurr_handle *h, int b, int *launches, float *out = nullptr)
{
  const int L = h->nl;
  for (int l = 0; l < L; ++l) {
    const bf16 *A = l == 0 ? h->xb : h->hs[l - 1];
    const int lda = in_width(h, l);
    ss::Epi e{};
    if (l < L - 1) {
      e.out_f = h->zs[l];
      e.ldo = h->fop[l];
      e.out_h = h->hs[l];
      e.ldh = h->fop[l];
      e.bias = h->w + h->boff[l];
      gemm<false, false, ss::EPI_BIAS_GELU>(h, b, h->fop[l], h->fip[l], A, lda, h->wb + h->woff[l], h->fop[l], e,
                                            launches);
    } else {
      e.out_f = out ? out : h->yh;
      e.ldo = out ? h->s.n_out : h->nout_p;
      e.bias = h->w + h->boff[l];
      gemm<false, false, ss::EPI_BIAS>(h, b, out ? h->s.n_out : h->fop[l], h->fip[l], A, lda, h->wb + h->woff[l],
                                       h->fop[l], e, launches);
    }
  }
}

// loss of the forward output against h->yb, gradient into h->g; loss value
// to *loss_dev.  defer: split weight gradients stay as slices in
// h->kpart_all (described in h->slices) for k_adam to sum
void loss_backward(sedsurr_handle *h, int b, float *loss_dev, int *launches, bool defer = false)
{
  const int L = h->nl;
  const float scale = 2.0f / ((float)b * (float)h->s.n_out);
  hipLaunchKernelGGL(ss::k_loss, dim3(kLossBlocks), dim3(256), 0, h->st, h->yh, h->yb, b, h->s.n_out, h->nout_p,
                     scale, h->dz0, h->part);
  hipLaunchKernelGGL(ss::k_sum, dim3(1), dim3(256), 0, h->st, h->part, kLossBlocks,
                     1.0f / ((float)b * (float)h->s.n_out), loss_dev);
  *launches += 2;
  bf16 *dz = h->dz0, *dzn = h->dz1;
  for (int l = L - 1; l >= 0; --l) {
    const bf16 *Hin = l == 0 ? h->xb : h->hs[l - 1];
    const int win = in_width(h, l), wout = h->fop[l];
    // [dW_l; db_l] = [H_in, 1]^T dZ_l in one product: the bias gradient
    // (the column sums of dZ_l) is the row of a ones column appended to
    // H_in, and b_l follows W_l in the parameter vector, so that row lands
    // on it; split over the batch into slices summed in order
    const int split = split_for(win + 1, wout, b);
    const long long welems = (long long)(win + 1) * wout;
    const int kc = ((b + split - 1) / split + ss::BK - 1) / ss::BK * ss::BK;
    const int nz = (b + kc - 1) / kc;
    float *part = defer ? h->kpart_all + h->kpofs[l] : h->kpart;
    ss::Epi e{};
    e.out_f = nz > 1 ? part : h->g + h->woff[l];
    e.ldo = wout;
    e.zstride = welems;
    gemm<true, false, ss::EPI_F32>(h, win + 1, wout, b, Hin, win, dz, wout, e, launches, split, win);
    h->slices.woff[l] = h->woff[l];
    h->slices.welems[l] = welems;
    h->slices.pofs[l] = h->kpofs[l];
    h->slices.nz[l] = defer ? nz : 1;
    if (nz > 1 && !defer) {
      hipLaunchKernelGGL(ss::k_sum_slices, dim3(512), dim3(256), 0, h->st, h->kpart, nz, welems, welems,
                         h->g + h->woff[l]);
      ++*launches;
    }
    if (l > 0) {
      // dZ_{l-1} = GELU'(Z_{l-1}) * (dZ_l W_l^T)
      ss::Epi d{};
      d.out_h = dzn;
      d.ldh = win;
      d.zin = h->zs[l - 1];
      d.ldz = win;
      gemm<false, true, ss::EPI_GELU_GRAD>(h, b, win, wout, dz, wout, h->wb + h->woff[l], wout, d, launches);
      std::swap(dz, dzn);
    }
  }
}

// host logical flat vector <-> padded device layout
void pad_params(const sedsurr_handle *h, const float *src, std::vector<float> &dst)
{
  dst.assign(h->np_pad, 0.0f);
  long long o = 0;
  for (int l = 0; l < h->nl; ++l) {
    for (int i = 0; i < h->fi[l]; ++i)
      for (int j = 0; j < h->fo[l]; ++j) dst[h->woff[l] + (long long)i * h->fop[l] + j] = src[o++];
    for (int j = 0; j < h->fo[l]; ++j) dst[h->boff[l] + j] = src[o++];
  }
}

void unpad_params(const sedsurr_handle *h, const std::vector<float> &src, float *dst)
{
  long long o = 0;
  for (int l = 0; l < h->nl; ++l) {
    for (int i = 0; i < h->fi[l]; ++i)
      for (int j = 0; j < h->fo[l]; ++j) dst[o++] = src[h->woff[l] + (long long)i * h->fop[l] + j];
    for (int j = 0; j < h->fo[l]; ++j) dst[o++] = src[h->boff[l] + j];
  }
}

}  // namespace

extern "C" {

const char *sedsurr_last_error(void) { return g_err.c_str(); }

int sedsurr_create(const sedsurr_shape *shape, sedsurr_handle **out)
{
  if (!shape || !out) return fail(SEDCOL_EINVAL, "null argument");
  *out = nullptr;
  const sedsurr_shape s = *shape;
  if (s.n_in < 1 || s.n_in > 64) return fail(SEDCOL_EINVAL, "n_in outside 1..64");
  if (s.n_hidden < 64 || s.n_hidden % 64 || s.n_hidden > 4096)
    return fail(SEDCOL_EINVAL, "n_hidden must be a multiple of 64 in 64..4096");
  if (s.n_layers < 1 || s.n_layers > 16) return fail(SEDCOL_EINVAL, "n_layers outside 1..16");
  if (s.n_out < 1 || s.n_out > 4096) return fail(SEDCOL_EINVAL, "n_out outside 1..4096");
  if (s.max_batch < 64 || s.max_batch % 64 || s.max_batch > 65536)
    return fail(SEDCOL_EINVAL, "max_batch must be a multiple of 64 in 64..65536");
  if (s.reserved != 0) return fail(SEDCOL_EINVAL, "reserved must be 0");
  sedsurr_handle *h = new sedsurr_handle();
  h->s = s;
  h->kin = round_up(s.n_in, 32);
  h->nout_p = round_up(s.n_out, 64);
  h->nl = s.n_layers + 1;
  long long off = 0;
  for (int l = 0; l < h->nl; ++l) {
    const int fi = l == 0 ? s.n_in : s.n_hidden, fo = l == h->nl - 1 ? s.n_out : s.n_hidden;
    const int fip = l == 0 ? h->kin : s.n_hidden, fop = l == h->nl - 1 ? h->nout_p : s.n_hidden;
    h->fi.push_back(fi);
    h->fo.push_back(fo);
    h->fip.push_back(fip);
    h->fop.push_back(fop);
    h->woff.push_back(off);
    off += (long long)fip * fop;
    h->boff.push_back(off);
    off += fop;
    h->np += (long long)fi * fo + fo;
  }
  h->np_pad = off;
  auto cleanup = [&](int code) {
    sedsurr_destroy(h);
    return code;
  };
#define HIPOK_C(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return cleanup(fail(SEDCOL_EHIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
  } while (0)
  HIPOK_C(hipGetDevice(&h->device));
  HIPOK_C(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  HIPOK_C(hipEventCreate(&h->e0));
  HIPOK_C(hipEventCreate(&h->e1));
  const size_t P = (size_t)h->np_pad, B = (size_t)s.max_batch;
  for (float **q : {&h->w, &h->g, &h->m, &h->v}) {
    HIPOK_C(hipMalloc(q, sizeof(float) * P));
    HIPOK_C(hipMemset(*q, 0, sizeof(float) * P));
  }
  HIPOK_C(hipMalloc(&h->wb, sizeof(bf16) * P));
  HIPOK_C(hipMemset(h->wb, 0, sizeof(bf16) * P));
  HIPOK_C(hipMalloc(&h->xb, sizeof(bf16) * B * h->kin));
  for (int l = 0; l < s.n_layers; ++l) {
    bf16 *hb = nullptr;
    float *zb = nullptr;
    HIPOK_C(hipMalloc(&hb, sizeof(bf16) * B * s.n_hidden));
    h->hs.push_back(hb);
    HIPOK_C(hipMalloc(&zb, sizeof(float) * B * s.n_hidden));
    h->zs.push_back(zb);
  }
  const size_t wmax = (size_t)std::max(s.n_hidden, h->nout_p);
  HIPOK_C(hipMalloc(&h->yh, sizeof(float) * B * h->nout_p));
  HIPOK_C(hipMalloc(&h->yb, sizeof(float) * B * h->nout_p));
  HIPOK_C(hipMalloc(&h->dz0, sizeof(bf16) * B * wmax));
  HIPOK_C(hipMalloc(&h->dz1, sizeof(bf16) * B * wmax));
  HIPOK_C(hipMalloc(&h->part, sizeof(float) * (kLossBlocks + 1)));  // + the loss of sedsurr_loss_grad
  long long wmax_el = 0;
  for (int l = 0; l < h->nl; ++l) wmax_el = std::max(wmax_el, (long long)(h->fip[l] + 1) * h->fop[l]);
  HIPOK_C(hipMalloc(&h->kpart, sizeof(float) * kMaxSplit * wmax_el));
  if (h->nl > ss::kMaxLayers) return cleanup(fail(SEDCOL_EINVAL, "too many layers"));
  long long kall = 0;
  for (int l = 0; l < h->nl; ++l) {
    h->kpofs.push_back(kall);
    kall += (long long)kMaxSplit * (h->fip[l] + 1) * h->fop[l];
  }
  HIPOK_C(hipMalloc(&h->kpart_all, sizeof(float) * kall));
  h->slices.nl = h->nl;
  HIPOK_C(hipMalloc(&h->xin, sizeof(float) * B * s.n_in));
  HIPOK_C(hipMalloc(&h->yin, sizeof(float) * B * s.n_out));
#undef HIPOK_C
  *out = h;
  return SEDCOL_OK;
}

int sedsurr_destroy(sedsurr_handle *h)
{
  if (!h) return SEDCOL_OK;
  const DeviceScope dscope(h->device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  std::vector<void *> ptrs = {h->w, h->g, h->m, h->v, h->wb, h->xb, h->yh, h->yb, h->dz0, h->dz1,
                              h->part, h->loss, h->X, h->Y, h->rows, h->xin, h->yin, h->kpart, h->kpart_all,
                              h->ydev};
  for (auto *q : h->hs) ptrs.push_back(q);
  for (auto *q : h->zs) ptrs.push_back(q);
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return SEDCOL_OK;
}

int sedsurr_num_params(sedsurr_handle *h, int64_t *n)
{
  if (!h || !n) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  *n = h->np;
  return SEDCOL_OK;
}

int sedsurr_set_params(sedsurr_handle *h, const float *params)
{
  if (!h || !params) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  for (long long i = 0; i < h->np; ++i)
    if (!std::isfinite(params[i])) return fail(SEDCOL_EDOMAIN, "non-finite parameter");
  std::vector<float> buf;
  pad_params(h, params, buf);
  const size_t P = (size_t)h->np_pad;
  HIPOK(hipMemcpyAsync(h->w, buf.data(), sizeof(float) * P, hipMemcpyHostToDevice, h->st));
  HIPOK(hipMemsetAsync(h->m, 0, sizeof(float) * P, h->st));
  HIPOK(hipMemsetAsync(h->v, 0, sizeof(float) * P, h->st));
  hipLaunchKernelGGL(ss::k_to_bf16, dim3(1024), dim3(256), 0, h->st, h->w, h->wb, (long long)P);
  HIPOK(hipGetLastError());
  HIPOK(hipStreamSynchronize(h->st));
  h->step = 0;
  return SEDCOL_OK;
}

int sedsurr_get_params(sedsurr_handle *h, float *params)
{
  if (!h || !params) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  std::vector<float> buf(h->np_pad);
  HIPOK(hipMemcpyAsync(buf.data(), h->w, sizeof(float) * h->np_pad, hipMemcpyDeviceToHost, h->st));
  HIPOK(hipStreamSynchronize(h->st));
  unpad_params(h, buf, params);
  return SEDCOL_OK;
}

int sedsurr_set_data(sedsurr_handle *h, const float *X, const float *Y, int64_t n)
{
  if (!h || !X || !Y) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (n < 1) return fail(SEDCOL_EINVAL, "need n >= 1");
  if (h->X) (void)hipFree(h->X);
  if (h->Y) (void)hipFree(h->Y);
  h->X = h->Y = nullptr;
  h->ndata = 0;
  HIPOK(hipMalloc(&h->X, sizeof(float) * (size_t)n * h->s.n_in));
  HIPOK(hipMalloc(&h->Y, sizeof(float) * (size_t)n * h->s.n_out));
  HIPOK(hipMemcpyAsync(h->X, X, sizeof(float) * (size_t)n * h->s.n_in, hipMemcpyHostToDevice, h->st));
  HIPOK(hipMemcpyAsync(h->Y, Y, sizeof(float) * (size_t)n * h->s.n_out, hipMemcpyHostToDevice, h->st));
  HIPOK(hipStreamSynchronize(h->st));
  h->ndata = n;
  return SEDCOL_OK;
}

int sedsurr_train(sedsurr_handle *h, int32_t steps, int32_t batch, const int32_t *rows, const sedsurr_adam *adam,
                  float *loss)
{
  if (!h || !rows || !adam) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (!h->X) return fail(SEDCOL_EINVAL, "no training set (sedsurr_set_data)");
  if (steps < 1) return fail(SEDCOL_EINVAL, "need steps >= 1");
  if (batch < 1 || batch > h->s.max_batch) return fail(SEDCOL_EINVAL, "batch outside 1..max_batch");
  if (!(adam->lr > 0.0f) || !(adam->beta1 >= 0.0f && adam->beta1 < 1.0f) ||
      !(adam->beta2 >= 0.0f && adam->beta2 < 1.0f) || !(adam->eps > 0.0f))
    return fail(SEDCOL_EINVAL, "Adam settings out of range");
  const long long nr = (long long)steps * batch;
  for (long long i = 0; i < nr; ++i)
    if (rows[i] < 0 || rows[i] >= h->ndata) return fail(SEDCOL_EINVAL, "row index outside the training set");
  if (nr > h->rows_cap) {
    if (h->rows) (void)hipFree(h->rows);
    h->rows = nullptr;
    h->rows_cap = 0;
    HIPOK(hipMalloc(&h->rows, sizeof(int) * nr));
    h->rows_cap = nr;
  }
  if (steps > h->loss_cap) {
    if (h->loss) (void)hipFree(h->loss);
    h->loss = nullptr;
    h->loss_cap = 0;
    HIPOK(hipMalloc(&h->loss, sizeof(float) * steps));
    h->loss_cap = steps;
  }
  HIPOK(hipMemcpyAsync(h->rows, rows, sizeof(int) * nr, hipMemcpyHostToDevice, h->st));
  int launches = 0;
  HIPOK(hipEventRecord(h->e0, h->st));
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(ss::k_gather, dim3(batch), dim3(64), 0, h->st, h->X, h->s.n_in, h->Y, h->s.n_out,
                       h->rows + (long long)s * batch, batch, h->kin, h->nout_p, h->xb, h->yb);
    ++launches;
    forward(h, batch, &launches);
    loss_backward(h, batch, h->loss + s, &launches, true);
    ++h->step;
    const float bc1 = 1.0f - std::pow(adam->beta1, (float)h->step);
    const float bc2 = 1.0f - std::pow(adam->beta2, (float)h->step);
    hipLaunchKernelGGL(ss::k_adam, dim3(1024), dim3(256), 0, h->st, h->w, h->g, h->m, h->v, h->wb, h->np_pad,
                       adam->lr, adam->beta1, adam->beta2, adam->eps, bc1, bc2, h->kpart_all, h->slices);
    ++launches;
  }
  HIPOK(hipGetLastError());
  HIPOK(hipEventRecord(h->e1, h->st));
  HIPOK(hipEventSynchronize(h->e1));
  float ms = 0.f;
  HIPOK(hipEventElapsedTime(&ms, h->e0, h->e1));
  h->last_ms = ms;
  h->last_launches = launches;
  if (loss) {
    HIPOK(hipMemcpyAsync(loss, h->loss, sizeof(float) * steps, hipMemcpyDeviceToHost, h->st));
    HIPOK(hipStreamSynchronize(h->st));
  }
  return SEDCOL_OK;
}

int sedsurr_loss_grad(sedsurr_handle *h, const float *X, const float *Y, int32_t n, float *loss, float *grad)
{
  if (!h || !X || !Y || !loss || !grad) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (n < 1 || n > h->s.max_batch) return fail(SEDCOL_EINVAL, "n outside 1..max_batch");
  HIPOK(hipMemcpyAsync(h->xin, X, sizeof(float) * (size_t)n * h->s.n_in, hipMemcpyHostToDevice, h->st));
  HIPOK(hipMemcpyAsync(h->yin, Y, sizeof(float) * (size_t)n * h->s.n_out, hipMemcpyHostToDevice, h->st));
  int launches = 0;
  hipLaunchKernelGGL(ss::k_gather, dim3(n), dim3(64), 0, h->st, h->xin, h->s.n_in, h->yin, h->s.n_out,
                     (const int *)nullptr, n, h->kin, h->nout_p, h->xb, h->yb);
  forward(h, n, &launches);
  loss_backward(h, n, h->part + kLossBlocks, &launches);
  HIPOK(hipGetLastError());
  std::vector<float> buf(h->np_pad);
  HIPOK(hipMemcpyAsync(buf.data(), h->g, sizeof(float) * h->np_pad, hipMemcpyDeviceToHost, h->st));
  HIPOK(hipMemcpyAsync(loss, h->part + kLossBlocks, sizeof(float), hipMemcpyDeviceToHost, h->st));
  HIPOK(hipStreamSynchronize(h->st));
  unpad_params(h, buf, grad);
  return SEDCOL_OK;
}

// prediction passes of max_batch rows; Y on the device (written by the last
// layer's product directly) or in host memory (copied per pass)
static int predict_rows(sedsurr_handle *h, const float *X, int64_t n, float *Y, bool y_on_device)
{
  int launches = 0;
  HIPOK(hipEventRecord(h->e0, h->st));
  const int B = h->s.max_batch, no = h->s.n_out;
  for (int64_t r0 = 0; r0 < n; r0 += B) {
    const int b = (int)std::min<int64_t>(B, n - r0);
    HIPOK(hipMemcpyAsync(h->xin, X + r0 * h->s.n_in, sizeof(float) * (size_t)b * h->s.n_in, hipMemcpyHostToDevice,
                         h->st));
    hipLaunchKernelGGL(ss::k_gather, dim3(b), dim3(64), 0, h->st, h->xin, h->s.n_in, (const float *)nullptr, no,
                       (const int *)nullptr, b, h->kin, h->nout_p, h->xb, (float *)nullptr);
    ++launches;
    // compact b x n_out output, in place on the device or staged for the host
    forward(h, b, &launches, y_on_device ? Y + r0 * no : h->yb);
    HIPOK(hipGetLastError());
    if (!y_on_device)
      HIPOK(hipMemcpyAsync(Y + r0 * no, h->yb, sizeof(float) * (size_t)b * no, hipMemcpyDeviceToHost, h->st));
  }
  HIPOK(hipEventRecord(h->e1, h->st));
  HIPOK(hipEventSynchronize(h->e1));
  float ms = 0.f;
  HIPOK(hipEventElapsedTime(&ms, h->e0, h->e1));
  h->last_ms = ms;
  h->last_launches = launches;
  return SEDCOL_OK;
}

int sedsurr_predict(sedsurr_handle *h, const float *X, int64_t n, float *Y)
{
  if (!h || !X || !Y) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (n < 0) return fail(SEDCOL_EINVAL, "need n >= 0");
  return predict_rows(h, X, n, Y, false);
}

int sedsurr_predict_device(sedsurr_handle *h, const float *X, int64_t n, const float **Y_device)
{
  if (!h || !X || !Y_device) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  if (n < 1) return fail(SEDCOL_EINVAL, "need n >= 1");
  if (n > h->ydev_rows) {
    HIPOK(hipStreamSynchronize(h->st));
    if (h->ydev) HIPOK(hipFree(h->ydev));
    h->ydev = nullptr;
    h->ydev_rows = 0;
    HIPOK(hipMalloc(&h->ydev, sizeof(float) * (size_t)n * h->s.n_out));
    h->ydev_rows = n;
  }
  const int rc = predict_rows(h, X, n, h->ydev, true);
  *Y_device = rc == SEDCOL_OK ? h->ydev : nullptr;
  return rc;
}

int sedsurr_last_timing(sedsurr_handle *h, double *ms, int32_t *launches)
{
  if (!h || !ms || !launches) return fail(SEDCOL_EINVAL, "null argument");
  const DeviceScope dscope(h->device);
  *ms = h->last_ms;
  *launches = h->last_launches;
  return SEDCOL_OK;
}

}  // extern "C"
