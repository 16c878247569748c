// Kernels of the Fourier-neural-operator surrogate for gfx950: a strided,
// batched GEMM on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulation) that carries every contraction of the network (pointwise
// channel mixing, the truncated DFTs as dense DFT-matrix products, the
// per-mode complex channel mixing in real 2x2 block form) and their
// transposes for the backward pass, plus the elementwise, reduction and
// optimizer kernels around it.  Operands are read as FP32 or BF16 and rounded
// to BF16 (round to nearest even) when staged into LDS.
#include "common.hpp"

namespace wfd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kGemmThreads = 256;  // 4 wavefronts, a 32x32 quadrant each
constexpr int kPad = 8;            // LDS row padding (bf16) against bank conflicts

__device__ __forceinline__ unsigned short to_bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_bits_to_float(unsigned short h) {
  return __uint_as_float(((unsigned int)h) << 16);
}

struct GemmArgs {
  const void* a;
  const void* b;
  float* c;
  long long sab, sam, sak;
  long long sbb, sbk, sbn;
  long long scb, scm, scn;
  int m, n, k;
  float alpha, beta;
};

template <bool ABF16>
__device__ __forceinline__ unsigned short load_bits(const void* p, long long i) {
  if (ABF16) return static_cast<const unsigned short*>(p)[i];
  return to_bf16_bits(static_cast<const float*>(p)[i]);
}

// C[b] = alpha * A[b] (m x k) * B[b] (k x n) + beta * C[b], any strides.
// Block: a 64x64 tile of one batch member; wave w owns rows 32*(w/2)..+32 and
// columns 32*(w%2)..+32 as 2x2 MFMA 16x16 tiles; K advances 32 per stage.
template <bool ABF16, bool BBF16>
__global__ __launch_bounds__(kGemmThreads) void k_gemm(GemmArgs g) {
  __shared__ unsigned short As[kBM][kBK + kPad];
  __shared__ unsigned short Bs[kBN][kBK + kPad];  // transposed: [n][k]
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
  const long long bat = blockIdx.z;
  const char* A = static_cast<const char*>(g.a) + bat * g.sab * (ABF16 ? 2 : 4);
  const char* B = static_cast<const char*>(g.b) + bat * g.sbb * (BBF16 ? 2 : 4);
  floatx4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  for (int k0 = 0; k0 < g.k; k0 += kBK) {
    // stage A (64 x 32) and B (32 x 64): 8 elements per thread each
    for (int e = tid; e < kBM * kBK; e += kGemmThreads) {
      const int r = e / kBK, kk = e % kBK;
      const int gm = m0 + r, gk = k0 + kk;
      As[r][kk] = (gm < g.m && gk < g.k)
                      ? load_bits<ABF16>(A, (long long)gm * g.sam + (long long)gk * g.sak)
                      : (unsigned short)0;
    }
    for (int e = tid; e < kBN * kBK; e += kGemmThreads) {
      const int c = e / kBK, kk = e % kBK;
      const int gn = n0 + c, gk = k0 + kk;
      Bs[c][kk] = (gn < g.n && gk < g.k)
                      ? load_bits<BBF16>(B, (long long)gk * g.sbk + (long long)gn * g.sbn)
                      : (unsigned short)0;
    }
    __syncthreads();
    const int row = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 av[2], bv[2];
    for (int i = 0; i < 2; ++i) {
      const unsigned short* pa = &As[wm + 16 * i + row][kq];
      const unsigned short* pb = &Bs[wn + 16 * i + row][kq];
      for (int q = 0; q < 8; ++q) {
        av[i][q] = __builtin_bit_cast(__bf16, pa[q]);
        bv[i][q] = __builtin_bit_cast(__bf16, pb[q]);
      }
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // lane holds rows 4*(lane/16)+q, column lane%16 of each 16x16 tile
  float* C = g.c + bat * g.scb;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) {
        const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
        const int gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < g.m && gn < g.n) {
          float* p = C + (long long)gm * g.scm + (long long)gn * g.scn;
          const float v = g.alpha * acc[i][j][q];
          *p = g.beta != 0.f ? v + g.beta * *p : v;
        }
      }
}

constexpr int kBlock = 256;

__global__ void k_to_bf16(const float* __restrict__ x, unsigned short* __restrict__ y,
                          long long n) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = to_bf16_bits(x[i]);
}

__global__ void k_round_bf16(const float* __restrict__ x, float* __restrict__ y, long long n) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = bf16_bits_to_float(to_bf16_bits(x[i]));
}

// GELU (erf form) and its derivative
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678f)) + z * 0.39894228f * expf(-0.5f * z * z);
}

// pre = y + bias[col]; out = act(pre) (act 0: identity, 1: GELU); rows of n.
__global__ void k_bias_act(const float* __restrict__ y, const float* __restrict__ bias,
                           float* __restrict__ pre, float* __restrict__ out, long long total,
                           int n, int act) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= total) return;
  const float z = y[i] + (bias ? bias[i % n] : 0.f);
  if (pre) pre[i] = z;
  out[i] = act ? gelu(z) : z;
}

__global__ void k_act_back(const float* __restrict__ dout, const float* __restrict__ pre,
                           float* __restrict__ dz, long long total, int act) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= total) return;
  dz[i] = act ? dout[i] * gelu_d(pre[i]) : dout[i];
}

// Column sums of x (rows x n, row-major): pass 1 writes one partial row per
// block over a fixed row range, pass 2 adds the partials in block order.
constexpr int kRowsPerBlock = 1024;

__global__ void k_colsum_partial(const float* __restrict__ x, long long rows, int n,
                                 float* __restrict__ partial) {
  // block (row range, group of 32 columns): 8 row lanes x 32 columns
  __shared__ float red[kBlock];
  const long long r0 = (long long)blockIdx.x * kRowsPerBlock;
  const long long r1 = min(rows, r0 + kRowsPerBlock);
  const int c = blockIdx.y * 32 + (threadIdx.x & 31), lanes = 8, rl = threadIdx.x >> 5;
  float s = 0.f;
  if (c < n)
    for (long long r = r0 + rl; r < r1; r += lanes) s += x[r * n + c];
  red[threadIdx.x] = s;
  __syncthreads();
  if (rl == 0 && c < n) {
    float t = 0.f;
    for (int q = 0; q < lanes; ++q) t += red[q * 32 + (threadIdx.x & 31)];
    partial[(long long)blockIdx.x * n + c] = t;
  }
}

__global__ void k_colsum_final(const float* __restrict__ partial, int parts, int n,
                               float* __restrict__ out, float beta) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= n) return;
  float s = 0.f;
  for (int p = 0; p < parts; ++p) s += partial[(long long)p * n + c];
  out[c] = beta != 0.f ? s + beta * out[c] : s;
}

// mean over the p pixels of each batch member: x (b, p, w) -> y (b, w).
// Pass 1: block (member, group of 32 channels, run of kPoolRows pixels) sums
// its run (8 pixel lanes x 32 channels); pass 2 adds the runs in order.
constexpr int kPoolRows = 512;

__global__ void k_pool_partial(const float* __restrict__ x, int p, int w,
                               float* __restrict__ partial) {
  __shared__ float red[kBlock];
  const int b = blockIdx.x, c = blockIdx.y * 32 + (threadIdx.x & 31), rl = threadIdx.x >> 5;
  const int i0 = blockIdx.z * kPoolRows, i1 = min(p, i0 + kPoolRows);
  float s = 0.f;
  if (c < w)
    for (int i = i0 + rl; i < i1; i += 8) s += x[((long long)b * p + i) * w + c];
  red[threadIdx.x] = s;
  __syncthreads();
  if (rl == 0 && c < w) {
    float t = 0.f;
    for (int q = 0; q < 8; ++q) t += red[q * 32 + (threadIdx.x & 31)];
    partial[((long long)b * gridDim.z + blockIdx.z) * w + c] = t;
  }
}

__global__ void k_pool_final(const float* __restrict__ partial, int b, int runs, int w, int p,
                             float* __restrict__ y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= b * w) return;
  const int m = i / w, c = i % w;
  float s = 0.f;
  for (int k = 0; k < runs; ++k) s += partial[((long long)m * runs + k) * w + c];
  y[i] = s / (float)p;
}

__global__ void k_pool_mean_back(const float* __restrict__ dy, long long total, int p, int w,
                                 float* __restrict__ dx, float beta) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= total) return;
  const long long b = i / ((long long)p * w);
  const int c = (int)(i % w);
  const float v = dy[b * w + c] / (float)p;
  dx[i] = beta != 0.f ? v + beta * dx[i] : v;
}

// d = scale * (pred - target); partial[block] = sum of (pred - target)^2
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                      long long n, float scale, float* __restrict__ d,
                      float* __restrict__ partial) {
  __shared__ float red[kBlock];
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  float e = 0.f;
  if (i < n) {
    const float r = pred[i] - target[i];
    d[i] = scale * r;
    e = r * r;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on FP32 master weights; w16 receives the BF16 copy the GEMMs read.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v,
                       unsigned short* __restrict__ w16, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float decay) {
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + decay * w[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (w16) w16[i] = to_bf16_bits(wi);
}

unsigned grid_of(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_gemm_bf16(const void* a, int a_bf16, long long sab, long long sam, long long sak,
                  const void* b, int b_bf16, long long sbb, long long sbk, long long sbn,
                  float* c, long long scb, long long scm, long long scn, int batch, int m, int n,
                  int k, float alpha, float beta, void* stream) {
  WFD_REQUIRE(a && b && c, "gemm: null operand");
  WFD_REQUIRE(batch >= 1 && m >= 1 && n >= 1 && k >= 1, "gemm: sizes must be positive");
  WFD_REQUIRE(batch <= 65535 && (m + kBM - 1) / kBM <= 2147483647 && (n + kBN - 1) / kBN <= 65535,
              "gemm: grid too large");
  GemmArgs g{a, b, c, sab, sam, sak, sbb, sbk, sbn, scb, scm, scn, m, n, k, alpha, beta};
  const dim3 grid((m + kBM - 1) / kBM, (n + kBN - 1) / kBN, batch);
  hipStream_t s = as_stream(stream);
  if (a_bf16 && b_bf16)
    hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(kGemmThreads), 0, s, g);
  else if (a_bf16)
    hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(kGemmThreads), 0, s, g);
  else if (b_bf16)
    hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(kGemmThreads), 0, s, g);
  else
    hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(kGemmThreads), 0, s, g);
  WFD_LAUNCH_CHECK("gemm_bf16");
  return WFD_OK;
}

int wfd_to_bf16(const float* x, void* y, long long n, void* stream) {
  WFD_REQUIRE(x && y && n >= 0, "to_bf16: bad argument");
  if (!n) return WFD_OK;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_of(n)), dim3(kBlock), 0, as_stream(stream), x,
                     static_cast<unsigned short*>(y), n);
  WFD_LAUNCH_CHECK("to_bf16");
  return WFD_OK;
}

int wfd_round_bf16(const float* x, float* y, long long n, void* stream) {
  WFD_REQUIRE(x && y && n >= 0, "round_bf16: bad argument");
  if (!n) return WFD_OK;
  hipLaunchKernelGGL(k_round_bf16, dim3(grid_of(n)), dim3(kBlock), 0, as_stream(stream), x, y,
                     n);
  WFD_LAUNCH_CHECK("round_bf16");
  return WFD_OK;
}

int wfd_bias_act(const float* y, const float* bias, float* pre, float* out, long long rows,
                 int n, int act, void* stream) {
  WFD_REQUIRE(y && out && rows >= 0 && n >= 1 && (act == 0 || act == 1), "bias_act: bad argument");
  const long long total = rows * n;
  if (!total) return WFD_OK;
  hipLaunchKernelGGL(k_bias_act, dim3(grid_of(total)), dim3(kBlock), 0, as_stream(stream), y,
                     bias, pre, out, total, n, act);
  WFD_LAUNCH_CHECK("bias_act");
  return WFD_OK;
}

int wfd_act_back(const float* dout, const float* pre, float* dz, long long total, int act,
                 void* stream) {
  WFD_REQUIRE(dout && dz && total >= 0 && (act == 0 || (act == 1 && pre)),
              "act_back: bad argument");
  if (!total) return WFD_OK;
  hipLaunchKernelGGL(k_act_back, dim3(grid_of(total)), dim3(kBlock), 0, as_stream(stream), dout,
                     pre, dz, total, act);
  WFD_LAUNCH_CHECK("act_back");
  return WFD_OK;
}

long long wfd_colsum_scratch(long long rows, int n) {
  if (rows < 1 || n < 1) return -1;
  return ((rows + kRowsPerBlock - 1) / kRowsPerBlock) * (long long)n;
}

int wfd_colsum(const float* x, long long rows, int n, float* out, float* scratch,
               long long scratch_floats, float beta, void* stream) {
  WFD_REQUIRE(x && out && scratch && rows >= 1 && n >= 1, "colsum: bad argument");
  const long long parts = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
  WFD_REQUIRE(scratch_floats >= parts * n, "colsum: scratch smaller than wfd_colsum_scratch");
  WFD_REQUIRE(parts <= 2147483647, "colsum: too many rows");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_colsum_partial, dim3((unsigned)parts, (n + 31) / 32), dim3(kBlock), 0, s,
                     x, rows, n, scratch);
  hipLaunchKernelGGL(k_colsum_final, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, scratch,
                     (int)parts, n, out, beta);
  WFD_LAUNCH_CHECK("colsum");
  return WFD_OK;
}

long long wfd_pool_mean_scratch(int batch, int p, int w) {
  if (batch < 1 || p < 1 || w < 1) return -1;
  return (long long)batch * ((p + kPoolRows - 1) / kPoolRows) * w;
}

int wfd_pool_mean(const float* x, int batch, int p, int w, float* y, float* scratch,
                  long long scratch_floats, void* stream) {
  WFD_REQUIRE(x && y && scratch && batch >= 1 && p >= 1 && w >= 1, "pool_mean: bad argument");
  const int runs = (p + kPoolRows - 1) / kPoolRows;
  WFD_REQUIRE(scratch_floats >= (long long)batch * runs * w,
              "pool_mean: scratch smaller than wfd_pool_mean_scratch");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(k_pool_partial, dim3(batch, (w + 31) / 32, runs), dim3(kBlock), 0, s, x, p,
                     w, scratch);
  hipLaunchKernelGGL(k_pool_final, dim3((batch * w + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                     scratch, batch, runs, w, p, y);
  WFD_LAUNCH_CHECK("pool_mean");
  return WFD_OK;
}

int wfd_pool_mean_back(const float* dy, int batch, int p, int w, float* dx, float beta,
                       void* stream) {
  WFD_REQUIRE(dy && dx && batch >= 1 && p >= 1 && w >= 1, "pool_mean_back: bad argument");
  const long long total = (long long)batch * p * w;
  hipLaunchKernelGGL(k_pool_mean_back, dim3(grid_of(total)), dim3(kBlock), 0, as_stream(stream),
                     dy, total, p, w, dx, beta);
  WFD_LAUNCH_CHECK("pool_mean_back");
  return WFD_OK;
}

long long wfd_mse_partials(long long n) { return n < 1 ? -1 : (long long)grid_of(n); }

int wfd_mse(const float* pred, const float* target, long long n, float scale, float* d,
            float* partial, void* stream) {
  WFD_REQUIRE(pred && target && d && partial && n >= 1, "mse: bad argument");
  hipLaunchKernelGGL(k_mse, dim3(grid_of(n)), dim3(kBlock), 0, as_stream(stream), pred, target,
                     n, scale, d, partial);
  WFD_LAUNCH_CHECK("mse");
  return WFD_OK;
}

int wfd_adam(float* w, const float* g, float* m, float* v, void* w16, long long n, float lr,
             float b1, float b2, float eps, int t, float decay, void* stream) {
  WFD_REQUIRE(w && g && m && v && n >= 0 && t >= 1, "adam: bad argument");
  if (!n) return WFD_OK;
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(grid_of(n)), dim3(kBlock), 0, as_stream(stream), w, g, m, v,
                     static_cast<unsigned short*>(w16), n, lr, b1, b2, eps, c1, c2, decay);
  WFD_LAUNCH_CHECK("adam");
  return WFD_OK;
}

}  // extern "C"
