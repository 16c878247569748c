// Design-parameter material map of warp_fdtd/material.py for gfx950: clamped
// 19x19 Gaussian filter and tanh subpixel projection, with a hand-derived
// reverse mode (the reference differentiates through wp.Tape).
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;
constexpr int kR = 9;  // filter radius (19x19 taps)

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// material.py:12-25
__global__ __launch_bounds__(kBlock) void k_smooth(const float* __restrict__ p,
                                                   const float* __restrict__ w,
                                                   float* __restrict__ out, int sx, int sy) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= sx * sy) return;
  const int x = i / sy, y = i % sy;
  float total = 0.f;
  for (int dx = -kR; dx <= kR; ++dx)
    for (int dy = -kR; dy <= kR; ++dy) {
      const int ix = clampi(x + dx, 0, sx - 1);
      const int iy = clampi(y + dy, 0, sy - 1);
      total += p[ix * sy + iy] * w[(dx + kR) * (2 * kR + 1) + dy + kR];
    }
  out[i] = total;
}

// material.py:28-32
__device__ __forceinline__ float project(float r, float beta) {
  return (tanhf(0.5f * beta) + tanhf(beta * (r - 0.5f))) / (2.f * tanhf(0.5f * beta));
}

__device__ __forceinline__ float project_d(float r, float beta) {
  const float th = tanhf(beta * (r - 0.5f));
  return beta * (1.f - th * th) / (2.f * tanhf(0.5f * beta));
}

struct Stencil {
  int xm, xp, ym, yp;
  float gx, gy, norm2, r;
};

__device__ __forceinline__ Stencil stencil(const float* __restrict__ rho, int x, int y, int sx,
                                           int sy) {
  Stencil s;
  s.xm = max(x - 1, 0);
  s.xp = min(x + 1, sx - 1);
  s.ym = max(y - 1, 0);
  s.yp = min(y + 1, sy - 1);
  s.gx = (rho[s.xp * sy + y] - rho[s.xm * sy + y]) / (float)(s.xp - s.xm);
  s.gy = (rho[x * sy + s.yp] - rho[x * sy + s.ym]) / (float)(s.yp - s.ym);
  s.norm2 = s.gx * s.gx + s.gy * s.gy;
  s.r = rho[x * sy + y];
  return s;
}

// material.py:35-55
__global__ __launch_bounds__(kBlock) void k_subpixel(const float* __restrict__ rho,
                                                     float* __restrict__ out, int sx, int sy,
                                                     float beta) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= sx * sy) return;
  const int x = i / sy, y = i % sy;
  const Stencil s = stencil(rho, x, y, sx, sy);
  const float r = s.r;
  float value = project(r, beta);
  if (s.norm2 > 0.f) {
    const float norm = sqrtf(s.norm2);
    const float d = (0.5f - r) / (0.55f * norm);
    if (fabsf(d) < 1.f) {
      const float fill =
          0.5f - 0.9375f * d + ((0.625f * d) * d) * d - ((((0.1875f * d) * d) * d) * d) * d;
      const float lo = project(r - (0.55f * norm) * fill, beta);
      const float hi = project(r + (0.55f * norm) * (1.f - fill), beta);
      value = (1.f - fill) * lo + fill * hi;
    }
  }
  out[i] = 1.f / (1.f + 11.25f * value);
}

// Reverse of k_subpixel, pass 1: per output cell the cotangent's share for
// its own rho (c_r) and for the two central differences (c_gx, c_gy).
__global__ __launch_bounds__(kBlock) void k_subpixel_back1(const float* __restrict__ rho,
                                                           const float* __restrict__ g_out,
                                                           float* __restrict__ scratch,
                                                           int sx, int sy, float beta) {
  const int m = sx * sy;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  const int x = i / sy, y = i % sy;
  const Stencil s = stencil(rho, x, y, sx, sy);
  const float r = s.r;
  float value = project(r, beta);
  float dv_dr = project_d(r, beta);
  float dv_dnorm = 0.f;
  float norm = 0.f;
  if (s.norm2 > 0.f) {
    norm = sqrtf(s.norm2);
    const float A = 0.55f * norm;
    const float d = (0.5f - r) / A;
    if (fabsf(d) < 1.f) {
      const float d2 = d * d;
      const float fill = 0.5f - 0.9375f * d + ((0.625f * d) * d) * d -
                         ((((0.1875f * d) * d) * d) * d) * d;
      const float dfill = -0.9375f + 1.875f * d2 - 0.9375f * d2 * d2;
      const float ulo = r - A * fill, uhi = r + A * (1.f - fill);
      const float lo = project(ulo, beta), hi = project(uhi, beta);
      value = (1.f - fill) * lo + fill * hi;
      const float c_lo = (1.f - fill) * project_d(ulo, beta);
      const float c_hi = fill * project_d(uhi, beta);
      const float c_f = (hi - lo) - A * (c_lo + c_hi);
      dv_dr = (c_lo + c_hi) - c_f * dfill / A;
      const float dv_dA = -c_f * dfill * d / A - c_lo * fill + c_hi * (1.f - fill);
      dv_dnorm = 0.55f * dv_dA;
    }
  }
  const float den = 1.f + 11.25f * value;
  const float G = g_out[i] * (-11.25f / (den * den));
  scratch[i] = G * dv_dr;
  if (dv_dnorm != 0.f) {
    scratch[m + i] = G * dv_dnorm * (s.gx / norm) / (float)(s.xp - s.xm);
    scratch[2 * m + i] = G * dv_dnorm * (s.gy / norm) / (float)(s.yp - s.ym);
  } else {
    scratch[m + i] = 0.f;
    scratch[2 * m + i] = 0.f;
  }
}

// Pass 2: gather the central-difference shares into each rho cell.
__global__ __launch_bounds__(kBlock) void k_subpixel_back2(const float* __restrict__ scratch,
                                                           float* __restrict__ g_rho, int sx,
                                                           int sy) {
  const int m = sx * sy;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m) return;
  const int u = i / sy, v = i % sy;
  const float* cgx = scratch + m;
  const float* cgy = scratch + 2 * m;
  float acc = scratch[i];
  // cells x whose xp(x) == u add +c_gx, whose xm(x) == u add -c_gx.
  if (u >= 1) acc += cgx[(u - 1) * sy + v];
  if (u == sx - 1) acc += cgx[u * sy + v];
  if (u + 1 <= sx - 1) acc -= cgx[(u + 1) * sy + v];
  if (u == 0) acc -= cgx[v];
  if (v >= 1) acc += cgy[u * sy + v - 1];
  if (v == sy - 1) acc += cgy[u * sy + v];
  if (v + 1 <= sy - 1) acc -= cgy[u * sy + v + 1];
  if (v == 0) acc -= cgy[u * sy];
  g_rho[i] = acc;
}

// Transpose of the clamped filter: every (x, dx) with clamp(x+dx) == i.
__global__ __launch_bounds__(kBlock) void k_smooth_back(const float* __restrict__ g,
                                                        const float* __restrict__ w,
                                                        float* __restrict__ g_p, int sx,
                                                        int sy) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= sx * sy) return;
  const int a = i / sy, b = i % sy;
  float total = 0.f;
  for (int dx = -kR; dx <= kR; ++dx) {
    int xlo = a - dx, xhi = a - dx;
    if (a == 0) xlo = 0;               // x + dx <= 0
    if (a == sx - 1) xhi = sx - 1;     // x + dx >= sx - 1
    xlo = max(xlo, 0);
    xhi = min(xhi, sx - 1);
    for (int x = xlo; x <= xhi; ++x) {
      if (clampi(x + dx, 0, sx - 1) != a) continue;
      for (int dy = -kR; dy <= kR; ++dy) {
        int ylo = b - dy, yhi = b - dy;
        if (b == 0) ylo = 0;
        if (b == sy - 1) yhi = sy - 1;
        ylo = max(ylo, 0);
        yhi = min(yhi, sy - 1);
        const float wt = w[(dx + kR) * (2 * kR + 1) + dy + kR];
        for (int y = ylo; y <= yhi; ++y) {
          if (clampi(y + dy, 0, sy - 1) != b) continue;
          total += g[x * sy + y] * wt;
        }
      }
    }
  }
  g_p[i] = total;
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_material_forward(const float* p, const float* weights, float* filtered,
                         float* inv_eps, int sx, int sy, float beta, void* stream) {
  WFD_REQUIRE(p && weights && filtered && inv_eps, "material: null pointer");
  WFD_REQUIRE(sx >= 2 && sy >= 2 && (int64_t)sx * sy <= kMaxCells, "material: bad shape");
  WFD_REQUIRE(beta > 0.f, "material: beta must be positive");
  const int m = sx * sy;
  const dim3 grid((m + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_smooth, grid, dim3(kBlock), 0, as_stream(stream), p, weights, filtered,
                     sx, sy);
  WFD_LAUNCH_CHECK("material smooth");
  hipLaunchKernelGGL(k_subpixel, grid, dim3(kBlock), 0, as_stream(stream), filtered, inv_eps,
                     sx, sy, beta);
  WFD_LAUNCH_CHECK("material subpixel");
  return WFD_OK;
}

int wfd_material_pullback(const float* weights, const float* filtered,
                          const float* grad_inv_eps, float* grad_p, float* scratch, int sx,
                          int sy, float beta, void* stream) {
  WFD_REQUIRE(weights && filtered && grad_inv_eps && grad_p && scratch,
              "material pullback: null pointer");
  WFD_REQUIRE(sx >= 2 && sy >= 2 && (int64_t)sx * sy * 4 <= kMaxCells,
              "material pullback: bad shape");
  const int m = sx * sy;
  const dim3 grid((m + kBlock - 1) / kBlock);
  // scratch: [0, 3m) the per-cell shares of pass 1, [3m, 4m) d/d(filtered).
  hipLaunchKernelGGL(k_subpixel_back1, grid, dim3(kBlock), 0, as_stream(stream), filtered,
                     grad_inv_eps, scratch, sx, sy, beta);
  WFD_LAUNCH_CHECK("material pullback (subpixel 1)");
  hipLaunchKernelGGL(k_subpixel_back2, grid, dim3(kBlock), 0, as_stream(stream), scratch,
                     scratch + 3 * m, sx, sy);
  WFD_LAUNCH_CHECK("material pullback (subpixel 2)");
  hipLaunchKernelGGL(k_smooth_back, grid, dim3(kBlock), 0, as_stream(stream), scratch + 3 * m,
                     weights, grad_p, sx, sy);
  WFD_LAUNCH_CHECK("material pullback (smooth)");
  return WFD_OK;
}

}  // extern "C"
