// Linear density design map for gfx950: eps = eps_lo + rho * (eps_hi - eps_lo)
// through the z-layers of a design box, and its reverse mode onto rho.
//
// forward: inv_eps[(x*ny + y)*nz + z] = (float)(1 / (eps_lo + rho[i, j] (eps_hi - eps_lo)))
//   for the box cells x = x0 + i, y = y0 + j, z in [z0, z0 + sz); the arithmetic is
//   float64 with one rounding to float32, so a bundle written by the same map on
//   the host holds the same bits.
// pullback: grad_rho[i, j] = sum_z grad_inv_eps[i, j, z] * (-(eps_hi - eps_lo) / eps^2),
//   summed over z in a fixed order (deterministic), float64 throughout.
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_density_forward(
    const double* __restrict__ rho, float* __restrict__ inv_eps, int ny, int nz, int x0, int y0,
    int z0, int sx, int sy, int sz, double eps_lo, double eps_hi) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t count = (int64_t)sx * sy * sz;
  if (i >= count) return;
  const int z = (int)(i % sz);
  const int j = (int)((i / sz) % sy);
  const int k = (int)(i / ((int64_t)sz * sy));
  const double e = eps_lo + rho[k * sy + j] * (eps_hi - eps_lo);
  inv_eps[((int64_t)(x0 + k) * ny + (y0 + j)) * nz + z0 + z] = (float)(1.0 / e);
}

__global__ __launch_bounds__(kBlock) void k_density_pullback(
    const double* __restrict__ rho, const double* __restrict__ grad, double* __restrict__ out,
    int sx, int sy, int sz, double eps_lo, double eps_hi) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= sx * sy) return;
  const double span = eps_hi - eps_lo;
  const double e = eps_lo + rho[i] * span;
  const double d = -span / (e * e);
  double total = 0.0;
  for (int z = 0; z < sz; ++z) total += grad[(int64_t)i * sz + z] * d;
  out[i] = total;
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_density_forward(const double* rho, float* inv_eps, int nx, int ny, int nz, int x0,
                        int y0, int z0, int sx, int sy, int sz, double eps_lo, double eps_hi,
                        void* stream) {
  WFD_REQUIRE(rho && inv_eps, "density_forward: null pointer");
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells,
              "density_forward: bad grid");
  WFD_REQUIRE(sx > 0 && sy > 0 && sz > 0 && x0 >= 0 && y0 >= 0 && z0 >= 0 && x0 + sx <= nx &&
                  y0 + sy <= ny && z0 + sz <= nz,
              "density_forward: design box outside the grid");
  WFD_REQUIRE(eps_lo > 0 && eps_hi > 0, "density_forward: permittivities must be positive");
  const int64_t count = (int64_t)sx * sy * sz;
  hipLaunchKernelGGL(k_density_forward, dim3((unsigned)((count + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, as_stream(stream), rho, inv_eps, ny, nz, x0, y0, z0, sx,
                     sy, sz, eps_lo, eps_hi);
  WFD_LAUNCH_CHECK("density_forward");
  return WFD_OK;
}

int wfd_density_pullback(const double* rho, const double* grad_inv_eps, double* grad_rho, int sx,
                         int sy, int sz, double eps_lo, double eps_hi, void* stream) {
  WFD_REQUIRE(rho && grad_inv_eps && grad_rho, "density_pullback: null pointer");
  WFD_REQUIRE(sx > 0 && sy > 0 && sz > 0 && (int64_t)sx * sy * sz <= kMaxCells,
              "density_pullback: bad design box");
  WFD_REQUIRE(eps_lo > 0 && eps_hi > 0, "density_pullback: permittivities must be positive");
  hipLaunchKernelGGL(k_density_pullback, dim3((sx * sy + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), rho, grad_inv_eps, grad_rho, sx, sy, sz, eps_lo, eps_hi);
  WFD_LAUNCH_CHECK("density_pullback");
  return WFD_OK;
}

}  // extern "C"
