// Sponge-bounded staggered Yee solver of warp_fdtd/driven_uniform.py for gfx950:
// flat (one thread per stored component) and xyz (one thread per lattice
// point, all three components) launches with identical arithmetic.
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;

struct I3 {
  int v[3];
};

// driven_uniform.py:10-19
__device__ __forceinline__ int e_index(int c, const int p[3], const int n[3]) {
  int dims[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
  dims[c] -= 1;
  int offset = 0;
  if (c >= 1) offset += n[0] * (n[1] + 1) * (n[2] + 1);
  if (c == 2) offset += (n[0] + 1) * n[1] * (n[2] + 1);
  return offset + (p[0] * dims[1] + p[1]) * dims[2] + p[2];
}

// driven_uniform.py:22-31
__device__ __forceinline__ int h_index(int c, const int p[3], const int n[3]) {
  int dims[3] = {n[0], n[1], n[2]};
  dims[c] += 1;
  int offset = 0;
  if (c >= 1) offset += (n[0] + 1) * n[1] * n[2];
  if (c == 2) offset += n[0] * (n[1] + 1) * n[2];
  return offset + (p[0] * dims[1] + p[1]) * dims[2] + p[2];
}

// driven_uniform.py:34-40
__device__ __forceinline__ float sigma(const float p[3], const float L[3], float width,
                                       float rate) {
  float result = 0.f;
  for (int a = 0; a < 3; ++a) {
    const float depth = fmaxf(0.f, 1.f - fminf(p[a], L[a] - p[a]) / width);
    result += ((rate * depth) * depth) * depth;
  }
  return result;
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.f), 1.f); }

// driven_uniform.py:43-59
__device__ __forceinline__ float epsilon(const float p[3], float h, const float L[3],
                                         const float core[2], float eps_core, float eps_clad) {
  float f = 1.f;
  for (int a = 1; a < 3; ++a) {
    const float lo = L[a] * 0.5f - core[a - 1] * 0.5f;
    const float hi = L[a] * 0.5f + core[a - 1] * 0.5f;
    f *= clamp01((fminf(p[a] + h * 0.5f, hi) - fmaxf(p[a] - h * 0.5f, lo)) / h);
  }
  return eps_clad + (eps_core - eps_clad) * f;
}

// driven_uniform.py:62-87
__device__ __forceinline__ float epsilon_normal(const float p[3], float h, const float L[3],
                                                const float core[2], float eps_core,
                                                float eps_clad, int component) {
  float fraction[2];
  for (int a = 1; a < 3; ++a) {
    const float lo = L[a] * 0.5f - core[a - 1] * 0.5f;
    const float hi = L[a] * 0.5f + core[a - 1] * 0.5f;
    fraction[a - 1] = clamp01((fminf(p[a] + h * 0.5f, hi) - fmaxf(p[a] - h * 0.5f, lo)) / h);
  }
  if (component == 0) return eps_clad + ((eps_core - eps_clad) * fraction[0]) * fraction[1];
  const int normal = component - 1;
  const int other = 1 - normal;
  const float inside = eps_clad + (eps_core - eps_clad) * fraction[other];
  return 1.f / ((1.f - fraction[normal]) / eps_clad + fraction[normal] / inside);
}

struct Guide {
  int n[3];
  float h, dt;
  float L[3];
  float core[2];
  float eps_core, eps_clad;
  float sw, sr;
  float src[3], sig[3];
  int normal_average;
};

// driven_uniform.py:90-130 / 192-225 body for component c at lattice point p.
__device__ __forceinline__ void magnetic_at(const Guide& g, const float* __restrict__ e,
                                            float* __restrict__ hf, int c, const int p[3]) {
  const int a = (c + 1) % 3, b = (c + 2) % 3;
  int pa[3] = {p[0], p[1], p[2]};
  pa[a] += 1;
  int pb[3] = {p[0], p[1], p[2]};
  pb[b] += 1;
  const float curl =
      ((e[e_index(b, pa, g.n)] - e[e_index(b, p, g.n)]) - e[e_index(a, pb, g.n)]) +
      e[e_index(a, p, g.n)];
  float xyz[3];
  for (int k = 0; k < 3; ++k) xyz[k] = ((float)p[k] + 0.5f) * g.h;
  xyz[c] -= 0.5f * g.h;
  const float q = (0.5f * g.dt) * sigma(xyz, g.L, g.sw, g.sr);
  const int i = h_index(c, p, g.n);
  hf[i] = ((1.f - q) * hf[i] - (g.dt / g.h) * curl) / (1.f + q);
}

// driven_uniform.py:133-189 / 228-274 body (interior test done by the caller).
__device__ __forceinline__ void electric_at(const Guide& g, const float* __restrict__ hf,
                                            float* __restrict__ e, int c, const int p[3],
                                            float w) {
  const int a = (c + 1) % 3, b = (c + 2) % 3;
  int pa[3] = {p[0], p[1], p[2]};
  pa[a] -= 1;
  int pb[3] = {p[0], p[1], p[2]};
  pb[b] -= 1;
  const float curl =
      ((hf[h_index(b, p, g.n)] - hf[h_index(b, pa, g.n)]) - hf[h_index(a, p, g.n)]) +
      hf[h_index(a, pb, g.n)];
  float xyz[3];
  for (int k = 0; k < 3; ++k) xyz[k] = (float)p[k] * g.h;
  xyz[c] += 0.5f * g.h;
  float ep = epsilon(xyz, g.h, g.L, g.core, g.eps_core, g.eps_clad);
  if (g.normal_average == 1)
    ep = epsilon_normal(xyz, g.h, g.L, g.core, g.eps_core, g.eps_clad, c);
  float current = 0.f;
  if (c == 1) {
    float d[3];
    for (int k = 0; k < 3; ++k) d[k] = (xyz[k] - g.src[k]) / g.sig[k];
    current = expf(-0.5f * ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])) * w;
  }
  const float q = (0.5f * g.dt) * sigma(xyz, g.L, g.sw, g.sr);
  const int i = e_index(c, p, g.n);
  e[i] = ((1.f - q) * e[i] + (g.dt * (curl / g.h + current)) / ep) / (1.f + q);
}

__global__ __launch_bounds__(kBlock) void k_magnetic_flat(Guide g, const float* __restrict__ e,
                                                          float* __restrict__ hf) {
  const int first = (g.n[0] + 1) * g.n[1] * g.n[2];
  const int second = g.n[0] * (g.n[1] + 1) * g.n[2];
  const int third = g.n[0] * g.n[1] * (g.n[2] + 1);
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= first + second + third) return;
  int c = 0, j = i;
  if (i >= first + second) {
    c = 2;
    j = i - first - second;
  } else if (i >= first) {
    c = 1;
    j = i - first;
  }
  int dims[3] = {g.n[0], g.n[1], g.n[2]};
  dims[c] += 1;
  const int p[3] = {j / (dims[1] * dims[2]), (j / dims[2]) % dims[1], j % dims[2]};
  magnetic_at(g, e, hf, c, p);
}

__global__ __launch_bounds__(kBlock) void k_electric_flat(Guide g, const float* __restrict__ hf,
                                                          float* __restrict__ e,
                                                          const float* __restrict__ wave,
                                                          const int* __restrict__ counter) {
  const int first = g.n[0] * (g.n[1] + 1) * (g.n[2] + 1);
  const int second = (g.n[0] + 1) * g.n[1] * (g.n[2] + 1);
  const int third = (g.n[0] + 1) * (g.n[1] + 1) * g.n[2];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= first + second + third) return;
  int c = 0, j = i;
  if (i >= first + second) {
    c = 2;
    j = i - first - second;
  } else if (i >= first) {
    c = 1;
    j = i - first;
  }
  int dims[3] = {g.n[0] + 1, g.n[1] + 1, g.n[2] + 1};
  dims[c] -= 1;
  const int p[3] = {j / (dims[1] * dims[2]), (j / dims[2]) % dims[1], j % dims[2]};
  const int a = (c + 1) % 3, b = (c + 2) % 3;
  if (p[a] > 0 && p[a] < g.n[a] && p[b] > 0 && p[b] < g.n[b])
    electric_at(g, hf, e, c, p, wave[counter[0]]);
}

// 3-D launch over (nx+1, ny+1, nz+1) lattice points, z fastest in the block.
__global__ __launch_bounds__(kBlock) void k_magnetic_xyz(Guide g, const float* __restrict__ e,
                                                         float* __restrict__ hf) {
  const int iz = blockIdx.x * 64 + (threadIdx.x & 63);
  const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int ix = blockIdx.z;
  if (iz > g.n[2] || iy > g.n[1]) return;
  const int p[3] = {ix, iy, iz};
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    if (p[c] <= g.n[c] && p[a] < g.n[a] && p[b] < g.n[b]) magnetic_at(g, e, hf, c, p);
  }
}

__global__ __launch_bounds__(kBlock) void k_electric_xyz(Guide g, const float* __restrict__ hf,
                                                         float* __restrict__ e,
                                                         const float* __restrict__ wave,
                                                         const int* __restrict__ counter) {
  const int iz = blockIdx.x * 64 + (threadIdx.x & 63);
  const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int ix = blockIdx.z;
  if (iz > g.n[2] || iy > g.n[1]) return;
  const int p[3] = {ix, iy, iz};
  const float w = wave[counter[0]];
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    if (p[c] < g.n[c] && p[a] > 0 && p[a] < g.n[a] && p[b] > 0 && p[b] < g.n[b])
      electric_at(g, hf, e, c, p, w);
  }
}

int to_guide(const wfd_guide* w, Guide& g) {
  WFD_REQUIRE(w, "driven: null parameters");
  WFD_REQUIRE(w->nx > 0 && w->ny > 0 && w->nz > 0, "driven: grid dimensions must be positive");
  WFD_REQUIRE((int64_t)(w->nx + 1) * (w->ny + 1) * (w->nz + 1) * 3 <= kMaxCells,
              "driven: grid too large");
  WFD_REQUIRE(w->spacing > 0.f && w->dt > 0.f && w->sponge_width > 0.f,
              "driven: spacing, dt and sponge width must be positive");
  g.n[0] = w->nx;
  g.n[1] = w->ny;
  g.n[2] = w->nz;
  g.h = w->spacing;
  g.dt = w->dt;
  g.L[0] = w->length;
  g.L[1] = w->width;
  g.L[2] = w->height;
  g.core[0] = w->core_width;
  g.core[1] = w->core_height;
  g.eps_core = w->core_epsilon;
  g.eps_clad = w->clad_epsilon;
  g.sw = w->sponge_width;
  g.sr = w->sponge_rate;
  g.src[0] = w->source_x;
  g.src[1] = w->source_y;
  g.src[2] = w->source_z;
  g.sig[0] = w->source_sigma_x;
  g.sig[1] = w->source_sigma_y;
  g.sig[2] = w->source_sigma_z;
  g.normal_average = w->normal_average;
  return WFD_OK;
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_driven_magnetic(const wfd_guide* w, const float* e, float* h, void* stream) {
  Guide g;
  if (int s = to_guide(w, g)) return s;
  WFD_REQUIRE(e && h, "driven_magnetic: null field");
  if (w->layout_xyz) {
    dim3 grid((g.n[2] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[0] + 1);
    hipLaunchKernelGGL(k_magnetic_xyz, grid, dim3(kBlock), 0, as_stream(stream), g, e, h);
  } else {
    const int total = (g.n[0] + 1) * g.n[1] * g.n[2] + g.n[0] * (g.n[1] + 1) * g.n[2] +
                      g.n[0] * g.n[1] * (g.n[2] + 1);
    hipLaunchKernelGGL(k_magnetic_flat, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       as_stream(stream), g, e, h);
  }
  WFD_LAUNCH_CHECK("driven_magnetic");
  return WFD_OK;
}

int wfd_driven_electric(const wfd_guide* w, const float* h, float* e, const float* wave,
                        const int* counter, void* stream) {
  Guide g;
  if (int s = to_guide(w, g)) return s;
  WFD_REQUIRE(e && h && wave && counter, "driven_electric: null pointer");
  if (w->layout_xyz) {
    dim3 grid((g.n[2] + 1 + 63) / 64, (g.n[1] + 1 + 3) / 4, g.n[0] + 1);
    hipLaunchKernelGGL(k_electric_xyz, grid, dim3(kBlock), 0, as_stream(stream), g, h, e, wave,
                       counter);
  } else {
    const int total = g.n[0] * (g.n[1] + 1) * (g.n[2] + 1) +
                      (g.n[0] + 1) * g.n[1] * (g.n[2] + 1) +
                      (g.n[0] + 1) * (g.n[1] + 1) * g.n[2];
    hipLaunchKernelGGL(k_electric_flat, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       as_stream(stream), g, h, e, wave, counter);
  }
  WFD_LAUNCH_CHECK("driven_electric");
  return WFD_OK;
}

}  // extern "C"
