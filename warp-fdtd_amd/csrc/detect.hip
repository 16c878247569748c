// Detectors for gfx950: the reference's per-monitor record (detectors.py),
// all-monitor record in one launch, FP64 multi-frequency DFT planes, the
// full-volume trilinear sampler (volume.py) and small diagnostics kernels.
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;

// detectors.py:10-23
__device__ __forceinline__ F3 ht(const float* __restrict__ h, const float* __restrict__ prev,
                                 int n, int x, int y, int z, int nx, int ny, int nz) {
  const F3 a = sample(h, n, x, y, z, nx, ny, nz, 0);
  const F3 b = sample(prev, n, x, y, z, nx, ny, nz, 0);
  return F3{0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
}

// detectors.py:26-62 — exact space/time colocation at the E_z point.
__device__ __forceinline__ void colocate(const float* __restrict__ e,
                                         const float* __restrict__ h,
                                         const float* __restrict__ prev, int n, int x, int y,
                                         int z, int nx, int ny, int nz, F3& ec, F3& hc) {
  const F3 e0 = sample(e, n, x, y, z, nx, ny, nz, 0);
  const F3 ex = sample(e, n, x - 1, y, z, nx, ny, nz, 0);
  const F3 ey = sample(e, n, x, y - 1, z, nx, ny, nz, 0);
  const F3 ez = sample(e, n, x, y, z + 1, nx, ny, nz, 0);
  const F3 exz = sample(e, n, x - 1, y, z + 1, nx, ny, nz, 0);
  const F3 eyz = sample(e, n, x, y - 1, z + 1, nx, ny, nz, 0);
  ec = F3{0.25f * (((e0.x + ex.x) + ez.x) + exz.x), 0.25f * (((e0.y + ey.y) + ez.y) + eyz.y),
          e0.z};
  const F3 h0 = ht(h, prev, n, x, y, z, nx, ny, nz);
  const F3 hx = ht(h, prev, n, x - 1, y, z, nx, ny, nz);
  const F3 hy = ht(h, prev, n, x, y - 1, z, nx, ny, nz);
  const F3 hxy = ht(h, prev, n, x - 1, y - 1, z, nx, ny, nz);
  const F3 hz = ht(h, prev, n, x, y, z + 1, nx, ny, nz);
  const F3 hxz = ht(h, prev, n, x - 1, y, z + 1, nx, ny, nz);
  const F3 hyz = ht(h, prev, n, x, y - 1, z + 1, nx, ny, nz);
  const F3 hxyz = ht(h, prev, n, x - 1, y - 1, z + 1, nx, ny, nz);
  hc = F3{0.5f * (h0.x + hy.x), 0.5f * (h0.y + hx.y),
          0.125f * ((((((h0.z + hx.z) + hy.z) + hxy.z) + hz.z) + hxz.z) + hyz.z + hxyz.z)};
}

__device__ __forceinline__ void plane_point(int i, int x0, int y0, int z0, int sy, int sz,
                                            int& x, int& y, int& z) {
  x = x0 + i / (sy * sz);
  y = y0 + (i / sz) % sy;
  z = z0 + i % sz;
}

__device__ __forceinline__ void fields_at(const float* e, const float* h, const float* prev,
                                          int n, int x, int y, int z, int nx, int ny, int nz,
                                          int exact, F3& ee, F3& hh) {
  if (exact) {
    colocate(e, h, prev, n, x, y, z, nx, ny, nz, ee, hh);
  } else {
    ee = sample(e, n, x, y, z, nx, ny, nz, 0);
    hh = sample(h, n, x, y, z, nx, ny, nz, 0);
  }
}

__device__ __forceinline__ float cross_axis(const F3& a, const F3& b, int axis) {
  if (axis == 0) return a.y * b.z - a.z * b.y;
  if (axis == 1) return a.z * b.x - a.x * b.z;
  return a.x * b.y - a.y * b.x;
}

struct MonitorArgs {
  wfd_monitor m[WFD_MAX_MONITORS];
  int begin[WFD_MAX_MONITORS + 1];  // first block of each monitor
  int count;
};

// Phasor branch of detectors.py:99-103 for plane cell i.
__device__ __forceinline__ void add_phasor(float* __restrict__ phasor, int count, int i,
                                           float wr, float wi, const F3& ee, const F3& hh) {
  const float v[6] = {ee.x, ee.y, ee.z, hh.x, hh.y, hh.z};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    float2* p = reinterpret_cast<float2*>(phasor) + (c * count + i);
    float2 cur = *p;
    cur.x += wr * v[c];
    cur.y += wi * v[c];
    *p = cur;
  }
}

// Flux branch of detectors.py:104-106: the whole plane reduced by one block
// in a fixed order (per-thread sequential partials, then an LDS tree).
__device__ void flux_block(const wfd_monitor& m, const float* e, const float* h,
                           const float* prev, int n, int nx, int ny, int nz, int index) {
  __shared__ float part[kBlock];
  const int count = m.sx * m.sy * m.sz;
  float acc = 0.f;
  for (int i = threadIdx.x; i < count; i += kBlock) {
    int x, y, z;
    plane_point(i, m.x0, m.y0, m.z0, m.sy, m.sz, x, y, z);
    F3 ee, hh;
    fields_at(e, h, prev, n, x, y, z, nx, ny, nz, m.exact, ee, hh);
    acc += cross_axis(ee, hh, m.axis) * m.area[i];
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) m.flux[index] += part[0];
}

__global__ __launch_bounds__(kBlock) void k_record_monitors(
    const float* __restrict__ e, const float* __restrict__ h, const float* __restrict__ prev,
    const int* __restrict__ step, MonitorArgs args, int nx, int ny, int nz) {
  const int b = blockIdx.x;
  int k = 0;
  while (k + 1 < args.count && b >= args.begin[k + 1]) ++k;
  const wfd_monitor& m = args.m[k];
  const int t = step[0];
  if (t < 0 || t >= m.steps) return;
  const int index = m.indices[t];
  if (index < 0) return;
  const int n = nx * ny * nz;
  if (!m.is_phasor) {
    flux_block(m, e, h, prev, n, nx, ny, nz, index);
    return;
  }
  const int count = m.sx * m.sy * m.sz;
  const int i = (b - args.begin[k]) * kBlock + threadIdx.x;
  if (i >= count) return;
  int x, y, z;
  plane_point(i, m.x0, m.y0, m.z0, m.sy, m.sz, x, y, z);
  F3 ee, hh;
  fields_at(e, h, prev, n, x, y, z, nx, ny, nz, m.exact, ee, hh);
  add_phasor(m.phasor, count, i, m.weights[2 * t], m.weights[2 * t + 1], ee, hh);
}

int launch_monitors(const float* e, const float* h, const float* prev, const int* step,
                    const wfd_monitor* mons, int count, int nx, int ny, int nz,
                    void* stream) {
  WFD_REQUIRE(count >= 1 && count <= WFD_MAX_MONITORS, "record: 1..%d monitors per launch",
              WFD_MAX_MONITORS);
  WFD_REQUIRE(e && h && step && mons, "record: null field, step or monitor pointer");
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells,
              "record: bad grid");
  MonitorArgs args;
  args.count = count;
  int blocks = 0;
  for (int k = 0; k < count; ++k) {
    const wfd_monitor& m = mons[k];
    WFD_REQUIRE(m.indices && m.sx > 0 && m.sy > 0 && m.sz > 0 && m.steps > 0,
                "record: bad monitor %d", k);
    WFD_REQUIRE(m.x0 >= 0 && m.y0 >= 0 && m.z0 >= 0 && m.x0 + m.sx <= nx &&
                    m.y0 + m.sy <= ny && m.z0 + m.sz <= nz,
                "record: monitor %d plane outside the grid", k);
    WFD_REQUIRE(!m.exact || prev, "record: exact colocation needs hprev");
    if (m.is_phasor) {
      WFD_REQUIRE(m.weights && m.phasor, "record: phasor monitor %d without arrays", k);
    } else {
      WFD_REQUIRE(m.flux && m.area && m.axis >= 0 && m.axis < 3,
                  "record: flux monitor %d without arrays", k);
    }
    args.m[k] = m;
    args.begin[k] = blocks;
    const int cnt = m.sx * m.sy * m.sz;
    blocks += m.is_phasor ? (cnt + kBlock - 1) / kBlock : 1;
  }
  args.begin[count] = blocks;
  hipLaunchKernelGGL(k_record_monitors, dim3(blocks), dim3(kBlock), 0, as_stream(stream), e,
                     h, prev, step, args, nx, ny, nz);
  WFD_LAUNCH_CHECK("record");
  return WFD_OK;
}

// ------------------------------------------------ multi-frequency DFT planes
__global__ __launch_bounds__(kBlock) void k_dft_stage(
    const float* __restrict__ e, const float* __restrict__ h, const float* __restrict__ prev,
    const int* __restrict__ step, float* __restrict__ stage, int slots, int nx, int ny,
    int nz, int x0, int y0, int z0, int sx, int sy, int sz, int exact) {
  const int count = sx * sy * sz;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const int slot = step[0] % slots;
  int x, y, z;
  plane_point(i, x0, y0, z0, sy, sz, x, y, z);
  F3 ee, hh;
  fields_at(e, h, prev, nx * ny * nz, x, y, z, nx, ny, nz, exact, ee, hh);
  float* s = stage + (size_t)slot * 6 * count + i;
  s[0] = ee.x;
  s[count] = ee.y;
  s[2 * count] = ee.z;
  s[3 * count] = hh.x;
  s[4 * count] = hh.y;
  s[5 * count] = hh.z;
}

constexpr int kFreqGroup = 8;

// phasor[f, c, i] += sum_t w[t, f] * stage[t % slots, c, i] for the m steps
// t = last-m+1 .. last (last = step[0], the most recently staged step).
// One thread per (c, i); frequencies in groups of 8 held in FP64 registers.
__global__ __launch_bounds__(kBlock) void k_dft_accumulate(
    const float* __restrict__ stage, int slots, const double* __restrict__ weights,
    int steps, int nf, const int* __restrict__ step, int m, double* __restrict__ phasor,
    int count) {
  const int ci = blockIdx.x * kBlock + threadIdx.x;
  if (ci >= 6 * count) return;
  const int f0 = blockIdx.y * kFreqGroup;
  const int last = step[0];
  double re[kFreqGroup], im[kFreqGroup];
#pragma unroll
  for (int k = 0; k < kFreqGroup; ++k) re[k] = im[k] = 0.0;
  for (int j = 0; j < m; ++j) {
    const int t = last - m + 1 + j;
    if (t < 0 || t >= steps) continue;
    const double v = (double)stage[(size_t)(t % slots) * 6 * count + ci];
    const double* w = weights + 2 * ((size_t)t * nf + f0);
#pragma unroll
    for (int k = 0; k < kFreqGroup; ++k) {
      if (f0 + k < nf) {
        re[k] += w[2 * k] * v;
        im[k] += w[2 * k + 1] * v;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kFreqGroup; ++k) {
    if (f0 + k < nf) {
      double2* p = reinterpret_cast<double2*>(phasor) + ((size_t)(f0 + k) * 6 * count + ci);
      double2 cur = *p;
      cur.x += re[k];
      cur.y += im[k];
      *p = cur;
    }
  }
}

// ------------------------------------------------------------ volume.py
__device__ __forceinline__ float trilinear(const float* __restrict__ f, int n, int comp,
                                           float px, float py, float pz, int nx, int ny,
                                           int nz) {
  const int ix = (int)floorf(px), iy = (int)floorf(py), iz = (int)floorf(pz);
  const float tx = px - (float)ix, ty = py - (float)iy, tz = pz - (float)iz;
  float value = 0.f;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int x = ix + a, y = iy + b, z = iz + c;
        if (x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz) {
          const float w =
              ((a ? tx : 1.f - tx) * (b ? ty : 1.f - ty)) * (c ? tz : 1.f - tz);
          value += w * f[comp * n + (x * ny + y) * nz + z];
        }
      }
  return value;
}

__global__ __launch_bounds__(kBlock) void k_record_volume(
    const float* __restrict__ e, const float* __restrict__ h, const float* __restrict__ prev,
    const float* __restrict__ positions, const float* __restrict__ weights,
    const int* __restrict__ step, float* __restrict__ phasor, int count, int nx, int ny,
    int nz, int first) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const int t = step[0];
  if (t < first) return;
  const float wr = weights[2 * t], wi = weights[2 * t + 1];
  const int n = nx * ny * nz;
  const float off[6][3] = {{0.5f, 0.f, 0.f}, {0.f, 0.5f, 0.f}, {0.f, 0.f, 0.5f},
                           {0.f, 0.5f, 0.5f}, {0.5f, 0.f, 0.5f}, {0.5f, 0.5f, 0.f}};
#pragma unroll
  for (int comp = 0; comp < 6; ++comp) {
    const float px = positions[3 * i] - off[comp][0];
    const float py = positions[3 * i + 1] - off[comp][1];
    const float pz = positions[3 * i + 2] - off[comp][2];
    float value;
    if (comp < 3) {
      value = trilinear(e, n, comp, px, py, pz, nx, ny, nz);
    } else {
      value = 0.5f * (trilinear(h, n, comp - 3, px, py, pz, nx, ny, nz) +
                      trilinear(prev, n, comp - 3, px, py, pz, nx, ny, nz));
    }
    float2* p = reinterpret_cast<float2*>(phasor) + ((size_t)comp * count + i);
    float2 cur = *p;
    cur.x += wr * value;
    cur.y += wi * value;
    *p = cur;
  }
}

__global__ void k_record_mode(const float* e, float* signal, int* step) {
  const int t = step[0];
  signal[t] = e[0];
  step[0] = t + 1;
}

__global__ __launch_bounds__(kBlock) void k_observer_gather(
    const int* __restrict__ indices, const float* __restrict__ weights,
    const float* __restrict__ field, const float* __restrict__ old, int center,
    float* __restrict__ out, int rows) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= rows) return;
  float value = 0.f;
  for (int j = 0; j < 8; ++j) {
    const int k = indices[row * 8 + j];
    float v = field[k];
    if (center == 1) v = 0.5f * (v + old[k]);
    value += weights[row * 8 + j] * v;
  }
  out[row] = value;
}

__global__ __launch_bounds__(kBlock) void k_count_nonfinite(const float* __restrict__ values,
                                                            int64_t n, int* count) {
  int local = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    local += !isfinite(values[i]);
  // wavefront reduction, then one atomic per wavefront with a nonzero count
  for (int off = 32; off > 0; off >>= 1) local += __shfl_down(local, off, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(count, local);
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_record(const float* e, const float* h, const float* prev, const int* step,
               const float* weights, const int* indices, float* phasor, float* flux,
               const float* area, int nx, int ny, int nz, int x0, int y0, int z0, int sx,
               int sy, int sz, int axis, int exact, int is_phasor, int steps,
               void* stream) {
  wfd_monitor m;
  m.steps = steps;
  m.indices = indices;
  m.weights = weights;
  m.phasor = phasor;
  m.flux = flux;
  m.area = area;
  m.x0 = x0;
  m.y0 = y0;
  m.z0 = z0;
  m.sx = sx;
  m.sy = sy;
  m.sz = sz;
  m.axis = axis;
  m.exact = exact;
  m.is_phasor = is_phasor;
  return launch_monitors(e, h, prev, step, &m, 1, nx, ny, nz, stream);
}

int wfd_record_monitors(const float* e, const float* h, const float* prev, const int* step,
                        const wfd_monitor* monitors, int count, int nx, int ny, int nz,
                        void* stream) {
  return launch_monitors(e, h, prev, step, monitors, count, nx, ny, nz, stream);
}

int wfd_dft_stage(const float* e, const float* h, const float* prev, const int* step,
                  float* stage, int slots, int nx, int ny, int nz, int x0, int y0, int z0,
                  int sx, int sy, int sz, int exact, void* stream) {
  WFD_REQUIRE(e && h && step && stage && slots > 0, "dft_stage: null pointer or slots < 1");
  WFD_REQUIRE(!exact || prev, "dft_stage: exact colocation needs hprev");
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells,
              "dft_stage: bad grid");
  WFD_REQUIRE(sx > 0 && sy > 0 && sz > 0 && x0 >= 0 && y0 >= 0 && z0 >= 0 &&
                  x0 + sx <= nx && y0 + sy <= ny && z0 + sz <= nz,
              "dft_stage: plane outside the grid");
  const int count = sx * sy * sz;
  hipLaunchKernelGGL(k_dft_stage, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), e, h, prev, step, stage, slots, nx, ny, nz, x0, y0,
                     z0, sx, sy, sz, exact);
  WFD_LAUNCH_CHECK("dft_stage");
  return WFD_OK;
}

int wfd_dft_accumulate(const float* stage, int slots, const double* weights, int steps,
                       int nf, const int* step, int m, double* phasor, int count,
                       void* stream) {
  WFD_REQUIRE(stage && weights && step && phasor, "dft_accumulate: null pointer");
  WFD_REQUIRE(slots > 0 && m >= 0 && m <= slots && nf > 0 && count > 0 && steps > 0,
              "dft_accumulate: bad sizes");
  if (m == 0) return WFD_OK;
  dim3 grid((6 * count + kBlock - 1) / kBlock, (nf + kFreqGroup - 1) / kFreqGroup);
  hipLaunchKernelGGL(k_dft_accumulate, grid, dim3(kBlock), 0, as_stream(stream), stage, slots,
                     weights, steps, nf, step, m, phasor, count);
  WFD_LAUNCH_CHECK("dft_accumulate");
  return WFD_OK;
}

int wfd_record_volume(const float* e, const float* h, const float* prev,
                      const float* positions, const float* weights, const int* step,
                      float* phasor, int count, int nx, int ny, int nz, int first,
                      void* stream) {
  WFD_REQUIRE(e && h && prev && positions && weights && step && phasor,
              "record_volume: null pointer");
  WFD_REQUIRE(count > 0 && nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells,
              "record_volume: bad sizes");
  hipLaunchKernelGGL(k_record_volume, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), e, h, prev, positions, weights, step, phasor, count,
                     nx, ny, nz, first);
  WFD_LAUNCH_CHECK("record_volume");
  return WFD_OK;
}

int wfd_record_mode(const float* e, float* signal, int* step, void* stream) {
  WFD_REQUIRE(e && signal && step, "record_mode: null pointer");
  hipLaunchKernelGGL(k_record_mode, dim3(1), dim3(1), 0, as_stream(stream), e, signal, step);
  WFD_LAUNCH_CHECK("record_mode");
  return WFD_OK;
}

int wfd_observer_gather(const int* indices, const float* weights, const float* field,
                        const float* old, int center, float* out, int rows, void* stream) {
  WFD_REQUIRE(indices && weights && field && out && rows > 0, "observer_gather: bad args");
  WFD_REQUIRE(center == 0 || old, "observer_gather: centering needs the old field");
  hipLaunchKernelGGL(k_observer_gather, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), indices, weights, field, old, center, out, rows);
  WFD_LAUNCH_CHECK("observer_gather");
  return WFD_OK;
}

int wfd_count_nonfinite(const float* values, int64_t n, int* count, void* stream) {
  WFD_REQUIRE(values && count && n >= 0, "count_nonfinite: bad args");
  if (n == 0) return WFD_OK;
  int64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_count_nonfinite, dim3((unsigned)blocks), dim3(kBlock), 0,
                     as_stream(stream), values, n, count);
  WFD_LAUNCH_CHECK("count_nonfinite");
  return WFD_OK;
}

}  // extern "C"
