// Exact discrete adjoint of the Yee + CPML step and detector colocation for
// gfx950 (warp_fdtd/adjoint.py), with an FP64 gradient accumulator variant.
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;

// adjoint.py:17-43
__global__ __launch_bounds__(kBlock) void k_record_curl(
    const float* __restrict__ h, float* __restrict__ hist, int hist_steps,
    const int* __restrict__ step, int nx, int ny, int nz, int x0, int y0, int z0, int sx,
    int sy, int sz, float courant) {
  const int dn = sx * sy * sz;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= dn) return;
  const int n = nx * ny * nz;
  const int x = x0 + i / (sy * sz);
  const int y = y0 + (i / sz) % sy;
  const int z = z0 + i % sz;
  const F3 v = sample(h, n, x, y, z, nx, ny, nz, 0);
  const F3 a = sample(h, n, x - 1, y, z, nx, ny, nz, 0);
  const F3 b = sample(h, n, x, y - 1, z, nx, ny, nz, 0);
  const F3 c = sample(h, n, x, y, z - 1, nx, ny, nz, 0);
  const F3 dx{v.x - a.x, v.y - a.y, v.z - a.z};
  const F3 dy{v.x - b.x, v.y - b.y, v.z - b.z};
  const F3 dz{v.x - c.x, v.y - c.y, v.z - c.z};
  const int64_t base = (int64_t)(step[0] % hist_steps) * 3 * dn + i;
  hist[base] = courant * (dy.z - dz.y);
  hist[base + dn] = courant * (dz.x - dx.z);
  hist[base + 2 * dn] = courant * (dx.y - dy.x);
}

// adjoint.py:46-88 — transposed CPML; d gets the six partial streams.
__global__ __launch_bounds__(kBlock) void k_prepare(
    const float* __restrict__ q, const float* __restrict__ eps, float* __restrict__ psi,
    const float* __restrict__ coeff, int rows, float* __restrict__ d, int nx, int ny, int nz,
    float courant, int electric) {
  const int n = nx * ny * nz;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int x = i / (ny * nz);
  const int y = (i / nz) % ny;
  const int z = i % nz;
  const float factor = electric ? courant * eps[i] : -courant;
  const float v[3] = {factor * q[i], factor * q[n + i], factor * q[2 * n + i]};
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    const int pos = axis == 0 ? x : (axis == 1 ? y : z);
    const float* co = coeff + (axis * rows + pos) * 3;
    const int a = (axis + 1) % 3;
    const int b = (axis + 2) % 3;
    const float q1 = -v[a];
    const float q2 = v[b];
    float d1 = co[2] * q1;
    float d2 = co[2] * q2;
    if (co[0] != 0.f) {
      const float p1 = psi[(2 * axis) * n + i] + q1;
      const float p2 = psi[(2 * axis + 1) * n + i] + q2;
      d1 += co[0] * p1;
      d2 += co[0] * p2;
      psi[(2 * axis) * n + i] = co[1] * p1;
      psi[(2 * axis + 1) * n + i] = co[1] * p2;
    }
    d[(2 * axis) * n + i] = d1;
    d[(2 * axis + 1) * n + i] = d2;
  }
}

// adjoint.py:91-104
__device__ __forceinline__ float read_d(const float* __restrict__ d, int n, int slot, int x,
                                        int y, int z, int nx, int ny, int nz) {
  if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return 0.f;
  return d[slot * n + (x * ny + y) * nz + z];
}

// adjoint.py:107-162
template <typename G>
__global__ __launch_bounds__(kBlock) void k_gather(
    float* __restrict__ q, const float* __restrict__ d, const float* __restrict__ hist,
    int hist_steps, G* __restrict__ grad, const int* __restrict__ step, int nx, int ny,
    int nz, int x0, int y0, int z0, int sx, int sy, int sz, int from_h) {
  const int n = nx * ny * nz;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int x = i / (ny * nz);
  const int y = (i / nz) % ny;
  const int z = i % nz;
  const int delta = from_h ? -1 : 1;
  float v0 = (read_d(d, n, 2, x, y + delta, z, nx, ny, nz) - d[2 * n + i]) +
             (read_d(d, n, 5, x, y, z + delta, nx, ny, nz) - d[5 * n + i]);
  float v1 = (read_d(d, n, 1, x + delta, y, z, nx, ny, nz) - d[n + i]) +
             (read_d(d, n, 4, x, y, z + delta, nx, ny, nz) - d[4 * n + i]);
  float v2 = (read_d(d, n, 0, x + delta, y, z, nx, ny, nz) - d[i]) +
             (read_d(d, n, 3, x, y + delta, z, nx, ny, nz) - d[3 * n + i]);
  if (!from_h) {
    v0 = -v0;
    v1 = -v1;
    v2 = -v2;
  }
  const float q0 = q[i] + v0, q1 = q[n + i] + v1, q2 = q[2 * n + i] + v2;
  q[i] = q0;
  q[n + i] = q1;
  q[2 * n + i] = q2;
  const int t = step[0];
  if (from_h && t >= 0 && x >= x0 && x < x0 + sx && y >= y0 && y < y0 + sy && z >= z0 &&
      z < z0 + sz) {
    const int dn = sx * sy * sz;
    const int j = ((x - x0) * sy + (y - y0)) * sz + (z - z0);
    const int64_t base = (int64_t)(t % hist_steps) * 3 * dn + j;
    const float contrib = hist[base] * q0 + hist[base + dn] * q1 + hist[base + 2 * dn] * q2;
    grad[j] += (G)contrib;
  }
}

// Slots 2*axis and 2*axis + 1 of k_prepare's d at cell j (coordinate `pos`
// along the axis), from q and psi before the update, with k_prepare's
// operations; `p1`, `p2` return the updated psi (used when `active`).
__device__ __forceinline__ void prepare_pair(const float* __restrict__ q,
                                             const float* __restrict__ eps,
                                             const float* __restrict__ psi,
                                             const float* __restrict__ coeff, int rows, int n,
                                             int j, int axis, int pos, float courant, int electric,
                                             float& d1, float& d2, float& p1, float& p2,
                                             bool& active) {
  const float factor = electric ? courant * eps[j] : -courant;
  const int a = (axis + 1) % 3, b = (axis + 2) % 3;
  const float q1 = -(factor * q[a * n + j]);
  const float q2 = factor * q[b * n + j];
  const float* co = coeff + (axis * rows + pos) * 3;
  d1 = co[2] * q1;
  d2 = co[2] * q2;
  active = co[0] != 0.f;
  if (active) {
    const float s1 = psi[(2 * axis) * n + j] + q1;
    const float s2 = psi[(2 * axis + 1) * n + j] + q2;
    d1 += co[0] * s1;
    d2 += co[0] * s2;
    p1 = co[1] * s1;
    p2 = co[1] * s2;
  }
}

// One k_prepare + k_gather pair in a single pass (adjoint.py:46-162): d is not
// stored; each cell computes its own six slots and, for each axis, the two
// slots of its neighbour along that axis (where the gather reads them), from
// q_src and psi_in with k_prepare's operations, then sums them in k_gather's
// order, so q_dst, psi and grad equal the two-kernel pass bit for bit.  The
// updated psi goes to psi_out (a second buffer: neighbours read psi_in), in
// the slots of the axes with CPML at the cell; other slots are never written
// in either buffer.
template <typename G>
__global__ __launch_bounds__(kBlock) void k_reverse_half(
    float* __restrict__ q_dst, const float* __restrict__ q_src, const float* __restrict__ eps,
    const float* __restrict__ psi_in, float* __restrict__ psi_out,
    const float* __restrict__ coeff, int rows, const float* __restrict__ hist, int hist_steps,
    G* __restrict__ grad, const int* __restrict__ step, int nx, int ny, int nz, int x0, int y0,
    int z0, int sx, int sy, int sz, float courant, int from_h) {
  const int n = nx * ny * nz;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int x = i / (ny * nz);
  const int y = (i / nz) % ny;
  const int z = i % nz;
  const int electric = from_h ? 0 : 1;
  const int delta = from_h ? -1 : 1;
  const int c[3] = {x, y, z};
  const int dim[3] = {nx, ny, nz};
  const int stride[3] = {ny * nz, nz, 1};
  float own[6], nb[6];
#pragma unroll
  for (int axis = 0; axis < 3; ++axis) {
    float p1 = 0.f, p2 = 0.f;
    bool active;
    prepare_pair(q_src, eps, psi_in, coeff, rows, n, i, axis, c[axis], courant, electric,
                 own[2 * axis], own[2 * axis + 1], p1, p2, active);
    if (active) {
      psi_out[(2 * axis) * n + i] = p1;
      psi_out[(2 * axis + 1) * n + i] = p2;
    }
    const int cn = c[axis] + delta;
    if (cn >= 0 && cn < dim[axis]) {
      float u1, u2;
      prepare_pair(q_src, eps, psi_in, coeff, rows, n, i + delta * stride[axis], axis, cn,
                   courant, electric, nb[2 * axis], nb[2 * axis + 1], u1, u2, active);
    } else {
      nb[2 * axis] = nb[2 * axis + 1] = 0.f;
    }
  }
  float v0 = (nb[2] - own[2]) + (nb[5] - own[5]);
  float v1 = (nb[1] - own[1]) + (nb[4] - own[4]);
  float v2 = (nb[0] - own[0]) + (nb[3] - own[3]);
  if (!from_h) {
    v0 = -v0;
    v1 = -v1;
    v2 = -v2;
  }
  const float q0 = q_dst[i] + v0, q1 = q_dst[n + i] + v1, q2 = q_dst[2 * n + i] + v2;
  q_dst[i] = q0;
  q_dst[n + i] = q1;
  q_dst[2 * n + i] = q2;
  const int t = step[0];
  if (from_h && t >= 0 && x >= x0 && x < x0 + sx && y >= y0 && y < y0 + sy && z >= z0 &&
      z < z0 + sz) {
    const int dn = sx * sy * sz;
    const int j = ((x - x0) * sy + (y - y0)) * sz + (z - z0);
    const int64_t base = (int64_t)(t % hist_steps) * 3 * dn + j;
    const float contrib = hist[base] * q0 + hist[base + dn] * q1 + hist[base + 2 * dn] * q2;
    grad[j] += (G)contrib;
  }
}

// adjoint.py:165-223 — transposed colocation.  The reference scatters each
// plane cell's six weighted values with atomics; here every target cell of
// the stencil footprint gathers the same contributions in a fixed order, so
// the reverse sweep is bit-reproducible (same sums, deterministic rounding).
struct InjectPlane {
  const float2* source;
  float w0, w1;
  int x0, y0, z0, sx, sy, sz, count;
};

// value of component c at plane cell (x, y, z); 0 outside the plane
__device__ __forceinline__ float plane_value(const InjectPlane& p, int c, int x, int y, int z) {
  if (x < p.x0 || y < p.y0 || z < p.z0 || x >= p.x0 + p.sx || y >= p.y0 + p.sy ||
      z >= p.z0 + p.sz)
    return 0.f;
  const int i = ((x - p.x0) * p.sy + (y - p.y0)) * p.sz + (z - p.z0);
  const float2 s = p.source[c * p.count + i];
  return s.x * p.w0 - s.y * p.w1;
}

// One plane's adjoint source for the reverse step (a wfd_inject_source).
struct InjectSource {
  const float* source;
  const float* weights;
  int x0, y0, z0, sy, sz, count, steps;
};

// The plane's stencil footprint (its target box): x in [x0-1, x0+sx),
// y in [y0-1, y0+sy), z in [z0, z0+sz+1).
__device__ __forceinline__ bool in_targets(const InjectSource& g, int x, int y, int z) {
  const int sx = g.count / (g.sy * g.sz);
  return x >= g.x0 - 1 && x < g.x0 + sx && y >= g.y0 - 1 && y < g.y0 + g.sy && z >= g.z0 &&
         z <= g.z0 + g.sz;
}

// Adds plane g's contributions at step t to target (x, y, z), in a fixed
// order: e[0..2] (unless old_h_only) and h[0..2].  A step with zero weights
// (or outside [0, steps)) adds nothing.
__device__ __forceinline__ void add_plane(const InjectSource& g, int t, int x, int y, int z,
                                          int old_h_only, float* e, float* h) {
  if (t < 0 || t >= g.steps) return;
  InjectPlane p;
  p.w0 = g.weights[2 * t];
  p.w1 = g.weights[2 * t + 1];
  if (!(fabsf(p.w0) + fabsf(p.w1) > 0.f)) return;
  p.source = reinterpret_cast<const float2*>(g.source);
  p.x0 = g.x0;
  p.y0 = g.y0;
  p.z0 = g.z0;
  p.sx = g.count / (g.sy * g.sz);
  p.sy = g.sy;
  p.sz = g.sz;
  p.count = g.count;
  // scatter target (x-dx, y, z+dz) of E_x <- plane cell (x+dx, y, z-dz), etc.
  if (!old_h_only) {
    float a0 = 0.f, a1 = 0.f;
    for (int dz = 0; dz < 2; ++dz) {
      for (int dx = 0; dx < 2; ++dx) a0 += 0.25f * plane_value(p, 0, x + dx, y, z - dz);
      for (int dy = 0; dy < 2; ++dy) a1 += 0.25f * plane_value(p, 1, x, y + dy, z - dz);
    }
    e[0] += a0;
    e[1] += a1;
    e[2] += plane_value(p, 2, x, y, z);
  }
  float b0 = 0.f, b1 = 0.f, b2 = 0.f;
  for (int dy = 0; dy < 2; ++dy) b0 += 0.25f * plane_value(p, 3, x, y + dy, z);
  for (int dx = 0; dx < 2; ++dx) b1 += 0.25f * plane_value(p, 4, x + dx, y, z);
  for (int dx = 0; dx < 2; ++dx)
    for (int dy = 0; dy < 2; ++dy)
      for (int dz = 0; dz < 2; ++dz) b2 += 0.0625f * plane_value(p, 5, x + dx, y + dy, z - dz);
  h[0] += b0;
  h[1] += b1;
  h[2] += b2;
}

// Planes g[0..count) in one launch.  Blocks are split per plane (begin: the
// first block of each plane) over its target box; a target that also lies in
// an earlier plane's box is left to that plane's thread, and the thread that
// owns a target adds every plane containing it in plane order.  Each cell is
// written by one thread, with the roundings of one launch per plane in order.
struct InjectMany {
  InjectSource g[WFD_MAX_INJECT];
  int begin[WFD_MAX_INJECT + 1];
  int count;
};

// Target cell of one thread of k_inject_many's grid (false: none, or a cell
// an earlier plane's thread owns).
__device__ __forceinline__ bool inject_target(const InjectMany& a, int nx, int ny, int nz,
                                              int& k, int& x, int& y, int& z) {
  const int b = blockIdx.x;
  k = 0;
  while (k + 1 < a.count && b >= a.begin[k + 1]) ++k;
  if (a.count == 0) return false;
  const InjectSource& g = a.g[k];
  const int sx = g.count / (g.sy * g.sz);
  const int tsy = g.sy + 1, tsz = g.sz + 1;
  const int r = (b - a.begin[k]) * kBlock + threadIdx.x;
  if (r >= (sx + 1) * tsy * tsz) return false;
  x = g.x0 - 1 + r / (tsy * tsz);
  y = g.y0 - 1 + (r / tsz) % tsy;
  z = g.z0 + r % tsz;
  if (x < 0 || y < 0 || x >= nx || y >= ny || z >= nz) return false;
  for (int j = 0; j < k; ++j)
    if (in_targets(a.g[j], x, y, z)) return false;
  return true;
}

__global__ __launch_bounds__(kBlock) void k_inject_many(InjectMany a, float* __restrict__ qe,
                                                        float* __restrict__ qh,
                                                        const int* __restrict__ step, int nx,
                                                        int ny, int nz, int old_h_only) {
  int k, x, y, z;
  if (!inject_target(a, nx, ny, nz, k, x, y, z)) return;
  const int t = step[0];
  const int n = nx * ny * nz;
  const int i = (x * ny + y) * nz + z;
  float e[3] = {0.f, 0.f, 0.f}, h[3];
  if (!old_h_only)
    for (int c = 0; c < 3; ++c) e[c] = qe[c * n + i];
  for (int c = 0; c < 3; ++c) h[c] = qh[c * n + i];
  for (int l = k; l < a.count; ++l)
    if (l == k || in_targets(a.g[l], x, y, z)) add_plane(a.g[l], t, x, y, z, old_h_only, e, h);
  if (!old_h_only)
    for (int c = 0; c < 3; ++c) qe[c * n + i] = e[c];
  for (int c = 0; c < 3; ++c) qh[c * n + i] = h[c];
}

// The turn between two reverse steps (adjoint.py:388-414 inject(1), retreat,
// then the next step's inject(0)): the H_old terms of step t = step[0] and
// then every term of step t - 1, added per target cell in that order (the
// additions of the two launches they replace), then step[0] = t - 1 by the
// last workgroup to finish (every workgroup read step[0] before its ticket).
__global__ __launch_bounds__(kBlock) void k_inject_turn(InjectMany a, float* __restrict__ qe,
                                                        float* __restrict__ qh,
                                                        int* __restrict__ step,
                                                        int* __restrict__ ticket, int nx,
                                                        int ny, int nz) {
  const int t = step[0];
  int k, x, y, z;
  if (inject_target(a, nx, ny, nz, k, x, y, z)) {
    const int n = nx * ny * nz;
    const int i = (x * ny + y) * nz + z;
    float e[3], h[3];
    for (int c = 0; c < 3; ++c) {
      e[c] = qe[c * n + i];
      h[c] = qh[c * n + i];
    }
    for (int l = k; l < a.count; ++l)
      if (l == k || in_targets(a.g[l], x, y, z)) add_plane(a.g[l], t, x, y, z, 1, e, h);
    for (int l = k; l < a.count; ++l)
      if (l == k || in_targets(a.g[l], x, y, z)) add_plane(a.g[l], t - 1, x, y, z, 0, e, h);
    for (int c = 0; c < 3; ++c) {
      qe[c * n + i] = e[c];
      qh[c * n + i] = h[c];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(ticket, 1) == (int)gridDim.x - 1) {
    step[0] = t - 1;
    *ticket = 0;
  }
}

// adjoint.py:472-490
__global__ __launch_bounds__(kBlock) void k_write_design(const float* __restrict__ plane,
                                                         float* __restrict__ eps, int ny,
                                                         int nz, int x0, int y0, int z0,
                                                         int sx, int sy, int sz) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= sx * sy * sz) return;
  const int x = i / (sy * sz);
  const int y = (i / sz) % sy;
  const int z = i % sz;
  eps[((x0 + x) * ny + y0 + y) * nz + z0 + z] = plane[x * sy + y];
}

int check_box(int nx, int ny, int nz, int x0, int y0, int z0, int sx, int sy, int sz) {
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells, "bad grid");
  WFD_REQUIRE(sx >= 0 && sy >= 0 && sz >= 0 && x0 >= 0 && y0 >= 0 && z0 >= 0 &&
                  x0 + sx <= nx && y0 + sy <= ny && z0 + sz <= nz,
              "box (%d,%d,%d)+(%d,%d,%d) outside the grid", x0, y0, z0, sx, sy, sz);
  return WFD_OK;
}

template <typename G>
int launch_gather(float* q, const float* d, const float* hist, int hist_steps, G* grad,
                  const int* step, int nx, int ny, int nz, int x0, int y0, int z0, int sx,
                  int sy, int sz, int from_h, void* stream) {
  WFD_REQUIRE(q && d && step, "gather: null pointer");
  if (int s = check_box(nx, ny, nz, x0, y0, z0, sx, sy, sz)) return s;
  const bool has_box = sx * sy * sz > 0;
  WFD_REQUIRE(!(from_h && has_box) || (hist && grad && hist_steps > 0),
              "gather: history and gradient required");
  const int n = nx * ny * nz;
  hipLaunchKernelGGL((k_gather<G>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), q, d, hist, hist_steps > 0 ? hist_steps : 1, grad,
                     step, nx, ny, nz, x0, y0, z0, sx, sy, sz, from_h);
  WFD_LAUNCH_CHECK("gather");
  return WFD_OK;
}

template <typename G>
int launch_reverse_half(float* q_dst, const float* q_src, const float* eps, const float* psi_in,
                        float* psi_out, const float* coeff, int coeff_rows, const float* hist,
                        int hist_steps, G* grad, const int* step, int nx, int ny, int nz, int x0,
                        int y0, int z0, int sx, int sy, int sz, float courant, int from_h,
                        void* stream) {
  WFD_REQUIRE(q_dst && q_src && psi_in && psi_out && coeff && step && (from_h || eps),
              "reverse_half: null pointer");
  WFD_REQUIRE(psi_in != psi_out && q_dst != q_src, "reverse_half: buffers must be distinct");
  if (int s = check_box(nx, ny, nz, x0, y0, z0, sx, sy, sz)) return s;
  WFD_REQUIRE(coeff_rows >= nx && coeff_rows >= ny && coeff_rows >= nz,
              "reverse_half: coeff_rows < max(shape)");
  const bool has_box = sx * sy * sz > 0;
  WFD_REQUIRE(!(from_h && has_box) || (hist && grad && hist_steps > 0),
              "reverse_half: history and gradient required");
  const int n = nx * ny * nz;
  hipLaunchKernelGGL((k_reverse_half<G>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), q_dst, q_src, eps, psi_in, psi_out, coeff, coeff_rows,
                     hist, hist_steps > 0 ? hist_steps : 1, grad, step, nx, ny, nz, x0, y0, z0,
                     sx, sy, sz, courant, from_h);
  WFD_LAUNCH_CHECK("reverse_half");
  return WFD_OK;
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_adjoint_reverse_half(float* q_dst, const float* q_src, const float* eps,
                             const float* psi_in, float* psi_out, const float* coeff,
                             int coeff_rows, const float* hist, int hist_steps, float* grad,
                             const int* step, int nx, int ny, int nz, int x0, int y0, int z0,
                             int sx, int sy, int sz, float courant, int from_h, void* stream) {
  return launch_reverse_half<float>(q_dst, q_src, eps, psi_in, psi_out, coeff, coeff_rows, hist,
                                    hist_steps, grad, step, nx, ny, nz, x0, y0, z0, sx, sy, sz,
                                    courant, from_h, stream);
}

int wfd_adjoint_reverse_half_f64(float* q_dst, const float* q_src, const float* eps,
                                 const float* psi_in, float* psi_out, const float* coeff,
                                 int coeff_rows, const float* hist, int hist_steps, double* grad,
                                 const int* step, int nx, int ny, int nz, int x0, int y0, int z0,
                                 int sx, int sy, int sz, float courant, int from_h,
                                 void* stream) {
  return launch_reverse_half<double>(q_dst, q_src, eps, psi_in, psi_out, coeff, coeff_rows,
                                     hist, hist_steps, grad, step, nx, ny, nz, x0, y0, z0, sx,
                                     sy, sz, courant, from_h, stream);
}


int wfd_record_curl(const float* h, float* hist, int hist_steps, const int* step, int nx,
                    int ny, int nz, int x0, int y0, int z0, int sx, int sy, int sz,
                    float courant, void* stream) {
  WFD_REQUIRE(h && hist && step && hist_steps > 0, "record_curl: bad args");
  if (int s = check_box(nx, ny, nz, x0, y0, z0, sx, sy, sz)) return s;
  const int dn = sx * sy * sz;
  if (dn == 0) return WFD_OK;
  hipLaunchKernelGGL(k_record_curl, dim3((dn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), h, hist, hist_steps, step, nx, ny, nz, x0, y0, z0, sx,
                     sy, sz, courant);
  WFD_LAUNCH_CHECK("record_curl");
  return WFD_OK;
}

int wfd_adjoint_prepare(const float* q, const float* eps, float* psi, const float* coeff,
                        int coeff_rows, float* d, int nx, int ny, int nz, float courant,
                        int electric, void* stream) {
  WFD_REQUIRE(q && psi && coeff && d && (!electric || eps), "prepare: null pointer");
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells, "prepare: bad grid");
  WFD_REQUIRE(coeff_rows >= nx && coeff_rows >= ny && coeff_rows >= nz,
              "prepare: coeff_rows < max(shape)");
  const int n = nx * ny * nz;
  hipLaunchKernelGGL(k_prepare, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), q, eps, psi, coeff, coeff_rows, d, nx, ny, nz, courant,
                     electric);
  WFD_LAUNCH_CHECK("prepare");
  return WFD_OK;
}

int wfd_adjoint_gather(float* q, const float* d, const float* hist, int hist_steps,
                       float* grad, const int* step, int nx, int ny, int nz, int x0, int y0,
                       int z0, int sx, int sy, int sz, int from_h, void* stream) {
  return launch_gather<float>(q, d, hist, hist_steps, grad, step, nx, ny, nz, x0, y0, z0, sx,
                              sy, sz, from_h, stream);
}

int wfd_adjoint_gather_f64(float* q, const float* d, const float* hist, int hist_steps,
                           double* grad, const int* step, int nx, int ny, int nz, int x0,
                           int y0, int z0, int sx, int sy, int sz, int from_h, void* stream) {
  return launch_gather<double>(q, d, hist, hist_steps, grad, step, nx, ny, nz, x0, y0, z0,
                               sx, sy, sz, from_h, stream);
}

int wfd_adjoint_inject(float* qe, float* qh, const float* source, const float* weights,
                       const int* step, int nx, int ny, int nz, int x0, int y0, int z0,
                       int sy, int sz, int count, int old_h_only, int steps, void* stream) {
  WFD_REQUIRE(qe && qh && source && weights && step, "inject: null pointer");
  WFD_REQUIRE(count > 0 && sy > 0 && sz > 0 && count % (sy * sz) == 0, "inject: bad plane");
  const wfd_inject_source one{source, weights, x0, y0, z0, sy, sz, count, steps};
  return wfd_adjoint_inject_many(&one, 1, qe, qh, step, nx, ny, nz, old_h_only, stream);
}

int wfd_adjoint_inject_many(const wfd_inject_source* sources, int count, float* qe,
                            float* qh, const int* step, int nx, int ny, int nz, int old_h_only,
                            void* stream) {
  WFD_REQUIRE(sources && qe && qh && step, "inject_many: null pointer");
  WFD_REQUIRE(count >= 0 && count <= WFD_MAX_INJECT, "inject_many: 0..%d planes",
              WFD_MAX_INJECT);
  InjectMany a;
  a.count = count;
  int blocks = 0;
  for (int k = 0; k < count; ++k) {
    const wfd_inject_source& d = sources[k];
    WFD_REQUIRE(d.source && d.weights && d.steps > 0, "inject_many: plane %d arrays", k);
    WFD_REQUIRE(d.count > 0 && d.sy > 0 && d.sz > 0 && d.count % (d.sy * d.sz) == 0,
                "inject_many: plane %d shape", k);
    const int sx = d.count / (d.sy * d.sz);
    if (int s = check_box(nx, ny, nz, d.x0, d.y0, d.z0, sx, d.sy, d.sz)) return s;
    a.g[k] = InjectSource{d.source, d.weights, d.x0, d.y0, d.z0, d.sy, d.sz, d.count, d.steps};
    a.begin[k] = blocks;
    blocks += ((sx + 1) * (d.sy + 1) * (d.sz + 1) + kBlock - 1) / kBlock;
  }
  a.begin[count] = blocks;
  if (count == 0) return WFD_OK;
  hipLaunchKernelGGL(k_inject_many, dim3(blocks), dim3(kBlock), 0, as_stream(stream), a, qe,
                     qh, step, nx, ny, nz, old_h_only);
  WFD_LAUNCH_CHECK("inject_many");
  return WFD_OK;
}

int wfd_adjoint_inject_turn(const wfd_inject_source* sources, int count, float* qe, float* qh,
                            int* step, int* ticket, int nx, int ny, int nz, void* stream) {
  WFD_REQUIRE((sources || count == 0) && qe && qh && step && ticket,
              "inject_turn: null pointer");
  WFD_REQUIRE(count >= 0 && count <= WFD_MAX_INJECT, "inject_turn: 0..%d planes",
              WFD_MAX_INJECT);
  InjectMany a;
  a.count = count;
  int blocks = 0;
  for (int k = 0; k < count; ++k) {
    const wfd_inject_source& d = sources[k];
    WFD_REQUIRE(d.source && d.weights && d.steps > 0, "inject_turn: plane %d arrays", k);
    WFD_REQUIRE(d.count > 0 && d.sy > 0 && d.sz > 0 && d.count % (d.sy * d.sz) == 0,
                "inject_turn: plane %d shape", k);
    const int sx = d.count / (d.sy * d.sz);
    if (int s = check_box(nx, ny, nz, d.x0, d.y0, d.z0, sx, d.sy, d.sz)) return s;
    a.g[k] = InjectSource{d.source, d.weights, d.x0, d.y0, d.z0, d.sy, d.sz, d.count, d.steps};
    a.begin[k] = blocks;
    blocks += ((sx + 1) * (d.sy + 1) * (d.sz + 1) + kBlock - 1) / kBlock;
  }
  a.begin[count] = blocks;
  if (blocks == 0) blocks = 1;  // still retreats the step
  hipLaunchKernelGGL(k_inject_turn, dim3(blocks), dim3(kBlock), 0, as_stream(stream), a, qe,
                     qh, step, ticket, nx, ny, nz);
  WFD_LAUNCH_CHECK("inject_turn");
  return WFD_OK;
}

int wfd_write_design(const float* plane, float* eps, int nx, int ny, int nz, int x0, int y0,
                     int z0, int sx, int sy, int sz, void* stream) {
  WFD_REQUIRE(plane && eps, "write_design: null pointer");
  if (int s = check_box(nx, ny, nz, x0, y0, z0, sx, sy, sz)) return s;
  const int dn = sx * sy * sz;
  if (dn == 0) return WFD_OK;
  hipLaunchKernelGGL(k_write_design, dim3((dn + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     as_stream(stream), plane, eps, ny, nz, x0, y0, z0, sx, sy, sz);
  WFD_LAUNCH_CHECK("write_design");
  return WFD_OK;
}

}  // extern "C"
