// Yee half-steps for gfx950: the per-cell kernel ABI of warp_fdtd/kernels.py
// and the fast column-marching step used by the solver.
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;

// kernels.py:50-74 cpml_packed: psi row offset of a packed slab plane.
__device__ __forceinline__ int64_t packed_index(int axis, int plane, int i, int nx, int ny,
                                                int nz) {
  const int plane_yz = ny * nz;
  if (axis == 0) return (int64_t)plane * plane_yz + i % plane_yz;
  if (axis == 1) return (int64_t)plane * nx * nz + (int64_t)(i / plane_yz) * nz + i % nz;
  return (int64_t)plane * nx * ny + i / nz;
}

template <bool PACKED>
__device__ __forceinline__ float cpml_ref(float d, float* __restrict__ psi, int64_t psi_stride,
                                          const float* __restrict__ coeff, int rows, int axis,
                                          int position, int slot, int i, int nx, int ny,
                                          int nz) {
  constexpr int W = PACKED ? 4 : 3;
  const float* c = coeff + ((int64_t)axis * rows + position) * W;
  const float a = c[0];
  if (a != 0.f) {
    const int64_t index =
        PACKED ? packed_index(axis, (int)c[3], i, nx, ny, nz) : (int64_t)i;
    float* at = psi + slot * psi_stride + index;
    const float p = c[1] * (*at) + a * d;
    *at = p;
    return c[2] * d + p;
  }
  return c[2] * d;
}

// warp_fdtd/kernels.py:77-142 — one thread per cell, identical operation order.
template <bool ELECTRIC, bool PACKED>
__global__ __launch_bounds__(kBlock) void k_update(
    float* __restrict__ f, const float* __restrict__ g, const float* __restrict__ inv_eps,
    float* __restrict__ psi, int64_t psi_stride, const float* __restrict__ coeff, int rows,
    const float* __restrict__ source, int64_t source_stride,
    const float* __restrict__ waves, int64_t waves_stride, const int* __restrict__ step,
    float* __restrict__ hprev, int save_h_from, int nx, int ny, int nz, int source_x,
    float courant, int periodic) {
  const int n = nx * ny * nz;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int z = i % nz;
  const int y = (i / nz) % ny;
  const int x = i / (ny * nz);
  const F3 v = sample(g, n, x, y, z, nx, ny, nz, periodic);
  F3 dx, dy, dz;
  if (ELECTRIC) {
    const F3 a = sample(g, n, x - 1, y, z, nx, ny, nz, periodic);
    const F3 b = sample(g, n, x, y - 1, z, nx, ny, nz, periodic);
    const F3 c = sample(g, n, x, y, z - 1, nx, ny, nz, periodic);
    dx = F3{v.x - a.x, v.y - a.y, v.z - a.z};
    dy = F3{v.x - b.x, v.y - b.y, v.z - b.z};
    dz = F3{v.x - c.x, v.y - c.y, v.z - c.z};
  } else {
    const F3 a = sample(g, n, x + 1, y, z, nx, ny, nz, periodic);
    const F3 b = sample(g, n, x, y + 1, z, nx, ny, nz, periodic);
    const F3 c = sample(g, n, x, y, z + 1, nx, ny, nz, periodic);
    dx = F3{a.x - v.x, a.y - v.y, a.z - v.z};
    dy = F3{b.x - v.x, b.y - v.y, b.z - v.z};
    dz = F3{c.x - v.x, c.y - v.y, c.z - v.z};
  }
  const float dxz = cpml_ref<PACKED>(dx.z, psi, psi_stride, coeff, rows, 0, x, 0, i, nx, ny, nz);
  const float dxy = cpml_ref<PACKED>(dx.y, psi, psi_stride, coeff, rows, 0, x, 1, i, nx, ny, nz);
  const float dyx = cpml_ref<PACKED>(dy.x, psi, psi_stride, coeff, rows, 1, y, 2, i, nx, ny, nz);
  const float dyz = cpml_ref<PACKED>(dy.z, psi, psi_stride, coeff, rows, 1, y, 3, i, nx, ny, nz);
  const float dzy = cpml_ref<PACKED>(dz.y, psi, psi_stride, coeff, rows, 2, z, 4, i, nx, ny, nz);
  const float dzx = cpml_ref<PACKED>(dz.x, psi, psi_stride, coeff, rows, 2, z, 5, i, nx, ny, nz);
  const F3 curl{dyz - dzy, dzx - dxz, dxy - dyx};
  const F3 old{f[i], f[n + i], f[2 * n + i]};
  const int t = step[0];
  // Steps past the wave table inject nothing (the table has waves_stride entries).
  const bool on = x == source_x && t >= 0 && t < waves_stride;
  F3 value;
  if (ELECTRIC) {
    const float ie = inv_eps[i];
    value = F3{old.x + (courant * curl.x) * ie, old.y + (courant * curl.y) * ie,
               old.z + (courant * curl.z) * ie};
    if (on) {
      const int s = y * nz + z;
      value.y += ((source[2 * source_stride + s] * waves[2 * waves_stride + t]) * courant) * ie;
      value.z -= ((source[source_stride + s] * waves[waves_stride + t]) * courant) * ie;
    }
  } else {
    value = F3{old.x - courant * curl.x, old.y - courant * curl.y, old.z - courant * curl.z};
    if (t >= save_h_from) {
      hprev[i] = old.x;
      hprev[n + i] = old.y;
      hprev[2 * n + i] = old.z;
    }
    if (on) {
      const int s = y * nz + z;
      value.y -= (source[2 * source_stride + s] * waves[2 * waves_stride + t]) * courant;
      value.z += (source[source_stride + s] * waves[waves_stride + t]) * courant;
    }
  }
  f[i] = value.x;
  f[n + i] = value.y;
  f[2 * n + i] = value.z;
}

__global__ void k_advance(int* step, int delta) { step[0] += delta; }

// ------------------------------------------------------------------------
// Fast step.  A block owns a contiguous 256-cell run of one (y, z) plane
// (offset o = y*nz + z, coalesced for any nz) and marches it through a chunk
// of x planes.  The x-neighbour of the difference stencil stays in registers
// across the march; y/z neighbours come from the block's own or adjacent
// rows through L1/L2.  Logical block ids are striped over the 8 XCDs so each
// XCD sweeps a contiguous range of x chunks and keeps its neighbours in its
// own L2.
struct YeeArgs {
  float* e;
  float* h;
  const float* inv_eps;
  float* psi_e;
  float* psi_h;
  const float* coeff_e;
  const float* coeff_h;
  const float* src_e;
  const float* src_h;
  const float* wave_e;
  const float* wave_h;
  int* counters;
  float* hprev;
  int hbox[6];
  int save_h_from;
  int rows;
  int steps;
  int nx, ny, nz;
  int n;
  int P;  // ny*nz
  float inv_nz;
  int source_x;
  float courant;
  int x_chunk;
  int segs;   // plane segments per x chunk
  int total;  // logical blocks
  int per;    // logical blocks per XCD stripe
};

__device__ __forceinline__ void load3(const float* __restrict__ f, int n, int i, float& a,
                                      float& b, float& c) {
  a = f[i];
  b = f[n + i];
  c = f[2 * n + i];
}

__global__ __launch_bounds__(kBlock) void k_yee_e(YeeArgs a) {
  const int bid = blockIdx.x;
  const int logical = (bid & 7) * a.per + (bid >> 3);
  if (logical >= a.total) return;
  const int t = a.counters[0];
  if (logical == 0 && threadIdx.x == 0) a.counters[1] = t;
  const int chunk = logical / a.segs;
  const int o = (logical - chunk * a.segs) * kBlock + threadIdx.x;
  if (o >= a.P) return;
  const int nz = a.nz, n = a.n, P = a.P;
  const int y = div_small(o, nz, a.inv_nz);
  const int z = o - y * nz;
  const int x0 = chunk * a.x_chunk;
  const int x1 = min(a.nx, x0 + a.x_chunk);
  const int rows = a.rows;
  const float* __restrict__ ce = a.coeff_e;
  const float ya = ce[(rows + y) * 3], yb = ce[(rows + y) * 3 + 1], yk = ce[(rows + y) * 3 + 2];
  const float za = ce[(2 * rows + z) * 3], zb = ce[(2 * rows + z) * 3 + 1],
              zk = ce[(2 * rows + z) * 3 + 2];
  const float S = a.courant;
  const float* __restrict__ H = a.h;
  float* __restrict__ E = a.e;
  float* __restrict__ psi = a.psi_e;
  float mx, my, mz;  // H(x-1) at this (y, z)
  if (x0 > 0) load3(H, n, (x0 - 1) * P + o, mx, my, mz);
  else mx = my = mz = 0.f;
  float wave_y = 0.f, wave_z = 0.f;
  if (a.source_x >= x0 && a.source_x < x1 && t < a.steps) {
    wave_y = a.wave_h[2 * a.steps + t];
    wave_z = a.wave_h[a.steps + t];
  }
  for (int x = x0; x < x1; ++x) {
    const int i = x * P + o;
    float hx, hy, hz;
    load3(H, n, i, hx, hy, hz);
    float yx = 0.f, yy = 0.f, yz = 0.f, zx = 0.f, zy = 0.f, zz = 0.f;
    if (y > 0) load3(H, n, i - nz, yx, yy, yz);
    if (z > 0) load3(H, n, i - 1, zx, zy, zz);
    const float dx1 = hy - my, dx2 = hz - mz;
    const float dy0 = hx - yx, dy2 = hz - yz;
    const float dz0 = hx - zx, dz1 = hy - zy;
    const float xa = ce[x * 3], xb = ce[x * 3 + 1], xk = ce[x * 3 + 2];
    const float dxz = cpml_term(dx2, xa, xb, xk, psi + i);
    const float dxy = cpml_term(dx1, xa, xb, xk, psi + n + i);
    const float dyx = cpml_term(dy0, ya, yb, yk, psi + 2 * n + i);
    const float dyz = cpml_term(dy2, ya, yb, yk, psi + 3 * n + i);
    const float dzy = cpml_term(dz1, za, zb, zk, psi + 4 * n + i);
    const float dzx = cpml_term(dz0, za, zb, zk, psi + 5 * n + i);
    const float ie = a.inv_eps[i];
    float ex, ey, ez;
    load3(E, n, i, ex, ey, ez);
    ex = ex + (S * (dyz - dzy)) * ie;
    ey = ey + (S * (dzx - dxz)) * ie;
    ez = ez + (S * (dxy - dyx)) * ie;
    if (x == a.source_x) {
      ey += ((a.src_h[2 * P + o] * wave_y) * S) * ie;
      ez -= ((a.src_h[P + o] * wave_z) * S) * ie;
    }
    E[i] = ex;
    E[n + i] = ey;
    E[2 * n + i] = ez;
    mx = hx;
    my = hy;
    mz = hz;
  }
}

__global__ __launch_bounds__(kBlock) void k_yee_h(YeeArgs a) {
  const int bid = blockIdx.x;
  const int logical = (bid & 7) * a.per + (bid >> 3);
  if (logical >= a.total) return;
  const int t = a.counters[1];
  if (logical == 0 && threadIdx.x == 0) a.counters[0] = t + 1;
  const int chunk = logical / a.segs;
  const int o = (logical - chunk * a.segs) * kBlock + threadIdx.x;
  if (o >= a.P) return;
  const int nz = a.nz, ny = a.ny, nx = a.nx, n = a.n, P = a.P;
  const int y = div_small(o, nz, a.inv_nz);
  const int z = o - y * nz;
  const int x0 = chunk * a.x_chunk;
  const int x1 = min(nx, x0 + a.x_chunk);
  const int rows = a.rows;
  const float* __restrict__ ch = a.coeff_h;
  const float ya = ch[(rows + y) * 3], yb = ch[(rows + y) * 3 + 1], yk = ch[(rows + y) * 3 + 2];
  const float za = ch[(2 * rows + z) * 3], zb = ch[(2 * rows + z) * 3 + 1],
              zk = ch[(2 * rows + z) * 3 + 2];
  const float S = a.courant;
  const float* __restrict__ E = a.e;
  float* __restrict__ H = a.h;
  float* __restrict__ psi = a.psi_h;
  const bool save = t >= a.save_h_from && y >= a.hbox[2] && y < a.hbox[3] &&
                    z >= a.hbox[4] && z < a.hbox[5];
  float sx, sy, sz;  // E(x) at this (y, z)
  load3(E, n, x0 * P + o, sx, sy, sz);
  float wave_y = 0.f, wave_z = 0.f;
  if (a.source_x >= x0 && a.source_x < x1 && t < a.steps) {
    wave_y = a.wave_e[2 * a.steps + t];
    wave_z = a.wave_e[a.steps + t];
  }
  for (int x = x0; x < x1; ++x) {
    const int i = x * P + o;
    float px = 0.f, py = 0.f, pz = 0.f, qx = 0.f, qy = 0.f, qz = 0.f, rx = 0.f, ry = 0.f,
          rz = 0.f;
    if (x + 1 < nx) load3(E, n, i + P, px, py, pz);
    if (y + 1 < ny) load3(E, n, i + nz, qx, qy, qz);
    if (z + 1 < nz) load3(E, n, i + 1, rx, ry, rz);
    const float dx1 = py - sy, dx2 = pz - sz;
    const float dy0 = qx - sx, dy2 = qz - sz;
    const float dz0 = rx - sx, dz1 = ry - sy;
    const float xa = ch[x * 3], xb = ch[x * 3 + 1], xk = ch[x * 3 + 2];
    const float dxz = cpml_term(dx2, xa, xb, xk, psi + i);
    const float dxy = cpml_term(dx1, xa, xb, xk, psi + n + i);
    const float dyx = cpml_term(dy0, ya, yb, yk, psi + 2 * n + i);
    const float dyz = cpml_term(dy2, ya, yb, yk, psi + 3 * n + i);
    const float dzy = cpml_term(dz1, za, zb, zk, psi + 4 * n + i);
    const float dzx = cpml_term(dz0, za, zb, zk, psi + 5 * n + i);
    float hx, hy, hz;
    load3(H, n, i, hx, hy, hz);
    if (save && x >= a.hbox[0] && x < a.hbox[1]) {
      a.hprev[i] = hx;
      a.hprev[n + i] = hy;
      a.hprev[2 * n + i] = hz;
    }
    hx = hx - S * (dyz - dzy);
    hy = hy - S * (dzx - dxz);
    hz = hz - S * (dxy - dyx);
    if (x == a.source_x) {
      hy -= (a.src_e[2 * P + o] * wave_y) * S;
      hz += (a.src_e[P + o] * wave_z) * S;
    }
    H[i] = hx;
    H[n + i] = hy;
    H[2 * n + i] = hz;
    sx = px;
    sy = py;
    sz = pz;
  }
}

template <bool ELECTRIC, bool PACKED>
int launch_update(float* f, const float* g, const float* inv_eps, float* psi,
                  int64_t psi_stride, const float* coeff, int rows, const float* source,
                  int64_t source_stride, const float* waves, int64_t waves_stride,
                  const int* step, float* hprev, int save_h_from, int nx, int ny, int nz,
                  int source_x, float courant, int periodic, void* stream) {
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0, "update: grid dimensions must be positive");
  WFD_REQUIRE(cells(nx, ny, nz) <= kMaxCells, "update: grid exceeds 2^31 cells");
  WFD_REQUIRE(f && g && coeff && step, "update: null field, coeff or step pointer");
  WFD_REQUIRE(rows >= nx && rows >= ny && rows >= nz, "update: coeff_rows < max(shape)");
  WFD_REQUIRE(!ELECTRIC || inv_eps, "update_e: null inv_eps");
  WFD_REQUIRE(source_x < 0 || (source && waves), "update: source plane without tables");
  WFD_REQUIRE(source_x < 0 || source_stride >= (int64_t)ny * nz,
              "update: source stride smaller than ny*nz");
  const int n = nx * ny * nz;
  const int blocks = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((k_update<ELECTRIC, PACKED>), dim3(blocks), dim3(kBlock), 0,
                     as_stream(stream), f, g, inv_eps, psi, psi_stride, coeff, rows, source,
                     source_stride, waves, waves_stride, step, hprev, save_h_from, nx, ny, nz,
                     source_x, courant, periodic);
  WFD_LAUNCH_CHECK(ELECTRIC ? "update_e" : "update_h");
  return WFD_OK;
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

#define WFD_UPDATE_ENTRY(NAME, E, PK)                                                      \
  int NAME(float* f, const float* g, const float* inv_eps, float* psi, int64_t psi_stride, \
           const float* coeff, int coeff_rows, const float* source, int64_t source_stride, \
           const float* waves, int64_t waves_stride, const int* step, float* hprev,        \
           int save_h_from, int nx, int ny, int nz, int source_x, float courant,           \
           int periodic, void* stream) {                                                   \
    return launch_update<E, PK>(f, g, inv_eps, psi, psi_stride, coeff, coeff_rows, source, \
                                source_stride, waves, waves_stride, step, hprev,           \
                                save_h_from, nx, ny, nz, source_x, courant, periodic,      \
                                stream);                                                   \
  }

WFD_UPDATE_ENTRY(wfd_update_e, true, false)
WFD_UPDATE_ENTRY(wfd_update_h, false, false)
WFD_UPDATE_ENTRY(wfd_update_e_packed, true, true)
WFD_UPDATE_ENTRY(wfd_update_h_packed, false, true)

int wfd_advance(int* step, int delta, void* stream) {
  WFD_REQUIRE(step, "advance: null step");
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, as_stream(stream), step, delta);
  WFD_LAUNCH_CHECK("advance");
  return WFD_OK;
}

int wfd_yee_step(const wfd_yee_plan* p, void* stream) {
  WFD_REQUIRE(p, "yee_step: null plan");
  WFD_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0, "yee_step: grid dimensions must be positive");
  WFD_REQUIRE(cells(p->nx, p->ny, p->nz) <= kMaxCells, "yee_step: grid exceeds 2^31 cells");
  WFD_REQUIRE((int64_t)p->ny * p->nz < (1 << 24), "yee_step: ny*nz must be < 2^24");
  WFD_REQUIRE(p->e && p->h && p->inv_eps && p->psi_e && p->psi_h && p->coeff_e &&
                  p->coeff_h && p->counters,
              "yee_step: null field, psi, coeff or counter pointer");
  WFD_REQUIRE(p->coeff_rows >= p->nx && p->coeff_rows >= p->ny && p->coeff_rows >= p->nz,
              "yee_step: coeff_rows < max(shape)");
  WFD_REQUIRE(p->source_x < p->nx, "yee_step: source plane outside the grid");
  WFD_REQUIRE(p->source_x < 0 ||
                  (p->src_e && p->src_h && p->wave_e && p->wave_h && p->steps > 0),
              "yee_step: source plane without profiles/waves");
  WFD_REQUIRE(p->save_h_from >= p->steps || p->hprev, "yee_step: hprev required");
  YeeArgs a;
  a.e = p->e;
  a.h = p->h;
  a.inv_eps = p->inv_eps;
  a.psi_e = p->psi_e;
  a.psi_h = p->psi_h;
  a.coeff_e = p->coeff_e;
  a.coeff_h = p->coeff_h;
  a.src_e = p->src_e;
  a.src_h = p->src_h;
  a.wave_e = p->wave_e;
  a.wave_h = p->wave_h;
  a.counters = p->counters;
  a.hprev = p->hprev;
  for (int k = 0; k < 6; ++k) a.hbox[k] = p->hbox[k];
  a.save_h_from = p->save_h_from;
  a.rows = p->coeff_rows;
  a.steps = p->steps;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.n = p->nx * p->ny * p->nz;
  a.P = p->ny * p->nz;
  a.inv_nz = 1.0f / (float)p->nz;
  a.source_x = p->source_x;
  a.courant = p->courant;
  a.x_chunk = p->x_chunk > 0 ? p->x_chunk : 4;
  a.segs = (a.P + kBlock - 1) / kBlock;
  const int chunks = (a.nx + a.x_chunk - 1) / a.x_chunk;
  a.total = a.segs * chunks;
  a.per = (a.total + 7) / 8;
  const int blocks = a.per * 8;
  hipLaunchKernelGGL(k_yee_e, dim3(blocks), dim3(kBlock), 0, as_stream(stream), a);
  WFD_LAUNCH_CHECK("yee_step(E)");
  hipLaunchKernelGGL(k_yee_h, dim3(blocks), dim3(kBlock), 0, as_stream(stream), a);
  WFD_LAUNCH_CHECK("yee_step(H)");
  return WFD_OK;
}

}  // extern "C"
