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
// (offset o = y*nz + z, coalesced for any nz) and a chunk of XC x planes.
// Every load of the chunk is issued before any arithmetic (fully unrolled),
// so a wave pays one memory round trip for XC planes instead of XC dependent
// ones; only the field components the difference stencil uses are loaded,
// and the x-neighbour of plane p is plane p-1 of the same registers.
// Logical block ids are striped over the 8 XCDs so each XCD sweeps a
// contiguous range of x chunks and keeps its neighbours in its own L2.
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
  int nstage;
  int slots;
  StageBox st[WFD_MAX_STAGES];
};

// Stage E (after the step), H (after) and H_old of a footprint cell.
__device__ __forceinline__ void stage_cell(const StageBox* st, int nstage, int slot, int x,
                                           int y, int z, float ex, float ey, float ez,
                                           float hx, float hy, float hz, float ox, float oy,
                                           float oz) {
  for (int b = 0; b < nstage; ++b) {
    const StageBox& s = st[b];
    if (x >= s.x0 && x < s.x1 && y >= s.y0 && y < s.y1 && z >= s.z0 && z < s.z1) {
      const int cnt = s.count;
      float* d = s.buf + (size_t)slot * 9 * cnt +
                 ((x - s.x0) * (s.y1 - s.y0) + (y - s.y0)) * (s.z1 - s.z0) + (z - s.z0);
      d[0] = ex;
      d[cnt] = ey;
      d[2 * cnt] = ez;
      d[3 * cnt] = hx;
      d[4 * cnt] = hy;
      d[5 * cnt] = hz;
      d[6 * cnt] = ox;
      d[7 * cnt] = oy;
      d[8 * cnt] = oz;
    }
  }
}

// Of the boxes in ``boxes`` (bit b for box b), those that can hold cells of
// planes [xa, xb) whose flattened (y, z) offsets lie in [A, B): a box spans
// the offsets from y0 nz + z0 up to (y1 - 1) nz + z1.  Uniform in the
// single-pass kernels, which march one plane at a time over one tile of
// offsets: a workgroup tests every box once against its whole chunk, then
// each plane against the boxes that passed, so planes and tiles that touch
// no box skip the per-cell test of stage_cell (and most workgroups read no
// box inside the march).
__device__ __forceinline__ unsigned stage_hits(const StageBox* st, unsigned boxes, int xa,
                                               int xb, int A, int B, int nz) {
  unsigned hits = 0;
  for (; boxes; boxes &= boxes - 1) {
    const int b = __builtin_ctz(boxes);
    const StageBox& s = st[b];
    if (xa < s.x1 && xb > s.x0 && s.y0 * nz + s.z0 < B && (s.y1 - 1) * nz + s.z1 > A)
      hits |= 1u << b;
  }
  return hits;
}

// stage_cell over the boxes of ``hits`` only.
__device__ __forceinline__ void stage_hit_cell(const StageBox* st, unsigned hits, int slot,
                                               int x, int y, int z, float ex, float ey, float ez,
                                               float hx, float hy, float hz, float ox, float oy,
                                               float oz) {
  for (; hits; hits &= hits - 1) {
    const int b = __builtin_ctz(hits);
    stage_cell(st + b, 1, slot, x, y, z, ex, ey, ez, hx, hy, hz, ox, oy, oz);
  }
}

struct Cell {
  int o, y, z, x0, nvalid;
};

// Decode the XCD-striped block id; false for padding blocks and idle lanes.
__device__ __forceinline__ bool locate(const YeeArgs& a, int bid, int xc, int& logical,
                                       Cell& c) {
  logical = (bid & 7) * a.per + (bid >> 3);
  if (logical >= a.total) return false;
  const int chunk = logical / a.segs;
  c.o = (logical - chunk * a.segs) * kBlock + threadIdx.x;
  c.x0 = chunk * xc;
  c.nvalid = min(xc, a.nx - c.x0);
  if (c.o >= a.P) return true;
  c.y = div_small(c.o, a.nz, a.inv_nz);
  c.z = c.o - c.y * a.nz;
  return true;
}

template <int XC>
__device__ __forceinline__ void yee_e(const YeeArgs& a, int bid) {
  int logical;
  Cell c;
  if (!locate(a, bid, XC, logical, c)) return;
  const int t = a.counters[0];
  if (logical == 0 && threadIdx.x == 0) a.counters[1] = t;
  if (c.o >= a.P) return;
  const int nz = a.nz, n = a.n, P = a.P, o = c.o, y = c.y, z = c.z, x0 = c.x0;
  const int rows = a.rows;
  const float* __restrict__ ce = a.coeff_e;
  const float* __restrict__ H = a.h;
  float* __restrict__ E = a.e;
  float* __restrict__ psi = a.psi_e;
  const float ya = ce[(rows + y) * 3], yb = ce[(rows + y) * 3 + 1], yk = ce[(rows + y) * 3 + 2];
  const float za = ce[(2 * rows + z) * 3], zb = ce[(2 * rows + z) * 3 + 1],
              zk = ce[(2 * rows + z) * 3 + 2];
  // phase 1: every load of the chunk
  float my = 0.f, mz = 0.f;  // H_y, H_z of plane x0-1
  if (x0 > 0) {
    my = H[n + (x0 - 1) * P + o];
    mz = H[2 * n + (x0 - 1) * P + o];
  }
  float hx[XC], hy[XC], hz[XC], yx[XC], yz[XC], zx[XC], zy[XC], ex[XC], ey[XC], ez[XC],
      ie[XC], ps[XC][6];
#pragma unroll
  for (int p = 0; p < XC; ++p) {
    if (p < c.nvalid) {
      const int i = (x0 + p) * P + o;
      hx[p] = H[i];
      hy[p] = H[n + i];
      hz[p] = H[2 * n + i];
      yx[p] = y > 0 ? H[i - nz] : 0.f;
      yz[p] = y > 0 ? H[2 * n + i - nz] : 0.f;
      zx[p] = z > 0 ? H[i - 1] : 0.f;
      zy[p] = z > 0 ? H[n + i - 1] : 0.f;
      ex[p] = E[i];
      ey[p] = E[n + i];
      ez[p] = E[2 * n + i];
      ie[p] = a.inv_eps[i];
      const bool ax = ce[(x0 + p) * 3] != 0.f;
      ps[p][0] = ax ? psi[i] : 0.f;
      ps[p][1] = ax ? psi[n + i] : 0.f;
      ps[p][2] = ya != 0.f ? psi[2 * n + i] : 0.f;
      ps[p][3] = ya != 0.f ? psi[3 * n + i] : 0.f;
      ps[p][4] = za != 0.f ? psi[4 * n + i] : 0.f;
      ps[p][5] = za != 0.f ? psi[5 * n + i] : 0.f;
    }
  }
  const bool src = a.source_x >= x0 && a.source_x < x0 + c.nvalid && t < a.steps;
  float wave_y = 0.f, wave_z = 0.f, sy = 0.f, sz = 0.f;
  if (src) {
    wave_y = a.wave_h[2 * a.steps + t];
    wave_z = a.wave_h[a.steps + t];
    sy = a.src_h[2 * P + o];
    sz = a.src_h[P + o];
  }
  const float S = a.courant;
  // phase 2: arithmetic in the reference's order, then the stores
#pragma unroll
  for (int p = 0; p < XC; ++p) {
    if (p < c.nvalid) {
      const int x = x0 + p;
      const int i = x * P + o;
      const float py = p == 0 ? my : hy[p - 1 < 0 ? 0 : p - 1];
      const float pz = p == 0 ? mz : hz[p - 1 < 0 ? 0 : p - 1];
      const float dx1 = hy[p] - py, dx2 = hz[p] - pz;
      const float dy0 = hx[p] - yx[p], dy2 = hz[p] - yz[p];
      const float dz0 = hx[p] - zx[p], dz1 = hy[p] - zy[p];
      const float xa = ce[x * 3], xb = ce[x * 3 + 1], xk = ce[x * 3 + 2];
      float dxz, dxy, dyx, dyz, dzy, dzx;
      if (xa != 0.f) {
        const float p0 = xb * ps[p][0] + xa * dx2;
        const float p1 = xb * ps[p][1] + xa * dx1;
        psi[i] = p0;
        psi[n + i] = p1;
        dxz = xk * dx2 + p0;
        dxy = xk * dx1 + p1;
      } else {
        dxz = xk * dx2;
        dxy = xk * dx1;
      }
      if (ya != 0.f) {
        const float p2 = yb * ps[p][2] + ya * dy0;
        const float p3 = yb * ps[p][3] + ya * dy2;
        psi[2 * n + i] = p2;
        psi[3 * n + i] = p3;
        dyx = yk * dy0 + p2;
        dyz = yk * dy2 + p3;
      } else {
        dyx = yk * dy0;
        dyz = yk * dy2;
      }
      if (za != 0.f) {
        const float p4 = zb * ps[p][4] + za * dz1;
        const float p5 = zb * ps[p][5] + za * dz0;
        psi[4 * n + i] = p4;
        psi[5 * n + i] = p5;
        dzy = zk * dz1 + p4;
        dzx = zk * dz0 + p5;
      } else {
        dzy = zk * dz1;
        dzx = zk * dz0;
      }
      float vx = ex[p] + (S * (dyz - dzy)) * ie[p];
      float vy = ey[p] + (S * (dzx - dxz)) * ie[p];
      float vz = ez[p] + (S * (dxy - dyx)) * ie[p];
      if (src && x == a.source_x) {
        vy += ((sy * wave_y) * S) * ie[p];
        vz -= ((sz * wave_z) * S) * ie[p];
      }
      E[i] = vx;
      E[n + i] = vy;
      E[2 * n + i] = vz;
    }
  }
}

template <int XC>
__device__ __forceinline__ void yee_h(const YeeArgs& a, int bid) {
  int logical;
  Cell c;
  if (!locate(a, bid, XC, logical, c)) return;
  const int t = a.counters[1];
  if (logical == 0 && threadIdx.x == 0) a.counters[0] = t + 1;
  if (c.o >= a.P) return;
  const int nz = a.nz, ny = a.ny, nx = a.nx, n = a.n, P = a.P, o = c.o, y = c.y, z = c.z,
            x0 = c.x0;
  const int rows = a.rows;
  const float* __restrict__ ch = a.coeff_h;
  const float* __restrict__ E = a.e;
  float* __restrict__ H = a.h;
  float* __restrict__ psi = a.psi_h;
  const float ya = ch[(rows + y) * 3], yb = ch[(rows + y) * 3 + 1], yk = ch[(rows + y) * 3 + 2];
  const float za = ch[(2 * rows + z) * 3], zb = ch[(2 * rows + z) * 3 + 1],
              zk = ch[(2 * rows + z) * 3 + 2];
  // phase 1: loads.  E planes x0..x0+XC (the last one is the x+1 neighbour).
  float ex[XC + 1], ey[XC + 1], ez[XC + 1];
  float qx[XC], qz[XC], rx[XC], ry[XC], hx[XC], hy[XC], hz[XC], ps[XC][6];
  {
    const int i = x0 * P + o;
    ex[0] = E[i];
    ey[0] = E[n + i];
    ez[0] = E[2 * n + i];
  }
#pragma unroll
  for (int p = 0; p < XC; ++p) {
    ex[p + 1] = ey[p + 1] = ez[p + 1] = 0.f;
    if (p < c.nvalid) {
      const int x = x0 + p;
      const int i = x * P + o;
      if (x + 1 < nx) {
        if (p + 1 < XC) ex[p + 1] = E[i + P];
        ey[p + 1] = E[n + i + P];
        ez[p + 1] = E[2 * n + i + P];
      }
      qx[p] = y + 1 < ny ? E[i + nz] : 0.f;
      qz[p] = y + 1 < ny ? E[2 * n + i + nz] : 0.f;
      rx[p] = z + 1 < nz ? E[i + 1] : 0.f;
      ry[p] = z + 1 < nz ? E[n + i + 1] : 0.f;
      hx[p] = H[i];
      hy[p] = H[n + i];
      hz[p] = H[2 * n + i];
      const bool ax = ch[x * 3] != 0.f;
      ps[p][0] = ax ? psi[i] : 0.f;
      ps[p][1] = ax ? psi[n + i] : 0.f;
      ps[p][2] = ya != 0.f ? psi[2 * n + i] : 0.f;
      ps[p][3] = ya != 0.f ? psi[3 * n + i] : 0.f;
      ps[p][4] = za != 0.f ? psi[4 * n + i] : 0.f;
      ps[p][5] = za != 0.f ? psi[5 * n + i] : 0.f;
    }
  }
  const bool src = a.source_x >= x0 && a.source_x < x0 + c.nvalid && t < a.steps;
  float wave_y = 0.f, wave_z = 0.f, sy = 0.f, sz = 0.f;
  if (src) {
    wave_y = a.wave_e[2 * a.steps + t];
    wave_z = a.wave_e[a.steps + t];
    sy = a.src_e[2 * P + o];
    sz = a.src_e[P + o];
  }
  const bool save = t >= a.save_h_from && y >= a.hbox[2] && y < a.hbox[3] &&
                    z >= a.hbox[4] && z < a.hbox[5];
  const float S = a.courant;
#pragma unroll
  for (int p = 0; p < XC; ++p) {
    if (p < c.nvalid) {
      const int x = x0 + p;
      const int i = x * P + o;
      const float dx1 = ey[p + 1] - ey[p], dx2 = ez[p + 1] - ez[p];
      const float dy0 = qx[p] - ex[p], dy2 = qz[p] - ez[p];
      const float dz0 = rx[p] - ex[p], dz1 = ry[p] - ey[p];
      const float xa = ch[x * 3], xb = ch[x * 3 + 1], xk = ch[x * 3 + 2];
      float dxz, dxy, dyx, dyz, dzy, dzx;
      if (xa != 0.f) {
        const float p0 = xb * ps[p][0] + xa * dx2;
        const float p1 = xb * ps[p][1] + xa * dx1;
        psi[i] = p0;
        psi[n + i] = p1;
        dxz = xk * dx2 + p0;
        dxy = xk * dx1 + p1;
      } else {
        dxz = xk * dx2;
        dxy = xk * dx1;
      }
      if (ya != 0.f) {
        const float p2 = yb * ps[p][2] + ya * dy0;
        const float p3 = yb * ps[p][3] + ya * dy2;
        psi[2 * n + i] = p2;
        psi[3 * n + i] = p3;
        dyx = yk * dy0 + p2;
        dyz = yk * dy2 + p3;
      } else {
        dyx = yk * dy0;
        dyz = yk * dy2;
      }
      if (za != 0.f) {
        const float p4 = zb * ps[p][4] + za * dz1;
        const float p5 = zb * ps[p][5] + za * dz0;
        psi[4 * n + i] = p4;
        psi[5 * n + i] = p5;
        dzy = zk * dz1 + p4;
        dzx = zk * dz0 + p5;
      } else {
        dzy = zk * dz1;
        dzx = zk * dz0;
      }
      if (save && x >= a.hbox[0] && x < a.hbox[1]) {
        a.hprev[i] = hx[p];
        a.hprev[n + i] = hy[p];
        a.hprev[2 * n + i] = hz[p];
      }
      float vx = hx[p] - S * (dyz - dzy);
      float vy = hy[p] - S * (dzx - dxz);
      float vz = hz[p] - S * (dxy - dyx);
      if (src && x == a.source_x) {
        vy -= (sy * wave_y) * S;
        vz += (sz * wave_z) * S;
      }
      H[i] = vx;
      H[n + i] = vy;
      H[2 * n + i] = vz;
      if (a.nstage)
        stage_cell(a.st, a.nstage, t % a.slots, x, y, z, ex[p], ey[p], ez[p], vx, vy, vz, hx[p],
                   hy[p], hz[p]);
    }
  }
}

template <int XC>
__global__ __launch_bounds__(kBlock) void k_yee_e(YeeArgs a) {
  yee_e<XC>(a, blockIdx.x);
}

template <int XC>
__global__ __launch_bounds__(kBlock) void k_yee_h(YeeArgs a) {
  yee_h<XC>(a, blockIdx.x);
}

// Batched step: the grid concatenates the (XCD-striped, padded) block ranges
// of `count` plans of one x-chunk class; first[d] is the first block of plan
// d (first[count] = the grid size).  The per-cell arithmetic is the single
// plan's, so every plan's result is bitwise its own wfd_yee_step's.
// The design of block `bid`: the largest d with first[d] <= bid (first is
// nondecreasing, first[0] = 0), by binary search: at most log2(count)
// dependent loads before the block starts.
__device__ __forceinline__ int batch_member(const int* __restrict__ first, int count, int bid) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= bid)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int XC>
__global__ __launch_bounds__(kBlock) void k_batch_e(const YeeArgs* __restrict__ table,
                                                    const int* __restrict__ first, int count) {
  const int d = batch_member(first, count, blockIdx.x);
  yee_e<XC>(table[d], blockIdx.x - first[d]);
}

template <int XC>
__global__ __launch_bounds__(kBlock) void k_batch_h(const YeeArgs* __restrict__ table,
                                                    const int* __restrict__ first, int count) {
  const int d = batch_member(first, count, blockIdx.x);
  yee_h<XC>(table[d], blockIdx.x - first[d]);
}

// The CPML coefficient rows of one cell (solver.py's coeff tables).
struct Rows {
  float xa, xb, xk, ya, yb, yk, za, zb, zk;
  int my, mz;  // lane masks: all ones where the y / z row has CPML
};

// The six CPML partials of one cell (kernels.py:29-47, the fused step's
// operations); ps[] becomes the updated psi where the axis has CPML.  Both
// forms of each term are computed and one selected, so no branch splits a
// wave whose cells straddle a CPML slab.
__device__ __forceinline__ void partials(const Rows& r, float dx1, float dx2, float dy0,
                                         float dy2, float dz0, float dz1, float (&ps)[6],
                                         float& dxz, float& dxy, float& dyx, float& dyz,
                                         float& dzy, float& dzx) {
  const float p0 = r.xb * ps[0] + r.xa * dx2;
  const float p1 = r.xb * ps[1] + r.xa * dx1;
  const float p2 = r.yb * ps[2] + r.ya * dy0;
  const float p3 = r.yb * ps[3] + r.ya * dy2;
  const float p4 = r.zb * ps[4] + r.za * dz1;
  const float p5 = r.zb * ps[5] + r.za * dz0;
  const bool cx = r.xa != 0.f;  // uniform: the x row of the plane
  dxz = cx ? r.xk * dx2 + p0 : r.xk * dx2;
  dxy = cx ? r.xk * dx1 + p1 : r.xk * dx1;
  dyx = pick(r.my, r.yk * dy0 + p2, r.yk * dy0);
  dyz = pick(r.my, r.yk * dy2 + p3, r.yk * dy2);
  dzy = pick(r.mz, r.zk * dz1 + p4, r.zk * dz1);
  dzx = pick(r.mz, r.zk * dz0 + p5, r.zk * dz0);
  ps[0] = cx ? p0 : ps[0];
  ps[1] = cx ? p1 : ps[1];
  ps[2] = pick(r.my, p2, ps[2]);
  ps[3] = pick(r.my, p3, ps[3]);
  ps[4] = pick(r.mz, p4, ps[4]);
  ps[5] = pick(r.mz, p5, ps[5]);
}

// psi of one cell at byte offset v (kAway: none), where its axis has CPML
__device__ __forceinline__ void load_psi6(const Field& psi, int v, const Rows& r,
                                          float (&ps)[6]) {
  const int vx = r.xa != 0.f ? v : kAway, vy = pick(r.my, v, kAway), vz = pick(r.mz, v, kAway);
  ps[0] = psi.at(0, vx);
  ps[1] = psi.at(1, vx);
  ps[2] = psi.at(2, vy);
  ps[3] = psi.at(3, vy);
  ps[4] = psi.at(4, vz);
  ps[5] = psi.at(5, vz);
}

__device__ __forceinline__ void store_psi6(const Field& psi, int v, const Rows& r,
                                           const float (&ps)[6]) {
  const int vx = r.xa != 0.f ? v : kAway, vy = pick(r.my, v, kAway), vz = pick(r.mz, v, kAway);
  psi.set(0, vx, ps[0]);
  psi.set(1, vx, ps[1]);
  psi.set(2, vy, ps[2]);
  psi.set(3, vy, ps[3]);
  psi.set(4, vz, ps[4]);
  psi.set(5, vz, ps[5]);
}

// The y and z CPML rows of one (y, z) offset, fixed along a march in x and so
// loaded once per workgroup (the x row is uniform per plane).
struct YZ {
  float ya, yb, yk, za, zb, zk;
  int my, mz;  // all ones where the row has CPML (ya, za nonzero)
};

__device__ __forceinline__ YZ yz_at(const float* __restrict__ c, int rows, int y, int z) {
  YZ q;
  q.ya = c[(rows + y) * 3];
  q.yb = c[(rows + y) * 3 + 1];
  q.yk = c[(rows + y) * 3 + 2];
  q.za = c[(2 * rows + z) * 3];
  q.zb = c[(2 * rows + z) * 3 + 1];
  q.zk = c[(2 * rows + z) * 3 + 2];
  q.my = q.ya != 0.f ? -1 : 0;
  q.mz = q.za != 0.f ? -1 : 0;
  return q;
}

__device__ __forceinline__ Rows rows_of(const float* __restrict__ c, int x, const YZ& q) {
  Rows r;
  r.xa = c[x * 3];
  r.xb = c[x * 3 + 1];
  r.xk = c[x * 3 + 2];
  r.ya = q.ya;
  r.yb = q.yb;
  r.yk = q.yk;
  r.za = q.za;
  r.zb = q.zb;
  r.zk = q.zk;
  r.my = q.my;
  r.mz = q.mz;
  return r;
}

// ------------------------------------------------------------------------
// Single-pass step: E and H of one step in one launch, reading the step's
// start (E, H, psi) from one buffer set and writing its result to another
// (the caller alternates the two), so that a workgroup can recompute the E
// values its H update needs from the neighbouring tile's cells while that
// tile writes its own results elsewhere.
//
// A workgroup owns a tile of W consecutive (y, z) offsets o in [A, A + W)
// (o = y*nz + z) over a chunk of x planes [x0, x1) and marches x = x0 .. x1:
// at plane x it computes the new E over [A, A + W + nz) -- the nz cells past
// the tile are the y + 1 neighbours of its last row, recomputed and not
// stored -- and puts it in LDS; then the new H of plane x - 1 over the tile
// from the new E of plane x - 1 (LDS: the y + 1 and z + 1 neighbours) and of
// plane x (registers: the x + 1 neighbour).  Plane x1, the next chunk's first,
// is recomputed the same way.  Each cell's E, H and 1/eps are read once and
// its E and H written once (52 B per cell-update plus CPML psi); the halo
// rows and planes are re-read by the neighbouring workgroups through L2.
// Thread t of a T-thread block handles the offsets A + t + T * j, j < J.
struct PassArgs {
  YeeArgs a;     // the step's start (e, h, psi_e, psi_h) and every table
  float* e_out;  // the step's result
  float* h_out;
  float* pe_out;
  float* ph_out;
  int* ticket;   // workgroups finished (the last one advances the counters)
  float* hist;   // adjoint curl history (null: none), as wfd_record_curl writes it
  int hist_steps;
  int hbox[6];   // its box x0, y0, z0, sx, sy, sz
  int W;         // tile cells (W + nz <= J * kBlock)
  int tiles;     // tiles per plane
  int xc;        // x planes per chunk
  int total;     // tiles * chunks
  int per;       // logical workgroups per XCD stripe
};

template <int T, int J>
// (one offset per thread: at most 64 VGPRs, 8 waves per SIMD; the compiler
// would otherwise take 65)
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(J == 1 ? 8 : 1)))
void k_yee_pass(PassArgs pa) {
  constexpr int SLOTS = J * T;
  __shared__ float lds[2][3][SLOTS];  // the new E of planes x - 1 and x
  const YeeArgs& a = pa.a;
  const int logical = (blockIdx.x & 7) * pa.per + (blockIdx.x >> 3);
  const int t = a.counters[0];
  if (logical < pa.total) {
    const int nx = a.nx, ny = a.ny, nz = a.nz, n = a.n, P = a.P, rows = a.rows;
    const int tile = logical % pa.tiles, chunk = logical / pa.tiles;
    const int A = tile * pa.W;
    const int x0 = chunk * pa.xc, x1 = min(x0 + pa.xc, nx);
    const int B = min(A + pa.W, P), Bx = min(A + pa.W + nz, P);
    const float* __restrict__ ce = a.coeff_e;
    const float* __restrict__ ch = a.coeff_h;
    const Field E0(a.e, 3, n), H0(a.h, 3, n), IE(a.inv_eps, 1, n), pe0(a.psi_e, 6, n),
        ph0(a.psi_h, 6, n), E1(pa.e_out, 3, n), H1(pa.h_out, 3, n), pe1(pa.pe_out, 6, n),
        ph1(pa.ph_out, 6, n);
    const float S = a.courant;
    const bool src_on = t < a.steps && a.source_x >= x0 && a.source_x <= x1;
    const float wave_hy = src_on ? a.wave_h[2 * a.steps + t] : 0.f;
    const float wave_hz = src_on ? a.wave_h[a.steps + t] : 0.f;
    const float wave_ey = src_on ? a.wave_e[2 * a.steps + t] : 0.f;
    const float wave_ez = src_on ? a.wave_e[a.steps + t] : 0.f;
    const bool save_t = t >= a.save_h_from;
    // the stage boxes this tile's planes [x0, x1) can touch
    const unsigned tile_boxes = stage_hits(a.st, (1u << a.nstage) - 1, x0, x1, A, B, nz);
    // per slot: offset, position and the values carried along x
    int o[J], y[J], z[J];
    float hp[J][3];  // H at the plane before (the step's start)
    float ep[J][3];  // new E at the plane before
    YZ ye[J], yh[J];
    float sh_y[J], sh_z[J], se_y[J], se_z[J];  // source profiles at the slot's (y, z)
#pragma unroll
    for (int j = 0; j < J; ++j) {
      o[j] = A + threadIdx.x + T * j;
      const int oc = o[j] < Bx ? o[j] : A;
      y[j] = div_small(oc, nz, a.inv_nz);
      z[j] = oc - y[j] * nz;
      ye[j] = yz_at(ce, rows, y[j], z[j]);
      yh[j] = yz_at(ch, rows, y[j], z[j]);
      sh_y[j] = src_on ? a.src_h[2 * P + oc] : 0.f;
      sh_z[j] = src_on ? a.src_h[P + oc] : 0.f;
      se_y[j] = src_on ? a.src_e[2 * P + oc] : 0.f;
      se_z[j] = src_on ? a.src_e[P + oc] : 0.f;
      const int v = o[j] < Bx && x0 > 0 ? ((x0 - 1) * P + o[j]) * 4 : kAway;
      hp[j][0] = 0.f;
      hp[j][1] = H0.at(1, v);
      hp[j][2] = H0.at(2, v);
      ep[j][0] = ep[j][1] = ep[j][2] = 0.f;
    }
    // Every global access of the march is unconditional, masked by an
    // out-of-range offset: with no loads or stores under branches the
    // compiler counts them exactly and waits for a load only, never for the
    // stores issued after it.  (Loading plane x + 1 during the arithmetic of
    // plane x measured no faster: more registers, fewer workgroups per CU.)
    struct In {
      float hn[3], yx, yz, zx, zy, ex[3], ie, ps[6], psh[6];
    };
    // the step's start at plane xl (for its new E) and the H psi of plane
    // xl - 1 (for its new H); zeros where the march does not reach
    auto load = [&](int xl, In (&q)[J]) {
      const bool e_on = xl < nx && xl <= x1, h_on = xl > x0 && xl <= x1;  // uniform
      const int xe = e_on ? xl : 0, xg = h_on ? xl - 1 : 0;
      const float xa = e_on ? ce[xe * 3] : 0.f, hxa = h_on ? ch[xg * 3] : 0.f;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool on = e_on && o[j] < Bx;
        const int v = on ? (xl * P + o[j]) * 4 : kAway;
        const int vy = on && y[j] > 0 ? v - 4 * nz : kAway;
        const int vz = on && z[j] > 0 ? v - 4 : kAway;
        for (int c = 0; c < 3; ++c) q[j].hn[c] = H0.at(c, v);
        q[j].yx = H0.at(0, vy);
        q[j].yz = H0.at(2, vy);
        q[j].zx = H0.at(0, vz);
        q[j].zy = H0.at(1, vz);
        for (int c = 0; c < 3; ++c) q[j].ex[c] = E0.at(c, v);
        q[j].ie = IE.at(0, v);
        const int vx_ = xa != 0.f ? v : kAway, vy_ = ye[j].ya != 0.f ? v : kAway,
                  vz_ = ye[j].za != 0.f ? v : kAway;
        q[j].ps[0] = pe0.at(0, vx_);
        q[j].ps[1] = pe0.at(1, vx_);
        q[j].ps[2] = pe0.at(2, vy_);
        q[j].ps[3] = pe0.at(3, vy_);
        q[j].ps[4] = pe0.at(4, vz_);
        q[j].ps[5] = pe0.at(5, vz_);
        const int w = h_on && o[j] < B ? ((xl - 1) * P + o[j]) * 4 : kAway;
        const int wx = hxa != 0.f ? w : kAway, wy = yh[j].ya != 0.f ? w : kAway,
                  wz = yh[j].za != 0.f ? w : kAway;
        q[j].psh[0] = ph0.at(0, wx);
        q[j].psh[1] = ph0.at(1, wx);
        q[j].psh[2] = ph0.at(2, wy);
        q[j].psh[3] = ph0.at(3, wy);
        q[j].psh[4] = ph0.at(4, wz);
        q[j].psh[5] = ph0.at(5, wz);
      }
    };
#pragma unroll 1
    for (int x = x0; x <= x1; ++x) {
      float* __restrict__ cur = &lds[x & 1][0][0];
      const float* __restrict__ prev = &lds[(x + 1) & 1][0][0];
      const int xh = x - 1;
      const bool e_plane = x < nx, h_plane = x > x0;  // uniform
      // x rows read unconditionally at a clamped plane (scalar loads)
      const int xe = e_plane ? x : nx - 1, xhc = h_plane ? xh : x0;
      const float xa = e_plane ? ce[xe * 3] : 0.f, xb = ce[xe * 3 + 1], xk = ce[xe * 3 + 2];
      const float hxa = h_plane ? ch[xhc * 3] : 0.f, hxb = ch[xhc * 3 + 1],
                  hxk = ch[xhc * 3 + 2];
      In in[J];
      load(x, in);
      float hn[J][3], yx[J], yz[J], zx[J], zy[J], ex[J][3], ie[J], ps[J][6], psh[J][6];
#pragma unroll
      for (int j = 0; j < J; ++j) {
        for (int c = 0; c < 3; ++c) {
          hn[j][c] = in[j].hn[c];
          ex[j][c] = in[j].ex[c];
        }
        yx[j] = in[j].yx;
        yz[j] = in[j].yz;
        zx[j] = in[j].zx;
        zy[j] = in[j].zy;
        ie[j] = in[j].ie;
        for (int c = 0; c < 6; ++c) {
          ps[j][c] = in[j].ps[c];
          psh[j][c] = in[j].psh[c];
        }
      }
      // the new E of plane x over the tile and its y + 1 rows (zero past Bx)
      float en[J][3];
#pragma unroll
      for (int j = 0; j < J; ++j) en[j][0] = en[j][1] = en[j][2] = 0.f;
      if (e_plane) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int v = x < x1 && o[j] < B ? (x * P + o[j]) * 4 : kAway;  // owned: stored
          const float ya = ye[j].ya, yb = ye[j].yb, yk = ye[j].yk;
          const float za = ye[j].za, zb = ye[j].zb, zk = ye[j].zk;
          const float dx1 = hn[j][1] - hp[j][1], dx2 = hn[j][2] - hp[j][2];
          const float dy0 = hn[j][0] - yx[j], dy2 = hn[j][2] - yz[j];
          const float dz0 = hn[j][0] - zx[j], dz1 = hn[j][1] - zy[j];
          // both forms of each CPML term, then a select (no branch holds a
          // load or store)
          const float p0 = xb * ps[j][0] + xa * dx2;
          const float p1 = xb * ps[j][1] + xa * dx1;
          const int vx_ = xa != 0.f ? v : kAway;
          pe1.set(0, vx_, p0);
          pe1.set(1, vx_, p1);
          const float dxz = xa != 0.f ? xk * dx2 + p0 : xk * dx2;
          const float dxy = xa != 0.f ? xk * dx1 + p1 : xk * dx1;
          const float p2 = yb * ps[j][2] + ya * dy0;
          const float p3 = yb * ps[j][3] + ya * dy2;
          const int vy_ = ya != 0.f ? v : kAway;
          pe1.set(2, vy_, p2);
          pe1.set(3, vy_, p3);
          const float dyx = ya != 0.f ? yk * dy0 + p2 : yk * dy0;
          const float dyz = ya != 0.f ? yk * dy2 + p3 : yk * dy2;
          const float p4 = zb * ps[j][4] + za * dz1;
          const float p5 = zb * ps[j][5] + za * dz0;
          const int vz_ = za != 0.f ? v : kAway;
          pe1.set(4, vz_, p4);
          pe1.set(5, vz_, p5);
          const float dzy = za != 0.f ? zk * dz1 + p4 : zk * dz1;
          const float dzx = za != 0.f ? zk * dz0 + p5 : zk * dz0;
          float vx = ex[j][0] + (S * (dyz - dzy)) * ie[j];
          float vy = ex[j][1] + (S * (dzx - dxz)) * ie[j];
          float vz = ex[j][2] + (S * (dxy - dyx)) * ie[j];
          if (src_on && x == a.source_x) {
            vy += ((sh_y[j] * wave_hy) * S) * ie[j];
            vz -= ((sh_z[j] * wave_hz) * S) * ie[j];
          }
          en[j][0] = vx;
          en[j][1] = vy;
          en[j][2] = vz;
          E1.set(0, v, vx);
          E1.set(1, v, vy);
          E1.set(2, v, vz);
          const int s = threadIdx.x + T * j;
          cur[s] = vx;
          cur[SLOTS + s] = vy;
          cur[2 * SLOTS + s] = vz;
        }
      }
      // the new H of plane x - 1 over the tile
      if (h_plane) {
        const unsigned hits = tile_boxes ? stage_hits(a.st, tile_boxes, xh, xh + 1, A, B, nz) : 0u;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const bool own = o[j] < B;
          const int w = own ? (xh * P + o[j]) * 4 : kAway;
          const int s = threadIdx.x + T * j;
          // (the clamps only touch slots past the tile, whose results are dropped)
          const int sy = min(s + nz, SLOTS - 1), sz = min(s + 1, SLOTS - 1);
          const bool up_y = y[j] + 1 < ny, up_z = z[j] + 1 < nz;
          const float qx = up_y ? prev[sy] : 0.f;
          const float qz = up_y ? prev[2 * SLOTS + sy] : 0.f;
          const float rx = up_z ? prev[sz] : 0.f;
          const float ry = up_z ? prev[SLOTS + sz] : 0.f;
          const float dx1 = en[j][1] - ep[j][1], dx2 = en[j][2] - ep[j][2];
          const float dy0 = qx - ep[j][0], dy2 = qz - ep[j][2];
          const float dz0 = rx - ep[j][0], dz1 = ry - ep[j][1];
          const float ya = yh[j].ya, yb = yh[j].yb, yk = yh[j].yk;
          const float za = yh[j].za, zb = yh[j].zb, zk = yh[j].zk;
          const float p0 = hxb * psh[j][0] + hxa * dx2;
          const float p1 = hxb * psh[j][1] + hxa * dx1;
          const int wx = hxa != 0.f ? w : kAway;
          ph1.set(0, wx, p0);
          ph1.set(1, wx, p1);
          const float dxz = hxa != 0.f ? hxk * dx2 + p0 : hxk * dx2;
          const float dxy = hxa != 0.f ? hxk * dx1 + p1 : hxk * dx1;
          const float p2 = yb * psh[j][2] + ya * dy0;
          const float p3 = yb * psh[j][3] + ya * dy2;
          const int wy = ya != 0.f ? w : kAway;
          ph1.set(2, wy, p2);
          ph1.set(3, wy, p3);
          const float dyx = ya != 0.f ? yk * dy0 + p2 : yk * dy0;
          const float dyz = ya != 0.f ? yk * dy2 + p3 : yk * dy2;
          const float p4 = zb * psh[j][4] + za * dz1;
          const float p5 = zb * psh[j][5] + za * dz0;
          const int wz = za != 0.f ? w : kAway;
          ph1.set(4, wz, p4);
          ph1.set(5, wz, p5);
          const float dzy = za != 0.f ? zk * dz1 + p4 : zk * dz1;
          const float dzx = za != 0.f ? zk * dz0 + p5 : zk * dz0;
          const float hx = hp[j][0], hy = hp[j][1], hz = hp[j][2];
          if (own && save_t && xh >= a.hbox[0] && xh < a.hbox[1] && y[j] >= a.hbox[2] &&
              y[j] < a.hbox[3] && z[j] >= a.hbox[4] && z[j] < a.hbox[5]) {
            const int i = xh * P + o[j];
            a.hprev[i] = hx;
            a.hprev[n + i] = hy;
            a.hprev[2 * n + i] = hz;
          }
          float vx = hx - S * (dyz - dzy);
          float vy = hy - S * (dzx - dxz);
          float vz = hz - S * (dxy - dyx);
          if (src_on && xh == a.source_x) {
            vy -= (se_y[j] * wave_ey) * S;
            vz += (se_z[j] * wave_ez) * S;
          }
          H1.set(0, w, vx);
          H1.set(1, w, vy);
          H1.set(2, w, vz);
          if (hits && own)
            stage_hit_cell(a.st, hits, t % a.slots, xh, y[j], z[j], ep[j][0], ep[j][1],
                           ep[j][2], vx, vy, vz, hx, hy, hz);
        }
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        for (int c = 0; c < 3; ++c) {
          hp[j][c] = hn[j][c];
          ep[j][c] = en[j][c];
        }
      }
      __syncthreads();
    }
  }
  // the last workgroup to finish advances the step counters (every workgroup
  // read counters[0] before it took its ticket)
  if (threadIdx.x == 0) {
    const int done = atomicAdd(pa.ticket, 1);
    if (done == (int)gridDim.x - 1) {
      a.counters[0] = t + 1;
      a.counters[1] = t;
      *pa.ticket = 0;
    }
  }
}

// ------------------------------------------------------------------------
// Two steps per pass (temporal blocking): the step above done twice in one
// march, the first step's E and H (E1, H1) and CPML psi never leaving the
// chip.  Per (y, z) offset, in flattened o = y*nz + z, a tile [A, B) of the
// second step's H needs the second step's E on [A, B + nz), the first step's
// H on [A - nz, B + nz) and its E on [A - nz, B + 2 nz); the slots of a
// workgroup cover [A - nz, B + 2 nz).  Along x, the march at plane x computes
// E1(x), H1(x - 1), E2(x - 1) and H2(x - 2); the chunk [x0, x1) of outputs
// needs E1 on [x0 - 1, x1 + 1].  Planes of E1 and E2 go through LDS (double
// buffered: the y + 1 and z + 1 neighbours of the plane before), the plane of
// H1 too (the y - 1 and z - 1 neighbours of its own plane); every value
// needed along x stays in the registers of the thread that holds its
// offset.  Reads E, H, 1/eps once and writes E, H once per two
// cell-updates; psi is read and written once per two steps as well.

// The step's-start values one new-E computation of the two-step pass reads
// at a cell: H and its y - 1 / z - 1 neighbours, E, 1/eps and the E psi.
struct PlaneIn {
  float h[3], yx, yz, zx, zy, e[3], ie, ps[6];
};

// The arrays of a single pass: the step's start and its result.
struct PassFields {
  Field e0, h0, ie, pe0, ph0, e1, h1, pe1, ph1;
  __device__ __forceinline__ PassFields(const PassArgs& pa)
      : e0(opaque(pa.a.e), 3, pa.a.n),
        h0(opaque(pa.a.h), 3, pa.a.n),
        ie(opaque(pa.a.inv_eps), 1, pa.a.n),
        pe0(opaque(pa.a.psi_e), 6, pa.a.n),
        ph0(opaque(pa.a.psi_h), 6, pa.a.n),
        e1(opaque(pa.e_out), 3, pa.a.n),
        h1(opaque(pa.h_out), 3, pa.a.n),
        pe1(opaque(pa.pe_out), 6, pa.a.n),
        ph1(opaque(pa.ph_out), 6, pa.a.n) {}
};

// Plane loads at byte offset v of the cell (kAway: zeros), branch-free.
__device__ __forceinline__ void load_plane(const YeeArgs& a, const PassFields& f, int x, int v,
                                           int y, int z, const YZ& ye, PlaneIn& q) {
  const int vy = v != kAway && y > 0 ? v - 4 * a.nz : kAway;
  const int vz = v != kAway && z > 0 ? v - 4 : kAway;
  q.h[0] = f.h0.at(0, v);
  q.h[1] = f.h0.at(1, v);
  q.h[2] = f.h0.at(2, v);
  q.yx = f.h0.at(0, vy);
  q.yz = f.h0.at(2, vy);
  q.zx = f.h0.at(0, vz);
  q.zy = f.h0.at(1, vz);
  q.e[0] = f.e0.at(0, v);
  q.e[1] = f.e0.at(1, v);
  q.e[2] = f.e0.at(2, v);
  q.ie = f.ie.at(0, v);
  load_psi6(f.pe0, v, rows_of(a.coeff_e, x, ye), q.ps);
}

// wfd_record_curl's value (adjoint.py:17-43) of one cell at the start of step
// t from the differences of H the new-E update uses, when the cell is in the
// history box.
__device__ __forceinline__ void record_curl_at(const PassArgs& pa, int t, int x, int y, int z,
                                               float dx1, float dx2, float dy0, float dy2,
                                               float dz0, float dz1) {
  const int* b = pa.hbox;
  if (x < b[0] || y < b[1] || z < b[2] || x >= b[0] + b[3] || y >= b[1] + b[4] ||
      z >= b[2] + b[5])
    return;
  const float S = pa.a.courant;
  const int dn = b[3] * b[4] * b[5];
  const int j = ((x - b[0]) * b[4] + (y - b[1])) * b[5] + (z - b[2]);
  float* h = pa.hist + (int64_t)(t % pa.hist_steps) * 3 * dn + j;
  h[0] = S * (dy2 - dz1);
  h[dn] = S * (dz0 - dx2);
  h[2 * dn] = S * (dx1 - dy0);
}

template <int T, int J>
__device__ __forceinline__ void pass2_body(const PassArgs& pa, int bid, int nblocks) {
  constexpr int SLOTS = J * T;
  __shared__ float le1[2][3][SLOTS];  // E1 of planes x - 1 and x
  __shared__ float lh1[3][SLOTS];     // H1 of plane x - 1
  __shared__ float le2[2][3][SLOTS];  // E2 of planes x - 2 and x - 1
  const YeeArgs& a = pa.a;
  const int logical = (bid & 7) * pa.per + (bid >> 3);
  const int t = a.counters[0];
  if (logical < pa.total) {
    const int nx = a.nx, ny = a.ny, nz = a.nz, P = a.P, rows = a.rows;
    const int tile = logical % pa.tiles, chunk = logical / pa.tiles;
    const int A = tile * pa.W;
    const int x0 = chunk * pa.xc, x1 = min(x0 + pa.xc, nx);
    const int B = min(A + pa.W, P);
    const int Bh = min(B + nz, P), Be = min(B + 2 * nz, P);  // H1 and E1 range ends
    const int lo = A - nz;                                   // offset of slot 0
    const float* __restrict__ ce = a.coeff_e;
    const float* __restrict__ ch = a.coeff_h;
    const PassFields f(pa);
    const float S = a.courant;
    // sources of the two steps
    const bool src1 = t < a.steps && a.source_x >= x0 - 1 && a.source_x <= x1 + 1;
    const bool src2 = t + 1 < a.steps && a.source_x >= x0 - 1 && a.source_x <= x1 + 1;
    const bool src = src1 || src2;
    const float wh1y = src1 ? a.wave_h[2 * a.steps + t] : 0.f;
    const float wh1z = src1 ? a.wave_h[a.steps + t] : 0.f;
    const float we1y = src1 ? a.wave_e[2 * a.steps + t] : 0.f;
    const float we1z = src1 ? a.wave_e[a.steps + t] : 0.f;
    const float wh2y = src2 ? a.wave_h[2 * a.steps + t + 1] : 0.f;
    const float wh2z = src2 ? a.wave_h[a.steps + t + 1] : 0.f;
    const float we2y = src2 ? a.wave_e[2 * a.steps + t + 1] : 0.f;
    const float we2z = src2 ? a.wave_e[a.steps + t + 1] : 0.f;
    const bool save = t + 1 >= a.save_h_from;
    // the stage boxes this tile's planes [x0, x1) can touch
    const unsigned tile_boxes = stage_hits(a.st, (1u << a.nstage) - 1, x0, x1, A, B, nz);
    const int xs = max(x0 - 1, 0);
    int o[J], y[J], z[J];
    float hp0[J][3], e1p[J][3], pse1[J][6], ie1[J], h1p[J][3], psh1[J][6], e2p[J][3];
    float sh_y[J], sh_z[J], se_y[J], se_z[J];  // source profiles at the slot's (y, z)
    YZ ye[J], yh[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      o[j] = lo + (int)threadIdx.x + T * j;
      const int oc = (o[j] >= 0 && o[j] < Be) ? o[j] : max(A, 0);
      y[j] = div_small(oc, nz, a.inv_nz);
      z[j] = oc - y[j] * nz;
      ye[j] = yz_at(ce, rows, y[j], z[j]);
      yh[j] = yz_at(ch, rows, y[j], z[j]);
      sh_y[j] = src ? a.src_h[2 * P + oc] : 0.f;
      sh_z[j] = src ? a.src_h[P + oc] : 0.f;
      se_y[j] = src ? a.src_e[2 * P + oc] : 0.f;
      se_z[j] = src ? a.src_e[P + oc] : 0.f;
      for (int c = 0; c < 3; ++c) hp0[j][c] = e1p[j][c] = h1p[j][c] = e2p[j][c] = 0.f;
      for (int c = 0; c < 6; ++c) pse1[j][c] = psh1[j][c] = 0.f;
      ie1[j] = 0.f;
      const int v = o[j] >= 0 && o[j] < Be && xs > 0 ? ((xs - 1) * P + o[j]) * 4 : kAway;
      hp0[j][1] = f.h0.at(1, v);
      hp0[j][2] = f.h0.at(2, v);
    }
    // As in the one-step pass, every global load and store of the march is
    // unconditional, masked by an out-of-range offset, and the CPML terms are
    // selected rather than branched on.
#pragma unroll 1
    for (int x = xs; x <= x1 + 1; ++x) {
      // The per-slot values below are loop-invariant; hiding that from the
      // compiler keeps it from hoisting the lane masks derived from them
      // (each a 64-bit scalar register pair) out of the march, which ran it
      // out of scalar registers into spills.
#pragma unroll
      for (int j = 0; j < J; ++j) {
        asm volatile("" : "+v"(o[j]), "+v"(y[j]), "+v"(z[j]));
        asm volatile("" : "+v"(ye[j].my), "+v"(ye[j].mz), "+v"(yh[j].my), "+v"(yh[j].mz));
      }
      const int xh = x - 1, xq = x - 2;
      // uniform: which of the four updates this plane of the march makes
      const bool do_a = x < nx;
      const bool do_b = xh >= x0 - 1 && xh >= 0 && xh < nx;
      const bool do_c = xh >= x0 && xh <= x1 && xh < nx;
      const bool do_d = xq >= x0 && xq < x1;
      const int cx = do_a ? x : 0, cxh = do_b || do_c ? xh : 0, cxq = do_d ? xq : 0;
      float en[J][3], hn[J][3], pse[J][6], iex[J], h1n[J][3], psh[J][6], e2n[J][3];
      PlaneIn in[J];
      PassFields f(pa);  // rebuilt per plane (see opaque)
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool on = do_a && o[j] >= 0 && o[j] < Be;
        load_plane(a, f, cx, on ? (x * P + o[j]) * 4 : kAway, y[j], z[j], ye[j], in[j]);
        // the H psi of plane x - 1, for (b)
        const bool hon = do_b && o[j] >= 0 && o[j] < Bh;
        load_psi6(f.ph0, hon ? (xh * P + o[j]) * 4 : kAway, rows_of(ch, cxh, yh[j]), psh[j]);
      }
      // (a) E1 of plane x over [A - nz, B + 2 nz)
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool on = do_a && o[j] >= 0 && o[j] < Be;
        const Rows r = rows_of(ce, cx, ye[j]);
        for (int c = 0; c < 3; ++c) hn[j][c] = in[j].h[c];
        for (int c = 0; c < 6; ++c) pse[j][c] = in[j].ps[c];
        iex[j] = in[j].ie;
        const float dx1 = hn[j][1] - hp0[j][1], dx2 = hn[j][2] - hp0[j][2];
        const float dy0 = hn[j][0] - in[j].yx, dy2 = hn[j][2] - in[j].yz;
        const float dz0 = hn[j][0] - in[j].zx, dz1 = hn[j][1] - in[j].zy;
        if (pa.hist && on && x >= x0 && x < x1 && o[j] >= A && o[j] < B)
          record_curl_at(pa, t, x, y[j], z[j], dx1, dx2, dy0, dy2, dz0, dz1);
        float dxz, dxy, dyx, dyz, dzy, dzx;
        partials(r, dx1, dx2, dy0, dy2, dz0, dz1, pse[j], dxz, dxy, dyx, dyz, dzy, dzx);
        float vx = in[j].e[0] + (S * (dyz - dzy)) * iex[j];
        float vy = in[j].e[1] + (S * (dzx - dxz)) * iex[j];
        float vz = in[j].e[2] + (S * (dxy - dyx)) * iex[j];
        if (src1 && x == a.source_x) {
          vy += ((sh_y[j] * wh1y) * S) * iex[j];
          vz -= ((sh_z[j] * wh1z) * S) * iex[j];
        }
        en[j][0] = on ? vx : 0.f;
        en[j][1] = on ? vy : 0.f;
        en[j][2] = on ? vz : 0.f;
        const int s = threadIdx.x + T * j;
        le1[x & 1][0][s] = en[j][0];
        le1[x & 1][1][s] = en[j][1];
        le1[x & 1][2][s] = en[j][2];
      }
      // (b) H1 of plane x - 1 over [A - nz, B + nz)
      const unsigned hits1 = tile_boxes && xh >= x0 && xh < x1
                                 ? stage_hits(a.st, tile_boxes, xh, xh + 1, A, B, nz)
                                 : 0u;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool on = do_b && o[j] >= 0 && o[j] < Bh;
        const int s = threadIdx.x + T * j;
        const Rows r = rows_of(ch, cxh, yh[j]);
        // (the clamps only touch slots past the range, whose results are zeroed)
        const int sy = min(s + nz, SLOTS - 1), sz = min(s + 1, SLOTS - 1);
        const bool up_y = y[j] + 1 < ny, up_z = z[j] + 1 < nz;
        const float(*prev)[SLOTS] = le1[xh & 1];
        const float qx = up_y ? prev[0][sy] : 0.f;
        const float qz = up_y ? prev[2][sy] : 0.f;
        const float rx = up_z ? prev[0][sz] : 0.f;
        const float ry = up_z ? prev[1][sz] : 0.f;
        float dxz, dxy, dyx, dyz, dzy, dzx;
        partials(r, en[j][1] - e1p[j][1], en[j][2] - e1p[j][2], qx - e1p[j][0],
                 qz - e1p[j][2], rx - e1p[j][0], ry - e1p[j][1], psh[j], dxz, dxy, dyx, dyz,
                 dzy, dzx);
        float vx = hp0[j][0] - S * (dyz - dzy);
        float vy = hp0[j][1] - S * (dzx - dxz);
        float vz = hp0[j][2] - S * (dxy - dyx);
        if (src1 && xh == a.source_x) {
          vy -= (se_y[j] * we1y) * S;
          vz += (se_z[j] * we1z) * S;
        }
        h1n[j][0] = on ? vx : 0.f;
        h1n[j][1] = on ? vy : 0.f;
        h1n[j][2] = on ? vz : 0.f;
        if (hits1 && on && o[j] >= A && o[j] < B)
          stage_hit_cell(a.st, hits1, t % a.slots, xh, y[j], z[j], e1p[j][0], e1p[j][1],
                         e1p[j][2], vx, vy, vz, hp0[j][0], hp0[j][1], hp0[j][2]);
        lh1[0][s] = h1n[j][0];
        lh1[1][s] = h1n[j][1];
        lh1[2][s] = h1n[j][2];
      }
      __syncthreads();
      // (c) E2 of plane x - 1 over [A, B + nz)
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool on = do_c && o[j] >= A && o[j] < Bh;
        const int v = on && o[j] < B && xh < x1 ? (xh * P + o[j]) * 4 : kAway;  // stored
        const int s = threadIdx.x + T * j;
        const Rows r = rows_of(ce, cxh, ye[j]);
        // (slots below A read clamped neighbours; their results are zeroed)
        const int sy = max(s - nz, 0), sz = max(s - 1, 0);
        const float yx = y[j] > 0 ? lh1[0][sy] : 0.f;
        const float yz = y[j] > 0 ? lh1[2][sy] : 0.f;
        const float zx = z[j] > 0 ? lh1[0][sz] : 0.f;
        const float zy = z[j] > 0 ? lh1[1][sz] : 0.f;
        const float dx1 = h1n[j][1] - h1p[j][1], dx2 = h1n[j][2] - h1p[j][2];
        const float dy0 = h1n[j][0] - yx, dy2 = h1n[j][2] - yz;
        const float dz0 = h1n[j][0] - zx, dz1 = h1n[j][1] - zy;
        if (pa.hist && on && xh < x1 && o[j] < B)
          record_curl_at(pa, t + 1, xh, y[j], z[j], dx1, dx2, dy0, dy2, dz0, dz1);
        float dxz, dxy, dyx, dyz, dzy, dzx;
        partials(r, dx1, dx2, dy0, dy2, dz0, dz1, pse1[j], dxz, dxy, dyx, dyz, dzy, dzx);
        float vx = e1p[j][0] + (S * (dyz - dzy)) * ie1[j];
        float vy = e1p[j][1] + (S * (dzx - dxz)) * ie1[j];
        float vz = e1p[j][2] + (S * (dxy - dyx)) * ie1[j];
        if (src2 && xh == a.source_x) {
          vy += ((sh_y[j] * wh2y) * S) * ie1[j];
          vz -= ((sh_z[j] * wh2z) * S) * ie1[j];
        }
        e2n[j][0] = on ? vx : 0.f;
        e2n[j][1] = on ? vy : 0.f;
        e2n[j][2] = on ? vz : 0.f;
        f.e1.set(0, v, vx);
        f.e1.set(1, v, vy);
        f.e1.set(2, v, vz);
        store_psi6(f.pe1, v, r, pse1[j]);
        le2[xh & 1][0][s] = e2n[j][0];
        le2[xh & 1][1][s] = e2n[j][1];
        le2[xh & 1][2][s] = e2n[j][2];
      }
      // (d) H2 of plane x - 2 over the tile
      const unsigned hits2 =
          tile_boxes && do_d ? stage_hits(a.st, tile_boxes, xq, xq + 1, A, B, nz) : 0u;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool on = do_d && o[j] >= A && o[j] < B;
        const int v = on ? (xq * P + o[j]) * 4 : kAway;
        const int s = threadIdx.x + T * j;
        const Rows r = rows_of(ch, cxq, yh[j]);
        const int sy = min(s + nz, SLOTS - 1), sz = min(s + 1, SLOTS - 1);
        const bool up_y = y[j] + 1 < ny, up_z = z[j] + 1 < nz;
        const float(*prev)[SLOTS] = le2[xq & 1];
        const float qx = up_y ? prev[0][sy] : 0.f;
        const float qz = up_y ? prev[2][sy] : 0.f;
        const float rx = up_z ? prev[0][sz] : 0.f;
        const float ry = up_z ? prev[1][sz] : 0.f;
        float dxz, dxy, dyx, dyz, dzy, dzx;
        partials(r, e2n[j][1] - e2p[j][1], e2n[j][2] - e2p[j][2], qx - e2p[j][0],
                 qz - e2p[j][2], rx - e2p[j][0], ry - e2p[j][1], psh1[j], dxz, dxy, dyx, dyz,
                 dzy, dzx);
        float vx = h1p[j][0] - S * (dyz - dzy);
        float vy = h1p[j][1] - S * (dzx - dxz);
        float vz = h1p[j][2] - S * (dxy - dyx);
        if (src2 && xq == a.source_x) {
          vy -= (se_y[j] * we2y) * S;
          vz += (se_z[j] * we2z) * S;
        }
        if (on && save && xq >= a.hbox[0] && xq < a.hbox[1] && y[j] >= a.hbox[2] &&
            y[j] < a.hbox[3] && z[j] >= a.hbox[4] && z[j] < a.hbox[5]) {
          const int i = xq * P + o[j];
          a.hprev[i] = h1p[j][0];
          a.hprev[a.n + i] = h1p[j][1];
          a.hprev[2 * a.n + i] = h1p[j][2];
        }
        f.h1.set(0, v, vx);
        f.h1.set(1, v, vy);
        f.h1.set(2, v, vz);
        store_psi6(f.ph1, v, r, psh1[j]);
        if (hits2 && on)
          stage_hit_cell(a.st, hits2, (t + 1) % a.slots, xq, y[j], z[j], e2p[j][0], e2p[j][1],
                         e2p[j][2], vx, vy, vz, h1p[j][0], h1p[j][1], h1p[j][2]);
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        for (int c = 0; c < 3; ++c) {
          hp0[j][c] = hn[j][c];
          e1p[j][c] = en[j][c];
          h1p[j][c] = h1n[j][c];
          e2p[j][c] = e2n[j][c];
        }
        for (int c = 0; c < 6; ++c) {
          pse1[j][c] = pse[j][c];
          psh1[j][c] = psh[j][c];
        }
        ie1[j] = iex[j];
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    const int done = atomicAdd(pa.ticket, 1);
    if (done == nblocks - 1) {
      a.counters[0] = t + 2;
      a.counters[1] = t + 1;
      *pa.ticket = 0;
    }
  }
}

template <int T, int J>
__global__ __launch_bounds__(T) void k_yee_pass2(PassArgs pa) {
  pass2_body<T, J>(pa, blockIdx.x, gridDim.x);
}

// Batched two-step pass: the grid concatenates the (XCD-striped, padded to
// multiples of 8) block ranges of `count` passes, first[d] the first block of
// pass d; each design's arithmetic, ticket and counters are its own pass's,
// so its result is bitwise that of its own wfd_yee_pass.
template <int T, int J>
__global__ __launch_bounds__(T) void k_batch_pass2(const PassArgs* __restrict__ table,
                                                   const int* __restrict__ first, int count) {
  const int d = batch_member(first, count, blockIdx.x);
  pass2_body<T, J>(table[d], blockIdx.x - first[d], first[d + 1] - first[d]);
}

// Validate a plan and fill the fused step's arguments (shared by the single
// and the batched step).
int build_args(const wfd_yee_plan* p, YeeArgs& a, const char* who) {
  WFD_REQUIRE(p, "%s: null plan", who);
  WFD_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0, "%s: grid dimensions must be positive", who);
  WFD_REQUIRE(cells(p->nx, p->ny, p->nz) <= kMaxCells, "%s: grid exceeds 2^31 cells", who);
  WFD_REQUIRE((int64_t)p->ny * p->nz < (1 << 24), "%s: ny*nz must be < 2^24", who);
  WFD_REQUIRE(p->e && p->h && p->inv_eps && p->psi_e && p->psi_h && p->coeff_e &&
                  p->coeff_h && p->counters,
              "%s: null field, psi, coeff or counter pointer", who);
  WFD_REQUIRE(p->coeff_rows >= p->nx && p->coeff_rows >= p->ny && p->coeff_rows >= p->nz,
              "%s: coeff_rows < max(shape)", who);
  WFD_REQUIRE(p->source_x < p->nx, "%s: source plane outside the grid", who);
  WFD_REQUIRE(p->source_x < 0 ||
                  (p->src_e && p->src_h && p->wave_e && p->wave_h && p->steps > 0),
              "%s: source plane without profiles/waves", who);
  WFD_REQUIRE(p->save_h_from >= p->steps || p->hprev, "%s: hprev required", who);
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
  a.x_chunk = p->x_chunk > 0 ? p->x_chunk : 2;
  WFD_REQUIRE(a.x_chunk == 1 || a.x_chunk == 2 || a.x_chunk == 4 || a.x_chunk == 8,
              "%s: x_chunk must be 1, 2, 4 or 8", who);
  WFD_REQUIRE(p->nstage >= 0 && p->nstage <= WFD_MAX_STAGES, "%s: 0..%d stages", who,
              WFD_MAX_STAGES);
  WFD_REQUIRE(p->nstage == 0 || p->slots > 0, "%s: stage slots must be positive", who);
  a.nstage = p->nstage;
  a.slots = p->slots > 0 ? p->slots : 1;
  for (int b = 0; b < WFD_MAX_STAGES; ++b) a.st[b] = StageBox{};
  for (int b = 0; b < p->nstage; ++b)
    WFD_REQUIRE(to_stage_box(p->stages[b], p->nx, p->ny, p->nz, a.st[b]) == WFD_OK,
                "%s: stage %d outside the grid or without a buffer", who, b);
  a.segs = (a.P + kBlock - 1) / kBlock;
  const int chunks = (a.nx + a.x_chunk - 1) / a.x_chunk;
  a.total = a.segs * chunks;
  a.per = (a.total + 7) / 8;
  return WFD_OK;
}

constexpr int kClasses[4] = {1, 2, 4, 8};

// Batch table order: plans grouped by x-chunk class (1, 2, 4, 8), each group in
// the caller's order.  order[k] = plan index of table entry k.
void batch_order(const YeeArgs* args, int count, int* order, int* class_start) {
  int k = 0;
  for (int c = 0; c < 4; ++c) {
    class_start[c] = k;
    for (int d = 0; d < count; ++d)
      if (args[d].x_chunk == kClasses[c]) order[k++] = d;
  }
  class_start[4] = k;
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

// x planes per workgroup for a pass shape: a workgroup marches its chunk
// with a barrier per plane, so the pass takes about (rounds of co-resident
// workgroups per XCD) x (planes a march visits: the chunk plus one halo plane
// per step).  The count minimising that, from the kernel's occupancy, for
// `count` grids launched together (tiles[d] tiles per plane, nx[d] planes).
int pass_planes(const void* kernel, int threads, const int* tiles, const int* nx, int count,
                int steps) {
  int dev = 0, cus = 0, occ = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess)
    return 8;
  const int64_t per_xcd = (int64_t)max(cus, 8) / 8 * max(occ, 1);
  int longest = 1;
  for (int d = 0; d < count; ++d) longest = max(longest, nx[d]);
  int best = 1;
  int64_t best_cost = INT64_MAX;
  for (int xc = 1; xc <= min(64, longest); ++xc) {
    int64_t blocks = 0;  // per XCD
    for (int d = 0; d < count; ++d)
      blocks += ((int64_t)tiles[d] * ((nx[d] + xc - 1) / xc) + 7) / 8;
    const int64_t cost = (blocks + per_xcd - 1) / per_xcd * (xc + steps);
    if (cost < best_cost) {
      best_cost = cost;
      best = xc;
    }
  }
  return best;
}

// Validate a pass (plan = the step's start, out = its targets) and fill its
// arguments but for the x chunking; the shape it launches in steps,
// per_thread and threads.
int pass_setup(const wfd_yee_plan* p, const wfd_yee_targets* out, PassArgs& pa, int& steps,
               int& per_thread, int& threads) {
  if (int status = build_args(p, pa.a, "yee_pass")) return status;
  WFD_REQUIRE(out && out->e && out->h && out->psi_e && out->psi_h && out->ticket,
              "yee_pass: null target or ticket pointer");
  // psi (6 components) is addressed through 32-bit buffer offsets
  WFD_REQUIRE((int64_t)24 * pa.a.n <= kFieldBytes,
              "yee_pass: grids above %d cells step with wfd_yee_step", (int)(kFieldBytes / 24));
  WFD_REQUIRE(out->e != p->e && out->h != p->h && out->psi_e != p->psi_e &&
                  out->psi_h != p->psi_h,
              "yee_pass: targets must be buffers other than the plan's");
  steps = out->steps ? out->steps : 1;
  WFD_REQUIRE(steps == 1 || steps == 2, "yee_pass: steps must be 1 or 2");
  // one step: 512 or 1024 threads x 1 offset, 256 threads x 2 or 4; two
  // steps: 512 threads x 1 or 2, 1024 threads x 1
  per_thread = out->slots_per_thread ? out->slots_per_thread : (steps == 1 ? 4 : 2);
  WFD_REQUIRE(steps == 1 ? (per_thread == 1 || per_thread == 2 || per_thread == 4)
                         : (per_thread == 1 || per_thread == 2),
              "yee_pass: slots_per_thread must be 1, 2 or 4 (one step), 1 or 2 (two steps)");
  threads = out->threads ? out->threads : (steps == 2 || per_thread == 1) ? 2 * kBlock : kBlock;
  WFD_REQUIRE(threads == 4 * kBlock ? per_thread == 1
                                    : threads == ((steps == 2 || per_thread == 1) ? 2 * kBlock
                                                                                  : kBlock),
              "yee_pass: %d threads with %d slots per thread is not a pass shape", threads,
              per_thread);
  const int slots = per_thread * threads;
  const int halo = steps == 1 ? p->nz : 3 * p->nz;
  WFD_REQUIRE(halo < slots, "yee_pass: nz = %d leaves no tile in %d slots per workgroup", p->nz,
              slots);
  WFD_REQUIRE(out->planes >= 0 && out->planes <= 64, "yee_pass: planes must be 0..64");
  pa.e_out = out->e;
  pa.h_out = out->h;
  pa.pe_out = out->psi_e;
  pa.ph_out = out->psi_h;
  pa.ticket = out->ticket;
  pa.hist = out->hist;
  pa.hist_steps = out->hist_steps > 0 ? out->hist_steps : 1;
  for (int k = 0; k < 6; ++k) pa.hbox[k] = out->hist_box[k];
  if (out->hist) {
    WFD_REQUIRE(steps == 2, "yee_pass: the curl history is recorded by two-step passes only");
    WFD_REQUIRE(out->hist_steps > 0, "yee_pass: hist_steps must be positive");
    const int* b = out->hist_box;
    WFD_REQUIRE(b[0] >= 0 && b[1] >= 0 && b[2] >= 0 && b[3] >= 0 && b[4] >= 0 && b[5] >= 0 &&
                    b[0] + b[3] <= p->nx && b[1] + b[4] <= p->ny && b[2] + b[5] <= p->nz,
                "yee_pass: history box outside the grid");
  }
  pa.W = slots - halo;
  pa.tiles = (pa.a.P + pa.W - 1) / pa.W;
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
  YeeArgs a;
  if (int status = build_args(p, a, "yee_step")) return status;
  const dim3 grid(a.per * 8), block(kBlock);
  hipStream_t s = as_stream(stream);
  switch (a.x_chunk) {
    case 1:
      hipLaunchKernelGGL(k_yee_e<1>, grid, block, 0, s, a);
      hipLaunchKernelGGL(k_yee_h<1>, grid, block, 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL(k_yee_e<2>, grid, block, 0, s, a);
      hipLaunchKernelGGL(k_yee_h<2>, grid, block, 0, s, a);
      break;
    case 4:
      hipLaunchKernelGGL(k_yee_e<4>, grid, block, 0, s, a);
      hipLaunchKernelGGL(k_yee_h<4>, grid, block, 0, s, a);
      break;
    default:
      hipLaunchKernelGGL(k_yee_e<8>, grid, block, 0, s, a);
      hipLaunchKernelGGL(k_yee_h<8>, grid, block, 0, s, a);
      break;
  }
  WFD_LAUNCH_CHECK("yee_step");
  return WFD_OK;
}

int wfd_yee_pass(const wfd_yee_plan* p, const wfd_yee_targets* out, void* stream) {
  PassArgs pa;
  int steps, per_thread, threads;
  if (int status = pass_setup(p, out, pa, steps, per_thread, threads)) return status;
  int xc = out->planes;
  if (xc == 0) {
    const void* kernel =
        steps == 2 ? (threads == 4 * kBlock ? (const void*)&k_yee_pass2<4 * kBlock, 1>
                      : per_thread == 1     ? (const void*)&k_yee_pass2<2 * kBlock, 1>
                                            : (const void*)&k_yee_pass2<2 * kBlock, 2>)
        : threads == 4 * kBlock ? (const void*)&k_yee_pass<4 * kBlock, 1>
        : per_thread == 1       ? (const void*)&k_yee_pass<2 * kBlock, 1>
        : per_thread == 2       ? (const void*)&k_yee_pass<kBlock, 2>
                                : (const void*)&k_yee_pass<kBlock, 4>;
    xc = pass_planes(kernel, threads, &pa.tiles, &p->nx, 1, steps);
  }
  pa.xc = xc;
  const int chunks = (p->nx + xc - 1) / xc;
  pa.total = pa.tiles * chunks;
  pa.per = (pa.total + 7) / 8;
  hipStream_t s = as_stream(stream);
  const dim3 grid(pa.per * 8);
  if (steps == 2) {
    if (threads == 4 * kBlock)
      hipLaunchKernelGGL((k_yee_pass2<4 * kBlock, 1>), grid, dim3(threads), 0, s, pa);
    else if (per_thread == 1)
      hipLaunchKernelGGL((k_yee_pass2<2 * kBlock, 1>), grid, dim3(threads), 0, s, pa);
    else
      hipLaunchKernelGGL((k_yee_pass2<2 * kBlock, 2>), grid, dim3(threads), 0, s, pa);
  } else if (threads == 4 * kBlock) {
    hipLaunchKernelGGL((k_yee_pass<4 * kBlock, 1>), grid, dim3(threads), 0, s, pa);
  } else if (per_thread == 1) {
    hipLaunchKernelGGL((k_yee_pass<2 * kBlock, 1>), grid, dim3(threads), 0, s, pa);
  } else if (per_thread == 2) {
    hipLaunchKernelGGL((k_yee_pass<kBlock, 2>), grid, dim3(threads), 0, s, pa);
  } else {
    hipLaunchKernelGGL((k_yee_pass<kBlock, 4>), grid, dim3(threads), 0, s, pa);
  }
  WFD_LAUNCH_CHECK("yee_pass");
  return WFD_OK;
}

int64_t wfd_batch_workspace_bytes(int count) {
  if (count < 1 || count > WFD_MAX_BATCH) return -1;
  return (int64_t)count * sizeof(YeeArgs) + (int64_t)(count + 4) * sizeof(int);
}

int wfd_batch_prepare(const wfd_yee_plan* plans, int count, void* workspace,
                      int64_t workspace_bytes, void* stream) {
  WFD_REQUIRE(plans && workspace, "batch_prepare: null argument");
  WFD_REQUIRE(count >= 1 && count <= WFD_MAX_BATCH, "batch_prepare: 1..%d plans",
              WFD_MAX_BATCH);
  WFD_REQUIRE(workspace_bytes >= wfd_batch_workspace_bytes(count),
              "batch_prepare: workspace smaller than wfd_batch_workspace_bytes");
  YeeArgs args[WFD_MAX_BATCH];
  for (int d = 0; d < count; ++d)
    if (int status = build_args(plans + d, args[d], "batch_prepare")) return status;
  int order[WFD_MAX_BATCH], start[5];
  batch_order(args, count, order, start);
  YeeArgs table[WFD_MAX_BATCH];
  int first[WFD_MAX_BATCH + 4];
  int f = 0;
  for (int c = 0; c < 4; ++c) {
    int at = 0;
    for (int k = start[c]; k < start[c + 1]; ++k) {
      table[k] = args[order[k]];
      first[f++] = at;
      at += table[k].per * 8;
    }
    first[f++] = at;  // class end
  }
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  WFD_HIP(hipMemcpyAsync(ws, table, sizeof(YeeArgs) * count, hipMemcpyHostToDevice, s));
  WFD_HIP(hipMemcpyAsync(ws + sizeof(YeeArgs) * count, first, sizeof(int) * (count + 4),
                         hipMemcpyHostToDevice, s));
  WFD_HIP(hipStreamSynchronize(s));
  return WFD_OK;
}

int wfd_batch_step(const wfd_yee_plan* plans, int count, const void* workspace, void* stream) {
  WFD_REQUIRE(plans && workspace, "batch_step: null argument");
  WFD_REQUIRE(count >= 1 && count <= WFD_MAX_BATCH, "batch_step: 1..%d plans", WFD_MAX_BATCH);
  YeeArgs args[WFD_MAX_BATCH];
  for (int d = 0; d < count; ++d)
    if (int status = build_args(plans + d, args[d], "batch_step")) return status;
  int order[WFD_MAX_BATCH], start[5];
  batch_order(args, count, order, start);
  const YeeArgs* table = static_cast<const YeeArgs*>(workspace);
  const int* first = reinterpret_cast<const int*>(table + count);
  hipStream_t s = as_stream(stream);
  const dim3 block(kBlock);
  // E half-steps of every class, then H half-steps (H of a plan needs only
  // that plan's E, so the class order is free)
  for (int half = 0; half < 2; ++half) {
    for (int c = 0; c < 4; ++c) {
      const int members = start[c + 1] - start[c];
      if (!members) continue;
      int blocks = 0;
      for (int k = start[c]; k < start[c + 1]; ++k) blocks += args[order[k]].per * 8;
      const YeeArgs* t = table + start[c];
      const int* fc = first + start[c] + c;
      const dim3 grid(blocks);
#define WFD_BATCH_LAUNCH(XC)                                                          \
  if (half == 0)                                                                      \
    hipLaunchKernelGGL(k_batch_e<XC>, grid, block, 0, s, t, fc, members);             \
  else                                                                                \
    hipLaunchKernelGGL(k_batch_h<XC>, grid, block, 0, s, t, fc, members);
      switch (kClasses[c]) {
        case 1: WFD_BATCH_LAUNCH(1) break;
        case 2: WFD_BATCH_LAUNCH(2) break;
        case 4: WFD_BATCH_LAUNCH(4) break;
        default: WFD_BATCH_LAUNCH(8) break;
      }
#undef WFD_BATCH_LAUNCH
    }
  }
  WFD_LAUNCH_CHECK("batch_step");
  return WFD_OK;
}


// ------------------------------------------------------------ batched pass
// Two-step passes of many grids in one launch (k_batch_pass2<1024, 1>).
namespace {
int batch_pass_args(const wfd_yee_plan* plans, const wfd_yee_targets* targets, int count,
                    PassArgs* table, int* first) {
  WFD_REQUIRE(plans && targets, "batch_pass: null argument");
  WFD_REQUIRE(count >= 1 && count <= WFD_MAX_BATCH, "batch_pass: 1..%d passes", WFD_MAX_BATCH);
  int tiles[WFD_MAX_BATCH], nx[WFD_MAX_BATCH];
  for (int d = 0; d < count; ++d) {
    wfd_yee_targets out = targets[d];
    WFD_REQUIRE(out.steps == 2 && (out.slots_per_thread == 0 || out.slots_per_thread == 1) &&
                    (out.threads == 0 || out.threads == 4 * kBlock),
                "batch_pass: pass %d is not a two-step pass of 1024 x 1 slots", d);
    WFD_REQUIRE(out.planes == targets[0].planes, "batch_pass: passes differ in planes");
    out.slots_per_thread = 1;
    out.threads = 4 * kBlock;
    int steps, per_thread, threads;
    if (int status = pass_setup(plans + d, &out, table[d], steps, per_thread, threads))
      return status;
    for (int e = 0; e < d; ++e)
      WFD_REQUIRE(targets[e].ticket != out.ticket && plans[e].counters != plans[d].counters,
                  "batch_pass: passes %d and %d share a ticket or counters", e, d);
    tiles[d] = table[d].tiles;
    nx[d] = plans[d].nx;
  }
  int xc = targets[0].planes;
  if (xc == 0)
    xc = pass_planes((const void*)&k_batch_pass2<4 * kBlock, 1>, 4 * kBlock, tiles, nx, count,
                     2);
  WFD_REQUIRE(xc >= 1 && xc <= 64, "batch_pass: planes must be 0..64");
  int at = 0;
  for (int d = 0; d < count; ++d) {
    PassArgs& pa = table[d];
    pa.xc = xc;
    pa.total = pa.tiles * ((nx[d] + xc - 1) / xc);
    pa.per = (pa.total + 7) / 8;
    first[d] = at;
    at += pa.per * 8;
  }
  first[count] = at;
  return WFD_OK;
}
}  // namespace

int64_t wfd_batch_pass_workspace_bytes(int count) {
  if (count < 1 || count > WFD_MAX_BATCH) return -1;
  return (int64_t)count * sizeof(PassArgs) + (int64_t)(count + 1) * sizeof(int);
}

int wfd_batch_pass_prepare(const wfd_yee_plan* plans, const wfd_yee_targets* targets, int count,
                           void* workspace, int64_t workspace_bytes, void* stream) {
  WFD_REQUIRE(workspace, "batch_pass_prepare: null workspace");
  WFD_REQUIRE(workspace_bytes >= wfd_batch_pass_workspace_bytes(count),
              "batch_pass_prepare: workspace smaller than wfd_batch_pass_workspace_bytes");
  PassArgs table[WFD_MAX_BATCH];
  int first[WFD_MAX_BATCH + 1];
  if (int status = batch_pass_args(plans, targets, count, table, first)) return status;
  hipStream_t s = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  WFD_HIP(hipMemcpyAsync(ws, table, sizeof(PassArgs) * count, hipMemcpyHostToDevice, s));
  WFD_HIP(hipMemcpyAsync(ws + sizeof(PassArgs) * count, first, sizeof(int) * (count + 1),
                         hipMemcpyHostToDevice, s));
  WFD_HIP(hipStreamSynchronize(s));
  return WFD_OK;
}

int wfd_batch_pass(const wfd_yee_plan* plans, const wfd_yee_targets* targets, int count,
                   const void* workspace, void* stream) {
  WFD_REQUIRE(workspace, "batch_pass: null workspace");
  PassArgs table[WFD_MAX_BATCH];
  int first[WFD_MAX_BATCH + 1];
  if (int status = batch_pass_args(plans, targets, count, table, first)) return status;
  const PassArgs* t = static_cast<const PassArgs*>(workspace);
  const int* f = reinterpret_cast<const int*>(t + count);
  hipLaunchKernelGGL((k_batch_pass2<4 * kBlock, 1>), dim3(first[count]), dim3(4 * kBlock), 0,
                     as_stream(stream), t, f, count);
  WFD_LAUNCH_CHECK("batch_pass");
  return WFD_OK;
}

}  // extern "C"
