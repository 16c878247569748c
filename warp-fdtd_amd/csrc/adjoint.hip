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

// The fused reverse half-step again, with a thread marching XC x planes of
// one (y, z) offset (the block: 256 consecutive offsets, blocks striped over
// the XCDs as in the forward step).  A cell's x-axis pair of prepare slots is
// the neighbouring plane's own pair, so the march computes it once per plane
// (plus one halo plane) instead of twice, and reads q, eps and psi of the x
// neighbour from registers; every access is a buffer access, masked by an
// out-of-range offset rather than a branch, and the CPML terms are selected.
// The arithmetic of each cell is k_reverse_half's, so results are the same
// bits.
struct RevArgs {
  float* q_dst;
  const float* q_src;
  const float* eps;
  const float* psi_in;
  float* psi_out;
  const float* coeff;
  void* grad;
  const float* hist;
  const int* step;
  int rows, hist_steps;
  int nx, ny, nz, n, P;
  float inv_nz;
  int bx0, by0, bz0, bsx, bsy, bsz;  // design box (history and gradient)
  float courant;
  int segs, total, per;  // (y, z) segments per chunk, logical blocks, per XCD
};

// One axis pair of prepare slots (prepare_pair's operations) from the loaded
// values: qa, qb the pair's q components, e the cell's eps (electric), sa, sb
// its psi slots, c0..c2 the axis row; p1, p2 the updated psi.
template <bool ELECTRIC>
__device__ __forceinline__ void rev_pair(float qa, float qb, float e, float sa, float sb,
                                         float c0, float c1, float c2, float courant,
                                         float& d1, float& d2, float& p1, float& p2) {
  const float factor = ELECTRIC ? courant * e : -courant;
  const float q1 = -(factor * qa);
  const float q2 = factor * qb;
  const float s1 = sa + q1, s2 = sb + q2;
  d1 = c0 != 0.f ? c2 * q1 + c0 * s1 : c2 * q1;
  d2 = c0 != 0.f ? c2 * q2 + c0 * s2 : c2 * q2;
  p1 = c1 * s1;
  p2 = c1 * s2;
}

template <typename G, int XC, bool FROM_H>
__global__ __launch_bounds__(kBlock) void k_reverse_march(RevArgs r) {
  constexpr bool E = !FROM_H;           // prepare is electric when q_src is qe
  constexpr int D = FROM_H ? -1 : 1;    // the gather's neighbour along each axis
  const int logical = (blockIdx.x & 7) * r.per + (blockIdx.x >> 3);
  if (logical >= r.total) return;
  const int chunk = logical / r.segs;
  const int o = (logical - chunk * r.segs) * kBlock + threadIdx.x;
  if (o >= r.P) return;
  const int nx = r.nx, ny = r.ny, nz = r.nz, n = r.n, P = r.P, rows = r.rows;
  const int x0 = chunk * XC;
  const int y = div_small(o, nz, r.inv_nz), z = o - y * nz;
  const Field qs(r.q_src, 3, n), qd(r.q_dst, 3, n), pin(r.psi_in, 6, n), pout(r.psi_out, 6, n);
  const Field ep(E ? r.eps : r.q_src, 1, n);  // (unread when FROM_H)
  const float* __restrict__ co = r.coeff;
  const float S = r.courant;
  // y and z rows of the cell and of its neighbours
  const int yn = y + D, zn = z + D;
  const bool yok = yn >= 0 && yn < ny, zok = zn >= 0 && zn < nz;
  const int ync = yok ? yn : y, znc = zok ? zn : z;
  const float cy0 = co[(rows + y) * 3], cy1 = co[(rows + y) * 3 + 1],
              cy2 = co[(rows + y) * 3 + 2];
  const float cz0 = co[(2 * rows + z) * 3], cz1 = co[(2 * rows + z) * 3 + 1],
              cz2 = co[(2 * rows + z) * 3 + 2];
  const float ny0 = co[(rows + ync) * 3], ny1 = co[(rows + ync) * 3 + 1],
              ny2 = co[(rows + ync) * 3 + 2];
  const float nz0 = co[(2 * rows + znc) * 3], nz1 = co[(2 * rows + znc) * 3 + 1],
              nz2 = co[(2 * rows + znc) * 3 + 2];
  // the x-axis pair of planes x0 - 1 + k (D < 0) or x0 + k (D > 0), k = 0..XC
  float xd1[XC + 1], xd2[XC + 1];
  constexpr int kbase = D < 0 ? -1 : 0;
#pragma unroll
  for (int k = 0; k <= XC; ++k) {
    const int x = x0 + kbase + k;
    const bool on = x >= 0 && x < nx;  // uniform
    const int xc = on ? x : 0;
    const int v = on ? (xc * P + o) * 4 : kAway;
    const float c0 = co[xc * 3], c1 = co[xc * 3 + 1], c2 = co[xc * 3 + 2];
    const int vp = c0 != 0.f ? v : kAway;
    float p1, p2;
    rev_pair<E>(qs.at(1, v), qs.at(2, v), E ? ep.at(0, v) : 0.f, pin.at(0, vp), pin.at(1, vp),
                c0, c1, c2, S, xd1[k], xd2[k], p1, p2);
    // planes of the chunk write their updated x psi; planes past the grid
    // contribute zero
    if (!on) xd1[k] = xd2[k] = 0.f;
    const bool mine = on && x >= x0 && x < x0 + XC;
    pout.set(0, mine ? vp : kAway, p1);
    pout.set(1, mine ? vp : kAway, p2);
  }
  const int t = r.step[0];
#pragma unroll
  for (int p = 0; p < XC; ++p) {
    const int x = x0 + p;
    if (x >= nx) break;  // uniform
    const int i = x * P + o;
    const int v = i * 4;
    const int ko = p - kbase, kn = ko + D;  // own and neighbour x pairs
    const float own0 = xd1[ko], own1 = xd2[ko], nb0 = xd1[kn], nb1 = xd2[kn];
    // y axis: components 2, 0; slots 2, 3
    float own2, own3, nb2 = 0.f, nb3 = 0.f, p1, p2, u1, u2;
    const float ev = E ? ep.at(0, v) : 0.f;
    const float q0 = qs.at(0, v), q1 = qs.at(1, v), q2 = qs.at(2, v);
    const int vy = cy0 != 0.f ? v : kAway;
    rev_pair<E>(q2, q0, ev, pin.at(2, vy), pin.at(3, vy), cy0, cy1, cy2, S, own2, own3, p1, p2);
    pout.set(2, vy, p1);
    pout.set(3, vy, p2);
    {
      const int w = yok ? v + D * nz * 4 : kAway;
      const int wp = ny0 != 0.f ? w : kAway;
      rev_pair<E>(qs.at(2, w), qs.at(0, w), E ? ep.at(0, w) : 0.f, pin.at(2, wp), pin.at(3, wp),
                  ny0, ny1, ny2, S, nb2, nb3, u1, u2);
      if (!yok) nb2 = nb3 = 0.f;
    }
    // z axis: components 0, 1; slots 4, 5
    float own4, own5, nb4 = 0.f, nb5 = 0.f;
    const int vz = cz0 != 0.f ? v : kAway;
    rev_pair<E>(q0, q1, ev, pin.at(4, vz), pin.at(5, vz), cz0, cz1, cz2, S, own4, own5, p1, p2);
    pout.set(4, vz, p1);
    pout.set(5, vz, p2);
    {
      const int w = zok ? v + D * 4 : kAway;
      const int wp = nz0 != 0.f ? w : kAway;
      rev_pair<E>(qs.at(0, w), qs.at(1, w), E ? ep.at(0, w) : 0.f, pin.at(4, wp), pin.at(5, wp),
                  nz0, nz1, nz2, S, nb4, nb5, u1, u2);
      if (!zok) nb4 = nb5 = 0.f;
    }
    // the gather (k_reverse_half's sums, in its order)
    float v0 = (nb2 - own2) + (nb5 - own5);
    float v1 = (nb1 - own1) + (nb4 - own4);
    float v2 = (nb0 - own0) + (nb3 - own3);
    if (!FROM_H) {
      v0 = -v0;
      v1 = -v1;
      v2 = -v2;
    }
    const float a0 = qd.at(0, v) + v0, a1 = qd.at(1, v) + v1, a2 = qd.at(2, v) + v2;
    qd.set(0, v, a0);
    qd.set(1, v, a1);
    qd.set(2, v, a2);
    if (FROM_H && t >= 0 && x >= r.bx0 && x < r.bx0 + r.bsx && y >= r.by0 &&
        y < r.by0 + r.bsy && z >= r.bz0 && z < r.bz0 + r.bsz) {
      const int dn = r.bsx * r.bsy * r.bsz;
      const int j = ((x - r.bx0) * r.bsy + (y - r.by0)) * r.bsz + (z - r.bz0);
      const int64_t base = (int64_t)(t % r.hist_steps) * 3 * dn + j;
      const float contrib =
          r.hist[base] * a0 + r.hist[base + dn] * a1 + r.hist[base + 2 * dn] * a2;
      static_cast<G*>(r.grad)[j] += (G)contrib;
    }
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

// value of component c at plane cell (x, y, z); 0 outside the plane.  The
// load is unconditional (a cell outside reads the plane's first cell and the
// value is dropped), so the stencil's loads issue together.
__device__ __forceinline__ float plane_value(const InjectPlane& p, int c, int x, int y, int z) {
  const bool in = x >= p.x0 && y >= p.y0 && z >= p.z0 && x < p.x0 + p.sx &&
                  y < p.y0 + p.sy && z < p.z0 + p.sz;
  const int i = in ? ((x - p.x0) * p.sy + (y - p.y0)) * p.sz + (z - p.z0) : 0;
  const float2 s = p.source[c * p.count + i];
  const float v = s.x * p.w0 - s.y * p.w1;
  return in ? v : 0.f;
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
  hipStream_t s = as_stream(stream);
  if ((int64_t)24 * n > kFieldBytes) {  // past 32-bit buffer offsets: one thread per cell
    hipLaunchKernelGGL((k_reverse_half<G>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       q_dst, q_src, eps, psi_in, psi_out, coeff, coeff_rows, hist,
                       hist_steps > 0 ? hist_steps : 1, grad, step, nx, ny, nz, x0, y0, z0, sx,
                       sy, sz, courant, from_h);
    WFD_LAUNCH_CHECK("reverse_half");
    return WFD_OK;
  }
  constexpr int XC = 4;
  RevArgs r;
  r.q_dst = q_dst;
  r.q_src = q_src;
  r.eps = eps;
  r.psi_in = psi_in;
  r.psi_out = psi_out;
  r.coeff = coeff;
  r.grad = grad;
  r.hist = hist;
  r.step = step;
  r.rows = coeff_rows;
  r.hist_steps = hist_steps > 0 ? hist_steps : 1;
  r.nx = nx;
  r.ny = ny;
  r.nz = nz;
  r.n = n;
  r.P = ny * nz;
  r.inv_nz = 1.0f / (float)nz;
  r.bx0 = x0;
  r.by0 = y0;
  r.bz0 = z0;
  r.bsx = sx;
  r.bsy = sy;
  r.bsz = sz;
  r.courant = courant;
  r.segs = (r.P + kBlock - 1) / kBlock;
  r.total = r.segs * ((nx + XC - 1) / XC);
  r.per = (r.total + 7) / 8;
  const dim3 grid(r.per * 8);
  if (from_h)
    hipLaunchKernelGGL((k_reverse_march<G, XC, true>), grid, dim3(kBlock), 0, s, r);
  else
    hipLaunchKernelGGL((k_reverse_march<G, XC, false>), grid, dim3(kBlock), 0, s, r);
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
