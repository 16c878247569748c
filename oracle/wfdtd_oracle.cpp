// CPU oracle for the uniform-grid FDTD path — TEST INFRASTRUCTURE ONLY.
//
// A literal, scalar restatement of the reference's Warp kernels, in the same
// operation order, instantiated for float and double.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library, and only as the checker (or as the timed CPU baseline); the product
// path never calls it.  Built with g++ -ffp-contract=off so that the float
// instantiation rounds every operation separately.
//
// Each function cites the reference file:line it restates.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

#include <omp.h>

namespace {

template <typename T>
struct V3 {
  T v[3];
  T& operator[](int c) { return v[c]; }
  T operator[](int c) const { return v[c]; }
};

template <typename T>
V3<T> make3(T a, T b, T c) {
  V3<T> r;
  r.v[0] = a;
  r.v[1] = b;
  r.v[2] = c;
  return r;
}

// warp_fdtd/kernels.py:8-26 — zero-padded (or periodic) fetch of a 3-vector
// from SoA (3, n) storage at flat index (x*ny + y)*nz + z.
template <typename T>
V3<T> sample(const T* f, int64_t n, int x, int y, int z, int nx, int ny, int nz,
             int periodic) {
  if (periodic != 0) {
    x = (x + nx) % nx;
    y = (y + ny) % ny;
    z = (z + nz) % nz;
  }
  if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz)
    return make3<T>(0, 0, 0);
  int64_t i = ((int64_t)x * ny + y) * nz + z;
  return make3<T>(f[i], f[n + i], f[2 * n + i]);
}

// warp_fdtd/kernels.py:29-47 (full psi) and 50-74 (packed psi).
// coeff rows are (a, b, 1/kappa[, plane rank]) with `width` floats per row.
template <typename T>
T cpml(T d, T* psi, int64_t psi_stride, const T* coeff, int width, int maxs,
       int axis, int position, int slot, int64_t i, int packed, int nx, int ny,
       int nz) {
  const T* c = coeff + ((int64_t)axis * maxs + position) * width;
  if (c[0] != T(0)) {
    int64_t index = i;
    if (packed) {
      int64_t plane = (int64_t)c[3];
      int64_t plane_yz = (int64_t)ny * nz;
      index = plane * plane_yz + i % plane_yz;
      if (axis == 1)
        index = plane * nx * nz + (i / plane_yz) * nz + i % nz;
      else if (axis == 2)
        index = plane * nx * ny + i / nz;
    }
    T p = c[1] * psi[slot * psi_stride + index] + c[0] * d;
    psi[slot * psi_stride + index] = p;
    return c[2] * d + p;
  }
  return c[2] * d;
}

// warp_fdtd/kernels.py:77-142 — one Yee half-step over all n cells.
template <typename T>
void update(int electric, int packed, T* f, const T* g, const T* inv_eps, T* psi,
            int64_t psi_stride, const T* coeff, int maxs, const T* source,
            int64_t source_stride, const T* waves, int64_t waves_stride,
            const int* step, T* hprev, int save_h_from, int nx, int ny, int nz,
            int source_x, T courant, int periodic) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int width = packed ? 4 : 3;
  const int t = step[0];
#pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; ++x) {
    for (int y = 0; y < ny; ++y) {
      for (int z = 0; z < nz; ++z) {
        const int64_t i = ((int64_t)x * ny + y) * nz + z;
        V3<T> v = sample(g, n, x, y, z, nx, ny, nz, periodic);
        V3<T> dx, dy, dz;
        if (electric) {
          V3<T> a = sample(g, n, x - 1, y, z, nx, ny, nz, periodic);
          V3<T> b = sample(g, n, x, y - 1, z, nx, ny, nz, periodic);
          V3<T> c = sample(g, n, x, y, z - 1, nx, ny, nz, periodic);
          for (int k = 0; k < 3; ++k) {
            dx[k] = v[k] - a[k];
            dy[k] = v[k] - b[k];
            dz[k] = v[k] - c[k];
          }
        } else {
          V3<T> a = sample(g, n, x + 1, y, z, nx, ny, nz, periodic);
          V3<T> b = sample(g, n, x, y + 1, z, nx, ny, nz, periodic);
          V3<T> c = sample(g, n, x, y, z + 1, nx, ny, nz, periodic);
          for (int k = 0; k < 3; ++k) {
            dx[k] = a[k] - v[k];
            dy[k] = b[k] - v[k];
            dz[k] = c[k] - v[k];
          }
        }
        // kernels.py:113-118 slot order.
        T dxz = cpml(dx[2], psi, psi_stride, coeff, width, maxs, 0, x, 0, i, packed, nx, ny, nz);
        T dxy = cpml(dx[1], psi, psi_stride, coeff, width, maxs, 0, x, 1, i, packed, nx, ny, nz);
        T dyx = cpml(dy[0], psi, psi_stride, coeff, width, maxs, 1, y, 2, i, packed, nx, ny, nz);
        T dyz = cpml(dy[2], psi, psi_stride, coeff, width, maxs, 1, y, 3, i, packed, nx, ny, nz);
        T dzy = cpml(dz[1], psi, psi_stride, coeff, width, maxs, 2, z, 4, i, packed, nx, ny, nz);
        T dzx = cpml(dz[0], psi, psi_stride, coeff, width, maxs, 2, z, 5, i, packed, nx, ny, nz);
        V3<T> curl = make3<T>(dyz - dzy, dzx - dxz, dxy - dyx);
        V3<T> old = make3<T>(f[i], f[n + i], f[2 * n + i]);
        V3<T> value;
        if (electric) {
          // kernels.py:123: old + courant*curl*inv_eps (left to right).
          for (int k = 0; k < 3; ++k) value[k] = old[k] + (courant * curl[k]) * inv_eps[i];
          if (x == source_x) {
            const int64_t s = (int64_t)y * nz + z;
            value[1] += ((source[2 * source_stride + s] * waves[2 * waves_stride + t]) * courant) * inv_eps[i];
            value[2] -= ((source[1 * source_stride + s] * waves[1 * waves_stride + t]) * courant) * inv_eps[i];
          }
        } else {
          for (int k = 0; k < 3; ++k) value[k] = old[k] - courant * curl[k];
          if (t >= save_h_from) {
            hprev[i] = old[0];
            hprev[n + i] = old[1];
            hprev[2 * n + i] = old[2];
          }
          if (x == source_x) {
            const int64_t s = (int64_t)y * nz + z;
            value[1] -= (source[2 * source_stride + s] * waves[2 * waves_stride + t]) * courant;
            value[2] += (source[1 * source_stride + s] * waves[1 * waves_stride + t]) * courant;
          }
        }
        f[i] = value[0];
        f[n + i] = value[1];
        f[2 * n + i] = value[2];
      }
    }
  }
}

// warp_fdtd/detectors.py:10-23
template <typename T>
V3<T> ht(const T* h, const T* prev, int64_t n, int x, int y, int z, int nx, int ny,
         int nz) {
  V3<T> a = sample(h, n, x, y, z, nx, ny, nz, 0);
  V3<T> b = sample(prev, n, x, y, z, nx, ny, nz, 0);
  return make3<T>(T(0.5) * (a[0] + b[0]), T(0.5) * (a[1] + b[1]), T(0.5) * (a[2] + b[2]));
}

// warp_fdtd/detectors.py:26-62 — exact space/time colocation at the E_z point.
template <typename T>
void colocate(const T* e, const T* h, const T* prev, int64_t n, int x, int y, int z,
              int nx, int ny, int nz, V3<T>& ec, V3<T>& hc) {
  V3<T> e0 = sample(e, n, x, y, z, nx, ny, nz, 0);
  V3<T> ex = sample(e, n, x - 1, y, z, nx, ny, nz, 0);
  V3<T> ey = sample(e, n, x, y - 1, z, nx, ny, nz, 0);
  V3<T> ez = sample(e, n, x, y, z + 1, nx, ny, nz, 0);
  V3<T> exz = sample(e, n, x - 1, y, z + 1, nx, ny, nz, 0);
  V3<T> eyz = sample(e, n, x, y - 1, z + 1, nx, ny, nz, 0);
  ec = make3<T>(T(0.25) * (((e0[0] + ex[0]) + ez[0]) + exz[0]),
                T(0.25) * (((e0[1] + ey[1]) + ez[1]) + eyz[1]), e0[2]);
  V3<T> h0 = ht(h, prev, n, x, y, z, nx, ny, nz);
  V3<T> hx = ht(h, prev, n, x - 1, y, z, nx, ny, nz);
  V3<T> hy = ht(h, prev, n, x, y - 1, z, nx, ny, nz);
  V3<T> hxy = ht(h, prev, n, x - 1, y - 1, z, nx, ny, nz);
  V3<T> hz = ht(h, prev, n, x, y, z + 1, nx, ny, nz);
  V3<T> hxz = ht(h, prev, n, x - 1, y, z + 1, nx, ny, nz);
  V3<T> hyz = ht(h, prev, n, x, y - 1, z + 1, nx, ny, nz);
  V3<T> hxyz = ht(h, prev, n, x - 1, y - 1, z + 1, nx, ny, nz);
  hc = make3<T>(T(0.5) * (h0[0] + hy[0]), T(0.5) * (h0[1] + hx[1]),
                T(0.125) * (((((((h0[2] + hx[2]) + hy[2]) + hxy[2]) + hz[2]) + hxz[2]) + hyz[2]) + hxyz[2]));
}

template <typename T>
V3<T> cross(const V3<T>& a, const V3<T>& b) {
  return make3<T>(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                  a[0] * b[1] - a[1] * b[0]);
}

// warp_fdtd/detectors.py:65-106 — one monitor's per-step record.  Phasors are
// interleaved (re, im) pairs of shape (6, count, 2); flux accumulates
// sequentially in plane order (the reference uses atomics, order unspecified).
template <typename T, typename A>
void record(const T* e, const T* h, const T* prev, const int* step, const T* weights,
            const int* indices, A* phasor, A* flux, const T* area, int nx, int ny,
            int nz, int x0, int y0, int z0, int sx, int sy, int sz, int axis, int exact,
            int is_phasor) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int t = step[0];
  const int index = indices[t];
  if (index < 0) return;
  const int64_t count = (int64_t)sx * sy * sz;
  for (int64_t i = 0; i < count; ++i) {
    int x = x0 + (int)(i / ((int64_t)sy * sz));
    int y = y0 + (int)((i / sz) % sy);
    int z = z0 + (int)(i % sz);
    V3<T> ee = sample(e, n, x, y, z, nx, ny, nz, 0);
    V3<T> hh = sample(h, n, x, y, z, nx, ny, nz, 0);
    if (exact) colocate(e, h, prev, n, x, y, z, nx, ny, nz, ee, hh);
    if (is_phasor) {
      const T wr = weights[2 * t], wi = weights[2 * t + 1];
      for (int c = 0; c < 3; ++c) {
        A* p = phasor + 2 * ((int64_t)c * count + i);
        p[0] += A(wr * ee[c]);
        p[1] += A(wi * ee[c]);
        A* q = phasor + 2 * ((int64_t)(c + 3) * count + i);
        q[0] += A(wr * hh[c]);
        q[1] += A(wi * hh[c]);
      }
    } else {
      T power = cross(ee, hh)[axis] * area[i];
      flux[index] += A(power);
    }
  }
}

// Multi-frequency running DFT of the colocated fields on a plane: for each of
// nf frequencies, phasor[f, c, i] += w[t, f] * field_c with FP64 accumulation.
// Same colocation and window rule as detectors.py:65-106, applied per frequency
// (the reference's single-frequency phasor detector is the nf = 1 case).
template <typename T>
void record_dft(const T* e, const T* h, const T* prev, int t, const double* weights,
                int nf, double* phasor, int nx, int ny, int nz, int x0, int y0, int z0,
                int sx, int sy, int sz, int exact) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t count = (int64_t)sx * sy * sz;
  // every (frequency, component, cell) accumulator belongs to one iteration:
  // the parallel loop adds the same terms in the same order as a serial one
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) {
    int x = x0 + (int)(i / ((int64_t)sy * sz));
    int y = y0 + (int)((i / sz) % sy);
    int z = z0 + (int)(i % sz);
    V3<T> ee = sample(e, n, x, y, z, nx, ny, nz, 0);
    V3<T> hh = sample(h, n, x, y, z, nx, ny, nz, 0);
    if (exact) colocate(e, h, prev, n, x, y, z, nx, ny, nz, ee, hh);
    for (int f = 0; f < nf; ++f) {
      const double wr = weights[2 * ((int64_t)t * nf + f)];
      const double wi = weights[2 * ((int64_t)t * nf + f) + 1];
      for (int c = 0; c < 6; ++c) {
        const double v = (double)(c < 3 ? ee[c] : hh[c - 3]);
        double* p = phasor + 2 * (((int64_t)f * 6 + c) * count + i);
        p[0] += wr * v;
        p[1] += wi * v;
      }
    }
  }
}

// warp_fdtd/adjoint.py:17-43
template <typename T>
void record_curl(const T* h, T* hist, const int* step, int nx, int ny, int nz, int x0,
                 int y0, int z0, int sx, int sy, int sz, T courant) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t dn = (int64_t)sx * sy * sz;
  const int t = step[0];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < dn; ++i) {
    int x = x0 + (int)(i / ((int64_t)sy * sz));
    int y = y0 + (int)((i / sz) % sy);
    int z = z0 + (int)(i % sz);
    V3<T> v = sample(h, n, x, y, z, nx, ny, nz, 0);
    V3<T> a = sample(h, n, x - 1, y, z, nx, ny, nz, 0);
    V3<T> b = sample(h, n, x, y - 1, z, nx, ny, nz, 0);
    V3<T> c = sample(h, n, x, y, z - 1, nx, ny, nz, 0);
    V3<T> dx, dy, dz;
    for (int k = 0; k < 3; ++k) {
      dx[k] = v[k] - a[k];
      dy[k] = v[k] - b[k];
      dz[k] = v[k] - c[k];
    }
    T value[3] = {courant * (dy[2] - dz[1]), courant * (dz[0] - dx[2]),
                  courant * (dx[1] - dy[0])};
    for (int k = 0; k < 3; ++k) hist[((int64_t)t * 3 + k) * dn + i] = value[k];
  }
}

// warp_fdtd/adjoint.py:46-88 — transposed CPML for one half-step.
template <typename T>
void prepare(const T* q, const T* eps, T* psi, const T* coeff, int maxs, T* d, int nx,
             int ny, int nz, T courant, int electric) {
  const int64_t n = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int x = (int)(i / ((int64_t)ny * nz));
    int y = (int)((i / nz) % ny);
    int z = (int)(i % nz);
    T factor = -courant;
    if (electric != 0) factor = courant * eps[i];
    T v[3] = {factor * q[i], factor * q[n + i], factor * q[2 * n + i]};
    for (int axis = 0; axis < 3; ++axis) {
      int pos = axis == 0 ? x : (axis == 1 ? y : z);
      const T* co = coeff + ((int64_t)axis * maxs + pos) * 3;
      int a = (axis + 1) % 3;
      int b = (axis + 2) % 3;
      T q1 = -v[a];
      T q2 = v[b];
      T d1 = co[2] * q1;
      T d2 = co[2] * q2;
      if (co[0] != T(0)) {
        T p1 = psi[(2 * axis) * n + i] + q1;
        T p2 = psi[(2 * axis + 1) * n + i] + q2;
        d1 += co[0] * p1;
        d2 += co[0] * p2;
        psi[(2 * axis) * n + i] = co[1] * p1;
        psi[(2 * axis + 1) * n + i] = co[1] * p2;
      }
      d[(2 * axis) * n + i] = d1;
      d[(2 * axis + 1) * n + i] = d2;
    }
  }
}

// warp_fdtd/adjoint.py:91-104
template <typename T>
T read_d(const T* d, int64_t n, int slot, int x, int y, int z, int nx, int ny, int nz) {
  if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return T(0);
  return d[slot * n + ((int64_t)x * ny + y) * nz + z];
}

// warp_fdtd/adjoint.py:107-162 — transposed differences; gradient accumulates
// into an accumulator of type G (float restates the reference; double is the
// FP64 accumulation used by the product's adjoint).
template <typename T, typename G>
void gather(T* q, const T* d, const T* hist, G* grad, const int* step, int nx, int ny,
            int nz, int x0, int y0, int z0, int sx, int sy, int sz, int from_h) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t dn = (int64_t)sx * sy * sz;
  const int t = step[0];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int x = (int)(i / ((int64_t)ny * nz));
    int y = (int)((i / nz) % ny);
    int z = (int)(i % nz);
    int delta = from_h != 0 ? -1 : 1;
    T v[3];
    v[0] = (read_d(d, n, 2, x, y + delta, z, nx, ny, nz) - d[2 * n + i]) +
           (read_d(d, n, 5, x, y, z + delta, nx, ny, nz) - d[5 * n + i]);
    v[1] = (read_d(d, n, 1, x + delta, y, z, nx, ny, nz) - d[1 * n + i]) +
           (read_d(d, n, 4, x, y, z + delta, nx, ny, nz) - d[4 * n + i]);
    v[2] = (read_d(d, n, 0, x + delta, y, z, nx, ny, nz) - d[0 * n + i]) +
           (read_d(d, n, 3, x, y + delta, z, nx, ny, nz) - d[3 * n + i]);
    if (from_h == 0)
      for (int k = 0; k < 3; ++k) v[k] = -v[k];
    T value[3];
    for (int k = 0; k < 3; ++k) {
      value[k] = q[k * n + i] + v[k];
      q[k * n + i] = value[k];
    }
    if (from_h != 0 && x >= x0 && x < x0 + sx && y >= y0 && y < y0 + sy && z >= z0 &&
        z < z0 + sz) {
      int64_t j = ((int64_t)(x - x0) * sy + (y - y0)) * sz + (z - z0);
      T contrib = hist[((int64_t)t * 3 + 0) * dn + j] * value[0] +
                  hist[((int64_t)t * 3 + 1) * dn + j] * value[1] +
                  hist[((int64_t)t * 3 + 2) * dn + j] * value[2];
      grad[j] += G(contrib);
    }
  }
}

// warp_fdtd/adjoint.py:165-178
template <typename T>
void add(T* q, int64_t n, int c, int x, int y, int z, int nx, int ny, int nz, T value) {
  if (x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz)
    q[c * n + ((int64_t)x * ny + y) * nz + z] += value;
}

// warp_fdtd/adjoint.py:181-223 — transposed detector colocation (sequential
// in place of the reference's atomics).
template <typename T>
void inject(T* qe, T* qh, const T* source, const T* weights, const int* step, int nx,
            int ny, int nz, int x0, int y0, int z0, int sy, int sz, int64_t count,
            int old_h_only) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int t = step[0];
  const T w0 = weights[2 * t], w1 = weights[2 * t + 1];
  if (!(std::fabs(w0) + std::fabs(w1) > T(0))) return;
  for (int64_t i = 0; i < count; ++i) {
    int x = x0 + (int)(i / ((int64_t)sy * sz));
    int y = y0 + (int)((i / sz) % sy);
    int z = z0 + (int)(i % sz);
    T v[6];
    for (int c = 0; c < 6; ++c) {
      const T* a = source + 2 * ((int64_t)c * count + i);
      v[c] = a[0] * w0 - a[1] * w1;
    }
    if (old_h_only == 0) {
      for (int dz = 0; dz < 2; ++dz) {
        for (int dx = 0; dx < 2; ++dx) add(qe, n, 0, x - dx, y, z + dz, nx, ny, nz, T(0.25) * v[0]);
        for (int dy = 0; dy < 2; ++dy) add(qe, n, 1, x, y - dy, z + dz, nx, ny, nz, T(0.25) * v[1]);
      }
      add(qe, n, 2, x, y, z, nx, ny, nz, v[2]);
    }
    for (int dy = 0; dy < 2; ++dy) add(qh, n, 0, x, y - dy, z, nx, ny, nz, T(0.25) * v[3]);
    for (int dx = 0; dx < 2; ++dx) add(qh, n, 1, x - dx, y, z, nx, ny, nz, T(0.25) * v[4]);
    for (int dx = 0; dx < 2; ++dx)
      for (int dy = 0; dy < 2; ++dy)
        for (int dz = 0; dz < 2; ++dz)
          add(qh, n, 2, x - dx, y - dy, z + dz, nx, ny, nz, T(0.0625) * v[5]);
  }
}

// warp_fdtd/adjoint.py:472-490
template <typename T>
void write_design(const T* plane, T* eps, int nx, int ny, int nz, int x0, int y0, int z0,
                  int sx, int sy, int sz) {
  (void)nx;
  for (int x = 0; x < sx; ++x)
    for (int y = 0; y < sy; ++y)
      for (int z = 0; z < sz; ++z)
        eps[((int64_t)(x0 + x) * ny + y0 + y) * nz + z0 + z] = plane[(int64_t)x * sy + y];
}

// ---------------------------------------------------------------------------
// warp_fdtd/driven_uniform.py — staggered-array Yee with a conductivity sponge.

struct I3 {
  int v[3];
};

// driven_uniform.py:10-19
inline int64_t e_index(int c, const int p[3], const int n[3]) {
  int dims[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
  dims[c] -= 1;
  int64_t offset = 0;
  if (c >= 1) offset += (int64_t)n[0] * (n[1] + 1) * (n[2] + 1);
  if (c == 2) offset += (int64_t)(n[0] + 1) * n[1] * (n[2] + 1);
  return offset + ((int64_t)p[0] * dims[1] + p[1]) * dims[2] + p[2];
}

// driven_uniform.py:22-31
inline int64_t h_index(int c, const int p[3], const int n[3]) {
  int dims[3] = {n[0], n[1], n[2]};
  dims[c] += 1;
  int64_t offset = 0;
  if (c >= 1) offset += (int64_t)(n[0] + 1) * n[1] * n[2];
  if (c == 2) offset += (int64_t)n[0] * (n[1] + 1) * n[2];
  return offset + ((int64_t)p[0] * dims[1] + p[1]) * dims[2] + p[2];
}

// driven_uniform.py:34-40
template <typename T>
T sigma(const T p[3], const T lengths[3], T width, T rate) {
  T result = 0;
  for (int a = 0; a < 3; ++a) {
    T depth = std::max(T(0), T(1) - std::min(p[a], lengths[a] - p[a]) / width);
    result += ((rate * depth) * depth) * depth;
  }
  return result;
}

template <typename T>
T clamp01(T v) {
  return std::min(std::max(v, T(0)), T(1));
}

// driven_uniform.py:43-59
template <typename T>
T epsilon(const T p[3], T h, const T lengths[3], const T core[2], T eps_core, T eps_clad) {
  T f = 1;
  for (int a = 1; a < 3; ++a) {
    T lo = lengths[a] * T(0.5) - core[a - 1] * T(0.5);
    T hi = lengths[a] * T(0.5) + core[a - 1] * T(0.5);
    f *= clamp01((std::min(p[a] + h * T(0.5), hi) - std::max(p[a] - h * T(0.5), lo)) / h);
  }
  return eps_clad + (eps_core - eps_clad) * f;
}

// driven_uniform.py:62-87
template <typename T>
T epsilon_normal(const T p[3], T h, const T lengths[3], const T core[2], T eps_core,
                 T eps_clad, int component) {
  T fraction[2];
  for (int a = 1; a < 3; ++a) {
    T lo = lengths[a] * T(0.5) - core[a - 1] * T(0.5);
    T hi = lengths[a] * T(0.5) + core[a - 1] * T(0.5);
    fraction[a - 1] =
        clamp01((std::min(p[a] + h * T(0.5), hi) - std::max(p[a] - h * T(0.5), lo)) / h);
  }
  if (component == 0) return eps_clad + ((eps_core - eps_clad) * fraction[0]) * fraction[1];
  int normal = component - 1;
  int other = 1 - normal;
  T inside = eps_clad + (eps_core - eps_clad) * fraction[other];
  return T(1) / ((T(1) - fraction[normal]) / eps_clad + fraction[normal] / inside);
}

// driven_uniform.py:90-130 (flat) == 192-225 (xyz): H update of every stored H entry.
template <typename T>
void driven_magnetic(const T* e, T* hfield, const int n[3], T h, T dt, const T lengths[3],
                     T sponge_width, T sponge_rate) {
  const int64_t first = (int64_t)(n[0] + 1) * n[1] * n[2];
  const int64_t second = (int64_t)n[0] * (n[1] + 1) * n[2];
  const int64_t third = (int64_t)n[0] * n[1] * (n[2] + 1);
  const int64_t total = first + second + third;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < total; ++i) {
    int c = 0;
    int64_t j = i;
    if (i >= first + second) {
      c = 2;
      j = i - first - second;
    } else if (i >= first) {
      c = 1;
      j = i - first;
    }
    int dims[3] = {n[0], n[1], n[2]};
    dims[c] += 1;
    int p[3] = {(int)(j / ((int64_t)dims[1] * dims[2])), (int)((j / dims[2]) % dims[1]),
                (int)(j % dims[2])};
    int a = (c + 1) % 3, b = (c + 2) % 3;
    int pa[3] = {p[0], p[1], p[2]};
    pa[a] += 1;
    int pb[3] = {p[0], p[1], p[2]};
    pb[b] += 1;
    T curl = ((e[e_index(b, pa, n)] - e[e_index(b, p, n)]) - e[e_index(a, pb, n)]) +
             e[e_index(a, p, n)];
    T xyz[3];
    for (int k = 0; k < 3; ++k) xyz[k] = ((T)p[k] + T(0.5)) * h;
    xyz[c] -= T(0.5) * h;
    T q = (T(0.5) * dt) * sigma(xyz, lengths, sponge_width, sponge_rate);
    hfield[i] = ((T(1) - q) * hfield[i] - (dt / h) * curl) / (T(1) + q);
  }
}

// driven_uniform.py:133-189 (flat) == 228-274 (xyz): E update with the
// Gaussian impressed current on E_y and PEC tangential boundary.
template <typename T>
void driven_electric(const T* hfield, T* e, const int n[3], T h, T dt, const T lengths[3],
                     const T core[2], T eps_core, T eps_clad, T sponge_width, T sponge_rate,
                     const T src[3], const T src_sigma[3], const T* wave, const int* counter,
                     int normal_average) {
  const int64_t first = (int64_t)n[0] * (n[1] + 1) * (n[2] + 1);
  const int64_t second = (int64_t)(n[0] + 1) * n[1] * (n[2] + 1);
  const int64_t third = (int64_t)(n[0] + 1) * (n[1] + 1) * n[2];
  const int64_t total = first + second + third;
  const T w = wave[counter[0]];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < total; ++i) {
    int c = 0;
    int64_t j = i;
    if (i >= first + second) {
      c = 2;
      j = i - first - second;
    } else if (i >= first) {
      c = 1;
      j = i - first;
    }
    int dims[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
    dims[c] -= 1;
    int p[3] = {(int)(j / ((int64_t)dims[1] * dims[2])), (int)((j / dims[2]) % dims[1]),
                (int)(j % dims[2])};
    int a = (c + 1) % 3, b = (c + 2) % 3;
    if (p[a] > 0 && p[a] < n[a] && p[b] > 0 && p[b] < n[b]) {
      int pa[3] = {p[0], p[1], p[2]};
      pa[a] -= 1;
      int pb[3] = {p[0], p[1], p[2]};
      pb[b] -= 1;
      T curl = ((hfield[h_index(b, p, n)] - hfield[h_index(b, pa, n)]) -
                hfield[h_index(a, p, n)]) +
               hfield[h_index(a, pb, n)];
      T xyz[3];
      for (int k = 0; k < 3; ++k) xyz[k] = (T)p[k] * h;
      xyz[c] += T(0.5) * h;
      T ep = epsilon(xyz, h, lengths, core, eps_core, eps_clad);
      if (normal_average == 1)
        ep = epsilon_normal(xyz, h, lengths, core, eps_core, eps_clad, c);
      T current = 0;
      if (c == 1) {
        T d[3];
        for (int k = 0; k < 3; ++k) d[k] = (xyz[k] - src[k]) / src_sigma[k];
        current = std::exp(T(-0.5) * ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])) * w;
      }
      T q = (T(0.5) * dt) * sigma(xyz, lengths, sponge_width, sponge_rate);
      e[i] = ((T(1) - q) * e[i] + (dt * (curl / h + current)) / ep) / (T(1) + q);
    }
  }
}

// ---------------------------------------------------------------------------
// warp_fdtd/volume.py:238-302 — trilinear, time-centred full-volume phasors.
template <typename T>
T trilinear(const T* f, int64_t n, int component, const T p[3], int nx, int ny, int nz) {
  int ix = (int)std::floor(p[0]), iy = (int)std::floor(p[1]), iz = (int)std::floor(p[2]);
  T tx = p[0] - (T)ix, ty = p[1] - (T)iy, tz = p[2] - (T)iz;
  T value = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        int x = ix + a, y = iy + b, z = iz + c;
        if (x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz) {
          T w = ((a ? tx : T(1) - tx) * (b ? ty : T(1) - ty)) * (c ? tz : T(1) - tz);
          value += w * f[component * n + ((int64_t)x * ny + y) * nz + z];
        }
      }
  return value;
}

template <typename T>
void record_volume(const T* e, const T* h, const T* prev, const T* positions,
                   const T* weights, const int* step, T* phasor, int64_t count, int nx,
                   int ny, int nz, int first) {
  static const T off[6][3] = {{0.5, 0, 0}, {0, 0.5, 0}, {0, 0, 0.5},
                              {0, 0.5, 0.5}, {0.5, 0, 0.5}, {0.5, 0.5, 0}};
  const int64_t n = (int64_t)nx * ny * nz;
  const int t = step[0];
  if (t < first) return;
  const T wr = weights[2 * t], wi = weights[2 * t + 1];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) {
    for (int comp = 0; comp < 6; ++comp) {
      T p[3];
      for (int k = 0; k < 3; ++k) p[k] = positions[3 * i + k] - off[comp][k];
      T value;
      if (comp < 3) {
        value = trilinear(e, n, comp, p, nx, ny, nz);
      } else {
        value = T(0.5) * (trilinear(h, n, comp - 3, p, nx, ny, nz) +
                          trilinear(prev, n, comp - 3, p, nx, ny, nz));
      }
      T* q = phasor + 2 * ((int64_t)comp * count + i);
      q[0] += wr * value;
      q[1] += wi * value;
    }
  }
}

// ---------------------------------------------------------------------------
// warp_fdtd/material.py:12-55 — Gaussian filter and subpixel projection.
template <typename T>
void material_smooth(const T* p, const T* weights, T* out, int sx, int sy) {
  for (int x = 0; x < sx; ++x)
    for (int y = 0; y < sy; ++y) {
      T total = 0;
      for (int dx = -9; dx < 10; ++dx)
        for (int dy = -9; dy < 10; ++dy) {
          int ix = std::min(std::max(x + dx, 0), sx - 1);
          int iy = std::min(std::max(y + dy, 0), sy - 1);
          total += p[ix * sy + iy] * weights[(dx + 9) * 19 + dy + 9];
        }
      out[x * sy + y] = total;
    }
}

template <typename T>
T project(T r, T beta) {
  return (std::tanh(T(0.5) * beta) + std::tanh(beta * (r - T(0.5)))) /
         (T(2) * std::tanh(T(0.5) * beta));
}

template <typename T>
void material_subpixel(const T* rho, T* out, int sx, int sy, T beta) {
  for (int x = 0; x < sx; ++x)
    for (int y = 0; y < sy; ++y) {
      int xm = std::max(x - 1, 0), xp = std::min(x + 1, sx - 1);
      int ym = std::max(y - 1, 0), yp = std::min(y + 1, sy - 1);
      T gx = (rho[xp * sy + y] - rho[xm * sy + y]) / (T)(xp - xm);
      T gy = (rho[x * sy + yp] - rho[x * sy + ym]) / (T)(yp - ym);
      T norm2 = gx * gx + gy * gy;
      T r = rho[x * sy + y];
      T value = project(r, beta);
      if (norm2 > T(0)) {
        T norm = std::sqrt(norm2);
        T d = (T(0.5) - r) / (T(0.55) * norm);
        if (std::fabs(d) < T(1)) {
          T fill = T(0.5) - T(0.9375) * d + ((T(0.625) * d) * d) * d -
                   ((((T(0.1875) * d) * d) * d) * d) * d;
          T lo = project(r - (T(0.55) * norm) * fill, beta);
          T hi = project(r + (T(0.55) * norm) * (T(1) - fill), beta);
          value = (T(1) - fill) * lo + fill * hi;
        }
      }
      out[x * sy + y] = T(1) / (T(1) + T(11.25) * value);
    }
}

}  // namespace

extern "C" {

int ora_version() { return 1; }

#define ORA_INSTANTIATE(T, SUF)                                                           \
  void ora_update_##SUF(int electric, int packed, T* f, const T* g, const T* inv_eps,   \
                        T* psi, int64_t psi_stride, const T* coeff, int maxs,            \
                        const T* source, int64_t source_stride, const T* waves,          \
                        int64_t waves_stride, const int* step, T* hprev,                 \
                        int save_h_from, int nx, int ny, int nz, int source_x,           \
                        T courant, int periodic) {                                       \
    update<T>(electric, packed, f, g, inv_eps, psi, psi_stride, coeff, maxs, source,    \
              source_stride, waves, waves_stride, step, hprev, save_h_from, nx, ny, nz,  \
              source_x, courant, periodic);                                              \
  }                                                                                      \
  void ora_record_##SUF(const T* e, const T* h, const T* prev, const int* step,         \
                        const T* weights, const int* indices, T* phasor, T* flux,        \
                        const T* area, int nx, int ny, int nz, int x0, int y0, int z0,   \
                        int sx, int sy, int sz, int axis, int exact, int is_phasor) {    \
    record<T, T>(e, h, prev, step, weights, indices, phasor, flux, area, nx, ny, nz, x0, \
                 y0, z0, sx, sy, sz, axis, exact, is_phasor);                            \
  }                                                                                      \
  void ora_record_dft_##SUF(const T* e, const T* h, const T* prev, int t,               \
                            const double* weights, int nf, double* phasor, int nx,       \
                            int ny, int nz, int x0, int y0, int z0, int sx, int sy,      \
                            int sz, int exact) {                                         \
    record_dft<T>(e, h, prev, t, weights, nf, phasor, nx, ny, nz, x0, y0, z0, sx, sy,    \
                  sz, exact);                                                            \
  }                                                                                      \
  void ora_record_curl_##SUF(const T* h, T* hist, const int* step, int nx, int ny,      \
                             int nz, int x0, int y0, int z0, int sx, int sy, int sz,     \
                             T courant) {                                                \
    record_curl<T>(h, hist, step, nx, ny, nz, x0, y0, z0, sx, sy, sz, courant);          \
  }                                                                                      \
  void ora_prepare_##SUF(const T* q, const T* eps, T* psi, const T* coeff, int maxs,    \
                         T* d, int nx, int ny, int nz, T courant, int electric) {        \
    prepare<T>(q, eps, psi, coeff, maxs, d, nx, ny, nz, courant, electric);              \
  }                                                                                      \
  void ora_gather_##SUF(T* q, const T* d, const T* hist, T* grad, const int* step,      \
                        int nx, int ny, int nz, int x0, int y0, int z0, int sx, int sy,  \
                        int sz, int from_h) {                                            \
    gather<T, T>(q, d, hist, grad, step, nx, ny, nz, x0, y0, z0, sx, sy, sz, from_h);    \
  }                                                                                      \
  void ora_gather_f64acc_##SUF(T* q, const T* d, const T* hist, double* grad,           \
                               const int* step, int nx, int ny, int nz, int x0, int y0,  \
                               int z0, int sx, int sy, int sz, int from_h) {             \
    gather<T, double>(q, d, hist, grad, step, nx, ny, nz, x0, y0, z0, sx, sy, sz,        \
                      from_h);                                                           \
  }                                                                                      \
  void ora_inject_##SUF(T* qe, T* qh, const T* source, const T* weights,                \
                        const int* step, int nx, int ny, int nz, int x0, int y0, int z0, \
                        int sy, int sz, int64_t count, int old_h_only) {                 \
    inject<T>(qe, qh, source, weights, step, nx, ny, nz, x0, y0, z0, sy, sz, count,      \
              old_h_only);                                                               \
  }                                                                                      \
  void ora_write_design_##SUF(const T* plane, T* eps, int nx, int ny, int nz, int x0,   \
                              int y0, int z0, int sx, int sy, int sz) {                  \
    write_design<T>(plane, eps, nx, ny, nz, x0, y0, z0, sx, sy, sz);                     \
  }                                                                                      \
  void ora_driven_magnetic_##SUF(const T* e, T* hfield, const int* n, T h, T dt,        \
                                 const T* lengths, T sponge_width, T sponge_rate) {      \
    driven_magnetic<T>(e, hfield, n, h, dt, lengths, sponge_width, sponge_rate);         \
  }                                                                                      \
  void ora_driven_electric_##SUF(const T* hfield, T* e, const int* n, T h, T dt,        \
                                 const T* lengths, const T* core, T eps_core,            \
                                 T eps_clad, T sponge_width, T sponge_rate,              \
                                 const T* src, const T* src_sigma, const T* wave,        \
                                 const int* counter, int normal_average) {               \
    driven_electric<T>(hfield, e, n, h, dt, lengths, core, eps_core, eps_clad,           \
                       sponge_width, sponge_rate, src, src_sigma, wave, counter,         \
                       normal_average);                                                  \
  }                                                                                      \
  T ora_epsilon_normal_##SUF(const T* p, T h, const T* lengths, const T* core,          \
                             T eps_core, T eps_clad, int component) {                    \
    return epsilon_normal<T>(p, h, lengths, core, eps_core, eps_clad, component);        \
  }                                                                                      \
  void ora_record_volume_##SUF(const T* e, const T* h, const T* prev,                   \
                               const T* positions, const T* weights, const int* step,    \
                               T* phasor, int64_t count, int nx, int ny, int nz,         \
                               int first) {                                              \
    record_volume<T>(e, h, prev, positions, weights, step, phasor, count, nx, ny, nz,    \
                     first);                                                             \
  }                                                                                      \
  void ora_material_smooth_##SUF(const T* p, const T* w, T* out, int sx, int sy) {      \
    material_smooth<T>(p, w, out, sx, sy);                                               \
  }                                                                                      \
  void ora_material_subpixel_##SUF(const T* rho, T* out, int sx, int sy, T beta) {      \
    material_subpixel<T>(rho, out, sx, sy, beta);                                        \
  }

ORA_INSTANTIATE(float, f32)
ORA_INSTANTIATE(double, f64)

// OpenMP threads of the oracle's parallel loops (the CPU baseline's 1-thread
// and all-core figures); returns the previous setting.
int ora_set_threads(int threads) {
  const int previous = omp_get_max_threads();
  if (threads > 0) omp_set_num_threads(threads);
  return previous;
}

}  // extern "C"
