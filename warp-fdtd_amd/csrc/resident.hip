// On-chip-resident Yee stepping for gfx950 (grids whose E, H and 1/eps fit
// in the LDS of the co-resident workgroups: 160 KB x one workgroup per CU).
//
// The grid is cut into sub-boxes of sx x sy x sz cells, one 1024-thread
// workgroup each, which holds the sub-box's E, H, 1/eps and CPML coefficient
// rows in LDS for a whole chunk of steps.  A half-step reads the stencil
// neighbours from LDS and the sub-box faces of the neighbouring workgroups
// from global face buffers.
//
// Face exchange: every face value travels as one 64-bit word, the float's
// bits in the low half and a step tag in the high half, written and read
// with relaxed agent-scope atomics.  A reader polls each word until its tag
// is the one it needs, so a face value and the proof that it is current
// arrive together: no flags, no release/acquire fences and no cache
// write-back or invalidation (those cost more than the arithmetic here).
// Tags: a launch starts from the workspace's epoch word b (read once) and
// tags the H faces of the state after k of its steps b + 2k + 1 and the E
// faces after the E half of its step k b + 2k + 2; the counter kernel that
// follows the launch advances the epoch past every tag of it (by 2n + 2 for
// n steps).  Every face word is rewritten in every step, so at the start of
// a launch the words hold tags of the launch before it, all below the new
// base by at most 2n' + 1 (n' < 2^30 steps): modulo 2^32 none equals a tag
// the new launch waits for, and the words need no clearing between
// launches -- only once, when the workspace is first used (zeroed).  A
// writer can only overwrite a word its reader has consumed: overwriting
// needs the reader's own next faces, which the reader writes after
// consuming.
//
// Inside the chunk only CPML psi (slab cells), the faces and detector staging
// touch memory; state is loaded at the start and written back at the end.
// The arithmetic is the fused step's (yee.hip) operation for operation, so
// results are bitwise those of the streaming kernels.
//
// Termination: every poll loop is bounded (spin_limit polls); a thread that
// times out raises the error word, which every poller checks, and its
// workgroup leaves; the launch is cooperative, which guarantees co-residency
// of the whole grid or fails.
#include <mutex>

#include "common.hpp"

namespace wfd {
namespace {

constexpr int kThreads = 1024;
constexpr int kLds = 40896;  // floats of LDS per workgroup (163,584 bytes)

typedef unsigned long long Word;

struct ResArgs {
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
  const int* counters;
  int nstage, slots;
  StageBox stages[WFD_MAX_STAGES];
  int rows, steps, nx, ny, nz, n, source_x;
  float courant;
  int sx, sy, sz, gx, gy, gz;
  Word* face;
  int* error;
  const unsigned* epoch;
  int spin_limit;
};

// LDS floats a sub-box needs: E, H, 1/eps per cell and (a, b, 1/kappa) rows
// of E and H along each axis.
__host__ __device__ inline long lds_floats(int sx, int sy, int sz) {
  return 7L * sx * sy * sz + 6L * (sx + sy + sz);
}

// Words of one workgroup's face buffer: H high faces then E low faces, two
// components each, laid out x (sy*sz), y (sx*sz), z (sx*sy).
__host__ __device__ inline size_t face_words(int sx, int sy, int sz) {
  return 4 * ((size_t)sy * sz + (size_t)sx * sz + (size_t)sx * sy);
}

__device__ __forceinline__ Word pack(float v, unsigned tag) {
  return ((Word)tag << 32) | (Word)__float_as_uint(v);
}

__device__ __forceinline__ void put(Word* p, float v, unsigned tag) {
  __hip_atomic_store(p, pack(v, tag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Poll a word until it carries `tag`; false once the error word is raised
// or the poll limit is reached (which raises it).
__device__ __forceinline__ bool take(const ResArgs& a, const Word* p, unsigned tag, float& v) {
  int spins = 0;
  for (;;) {
    const Word w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned)(w >> 32) == tag) {
      v = __uint_as_float((unsigned)w);
      return true;
    }
    if (++spins > a.spin_limit) {
      __hip_atomic_store(a.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((spins & 63) == 0 &&
        __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
      return false;
    if (spins < 8)
      __builtin_amdgcn_s_sleep(2);
    else
      __builtin_amdgcn_s_sleep(8);
  }
}

// One CPML partial with the psi value already loaded (kernels.py:29-47); the
// new psi goes to byte offset voff + soff of the buffer.
__device__ __forceinline__ float cpml_pre(float d, float a, float b, float k, float psi_old,
                                          Rsrc r, int voff, int soff) {
  if (a != 0.f) {
    const float p = b * psi_old + a * d;
    st(r, p, voff, soff);
    return k * d + p;
  }
  return k * d;
}

// An opaque copy: the compiler cannot see through it, so values derived from
// it inside the step loop are recomputed there instead of being hoisted out
// of the loop and held in registers for every step (the face addresses of a
// column would not fit, and their spills cost more than recomputing them).
template <class T>
__device__ __forceinline__ T opaque_v(T v) {
  asm volatile("" : "+v"(v));
  return v;
}
template <class T>
__device__ __forceinline__ T opaque_s(T v) {
  asm volatile("" : "+s"(v));
  return v;
}

__global__ __launch_bounds__(kThreads) void k_resident(ResArgs a, int nsteps) {
  __shared__ float lds[kLds];
  __shared__ int bail;
  const int tid = threadIdx.x;
  const int wg = blockIdx.x;
  // x fastest: workgroups are dealt to the XCDs round-robin, so consecutive
  // indices along x spread the CPML slabs (whose psi traffic loads the L2)
  // over all XCDs; with z or y fastest each y row of sub-boxes shared one
  // XCD, and the two CPML rows loaded two of the eight L2s (8 % slower)
  const int bxi = wg % a.gx, byi = (wg / a.gx) % a.gy, bzi = wg / (a.gx * a.gy);
  const int X0 = bxi * a.sx, Y0 = byi * a.sy, Z0 = bzi * a.sz;
  const int vx = min(a.sx, a.nx - X0), vy = min(a.sy, a.ny - Y0), vz = min(a.sz, a.nz - Z0);
  const int sx = a.sx, sy = a.sy, sz = a.sz;
  const int cells = sx * sy * sz;  // LDS slots (the box may be partly outside)
  const int fyz = sy * sz, fxz = sx * sz, fxy = sx * sy;
  const int R = sx + sy + sz;
  float* E = lds;                 // [3][cells]
  float* H = lds + 3 * cells;     // [3][cells]
  float* IE = lds + 6 * cells;    // [cells]
  float* co = lds + 7 * cells;    // [2][R][3]: E rows then H rows (x | y | z)
  const int ny = a.ny, nz = a.nz, P = ny * nz, n = a.n;
  const int sxy = a.gx * a.gy;
  const int wxl = bxi > 0 ? wg - 1 : -1, wxh = bxi + 1 < a.gx ? wg + 1 : -1;
  const int wyl = byi > 0 ? wg - a.gx : -1, wyh = byi + 1 < a.gy ? wg + a.gx : -1;
  const int wzl = bzi > 0 ? wg - sxy : -1, wzh = bzi + 1 < a.gz ? wg + sxy : -1;
  const size_t per = face_words(sx, sy, sz);
  const size_t half = per / 2;
  Word* myface = a.face + (size_t)wg * per;
  // a raised error word (an earlier run gave up) stops the launch, uniformly
  if (tid == 0) bail = __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (bail) return;
  const unsigned base = *a.epoch;
  const int t0 = a.counters[0];
  const float S = a.courant;
  const int rows = a.rows;
  for (int r = tid; r < 2 * R; r += kThreads) {
    const int f = r / R, q = r % R;
    const float* src = f ? a.coeff_h : a.coeff_e;
    int axis, pos;
    if (q < sx) {
      axis = 0;
      pos = min(X0 + q, a.nx - 1);
    } else if (q < sx + sy) {
      axis = 1;
      pos = min(Y0 + q - sx, ny - 1);
    } else {
      axis = 2;
      pos = min(Z0 + q - sx - sy, nz - 1);
    }
    for (int k = 0; k < 3; ++k) co[(f * R + q) * 3 + k] = src[(axis * rows + pos) * 3 + k];
  }
  const float* cE = co;
  const float* cH = co + 3 * R;

  // cell c = (lx*sy + ly)*sz + lz
  const float inv_sz = 1.0f / (float)sz, inv_sy = 1.0f / (float)sy;
  auto coords = [&](int c, int& lx, int& ly, int& lz) {
    const int r = div_small(c, sz, inv_sz);
    lz = c - r * sz;
    lx = div_small(r, sy, inv_sy);
    ly = r - lx * sy;
  };

#pragma unroll 1
  for (int c = tid; c < cells; c += kThreads) {
    int lx, ly, lz;
    coords(c, lx, ly, lz);
    float v[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (lx < vx && ly < vy && lz < vz) {
      const int i = ((X0 + lx) * ny + Y0 + ly) * nz + Z0 + lz;
      v[0] = a.e[i];
      v[1] = a.e[n + i];
      v[2] = a.e[2 * n + i];
      v[3] = a.h[i];
      v[4] = a.h[n + i];
      v[5] = a.h[2 * n + i];
      v[6] = a.inv_eps[i];
    }
    for (int k = 0; k < 3; ++k) {
      E[k * cells + c] = v[k];
      H[k * cells + c] = v[3 + k];
    }
    IE[c] = v[6];
  }
  __syncthreads();

  // The initial H faces (tag base + 1): components (1,2) on x, (0,2) on y, (0,1) on z.
  for (int t = tid; t < 2 * vy * vz; t += kThreads) {
    const int comp = t >= vy * vz, r = t - comp * vy * vz;
    const int ly = r / vz, lz = r - ly * vz;
    put(myface + comp * fyz + ly * sz + lz,
        H[(1 + comp) * cells + ((vx - 1) * sy + ly) * sz + lz], base + 1u);
  }
  for (int t = tid; t < 2 * vx * vz; t += kThreads) {
    const int comp = t >= vx * vz, r = t - comp * vx * vz;
    const int lx = r / vz, lz = r - lx * vz;
    put(myface + 2 * fyz + comp * fxz + lx * sz + lz,
        H[2 * comp * cells + (lx * sy + vy - 1) * sz + lz], base + 1u);
  }
  for (int t = tid; t < 2 * vx * vy; t += kThreads) {
    const int comp = t >= vx * vy, r = t - comp * vx * vy;
    const int lx = r / vy, ly = r - lx * vy;
    put(myface + 2 * fyz + 2 * fxz + comp * fxy + lx * sy + ly,
        H[comp * cells + (lx * sy + ly) * sz + vz - 1], base + 1u);
  }

  // Columns: thread t < vy*vz owns one (y, z) column per half-step and marches
  // x.  The columns that read neighbour faces in y or z are the first
  // vy + vz - 1 threads (one wavefront for sub-boxes with vy + vz <= 65), so the
  // other wavefronts never wait on them; every face word is polled by the
  // thread that uses it, right when it needs it, and sent by the thread that
  // computed it, right after computing it.  Per-word dependencies order every
  // overwrite after the read of the previous value (see the file header), so
  // the half-steps need no barrier beyond the LDS one between them.
  const int ncol = vy * vz;
  const bool active = tid < ncol;
  const int nb = vy + vz - 1, w1 = vz - 1;
  // the waves holding face columns (they also poll the neighbours' faces)
  // issue ahead of the others: the barrier that ends each half waits on them
  if ((tid & ~63) < nb) __builtin_amdgcn_s_setprio(2);
  int eyl, ezl, hyl, hzl;
  if (tid < vz) {
    eyl = 0;
    ezl = tid;
    hyl = vy - 1;
    hzl = tid;
  } else if (tid < nb) {
    eyl = tid - vz + 1;
    ezl = 0;
    hyl = tid - vz;
    hzl = vz - 1;
  } else {
    const int m = tid - nb;
    eyl = w1 > 0 ? 1 + m / w1 : 0;
    ezl = w1 > 0 ? 1 + m % w1 : 0;
    hyl = w1 > 0 ? m / w1 : 0;
    hzl = w1 > 0 ? m % w1 : 0;
  }
  const int plane = sy * sz;

  // psi of one plane of a column (zero where the axis has no CPML there)
  // psi component k of cell i sits at byte i * 4 + k * pb of the buffers
  const int pb = n * 4;
  const Rsrc rpe = buffer(a.psi_e, 6u * (unsigned)pb), rph = buffer(a.psi_h, 6u * (unsigned)pb);
  auto load_psi = [&](Rsrc r, const float* cf, float ya, float za, int i, int lx,
                      float (&p)[6]) {
    const bool ax = cf[lx * 3] != 0.f;
    const int v = i * 4;
    p[0] = ax ? ld(r, v, 0) : 0.f;
    p[1] = ax ? ld(r, v, pb) : 0.f;
    p[2] = ya != 0.f ? ld(r, v, 2 * pb) : 0.f;
    p[3] = ya != 0.f ? ld(r, v, 3 * pb) : 0.f;
    p[4] = za != 0.f ? ld(r, v, 4 * pb) : 0.f;
    p[5] = za != 0.f ? ld(r, v, 5 * pb) : 0.f;
  };  // a face value: the word already loaded if it carries the tag, else poll
  auto value = [&](const Word* p, Word w, unsigned tag) {
    if ((unsigned)(w >> 32) == tag) return __uint_as_float((unsigned)w);
    float v = 0.f;
    if (!take(a, p, tag, v)) bail = 1;
    return v;
  };
  auto peek = [&](const Word* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };

  // the plane source's profile at the thread's columns, loaded once per
  // launch (a load inside the march would stall the source slab's
  // workgroups, and their neighbours wait on them)
  float src_ey = 0.f, src_ez = 0.f, src_hy = 0.f, src_hz = 0.f;
  if (active && a.source_x >= X0 && a.source_x < X0 + vx) {
    const int oe = (Y0 + eyl) * nz + Z0 + ezl, oh = (Y0 + hyl) * nz + Z0 + hzl;
    src_ey = a.src_h[2 * P + oe];
    src_ez = a.src_h[P + oe];
    src_hy = a.src_e[2 * P + oh];
    src_hz = a.src_e[P + oh];
  }

  for (int k = 0; k < nsteps; ++k) {
    const int t = t0 + k;
    // ------------------------------------------------------------ E half
    if (active) {
      const unsigned tag = base + 2u * k + 1u, out = base + 2u * k + 2u;
      Word* const face = opaque_s(a.face);
      Word* const myface = face + (size_t)wg * per;
      const int ly = opaque_v(eyl), lz = opaque_v(ezl), col = ly * sz + lz;
      const int y = Y0 + ly, z = Z0 + lz, o = y * nz + z;
      const float ya = cE[(sx + ly) * 3], yb = cE[(sx + ly) * 3 + 1], yk = cE[(sx + ly) * 3 + 2];
      const float za = cE[(sx + sy + lz) * 3], zb = cE[(sx + sy + lz) * 3 + 1],
                  zk = cE[(sx + sy + lz) * 3 + 2];
      const bool fy = ly == 0 && wyl >= 0, fz = lz == 0 && wzl >= 0;
      const Word* px = wxl >= 0 ? face + (size_t)wxl * per + ly * sz + lz : nullptr;
      const Word* py = fy ? face + (size_t)wyl * per + 2 * fyz + lz : nullptr;
      const Word* pz = fz ? face + (size_t)wzl * per + 2 * fyz + 2 * fxz + ly : nullptr;
      const bool src_here = a.source_x >= X0 && a.source_x < X0 + vx && t < a.steps;
      const float wave_y = src_here ? a.wave_h[2 * a.steps + t] : 0.f;
      const float wave_z = src_here ? a.wave_h[a.steps + t] : 0.f;
      // march lx = 1, 2, ..., vx - 1, then 0 (the plane that needs the x face)
      float my = vx > 1 ? H[cells + col] : 0.f, mz = vx > 1 ? H[2 * cells + col] : 0.f;
      int lx = vx > 1 ? 1 : 0;
      float cur[6], nxt[6];
      Word wy[2] = {0, 0}, wz[2] = {0, 0};
      load_psi(rpe, cE, ya, za, (X0 + lx) * P + o, lx, cur);
      if (fy) {
        wy[0] = peek(py + lx * sz);
        wy[1] = peek(py + fxz + lx * sz);
      }
      if (fz) {
        wz[0] = peek(pz + lx * sy);
        wz[1] = peek(pz + fxy + lx * sy);
      }
      Word wx0 = 0, wx1 = 0;
      if (px) {
        wx0 = peek(px);
        wx1 = peek(px + fyz);
      }
#pragma unroll 1
      for (int q = 0; q < vx; ++q) {
        const int nlx = q + 2 < vx ? lx + 1 : 0;  // next plane of the march
        Word ny0 = 0, ny1 = 0, nz0 = 0, nz1 = 0;
        if (q + 1 < vx) {
          load_psi(rpe, cE, ya, za, (X0 + nlx) * P + o, nlx, nxt);
          if (fy) {
            ny0 = peek(py + nlx * sz);
            ny1 = peek(py + fxz + nlx * sz);
          }
          if (fz) {
            nz0 = peek(pz + nlx * sy);
            nz1 = peek(pz + fxy + nlx * sy);
          }
        }
        const int c = lx * plane + col;
        const int x = X0 + lx;
        const int i = x * P + o;
        const float hx = H[c], hy = H[cells + c], hz = H[2 * cells + c];
        float xm_y, xm_z;
        if (lx > 0) {
          xm_y = my;
          xm_z = mz;
        } else if (px) {
          xm_y = value(px, wx0, tag);
          xm_z = value(px + fyz, wx1, tag);
        } else {
          xm_y = xm_z = 0.f;
        }
        float yx, yz, zx, zy;
        if (ly > 0) {
          yx = H[c - sz];
          yz = H[2 * cells + c - sz];
        } else if (fy) {
          yx = value(py + lx * sz, wy[0], tag);
          yz = value(py + fxz + lx * sz, wy[1], tag);
        } else {
          yx = yz = 0.f;
        }
        if (lz > 0) {
          zx = H[c - 1];
          zy = H[cells + c - 1];
        } else if (fz) {
          zx = value(pz + lx * sy, wz[0], tag);
          zy = value(pz + fxy + lx * sy, wz[1], tag);
        } else {
          zx = zy = 0.f;
        }
        const float dx1 = hy - xm_y, dx2 = hz - xm_z;
        const float dy0 = hx - yx, dy2 = hz - yz;
        const float dz0 = hx - zx, dz1 = hy - zy;
        const float xa = cE[lx * 3], xb = cE[lx * 3 + 1], xk = cE[lx * 3 + 2];
        const int v = i * 4;
        const float dxz = cpml_pre(dx2, xa, xb, xk, cur[0], rpe, v, 0);
        const float dxy = cpml_pre(dx1, xa, xb, xk, cur[1], rpe, v, pb);
        const float dyx = cpml_pre(dy0, ya, yb, yk, cur[2], rpe, v, 2 * pb);
        const float dyz = cpml_pre(dy2, ya, yb, yk, cur[3], rpe, v, 3 * pb);
        const float dzy = cpml_pre(dz1, za, zb, zk, cur[4], rpe, v, 4 * pb);
        const float dzx = cpml_pre(dz0, za, zb, zk, cur[5], rpe, v, 5 * pb);
        const float ie = IE[c];
        float ex = E[c] + (S * (dyz - dzy)) * ie;
        float ey = E[cells + c] + (S * (dzx - dxz)) * ie;
        float ez = E[2 * cells + c] + (S * (dxy - dyx)) * ie;
        if (src_here && x == a.source_x) {
          ey += ((src_ey * wave_y) * S) * ie;
          ez -= ((src_ez * wave_z) * S) * ie;
        }
        E[c] = ex;
        E[cells + c] = ey;
        E[2 * cells + c] = ez;
        // low E faces for the low neighbours' H half
        Word* f = myface + half;
        if (lx == 0) {
          put(f + ly * sz + lz, ey, out);
          put(f + fyz + ly * sz + lz, ez, out);
        }
        if (ly == 0) {
          put(f + 2 * fyz + lx * sz + lz, ex, out);
          put(f + 2 * fyz + fxz + lx * sz + lz, ez, out);
        }
        if (lz == 0) {
          put(f + 2 * fyz + 2 * fxz + lx * sy + ly, ex, out);
          put(f + 2 * fyz + 2 * fxz + fxy + lx * sy + ly, ey, out);
        }
        my = hy;
        mz = hz;
        lx = nlx;
        wy[0] = ny0;
        wy[1] = ny1;
        wz[0] = nz0;
        wz[1] = nz1;
#pragma unroll
        for (int m = 0; m < 6; ++m) cur[m] = nxt[m];
      }
    }
    __syncthreads();
    if (bail) return;
    // ------------------------------------------------------------ H half
    if (active) {
      const unsigned tag = base + 2u * k + 2u, out = base + 2u * k + 3u;
      Word* const face = opaque_s(a.face);
      Word* const myface = face + (size_t)wg * per;
      const int ly = opaque_v(hyl), lz = opaque_v(hzl), col = ly * sz + lz;
      const int y = Y0 + ly, z = Z0 + lz, o = y * nz + z;
      const float ya = cH[(sx + ly) * 3], yb = cH[(sx + ly) * 3 + 1], yk = cH[(sx + ly) * 3 + 2];
      const float za = cH[(sx + sy + lz) * 3], zb = cH[(sx + sy + lz) * 3 + 1],
                  zk = cH[(sx + sy + lz) * 3 + 2];
      const bool fy = ly == vy - 1 && wyh >= 0, fz = lz == vz - 1 && wzh >= 0;
      const Word* px = wxh >= 0 ? face + (size_t)wxh * per + half + ly * sz + lz : nullptr;
      const Word* py = fy ? face + (size_t)wyh * per + half + 2 * fyz + lz : nullptr;
      const Word* pz = fz ? face + (size_t)wzh * per + half + 2 * fyz + 2 * fxz + ly : nullptr;
      const bool src_here = a.source_x >= X0 && a.source_x < X0 + vx && t < a.steps;
      const float wave_y = src_here ? a.wave_e[2 * a.steps + t] : 0.f;
      const float wave_z = src_here ? a.wave_e[a.steps + t] : 0.f;
      const int slot = a.nstage ? t % a.slots : 0;
      // march lx = 0, 1, ..., vx - 1 (the last plane needs the x face)
      float ex = E[col], ey = E[cells + col], ez = E[2 * cells + col];
      float cur[6], nxt[6];
      Word wy[2] = {0, 0}, wz[2] = {0, 0};
      load_psi(rph, cH, ya, za, X0 * P + o, 0, cur);
      if (fy) {
        wy[0] = peek(py);
        wy[1] = peek(py + fxz);
      }
      if (fz) {
        wz[0] = peek(pz);
        wz[1] = peek(pz + fxy);
      }
      Word wx0 = 0, wx1 = 0;
      if (px) {
        wx0 = peek(px);
        wx1 = peek(px + fyz);
      }
#pragma unroll 1
      for (int lx = 0; lx < vx; ++lx) {
        Word ny0 = 0, ny1 = 0, nz0 = 0, nz1 = 0;
        if (lx + 1 < vx) {
          load_psi(rph, cH, ya, za, (X0 + lx + 1) * P + o, lx + 1, nxt);
          if (fy) {
            ny0 = peek(py + (lx + 1) * sz);
            ny1 = peek(py + fxz + (lx + 1) * sz);
          }
          if (fz) {
            nz0 = peek(pz + (lx + 1) * sy);
            nz1 = peek(pz + fxy + (lx + 1) * sy);
          }
        }
        const int c = lx * plane + col;
        const int x = X0 + lx;
        const int i = x * P + o;
        float xp_x = 0.f, xp_y, xp_z;  // E of plane lx + 1
        if (lx + 1 < vx) {
          xp_x = E[c + plane];
          xp_y = E[cells + c + plane];
          xp_z = E[2 * cells + c + plane];
        } else if (px) {
          xp_y = value(px, wx0, tag);
          xp_z = value(px + fyz, wx1, tag);
        } else {
          xp_y = xp_z = 0.f;
        }
        float qx, qz, rx, ry;
        if (ly + 1 < vy) {
          qx = E[c + sz];
          qz = E[2 * cells + c + sz];
        } else if (fy) {
          qx = value(py + lx * sz, wy[0], tag);
          qz = value(py + fxz + lx * sz, wy[1], tag);
        } else {
          qx = qz = 0.f;
        }
        if (lz + 1 < vz) {
          rx = E[c + 1];
          ry = E[cells + c + 1];
        } else if (fz) {
          rx = value(pz + lx * sy, wz[0], tag);
          ry = value(pz + fxy + lx * sy, wz[1], tag);
        } else {
          rx = ry = 0.f;
        }
        const float dx1 = xp_y - ey, dx2 = xp_z - ez;
        const float dy0 = qx - ex, dy2 = qz - ez;
        const float dz0 = rx - ex, dz1 = ry - ey;
        const float xa = cH[lx * 3], xb = cH[lx * 3 + 1], xk = cH[lx * 3 + 2];
        const int v = i * 4;
        const float dxz = cpml_pre(dx2, xa, xb, xk, cur[0], rph, v, 0);
        const float dxy = cpml_pre(dx1, xa, xb, xk, cur[1], rph, v, pb);
        const float dyx = cpml_pre(dy0, ya, yb, yk, cur[2], rph, v, 2 * pb);
        const float dyz = cpml_pre(dy2, ya, yb, yk, cur[3], rph, v, 3 * pb);
        const float dzy = cpml_pre(dz1, za, zb, zk, cur[4], rph, v, 4 * pb);
        const float dzx = cpml_pre(dz0, za, zb, zk, cur[5], rph, v, 5 * pb);
        const float ox = H[c], oy = H[cells + c], oz = H[2 * cells + c];
        float hx = ox - S * (dyz - dzy);
        float hy = oy - S * (dzx - dxz);
        float hz = oz - S * (dxy - dyx);
        if (src_here && x == a.source_x) {
          hy -= (src_hy * wave_y) * S;
          hz += (src_hz * wave_z) * S;
        }
        H[c] = hx;
        H[cells + c] = hy;
        H[2 * cells + c] = hz;
        // high H faces for the high neighbours' next E half
        if (lx == vx - 1) {
          put(myface + ly * sz + lz, hy, out);
          put(myface + fyz + ly * sz + lz, hz, out);
        }
        if (ly == vy - 1) {
          put(myface + 2 * fyz + lx * sz + lz, hx, out);
          put(myface + 2 * fyz + fxz + lx * sz + lz, hz, out);
        }
        if (lz == vz - 1) {
          put(myface + 2 * fyz + 2 * fxz + lx * sy + ly, hx, out);
          put(myface + 2 * fyz + 2 * fxz + fxy + lx * sy + ly, hy, out);
        }
        for (int b = 0; b < a.nstage; ++b) {
          const StageBox& st = a.stages[b];
          if (x >= st.x0 && x < st.x1 && y >= st.y0 && y < st.y1 && z >= st.z0 && z < st.z1) {
            const int cnt = st.count;
            float* d = st.buf + (size_t)slot * 9 * cnt +
                       ((x - st.x0) * (st.y1 - st.y0) + (y - st.y0)) * (st.z1 - st.z0) +
                       (z - st.z0);
            // streaming stores: the ring is read once, by the replay
            __builtin_nontemporal_store(ex, d);
            __builtin_nontemporal_store(ey, d + cnt);
            __builtin_nontemporal_store(ez, d + 2 * cnt);
            __builtin_nontemporal_store(hx, d + 3 * cnt);
            __builtin_nontemporal_store(hy, d + 4 * cnt);
            __builtin_nontemporal_store(hz, d + 5 * cnt);
            __builtin_nontemporal_store(ox, d + 6 * cnt);
            __builtin_nontemporal_store(oy, d + 7 * cnt);
            __builtin_nontemporal_store(oz, d + 8 * cnt);
          }
        }
        ex = xp_x;
        ey = xp_y;
        ez = xp_z;
        wy[0] = ny0;
        wy[1] = ny1;
        wz[0] = nz0;
        wz[1] = nz1;
#pragma unroll
        for (int m = 0; m < 6; ++m) cur[m] = nxt[m];
      }
    }
    __syncthreads();
    if (bail) return;
  }
#pragma unroll 1
  for (int c = tid; c < cells; c += kThreads) {
    int lx, ly, lz;
    coords(c, lx, ly, lz);
    if (lx < vx && ly < vy && lz < vz) {
      const int i = ((X0 + lx) * ny + Y0 + ly) * nz + Z0 + lz;
      for (int k = 0; k < 3; ++k) {
        a.e[k * n + i] = E[k * cells + c];
        a.h[k * n + i] = H[k * cells + c];
      }
    }
  }
}

// After a launch of nsteps: the step counters, the face epoch past every
// tag of the launch and (when given) the error word copied to `status`.
__global__ void k_advance_counters(int* counters, unsigned* epoch, int nsteps, const int* error,
                                   int* status) {
  counters[0] += nsteps;
  counters[1] = counters[0] - 1;
  epoch[0] += 2u * (unsigned)nsteps + 2u;
  if (status) status[0] = error[0];
}

// Zero `count` 32-bit words (a kernel rather than hipMemsetAsync, so that the
// clearing is an ordinary kernel node when a caller captures the stream).
__global__ void k_fill_zero(unsigned* w, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    w[i] = 0u;
}

// Scratch of wfd_resident_run_noalloc: face words, error word and epoch of
// one launch, in device memory that the library owns (one copy per GPU,
// held until the process ends).  Host code never passes `scratch` itself to
// a kernel (in host code the name is a stand-in, not a device address): it
// uses the address hipGetSymbolAddress returns on each device.
constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call
// made with that device current.
int scratch_address(void** out) {
  int device = 0;
  WFD_HIP(hipGetDevice(&device));
  WFD_REQUIRE(device >= 0 && device < kMaxDevices, "resident_run_noalloc: device %d beyond %d",
              device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    WFD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    WFD_REQUIRE(p != nullptr, "resident_run_noalloc: no device address for the scratch array");
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return WFD_OK;
}

struct Layout {
  int sx, sy, sz, gx, gy, gz, blocks;
};

// Sub-box choice: one thread per (y, z) column (sy*sz <= 1024), the fewest x
// planes per column, then the smaller sub-box surface (face traffic), then
// fewer workgroups; every candidate fits the LDS (lds_floats <= kLds).
bool choose(int nx, int ny, int nz, int capacity, int want_sx, int want_sy, int want_sz,
            Layout& best) {
  bool found = false;
  long best_key[3] = {0, 0, 0};
  for (int sz = 1; sz <= nz; ++sz) {
    if (want_sz && sz != want_sz) continue;
    for (int sy = 1; sy <= ny; ++sy) {
      if (want_sy && sy != want_sy) continue;
      if (7L * sy * sz > kLds || sy * sz > kThreads) break;
      const int top = min(nx, (int)(kLds / (7L * sy * sz)));
      for (int sx = want_sx ? want_sx : 1; sx <= (want_sx ? want_sx : top); ++sx) {
        // even out the cells per workgroup along each axis without adding
        // workgroups: an evened size never exceeds the first one
        int gx = (nx + sx - 1) / sx, gy = (ny + sy - 1) / sy, gz = (nz + sz - 1) / sz;
        const int ex = (nx + gx - 1) / gx, ey = (ny + gy - 1) / gy, ez = (nz + gz - 1) / gz;
        if (lds_floats(ex, ey, ez) > kLds || ey * ez > kThreads) continue;
        gx = (nx + ex - 1) / ex;
        gy = (ny + ey - 1) / ey;
        gz = (nz + ez - 1) / ez;
        const long blocks = (long)gx * gy * gz;
        if (blocks > capacity) continue;
        const long key[3] = {ex, (long)ex * ey + (long)ex * ez + (long)ey * ez, blocks};
        const bool better = !found || key[0] < best_key[0] ||
                            (key[0] == best_key[0] &&
                             (key[1] < best_key[1] ||
                              (key[1] == best_key[1] && key[2] < best_key[2])));
        if (better) {
          found = true;
          for (int q = 0; q < 3; ++q) best_key[q] = key[q];
          best = Layout{ex, ey, ez, gx, gy, gz, (int)blocks};
        }
      }
    }
  }
  return found;
}

int64_t workspace_bytes(int sx, int sy, int sz, int blocks) {
  return (int64_t)(face_words(sx, sy, sz) * blocks * sizeof(Word)) + 64 * (int64_t)sizeof(int);
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_resident_layout_for(const wfd_yee_plan* p, int sx, int sy, int sz,
                            wfd_resident_layout* out) {
  WFD_REQUIRE(p && out, "resident_layout: null argument");
  WFD_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0 && cells(p->nx, p->ny, p->nz) <= kMaxCells,
              "resident_layout: bad grid");
  WFD_REQUIRE(sx >= 0 && sy >= 0 && sz >= 0, "resident_layout: negative sub-box");
  int device = 0, cus = 0, per_cu = 0, coop = 0;
  WFD_HIP(hipGetDevice(&device));
  WFD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  WFD_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
  WFD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&k_resident), kThreads, 0));
  const int capacity = per_cu * cus;
  wfd_resident_layout L = {};
  L.capacity = capacity;
  if (!coop) {
    *out = L;
    set_error("resident_layout: device without cooperative launch");
    return WFD_ERR_UNSUPPORTED;
  }
  Layout c;
  if (!choose(p->nx, p->ny, p->nz, capacity, sx, sy, sz, c)) {
    *out = L;
    set_error("resident_layout: %dx%dx%d does not fit %d co-resident workgroups of %d LDS floats",
              p->nx, p->ny, p->nz, capacity, kLds);
    return WFD_ERR_UNSUPPORTED;
  }
  L.sx = c.sx;
  L.sy = c.sy;
  L.sz = c.sz;
  L.gx = c.gx;
  L.gy = c.gy;
  L.gz = c.gz;
  L.blocks = c.blocks;
  L.workspace_bytes = workspace_bytes(c.sx, c.sy, c.sz, c.blocks);
  *out = L;
  return WFD_OK;
}

}  // extern "C"

namespace wfd {
namespace {

// Validate and launch one resident chunk with its face words, error word and
// epoch at `workspace` (wfd_resident_run: the caller's; the no-allocation
// variant: the library's scratch, zeroed first).  `status` (optional) gets
// the error word after the launch.
int resident_launch(const wfd_yee_plan* p, const wfd_resident_layout* L, void* workspace,
                    bool clear, int nsteps, int spin_limit, int* status, void* stream) {
  WFD_REQUIRE(p && L && workspace, "resident_run: null argument");
  WFD_REQUIRE(nsteps >= 0 && nsteps < (1 << 30), "resident_run: step count out of range");
  WFD_REQUIRE(p->e && p->h && p->inv_eps && p->psi_e && p->psi_h && p->coeff_e && p->coeff_h &&
                  p->counters,
              "resident_run: null field, psi, coeff or counter pointer");
  WFD_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0 && cells(p->nx, p->ny, p->nz) <= kMaxCells,
              "resident_run: bad grid");
  // psi is addressed through 32-bit buffer offsets (6 components of 4 bytes)
  WFD_REQUIRE(cells(p->nx, p->ny, p->nz) * 24 < (int64_t(1) << 31),
              "resident_run: grid too large for 32-bit psi offsets");
  WFD_REQUIRE(p->coeff_rows >= p->nx && p->coeff_rows >= p->ny && p->coeff_rows >= p->nz,
              "resident_run: coeff_rows < max(shape)");
  WFD_REQUIRE(p->source_x < p->nx, "resident_run: source plane outside the grid");
  WFD_REQUIRE(p->source_x < 0 ||
                  (p->src_e && p->src_h && p->wave_e && p->wave_h && p->steps > 0),
              "resident_run: source plane without profiles/waves");
  WFD_REQUIRE(p->nstage >= 0 && p->nstage <= WFD_MAX_STAGES, "resident_run: 0..%d stages",
              WFD_MAX_STAGES);
  WFD_REQUIRE(p->nstage == 0 || p->slots > 0, "resident_run: stage slots must be positive");
  WFD_REQUIRE(L->sx > 0 && L->sy > 0 && L->sz > 0 && lds_floats(L->sx, L->sy, L->sz) <= kLds &&
                  L->sy * L->sz <= kThreads &&
                  L->gx == (p->nx + L->sx - 1) / L->sx &&
                  L->gy == (p->ny + L->sy - 1) / L->sy &&
                  L->gz == (p->nz + L->sz - 1) / L->sz && L->blocks == L->gx * L->gy * L->gz &&
                  L->blocks <= L->capacity,
              "resident_run: layout does not match the plan");
  WFD_REQUIRE(L->workspace_bytes == workspace_bytes(L->sx, L->sy, L->sz, L->blocks),
              "resident_run: workspace size does not match the layout");
  WFD_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 7) == 0,
              "resident_run: workspace must be 8-byte aligned");
  if (nsteps == 0) return WFD_OK;
  ResArgs a;
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
  a.rows = p->coeff_rows;
  a.steps = p->steps;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.n = p->nx * p->ny * p->nz;
  a.source_x = p->source_x;
  a.courant = p->courant;
  a.sx = L->sx;
  a.sy = L->sy;
  a.sz = L->sz;
  a.gx = L->gx;
  a.gy = L->gy;
  a.gz = L->gz;
  const size_t words = face_words(L->sx, L->sy, L->sz) * L->blocks;
  a.face = reinterpret_cast<Word*>(workspace);
  a.error = reinterpret_cast<int*>(a.face + words);
  unsigned* epoch = reinterpret_cast<unsigned*>(a.error + 1);
  a.epoch = epoch;
  a.spin_limit = spin_limit > 0 ? spin_limit : (1 << 22);
  a.nstage = p->nstage;
  a.slots = p->slots > 0 ? p->slots : 1;
  for (int b = 0; b < WFD_MAX_STAGES; ++b) a.stages[b] = StageBox{};
  for (int b = 0; b < p->nstage; ++b)
    WFD_REQUIRE(to_stage_box(p->stages[b], p->nx, p->ny, p->nz, a.stages[b]) == WFD_OK,
                "resident_run: stage %d outside the grid or without a buffer", b);
  hipStream_t s = as_stream(stream);
  if (clear) {
    hipLaunchKernelGGL(k_fill_zero, dim3(256), dim3(256), 0, s,
                       reinterpret_cast<unsigned*>(workspace),
                       (size_t)(L->workspace_bytes / 4));
    WFD_LAUNCH_CHECK("resident_run(clear scratch)");
  }
  void* args[] = {&a, &nsteps};
  WFD_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_resident),
                                     dim3(L->blocks), dim3(kThreads), args, 0, s));
  hipLaunchKernelGGL(k_advance_counters, dim3(1), dim3(1), 0, s, p->counters, epoch, nsteps,
                     a.error, status);
  WFD_LAUNCH_CHECK("resident_run(counters)");
  return WFD_OK;
}

}  // namespace
}  // namespace wfd

extern "C" {

int wfd_resident_run(const wfd_yee_plan* p, const wfd_resident_layout* L, void* workspace,
                     int nsteps, int spin_limit, void* stream) {
  return resident_launch(p, L, workspace, false, nsteps, spin_limit, nullptr, stream);
}

int wfd_resident_run_noalloc(const wfd_yee_plan* p, const wfd_resident_layout* L, int nsteps,
                             int spin_limit, int* status, void* stream) {
  WFD_REQUIRE(p && L && status, "resident_run_noalloc: null argument");
  WFD_REQUIRE(L->workspace_bytes > 0 && (size_t)L->workspace_bytes <= kScratchBytes,
              "resident_run_noalloc: the layout needs %lld bytes of face words, more than the "
              "%zu-byte scratch (use wfd_resident_run with a caller workspace)",
              (long long)L->workspace_bytes, kScratchBytes);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != WFD_OK) return rc;
  return resident_launch(p, L, ws, true, nsteps, spin_limit, status, stream);
}

}  // extern "C"
