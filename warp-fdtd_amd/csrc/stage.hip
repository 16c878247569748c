// Deferred detector recording for gfx950.  The H half-step stages the raw E,
// H and H_old of every detector footprint each step (yee.hip); once per graph
// chunk one launch replays the steps, colocating each plane from the staged
// values with exactly the arithmetic of detectors.py:26-62 and then
// accumulating like detectors.py:99-106, so phasor and flux results equal
// per-step recording bit for bit while the time loop carries no detector
// launches.  FP64 DFT planes sum a chunk's steps (in a fixed order of step
// segments) before adding them: deterministic, and within rounding of a
// per-step running sum.
#include "common.hpp"

namespace wfd {
namespace {

constexpr int kBlock = 256;
constexpr int kFreqGroup = 8;
constexpr int kTile = 64;  // steps of DFT weights staged in LDS at a time
constexpr int kSeg = 8;    // step segments per (component, cell) of a DFT plane
constexpr int kPairs = kBlock / kSeg;

struct Replay {
  const float* stage;
  int fx0, fy0, fz0, fsy, fsz, fcount;  // footprint origin, y/z extents, cells
  int x0, y0, z0, sx, sy, sz, count;    // plane
  int kind, exact, axis, steps, nf;
  const int* indices;
  const float* weights;
  const double* weights64;
  float* phasor;
  double* phasor64;
  float* flux;
  const float* area;
};

struct ReplayArgs {
  Replay r[WFD_MAX_STAGES];
  int begin[WFD_MAX_STAGES + 1];
  int count;
};

__device__ __forceinline__ void plane_xyz(const Replay& r, int i, int& x, int& y, int& z) {
  x = r.x0 + i / (r.sy * r.sz);
  y = r.y0 + (i / r.sz) % r.sy;
  z = r.z0 + i % r.sz;
}

// The eight stencil points of a plane cell's colocation, as footprint
// indices (-1 outside the grid): 0 (x,y,z), 1 (x-1,y,z), 2 (x,y-1,z),
// 3 (x-1,y-1,z), 4-7 the same at z+1.  Computed once per thread, so a loop
// over steps only adds the slot and field offsets.
struct Points {
  int idx[8];
};

__device__ __forceinline__ Points make_points(const Replay& r, int x, int y, int z, int nx,
                                              int ny, int nz) {
  Points k;
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const int px = x - (p & 1), py = y - ((p >> 1) & 1), pz = z + (p >> 2);
    const bool in = px >= 0 && py >= 0 && pz >= 0 && px < nx && py < ny && pz < nz;
    k.idx[p] = in ? ((px - r.fx0) * r.fsy + (py - r.fy0)) * r.fsz + (pz - r.fz0) : -1;
  }
  return k;
}

// The colocation of component c (detectors.py:26-62, the same operations in
// the same order) from a step's slot `base` through the point table.
__device__ __forceinline__ float colocated_at(const Points& k, const float* __restrict__ base,
                                              int fcount, int c, int exact) {
  auto at = [&](int f, int p) { return k.idx[p] >= 0 ? base[f * fcount + k.idx[p]] : 0.f; };
  auto ht = [&](int cc, int p) { return 0.5f * (at(3 + cc, p) + at(6 + cc, p)); };
  if (!exact) return at(c, 0);
  switch (c) {
    case 0:
      return 0.25f * (((at(0, 0) + at(0, 1)) + at(0, 4)) + at(0, 5));
    case 1:
      return 0.25f * (((at(1, 0) + at(1, 2)) + at(1, 4)) + at(1, 6));
    case 2:
      return at(2, 0);
    case 3:
      return 0.5f * (ht(0, 0) + ht(0, 2));
    case 4:
      return 0.5f * (ht(1, 0) + ht(1, 1));
    default:
      return 0.125f * (((((((ht(2, 0) + ht(2, 1)) + ht(2, 2)) + ht(2, 3)) + ht(2, 4)) +
                         ht(2, 5)) +
                        ht(2, 6)) +
                       ht(2, 7));
  }
}

// The staged values one colocated component reads, as byte offsets into a
// step's slot (4 * (field * fcount + footprint index); past the slot's end
// for points outside the grid, which a buffer load reads as 0), in the order
// the colocation consumes them, so a loop over steps only loads and evaluates
// the same expression: form 0 one value, 1 the four-point E average, 2 the
// two-point H average of (H + H_old) / 2, 3 the eight-point one.
struct Terms {
  int off[16];
  int form;
};

constexpr int kOutside = 0x7ffffff0;  // beyond any slot: reads 0

__device__ __forceinline__ Terms make_terms(const Replay& r, int c, int x, int y, int z,
                                            int nx, int ny, int nz) {
  Terms k;
  int n = 0;
  auto add = [&](int f, int px, int py, int pz) {
    const bool in = px >= 0 && py >= 0 && pz >= 0 && px < nx && py < ny && pz < nz;
    k.off[n++] =
        in ? 4 * (f * r.fcount + ((px - r.fx0) * r.fsy + (py - r.fy0)) * r.fsz + (pz - r.fz0))
           : kOutside;
  };
  for (int j = 0; j < 16; ++j) k.off[j] = kOutside;
  if (!r.exact) {
    k.form = 0;
    add(c, x, y, z);
  } else if (c == 0) {
    k.form = 1;
    add(0, x, y, z), add(0, x - 1, y, z), add(0, x, y, z + 1), add(0, x - 1, y, z + 1);
  } else if (c == 1) {
    k.form = 1;
    add(1, x, y, z), add(1, x, y - 1, z), add(1, x, y, z + 1), add(1, x, y - 1, z + 1);
  } else if (c == 2) {
    k.form = 0;
    add(2, x, y, z);
  } else if (c == 3) {
    k.form = 2;
    add(3, x, y, z), add(6, x, y, z), add(3, x, y - 1, z), add(6, x, y - 1, z);
  } else if (c == 4) {
    k.form = 2;
    add(4, x, y, z), add(7, x, y, z), add(4, x - 1, y, z), add(7, x - 1, y, z);
  } else {
    k.form = 3;
    const int px[8] = {x, x - 1, x, x - 1, x, x - 1, x, x - 1};
    const int py[8] = {y, y, y - 1, y - 1, y, y, y - 1, y - 1};
    const int pz[8] = {z, z, z, z, z + 1, z + 1, z + 1, z + 1};
    for (int j = 0; j < 8; ++j) add(5, px[j], py[j], pz[j]), add(8, px[j], py[j], pz[j]);
  }
  return k;
}

// One component's colocation of one step through precomputed terms: the
// same operations in the same order.  `ring` is a buffer over the whole stage
// ring and `so` the step's slot offset in bytes (an outside term stays past
// the ring's end).
__device__ __forceinline__ float eval_terms(const Terms& k, Rsrc ring, unsigned so) {
  auto v = [&](int j) { return ld(ring, (int)((unsigned)k.off[j] + so), 0); };
  switch (k.form) {
    case 0:
      return v(0);
    case 1:
      return 0.25f * (((v(0) + v(1)) + v(2)) + v(3));
    case 2:
      return 0.5f * ((0.5f * (v(0) + v(1))) + (0.5f * (v(2) + v(3))));
    default: {
      float t[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) t[j] = 0.5f * (v(2 * j) + v(2 * j + 1));
      return 0.125f * (((((((t[0] + t[1]) + t[2]) + t[3]) + t[4]) + t[5]) + t[6]) + t[7]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_replay(ReplayArgs args, int slots,
                                                   const int* __restrict__ step, int m,
                                                   int nx, int ny, int nz) {
  const int b = blockIdx.x;
  int k = 0;
  while (k + 1 < args.count && b >= args.begin[k + 1]) ++k;
  const Replay& r = args.r[k];
  const int last = step[0];
  const int local = b - args.begin[k];
  if (r.kind == 1) {  // flux: one block, plane reduced per step in a fixed order
    __shared__ float part[kBlock];
    for (int j = 0; j < m; ++j) {
      const int t = last - m + 1 + j;
      if (t < 0 || t >= r.steps) continue;
      const int index = r.indices[t];
      if (index < 0) continue;
      const float* base = r.stage + (size_t)(t % slots) * 9 * r.fcount;
      float acc = 0.f;
      for (int i = threadIdx.x; i < r.count; i += kBlock) {
        int x, y, z;
        plane_xyz(r, i, x, y, z);
        const Points pts = make_points(r, x, y, z, nx, ny, nz);
        float e[3], h[3];
        for (int c = 0; c < 3; ++c) {
          e[c] = colocated_at(pts, base, r.fcount, c, r.exact);
          h[c] = colocated_at(pts, base, r.fcount, 3 + c, r.exact);
        }
        float cr;
        if (r.axis == 0) cr = e[1] * h[2] - e[2] * h[1];
        else if (r.axis == 1) cr = e[2] * h[0] - e[0] * h[2];
        else cr = e[0] * h[1] - e[1] * h[0];
        acc += cr * r.area[i];
      }
      part[threadIdx.x] = acc;
      __syncthreads();
      for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) r.flux[index] += part[0];
      __syncthreads();
    }
    return;
  }
  if (r.kind == 0) {  // float32 phasor, one thread per plane cell
    const int i = local * kBlock + threadIdx.x;
    if (i >= r.count) return;
    int x, y, z;
    plane_xyz(r, i, x, y, z);
    const Points pts = make_points(r, x, y, z, nx, ny, nz);
    float2 acc[6];
    float2* ph = reinterpret_cast<float2*>(r.phasor);
    for (int c = 0; c < 6; ++c) acc[c] = ph[c * r.count + i];
    for (int j = 0; j < m; ++j) {
      const int t = last - m + 1 + j;
      if (t < 0 || t >= r.steps) continue;
      if (r.indices[t] < 0) continue;
      const float* base = r.stage + (size_t)(t % slots) * 9 * r.fcount;
      const float wr = r.weights[2 * t], wi = r.weights[2 * t + 1];
      for (int c = 0; c < 6; ++c) {
        const float v = colocated_at(pts, base, r.fcount, c, r.exact);
        acc[c].x += wr * v;
        acc[c].y += wi * v;
      }
    }
    for (int c = 0; c < 6; ++c) ph[c * r.count + i] = acc[c];
    return;
  }
  // kind 2: FP64 multi-frequency DFT.  A block holds kPairs (component, cell)
  // pairs times kSeg step segments: thread (seg, pair) sums the chunk's steps
  // t with (t - t_lo) % kSeg == seg, in order, for kFreqGroup frequencies
  // (frequency groups along the block index); the segments' sums are then
  // added in segment order and the total added to the phasor.  The weights of
  // the block's frequencies are staged in LDS one tile of kTile steps at a
  // time (every thread of the block reads the same ones).
  __shared__ double wsh[kTile * kFreqGroup * 2];
  // segment sums of half the frequencies at a time (16 KB instead of 32 KB
  // keeps four blocks resident per CU)
  __shared__ double part[kSeg][kFreqGroup][kPairs];
  const int groups = (r.nf + kFreqGroup - 1) / kFreqGroup;
  const int per_group = (6 * r.count + kPairs - 1) / kPairs;
  const int g = local / per_group;
  const int seg = threadIdx.x / kPairs, lane = threadIdx.x % kPairs;
  const int ci = (local - g * per_group) * kPairs + lane;
  const bool valid = g < groups && ci < 6 * r.count;  // block-uniform g; ci per thread
  const int c = valid ? ci / r.count : 0, i = valid ? ci - c * r.count : 0;
  const int f0 = g * kFreqGroup, fn = min(kFreqGroup, r.nf - f0);
  int x, y, z;
  plane_xyz(r, i, x, y, z);
  const Terms terms = make_terms(r, c, x, y, z, nx, ny, nz);
  const unsigned slot_bytes = 36u * (unsigned)r.fcount;  // 9 fields of a step
  const Rsrc ring = buffer(r.stage, slot_bytes * (unsigned)slots);
  double re[kFreqGroup], im[kFreqGroup];
#pragma unroll
  for (int q = 0; q < kFreqGroup; ++q) re[q] = im[q] = 0.0;
  // steps t in [t_lo, t_hi]: the chunk's steps inside the recording window
  const int t_lo = max(last - m + 1, 0), t_hi = min(last, r.steps - 1);
  for (int t0 = t_lo; t0 <= t_hi; t0 += kTile) {
    const int len = min(kTile, t_hi - t0 + 1);
    const int s0 = t0 % slots;  // len <= m <= slots: a tile wraps the ring at most once
    __syncthreads();
    for (int e = threadIdx.x; e < len * kFreqGroup * 2; e += kBlock) {
      const int tt = e / (kFreqGroup * 2), q = (e / 2) % kFreqGroup, half = e % 2;
      wsh[e] = q < fn ? r.weights64[2 * ((size_t)(t0 + tt) * r.nf + f0 + q) + half] : 0.0;
    }
    __syncthreads();
    if (!valid) continue;
    // unrolled by 2 so the colocation loads of two steps are in flight at once
#pragma unroll 2
    for (int tt = seg; tt < len; tt += kSeg) {
      const int slot = s0 + tt < slots ? s0 + tt : s0 + tt - slots;  // (t0 + tt) % slots
      const double v = (double)eval_terms(terms, ring, (unsigned)slot * slot_bytes);
      const double* w = wsh + tt * kFreqGroup * 2;
#pragma unroll
      for (int q = 0; q < kFreqGroup; ++q) {
        re[q] += w[2 * q] * v;
        im[q] += w[2 * q + 1] * v;
      }
    }
  }
  constexpr int kHalf = kFreqGroup / 2;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    __syncthreads();  // the previous half's sums have been read
#pragma unroll
    for (int q = 0; q < kHalf; ++q) {
      part[seg][2 * q][lane] = re[h * kHalf + q];
      part[seg][2 * q + 1][lane] = im[h * kHalf + q];
    }
    __syncthreads();
    if (valid && seg == 0) {
#pragma unroll
      for (int q = 0; q < kHalf; ++q) {
        const int f = h * kHalf + q;
        if (f < fn) {
          double sr = part[0][2 * q][lane], si = part[0][2 * q + 1][lane];
#pragma unroll
          for (int k2 = 1; k2 < kSeg; ++k2) {
            sr += part[k2][2 * q][lane];
            si += part[k2][2 * q + 1][lane];
          }
          double2* p =
              reinterpret_cast<double2*>(r.phasor64) + ((size_t)(f0 + f) * 6 * r.count + ci);
          double2 cur = *p;
          cur.x += sr;
          cur.y += si;
          *p = cur;
        }
      }
    }
  }
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" {

int wfd_replay_monitors(const wfd_replay* monitors, int count, int slots, const int* step,
                        int m, int nx, int ny, int nz, void* stream) {
  WFD_REQUIRE(monitors && step && count >= 1 && count <= WFD_MAX_STAGES,
              "replay: 1..%d detectors per launch", WFD_MAX_STAGES);
  WFD_REQUIRE(slots > 0 && m >= 0 && m <= slots, "replay: need 0 <= m <= slots");
  WFD_REQUIRE(nx > 0 && ny > 0 && nz > 0 && cells(nx, ny, nz) <= kMaxCells, "replay: bad grid");
  if (m == 0) return WFD_OK;
  ReplayArgs args;
  args.count = count;
  int blocks = 0;
  for (int k = 0; k < count; ++k) {
    const wfd_replay& d = monitors[k];
    Replay& r = args.r[k];
    const int* f = d.box;
    WFD_REQUIRE(d.stage && f[0] >= 0 && f[2] >= 0 && f[4] >= 0 && f[1] <= nx && f[3] <= ny &&
                    f[5] <= nz && f[1] > f[0] && f[3] > f[2] && f[5] > f[4],
                "replay: detector %d footprint outside the grid", k);
    WFD_REQUIRE(d.sx > 0 && d.sy > 0 && d.sz > 0 && d.steps > 0, "replay: detector %d plane", k);
    // every in-grid cell the colocation stencil reads must lie in the footprint
    const int lox = d.exact ? d.x0 - 1 : d.x0, loy = d.exact ? d.y0 - 1 : d.y0;
    const int hiz = d.exact ? d.z0 + d.sz + 1 : d.z0 + d.sz;
    WFD_REQUIRE(f[0] <= (lox < 0 ? 0 : lox) && f[2] <= (loy < 0 ? 0 : loy) && f[4] <= d.z0 &&
                    f[1] >= d.x0 + d.sx && f[3] >= d.y0 + d.sy && f[5] >= (hiz > nz ? nz : hiz),
                "replay: detector %d footprint does not cover its stencil", k);
    WFD_REQUIRE(d.kind >= 0 && d.kind <= 2, "replay: detector %d kind", k);
    WFD_REQUIRE((int64_t)36 * (f[1] - f[0]) * (f[3] - f[2]) * (f[5] - f[4]) * slots <
                    (int64_t)kOutside - 16,
                "replay: detector %d stage ring exceeds 32-bit offsets", k);
    WFD_REQUIRE(d.kind == 2 || d.indices, "replay: detector %d without indices", k);
    WFD_REQUIRE(d.kind != 0 || (d.weights && d.phasor), "replay: phasor detector %d arrays", k);
    WFD_REQUIRE(d.kind != 1 || (d.flux && d.area && d.axis >= 0 && d.axis < 3),
                "replay: flux detector %d arrays", k);
    WFD_REQUIRE(d.kind != 2 || (d.weights64 && d.phasor64 && d.nf > 0),
                "replay: dft detector %d arrays", k);
    r.stage = d.stage;
    r.fx0 = f[0];
    r.fy0 = f[2];
    r.fz0 = f[4];
    r.fsy = f[3] - f[2];
    r.fsz = f[5] - f[4];
    r.fcount = (f[1] - f[0]) * r.fsy * r.fsz;
    r.x0 = d.x0;
    r.y0 = d.y0;
    r.z0 = d.z0;
    r.sx = d.sx;
    r.sy = d.sy;
    r.sz = d.sz;
    r.count = d.sx * d.sy * d.sz;
    r.kind = d.kind;
    r.exact = d.exact;
    r.axis = d.axis;
    r.steps = d.steps;
    r.nf = d.nf;
    r.indices = d.indices;
    r.weights = d.weights;
    r.weights64 = d.weights64;
    r.phasor = d.phasor;
    r.phasor64 = d.phasor64;
    r.flux = d.flux;
    r.area = d.area;
    args.begin[k] = blocks;
    if (d.kind == 1) blocks += 1;
    else if (d.kind == 0) blocks += (r.count + kBlock - 1) / kBlock;
    else blocks += ((6 * r.count + kPairs - 1) / kPairs) * ((d.nf + kFreqGroup - 1) / kFreqGroup);
  }
  args.begin[count] = blocks;
  hipLaunchKernelGGL(k_replay, dim3(blocks), dim3(kBlock), 0, as_stream(stream), args, slots,
                     step, m, nx, ny, nz);
  WFD_LAUNCH_CHECK("replay");
  return WFD_OK;
}

}  // extern "C"
