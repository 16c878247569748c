/*
 * libwfdtd — C ABI of the MI355X (gfx950) uniform-grid FDTD path.
 *
 * Every entry point takes plain device pointers (torch tensors' data_ptr())
 * and sizes, enqueues on the given HIP stream (void* = hipStream_t; NULL = the
 * null stream) and returns an int status: WFD_OK (0) or a negative WFD_ERR_*
 * code, with a description in wfd_last_error().  Nothing throws across the ABI
 * and no entry point allocates or frees caller buffers.
 *
 * Field storage follows the reference: SoA (3, n) float32 with flat index
 * i = (x*ny + y)*nz + z (z fastest), warp_fdtd/kernels.py:25,100-103.
 * Citations are to files of the reference (alecjacobson/warp-fdtd).
 */
#ifndef WFDTD_H
#define WFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFD_OK 0
#define WFD_ERR_ARGUMENT (-1)    /* bad pointer, size or shape -> ValueError */
#define WFD_ERR_HIP (-2)         /* HIP runtime error -> RuntimeError */
#define WFD_ERR_UNSUPPORTED (-3) /* valid but unsupported configuration -> ValueError */

#define WFD_MAX_MONITORS 16
#define WFD_MAX_STAGES 8
#define WFD_MAX_BATCH 64

/* ---------------------------------------------------------------- runtime */
int wfd_version(void);
/* Thread-local text of the last failure (empty when none). */
const char* wfd_last_error(void);
int wfd_device_synchronize(void);
int wfd_stream_synchronize(void* stream);
/* Stream capture into a HIP graph (replaces wp.ScopedCapture / wp.capture_launch,
 * warp_fdtd/solver.py:239-258).  begin/end bracket launches on `stream`. */
int wfd_stream_create(void** stream);
int wfd_stream_destroy(void* stream);
int wfd_graph_begin(void* stream);
int wfd_graph_end(void* stream, void** graph_exec);
int wfd_graph_launch(void* graph_exec, void* stream);
int wfd_graph_destroy(void* graph_exec);

/* ------------------------------------------------ Yee half-steps (kernel ABI)
 * warp_fdtd/kernels.py:77-148, make_update(electric, packed): one thread per
 * cell over n = nx*ny*nz.  Arguments mirror the Warp kernel's
 * (f, g, inv_eps, psi, coeff, source, waves, step, hprev, save_h_from,
 *  nx, ny, nz, source_x, courant, periodic), with the row strides of the 2-D
 * arrays made explicit.  coeff is (3, coeff_rows, 3) rows (a, b, 1/kappa) for
 * the full psi layout (6, psi_stride) and (3, coeff_rows, 4) rows
 * (a, b, 1/kappa, plane rank) for the packed layout (solver.py:16-29).
 * source/waves/hprev may be dummies when source_x < 0 / save_h_from > steps. */
int wfd_update_e(float* f, const float* g, const float* inv_eps, float* psi,
                 int64_t psi_stride, const float* coeff, int coeff_rows,
                 const float* source, int64_t source_stride, const float* waves,
                 int64_t waves_stride, const int* step, float* hprev, int save_h_from,
                 int nx, int ny, int nz, int source_x, float courant, int periodic,
                 void* stream);
int wfd_update_h(float* f, const float* g, const float* inv_eps, float* psi,
                 int64_t psi_stride, const float* coeff, int coeff_rows,
                 const float* source, int64_t source_stride, const float* waves,
                 int64_t waves_stride, const int* step, float* hprev, int save_h_from,
                 int nx, int ny, int nz, int source_x, float courant, int periodic,
                 void* stream);
int wfd_update_e_packed(float* f, const float* g, const float* inv_eps, float* psi,
                        int64_t psi_stride, const float* coeff, int coeff_rows,
                        const float* source, int64_t source_stride, const float* waves,
                        int64_t waves_stride, const int* step, float* hprev,
                        int save_h_from, int nx, int ny, int nz, int source_x,
                        float courant, int periodic, void* stream);
int wfd_update_h_packed(float* f, const float* g, const float* inv_eps, float* psi,
                        int64_t psi_stride, const float* coeff, int coeff_rows,
                        const float* source, int64_t source_stride, const float* waves,
                        int64_t waves_stride, const int* step, float* hprev,
                        int save_h_from, int nx, int ny, int nz, int source_x,
                        float courant, int periodic, void* stream);
/* warp_fdtd/kernels.py:151-153 (delta=+1) and adjoint.py:226-228 (delta=-1). */
int wfd_advance(int* step, int delta, void* stream);

/* ------------------------------------------------------ fused Yee time step
 * One full step (E half-step, then H half-step) of warp_fdtd/solver.py:194-234
 * with CPML, plane-source injection and hprev saving, as two column-marching
 * kernels.  Replaces the pair of wp.launch(update_e/update_h) calls.  The step
 * counter protocol needs no separate advance launch: counters[0] holds the
 * number of completed steps (the reference's `step`), counters[1] the step the
 * H half-step is applying (read by detectors recorded after the step). */
/* A detector footprint the step stages every step: the raw E, H (after the
 * step) and H_old (before it) of every cell of the half-open box, written to
 * buf[(t % slots), 9, count] with component order Ex Ey Ez Hx Hy Hz Hox Hoy
 * Hoz and cells in (x, y, z) order, z fastest (count = box volume). */
typedef struct wfd_stage {
  float* buf;
  int box[6]; /* x0, x1, y0, y1, z0, z1 */
} wfd_stage;

typedef struct wfd_yee_plan {
  float* e;               /* (3, n) */
  float* h;               /* (3, n) */
  const float* inv_eps;   /* (n) */
  float* psi_e;           /* (6, n) full-volume CPML state, touched in slabs only */
  float* psi_h;           /* (6, n) */
  const float* coeff_e;   /* (3, coeff_rows, 3) rows (a, b, 1/kappa) */
  const float* coeff_h;   /* (3, coeff_rows, 3) */
  const float* src_e;     /* (3, ny*nz) E profile (drives the H half-step) */
  const float* src_h;     /* (3, ny*nz) H profile (drives the E half-step) */
  const float* wave_e;    /* (3, steps) */
  const float* wave_h;    /* (3, steps) */
  int* counters;          /* int32[2] */
  float* hprev;           /* (3, n); written inside hbox once t >= save_h_from */
  int hbox[6];            /* x0, x1, y0, y1, z0, z1 (half-open) */
  int save_h_from;
  int coeff_rows;
  int steps;              /* length of the wave tables */
  int nx, ny, nz;
  int source_x;           /* < 0: no source */
  float courant;
  int x_chunk;            /* planes per thread column; 0 picks a default */
  int nstage;             /* footprints staged by the H half-step */
  int slots;              /* ring length of every stage buffer */
  wfd_stage stages[WFD_MAX_STAGES];
} wfd_yee_plan;

int wfd_yee_step(const wfd_yee_plan* plan, void* stream);

/* The same step (bitwise) as one launch that reads the plan's e, h, psi_e
 * and psi_h (the state after counters[0] steps) and writes the next state to
 * the targets (buffers of the same shapes, distinct from the plan's); the
 * caller alternates the two sets.  With steps = 2 the launch applies two
 * steps (the first step's fields stay on chip; counters advance by 2, both
 * steps are staged, and hprev gets H at the start of the second step).  Each workgroup recomputes the new E of the
 * row past its tile and of the plane past its chunk instead of waiting for
 * the workgroups that own them, so every cell's E, H and 1/eps are read once
 * and written once per step (the fused pair of wfd_yee_step reads E and H
 * twice).  Same step protocol as wfd_yee_step (counters, staging, hprev,
 * source); x_chunk is not used.  Only psi components on a CPML axis of a
 * cell are written to the targets (the others are neither read nor written).
 * The ticket is a device int that is zero before the first call and left
 * zero by every call; calls sharing a ticket must not overlap.  No reference
 * counterpart (replaces the same pair of launches as wfd_yee_step). */
typedef struct wfd_yee_targets {
  float* e;
  float* h;
  float* psi_e;
  float* psi_h;
  int* ticket;
  int planes;           /* x planes per workgroup (0: chosen) */
  int slots_per_thread; /* one step: 1 (512 threads) or 2 or 4 (256 threads;
                           0: 4), a workgroup covering threads x this - nz
                           (y, z) cells; two steps: 1 or 2 (0: 2), 512
                           threads covering 512 x this - 3 nz cells */
  int steps;            /* steps per launch: 1 or 2 (0: 1) */
  float* hist;          /* two steps only, optional: the adjoint curl history
                           of both steps, written as wfd_record_curl would
                           (slot step % hist_steps, box below) */
  int hist_steps;
  int hist_box[6];      /* x0, y0, z0, sx, sy, sz */
  int threads;          /* workgroup size: 0 (512 for one slot per thread or
                           two steps, else 256) or 1024 with one slot per
                           thread (a tile of 1024 - nz or 1024 - 3 nz cells) */
} wfd_yee_targets;

int wfd_yee_pass(const wfd_yee_plan* plan, const wfd_yee_targets* out, void* stream);

/* ------------------------------------------------------------ batched grids
 * Several independent grids (one wfd_yee_plan each, any shapes) stepped by
 * one launch per half-step and x_chunk class: the launch concatenates every
 * plan's block range and each block finds its plan in a table held in device
 * memory.  Each plan's result is bitwise that of its own wfd_yee_step.  No
 * reference counterpart (the reference steps one grid per Solver).
 * prepare writes the table for `plans` into the workspace (workspace_bytes
 * from wfd_batch_workspace_bytes, -1 for a bad count) and synchronizes the
 * stream: call it outside graph capture, again whenever a plan changes.
 * step must receive the same plans (pointers, shapes, x_chunk, stages) as the
 * prepare that wrote the workspace. */
int64_t wfd_batch_workspace_bytes(int count);
int wfd_batch_prepare(const wfd_yee_plan* plans, int count, void* workspace,
                      int64_t workspace_bytes, void* stream);
int wfd_batch_step(const wfd_yee_plan* plans, int count, const void* workspace, void* stream);

/* Two-step passes (wfd_yee_pass, steps = 2, 1024 threads x 1 slot) of
 * several grids in one launch: plans[d] is pass d's start and targets[d] its
 * result (the buffer set it writes, its own ticket), as for wfd_yee_pass;
 * planes (equal in every target; 0: chosen from the occupancy for the whole
 * launch) is common to all.  Each grid's result is bitwise that of its own
 * wfd_yee_pass.  prepare writes the argument table into the workspace
 * (workspace_bytes from wfd_batch_pass_workspace_bytes, -1 for a bad count)
 * and synchronizes the stream: call it outside graph capture; the launch
 * must receive the same plans and targets.  No reference counterpart. */
int64_t wfd_batch_pass_workspace_bytes(int count);
int wfd_batch_pass_prepare(const wfd_yee_plan* plans, const wfd_yee_targets* targets, int count,
                           void* workspace, int64_t workspace_bytes, void* stream);
int wfd_batch_pass(const wfd_yee_plan* plans, const wfd_yee_targets* targets, int count,
                   const void* workspace, void* stream);

/* ------------------------------------------------- on-chip-resident stepping
 * The same step as wfd_yee_step (bitwise), for grids whose E, H and 1/eps fit
 * in the LDS of co-resident workgroups: sub-boxes of sx x sy x sz cells (one
 * 1024-thread workgroup each, up to 160 KB of LDS) stay on chip for nsteps
 * steps; faces move between workgroups through the workspace as 64-bit words
 * (value + step tag) that readers poll.  layout_for picks the sub-box when
 * sx/sy/sz are 0.  One cooperative launch per call (replaces nsteps pairs of
 * half-step launches of solver.py:203-234); counters advance by nsteps.
 * layout_for validates the decomposition (WFD_ERR_UNSUPPORTED when the
 * workgroups cannot all be co-resident) and sizes the workspace (8-byte
 * aligned, allocated by the caller).  The caller zeroes the workspace before
 * its first launch and whenever the layout used with it changes; launches
 * of one layout need no clearing between them (the face words carry tags
 * derived from an epoch word that each launch advances).  The int after
 * the face words is an error word, nonzero after a bounded poll gave up
 * (the fields of that run are then invalid), which the caller reads and
 * clears (zeroing the whole workspace); the int after it is the epoch. */
typedef struct wfd_resident_layout {
  int sx, sy, sz;        /* sub-box cells along x, y, z */
  int gx, gy, gz;        /* workgroups along each axis */
  int blocks, capacity;  /* workgroups used / co-resident maximum */
  int64_t workspace_bytes;
} wfd_resident_layout;

int wfd_resident_layout_for(const wfd_yee_plan* plan, int sx, int sy, int sz,
                            wfd_resident_layout* layout);
int wfd_resident_run(const wfd_yee_plan* plan, const wfd_resident_layout* layout,
                     void* workspace, int nsteps, int spin_limit, void* stream);

/* wfd_resident_run without a workspace argument and without any allocation
 * (no reference counterpart: for callers that cannot allocate device memory,
 * such as code under HIP graph capture or a serving loop that forbids
 * hipMalloc).  The face words, error word and epoch live in a 16 MiB
 * device array the library owns, one per GPU, held until the process ends;
 * layouts whose workspace_bytes exceed it are rejected (WFD_ERR_ARGUMENT).
 * Each call clears the part it uses before the launch and, after it,
 * writes the error word to *status (a device int of the caller: 0, or
 * nonzero when a bounded poll gave up and the fields of that run are
 * invalid); with nsteps == 0 nothing is launched and *status is left as
 * is.  Every call on one GPU shares that array: calls that overlap on one
 * GPU (concurrent streams, other threads) race. */
int wfd_resident_run_noalloc(const wfd_yee_plan* plan, const wfd_resident_layout* layout,
                             int nsteps, int spin_limit, int* status, void* stream);

/* ------------------------------------------------------------- detectors
 * warp_fdtd/detectors.py:65-106 record(): one monitor, one launch over its
 * plane.  weights (steps, 2) float32 complex; phasor (6, count, 2) float32;
 * flux (max index + 1); area (count); indices/weights of length steps.
 * Reads t = step[0]; skips when t is outside [0, steps) or indices[t] < 0.
 * Flux sums are reduced per plane in a fixed order. */
int wfd_record(const float* e, const float* h, const float* prev, const int* step,
               const float* weights, const int* indices, float* phasor, float* flux,
               const float* area, int nx, int ny, int nz, int x0, int y0, int z0, int sx,
               int sy, int sz, int axis, int exact, int is_phasor, int steps,
               void* stream);

/* All monitors of a solver in one launch (solver.py:235-236 loop). */
typedef struct wfd_monitor {
  const int* indices;
  const float* weights;
  float* phasor;
  float* flux;
  const float* area;
  int x0, y0, z0, sx, sy, sz;
  int axis, exact, is_phasor;
  int steps; /* length of indices/weights; steps past it record nothing */
} wfd_monitor;

int wfd_record_monitors(const float* e, const float* h, const float* prev,
                        const int* step, const wfd_monitor* monitors, int count, int nx,
                        int ny, int nz, void* stream);

/* Deferred detector recording from staged footprints (wfd_stage): for the m
 * steps t = last-m+1 .. last (last = step[0]) in order, every detector
 * colocates its plane from the staged raw fields exactly as detectors.py:
 * 26-62 does from the live arrays, then applies detectors.py:99-106 (kind 0:
 * float32 phasor += w_t * v where indices[t] >= 0; kind 1: flux[indices[t]] +=
 * sum(cross * area), reduced in a fixed order) or the FP64 multi-frequency
 * DFT (kind 2: phasor (nf, 6, count, 2) += w[t, f] * v).  Bitwise the same
 * values as recording every step. */
typedef struct wfd_replay {
  const float* stage;     /* the footprint's stage buffer */
  int box[6];             /* its footprint box */
  int x0, y0, z0, sx, sy, sz;  /* detector plane */
  int kind;               /* 0 phasor, 1 flux, 2 dft */
  int exact, axis;
  int steps;              /* length of indices / weights */
  int nf;                 /* kind 2: number of frequencies */
  const int* indices;     /* kinds 0, 1 */
  const float* weights;   /* kind 0: (steps, 2) */
  const double* weights64;/* kind 2: (steps, nf, 2) */
  float* phasor;          /* kind 0: (6, count, 2) */
  double* phasor64;       /* kind 2: (nf, 6, count, 2) */
  float* flux;            /* kind 1 */
  const float* area;      /* kind 1: (count) */
} wfd_replay;

int wfd_replay_monitors(const wfd_replay* monitors, int count, int slots, const int* step,
                        int m, int nx, int ny, int nz, void* stream);

/* Multi-frequency running DFT (no reference counterpart; the single-frequency
 * case is detectors.py:65-106 with a per-frequency weight table).  stage writes
 * the colocated E,H of plane (x0..,sx..) for step t = step[0] into slot
 * t % slots of stage (slots, 6, count); accumulate adds, for the m steps
 * t = last-m+1 .. last with last = step[0] (the latest staged step),
 * w[t, f] * stage[t % slots] into the FP64 accumulators phasor
 * (nf, 6, count, 2); weights (steps, nf, 2) float64; steps outside
 * [0, steps) add nothing. */
int wfd_dft_stage(const float* e, const float* h, const float* prev, const int* step,
                  float* stage, int slots, int nx, int ny, int nz, int x0, int y0, int z0,
                  int sx, int sy, int sz, int exact, void* stream);
int wfd_dft_accumulate(const float* stage, int slots, const double* weights, int steps,
                       int nf, const int* step, int m, double* phasor, int count,
                       void* stream);

/* warp_fdtd/volume.py:238-302 record_volume(): trilinear, time-centred
 * phasors at count physical points (positions (count, 3) in cell units). */
int wfd_record_volume(const float* e, const float* h, const float* prev,
                      const float* positions, const float* weights, const int* step,
                      float* phasor, int count, int nx, int ny, int nz, int first,
                      void* stream);

/* ---------------------------------------------------------------- adjoint
 * warp_fdtd/adjoint.py:17-43 record_curl: hist[t % hist_steps, c, j] =
 * courant * curl_backward(H) in the box (x0,y0,z0,sx,sy,sz), t = step[0].
 * hist_steps = steps restates the reference's full history. */
int wfd_record_curl(const float* h, float* hist, int hist_steps, const int* step, int nx,
                    int ny, int nz, int x0, int y0, int z0, int sx, int sy, int sz,
                    float courant, void* stream);
/* adjoint.py:46-88 prepare (transposed CPML); d is (6, n). */
int wfd_adjoint_prepare(const float* q, const float* eps, float* psi, const float* coeff,
                        int coeff_rows, float* d, int nx, int ny, int nz, float courant,
                        int electric, void* stream);
/* adjoint.py:107-162 gather (transposed differences; grad += hist . q in the
 * box when from_h).  _f64 accumulates the gradient in double precision. */
int wfd_adjoint_gather(float* q, const float* d, const float* hist, int hist_steps,
                       float* grad, const int* step, int nx, int ny, int nz, int x0,
                       int y0, int z0, int sx, int sy, int sz, int from_h, void* stream);
int wfd_adjoint_gather_f64(float* q, const float* d, const float* hist, int hist_steps,
                           double* grad, const int* step, int nx, int ny, int nz, int x0,
                           int y0, int z0, int sx, int sy, int sz, int from_h,
                           void* stream);
/* adjoint.py:46-162 prepare followed by gather in one pass, without d:
 * from_h = 1 is prepare(q_src = qh, electric = 0) then gather into
 * q_dst = qe (with the gradient); from_h = 0 is prepare(q_src = qe,
 * electric = 1) then gather into q_dst = qh.  Reads psi_in, writes the
 * updated psi to psi_out (distinct buffers) in the slots of axes with CPML
 * at each cell; the other slots are never written, so callers alternating
 * psi_in / psi_out must start both buffers equal there.  Results equal
 * prepare + gather bit for bit. */
int wfd_adjoint_reverse_half(float* q_dst, const float* q_src, const float* eps,
                             const float* psi_in, float* psi_out, const float* coeff,
                             int coeff_rows, const float* hist, int hist_steps, float* grad,
                             const int* step, int nx, int ny, int nz, int x0, int y0, int z0,
                             int sx, int sy, int sz, float courant, int from_h, void* stream);
int wfd_adjoint_reverse_half_f64(float* q_dst, const float* q_src, const float* eps,
                                 const float* psi_in, float* psi_out, const float* coeff,
                                 int coeff_rows, const float* hist, int hist_steps, double* grad,
                                 const int* step, int nx, int ny, int nz, int x0, int y0, int z0,
                                 int sx, int sy, int sz, float courant, int from_h,
                                 void* stream);
/* adjoint.py:181-223 inject (the transposed detector colocation; the
 * reference scatters with atomics, here each target cell gathers its terms
 * in a fixed order); source (6, count, 2) float32 complex, weights
 * (steps, 2). */
int wfd_adjoint_inject(float* qe, float* qh, const float* source, const float* weights,
                       const int* step, int nx, int ny, int nz, int x0, int y0, int z0,
                       int sy, int sz, int count, int old_h_only, int steps,
                       void* stream);
/* The inject of several planes (adjoint.py:181-223, the loop over the
 * objective's monitors) in one launch, count <= WFD_MAX_INJECT.  A cell in
 * several planes' stencil footprints is updated by one thread, which adds
 * the planes in order, so the results equal one wfd_adjoint_inject per
 * plane, in order, bit for bit. */
#define WFD_MAX_INJECT 8
typedef struct wfd_inject_source {
  const float* source;   /* (6, count, 2) float32 complex */
  const float* weights;  /* (steps, 2) */
  int x0, y0, z0, sy, sz, count, steps;
} wfd_inject_source;

int wfd_adjoint_inject_many(const wfd_inject_source* sources, int count, float* qe,
                            float* qh, const int* step, int nx, int ny, int nz, int old_h_only,
                            void* stream);
/* The launches between two reverse steps of adjoint.py:388-414 in one: the
 * H_old terms of step t = step[0] (inject(1)), the retreat, and every term
 * of step t - 1 (the next step's inject(0)); per cell the same additions in
 * the same order, so the result equals wfd_adjoint_inject_many(old_h_only =
 * 1), wfd_advance(-1), wfd_adjoint_inject_many(old_h_only = 0) bit for bit.
 * ticket: a device int, zero before the first call and left zero by each
 * call (calls sharing it must not overlap). */
int wfd_adjoint_inject_turn(const wfd_inject_source* sources, int count, float* qe, float* qh,
                            int* step, int* ticket, int nx, int ny, int nz, void* stream);
/* adjoint.py:472-490 write_design: plane (sx, sy) into every z-layer of the box. */
int wfd_write_design(const float* plane, float* eps, int nx, int ny, int nz, int x0,
                     int y0, int z0, int sx, int sy, int sz, void* stream);

/* ------------------------------------------------------ DrivenUniform path
 * warp_fdtd/driven_uniform.py:90-274: staggered arrays (e: sum over c of the
 * E component grids, h likewise), conductivity sponge, Gaussian current on
 * E_y, PEC exterior.  layout_xyz selects the 3-D (nx+1, ny+1, nz+1) launch of
 * magnetic_xyz/electric_xyz; otherwise the flat launch of magnetic/electric. */
typedef struct wfd_guide {
  int nx, ny, nz;
  float spacing, dt;
  float length, width, height;      /* problem.lengths */
  float core_width, core_height;
  float core_epsilon, clad_epsilon;
  float sponge_width, sponge_rate;
  float source_x, source_y, source_z;             /* source centre */
  float source_sigma_x, source_sigma_y, source_sigma_z;
  int normal_average;
  int layout_xyz;
} wfd_guide;

int wfd_driven_magnetic(const wfd_guide* g, const float* e, float* h, void* stream);
int wfd_driven_electric(const wfd_guide* g, const float* h, float* e, const float* wave,
                        const int* counter, void* stream);

/* warp_fdtd/guide_observer.py:9-25 gather: out[row] = sum_j w[row,j] * v_j with
 * v = field (center=0) or 0.5*(field+old) (center=1); indices/weights (rows, 8). */
int wfd_observer_gather(const int* indices, const float* weights, const float* field,
                        const float* old, int center, float* out, int rows,
                        void* stream);
/* warp_fdtd/guide_validation.py:6-20 count_nonfinite: count += #non-finite. */
int wfd_count_nonfinite(const float* values, int64_t n, int* count, void* stream);
/* warp_fdtd/periodic.py:95-100 record_mode: signal[step] = e[0,0]; step += 1. */
int wfd_record_mode(const float* e, float* signal, int* step, void* stream);

/* --------------------------------------------------------------- material
 * warp_fdtd/material.py:12-55: 19x19 clamped Gaussian filter (weights (19,19))
 * then the tanh subpixel projection to inv_eps = 1/(1 + 11.25 value); all
 * (sx, sy) float32.  pullback is the hand-derived reverse mode of both maps
 * (material.py:93-102 gets it from wp.Tape): grad_p = S^T P^T grad_inv_eps;
 * scratch holds 4*sx*sy floats. */
int wfd_material_forward(const float* p, const float* weights, float* filtered,
                         float* inv_eps, int sx, int sy, float beta, void* stream);
int wfd_material_pullback(const float* weights, const float* filtered,
                          const float* grad_inv_eps, float* grad_p, float* scratch,
                          int sx, int sy, float beta, void* stream);

/* ------------------------------------------------ linear density design map
 * eps = eps_lo + rho * (eps_hi - eps_lo) through the sz z-layers of the design
 * box (x0, y0, z0, sx, sy, sz); rho is (sx, sy) float64 with values in [0, 1].
 * No reference counterpart: the reference's design map is material.py; this
 * is the linear map of the density-parameterised bend (eps 2.085 .. 12.08).
 * forward writes inv_eps = (float)(1 / eps) (float64 arithmetic, one rounding)
 * into the box of the (nx, ny, nz) grid; pullback maps d/d(inv_eps) of the box
 * cells, (sx, sy, sz) float64 as the adjoint accumulates it, to d/d(rho):
 * grad_rho = sum_z grad_inv_eps * (-(eps_hi - eps_lo) / eps^2). */
int wfd_density_forward(const double* rho, float* inv_eps, int nx, int ny, int nz, int x0,
                        int y0, int z0, int sx, int sy, int sz, double eps_lo, double eps_hi,
                        void* stream);
int wfd_density_pullback(const double* rho, const double* grad_inv_eps, double* grad_rho, int sx,
                         int sy, int sz, double eps_lo, double eps_hi, void* stream);

/* ------------------------------------------- Fourier-neural-operator kernels
 * No reference counterpart (the surrogate of the density-design loop).
 * gemm_bf16: C[b] = alpha * A[b] (m x k) * B[b] (k x n) + beta * C[b] for
 * b < batch, element (b, i, j) of each operand at base + b*s?b + i*s?(row) +
 * j*s?(col) (strides in elements); A and B are FP32 (a_bf16/b_bf16 = 0) or
 * BF16 bit patterns (1), rounded to BF16 and multiplied on the matrix cores
 * with FP32 accumulation; C is FP32.  Transposes are stride swaps. */
int wfd_gemm_bf16(const void* a, int a_bf16, long long sab, long long sam, long long sak,
                  const void* b, int b_bf16, long long sbb, long long sbk, long long sbn,
                  float* c, long long scb, long long scm, long long scn, int batch, int m, int n,
                  int k, float alpha, float beta, void* stream);
/* y = BF16 bits of x (round to nearest even); round_bf16 keeps FP32 storage. */
int wfd_to_bf16(const float* x, void* y, long long n, void* stream);
int wfd_round_bf16(const float* x, float* y, long long n, void* stream);
/* pre = y + bias[j] over rows of n (bias may be NULL, pre may be NULL);
 * out = pre (act 0) or GELU(pre) (act 1, erf form).  act_back: dz = dout * act'(pre). */
int wfd_bias_act(const float* y, const float* bias, float* pre, float* out, long long rows,
                 int n, int act, void* stream);
int wfd_act_back(const float* dout, const float* pre, float* dz, long long total, int act,
                 void* stream);
/* out[j] = sum_i x[i, j] (+ beta * out[j]) in a fixed order; scratch holds
 * wfd_colsum_scratch(rows, n) floats. */
long long wfd_colsum_scratch(long long rows, int n);
int wfd_colsum(const float* x, long long rows, int n, float* out, float* scratch,
               long long scratch_floats, float beta, void* stream);
/* y[b, j] = mean_i x[b, i, j] (runs of 512 pixels summed in parallel, then
 * in order; scratch holds wfd_pool_mean_scratch(batch, p, w) floats) and its
 * transpose (dx = dy / p + beta * dx). */
long long wfd_pool_mean_scratch(int batch, int p, int w);
int wfd_pool_mean(const float* x, int batch, int p, int w, float* y, float* scratch,
                  long long scratch_floats, void* stream);
int wfd_pool_mean_back(const float* dy, int batch, int p, int w, float* dx, float beta,
                       void* stream);
/* d = scale * (pred - target); partial[k] = sum of squared errors of block k
 * (wfd_mse_partials(n) blocks). */
long long wfd_mse_partials(long long n);
int wfd_mse(const float* pred, const float* target, long long n, float scale, float* d,
            float* partial, void* stream);
/* Adam step t >= 1 on FP32 master weights (decoupled weight decay `decay`),
 * writing the BF16 copy w16 when non-NULL. */
int wfd_adam(float* w, const float* g, float* m, float* v, void* w16, long long n, float lr,
             float b1, float b2, float eps, int t, float decay, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WFDTD_H */
