/*
 * hipfdtd.h — C ABI of the MI355X-native FDTD engine (libhipfdtd.so).
 *
 * Plain C types only: opaque handles, pointers and sizes.  Every call returns
 * an int status (HFDTD_OK = 0, negative on error); hfdtd_last_error() gives a
 * thread-local message for the last failing call on that thread.
 *
 * Each entry point names the reference interface it replaces
 * (r-bridge/opencl-fdtd-solver, paths relative to the repository root).
 * Host pointers are borrowed for the duration of the call only and must be
 * C-contiguous in the engine's dtype unless stated otherwise.  Calls on one
 * handle are not thread-safe; work is ordered on the handle's HIP stream.
 */
#ifndef HIPFDTD_H
#define HIPFDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFDTD_ABI_VERSION 1

/* status codes */
#define HFDTD_OK 0
#define HFDTD_EINVAL (-1)      /* bad argument (maps to ValueError)       */
#define HFDTD_ENOMEM (-2)      /* device allocation failed (MemoryError)  */
#define HFDTD_EHIP (-3)        /* HIP runtime error (RuntimeError)        */
#define HFDTD_EUNSUPPORTED (-4)/* unsupported configuration (ValueError)  */
#define HFDTD_ENODEV (-5)      /* no usable device (RuntimeError)         */

/* volume ids for read/write: fields Ex..Hz, then ce_x..z, ca_x..z
 * (engine.py:170-187 buffer set) */
#define HFDTD_EX 0
#define HFDTD_EY 1
#define HFDTD_EZ 2
#define HFDTD_HX 3
#define HFDTD_HY 4
#define HFDTD_HZ 5
#define HFDTD_CEX 6
#define HFDTD_CEY 7
#define HFDTD_CEZ 8
#define HFDTD_CAX 9
#define HFDTD_CAY 10
#define HFDTD_CAZ 11

/* source modes */
#define HFDTD_SRC_SOFT 0    /* F += amp                 (sources.cl:7-21)  */
#define HFDTD_SRC_CURRENT 1 /* F += -ce_c * amp * w     (sources.cl:26-49) */

/* N2F accumulator modes */
#define HFDTD_ACC_REAL 0  /* accumulate in the field dtype (reference default) */
#define HFDTD_ACC_FP64 1  /* accumulate in double (accum_fp64=True)            */
#define HFDTD_ACC_KAHAN 2 /* FP32 with compensated (Kahan) summation           */

/* plane-DFT modes */
#define HFDTD_DFT_DIRECT 0   /* per-step accumulation (FP32, Kahan or FP64)          */
#define HFDTD_DFT_BUFFERED 1 /* K steps buffered, one FP32 matrix-core contraction   */

/* step-path selection for hfdtd_step */
#define HFDTD_PATH_SPLIT 0 /* update_H, sources, update_E: two sweeps        */
#define HFDTD_PATH_FUSED 1 /* one fused H+sources+E sweep (ping-pong state)  */
#define HFDTD_PATH_MARCH 2 /* one register-pipelined fused sweep (ping-pong) */

typedef struct hfdtd_engine hfdtd_engine;
typedef struct hfdtd_n2f hfdtd_n2f;
typedef struct hfdtd_probe hfdtd_probe;
typedef struct hfdtd_batch hfdtd_batch;
typedef struct hfdtd_dft hfdtd_dft;
typedef struct hfdtd_adjoint hfdtd_adjoint;

/* One current / soft source on a box [i0,i1)x[j0,j1)x[k0,k1) of component
 * comp (0=x,1=y,2=z).  Bounds may extend past the grid (clipped); rim-taper
 * edge tests use the unclipped bounds, as add_source_Jx does
 * (engine.py:482-538, sources.cl:26-49).  taper: bit0 i-edges, bit1
 * j-edges, bit2 k-edges, each multiplying w by rim_edge.  amp is rounded to
 * the engine dtype before use (engine.py:529 self.real(jx)). */
typedef struct hfdtd_source {
  int comp;
  int mode;
  int i0, i1, j0, j1, k0, k1;
  int taper;
  double rim_edge;
  double amp;
} hfdtd_source;

/* ── library / device ─────────────────────────────────────────────────── */

/* thread-local message of the last failed call */
const char *hfdtd_last_error(void);
/* HFDTD_ABI_VERSION */
int hfdtd_version(void);
/* replaces _default_opencl_runtime device discovery (engine.py:40-82) */
int hfdtd_device_count(int *count);
/* replaces device.global_mem_size / device.name (engine.py:337-363) */
int hfdtd_device_info(int device, uint64_t *total_bytes, uint64_t *free_bytes, char *name,
                      int name_len);
/* Streaming-bandwidth probe (extension, no counterpart in the reference; for
 * reading a sweep's measured traffic against what plain streaming kernels
 * reach on the device): allocates its own buffers of `bytes` per array
 * (a multiple of 64, 4 KiB to 16 GiB), times `reps` launches of `blocks`
 * 256-thread workgroups and stores the mean seconds per launch.  kind: 0 copy
 * (bytes read + bytes written), 1 triad (2 + 1 arrays), 2 six arrays read and
 * six written, 3 read only (a sum). */
int hfdtd_stream_probe(int device, int kind, uint64_t bytes, int blocks, int reps,
                       double *seconds);

/* Host-only (no GPU needed): 1-D CFS-CPML profile b, c, kappa of length n in
 * float (fp64=0) or double (fp64=1); replaces cpml.build_cpml_1d
 * (cpml.py:84-124).  offset is 0 for E nodes and 0.5 for H nodes. */
int hfdtd_cpml_profile(int n, int npml, double dl, double dt, double offset, int fp64, void *b,
                       void *c, void *kappa);

/* ── engine ───────────────────────────────────────────────────────────── */

/* replaces OpenCLFDTD.__init__ (engine.py:99-210): allocates the 12 volume
 * arrays and the face-stripe psi buffers, fields = 0, ce = dt/eps0, ca = 1,
 * uploads the CPML profiles, folds dtm_dl, dtm, inv_dl (engine.py:543-544,
 * 634).  fp64 selects double fields. */
int hfdtd_create(int nx, int ny, int nz, int npml, double dl, double dt, int fp64, int device,
                 hfdtd_engine **out);
/* frees all device memory of the engine and its monitors */
int hfdtd_destroy(hfdtd_engine *e);
/* bytes the engine holds on the device */
int hfdtd_memory_bytes(const hfdtd_engine *e, uint64_t *bytes);

/* replaces set_epsilon (engine.py:212-254): cell-centred eps_r and optional
 * sigma (NULL = lossless) on the host, in float (src_fp64=0) or double
 * (src_fp64=1); edge averaging and ce/ca/cb are computed on the device in
 * FP64 (materials.py:28-89) and stored in the engine dtype. */
int hfdtd_set_materials(hfdtd_engine *e, const void *eps_r, const void *sigma, int src_fp64);

/* replaces enqueue_copy into a field/coefficient buffer */
int hfdtd_write_volume(hfdtd_engine *e, int which, const void *host);
/* replaces _read_field (engine.py:747-751): blocking device->host copy */
int hfdtd_read_volume(hfdtd_engine *e, int which, void *host);
/* replaces read_point (engine.py:753-765); index = (i*Ny + j)*Nz + k */
int hfdtd_read_point(hfdtd_engine *e, int which, int64_t index, double *out);
/* zero fields and psi, keep coefficients */
int hfdtd_reset_fields(hfdtd_engine *e);

/* replaces _update_H (engine.py:540-629; yee_update.cl:11-118) */
int hfdtd_update_H(hfdtd_engine *e);
/* replaces _update_E (engine.py:631-722; yee_update.cl:120-238) */
int hfdtd_update_E(hfdtd_engine *e);
/* replaces add_source_Ex / add_source_Jx (engine.py:448-538) for n sources */
int hfdtd_add_sources(hfdtd_engine *e, const hfdtd_source *src, int n);
/* one Yee step H -> sources -> E (engine.py:724-735 minus the callbacks);
 * path = HFDTD_PATH_SPLIT, HFDTD_PATH_FUSED (LDS-DMA tiles) or
 * HFDTD_PATH_MARCH (register-pipelined tiles).  n <= 16 sources. */
int hfdtd_step(hfdtd_engine *e, const hfdtd_source *src, int n, int path);
/* queue.finish() */
int hfdtd_synchronize(hfdtd_engine *e);
/* Fused-sweep launch shape: rows of j per block (1, 2, 4, 6, 8, 10 or 12;
 * FP64 engines use at most 4, lossy FP32 engines at most 10; a block is
 * 64 x (rows + 1) threads), planes of i per block (>= 1; -1 = chosen from the
 * grid: up to 64, two blocks per CU) and the number of
 * XCDs the work order is grouped for (1 = hardware order, 8 = MI355X); 0
 * keeps the current value.  Used by the autotuner. */
int hfdtd_set_fused_tiling(hfdtd_engine *e, int rows_j, int planes_i, int xcd_groups);
/* Split-sweep work order: each XCD runs runs of run_blocks consecutive
 * 256-cell blocks (1 = hardware round robin; 0 keeps the current value).
 * Used by the autotuner. */
int hfdtd_set_split_schedule(hfdtd_engine *e, int run_blocks);
/* March-sweep launch shape (HFDTD_PATH_MARCH): j rows per wavefront (1 or 2;
 * a block is 4 wavefronts of 64 k; FP64 engines use 1; until set, one on
 * grids with a material index and two otherwise), planes of i per block
 * (rounded up to even; -1 = chosen from the grid), consecutive blocks per XCD
 * run, and non-temporal field streams (1 on, 0 off, -1 when the slabs exceed
 * the last-level cache).  0 keeps rows / planes / run, -2 keeps nontemporal.
 * Used by the autotuner; no counterpart in the reference. */
int hfdtd_set_march_tiling(hfdtd_engine *e, int rows_per_wave, int planes_i, int run_blocks,
                           int nontemporal);
/* No counterpart in the reference.  Addressing of the march sweep: 0 (the
 * default) one buffer resource per slab, with 32-bit offsets, on slabs below
 * 4 GB, and one per volume on FP32 slabs of 4 GB and more (volumes below
 * 2 GB, about 810^3 cells); 1 one per volume on any FP32 grid (same results;
 * for testing that addressing on small grids). */
/* March-sweep tile width: 64 (default) or 32 columns of k, with 64 / width
 * rows of j per wavefront slot (0 = 64).  32-column tiles take fewer tiles
 * where nz modulo 64 lies in (0, 32] (thin grids: nz <= 32, nz = 96 ...);
 * they apply to FP32 grids with slabs below 4 GB and one row per wavefront
 * (FP64 and larger grids keep 64).  Results are the same for every width. */
int hfdtd_set_march_width(hfdtd_engine *e, int width);
int hfdtd_set_march_addressing(hfdtd_engine *e, int per_volume);
/* No counterpart in the reference.  1 (the default): a K-buffered
 * near-to-far monitor's per-step gather that does not fill its K-step block
 * is held by the engine, and the next march sweep (FP32, slabs below 4 GB)
 * runs it in blocks after its own, in the same launch: both only read the
 * state the step left (the sweep writes the other slab).  Any other call on
 * the engine runs a held gather on its own first.  0: every gather is
 * launched at once.  Results are the same either way. */
int hfdtd_set_overlap(hfdtd_engine *e, int on);

/* Row-padded runs (extension, no counterpart in the reference: its kernels
 * index the dense layout).  Where nz does not fill whole 128-byte lines
 * (FP32: nz not a multiple of 32; FP64: of 16), a march tile's rows start
 * mid-line and each 64-column row segment touches three cache lines instead
 * of two.  hfdtd_padded_begin copies the state, psi and the coefficient
 * source the march sweep reads into a copy with padded rows; HFDTD_PATH_MARCH
 * steps with at most 4 sources then run on the copy, with identical results.
 * Every other call on the engine, its monitors, probes, plane DFTs, adjoint
 * or batches first ends the run (the copy's state and psi written back), as
 * does hfdtd_padded_end.  A no-op (not live) where rows already fill whole
 * lines; HFDTD_EUNSUPPORTED while a monitor records z faces in the sweep or
 * when the copy would exceed the sweep's slab limits; HFDTD_ENOMEM without
 * memory for the copy (one more state and psi, both parities). */
int hfdtd_padded_begin(hfdtd_engine *e);
int hfdtd_padded_end(hfdtd_engine *e);
/* 1 while a row-padded run is live, else 0 */
int hfdtd_padded_live(const hfdtd_engine *e, int *live);
/* Two Yee steps in one launch (the engine.py:724-735 step twice, minus the
 * callbacks): src_a[0..na) are the sources of the first step, src_b[0..nb)
 * those of the second, at most 4 each.  Results equal two hfdtd_step calls
 * bit for bit.  FP32 engines with field and psi slabs below 2 GB; otherwise
 * HFDTD_EUNSUPPORTED.  No counterpart in the reference (an extension: the
 * reference's host loop cannot fuse steps). */
int hfdtd_step_pair(hfdtd_engine *e, const hfdtd_source *src_a, int na,
                    const hfdtd_source *src_b, int nb);
/* Cells whose state between the two steps of hfdtd_step_pair is kept for
 * observers: n <= 8 boxes, 7 ints each (i0, i1, j0, j1, k0, k1, half-open
 * and clipped to the grid, then a mask of the components kept, bit c for
 * Ex Ey Ez Hx Hy Hz, 0 = all).  n > 0 allocates a shadow field slab on
 * first use; n = 0 records nothing.  Extension, no counterpart in the
 * reference. */
int hfdtd_set_watch(hfdtd_engine *e, const int *boxes, int n);
/* 1: field reads (volumes, monitors, probes) see the shadow slab, i.e. the
 * state after the first step of the last hfdtd_step_pair at the watched
 * cells (other cells are undefined); 0: the current state.  Extension. */
int hfdtd_set_view(hfdtd_engine *e, int shadow);
/* Materials of the engine's material index: on a lossless FP32 grid with at
 * most 48 distinct eps_r values, hfdtd_set_materials also stores a byte of
 * material per cell and a table of ce per pair of materials, which the march
 * sweep and the split E sweeps read instead of the ce volumes (identical
 * values); 0 when the grid
 * has no index (more values, lossy or FP64 grids, a coefficient volume
 * written by hand).  Extension, no counterpart in the reference. */
int hfdtd_material_count(const hfdtd_engine *e, int *nmat);
/* Where the march sweep takes ce from: 0 the three ce volumes, 1 the
 * material index, 2 the eps_r volume (lossless FP32 grids without an index
 * whose eps_r values are exact in FP32: hfdtd_set_materials keeps them, and
 * the sweep recomputes each edge's ce in FP64 as the coefficients were made,
 * identical values).  Writing a coefficient volume returns it to 0 until the
 * next hfdtd_set_materials.  Extension, no counterpart in the reference. */
int hfdtd_ce_source(const hfdtd_engine *e, int *source);
/* i planes per block of hfdtd_step_pair (0 = 24, at most 1024).  Extension. */
int hfdtd_set_pair_planes(hfdtd_engine *e, int planes_i);

/* HIP-event timing of the Yee sweep kernels on the engine's stream: while
 * enabled, every n-th update_H / update_E / fused / march / pair sweep launch
 * (n = enable: 1 times every launch, 0 turns timing off) records a start and
 * a stop event around its kernel (hipExtLaunchKernelGGL).  A timed launch
 * costs the stream a few microseconds, so a sample (an odd n, which times
 * the H and E sweeps of the split path alike) keeps the timed run close to
 * an untimed one.  Enabling clears earlier records.  Extension, no
 * counterpart in the reference. */
int hfdtd_timing(hfdtd_engine *e, int enable);
/* synchronizes the stream; total milliseconds and number of the timed sweep
 * launches */
int hfdtd_timing_read(hfdtd_engine *e, double *total_ms, int *launches);
/* sweep launches while timing was on, timed or not.  Extension. */
int hfdtd_timing_launches(const hfdtd_engine *e, long long *launches);

/* ── near-to-far monitor (monitors.py:383-804, dft_farfield.cl:98-635) ── */

/* box = ix0,ix1,iy0,iy1,iz0,iz1 (inclusive, already snapped); nfreq >= 1 */
int hfdtd_n2f_create(hfdtd_engine *e, const int box[6], int nfreq, int acc_mode,
                     hfdtd_n2f **out);
int hfdtd_n2f_destroy(hfdtd_n2f *m);
/* n_face and the six packed face offsets (monitors.py:396-412) */
int hfdtd_n2f_layout(const hfdtd_n2f *m, int *n_face, int offsets[6]);
/* replaces accumulate_dft_faces_fused (dft_farfield.cl:116-230).  phase holds
 * nfreq x {e.re, e.im, h.re, h.im}, already including dt. */
int hfdtd_n2f_accumulate(hfdtd_n2f *m, const double *phase);
/* replaces snapshot_dft (monitors.py:498-503) */
int hfdtd_n2f_snapshot(hfdtd_n2f *m);
/* replaces dft_rel_change_partial + host sum (monitors.py:505-569): returns
 * sum|cur-prev|^2 and sum|cur|^2 over the 6 buffers of one frequency */
int hfdtd_n2f_rel_change(hfdtd_n2f *m, int freq, double *num, double *den);
/* face-packed DFT of one component (0..5 = Ex..Hz) and frequency as
 * n_face complex values in double (re, im interleaved) */
int hfdtd_n2f_read(hfdtd_n2f *m, int comp, int freq, double *host);
/* replaces farfield_accumulate_nl + farfield_nl_to_eh
 * (dft_farfield.cl:372-635; monitors.py:644-731) with a deterministic
 * two-stage reduction.  obs: n_obs x 3 metres; eh: n_obs x 6 complex (12
 * doubles each).  k_wave and obs are used as given (the caller rounds them
 * like monitors.py:653,669 if it wants the reference's rounding); l_sign is
 * -1 for eta0*N_t - r x L (dft_farfield.cl:617) or +1 (monitors.py:378).
 * geom_fp64 selects double (1) or the field dtype (0) for r.r' phases. */
int hfdtd_n2f_farfield(hfdtd_n2f *m, int freq, const double *obs, int n_obs, double k_wave,
                       double eta0, int l_sign, int geom_fp64, double *eh);
/* No counterpart in the reference.  hfdtd_n2f_farfield for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it): same arguments, kernels and results, but
 * the partial sums live in a library-owned static device array of 1048576
 * doubles (8 MiB, held on every device the library runs on until the process
 * ends) instead of a buffer allocated at the first call.  The call needs
 * n_obs * (12 * ceil(n_face / 2048) + 15) doubles of it and returns
 * HFDTD_EINVAL beyond that (the 256^3 benchmark's 362 points over its
 * 221184 face samples need 474552).  It calls no allocator.  All calls on one
 * device share the array: calls that overlap on one device (other engines'
 * streams, other threads) race and must be serialised by the caller. */
int hfdtd_n2f_farfield_noalloc(hfdtd_n2f *m, int freq, const double *obs, int n_obs,
                               double k_wave, double eta0, int l_sign, int geom_fp64,
                               double *eh);

/* K-step buffering of the face DFT (extension; per-step accumulation is the
 * reference's monitors.py:571-624).  K = 0: per-step sums; K even, 2..256:
 * the co-located samples of K steps are kept on the device and folded into
 * the sums by one FP32 matrix-core contraction per K steps (float32 engines,
 * at most 16 frequencies).  Readers (snapshot, rel change, read, farfield)
 * fold a partial block in first. */
int hfdtd_n2f_set_buffering(hfdtd_n2f *m, int K);
/* folds the samples buffered so far into the sums now (asynchronous, on the
 * engine's stream); the readers above do it themselves.  No counterpart in
 * the reference, whose per-step sums need none. */
int hfdtd_n2f_flush(hfdtd_n2f *m);
/* No counterpart in the reference.  1 in *on when this monitor's z faces are
 * recorded by the engine's march sweep: with K-step buffering on a float32
 * engine, the first such monitor of the engine whose z faces lie at k >= 1
 * and at least two cells apart; its gather then reads the two z faces from
 * that record (the same values, contiguous) after march steps, and from the
 * fields after any other change of the state. */
int hfdtd_n2f_zrecord(const hfdtd_n2f *m, int *on);

/* ── point probe (device-side DFT of one field sample) ──────────────────── */

/* Replaces the reference harness's per-step read_point + host DFT
 * (tests/meep_validation/cases.py:179-197): accumulates
 * sum_n F(i,j,k) * phase_n[f] in FP64 on the device, one launch per step.
 * comp: 0..5 = Ex..Hz. */
int hfdtd_probe_create(hfdtd_engine *e, int i, int j, int k, int comp, int nfreq,
                       hfdtd_probe **out);
int hfdtd_probe_destroy(hfdtd_probe *p);
/* phase: nfreq x {re, im} for this step (already including dt) */
int hfdtd_probe_accumulate(hfdtd_probe *p, const double *phase);
/* out: nfreq x {re, im} */
int hfdtd_probe_read(hfdtd_probe *p, double *out);

/* ── multi-frequency plane DFT ────────────────────────────────────────────
 * Replaces the host-side plane DFT of the reference's test harness
 * (tests/harness cases.py:179-197: acc += Ex(t_{n+1}) e^{jw t_{n+1}} dt over a
 * z-plane) with a device monitor over any axis-aligned plane and any subset
 * of Ex..Hz.  Samples p = c*(n0*n1) + a*n1 + b, c over the selected
 * components in Ex..Hz order; plane axes (a, b) = (j, k) for x-planes,
 * (i, k) for y-planes, (i, j) for z-planes. */
/* comp_mask: bit c selects component c (HFDTD_EX..HFDTD_HZ); 1..16
 * frequencies; mode HFDTD_DFT_DIRECT or HFDTD_DFT_BUFFERED (K even, 2..256,
 * FP32 engines); accum HFDTD_ACC_REAL / _FP64 / _KAHAN. */
int hfdtd_dft_plane_create(hfdtd_engine *e, int axis, int index, int comp_mask, int nfreq, int K,
                           int mode, int accum, hfdtd_dft **out);
int hfdtd_dft_destroy(hfdtd_dft *d);
/* one step: phase = nfreq x {cos, sin}(w t) * dt, applied to the current fields */
int hfdtd_dft_accumulate(hfdtd_dft *d, const double *phase);
/* contract a partially filled block now (read does this itself) */
int hfdtd_dft_flush(hfdtd_dft *d);
/* zero the sums */
int hfdtd_dft_reset(hfdtd_dft *d);
int hfdtd_dft_shape(const hfdtd_dft *d, int *ncomp, int *n0, int *n1, int *nfreq);
/* sums as double: [ncomp][n0][n1][nfreq] x {re, im} */
int hfdtd_dft_read(hfdtd_dft *d, double *out);

/* ── batched sweep of independent simulations ─────────────────────────── */

/* No counterpart in the reference (it runs one grid at a time).  Groups n
 * engines of any sizes (same dl, dtype and device) so one launch per step
 * advances all of them (path below); the engines stay usable
 * for reads afterwards.  Each engine may carry one current/soft source whose
 * amplitude per step comes from a table, and one point probe whose per-step
 * phases come from a table. */
int hfdtd_batch_create(hfdtd_engine *const *engines, int n, hfdtd_batch **out);
int hfdtd_batch_destroy(hfdtd_batch *b);
/* src: n descriptors (amp ignored); amp: n x n_steps table (row per engine) */
int hfdtd_batch_set_sources(hfdtd_batch *b, const hfdtd_source *src, const double *amp,
                            int n_steps);
/* ijk: n x 3 cells; comp 0..5; phase: n x n_steps x nfreq x {re, im} */
int hfdtd_batch_set_probes(hfdtd_batch *b, const int *ijk, int comp, int nfreq,
                           const double *phase, int n_steps);
/* advance every engine n_steps steps; graph_steps > 0 captures that many
 * steps in a HIP graph and replays it (remainder launched directly) */
int hfdtd_batch_run(hfdtd_batch *b, int n_steps, int graph_steps);
/* probe DFTs: n x nfreq x {re, im} */
int hfdtd_batch_read_probes(hfdtd_batch *b, double *out);
/* persistent grid size (0 = one block per tile), rows of j per block (2, 4, 6
 * or 8; FP64 uses at most 4) and planes of i per block */
int hfdtd_batch_set_launch(hfdtd_batch *b, int blocks, int rows_j, int planes_i);
/* HFDTD_PATH_FUSED (default): one persistent fused-tile launch per step;
 * HFDTD_PATH_SPLIT: persistent split sweeps over blocks of 256 cells of the
 * flattened (j, k) planes (H, sources, E, probes), which leave no lanes idle
 * on extents that are not multiples of 64;
 * HFDTD_PATH_MARCH: one launch per step of march tiles over every engine,
 * with each engine's source applied and its probe recorded inside the
 * sweep (field and psi slabs below 4 GB / 2 GB per engine) */
int hfdtd_batch_set_path(hfdtd_batch *b, int path);
/* HFDTD_PATH_SPLIT work items: runs of 256 cells of one plane per block (1..8;
 * 0 keeps the current value, default 2).  Longer items pay a block's set-up
 * once for more cells and leave more threads idle at each plane's end. */
int hfdtd_batch_set_split(hfdtd_batch *b, int runs);
/* HFDTD_PATH_MARCH launch shape: j rows per wavefront (1 or 2; FP64: 1),
 * i planes per chunk (2..256, rounded up to even), consecutive work items
 * per XCD run, tile width in k (64 or 32 columns; 32 with one row per
 * wavefront, FP32 only); 0 keeps the current value (defaults 1, 32, 16,
 * 64) */
int hfdtd_batch_set_march(hfdtd_batch *b, int rows, int planes, int run, int width);
/* HFDTD_PATH_MARCH row padding (1 = on, the default; 0 = off).  The march
 * path runs every engine on a copy in a batch-owned arena (state and psi of
 * both parities, coefficients, profiles, material index; filled before a
 * run, written back after it).  With padding on, the copy's rows are padded
 * to whole 128-byte lines where nz does not fill them: unaligned rows make a
 * 64-column tile row touch three cache lines instead of two.  Results are
 * identical either way. */
int hfdtd_batch_set_march_padding(hfdtd_batch *b, int on);

/* ── adjoint gradient of probe-DFT objectives ─────────────────────────────
 * No counterpart in the reference (it has no differentiation).  Objective:
 * any real L(F) of probe DFTs F_p = sum_{n=1..N} E_p(t_n) e^{j w t_n} dt,
 * t_n = n dt, on Ex/Ey/Ez cells.  The reverse sweep runs the transposed
 * Yee + CPML step (exact discrete adjoint of hfdtd_step on a lossless grid)
 * and sums G_c = sum_n E~_c^n (E_c^n - E_c^{n-1}) in FP64 over the edges of
 * the design box grown one cell toward lower indices; dL/d(eps_edge) =
 * -G/eps_edge.  Forward states are rebuilt per segment from checkpoints
 * (checkpoint-every-K recomputation); E over that box is recorded per step
 * into history slots. */
/* design_box: i0,i1,j0,j1,k0,k1 (half-open, inside the grid); hist_slots >= 2
 * (segment length + 1); n_ckpt full-state checkpoints.  Lossless grids only. */
int hfdtd_adjoint_create(hfdtd_engine *e, const int *design_box, int hist_slots, int n_ckpt,
                         hfdtd_adjoint **out);
int hfdtd_adjoint_destroy(hfdtd_adjoint *a);
int hfdtd_adjoint_memory_bytes(const hfdtd_adjoint *a, uint64_t *bytes);
/* box[6] of the gradient sums (design box grown toward lower indices) */
int hfdtd_adjoint_grad_box(const hfdtd_adjoint *a, int *box);
/* ijkc: n x {i, j, k, component 0..2}, distinct; omega: nfreq angular freqs */
int hfdtd_adjoint_set_probes(hfdtd_adjoint *a, int n, const int *ijkc, int nfreq,
                             const double *omega);
int hfdtd_adjoint_probe_reset(hfdtd_adjoint *a);
/* F: n x nfreq x {re, im} (FP64 sums) */
int hfdtd_adjoint_probe_read(hfdtd_adjoint *a, double *out);
/* conj(dL/dF*)-style coefficients a_pf (n x nfreq x {re, im}): the adjoint
 * source after step n is sum_f Re(a_pf e^{j w_f t_n}) dt; for torch's
 * gradient g of a real L, a = conj(g) */
int hfdtd_adjoint_set_objective_grad(hfdtd_adjoint *a, const double *coef);
/* full forward state (E, H, psi) to / from checkpoint c (device copies) */
int hfdtd_adjoint_save(hfdtd_adjoint *a, int c);
int hfdtd_adjoint_restore(hfdtd_adjoint *a, int c);
/* E over the gradient box into history slot */
int hfdtd_adjoint_record(hfdtd_adjoint *a, int slot);
/* n_steps forward steps of the engine (path HFDTD_PATH_FUSED or
 * HFDTD_PATH_MARCH when the grid allows it, else the split sweeps), step q
 * using src[r].amp =
 * amp[q*nsrc + r]; the step taking the grid to t_n, n = step0 + q + 1,
 * accumulates the probe DFT if probe != 0 and records E into slot
 * slot0 + q + 1 if slot0 >= 0 */
int hfdtd_adjoint_forward(hfdtd_adjoint *a, const hfdtd_source *src, int nsrc,
                          const double *amp, int n_steps, int64_t step0, int slot0, int probe,
                          int path);
/* zero E~, H~, psi~ and G; add the adjoint source of step n_last */
int hfdtd_adjoint_begin(hfdtd_adjoint *a, int64_t n_last);
/* reverse steps n = n_hi .. n_lo+1: G += E~ (E^n - E^{n-1}) from slots
 * slot(n) = slot_hi - (n_hi - n) and slot(n) - 1, transposed step, adjoint
 * source of step n - 1 */
int hfdtd_adjoint_reverse(hfdtd_adjoint *a, int64_t n_hi, int64_t n_lo, int slot_hi);
/* G: 3 x box cells (component-major, k fastest) */
int hfdtd_adjoint_read_grad(hfdtd_adjoint *a, double *out);
/* adjoint volumes 0..5 in the engine dtype: ce_c E~_c for c = x, y, z (the
 * stored, ce-scaled E adjoint), then H~x, H~y, H~z */
int hfdtd_adjoint_read_volume(hfdtd_adjoint *a, int which, void *host);

#ifdef __cplusplus
}
#endif

#endif /* HIPFDTD_H */
