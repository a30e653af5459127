/*
 * libbpmi -- MI355X (gfx950) SAR backprojection and FFBP, C ABI.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses;
 * "host" arrays are read during the call.  Every function returns 0 (BPMI_OK)
 * on success and a nonzero status otherwise; bpmi_last_error() returns the
 * message of the calling thread's last failure.  Work is enqueued on the
 * given stream (NULL = the null stream) and never synchronises unless the
 * function name says so.  The library allocates nothing implicitly: the
 * caller owns every buffer.
 *
 * Complex data are interleaved complex64 (float2) unless stated.  Node
 * geometry ``geo`` is a host array of 7 doubles (ox, oy, oz, r0, dr, u0, du):
 * origin, first slant range and step, first u = sin(angle) and step
 * (fast_bp/ffbp.py:98-111).  Polar images are [nr, nu] with u fastest.
 *
 * Python binding: bp_project_amd/_lib.py (ctypes).  The reference functions
 * each entry point replaces are cited beside it.
 */
#ifndef BPMI_H
#define BPMI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BPMI_VERSION 1

/* status codes */
#define BPMI_OK 0
#define BPMI_ERR_INVALID 1     /* bad argument (Python maps to ValueError) */
#define BPMI_ERR_HIP 2         /* HIP runtime error (RuntimeError) */
#define BPMI_ERR_UNSUPPORTED 3 /* valid but unsupported combination */

/* arithmetic modes (fast_bp/gpu_backend.py:57 MODES plus two additions) */
#define BPMI_MODE_FP64 0   /* FP64 geometry, interpolation, phase, accumulation */
#define BPMI_MODE_FP32 1   /* FP32 everything (fast_bp/gpu/kernels.cu fp32) */
#define BPMI_MODE_MIXED 2  /* FP64 range and phase reduction, FP32 interp/accum */
#define BPMI_MODE_DFLOAT 3 /* double-float (hi+lo FP32) range and phase */
#define BPMI_MODE_SPLIT 4  /* LDS-tiled: FP64 per-tile reference + FP32 residual */

/* interpolators */
#define BPMI_INTERP_LAGRANGE4 0 /* fast_bp/ffbp.py:17-40 */
#define BPMI_INTERP_LINEAR 1
#define BPMI_INTERP_KAISER8 2   /* 8-tap Kaiser(beta=6)-windowed sinc */
#define BPMI_INTERP_NONE (-1)   /* persistent FFBP without leaves: rejected (switched off) */

/* flags for BP outputs */
#define BPMI_ACCUMULATE 1 /* out += result (else out = result) */
#define BPMI_OUT_C128 2   /* out is complex128 (FP64 mode accumulator) */

/* ---- runtime ----------------------------------------------------------- */
int bpmi_version(void);
int bpmi_last_error(char* buf, size_t n);
int bpmi_device_count(int* count);
int bpmi_set_device(int device);
int bpmi_get_device(int* device);
int bpmi_device_info(int device, char* name, size_t name_len, char* arch, size_t arch_len,
                     int* cu_count, int* clock_khz, size_t* total_mem);
int bpmi_runtime_version(int* version);
int bpmi_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Rate of the device wall clock (s_memrealtime, the clock of the persistent
 * scheduler's bounded waits and completion stamps) of the current device, kHz. */
int bpmi_wall_clock_khz(int* khz);
int bpmi_malloc(void** ptr, size_t bytes);
int bpmi_free(void* ptr);
int bpmi_host_alloc(void** ptr, size_t bytes);
int bpmi_host_free(void* ptr);
int bpmi_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int bpmi_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int bpmi_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
/* Kernel-based zero fill (safe inside graph capture). ptr 16-byte aligned. */
int bpmi_zero(void* ptr, size_t bytes, void* stream);
int bpmi_stream_create(void** stream);
int bpmi_stream_destroy(void* stream);
int bpmi_stream_synchronize(void* stream);
int bpmi_device_synchronize(void);
int bpmi_event_create(void** event);
int bpmi_event_destroy(void* event);
int bpmi_event_record(void* event, void* stream);
int bpmi_event_synchronize(void* event);
int bpmi_event_elapsed(float* ms, void* start, void* end);
/* Graph capture of everything enqueued on ``stream`` between begin and end
 * (resident_backend.py:250-301 capture_graph/replay). */
int bpmi_graph_begin(void* stream);
int bpmi_graph_end(void* stream, void** graph_exec);
int bpmi_graph_launch(void* graph_exec, void* stream);
int bpmi_graph_destroy(void* graph_exec);

/* ---- direct backprojection --------------------------------------------- */
/* Common inputs: data complex64 [pulses, samples]; pos f64 [pulses, 3];
 * weight, phase f64 [pulses] (NULL = 1 and 0).  Pulses [pulse_begin,
 * pulse_end) are summed.  range_start/data_dr define the range of sample i
 * as range_start + i*data_dr; k = 4*pi*fc/c with signed fc. */

/* Polar grid of one node, demodulated by exp(-j k r) (pixels on z = 0).
 * Replaces bp_polar (fast_bp/ffbp.py:58-76), bp_polar_{mode} and
 * direct_chunk_{mode} (fast_bp/gpu/kernels.cu:139-339). */
int bpmi_bp_polar(const void* data, int samples, const double* pos, const double* weight,
                  const double* phase, int pulse_begin, int pulse_end, const double* geo,
                  int nr, int nu, int side, double range_start, double data_dr, double k,
                  int mode, int interp, int flags, void* out, void* stream);

/* bpmi_bp_polar with an explicit SPLIT-mode geometry reference
 * (origin_ref = 0: each tile's centre; 1: the node origin, for sub-apertures
 * short relative to the slant range such as FFBP leaves) and tile width
 * (tile_cols_log2 = 3..6, or -1 to choose from the grid spacing).  Other
 * modes ignore both. */
int bpmi_bp_polar_ex(const void* data, int samples, const double* pos, const double* weight,
                     const double* phase, int pulse_begin, int pulse_end, const double* geo,
                     int nr, int nu, int side, double range_start, double data_dr, double k,
                     int mode, int interp, int flags, void* out, void* stream, int origin_ref,
                     int tile_cols_log2);

/* Regular Cartesian grid on z = z0, image [ny, nx] with x fastest, not
 * demodulated: exp(+j (k d + phase)).  Cartesian direct BP
 * (backproject_points at grid pixels, fast_bp/ffbp.py:43-55). */
int bpmi_bp_grid(const void* data, int samples, const double* pos, const double* weight,
                 const double* phase, int pulse_begin, int pulse_end, double x0, double dx,
                 int nx, double y0, double dy, int ny, double z0, double range_start,
                 double data_dr, double k, int mode, int interp, int flags, void* out,
                 void* stream);

/* bpmi_bp_grid with a caller workspace (device, workspace_bytes long).  In
 * split mode a grid of few tiles (fewer than 4096 tiles of 1024 pixels)
 * has its pulse range split into chunks, each written by its own
 * workgroups to a partial image in the workspace, and the partials are
 * summed in chunk order into out (added to it with BPMI_ACCUMULATE): the
 * same image as bpmi_bp_grid up to FP32 summation order, reproducible from
 * call to call.  The chunk count is capped by what the workspace holds;
 * without a workspace, or when the grid fills the GPU, this is
 * bpmi_bp_grid.  Same reference function as bpmi_bp_grid
 * (backproject_points, fast_bp/ffbp.py:43-55). */
int bpmi_bp_grid_ws(const void* data, int samples, const double* pos, const double* weight,
                    const double* phase, int pulse_begin, int pulse_end, double x0, double dx,
                    int nx, double y0, double dy, int ny, double z0, double range_start,
                    double data_dr, double k, int mode, int interp, int flags, void* out,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Workspace bytes with which bpmi_bp_grid_ws uses its full chunk count for
 * an nx x ny grid and that many pulses (0: the grid needs none). */
int bpmi_bp_grid_workspace_bytes(int nx, int ny, int pulses, size_t* bytes);

/* bpmi_bp_grid_ws that allocates nothing and takes no workspace: the
 * pulse-split partial images live in a static 96 MiB device array of the
 * library (one per GPU, held until the process ends), so the output equals
 * bpmi_bp_grid_ws with bpmi_bp_grid_workspace_bytes() bytes bit for bit.
 * Partials above 96 MiB are rejected (BPMI_ERR_INVALID); by the chunk rule no
 * grid needs more (at most (8192 + T - 1) * 8 KiB for T < 4096 tiles of 1024
 * pixels; grids of 4096 tiles or more need none).  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Calls that overlap on one GPU share
 * the array and race; the first call on a device resolves the array's
 * address, so make it outside graph capture. */
int bpmi_bp_grid_noalloc(const void* data, int samples, const double* pos, const double* weight,
                         const double* phase, int pulse_begin, int pulse_end, double x0,
                         double dx, int nx, double y0, double dy, int ny, double z0,
                         double range_start, double data_dr, double k, int mode, int interp,
                         int flags, void* out, void* stream);

/* Arbitrary points xyz f64 [npoints, 3] (device).  Replaces
 * backproject_points (fast_bp/ffbp.py:43-55). */
int bpmi_bp_points(const void* data, int samples, const double* pos, const double* weight,
                   const double* phase, int pulse_begin, int pulse_end, const double* xyz,
                   long long npoints, double range_start, double data_dr, double k, int mode,
                   int interp, int flags, void* out, void* stream);

/* Per-(point, pulse) unweighted contributions out[npoints, pulses] complex64
 * (fast_bp/ffbp.py:201-211 target_contributions). */
int bpmi_target_contributions(const void* data, int samples, const double* pos, int pulses,
                              const double* xyz, long long npoints, double range_start,
                              double data_dr, double k, int mode, int interp, void* out,
                              void* stream);

/* ---- adjoint: image gradient -> per-pulse gradients --------------------- */
/* No reference counterpart (the reference autofocus, fast_bp/autofocus.py:37-79,
 * is a phase-gradient estimate built on target_contributions).  For the
 * Cartesian/points BP above (any mode's forward model), given grad = dL/dI
 * (complex64, PyTorch convention dL = Re sum conj(grad) dI), computes
 * grad_phase[n] = dL/dphase_n and grad_pos[n, 3] = dL/dpos_n (FP64, device)
 * for pulses [pulse_begin, pulse_end).  ``partials`` is caller-provided
 * scratch (bpmi_grid_adjoint_workspace_bytes() bytes for the grid adjoint,
 * bpmi_adjoint_workspace_bytes() for points).  Deterministic.  The grid
 * adjoint runs on the SPLIT forward's tiles (FP64 tile reference, FP32
 * residual, staged windows); the points adjoint in per-point FP64 geometry. */
int bpmi_adjoint_workspace_bytes(long long npixels, int pulses, size_t* bytes);
/* Workspace of bpmi_bp_grid_adjoint for an nx x ny grid (its 2-D tiles need
 * more partial rows than npixels / 1024 when the grid is not a multiple of
 * the tile shape); bpmi_adjoint_workspace_bytes sizes bpmi_bp_points_adjoint. */
int bpmi_grid_adjoint_workspace_bytes(int nx, int ny, int pulses, size_t* bytes);
int bpmi_bp_grid_adjoint(const void* data, int samples, const double* pos, const double* weight,
                         const double* phase, int pulse_begin, int pulse_end, double x0,
                         double dx, int nx, double y0, double dy, int ny, double z0,
                         double range_start, double data_dr, double k, int interp,
                         const void* grad_image, void* partials, size_t partials_bytes,
                         double* grad_phase, double* grad_pos, void* stream);
int bpmi_bp_points_adjoint(const void* data, int samples, const double* pos, const double* weight,
                           const double* phase, int pulse_begin, int pulse_end, const double* xyz,
                           long long npoints, double range_start, double data_dr, double k,
                           int interp, const void* grad_points, void* partials,
                           size_t partials_bytes, double* grad_phase, double* grad_pos,
                           void* stream);

/* ---- FFBP merge / regrid ----------------------------------------------- */
/* One child polar image -> destination grid, out = v exp(j k (rr - r)) or
 * out += that with BPMI_ACCUMULATE.  Replaces regrid (fast_bp/ffbp.py:79-95)
 * and regrid_{mode} (fast_bp/gpu/kernels.cu:341-429). */
int bpmi_regrid(const void* src, const double* src_geo, int snr, int snu,
                const double* dst_geo, int nr, int nu, int side, double k, int mode, int interp,
                int flags, void* out, void* stream);

/* Both children in one pass, each rounded to complex64 before the FP32 sum
 * (merge_regrid_{mode}, fast_bp/gpu/resident_kernels.cu:35-157). */
int bpmi_merge2(const void* left, const double* left_geo, int lnr, int lnu, const void* right,
                const double* right_geo, int rnr, int rnu, const double* dst_geo, int nr, int nu,
                int side, double k, int mode, int interp, void* out, void* stream);

/* left += right, complex64 (add_float2_inplace, kernels.cu:431-437). */
int bpmi_add_inplace(void* left, const void* right, long long n, void* stream);
/* complex128 -> complex64 (cast_double2_to_float2, kernels.cu:439-444). */
int bpmi_cast_c128_c64(const void* in, void* out, long long n, void* stream);
/* out[nr, nu] = in[r_begin + i, u_begin + j] with row length src_nu
 * (crop_float2, kernels.cu:446-454). */
int bpmi_crop(const void* in, int src_nu, int r_begin, int u_begin, int nr, int nu, void* out,
              void* stream);

/* Polar image (rows of length ld >= nu, so a crop is a view) -> points,
 * remodulated by exp(+j k r) (evaluate_image, fast_bp/ffbp.py:189-198). */
int bpmi_eval_points(const void* img, const double* geo, int nr, int nu, int ld,
                     const double* xyz, long long npoints, double k, int mode, int interp,
                     void* out, void* stream);
/* Same onto a regular grid on z = 0, image [ny, nx] x fastest
 * (the Cartesian ground image of fast_bp/run.py:118-125). */
int bpmi_eval_grid(const void* img, const double* geo, int nr, int nu, int ld, double x0,
                   double dx, int nx, double y0, double dy, int ny, double k, int mode, int interp,
                   void* out, void* stream);

/* ---- level-batched FFBP (batched_resident_backend.py:72-150) ----------- */
/* node_int int32 [nodes, 8] = (a, b, nr, nu, side, left, right, height);
 * node_geo f64 [nodes, 7]; offsets int64 [nodes] element offsets of each
 * node image in ``arena`` (complex64).  ``blocks`` int32 [nblocks, 2] =
 * (node, first_pixel) pairs; each block covers 256 pixels of one node.
 * In BPMI_MODE_SPLIT the leaf table instead holds (node, tile) pairs, with
 * bpmi_split_tiles() tiles of 1024 pixels per node.
 * All table pointers are device pointers. */
int bpmi_ffbp_leaves(const void* data, int samples, const double* pos, const double* weight,
                     const double* phase, const int* node_int, const double* node_geo,
                     const long long* offsets, void* arena, const int* blocks, int nblocks,
                     double range_start, double data_dr, double k, int mode, int interp,
                     void* stream);
int bpmi_ffbp_merges(const int* node_int, const double* node_geo, const long long* offsets,
                     void* arena, const int* blocks, int nblocks, double k, int mode, int interp,
                     void* stream);
/* Number of split-mode tiles (1024 pixels: 64 u-columns x 16 r-rows) of a
 * leaf node of nr x nu pixels. */
int bpmi_split_tiles(int nr, int nu, int* tiles);

/* Split-mode merges of one subtree height with 2-D tiles
 * (batched_resident_backend.py:72-150 merge stage, LDS-tiled).  ``blocks``
 * int32 [nblocks, 2] = (node, tile) pairs, bpmi_merge_tiles() tiles of
 * 32 r-rows x 32 u-columns per parent node (entries with node -1 are
 * skipped, e.g. padding of a table ordered for the 8 dies); tables as for
 * bpmi_ffbp_merges.
 * Bitwise equal to bpmi_ffbp_merges in BPMI_MODE_SPLIT. */
int bpmi_ffbp_merges_tiled(const int* node_int, const double* node_geo,
                           const long long* offsets, void* arena, const int* blocks,
                           int nblocks, double k, int interp, void* stream);
/* bpmi_ffbp_merges_tiled with an explicit launch shape, for the autotuner:
 * (waves, window_cap) = the occupancy hint of the launch bounds and the
 * per-child LDS window cap in samples, one of (4, 2176) (the default shape of
 * bpmi_ffbp_merges_tiled), (5, 1664), (5, 1920), (7, 1088); anything else
 * is BPMI_ERR_INVALID.  Every shape computes the same values
 * (a window above the cap is read from global memory with the same
 * arithmetic).  The reference has no counterpart: it fixed its launch shapes
 * by hand sweeps (resident_experiment.py:25-31). */
int bpmi_ffbp_merges_tiled_ex(const int* node_int, const double* node_geo,
                              const long long* offsets, void* arena, const int* blocks,
                              int nblocks, double k, int interp, int waves, int window_cap,
                              void* stream);
/* Window table of a tiled-merge block table: ``windows`` holds nblocks
 * 64-byte entries, each the LDS window bounds (int32 r0, c0, h, w) of the
 * (node, tile) entry's stencils in child a, then in child b, followed by
 * the children's reciprocal spacings (double 1/dr, 1/du of a, then of b) --
 * computed from the node geometry alone (no pulse data), exactly as each
 * merge workgroup would derive them.  Built once per block table; no
 * reference counterpart (the reference's merge gathers every tap from
 * global memory, resident_kernels.cu:35-157). */
int bpmi_ffbp_merge_windows(const int* node_int, const double* node_geo, const int* blocks,
                            int nblocks, int interp, void* windows, void* stream);
/* bpmi_ffbp_merges_tiled_ex reading each tile's entry of a
 * bpmi_ffbp_merge_windows table of the same block table instead of deriving
 * it in the workgroup (null: derive it); outputs bitwise equal.  Window
 * bounds are clipped to the child they are applied to, so a table of
 * another block table gives the same outputs, only slower; the reciprocal
 * spacings must be the children's (a table of this plan's nodes). */
int bpmi_ffbp_merges_tiled_win(const int* node_int, const double* node_geo,
                               const long long* offsets, void* arena, const int* blocks,
                               int nblocks, double k, int interp, int waves, int window_cap,
                               const void* windows, void* stream);
int bpmi_merge_tiles(int nr, int nu, int* tiles);

/* Required-tile marking for culled schedules (no reference counterpart: the
 * reference computes every pixel of every node, gpu_backend.py:121-224 sizes
 * each node to the whole root sector).  ``masks`` uint8, one byte per
 * bpmi_merge_tiles() tile of each node, node i's bytes at masks +
 * mask_offsets[i], tile = tile_row * ceil(nu / 32) + tile_col.
 * bpmi_ffbp_required_tiles: for every (node, tile) entry of ``blocks`` whose
 * byte is set, sets the bytes of both children's tiles that the split-mode
 * tiled merge of that tile reads (stencils widened by one cell).  Run it per
 * subtree height from the root down: the set bytes of every node then cover
 * all pixels that the set root tiles depend on.
 * bpmi_eval_required_tiles: sets the root bytes that bpmi_eval_grid in
 * BPMI_MODE_SPLIT reads for the Cartesian grid, for the nr x nu view at
 * (row_offset, col_offset) of an root_nr x root_nu root (geo: the view's
 * geometry, as passed to bpmi_eval_grid).  Bytes are only ever set. */
int bpmi_ffbp_required_tiles(const int* node_int, const double* node_geo,
                             const long long* mask_offsets, void* masks, const int* blocks,
                             int nblocks, int interp, void* stream);
int bpmi_eval_required_tiles(const double* geo, int nr, int nu, int row_offset, int col_offset,
                             int root_nr, int root_nu, double x0, double dx, int nx, double y0,
                             double dy, int ny, int interp, void* mask, void* stream);

/* ---- many independent FFBP patches in one launch per level -------------- */
/* Per-node input description for leaves of different acquisitions (no
 * reference counterpart: the reference batches the nodes of one image,
 * batched_resident_backend.py:72-150).  64 bytes, device array [nodes]. */
typedef struct bpmi_node_source {
    const void* data;       /* complex64 [pulses, samples] of the node's patch */
    const double* pos;      /* f64 [pulses, 3] */
    const double* weight;   /* f64 [pulses] or NULL */
    const double* phase;    /* f64 [pulses] or NULL */
    double range_start;
    double data_dr;
    double k;
    int samples;
    int reserved;
} bpmi_node_source;

/* Leaves of many patches (node tables as for bpmi_ffbp_leaves; pulse
 * indices a, b are relative to the node's own patch).  ``tile_cols_log2``
 * (3..6) sets the split-mode tile width; the block table then holds
 * (node, tile) pairs with bpmi_split_tiles_ex() tiles per node. */
int bpmi_ffbp_leaves_multi(const bpmi_node_source* sources, const int* node_int,
                           const double* node_geo, const long long* offsets, void* arena,
                           const int* blocks, int nblocks, int mode, int interp,
                           int tile_cols_log2, void* stream);
/* bpmi_ffbp_leaves_multi with the split kernel's pulse-block depth (16, 32
 * or 64 pulses staged in LDS per block; 32 is the default of
 * bpmi_ffbp_leaves_multi), for the autotuner; other values are
 * BPMI_ERR_INVALID, other modes ignore it.  The depth regroups the pulses a
 * tile accumulates per path, so outputs agree to FP32 rounding (bitwise when
 * every pulse of a tile takes the same path).  No counterpart in the
 * reference (hand-swept launch shapes, resident_experiment.py:25-31). */
int bpmi_ffbp_leaves_multi_ex(const bpmi_node_source* sources, const int* node_int,
                              const double* node_geo, const long long* offsets, void* arena,
                              const int* blocks, int nblocks, int mode, int interp,
                              int tile_cols_log2, int pulse_block, void* stream);
int bpmi_split_tiles_ex(int nr, int nu, int tile_cols_log2, int* tiles);

/* One Cartesian evaluation job per patch (64-bit fields first). */
typedef struct bpmi_eval_job {
    const void* img;        /* complex64 polar image, rows of length ld */
    void* out;              /* complex64 [ny, nx] */
    double geo[7];          /* ox, oy, oz, r0, dr, u0, du of the image */
    double x0, dx, y0, dy, k;
    int nr, nu, ld, nx, ny, reserved;
} bpmi_eval_job;
/* blocks = (job, first_pixel) pairs of 256 output pixels, device arrays. */
int bpmi_eval_grid_multi(const bpmi_eval_job* jobs, const int* blocks, int nblocks, int mode,
                         int interp, void* stream);
/* Split mode with LDS-staged stencil windows: blocks = (job, tile) pairs,
 * bpmi_eval_tiles() tiles of 64 x-columns by 4 y-rows per job grid, row-major.
 * Bitwise equal to bpmi_eval_grid_multi in BPMI_MODE_SPLIT (bpmi_eval_grid in
 * split mode runs the same tiled kernel). */
int bpmi_eval_grid_multi_tiled(const bpmi_eval_job* jobs, const int* blocks, int nblocks,
                               int interp, void* stream);
int bpmi_eval_tiles(int nx, int ny, int* tiles);

/* ---- persistent FFBP: one launch for a whole schedule ------------------- */
/* No single counterpart in the reference: it replaces the per-level launch
 * sequence of BatchedResidentProgram (batched_resident_backend.py:72-150:
 * batched leaves, one merge launch per subtree height, crop) -- and, for
 * many patches, one such sequence per patch -- by ONE launch in split
 * precision.  Workgroups claim work items from one ticket counter; a merge
 * tile starts once both children have completed all their tiles, an
 * evaluation block once its root has.  Outputs are bitwise equal to the
 * level-by-level launches (bpmi_ffbp_leaves[_multi] + bpmi_ffbp_merges_tiled
 * + bpmi_eval_grid_multi) on the same tables. */
enum { BPMI_WORK_LEAF = 0, BPMI_WORK_MERGE = 1, BPMI_WORK_EVAL = 2 };
typedef struct bpmi_persistent_schedule {
    /* node tables and arena as for bpmi_ffbp_leaves / bpmi_ffbp_merges_tiled */
    const int* node_int;
    const double* node_geo;
    const long long* offsets;
    void* arena;
    /* leaf inputs: per node (many patches), or NULL and the single patch below */
    const bpmi_node_source* sources;
    const void* data;               /* complex64 [pulses, samples] */
    const double* pos;              /* f64 [pulses, 3] */
    const double* weight;           /* f64 [pulses] or NULL */
    const double* phase;            /* f64 [pulses] or NULL */
    int samples;
    int tile_cols_log2;             /* leaf split-tile width, 3..6 */
    double range_start, data_dr;
    double k;                       /* carrier wavenumber of the merges */
    /* Cartesian evaluation jobs (may be empty) */
    const bpmi_eval_job* jobs;
    int njobs;
    /* work list: int32 [nitems, 4] = (kind, node or job, tile, 0), every
     * item after the items it depends on; an evaluation item is one
     * bpmi_eval_tiles() tile (64 x 4 pixels) of its job's grid */
    const int* items;
    int nitems;
    const int* node_tiles;          /* int32 [nodes]: items per node */
    int nodes;
    const int* job_root;            /* int32 [njobs]: root node of each job */
    const int* job_blocks;          /* int32 [njobs]: evaluation items per job */
    /* device workspace of bpmi_persistent_workspace_bytes(): the launch
     * arguments, the ticket / abort / completion counters and the time
     * stamps; written by the launch itself (a reset kernel enqueued first) */
    void* workspace;
    double max_wait_s;              /* bound of one dependency wait, seconds */
} bpmi_persistent_schedule;
int bpmi_persistent_workspace_bytes(int nodes, int njobs, size_t* bytes);
/* Workgroups that fit on the device at once for this kernel (the default
 * grid; any positive grid is correct, this one keeps every CU busy). */
int bpmi_persistent_grid(int leaf_interp, int merge_interp, int* grid);
/* Launch (grid <= 0: bpmi_persistent_grid()).  leaf_interp lagrange4 or
 * linear (BPMI_INTERP_NONE, a list without leaf items, returns
 * BPMI_ERR_UNSUPPORTED: that variant is switched off); merge_interp
 * lagrange4 or kaiser8.  Enqueues the workspace reset
 * and the launch on ``stream``; never synchronizes.  Launches that share a
 * workspace must not overlap. */
int bpmi_ffbp_persistent(const bpmi_persistent_schedule* schedule, int leaf_interp,
                         int merge_interp, int grid, void* stream);
/* After the launch has completed (synchronous device-to-host reads):
 * *aborted = 1 when a dependency wait ran out (the outputs are then
 * incomplete); stamps[0] = first workgroup start, stamps[1 + j] = completion
 * of evaluation job j, in ticks of the 100 MHz device wall clock. */
int bpmi_persistent_aborted(const void* workspace, int nodes, int njobs, int* aborted);
int bpmi_persistent_stamps(const void* workspace, int nodes, int njobs,
                           unsigned long long* stamps);

/* ---- phase-LUT direct BP (resident_precision.py:68-175) ---------------- */
/* lut[i] = exp(j k (range_start + i data_dr)), complex64, FP64 evaluation
 * (fill_range_phase_lut, precision_probe.cu:106-115). */
int bpmi_fill_range_phase_lut(int samples, double range_start, double data_dr, double k,
                              void* lut, void* stream);
/* Mixed direct BP with the LUT phase decomposition
 * (direct_chunk_mixed_phase_lut, precision_probe.cu:118-164); out += chunk. */
int bpmi_bp_polar_lut(const void* data, int samples, const double* pos, const double* weight,
                      const double* phase, int pulse_begin, int pulse_end, const double* geo,
                      int nr, int nu, int side, double range_start, double data_dr, double k,
                      const void* lut, void* out, void* stream);

/* ---- range compression and raw IQ decode (the step before BP) ---------- */
/* Matched filter of fast_bp/sar_io.py:116-139 (range_compress): each pulse
 * of raw complex64 [pulses, ld_raw] (first ``samples`` valid) minus its mean,
 * correlated with the centred LFM reference exp(j pi chirp_sign B/tp tr^2)
 * times a periodic Hamming window (L = round(tp fs) samples), through
 * zero-padded complex64 FFTs of length nfft (>= samples + L - 1; the
 * reference uses scipy.fft.next_fast_len(samples + L - 1)); oversample > 1
 * Fourier-interpolates the correlation (scipy.signal.resample) of that
 * length before the crop.  With oversample 1 the transforms use the
 * smallest 2^a m >= samples + L - 1 (m = 1, 3, 5, 9) instead, which gives
 * the same linear correlation (to FP32 rounding) faster; up to 9408 points
 * (a != 1) the whole filter then runs in one kernel with the row in LDS.  out complex64 [pulses, samples * oversample]: sample k is the pulse
 * centred at raw sample k / oversample.  FFTs run on hipFFT plans cached per
 * size (their twiddle tables are allocated when a plan is first made); the
 * FFT work area and every intermediate are in the caller's workspace of
 * bpmi_range_compress_workspace_bytes() bytes (which makes the plans). */
int bpmi_range_compress_workspace_bytes(int pulses, int samples, double tp, double fs, int nfft,
                                        int oversample, size_t* bytes);
int bpmi_range_compress(const void* raw, int pulses, int samples, long long ld_raw, double tp,
                        double fs, double bandwidth, int chirp_sign, int nfft, int oversample,
                        void* out, void* workspace, size_t workspace_bytes, void* stream);
/* Raw capture frames (fast_bp/sar_io.py:18-81 Capture.read): ``frames`` holds
 * nframes frames of frame_bytes bytes in device memory; row i of out
 * (complex64 [count, samples]) is the int16 I/Q block at byte iq_offset of
 * frame frame_index[i] (device int64 array; an index outside [0, nframes)
 * gives zeros). */
int bpmi_decode_iq16(const void* frames, long long nframes, long long frame_bytes,
                     long long iq_offset, int samples, const long long* frame_index,
                     long long count, void* out, void* stream);

/* ---- Kaiser table and probes ------------------------------------------- */
/* Upload the Kaiser-sinc kernel table used by BPMI_INTERP_KAISER8: host
 * array of BPMI_KAISER_TABLE_LEN doubles, the kernel at x = i/2048.  FP64
 * paths read it as is, the others its FP32 rounding (linear interpolation
 * between entries, 8 weights normalised to unit sum).  Must be called once
 * before any kaiser8 launch. */
#define BPMI_KAISER_TABLE_LEN (4 * 2048 + 2)
int bpmi_set_kaiser_table(const double* host_table, int len);
/* sin/cos(2*pi*t) with the hardware transcendental path (device arrays). */
int bpmi_probe_sincos_turns(const float* t, float* s, float* c, long long n, void* stream);
/* FP32 FMA issue-rate probe (measures the vector FP32 peak this build can
 * reach; packed FP32 is not used): ``blocks`` workgroups of 256 threads, each
 * thread 8 independent FMA chains of iters x 16 steps, one float stored per
 * thread into out[blocks * 256].  FLOPs = blocks * 256 * iters * 256. */
int bpmi_probe_fma_peak(float* out, int blocks, int iters, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BPMI_H */
