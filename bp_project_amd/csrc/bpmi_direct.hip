// Direct time-domain backprojection kernels for gfx950.
//
// Two families:
//
// 1. Per-pixel kernels (one thread per output pixel, pulses read from HBM
//    through the caches) for the reference precision policies FP64, FP32,
//    MIXED and the double-float DFLOAT study mode.  They restate
//    bp_polar / direct_chunk_{mode} (fast_bp/gpu/kernels.cu:139-339) and
//    backproject_points (fast_bp/ffbp.py:43-55).
//
// 2. The SPLIT kernel: a workgroup owns a 1024-pixel tile; pulses are
//    processed in blocks of 32 whose per-pulse geometry is reduced once per
//    tile in FP64 and whose range-window slices are staged in LDS as cubic
//    (or linear) polynomial coefficients pre-multiplied by the pulse weight.
//    Per pixel and pulse the residual geometry runs in FP32, the large linear
//    phase term and its reduction modulo one turn in FP64, the sine/cosine on
//    the FP32 hardware transcendental path, interpolation and accumulation in
//    FP32.  Pulses whose window or expansion error would exceed the bounds
//    fall back, per tile and pulse, to an exact FP64-geometry path.
#include "bpmi_split.h"

#include <mutex>

namespace bpmi {


// ---------------------------------------------------------------------------
// Host-side dispatch
// ---------------------------------------------------------------------------
extern const float* kaiser_table_device();

static int validate_common(const void* data, int samples, const double* pos, int pb, int pe,
                           int mode, int interp, int flags, void* out) {
    if (!data || !pos || !out) return set_error(BPMI_ERR_INVALID, "data, pos and out must be non-null");
    if (samples <= 0) return set_error(BPMI_ERR_INVALID, "samples must be positive");
    if (pb < 0 || pe < pb) return set_error(BPMI_ERR_INVALID, "invalid pulse range [%d, %d)", pb, pe);
    if (mode < 0 || mode > BPMI_MODE_SPLIT)
        return set_error(BPMI_ERR_INVALID, "mode must be one of fp64, fp32, mixed, dfloat, split");
    if (interp < 0 || interp > BPMI_INTERP_KAISER8) return set_error(BPMI_ERR_INVALID, "unknown interpolator");
    if ((flags & BPMI_OUT_C128) && mode != BPMI_MODE_FP64)
        return set_error(BPMI_ERR_INVALID, "complex128 output is only defined for fp64 mode");
    if (interp == BPMI_INTERP_KAISER8 && !kaiser_table_device())
        return set_error(BPMI_ERR_INVALID, "kaiser8 needs bpmi_set_kaiser_table() first");
    return BPMI_OK;
}

template <int MODE, int SRC>
static void launch_pixel_interp(int interp, dim3 g, hipStream_t st, const BpArgs& a, const SrcArgs& s) {
    const float* kt = kaiser_table_device();
    if (interp == BPMI_INTERP_LAGRANGE4)
        hipLaunchKernelGGL((bp_pixel_kernel<MODE, INTERP_LAGRANGE4, SRC>), g, dim3(256), 0, st, a, s, kt);
    else if (interp == BPMI_INTERP_LINEAR)
        hipLaunchKernelGGL((bp_pixel_kernel<MODE, INTERP_LINEAR, SRC>), g, dim3(256), 0, st, a, s, kt);
    else
        hipLaunchKernelGGL((bp_pixel_kernel<MODE, INTERP_KAISER8, SRC>), g, dim3(256), 0, st, a, s, kt);
}

template <int SRC>
static void launch_pixel(int mode, int interp, long long count, hipStream_t st, const BpArgs& a,
                         const SrcArgs& s) {
    const dim3 g(grid_1d(count, 256));
    switch (mode) {
        case BPMI_MODE_FP64: launch_pixel_interp<MODE_FP64, SRC>(interp, g, st, a, s); break;
        case BPMI_MODE_FP32: launch_pixel_interp<MODE_FP32, SRC>(interp, g, st, a, s); break;
        case BPMI_MODE_MIXED: launch_pixel_interp<MODE_MIXED, SRC>(interp, g, st, a, s); break;
        default: launch_pixel_interp<MODE_DFLOAT, SRC>(interp, g, st, a, s); break;
    }
}

// Tile shape for the SPLIT kernel: log2 of the tile width in columns
// (8..64, the tile holds 1024 pixels).  The LDS window of a tile and pulse
// spans ~2B/data_dr samples, B = |u_col| h_col + |u_row| h_row the
// projection of the tile's half extents onto the look direction; ``col_look``
// and ``row_look`` are the typical magnitudes of the look-vector components
// along the column and row axes.  The shape minimising B is chosen.
static int choose_tc_log2(double col_step_m, double row_step_m, double col_look, double row_look) {
    int best = 6;
    double best_b = 1e300;
    for (int l = 3; l <= 6; ++l) {
        const double hc = 0.5 * (double)(1 << l) * col_step_m;
        const double hr = 0.5 * (double)(SP_TILE >> l) * row_step_m;
        const double b = col_look * hc + row_look * hr;
        if (b < best_b * 0.999) {
            best_b = b;
            best = l;
        }
    }
    return best;
}

template <int SRC, bool ORIGIN_REF>
static int launch_split(int interp, long long tiles, int tc_log2, hipStream_t st, const BpArgs& a,
                        const SrcArgs& s, LeafTables lt = LeafTables{}, int chunks = 1) {
    if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
        return set_error(BPMI_ERR_UNSUPPORTED, "split mode supports lagrange4 and linear interpolation");
    // Grids are bounded by the launch limit; split very large images.
    const long long max_blocks = 1LL << 30;
    for (long long base = 0; base < tiles; base += max_blocks) {
        const unsigned nb = (unsigned)(tiles - base < max_blocks ? tiles - base : max_blocks);
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LAGRANGE4, SRC, ORIGIN_REF>),
                               dim3(nb, chunks), dim3(SP_THREADS), 0, st, a, s, tc_log2,
                               (int)base, lt);
        else
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LINEAR, SRC, ORIGIN_REF>),
                               dim3(nb, chunks), dim3(SP_THREADS), 0, st, a, s, tc_log2,
                               (int)base, lt);
    }
    return check_launch("bp_split_kernel");
}

static long long tiles_for(int rows, int cols, int tc_log2) {
    const int tc = 1 << tc_log2, tr = SP_TILE >> tc_log2;
    return (long long)((cols + tc - 1) / tc) * (long long)((rows + tr - 1) / tr);
}

// Pulse-split direct BP of small grids.  A grid of few 1024-pixel tiles
// leaves most of the 256 CUs idle (64 tiles at 256^2); splitting the pulses
// into chunks, each tile and chunk its own workgroup writing a partial image,
// fills the chip, and one pass sums the partials in chunk order (the result
// does not depend on scheduling).  Chunks keep at least SPLIT_MIN_PULSES
// pulses so the per-tile set-up stays amortised.  Grids below 4096 tiles
// split: 1024 tiles (1024^2) make one round of 4 workgroups per CU whose
// slowest workgroup sets the time; in 4 chunks the same forward takes 3.36
// instead of 3.87 ms (4096 pulses, 1024^2 at 0.5 m), in 8 chunks the
// autofocus gradient step (forward and adjoint) 9.37 instead of 9.67 ms.
// 32-pulse chunks were slower on 256^2 over 1024 pulses (0.079 vs 0.074 ms).
constexpr long long SPLIT_TARGET_BLOCKS = 8192;
constexpr int SPLIT_MIN_PULSES = 64;

static int pulse_chunks(long long tiles, int pulses) {
    if (tiles * 2 >= SPLIT_TARGET_BLOCKS) return 1;
    long long c = (SPLIT_TARGET_BLOCKS + tiles - 1) / tiles;
    c = std::min(c, (long long)(pulses / SPLIT_MIN_PULSES));
    return c < 2 ? 1 : (int)c;
}

// Partial images of bpmi_bp_grid_noalloc.  Each GPU the library runs on
// holds its own copy of this array until the process ends; host code never
// passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol), it uses the device address that grid_scratch_address()
// resolves once per device.  Size: a grid splits its pulses only when it has
// T < 4096 tiles of 1024 pixels, into at most ceil(8192 / T) chunks, so its
// partials take at most (8192 + T - 1) * 1024 * 8 bytes < 96 MiB; 96 MiB
// therefore holds every grid's partials (256 x 256 over 1024 pulses: 16
// partials, 8 MiB; 1024 x 1024 over 4096 pulses: 8 partials, 64 MiB), and
// the size check in bpmi_bp_grid_noalloc guards that bound.  Every chunk
// overwrites all pixels of its partial, so nothing in it has to start at
// zero.
constexpr size_t GRID_SCRATCH_BYTES = (size_t)96 << 20;
static __device__ __align__(256) unsigned char scratch[GRID_SCRATCH_BYTES];

constexpr int GRID_SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[GRID_SCRATCH_MAX_DEVICES] = {};

// Device address of ``scratch`` on the current device (resolved on the
// first call per device, with that device current).
static int grid_scratch_address(void** ptr) {
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st) return st;
    if (dev < 0 || dev >= GRID_SCRATCH_MAX_DEVICES)
        return set_error(BPMI_ERR_UNSUPPORTED, "device index %d beyond the %d devices with scratch",
                         dev, GRID_SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        st = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (st) return st;
        if (!p) return set_error(BPMI_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return BPMI_OK;
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const float2* __restrict__ part,
                                                           int parts, long long n, float2* out,
                                                           int accumulate) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 v = accumulate ? out[i] : make_float2(0.f, 0.f);
    for (int c = 0; c < parts; ++c) {
        const float2 p = part[(long long)c * n + i];
        v.x += p.x;
        v.y += p.y;
    }
    out[i] = v;
}

}  // namespace bpmi

using namespace bpmi;

extern "C" int bpmi_bp_polar_ex(const void* data, int samples, const double* pos,
                                const double* weight, const double* phase, int pulse_begin,
                                int pulse_end, const double* geo, int nr, int nu, int side,
                                double range_start, double data_dr, double k, int mode,
                                int interp, int flags, void* out, void* stream, int origin_ref,
                                int tile_cols_log2);

extern "C" int bpmi_bp_polar(const void* data, int samples, const double* pos, const double* weight,
                             const double* phase, int pulse_begin, int pulse_end, const double* geo,
                             int nr, int nu, int side, double range_start, double data_dr, double k,
                             int mode, int interp, int flags, void* out, void* stream) {
    return bpmi_bp_polar_ex(data, samples, pos, weight, phase, pulse_begin, pulse_end, geo, nr,
                            nu, side, range_start, data_dr, k, mode, interp, flags, out, stream,
                            0, -1);
}

extern "C" int bpmi_bp_polar_ex(const void* data, int samples, const double* pos,
                                const double* weight, const double* phase, int pulse_begin,
                                int pulse_end, const double* geo, int nr, int nu, int side,
                                double range_start, double data_dr, double k, int mode,
                                int interp, int flags, void* out, void* stream, int origin_ref,
                                int tile_cols_log2) {
    int st = validate_common(data, samples, pos, pulse_begin, pulse_end, mode, interp, flags, out);
    if (st) return st;
    if (!geo) return set_error(BPMI_ERR_INVALID, "geo must be non-null");
    if (nr <= 0 || nu <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    BpArgs a{(const float2*)data, pos, weight, phase, samples, pulse_begin, pulse_end,
             range_start, data_dr, k, flags, out};
    SrcArgs s{geo[0], geo[1], geo[2], geo[3], geo[4], geo[5], geo[6], nr, nu, side, nullptr, 0};
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (flags & BPMI_OUT_C128) return set_error(BPMI_ERR_INVALID, "split mode writes complex64");
        if (tile_cols_log2 != -1 && (tile_cols_log2 < 3 || tile_cols_log2 > 6))
            return set_error(BPMI_ERR_INVALID, "tile_cols_log2 must be -1 or 3..6");
        if (pulse_end == pulse_begin) {
            if (!(flags & BPMI_ACCUMULATE)) return bpmi_zero(out, (size_t)nr * nu * 8, stream);
            return BPMI_OK;
        }
        // rows = slant range (along the look direction), cols = u (across it)
        const double rmid = fabs(geo[3] + geo[4] * nr * 0.5);
        const int tcl = tile_cols_log2 >= 0 ? tile_cols_log2
                      : (origin_ref ? 6 : choose_tc_log2(fabs(geo[6]) * rmid, fabs(geo[4]), 0.3, 1.0));
        const long long tiles = tiles_for(nr, nu, tcl);
        if (origin_ref) return launch_split<SRC_POLAR, true>(interp, tiles, tcl, hs, a, s);
        return launch_split<SRC_POLAR, false>(interp, tiles, tcl, hs, a, s);
    }
    launch_pixel<SRC_POLAR>(mode, interp, (long long)nr * nu, hs, a, s);
    return check_launch("bp_pixel_kernel(polar)");
}

extern "C" int bpmi_bp_grid_ws(const void* data, int samples, const double* pos,
                               const double* weight, const double* phase, int pulse_begin,
                               int pulse_end, double x0, double dx, int nx, double y0, double dy,
                               int ny, double z0, double range_start, double data_dr, double k,
                               int mode, int interp, int flags, void* out, void* workspace,
                               size_t workspace_bytes, void* stream);

extern "C" int bpmi_bp_grid(const void* data, int samples, const double* pos, const double* weight,
                            const double* phase, int pulse_begin, int pulse_end, double x0,
                            double dx, int nx, double y0, double dy, int ny, double z0,
                            double range_start, double data_dr, double k, int mode, int interp,
                            int flags, void* out, void* stream) {
    return bpmi_bp_grid_ws(data, samples, pos, weight, phase, pulse_begin, pulse_end, x0, dx, nx,
                           y0, dy, ny, z0, range_start, data_dr, k, mode, interp, flags, out,
                           nullptr, 0, stream);
}

extern "C" int bpmi_bp_grid_workspace_bytes(int nx, int ny, int pulses, size_t* bytes) {
    if (!bytes) return set_error(BPMI_ERR_INVALID, "bytes must be non-null");
    if (nx <= 0 || ny <= 0 || pulses < 0) return set_error(BPMI_ERR_INVALID, "invalid grid or pulse count");
    // the tile shape depends on the spacings: the largest chunk count over the shapes
    int chunks = 1;
    for (int l = 3; l <= 6; ++l) chunks = std::max(chunks, pulse_chunks(tiles_for(ny, nx, l), pulses));
    *bytes = chunks > 1 ? (size_t)chunks * (size_t)nx * (size_t)ny * sizeof(float2) : 0;
    return BPMI_OK;
}

extern "C" int bpmi_bp_grid_ws(const void* data, int samples, const double* pos,
                               const double* weight, const double* phase, int pulse_begin,
                               int pulse_end, double x0, double dx, int nx, double y0, double dy,
                               int ny, double z0, double range_start, double data_dr, double k,
                               int mode, int interp, int flags, void* out, void* workspace,
                               size_t workspace_bytes, void* stream) {
    int st = validate_common(data, samples, pos, pulse_begin, pulse_end, mode, interp, flags, out);
    if (st) return st;
    if (nx <= 0 || ny <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    BpArgs a{(const float2*)data, pos, weight, phase, samples, pulse_begin, pulse_end,
             range_start, data_dr, k, flags, out};
    SrcArgs s{0.0, 0.0, z0, x0, dx, y0, dy, ny, nx, 1, nullptr, 0};
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (pulse_end == pulse_begin) {
            if (!(flags & BPMI_ACCUMULATE)) return bpmi_zero(out, (size_t)nx * ny * 8, stream);
            return BPMI_OK;
        }
        // cols = x (cross-track, along the look direction), rows = y
        const int tcl = choose_tc_log2(fabs(dx), fabs(dy), 1.0, 0.3);
        const long long tiles = tiles_for(ny, nx, tcl);
        const long long npix = (long long)nx * ny;
        const int pulses = pulse_end - pulse_begin;
        int chunks = pulse_chunks(tiles, pulses);
        if (workspace) {
            const size_t fit = workspace_bytes / ((size_t)npix * sizeof(float2));
            if ((size_t)chunks > fit) chunks = (int)fit;
        } else {
            chunks = 1;
        }
        if (chunks < 2 || (flags & BPMI_OUT_C128))
            return launch_split<SRC_GRID, false>(interp, tiles, tcl, hs, a, s);
        // equal chunks, none empty
        const int chunk = (pulses + chunks - 1) / chunks;
        chunks = (pulses + chunk - 1) / chunk;
        BpArgs ap = a;
        ap.chunk = chunk;
        ap.partial = workspace;
        st = launch_split<SRC_GRID, false>(interp, tiles, tcl, hs, ap, s, LeafTables{}, chunks);
        if (st) return st;
        hipLaunchKernelGGL(sum_partials_kernel, dim3(grid_1d(npix, 256)), dim3(256), 0, hs,
                           (const float2*)workspace, chunks, npix, (float2*)out,
                           (flags & BPMI_ACCUMULATE) ? 1 : 0);
        return check_launch("sum_partials_kernel");
    }
    launch_pixel<SRC_GRID>(mode, interp, (long long)nx * ny, hs, a, s);
    return check_launch("bp_pixel_kernel(grid)");
}

extern "C" int bpmi_bp_grid_noalloc(const void* data, int samples, const double* pos,
                                    const double* weight, const double* phase, int pulse_begin,
                                    int pulse_end, double x0, double dx, int nx, double y0,
                                    double dy, int ny, double z0, double range_start,
                                    double data_dr, double k, int mode, int interp, int flags,
                                    void* out, void* stream) {
    int st = validate_common(data, samples, pos, pulse_begin, pulse_end, mode, interp, flags, out);
    if (st) return st;
    if (nx <= 0 || ny <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    size_t need = 0;
    st = bpmi_bp_grid_workspace_bytes(nx, ny, pulse_end - pulse_begin, &need);
    if (st) return st;
    if (mode != BPMI_MODE_SPLIT) need = 0;
    if (need > GRID_SCRATCH_BYTES)
        return set_error(BPMI_ERR_INVALID,
                         "bpmi_bp_grid_noalloc: a %d x %d grid over %d pulses needs %zu bytes of "
                         "partial images, above the %zu-byte static scratch (use bpmi_bp_grid_ws)",
                         nx, ny, pulse_end - pulse_begin, need, GRID_SCRATCH_BYTES);
    void* ws = nullptr;
    if (need > 0) {
        st = grid_scratch_address(&ws);
        if (st) return st;
    }
    return bpmi_bp_grid_ws(data, samples, pos, weight, phase, pulse_begin, pulse_end, x0, dx, nx,
                           y0, dy, ny, z0, range_start, data_dr, k, mode, interp, flags, out, ws,
                           need, stream);
}

extern "C" int bpmi_bp_points(const void* data, int samples, const double* pos,
                              const double* weight, const double* phase, int pulse_begin,
                              int pulse_end, const double* xyz, long long npoints,
                              double range_start, double data_dr, double k, int mode, int interp,
                              int flags, void* out, void* stream) {
    int st = validate_common(data, samples, pos, pulse_begin, pulse_end, mode, interp, flags, out);
    if (st) return st;
    if (npoints <= 0) return BPMI_OK;
    if (!xyz) return set_error(BPMI_ERR_INVALID, "xyz must be non-null");
    BpArgs a{(const float2*)data, pos, weight, phase, samples, pulse_begin, pulse_end,
             range_start, data_dr, k, flags, out};
    SrcArgs s{0, 0, 0, 0, 0, 0, 0, 0, 0, 1, xyz, npoints};
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (pulse_end == pulse_begin) {
            if (!(flags & BPMI_ACCUMULATE)) return bpmi_zero(out, (size_t)npoints * 8, stream);
            return BPMI_OK;
        }
        return launch_split<SRC_POINTS, false>(interp, (npoints + SP_TILE - 1) / SP_TILE, 0, hs, a, s);
    }
    launch_pixel<SRC_POINTS>(mode, interp, npoints, hs, a, s);
    return check_launch("bp_pixel_kernel(points)");
}

extern "C" int bpmi_split_tiles_ex(int nr, int nu, int tile_cols_log2, int* tiles) {
    if (!tiles) return set_error(BPMI_ERR_INVALID, "tiles is null");
    if (nr <= 0 || nu <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    if (tile_cols_log2 < 3 || tile_cols_log2 > 6)
        return set_error(BPMI_ERR_INVALID, "tile_cols_log2 must be 3..6");
    const long long t = tiles_for(nr, nu, tile_cols_log2);
    if (t > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "node too large");
    *tiles = (int)t;
    return BPMI_OK;
}

extern "C" int bpmi_ffbp_leaves_multi(const bpmi_node_source* sources, const int* node_int,
                                      const double* node_geo, const long long* offsets,
                                      void* arena, const int* blocks, int nblocks, int mode,
                                      int interp, int tile_cols_log2, void* stream) {
    return bpmi_ffbp_leaves_multi_ex(sources, node_int, node_geo, offsets, arena, blocks, nblocks,
                                     mode, interp, tile_cols_log2, SP_NB, stream);
}

extern "C" int bpmi_ffbp_leaves_multi_ex(const bpmi_node_source* sources, const int* node_int,
                                         const double* node_geo, const long long* offsets,
                                         void* arena, const int* blocks, int nblocks, int mode,
                                         int interp, int tile_cols_log2, int pulse_block,
                                         void* stream) {
    if (pulse_block != 16 && pulse_block != 32 && pulse_block != 64)
        return set_error(BPMI_ERR_INVALID, "pulse_block must be 16, 32 or 64");
    if (nblocks <= 0) return BPMI_OK;
    if (!sources || !node_int || !node_geo || !offsets || !arena || !blocks)
        return set_error(BPMI_ERR_INVALID, "null table");
    if (mode < 0 || mode > BPMI_MODE_SPLIT)
        return set_error(BPMI_ERR_INVALID, "mode must be one of fp64, fp32, mixed, dfloat, split");
    if (interp < 0 || interp > BPMI_INTERP_KAISER8) return set_error(BPMI_ERR_INVALID, "unknown interpolator");
    if (interp == BPMI_INTERP_KAISER8 && !kaiser_table_device())
        return set_error(BPMI_ERR_INVALID, "kaiser8 needs bpmi_set_kaiser_table() first");
    BpArgs a{nullptr, nullptr, nullptr, nullptr, 0, 0, 0, 0.0, 1.0, 0.0, 0, nullptr};
    LeafTables lt{node_int, node_geo, offsets, (float2*)arena, blocks, sources};
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (tile_cols_log2 < 3 || tile_cols_log2 > 6)
            return set_error(BPMI_ERR_INVALID, "tile_cols_log2 must be 3..6");
        if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
            return set_error(BPMI_ERR_UNSUPPORTED, "split mode supports lagrange4 and linear interpolation");
        SrcArgs s{};
        auto launch = [&]<int NB>() {
            if (interp == BPMI_INTERP_LAGRANGE4)
                hipLaunchKernelGGL((bp_split_kernel<INTERP_LAGRANGE4, SRC_POLAR, true, NB>),
                                   dim3(nblocks), dim3(SP_THREADS), 0, hs, a, s, tile_cols_log2,
                                   0, lt);
            else
                hipLaunchKernelGGL((bp_split_kernel<INTERP_LINEAR, SRC_POLAR, true, NB>),
                                   dim3(nblocks), dim3(SP_THREADS), 0, hs, a, s, tile_cols_log2,
                                   0, lt);
        };
        if (pulse_block == 16)
            launch.template operator()<16>();
        else if (pulse_block == 64)
            launch.template operator()<64>();
        else
            launch.template operator()<32>();
        return check_launch("bp_split_kernel(multi-patch leaves)");
    }
    const float* kt = kaiser_table_device();
    const dim3 g(nblocks);
    auto l = [&]<int M>() {
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((batched_leaf_kernel<M, INTERP_LAGRANGE4>), g, dim3(256), 0, hs, a, lt, kt);
        else if (interp == BPMI_INTERP_LINEAR)
            hipLaunchKernelGGL((batched_leaf_kernel<M, INTERP_LINEAR>), g, dim3(256), 0, hs, a, lt, kt);
        else
            hipLaunchKernelGGL((batched_leaf_kernel<M, INTERP_KAISER8>), g, dim3(256), 0, hs, a, lt, kt);
    };
    switch (mode) {
        case BPMI_MODE_FP64: l.template operator()<MODE_FP64>(); break;
        case BPMI_MODE_FP32: l.template operator()<MODE_FP32>(); break;
        case BPMI_MODE_MIXED: l.template operator()<MODE_MIXED>(); break;
        default: l.template operator()<MODE_DFLOAT>(); break;
    }
    return check_launch("batched_leaf_kernel(multi-patch)");
}

extern "C" int bpmi_split_tiles(int nr, int nu, int* tiles) {
    if (!tiles) return set_error(BPMI_ERR_INVALID, "tiles is null");
    if (nr <= 0 || nu <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    const long long t = tiles_for(nr, nu, 6);
    if (t > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "node too large");
    *tiles = (int)t;
    return BPMI_OK;
}

extern "C" int bpmi_ffbp_leaves(const void* data, int samples, const double* pos,
                                const double* weight, const double* phase, const int* node_int,
                                const double* node_geo, const long long* offsets, void* arena,
                                const int* blocks, int nblocks, double range_start,
                                double data_dr, double k, int mode, int interp, void* stream) {
    int st = validate_common(data, samples, pos, 0, 0, mode, interp, 0, arena);
    if (st) return st;
    if (nblocks <= 0) return BPMI_OK;
    if (!node_int || !node_geo || !offsets || !blocks)
        return set_error(BPMI_ERR_INVALID, "null table");
    BpArgs a{(const float2*)data, pos, weight, phase, samples, 0, 0, range_start, data_dr, k, 0,
             nullptr};
    LeafTables lt{node_int, node_geo, offsets, (float2*)arena, blocks, nullptr};
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        // blocks = (node, tile) pairs; tiles are 64 u-columns x 16 r-rows and
        // the node origin is the geometry reference (leaf sub-apertures are short).
        SrcArgs s{};
        if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
            return set_error(BPMI_ERR_UNSUPPORTED, "split mode supports lagrange4 and linear interpolation");
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LAGRANGE4, SRC_POLAR, true>), dim3(nblocks),
                               dim3(SP_THREADS), 0, hs, a, s, 6, 0, lt);
        else
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LINEAR, SRC_POLAR, true>), dim3(nblocks),
                               dim3(SP_THREADS), 0, hs, a, s, 6, 0, lt);
        return check_launch("bp_split_kernel(batched leaves)");
    }
    const float* kt = kaiser_table_device();
    const dim3 g(nblocks);
    auto l = [&]<int M>() {
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((batched_leaf_kernel<M, INTERP_LAGRANGE4>), g, dim3(256), 0, hs, a, lt, kt);
        else if (interp == BPMI_INTERP_LINEAR)
            hipLaunchKernelGGL((batched_leaf_kernel<M, INTERP_LINEAR>), g, dim3(256), 0, hs, a, lt, kt);
        else
            hipLaunchKernelGGL((batched_leaf_kernel<M, INTERP_KAISER8>), g, dim3(256), 0, hs, a, lt, kt);
    };
    switch (mode) {
        case BPMI_MODE_FP64: l.template operator()<MODE_FP64>(); break;
        case BPMI_MODE_FP32: l.template operator()<MODE_FP32>(); break;
        case BPMI_MODE_MIXED: l.template operator()<MODE_MIXED>(); break;
        default: l.template operator()<MODE_DFLOAT>(); break;
    }
    return check_launch("batched_leaf_kernel");
}

// ---------------------------------------------------------------------------
// Adjoint (image -> per-pulse gradients) of the Cartesian / point BP
// ---------------------------------------------------------------------------
// Forward model: I(x) = sum_n w_n s_n(q_n(x)) exp(j (k d_n(x) + phase_n)),
// d_n = |x - p_n|, q_n = (d_n - range_start) / data_dr.  Given g = dL/dI in
// PyTorch's convention (dL = Re sum_x conj(g_x) dI_x), per pulse n:
//   dL/dphase_n = Re sum_x conj(g_x) j c_xn,          c_xn = w s e^{j theta}
//   dL/dp_n     = Re sum_x conj(g_x) w e^{j theta} (s'/data_dr + j k s) (p_n - x)/d
// Range and phase are recomputed (FP64 geometry, FP32 interpolation and
// products); nothing is stored per (pixel, pulse).  Each workgroup reduces
// its 1024 pixels per pulse into partials[tile][pulse][5] =
// (Re A, Im A, Gx, Gy, Gz) with A = sum conj(g) c; a second kernel sums the
// tiles in a fixed order in FP64, so results are bitwise reproducible.
constexpr int ADJ_THREADS = 256;
constexpr int ADJ_PPT = 4;
constexpr int ADJ_TILE = ADJ_THREADS * ADJ_PPT;

template <int INTERP>
__device__ __forceinline__ void sample1_with_derivative(const float2* __restrict__ row, int len,
                                                        int i, float t, float2* v, float2* dv) {
    *v = make_float2(0.f, 0.f);
    *dv = make_float2(0.f, 0.f);
    if (!InterpShape<INTERP>::valid(i, len)) return;
    if (INTERP == INTERP_LINEAR) {
        const float2 a = row[i], b = row[i + 1];
        v->x = a.x + t * (b.x - a.x);
        v->y = a.y + t * (b.y - a.y);
        dv->x = b.x - a.x;
        dv->y = b.y - a.y;
    } else {
        float w[4], dw[4];
        cubic_weights(t, w);
        const float t2 = t * t;
        dw[0] = -(3.f * t2 - 6.f * t + 2.f) * (1.f / 6.f);
        dw[1] = (3.f * t2 - 4.f * t - 1.f) * 0.5f;
        dw[2] = -(3.f * t2 - 2.f * t - 2.f) * 0.5f;
        dw[3] = (3.f * t2 - 1.f) * (1.f / 6.f);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float2 s = row[i + a - 1];
            v->x += w[a] * s.x;
            v->y += w[a] * s.y;
            dv->x += dw[a] * s.x;
            dv->y += dw[a] * s.y;
        }
    }
}

// Per pixel and pulse: FP64 range from the hardware reciprocal square root
// with one Newton step (which also gives 1/d), the sample coordinate by a
// multiplication with 1/data_dr, the phase reduced with v_fract_f64.  Per
// pulse the workgroup needs five sums over its 1024 pixels; rather than five
// cross-lane reductions per pulse, each thread keeps the sums of its four
// pixels for ADJ_NB pulses in registers and the workgroup reduces them
// through LDS once per ADJ_NB pulses (each of the ADJ_NB * 5 rows summed by
// four threads in a fixed order, then across those four lanes).
constexpr int ADJ_NB = 4;
constexpr int ADJ_VALS = 5;
constexpr int ADJ_ROWS = ADJ_NB * ADJ_VALS;

template <int INTERP, int SRC>
__global__ __launch_bounds__(ADJ_THREADS) void bp_adjoint_kernel(BpArgs a, SrcArgs s,
                                                                 const float2* __restrict__ grad,
                                                                 float* __restrict__ partials) {
    static_assert(ADJ_ROWS * 4 <= ADJ_THREADS, "four reducing threads per row");
    __shared__ float red[ADJ_ROWS][ADJ_THREADS + 1];
    const int t = threadIdx.x;
    const long long base = (long long)blockIdx.x * ADJ_TILE;
    const long long count = src_count(s, SRC);
    const int S = a.samples;
    const int P = a.pulse_end - a.pulse_begin;
    const double kt = a.k * INV_TWO_PI;
    const double inv_ddr = 1.0 / a.data_dr;
    const float k32 = (float)a.k;
    const float invdr = (float)inv_ddr;
    double px[ADJ_PPT], py[ADJ_PPT], pz[ADJ_PPT];
    float2 g[ADJ_PPT];
    bool val[ADJ_PPT];
#pragma unroll
    for (int k = 0; k < ADJ_PPT; ++k) {
        const long long p = base + t + (long long)k * ADJ_THREADS;
        val[k] = p < count;
        double r;
        src_pixel64<SRC>(s, val[k] ? p : 0, &px[k], &py[k], &pz[k], &r);
        g[k] = val[k] ? grad[p] : make_float2(0.f, 0.f);
    }
    for (int n0 = a.pulse_begin; n0 < a.pulse_end; n0 += ADJ_NB) {
        const int nb = min(ADJ_NB, a.pulse_end - n0);
        float acc[ADJ_NB][ADJ_VALS];
#pragma unroll
        for (int j = 0; j < ADJ_NB; ++j) {
#pragma unroll
            for (int v = 0; v < ADJ_VALS; ++v) acc[j][v] = 0.f;
            if (j >= nb) continue;
            const int n = n0 + j;
            const double ax = a.pos[3 * n], ay = a.pos[3 * n + 1], az = a.pos[3 * n + 2];
            const float2* row = a.data + (size_t)n * S;
            const double ph = pulse_phase(a.phase, n) * INV_TWO_PI;
            const float w = (float)pulse_weight(a.weight, n);
#pragma unroll
            for (int k = 0; k < ADJ_PPT; ++k) {
                if (!val[k]) continue;
                const double ex = px[k] - ax, ey = py[k] - ay, ez = pz[k] - az;
                const double s2 = ex * ex + ey * ey + ez * ez;
                const double rinv = rsqrt64(s2);
                const double d = s2 * rinv;     // NaN for s2 = 0: skipped below
                const double q = (d - a.range_start) * inv_ddr;
                int i;
                if (!floor_index(q, &i)) continue;
                float2 v, dv;
                sample1_with_derivative<INTERP>(row, S, i, (float)(q - (double)i), &v, &dv);
                float sn, cs;
                sincos_turns((float)hw_fract64(kt * d + ph), &sn, &cs);
                // conj(g) * w e^{j theta}
                const float hr = w * (g[k].x * cs + g[k].y * sn);
                const float hi = w * (g[k].x * sn - g[k].y * cs);
                // A += conj(g) w e^{j theta} v
                acc[j][0] += hr * v.x - hi * v.y;
                acc[j][1] += hr * v.y + hi * v.x;
                // b = conj(g) w e^{j theta} (dv/dr + j k v) / d ; only Re(b) is needed
                const float ur = dv.x * invdr - k32 * v.y;
                const float ui = dv.y * invdr + k32 * v.x;
                const float br = (hr * ur - hi * ui) * (float)rinv;
                // (p_n - x) = -(ex, ey, ez)
                acc[j][2] -= br * (float)ex;
                acc[j][3] -= br * (float)ey;
                acc[j][4] -= br * (float)ez;
            }
        }
#pragma unroll
        for (int j = 0; j < ADJ_NB; ++j)
#pragma unroll
            for (int v = 0; v < ADJ_VALS; ++v) red[j * ADJ_VALS + v][t] = acc[j][v];
        __syncthreads();
        if (t < 4 * ADJ_ROWS) {
            const int r = t >> 2, part = t & 3;
            float sum = 0.f;
            for (int c = part; c < ADJ_THREADS; c += 4) sum += red[r][c];
            sum += __shfl_xor(sum, 1, 4);
            sum += __shfl_xor(sum, 2, 4);
            const int j = r / ADJ_VALS;
            if (part == 0 && j < nb)
                partials[((size_t)blockIdx.x * P + (n0 + j - a.pulse_begin)) * ADJ_VALS +
                         (r - j * ADJ_VALS)] = sum;
        }
        __syncthreads();
    }
}

// ---- split-precision adjoint of the grid BP ---------------------------------
// The SPLIT forward's decomposition applied to the adjoint: a workgroup owns
// a 2-D tile of 1024 grid pixels; per block of ADJ_SNB pulses the FP64 tile
// reference and per-pulse parameters (split_pulse_params) are computed once
// and the range windows staged in LDS as weighted cubic (or line)
// coefficients (stage_pulse_block).  Per pixel and pulse: the FP64 linear
// range term and its phase modulo one turn, the FP32 range residual, the
// staged polynomial's value v and derivative dv at the sample coordinate,
// and the gradient products of bp_adjoint_kernel, with 1/d from the tile's
// 1/D by a first-order expansion in (d - D)/D (|d - D| <= tile radius rho,
// relative error below (rho/D)^2).  Pulses whose bound fails (PF_SLOW) take
// bp_adjoint_kernel's exact FP64 per-pixel path.  Per ADJ_NB pulses the
// workgroup reduces its sums through LDS and writes partials[tile][pulse][5]
// in bp_adjoint_kernel's layout.
constexpr int ADJ_SNB = 16;

template <int INTERP>
struct AdjSplitLds {
    float4 coef[ADJ_SNB * SP_W * SlotSize<INTERP>::F4];
    PulseParams prm[ADJ_SNB];
    float4 off[ADJ_SNB];  // (c - a) of each pulse in FP32
    TileRef ref;
    double red8[4][8];
    float4 red[ADJ_ROWS][ADJ_THREADS / 4 + 1];  // rows 16-byte aligned
};

// Workgroup (tile, chunk): the pulses [pulse_begin + chunk * chunk_pulses,
// + chunk_pulses) of one tile; every pulse's partials are written by exactly
// one workgroup per tile, whatever the chunk count.
template <int INTERP>
__global__ __launch_bounds__(ADJ_THREADS, 4) void bp_adjoint_split_kernel(
        BpArgs a, SrcArgs s, int tc_log2, const float2* __restrict__ grad,
        float* __restrict__ partials, int chunk_pulses) {
    static_assert(ADJ_THREADS == SP_THREADS && ADJ_PPT == SP_PPT, "split tile shape");
    static_assert(ADJ_SNB % ADJ_NB == 0, "whole pulse groups per block");
    constexpr int F4 = SlotSize<INTERP>::F4;
    __shared__ AdjSplitLds<INTERP> lds;
    const int t = threadIdx.x;
    const int tile = blockIdx.x;
    const int S = a.samples;
    const int P = a.pulse_end - a.pulse_begin;
    const double kt = a.k * INV_TWO_PI;
    const float kt32 = (float)kt;
    const float k32 = (float)a.k;
    const double inv_ddr = 1.0 / a.data_dr;
    const float invdr = (float)inv_ddr;

    // ---- pixels and the tile reference ---------------------------------------
    bool val[SP_PPT];
    double px[SP_PPT], py[SP_PPT];  // dead after the tile reference
    float2 g[SP_PPT];
    double lo_x = 1e300, lo_y = 1e300, hi_x = -1e300, hi_y = -1e300;
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        int row = 0, col = 0;
        const long long p = split_pixel<SRC_GRID>(s, tile, tc_log2, t, k, &val[k], &row, &col);
        if (!val[k]) row = col = 0;
        double z, r;
        split_pixel64<SRC_GRID>(s, val[k] ? p : 0, row, col, &px[k], &py[k], &z, &r);
        g[k] = val[k] ? grad[p] : make_float2(0.f, 0.f);
        if (val[k]) {
            lo_x = fmin(lo_x, px[k]); hi_x = fmax(hi_x, px[k]);
            lo_y = fmin(lo_y, py[k]); hi_y = fmax(hi_y, py[k]);
        }
    }
    // grid pixels lie in the plane z = z0: zero z extent, slant span unused
    tile_reference_bbox(a, lo_x, lo_y, s.a2, hi_x, hi_y, s.a2, 0.0, 0.0, lds.red8, lds.ref, t);
    const double cx = lds.ref.cx, cy = lds.ref.cy;
    double dxk[SP_PPT], dyk[SP_PPT];
    float ek[SP_PPT], ex32[SP_PPT], ey32[SP_PPT];
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        dxk[k] = px[k] - cx;
        dyk[k] = py[k] - cy;
        ek[k] = (float)(dxk[k] * dxk[k] + dyk[k] * dyk[k]);
        ex32[k] = (float)dxk[k];
        ey32[k] = (float)dyk[k];
    }
    const int vlo = InterpShape<INTERP>::FIRST < 0 ? -InterpShape<INTERP>::FIRST : 0;
    const int vhi = S - (InterpShape<INTERP>::TAPS + InterpShape<INTERP>::FIRST);

    const int cbegin = a.pulse_begin + (int)blockIdx.y * chunk_pulses;
    const int cend = min(a.pulse_end, cbegin + chunk_pulses);
    for (int nb0 = cbegin; nb0 < cend; nb0 += ADJ_SNB) {
        const int nb = min(ADJ_SNB, cend - nb0);
        if (t < nb) {
            const int n = nb0 + t;
            lds.prm[t] = split_pulse_params<false>(a, lds.ref, n, kt, inv_ddr, vlo, vhi);
            lds.off[t] = make_float4((float)(lds.ref.cx - a.pos[3 * n]),
                                     (float)(lds.ref.cy - a.pos[3 * n + 1]),
                                     (float)(lds.ref.cz - a.pos[3 * n + 2]), 0.f);
        }
        __syncthreads();
        stage_pulse_block<INTERP>(a, lds.prm, lds.coef, nb, vlo, vhi, t);
        __syncthreads();
        for (int j0 = 0; j0 < nb; j0 += ADJ_NB) {
            // per pulse: A = sum conj(g) c (re, im), then sum b (x - c) over
            // x and y, and sum b; G = -(sum b (x - c) + (c - a) sum b)
            float acc[ADJ_NB][ADJ_VALS];
#pragma unroll
            for (int jj = 0; jj < ADJ_NB; ++jj) {
#pragma unroll
                for (int v = 0; v < ADJ_VALS; ++v) acc[jj][v] = 0.f;
                const int b = j0 + jj;
                if (b >= nb) continue;
                const int flag = uniform_i32(lds.prm[b].flag);
                if (flag == PF_SKIP) continue;
                const PulseParams q = lds.prm[b];
                // Per pixel: h = conj(g) e^{j theta}; A += h v;
                // b = Re(h (dv / data_dr + j k v)) / d = (hr ur - hi ui) / (data_dr d)
                // with ur = dv.x - kd v.y, ui = dv.y + kd v.x, kd = k data_dr, and
                // 1 / (data_dr d) = (1 / (data_dr D)) (1 - (d - D) / D) to first
                // order (relative error below (rho / D)^2).
                const float kd = (float)(a.k * a.data_dr);
                const float r0 = invdr * q.invr, r1 = r0 * q.invr;
                auto fast = [&]<bool ORDER4>() {
                    const float4* cb = lds.coef + b * SP_W * F4;
#pragma unroll
                    for (int k = 0; k < SP_PPT; ++k) {
                        const double lin = dxk[k] * q.vx + dyk[k] * q.vy;
                        const double T = hw_fract64(q.P + lin * kt);
                        const float l32 = (float)lin;
                        const float h = ek[k] - l32 * l32;
                        const float m = l32 * q.invr;
                        const float hs = h * q.scale;
                        float corr;
                        if (ORDER4) {
                            float pz = fmaf(m, m - 1.f, 1.f);
                            pz = fmaf(-hs, q.scale, pz);
                            corr = hs * pz;
                        } else {
                            corr = hs - hs * m;
                        }
                        const float dl = l32 + corr;  // d - D
                        const float ql = q.qb + dl * invdr;
                        const float turns = (float)T + corr * kt32;
                        int j = (int)ql;
                        const float tt = hw_fract32(ql);
                        j = min(max(j, 0), SP_W - 1);
                        float2 v, dv;
                        if (INTERP == INTERP_LAGRANGE4) {
                            // p = ((c3 t + c2) t + c1) t + c0, p' = B + t (A + t c3)
                            const float4 lo = cb[2 * j], hi = cb[2 * j + 1];
                            const float ax_ = fmaf(hi.z, tt, hi.x), ay_ = fmaf(hi.w, tt, hi.y);
                            const float bx = fmaf(ax_, tt, lo.z), by = fmaf(ay_, tt, lo.w);
                            v.x = fmaf(bx, tt, lo.x);
                            v.y = fmaf(by, tt, lo.y);
                            dv.x = fmaf(tt, fmaf(tt, hi.z, ax_), bx);
                            dv.y = fmaf(tt, fmaf(tt, hi.w, ay_), by);
                        } else {
                            const float4 lo = cb[j];
                            v.x = fmaf(lo.z, tt, lo.x);
                            v.y = fmaf(lo.w, tt, lo.y);
                            dv.x = lo.z;
                            dv.y = lo.w;
                        }
                        float sn, cs;
                        sincos_turns(turns, &sn, &cs);
                        // (w is in the staged coefficients)
                        const float hr = g[k].x * cs + g[k].y * sn;
                        const float hi = g[k].x * sn - g[k].y * cs;
                        acc[jj][0] += hr * v.x - hi * v.y;
                        acc[jj][1] += hr * v.y + hi * v.x;
                        const float ur = fmaf(-kd, v.y, dv.x), ui = fmaf(kd, v.x, dv.y);
                        const float br = (hr * ur - hi * ui) * fmaf(-dl, r1, r0);
                        acc[jj][2] += br * ex32[k];
                        acc[jj][3] += br * ey32[k];
                        acc[jj][4] += br;
                    }
                };
                if (flag == PF_FAST4) fast.template operator()<true>();
                else if (flag == PF_FAST2) fast.template operator()<false>();
                else {
                    // exact FP64 geometry, samples from global memory
                    const int n = q.pulse;
                    const float2* row = a.data + (size_t)n * S;
                    const double ph = pulse_phase(a.phase, n) * INV_TWO_PI;
                    const double ux = cx - a.pos[3 * n], uy = cy - a.pos[3 * n + 1];
                    const double ez = s.a2 - a.pos[3 * n + 2];
                    const float w = q.w;
#pragma unroll
                    for (int k = 0; k < SP_PPT; ++k) {
                        if (!val[k]) continue;
                        const double ex = dxk[k] + ux, ey = dyk[k] + uy;
                        const double s2 = ex * ex + ey * ey + ez * ez;
                        const double rinv = rsqrt64(s2);
                        const double d = s2 * rinv;     // NaN for s2 = 0: skipped below
                        const double qq = (d - a.range_start) * inv_ddr;
                        int i;
                        if (!floor_index(qq, &i)) continue;
                        float2 v, dv;
                        sample1_with_derivative<INTERP>(row, S, i, (float)(qq - (double)i), &v,
                                                        &dv);
                        float sn, cs;
                        sincos_turns((float)hw_fract64(kt * d + ph), &sn, &cs);
                        const float hr = w * (g[k].x * cs + g[k].y * sn);
                        const float hi = w * (g[k].x * sn - g[k].y * cs);
                        acc[jj][0] += hr * v.x - hi * v.y;
                        acc[jj][1] += hr * v.y + hi * v.x;
                        const float ur = dv.x * invdr - k32 * v.y;
                        const float ui = dv.y * invdr + k32 * v.x;
                        const float br = (hr * ur - hi * ui) * (float)rinv;
                        acc[jj][2] += br * ex32[k];
                        acc[jj][3] += br * ey32[k];
                        acc[jj][4] += br;
                    }
                }
            }
#pragma unroll
            for (int jj = 0; jj < ADJ_NB; ++jj)
#pragma unroll
                for (int v = 0; v < ADJ_VALS; ++v)
                    reinterpret_cast<float*>(lds.red[jj * ADJ_VALS + v])[t] = acc[jj][v];
            __syncthreads();
            // Wave jj sums pulse jj's five rows, eight lanes per row (lane
            // 8 v + part adds columns 32 part .. 32 part + 31 in a fixed
            // order), then combines the rows within the wave.
            static_assert(ADJ_NB == ADJ_THREADS / 64, "one wave per pulse of a group");
            {
                const int jj = t >> 6, v = (t >> 3) & 7, part = t & 7;
                float sum = 0.f;
                if (v < ADJ_VALS) {
                    const float4* row = lds.red[jj * ADJ_VALS + v] + 8 * part;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const float4 x = row[c];
                        sum += (x.x + x.y) + (x.z + x.w);
                    }
                }
                sum += __shfl_xor(sum, 1, 8);
                sum += __shfl_xor(sum, 2, 8);
                sum += __shfl_xor(sum, 4, 8);
                const float sb = __shfl(sum, 8 * 4, 64);
                if (part == 0 && v < ADJ_VALS && j0 + jj < nb) {
                    const float4 off = lds.off[j0 + jj];
                    const float out = v < 2 ? sum
                                    : v == 2 ? -(sum + off.x * sb)
                                    : v == 3 ? -(sum + off.y * sb)
                                             : -off.z * sb;
                    partials[((size_t)tile * P + (nb0 + j0 + jj - a.pulse_begin)) * ADJ_VALS + v] =
                        out;
                }
            }
            __syncthreads();
        }
    }
}

// Fixed-order FP64 reduction over tiles: out_phase[n] = -Im A, out_pos[n] = G.
// A workgroup takes AR_PULSES pulses; wave c sums part c of the tile range
// (tiles in order, four loads in flight), then lane n of wave 0 adds the
// AR_PARTS part sums in order, so the result is the same on every call.
// (One thread per pulse over all tiles ran 16 workgroups, latency-bound:
// 0.32 ms for the 1024 x 4096 partials of a 1024^2 grid.)
constexpr int AR_PULSES = 64, AR_PARTS = 16;
__global__ __launch_bounds__(AR_PULSES * AR_PARTS) void adjoint_reduce_kernel(
        const float* __restrict__ partials, int tiles, int P, double* __restrict__ grad_phase,
        double* __restrict__ grad_pos) {
    __shared__ double part[AR_PARTS][ADJ_VALS][AR_PULSES];
    const int lane = threadIdx.x % AR_PULSES, c = threadIdx.x / AR_PULSES;
    const int n = blockIdx.x * AR_PULSES + lane;
    const int per = (tiles + AR_PARTS - 1) / AR_PARTS;
    const int b0 = min(tiles, c * per), b1 = min(tiles, b0 + per);
    double s[ADJ_VALS] = {0, 0, 0, 0, 0};
    if (n < P) {
#pragma unroll 4
        for (int b = b0; b < b1; ++b) {
            const float* p = partials + ((size_t)b * P + n) * ADJ_VALS;
#pragma unroll
            for (int j = 0; j < ADJ_VALS; ++j) s[j] += (double)p[j];
        }
    }
#pragma unroll
    for (int j = 0; j < ADJ_VALS; ++j) part[c][j][lane] = s[j];
    __syncthreads();
    if (c != 0 || n >= P) return;
    double t[ADJ_VALS] = {0, 0, 0, 0, 0};
    for (int q = 0; q < AR_PARTS; ++q)
#pragma unroll
        for (int j = 0; j < ADJ_VALS; ++j) t[j] += part[q][j][lane];
    // dL/dphase = Re(j A) = -Im A
    if (grad_phase) grad_phase[n] = -t[1];
    if (grad_pos) {
        grad_pos[3 * n] = t[2];
        grad_pos[3 * n + 1] = t[3];
        grad_pos[3 * n + 2] = t[4];
    }
}

static int adjoint_reduce(const BpArgs& a, long long tiles, const void* partials,
                          double* grad_phase, double* grad_pos, hipStream_t hs) {
    const int P = a.pulse_end - a.pulse_begin;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(grid_1d(P, AR_PULSES)),
                       dim3(AR_PULSES * AR_PARTS), 0, hs, (const float*)partials, (int)tiles, P,
                       grad_phase, grad_pos);
    return check_launch("adjoint_reduce_kernel");
}

static int adjoint_check(int interp, const void* grad, const void* partials) {
    if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
        return set_error(BPMI_ERR_UNSUPPORTED, "adjoint supports lagrange4 and linear");
    if (!grad || !partials) return set_error(BPMI_ERR_INVALID, "grad and partials must be non-null");
    return BPMI_OK;
}

static size_t adjoint_partials_bytes(long long tiles, int pulses) {
    return (size_t)tiles * (size_t)pulses * ADJ_VALS * sizeof(float);
}

// Points: tiles of 1024 consecutive points, per-pixel FP64 geometry.
static int adjoint_points_launch(int interp, const BpArgs& a, const SrcArgs& s, long long count,
                                 const void* grad, void* partials, size_t partials_bytes,
                                 double* grad_phase, double* grad_pos, hipStream_t hs) {
    int st = adjoint_check(interp, grad, partials);
    if (st) return st;
    const long long tiles = (count + ADJ_TILE - 1) / ADJ_TILE;
    const int P = a.pulse_end - a.pulse_begin;
    if (P <= 0) return BPMI_OK;
    const size_t need = adjoint_partials_bytes(tiles, P);
    if (partials_bytes < need)
        return set_error(BPMI_ERR_INVALID, "partials workspace too small: need %zu bytes", need);
    if (tiles > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "image too large");
    const dim3 g((unsigned)tiles);
    if (interp == BPMI_INTERP_LAGRANGE4)
        hipLaunchKernelGGL((bp_adjoint_kernel<INTERP_LAGRANGE4, SRC_POINTS>), g, dim3(ADJ_THREADS), 0, hs, a, s, (const float2*)grad, (float*)partials);
    else
        hipLaunchKernelGGL((bp_adjoint_kernel<INTERP_LINEAR, SRC_POINTS>), g, dim3(ADJ_THREADS), 0, hs, a, s, (const float2*)grad, (float*)partials);
    st = check_launch("bp_adjoint_kernel");
    if (st) return st;
    return adjoint_reduce(a, tiles, partials, grad_phase, grad_pos, hs);
}

// Grid: 2-D split tiles of the forward's shape (choose_tc_log2 as bpmi_bp_grid).
static int adjoint_grid_launch(int interp, const BpArgs& a, const SrcArgs& s, const void* grad,
                               void* partials, size_t partials_bytes, double* grad_phase,
                               double* grad_pos, hipStream_t hs) {
    int st = adjoint_check(interp, grad, partials);
    if (st) return st;
    const int P = a.pulse_end - a.pulse_begin;
    if (P <= 0) return BPMI_OK;
    const int tcl = choose_tc_log2(fabs(s.b1), fabs(s.b3), 1.0, 0.3);
    const long long tiles = tiles_for(s.rows, s.cols, tcl);
    const size_t need = adjoint_partials_bytes(tiles, P);
    if (partials_bytes < need)
        return set_error(BPMI_ERR_INVALID, "partials workspace too small: need %zu bytes", need);
    if (tiles > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "image too large");
    // pulse chunks by the forward's rule, whole pulse blocks each
    int chunks = pulse_chunks(tiles, P);
    int cp = (P + chunks - 1) / chunks;
    cp = (cp + ADJ_SNB - 1) / ADJ_SNB * ADJ_SNB;
    chunks = (P + cp - 1) / cp;
    const dim3 g((unsigned)tiles, (unsigned)chunks);
    if (interp == BPMI_INTERP_LAGRANGE4)
        hipLaunchKernelGGL((bp_adjoint_split_kernel<INTERP_LAGRANGE4>), g, dim3(ADJ_THREADS), 0, hs, a, s, tcl, (const float2*)grad, (float*)partials, cp);
    else
        hipLaunchKernelGGL((bp_adjoint_split_kernel<INTERP_LINEAR>), g, dim3(ADJ_THREADS), 0, hs, a, s, tcl, (const float2*)grad, (float*)partials, cp);
    st = check_launch("bp_adjoint_split_kernel");
    if (st) return st;
    return adjoint_reduce(a, tiles, partials, grad_phase, grad_pos, hs);
}

extern "C" int bpmi_adjoint_workspace_bytes(long long npixels, int pulses, size_t* bytes) {
    if (!bytes) return set_error(BPMI_ERR_INVALID, "bytes is null");
    if (npixels < 0 || pulses < 0) return set_error(BPMI_ERR_INVALID, "bad sizes");
    *bytes = adjoint_partials_bytes((npixels + ADJ_TILE - 1) / ADJ_TILE, pulses);
    return BPMI_OK;
}

extern "C" int bpmi_grid_adjoint_workspace_bytes(int nx, int ny, int pulses, size_t* bytes) {
    if (!bytes) return set_error(BPMI_ERR_INVALID, "bytes is null");
    if (nx < 0 || ny < 0 || pulses < 0) return set_error(BPMI_ERR_INVALID, "bad sizes");
    long long tiles = 0;
    for (int l = 3; l <= 6; ++l) {
        const long long t = tiles_for(ny, nx, l);
        tiles = t > tiles ? t : tiles;
    }
    *bytes = adjoint_partials_bytes(tiles, pulses);
    return BPMI_OK;
}

extern "C" int bpmi_bp_grid_adjoint(const void* data, int samples, const double* pos,
                                    const double* weight, const double* phase, int pulse_begin,
                                    int pulse_end, double x0, double dx, int nx, double y0,
                                    double dy, int ny, double z0, double range_start,
                                    double data_dr, double k, int interp, const void* grad_image,
                                    void* partials, size_t partials_bytes, double* grad_phase,
                                    double* grad_pos, void* stream) {
    int st = validate_common(data, samples, pos, pulse_begin, pulse_end, BPMI_MODE_MIXED, interp,
                             0, (void*)grad_image);
    if (st) return st;
    if (nx <= 0 || ny <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    BpArgs a{(const float2*)data, pos, weight, phase, samples, pulse_begin, pulse_end,
             range_start, data_dr, k, 0, nullptr};
    SrcArgs s{0.0, 0.0, z0, x0, dx, y0, dy, ny, nx, 1, nullptr, 0};
    return adjoint_grid_launch(interp, a, s, grad_image, partials, partials_bytes, grad_phase,
                               grad_pos, (hipStream_t)stream);
}

extern "C" int bpmi_bp_points_adjoint(const void* data, int samples, const double* pos,
                                      const double* weight, const double* phase, int pulse_begin,
                                      int pulse_end, const double* xyz, long long npoints,
                                      double range_start, double data_dr, double k, int interp,
                                      const void* grad_points, void* partials,
                                      size_t partials_bytes, double* grad_phase,
                                      double* grad_pos, void* stream) {
    int st = validate_common(data, samples, pos, pulse_begin, pulse_end, BPMI_MODE_MIXED, interp,
                             0, (void*)grad_points);
    if (st) return st;
    if (npoints <= 0) return BPMI_OK;
    if (!xyz) return set_error(BPMI_ERR_INVALID, "xyz must be non-null");
    BpArgs a{(const float2*)data, pos, weight, phase, samples, pulse_begin, pulse_end,
             range_start, data_dr, k, 0, nullptr};
    SrcArgs s{0, 0, 0, 0, 0, 0, 0, 0, 0, 1, xyz, npoints};
    return adjoint_points_launch(interp, a, s, npoints, grad_points, partials, partials_bytes,
                                 grad_phase, grad_pos, (hipStream_t)stream);
}
