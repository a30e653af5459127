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
//    processed in blocks of 16 whose per-pulse geometry is reduced once per
//    tile in FP64 and whose range-window slices are staged in LDS as cubic
//    (or linear) polynomial coefficients pre-multiplied by the pulse weight.
//    Per pixel and pulse the residual geometry runs in FP32, the large linear
//    phase term and its reduction modulo one turn in FP64, the sine/cosine on
//    the FP32 hardware transcendental path, interpolation and accumulation in
//    FP32.  Pulses whose window or expansion error would exceed the bounds
//    fall back, per tile and pulse, to an exact FP64-geometry path.
#include "bpmi_internal.h"

namespace bpmi {

// ---------------------------------------------------------------------------
// Pixel sources
// ---------------------------------------------------------------------------
enum Source { SRC_POLAR = 0, SRC_GRID = 1, SRC_POINTS = 2 };

struct SrcArgs {
    // polar: origin (a0,a1,a2), r0=b0, dr=b1, u0=b2, du=b3, rows=nr, cols=nu
    // grid:  x0=b0, dx=b1, y0=b2, dy=b3, z0=a2, rows=ny, cols=nx
    double a0, a1, a2;
    double b0, b1, b2, b3;
    int rows, cols, side;
    const double* xyz;
    long long count;
};

struct BpArgs {
    const float2* data;
    const double* pos;
    const double* weight;
    const double* phase;
    int samples;
    int pulse_begin, pulse_end;
    double range_start, data_dr, k;
    int flags;
    void* out;
};

__device__ __forceinline__ long long src_count(const SrcArgs& s, int SRC) {
    return SRC == SRC_POINTS ? s.count : (long long)s.rows * (long long)s.cols;
}

// FP64 coordinates of pixel p; r_out = polar slant range (polar only).
template <int SRC>
__device__ __forceinline__ void src_pixel64(const SrcArgs& s, long long p, double* x, double* y,
                                            double* z, double* r_out) {
    if (SRC == SRC_POLAR) {
        const int ir = (int)(p / s.cols);
        const int iu = (int)(p - (long long)ir * s.cols);
        const double r = s.b0 + (double)ir * s.b1;
        const double u = s.b2 + (double)iu * s.b3;
        polar_pixel(s.a0, s.a1, s.a2, r, u, s.side, x, y);
        *z = 0.0;
        *r_out = r;
    } else if (SRC == SRC_GRID) {
        const int iy = (int)(p / s.cols);
        const int ix = (int)(p - (long long)iy * s.cols);
        *x = s.b0 + (double)ix * s.b1;
        *y = s.b2 + (double)iy * s.b3;
        *z = s.a2;
        *r_out = 0.0;
    } else {
        *x = s.xyz[3 * p];
        *y = s.xyz[3 * p + 1];
        *z = s.xyz[3 * p + 2];
        *r_out = 0.0;
    }
}

__device__ __forceinline__ double pulse_weight(const double* w, int n) { return w ? w[n] : 1.0; }
__device__ __forceinline__ double pulse_phase(const double* ph, int n) { return ph ? ph[n] : 0.0; }

// ---------------------------------------------------------------------------
// Per-pixel kernels
// ---------------------------------------------------------------------------
template <int MODE, int INTERP, int SRC>
__device__ __forceinline__ void bp_pixel_body(const BpArgs& a, const SrcArgs& s, long long p,
                                              const float* ktab) {
    const bool demod = (SRC == SRC_POLAR);
    const int S = a.samples;

    if (MODE == MODE_FP64) {
        double x, y, z, r;
        src_pixel64<SRC>(s, p, &x, &y, &z, &r);
        double tr = 0.0, ti = 0.0;
        if (a.flags & BPMI_ACCUMULATE) {
            if (a.flags & BPMI_OUT_C128) {
                const double2 v = ((const double2*)a.out)[p];
                tr = v.x;
                ti = v.y;
            } else {
                const float2 v = ((const float2*)a.out)[p];
                tr = v.x;
                ti = v.y;
            }
        }
        for (int n = a.pulse_begin; n < a.pulse_end; ++n) {
            const double dx = x - a.pos[3 * n];
            const double dy = y - a.pos[3 * n + 1];
            const double dz = z - a.pos[3 * n + 2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const double2 v = sample1_f64<INTERP>(a.data + (size_t)n * S, S,
                                                  (d - a.range_start) / a.data_dr,
                                                  INTERP == INTERP_KAISER8 ? kaiser_table64(ktab) : nullptr);
            double sn, cs;
            sincos(a.k * (demod ? d - r : d) + pulse_phase(a.phase, n), &sn, &cs);
            const double w = pulse_weight(a.weight, n);
            const double wr = w * v.x, wi = w * v.y;
            tr = tr + (wr * cs - wi * sn);
            ti = ti + (wr * sn + wi * cs);
        }
        if (a.flags & BPMI_OUT_C128)
            ((double2*)a.out)[p] = make_double2(tr, ti);
        else
            ((float2*)a.out)[p] = make_float2((float)tr, (float)ti);
        return;
    }

    if (MODE == MODE_FP32) {
        // Every input rounded to FP32 first, as the reference's fp32 policy.
        float x, y, z = 0.f, r = 0.f;
        if (SRC == SRC_POLAR) {
            const int ir = (int)(p / s.cols);
            const int iu = (int)(p - (long long)ir * s.cols);
            r = (float)s.b0 + (float)ir * (float)s.b1;
            const float u = (float)s.b2 + (float)iu * (float)s.b3;
            polar_pixel_f32((float)s.a0, (float)s.a1, (float)s.a2, r, u, s.side, &x, &y);
        } else if (SRC == SRC_GRID) {
            const int iy = (int)(p / s.cols);
            const int ix = (int)(p - (long long)iy * s.cols);
            x = (float)s.b0 + (float)ix * (float)s.b1;
            y = (float)s.b2 + (float)iy * (float)s.b3;
            z = (float)s.a2;
        } else {
            x = (float)s.xyz[3 * p];
            y = (float)s.xyz[3 * p + 1];
            z = (float)s.xyz[3 * p + 2];
        }
        const float rs = (float)a.range_start, ddr = (float)a.data_dr, k = (float)a.k;
        float2 total = make_float2(0.f, 0.f);
        if (a.flags & BPMI_ACCUMULATE) total = ((const float2*)a.out)[p];
        for (int n = a.pulse_begin; n < a.pulse_end; ++n) {
            const float dx = x - (float)a.pos[3 * n];
            const float dy = y - (float)a.pos[3 * n + 1];
            const float dz = z - (float)a.pos[3 * n + 2];
            const float d = sqrtf(dx * dx + dy * dy + dz * dz);
            const float q = (d - rs) / ddr;
            int i;
            float2 v = make_float2(0.f, 0.f);
            if (floor_index_f32(q, &i))
                v = sample1_f32w<INTERP>(a.data + (size_t)n * S, S, i, q - (float)i, ktab);
            float sn, cs;
            sincosf(k * (demod ? d - r : d) + (float)pulse_phase(a.phase, n), &sn, &cs);
            const float w = (float)pulse_weight(a.weight, n);
            total.x += w * (v.x * cs - v.y * sn);
            total.y += w * (v.x * sn + v.y * cs);
        }
        ((float2*)a.out)[p] = total;
        return;
    }

    // MIXED and DFLOAT: FP32 interpolation and accumulation.
    double x, y, z, r;
    src_pixel64<SRC>(s, p, &x, &y, &z, &r);
    const double kt = a.k * INV_TWO_PI;
    float2 total = make_float2(0.f, 0.f);
    if (a.flags & BPMI_ACCUMULATE) total = ((const float2*)a.out)[p];
    if (MODE == MODE_MIXED) {
        for (int n = a.pulse_begin; n < a.pulse_end; ++n) {
            const double dx = x - a.pos[3 * n];
            const double dy = y - a.pos[3 * n + 1];
            const double dz = z - a.pos[3 * n + 2];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const double q = (d - a.range_start) / a.data_dr;
            int i;
            float2 v = make_float2(0.f, 0.f);
            if (floor_index(q, &i))
                v = sample1_f32w<INTERP>(a.data + (size_t)n * S, S, i, (float)(q - (double)i), ktab);
            const double turns = kt * (demod ? d - r : d) + pulse_phase(a.phase, n) * INV_TWO_PI;
            float sn, cs;
            sincos_turns((float)fract64(turns), &sn, &cs);
            const float w = (float)pulse_weight(a.weight, n);
            total.x += w * (v.x * cs - v.y * sn);
            total.y += w * (v.x * sn + v.y * cs);
        }
    } else {
        // Double-float range: pixel and antenna coordinates split hi + lo,
        // distance, sample coordinate and phase carried as hi + lo FP32 pairs.
        const df px = df_from(x), py = df_from(y), pz = df_from(z);
        const df rs = df_from(a.range_start);
        const df inv_dr = df_from(1.0 / a.data_dr);
        const df ktd = df_from(kt);
        const df rr = df_from(r);
        for (int n = a.pulse_begin; n < a.pulse_end; ++n) {
            const df dx = df_sub(px, df_from(a.pos[3 * n]));
            const df dy = df_sub(py, df_from(a.pos[3 * n + 1]));
            const df dz = df_sub(pz, df_from(a.pos[3 * n + 2]));
            const df d2 = df_add(df_add(df_sqr(dx), df_sqr(dy)), df_sqr(dz));
            const df d = df_sqrt(d2);
            const df qd = df_mul(df_sub(d, rs), inv_dr);
            const float qf = floorf(qd.hi);
            float2 v = make_float2(0.f, 0.f);
            if (qf > -2.0e9f && qf < 2.0e9f) {
                const int i = (int)qf;
                const float t = (qd.hi - qf) + qd.lo;
                // t may leave [0,1) by one ulp of the low word: re-normalise.
                const int i2 = t >= 1.f ? i + 1 : (t < 0.f ? i - 1 : i);
                const float t2 = t >= 1.f ? t - 1.f : (t < 0.f ? t + 1.f : t);
                v = sample1_f32w<INTERP>(a.data + (size_t)n * S, S, i2, t2, ktab);
            }
            const df arg = demod ? df_sub(d, rr) : d;
            const df ph = df_mul(ktd, arg);
            const float fh = ph.hi - floorf(ph.hi);
            const float turns = fh + ph.lo + (float)(pulse_phase(a.phase, n) * INV_TWO_PI);
            float sn, cs;
            sincos_turns(turns - floorf(turns), &sn, &cs);
            const float w = (float)pulse_weight(a.weight, n);
            total.x += w * (v.x * cs - v.y * sn);
            total.y += w * (v.x * sn + v.y * cs);
        }
    }
    ((float2*)a.out)[p] = total;
}

template <int MODE, int INTERP, int SRC>
__global__ __launch_bounds__(256) void bp_pixel_kernel(BpArgs a, SrcArgs s, const float* ktab) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= src_count(s, SRC)) return;
    bp_pixel_body<MODE, INTERP, SRC>(a, s, p, ktab);
}

// Level-batched FFBP leaves: the node tables give each block its node and
// pixel range (node_int row = a, b, nr, nu, side, left, right, height;
// node_geo row = ox, oy, oz, r0, dr, u0, du).
struct LeafTables {
    const int* node_int;
    const double* node_geo;
    const long long* offsets;
    float2* arena;
    const int* blocks;
    const bpmi_node_source* sources;  // optional per-node inputs (many patches)
};

__device__ __forceinline__ void leaf_args(const LeafTables& lt, int node, BpArgs* a, SrcArgs* s) {
    const int* ni = lt.node_int + 8 * node;
    const double* g = lt.node_geo + 7 * node;
    a->pulse_begin = ni[0];
    a->pulse_end = ni[1];
    a->flags = 0;
    if (lt.sources) {
        const bpmi_node_source& src = lt.sources[node];
        a->data = (const float2*)src.data;
        a->pos = src.pos;
        a->weight = src.weight;
        a->phase = src.phase;
        a->range_start = src.range_start;
        a->data_dr = src.data_dr;
        a->k = src.k;
        a->samples = src.samples;
    }
    a->out = lt.arena + lt.offsets[node];
    s->a0 = g[0]; s->a1 = g[1]; s->a2 = g[2];
    s->b0 = g[3]; s->b1 = g[4]; s->b2 = g[5]; s->b3 = g[6];
    s->rows = ni[2];
    s->cols = ni[3];
    s->side = ni[4];
}

template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void batched_leaf_kernel(BpArgs a, LeafTables lt,
                                                           const float* ktab) {
    const int node = lt.blocks[2 * blockIdx.x];
    const long long p = (long long)lt.blocks[2 * blockIdx.x + 1] + threadIdx.x;
    SrcArgs s{};
    leaf_args(lt, node, &a, &s);
    if (p >= (long long)s.rows * s.cols) return;
    bp_pixel_body<MODE, INTERP, SRC_POLAR>(a, s, p, ktab);
}

// ---------------------------------------------------------------------------
// SPLIT kernel
// ---------------------------------------------------------------------------
constexpr int SP_THREADS = 256;
constexpr int SP_PPT = 4;
constexpr int SP_TILE = SP_THREADS * SP_PPT;
constexpr int SP_NB = 32;
constexpr int SP_W = 32;

// Per-pulse path of a tile: staged polynomial window with the fourth- or the
// second-order range residual, skipped (window outside the data), or exact
// FP64 geometry.
enum { PF_FAST4 = 0, PF_FAST2 = 1, PF_SKIP = 2, PF_SLOW = 3 };

struct PulseParams {
    double vx, vy, vz;  // tile reference: unit vector (c-a)/D; origin reference: o - a
    double P;           // constant phase in turns
    float qb;           // window coordinate offset
    float q2;           // origin reference: |o - a|^2
    float scale;        // tile reference: 1/(2D)
    float invr;         // tile reference: 1/D
    int flag;
    int ilo;
    int pulse;
    float w;
};

struct TileRef {
    double cx, cy, cz;     // reference point (tile centre or node origin)
    double hx, hy, hz;     // half extents of the tile's bounding box
    double rho;            // tile radius about the centre
    double rmin, rmax;     // origin reference: slant range span
    int itile;             // origin reference: integer sample base
};

template <int INTERP> struct SlotSize;
template <> struct SlotSize<INTERP_LAGRANGE4> { static constexpr int F4 = 2; };
template <> struct SlotSize<INTERP_LINEAR> { static constexpr int F4 = 1; };

__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Pixel index of slot k of thread t in a tile (row/col of grid sources).
template <int SRC>
__device__ __forceinline__ long long split_pixel(const SrcArgs& s, int tile, int tc_log2, int t,
                                                 int k, bool* valid, int* prow = nullptr,
                                                 int* pcol = nullptr) {
    if (SRC == SRC_POINTS) {
        const long long p = (long long)tile * SP_TILE + t + k * SP_THREADS;
        *valid = p < s.count;
        return p;
    }
    const int tc = 1 << tc_log2;
    const int tiles_c = (s.cols + tc - 1) >> tc_log2;
    const int trow = tile / tiles_c;
    const int tcol = tile - trow * tiles_c;
    const int rows_per_k = SP_THREADS >> tc_log2;
    const int tr = SP_TILE >> tc_log2;
    const int col = (tcol << tc_log2) + (t & (tc - 1));
    const int row = trow * tr + (t >> tc_log2) + k * rows_per_k;
    *valid = col < s.cols && row < s.rows;
    if (prow) {
        *prow = row;
        *pcol = col;
    }
    return (long long)row * s.cols + col;
}

// First and last row of a split tile of a grid source (clipped to the grid).
__device__ __forceinline__ void split_tile_rows(const SrcArgs& s, int tile, int tc_log2, int* lo,
                                                int* hi) {
    const int tiles_c = (s.cols + (1 << tc_log2) - 1) >> tc_log2;
    const int tr = SP_TILE >> tc_log2;
    *lo = (tile / tiles_c) * tr;
    *hi = min(*lo + tr - 1, s.rows - 1);
}

// FP64 coordinates of the pixel at (row, col) of a grid source, square roots
// from rsqrt64 (relative error <= 4.2e-15); slot pixels of the split kernel.
template <int SRC>
__device__ __forceinline__ void split_pixel64(const SrcArgs& s, long long p, int row, int col,
                                              double* x, double* y, double* z, double* r_out) {
    if (SRC == SRC_POLAR) {
        const double r = s.b0 + (double)row * s.b1;
        const double u = s.b2 + (double)col * s.b3;
        const double ground = sqrt64_fast(r * r - s.a2 * s.a2);
        *x = s.a0 + (double)s.side * ground * sqrt64_fast(1.0 - u * u);
        *y = s.a1 + ground * u;
        *z = 0.0;
        *r_out = r;
    } else if (SRC == SRC_GRID) {
        *x = s.b0 + (double)col * s.b1;
        *y = s.b2 + (double)row * s.b3;
        *z = s.a2;
        *r_out = 0.0;
    } else {
        src_pixel64<SRC>(s, p, x, y, z, r_out);
    }
}

template <int INTERP, int SRC, bool ORIGIN_REF>
__global__ __launch_bounds__(SP_THREADS) void bp_split_kernel(BpArgs a, SrcArgs s, int tc_log2,
                                                              int tile_base, LeafTables lt) {
    constexpr int F4 = SlotSize<INTERP>::F4;
    __shared__ float4 coef[SP_NB * SP_W * F4];
    __shared__ PulseParams prm[SP_NB];
    __shared__ TileRef ref;
    __shared__ double red[4][8];

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    int tile = blockIdx.x + tile_base;
    if (lt.blocks) {
        // batched leaves: (node, tile) from the block table
        leaf_args(lt, lt.blocks[2 * blockIdx.x], &a, &s);
        tile = lt.blocks[2 * blockIdx.x + 1];
    }
    const int S = a.samples;
    const bool demod_end = (SRC == SRC_POLAR) && !ORIGIN_REF;
    const double kt = a.k * INV_TWO_PI;
    const float kt32 = (float)kt;
    const float invdr = (float)(1.0 / a.data_dr);
    const double inv_ddr = 1.0 / a.data_dr;

    // ---- pixel setup -------------------------------------------------------
    long long pix[SP_PPT];
    bool val[SP_PPT];
    double px[SP_PPT], py[SP_PPT], pz[SP_PPT], pr[SP_PPT];
    double lo_x = 1e300, lo_y = 1e300, lo_z = 1e300, hi_x = -1e300, hi_y = -1e300, hi_z = -1e300;
    double rmin = 1e300, rmax = -1e300;
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        int row = 0, col = 0;
        pix[k] = split_pixel<SRC>(s, tile, tc_log2, t, k, &val[k], &row, &col);
        if (!val[k]) row = col = 0;
        split_pixel64<SRC>(s, val[k] ? pix[k] : 0, row, col, &px[k], &py[k], &pz[k], &pr[k]);
        if (!ORIGIN_REF && val[k]) {
            lo_x = fmin(lo_x, px[k]); hi_x = fmax(hi_x, px[k]);
            lo_y = fmin(lo_y, py[k]); hi_y = fmax(hi_y, py[k]);
            lo_z = fmin(lo_z, pz[k]); hi_z = fmax(hi_z, pz[k]);
            rmin = fmin(rmin, pr[k]); rmax = fmax(rmax, pr[k]);
        }
    }
    if (ORIGIN_REF) {
        // The node origin is the reference; the tile's slant-range span is
        // that of its first and last rows (every row holds a valid pixel).
        if (t == 0) {
            int lo, hi;
            split_tile_rows(s, tile, tc_log2, &lo, &hi);
            const double r_lo = s.b0 + (double)lo * s.b1, r_hi = s.b0 + (double)hi * s.b1;
            ref.cx = s.a0; ref.cy = s.a1; ref.cz = s.a2;
            ref.hx = ref.hy = ref.hz = ref.rho = 0.0;
            ref.rmin = fmin(r_lo, r_hi);
            ref.rmax = fmax(r_lo, r_hi);
            const double it = floor((ref.rmin - a.range_start) / a.data_dr);
            ref.itile = (it > -1.0e9 && it < 1.0e9) ? (int)it : 0;
        }
        __syncthreads();
    } else {
        double v[8] = {lo_x, lo_y, lo_z, -hi_x, -hi_y, -hi_z, rmin, -rmax};
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = wave_min(v[j]);
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) red[wave][j] = v[j];
        }
        __syncthreads();
        if (t == 0) {
            double m[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                m[j] = fmin(fmin(red[0][j], red[1][j]), fmin(red[2][j], red[3][j]));
            ref.cx = 0.5 * (m[0] - m[3]);
            ref.cy = 0.5 * (m[1] - m[4]);
            ref.cz = 0.5 * (m[2] - m[5]);
            ref.hx = 0.5 * (-m[3] - m[0]);
            ref.hy = 0.5 * (-m[4] - m[1]);
            ref.hz = 0.5 * (-m[5] - m[2]);
            ref.rho = sqrt(ref.hx * ref.hx + ref.hy * ref.hy + ref.hz * ref.hz) * (1.0 + 1e-12) + 1e-9;
            ref.rmin = m[6];
            ref.rmax = -m[7];
            const double it = floor((m[6] - a.range_start) / a.data_dr);
            ref.itile = (it > -1.0e9 && it < 1.0e9) ? (int)it : 0;
        }
        __syncthreads();
    }
    const double cx = ref.cx, cy = ref.cy, cz = ref.cz;
    // Per-pixel state.
    double dxk[SP_PPT], dyk[SP_PPT], dzk[SP_PPT];
    float ek[SP_PPT], qpk[SP_PPT], sck[SP_PPT], irk[SP_PPT];
    float2 acc[SP_PPT];
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        acc[k] = make_float2(0.f, 0.f);
        if (ORIGIN_REF) {
            // unit vector from the node origin to the pixel; |p - o| = r
            const double r = pr[k] > 0.0 ? pr[k] : 1.0;
            const double ir = rcp64(r);
            dxk[k] = (px[k] - cx) * ir;
            dyk[k] = (py[k] - cy) * ir;
            dzk[k] = (pz[k] - cz) * ir;
            ek[k] = 0.f;
            qpk[k] = (float)((pr[k] - a.range_start) * inv_ddr - (double)ref.itile);
            sck[k] = (float)(0.5 * ir);
            irk[k] = (float)ir;
        } else {
            dxk[k] = px[k] - cx;
            dyk[k] = py[k] - cy;
            dzk[k] = pz[k] - cz;
            ek[k] = (float)(dxk[k] * dxk[k] + dyk[k] * dyk[k] + dzk[k] * dzk[k]);
            qpk[k] = 0.f;
            sck[k] = 0.f;
            irk[k] = 0.f;
        }
    }

    const int vlo = InterpShape<INTERP>::FIRST < 0 ? -InterpShape<INTERP>::FIRST : 0;
    const int vhi = S - (InterpShape<INTERP>::TAPS + InterpShape<INTERP>::FIRST);  // last valid i

    for (int nb0 = a.pulse_begin; nb0 < a.pulse_end; nb0 += SP_NB) {
        const int nb = min(SP_NB, a.pulse_end - nb0);
        // ---- per-pulse parameters (FP64, one thread per pulse) ------------
        if (t < nb) {
            const int n = nb0 + t;
            PulseParams q;
            const double ax = a.pos[3 * n], ay = a.pos[3 * n + 1], az = a.pos[3 * n + 2];
            const double ux = cx - ax, uy = cy - ay, uz = cz - az;
            const double D = sqrt64_fast(ux * ux + uy * uy + uz * uz);
            const double ph = pulse_phase(a.phase, n) * INV_TWO_PI;
            q.w = (float)pulse_weight(a.weight, n);
            q.pulse = n;
            double qlo, qhi, err, err2 = 1e30;
            bool near;
            if (ORIGIN_REF) {
                q.vx = ux; q.vy = uy; q.vz = uz;
                q.P = ph;
                q.q2 = (float)(D * D);
                q.scale = 0.f;
                q.invr = 0.f;
                qlo = (ref.rmin - D - a.range_start) * inv_ddr;
                qhi = (ref.rmax + D - a.range_start) * inv_ddr;
                const double rm = ref.rmin - D;
                near = rm <= D || ref.rmin <= 0.0;
                const double mm = D / fmax(ref.rmin, 1e-30);
                err = near ? 1e30 : D * D * D * D / (8.0 * rm * rm * rm) + mm * mm * D * D / (2.0 * ref.rmin * (1.0 - mm));
            } else {
                const double inv = rcp64(D);
                q.vx = ux * inv; q.vy = uy * inv; q.vz = uz * inv;
                q.P = fract64(kt * D + ph);
                q.q2 = 0.f;
                q.scale = (float)(0.5 * inv);
                q.invr = (float)inv;
                const double B = fabs(q.vx) * ref.hx + fabs(q.vy) * ref.hy + fabs(q.vz) * ref.hz;
                const double rho = ref.rho;
                near = D <= 2.0 * rho;
                const double extra = near ? 0.0 : rho * rho / (2.0 * (D - rho));
                qlo = (D - B - 1e-6 - a.range_start) * inv_ddr;
                qhi = (D + B + extra + 1e-6 - a.range_start) * inv_ddr;
                // Truncation bound of the fourth-order range expansion used below:
                // sqrt tail h^3 / (16 L^5), 1/(1+m) tail after m^2, and the
                // (1+m)^-3 factor dropped from the h^2 term.
                const double dm = D - rho, mm = rho * inv, r2 = rho * rho;
                const double om = 1.0 - mm;
                err = near ? 1e30
                           : r2 * r2 * r2 / (16.0 * dm * dm * dm * dm * dm) +
                                 r2 * 0.5 * inv * mm * mm * mm / om +
                                 r2 * r2 * 0.125 * inv * inv * inv * 3.0 * mm / (om * om * om * om);
                // second-order residual hs (1 - m): sqrt tail h^2 / (8 L^3) and
                // the 1/(1+m) tail after m
                err2 = near ? 1e30 : r2 * r2 / (8.0 * dm * dm * dm) + r2 * 0.5 * inv * mm * mm / om;
            }
            int flag = PF_FAST4;
            int ilo = 0;
            if (!(qlo > -1.0e9 && qhi < 1.0e9)) {
                flag = PF_SLOW;
            } else {
                ilo = (int)floor(qlo) - 1;
                const int ihi = (int)floor(qhi) + 1;
                if (ihi < vlo || ilo > vhi) flag = PF_SKIP;
                else if (ihi - ilo + 1 > SP_W || near || fabs(kt) * err > 1e-6) flag = PF_SLOW;
                else if (ORIGIN_REF || fabs(kt) * err2 <= 1e-6) flag = PF_FAST2;
            }
            if (flag <= PF_FAST2) {
                if (ORIGIN_REF) q.qb = (float)(ref.itile - ilo);
                else q.qb = (float)((D - a.range_start) / a.data_dr - (double)ilo);
            } else {
                q.qb = 0.f;
            }
            q.flag = flag;
            q.ilo = ilo;
            prm[t] = q;
        }
        __syncthreads();
        // ---- stage range windows as polynomial coefficients ----------------
        for (int sidx = t; sidx < nb * SP_W; sidx += SP_THREADS) {
            const int b = sidx / SP_W, j = sidx - b * SP_W;
            const PulseParams& q = prm[b];
            float4 c01 = make_float4(0.f, 0.f, 0.f, 0.f), c23 = c01;
            const int i = q.ilo + j;
            if (q.flag > PF_FAST2) continue;
            if (i >= vlo && i <= vhi) {
                const float2* row = a.data + (size_t)q.pulse * S;
                const float w = q.w;
                if (INTERP == INTERP_LAGRANGE4) {
                    const float2 vm = row[i - 1], v0 = row[i], v1 = row[i + 1], v2 = row[i + 2];
                    // p(t) = v0 + c1 t + c2 t^2 + c3 t^3 through taps at -1, 0, 1, 2
                    const float c1x = -vm.x * (1.f / 3.f) - 0.5f * v0.x + v1.x - v2.x * (1.f / 6.f);
                    const float c1y = -vm.y * (1.f / 3.f) - 0.5f * v0.y + v1.y - v2.y * (1.f / 6.f);
                    const float c2x = 0.5f * vm.x - v0.x + 0.5f * v1.x;
                    const float c2y = 0.5f * vm.y - v0.y + 0.5f * v1.y;
                    const float c3x = (v2.x - vm.x) * (1.f / 6.f) + 0.5f * (v0.x - v1.x);
                    const float c3y = (v2.y - vm.y) * (1.f / 6.f) + 0.5f * (v0.y - v1.y);
                    c01 = make_float4(w * v0.x, w * v0.y, w * c1x, w * c1y);
                    c23 = make_float4(w * c2x, w * c2y, w * c3x, w * c3y);
                } else {
                    const float2 v0 = row[i], v1 = row[i + 1];
                    c01 = make_float4(w * v0.x, w * v0.y, w * (v1.x - v0.x), w * (v1.y - v0.y));
                }
            }
            coef[(b * SP_W + j) * F4] = c01;
            if (F4 == 2) coef[(b * SP_W + j) * F4 + 1] = c23;
        }
        __syncthreads();
        // ---- accumulate ------------------------------------------------------
        // One pass per path (staged window with the second- or fourth-order
        // residual, exact geometry) over that path's pulses, found from
        // wave-wide ballots of the pulse flags: the loops run on scalar bit
        // masks and the accumulators stay in place across pulses.
        const int my_flag = lane < nb ? prm[lane].flag : PF_SKIP;
        const unsigned long long pulses2 = __ballot(my_flag == PF_FAST2);
        const unsigned long long pulses4 = __ballot(my_flag == PF_FAST4);
        const unsigned long long pulses_slow = __ballot(my_flag == PF_SLOW);
        auto fast_pass = [&]<bool ORDER4>() {
            for (unsigned long long m = ORDER4 ? pulses4 : pulses2; m; m &= m - 1) {
                const int b = __builtin_ctzll(m);
                const PulseParams q = prm[b];
                const float4* cb = coef + b * SP_W * F4;
#pragma unroll
                for (int k = 0; k < SP_PPT; ++k) {
                    // linear range term in FP64 and its phase reduced modulo one turn
                    // (grid tiles lie in one z plane through the tile centre:
                    // their z offset is exactly zero)
                    const double lin = SRC == SRC_GRID ? dxk[k] * q.vx + dyk[k] * q.vy
                                                       : dxk[k] * q.vx + dyk[k] * q.vy + dzk[k] * q.vz;
                    const double T = hw_fract64(q.P + lin * kt);
                    const float l32 = (float)lin;
                    // Range residual d - (R + lin) = sqrt((R + lin)^2 + h) - (R + lin)
                    // with h = Q - lin^2, m = lin / R and hs = h / (2R): hs (1 - m),
                    // plus the fourth-order terms hs (m^2 - h / (4R^2)) where the
                    // tile's bound needs them.
                    float ql, corr;
                    if (ORIGIN_REF) {
                        const float h = q.q2 - l32 * l32;
                        const float m = l32 * irk[k];
                        const float hs = h * sck[k];
                        corr = hs - hs * m;
                    } else if (ORDER4) {
                        const float h = ek[k] - l32 * l32;
                        const float m = l32 * q.invr;
                        const float hs = h * q.scale;
                        float p = fmaf(m, m - 1.f, 1.f);
                        p = fmaf(-hs, q.scale, p);
                        corr = hs * p;
                    } else {
                        const float h = ek[k] - l32 * l32;
                        const float m = l32 * q.invr;
                        const float hs = h * q.scale;
                        corr = hs - hs * m;
                    }
                    if (ORIGIN_REF)
                        ql = (qpk[k] + q.qb) + (l32 + corr) * invdr;
                    else
                        ql = q.qb + (l32 + corr) * invdr;
                    const float turns = (float)T + corr * kt32;
                    int j = (int)ql;
                    const float tt = hw_fract32(ql);
                    j = min(max(j, 0), SP_W - 1);
                    float2 v;
                    if (INTERP == INTERP_LAGRANGE4) {
                        const float4 lo = cb[2 * j], hi = cb[2 * j + 1];
                        v.x = ((hi.z * tt + hi.x) * tt + lo.z) * tt + lo.x;
                        v.y = ((hi.w * tt + hi.y) * tt + lo.w) * tt + lo.y;
                    } else {
                        const float4 lo = cb[j];
                        v.x = lo.z * tt + lo.x;
                        v.y = lo.w * tt + lo.y;
                    }
                    float sn, cs;
                    sincos_turns(turns, &sn, &cs);
                    acc[k].x = fmaf(v.x, cs, acc[k].x);
                    acc[k].x = fmaf(-v.y, sn, acc[k].x);
                    acc[k].y = fmaf(v.x, sn, acc[k].y);
                    acc[k].y = fmaf(v.y, cs, acc[k].y);
                }
            }
        };
        fast_pass.template operator()<false>();
        if (!ORIGIN_REF) fast_pass.template operator()<true>();
        for (unsigned long long m = pulses_slow; m; m &= m - 1) {
            const int b = __builtin_ctzll(m);
            const PulseParams q = prm[b];
            const int n = q.pulse;
            const double ax = a.pos[3 * n], ay = a.pos[3 * n + 1], az = a.pos[3 * n + 2];
            const float2* row = a.data + (size_t)n * S;
            const double ph = pulse_phase(a.phase, n) * INV_TWO_PI;
#pragma unroll
            for (int k = 0; k < SP_PPT; ++k) {
                double ex, ey, ez;
                if (ORIGIN_REF) {
                    // p - a = r * unit(p - o) + (o - a)
                    ex = pr[k] * dxk[k] + q.vx;
                    ey = pr[k] * dyk[k] + q.vy;
                    ez = pr[k] * dzk[k] + q.vz;
                } else {
                    ex = (cx - ax) + dxk[k];
                    ey = (cy - ay) + dyk[k];
                    ez = (cz - az) + dzk[k];
                }
                const double d = sqrt(ex * ex + ey * ey + ez * ez);
                const double qq = (d - a.range_start) / a.data_dr;
                int i;
                if (!floor_index(qq, &i) || i < vlo || i > vhi) continue;
                const float2 v = sample1_f32w<INTERP>(row, S, i, (float)(qq - (double)i), nullptr);
                const double turns = kt * (ORIGIN_REF ? d - pr[k] : d) + ph;
                float sn, cs;
                sincos_turns((float)fract64(turns), &sn, &cs);
                acc[k].x += q.w * (v.x * cs - v.y * sn);
                acc[k].y += q.w * (v.x * sn + v.y * cs);
            }
        }
        __syncthreads();
    }

    // ---- output ----------------------------------------------------------------
    float2* out = (float2*)a.out;
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        if (!val[k]) continue;
        float2 v = acc[k];
        if (demod_end) {
            double x, y, z, r;
            src_pixel64<SRC>(s, pix[k], &x, &y, &z, &r);
            const float2 rot = cis_turns64(-kt * r);
            v = make_float2(v.x * rot.x - v.y * rot.y, v.x * rot.y + v.y * rot.x);
        }
        if (a.flags & BPMI_ACCUMULATE) {
            const float2 o = out[pix[k]];
            v.x += o.x;
            v.y += o.y;
        }
        out[pix[k]] = v;
    }
}

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
                        const SrcArgs& s, LeafTables lt = LeafTables{}) {
    if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
        return set_error(BPMI_ERR_UNSUPPORTED, "split mode supports lagrange4 and linear interpolation");
    // Grids are bounded by the launch limit; split very large images.
    const long long max_blocks = 1LL << 30;
    for (long long base = 0; base < tiles; base += max_blocks) {
        const unsigned nb = (unsigned)(tiles - base < max_blocks ? tiles - base : max_blocks);
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LAGRANGE4, SRC, ORIGIN_REF>), dim3(nb),
                               dim3(SP_THREADS), 0, st, a, s, tc_log2, (int)base, lt);
        else
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LINEAR, SRC, ORIGIN_REF>), dim3(nb),
                               dim3(SP_THREADS), 0, st, a, s, tc_log2, (int)base, lt);
    }
    return check_launch("bp_split_kernel");
}

static long long tiles_for(int rows, int cols, int tc_log2) {
    const int tc = 1 << tc_log2, tr = SP_TILE >> tc_log2;
    return (long long)((cols + tc - 1) / tc) * (long long)((rows + tr - 1) / tr);
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

extern "C" int bpmi_bp_grid(const void* data, int samples, const double* pos, const double* weight,
                            const double* phase, int pulse_begin, int pulse_end, double x0,
                            double dx, int nx, double y0, double dy, int ny, double z0,
                            double range_start, double data_dr, double k, int mode, int interp,
                            int flags, void* out, void* stream) {
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
        return launch_split<SRC_GRID, false>(interp, tiles_for(ny, nx, tcl), tcl, hs, a, s);
    }
    launch_pixel<SRC_GRID>(mode, interp, (long long)nx * ny, hs, a, s);
    return check_launch("bp_pixel_kernel(grid)");
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
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LAGRANGE4, SRC_POLAR, true>), dim3(nblocks),
                               dim3(SP_THREADS), 0, hs, a, s, tile_cols_log2, 0, lt);
        else
            hipLaunchKernelGGL((bp_split_kernel<INTERP_LINEAR, SRC_POLAR, true>), dim3(nblocks),
                               dim3(SP_THREADS), 0, hs, a, s, tile_cols_log2, 0, lt);
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

template <int INTERP, int SRC>
__global__ __launch_bounds__(ADJ_THREADS) void bp_adjoint_kernel(BpArgs a, SrcArgs s,
                                                                 const float2* __restrict__ grad,
                                                                 float* __restrict__ partials) {
    __shared__ float red[4][8];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const long long base = (long long)blockIdx.x * ADJ_TILE;
    const long long count = src_count(s, SRC);
    const int S = a.samples;
    const int P = a.pulse_end - a.pulse_begin;
    const double kt = a.k * INV_TWO_PI;
    const float k32 = (float)a.k;
    const float invdr = (float)(1.0 / a.data_dr);
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
    for (int n = a.pulse_begin; n < a.pulse_end; ++n) {
        const double ax = a.pos[3 * n], ay = a.pos[3 * n + 1], az = a.pos[3 * n + 2];
        const float2* row = a.data + (size_t)n * S;
        const double ph = pulse_phase(a.phase, n) * INV_TWO_PI;
        const float w = (float)pulse_weight(a.weight, n);
        float are = 0.f, aim = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int k = 0; k < ADJ_PPT; ++k) {
            if (!val[k]) continue;
            const double ex = px[k] - ax, ey = py[k] - ay, ez = pz[k] - az;
            const double d = sqrt(ex * ex + ey * ey + ez * ez);
            const double q = (d - a.range_start) / a.data_dr;
            int i;
            if (!floor_index(q, &i)) continue;
            float2 v, dv;
            sample1_with_derivative<INTERP>(row, S, i, (float)(q - (double)i), &v, &dv);
            float sn, cs;
            sincos_turns((float)fract64(kt * d + ph), &sn, &cs);
            // conj(g) * w e^{j theta}
            const float hr = w * (g[k].x * cs + g[k].y * sn);
            const float hi = w * (g[k].x * sn - g[k].y * cs);
            // A += conj(g) w e^{j theta} v
            are += hr * v.x - hi * v.y;
            aim += hr * v.y + hi * v.x;
            // b = conj(g) w e^{j theta} (dv/dr + j k v) / d ; only Re(b) is needed
            const float ur = dv.x * invdr - k32 * v.y;
            const float ui = dv.y * invdr + k32 * v.x;
            const float invd = (float)(1.0 / d);
            const float br = (hr * ur - hi * ui) * invd;
            // (p_n - x) = -(ex, ey, ez)
            gx -= br * (float)ex;
            gy -= br * (float)ey;
            gz -= br * (float)ez;
        }
        float vals[5] = {are, aim, gx, gy, gz};
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            float v = vals[j];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            vals[j] = v;
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < 5; ++j) red[wave][j] = vals[j];
        }
        __syncthreads();
        if (t < 5) {
            const float sum = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
            partials[((size_t)blockIdx.x * P + (n - a.pulse_begin)) * 5 + t] = sum;
        }
        __syncthreads();
    }
}

// Fixed-order FP64 reduction over tiles: out_phase[n] = -Im A, out_pos[n] = G.
__global__ void adjoint_reduce_kernel(const float* __restrict__ partials, int tiles, int P,
                                      double* __restrict__ grad_phase,
                                      double* __restrict__ grad_pos) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= P) return;
    double s[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < tiles; ++b) {
        const float* p = partials + ((size_t)b * P + n) * 5;
#pragma unroll
        for (int j = 0; j < 5; ++j) s[j] += (double)p[j];
    }
    // dL/dphase = Re(j A) = -Im A
    if (grad_phase) grad_phase[n] = -s[1];
    if (grad_pos) {
        grad_pos[3 * n] = s[2];
        grad_pos[3 * n + 1] = s[3];
        grad_pos[3 * n + 2] = s[4];
    }
}

static int adjoint_launch(int interp, int src, const BpArgs& a, const SrcArgs& s, long long count,
                          const void* grad, void* partials, size_t partials_bytes,
                          double* grad_phase, double* grad_pos, hipStream_t hs) {
    if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
        return set_error(BPMI_ERR_UNSUPPORTED, "adjoint supports lagrange4 and linear");
    if (!grad || !partials) return set_error(BPMI_ERR_INVALID, "grad and partials must be non-null");
    const long long tiles = (count + ADJ_TILE - 1) / ADJ_TILE;
    const int P = a.pulse_end - a.pulse_begin;
    if (P <= 0) return BPMI_OK;
    const size_t need = (size_t)tiles * (size_t)P * 5 * sizeof(float);
    if (partials_bytes < need)
        return set_error(BPMI_ERR_INVALID, "partials workspace too small: need %zu bytes", need);
    if (tiles > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "image too large");
    const dim3 g((unsigned)tiles);
    if (src == SRC_GRID) {
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((bp_adjoint_kernel<INTERP_LAGRANGE4, SRC_GRID>), g, dim3(ADJ_THREADS), 0, hs, a, s, (const float2*)grad, (float*)partials);
        else
            hipLaunchKernelGGL((bp_adjoint_kernel<INTERP_LINEAR, SRC_GRID>), g, dim3(ADJ_THREADS), 0, hs, a, s, (const float2*)grad, (float*)partials);
    } else {
        if (interp == BPMI_INTERP_LAGRANGE4)
            hipLaunchKernelGGL((bp_adjoint_kernel<INTERP_LAGRANGE4, SRC_POINTS>), g, dim3(ADJ_THREADS), 0, hs, a, s, (const float2*)grad, (float*)partials);
        else
            hipLaunchKernelGGL((bp_adjoint_kernel<INTERP_LINEAR, SRC_POINTS>), g, dim3(ADJ_THREADS), 0, hs, a, s, (const float2*)grad, (float*)partials);
    }
    int st = check_launch("bp_adjoint_kernel");
    if (st) return st;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(grid_1d(P, 256)), dim3(256), 0, hs,
                       (const float*)partials, (int)tiles, P, grad_phase, grad_pos);
    return check_launch("adjoint_reduce_kernel");
}

extern "C" int bpmi_adjoint_workspace_bytes(long long npixels, int pulses, size_t* bytes) {
    if (!bytes) return set_error(BPMI_ERR_INVALID, "bytes is null");
    if (npixels < 0 || pulses < 0) return set_error(BPMI_ERR_INVALID, "bad sizes");
    *bytes = (size_t)((npixels + ADJ_TILE - 1) / ADJ_TILE) * (size_t)pulses * 5 * sizeof(float);
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
    return adjoint_launch(interp, SRC_GRID, a, s, (long long)nx * ny, grad_image, partials,
                          partials_bytes, grad_phase, grad_pos, (hipStream_t)stream);
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
    return adjoint_launch(interp, SRC_POINTS, a, s, npoints, grad_points, partials,
                          partials_bytes, grad_phase, grad_pos, (hipStream_t)stream);
}
