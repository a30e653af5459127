// Device code of the direct-BP kernels (per-pixel policy kernels, level-
// batched leaves and the SPLIT tile body), shared by bpmi_direct.hip and
// the persistent FFBP kernel in bpmi_ffbp.hip.
#pragma once

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
    // Pulse-split launch (split kernel): chunk > 0 gives workgroup row
    // blockIdx.y the pulses [pulse_begin + y chunk, + chunk) and its own
    // partial image at partial + y * (pixels of the image).
    int chunk = 0;
    void* partial = nullptr;
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
    const BPMI_CONST_AS int* ni = as_const(lt.node_int) + 8 * node;
    const BPMI_CONST_AS double* g = as_const(lt.node_geo) + 7 * node;
    a->pulse_begin = ni[0];
    a->pulse_end = ni[1];
    a->flags = 0;
    if (lt.sources) {
        const BPMI_CONST_AS bpmi_node_source& src = as_const(lt.sources)[node];
        a->data = (const float2*)src.data;
        a->pos = src.pos;
        a->weight = src.weight;
        a->phase = src.phase;
        a->range_start = src.range_start;
        a->data_dr = src.data_dr;
        a->k = src.k;
        a->samples = src.samples;
    }
    a->out = lt.arena + as_const(lt.offsets)[node];
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

// Tile reference from the bounding box of the tile's pixels (wave minima,
// then thread 0 of the workgroup): centre, half extents, radius, slant-range
// span and integer sample base.  All threads; two barriers.
__device__ __forceinline__ void tile_reference_bbox(const BpArgs& a, double lo_x, double lo_y,
                                                    double lo_z, double hi_x, double hi_y,
                                                    double hi_z, double rmin, double rmax,
                                                    double (*red)[8], TileRef& ref, int t) {
    const int lane = t & 63, wave = t >> 6;
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

// Per-pulse parameters of pulse n for a tile with reference ``ref`` (FP64;
// one thread per pulse): path flag, window base and the reference-point
// range and phase terms.  vlo / vhi: first and last valid stencil index.
template <bool ORIGIN_REF>
__device__ __forceinline__ PulseParams split_pulse_params(const BpArgs& a, const TileRef& ref,
                                                          int n, double kt, double inv_ddr,
                                                          int vlo, int vhi) {
    const double cx = ref.cx, cy = ref.cy, cz = ref.cz;
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
    return q;
}

// Stage the range windows of a pulse block (nb pulses, parameters prm) as
// weighted polynomial coefficients: slot j of pulse b holds the cubic (or
// line) through the taps around sample prm[b].ilo + j, zero where the
// stencil leaves the valid range [vlo, vhi].  All threads of the workgroup.
template <int INTERP>
__device__ __forceinline__ void stage_pulse_block(const BpArgs& a, const PulseParams* prm,
                                                  float4* coef, int nb, int vlo, int vhi,
                                                  int t) {
    constexpr int F4 = SlotSize<INTERP>::F4;
    const int S = a.samples;
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
}

// LDS of one split tile: staged window coefficients of a pulse block of NB
// pulses, the block's per-pulse parameters, the tile reference and a
// reduction scratch.  NB is a launch-shape choice (16, 32 or 64: one wave's
// ballot covers a block); deeper blocks amortise the two barriers per block
// over more pulses, shallower ones need less LDS per workgroup.
template <int INTERP, int NB = SP_NB>
struct SplitLds {
    static_assert(NB >= 1 && NB <= 64, "a pulse block is at most one wave of flags");
    float4 coef[NB * SP_W * SlotSize<INTERP>::F4];
    PulseParams prm[NB];
    TileRef ref;
    double red[4][8];
};

// One tile of the SPLIT kernel (all SP_THREADS threads of the workgroup
// call it; every thread reaches every barrier, no early return).  A caller
// that runs several tiles in a row must place a barrier between them, since
// the first barrier here is reached after LDS writes of the tile setup.
// COH: the output is read by other workgroups of the same launch (device-
// scope stores, see ld_c64/st_c64).
template <int INTERP, int SRC, bool ORIGIN_REF, bool COH = false, int NB = SP_NB>
__device__ __forceinline__ void bp_split_tile(BpArgs a, SrcArgs s, int tc_log2, int tile,
                                              SplitLds<INTERP, NB>& lds) {
    constexpr int F4 = SlotSize<INTERP>::F4;
    float4* const coef = lds.coef;
    PulseParams* const prm = lds.prm;
    TileRef& ref = lds.ref;
    double (*const red)[8] = lds.red;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
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
    int col0 = 0;  // the thread's column (grid and polar sources: the same for every slot)
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        int row = 0, col = 0;
        pix[k] = split_pixel<SRC>(s, tile, tc_log2, t, k, &val[k], &row, &col);
        if (!val[k]) row = col = 0;
        if (k == 0) col0 = col;
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
        tile_reference_bbox(a, lo_x, lo_y, lo_z, hi_x, hi_y, hi_z, rmin, rmax, red, ref, t);
    }
    const double cx = ref.cx, cy = ref.cy, cz = ref.cz;
    // Per-pixel state.
    double dxk[SP_PPT], dyk[SP_PPT], dzk[SP_PPT];
    float ek[SP_PPT], qpk[SP_PPT], sck[SP_PPT], irk[SP_PPT];
    float2 acc[SP_PPT];
    // A polar image about the reference origin: the unit vector to pixel
    // (row, col) is (side sqrt(1 - u^2) g / r, u g / r, -z / r) with the
    // ground range g = sqrt(r^2 - z^2) -- a column factor (the thread's
    // column is the same for all its pixels) times row terms, instead of
    // the pixel's coordinates minus the origin.
    double col_x = 0.0, col_y = 0.0;
    if (ORIGIN_REF && SRC == SRC_POLAR) {
        // (slot 0 invalid: every slot of the thread is, and nothing uses these)
        const double u = s.b2 + (double)col0 * s.b3;
        col_x = (double)s.side * sqrt64_fast(1.0 - u * u);
        col_y = u;
    }
#pragma unroll
    for (int k = 0; k < SP_PPT; ++k) {
        acc[k] = make_float2(0.f, 0.f);
        if (ORIGIN_REF && SRC == SRC_POLAR) {
            const double r = pr[k] > 0.0 ? pr[k] : 1.0;
            const double ir = rcp64(r);
            const double gi = sqrt64_fast(pr[k] * pr[k] - s.a2 * s.a2) * ir;
            dxk[k] = col_x * gi;
            dyk[k] = col_y * gi;
            dzk[k] = -s.a2 * ir;
            ek[k] = 0.f;
            qpk[k] = (float)((pr[k] - a.range_start) * inv_ddr - (double)ref.itile);
            sck[k] = (float)(0.5 * ir);
            irk[k] = (float)ir;
        } else if (ORIGIN_REF) {
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

    for (int nb0 = a.pulse_begin; nb0 < a.pulse_end; nb0 += NB) {
        const int nb = min(NB, a.pulse_end - nb0);
        // ---- per-pulse parameters (FP64, one thread per pulse) ------------
        if (t < nb) prm[t] = split_pulse_params<ORIGIN_REF>(a, ref, nb0 + t, kt, inv_ddr, vlo, vhi);
        __syncthreads();
        // ---- stage range windows as polynomial coefficients ----------------
        stage_pulse_block<INTERP>(a, prm, coef, nb, vlo, vhi, t);
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
            const float2 o = ld_c64<COH>(out + pix[k]);
            v.x += o.x;
            v.y += o.y;
        }
        st_c64<COH>(out + pix[k], v);
    }
}

template <int INTERP, int SRC, bool ORIGIN_REF, int NB = SP_NB>
__global__ __launch_bounds__(SP_THREADS) void bp_split_kernel(BpArgs a, SrcArgs s, int tc_log2,
                                                              int tile_base, LeafTables lt) {
    __shared__ SplitLds<INTERP, NB> lds;
    int tile = blockIdx.x + tile_base;
    if (lt.blocks) {
        // batched leaves: (node, tile) from the block table
        leaf_args(lt, lt.blocks[2 * blockIdx.x], &a, &s);
        tile = lt.blocks[2 * blockIdx.x + 1];
    }
    if (a.chunk > 0) {
        const int pb = a.pulse_begin + (int)blockIdx.y * a.chunk;
        a.pulse_begin = pb;
        a.pulse_end = min(a.pulse_end, pb + a.chunk);
        a.out = (float2*)a.partial + (long long)blockIdx.y * src_count(s, SRC);
        a.flags &= ~BPMI_ACCUMULATE;
    }
    bp_split_tile<INTERP, SRC, ORIGIN_REF, false, NB>(a, s, tc_log2, tile, lds);
}

}  // namespace bpmi
