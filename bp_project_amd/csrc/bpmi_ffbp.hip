// FFBP merge/regrid, polar->Cartesian evaluation, per-pulse contributions,
// the phase-LUT direct variant and the small utility kernels.
//
// Precision policies per kernel:
//   FP64   : FP64 geometry, weights, accumulation and sincos; complex64 store
//            (regrid_fp64, fast_bp/gpu/kernels.cu:341-369).
//   FP32   : every input rounded to FP32, FP32 arithmetic, sincosf
//            (regrid_fp32, kernels.cu:371-398).
//   others : FP64 geometry and sample coordinates, FP32 weights and
//            accumulation, FP64 phase reduced to turns + FP32 sincos.
#include "bpmi_internal.h"
#include "bpmi_split.h"

#include <cstring>

namespace bpmi {

static float* g_kaiser_table = nullptr;

const float* kaiser_table_device() { return g_kaiser_table; }

struct Geo {
    double ox, oy, oz, r0, dr, u0, du;
};

__host__ __device__ inline Geo make_geo(const double* g) {
    Geo r;
    r.ox = g[0]; r.oy = g[1]; r.oz = g[2]; r.r0 = g[3]; r.dr = g[4]; r.u0 = g[5]; r.du = g[6];
    return r;
}

// ---- 2-D sampling of a polar image [nr, ld] (first nu columns valid) -------
template <int INTERP>
__device__ __forceinline__ float2 sample2_f32w(const float2* __restrict__ img, int nr, int nu,
                                               int ld, int ir, int iu, float tr, float tu,
                                               const float* ktab) {
    float2 z = make_float2(0.f, 0.f);
    if (!InterpShape<INTERP>::valid(ir, nr) || !InterpShape<INTERP>::valid(iu, nu)) return z;
    if (INTERP == INTERP_KAISER8) {
        float wr[8], wu[8];
        kaiser_weights_f32(ktab, tr, wr);
        kaiser_weights_f32(ktab, tu, wu);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float2* row = img + (size_t)(ir + a - 3) * ld + (iu - 3);
            float sx = 0.f, sy = 0.f;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const float2 v = row[b];
                sx += wu[b] * v.x;
                sy += wu[b] * v.y;
            }
            z.x += wr[a] * sx;
            z.y += wr[a] * sy;
        }
        return z;
    }
    float wr[4], wu[4];
    cubic_weights(tr, wr);
    cubic_weights(tu, wu);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float2* row = img + (size_t)(ir + a - 1) * ld + (iu - 1);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float2 v = row[b];
            const float w = wr[a] * wu[b];
            z.x += w * v.x;
            z.y += w * v.y;
        }
    }
    return z;
}

template <int INTERP>
__device__ __forceinline__ double2 sample2_f64(const float2* __restrict__ img, int nr, int nu,
                                               int ld, double qr, double qu, const float* ktab) {
    double2 z = make_double2(0.0, 0.0);
    int ir, iu;
    if (!floor_index(qr, &ir) || !floor_index(qu, &iu)) return z;
    if (!InterpShape<INTERP>::valid(ir, nr) || !InterpShape<INTERP>::valid(iu, nu)) return z;
    const double tr = qr - (double)ir, tu = qu - (double)iu;
    if (INTERP == INTERP_KAISER8) {
        double wr[8], wu[8];
        kaiser_weights_f64(kaiser_table64(ktab), tr, wr);
        kaiser_weights_f64(kaiser_table64(ktab), tu, wu);
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                const float2 v = img[(size_t)(ir + a - 3) * ld + iu + b - 3];
                const double w = wr[a] * wu[b];
                z.x += w * (double)v.x;
                z.y += w * (double)v.y;
            }
        return z;
    }
    double wr[4], wu[4];
    cubic_weights(tr, wr);
    cubic_weights(tu, wu);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float2 v = img[(size_t)(ir + a - 1) * ld + iu + b - 1];
            const double w = wr[a] * wu[b];
            z.x += w * (double)v.x;
            z.y += w * (double)v.y;
        }
    return z;
}

// ---- one child regrid at destination pixel p (fast_bp/ffbp.py:79-95) -------
template <int MODE, int INTERP>
__device__ __forceinline__ float2 regrid_pixel(const float2* __restrict__ src, const Geo& sg,
                                               int snr, int snu, const Geo& dg, int nu,
                                               long long p, int side, double k,
                                               const float* ktab) {
    const int ir = (int)(p / nu);
    const int iu = (int)(p - (long long)ir * nu);
    if (MODE == MODE_FP32) {
        const float r = (float)dg.r0 + (float)ir * (float)dg.dr;
        const float u = (float)dg.u0 + (float)iu * (float)dg.du;
        float x, y;
        polar_pixel_f32((float)dg.ox, (float)dg.oy, (float)dg.oz, r, u, side, &x, &y);
        const float dx = x - (float)sg.ox, dy = y - (float)sg.oy;
        const float g = sqrtf(dx * dx + dy * dy);
        if (g == 0.0f) return make_float2(0.f, 0.f);
        const float soz = (float)sg.oz;
        const float rr = sqrtf(g * g + soz * soz);
        const float uu = dy / g;
        const float qr = (rr - (float)sg.r0) / (float)sg.dr;
        const float qu = (uu - (float)sg.u0) / (float)sg.du;
        int i0, i1;
        float2 v = make_float2(0.f, 0.f);
        if (floor_index_f32(qr, &i0) && floor_index_f32(qu, &i1))
            v = sample2_f32w<INTERP>(src, snr, snu, snu, i0, i1, qr - (float)i0, qu - (float)i1, ktab);
        float s, c;
        sincosf((float)k * (rr - r), &s, &c);
        return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
    }
    const double r = dg.r0 + (double)ir * dg.dr;
    const double u = dg.u0 + (double)iu * dg.du;
    double x, y;
    polar_pixel(dg.ox, dg.oy, dg.oz, r, u, side, &x, &y);
    const double dx = x - sg.ox, dy = y - sg.oy;
    const double g = sqrt(dx * dx + dy * dy);
    if (g == 0.0) return make_float2(0.f, 0.f);
    const double rr = sqrt(g * g + sg.oz * sg.oz);
    const double uu = dy / g;
    const double qr = (rr - sg.r0) / sg.dr;
    const double qu = (uu - sg.u0) / sg.du;
    if (MODE == MODE_FP64) {
        const double2 v = sample2_f64<INTERP>(src, snr, snu, snu, qr, qu, ktab);
        double s, c;
        sincos(k * (rr - r), &s, &c);
        return make_float2((float)(v.x * c - v.y * s), (float)(v.x * s + v.y * c));
    }
    int i0, i1;
    float2 v = make_float2(0.f, 0.f);
    if (floor_index(qr, &i0) && floor_index(qu, &i1))
        v = sample2_f32w<INTERP>(src, snr, snu, snu, i0, i1, (float)(qr - (double)i0),
                                 (float)(qu - (double)i1), ktab);
    const float2 rot = cis_turns64(k * INV_TWO_PI * (rr - r));
    return make_float2(v.x * rot.x - v.y * rot.y, v.x * rot.y + v.y * rot.x);
}

template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void regrid_kernel(const float2* __restrict__ src, Geo sg,
                                                     int snr, int snu, Geo dg, int nr, int nu,
                                                     int side, double k, int accumulate,
                                                     float2* __restrict__ out, const float* ktab) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nr * nu) return;
    float2 v = regrid_pixel<MODE, INTERP>(src, sg, snr, snu, dg, nu, p, side, k, ktab);
    if (accumulate) {
        const float2 o = out[p];
        v = make_float2(__fadd_rn(o.x, v.x), __fadd_rn(o.y, v.y));
    }
    out[p] = v;
}

// Both children rounded to complex64, then an unfused FP32 add
// (fast_bp/gpu/resident_kernels.cu:1-4,53-54).
template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void merge2_kernel(const float2* __restrict__ left, Geo lg,
                                                     int lnr, int lnu,
                                                     const float2* __restrict__ right, Geo rg,
                                                     int rnr, int rnu, Geo dg, int nr, int nu,
                                                     int side, double k, float2* __restrict__ out,
                                                     const float* ktab) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nr * nu) return;
    const float2 a = regrid_pixel<MODE, INTERP>(left, lg, lnr, lnu, dg, nu, p, side, k, ktab);
    const float2 b = regrid_pixel<MODE, INTERP>(right, rg, rnr, rnu, dg, nu, p, side, k, ktab);
    out[p] = make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}

// Level-batched merges: blocks = (node, first_pixel) pairs, 256 pixels each.
template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void batched_merge_kernel(const int* __restrict__ node_int,
                                                            const double* __restrict__ node_geo,
                                                            const long long* __restrict__ offsets,
                                                            float2* arena,
                                                            const int* __restrict__ blocks,
                                                            double k, const float* ktab) {
    const int node = blocks[2 * blockIdx.x];
    const long long p = (long long)blocks[2 * blockIdx.x + 1] + threadIdx.x;
    const int* ni = node_int + 8 * node;
    const int nr = ni[2], nu = ni[3], side = ni[4], l = ni[5], r = ni[6];
    if (p >= (long long)nr * nu) return;
    const Geo dg = make_geo(node_geo + 7 * node);
    const Geo lg = make_geo(node_geo + 7 * l);
    const Geo rg = make_geo(node_geo + 7 * r);
    const int* li = node_int + 8 * l;
    const int* ri = node_int + 8 * r;
    const float2 a = regrid_pixel<MODE, INTERP>(arena + offsets[l], lg, li[2], li[3], dg, nu, p,
                                                side, k, ktab);
    const float2 b = regrid_pixel<MODE, INTERP>(arena + offsets[r], rg, ri[2], ri[3], dg, nu, p,
                                                side, k, ktab);
    arena[offsets[node] + p] = make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}

// ---- polar -> points (fast_bp/ffbp.py:189-198); point height ignored ------
template <int MODE, int INTERP>
__device__ __forceinline__ float2 eval_pixel(const float2* __restrict__ img, const Geo& g, int nr,
                                             int nu, int ld, double x, double y, double k,
                                             const float* ktab) {
    if (MODE == MODE_FP32) {
        const float dx = (float)x - (float)g.ox, dy = (float)y - (float)g.oy;
        const float gg = sqrtf(dx * dx + dy * dy);
        if (gg == 0.f) return make_float2(0.f, 0.f);
        const float oz = (float)g.oz;
        const float r = sqrtf(gg * gg + oz * oz);
        const float u = dy / gg;
        const float qr = (r - (float)g.r0) / (float)g.dr, qu = (u - (float)g.u0) / (float)g.du;
        int i0, i1;
        float2 v = make_float2(0.f, 0.f);
        if (floor_index_f32(qr, &i0) && floor_index_f32(qu, &i1))
            v = sample2_f32w<INTERP>(img, nr, nu, ld, i0, i1, qr - (float)i0, qu - (float)i1, ktab);
        float s, c;
        sincosf((float)k * r, &s, &c);
        return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
    }
    const double dx = x - g.ox, dy = y - g.oy;
    const double gg = sqrt(dx * dx + dy * dy);
    if (gg == 0.0) return make_float2(0.f, 0.f);
    const double r = sqrt(gg * gg + g.oz * g.oz);
    const double u = dy / gg;
    const double qr = (r - g.r0) / g.dr, qu = (u - g.u0) / g.du;
    if (MODE == MODE_FP64) {
        const double2 v = sample2_f64<INTERP>(img, nr, nu, ld, qr, qu, ktab);
        double s, c;
        sincos(k * r, &s, &c);
        return make_float2((float)(v.x * c - v.y * s), (float)(v.x * s + v.y * c));
    }
    int i0, i1;
    float2 v = make_float2(0.f, 0.f);
    if (floor_index(qr, &i0) && floor_index(qu, &i1))
        v = sample2_f32w<INTERP>(img, nr, nu, ld, i0, i1, (float)(qr - (double)i0),
                                 (float)(qu - (double)i1), ktab);
    const float2 rot = cis_turns64(k * INV_TWO_PI * r);
    return make_float2(v.x * rot.x - v.y * rot.y, v.x * rot.y + v.y * rot.x);
}

template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void eval_points_kernel(const float2* __restrict__ img, Geo g,
                                                          int nr, int nu, int ld,
                                                          const double* __restrict__ xyz,
                                                          long long n, double k,
                                                          float2* __restrict__ out,
                                                          const float* ktab) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    out[p] = eval_pixel<MODE, INTERP>(img, g, nr, nu, ld, xyz[3 * p], xyz[3 * p + 1], k, ktab);
}

template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void eval_grid_multi_kernel(const bpmi_eval_job* __restrict__ jobs,
                                                              const int* __restrict__ blocks,
                                                              const float* ktab) {
    const bpmi_eval_job& j = jobs[blocks[2 * blockIdx.x]];
    const long long p = (long long)blocks[2 * blockIdx.x + 1] + threadIdx.x;
    if (p >= (long long)j.nx * j.ny) return;
    const int iy = (int)(p / j.nx), ix = (int)(p - (long long)iy * j.nx);
    const Geo g = make_geo(j.geo);
    ((float2*)j.out)[p] = eval_pixel<MODE, INTERP>((const float2*)j.img, g, j.nr, j.nu, j.ld,
                                                   j.x0 + ix * j.dx, j.y0 + iy * j.dy, j.k, ktab);
}

template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void eval_grid_kernel(const float2* __restrict__ img, Geo g,
                                                        int nr, int nu, int ld, double x0,
                                                        double dx, int nx, double y0, double dy,
                                                        int ny, double k, float2* __restrict__ out,
                                                        const float* ktab) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nx * ny) return;
    const int iy = (int)(p / nx), ix = (int)(p - (long long)iy * nx);
    out[p] = eval_pixel<MODE, INTERP>(img, g, nr, nu, ld, x0 + ix * dx, y0 + iy * dy, k, ktab);
}

// ---- per-(point, pulse) contributions (fast_bp/ffbp.py:201-211) ------------
template <int MODE, int INTERP>
__global__ __launch_bounds__(256) void contrib_kernel(const float2* __restrict__ data, int S,
                                                      const double* __restrict__ pos, int P,
                                                      const double* __restrict__ xyz,
                                                      long long M, double rs, double ddr,
                                                      double k, float2* __restrict__ out,
                                                      const float* ktab) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * (long long)P) return;
    const long long p = idx / P;
    const int n = (int)(idx - p * P);
    const double dx = xyz[3 * p] - pos[3 * n], dy = xyz[3 * p + 1] - pos[3 * n + 1],
                 dz = xyz[3 * p + 2] - pos[3 * n + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double q = (d - rs) / ddr;
    const float2* row = data + (size_t)n * S;
    if (MODE == MODE_FP64) {
        const double2 v = sample1_f64<INTERP>(row, S, q);
        double s, c;
        sincos(k * d, &s, &c);
        out[idx] = make_float2((float)(v.x * c - v.y * s), (float)(v.x * s + v.y * c));
        return;
    }
    int i;
    float2 v = make_float2(0.f, 0.f);
    if (floor_index(q, &i)) v = sample1_f32w<INTERP>(row, S, i, (float)(q - (double)i), ktab);
    const float2 rot = cis_turns64(k * INV_TWO_PI * d);
    out[idx] = make_float2(v.x * rot.x - v.y * rot.y, v.x * rot.y + v.y * rot.x);
}


// ---------------------------------------------------------------------------
// Fast (split-mode) regrid: FP64 geometry without IEEE square roots or
// divisions (hardware reciprocal square root / reciprocal + Newton steps),
// the destination pixel evaluated once per parent pixel, and Kaiser-sinc
// weights from per-tap polynomials (degree 7 in s = 2t - 1, fitted by the
// library to the tabulated kernel at upload; max weight error 1.3e-6, and
// the eight weights sum to 1 up to FP32 rounding).
//
// The 8-tap kernel is mirror-symmetric, w_{7-a}(s) = w_a(-s), so each pair
// (a, 7 - a) shares an even part E_a(s^2) (4 coefficients) and an odd part
// s O_a(s^2) (4 coefficients): w_a = E + s O, w_{7-a} = E - s O.
// ---------------------------------------------------------------------------
constexpr int KPOLY_EVEN = 4, KPOLY_ODD = 4, KPOLY_STRIDE = KPOLY_EVEN + KPOLY_ODD;
__constant__ float c_kaiser_poly[4][KPOLY_STRIDE];

// kp: the coefficient table (c_kaiser_poly, or a workgroup's LDS copy of it).
__device__ __forceinline__ void kaiser_weights_poly(const float* kp, float t, float w[8]) {
    const float s = 2.f * t - 1.f;
    const float x = s * s;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float* c = kp + a * KPOLY_STRIDE;
        float e = c[KPOLY_EVEN - 1];
#pragma unroll
        for (int j = KPOLY_EVEN - 2; j >= 0; --j) e = fmaf(e, x, c[j]);
        float o = c[KPOLY_STRIDE - 1];
#pragma unroll
        for (int j = KPOLY_ODD - 2; j >= 0; --j) o = fmaf(o, x, c[KPOLY_EVEN + j]);
        w[a] = fmaf(s, o, e);
        w[7 - a] = fmaf(-s, o, e);
    }
}


struct FastSrc {
    const float2* img;
    double ox, oy, oz2, r0, inv_dr, u0, inv_du;
    int nr, nu, ld;
};

__device__ __forceinline__ FastSrc make_fast_src_inv(const float2* img, const double* g, int nr,
                                                     int nu, int ld, double inv_dr, double inv_du) {
    FastSrc f;
    f.img = img;
    f.ox = g[0]; f.oy = g[1]; f.oz2 = g[2] * g[2];
    f.r0 = g[3]; f.inv_dr = inv_dr;
    f.u0 = g[5]; f.inv_du = inv_du;
    f.nr = nr; f.nu = nu; f.ld = ld;
    return f;
}

__device__ __forceinline__ FastSrc make_fast_src(const float2* img, const double* g, int nr, int nu,
                                                 int ld) {
    return make_fast_src_inv(img, g, nr, nu, ld, rcp64(g[4]), rcp64(g[6]));
}

template <int INTERP> struct FastShape;
template <> struct FastShape<INTERP_KAISER8> { static constexpr int TAPS = 8, FIRST = 3; };
template <> struct FastShape<INTERP_LAGRANGE4> { static constexpr int TAPS = 4, FIRST = 1; };

// Where a destination point falls in a source polar image: stencil origin
// (floor indices), fractional offsets and the phase rotation exp(j k (rr - r))
// (r = 0 remodulates by exp(j k rr)).  ok = false when the point lies outside
// the image's valid stencil range or on the origin; the sample is then 0.
// ir/iu hold the floor indices whenever they are defined (also when !ok),
// else INT_MAX.
struct ChildTap {
    int ir, iu;
    float tr, tu;
    float2 rot;
    bool ok;
};

template <int INTERP>
__device__ __forceinline__ ChildTap child_tap(const FastSrc& f, double x, double y, double r,
                                              double kt) {
    // Branch-free: every quantity is computed, then masked by the validity
    // conditions (an index beyond int range, a stencil leaving the image),
    // which keeps the geometry in straight-line code.  A point on the origin
    // (s2 = 0) gives 0 * rsqrt(0) = NaN in uu, and NaN fails the range test.
    const double dx = x - f.ox, dy = y - f.oy;
    const double s2 = dx * dx + dy * dy;
    const double uu = dy * rsqrt64(s2);
    const double rr = sqrt64_newton(s2 + f.oz2);
    const double qr = (rr - f.r0) * f.inv_dr, qu = (uu - f.u0) * f.inv_du;
    const double fr = floor(qr), fu = floor(qu);
    const bool defined = fabs(fr) < 2147483000.0 && fabs(fu) < 2147483000.0;
    ChildTap c;
    c.ir = defined ? (int)fr : 0x7fffffff;
    c.iu = defined ? (int)fu : 0x7fffffff;
    c.ok = defined && InterpShape<INTERP>::valid(c.ir, f.nr) && InterpShape<INTERP>::valid(c.iu, f.nu);
    c.tr = (float)(qr - fr);
    c.tu = (float)(qu - fu);
    c.rot = cis_turns64(fma(kt, rr, -kt * r));
    return c;
}

// The tiled pixel loops' form of child_tap (3 instead of 6 VALU per axis):
// the fraction is v_fract_f64 (q - floor(q), the same value) and the index
// the saturating truncation of q, which equals floor(q) for q >= 0; for a
// negative, NaN (origin) or out-of-range q it is 0, INT_MIN or INT_MAX, and
// no stencil with such an index is valid (FIRST >= 1 taps before it), so the
// sample is 0 as with child_tap.  ok is not computed: the LDS window path
// needs no validity test and tile_sample's global-memory path tests it.
template <int INTERP>
__device__ __forceinline__ ChildTap child_tap_px(const FastSrc& f, double x, double y, double r,
                                                 double kt) {
    static_assert(FastShape<INTERP>::FIRST >= 1, "a truncated negative index must be invalid");
    const double dx = x - f.ox, dy = y - f.oy;
    const double s2 = dx * dx + dy * dy;
    const double uu = dy * rsqrt64(s2);
    const double rr = sqrt64_newton(s2 + f.oz2);
    // (q - origin) * (1 / spacing): one scalar operand per instruction (an
    // FMA with the origin folded in would take two and hold one of them in
    // vector registers)
    const double qr = (rr - f.r0) * f.inv_dr, qu = (uu - f.u0) * f.inv_du;
    ChildTap c;
    c.ir = cvt_i32_sat(qr);
    c.iu = cvt_i32_sat(qu);
    c.tr = (float)hw_fract64(qr);
    c.tu = (float)hw_fract64(qu);
    c.ok = false;
    c.rot = cis_turns64(fma(kt, rr, -kt * r));
    return c;
}

// Separable stencil sum times the phase rotation.  row(a) points at tap
// (a, 0) of the stencil (row ir - FIRST + a, column iu - FIRST), in global
// memory (GLOBAL) or in an LDS copy: the arithmetic is identical either way.
template <int INTERP, typename Row, bool GLOBAL = false>
__device__ __forceinline__ float2 stencil(const ChildTap& c, Row row,
                                          const float* kp = &c_kaiser_poly[0][0]) {
    constexpr int TAPS = FastShape<INTERP>::TAPS;
    float wr[TAPS], wu[TAPS];
    if (INTERP == INTERP_KAISER8) {
        kaiser_weights_poly(kp, c.tr, wr);
        kaiser_weights_poly(kp, c.tu, wu);
    } else {
        cubic_weights(c.tr, wr);
        cubic_weights(c.tu, wu);
    }
    float2 z = make_float2(0.f, 0.f);
#pragma unroll
    for (int a = 0; a < TAPS; ++a) {
        const float2* rp = row(a);
        float sx = 0.f, sy = 0.f;
#pragma unroll
        for (int b = 0; b < TAPS; ++b) {
            const float2 v = GLOBAL ? ld_global_c64(rp + b) : rp[b];
            sx += wu[b] * v.x;
            sy += wu[b] * v.y;
        }
        z.x += wr[a] * sx;
        z.y += wr[a] * sy;
    }
    return make_float2(z.x * c.rot.x - z.y * c.rot.y, z.x * c.rot.y + z.y * c.rot.x);
}

template <int INTERP>
__device__ __forceinline__ float2 stencil_global(const FastSrc& f, const ChildTap& c,
                                                 const float* kp = &c_kaiser_poly[0][0]) {
    constexpr int FIRST = FastShape<INTERP>::FIRST;
    const float2* base = f.img + (long long)(c.ir - FIRST) * f.ld + (c.iu - FIRST);
    const long long ld = f.ld;
    auto row = [&](int a) { return base + a * ld; };
    return stencil<INTERP, decltype(row), true>(c, row, kp);
}

// Destination polar pixel p of node geometry g (fast_bp/ffbp.py:62-67).
// Column terms of a destination pixel: u and side * sqrt(1 - u^2).
struct DestColumn {
    double u, scu;
};

__device__ __forceinline__ DestColumn dest_column(const double* g, int side, int iu) {
    DestColumn c;
    c.u = g[5] + (double)iu * g[6];
    const double c2 = 1.0 - c.u * c.u;
    c.scu = (double)side * (c2 > 0.0 ? c2 * rsqrt64(c2) : 0.0);
    return c;
}

__device__ __forceinline__ void dest_pixel_col(const double* g, const DestColumn& c, int ir,
                                               double* x, double* y, double* r) {
    const double rr = g[3] + (double)ir * g[4];
    const double g2 = rr * rr - g[2] * g[2];
    const double ground = g2 > 0.0 ? g2 * rsqrt64(g2) : 0.0;
    *x = g[0] + ground * c.scu;
    *y = g[1] + ground * c.u;
    *r = rr;
}

__device__ __forceinline__ void dest_pixel_rc(const double* g, int side, int ir, int iu, double* x,
                                              double* y, double* r) {
    dest_pixel_col(g, dest_column(g, side, iu), ir, x, y, r);
}

__device__ __forceinline__ void dest_pixel_fast(const double* g, int side, long long p, int nu,
                                                double* x, double* y, double* r) {
    const int ir = (int)(p / nu);
    dest_pixel_rc(g, side, ir, (int)(p - (long long)ir * nu), x, y, r);
}

// Child regrid at destination point (x, y) with destination slant range r.
template <int INTERP>
__device__ __forceinline__ float2 regrid_fast(const FastSrc& f, double x, double y, double r,
                                              double kt) {
    const ChildTap c = child_tap<INTERP>(f, x, y, r, kt);
    return c.ok ? stencil_global<INTERP>(f, c) : make_float2(0.f, 0.f);
}

// Polar image -> point (x, y), remodulated by exp(+j k r) (ffbp.py:189-198).
template <int INTERP>
__device__ __forceinline__ float2 eval_fast(const FastSrc& f, double x, double y, double kt) {
    return regrid_fast<INTERP>(f, x, y, 0.0, kt);
}

template <int INTERP>
__global__ __launch_bounds__(256) void regrid_fast_kernel(const float2* __restrict__ src, Geo sg,
                                                          int snr, int snu, Geo dg, int nr, int nu,
                                                          int side, double k, int accumulate,
                                                          float2* __restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nr * nu) return;
    const double sgv[7] = {sg.ox, sg.oy, sg.oz, sg.r0, sg.dr, sg.u0, sg.du};
    const double dgv[7] = {dg.ox, dg.oy, dg.oz, dg.r0, dg.dr, dg.u0, dg.du};
    double x, y, r;
    dest_pixel_fast(dgv, side, p, nu, &x, &y, &r);
    const FastSrc f = make_fast_src(src, sgv, snr, snu, snu);
    float2 v = regrid_fast<INTERP>(f, x, y, r, k * INV_TWO_PI);
    if (accumulate) {
        const float2 o = out[p];
        v = make_float2(__fadd_rn(o.x, v.x), __fadd_rn(o.y, v.y));
    }
    out[p] = v;
}

template <int INTERP>
__global__ __launch_bounds__(256) void merge2_fast_kernel(const float2* __restrict__ left, Geo lg,
                                                          int lnr, int lnu,
                                                          const float2* __restrict__ right,
                                                          Geo rg, int rnr, int rnu, Geo dg,
                                                          int nr, int nu, int side, double k,
                                                          float2* __restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nr * nu) return;
    const double lgv[7] = {lg.ox, lg.oy, lg.oz, lg.r0, lg.dr, lg.u0, lg.du};
    const double rgv[7] = {rg.ox, rg.oy, rg.oz, rg.r0, rg.dr, rg.u0, rg.du};
    const double dgv[7] = {dg.ox, dg.oy, dg.oz, dg.r0, dg.dr, dg.u0, dg.du};
    double x, y, r;
    dest_pixel_fast(dgv, side, p, nu, &x, &y, &r);
    const double kt = k * INV_TWO_PI;
    const float2 a = regrid_fast<INTERP>(make_fast_src(left, lgv, lnr, lnu, lnu), x, y, r, kt);
    const float2 b = regrid_fast<INTERP>(make_fast_src(right, rgv, rnr, rnu, rnu), x, y, r, kt);
    out[p] = make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}

template <int INTERP>
__global__ __launch_bounds__(256) void batched_merge_fast_kernel(
        const int* __restrict__ node_int, const double* __restrict__ node_geo,
        const long long* __restrict__ offsets, float2* arena, const int* __restrict__ blocks,
        double k) {
    const int node = blocks[2 * blockIdx.x];
    const long long p = (long long)blocks[2 * blockIdx.x + 1] + threadIdx.x;
    const int* ni = node_int + 8 * node;
    const int nr = ni[2], nu = ni[3], side = ni[4], l = ni[5], r = ni[6];
    if (p >= (long long)nr * nu) return;
    double x, y, rr;
    dest_pixel_fast(node_geo + 7 * node, side, p, nu, &x, &y, &rr);
    const int* li = node_int + 8 * l;
    const int* ri = node_int + 8 * r;
    const double kt = k * INV_TWO_PI;
    const float2 a = regrid_fast<INTERP>(
        make_fast_src(arena + offsets[l], node_geo + 7 * l, li[2], li[3], li[3]), x, y, rr, kt);
    const float2 b = regrid_fast<INTERP>(
        make_fast_src(arena + offsets[r], node_geo + 7 * r, ri[2], ri[3], ri[3]), x, y, rr, kt);
    arena[offsets[node] + p] = make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}

// ---- LDS-tiled batched merge (split mode) -----------------------------------
// A workgroup owns a 32 x 32 tile of one parent node (4 pixels per thread, in
// rows t/32, t/32 + 8, ..).  The child-coordinate map is smooth over a tile,
// so the stencil origins of the tile's four corner pixels, widened by one
// cell, bound the tile's footprint in each child (~1050 samples at the
// lower levels, up to ~2050 at the root of an 8192-pulse aperture, where the
// children's origins are far apart); the workgroup copies that window into
// LDS once and the stencil
// sums gather from LDS instead of re-fetching ~1 KB per parent pixel through
// the vector cache.  A pixel whose stencil is not inside the window, or a
// window above MT_CAP samples, reads global memory with identical
// arithmetic, so results never depend on the path taken.
// The default launch shape: launch bounds (256, 3) and a window cap of
// MT_CAP samples (~35 KiB of LDS per workgroup, so 4 waves per SIMD).  Other
// shapes -- a smaller cap (less LDS, more resident workgroups, more tiles on
// the global-memory path) and other occupancy hints -- are instantiated for
// the autotuner; results do not depend on the shape (same arithmetic on
// either path).
constexpr int MT_ROWS = 32, MT_COLS = 32, MT_THREADS = 256;
constexpr int MT_PPT = MT_ROWS * MT_COLS / MT_THREADS;
constexpr int MT_CAP = 2176;
constexpr int MT_WAVES = 4;


struct Window {
    int r0, c0, h, w;  // rows [r0, r0 + h), columns [c0, c0 + w) of the child
    int pitch;         // LDS row pitch in samples (w, or NARROW_PITCH)
    bool lds;
};

// Narrow windows (at most NARROW_COLS columns: every merge level of the
// 8192-pulse stripmap plan below the root, whose windows are 41-42 x 25-32)
// are staged with a fixed row pitch of NARROW_PITCH samples, so that all 64
// taps of a stencil lie at compile-time offsets of one LDS address (7 x 33 +
// 7 = 238 units of 8 bytes, within ds_read2_b64's 8-bit offsets) and the
// per-row address additions disappear; the odd pitch keeps the two parent
// rows of a wave, one child row apart, off a common bank.
constexpr int NARROW_COLS = 32, NARROW_PITCH = 33;

// Window bounds from the corner stencil origins ir[4], iu[4] (INT_MAX if
// undefined), before any LDS cap: rows [r0, r0 + h), columns [c0, c0 + w)
// of the child, the origins' box widened by one cell, covering the stencils
// and clipped to the image; h = w = 0 when an origin is undefined or the
// window is narrower than one stencil (it holds no stencil).
struct WinBounds {
    int r0, c0, h, w;
};

template <int INTERP>
__device__ __forceinline__ WinBounds window_bounds(const int* ir, const int* iu, int nr, int nu) {
    constexpr int TAPS = FastShape<INTERP>::TAPS, FIRST = FastShape<INTERP>::FIRST;
    constexpr int LIM = 0x3fffffff;  // stencil origins of any real image are far below
    int rlo = ir[0], rhi = ir[0], clo = iu[0], chi = iu[0];
#pragma unroll
    for (int j = 1; j < 4; ++j) {
        rlo = min(rlo, ir[j]); rhi = max(rhi, ir[j]);
        clo = min(clo, iu[j]); chi = max(chi, iu[j]);
    }
    const bool defined = rlo > -LIM && rhi < LIM && clo > -LIM && chi < LIM;
    const int r0 = max(rlo - 1 - FIRST, 0), r1 = min(rhi + 1 + (TAPS - 1 - FIRST), nr - 1);
    const int c0 = max(clo - 1 - FIRST, 0), c1 = min(chi + 1 + (TAPS - 1 - FIRST), nu - 1);
    const int h = r1 - r0 + 1, wd = c1 - c0 + 1;
    const bool holds = defined && h >= TAPS && wd >= TAPS;
    WinBounds b;
    b.r0 = r0;
    b.c0 = c0;
    b.h = holds ? h : 0;
    b.w = holds ? wd : 0;
    return b;
}

// One entry of a merge window table (merge_windows_kernel): both children's
// window bounds and reciprocal spacings (1/dr, 1/du of child a, then b).
struct MergeWindowEntry {
    WinBounds a, b;
    double inv[4];
};
static_assert(sizeof(MergeWindowEntry) == 64, "64-byte table entries");

// Bounds read from a table, made safe for the child they are applied to
// (a table built for another block table then costs speed, not
// correctness): clipped to the nr x nu image and emptied when that leaves
// less than one stencil.  Scalar arithmetic.
template <int INTERP>
__device__ __forceinline__ WinBounds clip_bounds(WinBounds b, int nr, int nu) {
    constexpr int TAPS = FastShape<INTERP>::TAPS;
    b.r0 = max(b.r0, 0);
    b.c0 = max(b.c0, 0);
    b.h = min(b.h, nr - b.r0);
    b.w = min(b.w, nu - b.c0);
    if (b.h < TAPS || b.w < TAPS) b.h = b.w = 0;
    return b;
}

// The window as a launch with CAP samples of LDS per child stages it: in LDS
// when the bounds fit the cap, else read from global memory (lds false).
template <int CAP>
__device__ __forceinline__ Window cap_window(const WinBounds& b) {
    Window w;
    // the product is formed only once both sides are at most CAP
    w.lds = b.h > 0 && b.h <= CAP && b.w <= CAP && b.h * b.w <= CAP;
    w.r0 = b.r0;
    w.c0 = b.c0;
    w.h = w.lds ? b.h : 0;
    w.w = w.lds ? b.w : 0;
    w.pitch = w.w;
    return w;
}

template <int INTERP, int CAP = MT_CAP>
__device__ __forceinline__ Window corner_window(const int* ir, const int* iu, int nr, int nu) {
    return cap_window<CAP>(window_bounds<INTERP>(ir, iu, nr, nu));
}

// Switches both windows of a merge tile to the narrow layout when both are
// staged, at most NARROW_COLS wide and fit CAP samples at NARROW_PITCH.
template <int CAP>
__device__ __forceinline__ bool narrow_windows(Window& a, Window& b) {
    const bool n = a.lds && b.lds && a.w <= NARROW_COLS && b.w <= NARROW_COLS &&
                   a.h * NARROW_PITCH <= CAP && b.h * NARROW_PITCH <= CAP;
    if (n) a.pitch = b.pitch = NARROW_PITCH;
    return n;
}

// Stencil offsets (dr, du) of c in the window and whether the whole stencil
// lies in it: one unsigned compare per axis (w.h, w.w >= TAPS when w.lds;
// an undefined origin, INT_MAX or INT_MIN, lands far outside: the offsets
// are formed in unsigned arithmetic, which wraps).
template <int INTERP>
__device__ __forceinline__ bool in_window(const ChildTap& c, const Window& w, int* dr, int* du) {
    constexpr int TAPS = FastShape<INTERP>::TAPS, FIRST = FastShape<INTERP>::FIRST;
    const unsigned ur = (unsigned)c.ir - (unsigned)(FIRST + w.r0);
    const unsigned uc = (unsigned)c.iu - (unsigned)(FIRST + w.c0);
    *dr = (int)ur;
    *du = (int)uc;
    return w.lds && ur <= (unsigned)(w.h - TAPS) && uc <= (unsigned)(w.w - TAPS);
}

// Copy the window into LDS: wave v takes rows v, v + 4, .., lane j columns
// j, j + 64, ..; per-row steps are uniform pointer increments.  Windows up
// to 64 columns (the usual case) load STAGE_ROWS rows before storing any, so
// a wave waits for one batch of loads instead of once per row.
constexpr int STAGE_ROWS = 4;

// Window loads are global loads (ld_global_c64) also where the image pointer
// was read from a table in memory: as flat loads they could alias LDS, and
// the compiler would order them behind every LDS access of the wave.
// ROWS: rows per load batch (the evaluation, with one window and one pixel
// per thread, stages a whole window in one batch).
template <int ROWS = STAGE_ROWS>
__device__ __forceinline__ void stage_window(const FastSrc& f, const Window& w, float2* dst,
                                             int wave, int lane) {
    constexpr int WAVES = MT_THREADS / 64;
    wave = uniform_i32(wave);  // the row tests below are then scalar branches
    if (wave >= w.h) return;
    const float2* src = f.img + (long long)(w.r0 + wave) * f.ld + w.c0;
    const long long src_step = (long long)WAVES * f.ld;
    float2* d = dst + wave * w.pitch;
    const int dst_step = WAVES * w.pitch;
    if (w.w <= 64) {
        if (lane < w.w) {
            for (int i = wave; i < w.h; i += ROWS * WAVES) {
                // rows i + b * WAVES of this batch: nb of them lie in the
                // window; the loads of the others repeat its last row (no
                // branch around a load), only the nb rows are stored
                const int nb = min(ROWS, (w.h - i + WAVES - 1) / WAVES);
                float2 v[ROWS];
#pragma unroll
                for (int b = 0; b < ROWS; ++b)
                    v[b] = ld_global_c64(src + min(b, nb - 1) * src_step + lane);
#pragma unroll
                for (int b = 0; b < ROWS; ++b)
                    if (b < nb) d[b * dst_step + lane] = v[b];
                src += ROWS * src_step;
                d += ROWS * dst_step;
            }
        }
        return;
    }
    for (int i = wave; i < w.h; i += WAVES) {
        for (int j = lane; j < w.w; j += 64) d[j] = ld_global_c64(src + j);
        src += src_step;
        d += dst_step;
    }
}

// Both child windows of a merge tile with their loads in flight together:
// each batch issues STAGE_ROWS rows of window a and of window b before
// storing any (1.3% faster merges than staging one window after the other).
// Lanes beyond a window's width and rows past its height load its last
// column and row (valid addresses) and store into ``dump`` (one slot per
// lane) instead of the window: every store is unconditional, so the
// compiler moves no load into a branch where it would be waited for alone.
// Windows wider than 64 columns take stage_window.
//
// In the narrow layout (narrow_windows: both windows staged, at most
// NARROW_COLS wide, row pitch NARROW_PITCH) one load covers two window rows:
// lanes 0-31 take the columns of row u, lanes 32-63 those of row u + 1.
// Lanes past a window's width load its last column and store into the
// row's padding (columns w .. NARROW_PITCH - 1, which no stencil reads), and
// u is clamped to h - 2 so that both rows exist (a pair past the end stores
// rows h - 2 and h - 1 again, with the same values).  Row offsets and clamps
// are then wave-uniform: no per-lane address arithmetic besides the lane's
// fixed offsets, and half the load and store instructions.
__device__ __forceinline__ void stage_two_windows(const FastSrc& fa, const Window& wa, float2* da,
                                                  const FastSrc& fb, const Window& wb, float2* db,
                                                  float2* dump, bool narrow, int wave, int lane) {
    constexpr int WAVES = MT_THREADS / 64;
    static_assert(2 * NARROW_COLS == 64 && NARROW_PITCH > NARROW_COLS, "two rows per wave load");
    if (narrow) {
        wave = uniform_i32(wave);
        const int half = lane >> 5, col = lane & (NARROW_COLS - 1);
        const float2* const ia = fa.img + (long long)wa.r0 * fa.ld + wa.c0;
        const float2* const ib = fb.img + (long long)wb.r0 * fb.ld + wb.c0;
        const int oa = half * fa.ld + min(col, wa.w - 1), ob = half * fb.ld + min(col, wb.w - 1);
        const int os = half * NARROW_PITCH + col;
        const int rows = max(wa.h, wb.h);
        for (int i = 2 * wave; i < rows; i += 2 * STAGE_ROWS * WAVES) {
            float2 va[STAGE_ROWS], vb[STAGE_ROWS];
#pragma unroll
            for (int b = 0; b < STAGE_ROWS; ++b) {
                const int ua = min(i + 2 * b * WAVES, wa.h - 2), ub = min(i + 2 * b * WAVES, wb.h - 2);
                va[b] = ld_global_c64(ia + (long long)ua * fa.ld + oa);
                vb[b] = ld_global_c64(ib + (long long)ub * fb.ld + ob);
            }
#pragma unroll
            for (int b = 0; b < STAGE_ROWS; ++b) {
                const int ua = min(i + 2 * b * WAVES, wa.h - 2), ub = min(i + 2 * b * WAVES, wb.h - 2);
                da[ua * NARROW_PITCH + os] = va[b];
                db[ub * NARROW_PITCH + os] = vb[b];
            }
        }
        return;
    }
    if (!(wa.lds && wb.lds && wa.w <= 64 && wb.w <= 64)) {
        if (wa.lds) stage_window(fa, wa, da, wave, lane);
        if (wb.lds) stage_window(fb, wb, db, wave, lane);
        return;
    }
    wave = uniform_i32(wave);
    const float2* sa = fa.img + (long long)wa.r0 * fa.ld + wa.c0 + min(lane, wa.w - 1);
    const float2* sb = fb.img + (long long)wb.r0 * fb.ld + wb.c0 + min(lane, wb.w - 1);
    const int rows = max(wa.h, wb.h);
    for (int i = wave; i < rows; i += STAGE_ROWS * WAVES) {
        float2 va[STAGE_ROWS], vb[STAGE_ROWS];
#pragma unroll
        for (int b = 0; b < STAGE_ROWS; ++b) {
            const int ra = min(i + b * WAVES, wa.h - 1), rb = min(i + b * WAVES, wb.h - 1);
            va[b] = ld_global_c64(sa + (long long)ra * fa.ld);
            vb[b] = ld_global_c64(sb + (long long)rb * fb.ld);
        }
#pragma unroll
        for (int b = 0; b < STAGE_ROWS; ++b) {
            const int r = i + b * WAVES;
            *(r < wa.h && lane < wa.w ? da + r * wa.pitch + lane : dump + lane) = va[b];
            *(r < wb.h && lane < wb.w ? db + r * wb.pitch + lane : dump + lane) = vb[b];
        }
    }
}

// PITCH: the window's LDS row pitch when it is known at compile time (the
// narrow layout), else 0 (w.pitch).
template <int INTERP, int PITCH = 0>
__device__ __forceinline__ float2 tile_sample(const FastSrc& f, const ChildTap& c, const Window& w,
                                              const float2* sh, const float* kp) {
    // The window is clipped to the image, so a stencil inside it is valid:
    // only the global-memory path needs the validity test, which is restated
    // here for the indices of child_tap (INT_MAX on both axes when undefined)
    // and of child_tap_px (truncated, saturated on each axis).
    int dr, du;
    if (in_window<INTERP>(c, w, &dr, &du)) {
        const int pitch = PITCH > 0 ? PITCH : w.pitch;
        const float2* base = sh + dr * pitch + du;
        return stencil<INTERP>(c, [&](int a) { return base + a * pitch; }, kp);
    }
    const bool ok = c.ir != 0x7fffffff && c.iu != 0x7fffffff &&
                    InterpShape<INTERP>::valid(c.ir, f.nr) && InterpShape<INTERP>::valid(c.iu, f.nu);
    if (!ok) return make_float2(0.f, 0.f);
    return stencil_global<INTERP>(f, c, kp);
}

// LDS of one merge tile: both child windows (CAP samples each), the
// destination row and column terms, both windows' placement, the children's
// reciprocal spacings and a copy of the Kaiser weight polynomials (read as
// vector operands, they free the 36 scalar registers whose pressure
// otherwise spills scalars into vector lanes in the pixel loop).
template <int CAP>
struct MergeLdsCap {
    float2 sh[2 * CAP];
    double2 row[MT_ROWS];  // per tile row: slant range r, ground range sqrt(r^2 - z^2)
    double2 col[MT_COLS];  // per tile column: u, side * sqrt(1 - u^2) (dest_column)
    int win[2][6];         // windows a, b: r0, c0, h, w, pitch, staged
    int narrow;            // both windows in the narrow layout
    double inv[4];         // 1/dr, 1/du of both children
    float kpoly[4 * KPOLY_STRIDE];
    float2 dump[64];       // staging stores of lanes outside a window
};
using MergeLds = MergeLdsCap<MT_CAP>;

// One 32 x 32 tile of parent node ``node`` (all MT_THREADS threads call it;
// every thread reaches both barriers).  kp_src: the Kaiser weight
// polynomials (c_kaiser_poly).  win: the tile's entry of a
// merge_windows_kernel table (window bounds and reciprocal spacings of both
// children), or null to derive them here.  A caller that runs several tiles in a row must place a
// barrier between them.
template <int INTERP, bool COH = false, int CAP = MT_CAP>
__device__ __forceinline__ void merge_tile(const int* __restrict__ node_int,
                                           const double* __restrict__ node_geo,
                                           const long long* __restrict__ offsets, float2* arena,
                                           int node, int tile, double k, const float* kp_src,
                                           const MergeWindowEntry* __restrict__ win,
                                           MergeLdsCap<CAP>& lds) {
    // The second window's base offset is opaque to the compiler, so the
    // stencil reads address it from a register rather than through
    // per-instruction offset additions (ds_read2 offsets reach 2 KiB).
    int second = CAP;
    asm volatile("" : "+s"(second));
    float2* const sha = lds.sh;
    float2* const shb = lds.sh + second;
    const float* const kpoly = lds.kpoly;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // Node table entries as wave-uniform values (scalar loads and registers).
    const BPMI_CONST_AS int* tni = as_const(node_int);
    const BPMI_CONST_AS double* tgeo = as_const(node_geo);
    const BPMI_CONST_AS long long* toff = as_const(offsets);
    const int nr = uniform_i32(tni[8 * node + 2]), nu = uniform_i32(tni[8 * node + 3]);
    const int side = uniform_i32(tni[8 * node + 4]);
    const int l = uniform_i32(tni[8 * node + 5]), r = uniform_i32(tni[8 * node + 6]);
    const int lnr = uniform_i32(tni[8 * l + 2]), lnu = uniform_i32(tni[8 * l + 3]);
    const int rnr = uniform_i32(tni[8 * r + 2]), rnu = uniform_i32(tni[8 * r + 3]);
    const float2* const img_a = arena + uniform_i64(toff[l]);
    const float2* const img_b = arena + uniform_i64(toff[r]);
    float2* const out = arena + uniform_i64(toff[node]);
    double g[7], ga[7], gb[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        g[j] = uniform_f64(tgeo[7 * node + j]);
        ga[j] = uniform_f64(tgeo[7 * l + j]);
        gb[j] = uniform_f64(tgeo[7 * r + j]);
    }
    const int tiles_u = (nu + MT_COLS - 1) / MT_COLS;
    const int tr = tile / tiles_u, tu = tile - tr * tiles_u;
    const int row0 = tr * MT_ROWS + t / MT_COLS, col = tu * MT_COLS + (t % MT_COLS);
    constexpr int ROW_STEP = MT_THREADS / MT_COLS;

    const double kt = k * INV_TWO_PI;
    // The first wave alone computes the children's reciprocal spacings and,
    // without a window table, places both windows from the stencil origins
    // of the tile's corner pixels (clamped to the node; lanes 0..3 take one
    // corner each).  The second wave computes the destination row and column
    // terms.  The other waves then read the windows' placement instead of
    // deriving it.
    const bool table = win != nullptr;  // uniform
    if (uniform_i32(wave) == 0) {
        if (t < 4 * KPOLY_STRIDE) lds.kpoly[t] = kp_src[t];
        if (!table) {
            const FastSrc fa = make_fast_src(img_a, ga, lnr, lnu, lnu);
            const FastSrc fb = make_fast_src(img_b, gb, rnr, rnu, rnu);
            if (t == 0) {
                lds.inv[0] = fa.inv_dr; lds.inv[1] = fa.inv_du;
                lds.inv[2] = fb.inv_dr; lds.inv[3] = fb.inv_du;
            }
            int car = 0, cau = 0, cbr = 0, cbu = 0;
            if (t < 4) {
                const int crow = min(tr * MT_ROWS + (t >> 1) * (MT_ROWS - 1), nr - 1);
                const int ccol = min(tu * MT_COLS + (t & 1) * (MT_COLS - 1), nu - 1);
                double x, y, rr;
                dest_pixel_rc(g, side, crow, ccol, &x, &y, &rr);
                const ChildTap ca = child_tap<INTERP>(fa, x, y, rr, kt);
                const ChildTap cb = child_tap<INTERP>(fb, x, y, rr, kt);
                car = ca.ir; cau = ca.iu;
                cbr = cb.ir; cbu = cb.iu;
            }
            int ira[4], iua[4], irb[4], iub[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ira[j] = __builtin_amdgcn_readlane(car, j);
                iua[j] = __builtin_amdgcn_readlane(cau, j);
                irb[j] = __builtin_amdgcn_readlane(cbr, j);
                iub[j] = __builtin_amdgcn_readlane(cbu, j);
            }
            Window wa = corner_window<INTERP, CAP>(ira, iua, fa.nr, fa.nu);
            Window wb = corner_window<INTERP, CAP>(irb, iub, fb.nr, fb.nu);
            const int narrow = narrow_windows<CAP>(wa, wb) ? 1 : 0;
            if (t == 0) {
                const Window* w2[2] = {&wa, &wb};
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    lds.win[i][0] = w2[i]->r0; lds.win[i][1] = w2[i]->c0;
                    lds.win[i][2] = w2[i]->h; lds.win[i][3] = w2[i]->w;
                    lds.win[i][4] = w2[i]->pitch; lds.win[i][5] = w2[i]->lds ? 1 : 0;
                }
                lds.narrow = narrow;
            }
        }
    } else if (uniform_i32(wave) == 1) {
        static_assert(MT_ROWS + MT_COLS <= 64, "one wave computes the row and column terms");
        if (lane < MT_ROWS) {
            // the row terms of dest_pixel_col, once per tile row
            const double rr = g[3] + (double)(tr * MT_ROWS + lane) * g[4];
            const double g2 = rr * rr - g[2] * g[2];
            lds.row[lane] = make_double2(rr, g2 > 0.0 ? g2 * rsqrt64(g2) : 0.0);
        } else if (lane < MT_ROWS + MT_COLS) {
            // dest_column, once per tile column (columns past the node give
            // values no thread uses)
            const DestColumn c = dest_column(g, side, tu * MT_COLS + lane - MT_ROWS);
            lds.col[lane - MT_ROWS] = make_double2(c.u, c.scu);
        }
    }
    // With a table nothing below reads LDS before the staging barrier (the
    // windows come from the table), so the waves start staging at once and
    // one barrier orders the Kaiser polynomials, row and column terms and
    // windows; without one, the first wave's window placement is waited for.
    if (!table) __syncthreads();
    // the table entry through scalar loads (the cap test then runs in
    // scalar registers), else what the first wave left in LDS
    const BPMI_CONST_AS MergeWindowEntry* e = as_const(win);
    const FastSrc fa = make_fast_src_inv(img_a, ga, lnr, lnu, lnu,
                                         table ? e->inv[0] : uniform_f64(lds.inv[0]),
                                         table ? e->inv[1] : uniform_f64(lds.inv[1]));
    const FastSrc fb = make_fast_src_inv(img_b, gb, rnr, rnu, rnu,
                                         table ? e->inv[2] : uniform_f64(lds.inv[2]),
                                         table ? e->inv[3] : uniform_f64(lds.inv[3]));
    Window wa, wb;
    bool narrow;
    if (table) {
        WinBounds ba, bb;
        ba.r0 = e->a.r0; ba.c0 = e->a.c0; ba.h = e->a.h; ba.w = e->a.w;
        bb.r0 = e->b.r0; bb.c0 = e->b.c0; bb.h = e->b.h; bb.w = e->b.w;
        wa = cap_window<CAP>(clip_bounds<INTERP>(ba, lnr, lnu));
        wb = cap_window<CAP>(clip_bounds<INTERP>(bb, rnr, rnu));
        narrow = narrow_windows<CAP>(wa, wb);
    } else {
        Window* const w2[2] = {&wa, &wb};
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            w2[i]->r0 = uniform_i32(lds.win[i][0]); w2[i]->c0 = uniform_i32(lds.win[i][1]);
            w2[i]->h = uniform_i32(lds.win[i][2]); w2[i]->w = uniform_i32(lds.win[i][3]);
            w2[i]->pitch = uniform_i32(lds.win[i][4]);
            w2[i]->lds = uniform_i32(lds.win[i][5]) != 0;
        }
        narrow = uniform_i32(lds.narrow) != 0;
    }
    stage_two_windows(fa, wa, sha, fb, wb, shb, lds.dump, narrow, wave, lane);
    __syncthreads();
    if (col >= nu) return;
    const double2 cterm = lds.col[t % MT_COLS];
    DestColumn dc;
    dc.u = cterm.x;
    dc.scu = cterm.y;
    // The pixel loop, with the windows' LDS pitch fixed (narrow layout) or
    // read from the window: the same arithmetic either way.
    // output address of the thread's first pixel, stepped by ROW_STEP rows
    // (no per-pixel 64-bit multiply)
    float2* const out_first = out + (long long)row0 * nu + col;
    const long long out_step = (long long)ROW_STEP * nu;
    auto pixels = [&]<int PITCH>() {
        float2* o = out_first;
#pragma unroll 1
        for (int j = 0; j < MT_PPT; ++j, o += out_step) {
            const int row = row0 + j * ROW_STEP;
            if (row >= nr) break;
            // dest_pixel_col with the row terms from LDS (the same values)
            const double2 rg = lds.row[t / MT_COLS + j * ROW_STEP];
            const double rr = rg.x;
            const double x = g[0] + rg.y * dc.scu, y = g[1] + rg.y * dc.u;
            const ChildTap ca = child_tap_px<INTERP>(fa, x, y, rr, kt);
            const ChildTap cb = child_tap_px<INTERP>(fb, x, y, rr, kt);
            const float2 va = tile_sample<INTERP, PITCH>(fa, ca, wa, sha, kpoly);
            const float2 vb = tile_sample<INTERP, PITCH>(fb, cb, wb, shb, kpoly);
            st_c64<COH>(o, make_float2(__fadd_rn(va.x, vb.x), __fadd_rn(va.y, vb.y)));
        }
    };
    if (narrow)
        pixels.template operator()<NARROW_PITCH>();
    else
        pixels.template operator()<0>();
}

// wins: null, or the block table's window table (merge_windows_kernel).
template <int INTERP, int WAVES = MT_WAVES, int CAP = MT_CAP>
__global__ __launch_bounds__(MT_THREADS, WAVES) void batched_merge_tiled_kernel(
        const int* __restrict__ node_int, const double* __restrict__ node_geo,
        const long long* __restrict__ offsets, float2* arena, const int* __restrict__ blocks,
        double k, const MergeWindowEntry* __restrict__ wins) {
    __shared__ MergeLdsCap<CAP> lds;
    const int node = blocks[2 * blockIdx.x], tile = blocks[2 * blockIdx.x + 1];
    if (node < 0) return;  // padding entry of a die-ordered table (uniform)
    merge_tile<INTERP, false, CAP>(node_int, node_geo, offsets, arena, node, tile, k,
                                   &c_kaiser_poly[0][0], wins ? wins + blockIdx.x : nullptr,
                                   lds);
}

// Window table of a tiled-merge block table: for every (node, tile) entry
// the same corner-stencil window placement and the same reciprocal child
// spacings as merge_tile derives without a table (they depend on the plan's
// geometry only, not on the pulses).  Padding entries (node < 0) get empty
// windows.  One thread per entry.
template <int INTERP>
__global__ __launch_bounds__(256) void merge_windows_kernel(const int* __restrict__ node_int,
                                                            const double* __restrict__ node_geo,
                                                            const int* __restrict__ blocks,
                                                            int nblocks,
                                                            MergeWindowEntry* __restrict__ table) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nblocks) return;
    MergeWindowEntry& out = table[e];
    const int node = blocks[2 * e], tile = blocks[2 * e + 1];
    if (node < 0) {
        out.a = out.b = WinBounds{0, 0, 0, 0};
        out.inv[0] = out.inv[1] = out.inv[2] = out.inv[3] = 0.0;
        return;
    }
    const int* const ni = node_int + 8 * node;
    const int nr = ni[2], nu = ni[3], side = ni[4], l = ni[5], r = ni[6];
    const int lnr = node_int[8 * l + 2], lnu = node_int[8 * l + 3];
    const int rnr = node_int[8 * r + 2], rnu = node_int[8 * r + 3];
    double g[7], ga[7], gb[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        g[j] = node_geo[7 * node + j];
        ga[j] = node_geo[7 * l + j];
        gb[j] = node_geo[7 * r + j];
    }
    const FastSrc fa = make_fast_src(nullptr, ga, lnr, lnu, lnu);
    const FastSrc fb = make_fast_src(nullptr, gb, rnr, rnu, rnu);
    const int tiles_u = (nu + MT_COLS - 1) / MT_COLS;
    const int tr = tile / tiles_u, tu = tile - tr * tiles_u;
    int ira[4], iua[4], irb[4], iub[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int crow = min(tr * MT_ROWS + (j >> 1) * (MT_ROWS - 1), nr - 1);
        const int ccol = min(tu * MT_COLS + (j & 1) * (MT_COLS - 1), nu - 1);
        double x, y, rr;
        dest_pixel_rc(g, side, crow, ccol, &x, &y, &rr);
        // the stencil origins do not depend on the phase term (kt = 0)
        const ChildTap ca = child_tap<INTERP>(fa, x, y, rr, 0.0);
        const ChildTap cb = child_tap<INTERP>(fb, x, y, rr, 0.0);
        ira[j] = ca.ir; iua[j] = ca.iu;
        irb[j] = cb.ir; iub[j] = cb.iu;
    }
    out.a = window_bounds<INTERP>(ira, iua, lnr, lnu);
    out.b = window_bounds<INTERP>(irb, iub, rnr, rnu);
    out.inv[0] = fa.inv_dr; out.inv[1] = fa.inv_du;
    out.inv[2] = fb.inv_dr; out.inv[3] = fb.inv_du;
}

// ---- required-tile marking ----------------------------------------------------
// A parent pixel reads a child image only through its stencil, so the child
// pixels that a set of parent tiles can reach lie in the union of those
// tiles' stencil boxes.  The marking kernels locate the stencils with the
// same geometry (dest_pixel_col + child_tap, or the evaluation's child_tap
// at r = 0) as the kernels that read them, widen each stencil by one cell on
// every side and set the mask byte of every 32 x 32 tile the box touches.
// Every pixel with a defined stencil origin counts, also those whose stencil
// leaves the image (they read nothing; counting them only over-marks).
// Mask bytes are only ever set to 1, so concurrent writers agree.
constexpr int RT_FAR = 1 << 29;  // stencil origins of any real image are far below

// Stencil box of one defined origin (ir, iu), widened by one cell and clipped
// to an nr x nu image, merged into [rlo, rhi] x [clo, chi].
template <int INTERP>
__device__ __forceinline__ void box_add(int ir, int iu, int nr, int nu, int& rlo, int& rhi,
                                        int& clo, int& chi) {
    constexpr int TAPS = FastShape<INTERP>::TAPS, FIRST = FastShape<INTERP>::FIRST;
    if (ir < -RT_FAR || ir > RT_FAR || iu < -RT_FAR || iu > RT_FAR) return;
    const int r0 = max(ir - FIRST - 1, 0), r1 = min(ir - FIRST + TAPS, nr - 1);
    const int c0 = max(iu - FIRST - 1, 0), c1 = min(iu - FIRST + TAPS, nu - 1);
    if (r0 > r1 || c0 > c1) return;
    rlo = min(rlo, r0); rhi = max(rhi, r1);
    clo = min(clo, c0); chi = max(chi, c1);
}

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Sets the mask bytes of the tiles of an image with tiles_u tile columns
// that cover rows [rlo, rhi] x columns [clo, chi] (threads t0, t0 + step, ..).
__device__ __forceinline__ void mark_box(unsigned char* mask, int tiles_u, int rlo, int rhi,
                                         int clo, int chi, int t0, int step) {
    if (rlo > rhi || clo > chi) return;
    const int tr0 = rlo / MT_ROWS, tr1 = rhi / MT_ROWS;
    const int tc0 = clo / MT_COLS, tc1 = chi / MT_COLS;
    const int ntc = tc1 - tc0 + 1, n = (tr1 - tr0 + 1) * ntc;
    for (int i = t0; i < n; i += step) {
        const int a = i / ntc;
        mask[(long long)(tr0 + a) * tiles_u + tc0 + (i - a * ntc)] = 1;
    }
}

// One workgroup per (parent node, tile) entry; entries whose own mask byte is
// clear return at once.  Pixel placement as in merge_tile.
template <int INTERP>
__global__ __launch_bounds__(MT_THREADS) void required_merge_tiles_kernel(
        const int* __restrict__ node_int, const double* __restrict__ node_geo,
        const long long* __restrict__ mask_offsets, unsigned char* masks,
        const int* __restrict__ blocks) {
    __shared__ int box[2][4];
    const int node = blocks[2 * blockIdx.x], tile = blocks[2 * blockIdx.x + 1];
    if (node < 0 || masks[mask_offsets[node] + tile] == 0) return;
    const int t = threadIdx.x, lane = t & 63;
    if (t < 8) box[t >> 2][t & 3] = (t & 1) ? -1 : 0x7fffffff;
    __syncthreads();
    const int* ni = node_int + 8 * node;
    const int nr = ni[2], nu = ni[3], side = ni[4];
    const int child[2] = {ni[5], ni[6]};
    double g[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) g[j] = node_geo[7 * node + j];
    int rlo[2], rhi[2], clo[2], chi[2];
    FastSrc f[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int* ci = node_int + 8 * child[c];
        f[c] = make_fast_src(nullptr, node_geo + 7 * child[c], ci[2], ci[3], ci[3]);
        rlo[c] = clo[c] = 0x7fffffff;
        rhi[c] = chi[c] = -1;
    }
    const int tiles_u = (nu + MT_COLS - 1) / MT_COLS;
    const int tr = tile / tiles_u, tu = tile - tr * tiles_u;
    const int row0 = tr * MT_ROWS + t / MT_COLS, col = tu * MT_COLS + (t % MT_COLS);
    constexpr int ROW_STEP = MT_THREADS / MT_COLS;
    if (col < nu) {
        const DestColumn dc = dest_column(g, side, col);
        for (int j = 0; j < MT_PPT; ++j) {
            const int row = row0 + j * ROW_STEP;
            if (row >= nr) break;
            double x, y, rr;
            dest_pixel_col(g, dc, row, &x, &y, &rr);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const ChildTap ct = child_tap<INTERP>(f[c], x, y, rr, 0.0);
                box_add<INTERP>(ct.ir, ct.iu, f[c].nr, f[c].nu, rlo[c], rhi[c], clo[c], chi[c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int a = wave_min_i(rlo[c]), b = wave_max_i(rhi[c]);
        const int d = wave_min_i(clo[c]), e = wave_max_i(chi[c]);
        if (lane == 0) {
            atomicMin(&box[c][0], a); atomicMax(&box[c][1], b);
            atomicMin(&box[c][2], d); atomicMax(&box[c][3], e);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; ++c)
        mark_box(masks + mask_offsets[child[c]], (f[c].nu + MT_COLS - 1) / MT_COLS, box[c][0],
                 box[c][1], box[c][2], box[c][3], t, MT_THREADS);
}

// One thread per Cartesian output pixel of the split-mode grid evaluation of
// an nr x nu view of a root image (row_offset, col_offset into the root, whose
// mask has root_tiles_u tile columns); each wave marks its pixels' box.
template <int INTERP>
__global__ __launch_bounds__(256) void required_eval_tiles_kernel(
        Geo g, int nr, int nu, int row_offset, int col_offset, int root_tiles_u, double x0,
        double dx, int nx, double y0, double dy, int ny, unsigned char* mask) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int rlo = 0x7fffffff, rhi = -1, clo = 0x7fffffff, chi = -1;
    if (p < (long long)nx * ny) {
        const int iy = (int)(p / nx), ix = (int)(p - (long long)iy * nx);
        const double gv[7] = {g.ox, g.oy, g.oz, g.r0, g.dr, g.u0, g.du};
        const FastSrc f = make_fast_src(nullptr, gv, nr, nu, nu);
        const ChildTap ct = child_tap<INTERP>(f, x0 + ix * dx, y0 + iy * dy, 0.0, 0.0);
        box_add<INTERP>(ct.ir, ct.iu, nr, nu, rlo, rhi, clo, chi);
    }
    rlo = wave_min_i(rlo); rhi = wave_max_i(rhi);
    clo = wave_min_i(clo); chi = wave_max_i(chi);
    if (rlo <= rhi)
        mark_box(mask, root_tiles_u, rlo + row_offset, rhi + row_offset, clo + col_offset,
                 chi + col_offset, lane, 64);
}

template <int INTERP>
__global__ __launch_bounds__(256) void eval_points_fast_kernel(const float2* __restrict__ img, Geo g,
                                                               int nr, int nu, int ld,
                                                               const double* __restrict__ xyz,
                                                               long long n, double k,
                                                               float2* __restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double gv[7] = {g.ox, g.oy, g.oz, g.r0, g.dr, g.u0, g.du};
    out[p] = eval_fast<INTERP>(make_fast_src(img, gv, nr, nu, ld), xyz[3 * p], xyz[3 * p + 1],
                               k * INV_TWO_PI);
}

template <int INTERP>
__global__ __launch_bounds__(256) void eval_grid_fast_kernel(const float2* __restrict__ img, Geo g,
                                                             int nr, int nu, int ld, double x0,
                                                             double dx, int nx, double y0,
                                                             double dy, int ny, double k,
                                                             float2* __restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nx * ny) return;
    const int iy = (int)(p / nx), ix = (int)(p - (long long)iy * nx);
    const double gv[7] = {g.ox, g.oy, g.oz, g.r0, g.dr, g.u0, g.du};
    out[p] = eval_fast<INTERP>(make_fast_src(img, gv, nr, nu, ld), x0 + ix * dx, y0 + iy * dy,
                               k * INV_TWO_PI);
}

template <int INTERP>
__global__ __launch_bounds__(256) void eval_grid_multi_fast_kernel(
        const bpmi_eval_job* __restrict__ jobs, const int* __restrict__ blocks) {
    const bpmi_eval_job& j = jobs[blocks[2 * blockIdx.x]];
    const long long p = (long long)blocks[2 * blockIdx.x + 1] + threadIdx.x;
    if (p >= (long long)j.nx * j.ny) return;
    const int iy = (int)(p / j.nx), ix = (int)(p - (long long)iy * j.nx);
    ((float2*)j.out)[p] = eval_fast<INTERP>(make_fast_src((const float2*)j.img, j.geo, j.nr, j.nu, j.ld),
                                            j.x0 + ix * j.dx, j.y0 + iy * j.dy, j.k * INV_TWO_PI);
}

// ---- LDS-tiled Cartesian evaluation (split mode) ------------------------------
// A workgroup owns ET_X x ET_Y output pixels (one per thread, x fastest: 64
// pixels along x, the across-track axis the polar range runs along, so the
// tile spans few angular cells).  As in the tiled merge, the stencil origins
// of the tile's corners, widened by one cell, bound its footprint in the
// polar image; that window (up to ET_CAP samples) is staged in LDS and each
// pixel's 8x8 (4x4) stencil reads it; a pixel outside the window, or a
// window above ET_CAP, reads global memory with identical arithmetic, so the
// output equals eval_grid_fast_kernel's bit for bit.
constexpr int ET_X = 64, ET_Y = 4, ET_CAP = 4096;
// rows per staging batch: a 64-column window of ET_CAP samples has 64 rows,
// 16 per wave, all in one batch
constexpr int ET_STAGE_ROWS = 16;
static_assert(ET_X * ET_Y == MT_THREADS, "one pixel per thread");

struct EvalLds {
    float2 sh[ET_CAP];
    int corner[4][2];
    float kpoly[4 * KPOLY_STRIDE];
};

// Tile (tx, ty) of an nx x ny grid over the polar view f; every thread of
// the workgroup calls it (two barriers).  SROWS: window rows per staging
// batch (the persistent kernel, whose registers its other bodies share,
// keeps the merge's batches).
template <int INTERP, int SROWS = ET_STAGE_ROWS>
__device__ __forceinline__ void eval_tile(const FastSrc& f, double x0, double dx, int nx, double y0,
                                          double dy, int ny, double kt, int tx, int ty,
                                          float2* __restrict__ out, EvalLds& lds) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < 4) {
        const int cx = min(tx * ET_X + (t & 1) * (ET_X - 1), nx - 1);
        const int cy = min(ty * ET_Y + (t >> 1) * (ET_Y - 1), ny - 1);
        const ChildTap c = child_tap<INTERP>(f, x0 + cx * dx, y0 + cy * dy, 0.0, kt);
        lds.corner[t][0] = c.ir;
        lds.corner[t][1] = c.iu;
    }
    if (t < 4 * KPOLY_STRIDE) lds.kpoly[t] = c_kaiser_poly[0][t];
    __syncthreads();
    int ir[4], iu[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ir[j] = uniform_i32(lds.corner[j][0]);
        iu[j] = uniform_i32(lds.corner[j][1]);
    }
    const Window w = corner_window<INTERP, ET_CAP>(ir, iu, f.nr, f.nu);
    if (w.lds) stage_window<SROWS>(f, w, lds.sh, wave, lane);
    __syncthreads();
    const int ix = tx * ET_X + (t % ET_X), iy = ty * ET_Y + t / ET_X;
    if (ix < nx && iy < ny) {
        const ChildTap c = child_tap_px<INTERP>(f, x0 + ix * dx, y0 + iy * dy, 0.0, kt);
        out[(long long)iy * nx + ix] = tile_sample<INTERP>(f, c, w, lds.sh, lds.kpoly);
    }
}

template <int INTERP>
__global__ __launch_bounds__(MT_THREADS) void eval_grid_tiled_kernel(
        const float2* __restrict__ img, Geo g, int nr, int nu, int ld, double x0, double dx, int nx,
        double y0, double dy, int ny, double k, float2* __restrict__ out) {
    __shared__ EvalLds lds;
    const int tiles_x = (nx + ET_X - 1) / ET_X;
    const int ty = blockIdx.x / tiles_x, tx = blockIdx.x - ty * tiles_x;
    const double gv[7] = {g.ox, g.oy, g.oz, g.r0, g.dr, g.u0, g.du};
    eval_tile<INTERP>(make_fast_src(img, gv, nr, nu, ld), x0, dx, nx, y0, dy, ny, k * INV_TWO_PI,
                      tx, ty, out, lds);
}

// Many jobs: blocks[b] = (job, tile) with tiles numbered row-major over the
// job's grid (bpmi_eval_tiles).
template <int INTERP>
__global__ __launch_bounds__(MT_THREADS) void eval_grid_multi_tiled_kernel(
        const bpmi_eval_job* __restrict__ jobs, const int* __restrict__ blocks) {
    __shared__ EvalLds lds;
    const bpmi_eval_job& j = jobs[blocks[2 * blockIdx.x]];
    const int tile = blocks[2 * blockIdx.x + 1];
    const int tiles_x = (j.nx + ET_X - 1) / ET_X;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    eval_tile<INTERP>(make_fast_src((const float2*)j.img, j.geo, j.nr, j.nu, j.ld), j.x0, j.dx,
                      j.nx, j.y0, j.dy, j.ny, j.k * INV_TWO_PI, tx, ty, (float2*)j.out, lds);
}

// Chebyshev interpolation (8 nodes) of the normalised 8-tap Kaiser weights
// w_a(t), a = 0..3, t in [0,1), read from the FP64 table exactly as the FP64
// paths do, converted to a power series in s = 2t - 1 and split into even
// coefficients (s^0, s^2, s^4, s^6) and odd coefficients (s^1, .., s^7).
// Taps 4..7 are the mirror images w_{7-a}(s) = w_a(-s).  The weights are
// normalised at every node, and interpolation on one node set is linear, so
// the eight polynomials sum to exactly 1 before their coefficients are
// rounded to FP32: the error is no gain error but a zero-sum ripple of at
// most 1.3e-6 per weight (degree 8 would give 1.9e-7 at one more FMA per tap
// pair and axis; degree 6, 1.3e-5).
static void fit_kaiser_poly(const double* table, float out[4][KPOLY_STRIDE]) {
    constexpr int N = 2 * KPOLY_ODD;
    static_assert(KPOLY_EVEN == KPOLY_ODD, "degree 2 * KPOLY_ODD - 1: N = 2 * KPOLY_ODD nodes");
    double f[4][N];
    double sn[N];
    for (int j = 0; j < N; ++j) {
        sn[j] = cos(M_PI * (j + 0.5) / N);
        const double t = 0.5 * (sn[j] + 1.0);
        double w[8], sum = 0.0;
        for (int a = 0; a < 8; ++a) {
            double x = fabs(t - (double)(a - 3)) * KAISER_TABLE_RES;
            int i = (int)x;
            double fr = x - (double)i;
            if (i > KAISER_TABLE_LEN - 2) i = KAISER_TABLE_LEN - 2;
            w[a] = table[i] + fr * (table[i + 1] - table[i]);
            sum += w[a];
        }
        for (int a = 0; a < 4; ++a) f[a][j] = w[a] / sum;
    }
    // power-basis coefficients of the Chebyshev polynomials T_0..T_8
    double T[N][N] = {};
    T[0][0] = 1.0;
    T[1][1] = 1.0;
    for (int k = 2; k < N; ++k)
        for (int m = 0; m < N; ++m)
            T[k][m] = (m > 0 ? 2.0 * T[k - 1][m - 1] : 0.0) - T[k - 2][m];
    for (int a = 0; a < 4; ++a) {
        double poly[N] = {};
        for (int k = 0; k < N; ++k) {
            double c = 0.0;
            for (int j = 0; j < N; ++j) c += f[a][j] * cos(M_PI * k * (j + 0.5) / N);
            c *= (k == 0 ? 1.0 : 2.0) / N;
            for (int m = 0; m < N; ++m) poly[m] += c * T[k][m];
        }
        for (int i = 0; i < KPOLY_EVEN; ++i) out[a][i] = (float)poly[2 * i];
        for (int i = 0; i < KPOLY_ODD; ++i) out[a][KPOLY_EVEN + i] = (float)poly[2 * i + 1];
    }
}

// ---- phase-LUT direct BP (precision_probe.cu:106-164) ----------------------
__global__ void fill_lut_kernel(int samples, double rs, double ddr, double k, float2* lut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= samples) return;
    double s, c;
    sincos(k * (rs + (double)i * ddr), &s, &c);
    lut[i] = make_float2((float)c, (float)s);
}

__global__ __launch_bounds__(256) void bp_polar_lut_kernel(const float2* __restrict__ data, int S,
                                                           const double* __restrict__ pos,
                                                           const double* __restrict__ weight,
                                                           const double* __restrict__ phase,
                                                           int pb, int pe, Geo g, int nr, int nu,
                                                           int side, double rs, double ddr,
                                                           double k, const float2* __restrict__ lut,
                                                           float2* __restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nr * nu) return;
    const int ir = (int)(p / nu), iu = (int)(p - (long long)ir * nu);
    const double r = g.r0 + ir * g.dr, u = g.u0 + iu * g.du;
    double x, y;
    polar_pixel(g.ox, g.oy, g.oz, r, u, side, &x, &y);
    double rsd, rcd;
    sincos(-k * r, &rsd, &rcd);
    const float ref_s = (float)rsd, ref_c = (float)rcd;
    float2 total = out[p];
    for (int n = pb; n < pe; ++n) {
        const double dx = x - pos[3 * n], dy = y - pos[3 * n + 1], dz = pos[3 * n + 2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double q = (d - rs) / ddr;
        int i;
        if (!floor_index(q, &i) || i < 1 || i + 2 >= S) continue;
        const double f = q - (double)i;
        float w4[4];
        cubic_weights((float)f, w4);
        const float2* row = data + (size_t)n * S;
        float2 v = make_float2(0.f, 0.f);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float2 t = row[i + a - 1];
            v.x += w4[a] * t.x;
            v.y += w4[a] * t.y;
        }
        float is, ic;
        sincosf((float)(k * ddr * f + (phase ? phase[n] : 0.0)), &is, &ic);
        const float2 base = lut[i];
        const float bc = base.x * ic - base.y * is;
        const float bs = base.x * is + base.y * ic;
        const float c = bc * ref_c - bs * ref_s;
        const float s = bc * ref_s + bs * ref_c;
        const float w = weight ? (float)weight[n] : 1.f;
        total.x += w * (v.x * c - v.y * s);
        total.y += w * (v.x * s + v.y * c);
    }
    out[p] = total;
}

// ---- utilities ---------------------------------------------------------------
__global__ void add_kernel(float2* __restrict__ left, const float2* __restrict__ right, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float2 a = left[p], b = right[p];
    left[p] = make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y));
}

__global__ void cast_kernel(const double2* __restrict__ in, float2* __restrict__ out, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double2 z = in[p];
    out[p] = make_float2((float)z.x, (float)z.y);
}

__global__ void crop_kernel(const float2* __restrict__ in, int src_nu, int rb, int ub, int nr,
                            int nu, float2* __restrict__ out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)nr * nu) return;
    const int ir = (int)(p / nu), iu = (int)(p - (long long)ir * nu);
    out[p] = in[(size_t)(ir + rb) * src_nu + iu + ub];
}

__global__ void sincos_probe_kernel(const float* t, float* s, float* c, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    sincos_turns(t[p], &s[p], &c[p]);
}

// FP32 FMA issue-rate probe: each thread runs FMA_CHAINS independent chains
// x <- x * m + c, 16 steps per iteration, and stores their sum (so nothing
// is folded away).  FLOPs per launch: blocks * 256 * iters * 16 * chains * 2.
constexpr int FMA_CHAINS = 8;
__global__ __launch_bounds__(256) void fma_peak_kernel(float* out, int iters, float m, float c) {
    float a[FMA_CHAINS];
#pragma unroll
    for (int j = 0; j < FMA_CHAINS; ++j) a[j] = (float)(threadIdx.x + j) * 1e-3f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int j = 0; j < FMA_CHAINS; ++j) a[j] = fmaf(a[j], m, c);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < FMA_CHAINS; ++j) s += a[j];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// ---- persistent FFBP: one launch for leaves, every merge level and the
// Cartesian evaluation ---------------------------------------------------------
// Workgroups claim work items in list order from one ticket counter.  The
// list is topologically ordered (every item's inputs come from items earlier
// in the list), so a claimed item only ever waits for items already claimed
// by running workgroups: the schedule cannot deadlock, whatever the number
// of resident workgroups.  A merge tile waits until both children have
// counted all their tiles complete; an evaluation block waits for its root.
// Node images pass between items through device-scope (sc1) stores, which
// write through to memory without the cache-wide L2 write-back / invalidate
// of an agent-scope fence (measured 5x slower on the 8192-pulse FFBP), and
// ordinary loads (ld_c64<true>: why they are coherent here): every wave waits
// for its stores to complete before the item's completion is counted, and a
// consumer reads a node only after it has seen the node's count.  Every wait
// is bounded by a wall-clock limit: a wait that runs out sets the abort flag,
// after which every remaining item is skipped and all workgroups drain, so
// the launch always ends; the host reads the flag.
enum { PW_LEAF = 0, PW_MERGE = 1, PW_EVAL = 2 };
enum { PS_TICKET = 0, PS_ABORT = 1, PS_NODES = 2 };

struct PersistentArgs {
    const int* node_int;
    const double* node_geo;
    const long long* offsets;
    float2* arena;
    const bpmi_node_source* sources;
    BpArgs leaf;  // single-patch leaf inputs (when sources is null)
    int tc_log2;
    const bpmi_eval_job* jobs;
    const int4* items;
    int nitems;
    const int* node_tiles;
    const int* job_root;
    const int* job_blocks;
    int* sync;  // ticket, abort flag, done counts of the nodes, then of the jobs
    int nodes;
    unsigned long long* stamps;  // [0] first start, [1 + j] completion of job j
    unsigned long long spin_ticks;
    double k;
};

// Counters of the sync array, addressed as global memory (a flat atomic
// would also wait for the wave's LDS accesses).
typedef __attribute__((address_space(1))) int global_i32;
__device__ __forceinline__ int load_count(const int* p) {
    return __hip_atomic_load((global_i32*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int add_count(int* p, int v) {
    return __hip_atomic_fetch_add((global_i32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ BpArgs uniform_bp_args(const BpArgs& v) {
    BpArgs a;
    a.data = uniform_ptr(v.data);
    a.pos = uniform_ptr(v.pos);
    a.weight = uniform_ptr(v.weight);
    a.phase = uniform_ptr(v.phase);
    a.samples = uniform_i32(v.samples);
    a.pulse_begin = uniform_i32(v.pulse_begin);
    a.pulse_end = uniform_i32(v.pulse_end);
    a.range_start = uniform_f64(v.range_start);
    a.data_dr = uniform_f64(v.data_dr);
    a.k = uniform_f64(v.k);
    a.flags = uniform_i32(v.flags);
    a.out = uniform_ptr(v.out);
    return a;
}

__device__ __forceinline__ SrcArgs uniform_src_args(const SrcArgs& v) {
    SrcArgs s;
    s.a0 = uniform_f64(v.a0); s.a1 = uniform_f64(v.a1); s.a2 = uniform_f64(v.a2);
    s.b0 = uniform_f64(v.b0); s.b1 = uniform_f64(v.b1); s.b2 = uniform_f64(v.b2);
    s.b3 = uniform_f64(v.b3);
    s.rows = uniform_i32(v.rows); s.cols = uniform_i32(v.cols); s.side = uniform_i32(v.side);
    s.xyz = uniform_ptr(v.xyz);
    s.count = uniform_i64(v.count);
    return s;
}

// The tables live in device memory (``tab``) and are read after each item
// starts; the loop's own state is kept in LDS and re-read where needed, so
// nothing stays live in scalar registers across an item body (the merge
// body alone fills the scalar register file).
//
// Contention: a device-scope atomic on one address is serialised (~50 ns
// each), so one claim and one completion count per item would cost more than
// the items themselves.  Workgroups therefore claim PW_CHUNK consecutive
// items per ticket and keep completion counts pending in LDS, adding them
// to the node's counter when the next item belongs to another counter, at
// the end of the chunk, and before the workgroup itself has to wait (so a
// count that a waiter needs is never held back by the waiter).  A
// dependency seen complete is remembered, so a run of tiles of one parent
// checks its children once.
constexpr int PW_CHUNK = 16;

struct PersistentLoop {
    const PersistentArgs* tab;
    const int4* items;
    int* sync;
    unsigned long long* stamps;
    int nitems;
    int item, last;
    int go;
    int pend_idx, pend_cnt;  // pending completion count and its counter index
    int ready0, ready1;      // last dependency pair seen complete
    int ready_kind, ready_idx;  // the last item found runnable (kind, node or job)
};

// A fresh read of loop state in LDS (v must live in LDS): addressed as LDS
// explicitly, since a volatile read through a generic pointer is a flat,
// system-coherent load.
template <typename T>
__device__ __forceinline__ T lds_reload(const T& v) {
    return *(const volatile __attribute__((address_space(3))) T*)&v;
}


// Thread 0: add the pending completion count to its counter (a job's
// counter reaching its block count stamps the job's completion time).
// A node's count is added without waiting for the old value (nothing reads
// it here).
__device__ __forceinline__ void flush_pending(PersistentLoop& st) {
    if (st.pend_cnt <= 0) return;
    const BPMI_CONST_AS PersistentArgs& pa = *as_const(st.tab);
    int* counter = st.sync + PS_NODES + st.pend_idx;
    const int job = st.pend_idx - pa.nodes;
    if (job >= 0) {
        const int old = add_count(counter, st.pend_cnt);
        if (old + st.pend_cnt == as_const(pa.job_blocks)[job]) st.stamps[1 + job] = wall_clock64();
    } else {
        add_count(counter, st.pend_cnt);
    }
    st.pend_cnt = 0;
}

// Thread 0: true when item w may run, false after an abort.
__device__ __forceinline__ bool persistent_wait(PersistentLoop& st, int4 w) {
    // another tile of the item found runnable last: no table reads
    if (w.x == st.ready_kind && w.y == st.ready_idx) return true;
    const BPMI_CONST_AS PersistentArgs& pa = *as_const(st.tab);
    int n0 = -1, n1 = -1;
    if (w.x == PW_MERGE) {
        n0 = as_const(pa.node_int)[8 * w.y + 5];
        n1 = as_const(pa.node_int)[8 * w.y + 6];
    } else if (w.x == PW_EVAL) {
        n0 = as_const(pa.job_root)[w.y];
    }
    if (n0 < 0 || (n0 == st.ready0 && n1 == st.ready1)) {
        st.ready_kind = w.x;
        st.ready_idx = w.y;
        return true;
    }
    const int* done = st.sync + PS_NODES;
    const int need0 = as_const(pa.node_tiles)[n0];
    const int need1 = n1 >= 0 ? as_const(pa.node_tiles)[n1] : 0;
    const unsigned long long t0 = wall_clock64();
    for (bool first = true;; first = false) {
        if (load_count(done + n0) >= need0 && (n1 < 0 || load_count(done + n1) >= need1)) {
            st.ready0 = n0;
            st.ready1 = n1;
            st.ready_kind = w.x;
            st.ready_idx = w.y;
            return true;
        }
        if (first) flush_pending(st);
        if (load_count(st.sync + PS_ABORT)) return false;
        if (wall_clock64() - t0 > pa.spin_ticks) {
            __hip_atomic_store((global_i32*)(st.sync + PS_ABORT), 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(8);
    }
}

// 4 waves per SIMD: the leaf and merge bodies share 128 VGPRs (3 waves, with
// no spill, measured slower; 5 waves spill heavily).
constexpr int PW_WAVES = 4;

template <int LEAF_INTERP, int MERGE_INTERP>
__global__ __launch_bounds__(MT_THREADS, PW_WAVES) void ffbp_persistent_kernel(
        const PersistentArgs* tab_arg, const int4* items_arg, int nitems_arg, int* sync_arg,
        unsigned long long* stamps_arg) {
    static_assert(SP_THREADS == MT_THREADS, "one workgroup shape for every item kind");
    __shared__ union {
        SplitLds<LEAF_INTERP> split;
        MergeLds merge;
        EvalLds eval;
    } lds;
    static_assert(sizeof(EvalLds) <= sizeof(MergeLds), "the evaluation tile adds no LDS");
    __shared__ PersistentLoop st;
    const int t = threadIdx.x;
    if (t == 0) {
        st.tab = tab_arg;
        st.items = items_arg;
        st.sync = sync_arg;
        st.stamps = stamps_arg;
        st.nitems = nitems_arg;
        st.pend_idx = -1;
        st.pend_cnt = 0;
        st.ready0 = st.ready1 = -2;
        st.ready_kind = st.ready_idx = -1;
        atomicMin(stamps_arg, wall_clock64());
    }
    // Per item: thread 0 has checked the item's dependencies (st.go) before
    // a barrier; every wave runs the body; a barrier; thread 0 counts the
    // item (pending), moves on and checks the next item's dependencies; a
    // barrier.  A wave waits for its stores (wait_vector_memory) only where
    // thread 0 may publish counts next: before the last item of a chunk
    // (the chunk's counts are flushed) and before an item whose successor
    // has another node or kind (the pending count is flushed when the
    // counter changes, and only such a successor can have to wait, which
    // flushes too); items of one node keep their stores in flight.
    for (;;) {
        // claim the next chunk; an aborted launch claims nothing more
        if (t == 0) {
            const bool abort = load_count(st.sync + PS_ABORT) != 0;
            const int first = abort ? st.nitems : add_count(st.sync + PS_TICKET, PW_CHUNK);
            st.item = first;
            st.last = min(first + PW_CHUNK, st.nitems);
            if (first < st.nitems) {
                const BPMI_CONST_AS int* wi = as_const((const int*)st.items) + 4 * first;
                st.go = persistent_wait(st, make_int4(wi[0], wi[1], wi[2], 0)) ? 1 : 0;
            }
        }
        __syncthreads();
        if (uniform_i32(lds_reload(st.item)) >= uniform_i32(lds_reload(st.nitems))) break;
        for (;;) {
            const int item = uniform_i32(lds_reload(st.item));
            const int last = uniform_i32(lds_reload(st.last));
            const BPMI_CONST_AS PersistentArgs* tab = as_const(uniform_ptr(lds_reload(st.tab)));
            const BPMI_CONST_AS int* wv =
                as_const((const int*)uniform_ptr(lds_reload(st.items))) + 4 * item;
            const int4 w = make_int4(uniform_i32(wv[0]), uniform_i32(wv[1]), uniform_i32(wv[2]), 0);
            const bool go = uniform_i32(lds_reload(st.go)) != 0;
            if (go) {
                const int* node_int = uniform_ptr(tab->node_int);
                const double* node_geo = uniform_ptr(tab->node_geo);
                const long long* offsets = uniform_ptr(tab->offsets);
                float2* arena = uniform_ptr(tab->arena);
                if (w.x == PW_LEAF) {
                    BpArgs a;
                    SrcArgs s{};
                    const bpmi_node_source* sources = uniform_ptr(tab->sources);
                    const LeafTables lt{node_int, node_geo, offsets, arena, nullptr, sources};
                    if (!sources) {
                        const BPMI_CONST_AS BpArgs& la = tab->leaf;
                        a.data = la.data; a.pos = la.pos; a.weight = la.weight; a.phase = la.phase;
                        a.samples = la.samples; a.range_start = la.range_start;
                        a.data_dr = la.data_dr; a.k = la.k;
                    }
                    leaf_args(lt, w.y, &a, &s);
                    a = uniform_bp_args(a);
                    s = uniform_src_args(s);
                    bp_split_tile<LEAF_INTERP, SRC_POLAR, true, true>(
                        a, s, uniform_i32(tab->tc_log2), w.z, lds.split);
                } else if (w.x == PW_MERGE) {
                    merge_tile<MERGE_INTERP, true>(node_int, node_geo, offsets, arena, w.y, w.z,
                                                   uniform_f64(tab->k), &c_kaiser_poly[0][0],
                                                   nullptr, lds.merge);
                } else {
                    // one 64 x 4-pixel tile of the job's grid, as the level
                    // programs' tiled evaluation computes it
                    const BPMI_CONST_AS bpmi_eval_job* jp = as_const(tab->jobs) + w.y;
                    const int nx = uniform_i32(jp->nx), ny = uniform_i32(jp->ny);
                    const int tiles_x = (nx + ET_X - 1) / ET_X;
                    const int ty = w.z / tiles_x, tx = w.z - ty * tiles_x;
                    double geo[7];
#pragma unroll
                    for (int j = 0; j < 7; ++j) geo[j] = uniform_f64(jp->geo[j]);
                    const FastSrc f = make_fast_src(uniform_ptr((const float2*)jp->img), geo,
                                                    uniform_i32(jp->nr), uniform_i32(jp->nu),
                                                    uniform_i32(jp->ld));
                    eval_tile<MERGE_INTERP, STAGE_ROWS>(f, uniform_f64(jp->x0), uniform_f64(jp->dx), nx,
                                            uniform_f64(jp->y0), uniform_f64(jp->dy), ny,
                                            uniform_f64(jp->k) * INV_TWO_PI, tx, ty,
                                            uniform_ptr((float2*)jp->out), lds.eval);
                }
            }
            // the successor, and whether this item's stores must complete in
            // every wave before the barrier (see above)
            const int next = item + 1;
            int4 wn = make_int4(-1, -1, -1, 0);
            if (next < last)
                wn = make_int4(uniform_i32(wv[4]), uniform_i32(wv[5]), uniform_i32(wv[6]), 0);
            if (go && (next >= last || wn.x != w.x || wn.y != w.y)) wait_vector_memory();
            __syncthreads();
            if (t == 0) {
                if (go) {
                    const int idx = w.x == PW_EVAL ? as_const(st.tab)->nodes + w.y : w.y;
                    if (idx != st.pend_idx) {
                        flush_pending(st);
                        st.pend_idx = idx;
                    }
                    st.pend_cnt += 1;
                }
                st.item = next;
                if (next < last) st.go = persistent_wait(st, wn) ? 1 : 0;
            }
            __syncthreads();
            if (next >= last) break;
        }
        // the chunk's last item has drained its stores in every wave
        if (t == 0) flush_pending(st);
    }
}

// Workspace layout: the argument block, then the int32 counters (ticket,
// abort flag, node completion counts, job completion counts), then the u64
// time stamps (first start, completion of each job).
constexpr size_t PW_ARGS_BYTES = 512;
static_assert(sizeof(PersistentArgs) <= PW_ARGS_BYTES, "argument block");
__host__ __device__ inline size_t pw_sync_ints(int nodes, int njobs) {
    return (size_t)PS_NODES + (size_t)nodes + (size_t)njobs;
}
__host__ __device__ inline size_t pw_stamps_offset(int nodes, int njobs) {
    return (PW_ARGS_BYTES + 4 * pw_sync_ints(nodes, njobs) + 7) & ~(size_t)7;
}

// Writes the argument block and clears the counters and time stamps.
__global__ void persistent_reset_kernel(PersistentArgs args, PersistentArgs* tab, int nsync,
                                        int nstamps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *tab = args;
    if (i < nsync) args.sync[i] = 0;
    if (i < nstamps) args.stamps[i] = i == 0 ? ~0ull : 0ull;
}

template <int L, int M>
static int persistent_grid_for(int* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st) return st;
    st = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                   "hipDeviceGetAttribute");
    if (st) return st;
    st = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                       &per_cu, reinterpret_cast<const void*>(ffbp_persistent_kernel<L, M>),
                       MT_THREADS, 0),
                   "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (st) return st;
    *grid = (per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
    return BPMI_OK;
}

static int persistent_interps(int leaf_interp, int merge_interp) {
    if (leaf_interp == BPMI_INTERP_NONE)
        return set_error(BPMI_ERR_UNSUPPORTED,
                         "a persistent work list without leaves is switched off (its one GPU run "
                         "did not finish; cause not found)");
    if (leaf_interp != BPMI_INTERP_LAGRANGE4 && leaf_interp != BPMI_INTERP_LINEAR)
        return set_error(BPMI_ERR_UNSUPPORTED, "persistent leaves support lagrange4 and linear");
    if (merge_interp != BPMI_INTERP_LAGRANGE4 && merge_interp != BPMI_INTERP_KAISER8)
        return set_error(BPMI_ERR_UNSUPPORTED, "persistent merges support lagrange4 and kaiser8");
    if (merge_interp == BPMI_INTERP_KAISER8 && !g_kaiser_table)
        return set_error(BPMI_ERR_INVALID, "kaiser8 needs bpmi_set_kaiser_table() first");
    return BPMI_OK;
}

// Calls f.template operator()<LEAF, MERGE>() for the runtime interpolators.
template <typename F>
static int dispatch_persistent(int leaf_interp, int merge_interp, F&& f) {
    const bool lin = leaf_interp == BPMI_INTERP_LINEAR, kai = merge_interp == BPMI_INTERP_KAISER8;
    if (lin) return kai ? f.template operator()<INTERP_LINEAR, INTERP_KAISER8>()
                        : f.template operator()<INTERP_LINEAR, INTERP_LAGRANGE4>();
    return kai ? f.template operator()<INTERP_LAGRANGE4, INTERP_KAISER8>()
               : f.template operator()<INTERP_LAGRANGE4, INTERP_LAGRANGE4>();
}

// ---- dispatch helpers --------------------------------------------------------
static int check_mode_interp(int mode, int interp) {
    if (mode < 0 || mode > BPMI_MODE_SPLIT)
        return set_error(BPMI_ERR_INVALID, "mode must be one of fp64, fp32, mixed, dfloat, split");
    if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_KAISER8)
        return set_error(BPMI_ERR_INVALID, "2-D interpolation supports lagrange4 and kaiser8");
    if (interp == BPMI_INTERP_KAISER8 && !g_kaiser_table)
        return set_error(BPMI_ERR_INVALID, "kaiser8 needs bpmi_set_kaiser_table() first");
    return BPMI_OK;
}

// Calls F.template operator()<MODE, INTERP>() for the runtime (mode, interp).
template <typename F>
static void dispatch2(int mode, int interp, F&& f) {
    const bool kai = interp == BPMI_INTERP_KAISER8;
    switch (mode) {
        case BPMI_MODE_FP64:
            kai ? f.template operator()<MODE_FP64, INTERP_KAISER8>()
                : f.template operator()<MODE_FP64, INTERP_LAGRANGE4>();
            break;
        case BPMI_MODE_FP32:
            kai ? f.template operator()<MODE_FP32, INTERP_KAISER8>()
                : f.template operator()<MODE_FP32, INTERP_LAGRANGE4>();
            break;
        default:
            kai ? f.template operator()<MODE_MIXED, INTERP_KAISER8>()
                : f.template operator()<MODE_MIXED, INTERP_LAGRANGE4>();
            break;
    }
}

}  // namespace bpmi

using namespace bpmi;

// The merge launch shapes: (occupancy hint in waves per SIMD, window cap)
// pairs.  The cap sets the LDS per workgroup (16 B per cap sample plus 752
// B): 2176 -> 35568 B (4 workgroups per CU), 1664 -> 27376 B and 1920 ->
// 31472 B (5), 1088 -> 18160 B (8); the hint caps the registers (Kaiser-8
// keeps its 96 VGPRs up to 5 waves and spills beyond; Lagrange-4 needs 69,
// which 7 waves allow), and the smaller of the two limits is the occupancy.
// On the 8192-pulse scene the windows (samples per child) are ~1025-1092 at
// the lower levels, up to 1344 at the last but one and 1804 (median) to 2142
// at the root, so the 5-wave shapes win below the root (~10% per launch) and
// the default shape at the root, where a tile whose window exceeds the cap
// reads global memory (a 32752 B workgroup, cap 2000, measured like 4
// waves: it does not fit 5 per CU).
#define BPMI_MERGE_SHAPE_LIST(X) X(4, 2176) X(5, 1664) X(5, 1920) X(7, 1088)

static bool merge_shape_known(int waves, int cap) {
#define BPMI_MERGE_SHAPE(W, C) \
    if (waves == W && cap == C) return true;
    BPMI_MERGE_SHAPE_LIST(BPMI_MERGE_SHAPE)
#undef BPMI_MERGE_SHAPE
    return false;
}

static const char* merge_shape_names() {
#define BPMI_MERGE_SHAPE(W, C) " (" #W ", " #C ")"
    return BPMI_MERGE_SHAPE_LIST(BPMI_MERGE_SHAPE);
#undef BPMI_MERGE_SHAPE
}

template <int INTERP>
static void launch_merge_shape(int waves, int cap, int nblocks, hipStream_t hs,
                               const int* node_int, const double* node_geo,
                               const long long* offsets, float2* ar, const int* blocks, double k,
                               const MergeWindowEntry* wins) {
#define BPMI_MERGE_SHAPE(W, C)                                                         \
    if (waves == W && cap == C) {                                                      \
        hipLaunchKernelGGL((batched_merge_tiled_kernel<INTERP, W, C>), dim3(nblocks),  \
                           dim3(MT_THREADS), 0, hs, node_int, node_geo, offsets, ar,   \
                           blocks, k, wins);                                           \
        return;                                                                        \
    }
    BPMI_MERGE_SHAPE_LIST(BPMI_MERGE_SHAPE)
#undef BPMI_MERGE_SHAPE
}

extern "C" {

int bpmi_set_kaiser_table(const double* host_table, int len) {
    if (!host_table || len != KAISER_TABLE_LEN)
        return set_error(BPMI_ERR_INVALID, "kaiser table must hold %d doubles", KAISER_TABLE_LEN);
    // one allocation: FP32 rounding of the table, then the FP64 table
    const size_t f32_bytes = sizeof(float) * KAISER_TABLE_LEN;
    const size_t bytes = f32_bytes + sizeof(double) * KAISER_TABLE_LEN;
    unsigned char host[sizeof(float) * KAISER_TABLE_LEN + sizeof(double) * KAISER_TABLE_LEN];
    float* h32 = (float*)host;
    for (int i = 0; i < KAISER_TABLE_LEN; ++i) h32[i] = (float)host_table[i];
    memcpy(host + f32_bytes, host_table, sizeof(double) * KAISER_TABLE_LEN);
    if (!g_kaiser_table) {
        int st = check_hip(hipMalloc((void**)&g_kaiser_table, bytes), "hipMalloc(kaiser table)");
        if (st) return st;
    }
    float poly[4][KPOLY_STRIDE];
    fit_kaiser_poly(host_table, poly);
    int st = check_hip(hipMemcpyToSymbol(HIP_SYMBOL(c_kaiser_poly), poly, sizeof(poly)),
                       "hipMemcpyToSymbol(kaiser polynomials)");
    if (st) return st;
    return check_hip(hipMemcpy(g_kaiser_table, host, bytes, hipMemcpyHostToDevice),
                     "hipMemcpy(kaiser table)");
}

int bpmi_regrid(const void* src, const double* src_geo, int snr, int snu, const double* dst_geo,
                int nr, int nu, int side, double k, int mode, int interp, int flags, void* out,
                void* stream) {
    int st = check_mode_interp(mode, interp);
    if (st) return st;
    if (!src || !out || !src_geo || !dst_geo) return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0 || snr <= 0 || snu <= 0)
        return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    const Geo sg = make_geo(src_geo), dg = make_geo(dst_geo);
    const dim3 grid(grid_1d((long long)nr * nu, 256));
    const int acc = (flags & BPMI_ACCUMULATE) ? 1 : 0;
    hipStream_t hs = (hipStream_t)stream;
    const float* kt = g_kaiser_table;
    const float2* s = (const float2*)src;
    float2* o = (float2*)out;
    if (mode == BPMI_MODE_SPLIT) {
        if (interp == BPMI_INTERP_KAISER8)
            hipLaunchKernelGGL(regrid_fast_kernel<INTERP_KAISER8>, grid, dim3(256), 0, hs, s, sg,
                               snr, snu, dg, nr, nu, side, k, acc, o);
        else
            hipLaunchKernelGGL(regrid_fast_kernel<INTERP_LAGRANGE4>, grid, dim3(256), 0, hs, s, sg,
                               snr, snu, dg, nr, nu, side, k, acc, o);
        return check_launch("regrid_fast_kernel");
    }
    auto l = [&]<int M, int I>() {
        hipLaunchKernelGGL((regrid_kernel<M, I>), grid, dim3(256), 0, hs, s, sg, snr, snu, dg, nr,
                           nu, side, k, acc, o, kt);
    };
    dispatch2(mode, interp, l);
    return check_launch("regrid_kernel");
}

int bpmi_merge2(const void* left, const double* left_geo, int lnr, int lnu, const void* right,
                const double* right_geo, int rnr, int rnu, const double* dst_geo, int nr, int nu,
                int side, double k, int mode, int interp, void* out, void* stream) {
    int st = check_mode_interp(mode, interp);
    if (st) return st;
    if (!left || !right || !out || !left_geo || !right_geo || !dst_geo)
        return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0 || lnr <= 0 || lnu <= 0 || rnr <= 0 || rnu <= 0)
        return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    const float2* a = (const float2*)left;
    const float2* b = (const float2*)right;
    const Geo ag = make_geo(left_geo), bg = make_geo(right_geo), dg = make_geo(dst_geo);
    float2* o = (float2*)out;
    const float* kt = g_kaiser_table;
    const dim3 g(grid_1d((long long)nr * nu, 256));
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (interp == BPMI_INTERP_KAISER8)
            hipLaunchKernelGGL(merge2_fast_kernel<INTERP_KAISER8>, g, dim3(256), 0, hs, a, ag, lnr,
                               lnu, b, bg, rnr, rnu, dg, nr, nu, side, k, o);
        else
            hipLaunchKernelGGL(merge2_fast_kernel<INTERP_LAGRANGE4>, g, dim3(256), 0, hs, a, ag,
                               lnr, lnu, b, bg, rnr, rnu, dg, nr, nu, side, k, o);
        return check_launch("merge2_fast_kernel");
    }
    auto l = [&]<int M, int I>() {
        hipLaunchKernelGGL((merge2_kernel<M, I>), g, dim3(256), 0, hs, a, ag, lnr, lnu, b, bg,
                           rnr, rnu, dg, nr, nu, side, k, o, kt);
    };
    dispatch2(mode, interp, l);
    return check_launch("merge2_kernel");
}

int bpmi_ffbp_merges(const int* node_int, const double* node_geo, const long long* offsets,
                     void* arena, const int* blocks, int nblocks, double k, int mode, int interp,
                     void* stream) {
    int st = check_mode_interp(mode, interp);
    if (st) return st;
    if (nblocks <= 0) return BPMI_OK;
    if (!node_int || !node_geo || !offsets || !arena || !blocks)
        return set_error(BPMI_ERR_INVALID, "null table");
    float2* ar = (float2*)arena;
    const float* kt = g_kaiser_table;
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (interp == BPMI_INTERP_KAISER8)
            hipLaunchKernelGGL(batched_merge_fast_kernel<INTERP_KAISER8>, dim3(nblocks), dim3(256),
                               0, hs, node_int, node_geo, offsets, ar, blocks, k);
        else
            hipLaunchKernelGGL(batched_merge_fast_kernel<INTERP_LAGRANGE4>, dim3(nblocks),
                               dim3(256), 0, hs, node_int, node_geo, offsets, ar, blocks, k);
        return check_launch("batched_merge_fast_kernel");
    }
    auto l = [&]<int M, int I>() {
        hipLaunchKernelGGL((batched_merge_kernel<M, I>), dim3(nblocks), dim3(256), 0, hs, node_int,
                           node_geo, offsets, ar, blocks, k, kt);
    };
    dispatch2(mode, interp, l);
    return check_launch("batched_merge_kernel");
}

int bpmi_merge_tiles(int nr, int nu, int* tiles) {
    if (!tiles || nr <= 0 || nu <= 0) return set_error(BPMI_ERR_INVALID, "bad node shape");
    const long long t = (long long)((nr + MT_ROWS - 1) / MT_ROWS) * ((nu + MT_COLS - 1) / MT_COLS);
    if (t > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "node too large");
    *tiles = (int)t;
    return BPMI_OK;
}

int bpmi_ffbp_merges_tiled(const int* node_int, const double* node_geo, const long long* offsets,
                           void* arena, const int* blocks, int nblocks, double k, int interp,
                           void* stream) {
    return bpmi_ffbp_merges_tiled_ex(node_int, node_geo, offsets, arena, blocks, nblocks, k,
                                     interp, MT_WAVES, MT_CAP, stream);
}

int bpmi_ffbp_merges_tiled_ex(const int* node_int, const double* node_geo,
                              const long long* offsets, void* arena, const int* blocks,
                              int nblocks, double k, int interp, int waves, int window_cap,
                              void* stream) {
    return bpmi_ffbp_merges_tiled_win(node_int, node_geo, offsets, arena, blocks, nblocks, k,
                                      interp, waves, window_cap, nullptr, stream);
}

int bpmi_ffbp_merge_windows(const int* node_int, const double* node_geo, const int* blocks,
                            int nblocks, int interp, void* windows, void* stream) {
    int st = check_mode_interp(BPMI_MODE_SPLIT, interp);
    if (st) return st;
    if (nblocks <= 0) return BPMI_OK;
    if (!node_int || !node_geo || !blocks || !windows)
        return set_error(BPMI_ERR_INVALID, "null table");
    hipStream_t hs = (hipStream_t)stream;
    const dim3 grid(grid_1d(nblocks, 256));
    if (interp == BPMI_INTERP_KAISER8)
        hipLaunchKernelGGL(merge_windows_kernel<INTERP_KAISER8>, grid, dim3(256), 0, hs, node_int,
                           node_geo, blocks, nblocks, (MergeWindowEntry*)windows);
    else
        hipLaunchKernelGGL(merge_windows_kernel<INTERP_LAGRANGE4>, grid, dim3(256), 0, hs,
                           node_int, node_geo, blocks, nblocks, (MergeWindowEntry*)windows);
    return check_launch("merge_windows_kernel");
}

int bpmi_ffbp_merges_tiled_win(const int* node_int, const double* node_geo,
                               const long long* offsets, void* arena, const int* blocks,
                               int nblocks, double k, int interp, int waves, int window_cap,
                               const void* windows, void* stream) {
    int st = check_mode_interp(BPMI_MODE_SPLIT, interp);
    if (st) return st;
    if (!merge_shape_known(waves, window_cap))
        return set_error(BPMI_ERR_INVALID, "merge shape (waves, window_cap) must be one of%s",
                         merge_shape_names());
    if (nblocks <= 0) return BPMI_OK;
    if (!node_int || !node_geo || !offsets || !arena || !blocks)
        return set_error(BPMI_ERR_INVALID, "null table");
    float2* ar = (float2*)arena;
    hipStream_t hs = (hipStream_t)stream;
    if (interp == BPMI_INTERP_KAISER8)
        launch_merge_shape<INTERP_KAISER8>(waves, window_cap, nblocks, hs, node_int, node_geo,
                                           offsets, ar, blocks, k,
                                           (const MergeWindowEntry*)windows);
    else
        launch_merge_shape<INTERP_LAGRANGE4>(waves, window_cap, nblocks, hs, node_int, node_geo,
                                             offsets, ar, blocks, k,
                                             (const MergeWindowEntry*)windows);
    return check_launch("batched_merge_tiled_kernel");
}

int bpmi_ffbp_required_tiles(const int* node_int, const double* node_geo,
                             const long long* mask_offsets, void* masks, const int* blocks,
                             int nblocks, int interp, void* stream) {
    int st = check_mode_interp(BPMI_MODE_SPLIT, interp);
    if (st) return st;
    if (nblocks <= 0) return BPMI_OK;
    if (!node_int || !node_geo || !mask_offsets || !masks || !blocks)
        return set_error(BPMI_ERR_INVALID, "null table");
    unsigned char* m = (unsigned char*)masks;
    hipStream_t hs = (hipStream_t)stream;
    if (interp == BPMI_INTERP_KAISER8)
        hipLaunchKernelGGL(required_merge_tiles_kernel<INTERP_KAISER8>, dim3(nblocks),
                           dim3(MT_THREADS), 0, hs, node_int, node_geo, mask_offsets, m, blocks);
    else
        hipLaunchKernelGGL(required_merge_tiles_kernel<INTERP_LAGRANGE4>, dim3(nblocks),
                           dim3(MT_THREADS), 0, hs, node_int, node_geo, mask_offsets, m, blocks);
    return check_launch("required_merge_tiles_kernel");
}

int bpmi_eval_required_tiles(const double* geo, int nr, int nu, int row_offset, int col_offset,
                             int root_nr, int root_nu, double x0, double dx, int nx, double y0,
                             double dy, int ny, int interp, void* mask, void* stream) {
    int st = check_mode_interp(BPMI_MODE_SPLIT, interp);
    if (st) return st;
    if (!geo || !mask) return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0 || nx <= 0 || ny <= 0 || row_offset < 0 || col_offset < 0 ||
        row_offset + nr > root_nr || col_offset + nu > root_nu)
        return set_error(BPMI_ERR_INVALID, "bad image shape");
    const Geo g = make_geo(geo);
    const int tiles_u = (root_nu + MT_COLS - 1) / MT_COLS;
    unsigned char* m = (unsigned char*)mask;
    const dim3 gr(grid_1d((long long)nx * ny, 256));
    hipStream_t hs = (hipStream_t)stream;
    if (interp == BPMI_INTERP_KAISER8)
        hipLaunchKernelGGL(required_eval_tiles_kernel<INTERP_KAISER8>, gr, dim3(256), 0, hs, g, nr,
                           nu, row_offset, col_offset, tiles_u, x0, dx, nx, y0, dy, ny, m);
    else
        hipLaunchKernelGGL(required_eval_tiles_kernel<INTERP_LAGRANGE4>, gr, dim3(256), 0, hs, g,
                           nr, nu, row_offset, col_offset, tiles_u, x0, dx, nx, y0, dy, ny, m);
    return check_launch("required_eval_tiles_kernel");
}

int bpmi_persistent_workspace_bytes(int nodes, int njobs, size_t* bytes) {
    if (!bytes || nodes < 0 || njobs < 0) return set_error(BPMI_ERR_INVALID, "bad sizes");
    *bytes = pw_stamps_offset(nodes, njobs) + 8 * (size_t)(1 + njobs);
    return BPMI_OK;
}

int bpmi_persistent_grid(int leaf_interp, int merge_interp, int* grid) {
    if (!grid) return set_error(BPMI_ERR_INVALID, "null grid");
    int st = persistent_interps(leaf_interp, merge_interp);
    if (st) return st;
    return dispatch_persistent(leaf_interp, merge_interp,
                               [&]<int L, int M>() { return persistent_grid_for<L, M>(grid); });
}

int bpmi_ffbp_persistent(const bpmi_persistent_schedule* sc, int leaf_interp, int merge_interp,
                         int grid, void* stream) {
    if (!sc) return set_error(BPMI_ERR_INVALID, "null schedule");
    int st = persistent_interps(leaf_interp, merge_interp);
    if (st) return st;
    if (sc->nitems <= 0) return BPMI_OK;
    if (!sc->node_int || !sc->node_geo || !sc->offsets || !sc->arena || !sc->items ||
        !sc->node_tiles || !sc->workspace)
        return set_error(BPMI_ERR_INVALID, "null table or workspace");
    if (sc->nodes <= 0 || sc->njobs < 0) return set_error(BPMI_ERR_INVALID, "bad node/job counts");
    if (sc->njobs > 0 && (!sc->jobs || !sc->job_root || !sc->job_blocks))
        return set_error(BPMI_ERR_INVALID, "null evaluation table");
    if (!sc->sources && (!sc->data || !sc->pos || sc->samples <= 0))
        return set_error(BPMI_ERR_INVALID, "leaf inputs missing (sources or data/pos/samples)");
    if (sc->tile_cols_log2 < 3 || sc->tile_cols_log2 > 6)
        return set_error(BPMI_ERR_INVALID, "tile_cols_log2 must be 3..6");
    if (!(sc->max_wait_s > 0.0) || sc->max_wait_s > 1.0e6)
        return set_error(BPMI_ERR_INVALID, "max_wait_s must be in (0, 1e6]");
    if (grid <= 0) {
        st = bpmi_persistent_grid(leaf_interp, merge_interp, &grid);
        if (st) return st;
    }
    char* ws = (char*)sc->workspace;
    PersistentArgs pa;
    pa.node_int = sc->node_int;
    pa.node_geo = sc->node_geo;
    pa.offsets = sc->offsets;
    pa.arena = (float2*)sc->arena;
    pa.sources = sc->sources;
    pa.leaf = BpArgs{(const float2*)sc->data, sc->pos, sc->weight, sc->phase, sc->samples, 0, 0,
                     sc->range_start, sc->data_dr, sc->k, 0, nullptr};
    pa.tc_log2 = sc->tile_cols_log2;
    pa.jobs = sc->jobs;
    pa.items = (const int4*)sc->items;
    pa.nitems = sc->nitems;
    pa.node_tiles = sc->node_tiles;
    pa.job_root = sc->job_root;
    pa.job_blocks = sc->job_blocks;
    pa.sync = (int*)(ws + PW_ARGS_BYTES);
    pa.nodes = sc->nodes;
    pa.stamps = (unsigned long long*)(ws + pw_stamps_offset(sc->nodes, sc->njobs));
    // the wait bound in ticks of the device wall clock, at the rate the
    // device reports for it
    int khz = 0;
    st = bpmi_wall_clock_khz(&khz);
    if (st) return st;
    pa.spin_ticks = (unsigned long long)(sc->max_wait_s * 1.0e3 * (double)khz);
    pa.k = sc->k;
    hipStream_t hs = (hipStream_t)stream;
    const int nsync = (int)pw_sync_ints(sc->nodes, sc->njobs);
    const int nstamps = 1 + sc->njobs;
    const int nreset = nsync > nstamps ? nsync : nstamps;
    PersistentArgs* tab = (PersistentArgs*)ws;
    hipLaunchKernelGGL(persistent_reset_kernel, dim3(grid_1d(nreset, 256)), dim3(256), 0, hs, pa,
                       tab, nsync, nstamps);
    st = check_launch("persistent_reset_kernel");
    if (st) return st;
    return dispatch_persistent(leaf_interp, merge_interp, [&]<int L, int M>() {
        hipLaunchKernelGGL((ffbp_persistent_kernel<L, M>), dim3(grid), dim3(MT_THREADS), 0, hs,
                           (const PersistentArgs*)tab, pa.items, pa.nitems, pa.sync, pa.stamps);
        return check_launch("ffbp_persistent_kernel");
    });
}

int bpmi_persistent_aborted(const void* workspace, int nodes, int njobs, int* aborted) {
    if (!workspace || !aborted || nodes < 0 || njobs < 0)
        return set_error(BPMI_ERR_INVALID, "bad argument");
    int v = 0;
    const int* sync = (const int*)((const char*)workspace + PW_ARGS_BYTES);
    int st = check_hip(hipMemcpy(&v, sync + PS_ABORT, sizeof(int), hipMemcpyDeviceToHost),
                       "hipMemcpy(abort flag)");
    if (st) return st;
    *aborted = v;
    return BPMI_OK;
}

int bpmi_persistent_stamps(const void* workspace, int nodes, int njobs,
                           unsigned long long* stamps) {
    if (!workspace || !stamps || nodes < 0 || njobs < 0)
        return set_error(BPMI_ERR_INVALID, "bad argument");
    const char* src = (const char*)workspace + pw_stamps_offset(nodes, njobs);
    return check_hip(hipMemcpy(stamps, src, 8 * (size_t)(1 + njobs), hipMemcpyDeviceToHost),
                     "hipMemcpy(stamps)");
}

int bpmi_eval_points(const void* img, const double* geo, int nr, int nu, int ld, const double* xyz,
                     long long npoints, double k, int mode, int interp, void* out, void* stream) {
    int st = check_mode_interp(mode, interp);
    if (st) return st;
    if (npoints <= 0) return BPMI_OK;
    if (!img || !geo || !xyz || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0 || ld < nu) return set_error(BPMI_ERR_INVALID, "bad image shape");
    const float2* im = (const float2*)img;
    const Geo g = make_geo(geo);
    float2* o = (float2*)out;
    const float* kt = g_kaiser_table;
    const dim3 gr(grid_1d(npoints, 256));
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (interp == BPMI_INTERP_KAISER8)
            hipLaunchKernelGGL(eval_points_fast_kernel<INTERP_KAISER8>, gr, dim3(256), 0, hs, im, g,
                               nr, nu, ld, xyz, npoints, k, o);
        else
            hipLaunchKernelGGL(eval_points_fast_kernel<INTERP_LAGRANGE4>, gr, dim3(256), 0, hs, im,
                               g, nr, nu, ld, xyz, npoints, k, o);
        return check_launch("eval_points_fast_kernel");
    }
    auto l = [&]<int M, int I>() {
        hipLaunchKernelGGL((eval_points_kernel<M, I>), gr, dim3(256), 0, hs, im, g, nr, nu, ld,
                           xyz, npoints, k, o, kt);
    };
    dispatch2(mode, interp, l);
    return check_launch("eval_points_kernel");
}

int bpmi_eval_grid(const void* img, const double* geo, int nr, int nu, int ld, double x0,
                   double dx, int nx, double y0, double dy, int ny, double k, int mode, int interp,
                   void* out, void* stream) {
    int st = check_mode_interp(mode, interp);
    if (st) return st;
    if (!img || !geo || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0 || ld < nu || nx <= 0 || ny <= 0)
        return set_error(BPMI_ERR_INVALID, "bad image shape");
    const float2* im = (const float2*)img;
    const Geo g = make_geo(geo);
    float2* o = (float2*)out;
    const float* kt = g_kaiser_table;
    const dim3 gr(grid_1d((long long)nx * ny, 256));
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        const long long tiles = (long long)((nx + ET_X - 1) / ET_X) * ((ny + ET_Y - 1) / ET_Y);
        if (tiles > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "grid too large");
        const dim3 gt((unsigned)tiles);
        if (interp == BPMI_INTERP_KAISER8)
            hipLaunchKernelGGL(eval_grid_tiled_kernel<INTERP_KAISER8>, gt, dim3(MT_THREADS), 0, hs,
                               im, g, nr, nu, ld, x0, dx, nx, y0, dy, ny, k, o);
        else
            hipLaunchKernelGGL(eval_grid_tiled_kernel<INTERP_LAGRANGE4>, gt, dim3(MT_THREADS), 0,
                               hs, im, g, nr, nu, ld, x0, dx, nx, y0, dy, ny, k, o);
        return check_launch("eval_grid_tiled_kernel");
    }
    auto l = [&]<int M, int I>() {
        hipLaunchKernelGGL((eval_grid_kernel<M, I>), gr, dim3(256), 0, hs, im, g, nr, nu, ld, x0,
                           dx, nx, y0, dy, ny, k, o, kt);
    };
    dispatch2(mode, interp, l);
    return check_launch("eval_grid_kernel");
}

int bpmi_eval_grid_multi(const bpmi_eval_job* jobs, const int* blocks, int nblocks, int mode,
                         int interp, void* stream) {
    int st = check_mode_interp(mode, interp);
    if (st) return st;
    if (nblocks <= 0) return BPMI_OK;
    if (!jobs || !blocks) return set_error(BPMI_ERR_INVALID, "null table");
    const float* kt = g_kaiser_table;
    hipStream_t hs = (hipStream_t)stream;
    if (mode == BPMI_MODE_SPLIT) {
        if (interp == BPMI_INTERP_KAISER8)
            hipLaunchKernelGGL(eval_grid_multi_fast_kernel<INTERP_KAISER8>, dim3(nblocks),
                               dim3(256), 0, hs, jobs, blocks);
        else
            hipLaunchKernelGGL(eval_grid_multi_fast_kernel<INTERP_LAGRANGE4>, dim3(nblocks),
                               dim3(256), 0, hs, jobs, blocks);
        return check_launch("eval_grid_multi_fast_kernel");
    }
    auto l = [&]<int M, int I>() {
        hipLaunchKernelGGL((eval_grid_multi_kernel<M, I>), dim3(nblocks), dim3(256), 0, hs, jobs,
                           blocks, kt);
    };
    dispatch2(mode, interp, l);
    return check_launch("eval_grid_multi_kernel");
}

int bpmi_eval_tiles(int nx, int ny, int* tiles) {
    if (!tiles || nx <= 0 || ny <= 0) return set_error(BPMI_ERR_INVALID, "bad grid shape");
    const long long t = (long long)((nx + ET_X - 1) / ET_X) * ((ny + ET_Y - 1) / ET_Y);
    if (t > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "grid too large");
    *tiles = (int)t;
    return BPMI_OK;
}

int bpmi_eval_grid_multi_tiled(const bpmi_eval_job* jobs, const int* blocks, int nblocks,
                               int interp, void* stream) {
    int st = check_mode_interp(BPMI_MODE_SPLIT, interp);
    if (st) return st;
    if (nblocks <= 0) return BPMI_OK;
    if (!jobs || !blocks) return set_error(BPMI_ERR_INVALID, "null table");
    hipStream_t hs = (hipStream_t)stream;
    if (interp == BPMI_INTERP_KAISER8)
        hipLaunchKernelGGL(eval_grid_multi_tiled_kernel<INTERP_KAISER8>, dim3(nblocks),
                           dim3(MT_THREADS), 0, hs, jobs, blocks);
    else
        hipLaunchKernelGGL(eval_grid_multi_tiled_kernel<INTERP_LAGRANGE4>, dim3(nblocks),
                           dim3(MT_THREADS), 0, hs, jobs, blocks);
    return check_launch("eval_grid_multi_tiled_kernel");
}

int bpmi_target_contributions(const void* data, int samples, const double* pos, int pulses,
                              const double* xyz, long long npoints, double range_start,
                              double data_dr, double k, int mode, int interp, void* out,
                              void* stream) {
    if (!data || !pos || !xyz || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (samples <= 0 || pulses < 0 || npoints < 0) return set_error(BPMI_ERR_INVALID, "bad sizes");
    if (interp != BPMI_INTERP_LAGRANGE4 && interp != BPMI_INTERP_LINEAR)
        return set_error(BPMI_ERR_INVALID, "contributions support lagrange4 and linear");
    const long long n = npoints * (long long)pulses;
    if (n == 0) return BPMI_OK;
    const dim3 g(grid_1d(n, 256));
    hipStream_t hs = (hipStream_t)stream;
    const bool f64 = mode == BPMI_MODE_FP64;
    if (interp == BPMI_INTERP_LAGRANGE4) {
        if (f64)
            hipLaunchKernelGGL((contrib_kernel<MODE_FP64, INTERP_LAGRANGE4>), g, dim3(256), 0, hs,
                               (const float2*)data, samples, pos, pulses, xyz, npoints,
                               range_start, data_dr, k, (float2*)out, nullptr);
        else
            hipLaunchKernelGGL((contrib_kernel<MODE_MIXED, INTERP_LAGRANGE4>), g, dim3(256), 0,
                               hs, (const float2*)data, samples, pos, pulses, xyz, npoints,
                               range_start, data_dr, k, (float2*)out, nullptr);
    } else {
        if (f64)
            hipLaunchKernelGGL((contrib_kernel<MODE_FP64, INTERP_LINEAR>), g, dim3(256), 0, hs,
                               (const float2*)data, samples, pos, pulses, xyz, npoints,
                               range_start, data_dr, k, (float2*)out, nullptr);
        else
            hipLaunchKernelGGL((contrib_kernel<MODE_MIXED, INTERP_LINEAR>), g, dim3(256), 0, hs,
                               (const float2*)data, samples, pos, pulses, xyz, npoints,
                               range_start, data_dr, k, (float2*)out, nullptr);
    }
    return check_launch("contrib_kernel");
}

int bpmi_fill_range_phase_lut(int samples, double range_start, double data_dr, double k, void* lut,
                              void* stream) {
    if (samples <= 0 || !lut) return set_error(BPMI_ERR_INVALID, "bad LUT arguments");
    hipLaunchKernelGGL(fill_lut_kernel, dim3(grid_1d(samples, 256)), dim3(256), 0,
                       (hipStream_t)stream, samples, range_start, data_dr, k, (float2*)lut);
    return check_launch("fill_lut_kernel");
}

int bpmi_bp_polar_lut(const void* data, int samples, const double* pos, const double* weight,
                      const double* phase, int pulse_begin, int pulse_end, const double* geo,
                      int nr, int nu, int side, double range_start, double data_dr, double k,
                      const void* lut, void* out, void* stream) {
    if (!data || !pos || !geo || !lut || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    if (pulse_begin < 0 || pulse_end < pulse_begin) return set_error(BPMI_ERR_INVALID, "bad pulse range");
    hipLaunchKernelGGL(bp_polar_lut_kernel, dim3(grid_1d((long long)nr * nu, 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float2*)data, samples, pos, weight, phase,
                       pulse_begin, pulse_end, make_geo(geo), nr, nu, side, range_start, data_dr,
                       k, (const float2*)lut, (float2*)out);
    return check_launch("bp_polar_lut_kernel");
}

int bpmi_add_inplace(void* left, const void* right, long long n, void* stream) {
    if (n <= 0) return BPMI_OK;
    if (!left || !right) return set_error(BPMI_ERR_INVALID, "null argument");
    hipLaunchKernelGGL(add_kernel, dim3(grid_1d(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       (float2*)left, (const float2*)right, n);
    return check_launch("add_kernel");
}

int bpmi_cast_c128_c64(const void* in, void* out, long long n, void* stream) {
    if (n <= 0) return BPMI_OK;
    if (!in || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    hipLaunchKernelGGL(cast_kernel, dim3(grid_1d(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const double2*)in, (float2*)out, n);
    return check_launch("cast_kernel");
}

int bpmi_crop(const void* in, int src_nu, int r_begin, int u_begin, int nr, int nu, void* out,
              void* stream) {
    if (!in || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (nr <= 0 || nu <= 0) return set_error(BPMI_ERR_INVALID, "empty image grids are unsupported");
    if (r_begin < 0 || u_begin < 0 || u_begin + nu > src_nu)
        return set_error(BPMI_ERR_INVALID, "crop window outside the source");
    hipLaunchKernelGGL(crop_kernel, dim3(grid_1d((long long)nr * nu, 256)), dim3(256), 0,
                       (hipStream_t)stream, (const float2*)in, src_nu, r_begin, u_begin, nr, nu,
                       (float2*)out);
    return check_launch("crop_kernel");
}

int bpmi_probe_sincos_turns(const float* t, float* s, float* c, long long n, void* stream) {
    if (n <= 0) return BPMI_OK;
    if (!t || !s || !c) return set_error(BPMI_ERR_INVALID, "null argument");
    hipLaunchKernelGGL(sincos_probe_kernel, dim3(grid_1d(n, 256)), dim3(256), 0,
                       (hipStream_t)stream, t, s, c, n);
    return check_launch("sincos_probe_kernel");
}

int bpmi_probe_fma_peak(float* out, int blocks, int iters, void* stream) {
    if (!out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (blocks <= 0 || iters <= 0) return set_error(BPMI_ERR_INVALID, "blocks and iters must be positive");
    hipLaunchKernelGGL(fma_peak_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out,
                       iters, 0.999f, 1e-3f);
    return check_launch("fma_peak_kernel");
}

}  // extern "C"
