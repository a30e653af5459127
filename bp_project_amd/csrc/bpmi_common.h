// Shared device helpers for the gfx950 SAR backprojection library.
//
// Complex samples are float2 (complex64) in HBM.  Interpolators:
//   LAGRANGE4 : the reference's 4-point Lagrange on taps i-1..i+2
//               (fast_bp/ffbp.py:17-29), zero unless 1 <= i <= len-3;
//   LINEAR    : (1-t) v[i] + t v[i+1], zero unless 0 <= i <= len-2;
//   KAISER8   : 8-tap truncated sinc with a Kaiser (beta = 6) window on taps
//               i-3..i+4, weights normalised to sum 1, zero unless
//               3 <= i <= len-5.
// Phase is handled in revolutions ("turns"): the FP64 argument is reduced by
// subtracting its floor, and the FP32 hardware sine/cosine (whose input is in
// revolutions) evaluates the reduced value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bpmi {

enum Mode { MODE_FP64 = 0, MODE_FP32 = 1, MODE_MIXED = 2, MODE_DFLOAT = 3, MODE_SPLIT = 4 };
enum Interp { INTERP_LAGRANGE4 = 0, INTERP_LINEAR = 1, INTERP_KAISER8 = 2 };

constexpr double TWO_PI = 6.283185307179586476925286766559;
constexpr double INV_TWO_PI = 0.15915494309189533576888376337251;

// Kaiser-sinc table: kernel value at x = i / KAISER_TABLE_RES for i in
// [0, 4 * KAISER_TABLE_RES + 1]; symmetric, so only x >= 0 is stored.  The
// interpolator reads it by linear interpolation (FP64 table in the FP64
// paths, its FP32 rounding elsewhere) and normalises the 8 weights.
constexpr int KAISER_TABLE_RES = 2048;
constexpr int KAISER_TABLE_LEN = 4 * KAISER_TABLE_RES + 2;

// The device table allocation holds the FP32 table followed by the FP64 one
// (KAISER_TABLE_LEN floats is a multiple of 8 bytes).
__device__ __forceinline__ const double* kaiser_table64(const float* t32) {
    return reinterpret_cast<const double*>(t32 + KAISER_TABLE_LEN);
}

__device__ __forceinline__ double fract64(double x) { return x - floor(x); }
__device__ __forceinline__ float fract32(float x) { return x - floorf(x); }
// Hardware fraction (v_fract_f64 / v_fract_f32): x - floor(x) in one
// instruction for finite x.
__device__ __forceinline__ double hw_fract64(double x) { return __builtin_amdgcn_fract(x); }
__device__ __forceinline__ float hw_fract32(float x) { return __builtin_amdgcn_fractf(x); }

// FP64 -> int32 rounded toward zero, saturating: v_cvt_i32_f64 gives INT_MAX
// above and INT_MIN below the int range and 0 for NaN.  Written as the
// instruction because a C++ conversion of an out-of-range value is undefined.
__device__ __forceinline__ int cvt_i32_sat(double x) {
    int i;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

// sin/cos of 2*pi*t with the hardware transcendental unit (input in turns).
__device__ __forceinline__ void sincos_turns(float t, float* s, float* c) {
    *s = __builtin_amdgcn_sinf(t);
    *c = __builtin_amdgcn_cosf(t);
}

// exp(j 2 pi T) for an FP64 phase in turns: exact FP64 reduction, FP32 sincos.
__device__ __forceinline__ float2 cis_turns64(double turns) {
    float s, c;
    // v_fract_f64 equals turns - floor(turns) except that it stays below 1
    // where the subtraction rounds up to 1; both round to 1.0f here.
    sincos_turns((float)hw_fract64(turns), &s, &c);
    return make_float2(c, s);
}

// 1/sqrt(x): hardware estimate (v_rsq_f64, relative error <= 5.2e-8 over
// [1e-3, 1e9] on gfx950) and one Newton step, relative error <= 4.2e-15 --
// 2.5e-11 m on a 6 km range, a 1.6e-9-turn phase error at X band -- for
// x > 0 and normal.  Contract at x = 0: the estimate is +inf and the Newton
// step forms 0 * inf, so the result is NaN (not inf), and so is x * rsqrt64(x).
// Callers rely on that NaN to reject a point on an origin without a
// separate test (child_tap, and through it the merge, evaluation and
// tile-marking kernels; the split and adjoint direct-BP pixel loops, where
// the NaN range fails floor_index): clamping x here would turn such points
// into defined taps.
__device__ __forceinline__ double rsqrt64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-0.5 * x, y * y, 0.5);
    return fma(y, e, y);
}

// sqrt(x) for x > 0: the hardware reciprocal square root estimate, then one
// Newton step on the square root itself, s + (y / 2) (x - s^2) with s = x y
// (relative error of the order of 1e-15, like rsqrt64; one multiplication
// fewer than x * rsqrt64(x)).  NaN at x = 0 (0 * inf), as x * rsqrt64(x).
__device__ __forceinline__ double sqrt64_newton(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double s = x * y;
    return fma(0.5 * y, fma(-s, s, x), s);
}

// 1/x: hardware estimate (v_rcp_f64, relative error <= 4.6e-8) and two Newton
// steps (below 1 ulp in practice; not IEEE-rounded).
__device__ __forceinline__ double rcp64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Wave-uniform copies (v_readfirstlane) of values every lane holds equally:
// they move table reads the compiler cannot prove uniform into scalar
// registers; on values already in scalar registers they compile to nothing.
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uniform_i64(long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double uniform_f64(double v) {
    return __longlong_as_double(uniform_i64(__double_as_longlong(v)));
}
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    return (T*)(uintptr_t)uniform_i64((long long)(uintptr_t)p);
}

// Read-only views of tables a kernel never writes (node tables, work
// lists, argument blocks): loads through the constant address space are
// scalar loads (scalar cache) when the address is wave-uniform, where a
// generic pointer that the compiler cannot prove unwritten gets vector loads.
#define BPMI_CONST_AS __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const BPMI_CONST_AS T* as_const(const T* p) {
    return (const BPMI_CONST_AS T*)p;
}

// complex64 loads/stores of node images that other workgroups of the same
// launch write or read (COH = true; COH = false is an ordinary access).
// Stores are device-scope (sc1): they are written through to memory, where
// every die sees them, and a wave waits for them (wait_vector_memory) before
// its item is counted complete.  Loads are ordinary L2-cached loads of the
// global address space, which is coherent here because of how node images
// are used within one launch: a node's image is read only after its
// completion count is seen, no workgroup reads it before that (producers
// read only the children of the node they write, and images never share a
// cache line: arena slots are 512-byte aligned), and the launch starts with
// no stale line cached, so the first read of a line on any die misses its
// caches and fetches the completed data.  (Device-scope loads bypass the
// die's L2, and neighbouring tiles' windows overlap: measured on the
// 64-patch batch, they made the one-launch FFBP 20% slower.)
//
// Both address global memory explicitly: through a generic pointer they
// would be flat accesses, which the compiler orders behind every LDS access
// of the wave (a flat access may touch LDS), so a window staged from them
// would wait out each load's round trip.
typedef __attribute__((address_space(1))) unsigned long long global_u64_rw;
// One 64-bit load of a complex64 sample in global memory (never LDS),
// addressed as global memory whatever the pointer's origin (a float2 copy
// would go through a generic reference and lose the address space again).
__device__ __forceinline__ float2 ld_global_c64(const float2* p) {
    const unsigned long long v = *(const global_u64_rw*)p;
    return make_float2(__uint_as_float((unsigned)v), __uint_as_float((unsigned)(v >> 32)));
}
template <bool COH>
__device__ __forceinline__ float2 ld_c64(const float2* p) {
    if constexpr (COH) {
        return ld_global_c64(p);
    } else {
        return *p;
    }
}
template <bool COH>
__device__ __forceinline__ void st_c64(float2* p, float2 v) {
    if constexpr (COH) {
        const unsigned long long u = (unsigned long long)__float_as_uint(v.x) |
                                     ((unsigned long long)__float_as_uint(v.y) << 32);
        __hip_atomic_store((global_u64_rw*)p, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}
// Waits until every vector-memory access this wave issued has completed
// (s_waitcnt vmcnt(0)): device-scope stores are then visible on every die.
__device__ __forceinline__ void wait_vector_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// sqrt(max(x, 0)) from the reciprocal square root (0 for x <= 0).
__device__ __forceinline__ double sqrt64_fast(double x) { return x > 0.0 ? x * rsqrt64(x) : 0.0; }

// Reference association order of cubic_weights (fast_bp/ffbp.py:17-21).
// FP64: the reference's divisions; FP32: multiplication by the rounded
// reciprocal (a 1-ulp weight difference, no IEEE division sequence).
__device__ __forceinline__ void cubic_weights(double t, double w[4]) {
    w[0] = -t * (t - 1.0) * (t - 2.0) / 6.0;
    w[1] = (t + 1.0) * (t - 1.0) * (t - 2.0) / 2.0;
    w[2] = -(t + 1.0) * t * (t - 2.0) / 2.0;
    w[3] = (t + 1.0) * t * (t - 1.0) / 6.0;
}

__device__ __forceinline__ void cubic_weights(float t, float w[4]) {
    constexpr float sixth = 1.0f / 6.0f;
    w[0] = -t * (t - 1.0f) * (t - 2.0f) * sixth;
    w[1] = (t + 1.0f) * (t - 1.0f) * (t - 2.0f) * 0.5f;
    w[2] = -(t + 1.0f) * t * (t - 2.0f) * 0.5f;
    w[3] = (t + 1.0f) * t * (t - 1.0f) * sixth;
}

// Kaiser-sinc weights from the device table (linear interpolation of the
// tabulated kernel), normalised to unit sum.
__device__ __forceinline__ void kaiser_weights_f32(const float* table, float t, float w[8]) {
    float sum = 0.f;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        float x = fabsf(t - (float)(a - 3)) * (float)KAISER_TABLE_RES;
        int i = (int)x;
        float f = x - (float)i;
        i = min(i, KAISER_TABLE_LEN - 2);
        float v = table[i] + f * (table[i + 1] - table[i]);
        w[a] = v;
        sum += v;
    }
    float inv = 1.f / sum;
#pragma unroll
    for (int a = 0; a < 8; ++a) w[a] *= inv;
}

__device__ __forceinline__ void kaiser_weights_f64(const double* table, double t, double w[8]) {
    double sum = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double x = fabs(t - (double)(a - 3)) * (double)KAISER_TABLE_RES;
        int i = (int)x;
        double f = x - (double)i;
        i = min(i, KAISER_TABLE_LEN - 2);
        w[a] = table[i] + f * (table[i + 1] - table[i]);
        sum += w[a];
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) w[a] /= sum;
}

// Interpolator geometry: first tap offset, tap count, validity of index i.
template <int INTERP> struct InterpShape;
template <> struct InterpShape<INTERP_LAGRANGE4> {
    static constexpr int FIRST = -1, TAPS = 4;
    __device__ static bool valid(int i, int len) { return i >= 1 && i + 2 < len; }
};
template <> struct InterpShape<INTERP_LINEAR> {
    static constexpr int FIRST = 0, TAPS = 2;
    __device__ static bool valid(int i, int len) { return i >= 0 && i + 1 < len; }
};
template <> struct InterpShape<INTERP_KAISER8> {
    static constexpr int FIRST = -3, TAPS = 8;
    __device__ static bool valid(int i, int len) { return i >= 3 && i + 4 < len; }
};

// 1-D sample of a complex64 row, FP32 weights and accumulation.
template <int INTERP>
__device__ __forceinline__ float2 sample1_f32w(const float2* __restrict__ row, int len, int i,
                                               float t, const float* ktab) {
    float2 z = make_float2(0.f, 0.f);
    if (!InterpShape<INTERP>::valid(i, len)) return z;
    if (INTERP == INTERP_LINEAR) {
        const float2 a = row[i], b = row[i + 1];
        const float w0 = 1.f - t;
        z.x = w0 * a.x + t * b.x;
        z.y = w0 * a.y + t * b.y;
    } else if (INTERP == INTERP_LAGRANGE4) {
        float w[4];
        cubic_weights(t, w);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float2 v = row[i + a - 1];
            z.x += w[a] * v.x;
            z.y += w[a] * v.y;
        }
    } else {
        float w[8];
        kaiser_weights_f32(ktab, t, w);
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float2 v = row[i + a - 3];
            z.x += w[a] * v.x;
            z.y += w[a] * v.y;
        }
    }
    return z;
}

// 1-D sample in FP64 (weights and accumulation), reference association order.
template <int INTERP>
__device__ __forceinline__ double2 sample1_f64(const float2* __restrict__ row, int len, double q,
                                               const double* ktab64 = nullptr) {
    double2 z = make_double2(0.0, 0.0);
    const double fi = floor(q);
    if (!(fi > -2147483000.0 && fi < 2147483000.0)) return z;
    const int i = (int)fi;
    if (!InterpShape<INTERP>::valid(i, len)) return z;
    const double t = q - (double)i;
    if (INTERP == INTERP_LINEAR) {
        const float2 a = row[i], b = row[i + 1];
        const double w0 = 1.0 - t;
        z.x = w0 * (double)a.x + t * (double)b.x;
        z.y = w0 * (double)a.y + t * (double)b.y;
    } else if (INTERP == INTERP_LAGRANGE4) {
        double w[4];
        cubic_weights(t, w);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float2 v = row[i + a - 1];
            z.x = z.x + w[a] * (double)v.x;
            z.y = z.y + w[a] * (double)v.y;
        }
    } else {
        double w[8];
        kaiser_weights_f64(ktab64, t, w);
        for (int a = 0; a < 8; ++a) {
            const float2 v = row[i + a - 3];
            z.x += w[a] * (double)v.x;
            z.y += w[a] * (double)v.y;
        }
    }
    return z;
}

// floor of an FP64 coordinate as int with a finite-range guard.
__device__ __forceinline__ bool floor_index(double q, int* i) {
    const double f = floor(q);
    if (!(f > -2147483000.0 && f < 2147483000.0)) return false;
    *i = (int)f;
    return true;
}

__device__ __forceinline__ bool floor_index_f32(float q, int* i) {
    const float f = floorf(q);
    if (!(f > -2.0e9f && f < 2.0e9f)) return false;
    *i = (int)f;
    return true;
}

// Pixel of a polar node grid (fast_bp/ffbp.py:62-67): slant range r from the
// node origin, u = sin(horizontal angle), pixel on z = 0.
__device__ __forceinline__ void polar_pixel(double ox, double oy, double oz, double r, double u,
                                            int side, double* x, double* y) {
    const double ground = sqrt(fmax(0.0, r * r - oz * oz));
    *x = ox + (double)side * ground * sqrt(fmax(0.0, 1.0 - u * u));
    *y = oy + ground * u;
}

__device__ __forceinline__ void polar_pixel_f32(float ox, float oy, float oz, float r, float u,
                                                int side, float* x, float* y) {
    const float ground = sqrtf(fmaxf(0.0f, r * r - oz * oz));
    *x = ox + (float)side * ground * sqrtf(fmaxf(0.0f, 1.0f - u * u));
    *y = oy + ground * u;
}

// ---------------------------------------------------------------------------
// Double-float (hi + lo FP32) arithmetic, error-free transformations with FMA
// (Dekker / Knuth two-sum and FMA two-product).
// ---------------------------------------------------------------------------
struct df {
    float hi, lo;
};

// The error-free transformations below must not be contracted into FMAs.
#pragma clang fp contract(off)
__device__ __forceinline__ df df_from(double x) {
    df r;
    r.hi = (float)x;
    r.lo = (float)(x - (double)r.hi);
    return r;
}

__device__ __forceinline__ df two_sum(float a, float b) {
    df r;
    r.hi = a + b;
    const float bb = r.hi - a;
    r.lo = (a - (r.hi - bb)) + (b - bb);
    return r;
}

__device__ __forceinline__ df quick_two_sum(float a, float b) {
    df r;
    r.hi = a + b;
    r.lo = b - (r.hi - a);
    return r;
}

__device__ __forceinline__ df df_add(df a, df b) {
    df s = two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return quick_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ df df_sub(df a, df b) {
    df nb;
    nb.hi = -b.hi;
    nb.lo = -b.lo;
    return df_add(a, nb);
}

__device__ __forceinline__ df df_mul(df a, df b) {
    const float p = a.hi * b.hi;
    const float e = fmaf(a.hi, b.hi, -p);
    const float lo = e + (a.hi * b.lo + a.lo * b.hi);
    return quick_two_sum(p, lo);
}

__device__ __forceinline__ df df_sqr(df a) { return df_mul(a, a); }
#pragma clang fp contract(on)

// sqrt by one Newton step from the FP32 root; the residual is formed exactly
// with an FMA.
__device__ __forceinline__ df df_sqrt(df a) {
    if (a.hi <= 0.f) {
        df z;
        z.hi = 0.f;
        z.lo = 0.f;
        return z;
    }
    const float s = sqrtf(a.hi);
    const float res = (fmaf(-s, s, a.hi) + a.lo);
    const float corr = res / (2.f * s);
    return quick_two_sum(s, corr);
}

}  // namespace bpmi
