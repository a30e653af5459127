// Workgroup FFT in LDS for gfx950: one complex64 sequence of N points per
// workgroup, N = 2^a 3^b 5^c up to RF_NMAX, decomposed into Stockham passes
// of radix 8, 4, 3 and 5 (natural order in and out, no bit reversal).
//
// Pass with radix R after passes whose radices multiply to Ns: butterfly j
// (j < N/R) reads x[j + r N/R], r < R, multiplies input r by the twiddle
// w^(k r), k = j mod Ns, w = exp(-+2 pi i / (Ns R)), takes the R-point DFT
// and writes output r to x[(j - k) R + k + r Ns].  Each pass stages its
// inputs in registers between two barriers, so the data needs one LDS buffer
// of N points, stored with one pad slot after every 16 (lds_index) so that
// the strided accesses of the early passes spread over the LDS banks.
// Twiddles come from a table tw[m] = exp(-2 pi i m / N) (FP64
// rounded to FP32) with w^(kr) = tw[k r N / (Ns R)]: powers 1, 2, 4 are
// read, the others formed by at most two complex products.
#pragma once

#include "bpmi_common.h"

namespace bpmi {

constexpr int RF_THREADS = 1024;
constexpr int RF_NMAX = 9408;     // points; 9408 * 17 / 16 padded slots = 80 KB of LDS
constexpr int RF_SLOTS = RF_NMAX + RF_NMAX / 16;
constexpr int RF_MAX_PASSES = 8;

// LDS slot of point i: one pad slot after every 16 points.
__device__ __forceinline__ int lds_index(int i) { return i + (i >> 4); }

struct FftPlan1 {
    int n;
    int passes;
    int radix[RF_MAX_PASSES];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot_i(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const float2 t2 = cadd(a1, a3), t3 = rot_i<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2* a) {
    constexpr float H = 0.86602540378443864676f;  // sin(2 pi / 3)
    const float2 s = cadd(a[1], a[2]), d = csub(a[1], a[2]);
    const float2 m = make_float2(a[0].x - 0.5f * s.x, a[0].y - 0.5f * s.y);
    const float2 q = rot_i<INV>(make_float2(H * d.x, H * d.y));
    a[0] = cadd(a[0], s);
    a[1] = cadd(m, q);
    a[2] = csub(m, q);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* a) {
    constexpr float C1 = 0.30901699437494742410f, C2 = -0.80901699437494742410f;
    constexpr float S1 = 0.95105651629515357212f, S2 = 0.58778525229247312917f;
    const float2 b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
    const float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
    const float2 e1 = make_float2(a[0].x + C1 * b1.x + C2 * b2.x, a[0].y + C1 * b1.y + C2 * b2.y);
    const float2 e2 = make_float2(a[0].x + C2 * b1.x + C1 * b2.x, a[0].y + C2 * b1.y + C1 * b2.y);
    const float2 f1 = rot_i<INV>(make_float2(S1 * d1.x + S2 * d2.x, S1 * d1.y + S2 * d2.y));
    const float2 f2 = rot_i<INV>(make_float2(S2 * d1.x - S1 * d2.x, S2 * d1.y - S1 * d2.y));
    a[0] = cadd(a[0], cadd(b1, b2));
    a[1] = cadd(e1, f1);
    a[4] = csub(e1, f1);
    a[2] = cadd(e2, f2);
    a[3] = csub(e2, f2);
}

// 8 points: even/odd halves by dft4, combined with W8^k.
template <bool INV>
__device__ __forceinline__ void dft8(float2* a) {
    constexpr float R = 0.70710678118654752440f;
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    // W8^1 = (1 -+ i) R, W8^2 = -+i, W8^3 = (-1 -+ i) R
    const float2 o1 = INV ? make_float2(R * (a[3].x - a[3].y), R * (a[3].x + a[3].y))
                          : make_float2(R * (a[3].x + a[3].y), R * (a[3].y - a[3].x));
    const float2 o2 = rot_i<INV>(a[5]);
    const float2 o3 = INV ? make_float2(-R * (a[7].x + a[7].y), R * (a[7].x - a[7].y))
                          : make_float2(R * (a[7].y - a[7].x), -R * (a[7].x + a[7].y));
    const float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6], o0 = a[1];
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1);
    a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2);
    a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3);
    a[7] = csub(e3, o3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a) {
    static_assert(R == 3 || R == 4 || R == 5 || R == 8, "radix 3, 4, 5 or 8");
    if constexpr (R == 3) dft3<INV>(a);
    else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 5) dft5<INV>(a);
    else dft8<INV>(a);
}

template <bool INV>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int m) {
    const float2 w = tw[m];
    return INV ? make_float2(w.x, -w.y) : w;
}

// One Stockham pass over buf[0, n) by all RF_THREADS threads (both barriers
// reached by every thread).
template <int R, bool INV>
__device__ __forceinline__ void fft_pass(float2* buf, int n, int ns, const float2* __restrict__ tw,
                                         int t) {
    constexpr int MAXB = (RF_NMAX / R + RF_THREADS - 1) / RF_THREADS;
    const int nr = n / R;
    const int stride = n / (ns * R);
    // j / ns by a multiply-shift, exact for j < 2^14 and ns <= RF_NMAX
    const unsigned magic = ((1u << 28) + (unsigned)ns - 1u) / (unsigned)ns;
    float2 v[MAXB][R];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int j = t + i * RF_THREADS;
        if (j < nr) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = buf[lds_index(j + r * nr)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int j = t + i * RF_THREADS;
        if (j < nr) {
            const int q = (int)(((unsigned long long)(unsigned)j * magic) >> 28);
            const int k = j - q * ns;
            if (ns > 1) {
                const int m = k * stride;
                float2 w[R];
                w[0] = make_float2(1.f, 0.f);
                w[1] = twiddle<INV>(tw, m);
                if constexpr (R > 2) w[2] = twiddle<INV>(tw, 2 * m);
                if constexpr (R > 3) w[3] = cmul(w[1], w[2]);
                if constexpr (R > 4) w[4] = twiddle<INV>(tw, 4 * m);
                if constexpr (R > 5) {
#pragma unroll
                    for (int r = 5; r < R; ++r) w[r] = cmul(w[r - 4], w[4]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], w[r]);
            }
            dft<R, INV>(v[i]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lds_index(base + r * ns)] = v[i][r];
        }
    }
    __syncthreads();
}

// Whole transform (plan radices in order); buf holds the input and, on
// return, the output in natural order (point i at lds_index(i)).
template <bool INV>
__device__ __forceinline__ void fft_lds(float2* buf, const FftPlan1& plan,
                                        const float2* __restrict__ tw, int t) {
    int ns = 1;
    for (int p = 0; p < plan.passes; ++p) {
        const int r = plan.radix[p];
        switch (r) {
            case 8: fft_pass<8, INV>(buf, plan.n, ns, tw, t); break;
            case 4: fft_pass<4, INV>(buf, plan.n, ns, tw, t); break;
            case 3: fft_pass<3, INV>(buf, plan.n, ns, tw, t); break;
            default: fft_pass<5, INV>(buf, plan.n, ns, tw, t); break;
        }
        ns *= r;
    }
}

// ---- plans -------------------------------------------------------------------
// Radices for n points: the power of two 2^a (a >= 2) by passes of 8 and 4,
// then the factors 3 and 5.  passes = 0 when n is out of range (above RF_NMAX,
// a single factor 2, another prime factor, or more than RF_MAX_PASSES passes).
constexpr FftPlan1 make_fft_plan(int n) {
    FftPlan1 plan{n, 0, {0, 0, 0, 0, 0, 0, 0, 0}};
    if (n < 2 || n > RF_NMAX) return plan;
    int m = n, a = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++a;
    }
    if (a == 1) return plan;
    int fours = a % 3 == 1 ? 2 : a % 3 == 2 ? 1 : 0;
    int radix[16] = {};
    int np = 0;
    for (int e = a - 2 * fours; e > 0; e -= 3) radix[np++] = 8;
    for (; fours > 0; --fours) radix[np++] = 4;
    while (m % 3 == 0 && np < 16) {
        radix[np++] = 3;
        m /= 3;
    }
    while (m % 5 == 0 && np < 16) {
        radix[np++] = 5;
        m /= 5;
    }
    if (m != 1 || np > RF_MAX_PASSES) return plan;
    for (int i = 0; i < np; ++i) plan.radix[i] = radix[i];
    plan.passes = np;
    return plan;
}

// ---- passes with compile-time sizes ---------------------------------------------
// The same pass as fft_pass for a transform length N known at compile time:
// strides, the j mod NS split and twiddle indices fold into constants, and
// where the padded layout allows (a stride that is a multiple of 16 points, or
// the unit-stride outputs of the first pass for R in {4, 8}) the R accesses
// of a butterfly are one base slot plus immediate offsets.
template <int R, bool INV, int N, int NS>
__device__ __forceinline__ void fft_pass_c(float2* buf, const float2* __restrict__ tw, int t) {
    constexpr int NR = N / R;
    constexpr int STRIDE = N / (NS * R);
    constexpr int MAXB = (NR + RF_THREADS - 1) / RF_THREADS;
    float2 v[MAXB][R];
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int j = t + i * RF_THREADS;
        if (MAXB * RF_THREADS == NR || j < NR) {
            if constexpr (NR % 16 == 0) {
                const int b0 = lds_index(j);
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = buf[b0 + r * (NR + NR / 16)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = buf[lds_index(j + r * NR)];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXB; ++i) {
        const int j = t + i * RF_THREADS;
        if (MAXB * RF_THREADS == NR || j < NR) {
            const int k = j % NS;
            if constexpr (NS > 1) {
                const int m = k * STRIDE;
                float2 w[R];
                w[0] = make_float2(1.f, 0.f);
                w[1] = twiddle<INV>(tw, m);
                if constexpr (R > 2) w[2] = twiddle<INV>(tw, 2 * m);
                if constexpr (R > 3) w[3] = cmul(w[1], w[2]);
                if constexpr (R > 4) w[4] = twiddle<INV>(tw, 4 * m);
                if constexpr (R > 5) {
#pragma unroll
                    for (int r = 5; r < R; ++r) w[r] = cmul(w[r - 4], w[4]);
                }
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], w[r]);
            }
            dft<R, INV>(v[i]);
            const int base = (j - k) * R + k;
            if constexpr (NS % 16 == 0) {
                const int b0 = lds_index(base);
#pragma unroll
                for (int r = 0; r < R; ++r) buf[b0 + r * (NS + NS / 16)] = v[i][r];
            } else if constexpr (NS == 1 && 16 % R == 0) {
                // base = j R: the R outputs share one block of 16 points
                const int b0 = lds_index(base);
#pragma unroll
                for (int r = 0; r < R; ++r) buf[b0 + r] = v[i][r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) buf[lds_index(base + r * NS)] = v[i][r];
            }
        }
    }
    __syncthreads();
}

template <bool INV, int N, int P = 0, int NS = 1>
__device__ __forceinline__ void fft_lds_c(float2* buf, const float2* __restrict__ tw, int t) {
    constexpr FftPlan1 plan = make_fft_plan(N);
    static_assert(plan.passes > 0, "transform length without a plan");
    if constexpr (P < plan.passes) {
        constexpr int R = plan.radix[P];
        fft_pass_c<R, INV, N, NS>(buf, tw, t);
        fft_lds_c<INV, N, P + 1, NS * R>(buf, tw, t);
    }
}

}  // namespace bpmi
