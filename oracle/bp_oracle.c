/*
 * FP64 CPU oracle for the SAR backprojection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The HIP library never calls it.
 *
 * It restates, in C with explicit real/imaginary FP64 arithmetic and with
 * floating-point contraction disabled (-ffp-contract=off), the Numba kernels
 * of the reference:
 *   cubic_weights        fast_bp/ffbp.py:17-21
 *   sample1              fast_bp/ffbp.py:24-29
 *   sample2              fast_bp/ffbp.py:32-40
 *   backproject_points   fast_bp/ffbp.py:43-55
 *   bp_polar             fast_bp/ffbp.py:58-76
 *   regrid               fast_bp/ffbp.py:79-95
 *   evaluate_image       fast_bp/ffbp.py:189-198
 *   target_contributions fast_bp/ffbp.py:201-211
 * The order of every operation follows the Python expressions in those
 * lines (left-to-right evaluation, complex128 promotion of complex64 samples).
 * Loops over output pixels run under OpenMP, like the reference's prange.
 *
 * Two interpolators have no reference counterpart and are specified here:
 *   linear      : (1-t)*row[i] + t*row[i+1], valid when 0 <= i <= len-2.
 *   kaiser-sinc : 8 taps at i-3..i+4.  The kernel
 *                 K(x) = sinc(x)*I0(beta*sqrt(1-(x/4)^2))/I0(beta), beta = 6,
 *                 is tabulated in FP64 at x = j/2048 (j = 0..8193) and read by
 *                 linear interpolation; the 8 weights are normalised by their
 *                 sum.  Valid when 3 <= i and i+4 <= len-1, else 0.
 * One documented deviation: regrid/evaluate_image return 0 where the source
 * ground distance g is exactly 0 (Numba would raise ZeroDivisionError; the
 * reference GPU kernels return 0, fast_bp/gpu/kernels.cu:359).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_LAGRANGE4 = 0, ORC_LINEAR = 1, ORC_KAISER8 = 2 };

#define KAISER_BETA 6.0
#define KAISER_HALF 4

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ffbp.py:17-21, same association order. */
void orc_cubic_weights(double t, double w[4]) {
    w[0] = -t * (t - 1) * (t - 2) / 6;
    w[1] = (t + 1) * (t - 1) * (t - 2) / 2;
    w[2] = -(t + 1) * t * (t - 2) / 2;
    w[3] = (t + 1) * t * (t - 1) / 6;
}

/* Modified Bessel function I0 by its power series (converges for all x). */
static double bessel_i0(double x) {
    double sum = 1.0, term = 1.0, q = 0.25 * x * x;
    for (int k = 1; k < 200; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

double orc_kaiser_sinc(double x) {
    if (fabs(x) >= KAISER_HALF) return 0.0;
    double s = (x == 0.0) ? 1.0 : sin(M_PI * x) / (M_PI * x);
    double a = x / KAISER_HALF;
    double win = bessel_i0(KAISER_BETA * sqrt(1.0 - a * a)) / bessel_i0(KAISER_BETA);
    return s * win;
}

#define KAISER_RES 2048
#define KAISER_LEN (4 * KAISER_RES + 2)
static double g_kaiser_table[KAISER_LEN];
static int g_kaiser_ready = 0;

static void kaiser_table_init(void) {
    #pragma omp critical(orc_kaiser_init)
    {
        if (!g_kaiser_ready) {
            for (int j = 0; j < KAISER_LEN; ++j) g_kaiser_table[j] = orc_kaiser_sinc((double)j / KAISER_RES);
            g_kaiser_ready = 1;
        }
    }
}

double orc_kaiser_table_value(int j) {
    if (!g_kaiser_ready) kaiser_table_init();
    return (j >= 0 && j < KAISER_LEN) ? g_kaiser_table[j] : 0.0;
}

/* Weights of the 8-tap Kaiser-sinc at fraction t; taps i-3..i+4. */
void orc_kaiser_weights(double t, double w[8]) {
    if (!g_kaiser_ready) kaiser_table_init();
    double sum = 0.0;
    for (int a = 0; a < 8; ++a) {
        double x = fabs(t - (double)(a - 3)) * KAISER_RES;
        int j = (int)x;
        double f = x - (double)j;
        if (j > KAISER_LEN - 2) j = KAISER_LEN - 2;
        w[a] = g_kaiser_table[j] + f * (g_kaiser_table[j + 1] - g_kaiser_table[j]);
        sum += w[a];
    }
    for (int a = 0; a < 8; ++a) w[a] /= sum;
}

static inline void load_c(const void* base, int is_c128, size_t idx, double* re, double* im) {
    if (is_c128) {
        const double* p = (const double*)base;
        *re = p[2 * idx];
        *im = p[2 * idx + 1];
    } else {
        const float* p = (const float*)base;
        *re = (double)p[2 * idx];
        *im = (double)p[2 * idx + 1];
    }
}

/* ffbp.py:24-29 (lagrange4) plus the linear and kaiser options. */
void orc_sample1(const float* row, long len, double q, int interp, double out[2]) {
    double fi = floor(q);
    out[0] = 0.0;
    out[1] = 0.0;
    if (!(fi > -2147483648.0 && fi < 2147483647.0)) return;
    long i = (long)fi;
    double t = q - (double)i;
    if (interp == ORC_LINEAR) {
        if (i < 0 || i + 1 >= len) return;
        double w0 = 1.0 - t, w1 = t;
        out[0] = w0 * (double)row[2 * i] + w1 * (double)row[2 * (i + 1)];
        out[1] = w0 * (double)row[2 * i + 1] + w1 * (double)row[2 * (i + 1) + 1];
        return;
    }
    if (interp == ORC_KAISER8) {
        if (i < 3 || i + 4 >= len) return;
        double w[8];
        orc_kaiser_weights(t, w);
        double zr = 0.0, zi = 0.0;
        for (int a = 0; a < 8; ++a) {
            zr += w[a] * (double)row[2 * (i + a - 3)];
            zi += w[a] * (double)row[2 * (i + a - 3) + 1];
        }
        out[0] = zr;
        out[1] = zi;
        return;
    }
    if (i < 1 || i + 2 >= len) return;
    double w[4];
    orc_cubic_weights(t, w);
    /* w0*row[i-1] + w1*row[i] + w2*row[i+1] + w3*row[i+2], left to right. */
    double zr = w[0] * (double)row[2 * (i - 1)];
    double zi = w[0] * (double)row[2 * (i - 1) + 1];
    zr = zr + w[1] * (double)row[2 * i];
    zi = zi + w[1] * (double)row[2 * i + 1];
    zr = zr + w[2] * (double)row[2 * (i + 1)];
    zi = zi + w[2] * (double)row[2 * (i + 1) + 1];
    zr = zr + w[3] * (double)row[2 * (i + 2)];
    zi = zi + w[3] * (double)row[2 * (i + 2) + 1];
    out[0] = zr;
    out[1] = zi;
}

/* ffbp.py:32-40 (lagrange4) plus the kaiser option; img is [nr, nu]. */
void orc_sample2(const void* img, int is_c128, long nr, long nu, double r, double u,
                 int interp, double out[2]) {
    out[0] = 0.0;
    out[1] = 0.0;
    double fr = floor(r), fu = floor(u);
    if (!(fr > -2147483648.0 && fr < 2147483647.0)) return;
    if (!(fu > -2147483648.0 && fu < 2147483647.0)) return;
    long ir = (long)fr, iu = (long)fu;
    if (interp == ORC_KAISER8) {
        if (ir < 3 || iu < 3 || ir + 4 >= nr || iu + 4 >= nu) return;
        double wr[8], wu[8];
        orc_kaiser_weights(r - (double)ir, wr);
        orc_kaiser_weights(u - (double)iu, wu);
        double zr = 0.0, zi = 0.0;
        for (int a = 0; a < 8; ++a) {
            for (int b = 0; b < 8; ++b) {
                double vr, vi;
                load_c(img, is_c128, (size_t)(ir + a - 3) * (size_t)nu + (size_t)(iu + b - 3), &vr, &vi);
                zr += vr * wr[a] * wu[b];
                zi += vi * wr[a] * wu[b];
            }
        }
        out[0] = zr;
        out[1] = zi;
        return;
    }
    if (ir < 1 || iu < 1 || ir + 2 >= nr || iu + 2 >= nu) return;
    double wr[4], wu[4];
    orc_cubic_weights(r - (double)ir, wr);
    orc_cubic_weights(u - (double)iu, wu);
    double zr = 0.0, zi = 0.0;
    for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) {
            double vr, vi;
            load_c(img, is_c128, (size_t)(ir + a - 1) * (size_t)nu + (size_t)(iu + b - 1), &vr, &vi);
            /* z += img[...]*wr[a]*wu[b]: (img*wr[a])*wu[b] */
            zr = zr + (vr * wr[a]) * wu[b];
            zi = zi + (vi * wr[a]) * wu[b];
        }
    }
    out[0] = zr;
    out[1] = zi;
}

/* total += weight*v*complex(cos(a), sin(a)): (weight*v) then complex product. */
static inline void mac_rot(double* tr, double* ti, double w, const double v[2], double a) {
    double wr = w * v[0], wi = w * v[1];
    double c = cos(a), s = sin(a);
    *tr = *tr + (wr * c - wi * s);
    *ti = *ti + (wr * s + wi * c);
}

/* ffbp.py:43-55: direct BP at arbitrary xyz, not demodulated. */
void orc_backproject_points(const float* data, long P, long S, const double* pos,
                            const double* xyz, long M, double r_start, double dr, double k,
                            const double* weight, const double* phase, int interp,
                            double* out) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (long p = 0; p < M; ++p) {
        double tr = 0.0, ti = 0.0;
        for (long n = 0; n < P; ++n) {
            double dx = xyz[3 * p] - pos[3 * n];
            double dy = xyz[3 * p + 1] - pos[3 * n + 1];
            double dz = xyz[3 * p + 2] - pos[3 * n + 2];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double v[2];
            orc_sample1(data + 2 * (size_t)n * (size_t)S, S, (r - r_start) / dr, interp, v);
            double a = k * r + phase[n];
            mac_rot(&tr, &ti, weight[n], v, a);
        }
        out[2 * p] = tr;
        out[2 * p + 1] = ti;
    }
}

/* ffbp.py:58-76: one pixel (ir, iu) of the polar BP, demodulated by
 * exp(-j k r); pixels on z = 0. */
static inline void bp_polar_pixel(const float* data, long P, long S, const double* pos,
                                  const double* origin, double r0, double dr, double u0,
                                  double du, long ir, long iu, double r_start, double data_dr,
                                  double k, const double* weight, const double* phase, int side,
                                  int interp, double out[2]) {
    double r = r0 + (double)ir * dr;
    double ground = sqrt(fmax(0.0, r * r - origin[2] * origin[2]));
    double u = u0 + (double)iu * du;
    double x = origin[0] + (double)side * ground * sqrt(fmax(0.0, 1 - u * u));
    double y = origin[1] + ground * u;
    double tr = 0.0, ti = 0.0;
    for (long n = 0; n < P; ++n) {
        double dx = x - pos[3 * n];
        double dy = y - pos[3 * n + 1];
        double d = sqrt(dx * dx + dy * dy + pos[3 * n + 2] * pos[3 * n + 2]);
        double v[2];
        orc_sample1(data + 2 * (size_t)n * (size_t)S, S, (d - r_start) / data_dr, interp, v);
        double a = k * (d - r) + phase[n];
        mac_rot(&tr, &ti, weight[n], v, a);
    }
    out[0] = tr;
    out[1] = ti;
}

void orc_bp_polar(const float* data, long P, long S, const double* pos, const double* origin,
                  double r0, double dr, long nr, double u0, double du, long nu,
                  double r_start, double data_dr, double k, const double* weight,
                  const double* phase, int side, int interp, double* out) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (long ir = 0; ir < nr; ++ir) {
        for (long iu = 0; iu < nu; ++iu) {
            size_t o = (size_t)ir * (size_t)nu + (size_t)iu;
            bp_polar_pixel(data, P, S, pos, origin, r0, dr, u0, du, ir, iu, r_start, data_dr, k,
                           weight, phase, side, interp, out + 2 * o);
        }
    }
}

/* ffbp.py:79-95: one destination pixel (ir, iu) of a child -> parent regrid. */
static inline void regrid_pixel(const void* img, int is_c128, long snr, long snu,
                                const double* src_origin, double sr0, double sdr, double su0,
                                double sdu, const double* dst_origin, double r0, double dr,
                                double u0, double du, long ir, long iu, double k, int side,
                                int interp, double out[2]) {
    double r = r0 + (double)ir * dr;
    double gr = sqrt(fmax(0.0, r * r - dst_origin[2] * dst_origin[2]));
    double u = u0 + (double)iu * du;
    double x = dst_origin[0] + (double)side * gr * sqrt(fmax(0.0, 1 - u * u));
    double y = dst_origin[1] + gr * u;
    double dx = x - src_origin[0];
    double dy = y - src_origin[1];
    double g = sqrt(dx * dx + dy * dy);
    if (g == 0.0) {
        out[0] = 0.0;
        out[1] = 0.0;
        return;
    }
    double rr = sqrt(g * g + src_origin[2] * src_origin[2]);
    double uu = dy / g;
    double v[2];
    orc_sample2(img, is_c128, snr, snu, (rr - sr0) / sdr, (uu - su0) / sdu, interp, v);
    double a = k * (rr - r);
    double c = cos(a), s = sin(a);
    out[0] = v[0] * c - v[1] * s;
    out[1] = v[0] * s + v[1] * c;
}

void orc_regrid(const void* img, int is_c128, long snr, long snu, const double* src_origin,
                double sr0, double sdr, double su0, double sdu, const double* dst_origin,
                double r0, double dr, long nr, double u0, double du, long nu, double k,
                int side, int interp, double* out) {
    #pragma omp parallel for schedule(dynamic, 4)
    for (long ir = 0; ir < nr; ++ir) {
        for (long iu = 0; iu < nu; ++iu) {
            size_t o = (size_t)ir * (size_t)nu + (size_t)iu;
            regrid_pixel(img, is_c128, snr, snu, src_origin, sr0, sdr, su0, sdu, dst_origin, r0,
                         dr, u0, du, ir, iu, k, side, interp, out + 2 * o);
        }
    }
}

/* Block-restricted forms (test infrastructure for sampled full-size checks):
 * only the pixels of the listed blocks -- (block row, block column) pairs of
 * a grid of bs x bs-pixel blocks -- are computed, with the same per-pixel
 * arithmetic as above, and stored into ``out`` as complex64 (out_c64 = 1,
 * the reference's node storage) or complex128.  accumulate = 1 adds to what
 * ``out`` holds in that storage type (ffbp.py:165-167, im += regrid(...)).
 * The other pixels of ``out`` are left untouched. */
static inline void store_pixel(void* out, int out_c64, int accumulate, size_t o, const double v[2]) {
    if (out_c64) {
        float* p = (float*)out + 2 * o;
        float re = (float)v[0], im = (float)v[1];
        if (accumulate) {
            re = p[0] + re;
            im = p[1] + im;
        }
        p[0] = re;
        p[1] = im;
    } else {
        double* p = (double*)out + 2 * o;
        if (accumulate) {
            p[0] = p[0] + v[0];
            p[1] = p[1] + v[1];
        } else {
            p[0] = v[0];
            p[1] = v[1];
        }
    }
}

void orc_bp_polar_blocks(const float* data, long P, long S, const double* pos,
                         const double* origin, double r0, double dr, long nr, double u0,
                         double du, long nu, double r_start, double data_dr, double k,
                         const double* weight, const double* phase, int side, int interp,
                         const int32_t* blocks, long nblocks, int bs, int out_c64, void* out) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < nblocks; ++b) {
        long r_end = (long)blocks[2 * b] * bs + bs, u_end = (long)blocks[2 * b + 1] * bs + bs;
        if (r_end > nr) r_end = nr;
        if (u_end > nu) u_end = nu;
        for (long ir = (long)blocks[2 * b] * bs; ir < r_end; ++ir)
            for (long iu = (long)blocks[2 * b + 1] * bs; iu < u_end; ++iu) {
                double v[2];
                bp_polar_pixel(data, P, S, pos, origin, r0, dr, u0, du, ir, iu, r_start, data_dr,
                               k, weight, phase, side, interp, v);
                store_pixel(out, out_c64, 0, (size_t)ir * (size_t)nu + (size_t)iu, v);
            }
    }
}

void orc_regrid_blocks(const void* img, int is_c128, long snr, long snu, const double* src_origin,
                       double sr0, double sdr, double su0, double sdu, const double* dst_origin,
                       double r0, double dr, long nr, double u0, double du, long nu, double k,
                       int side, int interp, const int32_t* blocks, long nblocks, int bs,
                       int out_c64, int accumulate, void* out) {
    #pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < nblocks; ++b) {
        long r_end = (long)blocks[2 * b] * bs + bs, u_end = (long)blocks[2 * b + 1] * bs + bs;
        if (r_end > nr) r_end = nr;
        if (u_end > nu) u_end = nu;
        for (long ir = (long)blocks[2 * b] * bs; ir < r_end; ++ir)
            for (long iu = (long)blocks[2 * b + 1] * bs; iu < u_end; ++iu) {
                double v[2];
                regrid_pixel(img, is_c128, snr, snu, src_origin, sr0, sdr, su0, sdu, dst_origin,
                             r0, dr, u0, du, ir, iu, k, side, interp, v);
                store_pixel(out, out_c64, accumulate, (size_t)ir * (size_t)nu + (size_t)iu, v);
            }
    }
}

/* Support of a regrid: for every pixel of the listed parent blocks, the
 * child position regrid_pixel samples (same geometry), its stencil box
 * (taps from floor(q) + first), widened by ``margin`` cells and clipped to
 * the child, sets the child blocks (cmask, [ceil(cnr/bs), ceil(cnu/bs)]) it
 * touches.  Points on the child's origin (ground distance 0) read nothing. */
void orc_mark_support(const double* pgeo, long pnr, long pnu, const double* cgeo, long cnr,
                      long cnu, int side, const int32_t* blocks, long nblocks, int bs, int first,
                      int taps, int margin, uint8_t* cmask) {
    const long cbu = (cnu + bs - 1) / bs;
    #pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < nblocks; ++b) {
        long r_end = (long)blocks[2 * b] * bs + bs, u_end = (long)blocks[2 * b + 1] * bs + bs;
        if (r_end > pnr) r_end = pnr;
        if (u_end > pnu) u_end = pnu;
        for (long ir = (long)blocks[2 * b] * bs; ir < r_end; ++ir) {
            double r = pgeo[3] + (double)ir * pgeo[4];
            double gr = sqrt(fmax(0.0, r * r - pgeo[2] * pgeo[2]));
            for (long iu = (long)blocks[2 * b + 1] * bs; iu < u_end; ++iu) {
                double u = pgeo[5] + (double)iu * pgeo[6];
                double x = pgeo[0] + (double)side * gr * sqrt(fmax(0.0, 1 - u * u));
                double y = pgeo[1] + gr * u;
                double dx = x - cgeo[0], dy = y - cgeo[1];
                double g = sqrt(dx * dx + dy * dy);
                if (!(g > 0.0)) continue;
                double rr = sqrt(g * g + cgeo[2] * cgeo[2]);
                double qr = floor((rr - cgeo[3]) / cgeo[4]), qu = floor((dy / g - cgeo[5]) / cgeo[6]);
                double lo_r = qr + first - margin, hi_r = qr + first + taps - 1 + margin;
                double lo_u = qu + first - margin, hi_u = qu + first + taps - 1 + margin;
                if (hi_r < 0 || lo_r > (double)(cnr - 1) || hi_u < 0 || lo_u > (double)(cnu - 1))
                    continue;
                long r0 = lo_r < 0 ? 0 : (long)lo_r, r1 = hi_r > cnr - 1 ? cnr - 1 : (long)hi_r;
                long u0 = lo_u < 0 ? 0 : (long)lo_u, u1 = hi_u > cnu - 1 ? cnu - 1 : (long)hi_u;
                for (long br = r0 / bs; br <= r1 / bs; ++br)
                    for (long bu = u0 / bs; bu <= u1 / bs; ++bu)
                        __atomic_store_n(&cmask[br * cbu + bu], (uint8_t)1, __ATOMIC_RELAXED);
            }
        }
    }
}

/* ffbp.py:189-198: polar image -> arbitrary points, remodulated by exp(+j k r).
 * The point height is ignored, exactly as in the reference (pixels on z=0). */
void orc_evaluate_image(const void* img, int is_c128, long nr, long nu, const double* origin,
                        double r0, double dr, double u0, double du, const double* xyz, long M,
                        double k, int interp, double* out) {
    #pragma omp parallel for schedule(dynamic, 64)
    for (long p = 0; p < M; ++p) {
        double dx = xyz[3 * p] - origin[0];
        double dy = xyz[3 * p + 1] - origin[1];
        double g = sqrt(dx * dx + dy * dy);
        if (g == 0.0) {
            out[2 * p] = 0.0;
            out[2 * p + 1] = 0.0;
            continue;
        }
        double r = sqrt(g * g + origin[2] * origin[2]);
        double u = dy / g;
        double v[2];
        orc_sample2(img, is_c128, nr, nu, (r - r0) / dr, (u - u0) / du, interp, v);
        double c = cos(k * r), s = sin(k * r);
        out[2 * p] = v[0] * c - v[1] * s;
        out[2 * p + 1] = v[0] * s + v[1] * c;
    }
}

/* ffbp.py:201-211: per-(point, pulse) unweighted contributions, [M, P]. */
void orc_target_contributions(const float* data, long P, long S, const double* pos,
                              const double* xyz, long M, double r_start, double dr, double k,
                              int interp, double* out) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (long p = 0; p < M; ++p) {
        for (long n = 0; n < P; ++n) {
            double dx = xyz[3 * p] - pos[3 * n];
            double dy = xyz[3 * p + 1] - pos[3 * n + 1];
            double dz = xyz[3 * p + 2] - pos[3 * n + 2];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double v[2];
            orc_sample1(data + 2 * (size_t)n * (size_t)S, S, (r - r_start) / dr, interp, v);
            double a = k * r;
            double c = cos(a), s = sin(a);
            size_t o = (size_t)p * (size_t)P + (size_t)n;
            out[2 * o] = v[0] * c - v[1] * s;
            out[2 * o + 1] = v[0] * s + v[1] * c;
        }
    }
}

/* FP64 adjoint of backproject_points (no reference counterpart; checked by
 * central finite differences of orc_backproject_points).  With g = dL/dI in
 * the convention dL = Re sum_x conj(g_x) dI_x (complex128 [M]):
 *   gphi[n]   = Re sum_x conj(g_x) j c_xn,   c_xn = w_n s_n(q) e^{j(k d + phase_n)}
 *   gpos[n,:] = Re sum_x conj(g_x) w_n e^{j theta} (s'_n(q)/dr + j k s_n(q)) (p_n - x)/d
 * s' is the derivative of the interpolant with respect to q. */
static void sample1_deriv(const float* row, long len, double q, int interp, double v[2], double dv[2]) {
    v[0] = v[1] = dv[0] = dv[1] = 0.0;
    double fi = floor(q);
    if (!(fi > -2147483648.0 && fi < 2147483647.0)) return;
    long i = (long)fi;
    double t = q - (double)i;
    if (interp == ORC_LINEAR) {
        if (i < 0 || i + 1 >= len) return;
        double ar = row[2 * i], ai = row[2 * i + 1], br = row[2 * i + 2], bi = row[2 * i + 3];
        v[0] = ar + t * (br - ar);
        v[1] = ai + t * (bi - ai);
        dv[0] = br - ar;
        dv[1] = bi - ai;
        return;
    }
    if (i < 1 || i + 2 >= len) return;
    double w[4], dw[4];
    orc_cubic_weights(t, w);
    dw[0] = -(3 * t * t - 6 * t + 2) / 6;
    dw[1] = (3 * t * t - 4 * t - 1) / 2;
    dw[2] = -(3 * t * t - 2 * t - 2) / 2;
    dw[3] = (3 * t * t - 1) / 6;
    for (int a = 0; a < 4; ++a) {
        double sr = row[2 * (i + a - 1)], si = row[2 * (i + a - 1) + 1];
        v[0] += w[a] * sr;
        v[1] += w[a] * si;
        dv[0] += dw[a] * sr;
        dv[1] += dw[a] * si;
    }
}

void orc_backproject_points_adjoint(const float* data, long P, long S, const double* pos,
                                    const double* xyz, long M, double r_start, double dr, double k,
                                    const double* weight, const double* phase, int interp,
                                    const double* grad, double* gphi, double* gpos) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (long n = 0; n < P; ++n) {
        double are = 0, aim = 0, gx = 0, gy = 0, gz = 0;
        for (long p = 0; p < M; ++p) {
            double ex = xyz[3 * p] - pos[3 * n];
            double ey = xyz[3 * p + 1] - pos[3 * n + 1];
            double ez = xyz[3 * p + 2] - pos[3 * n + 2];
            double d = sqrt(ex * ex + ey * ey + ez * ez);
            double v[2], dv[2];
            sample1_deriv(data + 2 * (size_t)n * (size_t)S, S, (d - r_start) / dr, interp, v, dv);
            double th = k * d + phase[n];
            double c = cos(th), s = sin(th);
            double gr = grad[2 * p], gi = grad[2 * p + 1];
            /* h = conj(g) w e^{j th} */
            double hr = weight[n] * (gr * c + gi * s);
            double hi = weight[n] * (gr * s - gi * c);
            are += hr * v[0] - hi * v[1];
            aim += hr * v[1] + hi * v[0];
            double ur = dv[0] / dr - k * v[1];
            double ui = dv[1] / dr + k * v[0];
            double br = (hr * ur - hi * ui) / d;
            gx -= br * ex;
            gy -= br * ey;
            gz -= br * ez;
        }
        gphi[n] = -aim;
        gpos[3 * n] = gx;
        gpos[3 * n + 1] = gy;
        gpos[3 * n + 2] = gz;
    }
}
