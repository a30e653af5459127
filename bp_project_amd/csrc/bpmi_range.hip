// Range compression and raw IQ decode: the step before backprojection
// (fast_bp/sar_io.py:18-81 frame decode, :116-139 range_compress).
//
// Range compression is a linear correlation of each pulse with a centred,
// Hamming-windowed LFM reference, computed with zero-padded complex64 FFTs
// (hipFFT / rocFFT, batched over pulses) exactly as the reference's scipy
// code: subtract the row mean, FFT to nfft = next_fast_len(samples + L - 1),
// multiply by the FFT of conj(reversed reference), inverse FFT, crop
// [L - 1 - L/2, + samples).  With oversample 1 the transforms run at a
// length 2^a m >= samples + L - 1 (m = 1, 3, 5, 9; the same correlation)
// and, up to RF_NMAX points, in one kernel: a workgroup per pulse holds the
// row in LDS through both transforms (bpmi_fft.h), so HBM sees one read of
// the raw row and one write of the result.  With oversample > 1 the length-nfft correlation
// is Fourier-interpolated to nfft * oversample first (scipy.signal.resample:
// spectrum copied to both ends, the +N/2 bin halved and mirrored to -N/2),
// which here is done on the product spectrum directly (resample's forward
// FFT of the inverse FFT is the identity).  The hand-written kernels do the
// row mean and padding, the reference spectrum's samples, the spectrum
// product / upsampling layout and the crop; the FFTs of the other cases are
// hipFFT plans cached per (size, batch) whose work area is the caller's
// workspace.
//
// The IQ decode gathers the int16 I/Q pairs of selected frames of a raw
// capture (already in device memory) into complex64 rows.
#include "bpmi_internal.h"
#include "bpmi_fft.h"

#include <hipfft/hipfft.h>

#include <map>
#include <mutex>
#include <utility>

namespace bpmi {

constexpr int RC_THREADS = 256;

// One workgroup per pulse: FP64 row mean, then raw - mean into a row of
// nfft samples padded with zeros.  Rows whose start is 16-byte aligned are
// read and written two samples at a time.
__global__ __launch_bounds__(RC_THREADS) void rc_center_pad_kernel(const float2* __restrict__ raw,
                                                                    long long ld_raw, int samples,
                                                                    int nfft,
                                                                    float2* __restrict__ out) {
    __shared__ double part[2][RC_THREADS / 64];
    const long long p = blockIdx.x;
    const float2* row = raw + p * ld_raw;
    float2* dst = out + p * nfft;
    const bool pairs = ((((uintptr_t)row) | ((uintptr_t)dst)) & 15) == 0;
    const int np = pairs ? samples / 2 : 0;  // whole pairs read as float4
    double sx = 0.0, sy = 0.0;
    for (int i = threadIdx.x; i < np; i += RC_THREADS) {
        const float4 v = ((const float4*)row)[i];
        sx += (double)v.x + (double)v.z;
        sy += (double)v.y + (double)v.w;
    }
    for (int i = 2 * np + threadIdx.x; i < samples; i += RC_THREADS) {
        const float2 v = row[i];
        sx += v.x;
        sy += v.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][wave] = sx;
        part[1][wave] = sy;
    }
    __syncthreads();
    double mx = 0.0, my = 0.0;
#pragma unroll
    for (int w = 0; w < RC_THREADS / 64; ++w) {
        mx += part[0][w];
        my += part[1][w];
    }
    const float fx = (float)(mx / samples), fy = (float)(my / samples);
    for (int i = threadIdx.x; i < np; i += RC_THREADS) {
        const float4 v = ((const float4*)row)[i];
        ((float4*)dst)[i] = make_float4(v.x - fx, v.y - fy, v.z - fx, v.w - fy);
    }
    for (int i = 2 * np + threadIdx.x; i < nfft; i += RC_THREADS) {
        if (i < samples) {
            const float2 v = row[i];
            dst[i] = make_float2(v.x - fx, v.y - fy);
        } else {
            dst[i] = make_float2(0.f, 0.f);
        }
    }
}

// h[n] = conj(ref[L - 1 - n]) for n < L, zero up to nfft, where
// ref[m] = exp(j pi s B / tp tr^2) w[m], tr = (m - floor(L/2)) / fs and w the
// periodic Hamming window 0.54 - 0.46 cos(2 pi m / L) (sar_io.py:124-128);
// phase and window in FP64, rounded to complex64 like the reference's arrays.
__global__ void rc_filter_kernel(int L, double chirp_rate, double fs, int nfft,
                                 float2* __restrict__ h) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nfft) return;
    if (n >= L) {
        h[n] = make_float2(0.f, 0.f);
        return;
    }
    const int m = L - 1 - n;
    const double tr = (double)(m - L / 2) / fs;
    double s, c;
    sincos(M_PI * chirp_rate * tr * tr, &s, &c);
    const float w = (float)(0.54 - 0.46 * cos(2.0 * M_PI * (double)m / (double)L));
    const float2 ref = make_float2((float)c * w, (float)s * w);
    h[n] = make_float2(ref.x, -ref.y);
}

// Product spectrum X[k] H[k] laid out for an inverse FFT of length
// num = N * os (os = 1: in place).  For os > 1: bins [0, N/2] at the start,
// bins (N/2, N) at the end, the +N/2 bin (N even) halved and mirrored to
// num - N/2, zeros between (scipy.signal.resample upsampling).
// Grid: x over the num output bins of a row, y over pulses.
__global__ void rc_multiply_kernel(const float2* __restrict__ X, const float2* __restrict__ H,
                                   int N, int os, float2* __restrict__ Y) {
    const long long num = (long long)N * os;
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= num) return;
    const long long p = blockIdx.y;
    const float2* x = X + p * N;
    long long src = -1;
    float scale = 1.f;
    if (os == 1) {
        src = j;
    } else {
        const int nyq = N / 2 + 1;
        if (j < nyq) {
            src = j;
            if ((N & 1) == 0 && j == N / 2) scale = 0.5f;
        } else if (j >= num - (N - nyq)) {
            src = j - num + N;
        } else if ((N & 1) == 0 && j == num - N / 2) {
            src = N / 2;
            scale = 0.5f;
        }
    }
    float2 v = make_float2(0.f, 0.f);
    if (src >= 0) {
        const float2 a = x[src], b = H[src];
        v = make_float2((a.x * b.x - a.y * b.y) * scale, (a.x * b.y + a.y * b.x) * scale);
    }
    Y[p * num + j] = v;
}

// out[p, i] = y[p, start + i] * scale, i < count.  Grid: x over i, y over p.
__global__ void rc_crop_kernel(const float2* __restrict__ y, long long ldy, int start, int count,
                               float scale, float2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const long long p = blockIdx.y;
    const float2 v = y[p * ldy + start + i];
    out[p * count + i] = make_float2(v.x * scale, v.y * scale);
}

// ---- fused path (oversample 1, n <= RF_NMAX points) ---------------------------
// One workgroup per pulse does the whole matched filter in LDS: row mean
// (FP64), the mean-free row zero-padded to n points, forward FFT, product
// with H = FFT(h) / n, inverse FFT, crop -- one read of the raw row and one
// write of the output row from HBM.

// tw[m] = exp(-2 pi i m / n) in FP64, rounded to complex64.
__global__ void rf_twiddle_kernel(int n, float2* __restrict__ tw) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
}

// H = FFT(h) / n, h as in rc_filter_kernel (one workgroup).
__global__ __launch_bounds__(RF_THREADS) void rf_filter_kernel(FftPlan1 plan, int L,
                                                               double chirp_rate, double fs,
                                                               const float2* __restrict__ tw,
                                                               float2* __restrict__ H) {
    __shared__ float2 buf[RF_SLOTS];
    const int t = threadIdx.x, n = plan.n;
    for (int i = t; i < n; i += RF_THREADS) {
        float2 h = make_float2(0.f, 0.f);
        if (i < L) {
            const int m = L - 1 - i;
            const double tr = (double)(m - L / 2) / fs;
            double sn, cs;
            sincos(M_PI * chirp_rate * tr * tr, &sn, &cs);
            const float w = (float)(0.54 - 0.46 * cos(2.0 * M_PI * (double)m / (double)L));
            h = make_float2((float)cs * w, -((float)sn * w));
        }
        buf[lds_index(i)] = h;
    }
    __syncthreads();
    fft_lds<false>(buf, plan, tw, t);
    const float scale = 1.f / (float)n;
    for (int i = t; i < n; i += RF_THREADS) {
        const float2 v = buf[lds_index(i)];
        H[i] = make_float2(v.x * scale, v.y * scale);
    }
}

// N > 0: transform length fixed at compile time (fft_lds_c); N = 0: plan.n.
template <int N>
__global__ __launch_bounds__(RF_THREADS) void rf_compress_kernel(
        const float2* __restrict__ raw, long long ld_raw, int samples, FftPlan1 plan,
        const float2* __restrict__ H, const float2* __restrict__ tw, int offset,
        float2* __restrict__ out) {
    __shared__ float2 buf[RF_SLOTS];
    __shared__ double part[2][RF_THREADS / 64];
    const int t = threadIdx.x, n = N > 0 ? N : plan.n;
    const long long p = blockIdx.x;
    const float2* row = raw + p * ld_raw;
    double sx = 0.0, sy = 0.0;
    for (int i = t; i < n; i += RF_THREADS) {
        float2 v = make_float2(0.f, 0.f);
        if (i < samples) {
            v = row[i];
            sx += v.x;
            sy += v.y;
        }
        buf[lds_index(i)] = v;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
    }
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0) {
        part[0][wave] = sx;
        part[1][wave] = sy;
    }
    __syncthreads();
    double mx = 0.0, my = 0.0;
#pragma unroll
    for (int w = 0; w < RF_THREADS / 64; ++w) {
        mx += part[0][w];
        my += part[1][w];
    }
    const float fx = (float)(mx / samples), fy = (float)(my / samples);
    for (int i = t; i < samples; i += RF_THREADS) {
        const float2 v = buf[lds_index(i)];
        buf[lds_index(i)] = make_float2(v.x - fx, v.y - fy);
    }
    __syncthreads();
    if constexpr (N > 0) fft_lds_c<false, N>(buf, tw, t);
    else fft_lds<false>(buf, plan, tw, t);
    for (int i = t; i < n; i += RF_THREADS) buf[lds_index(i)] = cmul(buf[lds_index(i)], H[i]);
    __syncthreads();
    if constexpr (N > 0) fft_lds_c<true, N>(buf, tw, t);
    else fft_lds<true>(buf, plan, tw, t);
    float2* dst = out + p * samples;
    for (int i = t; i < samples; i += RF_THREADS) dst[i] = buf[lds_index(offset + i)];
}

// Transform lengths with a compiled kernel (2^a m, m = 1, 3, 5, 9, 1024 to
// 9216 points: the lengths range compression picks for 600 to 8800 samples
// of a 400-tap reference); others run the runtime-plan kernel.
template <int N>
static void rf_launch(dim3 grid, hipStream_t hs, const float2* raw, long long ld_raw, int samples,
                      const FftPlan1& plan, const float2* H, const float2* tw, int offset,
                      float2* out) {
    hipLaunchKernelGGL(rf_compress_kernel<N>, grid, dim3(RF_THREADS), 0, hs, raw, ld_raw, samples,
                       plan, H, tw, offset, out);
}

template <int... Ns>
struct RfLengths {
    static void launch(int n, dim3 grid, hipStream_t hs, const float2* raw, long long ld_raw,
                       int samples, const FftPlan1& plan, const float2* H, const float2* tw,
                       int offset, float2* out) {
        const bool done = ((n == Ns ? (rf_launch<Ns>(grid, hs, raw, ld_raw, samples, plan, H, tw,
                                                      offset, out),
                                       true)
                                    : false) ||
                           ...);
        if (!done) rf_launch<0>(grid, hs, raw, ld_raw, samples, plan, H, tw, offset, out);
    }
};
using RfCompiled = RfLengths<1024, 1152, 1280, 1536, 2048, 2304, 2560, 3072, 4096, 4608, 5120,
                             6144, 8192, 9216>;

// Radices of the fused path for n points (make_fft_plan), or false.
static bool fused_plan(int n, FftPlan1* plan) {
    *plan = make_fft_plan(n);
    return plan->passes > 0;
}

// Frame decode: out[i, s] = (I, Q) int16 pair s of frame index[i]; an index
// outside [0, nframes) gives a zero row.
__global__ void decode_iq16_kernel(const unsigned char* __restrict__ frames, long long nframes,
                                   long long frame_bytes, long long iq_offset, int samples,
                                   const long long* __restrict__ index, long long count,
                                   float2* __restrict__ out) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count * samples) return;
    const long long i = idx / samples;
    const long long s = idx - i * samples;
    const long long f = index[i];
    float2 v = make_float2(0.f, 0.f);
    if (f >= 0 && f < nframes) {
        const short2 q = *(const short2*)(frames + f * frame_bytes + iq_offset + 4 * s);
        v = make_float2((float)q.x, (float)q.y);
    }
    out[idx] = v;
}

// ---- hipFFT plans, cached per (length, batch) --------------------------------
struct FftPlan {
    hipfftHandle handle = nullptr;
    size_t work = 0;
};

static std::mutex g_fft_mutex;
static std::map<std::pair<int, int>, FftPlan> g_fft_plans;

static int check_fft(hipfftResult r, const char* what) {
    if (r == HIPFFT_SUCCESS) return BPMI_OK;
    return set_error(BPMI_ERR_HIP, "%s failed (hipfft status %d)", what, (int)r);
}

// Batched in-place C2C plan of length n; the work area is set per call.
static int fft_plan(int n, int batch, FftPlan** out) {
    std::lock_guard<std::mutex> lock(g_fft_mutex);
    auto key = std::make_pair(n, batch);
    auto it = g_fft_plans.find(key);
    if (it != g_fft_plans.end()) {
        *out = &it->second;
        return BPMI_OK;
    }
    FftPlan plan;
    int st = check_fft(hipfftCreate(&plan.handle), "hipfftCreate");
    if (st) return st;
    st = check_fft(hipfftSetAutoAllocation(plan.handle, 0), "hipfftSetAutoAllocation");
    int dims[1] = {n};
    if (!st)
        st = check_fft(hipfftMakePlanMany(plan.handle, 1, dims, nullptr, 1, n, nullptr, 1, n,
                                          HIPFFT_C2C, batch, &plan.work),
                       "hipfftMakePlanMany");
    if (st) {
        hipfftDestroy(plan.handle);
        return st;
    }
    auto res = g_fft_plans.emplace(key, plan);
    *out = &res.first->second;
    return BPMI_OK;
}

static int fft_exec(FftPlan* plan, void* data, void* work, int direction, hipStream_t stream) {
    std::lock_guard<std::mutex> lock(g_fft_mutex);
    int st = check_fft(hipfftSetStream(plan->handle, stream), "hipfftSetStream");
    if (!st) st = check_fft(hipfftSetWorkArea(plan->handle, work), "hipfftSetWorkArea");
    if (!st)
        st = check_fft(hipfftExecC2C(plan->handle, (hipfftComplex*)data, (hipfftComplex*)data,
                                     direction),
                       "hipfftExecC2C");
    return st;
}

struct RcLayout {
    int n;                        // transform length used
    bool fused;                   // one-kernel LDS path (no hipFFT plans)
    FftPlan1 plan;
    size_t h, x, y, work, total;  // byte offsets in the workspace (fused: h, x = twiddles)
    FftPlan *filt, *fwd, *inv;
};

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Transform length actually used.  With oversample 1 the output is a crop
// of the linear correlation, which every length >= samples + L - 1 gives
// alike, so the smallest 2^a m (m = 1, 3, 5, 9) of at least that length
// replaces the reference's next_fast_len: rocFFT runs these in its fast
// kernels (8192 rows: 1.49 ms at 9216 against 3.56 ms at 8624 = 2^4 7^2 11).
// With oversample > 1 the resampled length nfft * oversample defines the
// result (scipy.signal.resample interpolates the length-nfft period), so
// nfft is kept.
static int transform_length(int samples, int L, int nfft, int os) {
    if (os != 1) return nfft;
    const long long need = (long long)samples + L - 1;
    long long best = 0;
    for (int m : {1, 3, 5, 9}) {
        long long n = m;
        while (n < need) n *= 2;
        if (!best || n < best) best = n;
    }
    return best > 0x7fffffffLL ? nfft : (int)best;
}

static int rc_layout(int pulses, int samples, int L, int nfft, int os, RcLayout* lay) {
    if (pulses <= 0 || samples <= 0) return set_error(BPMI_ERR_INVALID, "empty raw data");
    if (L <= 0) return set_error(BPMI_ERR_INVALID, "reference length must be positive");
    if (os < 1 || os > 64) return set_error(BPMI_ERR_INVALID, "oversample must be 1..64");
    if (nfft < samples + L - 1)
        return set_error(BPMI_ERR_INVALID, "nfft %d below samples + L - 1 = %d (circular wrap)",
                         nfft, samples + L - 1);
    if ((long long)nfft * os > 0x7fffffffLL) return set_error(BPMI_ERR_INVALID, "nfft too large");
    nfft = lay->n = transform_length(samples, L, nfft, os);
    lay->fused = os == 1 && fused_plan(nfft, &lay->plan);
    if (lay->fused) {
        lay->h = 0;
        lay->x = align256((size_t)nfft * 8);
        lay->y = lay->work = lay->x + align256((size_t)nfft * 8);
        lay->total = lay->work;
        lay->filt = lay->fwd = lay->inv = nullptr;
        return BPMI_OK;
    }
    int st = fft_plan(nfft, 1, &lay->filt);
    if (!st) st = fft_plan(nfft, pulses, &lay->fwd);
    if (!st) st = fft_plan(nfft * os, pulses, &lay->inv);
    if (st) return st;
    const size_t c64 = 8;
    lay->h = 0;
    lay->x = align256((size_t)nfft * c64);
    lay->y = lay->x + align256((size_t)pulses * nfft * c64);
    lay->work = lay->y + (os > 1 ? align256((size_t)pulses * nfft * os * c64) : 0);
    size_t w = lay->filt->work;
    if (lay->fwd->work > w) w = lay->fwd->work;
    if (lay->inv->work > w) w = lay->inv->work;
    lay->total = lay->work + align256(w);
    return BPMI_OK;
}

static int reference_length(double tp, double fs, int* L) {
    const double l = tp * fs;
    if (!(l >= 0.5) || l > 1.0e8) return set_error(BPMI_ERR_INVALID, "bad pulse width x rate");
    *L = (int)nearbyint(l);
    return BPMI_OK;
}

}  // namespace bpmi

using namespace bpmi;

extern "C" {

int bpmi_range_compress_workspace_bytes(int pulses, int samples, double tp, double fs, int nfft,
                                        int oversample, size_t* bytes) {
    if (!bytes) return set_error(BPMI_ERR_INVALID, "null bytes");
    int L = 0;
    int st = reference_length(tp, fs, &L);
    if (st) return st;
    RcLayout lay;
    st = rc_layout(pulses, samples, L, nfft, oversample, &lay);
    if (st) return st;
    *bytes = lay.total;
    return BPMI_OK;
}

int bpmi_range_compress(const void* raw, int pulses, int samples, long long ld_raw, double tp,
                        double fs, double bandwidth, int chirp_sign, int nfft, int oversample,
                        void* out, void* workspace, size_t workspace_bytes, void* stream) {
    if (!raw || !out || !workspace) return set_error(BPMI_ERR_INVALID, "null argument");
    if (ld_raw < samples) return set_error(BPMI_ERR_INVALID, "ld_raw below samples");
    if (chirp_sign != 1 && chirp_sign != -1)
        return set_error(BPMI_ERR_INVALID, "chirp_sign must be +1 or -1");
    if (!(bandwidth > 0.0) || !(fs > 0.0) || !(tp > 0.0))
        return set_error(BPMI_ERR_INVALID, "tp, fs and bandwidth must be positive");
    int L = 0;
    int st = reference_length(tp, fs, &L);
    if (st) return st;
    RcLayout lay;
    st = rc_layout(pulses, samples, L, nfft, oversample, &lay);
    if (st) return st;
    if (workspace_bytes < lay.total)
        return set_error(BPMI_ERR_INVALID, "workspace of %zu bytes below the %zu needed",
                         workspace_bytes, lay.total);
    nfft = lay.n;
    char* ws = (char*)workspace;
    hipStream_t hs = (hipStream_t)stream;
    if (lay.fused) {
        float2* H = (float2*)(ws + lay.h);
        float2* tw = (float2*)(ws + lay.x);
        hipLaunchKernelGGL(rf_twiddle_kernel, dim3(grid_1d(nfft, 256)), dim3(256), 0, hs, nfft, tw);
        hipLaunchKernelGGL(rf_filter_kernel, dim3(1), dim3(RF_THREADS), 0, hs, lay.plan, L,
                           (double)chirp_sign * bandwidth / tp, fs, (const float2*)tw, H);
        st = check_launch("rf_filter_kernel");
        if (st) return st;
        RfCompiled::launch(nfft, dim3(pulses), hs, (const float2*)raw, ld_raw, samples, lay.plan,
                           (const float2*)H, (const float2*)tw, L - 1 - L / 2, (float2*)out);
        return check_launch("rf_compress_kernel");
    }
    float2* H = (float2*)(ws + lay.h);
    float2* X = (float2*)(ws + lay.x);
    float2* Y = oversample > 1 ? (float2*)(ws + lay.y) : X;
    void* work = ws + lay.work;
    const long long num = (long long)nfft * oversample;
    hipLaunchKernelGGL(rc_filter_kernel, dim3(grid_1d(nfft, 256)), dim3(256), 0, hs, L,
                       (double)chirp_sign * bandwidth / tp, fs, nfft, H);
    st = check_launch("rc_filter_kernel");
    if (!st) st = fft_exec(lay.filt, H, work, HIPFFT_FORWARD, hs);
    if (st) return st;
    hipLaunchKernelGGL(rc_center_pad_kernel, dim3(pulses), dim3(RC_THREADS), 0, hs,
                       (const float2*)raw, ld_raw, samples, nfft, X);
    st = check_launch("rc_center_pad_kernel");
    if (!st) st = fft_exec(lay.fwd, X, work, HIPFFT_FORWARD, hs);
    if (st) return st;
    constexpr int ROWS_PER_LAUNCH = 65535;  // grid y limit
    for (int p0 = 0; p0 < pulses; p0 += ROWS_PER_LAUNCH) {
        const int rows = pulses - p0 < ROWS_PER_LAUNCH ? pulses - p0 : ROWS_PER_LAUNCH;
        hipLaunchKernelGGL(rc_multiply_kernel, dim3(grid_1d(num, 256), rows), dim3(256), 0, hs,
                           X + (size_t)p0 * nfft, H, nfft, oversample, Y + (size_t)p0 * num);
    }
    st = check_launch("rc_multiply_kernel");
    if (!st) st = fft_exec(lay.inv, Y, work, HIPFFT_BACKWARD, hs);
    if (st) return st;
    const int offset = (L - 1 - L / 2) * oversample;
    const int count = samples * oversample;
    for (int p0 = 0; p0 < pulses; p0 += ROWS_PER_LAUNCH) {
        const int rows = pulses - p0 < ROWS_PER_LAUNCH ? pulses - p0 : ROWS_PER_LAUNCH;
        hipLaunchKernelGGL(rc_crop_kernel, dim3(grid_1d(count, 256), rows), dim3(256), 0, hs,
                           Y + (size_t)p0 * num, num, offset, count, (float)(1.0 / nfft),
                           (float2*)out + (size_t)p0 * count);
    }
    return check_launch("rc_crop_kernel");
}

int bpmi_decode_iq16(const void* frames, long long nframes, long long frame_bytes,
                     long long iq_offset, int samples, const long long* frame_index,
                     long long count, void* out, void* stream) {
    if (!frames || !frame_index || !out) return set_error(BPMI_ERR_INVALID, "null argument");
    if (samples <= 0 || count < 0 || nframes < 0) return set_error(BPMI_ERR_INVALID, "bad sizes");
    if (frame_bytes % 4 || iq_offset % 4 || iq_offset < 0 ||
        iq_offset + 4LL * samples > frame_bytes)
        return set_error(BPMI_ERR_INVALID, "IQ block must be 4-byte aligned inside the frame");
    if (count == 0) return BPMI_OK;
    hipLaunchKernelGGL(decode_iq16_kernel, dim3(grid_1d(count * samples, 256)), dim3(256), 0,
                       (hipStream_t)stream, (const unsigned char*)frames, nframes, frame_bytes,
                       iq_offset, samples, frame_index, count, (float2*)out);
    return check_launch("decode_iq16_kernel");
}

}  // extern "C"
