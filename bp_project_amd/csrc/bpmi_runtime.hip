// C-ABI runtime of libbpmi: status codes, device memory, copies, streams,
// events and HIP graph capture.  Every function returns 0 on success and a
// nonzero status otherwise; bpmi_last_error() returns the message of the
// last failure on the calling thread.
#include "bpmi_internal.h"

#include <cstdio>
#include <cstring>
#include <string>

namespace bpmi {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return BPMI_OK;
    return set_error(BPMI_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int check_launch(const char* what) {
    return check_hip(hipGetLastError(), what);
}

__global__ void zero_fill_kernel(uint4* __restrict__ p, size_t n16, unsigned char* tail, size_t ntail) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = i; j < n16; j += stride) p[j] = make_uint4(0, 0, 0, 0);
    if (i < ntail) tail[i] = 0;
}

}  // namespace bpmi

using namespace bpmi;

extern "C" {

int bpmi_version(void) { return BPMI_VERSION; }

int bpmi_last_error(char* buf, size_t n) {
    if (buf && n) {
        std::strncpy(buf, g_last_error.c_str(), n - 1);
        buf[n - 1] = 0;
    }
    return (int)g_last_error.size();
}

int bpmi_device_count(int* count) {
    if (!count) return set_error(BPMI_ERR_INVALID, "count is null");
    return check_hip(hipGetDeviceCount(count), "hipGetDeviceCount");
}

int bpmi_set_device(int device) { return check_hip(hipSetDevice(device), "hipSetDevice"); }

int bpmi_get_device(int* device) {
    if (!device) return set_error(BPMI_ERR_INVALID, "device is null");
    return check_hip(hipGetDevice(device), "hipGetDevice");
}

int bpmi_device_info(int device, char* name, size_t name_len, char* arch, size_t arch_len,
                     int* cu_count, int* clock_khz, size_t* total_mem) {
    hipDeviceProp_t p;
    int st = check_hip(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    if (st) return st;
    if (name && name_len) {
        std::strncpy(name, p.name, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (arch && arch_len) {
        std::strncpy(arch, p.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    if (total_mem) *total_mem = p.totalGlobalMem;
    return BPMI_OK;
}

int bpmi_runtime_version(int* version) {
    if (!version) return set_error(BPMI_ERR_INVALID, "version is null");
    return check_hip(hipRuntimeGetVersion(version), "hipRuntimeGetVersion");
}

int bpmi_wall_clock_khz(int* khz) {
    if (!khz) return set_error(BPMI_ERR_INVALID, "khz is null");
    int dev = 0;
    int st = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (st) return st;
    st = check_hip(hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev),
                   "hipDeviceGetAttribute(wall clock rate)");
    if (st) return st;
    if (*khz <= 0) return set_error(BPMI_ERR_HIP, "device reports no wall clock rate");
    return BPMI_OK;
}

int bpmi_mem_info(size_t* free_bytes, size_t* total_bytes) {
    return check_hip(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int bpmi_malloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(BPMI_ERR_INVALID, "ptr is null");
    *ptr = nullptr;
    if (bytes == 0) return BPMI_OK;
    return check_hip(hipMalloc(ptr, bytes), "hipMalloc");
}

int bpmi_free(void* ptr) {
    if (!ptr) return BPMI_OK;
    return check_hip(hipFree(ptr), "hipFree");
}

int bpmi_host_alloc(void** ptr, size_t bytes) {
    if (!ptr) return set_error(BPMI_ERR_INVALID, "ptr is null");
    *ptr = nullptr;
    if (bytes == 0) return BPMI_OK;
    return check_hip(hipHostMalloc(ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
}

int bpmi_host_free(void* ptr) {
    if (!ptr) return BPMI_OK;
    return check_hip(hipHostFree(ptr), "hipHostFree");
}

int bpmi_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return BPMI_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                     "hipMemcpyAsync(H2D)");
}

int bpmi_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return BPMI_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                     "hipMemcpyAsync(D2H)");
}

int bpmi_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return BPMI_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream),
                     "hipMemcpyAsync(D2D)");
}

int bpmi_zero(void* ptr, size_t bytes, void* stream) {
    if (bytes == 0) return BPMI_OK;
    if (!ptr) return set_error(BPMI_ERR_INVALID, "zero: null pointer");
    if (((uintptr_t)ptr & 15) != 0)
        return set_error(BPMI_ERR_INVALID, "zero: pointer must be 16-byte aligned");
    const size_t n16 = bytes / 16, ntail = bytes - n16 * 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (uint4*)ptr, n16, (unsigned char*)ptr + n16 * 16, ntail);
    return check_launch("zero_fill_kernel");
}

int bpmi_stream_create(void** stream) {
    if (!stream) return set_error(BPMI_ERR_INVALID, "stream is null");
    hipStream_t s;
    int st = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    *stream = st ? nullptr : (void*)s;
    return st;
}

int bpmi_stream_destroy(void* stream) {
    if (!stream) return BPMI_OK;
    return check_hip(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}

int bpmi_stream_synchronize(void* stream) {
    return check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int bpmi_device_synchronize(void) {
    return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int bpmi_event_create(void** event) {
    if (!event) return set_error(BPMI_ERR_INVALID, "event is null");
    hipEvent_t e;
    int st = check_hip(hipEventCreate(&e), "hipEventCreate");
    *event = st ? nullptr : (void*)e;
    return st;
}

int bpmi_event_destroy(void* event) {
    if (!event) return BPMI_OK;
    return check_hip(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}

int bpmi_event_record(void* event, void* stream) {
    return check_hip(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

int bpmi_event_synchronize(void* event) {
    return check_hip(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

int bpmi_event_elapsed(float* ms, void* start, void* end) {
    if (!ms) return set_error(BPMI_ERR_INVALID, "ms is null");
    return check_hip(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end),
                     "hipEventElapsedTime");
}

int bpmi_graph_begin(void* stream) {
    if (!stream) return set_error(BPMI_ERR_INVALID, "graph capture needs a non-null stream");
    return check_hip(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
}

int bpmi_graph_end(void* stream, void** graph_exec) {
    if (!graph_exec) return set_error(BPMI_ERR_INVALID, "graph_exec is null");
    *graph_exec = nullptr;
    hipGraph_t g = nullptr;
    int st = check_hip(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
    if (st) return st;
    hipGraphExec_t ge = nullptr;
    st = check_hip(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    hipGraphDestroy(g);
    if (st) return st;
    *graph_exec = (void*)ge;
    return BPMI_OK;
}

int bpmi_graph_launch(void* graph_exec, void* stream) {
    if (!graph_exec) return set_error(BPMI_ERR_INVALID, "graph_exec is null");
    return check_hip(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream),
                     "hipGraphLaunch");
}

int bpmi_graph_destroy(void* graph_exec) {
    if (!graph_exec) return BPMI_OK;
    return check_hip(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

}  // extern "C"
