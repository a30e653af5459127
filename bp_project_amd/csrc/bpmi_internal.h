// Host-side internals shared by the libbpmi translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "../../include/bpmi.h"
#include "bpmi_common.h"

namespace bpmi {

int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);
int check_launch(const char* what);

inline unsigned grid_1d(long long count, int block) {
    long long g = (count + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace bpmi
