// Kernel launch wrapper with optional per-kernel HIP-event timing on the
// launch stream (enabled through agpu_profile_enable).
#pragma once
#include <hip/hip_runtime.h>

namespace agpu {
bool profiling();
// records a start event on stream s; returns a token for prof_end
int prof_begin(const char *name, hipStream_t s);
void prof_end(int token, hipStream_t s);
}  // namespace agpu

#define AGPU_LAUNCH(name, kernel, grid, block, shmem, strm, ...)                       \
    do {                                                                               \
        int tok_ = agpu::profiling() ? agpu::prof_begin(name, strm) : -1;             \
        hipLaunchKernelGGL(kernel, grid, block, shmem, strm, __VA_ARGS__);            \
        hipError_t e_ = hipGetLastError();                                             \
        if (tok_ >= 0) agpu::prof_end(tok_, strm);                                    \
        if (e_ != hipSuccess) return agpu::hip_fail(e_, name, __FILE__, __LINE__);    \
    } while (0)
