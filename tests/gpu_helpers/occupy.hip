// Test helper (not part of librlgpu): a kernel that holds CUs for a fixed
// time, so that kernels launched meanwhile on another stream run on the few
// CUs left free and their workgroups start in several waves.  Each
// workgroup takes 100 KiB of LDS, so no other workgroup needing more than
// the remaining 60 KiB fits beside it on its CU.  The wait is a sleep loop on
// the constant-rate wall clock with an iteration cap, so every workgroup
// ends after at most about a second whatever the clock reads.
#include <hip/hip_runtime.h>

constexpr int PAD = 100 * 1024 / 4;

__global__ __launch_bounds__(64) void k_occupy(long long ticks, float *sink) {
    __shared__ float pad[PAD];
    for (int i = threadIdx.x; i < PAD; i += 64) pad[i] = (float)i;
    __syncthreads();
    const long long t0 = wall_clock64();
    for (unsigned spins = 0; spins < (1u << 22) && wall_clock64() - t0 < ticks; ++spins)
        __builtin_amdgcn_s_sleep(8);
    __syncthreads();
    if (threadIdx.x == 0) sink[blockIdx.x] = pad[(blockIdx.x * 37) % PAD];
}

extern "C" {

// Creates a stream of the current device (default priority).
int occ_stream_create(void **out) {
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    *out = s;
    return (int)e;
}

int occ_stream_destroy(void *s) { return (int)hipStreamDestroy(static_cast<hipStream_t>(s)); }

int occ_stream_sync(void *s) { return (int)hipStreamSynchronize(static_cast<hipStream_t>(s)); }

// `blocks` workgroups that each hold a CU for `ms` milliseconds; `sink`
// receives one float per workgroup.
int occ_launch(void *stream, int blocks, double ms, float *sink) {
    int dev = 0, khz = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (e != hipSuccess) return (int)e;
    if (blocks < 1 || khz <= 0 || !(ms > 0) || ms > 500) return (int)hipErrorInvalidValue;
    const long long ticks = (long long)(ms * khz);
    hipLaunchKernelGGL(k_occupy, dim3(blocks), dim3(64), 0, static_cast<hipStream_t>(stream), ticks, sink);
    return (int)hipGetLastError();
}

}  // extern "C"
