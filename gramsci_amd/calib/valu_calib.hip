// valu_calib.hip -- calibration kernel for the vector-ALU issue ceiling.
//
// Every lane runs 8 independent chains of FP32 fused multiply-adds, so the
// kernel issues nothing but full-rate VALU instructions (plus a few loop
// instructions) with no memory traffic. Its measured rate is the practical
// ceiling against which the instruction issue of the counting kernels is read:
// FP32 lane-operations per second and VALU wave-instructions per CU cycle
// (the latter from the SQ_INSTS_VALU and GRBM_GUI_ACTIVE counters of the same
// launch). Built with packed FP32 instructions switched off, like the library.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

constexpr int kThreads = 256;
constexpr int kChains = 8;
constexpr int kUnroll = 16;

__global__ __launch_bounds__(kThreads) void k_valu_fma32(float* out, int iters, float a, float b)
{
    float x[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) x[k] = (float)threadIdx.x * 1e-6f + (float)k;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int r = 0; r < kUnroll; ++r) {
#pragma unroll
            for (int k = 0; k < kChains; ++k) x[k] = fmaf(x[k], a, b);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += x[k];
    // keeps the chains live; out holds one float per thread of the grid
    if (s == 12345.0f) out[(size_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// FMAs per thread per outer iteration
int gs_calib_fma_per_iter(void) { return kChains * kUnroll; }
int gs_calib_threads(void) { return kThreads; }

// Launches k_valu_fma32 on `blocks` work-groups of 256 threads, `iters` outer
// iterations, on the null stream, and returns its duration in *ms (HIP events).
// out: device buffer of at least blocks * 256 floats. Returns 0 or a hipError_t.
int gs_calib_fma32(int blocks, int iters, float* out, int64_t out_floats, float* ms)
{
    if (blocks <= 0 || iters <= 0 || !out || !ms || out_floats < (int64_t)blocks * kThreads) return -1;
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return (int)e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) { (void)hipEventDestroy(e0); return (int)e; }
    e = hipEventRecord(e0, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_valu_fma32, dim3(blocks), dim3(kThreads), 0, 0, out, iters, 0.999999f, 1e-7f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (int)e;
}

}  // extern "C"
