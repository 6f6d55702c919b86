// valu_peak.hip -- VALU throughput probe for gfx950 (measurement support).
//
// Measures the sustained issue rate of three VALU instructions the disparity
// kernels are built from, with no memory traffic inside the timed loop:
//   op 0  v_add_u32          one 32-bit lane-op per lane
//   op 1  v_pk_add_u16       two 16-bit lane-ops per lane (the SGM's packed state)
//   op 2  v_bcnt_u32_b32     popcount with addend (the Hamming costs)
// Each lane keeps 16 independent accumulators, so the dependent-issue latency
// is hidden and the SIMDs issue back to back.  The instructions are written
// as inline asm so that the compiler can neither fold the chains nor change
// their count.  The grid covers every CU many times over.
//
// extern "C" int valu_peak_run(int op, int blocks, int iters, double* ms, double* lane_ops)
//   runs the kernel once to warm up, then times one launch with HIP events;
//   lane_ops = issued instructions x 64 lanes (x2 for the packed op).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kThreads = 256;
constexpr int kAcc = 16;

template <int OP>
__global__ __launch_bounds__(kThreads) void k_valu(uint32_t* out, int iters, uint32_t seed)
{
    const uint32_t tid = blockIdx.x * kThreads + threadIdx.x;
    uint32_t a[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; k++) a[k] = tid * 2654435761u + seed + k;
    const uint32_t b = tid ^ seed;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < kAcc; k++) {
            if (OP == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            else if (OP == 1) asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            else asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(a[k]) : "v"(b));
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kAcc; k++) s ^= a[k];
    out[tid] = s;
}

}  // namespace

extern "C" __attribute__((visibility("default"))) int valu_peak_run(int op, int blocks, int iters, double* ms,
                                                                    double* lane_ops)
{
    if (op < 0 || op > 2 || blocks <= 0 || iters <= 0 || !ms || !lane_ops) return 1;
    uint32_t* out = nullptr;
    const size_t n = (size_t)blocks * kThreads;
    if (hipMalloc((void**)&out, n * sizeof(uint32_t)) != hipSuccess) return 2;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto launch = [&]() {
        if (op == 0) hipLaunchKernelGGL(k_valu<0>, dim3(blocks), dim3(kThreads), 0, 0, out, iters, 12345u);
        else if (op == 1) hipLaunchKernelGGL(k_valu<1>, dim3(blocks), dim3(kThreads), 0, 0, out, iters, 12345u);
        else hipLaunchKernelGGL(k_valu<2>, dim3(blocks), dim3(kThreads), 0, 0, out, iters, 12345u);
    };
    launch();
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(e0, 0);
    launch();
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return 3;
    *ms = t;
    *lane_ops = (double)n * iters * kAcc * (op == 1 ? 2.0 : 1.0);
    return 0;
}
