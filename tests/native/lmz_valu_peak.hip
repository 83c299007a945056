// Attainable vector-ALU rates of this build on the GPU: independent FMA
// chains held in registers, FP32 and FP64, compiled with the library's
// flags (no packed FP32).  Measurement instrument for the roofline of the
// pixel kernels (bench.py --extra valu_peak); not part of liblmz.so.
#include <hip/hip_runtime.h>

namespace {

constexpr int kThreads = 256;
constexpr int kChains = 16;

// OP 0: acc = fma(acc, a, b) with a, b uniform (VOP3, SGPR operands);
// 1: acc = fma(va, vb, acc) with va, vb per lane (the VOP2 fmac form);
// 2: acc = acc * va;  3: acc = acc + vb.
template <typename T, int OP>
__global__ __launch_bounds__(kThreads) void k_fma_peak(T *out, int iters, T a,
                                                       T b) {
    T acc[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc[c] = (T)(threadIdx.x + c) * (T)1e-3;
    const T va = a + (T)threadIdx.x * (T)1e-12;
    const T vb = b + (T)threadIdx.x * (T)1e-15;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if (OP == 0) acc[c] = fma(acc[c], a, b);
            else if (OP == 1) acc[c] = fma(va, vb, acc[c]);
            else if (OP == 2) acc[c] = acc[c] * va;
            else acc[c] = acc[c] + vb;
        }
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += acc[c];
    out[(long long)blockIdx.x * kThreads + threadIdx.x] = s;
}

template <typename T, int OP>
int run(int blocks, int iters, void *out, float *ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 1;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        T *o = static_cast<T *>(out);
        // warm launch, then the timed one
        hipLaunchKernelGGL((k_fma_peak<T, OP>), dim3(blocks), dim3(kThreads), 0,
                           0, o, iters, (T)0.999999, (T)1e-7);
        bool ok = hipEventRecord(e0, 0) == hipSuccess;
        hipLaunchKernelGGL((k_fma_peak<T, OP>), dim3(blocks), dim3(kThreads), 0,
                           0, o, iters, (T)0.999999, (T)1e-7);
        ok = ok && hipGetLastError() == hipSuccess;
        ok = ok && hipEventRecord(e1, 0) == hipSuccess;
        ok = ok && hipEventSynchronize(e1) == hipSuccess;
        ok = ok && hipEventElapsedTime(ms, e0, e1) == hipSuccess;
        rc = ok ? 0 : 1;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

template <typename T>
int run_op(int op, int blocks, int iters, void *out, float *ms) {
    switch (op) {
    case 0: return run<T, 0>(blocks, iters, out, ms);
    case 1: return run<T, 1>(blocks, iters, out, ms);
    case 2: return run<T, 2>(blocks, iters, out, ms);
    case 3: return run<T, 3>(blocks, iters, out, ms);
    default: return 1;
    }
}

}  // namespace

extern "C" {

// One launch of blocks * 256 threads, each running 16 independent chains
// of iters instructions of kind op (0: FMA with uniform operands, 1: FMA
// with per-lane operands, 2: multiply, 3: add) on dtype (0: float, 1:
// double).  out: device buffer of blocks * 256 elements of the type.
// Returns 0 on success; ms = the timed launch's duration.
int valu_peak(int dtype, int op, int blocks, int iters, void *out, float *ms) {
    if (blocks <= 0 || iters <= 0 || !out || !ms) return 1;
    return dtype == 0 ? run_op<float>(op, blocks, iters, out, ms)
                      : run_op<double>(op, blocks, iters, out, ms);
}

int valu_peak_chains(void) { return kChains; }
int valu_peak_threads(void) { return kThreads; }

}  // extern "C"
