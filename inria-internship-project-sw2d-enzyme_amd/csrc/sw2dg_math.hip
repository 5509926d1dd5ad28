// sw2dg_math.hip -- the device's exp, log and cbrt (sw_math.h) on arrays,
// so tests can compare them with a host build of the same code.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/sw2dg.h"
#include "sw_math.h"

namespace {

__global__ void k_math(int fn, int64_t n, const double *__restrict__ x, double *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = fn == 0 ? sw_exp(x[i]) : fn == 1 ? sw_log(x[i]) : sw_cbrt(x[i]);
}

}  // namespace

extern "C" int sw2dg_math_eval(int32_t fn, int64_t n, const double *x, double *y, int32_t device) {
    if (fn < 0 || fn > 2 || n < 0 || (n > 0 && (!x || !y))) return SW2DG_ERR_ARG;
    if (n == 0) return SW2DG_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SW2DG_ERR_NO_DEVICE;
    if (device >= 0 && hipSetDevice(device) != hipSuccess) return SW2DG_ERR_HIP;
    double *dx = nullptr, *dy = nullptr;
    int st = SW2DG_OK;
    if (hipMalloc(&dx, n * sizeof(double)) != hipSuccess || hipMalloc(&dy, n * sizeof(double)) != hipSuccess) {
        st = SW2DG_ERR_HIP;
    } else if (hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        st = SW2DG_ERR_HIP;
    } else {
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(k_math, dim3((unsigned)blocks), dim3(256), 0, 0, fn, n, dx, dy);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            st = SW2DG_ERR_HIP;
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return st;
}
