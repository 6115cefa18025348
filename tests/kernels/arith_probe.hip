// Test helper, not part of the library: evaluates the solver's own FP64 divide and square root
// (cfd::ddiv, cfd::dsqrt from cfd-art_amd/csrc/physics.h) and the compiler's IEEE a / b and
// sqrt on the GPU for given operands, so the tests can compare them bit for bit.
#include <hip/hip_runtime.h>

#include "physics.h"

__global__ void arith_kernel(const double* a, const double* b, double* out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[4 * i + 0] = cfd::ddiv(a[i], b[i]);
    out[4 * i + 1] = a[i] / b[i];
    out[4 * i + 2] = cfd::dsqrt(a[i]);
    out[4 * i + 3] = sqrt(a[i]);
}

// out: 4 doubles per operand pair (ddiv, a / b, dsqrt(a), sqrt(a)). Returns 0, or -1 on a HIP error.
extern "C" int arith_probe(const double* a, const double* b, double* out, int n)
{
    if (n <= 0) return 0;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n;
    int rc = -1;
    if (hipMalloc(&da, bytes) == hipSuccess && hipMalloc(&db, bytes) == hipSuccess &&
        hipMalloc(&dout, 4 * bytes) == hipSuccess && hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(arith_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, dout, n);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(out, dout, 4 * bytes, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    return rc;
}
