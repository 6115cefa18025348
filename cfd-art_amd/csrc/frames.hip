// Frame products on the device: the schlieren image of the reference's gnuplot_schlieren
// (euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:1142-1185,
// allaire_2d/sourcecode/allaire_diffuse_io.cpp:1109-1150) computed where the state lives, so
// a frame moves nx*ny doubles to the host instead of the state's ncomp planes plus a host pass.
#include "frames.h"

#include "sweep_device.h"

namespace cfd {

namespace {

__device__ __forceinline__ double density(const GridView& g, int y, int x, bool allaire)
{
    const int ys = mirror(y, g.ny), xs = mirror(x, g.nx);
    const double* b = g.cur + (long long)ys * g.pitch + xs;
    return allaire ? b[0] + b[g.plane] : b[0];
}

// |grad rho| by central differences per interior cell (ghost cells: the mirror of the current
// state, density being even under every boundary condition), packed row-major into out, and
// the maximum (bit pattern of a non-negative double) into *mx.
__global__ __launch_bounds__(256) void schlieren_kernel(GridView g, int allaire, double* out, unsigned long long* mx)
{
    const long long cells = (long long)g.nx * g.ny;
    double m = 0.0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells;
         c += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(c % g.nx), i = (int)(c / g.nx);
        const double rho_L = density(g, i, j - 1, allaire);
        const double rho_R = density(g, i, j + 1, allaire);
        const double rho_T = density(g, i + 1, j, allaire);
        const double rho_B = density(g, i - 1, j, allaire);
        const double gradx = (rho_R - rho_L) / (2.0 * g.dx);
        const double grady = (rho_T - rho_B) / (2.0 * g.dy);
        const double v = sqrt(gradx * gradx + grady * grady);
        out[c] = v;
        m = smax(m, v);
    }
    m = wave_max(m);
    __shared__ double s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s[0] = smax(s[0], s[w]);
        atomicMax(mx, (unsigned long long)__double_as_longlong(s[0]));
    }
}

__global__ __launch_bounds__(256) void normalise_kernel(double* out, long long n, const unsigned long long* mx)
{
    const double m = __longlong_as_double((long long)*mx);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        out[c] = out[c] / m;
}

int blocks_for(long long items)
{
    long long b = (items + 255) / 256;
    return (int)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

}  // namespace

hipError_t launch_schlieren(const GridView& g, int model, double* out, unsigned long long* mx, hipStream_t st)
{
    hipError_t e = hipMemsetAsync(mx, 0, sizeof(unsigned long long), st);
    if (e != hipSuccess) return e;
    const long long n = (long long)g.nx * g.ny;
    hipLaunchKernelGGL(schlieren_kernel, dim3(blocks_for(n)), dim3(256), 0, st, g, model == ALLAIRE ? 1 : 0, out, mx);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(normalise_kernel, dim3(blocks_for(n)), dim3(256), 0, st, out, n, mx);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t launch_copy(const double* src, double* dst, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

}  // namespace cfd
