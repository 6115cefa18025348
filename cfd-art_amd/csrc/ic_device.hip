// Device-side initial conditions and state checksums.
//
// The generators fill the interior SoA planes of a grid directly in HBM, bitwise equal to the
// host generators of host/ic.cpp (same disc / bubble draws, same per-cell arithmetic from
// ic_common.h, applied in the same order), so that admitting a request does not need a host
// loop over its cells or a host-to-device copy of its state. The cell fraction and blends
// follow euler_equations/euler_simulation_code/sourcecode/euler_io.cpp:344-389 (sub-sampled
// disc fraction) and allaire_2d/sourcecode/allaire_diffuse_io.cpp:312-330,365-399
// (half-space and circular states).
#include "ic_device.h"

#include "ic_common.h"

#include <cstring>
#include <string>

namespace cfd {

namespace {

// A disc / bubble contributes only where a sample point of the cell can fall inside it: the
// 10x10 samples lie within half a cell of the centre, so a cell whose centre is more than
// R + one cell away along either axis has fraction 0 (the host skips the same cells through
// its bounding box; both only skip cells whose fraction is 0).
__device__ __forceinline__ bool outside(double x, double y, double cx, double cy, double R, double dx, double dy)
{
    return fabs(x - cx) > R + dx || fabs(y - cy) > R + dy;
}

__global__ __launch_bounds__(256) void ic_discs_kernel(double* buf, int pitch, long long plane, int nx, int ny,
                                                       double x0, double y0, double dx, double dy, Ambient4 amb,
                                                       const Disc* discs, int n)
{
    const long long cells = (long long)nx * ny;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells;
         c += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(c % nx), i = (int)(c / nx);
        const double x = cell_centre(x0, dx, j), y = cell_centre(y0, dy, i);
        double u[4] = {amb.U[0], amb.U[1], amb.U[2], amb.U[3]};
        for (int k = 0; k < n; ++k) {
            const Disc& d = discs[k];
            if (outside(x, y, d.cx, d.cy, d.R, dx, dy)) continue;
            const double f = circle_fraction(x, y, dx, dy, d.cx, d.cy, d.R, 10, true);
            if (f > 0.0)
                for (int q = 0; q < 4; ++q) u[q] = (1.0 - f) * u[q] + f * d.U[q];
        }
        for (int q = 0; q < 4; ++q) buf[q * plane + (long long)i * pitch + j] = u[q];
    }
}

__global__ __launch_bounds__(256) void ic_bubbles_kernel(double* buf, int pitch, long long plane, int nx, int ny,
                                                         double x0, double y0, double dx, double dy, Material m,
                                                         BubbleStates bs, const Bubble* bub, int n)
{
    const long long cells = (long long)nx * ny;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells;
         c += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(c % nx), i = (int)(c / nx);
        const double x = cell_centre(x0, dx, j), y = cell_centre(y0, dy, i);
        double w[6];
        for (int q = 0; q < 6; ++q) w[q] = bs.water[q];
        for (int k = 0; k < n; ++k) {
            const Bubble& b = bub[k];
            if (outside(x, y, b.cx, b.cy, b.R, dx, dy)) continue;
            const double f = circle_fraction(x, y, dx, dy, b.cx, b.cy, b.R, 10, false);
            if (f > 0.0)
                for (int q = 0; q < 6; ++q) w[q] = f * bs.air[q] + (1.0 - f) * w[q];
        }
        if (x <= bs.shock_x)
            for (int q = 0; q < 6; ++q) w[q] = bs.post[q];
        double u[6];
        p2c<ALLAIRE>(m, w, u);
        for (int q = 0; q < 6; ++q) buf[q * plane + (long long)i * pitch + j] = u[q];
    }
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Sum (mod 2^64) over every interior value of mix64(bits ^ (index * golden)), index = the
// value's position in the packed [component][y][x] order: independent of the pitch and of
// the order of summation, so a grid alone and the same grid in a batch compare directly.
__device__ __forceinline__ void checksum_body(const double* buf, int pitch, long long plane, int nx, int ny, int nc,
                                              unsigned long long* out)
{
    const long long per = (long long)nx * ny, total = per * nc;
    unsigned long long h = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long k = t / per, c = t % per;
        const long long y = c / nx, x = c % nx;
        const unsigned long long bits = (unsigned long long)__double_as_longlong(buf[k * plane + y * pitch + x]);
        h += mix64(bits ^ ((unsigned long long)t * 0x9e3779b97f4a7c15ULL));
    }
    __shared__ unsigned long long s[256];
    s[threadIdx.x] = h;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

__global__ __launch_bounds__(256) void checksum_kernel(const double* buf, int pitch, long long plane, int nx, int ny,
                                                       int nc, unsigned long long* out)
{
    checksum_body(buf, pitch, plane, nx, ny, nc, out);
}

// The same over bufs[*sel], chosen on the device when the kernel runs.
__global__ __launch_bounds__(256) void checksum_sel_kernel(const double* const* bufs, const int* sel, int pitch,
                                                           long long plane, int nx, int ny, int nc,
                                                           unsigned long long* out)
{
    checksum_body(bufs[*sel & 1], pitch, plane, nx, ny, nc, out);
}

int blocks_for(long long items)
{
    long long b = (items + 255) / 256;
    return (int)(b < 2048 ? (b < 1 ? 1 : b) : 2048);
}

}  // namespace

hipError_t launch_ic_discs(double* buf, int pitch, long long plane, const cfdart_desc& d, const double* ambient_U,
                           const Disc* d_discs, int n, hipStream_t st)
{
    Ambient4 amb;
    for (int q = 0; q < 4; ++q) amb.U[q] = ambient_U[q];
    hipLaunchKernelGGL(ic_discs_kernel, dim3(blocks_for((long long)d.nx * d.ny)), dim3(256), 0, st, buf, pitch,
                       plane, d.nx, d.ny, d.x0, d.y0, d.dx, d.dy, amb, d_discs, n);
    return hipGetLastError();
}

hipError_t launch_ic_bubbles(double* buf, int pitch, long long plane, const cfdart_desc& d,
                             const cfdart_bubble_params& p, const Bubble* d_bub, int n, hipStream_t st)
{
    BubbleStates bs;
    for (int q = 0; q < 6; ++q) {
        bs.water[q] = p.water[q];
        bs.air[q] = p.air[q];
        bs.post[q] = p.postshock[q];
    }
    bs.shock_x = p.shock_x;
    const Material m = make_material(d.gamma1, d.pinf1, d.gamma2, d.pinf2);
    hipLaunchKernelGGL(ic_bubbles_kernel, dim3(blocks_for((long long)d.nx * d.ny)), dim3(256), 0, st, buf, pitch,
                       plane, d.nx, d.ny, d.x0, d.y0, d.dx, d.dy, m, bs, d_bub, n);
    return hipGetLastError();
}

hipError_t launch_checksum(const double* buf, int pitch, long long plane, int nx, int ny, int nc,
                           unsigned long long* out, hipStream_t st)
{
    hipLaunchKernelGGL(checksum_kernel, dim3(blocks_for((long long)nx * ny * nc)), dim3(256), 0, st, buf, pitch,
                       plane, nx, ny, nc, out);
    return hipGetLastError();
}

hipError_t launch_checksum_sel(const double* const* bufs, const int* sel, int pitch, long long plane, int nx, int ny,
                               int nc, unsigned long long* out, hipStream_t st)
{
    hipLaunchKernelGGL(checksum_sel_kernel, dim3(blocks_for((long long)nx * ny * nc)), dim3(256), 0, st, bufs, sel,
                       pitch, plane, nx, ny, nc, out);
    return hipGetLastError();
}

int generate_state(const cfdart_desc& d, const cfdart_disc_params* dp, const cfdart_bubble_params* bp, double* buf,
                   int pitch, long long plane, void* d_draws, hipStream_t st, std::string& err, void* h_stage)
{
    if ((dp != nullptr) == (bp != nullptr)) {
        err = "give exactly one of the disc and bubble parameters";
        return CFDART_E_ARG;
    }
    if (dp && d.model != CFDART_EULER) {
        err = "random discs are an Euler initial condition";
        return CFDART_E_ARG;
    }
    if (bp && d.model != CFDART_ALLAIRE) {
        err = "shocked bubbles are an Allaire initial condition";
        return CFDART_E_ARG;
    }
    hipError_t e;
    if (dp) {
        Disc h[IC_MAX_DISCS];
        if (dp->n < 0 || dp->n > IC_MAX_DISCS || draw_discs(&d, dp, h, IC_MAX_DISCS) != 0) {
            err = "bad disc parameters";
            return CFDART_E_ARG;
        }
        double Wa[4] = {dp->ambient[0], dp->ambient[1], dp->ambient[2], dp->ambient[3]}, Ua[4];
        p2c<EULER>(make_material(d.gamma1, d.pinf1, d.gamma2, d.pinf2), Wa, Ua);
        const void* src = h;
        if (h_stage) src = std::memcpy(h_stage, h, sizeof(Disc) * (size_t)dp->n);
        e = hipMemcpyAsync(d_draws, src, sizeof(Disc) * (size_t)dp->n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_ic_discs(buf, pitch, plane, d, Ua, (const Disc*)d_draws, dp->n, st);
        if (e == hipSuccess && !h_stage) e = hipStreamSynchronize(st);  // h goes out of scope
    } else {
        Bubble h[IC_MAX_DISCS];
        if (bp->n < 0 || bp->n > IC_MAX_DISCS || draw_bubbles(bp, h, IC_MAX_DISCS) != 0) {
            err = "bad bubble parameters";
            return CFDART_E_ARG;
        }
        const void* src = h;
        if (h_stage) src = std::memcpy(h_stage, h, sizeof(Bubble) * (size_t)bp->n);
        e = hipMemcpyAsync(d_draws, src, sizeof(Bubble) * (size_t)bp->n, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_ic_bubbles(buf, pitch, plane, d, *bp, (const Bubble*)d_draws, bp->n, st);
        if (e == hipSuccess && !h_stage) e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) {
        err = std::string("initial-condition kernel: ") + hipGetErrorString(e);
        return CFDART_E_HIP;
    }
    return 0;
}

}  // namespace cfd
