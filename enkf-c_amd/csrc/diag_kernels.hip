// The ensemble spread of das_writespread (diag/diags.c:79-168) on the device.
//
// Per cell: v1 = sum_e v_e and v2 = sum_e (v_e * v_e) in double, member order,
// the square a float product; mean = v1 / nmem; spread = sqrt(max(0,
// v2 / nmem - mean^2)), 0 where |mean| > MAXOBSVAL (common/global.h:119);
// stored as float.  The same operations in the same order as the reference,
// so the spread matches it bit for bit.
//
// HBM-bound: 4 m bytes read and 4 bytes written per cell.
#include <hip/hip_runtime.h>

#include "enkfg_internal.h"

namespace enkfg {

namespace {

constexpr double kMaxObsVal = 999999.0;   // MAXOBSVAL, common/global.h:119
constexpr int kCells = 64;                // cells per workgroup (one wave)

__device__ inline float finish_spread(double v1, double v2, int m)
{
    v1 /= (double)m;
    v2 = v2 / (double)m - v1 * v1;
    v2 = (v2 < 0.0) ? 0.0 : sqrt(v2);
    if (fabs(v1) > kMaxObsVal)
        v2 = 0.0;
    return (float)v2;
}

// E [field][cell][m], members contiguous.  One wave stages 64 cells in LDS
// with coalesced loads (float4 when VEC), rows padded to m + 1 words so that
// the lanes' reads of one member fall in distinct banks; each lane then sums
// its own cell in member order.
template <bool VEC>
__global__ __launch_bounds__(kCells) void spread_cells_kernel(const float* __restrict__ E, long long ncell, int m,
                                                              float* __restrict__ out)
{
    extern __shared__ float tile[];
    const int f = blockIdx.y;
    const long long c0 = (long long)blockIdx.x * kCells;
    const int nc = (int)((ncell - c0) < kCells ? (ncell - c0) : kCells);
    const float* src = E + ((long long)f * ncell + c0) * m;
    const int ld = m + 1;
    const int n = nc * m;
    if (VEC) {
        const float4* src4 = reinterpret_cast<const float4*>(src);
        for (int k4 = threadIdx.x; k4 < n / 4; k4 += kCells) {
            const float4 x = src4[k4];
            const int k = 4 * k4;
            const int c = k / m;
            float* d = tile + c * ld + (k - c * m);   // m % 4 == 0: the 4 values share a cell
            d[0] = x.x;
            d[1] = x.y;
            d[2] = x.z;
            d[3] = x.w;
        }
    } else {
        for (int k = threadIdx.x; k < n; k += kCells) {
            const int c = k / m;
            tile[c * ld + (k - c * m)] = src[k];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x >= nc)
        return;
    const float* row = tile + threadIdx.x * ld;
    double v1 = 0.0, v2 = 0.0;
    for (int e = 0; e < m; ++e) {
        const float x = row[e];
        const float x2 = x * x;
        v1 += (double)x;
        v2 += (double)x2;
    }
    out[(long long)f * ncell + c0 + threadIdx.x] = finish_spread(v1, v2, m);
}

// E [field][m][cell] (the reference's field buffers, members outermost): one
// lane per cell, coalesced across lanes for every member.
__global__ __launch_bounds__(256) void spread_members_kernel(const float* __restrict__ E, long long ncell, int m,
                                                             float* __restrict__ out)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell)
        return;
    const int f = blockIdx.y;
    const float* src = E + (long long)f * m * ncell + c;
    double v1 = 0.0, v2 = 0.0;
    for (int e = 0; e < m; ++e) {
        const float x = src[(long long)e * ncell];
        const float x2 = x * x;
        v1 += (double)x;
        v2 += (double)x2;
    }
    out[(long long)f * ncell + c] = finish_spread(v1, v2, m);
}

}  // namespace

hipError_t launch_spread(hipStream_t s, const float* E, int nfield, long long ncell, int m, bool cells_major,
                         float* out)
{
    if (nfield <= 0 || ncell <= 0)
        return hipSuccess;
    if (m <= 0 || nfield > 65535)
        return hipErrorInvalidValue;
    if (cells_major) {
        const long long nb = (ncell + kCells - 1) / kCells;
        if (nb > 0x7fffffffLL)
            return hipErrorInvalidValue;
        const size_t lds = (size_t)kCells * (m + 1) * sizeof(float);
        if (lds > 64 * 1024)
            return hipErrorInvalidValue;
        const bool vec = (m % 4 == 0) && (reinterpret_cast<uintptr_t>(E) % 16 == 0);
        if (vec)
            hipLaunchKernelGGL(spread_cells_kernel<true>, dim3((unsigned)nb, nfield), dim3(kCells), lds, s, E, ncell,
                               m, out);
        else
            hipLaunchKernelGGL(spread_cells_kernel<false>, dim3((unsigned)nb, nfield), dim3(kCells), lds, s, E, ncell,
                               m, out);
    } else {
        const long long nb = (ncell + 255) / 256;
        if (nb > 0x7fffffffLL)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(spread_members_kernel, dim3((unsigned)nb, nfield), dim3(256), 0, s, E, ncell, m, out);
    }
    return hipGetLastError();
}

}  // namespace enkfg
