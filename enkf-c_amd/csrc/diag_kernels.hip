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

// The same for m % 4 == 0, as a loop of each wave over chunks of 64 cells
// (chunk c of field c / nchunk): the float4 records of the next chunk are
// loaded into registers while the lanes sum the current one from LDS, so a
// wave keeps its loads in flight through its sums (the LDS per wave allows
// only about six waves per CU).  NQ float4 per lane hold a chunk
// (64 m / 256 <= NQ).
template <int NQ>
__global__ __launch_bounds__(kCells) void spread_cells_pipe_kernel(const float* __restrict__ E, long long ncell,
                                                                   int m, int nfield, float* __restrict__ out)
{
    extern __shared__ float tile[];
    const long long nchunk = (ncell + kCells - 1) / kCells;
    const long long total = nchunk * nfield;
    const int ld = m + 1;
    const int lane = threadIdx.x;
    float4 q[NQ];
    long long base = 0;   // first record of the chunk in q
    int nc = 0;           // its cells
    auto load = [&](long long c) {
        const long long f = c / nchunk;
        const long long c0 = (c - f * nchunk) * kCells;
        nc = (int)((ncell - c0) < kCells ? (ncell - c0) : kCells);
        base = f * ncell + c0;
        const float4* src4 = reinterpret_cast<const float4*>(E + base * m);
        const int n4 = nc * m / 4;
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int k4 = lane + kCells * j;
            q[j] = (k4 < n4) ? src4[k4] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    long long c = blockIdx.x;
    if (c >= total)
        return;
    load(c);
    for (;;) {
        // the chunk in q to LDS (m % 4 == 0: the four values share a cell)
        const int n4 = nc * m / 4;
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const int k4 = lane + kCells * j;
            if (k4 < n4) {
                const int k = 4 * k4, r = k / m;
                float* d = tile + r * ld + (k - r * m);
                d[0] = q[j].x;
                d[1] = q[j].y;
                d[2] = q[j].z;
                d[3] = q[j].w;
            }
        }
        const long long cur = base;
        const int ncur = nc;
        __syncthreads();
        const long long next = c + gridDim.x;
        if (next < total)
            load(next);   // in flight during the sums
        if (lane < ncur) {
            const float* row = tile + lane * ld;
            double v1 = 0.0, v2 = 0.0;
            for (int e = 0; e < m; ++e) {
                const float x = row[e];
                const float x2 = x * x;
                v1 += (double)x;
                v2 += (double)x2;
            }
            out[cur + lane] = finish_spread(v1, v2, m);
        }
        __syncthreads();   // the sums have read the tile before it is refilled
        if (next >= total)
            break;
        c = next;
    }
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
        // the looping kernel: enough waves to fill the CUs at the LDS limit
        const long long total = nb * nfield;
        const unsigned pgrid = (unsigned)(total < 4096 ? total : 4096);
        if (vec && m <= 96)
            hipLaunchKernelGGL(spread_cells_pipe_kernel<24>, dim3(pgrid), dim3(kCells), lds, s, E, ncell, m, nfield,
                               out);
        else if (vec && m <= 160)
            hipLaunchKernelGGL(spread_cells_pipe_kernel<40>, dim3(pgrid), dim3(kCells), lds, s, E, ncell, m, nfield,
                               out);
        else if (vec)
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
