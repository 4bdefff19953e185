// adjoint.hip -- the adjoint of the analysis mean with respect to the
// observation values and the observation-error variances.
//
// Forward (reference): per stride node n, w_n = M_n^-1 S_n^T s_n with
// M_n = I + S_n^T S_n, S_n and s_n the tapered, standardised local anomalies
// and innovations (calc/calcs.c:189-263, calc/ensobs.c:849-898); per model
// cell c, w_c is the bilinear interpolation of the four surrounding node
// vectors (update/update.c:195-272) and the analysis mean of field f is
// x^a = x^f + A^f w_c (update.c:381-410).  Inflation scales the analysis
// anomalies only (update.c:432-458), so the mean does not depend on it.
//
// Given dJ/dx^a for every field and cell, the adjoint is
//   g_c   = sum_f dJ/dx^a[f][c] A^f[f][c]                     (m-vector)
//   g_n   = sum_c beta(c, n) g_c       (beta: the interpolation weights)
//   z_n   = M_n^-1 g_n                 (recomputed M, never stored)
// and for every local observation o of node n, with a = S_o z_n,
// r = s_o - S_o w_n (S_o, s_o tapered and standardised, d_o = estd_o
// sqrt(rfactor) sqrt(m - 1) the standardisation divisor):
//   dJ/dy_o += a lc_o / d_o
//   dJ/dR_o += -a r / R_o              (R_o = estd_o^2; S_o, s_o ~ R_o^-1/2)
// The FP32 rounding of S in the standardisation is treated as the identity.
#include "enkfg_internal.h"

namespace enkfg {

namespace {

constexpr float kMaxVal = 999999.0f;   // MAXOBSVAL, common/global.h:114

// one wave per cell: g_c, then scattered to the four interpolation nodes
__global__ void __launch_bounds__(64) adj_cell_kernel(UpdArgs a, const float* __restrict__ dJ, double* __restrict__ g)
{
    const long long cell = blockIdx.x;
    if (cell >= a.ncell)
        return;
    const int lane = threadIdx.x;
    const int m = a.m, md = a.md;
    const int jl = (int)(cell / a.mni);
    const int i = (int)(cell % a.mni);
    const int j = a.j1 + jl;
    const int nl = a.nlevels[(long long)j * a.mni + i];
    // any sensitivity at this cell?
    bool any = false;
    for (int f = lane; f < a.nfield; f += 64)
        any |= dJ[(long long)f * a.ncell + cell] != 0.0f;
    if (!__any(any))
        return;
    constexpr int ME = 3;   // members per lane (m <= 192)
    double gc[ME] = {0.0, 0.0, 0.0};
    for (int f = 0; f < a.nfield; ++f) {
        const float dj = dJ[(long long)f * a.ncell + cell];
        if (dj == 0.0f)
            continue;
        // masked cells and cells with a missing dynamic member are set to
        // zero by the update (update.c:298-371): no dependence on w
        const bool msk = (a.surfk == 0) ? nl <= a.flevel[f] : nl <= (a.fsurf[f] ? 0 : a.surfk - a.flevel[f]);
        if (msk)
            continue;
        const float* rec = a.E + ((long long)f * a.ncell + cell) * m;
        float v[ME];
        int eb = m;
#pragma unroll
        for (int u = 0; u < ME; ++u) {
            const int e = lane + 64 * u;
            v[u] = (e < m) ? rec[e] : 0.0f;
            if (e < m && (!isfinite(v[u]) || fabsf(v[u]) > kMaxVal))
                eb = min(eb, e);
        }
        for (int o = 1; o < 64; o <<= 1)
            eb = min(eb, __shfl_xor(eb, o));
        if (eb < md)
            continue;
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < ME; ++u) {
            const int e = lane + 64 * u;
            if (e >= md || e >= m)
                v[u] = (e < m && eb < m) ? 0.0f : v[u];   // static members zeroed (update.c:360-365)
            if (e < md)
                s += (double)v[u];
        }
        for (int o = 1; o < 64; o <<= 1)
            s += __shfl_xor(s, o);
        const double mean = s / (double)md;
#pragma unroll
        for (int u = 0; u < ME; ++u) {
            const int e = lane + 64 * u;
            if (e < m)
                gc[u] += (double)dj * (double)(float)((double)v[u] - mean);
        }
    }
    // interpolation weights of update.c:195-272 as real numbers
    const int jj = j / a.stride, stepj = j % a.stride;
    const int ii = i / a.stride, stepi = i % a.stride;
    const int jj2 = (jj < a.nj_s - 1) ? jj + 1 : jj;
    const int ii2 = (ii < a.ni_s - 1) ? ii + 1 : (a.periodic_i ? (ii + 1) % a.ni_s : ii);
    const bool jcopy = (a.stride == 1 || stepj == 0);
    const bool icopy = (a.stride == 1 || stepi == 0);
    const float wj2 = (float)stepj / (float)a.stride, wj1 = 1.0f - wj2;
    const float wi2 = (float)stepi / (float)a.stride, wi1 = 1.0f - wi2;
    const double bj1 = jcopy ? 1.0 : wj1, bj2 = jcopy ? 0.0 : wj2;
    const double bi1 = icopy ? 1.0 : wi1, bi2 = icopy ? 0.0 : wi2;
    const int nodes[4] = {jj * a.ni_s + ii, jj2 * a.ni_s + ii, jj * a.ni_s + ii2, jj2 * a.ni_s + ii2};
    const double beta[4] = {bj1 * bi1, bj2 * bi1, bj1 * bi2, bj2 * bi2};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (beta[q] == 0.0)
            continue;
#pragma unroll
        for (int u = 0; u < ME; ++u) {
            const int e = lane + 64 * u;
            if (e < m)
                atomicAdd(&g[(size_t)nodes[q] * m + e], beta[q] * gc[u]);
        }
    }
}

// one workgroup per listed node: the node's local obs in parallel
__global__ void __launch_bounds__(256) adj_obs_kernel(CalcArgs a, int nlist, const int* __restrict__ perm,
                                                      const short* __restrict__ type, const float* __restrict__ estd,
                                                      const double* __restrict__ sqrt_rf, double* __restrict__ gy,
                                                      double* __restrict__ gR)
{
    __shared__ double zs[160], ws[160];
    const int k = blockIdx.x;
    if (k >= nlist)
        return;
    const int node = a.node_list[k];
    const int m = a.m;
    for (int e = threadIdx.x; e < m; e += 256) {
        zs[e] = a.adj_z[(size_t)k * m + e];
        ws[e] = a.adj_w[(size_t)k * m + e];
    }
    __syncthreads();
    const long long off = a.offsets[(size_t)node * a.ntypes];
    const long long end = a.offsets[(size_t)(node + 1) * a.ntypes];
    const double sqm = sqrt((double)(m - 1));
    for (long long q = off + threadIdx.x; q < end; q += 256) {
        const int row = a.ids[q];
        const double lc = a.lc[q];
        const float* Sr = a.S + (long long)row * a.ms;
        double az = 0.0, aw = 0.0;
        for (int e = 0; e < m; ++e) {
            const double sv = (double)Sr[e] * lc;
            az = fma(sv, zs[e], az);
            aw = fma(sv, ws[e], aw);
        }
        const double r = a.sv[row] * lc - aw;
        const int o = perm[row];
        const double sd = (double)estd[o];
        const double d = sd * sqrt_rf[type[o]] * sqm;
        atomicAdd(&gy[row], az * lc / d);
        atomicAdd(&gR[row], -az * r / (sd * sd));
    }
}

__global__ void nonzero_rows_kernel(const double* __restrict__ g, int nrow, int width, int* __restrict__ flags)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow)
        return;
    int nz = 0;
    for (int e = 0; e < width; ++e)
        nz |= g[(size_t)r * width + e] != 0.0;
    flags[r] = nz;
}

}  // namespace

hipError_t launch_nonzero_rows(hipStream_t s, const double* g, int nrow, int width, int* flags)
{
    if (nrow == 0)
        return hipSuccess;
    hipLaunchKernelGGL(nonzero_rows_kernel, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, s, g, nrow, width,
                       flags);
    return hipGetLastError();
}

hipError_t launch_adj_cells(hipStream_t s, const UpdArgs& a, const float* dJ_dmean, double* g_nodes)
{
    if (a.ncell == 0 || a.nfield == 0)
        return hipSuccess;
    if (a.m > 192 || a.ncell > 0x7fffffffLL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_cell_kernel, dim3((unsigned)a.ncell), dim3(64), 0, s, a, dJ_dmean, g_nodes);
    return hipGetLastError();
}

hipError_t launch_adj_obs(hipStream_t s, const CalcArgs& a, int nlist, const int* perm, const short* type,
                          const float* estd, const double* sqrt_rfactor, double* dJ_dy, double* dJ_dR)
{
    if (nlist == 0)
        return hipSuccess;
    if (a.m > 160)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_obs_kernel, dim3((unsigned)nlist), dim3(256), 0, s, a, nlist, perm, type, estd,
                       sqrt_rfactor, dJ_dy, dJ_dR);
    return hipGetLastError();
}

}  // namespace enkfg
