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
//
// The adjoint of the analysis ensemble (DEnKF, calc/calcs.c:232-248 with the
// relaxation of transforms.c:1270-1283) adds, given dJ/dX^a for the dynamic
// members of every field and cell:
//   per field and cell (update.c:381-458): v_a = A^f T_c^T (/ k_d when
//   hybrid), X^a = infl v_a + x^a with the capped inflation
//   infl = min(infl0, sqrt(v2f / v2a) rho + 1 - rho) applied iff infl - 1 > EPSF;
//     dJ/dx^a   = sum_e dJ/dX^a[e]                       (-> w, as above)
//     dJ/dinfl  = sum_e dJ/dX^a[e] v_a[e]                 (-> infl0 when capped
//                                                          or plain)
//     dJ/dv_a   = infl dJ/dX^a + dJ/dinfl d infl/d v2a 2 v_a   (uncapped ratio)
//     dJ/dT_c   = (dJ/dv_a / k_d) (A^f)^T,  summed over fields
//   per node: D_n = sum_c beta(c, n) dJ/dT_c; with T = alpha / 2 M^-1 + const,
//     dJ/dM = -B, B = alpha / 2 M^-1 D_n M^-1, and since M = I + S^T S with
//     S_o ~ R_o^-1/2:  dJ/dR_o += S_o B S_o^T / R_o  (S_o the tapered row).
//   The inflation's dependence on the forecast spread involves no
//   observation.
#include "enkfg_internal.h"
#include "update_interp.h"

namespace enkfg {

namespace {

constexpr float kMaxVal = 999999.0f;   // MAXOBSVAL, common/global.h:119
constexpr float kEps = 1.0e-6f;        // EPSF, update/update.c:41

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

// one workgroup per listed node of the batch: the node's local obs in
// parallel; with B (the batch's [bnode][MP][MP]), the term of T as well
template <int MT>
__global__ void __launch_bounds__(256) adj_obs_kernel(CalcArgs a, const int* __restrict__ perm,
                                                      const short* __restrict__ type, const float* __restrict__ estd,
                                                      const double* __restrict__ sqrt_rf,
                                                      const double* __restrict__ B, double* __restrict__ gy,
                                                      double* __restrict__ gR)
{
    constexpr int MP = 16 * MT;
    __shared__ double zs[MP], ws[MP];
    const int k = blockIdx.x;
    if (k >= a.nbatch)
        return;
    const size_t kk = (size_t)(a.node0 + k);
    const int node = a.node_list[kk];
    const int m = a.m;
    for (int e = threadIdx.x; e < MP; e += 256) {
        zs[e] = (e < m) ? a.adj_z[kk * m + e] : 0.0;
        ws[e] = (e < m) ? a.adj_w[kk * m + e] : 0.0;
    }
    __syncthreads();
    const double* Bk = B ? B + (size_t)k * MP * MP : nullptr;
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
        double qf = 0.0;   // S_o B S_o^T, in blocks of 16 members
        if (Bk != nullptr) {
#pragma unroll 1
            for (int ib = 0; ib < MT; ++ib) {
                double si[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    si[u] = (16 * ib + u < m) ? (double)Sr[16 * ib + u] * lc : 0.0;
#pragma unroll 1
                for (int jb = 0; jb < MT; ++jb) {
                    double sj[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        sj[u] = (16 * jb + u < m) ? (double)Sr[16 * jb + u] * lc : 0.0;
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        const double* br = Bk + (size_t)(16 * ib + u) * MP + 16 * jb;
                        double t = 0.0;
#pragma unroll
                        for (int v = 0; v < 16; ++v)
                            t = fma(br[v], sj[v], t);
                        qf = fma(si[u], t, qf);
                    }
                }
            }
        }
        const double r = a.sv[row] * lc - aw;
        const int o = perm[row];
        const double sd = (double)estd[o];
        const double d = sd * sqrt_rf[type[o]] * sqm;
        atomicAdd(&gy[row], az * lc / d);
        atomicAdd(&gR[row], (-az * r + qf) / (sd * sd));
    }
}

__global__ void nonzero_rows_kernel(const double* __restrict__ g, int nrow, int width, int* __restrict__ flags,
                                    int accumulate)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow)
        return;
    int nz = 0;
    for (int e = 0; e < width; ++e)
        nz |= g[(size_t)r * width + e] != 0.0;
    flags[r] = accumulate ? (flags[r] | nz) : nz;
}

// sum over the 256 threads of a workgroup (fixed order: waves, then the
// four wave totals in order); every thread receives it
__device__ inline double block_sum(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double s = ((red[0] + red[1]) + red[2]) + red[3];
    __syncthreads();
    return s;
}

// one workgroup per cell: the forward of update.c:381-458 recomputed for
// every field with a sensitivity, dJ/dT_c accumulated per thread (entries
// q = tid + 256 k of the padded MP x MP matrix), then scattered with the
// interpolation weights to the four nodes
template <int MT>
__global__ void __launch_bounds__(256) adj_ens_cell_kernel(UpdArgs a, const float* __restrict__ dJ,
                                                           double* __restrict__ g, double* __restrict__ D,
                                                           double* __restrict__ ginfl)
{
    constexpr int MP = 16 * MT;
    constexpr int NE = MP * MP / 256;
    __shared__ float tc[MP][MP + 1];   // T_c[e1][e2], zero outside md x m
    __shared__ float av[MP];           // A^f of the field, zero past m
    __shared__ double dva[MP];         // dJ/d(v_a before the hybrid rescaling)
    __shared__ double red[4];
    const long long cell = blockIdx.x;
    if (cell >= a.ncell)
        return;
    const int tid = threadIdx.x;
    const int m = a.m, md = a.md;
    bool any = false;
    for (int q = tid; q < a.nfield * md; q += 256) {
        const int f = q / md, e = q - f * md;
        any |= dJ[((long long)f * a.ncell + cell) * m + e] != 0.0f;
    }
    if (!__syncthreads_or(any))
        return;
    const int jl = (int)(cell / a.mni);
    const int i = (int)(cell % a.mni);
    const int j = a.j1 + jl;
    const int nl = a.nlevels[(long long)j * a.mni + i];
    const bool hybrid = (a.mode == ENKFG_MODE_HYBRID);
    const double k_d = (md > 1) ? sqrt((double)(m - 1) / (double)(md - 1)) : 1.0;
    // T_c as the update forms it (update.c:195-272)
    {
        const Bilin b = make_bilin(a, j, i, (long long)md * m);
        for (int q = tid; q < MP * MP; q += 256) {
            const int e1 = q / MP, e2 = q - e1 * MP;
            float t = 0.0f;
            if (e1 < md && e2 < m) {
                const long long o = (long long)e1 * m + e2;
                t = bilin(b, a.T[b.o00 + o], a.T[b.o10 + o], a.T[b.o01 + o], a.T[b.o11 + o]);
            }
            tc[e1][e2] = t;
        }
    }
    double dc[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k)
        dc[k] = 0.0;
    double gc = 0.0;
    bool touched = false;
    __syncthreads();
    for (int f = 0; f < a.nfield; ++f) {
        const float* G = dJ + ((long long)f * a.ncell + cell) * m;
        const float gv = (tid < md) ? G[tid] : 0.0f;
        if (!__syncthreads_or(gv != 0.0f))
            continue;
        // a masked field, or one with a missing dynamic member, is written as
        // zeros (update.c:298-371): no dependence on the transforms
        const bool msk = (a.surfk == 0) ? nl <= a.flevel[f] : nl <= (a.fsurf[f] ? 0 : a.surfk - a.flevel[f]);
        if (msk)
            continue;
        const float* rec = a.E + ((long long)f * a.ncell + cell) * m;
        const float xv = (tid < m) ? rec[tid] : 0.0f;
        const bool bad = tid < m && !(fabsf(xv) <= kMaxVal);
        if (__syncthreads_or(bad && tid < md))
            continue;
        const bool bads = __syncthreads_or(bad);
        // FP64 mean of the dynamic members, in member order
        if (tid < MP)
            av[tid] = (tid < m) ? xv : 0.0f;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int e = 0; e < md; ++e)
                s += (double)av[e];
            red[0] = s / (double)md;
        }
        __syncthreads();
        const double mean = red[0];
        __syncthreads();
        if (tid < MP) {
            const float x = (bads && tid >= md) ? 0.0f : xv;
            av[tid] = (tid < m) ? (float)((double)x - mean) : 0.0f;
        }
        __syncthreads();
        // v_a (the real-valued product; the update's rounding is not differentiated)
        double va = 0.0;
        if (tid < md) {
            for (int e2 = 0; e2 < m; ++e2)
                va = fma((double)av[e2], (double)tc[tid][e2], va);
            if (hybrid)
                va /= k_d;
        }
        double vh = 0.0;
        if (tid < md) {
            const float v = hybrid ? (float)((double)av[tid] / k_d) : av[tid];
            vh = (double)(v * v);
        }
        const double v2f = block_sum(vh, red);
        const double v2a = block_sum(tid < md ? va * va : 0.0, red);
        const double gbar = block_sum((double)gv, red);
        const double dinfl = block_sum((double)gv * va, red);
        const float infl0 = a.finfl[f];
        const double ratio = a.finfratio[f];
        double infl = infl0, dinfl_dv2a = 0.0;
        bool to_infl0 = true;   // infl = infl0 (plain, or capped)
        if (!isnan(ratio) && v2a > 0.0) {
            const double raw = sqrt(v2f / v2a) * ratio + 1.0 - ratio;
            if ((float)raw < infl0) {
                infl = raw;
                to_infl0 = false;
                dinfl_dv2a = -0.5 * ratio * sqrt(v2f) / (v2a * sqrt(v2a));
            }
        }
        if (!((float)infl - 1.0f > kEps)) {   // not applied: a constant factor 1
            infl = 1.0;
            dinfl_dv2a = 0.0;
            to_infl0 = false;
        }
        if (tid == 0 && to_infl0)
            atomicAdd(&ginfl[f], dinfl);
        if (tid < MP) {
            double d = 0.0;
            if (tid < md) {
                d = infl * (double)gv + dinfl * dinfl_dv2a * 2.0 * va;
                if (hybrid)
                    d /= k_d;
            }
            dva[tid] = d;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int q = tid + 256 * k;
            const int e1 = q / MP, e2 = q - e1 * MP;
            dc[k] = fma(dva[e1], (double)av[e2], dc[k]);
        }
        if (tid < MP)
            gc = fma(gbar, (double)av[tid], gc);
        touched = true;
        __syncthreads();
    }
    if (!touched)
        return;
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
    for (int nq = 0; nq < 4; ++nq) {
        if (beta[nq] == 0.0)
            continue;
        double* Dn = D + (size_t)nodes[nq] * md * m;
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int q = tid + 256 * k;
            const int e1 = q / MP, e2 = q - e1 * MP;
            if (e1 < md && e2 < m && dc[k] != 0.0)
                atomicAdd(&Dn[(size_t)e1 * m + e2], beta[nq] * dc[k]);
        }
        if (tid < m && gc != 0.0)
            atomicAdd(&g[(size_t)nodes[nq] * m + tid], beta[nq] * gc);
    }
}

// B = alpha / 2 M^-1 D_pad M^-1 for one batch node per workgroup; Y = D_pad
// M^-1 in the scratch `y` ([bnode][MP][MP])
template <int MT>
__global__ void __launch_bounds__(256) adj_B_kernel(CalcArgs a, const double* __restrict__ D, double* __restrict__ y,
                                                    double* __restrict__ B)
{
    constexpr int MP = 16 * MT;
    const int k = blockIdx.x;
    if (k >= a.nbatch)
        return;
    const int node = a.node_list[a.node0 + k];
    const int m = a.m, md = a.md;
    const double* mi = a.adj_minv + (size_t)k * MP * MP;
    const double* Dn = D + (size_t)node * md * m;
    double* Y = y + (size_t)k * MP * MP;
    double* Bk = B + (size_t)k * MP * MP;
    for (int q = threadIdx.x; q < MP * MP; q += 256) {
        const int r = q / MP, c = q - r * MP;
        double s = 0.0;
        if (r < md && c < m)
            for (int b = 0; b < m; ++b)
                s = fma(Dn[(size_t)r * m + b], mi[(size_t)b * MP + c], s);
        Y[q] = s;
    }
    __syncthreads();
    const double h = 0.5 * a.alpha;
    for (int q = threadIdx.x; q < MP * MP; q += 256) {
        const int r = q / MP, c = q - r * MP;
        double s = 0.0;
        if (r < m && c < m)
            for (int e = 0; e < md; ++e)
                s = fma(mi[(size_t)r * MP + e], Y[(size_t)e * MP + c], s);
        Bk[q] = h * s;
    }
}

}  // namespace

hipError_t launch_nonzero_rows(hipStream_t s, const double* g, int nrow, int width, int* flags, bool accumulate)
{
    if (nrow == 0)
        return hipSuccess;
    hipLaunchKernelGGL(nonzero_rows_kernel, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, s, g, nrow, width,
                       flags, accumulate ? 1 : 0);
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

hipError_t launch_adj_ens_cells(hipStream_t s, const UpdArgs& a, const float* dJ_dXa, double* g_nodes,
                                double* D_nodes, double* g_infl)
{
    if (a.ncell == 0 || a.nfield == 0)
        return hipSuccess;
    if (a.ncell > 0x7fffffffLL)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)a.ncell);
    switch ((a.m + 15) / 16) {
    case 1: hipLaunchKernelGGL(adj_ens_cell_kernel<1>, grid, dim3(256), 0, s, a, dJ_dXa, g_nodes, D_nodes, g_infl); break;
    case 2: hipLaunchKernelGGL(adj_ens_cell_kernel<2>, grid, dim3(256), 0, s, a, dJ_dXa, g_nodes, D_nodes, g_infl); break;
    case 3: hipLaunchKernelGGL(adj_ens_cell_kernel<3>, grid, dim3(256), 0, s, a, dJ_dXa, g_nodes, D_nodes, g_infl); break;
    case 4: hipLaunchKernelGGL(adj_ens_cell_kernel<4>, grid, dim3(256), 0, s, a, dJ_dXa, g_nodes, D_nodes, g_infl); break;
    case 5: hipLaunchKernelGGL(adj_ens_cell_kernel<5>, grid, dim3(256), 0, s, a, dJ_dXa, g_nodes, D_nodes, g_infl); break;
    case 6: hipLaunchKernelGGL(adj_ens_cell_kernel<6>, grid, dim3(256), 0, s, a, dJ_dXa, g_nodes, D_nodes, g_infl); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adj_B(hipStream_t s, const CalcArgs& a, const double* D_nodes, double* B)
{
    if (a.nbatch == 0)
        return hipSuccess;
    const dim3 grid((unsigned)a.nbatch);
    switch ((a.m + 15) / 16) {
    case 1: hipLaunchKernelGGL(adj_B_kernel<1>, grid, dim3(256), 0, s, a, D_nodes, a.gram, B); break;
    case 2: hipLaunchKernelGGL(adj_B_kernel<2>, grid, dim3(256), 0, s, a, D_nodes, a.gram, B); break;
    case 3: hipLaunchKernelGGL(adj_B_kernel<3>, grid, dim3(256), 0, s, a, D_nodes, a.gram, B); break;
    case 4: hipLaunchKernelGGL(adj_B_kernel<4>, grid, dim3(256), 0, s, a, D_nodes, a.gram, B); break;
    case 5: hipLaunchKernelGGL(adj_B_kernel<5>, grid, dim3(256), 0, s, a, D_nodes, a.gram, B); break;
    case 6: hipLaunchKernelGGL(adj_B_kernel<6>, grid, dim3(256), 0, s, a, D_nodes, a.gram, B); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adj_obs(hipStream_t s, const CalcArgs& a, const int* perm, const short* type, const float* estd,
                          const double* sqrt_rfactor, const double* B, double* dJ_dy, double* dJ_dR)
{
    if (a.nbatch == 0)
        return hipSuccess;
    const dim3 grid((unsigned)a.nbatch);
    switch ((a.m + 15) / 16) {
    case 1: hipLaunchKernelGGL(adj_obs_kernel<1>, grid, dim3(256), 0, s, a, perm, type, estd, sqrt_rfactor, B, dJ_dy, dJ_dR); break;
    case 2: hipLaunchKernelGGL(adj_obs_kernel<2>, grid, dim3(256), 0, s, a, perm, type, estd, sqrt_rfactor, B, dJ_dy, dJ_dR); break;
    case 3: hipLaunchKernelGGL(adj_obs_kernel<3>, grid, dim3(256), 0, s, a, perm, type, estd, sqrt_rfactor, B, dJ_dy, dJ_dR); break;
    case 4: hipLaunchKernelGGL(adj_obs_kernel<4>, grid, dim3(256), 0, s, a, perm, type, estd, sqrt_rfactor, B, dJ_dy, dJ_dR); break;
    case 5: hipLaunchKernelGGL(adj_obs_kernel<5>, grid, dim3(256), 0, s, a, perm, type, estd, sqrt_rfactor, B, dJ_dy, dJ_dR); break;
    case 6: hipLaunchKernelGGL(adj_obs_kernel<6>, grid, dim3(256), 0, s, a, perm, type, estd, sqrt_rfactor, B, dJ_dy, dJ_dR); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace enkfg
