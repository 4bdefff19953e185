// update_kernels.hip -- das_updatefields (update/update.c:92-564),
// das_updatebg (update.c:568-814) and das_sethybridensemble
// (common/dasystem.c:775-819) on a device-resident ensemble.
//
// Layout: E[field][cell][member], cell = j * mni + i (members contiguous, so
// one cell of one field is a 4*m-byte contiguous record).  One block of
// 64*ceil(m/64) threads handles one grid cell for all fields: thread e owns
// member e.  The transform for the cell is interpolated from the stride grid
// exactly as update.c:195-272 does (FP32, j first then i, no contraction),
// once per cell, and kept in registers (thread e1 holds row e1 of T); it is
// then applied to every field of the column.  The per-cell arithmetic follows
// update.c:282-527: FP64 ensemble mean and increment sums, FP32 anomalies and
// transform product, capped inflation, masks and missing values.
#include "enkfg_internal.h"

namespace enkfg {

namespace {

constexpr float kMaxObsVal = 999999.0f;   // MAXOBSVAL, common/global.h:114
constexpr float kEpsF = 1.0e-6f;          // EPSF, update/update.c:41

// cell index for block b: each of the 8 XCDs (blocks are dispatched to XCDs
// round-robin) gets a contiguous range of cells, so neighbouring cells, which
// share stride-grid transforms, hit the same L2.
__device__ inline long long xcd_cell(long long ncell)
{
    const long long chunk = (ncell + 7) / 8;
    return (long long)(blockIdx.x % 8) * chunk + blockIdx.x / 8;
}

template <int NW>
struct RedShared {
    double part[NW][2];
};

// deterministic block sum of two doubles (each wave reduces, then waves in order)
template <int NW>
__device__ inline void block_sum2(double& a, double& b, RedShared<NW>& sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sh.part[w][0] = a;
        sh.part[w][1] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        a += sh.part[k][0];
        b += sh.part[k][1];
    }
}

// bilinear stride-grid interpolation of one length-n row r of the node data
// (rows of `stride_elems` floats per node) into out[0..n), update.c:195-272
struct Interp {
    int jj, jj2, ii, ii2;
    bool jcopy, icopy;
    float wj1, wj2, wi1, wi2;
};

__device__ inline Interp make_interp(int j, int i, int stride, int nj_s, int ni_s, int periodic_i)
{
    Interp p;
    p.jj = j / stride;
    const int stepj = j % stride;
    p.ii = i / stride;
    const int stepi = i % stride;
    p.jj2 = (p.jj < nj_s - 1) ? p.jj + 1 : p.jj;
    p.ii2 = (p.ii < ni_s - 1) ? p.ii + 1 : (periodic_i ? (p.ii + 1) % ni_s : p.ii);
    p.jcopy = (stride == 1 || stepj == 0);
    p.icopy = (stride == 1 || stepi == 0);
    p.wj2 = (float)stepj / (float)stride;
    p.wj1 = 1.0f - p.wj2;
    p.wi2 = (float)stepi / (float)stride;
    p.wi1 = 1.0f - p.wi2;
    return p;
}

// value k of the interpolated record from node records base[(jj*ni_s+ii)*rec]
__device__ inline float interp_at(const float* __restrict__ base, long long rec, int ni_s, const Interp& p, int k)
{
    const float a00 = base[((long long)p.jj * ni_s + p.ii) * rec + k];
    float ta;
    if (p.jcopy)
        ta = a00;
    else {
        const float a10 = base[((long long)p.jj2 * ni_s + p.ii) * rec + k];
        ta = a00 * p.wj1 + a10 * p.wj2;
    }
    if (p.icopy)
        return ta;
    const float a01 = base[((long long)p.jj * ni_s + p.ii2) * rec + k];
    float tb;
    if (p.jcopy)
        tb = a01;
    else {
        const float a11 = base[((long long)p.jj2 * ni_s + p.ii2) * rec + k];
        tb = a01 * p.wj1 + a11 * p.wj2;
    }
    return ta * p.wi1 + tb * p.wi2;
}

__device__ inline bool masked(int nl, int level, int issurf, int surfk, int surf_thresh)
{
    if (surfk == 0)
        return nl <= level;
    return nl <= (issurf ? surf_thresh : surfk - level);
}

// ---------------------------------------------------------------- das_updatefields
// One block of 256 threads per grid cell (and per group of up to FG fields).
// The cell's transform is interpolated once into LDS and applied to every
// field of the group as one small GEMM, V^a (fields x m_d) = V^f (fields x m)
// T^T, register-blocked FPT fields x MT members per thread.  The per-field
// sums (means, A^f w, spread sums) are sequential FP64 loops over the members
// in the reference's order, one thread per field, so they are bit-exact;
// only the FP32 transform product is summed in a different order than sgemv.
template <int MT>
struct UpdCfg {
    static constexpr int MP = 16 * MT;
    static constexpr int FG = (MT <= 9) ? 112 : 64;   // fields per group
    static constexpr int FPT = FG / 16;               // fields per thread
};

template <int MT>
struct UpdShared {
    static constexpr int MP = UpdCfg<MT>::MP;
    static constexpr int FG = UpdCfg<MT>::FG;
    float at[MP][FG + 1];     // at[e][f]: member e of field f (values, anomalies, then v_a)
    float tt[MP][MP + 1];     // tt[e2][e1] = T_cell[e1][e2]
    float wc[MP];             // w_cell
    double v1f[FG], v1a[FG], v2f[FG];
    float infl[FG];
    int zero[FG];             // 0 update, 1 all members zero, 2 static members zero
};

template <int MT>
__global__ void __launch_bounds__(256) update_cell_kernel(UpdArgs a)
{
    constexpr int MP = UpdCfg<MT>::MP;
    constexpr int FG = UpdCfg<MT>::FG;
    constexpr int FPT = UpdCfg<MT>::FPT;
    __shared__ UpdShared<MT> sh;

    const long long cell = xcd_cell(a.ncell);
    if (cell >= a.ncell)
        return;
    const int tid = threadIdx.x;
    const int jl = (int)(cell / a.mni);
    const int i = (int)(cell % a.mni);
    const int j = a.j1 + jl;
    const int m = a.m, md = a.md;
    const bool hybrid = (a.mode == ENKFG_MODE_HYBRID);
    const double k_d = (md > 1) ? sqrt((double)(m - 1) / (double)(md - 1)) : 1.0;
    const Interp ip = make_interp(j, i, a.stride, a.nj_s, a.ni_s, a.periodic_i);
    const int nl = a.nlevels[(long long)j * a.mni + i];

    // ---- transform and w of the cell (interpolated once per cell); the node
    // values of UB elements per thread are loaded before any is used, so that
    // the loads are in flight together
    {
        constexpr int UB = 12;
        const long long rec = (long long)md * m;
        const float* p00 = a.T + ((long long)ip.jj * a.ni_s + ip.ii) * rec;
        const float* p10 = a.T + ((long long)ip.jj2 * a.ni_s + ip.ii) * rec;
        const float* p01 = a.T + ((long long)ip.jj * a.ni_s + ip.ii2) * rec;
        const float* p11 = a.T + ((long long)ip.jj2 * a.ni_s + ip.ii2) * rec;
        const int n = md * m;
        for (int q0 = tid; q0 < n; q0 += 256 * UB) {
            float v00[UB], v10[UB], v01[UB], v11[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int q = q0 + 256 * u;
                const bool ok = q < n;
                v00[u] = ok ? p00[q] : 0.0f;
                v10[u] = (ok && !ip.jcopy) ? p10[q] : 0.0f;
                v01[u] = (ok && !ip.icopy) ? p01[q] : 0.0f;
                v11[u] = (ok && !ip.jcopy && !ip.icopy) ? p11[q] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int q = q0 + 256 * u;
                if (q < n) {
                    // update.c:224-271: j first, then i, FP32, no contraction
                    const float ta = ip.jcopy ? v00[u] : v00[u] * ip.wj1 + v10[u] * ip.wj2;
                    const float tb = ip.jcopy ? v01[u] : v01[u] * ip.wj1 + v11[u] * ip.wj2;
                    const int e1 = q / m, e2 = q - e1 * m;
                    sh.tt[e2][e1] = ip.icopy ? ta : ta * ip.wi1 + tb * ip.wi2;
                }
            }
        }
        for (int q = tid; q < MP * (MP + 1); q += 256) {
            const int e2 = q / (MP + 1), e1 = q - e2 * (MP + 1);
            if (e1 >= md || e2 >= m)
                sh.tt[e2][e1] = 0.0f;
        }
        if (tid < MP)
            sh.wc[tid] = (tid < m) ? interp_at(a.w, m, a.ni_s, ip, tid) : 0.0f;
    }

    for (int f0 = 0; f0 < a.nfield; f0 += FG) {
        const int nf = min(FG, a.nfield - f0);
        __syncthreads();   // previous group's reads of sh are done
        // ---- load the group's fields of this cell: at[e][f] (all loads of a
        // thread issued before the stores)
        {
            constexpr int NL = (FG * MP + 255) / 256;
            float v[NL];
#pragma unroll
            for (int u = 0; u < NL; ++u) {
                const int q = tid + 256 * u;
                const int f = q / MP, e = q - f * MP;
                v[u] = 0.0f;
                if (f < nf && e < m && !masked(nl, a.flevel[f0 + f], a.fsurf[f0 + f], a.surfk, 0))
                    v[u] = a.E[((long long)(f0 + f) * a.ncell + cell) * m + e];
            }
#pragma unroll
            for (int u = 0; u < NL; ++u) {
                const int q = tid + 256 * u;
                const int f = q / MP, e = q - f * MP;
                if (f < FG)
                    sh.at[e][f] = v[u];
            }
        }
        __syncthreads();
        // ---- per-field stats (update.c:298-410), one thread per field
        if (tid < nf) {
            const int f = tid;
            int z = 0;
            if (masked(nl, a.flevel[f0 + f], a.fsurf[f0 + f], a.surfk, 0)) {
                z = 1;
            } else {
                int eb = m;
                for (int e = 0; e < m; ++e) {
                    const float v = sh.at[e][f];
                    if (!isfinite(v) || fabsf(v) > kMaxObsVal) {
                        eb = e;
                        break;
                    }
                }
                if (eb < md) {
                    z = 1;
                } else if (eb < m) {
                    z = 2;
                    for (int e = md; e < m; ++e)
                        sh.at[e][f] = 0.0f;
                }
            }
            sh.zero[f] = z;
            if (z != 1) {
                double v1_f = 0.0;
                for (int e = 0; e < md; ++e)
                    v1_f += sh.at[e][f];
                v1_f /= (double)md;
                double v1_a = 0.0, v2_f = 0.0;
                for (int e = 0; e < m; ++e) {
                    const float vf = (float)((double)sh.at[e][f] - v1_f);
                    sh.at[e][f] = vf;
                    v1_a += vf * sh.wc[e];
                    if (e < md) {
                        const float vh = hybrid ? (float)((double)vf / k_d) : vf;
                        v2_f += vh * vh;
                    }
                }
                sh.v1f[f] = v1_f;
                sh.v1a[f] = v1_a + v1_f;
                sh.v2f[f] = v2_f;
            }
        }
        __syncthreads();
        // ---- V^a = V^f T^T (sgemv('T', m, md, 1, T, m, v_f, ...), update.c:399)
        const int tf = tid >> 4, te = tid & 15;
        float acc[FPT][MT];
#pragma unroll
        for (int x = 0; x < FPT; ++x)
#pragma unroll
            for (int y = 0; y < MT; ++y)
                acc[x][y] = 0.0f;
        for (int k = 0; k < m; ++k) {
            float av[FPT], tv[MT];
#pragma unroll
            for (int x = 0; x < FPT; ++x)
                av[x] = sh.at[k][tf * FPT + x];
#pragma unroll
            for (int y = 0; y < MT; ++y)
                tv[y] = sh.tt[k][te * MT + y];
#pragma unroll
            for (int x = 0; x < FPT; ++x)
#pragma unroll
                for (int y = 0; y < MT; ++y)
                    acc[x][y] = fmaf(av[x], tv[y], acc[x][y]);
        }
        __syncthreads();   // all reads of the anomalies are done
#pragma unroll
        for (int x = 0; x < FPT; ++x)
#pragma unroll
            for (int y = 0; y < MT; ++y)
                sh.at[te * MT + y][tf * FPT + x] = acc[x][y];
        __syncthreads();
        // ---- capped inflation (update.c:432-458), one thread per field
        if (tid < nf) {
            const int f = tid;
            const float inflation0 = a.finfl[f0 + f];
            const double inf_ratio = a.finfratio[f0 + f];
            float inflation = inflation0;
            if (sh.zero[f] != 1 && !isnan(inf_ratio)) {
                double v2_a = 0.0;
                for (int e = 0; e < md; ++e) {
                    const float va = hybrid ? (float)((double)sh.at[e][f] / k_d) : sh.at[e][f];
                    v2_a += va * va;
                }
                if (v2_a > 0.0) {
                    inflation = (float)(sqrt(sh.v2f[f] / v2_a) * inf_ratio + 1.0 - inf_ratio);
                    if (inflation >= inflation0)
                        inflation = inflation0;
                }
            }
            if (!(inflation - 1.0f > kEpsF))
                inflation = 1.0f;
            sh.infl[f] = inflation;
        }
        __syncthreads();
        // ---- outputs (update.c:459-526)
        for (int q = tid; q < nf * m; q += 256) {
            const int f = q / m, e = q - f * m;
            float* rec = a.E + ((long long)(f0 + f) * a.ncell + cell) * m;
            const int z = sh.zero[f];
            if (z == 1) {
                rec[e] = 0.0f;
                continue;
            }
            if (e >= md) {
                if (z == 2)
                    rec[e] = 0.0f;
                continue;
            }
            float va = sh.at[e][f];
            if (hybrid)
                va = (float)((double)va / k_d);
            va = va * sh.infl[f];   // inflation is 1 when not applied
            va = (float)((double)va + sh.v1a[f]);
            float out = va;
            if (a.outputinc) {
                float vf = (float)((double)rec[e] - sh.v1f[f]);
                if (hybrid)
                    vf = (float)((double)vf / k_d);
                vf = (float)((double)vf + sh.v1f[f]);
                if (!a.fapplylog[f0 + f])
                    out = va - vf;
                else
                    out = isnormal(vf) ? (float)(exp10((double)va) - exp10((double)vf)) : 0.0f;
            }
            rec[e] = out;
        }
        if (a.infl != nullptr && tid < nf)
            a.infl[(long long)(f0 + tid) * a.ncell + cell] = (sh.zero[tid] == 1) ? 0.0f : sh.infl[tid];
    }
}

// das_updatebg: E is [field][cell][m + 1], member m the background
template <int NWT>
__global__ void __launch_bounds__(64 * NWT) update_bg_kernel(UpdArgs a)
{
    __shared__ float xs[64 * NWT];
    __shared__ float xmean_s;
    __shared__ int bad;
    const long long cell = xcd_cell(a.ncell);
    if (cell >= a.ncell)
        return;
    const int jl = (int)(cell / a.mni);
    const int i = (int)(cell % a.mni);
    const int j = a.j1 + jl;
    const int e = threadIdx.x;
    const int m = a.m;
    const Interp ip = make_interp(j, i, a.stride, a.nj_s, a.ni_s, a.periodic_i);
    const float wv = (e < m) ? interp_at(a.w, m, a.ni_s, ip, e) : 0.0f;
    const int nl = a.nlevels[(long long)j * a.mni + i];
    for (int f = 0; f < a.nfield; ++f) {
        float* rec = a.E + ((long long)f * a.ncell + cell) * (m + 1);
        // the surface threshold is 1 here (update.c:730), 0 in das_updatefields
        if (masked(nl, a.flevel[f], a.fsurf[f], a.surfk, 1)) {
            if (e == 0)
                rec[m] = 0.0f;
            continue;
        }
        const float bg = rec[m];
        if (fabsf(bg) > kMaxObsVal) {
            if (e == 0)
                rec[m] = 0.0f;
            continue;
        }
        const float v = (e < m) ? rec[e] : 0.0f;
        __syncthreads();   // the previous field's reads of shared data are done
        if (e == 0)
            bad = 0;
        __syncthreads();
        if (e < m && fabsf(v) > kMaxObsVal)
            bad = 1;
        xs[e] = v;
        __syncthreads();
        if (bad) {
            if (e == 0)
                rec[m] = 0.0f;
            continue;
        }
        // float mean and float accumulation in member order (update.c:751-761)
        if (e == 0) {
            float xm = 0.0f;
            for (int k = 0; k < m; ++k)
                xm += xs[k];
            xm /= (float)m;
            xmean_s = xm;
        }
        __syncthreads();
        const float xmean = xmean_s;
        const float contrib = (e < m) ? (v - xmean) * wv : 0.0f;
        __syncthreads();
        xs[e] = contrib;
        __syncthreads();
        if (e == 0) {
            float b = bg;
            for (int k = 0; k < m; ++k)
                b += xs[k];
            rec[m] = b;
        }
    }
}

// das_sethybridensemble on E[field][cell][m]
__global__ void hybrid_condition_kernel(UpdArgs a, double k_d, double k_s)
{
    const long long n = (long long)a.nfield * a.ncell;
    for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long long)gridDim.x * blockDim.x) {
        float* rec = a.E + r * a.m;
        double mean = 0.0;
        for (int e = 0; e < a.md; ++e)
            mean += (double)rec[e];
        mean /= (double)a.md;
        for (int e = 0; e < a.md; ++e)
            rec[e] = (float)(((double)rec[e] - mean) * k_d + mean);
        for (int e = a.md; e < a.m; ++e)
            rec[e] = (float)((double)rec[e] * k_s + mean);
    }
}

__global__ void transpose_tiles_kernel(const float* __restrict__ in, float* __restrict__ out, long long rows,
                                       long long cols)
{
    __shared__ float tile[32][33];
    const long long ntc = (cols + 31) / 32;
    const long long t = blockIdx.x;
    const long long r0 = (t / ntc) * 32, c0 = (t % ntc) * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        long long r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols)
            tile[k][threadIdx.x] = in[r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        long long c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols)
            out[c * rows + r] = tile[threadIdx.x][k];
    }
}

hipError_t transpose(hipStream_t s, const float* in, float* out, long long rows, long long cols)
{
    if (rows == 0 || cols == 0)
        return hipSuccess;
    const long long ntiles = ((rows + 31) / 32) * ((cols + 31) / 32);
    if (ntiles > 0x7fffffffLL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(transpose_tiles_kernel, dim3((unsigned)ntiles), dim3(32, 8), 0, s, in, out, rows, cols);
    return hipGetLastError();
}

unsigned cell_blocks(long long ncell)
{
    return (unsigned)(8 * ((ncell + 7) / 8));
}

}  // namespace

hipError_t launch_update(hipStream_t s, const UpdArgs& a)
{
    if (a.ncell == 0 || a.nfield == 0)
        return hipSuccess;
    const int MT = (a.m + 15) / 16;
    const dim3 grid(cell_blocks(a.ncell));
    const dim3 block(256);
    switch (MT) {
    case 1: hipLaunchKernelGGL(update_cell_kernel<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(update_cell_kernel<2>, grid, block, 0, s, a); break;
    case 3: hipLaunchKernelGGL(update_cell_kernel<3>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(update_cell_kernel<4>, grid, block, 0, s, a); break;
    case 5: hipLaunchKernelGGL(update_cell_kernel<5>, grid, block, 0, s, a); break;
    case 6: hipLaunchKernelGGL(update_cell_kernel<6>, grid, block, 0, s, a); break;
    case 7: hipLaunchKernelGGL(update_cell_kernel<7>, grid, block, 0, s, a); break;
    case 8: hipLaunchKernelGGL(update_cell_kernel<8>, grid, block, 0, s, a); break;
    case 9: hipLaunchKernelGGL(update_cell_kernel<9>, grid, block, 0, s, a); break;
    case 10: hipLaunchKernelGGL(update_cell_kernel<10>, grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_update_bg(hipStream_t s, const UpdArgs& a)
{
    if (a.ncell == 0 || a.nfield == 0)
        return hipSuccess;
    const dim3 grid(cell_blocks(a.ncell));
    if (a.m < 64)
        hipLaunchKernelGGL(update_bg_kernel<1>, grid, dim3(64), 0, s, a);
    else if (a.m < 128)
        hipLaunchKernelGGL(update_bg_kernel<2>, grid, dim3(128), 0, s, a);
    else if (a.m < 192)
        hipLaunchKernelGGL(update_bg_kernel<3>, grid, dim3(192), 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_hybrid_condition(hipStream_t s, const UpdArgs& a, double gamma)
{
    const int ms = a.m - a.md;
    if (ms < 2 || a.md < 1)
        return hipErrorInvalidValue;
    const double k_d = (a.md > 1) ? sqrt((double)(a.m - 1) / (double)(a.md - 1)) : 0.0;
    const double k_s = sqrt(gamma * (double)(a.m - 1) / (double)(ms - 1));
    const long long n = (long long)a.nfield * a.ncell;
    long long nb = (n + 255) / 256;
    if (nb > 65536)
        nb = 65536;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(hybrid_condition_kernel, dim3((unsigned)nb), dim3(256), 0, s, a, k_d, k_s);
    return hipGetLastError();
}

hipError_t launch_members_to_cells(hipStream_t s, const float* in, float* out, int nmem, long long ncell)
{
    return transpose(s, in, out, nmem, ncell);
}

hipError_t launch_cells_to_members(hipStream_t s, const float* in, float* out, int nmem, long long ncell)
{
    return transpose(s, in, out, ncell, nmem);
}

}  // namespace enkfg
