// update_kernels.hip -- das_updatebg (update/update.c:568-814),
// das_sethybridensemble (common/dasystem.c:775-819) and the layout
// transposes of the update, on a device-resident ensemble (das_updatefields
// itself is in update_mfma.hip).
//
// Layout: E[field][cell][member], cell = j * mni + i (members contiguous, so
// one cell of one field is a 4*m-byte contiguous record).  w is interpolated
// from the stride grid exactly as update.c:195-272 does (FP32, j first then
// i, no contraction).
#include "enkfg_internal.h"

namespace enkfg {

namespace {

constexpr float kMaxObsVal = 999999.0f;   // MAXOBSVAL, common/global.h:114

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

// das_sethybridensemble on E[field][cell][m]: one wave stages kHybRecs
// consecutive records (one record = one cell of one field, m contiguous
// floats) in LDS with coalesced loads (float4 when VEC), rows padded to
// m + 1 words; lane r < kHybRecs forms the FP64 mean of record r's dynamic
// members in member order, as the reference does; then the lanes rewrite the
// staged span with coalesced stores:
//   dynamic: (x - mean) * k_d + mean,  static: x * k_s + mean  (in double)
constexpr int kHybRecs = 32;

template <bool VEC>
__global__ void __launch_bounds__(64) hybrid_condition_kernel(float* __restrict__ E, long long nrec, int m, int md,
                                                              double k_d, double k_s)
{
    extern __shared__ float tile[];
    __shared__ double smean[kHybRecs];
    const long long r0 = (long long)blockIdx.x * kHybRecs;
    const int nr = (int)((nrec - r0) < kHybRecs ? (nrec - r0) : kHybRecs);
    float* span = E + r0 * m;
    const int ld = m + 1;
    const int n = nr * m;
    const int lane = threadIdx.x;
    if (VEC) {
        const float4* s4 = reinterpret_cast<const float4*>(span);
        for (int k4 = lane; k4 < n / 4; k4 += 64) {
            const float4 x = s4[k4];
            const int k = 4 * k4, r = k / m, e = k - r * m;   // m % 4 == 0: the four stay in record r
            float* t = tile + r * ld + e;
            t[0] = x.x;
            t[1] = x.y;
            t[2] = x.z;
            t[3] = x.w;
        }
    } else {
        for (int k = lane; k < n; k += 64) {
            const int r = k / m;
            tile[r * ld + (k - r * m)] = span[k];
        }
    }
    __syncthreads();
    if (lane < nr) {
        const float* t = tile + lane * ld;
        double mean = 0.0;
        for (int e = 0; e < md; ++e)
            mean += (double)t[e];
        smean[lane] = mean / (double)md;
    }
    __syncthreads();
    if (VEC) {
        float4* s4 = reinterpret_cast<float4*>(span);
        for (int k4 = lane; k4 < n / 4; k4 += 64) {
            const int k = 4 * k4, r = k / m, e = k - r * m;
            const float* t = tile + r * ld + e;
            const double mean = smean[r];
            float v[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                v[c] = (e + c < md) ? (float)(((double)t[c] - mean) * k_d + mean)
                                    : (float)((double)t[c] * k_s + mean);
            s4[k4] = make_float4(v[0], v[1], v[2], v[3]);
        }
    } else {
        for (int k = lane; k < n; k += 64) {
            const int r = k / m, e = k - r * m;
            const double x = (double)tile[r * ld + e], mean = smean[r];
            span[k] = (e < md) ? (float)((x - mean) * k_d + mean) : (float)(x * k_s + mean);
        }
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

// addnoise (update/update.c:46-87): v <- v * d + mult * r_e, in the
// reference's arithmetic: the product v * d in float, the sum in double
// with mult * r_e (float times double), rounded to float on the store
__global__ void __launch_bounds__(256) addnoise_kernel(float* __restrict__ E, long long n, int m, float defl,
                                                       float mult, const double* __restrict__ r)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(i % m);
        E[i] = (float)((double)(E[i] * defl) + (double)mult * r[e]);
    }
}

}  // namespace

hipError_t launch_addnoise(hipStream_t s, float* E, long long ncell, int m, float defl, float mult, const double* r)
{
    const long long n = ncell * m;
    if (n == 0)
        return hipSuccess;
    long long nb = (n + 255) / 256;
    if (nb > 65536)
        nb = 65536;
    hipLaunchKernelGGL(addnoise_kernel, dim3((unsigned)nb), dim3(256), 0, s, E, n, m, defl, mult, r);
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
    if (n == 0)
        return hipSuccess;
    const long long nb = (n + kHybRecs - 1) / kHybRecs;
    if (nb > 0x7fffffffLL)
        return hipErrorInvalidValue;
    const size_t lds = (size_t)kHybRecs * (a.m + 1) * sizeof(float);
    // float4 staging needs 16-byte aligned spans: m % 4 == 0 and E aligned
    if (a.m % 4 == 0 && (reinterpret_cast<uintptr_t>(a.E) & 15) == 0)
        hipLaunchKernelGGL(hybrid_condition_kernel<true>, dim3((unsigned)nb), dim3(64), lds, s, a.E, n, a.m, a.md,
                           k_d, k_s);
    else
        hipLaunchKernelGGL(hybrid_condition_kernel<false>, dim3((unsigned)nb), dim3(64), lds, s, a.E, n, a.m, a.md,
                           k_d, k_s);
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
