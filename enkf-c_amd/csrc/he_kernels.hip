// he_kernels.hip -- the ensemble observations HE on the device (das_getHE,
// calc/ensobs.c:46-508), from the device-resident ensemble, followed by the
// hybrid scaling (ensobs.c:434-486) and das_calcinnandspread
// (ensobs.c:541-620).
//
// One wavefront per observation, the members on the lanes (member e on lane
// e % 64, slot e / 64): the H function of the observation's type reads the
// members of the 4 (surface, H_surf_standard) or 8 (subsurface,
// H_subsurf_standard) surrounding grid values -- one coalesced record each,
// since the ensemble is [field][cell][member] -- and forms the weighted sum
// in FP64 in the reference's order (interpolate2d_structured /
// interpolate3d_structured, common/utils.c:1140-1367), so HE is the
// reference's bit for bit.  The sums over the members (dynamic and static
// means, the ensemble mean and the spread) are taken in member order, as the
// reference's loops take them, by broadcasting each member's value from its
// lane.  Observations whose status is not OK are not evaluated
// (obs_find_bytype): their HE stays NaN.
//
// Outputs: HA = HE - mean (FP32, [obs][m], observation-major), the
// innovations s = y - mean and the spread std (FP64); a non-finite forecast
// or innovation, or one beyond MAXOBSVAL, sets the error flag (the reference
// calls enkf_quit there, model2obs.c:111-119, ensobs.c:607-615).
#include "enkfg_internal.h"

namespace enkfg {

namespace {

constexpr double kMaxObsVal = 999999.0;   // MAXOBSVAL, common/global.h:119
constexpr double kEpsFloat = 1.0e-7;      // EPS_FLOAT, common/utils.c:51

// lane l's value of v, on every lane
__device__ inline double bcast(double v, int l)
{
    return __shfl(v, l);
}

template <int NE>
__device__ inline double member_sum(const double (&v)[NE], int e0, int e1)
{
    // sum of members e0 .. e1 - 1 in member order (every lane gets it)
    double s = 0.0;
    for (int e = e0; e < e1; ++e) {
        double x = 0.0;
#pragma unroll
        for (int q = 0; q < NE; ++q)
            if ((e >> 6) == q)
                x = bcast(v[q], e & 63);
        s += x;
    }
    return s;
}

template <int NE>
__global__ void __launch_bounds__(256) he_kernel(HeGetArgs a)
{
    const int o = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (o >= a.nobs)
        return;
    const int lane = threadIdx.x & 63;
    const int m = a.m;
    float* out = a.HA + (size_t)o * m;
    const bool ok = a.status[o] == ENKFG_STATUS_OK;
    double v[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q)
        v[q] = __builtin_nan("");
    if (ok) {
        const int t = a.type[o];
        const int hk = a.hkind[t];
        const int ni = a.ni, nj = a.nj;
        double fi = a.fi[o], fj = a.fj[o];
        // interpolate2d_structured / interpolate3d_structured index set-up
        if (hk == 0 ? fi >= (double)ni : fi > (double)ni)
            fi = (double)ni - kEpsFloat;
        if (fj >= (double)nj)
            fj = (double)nj - kEpsFloat;
        int i1 = (int)floor(fi), i2 = (int)ceil(fi), j1 = (int)floor(fj), j2 = (int)ceil(fj);
        double wi1 = ceil(fi) - fi, wi2 = fi - floor(fi), wj1 = ceil(fj) - fj, wj2 = fj - floor(fj);
        int k1 = 0, k2 = 0, k1top = 0, k2top = 0;
        double wk1 = 1.0, wk2 = 0.0;
        const int nk = a.nk[t];
        if (hk != 0) {
            double fk = (double)a.fk[o];
            if (fk < 0.0)
                fk = 0.0;
            if (fk > (double)(nk - 1))
                fk = (double)(nk - 1);
            k1 = (int)floor(fk);
            k1top = abs(a.ktop - k1);
            wk1 = ceil(fk) - fk;
            k2 = (int)ceil(fk);
            k2top = abs(a.ktop - k2);
            wk2 = fk - floor(fk);
            if (wk1 == 0.0 && wk2 == 0.0)
                wk1 = 1.0;
        }
        if (i1 == i2)
            wi1 = 1.0;
        if (j1 == j2)
            wj1 = 1.0;
        if (hk != 0 && k1 == k2)
            wk1 = 1.0;
        if (i1 == -1)
            i1 = a.periodic_i ? ni - 1 : i2;
        if (i2 == ni)
            i2 = a.periodic_i ? 0 : i1;
        if (j1 == -1)
            j1 = j2;
        if (j2 == nj)
            j2 = j1;
        const int* fl = a.fields + a.field0[t];
        const long long c11 = (long long)j1 * ni + i1, c12 = (long long)j1 * ni + i2;
        const long long c21 = (long long)j2 * ni + i1, c22 = (long long)j2 * ni + i2;
        const long long cells[4] = {c11, c12, c21, c22};
        const double wij[4] = {wj1 * wi1, wj1 * wi2, wj2 * wi1, wj2 * wi2};
        double sum[NE], w = 0.0;
#pragma unroll
        for (int q = 0; q < NE; ++q)
            sum[q] = 0.0;
        const int nlay = hk == 0 ? 1 : 2;
        for (int L = 0; L < nlay; ++L) {
            // H_surf_standard reads the variable's surface layer (ktop; the
            // only layer of a 2D variable)
            const int f = hk == 0 ? fl[nk > 1 ? a.ktop : 0] : fl[L == 0 ? k1 : k2];
            const int ktp = L == 0 ? k1top : k2top;
            const double wk = L == 0 ? wk1 : wk2;
            for (int c = 0; c < 4; ++c) {
                const int nl = a.nlevels[cells[c]];
                if (hk == 0 ? nl == 0 : nl <= ktp)
                    continue;
                const double ww = hk == 0 ? wij[c] : wij[c] * wk;
                const float* rec = a.E + ((size_t)f * a.ncell + cells[c]) * m;
#pragma unroll
                for (int q = 0; q < NE; ++q) {
                    const int e = lane + 64 * q;
                    if (e < m)
                        sum[q] += (double)rec[e] * ww;
                }
                w += ww;
            }
        }
        bool bad = false;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            const float x = (float)(sum[q] / w);
            v[q] = (double)x;
            if (e < m && (!isfinite(x) || fabs((double)x) > kMaxObsVal))
                bad = true;
        }
        if (__any(bad) && lane == 0)
            atomicExch(a.err, 1);
    }
    // hybrid scaling (ensobs.c:434-486): dynamic anomalies by k_d, static by
    // k_s about the dynamic mean
    if (a.mode == ENKFG_MODE_HYBRID) {
        const int md = a.md, ms = m - a.md;
        const double k_d = (md > 1) ? sqrt((double)(m - 1) / (double)(md - 1)) : 0.0;
        const double k_s = sqrt(a.gamma * (double)(m - 1) / (double)(ms - 1));
        const double mean_d = member_sum<NE>(v, 0, md) / (double)md;
        const double mean_s = member_sum<NE>(v, md, m) / (double)ms;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            float x = (float)v[q];
            if (e < md)
                x = (float)(((double)x - mean_d) * k_d + mean_d);
            else
                x = (float)(((double)x - mean_s) * k_s + mean_d);
            v[q] = (double)x;
        }
    }
    // das_calcinnandspread: mean over all members, HA = HE - (float) mean,
    // spread from the FP32 squares
    const double mean = member_sum<NE>(v, 0, m) / (double)m;
    double sq[NE];
#pragma unroll
    for (int q = 0; q < NE; ++q) {
        const int e = lane + 64 * q;
        const float x = (float)v[q] - (float)mean;
        if (e < m)
            out[e] = x;
        sq[q] = (double)(x * x);
    }
    double sp = member_sum<NE>(sq, 0, m);
    double s = mean;
    if (ok) {
        sp = sqrt(sp / (double)(m - 1));
        s = a.value[o] - mean;
        if ((!isfinite(s) || fabs(s) > kMaxObsVal) && lane == 0)
            atomicExch(a.err, 2);
    }
    if (lane == 0) {
        a.s[o] = s;
        a.std[o] = sp;
    }
}

}  // namespace

// rows of obs that are not OK: HA and s set to 0 for the calc, which never
// reads them (such obs are in no local list)
__global__ void he_clear_bad_kernel(const short* __restrict__ status, int nobs, int m, float* HA, double* s)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nobs * m)
        return;
    const int o = (int)(idx / m);
    if (status[o] != ENKFG_STATUS_OK) {
        HA[idx] = 0.0f;
        if (idx % m == 0)
            s[o] = 0.0;
    }
}

hipError_t launch_he(hipStream_t s, const HeGetArgs& a)
{
    if (a.nobs == 0)
        return hipSuccess;
    const dim3 grid((unsigned)((a.nobs + 3) / 4));
    if (a.m <= 64)
        hipLaunchKernelGGL(he_kernel<1>, grid, dim3(256), 0, s, a);
    else if (a.m <= 128)
        hipLaunchKernelGGL(he_kernel<2>, grid, dim3(256), 0, s, a);
    else if (a.m <= 192)
        hipLaunchKernelGGL(he_kernel<3>, grid, dim3(256), 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_he_clear_bad(hipStream_t s, const short* status, int nobs, int m, float* HA, double* sv)
{
    const long long n = (long long)nobs * m;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(he_clear_bad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, status, nobs, m, HA,
                       sv);
    return hipGetLastError();
}

// das_moderateobs (calc/ensobs.c:749-799): the K-factor moderation of the
// error of every good observation from its forecast spread and innovation,
//   estd' = sqrt(sqrt((svar + ovar)^2 + svar inn^2 / K / K) - svar),
// in double with every operation rounded separately (no contraction into
// fma, as the reference's host build evaluates it), stored as float.  The
// obs whose error variance more than doubles are counted per type.
__global__ void moderate_kernel(int nobs, double kfactor, const short* status, const short* type, const double* std,
                                const double* inn, float* estd, int* nmod)
{
#pragma clang fp contract(off)
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nobs || status[o] != ENKFG_STATUS_OK)
        return;
    const double svar = std[o] * std[o];
    const double ov = (double)estd[o];
    const double ovar = ov * ov;
    const double x = inn[o];
    const double sum = svar + ovar;
    const double e = sqrt(sqrt(sum * sum + svar * x * x / kfactor / kfactor) - svar);
    if (svar > 0.0 && e * e / ovar > 2.0)
        atomicAdd(&nmod[type[o]], 1);
    estd[o] = (float)e;
}

hipError_t launch_moderate(hipStream_t s, int nobs, double kfactor, const short* status, const short* type,
                           const double* std, const double* inn, float* estd, int* nmod)
{
    if (nobs == 0)
        return hipSuccess;
    hipLaunchKernelGGL(moderate_kernel, dim3((unsigned)((nobs + 255) / 256)), dim3(256), 0, s, nobs, kfactor, status,
                       type, std, inn, estd, nmod);
    return hipGetLastError();
}

}  // namespace enkfg
