// obs_update.hip -- the analysed ensemble observations (das_updateHE,
// calc/ensobs.c:2116-2129): the update of update/update.c applied to the
// ensemble observation vectors at the observations' nearest grid cells.
//
// update_HE_structured (ensobs.c:1228-1553), per observation o of a
// structured grid with status OK whose nearest row (int)(fj + 0.5) is a row
// of the grid, at the cell (i, j) = ((int)(fi + 0.5), (int)(fj + 0.5)) (the
// last column when i rounds to mni):
//   x^f = sum_{e < m_d} HE^f_e / m_d (FP64);  HA^f = HE^f - x^f (FP32)
//   HA^a = HA^f T_cell (FP32 products and sums);  x^a = x^f + sum_e HA^f_e w_e
//   HE^a_e = HA^a_e + x^a (e < m_d), capped or plain inflation about x^a
//   (ensobs.c:1475-1520), and the static members shifted by x^a - x^f;
// T_cell and w_cell are interpolated from the stride grid exactly as the
// field update does (update_interp.h).  EnOI (update_Hx_structured,
// ensobs.c:1933-1962): every member moves by (HE - mean HE) . w_cell.
//
// Layout: one workgroup per grid cell that holds observations (the host
// groups them by a counting sort on the cell); T_cell is interpolated once
// into LDS and each of the four waves then takes the cell's observations in
// turn, one observation per wave, its members on the lanes.  HE is
// observation-major on the device ([obs][m], one observation = one row).
#include "enkfg_internal.h"
#include "update_interp.h"

namespace enkfg {

namespace {

constexpr float kEps = 1.0e-6f;   // EPSF, update/update.c:41

template <int MT>
struct HeLds {
    static constexpr int MP = 16 * MT;
    float tc[MP][MP + 1];          // T_cell[e1][e2] (row stride m + 1 against bank conflicts)
    float wc[MP];
    float hf[4][MP];               // HA^f of each wave's observation
};

// full-wave sum in a fixed order (butterfly)
__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

template <int MT>
__global__ void __launch_bounds__(256) update_he_kernel(UpdArgs u, HeArgs h)
{
    constexpr int MP = 16 * MT;
    constexpr int NE = (MP + 63) / 64;     // members per lane
    __shared__ HeLds<MT> sh;
    const int g = blockIdx.x;
    if (g >= h.ngroup)
        return;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int m = u.m, md = u.md;
    const bool enoi = u.mode == ENKFG_MODE_ENOI;
    const int cell = h.gcell[g];
    const int j = cell / u.mni, i = cell % u.mni;

    // T_cell and w_cell, update.c:195-272 order
    if (!enoi) {
        const Bilin b = make_bilin(u, j, i, (long long)md * m);
        for (int idx = tid; idx < md * m; idx += 256) {
            const int e1 = idx / m, e2 = idx - e1 * m;
            sh.tc[e1][e2] = bilin(b, u.T[b.o00 + idx], u.T[b.o10 + idx], u.T[b.o01 + idx], u.T[b.o11 + idx]);
        }
    }
    {
        const Bilin b = make_bilin(u, j, i, m);
        for (int e = tid; e < m; e += 256)
            sh.wc[e] = bilin(b, u.w[b.o00 + e], u.w[b.o10 + e], u.w[b.o01 + e], u.w[b.o11 + e]);
    }
    __syncthreads();

    const int o0 = h.goff[g], o1 = h.goff[g + 1];
    for (int k = o0 + wv; k < o1; k += 4) {
        const int o = h.order[k];
        float* row = h.HE + (size_t)o * m;
        float v[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            v[q] = (e < m) ? row[e] : 0.0f;
        }
        if (enoi) {
            // update_Hx_structured: Hx over all members, dHx = sum (HE - Hx) w
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < NE; ++q)
                s += (double)v[q];
            const double Hx = wave_sum(s) / (double)m;
            double d = 0.0;
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                const int e = lane + 64 * q;
                if (e < m)
                    d += ((double)v[q] - Hx) * (double)sh.wc[e];
            }
            const double dHx = wave_sum(d);
#pragma unroll
            for (int q = 0; q < NE; ++q) {
                const int e = lane + 64 * q;
                if (e < m)
                    row[e] = (float)((double)v[q] + dHx);
            }
            continue;
        }
        // x^f over the dynamic members
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            if (e < md)
                s += (double)v[q];
        }
        const double v1f = wave_sum(s) / (double)md;
        float af[NE];
        double dx = 0.0;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            af[q] = (e < m) ? (float)((double)v[q] - v1f) : 0.0f;
            if (e < m) {
                sh.hf[wv][e] = af[q];
                dx += (double)(af[q] * sh.wc[e]);
            }
        }
        const double v1a = wave_sum(dx) + v1f;
        // HA^a[e1] = sum_e2 T_cell[e1][e2] HA^f[e2]: the wave's own LDS row
        // (written above by this wave's lanes, read after in program order)
        __builtin_amdgcn_wave_barrier();
        float aa[NE];
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e1 = lane + 64 * q;
            float acc = 0.0f;
            if (e1 < md)
                for (int e2 = 0; e2 < m; ++e2)
                    acc += sh.tc[e1][e2] * sh.hf[wv][e2];
            aa[q] = acc;
        }
        // HE^a = HA^a + x^a; HE^f restored as HA^f + x^f
        float ha[NE], hfr[NE];
        double sf = 0.0, sa = 0.0;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            ha[q] = (float)((double)aa[q] + v1a);
            hfr[q] = (float)((double)af[q] + v1f);
            if (e < md) {
                sf += (double)hfr[q] * (double)hfr[q];
                sa += (double)ha[q] * (double)ha[q];
            }
        }
        const double ratio = h.ratio[o];
        const float infl0 = h.infl0[o];
        float infl = NAN;
        if (!isnan(ratio)) {
            const double var_f = wave_sum(sf) / (double)md - v1f * v1f;
            const double var_a = wave_sum(sa) / (double)md - v1a * v1a;
            if (var_a > 0.0) {
                infl = (float)(sqrt(var_f / var_a) * ratio + 1.0 - ratio);
                if (infl >= infl0)
                    infl = infl0;
            }
        } else {
            infl = infl0;
        }
        const bool inflate = fabsf(infl - 1.0f) > kEps;
        const float v1af = (float)v1a;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = lane + 64 * q;
            if (e < md) {
                float r = ha[q];
                if (inflate)
                    r = (float)((double)((r - v1af) * infl) + v1a);
                row[e] = r;
            } else if (e < m) {
                row[e] = (float)((double)v[q] + (v1a - v1f));
            }
        }
        // the wave's next observation rewrites hf[wv] only after every lane
        // has read it (the reads above are complete in program order)
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

hipError_t launch_update_he(hipStream_t s, const UpdArgs& u, const HeArgs& h)
{
    if (h.ngroup == 0)
        return hipSuccess;
    const int MT = (u.m + 15) / 16;
    const dim3 grid((unsigned)h.ngroup);
    switch (MT) {
    case 1: hipLaunchKernelGGL(update_he_kernel<1>, grid, dim3(256), 0, s, u, h); break;
    case 2: hipLaunchKernelGGL(update_he_kernel<2>, grid, dim3(256), 0, s, u, h); break;
    case 3: hipLaunchKernelGGL(update_he_kernel<3>, grid, dim3(256), 0, s, u, h); break;
    case 4: hipLaunchKernelGGL(update_he_kernel<4>, grid, dim3(256), 0, s, u, h); break;
    case 5: hipLaunchKernelGGL(update_he_kernel<5>, grid, dim3(256), 0, s, u, h); break;
    case 6: hipLaunchKernelGGL(update_he_kernel<6>, grid, dim3(256), 0, s, u, h); break;
    case 7: hipLaunchKernelGGL(update_he_kernel<7>, grid, dim3(256), 0, s, u, h); break;
    case 8: hipLaunchKernelGGL(update_he_kernel<8>, grid, dim3(256), 0, s, u, h); break;
    case 9: hipLaunchKernelGGL(update_he_kernel<9>, grid, dim3(256), 0, s, u, h); break;
    case 10: hipLaunchKernelGGL(update_he_kernel<10>, grid, dim3(256), 0, s, u, h); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace enkfg
