// update_mfma.hip -- das_updatefields (update/update.c:92-564) on the matrix
// cores.
//
// One workgroup (4 waves) per grid cell; the cell's fields are processed in
// groups of 64, wave w owning fields 16w .. 16w + 15 of the group.  Per cell:
//
//   1. T and w are interpolated from the stride grid into LDS exactly as
//      update.c:195-272 does (FP32, j first then i, no contraction).
//   2. Per group: the 64 field records E[f][cell][0..m) are loaded, and each
//      field's statistics are formed by 4 lanes (a quarter of the members
//      each): first bad member (missing values, update.c:337-371), FP64 mean
//      of the dynamic members, FP32 anomalies, x_a = x_f + sum v_f w
//      (FP32 products, FP64 sums), and the spread sum for the capped
//      inflation.  The FP64 sums are combined in a fixed order; they equal
//      the reference's sequential sums whenever those are exact (members
//      within 2^29 of each other in magnitude), and differ by rounding only
//      otherwise.
//   3. V^a = V^f T_cell^T as 16x16 MFMA tiles: FP32 inputs with FP32
//      accumulation (v_mfma_f32_16x16x4_f32, the arithmetic class of the
//      reference's sgemv), or, in mixed precision, FP16 hi/lo splits of both
//      operands with FP32 accumulation (v_mfma_f32_16x16x32_f16, three
//      products: hi*hi + hi*lo + lo*hi), each field's anomalies scaled by a
//      power of two into FP16's normal range first.
//   4. Capped inflation (update.c:432-458) from the accumulators (the spread
//      sum is a 16-lane reduction), then the analysis / increment / log
//      increment outputs (update.c:459-526).
#include "enkfg_internal.h"
#include "update_interp.h"

namespace enkfg {

namespace {

constexpr float kMaxVal = 999999.0f;   // MAXOBSVAL, common/global.h:114
constexpr float kEps = 1.0e-6f;        // EPSF, update/update.c:41
constexpr int FG = 64;                 // fields per group (16 per wave)

typedef float flt4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef _Float16 half8 __attribute__((__vector_size__(8 * sizeof(_Float16))));

template <int MT, bool MIX>
struct UpdGeom {
    static constexpr int MP = 16 * MT;                          // padded members
    static constexpr int KP = MIX ? 32 * ((MP + 31) / 32) : MP; // padded K of the product
    // row strides == 16 or 48 mod 64 floats (the 4 rows of an operand read
    // fall in disjoint banks), tighter
    // for the largest ensembles so that the workgroup fits in LDS
    static constexpr int AS = (MT <= 9) ? FG + 16 : FG + 4;                          // at
    static constexpr int TS = (MT <= 9) ? MP + ((16 - MP % 32) + 32) % 32 : MP + 4;  // tt
    static constexpr int HS = KP + 8;                           // th/tl row stride (halves, 16-byte rows)
};

template <int MT, bool MIX>
struct UpdLds {
    using G = UpdGeom<MT, MIX>;
    float at[G::KP][G::AS];                        // at[e][f]: values, then anomalies (0 for e >= m)
    float tt[MIX ? 1 : G::KP][MIX ? 1 : G::TS];    // tt[e2][e1 ^ ((e2 / 4) % 16)] = T_cell[e1][e2] (FP32 path)
    alignas(16) _Float16 th[MIX ? G::MP : 1][MIX ? G::HS : 8];   // th[e1][e2] = hi(T_cell[e1][e2]) (mixed)
    alignas(16) _Float16 tl[MIX ? G::MP : 1][MIX ? G::HS : 8];   // lo part
    float wc[G::KP];
    double v1f[FG], v1a[FG], v2f[FG];
    float scale[FG];
    int zero[FG];              // 0 update, 1 all members zero, 2 static members zero
};

__device__ inline long long cell_of_block(long long ncell)
{
    // each XCD (blocks go to the 8 XCDs round-robin) gets a contiguous range
    // of cells, so the stride-grid transforms they share stay in its L2
    const long long chunk = (ncell + 7) / 8;
    return (long long)(blockIdx.x % 8) * chunk + blockIdx.x / 8;
}

__device__ inline bool is_masked(int nl, int level, int issurf, int surfk)
{
    if (surfk == 0)
        return nl <= level;
    return nl <= (issurf ? 0 : surfk - level);
}

// sum over the 16 lanes of a row (DPP: quad swaps, then row rotations by 4
// and 8); every lane of the row receives the total
template <int CTRL>
__device__ __attribute__((always_inline)) inline double dpp_f64(double x)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __attribute__((always_inline)) inline double row16_sum(double x)
{
    x += dpp_f64<0xb1>(x);    // quad_perm(1, 0, 3, 2)
    x += dpp_f64<0x4e>(x);    // quad_perm(2, 3, 0, 1)
    x += dpp_f64<0x124>(x);   // row_ror:4
    x += dpp_f64<0x128>(x);   // row_ror:8
    return x;
}

// the records of field group f0 into registers, coalesced: float q of field
// f is lane-contiguous; fields past the end re-read the last one
template <int MT, int NL>
__device__ __attribute__((always_inline)) inline void load_group(const UpdArgs& a, long long cell, int f0,
                                                                 float (&v)[NL])
{
    constexpr int MP = 16 * MT;
#pragma unroll
    for (int u = 0; u < NL; ++u) {
        const int q = (int)threadIdx.x + 256 * u;
        const int f = q / MP, e = q - f * MP;
        const int fg = min(f0 + min(f, FG - 1), a.nfield - 1);
        v[u] = a.E[((long long)fg * a.ncell + cell) * a.m + min(e, a.m - 1)];
    }
}

template <int MT, bool MIX>
__device__ inline void store_t(UpdLds<MT, MIX>& sh, int e1, int e2, float t)
{
    if constexpr (MIX) {
        const _Float16 h = (_Float16)t;
        sh.th[e1][e2] = h;
        sh.tl[e1][e2] = (_Float16)(t - (float)h);
    } else {
        // column swizzle within 16-member blocks: the lanes of a store hold
        // e2 four apart, all in one bank without it
        sh.tt[e2][e1 ^ ((e2 >> 2) & 15)] = t;
    }
}

// two workgroups per CU (LDS permitting): at most 256 registers per lane
template <int MT, bool MIX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) update_mfma_kernel(UpdArgs a)
{
    using G = UpdGeom<MT, MIX>;
    constexpr int MP = G::MP, KP = G::KP;
    __shared__ UpdLds<MT, MIX> sh;

    const long long cell = cell_of_block(a.ncell);
    if (cell >= a.ncell)
        return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int jl = (int)(cell / a.mni);
    const int i = (int)(cell % a.mni);
    const int j = a.j1 + jl;
    const int m = a.m, md = a.md;
    const bool hybrid = (a.mode == ENKFG_MODE_HYBRID);
    const double k_d = (md > 1) ? sqrt((double)(m - 1) / (double)(md - 1)) : 1.0;
    const int nl = a.nlevels[(long long)j * a.mni + i];
    // the first field group's records are loaded while T_cell is formed; each
    // later group's while the previous one is processed
    constexpr int NL = (FG * MP + 255) / 256;
    float vnext[NL];
    load_group<MT, NL>(a, cell, 0, vnext);

    // ---- 1. T_cell and w_cell (zero outside md x m)
    {
        const long long rec = (long long)md * m;
        const Bilin b = make_bilin(a, j, i, rec);
        if ((m & 3) == 0) {
            // quads of 4 consecutive e2: row e1 = q / (m / 4) by a multiply-high
            // (exact for these small q)
            constexpr int UB = 5;   // md*m/4 = 2304 quads at m = 96: two rounds of loads
            const int d = m >> 2;
            const unsigned magic = 0xffffffffu / (unsigned)d + 1u;
            const int n = md * d;
            for (int q0 = tid; q0 < n; q0 += 256 * UB) {
                float4 v00[UB], v10[UB], v01[UB], v11[UB];
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    const int q = 4 * min(q0 + 256 * u, n - 1);   // in range; duplicates are not stored
                    v00[u] = *reinterpret_cast<const float4*>(a.T + b.o00 + q);
                    v10[u] = *reinterpret_cast<const float4*>(a.T + b.o10 + q);
                    v01[u] = *reinterpret_cast<const float4*>(a.T + b.o01 + q);
                    v11[u] = *reinterpret_cast<const float4*>(a.T + b.o11 + q);
                }
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    const int q = q0 + 256 * u;
                    if (q < n) {
                        const int e1 = (int)__umulhi((unsigned)q, magic);
                        const int e2 = 4 * (q - e1 * d);
                        store_t<MT, MIX>(sh, e1, e2, bilin(b, v00[u].x, v10[u].x, v01[u].x, v11[u].x));
                        store_t<MT, MIX>(sh, e1, e2 + 1, bilin(b, v00[u].y, v10[u].y, v01[u].y, v11[u].y));
                        store_t<MT, MIX>(sh, e1, e2 + 2, bilin(b, v00[u].z, v10[u].z, v01[u].z, v11[u].z));
                        store_t<MT, MIX>(sh, e1, e2 + 3, bilin(b, v00[u].w, v10[u].w, v01[u].w, v11[u].w));
                    }
                }
            }
        } else {
            constexpr int UB = 8;
            const unsigned magic = 0xffffffffu / (unsigned)m + 1u;
            const int n = md * m;
            for (int q0 = tid; q0 < n; q0 += 256 * UB) {
                float v00[UB], v10[UB], v01[UB], v11[UB];
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    const int q = min(q0 + 256 * u, n - 1);   // in range; duplicates are not stored
                    v00[u] = a.T[b.o00 + q];
                    v10[u] = a.T[b.o10 + q];
                    v01[u] = a.T[b.o01 + q];
                    v11[u] = a.T[b.o11 + q];
                }
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    const int q = q0 + 256 * u;
                    if (q < n) {
                        const int e1 = (int)__umulhi((unsigned)q, magic), e2 = q - e1 * m;
                        store_t<MT, MIX>(sh, e1, e2, bilin(b, v00[u], v10[u], v01[u], v11[u]));
                    }
                }
            }
        }
        // zero padding of the operand: e1 in [md, MP), e2 in [m, KP)
        for (int q = tid; q < MP * KP; q += 256) {
            const int e1 = q / KP, e2 = q - e1 * KP;
            if (e1 >= md || e2 >= m)
                store_t<MT, MIX>(sh, e1, e2, 0.0f);
        }
        const Bilin bw = make_bilin(a, j, i, m);
        if (tid < KP) {
            const int e = min(tid, m - 1);
            const float v = bilin(bw, a.w[bw.o00 + e], a.w[bw.o10 + e], a.w[bw.o01 + e], a.w[bw.o11 + e]);
            sh.wc[tid] = (tid < m) ? v : 0.0f;
        }
    }

    for (int f0 = 0; f0 < a.nfield; f0 += FG) {
        // per-field parameters of this group, loaded ahead of the next
        // group's records (so that waiting for them does not wait for those)
        const int fst = min(f0 + 16 * w + (lane & 15), a.nfield - 1);   // statistics field of this lane
        const int lev_s = a.flevel[fst], surf_s = a.fsurf[fst];
        float infl0_r[4];
        double ratio_r[4];
        int log_r[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int fo = min(f0 + 16 * w + 4 * (lane >> 4) + r, a.nfield - 1);   // output fields of this lane
            infl0_r[r] = a.finfl[fo];
            ratio_r[r] = a.finfratio[fo];
            log_r[r] = a.fapplylog[fo];
        }
        __syncthreads();   // T is in LDS; the previous group's reads are done
        // ---- 2a. the group's records (loaded earlier) into LDS
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int q = tid + 256 * u;
            const int f = q / MP, e = q - f * MP;
            if (f < FG)
                sh.at[e][f] = (e < m) ? vnext[u] : 0.0f;
        }
        for (int q = tid; q < (KP - MP) * FG; q += 256)
            sh.at[MP + q / FG][q % FG] = 0.0f;
        if (f0 + FG < a.nfield)
            load_group<MT, NL>(a, cell, f0 + FG, vnext);
        __syncthreads();
        // ---- 2b. per-field statistics: lane (f, g) covers members
        // e = g (mod 4); the 4 lanes of a field are lanes f, f + 16, f + 32,
        // f + 48 of wave f / 16 (and read 4 consecutive rows of at[][], which
        // fall in disjoint banks)
        {
            const int fl = 16 * w + (lane & 15);
            const int g = lane >> 4;
            const int fgl = f0 + fl;
            const bool live = fgl < a.nfield;
            // first bad member and the FP64 sum of the dynamic members in one
            // pass (the sum is used only when no dynamic member is bad)
            int eb = m;
            double p = 0.0;
            for (int e = g; e < m; e += 4) {
                const float v = sh.at[e][fl];
                if (eb == m && (!isfinite(v) || fabsf(v) > kMaxVal))
                    eb = e;
                if (e < md)
                    p += (double)v;
            }
            eb = min(eb, __shfl_xor(eb, 16));
            eb = min(eb, __shfl_xor(eb, 32));
            int z = 0;
            if (!live || is_masked(nl, lev_s, surf_s, a.surfk))
                z = 1;
            else if (eb < md)
                z = 1;
            else if (eb < m)
                z = 2;
            // FP64 mean of the dynamic members (partials combined in lane order)
            double s0 = __shfl(p, lane & 15), s1 = __shfl(p, (lane & 15) + 16);
            double s2 = __shfl(p, (lane & 15) + 32), s3 = __shfl(p, (lane & 15) + 48);
            const double v1_f = (((s0 + s1) + s2) + s3) / (double)md;
            double pa = 0.0, pf = 0.0;
            float amax = 0.0f;
            for (int e = g; e < m; e += 4) {
                const float raw = (z == 2 && e >= md) ? 0.0f : sh.at[e][fl];
                const float vf = (z == 1) ? 0.0f : (float)((double)raw - v1_f);
                sh.at[e][fl] = vf;
                pa += (double)(vf * sh.wc[e]);
                if (e < md) {
                    const float vh = hybrid ? (float)((double)vf / k_d) : vf;
                    pf += (double)(vh * vh);
                }
                amax = fmaxf(amax, fabsf(vf));
            }
            s0 = __shfl(pa, lane & 15);
            s1 = __shfl(pa, (lane & 15) + 16);
            s2 = __shfl(pa, (lane & 15) + 32);
            s3 = __shfl(pa, (lane & 15) + 48);
            const double v1_a = (((s0 + s1) + s2) + s3) + v1_f;
            s0 = __shfl(pf, lane & 15);
            s1 = __shfl(pf, (lane & 15) + 16);
            s2 = __shfl(pf, (lane & 15) + 32);
            s3 = __shfl(pf, (lane & 15) + 48);
            const double v2_f = ((s0 + s1) + s2) + s3;
            amax = fmaxf(amax, __shfl_xor(amax, 16));
            amax = fmaxf(amax, __shfl_xor(amax, 32));
            if (g == 0) {
                sh.zero[fl] = z;
                sh.v1f[fl] = v1_f;
                sh.v1a[fl] = v1_a;
                sh.v2f[fl] = v2_f;
                // power of two that brings the largest anomaly to [2^9, 2^10)
                // (FP16 normal range for the split; exact scaling)
                sh.scale[fl] = (amax > 0.0f && isfinite(amax)) ? ldexpf(1.0f, max(-100, min(100, 9 - ilogbf(amax))))
                                                               : 1.0f;
            }
        }
        __syncthreads();
        // ---- 3. V^a (16 fields of wave w) x (MT member tiles)
        flt4 acc[MT];
#pragma unroll
        for (int J = 0; J < MT; ++J)
            acc[J] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
        const int col = lane & 15, kq = lane >> 4;
        const int fa = 16 * w + col;
        if constexpr (!MIX) {
#pragma unroll 4
            for (int s = 0; s < KP / 4; ++s) {
                const int k = 4 * s + kq;
                const float av = sh.at[k][fa];
#pragma unroll
                for (int J = 0; J < MT; ++J)
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, sh.tt[k][16 * J + (col ^ (s & 15))], acc[J], 0,
                                                                  0, 0);
            }
        } else {
            const float sc = sh.scale[fa];
#pragma unroll
            for (int s = 0; s < KP / 32; ++s) {
                const int k0 = 32 * s + 8 * kq;
                half8 ah, al;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float x = sh.at[k0 + q][fa] * sc;
                    const _Float16 h = (_Float16)x;
                    ah[q] = h;
                    al[q] = (_Float16)(x - (float)h);
                }
#pragma unroll
                for (int J = 0; J < MT; ++J) {
                    const half8 bh = *reinterpret_cast<const half8*>(&sh.th[16 * J + col][k0]);
                    const half8 bl = *reinterpret_cast<const half8*>(&sh.tl[16 * J + col][k0]);
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[J], 0, 0, 0);
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[J], 0, 0, 0);
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[J], 0, 0, 0);
                }
            }
        }
        // ---- 4. result element (field 16w + 4 kq + r, member 16 J + col) in
        // acc[J][r]; inflation per field from a 16-lane reduction
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int fl = 16 * w + 4 * kq + r;
            const int fgl = f0 + fl;
            const int z = sh.zero[fl];
            const float unscale = MIX ? 1.0f / sh.scale[fl] : 1.0f;
            double v2a = 0.0;
#pragma unroll
            for (int J = 0; J < MT; ++J) {
                const int e = 16 * J + col;
                float va = acc[J][r] * unscale;
                if (hybrid)
                    va = (float)((double)va / k_d);
                acc[J][r] = va;
                if (e < md)
                    v2a += (double)(va * va);
            }
            v2a = row16_sum(v2a);
            if (fgl >= a.nfield)
                continue;
            const float inflation0 = infl0_r[r];
            const double inf_ratio = ratio_r[r];
            float inflation = inflation0;
            if (!isnan(inf_ratio) && v2a > 0.0) {
                inflation = (float)(sqrt(sh.v2f[fl] / v2a) * inf_ratio + 1.0 - inf_ratio);
                if (inflation >= inflation0)
                    inflation = inflation0;
            }
            if (!(inflation - 1.0f > kEps))
                inflation = 1.0f;
            const double v1a = sh.v1a[fl], v1f = sh.v1f[fl];
            float* rec = a.E + ((long long)fgl * a.ncell + cell) * m;
            const bool outinc = a.outputinc != 0;
            const bool applylog = log_r[r] != 0;
#pragma unroll
            for (int J = 0; J < MT; ++J) {
                const int e = 16 * J + col;
                if (e >= m)
                    continue;
                if (z == 1) {   // masked or a dynamic member missing: all members 0
                    rec[e] = 0.0f;
                    continue;
                }
                if (e >= md) {
                    if (z == 2)
                        rec[e] = 0.0f;
                    continue;
                }
                float va = acc[J][r];
                if (inflation != 1.0f)
                    va = va * inflation;
                va = (float)((double)va + v1a);
                float out = va;
                if (outinc) {
                    float vf = sh.at[e][fl];
                    if (hybrid)
                        vf = (float)((double)vf / k_d);
                    vf = (float)((double)vf + v1f);
                    out = applylog ? (isnormal(vf) ? (float)(exp10((double)va) - exp10((double)vf)) : 0.0f)
                                   : va - vf;
                }
                rec[e] = out;
            }
            if (a.infl != nullptr && col == 0)
                a.infl[(long long)fgl * a.ncell + cell] = (z == 1) ? 0.0f : inflation;
        }
    }
}

}  // namespace

hipError_t launch_update_mfma(hipStream_t s, const UpdArgs& a, bool mixed)
{
    if (a.ncell == 0 || a.nfield == 0)
        return hipSuccess;
    const int MT = (a.m + 15) / 16;
    const long long chunk = (a.ncell + 7) / 8;
    if (8 * chunk > 0x7fffffffLL)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)(8 * chunk)), block(256);
#define ENKFG_UPD_CASE(N)                                                                      \
    case N:                                                                                    \
        if (mixed)                                                                             \
            hipLaunchKernelGGL((update_mfma_kernel<N, true>), grid, block, 0, s, a);           \
        else                                                                                   \
            hipLaunchKernelGGL((update_mfma_kernel<N, false>), grid, block, 0, s, a);          \
        break;
    switch (MT) {
        ENKFG_UPD_CASE(1)
        ENKFG_UPD_CASE(2)
        ENKFG_UPD_CASE(3)
        ENKFG_UPD_CASE(4)
        ENKFG_UPD_CASE(5)
        ENKFG_UPD_CASE(6)
        ENKFG_UPD_CASE(7)
        ENKFG_UPD_CASE(8)
        ENKFG_UPD_CASE(9)
        ENKFG_UPD_CASE(10)
    default:
        return hipErrorInvalidValue;
    }
#undef ENKFG_UPD_CASE
    return hipGetLastError();
}

}  // namespace enkfg
