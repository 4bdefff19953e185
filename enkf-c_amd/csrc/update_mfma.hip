// update_mfma.hip -- das_updatefields (update/update.c:92-564) on the matrix
// cores.
//
// One workgroup of 8 waves per grid cell.  The cell's fields are cut into
// tiles of 16; wave w takes tile w (and w + 8, ... when there are more than
// 128 fields).  Per cell:
//
//   1. Each wave starts loading its tile's records; meanwhile the eight waves
//      interpolate T and w from the stride grid into LDS exactly as
//      update.c:195-272 does (FP32, j first then i, no contraction), with the
//      corner loads and arithmetic of the cell's case (node, j only, i only,
//      both).
//   2. In the wave's registers: lane (col, kq) = (lane % 16, lane / 16)
//      holds field 16 t + col at members 16 s + 4 kq + c (s < KT, c < 4),
//      loaded as float4.  Those registers are the A operand of the product as
//      they stand: the product's summation runs over the members in this
//      order, and T is stored in LDS in the same order.  The lane forms its
//      share of the field's statistics there: missing values (update.c:
//      337-371), FP64 mean of the dynamic members, FP32 anomalies,
//      x_a = x_f + sum v_f w (FP32 products, FP64 sums) and the spread sum of
//      the capped inflation; the four lanes of a field combine their partial
//      sums in a fixed order.  The FP64 sums equal the reference's sequential
//      sums whenever those are exact (members within 2^29 of each other in
//      magnitude), and differ by rounding only otherwise.
//   3. V^a = V^f T_cell^T as 16 x 16 MFMA tiles: FP32 inputs with FP32
//      accumulation (v_mfma_f32_16x16x4_f32, the arithmetic class of the
//      reference's sgemv), or, in mixed precision, FP16 hi/lo splits of both
//      operands with FP32 accumulation (v_mfma_f32_16x16x32_f16, three
//      products: hi*hi + hi*lo + lo*hi), each field's anomalies scaled by a
//      power of two into FP16's normal range first.
//   4. Capped inflation (update.c:432-458) from the accumulators (the spread
//      sum is a 16-lane reduction), then the analysis / increment / log
//      increment outputs (update.c:459-526); each store writes 16
//      consecutive members of 4 fields.
#include "enkfg_internal.h"
#include "update_interp.h"

namespace enkfg {

namespace {

constexpr float kMaxVal = 999999.0f;   // MAXOBSVAL, common/global.h:119
constexpr float kEps = 1.0e-6f;        // EPSF, update/update.c:41
// waves per workgroup (one cell), one field tile each: 4 up to m = 96, so
// that four workgroups (cells) share a CU (128 VGPRs, 40 KB of LDS mixed and
// 37 KB FP32 at m = 96): mixed 85.4 -> 75.4 ms on the tile.  The FP32 T_cell
// rows, padded to 112 floats, took 43 KB and allowed three.  Larger
// ensembles: 8 (their LDS allows one or two workgroups).
template <int KT, int OT, bool MIX>
constexpr int upd_waves()
{
    return (KT <= 6 && (MIX || OT % 2 == 0)) ? 4 : 8;
}

typedef float flt4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef _Float16 half8 __attribute__((__vector_size__(8 * sizeof(_Float16))));
typedef _Float16 half4 __attribute__((__vector_size__(4 * sizeof(_Float16))));

// KT: member tiles of the summation (16 KT >= m); OT: output member tiles
// (16 OT >= m_d)
template <int KT, int OT, bool MIX>
struct TileGeom {
    static constexpr int NK = 16 * KT;                        // summation length, padded
    static constexpr int NKH = 32 * ((KT + 1) / 2);           // ... for the K = 32 half-precision steps
    static constexpr int NO = 16 * OT;                        // output members, padded
    // FP32 rows of tk: with an even OT the 16-column blocks of rows k with
    // bit 1 of k set are swapped in pairs (block J at J ^ 1), so that the
    // four rows a B-operand read touches fall in disjoint banks with no
    // padding; with an odd OT the row stride is 16 or 48 (mod 64) floats
    static constexpr bool SWAP = OT % 2 == 0;
    static constexpr int LD = SWAP ? NO : NO + ((16 - NO % 32) + 32) % 32;
    static constexpr int HS = NKH + 8;                        // half row stride of th/tl (16-byte rows)
};

template <int KT, int OT, bool MIX>
struct TileLds {
    using G = TileGeom<KT, OT, MIX>;
    // FP32: tk[k][tk_col(k, e1)] = T_cell[e1][e2], k = kpos32(e2) (the column
    // swizzle within blocks of 16 spreads the rows that the formation's
    // stores touch together over the banks; a B-operand read still takes 16
    // consecutive columns)
    alignas(16) float tk[MIX ? 1 : G::NK][MIX ? 1 : G::LD];
    // mixed: th/tl[e1][kpos16(e2)] = hi/lo(T_cell[e1][e2])
    alignas(16) _Float16 th[MIX ? G::NO : 1][MIX ? G::HS : 8];
    alignas(16) _Float16 tl[MIX ? G::NO : 1][MIX ? G::HS : 8];
    float wc[MIX ? G::NKH : G::NK];                           // w_cell[e2], 0 for e2 >= m
};

// position of member e in the summation order of the FP32 product (4-step
// MFMA): step s = 4 (e / 16) + e % 4 is taken by lane group kq = (e / 4) % 4
__device__ inline int kpos32(int e)
{
    return 16 * (e >> 4) + 4 * (e & 3) + ((e >> 2) & 3);
}

// ... of the half-precision product (32-step MFMA): lane group kq = (e / 4) % 4
// supplies 8 consecutive positions, members 32 t + 4 kq + c and
// 32 t + 16 + 4 kq + c
__device__ inline int kpos16(int e)
{
    return 32 * (e >> 5) + 8 * ((e >> 2) & 3) + 4 * ((e >> 4) & 1) + (e & 3);
}

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

// the four lanes of a field (col, col + 16, col + 32, col + 48) combine
// their partial sums, in lane order
__device__ __attribute__((always_inline)) inline double sum4(double p, int col)
{
    const double s0 = __shfl(p, col), s1 = __shfl(p, col + 16);
    const double s2 = __shfl(p, col + 32), s3 = __shfl(p, col + 48);
    return ((s0 + s1) + s2) + s3;
}

// a field (col) is flagged when any of its four lanes is
__device__ __attribute__((always_inline)) inline bool any4(bool b, int col)
{
    const unsigned long long m = __ballot(b);
    return ((m >> col) | (m >> (col + 16)) | (m >> (col + 32)) | (m >> (col + 48))) & 1ull;
}

// records of tile t: lane (col, kq) reads field 16 t + col (fields past the
// end re-read the last one) at members 16 s + 4 kq + c
template <int KT, bool VEC>
__device__ __attribute__((always_inline)) inline void load_tile(const UpdArgs& a, long long cell, int t, int col,
                                                                int kq, float (&x)[KT][4])
{
    const int f = min(16 * t + col, a.nfield - 1);
    const float* rec = a.E + ((long long)f * a.ncell + cell) * a.m;
#pragma unroll
    for (int s = 0; s < KT; ++s) {
        const int e = 16 * s + 4 * kq;
        if constexpr (VEC) {
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (16 * (s + 1) <= a.m || e < a.m)   // m % 4 == 0: the whole quad is in the record
                v = *reinterpret_cast<const float4*>(rec + e);
            x[s][0] = v.x;
            x[s][1] = v.y;
            x[s][2] = v.z;
            x[s][3] = v.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                x[s][c] = (e + c < a.m) ? rec[e + c] : 0.0f;
        }
    }
}

// column of T_cell row e1 in FP32 row k of tk
template <int KT, int OT, bool MIX>
__device__ inline int tk_col(int k, int e1)
{
    return e1 ^ (k >> 4) ^ (TileGeom<KT, OT, MIX>::SWAP ? 16 * ((k >> 1) & 1) : 0);
}

template <int KT, int OT, bool MIX>
__device__ inline void store_t(TileLds<KT, OT, MIX>& sh, int e1, int e2, float t)
{
    if constexpr (MIX) {
        const _Float16 h = (_Float16)t;
        sh.th[e1][kpos16(e2)] = h;
        sh.tl[e1][kpos16(e2)] = (_Float16)(t - (float)h);
    } else {
        const int k = kpos32(e2);
        sh.tk[k][tk_col<KT, OT, MIX>(k, e1)] = t;
    }
}

// four consecutive members e2 .. e2 + 3 (e2 % 4 == 0) of row e1
template <int KT, int OT, bool MIX>
__device__ inline void store_t4(TileLds<KT, OT, MIX>& sh, int e1, int e2, const float (&t)[4])
{
    if constexpr (MIX) {
        const int k = kpos16(e2);   // the four are consecutive positions
        half4 h, l;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            h[c] = (_Float16)t[c];
            l[c] = (_Float16)(t[c] - (float)h[c]);
        }
        *reinterpret_cast<half4*>(&sh.th[e1][k]) = h;
        *reinterpret_cast<half4*>(&sh.tl[e1][k]) = l;
    } else {
        const int k = kpos32(e2);   // positions k, k + 4, k + 8, k + 12 (one block of 16)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            sh.tk[k + 4 * c][tk_col<KT, OT, MIX>(k + 4 * c, e1)] = t[c];
    }
}

// one interpolated value in the cell's case: CASE bit 0 = interpolate in j,
// bit 1 = in i (update.c:224-271; the operations of bilin() for that case)
template <int CASE>
__device__ __attribute__((always_inline)) inline float interp(const Bilin& b, float v00, float v10, float v01,
                                                              float v11)
{
    if constexpr (CASE == 0)
        return v00;
    if constexpr (CASE == 1)
        return v00 * b.wj1 + v10 * b.wj2;
    if constexpr (CASE == 2)
        return v00 * b.wi1 + v01 * b.wi2;
    const float ta = v00 * b.wj1 + v10 * b.wj2;
    const float tb = v01 * b.wj1 + v11 * b.wj2;
    return ta * b.wi1 + tb * b.wi2;
}

// T_cell into LDS, loading only the corners the case uses
template <int KT, int OT, bool MIX, bool VEC, int CASE>
__device__ inline void form_t(const UpdArgs& a, const Bilin& b, TileLds<KT, OT, MIX>& sh, int tid)
{
    constexpr int NT = 64 * upd_waves<KT, OT, MIX>();
    constexpr bool J = (CASE & 1) != 0, I = (CASE & 2) != 0;
    const int m = a.m, md = a.md;
    if constexpr (VEC) {
        // quads of 4 consecutive e2: row e1 = q / (m / 4) by a multiply-high
        // (exact for these small q)
        constexpr int UB = 3;
        const int d = m >> 2;
        const unsigned magic = 0xffffffffu / (unsigned)d + 1u;
        const int n = md * d;
        for (int q0 = tid; q0 < n; q0 += NT * UB) {
            float4 v00[UB], v10[UB], v01[UB], v11[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int q = 4 * min(q0 + NT * u, n - 1);   // in range; duplicates are not stored
                v00[u] = *reinterpret_cast<const float4*>(a.T + b.o00 + q);
                if (J)
                    v10[u] = *reinterpret_cast<const float4*>(a.T + b.o10 + q);
                if (I)
                    v01[u] = *reinterpret_cast<const float4*>(a.T + b.o01 + q);
                if (J && I)
                    v11[u] = *reinterpret_cast<const float4*>(a.T + b.o11 + q);
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int q = q0 + NT * u;
                if (q < n) {
                    const int e1 = (int)__umulhi((unsigned)q, magic);
                    const int e2 = 4 * (q - e1 * d);
                    const float t[4] = {interp<CASE>(b, v00[u].x, v10[u].x, v01[u].x, v11[u].x),
                                        interp<CASE>(b, v00[u].y, v10[u].y, v01[u].y, v11[u].y),
                                        interp<CASE>(b, v00[u].z, v10[u].z, v01[u].z, v11[u].z),
                                        interp<CASE>(b, v00[u].w, v10[u].w, v01[u].w, v11[u].w)};
                    store_t4<KT, OT, MIX>(sh, e1, e2, t);
                }
            }
        }
    } else {
        constexpr int UB = 4;
        const unsigned magic = 0xffffffffu / (unsigned)m + 1u;
        const int n = md * m;
        for (int q0 = tid; q0 < n; q0 += NT * UB) {
            float v00[UB], v10[UB], v01[UB], v11[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int q = min(q0 + NT * u, n - 1);   // in range; duplicates are not stored
                v00[u] = a.T[b.o00 + q];
                v10[u] = J ? a.T[b.o10 + q] : 0.0f;
                v01[u] = I ? a.T[b.o01 + q] : 0.0f;
                v11[u] = (J && I) ? a.T[b.o11 + q] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int q = q0 + NT * u;
                if (q < n) {
                    const int e1 = (int)__umulhi((unsigned)q, magic), e2 = q - e1 * m;
                    store_t<KT, OT, MIX>(sh, e1, e2, interp<CASE>(b, v00[u], v10[u], v01[u], v11[u]));
                }
            }
        }
    }
}

// the log increment of an APPLYLOG variable (update.c:472-516), kept out of
// line: it is rare, and inlined it would take registers from the common path
__device__ __attribute__((noinline)) float log_increment(float va, float vf)
{
    return isnormal(vf) ? (float)(exp10((double)va) - exp10((double)vf)) : 0.0f;
}

// waves per SIMD the register budget allows (the LDS tiles of the larger
// ensembles allow fewer workgroups anyway)
template <int KT>
constexpr int tile_waves()
{
    return KT <= 6 ? 4 : (KT <= 8 ? 3 : 2);
}

// FULL: m == 16 KT and m_d == 16 OT, so that every member test is a
// compile-time constant (held per member, the tests would take two scalar
// registers each)
template <int KT, int OT, bool MIX, bool VEC, bool FULL>
__global__ void __launch_bounds__((64 * upd_waves<KT, OT, MIX>())) __attribute__((amdgpu_waves_per_eu(tile_waves<KT>())))
update_tile_kernel(UpdArgs a)
{
    using G = TileGeom<KT, OT, MIX>;
    __shared__ TileLds<KT, OT, MIX> sh;

    const long long cell = cell_of_block(a.ncell);
    if (cell >= a.ncell)
        return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const int jl = (int)(cell / a.mni);
    const int i = (int)(cell % a.mni);
    const int j = a.j1 + jl;
    const int m = a.m, md = a.md;
    // with m_d == m (FULL, OT == KT) the hybrid factor k_d is 1 and its
    // scaling the identity: left out at compile time
    const bool hybrid = (FULL && OT == KT) ? false : (a.mode == ENKFG_MODE_HYBRID);
    const double k_d = (md > 1) ? sqrt((double)(m - 1) / (double)(md - 1)) : 1.0;
    // the hybrid rescaling v /= k_d (update.c:420-427) as a product with the
    // reciprocal: the FP64 results differ by at most an ulp of FP64, so the
    // FP32 values stored agree except at rounding ties
    const double inv_kd = 1.0 / k_d;
    const int nl = a.nlevels[(long long)j * a.mni + i];
    const int ntiles = (a.nfield + 15) / 16;

    // the wave's first tile is loaded while T_cell is formed
    float x[KT][4];
    if (w < ntiles)
        load_tile<KT, VEC>(a, cell, w, col, kq, x);

    // ---- 1. T_cell and w_cell (zero outside md x m)
    {
        const Bilin b = make_bilin(a, j, i, (long long)md * m);
        const int ca = (b.jcopy ? 0 : 1) | (b.icopy ? 0 : 2);
        if (ca == 0)
            form_t<KT, OT, MIX, VEC, 0>(a, b, sh, tid);
        else if (ca == 1)
            form_t<KT, OT, MIX, VEC, 1>(a, b, sh, tid);
        else if (ca == 2)
            form_t<KT, OT, MIX, VEC, 2>(a, b, sh, tid);
        else
            form_t<KT, OT, MIX, VEC, 3>(a, b, sh, tid);
        // zero padding: members e2 in [m, NK(H)) of every output row, and output
        // rows e1 in [md, NO) of every member
        constexpr int NKX = MIX ? G::NKH : G::NK;
        if (m < NKX) {
            const int np = NKX - m;
            for (int q = tid; q < G::NO * np; q += 64 * upd_waves<KT, OT, MIX>())
                store_t<KT, OT, MIX>(sh, q / np, m + q % np, 0.0f);
        }
        if (md < G::NO) {
            const int nr = G::NO - md;
            for (int q = tid; q < nr * m; q += 64 * upd_waves<KT, OT, MIX>())
                store_t<KT, OT, MIX>(sh, md + q / m, q % m, 0.0f);
        }
        const Bilin bw = make_bilin(a, j, i, m);
        for (int e = tid; e < NKX; e += 64 * upd_waves<KT, OT, MIX>()) {
            const int ec = min(e, m - 1);
            const float v = bilin(bw, a.w[bw.o00 + ec], a.w[bw.o10 + ec], a.w[bw.o01 + ec], a.w[bw.o11 + ec]);
            sh.wc[e] = (e < m) ? v : 0.0f;
        }
    }
    __syncthreads();

    for (int t = w; t < ntiles; t += upd_waves<KT, OT, MIX>()) {
        if (t != w)
            load_tile<KT, VEC>(a, cell, t, col, kq, x);

        // ---- 2. statistics of field 16 t + col over the lane's members
        const int fs = 16 * t + col;
        const bool live = fs < a.nfield;
        const int fsc = min(fs, a.nfield - 1);
        // a member is below m / m_d when its block of 16 is (a uniform test;
        // always, for FULL), or by a per-lane test in the block that m / m_d
        // cuts
        bool badd = false, bads = false;
        double p = 0.0;
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int e = 16 * s + 4 * kq + c;
                const bool inm = FULL || (16 * (s + 1) <= m) || e < m;
                const bool inmd = (s < OT) && (FULL || (16 * (s + 1) <= md) || e < md);
                const float v = x[s][c];
                const bool bad = !(fabsf(v) <= kMaxVal);   // not finite, or beyond MAXOBSVAL
                if (s < OT)
                    badd = badd || (bad && inmd);
                if (s >= OT - 1)
                    bads = bads || (bad && inm && !inmd);
                if (s < OT)
                    p += inmd ? (double)v : 0.0;
            }
        const bool bad_d = any4(badd, col), bad_s = any4(bads, col);
        int z = 0;   // 0 update, 1 all members 0, 2 static members 0
        if (!live || bad_d || is_masked(nl, a.flevel[fsc], a.fsurf[fsc], a.surfk))
            z = 1;
        else if (bad_s)
            z = 2;
        const bool z2 = (z == 2);
        const double v1_f = sum4(p, col) / (double)md;
        // anomalies (a field with z = 1 is written as zeros whatever they are;
        // its row of the product involves no other field)
        double pa = 0.0, pf = 0.0;
        float amax = 0.0f;
#pragma unroll
        for (int s = 0; s < KT; ++s)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int e = 16 * s + 4 * kq + c;
                const bool inm = FULL || (16 * (s + 1) <= m) || e < m;
                const bool inmd = (s < OT) && (FULL || (16 * (s + 1) <= md) || e < md);
                // a static member of a field with a bad static member is 0
                // before the mean is taken off (update.c:337-371)
                float xin = x[s][c];
                if (s >= OT - 1)
                    xin = (z2 && !inmd) ? 0.0f : xin;
                float vf = (float)((double)xin - v1_f);
                if (s == KT - 1)
                    vf = inm ? vf : 0.0f;
                x[s][c] = vf;
                pa += (double)(vf * sh.wc[e]);   // wc is 0 past m
                if (s < OT) {
                    const float vh = hybrid ? (float)((double)vf * inv_kd) : vf;
                    pf += inmd ? (double)(vh * vh) : 0.0;
                }
                if (MIX)
                    amax = fmaxf(amax, fabsf(vf));
            }
        const double v1_a = sum4(pa, col) + v1_f;
        const double v2_f = sum4(pf, col);
        float scale = 1.0f;
        if constexpr (MIX) {
            amax = fmaxf(amax, __shfl_xor(amax, 16));
            amax = fmaxf(amax, __shfl_xor(amax, 32));
            // power of two that brings the largest anomaly to [2^9, 2^10) (FP16
            // normal range for the split; exact scaling)
            if (amax > 0.0f && isfinite(amax))
                scale = ldexpf(1.0f, max(-100, min(100, 9 - ilogbf(amax))));
        }

        // ---- 3. V^a for the tile's 16 fields x 16 OT members; the LDS
        // offset of the lane's B values is opaque to the compiler, so that
        // T_cell is read per tile rather than held in registers
        flt4 acc[OT];
#pragma unroll
        for (int J = 0; J < OT; ++J)
            acc[J] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
        int boff = col;
        asm volatile("" : "+v"(boff));
        if constexpr (!MIX) {
            // block J of the lane's row at 16 (J ^ f), f = bit 1 of kq when
            // the blocks are swapped (tk_col)
            const int f = G::SWAP ? (kq >> 1) & 1 : 0;
#pragma unroll
            for (int s = 0; s < KT; ++s)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float* brow = &sh.tk[16 * s + 4 * c + kq][0] + (boff ^ s);
                    const float* beven = brow + 16 * f;   // blocks J even: 16 J + 16 f
                    const float* bodd = brow - 16 * f;    // J odd: 16 J - 16 f
#pragma unroll
                    for (int J = 0; J < OT; ++J)
                        acc[J] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[s][c], (J & 1) ? bodd[16 * J] : beven[16 * J],
                                                                      acc[J], 0, 0, 0);
                }
        } else {
#pragma unroll
            for (int tt = 0; tt < (KT + 1) / 2; ++tt) {
                half8 ah, al;
#pragma unroll
                for (int pp = 0; pp < 2; ++pp)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int s = 2 * tt + pp;
                        const float xs = (s < KT) ? x[s < KT ? s : 0][c] * scale : 0.0f;
                        const _Float16 h = (_Float16)xs;
                        ah[4 * pp + c] = h;
                        al[4 * pp + c] = (_Float16)(xs - (float)h);
                    }
#pragma unroll
                for (int J = 0; J < OT; ++J) {
                    const half8 bh = *reinterpret_cast<const half8*>(&sh.th[16 * J][32 * tt + 8 * kq] + boff * G::HS);
                    const half8 bl = *reinterpret_cast<const half8*>(&sh.tl[16 * J][32 * tt + 8 * kq] + boff * G::HS);
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[J], 0, 0, 0);
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[J], 0, 0, 0);
                    acc[J] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[J], 0, 0, 0);
                }
            }
        }

        // ---- 4. result (field 16 t + 4 kq + r, member 16 J + col) in
        // acc[J][r]; the field's statistics from its stats lane (4 kq + r)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int fl = 4 * kq + r;
            const int fo = 16 * t + fl;
            const int fz = __shfl(z, fl);
            const double f1f = __shfl(v1_f, fl), f1a = __shfl(v1_a, fl), f2f = __shfl(v2_f, fl);
            const float unscale = MIX ? 1.0f / __shfl(scale, fl) : 1.0f;
            float va[OT];
            double v2a = 0.0;
#pragma unroll
            for (int J = 0; J < OT; ++J) {
                const int e = 16 * J + col;
                float v = acc[J][r];
                if (MIX)
                    v *= unscale;
                if (hybrid)
                    v = (float)((double)v * inv_kd);
                va[J] = v;
                const bool in = FULL || (16 * (J + 1) <= md) || e < md;
                v2a += in ? (double)(v * v) : 0.0;
            }
            v2a = row16_sum(v2a);
            if (fo >= a.nfield)
                continue;
            const float inflation0 = a.finfl[fo];
            const double inf_ratio = a.finfratio[fo];
            float inflation = inflation0;
            if (!isnan(inf_ratio) && v2a > 0.0) {
                inflation = (float)(sqrt(f2f / v2a) * inf_ratio + 1.0 - inf_ratio);
                if (inflation >= inflation0)
                    inflation = inflation0;
            }
            if (!(inflation - 1.0f > kEps))
                inflation = 1.0f;
            float* rec = a.E + ((long long)fo * a.ncell + cell) * m;
            if (fz == 1) {   // masked or a dynamic member missing: all members 0
                for (int e = col; e < m; e += 16)
                    rec[e] = 0.0f;
            } else {
                if (fz == 2)   // a static member missing: the static members 0
                    for (int e = md + col; e < m; e += 16)
                        rec[e] = 0.0f;
                const bool outinc = a.outputinc != 0;
                const bool applylog = outinc && a.fapplylog[fo] != 0;
#pragma unroll
                for (int J = 0; J < OT; ++J) {
                    const int e = 16 * J + col;
                    if (FULL || 16 * (J + 1) <= md || e < md) {
                        float v = va[J];
                        if (inflation != 1.0f)
                            v = v * inflation;
                        v = (float)((double)v + f1a);
                        if (outinc) {
                            float vf = (float)((double)rec[e] - f1f);
                            if (hybrid)
                                vf = (float)((double)vf * inv_kd);
                            vf = (float)((double)vf + f1f);
                            v = applylog ? log_increment(v, vf) : v - vf;
                        }
                        rec[e] = v;
                    }
                }
            }
            if (a.infl != nullptr && col == 0)
                a.infl[(long long)fo * a.ncell + cell] = (fz == 1) ? 0.0f : inflation;
        }
    }
}

template <int KT, int OT, bool MIX>
void launch_ktm(hipStream_t s, const UpdArgs& a, dim3 grid)
{
    const dim3 block(64 * upd_waves<KT, OT, MIX>());
    if (a.m == 16 * KT && a.md == 16 * OT)
        hipLaunchKernelGGL((update_tile_kernel<KT, OT, MIX, true, true>), grid, block, 0, s, a);
    else if (a.m % 4 == 0)
        hipLaunchKernelGGL((update_tile_kernel<KT, OT, MIX, true, false>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((update_tile_kernel<KT, OT, MIX, false, false>), grid, block, 0, s, a);
}

template <int KT, int OT>
hipError_t launch_kt(hipStream_t s, const UpdArgs& a, bool mixed, dim3 grid)
{
    if (mixed)
        launch_ktm<KT, OT, true>(s, a, grid);
    else
        launch_ktm<KT, OT, false>(s, a, grid);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_update_mfma(hipStream_t s, const UpdArgs& a, bool mixed)
{
    if (a.ncell == 0 || a.nfield == 0)
        return hipSuccess;
    if (a.m < 1 || a.m > 160 || a.md < 1 || a.md > a.m)
        return hipErrorInvalidValue;
    const int KT = (a.m + 15) / 16, OT = (a.md + 15) / 16;
    const long long chunk = (a.ncell + 7) / 8;
    if (8 * chunk > 0x7fffffffLL)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)(8 * chunk));
    // the output tiles cover m_d: one kernel per KT with OT = KT, and the
    // hybrid 96 + 48 shape; any other hybrid shape computes the zero rows of
    // T past m_d in the OT = KT kernel
    if (KT == 9 && OT == 6)
        return launch_kt<9, 6>(s, a, mixed, grid);
    switch (KT) {
    case 1: return launch_kt<1, 1>(s, a, mixed, grid);
    case 2: return launch_kt<2, 2>(s, a, mixed, grid);
    case 3: return launch_kt<3, 3>(s, a, mixed, grid);
    case 4: return launch_kt<4, 4>(s, a, mixed, grid);
    case 5: return launch_kt<5, 5>(s, a, mixed, grid);
    case 6: return launch_kt<6, 6>(s, a, mixed, grid);
    case 7: return launch_kt<7, 7>(s, a, mixed, grid);
    case 8: return launch_kt<8, 8>(s, a, mixed, grid);
    case 9: return launch_kt<9, 9>(s, a, mixed, grid);
    case 10: return launch_kt<10, 10>(s, a, mixed, grid);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace enkfg
