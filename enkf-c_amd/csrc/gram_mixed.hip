// gram_mixed.hip -- G0 = S^T S and S^T s of the local analyses in mixed
// precision: BF16 MFMA with FP32 accumulation on a hi/lo split of the tapered
// anomalies.
//
// Each tapered value x = S * (float) lc is split as x = hi + lo + r with
// hi = bf16(x), lo = bf16(x - hi) and |r| <= 2^-18 |x|; the product x_a x_b
// is taken as hi_a hi_b + hi_a lo_b + lo_a hi_b, which leaves out lo_a lo_b
// (<= 2^-18 relative), so every term of S^T S carries a relative error of
// about 2^-17 before the FP32 accumulation.  S^T s accumulates in FP64 from
// hi + lo.  The result is written in the FP64 Gram layout and the
// factorisation that follows stays in FP64 (factor_wave.hip).
//
// Work items, chunk pipeline and XCD placement are those of the FP64 item
// Gram kernel (calc_kernels.hip, gram_stage.h, gram_layout.h).
#include <algorithm>

#include "enkfg_internal.h"
#include "gram_layout.h"
#include "gram_stage.h"

namespace enkfg {

namespace {

typedef float flt4 __attribute__((__vector_size__(4 * sizeof(float))));
typedef __bf16 bf16x4 __attribute__((__vector_size__(4 * sizeof(__bf16))));

// ------------------------------------------------------------------ one wavefront per work item
//
// For m <= 160: independent wavefronts, no workgroup barrier.  A wavefront
// computes one of NG groups of the tri_count(MT) tiles (f = g, g + NG, ...;
// NG = 2: 11 and 10 tiles for m = 96, NG = 3: 7 each, FP32 accumulators;
// NG is chosen per obs-count bucket by the host's autotuner) of one (node,
// obstype segment).  The segment's list entries and S rows are copied into
// the wave's private LDS area in chunks of kChunk = 16 local obs, MKB chunks
// deep, with no VGPRs held for data in flight (gram_stage.h, the same
// pipeline as the FP64 item kernel; the other waves of a segment find the
// rows in L2).  A chunk is one K = 16 step of v_mfma_f32_16x16x16_bf16: lane
// (col, kg) forms its operand of block X -- obs 4 kg .. 4 kg + 3 of member
// 16 X + col -- from four FP32 values of the block-major rows, x = S *
// (float) lc, split into BF16 hi = bf16(x) and lo = bf16(x - hi) (the split
// carries about 2^-17 relative error); each element is converted once per
// wave.  S^T s is summed in FP32 over a chunk and in FP64 over chunks.
// Against the earlier kernel, which gathered one chunk ahead into registers
// and staged BF16 planes in LDS, the results are the same bit for bit.
constexpr int MKB = 3;
static_assert(kListPad >= (MKB + 1) * kChunk, "list padding too small for the mixed item Gram pipeline");
template <int MT>
using MixItemLds = ChunkLds<MT, MKB>;
typedef short sh4 __attribute__((ext_vector_type(4)));
typedef float flt2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

template <int MT, int NG, int G>
constexpr int mix_item_ntile()
{
    return (tri_count(MT) - G + NG - 1) / NG;
}

// the tiles f = G + NG t of group G
template <int MT, int NG, int G, int T>
__device__ __attribute__((always_inline)) inline void mix_item_tiles(const sh4 (&h)[MT], const sh4 (&l)[MT],
                                                                     flt4* acc)
{
    if constexpr (G + NG * T < tri_count(MT)) {
        constexpr int f = G + NG * T;
        constexpr int I = tri_I(MT, f), J = tri_J(MT, f);
        acc[T] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h[I], h[J], acc[T], 0, 0, 0);
        acc[T] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(h[I], l[J], acc[T], 0, 0, 0);
        acc[T] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(l[I], h[J], acc[T], 0, 0, 0);
        mix_item_tiles<MT, NG, G, T + 1>(h, l, acc);
    }
}

template <int MT, int NG, int G>
__device__ __attribute__((always_inline)) inline void mix_item_wave(const CalcArgs& a, int b, long long off, int cnt,
                                                                    MixItemLds<MT>& sh)
{
    constexpr int MP = 16 * MT;
    constexpr int NTG = mix_item_ntile<MT, NG, G>();
    const int lane = threadIdx.x & 63;
    const int kg = lane >> 4;      // which 4 obs of the K = 16 step
    const int col = lane & 15;

    flt4 acc[NTG > 0 ? NTG : 1];
#pragma unroll
    for (int f = 0; f < NTG; ++f)
        acc[f] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
    double sts[MT];
#pragma unroll
    for (int X = 0; X < MT; ++X)
        sts[X] = 0.0;

    const int nchunk = (cnt + kChunk - 1) / kChunk;
    chunk_prologue<MT, MKB>(a, sh, off, cnt);
    int buf = 0;    // c % MKB
    int slot = 0;   // c % (MKB + 1)
    const int k0 = 4 * kg;
    for (int c = 0; c < nchunk; ++c) {
        chunk_advance<MT, MKB>(a, sh, off, cnt, c, buf, slot);
        const double* svc = reinterpret_cast<const double*>(&sh.sv[buf][0]);
        const double* lcc = reinterpret_cast<const double*>(&sh.meta[slot][16]);
        // tapers of the lane's 4 obs (0 past the list's end) and s * lc
        float lcf[4], slv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double lr = lcc[k0 + q];
            const double l = c * kChunk + k0 + q < cnt ? lr : 0.0;
            lcf[q] = (float)l;
            slv[q] = (float)(svc[k0 + q] * l);
        }
        // operand of v_mfma_f32_16x16x16_bf16: lane l supplies 4 consecutive
        // obs (k = 4 (l / 16) ..) of member row l % 16 of a tile
        // (pairs converted together: v_cvt_pk_bf16_f32)
        sh4 h[MT], l[MT];
#pragma unroll
        for (int X = 0; X < MT; ++X) {
            flt2v x[2], r[2];
            bf16x2v hp[2], lp[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                x[p][0] = sh.rows[buf][X][k0 + 2 * p][col] * lcf[2 * p];
                x[p][1] = sh.rows[buf][X][k0 + 2 * p + 1][col] * lcf[2 * p + 1];
                hp[p] = __builtin_convertvector(x[p], bf16x2v);
                r[p] = x[p] - __builtin_convertvector(hp[p], flt2v);
                lp[p] = __builtin_convertvector(r[p], bf16x2v);
            }
            const bf16x4 hb = {hp[0][0], hp[0][1], hp[1][0], hp[1][1]};
            const bf16x4 lb = {lp[0][0], lp[0][1], lp[1][0], lp[1][1]};
            h[X] = __builtin_bit_cast(sh4, hb);
            l[X] = __builtin_bit_cast(sh4, lb);
        }
        mix_item_tiles<MT, NG, G, 0>(h, l, acc);
        if (G == 0) {
#pragma unroll
            for (int X = 0; X < MT; ++X) {
                const bf16x4 hb = __builtin_bit_cast(bf16x4, h[X]);
                const bf16x4 lb = __builtin_bit_cast(bf16x4, l[X]);
                float t = 0.0f;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    t = fmaf((float)hb[q] + (float)lb[q], slv[q], t);
                sts[X] += (double)t;
            }
        }
        chunk_fence();
        buf = buf == MKB - 1 ? 0 : buf + 1;
        slot = slot == MKB ? 0 : slot + 1;
    }
    // the copies of the chunks past the end land before the wave ends
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // v_mfma_f32_16x16x16_bf16 result: register i of lane l holds
    // D[4 (l / 16) + i][l % 16]
    double* gout = a.gram + (size_t)b * MP * MP;
#pragma unroll
    for (int t = 0; t < NTG; ++t) {
        const int f = G + NG * t;
        const int I = tri_I(MT, f), J = tri_J(MT, f);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * I + 4 * kg + i;
            const int cc = 16 * J + col;
            const double v = (double)acc[t][i];
            gout[gram_index<MT>(r, cc)] = v;
            if (I != J)
                gout[gram_index<MT>(cc, r)] = v;
        }
    }
    if (G == 0) {
        // the four lanes of a member (kg = 0..3) hold partial sums over
        // their 4 obs of each chunk
        double* stsout = a.sts + (size_t)b * MP;
#pragma unroll
        for (int X = 0; X < MT; ++X) {
            double v = sts[X];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (kg == 0)
                stsout[16 * X + col] = v;
        }
    }
}

template <int MT, int NG, int G>
__device__ __attribute__((always_inline)) inline void mix_item_dispatch(int g, const CalcArgs& a, int b, long long off,
                                                                        int cnt, MixItemLds<MT>& sh)
{
    if constexpr (G < NG) {
        if (g == G) {
            if constexpr (mix_item_ntile<MT, NG, G>() > 0)
                mix_item_wave<MT, NG, G>(a, b, off, cnt, sh);
        } else {
            mix_item_dispatch<MT, NG, G + 1>(g, a, b, off, cnt, sh);
        }
    }
}

// item `it` = (segment b, group g) of the batch
template <int MT, int NG>
__device__ __attribute__((always_inline)) inline void mix_item_run(const CalcArgs& a, int it, MixItemLds<MT>& sh)
{
    const ItemRef r = item_ref(a, it, NG);
    const int g = r.g, seg = r.seg;
    const int b = r.bn * a.nseg + seg;   // the (node, segment) slot of the Gram scratch
    const int node = a.node_list ? a.node_list[a.node0 + r.bn] : a.node0 + r.bn;
    long long off;
    int cnt;
    if (a.pertype) {
        off = a.offsets[(size_t)node * a.ntypes + seg];
        cnt = a.counts[(size_t)node * a.ntypes + seg];
    } else {
        off = a.offsets[(size_t)node * a.ntypes];
        cnt = (int)(a.offsets[(size_t)(node + 1) * a.ntypes] - off);
    }
    if (cnt > 0)
        mix_item_dispatch<MT, NG, 0>(g, a, b, off, cnt, sh);
}

template <int MT, int NG>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) gram_mixed_item_kernel(CalcArgs a)
{
    // one item per wave.  Blocks are handed out as earlier ones finish (the
    // hardware dispatcher is the work queue); each XCD takes runs of
    // kItemRun consecutive (segment, group) items, the NG groups of a
    // segment adjacent, so neighbouring nodes and the groups of one segment
    // share that XCD's L2
    __shared__ MixItemLds<MT> sh;
    const int nitem = a.nbatch * a.nseg * NG;
    const int it = item_of_block();
    if (it < nitem)
        mix_item_run<MT, NG>(a, it, sh);
}

}  // namespace

template <int M>
void launch_mix_item(hipStream_t s, dim3 grid, const CalcArgs& a, int ng)
{
    constexpr int lo = item_groups_lo(M);
    if (ng == lo + 1)
        hipLaunchKernelGGL((gram_mixed_item_kernel<M, lo + 1>), grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((gram_mixed_item_kernel<M, lo>), grid, dim3(64), 0, s, a);
}

hipError_t launch_gram_mixed(hipStream_t s, const CalcArgs& a)
{
    const int MT = (a.m + 15) / 16;
    if (a.nbatch == 0)
        return hipSuccess;
    if (MT >= 1 && MT <= 10) {
        // tile groups per segment: one of the two candidates of this MT (the
        // smaller by default)
        const int lo = item_groups_lo(MT);
        const int ng = (a.tile_groups == lo + 1) ? lo + 1 : lo;
        const long long nitem = (long long)a.nbatch * a.nseg * ng;
        const dim3 igrid(item_grid(nitem));
        switch (MT) {
        case 1: launch_mix_item<1>(s, igrid, a, ng); break;
        case 2: launch_mix_item<2>(s, igrid, a, ng); break;
        case 3: launch_mix_item<3>(s, igrid, a, ng); break;
        case 4: launch_mix_item<4>(s, igrid, a, ng); break;
        case 5: launch_mix_item<5>(s, igrid, a, ng); break;
        case 6: launch_mix_item<6>(s, igrid, a, ng); break;
        case 7: launch_mix_item<7>(s, igrid, a, ng); break;
        case 8: launch_mix_item<8>(s, igrid, a, ng); break;
        case 9: launch_mix_item<9>(s, igrid, a, ng); break;
        default: launch_mix_item<10>(s, igrid, a, ng); break;
        }
        return hipGetLastError();
    }
    return hipErrorInvalidValue;   // m > 160 is refused at configuration
}

}  // namespace enkfg
