// gram_mixed.hip -- G0 = S^T S and S^T s of the local analyses in mixed
// precision: BF16 MFMA with FP32 accumulation on a hi/lo split of the tapered
// anomalies.
//
// Each tapered value x = (float)((double)S * lc) is split as x = hi + lo + r
// with hi = bf16(x), lo = bf16(x - hi) and |r| <= 2^-18 |x|; the product
// x_a x_b is taken as hi_a hi_b + hi_a lo_b + lo_a hi_b, which leaves out
// lo_a lo_b (<= 2^-18 relative), so every term of S^T S carries a relative
// error of about 2^-17 before the FP32 accumulation.  S^T s accumulates in
// FP64 from hi + lo.  The result is written in the FP64 Gram layout and the
// factorisation that follows stays in FP64 (calc_kernels.hip).
//
// Work item, chunk pipeline and XCD-aware mapping are those of the FP64 Gram
// kernel: one workgroup of 4 waves per (node, obstype segment), local obs
// streamed in chunks of 64 (the gathers of chunk c+1 are in flight while chunk
// c feeds the MFMAs).  A chunk is stored in LDS member-major, 64 obs of one
// member contiguous as bf16, so that an operand of
// v_mfma_f32_16x16x32_bf16 -- 8 consecutive obs of one member -- is one
// 16-byte LDS read; the gathers are arranged so that each lane holds 4
// consecutive obs of one member and stores them as one 8-byte write.
#include <algorithm>

#include "enkfg_internal.h"
#include "gram_layout.h"
#include "gram_stage.h"

namespace enkfg {

namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float flt4 __attribute__((__vector_size__(4 * sizeof(float))));

constexpr int KC = kGramKC;    // obs per chunk (two K = 32 MFMA steps)
constexpr int KR = KC + 8;     // bf16 per member row in LDS (144 B: 16-byte aligned, staggered banks)

template <int MT>
struct MixLds {
    static constexpr int MP = 16 * MT;
    alignas(16) __bf16 hi[2][MP][KR];
    alignas(16) __bf16 lo[2][MP][KR];
    double sl[2][KC];          // s_o * lc_o
    double lc[3][KC];
    int id[3][KC];
};

template <int MT>
__device__ inline void meta_store(MixLds<MT>& sh, int buf, int id, double c)
{
    if (threadIdx.x < KC) {
        sh.id[buf][threadIdx.x] = id;
        sh.lc[buf][threadIdx.x] = c;
    }
}

typedef __bf16 bf16x4 __attribute__((__vector_size__(4 * sizeof(__bf16))));

// One chunk in registers, member-contiguous: item u = tid + 256 j holds
// member u % MP of rows 4 (u / MP) .. 4 (u / MP) + 3, so a wave's loads of
// one row are coalesced and its LDS stores are 4 consecutive obs (8 bytes)
// of one member.  Loads are unconditional (zero row for ids -1, the last
// member for padding members, which are not stored).
template <int MT>
struct MixRegs {
    float v[MT][4];
    double s;                  // s of row tid % KC
};

template <int MT>
__device__ inline void mix_gather(const CalcArgs& a, const int* ids_lds, MixRegs<MT>& L)
{
    constexpr int MP = 16 * MT;
#pragma unroll
    for (int j = 0; j < MT; ++j) {
        const int u = (int)threadIdx.x + 256 * j;
        const int rg = u / MP;               // < KC / 4 since 256 MT = 16 MP
        const int e = min(u - rg * MP, a.m - 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int id = ids_lds[4 * rg + k];
            const long long r = id >= 0 ? id : a.nobs;
            L.v[j][k] = a.S[r * a.ms + e];
        }
    }
    const int id = ids_lds[threadIdx.x & (KC - 1)];
    L.s = a.sv[id >= 0 ? id : a.nobs];
}

// taper, split, store 4 obs of one member into the member-major rows of
// buffer rbuf
template <int MT>
__device__ inline void vals_store(const CalcArgs& a, MixLds<MT>& sh, int rbuf, int mbuf, const MixRegs<MT>& L)
{
    constexpr int MP = 16 * MT;
#pragma unroll
    for (int j = 0; j < MT; ++j) {
        const int u = (int)threadIdx.x + 256 * j;
        const int rg = u / MP;
        const int e = u - rg * MP;
        if (e < a.m) {
            bf16x4 h, l;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float x = (float)((double)L.v[j][k] * sh.lc[mbuf][4 * rg + k]);
                h[k] = (__bf16)x;
                l[k] = (__bf16)(x - (float)h[k]);
            }
            *reinterpret_cast<bf16x4*>(&sh.hi[rbuf][e][4 * rg]) = h;
            *reinterpret_cast<bf16x4*>(&sh.lo[rbuf][e][4 * rg]) = l;
        }
    }
    if (threadIdx.x < KC)
        sh.sl[rbuf][threadIdx.x] = L.s * sh.lc[mbuf][threadIdx.x];
}

// the three split products for the tiles (I, J) of wave W
template <int MT, int W, int T>
__device__ inline void mix_tiles(const bf16x8 (&h)[MT], const bf16x8 (&l)[MT], flt4* acc)
{
    if constexpr (W + 4 * T < tri_count(MT)) {
        constexpr int f = W + 4 * T;
        constexpr int I = tri_I(MT, f), J = tri_J(MT, f);
        acc[T] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h[I], h[J], acc[T], 0, 0, 0);
        acc[T] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h[I], l[J], acc[T], 0, 0, 0);
        acc[T] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l[I], h[J], acc[T], 0, 0, 0);
        mix_tiles<MT, W, T + 1>(h, l, acc);
    }
}

template <int MT>
constexpr int ntile_of(int W)
{
    return (tri_count(MT) - W + 3) / 4 > 0 ? (tri_count(MT) - W + 3) / 4 : 1;
}

template <int MT, int W>
__device__ inline void mix_wave_chunk(const MixLds<MT>& sh, int buf, flt4 (&acc)[ntile_of<MT>(W)],
                                      double (&sts)[(MT + 3) / 4])
{
    const int lane = threadIdx.x & 63;
    const int kg = lane >> 4;      // which 4 obs of the K = 16 step
    const int col = lane & 15;
#pragma unroll
    for (int ks = 0; ks < KC / 32; ++ks) {
        const int k0 = 32 * ks + 8 * kg;
        bf16x8 h[MT], l[MT];
#pragma unroll
        for (int X = 0; X < MT; ++X) {
            h[X] = *reinterpret_cast<const bf16x8*>(&sh.hi[buf][16 * X + col][k0]);
            l[X] = *reinterpret_cast<const bf16x8*>(&sh.lo[buf][16 * X + col][k0]);
        }
        mix_tiles<MT, W, 0>(h, l, acc);
        // S^T s for the member tiles X = W, W + 4, ... (FP64)
#pragma unroll
        for (int u = 0; u < (MT + 3) / 4; ++u) {
            const int X = W + 4 * u;
            if (X < MT) {
                double t = 0.0;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    t += ((double)(float)h[X][q] + (double)(float)l[X][q]) * sh.sl[buf][k0 + q];
                sts[u] += t;
            }
        }
    }
}

template <int MT, int W>
__device__ inline void mix_wave(const CalcArgs& a, long long off, int cnt, double* __restrict__ gout,
                                double* __restrict__ stsout, MixLds<MT>& sh)
{
    constexpr int NTW = ntile_of<MT>(W);
    constexpr int NU = (MT + 3) / 4;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int kg = lane >> 4;
    const int col = lane & 15;

    flt4 acc[NTW];
#pragma unroll
    for (int t = 0; t < NTW; ++t)
        acc[t] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
    double sts[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u)
        sts[u] = 0.0;

    // zero both buffers once: members >= m stay zero
    {
        unsigned int* z = reinterpret_cast<unsigned int*>(&sh.hi[0][0][0]);
        const int nw = (int)(2 * sizeof(sh.hi) / sizeof(unsigned int));   // hi and lo are adjacent
        for (int q = tid; q < nw; q += 256)
            z[q] = 0u;
    }
    MixRegs<MT> L;
    const int nchunk = (cnt + KC - 1) / KC;
    int mid;
    double mc;
    gram_meta_load<KC>(a, off, cnt, 0, mid, mc);
    meta_store<MT>(sh, 0, mid, mc);
    gram_meta_load<KC>(a, off, cnt, KC, mid, mc);
    meta_store<MT>(sh, 1, mid, mc);
    meta_store<MT>(sh, 2, -1, 0.0);   // every id buffer holds valid rows or -1
    __syncthreads();
    mix_gather<MT>(a, sh.id[0], L);
    vals_store<MT>(a, sh, 0, 0, L);
    __syncthreads();
    for (int c = 0; c < nchunk; ++c) {
        const bool more = c + 1 < nchunk;
        const bool more2 = c + 2 < nchunk;
        // unconditional loads (list padding, zero row), see gram_layout.h
        mix_gather<MT>(a, sh.id[(c + 1) % 3], L);
        gram_meta_load<KC>(a, off, cnt, (c + 2) * KC, mid, mc);
        __builtin_amdgcn_sched_barrier(0);   // loads stay ahead of the MFMAs
        mix_wave_chunk<MT, W>(sh, c & 1, acc, sts);
        if (more)
            vals_store<MT>(a, sh, (c + 1) & 1, (c + 1) % 3, L);
        if (more2)
            meta_store<MT>(sh, (c + 2) % 3, mid, mc);
        __syncthreads();
    }
    // v_mfma_f32_16x16x32_bf16 result: register i of lane l holds
    // D[4 (l / 16) + i][l % 16]
#pragma unroll
    for (int t = 0; t < NTW; ++t) {
        const int f = W + 4 * t;
        if (f >= tri_count(MT))
            break;
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
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int X = W + 4 * u;
        double v = sts[u];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        if (X < MT && kg == 0)
            stsout[16 * X + col] = v;
    }
}

template <int MT>
__global__ void __launch_bounds__(256) gram_mixed_kernel(CalcArgs a)
{
    constexpr int MP = 16 * MT;
    const int nwork = a.nbatch * a.nseg;
    const int per = (nwork + 7) / 8;
    const int b = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (b >= nwork)
        return;
    const int node = a.node_list ? a.node_list[a.node0 + b / a.nseg] : a.node0 + b / a.nseg;
    const int seg = b % a.nseg;
    long long off;
    int cnt;
    if (a.pertype) {
        off = a.offsets[(size_t)node * a.ntypes + seg];
        cnt = a.counts[(size_t)node * a.ntypes + seg];
    } else {
        off = a.offsets[(size_t)node * a.ntypes];
        cnt = (int)(a.offsets[(size_t)(node + 1) * a.ntypes] - off);
    }
    if (cnt == 0)
        return;
    __shared__ MixLds<MT> sh;
    double* gout = a.gram + (size_t)b * MP * MP;
    double* stsout = a.sts + (size_t)b * MP;
    switch (threadIdx.x >> 6) {
    case 0: mix_wave<MT, 0>(a, off, cnt, gout, stsout, sh); break;
    case 1: mix_wave<MT, 1>(a, off, cnt, gout, stsout, sh); break;
    case 2: mix_wave<MT, 2>(a, off, cnt, gout, stsout, sh); break;
    default: mix_wave<MT, 3>(a, off, cnt, gout, stsout, sh); break;
    }
}

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
    dim3 grid((unsigned)(8 * ((a.nbatch * a.nseg + 7) / 8)));
    if (a.nbatch == 0)
        return hipSuccess;
    // the item kernel addresses S with 32-bit byte offsets
    const bool small_s = (unsigned long long)(a.nobs + 1) * (unsigned long long)a.ms * 4ull < (1ull << 32);
    if (MT <= 10 && small_s) {
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
    switch (MT) {
    case 1: hipLaunchKernelGGL(gram_mixed_kernel<1>, grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(gram_mixed_kernel<2>, grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(gram_mixed_kernel<3>, grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(gram_mixed_kernel<4>, grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL(gram_mixed_kernel<5>, grid, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(gram_mixed_kernel<6>, grid, dim3(256), 0, s, a); break;
    case 7: hipLaunchKernelGGL(gram_mixed_kernel<7>, grid, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL(gram_mixed_kernel<8>, grid, dim3(256), 0, s, a); break;
    case 9: hipLaunchKernelGGL(gram_mixed_kernel<9>, grid, dim3(256), 0, s, a); break;
    case 10: hipLaunchKernelGGL(gram_mixed_kernel<10>, grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace enkfg
