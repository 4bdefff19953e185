// calc_kernels.hip -- per-column ensemble transforms for the FP64 path.
//
// For each stride-grid node the reference (calc/transforms.c:794-879 and
// calc/calcs.c) gathers the local standardised anomalies S (p x m, tapered,
// FP64), forms M = I + S^T S, factorises it and builds G = M^-1 S^T (m x p),
// T and w.  Here the p x m matrices are never formed:
//
//   gram   (MFMA FP64)  G0 = S^T S and S^T s, streamed from the gathered
//                       FP32 rows of S, tapered in FP64 on their way to LDS;
//   factor (LDS/VGPR)   for ETKF with m > 96 (m <= 96, and DEnKF and EnOI:
//                       factor_wave.hip), M = I + G0 inverted by in-register
//                       Gauss-Jordan, then
//                         T = M^-1/2                (= I - (M+M^1/2)^-1 S^T S)
//                         w = M^-1 S^T s,  dfs = tr(M^-1 G0) (= tr(G S)),
//                         srf = sqrt(tr(G0) / dfs) - 1
//                       followed by relaxation to prior and the FP32 cast.
//   M^-1/2 comes from the scaled, coupled Denman-Beavers iteration
//   (Y0 = M, Z0 = I, Y <- (mu Y + Z^-1/mu)/2, Z <- (mu Z + Y^-1/mu)/2, with
//   determinant scaling mu = |det Y det Z|^(-1/(2m)); Higham, "Functions of
//   Matrices", SIAM 2008, eq. 6.15 and sec. 6.3).
//
// The identities used are those of the User Guide (doc/enkf-userguide.tex:
// 440-451, 490-522); the per-type diagnostics, which re-solve the analysis
// with only the obs of one type as the USE_MPIQUEUE build does
// (transforms.c:853-879), are factor_wave.hip's factor_trace_kernel for
// every scheme.
#include <algorithm>

#include "enkfg_internal.h"
#include "gram_layout.h"
#include "gram_stage.h"

namespace enkfg {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ utilities
__global__ void standardise_kernel(int nmem, int nobs, const short* __restrict__ type,
                                   const float* __restrict__ estd, const double* __restrict__ sqrt_rfactor,
                                   double mult, float* __restrict__ S, double* __restrict__ sv)
{
    // calc/ensobs.c:849-898: S /= estd * sqrt(rfactor) * sqrt(m - 1)
    const long long n = (long long)nobs * nmem;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const long long o = k / nmem;
        const double div = (double)estd[o] * sqrt_rfactor[type[o]] * mult;
        S[k] = (float)((double)S[k] / div);
        if (sv != nullptr && k % nmem == 0)
            sv[o] = sv[o] / div;
    }
}

__global__ void transpose_kernel(const float* __restrict__ in, float* __restrict__ out, long long rows, long long cols)
{
    __shared__ float tile[32][33];
    const long long c0 = (long long)blockIdx.x * 32, r0 = (long long)blockIdx.y * 32;
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

// Gather: dst row r (stride `stride`) = src row perm[r] (stride `width`)
// for r < nrow, zero padded to the stride, followed by one zero row.
// Scatter: dst row perm[r] (stride `width`) = the first `width` elements of
// src row r (stride `stride`).
template <class T, bool SCATTER>
__global__ void permute_rows_kernel(const T* __restrict__ src, const int* __restrict__ perm, T* __restrict__ dst,
                                    long long nrow, int width, int stride)
{
    const long long n = (SCATTER ? nrow : nrow + 1) * stride;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const long long r = k / stride;
        const int c = (int)(k - r * stride);
        if (SCATTER) {
            if (c < width)
                dst[(long long)perm[r] * width + c] = src[k];
        } else {
            dst[k] = (r < nrow && c < width) ? src[(long long)perm[r] * width + c] : T(0);
        }
    }
}

// ------------------------------------------------------------------ Gram, independent wavefronts
//
// Each wavefront computes one group of upper-triangle 16 x 16 tiles (tiles
// f = g, g + NG, ... of tri_count(MT); NG = 3 groups of 7 tiles or NG = 2
// groups of 11 and 10 for m = 96, 4 or 5 groups of the 45 tiles for m = 144:
// item_groups_lo, chosen per obs-count bucket by the host's autotuner) of
// one (node, obstype segment), with no workgroup barrier: the wave
// stages its own chunks of IKC local obs through a private LDS area, so the
// waves of one segment share nothing but L2 (the other waves' copies of the
// same rows hit there).  Per chunk: the copies of chunks c+1 and c+2 and the
// list entries of chunk c+3 are in flight while chunk c feeds the MFMAs.
// The taper is applied in FP64 as an operand is formed, (double) S * lc, as
// the reference forms Sloc.
// obs per chunk (kChunk, gram_stage.h): four k-steps.  With
// register-staged chunks, 16 rather than 8 gave a calc pass on the tile of
// 888 instead of 904 ms: at 8 the Gram took so many waves per CU that the
// factor kernels of the previous batch, on the other stream, overlapped it
// worse (32: 984 ms, 222 VGPRs).
constexpr int IKC = kChunk;
// chunk buffers (four measured 712.7 against 679.9 ms: six waves per CU)
constexpr int IKB = 3;
static_assert(kListPad >= (IKB + 1) * IKC, "list padding too small for the item Gram pipeline");
template <int MT>
using ItemLds = ChunkLds<MT, IKB>;

// number of tiles of group G of NG
template <int MT, int NG, int G>
constexpr int item_ntile()
{
    return (tri_count(MT) - G + NG - 1) / NG;
}

// the tiles f = G + NG t of group G: acc[t] += frag[I] frag[J]^T
template <int MT, int NG, int G, int T>
__device__ __attribute__((always_inline)) inline void item_tiles(const double (&frag)[MT], dbl4* acc)
{
    if constexpr (G + NG * T < tri_count(MT)) {
        constexpr int f = G + NG * T;
        constexpr int I = tri_I(MT, f), J = tri_J(MT, f);
        acc[T] = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[I], frag[J], acc[T], 0, 0, 0);
        item_tiles<MT, NG, G, T + 1>(frag, acc);
    }
}

template <int MT, int NG, int G>
__device__ __attribute__((always_inline)) inline void item_wave(const CalcArgs& a, int b, long long off, int cnt, ItemLds<MT>& sh)
{
    constexpr int MP = 16 * MT;
    constexpr int NTG = item_ntile<MT, NG, G>();
    const int lane = threadIdx.x & 63;
    const int kq = lane >> 4;
    const int col = lane & 15;

    dbl4 acc[NTG > 0 ? NTG : 1];
#pragma unroll
    for (int f = 0; f < NTG; ++f)
        acc[f] = dbl4{0.0, 0.0, 0.0, 0.0};
    double sts[MT];
#pragma unroll
    for (int X = 0; X < MT; ++X)
        sts[X] = 0.0;

    const int nchunk = (cnt + IKC - 1) / IKC;
    chunk_prologue<MT, IKB>(a, sh, off, cnt);
    int buf = 0;    // c % IKB
    int slot = 0;   // c % (IKB + 1)
    for (int c = 0; c < nchunk; ++c) {
        chunk_advance<MT, IKB>(a, sh, off, cnt, c, buf, slot);
        const double* svc = reinterpret_cast<const double*>(&sh.sv[buf][0]);
        const double* lcc = reinterpret_cast<const double*>(&sh.meta[slot][16]);
#pragma unroll
        for (int s = 0; s < IKC / 4; ++s) {
            const int r = 4 * s + kq;
            const double lr = lcc[r];
            const double l = c * IKC + r < cnt ? lr : 0.0;
            double frag[MT];
#pragma unroll
            for (int X = 0; X < MT; ++X)
                frag[X] = (double)sh.rows[buf][X][r][col] * l;
            item_tiles<MT, NG, G, 0>(frag, acc);
            if (G == 0) {
                const double sl = svc[r] * l;
#pragma unroll
                for (int X = 0; X < MT; ++X)
                    sts[X] += frag[X] * sl;
            }
        }
        chunk_fence();
        buf = buf == IKB - 1 ? 0 : buf + 1;
        slot = slot == IKB ? 0 : slot + 1;
    }
    // the copies of the chunks past the end land before the wave ends
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // outputs in the factor kernel's register-block layout, both triangles
    // (accumulator register i of lane l holds D[4 i + l / 16][l % 16])
    double* gout = a.gram + (size_t)b * MP * MP;
#pragma unroll
    for (int t = 0; t < NTG; ++t) {
        const int f = G + NG * t;
        const int I = tri_I(MT, f), J = tri_J(MT, f);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 16 * I + 4 * i + kq;
            const int cc = 16 * J + col;
            gout[gram_index<MT>(r, cc)] = acc[t][i];
            if (I != J)
                gout[gram_index<MT>(cc, r)] = acc[t][i];
        }
    }
    if (G == 0) {
        double* stsout = a.sts + (size_t)b * MP;
#pragma unroll
        for (int X = 0; X < MT; ++X) {
            double v = sts[X];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (kq == 0)
                stsout[16 * X + col] = v;
        }
    }
}

// wave of group g (the groups with no tile do nothing)
template <int MT, int NG, int G>
__device__ __attribute__((always_inline)) inline void item_dispatch(int g, const CalcArgs& a, int b, long long off,
                                                                    int cnt, ItemLds<MT>& sh)
{
    if constexpr (G < NG) {
        if (g == G) {
            if constexpr (item_ntile<MT, NG, G>() > 0)
                item_wave<MT, NG, G>(a, b, off, cnt, sh);
        } else {
            item_dispatch<MT, NG, G + 1>(g, a, b, off, cnt, sh);
        }
    }
}

// item `it` = (segment b, group g) of the batch
template <int MT, int NG>
__device__ __attribute__((always_inline)) inline void item_run(const CalcArgs& a, int it, ItemLds<MT>& sh)
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
        item_dispatch<MT, NG, 0>(g, a, b, off, cnt, sh);
}

template <int MT, int NG>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) gram_item_kernel(CalcArgs a)
{
    // one item per wave.  Blocks are handed out as earlier ones finish (the
    // hardware dispatcher is the work queue); each XCD takes runs of
    // kItemRun consecutive (segment, group) items, the NG groups of a
    // segment adjacent, so neighbouring nodes and the groups of one segment
    // share that XCD's L2
    __shared__ ItemLds<MT> sh;
    const int nitem = a.nbatch * a.nseg * NG;
    const int it = item_of_block();
    if (it < nitem)
        item_run<MT, NG>(a, it, sh);
}

// ------------------------------------------------------------------ factor
// 256 threads as a 16 x 16 grid; thread (ti, tj) owns the B x B block of rows
// ti*B.., columns tj*B.. of the padded 16B x 16B matrix (B = ceil(m / 16));
// the padding is kept as an identity block, which inversion leaves intact.
template <int B>
struct FactorShared {
    double piv[2][B * B];          // inverse of the pivot block (double buffered)
    double rowb[2][16][B * B];     // scaled pivot block-row, by block column
    double colb[2][16][B * B];     // pivot block-column before the step, by block row
    double red[4];
    double vec[16][16 * B + 1];
    int flag;
};

// Inverse of a small SPD block held by one thread (scalar Gauss-Jordan,
// compile-time indices).  Returns sum(log(pivots)); sets *bad on a pivot <= 0.
template <int B, bool LOGDET>
__device__ inline double small_invert(double (&p)[B][B], int& bad)
{
    double logdet = 0.0;
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const double d = p[k][k];
        if (!(d > 0.0))
            bad = 1;
        if constexpr (LOGDET)
            logdet += log(d);
        const double rd = 1.0 / d;
#pragma unroll
        for (int i = 0; i < B; ++i) {
            if (i == k)
                continue;
            const double f = p[i][k] * rd;
#pragma unroll
            for (int j = 0; j < B; ++j)
                if (j != k)
                    p[i][j] = fma(-f, p[k][j], p[i][j]);
        }
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (j != k)
                p[k][j] *= rd;
#pragma unroll
        for (int i = 0; i < B; ++i)
            if (i != k)
                p[i][k] = -p[i][k] * rd;
        p[k][k] = rd;
    }
    return logdet;
}

// In-place inversion of the SPD matrix held in blocks `x` by blocked
// Gauss-Jordan elimination without pivoting (stable for SPD matrices, whose
// pivot blocks are the SPD Schur complements).  Block step K:
//   P = A_KK^-1;  A_KJ <- P A_KJ;  A_IJ <- A_IJ - A_IK A_KJ;  A_IK <- -A_IK P;
//   A_KK <- P.
// Returns sum(log(pivots)) when LOGDET; sets sh.flag on a non-positive pivot.
template <int B, bool LOGDET = false>
__device__ inline double gj_invert(double (&x)[B][B], int m, FactorShared<B>& sh)
{
    (void)m;
    const int tid = threadIdx.x;
    const int ti = tid >> 4, tj = tid & 15;
    double logdet = 0.0;
    for (int K = 0; K < 16; ++K) {
        const int buf = K & 1;
        if (ti == K && tj == K) {
            int bad = 0;
            logdet = small_invert<B, LOGDET>(x, bad);
            if (bad)
                sh.flag = 1;
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c)
                    sh.piv[buf][r * B + c] = x[r][c];
        }
        __syncthreads();
        if (ti == K && tj != K) {
            // A_KJ <- P A_KJ, one column at a time
#pragma unroll
            for (int c = 0; c < B; ++c) {
                double t[B];
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    double s = 0.0;
#pragma unroll
                    for (int q = 0; q < B; ++q)
                        s = fma(sh.piv[buf][r * B + q], x[q][c], s);
                    t[r] = s;
                }
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    x[r][c] = t[r];
                    sh.rowb[buf][tj][r * B + c] = t[r];
                }
            }
        }
        if (tj == K && ti != K) {
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c)
                    sh.colb[buf][ti][r * B + c] = x[r][c];
        }
        __syncthreads();
        if (ti != K && tj != K) {
            // A_IJ <- A_IJ - A_IK A_KJ as B rank-1 updates
#pragma unroll
            for (int q = 0; q < B; ++q) {
                double aq[B], bq[B];
#pragma unroll
                for (int r = 0; r < B; ++r)
                    aq[r] = sh.colb[buf][ti][r * B + q];
#pragma unroll
                for (int c = 0; c < B; ++c)
                    bq[c] = sh.rowb[buf][tj][q * B + c];
#pragma unroll
                for (int r = 0; r < B; ++r)
#pragma unroll
                    for (int c = 0; c < B; ++c)
                        x[r][c] = fma(-aq[r], bq[c], x[r][c]);
            }
        } else if (ti != K && tj == K) {
            // A_IK <- -A_IK P, one row at a time
#pragma unroll
            for (int r = 0; r < B; ++r) {
                double t[B];
#pragma unroll
                for (int c = 0; c < B; ++c) {
                    double s = 0.0;
#pragma unroll
                    for (int q = 0; q < B; ++q)
                        s = fma(x[r][q], sh.piv[buf][q * B + c], s);
                    t[c] = -s;
                }
#pragma unroll
                for (int c = 0; c < B; ++c)
                    x[r][c] = t[c];
            }
        }
    }
    __syncthreads();
    if constexpr (LOGDET) {
        // the log-determinant accumulates in the diagonal-block owners; sum it
        double v = (ti == tj) ? logdet : 0.0;
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o);
        if ((tid & 63) == 0)
            sh.red[tid >> 6] = v;
        __syncthreads();
        logdet = sh.red[0] + sh.red[1] + sh.red[2] + sh.red[3];
        __syncthreads();
    }
    return logdet;
}

template <int B>
__device__ inline double block_sum(double v, FactorShared<B>& sh)
{
    // deterministic block reduction: wave reduction then 4 partials in order
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        sh.red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = sh.red[0] + sh.red[1] + sh.red[2] + sh.red[3];
    __syncthreads();
    return s;
}

// Loads G0 = sum of the node's segment Gram matrices (one segment if seg >= 0)
// into x (plus `diag` on the diagonal); returns the trace of G0.
template <int B>
__device__ inline double load_gram(const CalcArgs& a, int bnode, int seg, double diag, double (&x)[B][B], FactorShared<B>& sh)
{
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int mp = 16 * ((a.m + 15) / 16);
    const int m = a.m;
    double tr = 0.0;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            x[r][c] = 0.0;
    for (int s = 0; s < a.nseg; ++s) {
        if (seg >= 0 && s != seg)
            continue;
        const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
        const int cnt = a.pertype ? a.counts[(size_t)node * a.ntypes + s] : 1;
        if (cnt == 0)
            continue;
        const double* g = a.gram + ((size_t)bnode * a.nseg + s) * mp * mp;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int i = ti * B + r;
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int j = tj * B + c;
                if (i < m && j < m)
                    x[r][c] += g[gram_index<B>(i, j)];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const int i = ti * B + r;
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int j = tj * B + c;
            if (i == j && i < m) {
                tr += x[r][c];
                x[r][c] += diag;
            } else if (i == j) {
                x[r][c] = 1.0;   // identity padding, left intact by the inversion
            }
        }
    }
    return block_sum(tr, sh);
}

// <X, G0> for the inverse in x against the node's Gram matrices
template <int B>
__device__ inline double dot_gram(const CalcArgs& a, int bnode, int seg, const double (&x)[B][B], FactorShared<B>& sh)
{
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int mp = 16 * ((a.m + 15) / 16);
    const int m = a.m;
    double acc = 0.0;
    for (int s = 0; s < a.nseg; ++s) {
        if (seg >= 0 && s != seg)
            continue;
        const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
        const int cnt = a.pertype ? a.counts[(size_t)node * a.ntypes + s] : 1;
        if (cnt == 0)
            continue;
        const double* g = a.gram + ((size_t)bnode * a.nseg + s) * mp * mp;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int i = ti * B + r;
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int j = tj * B + c;
                if (i < m && j < m)
                    acc += x[r][c] * g[gram_index<B>(i, j)];
            }
        }
    }
    return block_sum(acc, sh);
}

// y = X v for the node's S^T s (summed over segments); result in sh.vec[0][0..m)
template <int B>
__device__ inline void matvec_sts(const CalcArgs& a, int bnode, const double (&x)[B][B], FactorShared<B>& sh)
{
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int mp = 16 * ((a.m + 15) / 16);
    const int m = a.m;
    double v[B];
#pragma unroll
    for (int c = 0; c < B; ++c)
        v[c] = 0.0;
    for (int s = 0; s < a.nseg; ++s) {
        const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
        const int cnt = a.pertype ? a.counts[(size_t)node * a.ntypes + s] : 1;
        if (cnt == 0)
            continue;
        const double* st = a.sts + ((size_t)bnode * a.nseg + s) * mp;
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int j = tj * B + c;
            if (j < m)
                v[c] += st[j];
        }
    }
#pragma unroll
    for (int r = 0; r < B; ++r) {
        double part = 0.0;
#pragma unroll
        for (int c = 0; c < B; ++c)
            part += x[r][c] * v[c];
        sh.vec[tj][ti * B + r] = part;
    }
    __syncthreads();
    if (tid < m) {
        double sum = 0.0;
        for (int q = 0; q < 16; ++q)
            sum += sh.vec[q][tid];
        sh.vec[0][tid] = sum;
    }
    __syncthreads();
}

template <int B>
__device__ inline void write_T(const CalcArgs& a, int node, const double (&x)[B][B], double scale, double shift, FactorShared<B>& sh)
{
    // T[e1][e2] = shift*delta + scale*X(e1, e2) for e1 < md, then relaxation
    // to prior T <- (T - I) alpha + I (transforms.c:814-828) and FP32 cast.
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int m = a.m, md = a.md;
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const int e1 = ti * B + r;
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int e2 = tj * B + c;
            if (e1 < md && e2 < m) {
                const double dlt = (e1 == e2) ? 1.0 : 0.0;
                double t = scale * x[r][c] + shift * dlt;
                t = (t - dlt) * a.alpha + dlt;
                const size_t o = ((size_t)node * md + e1) * m + e2;
                a.T[o] = (float)t;
                if (a.T64 != nullptr)
                    a.T64[o] = t;
            }
        }
    }
}

// The scaled coupled Denman-Beavers iteration of factor_kernel for B > 6,
// where Y, Z and Y^-1 do not fit in registers beside x: they live in the
// batch node's slice of a.etkf_ws, each thread's B x B block at
// ws[(k * B * B + r * B + c) * 256 + tid] for matrix k (0 Y, 1 Z, 2 Y^-1),
// and only the matrix being inverted or combined is held in x.  On entry x
// holds M^-1 = Y0^-1 and logdetA = log det M.  Writes T = Z (-> M^-1/2).
template <int B>
__device__ inline void etkf_ws_iterate(const CalcArgs& a, int bnode, int node, double logdetA, double (&x)[B][B],
                                       FactorShared<B>& sh)
{
    const int tid = threadIdx.x;
    const int ti = tid >> 4, tj = tid & 15;
    const int m = a.m;
    double* ws = a.etkf_ws + (size_t)bnode * etkf_ws_per_node(m);
    auto at = [&](int k, int r, int c) -> double& { return ws[((size_t)(k * B * B + r * B + c)) * 256 + tid]; };
    // Y^-1 = x; Y = M (= I + G0); Z = I
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            at(2, r, c) = x[r][c];
            at(1, r, c) = (ti * B + r == tj * B + c) ? 1.0 : 0.0;
        }
    load_gram<B>(a, bnode, -1, 1.0, x, sh);
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            at(0, r, c) = x[r][c];
    double logdetY = logdetA, logdetZ = 0.0;
    bool last = false;
    for (int it = 0; it < 60; ++it) {
        if (it > 0) {
            // Y^-1
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c)
                    x[r][c] = at(0, r, c);
            logdetY = gj_invert<B, true>(x, m, sh);
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c)
                    at(2, r, c) = x[r][c];
        }
        // x = Z^-1 (Z = I at the first step)
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                x[r][c] = at(1, r, c);
        if (it > 0)
            logdetZ = gj_invert<B, true>(x, m, sh);
        const double mu = last ? 1.0 : exp(-(logdetY + logdetZ) / (2.0 * m));
        const double hm = 0.5 * mu, hi = 0.5 / mu;
        double dn = 0.0, nn = 0.0;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const double y = at(0, r, c), z = at(1, r, c);
                const double yn = hm * y + hi * x[r][c];
                const double zn = hm * z + hi * at(2, r, c);
                dn += (zn - z) * (zn - z);
                nn += zn * zn;
                at(0, r, c) = yn;
                at(1, r, c) = zn;
            }
        dn = block_sum(dn, sh);
        nn = block_sum(nn, sh);
        if (last)
            break;
        if (dn <= 1e-20 * nn)
            last = true;   // converged quadratically: one more unscaled step
    }
    if (!last && tid == 0)
        sh.flag = 1;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            x[r][c] = at(1, r, c);
    write_T<B>(a, node, x, 1.0, 0.0, sh);
}

template <int B>
__global__ void __launch_bounds__(256) factor_kernel(CalcArgs a)
{
    __shared__ FactorShared<B> sh;
    const int bnode = blockIdx.x;
    const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
    const int tid = threadIdx.x;
    const int m = a.m, md = a.md;
    const bool dotT = (a.mode == ENKFG_MODE_ENKF || a.mode == ENKFG_MODE_HYBRID);
    const int nt = a.ntypes;

    const long long off = a.offsets[(size_t)node * nt];
    const int p = (int)(a.offsets[(size_t)(node + 1) * nt] - off);

    if (tid == 0)
        sh.flag = 0;
    if (p == 0) {
        // empty column: T = I, w = 0, diagnostics 0 (transforms.c:766-780)
        if (dotT)
            for (int k = tid; k < md * m; k += 256) {
                const int e1 = k / m, e2 = k % m;
                a.T[(size_t)node * md * m + k] = (e1 == e2) ? 1.0f : 0.0f;
                if (a.T64 != nullptr)
                    a.T64[(size_t)node * md * m + k] = (e1 == e2) ? 1.0 : 0.0;
            }
        for (int e = tid; e < m; e += 256) {
            a.w[(size_t)node * m + e] = 0.0f;
            if (a.w64 != nullptr)
                a.w64[(size_t)node * m + e] = 0.0;
        }
        for (int t = tid; t <= nt; t += 256) {
            a.pnlobs[(size_t)t * a.nnode + node] = 0;
            a.pdfs[(size_t)t * a.nnode + node] = 0.0f;
            a.psrf[(size_t)t * a.nnode + node] = 0.0f;
        }
        return;
    }

    double x[B][B];
    const double trG = load_gram<B>(a, bnode, -1, 1.0, x, sh);
    const double logdetA = gj_invert<B, true>(x, m, sh);
    // w = M^-1 S^T s
    matvec_sts<B>(a, bnode, x, sh);
    if (tid < m) {
        const double wv = sh.vec[0][tid];
        a.w[(size_t)node * m + tid] = (float)wv;
        if (a.w64 != nullptr)
            a.w64[(size_t)node * m + tid] = wv;
    }
    __syncthreads();
    // dfs = tr(M^-1 G0), srf = sqrt(tr(G0) / (float) dfs) - 1
    const double dfs = dot_gram<B>(a, bnode, -1, x, sh);
    if (tid == 0) {
        const float fd = (float)dfs;
        a.pnlobs[(size_t)nt * a.nnode + node] = p;
        a.pdfs[(size_t)nt * a.nnode + node] = fd;
        a.psrf[(size_t)nt * a.nnode + node] = (float)(sqrt(trG / (double)fd) - 1.0);
    }

    if (dotT) {
        etkf_ws_iterate<B>(a, bnode, node, logdetA, x, sh);
    }

    // the per-type diagnostics (the analysis re-solved with the obs of one
    // type) are factor_trace_kernel's work items, launched after this kernel
    if (!a.pertype && tid < nt) {
        a.pnlobs[(size_t)tid * a.nnode + node] = a.counts[(size_t)node * nt + tid];
        a.pdfs[(size_t)tid * a.nnode + node] = 0.0f;
        a.psrf[(size_t)tid * a.nnode + node] = 0.0f;
    }
    __syncthreads();
    if (tid == 0 && sh.flag)
        atomicExch(a.err, ENKFG_ERR_LINALG);
}

}  // namespace

hipError_t launch_standardise(hipStream_t s, int nmem, int nobs, const short* type, const float* estd,
                              const double* sqrt_rfactor, float* S_obsmajor, double* sv)
{
    const long long n = (long long)nmem * nobs;
    if (n == 0)
        return hipSuccess;
    const double mult = sqrt((double)nmem - 1);
    long long nb = (n + 255) / 256;
    if (nb > 65536)
        nb = 65536;
    hipLaunchKernelGGL(standardise_kernel, dim3((unsigned)nb), dim3(256), 0, s, nmem, nobs, type, estd, sqrt_rfactor,
                       mult, S_obsmajor, sv);
    return hipGetLastError();
}

hipError_t launch_transpose_f32(hipStream_t s, const float* in, float* out, long long rows, long long cols)
{
    if (rows == 0 || cols == 0)
        return hipSuccess;
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    if (grid.y > 65535)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, s, in, out, rows, cols);
    return hipGetLastError();
}

template <class T, bool SCATTER>
hipError_t launch_permute(hipStream_t s, const T* src, const int* perm, T* dst, long long nrow, int width, int stride)
{
    const long long n = (SCATTER ? nrow : nrow + 1) * stride;
    if (n == 0)
        return hipSuccess;
    const long long nb = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL((permute_rows_kernel<T, SCATTER>), dim3((unsigned)nb), dim3(256), 0, s, src, perm, dst, nrow,
                       width, stride);
    return hipGetLastError();
}

hipError_t launch_gather_rows(hipStream_t s, const float* src, const int* perm, float* dst, long long nrow, int width,
                              int stride)
{
    return launch_permute<float, false>(s, src, perm, dst, nrow, width, stride);
}
hipError_t launch_gather_rows(hipStream_t s, const double* src, const int* perm, double* dst, long long nrow,
                              int width, int stride)
{
    return launch_permute<double, false>(s, src, perm, dst, nrow, width, stride);
}
hipError_t launch_scatter_rows(hipStream_t s, const float* src, const int* perm, float* dst, long long nrow, int width,
                               int stride)
{
    return launch_permute<float, true>(s, src, perm, dst, nrow, width, stride);
}
hipError_t launch_scatter_rows(hipStream_t s, const double* src, const int* perm, double* dst, long long nrow,
                               int width, int stride)
{
    return launch_permute<double, true>(s, src, perm, dst, nrow, width, stride);
}

template <int M>
void launch_gram_item(hipStream_t s, dim3 grid, const CalcArgs& a, int ng)
{
    constexpr int lo = item_groups_lo(M);
    if (ng == lo)
        hipLaunchKernelGGL((gram_item_kernel<M, lo>), grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((gram_item_kernel<M, lo + 1>), grid, dim3(64), 0, s, a);
}

hipError_t launch_gram(hipStream_t s, const CalcArgs& a)
{
    const int MT = (a.m + 15) / 16;
    if (a.nbatch == 0)
        return hipSuccess;
    if (MT < 1 || MT > 10)
        return hipErrorInvalidValue;
    // tile groups per segment: one of the two candidates of this MT (the
    // larger by default)
    const int lo = item_groups_lo(MT);
    const int ng = (a.tile_groups == lo) ? lo : lo + 1;
    const long long nitem = (long long)a.nbatch * a.nseg * ng;
    const dim3 igrid(item_grid(nitem));
    switch (MT) {
    case 1: launch_gram_item<1>(s, igrid, a, ng); break;
    case 2: launch_gram_item<2>(s, igrid, a, ng); break;
    case 3: launch_gram_item<3>(s, igrid, a, ng); break;
    case 4: launch_gram_item<4>(s, igrid, a, ng); break;
    case 5: launch_gram_item<5>(s, igrid, a, ng); break;
    case 6: launch_gram_item<6>(s, igrid, a, ng); break;
    case 7: launch_gram_item<7>(s, igrid, a, ng); break;
    case 8: launch_gram_item<8>(s, igrid, a, ng); break;
    case 9: launch_gram_item<9>(s, igrid, a, ng); break;
    default: launch_gram_item<10>(s, igrid, a, ng); break;
    }
    return hipGetLastError();
}

template <int B>
void launch_factor_b(hipStream_t s, const CalcArgs& a)
{
    hipLaunchKernelGGL((factor_kernel<B>), dim3((unsigned)a.nbatch), dim3(256), 0, s, a);
}

hipError_t launch_factor(hipStream_t s, const CalcArgs& a)
{
    if (factor_wave_supported(a))
        return launch_factor_wave(s, a);   // DEnKF, EnOI: the symmetric sweep (factor_wave.hip)
    // ETKF: by the symmetric sweep up to m = 96; beyond, the Gauss-Jordan
    // kernel with its iterates in device memory
    if (a.m <= 96)
        return launch_factor_etkf(s, a);
    const int B = (a.m + 15) / 16;
    if (a.nbatch == 0)
        return hipSuccess;
    switch (B) {
    case 7: launch_factor_b<7>(s, a); break;
    case 8: launch_factor_b<8>(s, a); break;
    case 9: launch_factor_b<9>(s, a); break;
    case 10: launch_factor_b<10>(s, a); break;
    default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.pertype)
        return e;
    return launch_factor_trace(s, a);   // per-type diagnostics, after the node's own
}

}  // namespace enkfg
