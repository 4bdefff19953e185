// factor_wave.hip -- the per-column factorisation of the DEnKF and EnOI
// analyses by unblocked Gauss-Jordan elimination.
//
// A work item of factor_wave_kernel is the symmetric positive definite
// M = I + G0 of a stride-grid node, G0 its full Gram matrix (transforms, w,
// total diagnostics).  The per-obstype diagnostics (calc/transforms.c:843-879),
// which need only tr(M_t^-1) and tr(G_t), are factor_trace_kernel's work
// items, one per (node, obstype).  The 256 lanes of the workgroup
// form a 16 x 16 grid; lane (lr, lc) holds the MT x MT block of rows lr MT..,
// columns lc MT.. of the padded 16 MT x 16 MT matrix in registers (padding
// kept as identity, which every step leaves intact): 36 doubles per lane for
// m = 96, so that four matrices fit on a CU.  The symmetric-sweep kernels
// below take m <= 160 (MT <= 10); the Gauss-Jordan kernel, which only the
// adjoint uses, m <= 96.
//
// M^-1 comes from unblocked Gauss-Jordan elimination with diagonal pivots
// (stable for SPD matrices, whose pivots are positive Schur complements):
// the owners of pivot k's row and column publish them in LDS (double
// buffered, one barrier per pivot) and every lane applies the rank-1 update
// to its block.  The serial chain of the elimination (m pivots) is hidden by
// the other workgroups of the CU.  Then
//
//   DEnKF  T = (I + M^-1) / 2 (first m_d rows), relaxed to the prior and cast
//          to FP32 (transforms.c:814-842);
//   w      = M^-1 S^T s;
//   dfs    = tr(M^-1 G0) = m - tr(M^-1)   (as M^-1 G0 = I - M^-1),
//   srf    = sqrt(tr(G0) / (float) dfs) - 1, per type with the DFS_MIN guard.
#include "enkfg_internal.h"
#include "gram_layout.h"

namespace enkfg {

namespace {

constexpr double kDfsMin = 1.0e-6;   // DFS_MIN, calc/transforms.c:32

// sum over the 256 lanes: wave reductions, then the four partials in order
// through `tmp` (which must not be in use)
__device__ inline double group_sum(double v, double* tmp, int lane)
{
    for (int o = 1; o < 64; o <<= 1)
        v += __shfl_xor(v, o);
    __syncthreads();
    if ((lane & 63) == 0)
        tmp[lane >> 6] = v;
    __syncthreads();
    const double s = ((tmp[0] + tmp[1]) + tmp[2]) + tmp[3];
    __syncthreads();
    return s;
}

template <int MT>
struct PivotLds {
    double prow[2][16 * MT];
    double pcol[2][16 * MT];
};

template <int MT>
__global__ void __launch_bounds__(256) factor_wave_kernel(CalcArgs a)
{
    constexpr int MP = 16 * MT;
    constexpr int RR = MT;       // rows per lane (16 lane rows)
    constexpr int RC = MT;       // columns per lane (16 lane columns)
    __shared__ PivotLds<MT> sh;
    const int lane = threadIdx.x;
    const int lr = lane >> 4, lc = lane & 15;
    const int nt = a.ntypes;
    const int bnode = (int)blockIdx.x;
    if (bnode >= a.nbatch)
        return;
    const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
    const bool adjoint = a.adj_g != nullptr;   // z = M^-1 g and w only
    const int m = a.m, md = a.md;
    const bool dotT = (a.mode == ENKFG_MODE_ENKF || a.mode == ENKFG_MODE_HYBRID);
    const long long off = a.offsets[(size_t)node * nt];
    const int p = (int)(a.offsets[(size_t)(node + 1) * nt] - off);

    if (adjoint && p == 0) {
        for (int e = lane; e < m; e += 256) {
            a.adj_z[(size_t)(a.node0 + bnode) * m + e] = 0.0;
            a.adj_w[(size_t)(a.node0 + bnode) * m + e] = 0.0;
        }
        if (a.adj_minv != nullptr)   // no local obs: T = I whatever M (nothing depends on it)
            for (int k = lane; k < MP * MP; k += 256)
                a.adj_minv[(size_t)bnode * MP * MP + k] = 0.0;
        return;
    }
    if (!a.pertype && !adjoint) {
        if (lane < nt) {
            a.pnlobs[(size_t)lane * a.nnode + node] = a.counts[(size_t)node * nt + lane];
            a.pdfs[(size_t)lane * a.nnode + node] = 0.0f;
            a.psrf[(size_t)lane * a.nnode + node] = 0.0f;
        }
    }
    if (p == 0) {
        if (lane == 0) {
            a.pnlobs[(size_t)nt * a.nnode + node] = 0;
            a.pdfs[(size_t)nt * a.nnode + node] = 0.0f;
            a.psrf[(size_t)nt * a.nnode + node] = 0.0f;
        }
        // empty column: T = I, w = 0 (transforms.c:766-780)
        if (dotT)
            for (int k = lane; k < md * m; k += 256) {
                const int e1 = k / m, e2 = k - e1 * m;
                a.T[(size_t)node * md * m + k] = (e1 == e2) ? 1.0f : 0.0f;
                if (a.T64 != nullptr)
                    a.T64[(size_t)node * md * m + k] = (e1 == e2) ? 1.0 : 0.0;
            }
        for (int e = lane; e < m; e += 256) {
            a.w[(size_t)node * m + e] = 0.0f;
            if (a.w64 != nullptr)
                a.w64[(size_t)node * m + e] = 0.0;
        }
        return;
    }

    // ---- M = I + G0 (G0 summed over the node's segments), tr(G0)
    double x[RR][RC];
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
        for (int c = 0; c < RC; ++c)
            x[r][c] = (lr * RR + r == lc * RC + c) ? 1.0 : 0.0;
    double trg = 0.0;
    for (int s = 0; s < a.nseg; ++s) {
        if (a.pertype && a.counts[(size_t)node * nt + s] == 0)
            continue;
        // this lane's rows and columns are block (lr, lc) of the gram_index
        // layout: one run of MT^2 doubles
        const double* g = a.gram + ((size_t)bnode * a.nseg + s) * MP * MP;
        const double* g0 = g + gram_index<MT>(lr * RR, lc * RC);
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int i = lr * RR + r;
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                const int j = lc * RC + c;
                const double v = g0[r * MT + c];
                if (i < m && j < m) {
                    x[r][c] += v;
                    if (i == j)
                        trg += v;
                }
            }
        }
    }
    trg = group_sum(trg, sh.prow[0], lane);

    // ---- Gauss-Jordan: x <- M^-1.  Pivot k = kb MT + rk lies in lane row
    // and lane column kb, local row and column rk.
    bool bad = false;
    int buf = 0;
    for (int kb = 0; kb < 16; ++kb) {
#pragma unroll
        for (int rk = 0; rk < RR; ++rk) {
            const int k = kb * RR + rk;
            if (k >= m)
                break;   // the remaining pivots are identity padding
            const int kc = k / RC;   // lane column owning column k
            if (lr == kb) {
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    sh.prow[buf][lc * RC + c] = x[rk][c];
            }
            if (lc == kc) {
#pragma unroll
                for (int r = 0; r < RR; ++r)
                    sh.pcol[buf][lr * RR + r] = x[r][rk % RC];
            }
            __syncthreads();
            const double d = sh.prow[buf][k];
            bad |= !(d > 0.0);
            const double pinv = 1.0 / d;
            double prow[RC], pcol[RR];
#pragma unroll
            for (int c = 0; c < RC; ++c)
                prow[c] = sh.prow[buf][lc * RC + c] * pinv;
#pragma unroll
            for (int r = 0; r < RR; ++r)
                pcol[r] = sh.pcol[buf][lr * RR + r];
#pragma unroll
            for (int r = 0; r < RR; ++r)
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    x[r][c] = fma(-pcol[r], prow[c], x[r][c]);
            if (lr == kb) {
#pragma unroll
                for (int c = 0; c < RC; ++c)
                    x[rk][c] = prow[c];                 // A_kc / d
            }
            if (lc == kc) {
#pragma unroll
                for (int r = 0; r < RR; ++r)
                    x[r][rk % RC] = -pcol[r] * pinv;    // -A_rk / d
            }
            if (lr == kb && lc == kc)
                x[rk][rk % RC] = pinv;
            buf ^= 1;
        }
    }
    if (bad)
        atomicExch(a.err, ENKFG_ERR_LINALG);

    if (adjoint) {
        // z = M^-1 g and w = M^-1 S^T s for the listed node
        const size_t k = (size_t)(a.node0 + bnode);
        double gv[RC], sv[RC];
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const int j = lc * RC + c;
            gv[c] = (j < m) ? a.adj_g[(size_t)node * m + j] : 0.0;
            sv[c] = (j < m) ? a.sts[(size_t)bnode * MP + j] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            double yz = 0.0, yw = 0.0;
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                yz = fma(x[r][c], gv[c], yz);
                yw = fma(x[r][c], sv[c], yw);
            }
            for (int o = 1; o < 16; o <<= 1) {
                yz += __shfl_xor(yz, o);
                yw += __shfl_xor(yw, o);
            }
            const int e = lr * RR + r;
            if (lc == 0 && e < m) {
                a.adj_z[k * m + e] = yz;
                a.adj_w[k * m + e] = yw;
            }
        }
        if (a.adj_minv != nullptr) {
            double* mi = a.adj_minv + (size_t)bnode * MP * MP;
#pragma unroll
            for (int r = 0; r < RR; ++r)
#pragma unroll
                for (int c = 0; c < RC; ++c) {
                    const int i = lr * RR + r, j = lc * RC + c;
                    mi[(size_t)i * MP + j] = (i < m && j < m) ? x[r][c] : 0.0;
                }
        }
        return;
    }

    // ---- dfs = m - tr(M^-1)
    double trx = 0.0;
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const int i = lr * RR + r;
            if (i < m && i == lc * RC + c)
                trx += x[r][c];
        }
    trx = group_sum(trx, sh.prow[0], lane);
    const double dfs = (double)m - trx;

    if (lane == 0) {
        const float fd = (float)dfs;
        a.pnlobs[(size_t)nt * a.nnode + node] = p;
        a.pdfs[(size_t)nt * a.nnode + node] = fd;
        a.psrf[(size_t)nt * a.nnode + node] = (float)(sqrt(trg / (double)fd) - 1.0);
    }

    // ---- w = M^-1 S^T s (row partials reduced over the 16 lanes of a row)
    double v[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c)
        v[c] = 0.0;
    for (int s = 0; s < a.nseg; ++s) {
        if (a.pertype && a.counts[(size_t)node * nt + s] == 0)
            continue;
        const double* st = a.sts + ((size_t)bnode * a.nseg + s) * MP;
#pragma unroll
        for (int c = 0; c < RC; ++c) {
            const int j = lc * RC + c;
            if (j < m)
                v[c] += st[j];
        }
    }
#pragma unroll
    for (int r = 0; r < RR; ++r) {
        double y = 0.0;
#pragma unroll
        for (int c = 0; c < RC; ++c)
            y = fma(x[r][c], v[c], y);
        y += __shfl_xor(y, 1);
        y += __shfl_xor(y, 2);
        y += __shfl_xor(y, 4);
        y += __shfl_xor(y, 8);
        const int e = lr * RR + r;
        if (lc == 0 && e < m) {
            a.w[(size_t)node * m + e] = (float)y;
            if (a.w64 != nullptr)
                a.w64[(size_t)node * m + e] = y;
        }
    }

    // ---- DEnKF T = (I + M^-1) / 2, relaxation T <- (T - I) alpha + I, FP32
    if (dotT) {
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int e1 = lr * RR + r;
            if (e1 >= md)
                continue;
#pragma unroll
            for (int c = 0; c < RC; ++c) {
                const int e2 = lc * RC + c;
                if (e2 < m) {
                    const double dlt = (e1 == e2) ? 1.0 : 0.0;
                    double t = 0.5 * x[r][c] + 0.5 * dlt;
                    t = (t - dlt) * a.alpha + dlt;
                    const size_t o = ((size_t)node * md + e1) * m + e2;
                    a.T[o] = (float)t;
                    if (a.T64 != nullptr)
                        a.T64[o] = t;
                }
            }
        }
    }
}


// ---- per-obstype diagnostics: dfs_t = m - tr(M_t^-1), M_t = I + G_t
//
// Only the trace of the inverse is needed (calc/transforms.c:865-871), and M_t
// is symmetric, as is every intermediate matrix of the sweep operator: pivot
// k sets A_ij -= A_ik A_kj / d for i, j != k, A_ik = A_ik / d, A_kj = A_kj / d
// and A_kk = -1 / d, which ends, after all pivots, with -M^-1.  So only the
// 136 blocks on and below the diagonal of the 16 x 16 grid of MT x MT blocks
// are kept, one per lane of a 3-wave workgroup (lanes 136-191 hold none), and
// each pivot publishes one vector in LDS: column k, read by a block (br, bc)
// at rows br MT.. (A_ik) and at rows bc MT.. (A_kj = A_jk).  Three waves and
// one vector per matrix, against four waves and two vectors for the full
// elimination.
constexpr int kTriBlocks = 136;   // 16 * 17 / 2

// waves per SIMD that the sweep kernels' registers allow: a lane holds its
// MT x MT block and two MT-vectors in doubles (about 2 MT^2 + 4 MT + 20
// VGPRs): 4 waves up to MT = 6 (m <= 96), 3 at MT = 7, 2 beyond
__host__ __device__ constexpr int sweep_waves(int MT) { return MT <= 6 ? 4 : (MT == 7 ? 3 : 2); }

// 1/d for a pivot d >= 1 (the pivots of I + S^T S are Schur complements of
// at least 1): the hardware estimate refined by two Newton-Raphson steps
// r <- r + r (1 - d r), within an ulp of 1/d
__device__ inline double recip_pivot(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// sum over the 192 lanes (three waves), in order through `tmp`
__device__ inline double group_sum3(double v, double* tmp, int lane)
{
    for (int o = 1; o < 64; o <<= 1)
        v += __shfl_xor(v, o);
    __syncthreads();
    if ((lane & 63) == 0)
        tmp[lane >> 6] = v;
    __syncthreads();
    const double s = (tmp[0] + tmp[1]) + tmp[2];
    __syncthreads();
    return s;
}

template <int MT>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(sweep_waves(MT)))) factor_trace_kernel(CalcArgs a)
{
    constexpr int MP = 16 * MT;
    __shared__ double col[2][MP];
    __shared__ double tmp[4];
    const int lane = threadIdx.x;
    const int nt = a.ntypes;
    const int bnode = (int)(blockIdx.x / nt);
    const int t = (int)(blockIdx.x - (unsigned)bnode * nt);
    if (bnode >= a.nbatch)
        return;
    const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
    const int m = a.m;
    const int pt = a.counts[(size_t)node * nt + t];
    const size_t o = (size_t)t * a.nnode + node;
    if (pt == 0) {
        if (lane == 0) {
            a.pnlobs[o] = 0;
            a.pdfs[o] = 0.0f;
            a.psrf[o] = 0.0f;
        }
        return;
    }
    if (lane == 0)
        a.pnlobs[o] = pt;
    // every local obs of the node is of this type: M_t is the node's own M,
    // whose diagnostics factor_sym_kernel (earlier on this stream) has
    // written; the reference's re-solve works on the same matrix
    const long long p = a.offsets[(size_t)(node + 1) * nt] - a.offsets[(size_t)node * nt];
    if (pt == p) {
        if (lane == 0) {
            const size_t ot = (size_t)nt * a.nnode + node;
            const float fd = a.pdfs[ot];
            a.pdfs[o] = fd;
            a.psrf[o] = ((double)fd > kDfsMin) ? a.psrf[ot] : 0.0f;
        }
        return;
    }

    // block (br, bc), br >= bc, of lane br (br + 1) / 2 + bc; lanes past the
    // last block compute on block (0, 0)'s column data but never publish,
    // load or count
    const bool act = lane < kTriBlocks;
    int br = 0;
    if (act)
        while ((br + 1) * (br + 2) / 2 <= lane)
            ++br;
    const int bc = act ? lane - br * (br + 1) / 2 : 0;

    // ---- M_t = I + G_t on this lane's block, tr(G_t)
    double x[MT][MT];
    double trg = 0.0;
    const double* g = a.gram + ((size_t)bnode * a.nseg + t) * MP * MP + gram_index<MT>(br * MT, bc * MT);
#pragma unroll
    for (int r = 0; r < MT; ++r) {
        const int i = br * MT + r;
#pragma unroll
        for (int c = 0; c < MT; ++c) {
            const int j = bc * MT + c;
            double v = (i == j) ? 1.0 : 0.0;
            if (act && i < m && j < m) {
                const double gv = g[r * MT + c];
                v += gv;
                if (i == j)
                    trg += gv;
            }
            x[r][c] = v;
        }
    }
    trg = group_sum3(trg, tmp, lane);

    // column k of the current matrix, from the blocks (br >= kb, bc == kb)
    // (their column rk) and (br == kb, bc < kb) (their row rk)
    auto publish = [&](double* dst, int kb, int rk) {
        if (!act)
            return;
        if (bc == kb) {
#pragma unroll
            for (int r = 0; r < MT; ++r)
                dst[br * MT + r] = x[r][rk];
        } else if (br == kb) {
#pragma unroll
            for (int c = 0; c < MT; ++c)
                dst[bc * MT + c] = x[rk][c];
        }
    };

    bool bad = false;
    publish(col[0], 0, 0);
    for (int kb = 0; kb < 16; ++kb) {
#pragma unroll
        for (int rk = 0; rk < MT; ++rk) {
            const int k = kb * MT + rk;
            if (k >= m) {
                // the remaining pivots are identity padding.  Past 8 steps a
                // loop with an early exit is not unrolled (and the blocks
                // would go to scratch): such loops skip instead
                if (MT <= 8)
                    break;
                continue;
            }
            __syncthreads();
            const int buf = k & 1;   // pivot k reads buffer k % 2 (a constant per unrolled step for even MT)
            const double d = col[buf][k];
            bad |= !(d > 0.0);
            const double pinv = recip_pivot(d);
            double u[MT], ws[MT];
#pragma unroll
            for (int r = 0; r < MT; ++r)
                u[r] = col[buf][br * MT + r];             // A_ik
#pragma unroll
            for (int c = 0; c < MT; ++c)
                ws[c] = col[buf][bc * MT + c] * pinv;     // A_kj / d
            // row k and column k by the same rank-1 update: with u_k = d - 1,
            // A_kj - (d - 1) A_kj / d = A_kj / d; with A_kk / d taken as
            // (d - 1) / d, A_ik - A_ik (d - 1) / d = A_ik / d; then A_kk = -1/d
            if (br == kb)
                u[rk] = d - 1.0;
            if (bc == kb)
                ws[rk] = (d - 1.0) * pinv;
#pragma unroll
            for (int r = 0; r < MT; ++r)
#pragma unroll
                for (int c = 0; c < MT; ++c)
                    x[r][c] = fma(-u[r], ws[c], x[r][c]);
            if (br == kb && bc == kb)
                x[rk][rk] = -pinv;
            // column k + 1 into the other buffer: every wave has passed this
            // pivot's barrier, so none still reads it
            if (k + 1 < m)
                publish(col[buf ^ 1], (k + 1) / MT, (rk + 1) % MT);
        }
    }
    if (bad)
        atomicExch(a.err, ENKFG_ERR_LINALG);

    // ---- dfs = m - tr(M^-1) = m + tr(-M^-1), srf with the DFS_MIN guard
    double s = 0.0;
    if (act && br == bc) {
#pragma unroll
        for (int r = 0; r < MT; ++r)
            if (br * MT + r < m)
                s += x[r][r];
    }
    s = group_sum3(s, tmp, lane);
    if (lane == 0) {
        const float fd = (float)((double)m + s);
        a.pdfs[o] = fd;
        a.psrf[o] = ((double)fd > kDfsMin) ? (float)(sqrt(trg / (double)fd) - 1.0) : 0.0f;
    }
}


// ---- a node's own matrix by the same symmetric sweep (no adjoint)
//
// As factor_trace_kernel, on the 136 lower-triangle blocks of M = I + G0 (G0
// summed over the node's segments), with an IEEE division of the pivot.
// Then, with X = -M^-1 in the blocks:
//   T   = (I + M^-1) / 2 on rows < m_d, relaxed and cast (transforms.c:814-842),
//         staged one block row at a time in LDS and written as whole rows;
//   w   = M^-1 S^T s: each block's row products and, off the diagonal, its
//         transpose's, staged in LDS by (row block, other block) and summed in
//         block order by one lane per row;
//   dfs = m - tr(M^-1), srf = sqrt(tr(G0) / dfs) - 1.
template <int MT>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(sweep_waves(MT)))) factor_sym_kernel(CalcArgs a)
{
    constexpr int MP = 16 * MT;
    __shared__ double col[2][MP];
    __shared__ double part[16][16][MT];   // [row block][other block][row in block]
    __shared__ double vs[MP];             // S^T s over the node's segments
    __shared__ double slab[MT][MP];       // MT rows of M^-1 for the T output
    __shared__ double tmp[4];
    const int lane = threadIdx.x;
    const int nt = a.ntypes;
    const int bnode = (int)blockIdx.x;
    if (bnode >= a.nbatch)
        return;
    const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
    const int m = a.m, md = a.md;
    const bool dotT = (a.mode == ENKFG_MODE_ENKF || a.mode == ENKFG_MODE_HYBRID);
    const long long off = a.offsets[(size_t)node * nt];
    const int p = (int)(a.offsets[(size_t)(node + 1) * nt] - off);
    if (!a.pertype) {
        if (lane < nt) {
            a.pnlobs[(size_t)lane * a.nnode + node] = a.counts[(size_t)node * nt + lane];
            a.pdfs[(size_t)lane * a.nnode + node] = 0.0f;
            a.psrf[(size_t)lane * a.nnode + node] = 0.0f;
        }
    }
    if (p == 0) {
        if (lane == 0) {
            a.pnlobs[(size_t)nt * a.nnode + node] = 0;
            a.pdfs[(size_t)nt * a.nnode + node] = 0.0f;
            a.psrf[(size_t)nt * a.nnode + node] = 0.0f;
        }
        // empty column: T = I, w = 0 (transforms.c:766-780)
        if (dotT)
            for (int k = lane; k < md * m; k += 192) {
                const int e1 = k / m, e2 = k - e1 * m;
                a.T[(size_t)node * md * m + k] = (e1 == e2) ? 1.0f : 0.0f;
                if (a.T64 != nullptr)
                    a.T64[(size_t)node * md * m + k] = (e1 == e2) ? 1.0 : 0.0;
            }
        for (int e = lane; e < m; e += 192) {
            a.w[(size_t)node * m + e] = 0.0f;
            if (a.w64 != nullptr)
                a.w64[(size_t)node * m + e] = 0.0;
        }
        return;
    }

    const bool act = lane < kTriBlocks;
    int br = 0;
    if (act)
        while ((br + 1) * (br + 2) / 2 <= lane)
            ++br;
    const int bc = act ? lane - br * (br + 1) / 2 : 0;

    // ---- M = I + G0 on this lane's block, tr(G0)
    double x[MT][MT];
#pragma unroll
    for (int r = 0; r < MT; ++r)
#pragma unroll
        for (int c = 0; c < MT; ++c)
            x[r][c] = (br * MT + r == bc * MT + c) ? 1.0 : 0.0;
    double trg = 0.0;
    // every lane reads whole blocks of the padded scratch (lanes past the
    // last block read block (0, 0)); the entries outside m x m, or of those
    // lanes, are selected away.  Block rows outer, segments inner (each
    // element sums its segments in order), so that one row of loads is in
    // flight beside the accumulators.
    const double* gb = a.gram + (size_t)bnode * a.nseg * MP * MP + gram_index<MT>(br * MT, bc * MT);
#pragma unroll
    for (int r = 0; r < MT; ++r) {
        const int i = br * MT + r;
        for (int sg = 0; sg < a.nseg; ++sg) {
            if (a.pertype && a.counts[(size_t)node * nt + sg] == 0)
                continue;
            const double* g = gb + (size_t)sg * MP * MP + r * MT;
#pragma unroll
            for (int c = 0; c < MT; ++c) {
                const int j = bc * MT + c;
                const double v = g[c];
                const bool in = act && i < m && j < m;
                x[r][c] += in ? v : 0.0;
                if (r == c)
                    trg += (in && i == j) ? v : 0.0;
            }
        }
    }
    trg = group_sum3(trg, tmp, lane);

    auto publish = [&](double* dst, int kb, int rk) {
        if (!act)
            return;
        if (bc == kb) {
#pragma unroll
            for (int r = 0; r < MT; ++r)
                dst[br * MT + r] = x[r][rk];
        } else if (br == kb) {
#pragma unroll
            for (int c = 0; c < MT; ++c)
                dst[bc * MT + c] = x[rk][c];
        }
    };

    // ---- the sweep: X <- -M^-1
    bool bad = false;
    publish(col[0], 0, 0);
    for (int kb = 0; kb < 16; ++kb) {
#pragma unroll
        for (int rk = 0; rk < MT; ++rk) {
            const int k = kb * MT + rk;
            if (k >= m) {
                // the remaining pivots are identity padding.  Past 8 steps a
                // loop with an early exit is not unrolled (and the blocks
                // would go to scratch): such loops skip instead
                if (MT <= 8)
                    break;
                continue;
            }
            __syncthreads();
            const int buf = k & 1;   // pivot k reads buffer k % 2 (a constant per unrolled step for even MT)
            const double d = col[buf][k];
            bad |= !(d > 0.0);
            const double pinv = 1.0 / d;
            double u[MT], ws[MT];
#pragma unroll
            for (int r = 0; r < MT; ++r)
                u[r] = col[buf][br * MT + r];             // A_ik
#pragma unroll
            for (int c = 0; c < MT; ++c)
                ws[c] = col[buf][bc * MT + c] * pinv;     // A_kj / d
            // row and column k folded into the rank-1 update as in
            // factor_trace_kernel (about d eps of extra absolute rounding in
            // them); the diagonal, whose folded form would lose d^2 eps, is set
            if (br == kb)
                u[rk] = d - 1.0;
            if (bc == kb)
                ws[rk] = (d - 1.0) * pinv;
#pragma unroll
            for (int r = 0; r < MT; ++r)
#pragma unroll
                for (int c = 0; c < MT; ++c)
                    x[r][c] = fma(-u[r], ws[c], x[r][c]);
            if (br == kb && bc == kb)
                x[rk][rk] = -pinv;
            if (k + 1 < m)
                publish(col[buf ^ 1], (k + 1) / MT, (rk + 1) % MT);
        }
    }
    if (bad)
        atomicExch(a.err, ENKFG_ERR_LINALG);

    // S^T s into LDS (ordered before its reads by group_sum3's barriers)
    if (lane < MP) {
        double v = 0.0;
        if (lane < m)
            for (int sg = 0; sg < a.nseg; ++sg) {
                if (a.pertype && a.counts[(size_t)node * nt + sg] == 0)
                    continue;
                v += a.sts[((size_t)bnode * a.nseg + sg) * MP + lane];
            }
        vs[lane] = v;
    }

    // ---- dfs = m - tr(M^-1) = m + tr(X), srf
    double sd = 0.0;
    if (act && br == bc) {
#pragma unroll
        for (int r = 0; r < MT; ++r)
            if (br * MT + r < m)
                sd += x[r][r];
    }
    sd = group_sum3(sd, tmp, lane);
    if (lane == 0) {
        const float fd = (float)((double)m + sd);
        a.pnlobs[(size_t)nt * a.nnode + node] = p;
        a.pdfs[(size_t)nt * a.nnode + node] = fd;
        a.psrf[(size_t)nt * a.nnode + node] = (float)(sqrt(trg / (double)fd) - 1.0);
    }

    // ---- w = M^-1 S^T s
    if (act) {
#pragma unroll
        for (int r = 0; r < MT; ++r) {
            double y = 0.0;
#pragma unroll
            for (int c = 0; c < MT; ++c)
                y = fma(-x[r][c], vs[bc * MT + c], y);
            part[br][bc][r] = y;
        }
        if (br != bc) {
#pragma unroll
            for (int c = 0; c < MT; ++c) {
                double y = 0.0;
#pragma unroll
                for (int r = 0; r < MT; ++r)
                    y = fma(-x[r][c], vs[br * MT + r], y);
                part[bc][br][c] = y;
            }
        }
    }
    __syncthreads();
    if (lane < m) {
        const int rb = lane / MT, rr = lane - rb * MT;
        double y = 0.0;
        for (int sl = 0; sl < 16; ++sl)
            y += part[rb][sl][rr];
        a.w[(size_t)node * m + lane] = (float)y;
        if (a.w64 != nullptr)
            a.w64[(size_t)node * m + lane] = y;
    }

    // ---- DEnKF T = (I + M^-1) / 2, relaxation T <- (T - I) alpha + I, FP32:
    // M^-1 one block row at a time through LDS (rows R MT.., from the blocks
    // (R, bc) and, transposed, (br, R)), then written as whole rows of T
    if (dotT) {
        for (int R = 0; R < 16 && R * MT < md; ++R) {
            __syncthreads();   // the previous slab has been written out
            if (act && br == R) {
#pragma unroll
                for (int r = 0; r < MT; ++r)
#pragma unroll
                    for (int c = 0; c < MT; ++c)
                        slab[r][bc * MT + c] = -x[r][c];
            }
            if (act && bc == R && br != R) {
#pragma unroll
                for (int c = 0; c < MT; ++c)
#pragma unroll
                    for (int r = 0; r < MT; ++r)
                        slab[c][br * MT + r] = -x[r][c];
            }
            __syncthreads();
            for (int k = lane; k < MT * m; k += 192) {
                const int rr = k / m, e2 = k - rr * m;
                const int e1 = R * MT + rr;
                if (e1 < md) {
                    const double dlt = (e1 == e2) ? 1.0 : 0.0;
                    double t = 0.5 * slab[rr][e2] + 0.5 * dlt;
                    t = (t - dlt) * a.alpha + dlt;
                    const size_t o = ((size_t)node * md + e1) * m + e2;
                    a.T[o] = (float)t;
                    if (a.T64 != nullptr)
                        a.T64[o] = t;
                }
            }
        }
    }
}

// ---- ETKF, m <= 96: T = M^-1/2 by the symmetric sweep
//
// M = I + G0 and every iterate below are symmetric, so as in
// factor_sym_kernel only the blocks on and below the diagonal are held, one
// per lane, here 4 x 4 blocks of a G x G grid (4 G >= m; 300 lanes for
// m = 96), so that four matrices fit a lane's registers.  M^-1/2 comes
// from the scaled product form of the Denman-Beavers iteration (Higham,
// "Functions of Matrices", SIAM 2008, sec. 6.3):
//   W_k = M_k^-1,  mu_k = |det M_k|^(-1/(2m)),
//   Z_{k+1} = mu_k / 2 Z_k (I + mu_k^-2 W_k),  Z_0 = I   (-> M^-1/2)
//   M_{k+1} = (I + (mu_k^2 M_k + mu_k^-2 W_k) / 2) / 2,   M_0 = M   (-> I)
// one sweep inverse and one product per step, against two inverses per
// step for the coupled form.  The product Z_k W_k (the two commute, so it
// is symmetric) is formed block by block: for each block column k the
// lanes holding its blocks publish them (and, from the upper triangle,
// their transposes) in LDS, and every lane adds Z(br, k) W(k, bc).  The
// first inverse gives w, dfs and srf as in factor_sym_kernel; the step
// after the change of Z drops below 1e-10 of its norm is taken unscaled
// and is the last, as in factor_kernel.  Entries outside m x m are kept as
// identity.
constexpr int kEtkfBS = 4;

template <int G>
struct EtkfGeom {
    static constexpr int MP = kEtkfBS * G;
    static constexpr int NB = G * (G + 1) / 2;      // lower-triangle blocks
    static constexpr int NT = 64 * ((NB + 63) / 64);
};

// element (i, j) of a segment's Gram matrix in the scratch layout of the
// Gram kernels (gram_index with B = MT blocks of a 16 x 16 grid)
__device__ inline size_t gram_at(int MT, int i, int j)
{
    return ((size_t)((i / MT) * 16 + j / MT) * MT + (i % MT)) * MT + (j % MT);
}

// sum over the NT threads in a fixed order (waves, then the wave totals in
// order); every thread receives it
template <int NT>
__device__ inline double group_sum(double v, double* tmp, int lane)
{
    for (int o = 1; o < 64; o <<= 1)
        v += __shfl_xor(v, o);
    __syncthreads();
    if ((lane & 63) == 0)
        tmp[lane >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < NT / 64; ++q)
        s += tmp[q];
    __syncthreads();
    return s;
}

// x <- -x^-1 by the sweep (x symmetric positive definite in its m x m part,
// identity outside); returns log det x and flags a pivot that is not
// positive
template <int G>
__device__ inline double etkf_sweep(double (&x)[kEtkfBS][kEtkfBS], double (*col)[EtkfGeom<G>::MP], int m, int br,
                                    int bc, bool act, bool& bad)
{
    constexpr int BS = kEtkfBS;
    auto publish = [&](double* dst, int kb, int rk) {
        if (!act)
            return;
        if (bc == kb) {
#pragma unroll
            for (int r = 0; r < BS; ++r)
                dst[br * BS + r] = x[r][rk];
        } else if (br == kb) {
#pragma unroll
            for (int c = 0; c < BS; ++c)
                dst[bc * BS + c] = x[rk][c];
        }
    };
    double mant = 1.0;   // det = mant * 2^ex
    int ex = 0;
    __syncthreads();     // the previous sweep's last column has been read
    publish(col[0], 0, 0);
#pragma unroll 1
    for (int kb = 0; kb < G; ++kb) {
        if (kb * BS >= m)
            break;
#pragma unroll
        for (int rk = 0; rk < BS; ++rk) {
            const int k = kb * BS + rk;
            if (k >= m)
                break;
            __syncthreads();
            const int buf = k & 1;
            const double d = col[buf][k];
            bad |= !(d > 0.0);
            int e;
            mant = frexp(mant * d, &e);
            ex += e;
            const double pinv = 1.0 / d;
            double u[BS], ws[BS];
#pragma unroll
            for (int r = 0; r < BS; ++r)
                u[r] = col[buf][br * BS + r];
#pragma unroll
            for (int c = 0; c < BS; ++c)
                ws[c] = col[buf][bc * BS + c] * pinv;
            if (br == kb)
                u[rk] = d - 1.0;
            if (bc == kb)
                ws[rk] = (d - 1.0) * pinv;
#pragma unroll
            for (int r = 0; r < BS; ++r)
#pragma unroll
                for (int c = 0; c < BS; ++c)
                    x[r][c] = fma(-u[r], ws[c], x[r][c]);
            if (br == kb && bc == kb)
                x[rk][rk] = -pinv;
            if (k + 1 < m)   // a compile-time block column per unrolled step
                publish(col[buf ^ 1], rk + 1 < BS ? kb : kb + 1, (rk + 1) % BS);
        }
    }
    return log(mant) + (double)ex * 0.69314718055994530942;
}

template <int G>
__global__ void __launch_bounds__(EtkfGeom<G>::NT) __attribute__((amdgpu_waves_per_eu(2)))
factor_etkf_kernel(CalcArgs a)
{
    constexpr int BS = kEtkfBS;
    constexpr int MP = EtkfGeom<G>::MP, NB = EtkfGeom<G>::NB, NT = EtkfGeom<G>::NT;
    __shared__ double col[2][MP];
    __shared__ double zc[2][G][BS * BS];   // block column k of Z, block i at [i]
    __shared__ double wc[2][G][BS * BS];   // ... of W
    __shared__ double part[G][G][BS];      // w: [row block][other block][row in block]
    __shared__ double vs[MP];
    __shared__ double slab[BS][MP];        // BS rows of Z for the T output
    __shared__ double mst[BS * BS][NT];    // M_{k+1} during the product, element-major
    __shared__ double tmp[NT / 64];
    __shared__ int flag;
    const int lane = threadIdx.x;
    const int nt = a.ntypes;
    const int bnode = (int)blockIdx.x;
    if (bnode >= a.nbatch)
        return;
    const int node = a.node_list ? a.node_list[a.node0 + bnode] : a.node0 + bnode;
    const int m = a.m, md = a.md;
    const int MT = (m + 15) / 16;   // the Gram scratch's block size
    const int MPG = 16 * MT;
    const long long off = a.offsets[(size_t)node * nt];
    const int p = (int)(a.offsets[(size_t)(node + 1) * nt] - off);
    if (!a.pertype) {
        if (lane < nt) {
            a.pnlobs[(size_t)lane * a.nnode + node] = a.counts[(size_t)node * nt + lane];
            a.pdfs[(size_t)lane * a.nnode + node] = 0.0f;
            a.psrf[(size_t)lane * a.nnode + node] = 0.0f;
        }
    }
    if (p == 0) {
        if (lane == 0) {
            a.pnlobs[(size_t)nt * a.nnode + node] = 0;
            a.pdfs[(size_t)nt * a.nnode + node] = 0.0f;
            a.psrf[(size_t)nt * a.nnode + node] = 0.0f;
        }
        // empty column: T = I, w = 0 (transforms.c:766-780)
        for (int k = lane; k < md * m; k += NT) {
            const int e1 = k / m, e2 = k - e1 * m;
            a.T[(size_t)node * md * m + k] = (e1 == e2) ? 1.0f : 0.0f;
            if (a.T64 != nullptr)
                a.T64[(size_t)node * md * m + k] = (e1 == e2) ? 1.0 : 0.0;
        }
        for (int e = lane; e < m; e += NT) {
            a.w[(size_t)node * m + e] = 0.0f;
            if (a.w64 != nullptr)
                a.w64[(size_t)node * m + e] = 0.0;
        }
        return;
    }
    if (lane == 0)
        flag = 0;

    const bool act = lane < NB;
    int br = 0;
    if (act)
        while ((br + 1) * (br + 2) / 2 <= lane)
            ++br;
    const int bc = act ? lane - br * (br + 1) / 2 : 0;
    // entry (r, c) of this lane's block lies in the m x m part
    auto inside = [&](int r, int c) { return act && br * BS + r < m && bc * BS + c < m; };

    // ---- M = I + G0 (summed over the node's segments), tr(G0)
    double mo[BS][BS];
    double trg = 0.0;
    const double* gb = a.gram + (size_t)bnode * a.nseg * MPG * MPG;
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int i = br * BS + r, j = bc * BS + c;
            double v = (i == j) ? 1.0 : 0.0;
            if (inside(r, c))
                for (int sg = 0; sg < a.nseg; ++sg) {
                    if (a.pertype && a.counts[(size_t)node * nt + sg] == 0)
                        continue;
                    const double g = gb[(size_t)sg * MPG * MPG + gram_at(MT, i, j)];
                    v += g;
                    if (i == j)
                        trg += g;
                }
            mo[r][c] = v;
        }
    trg = group_sum<NT>(trg, tmp, lane);

    // ---- x = -M^-1
    double x[BS][BS];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c)
            x[r][c] = mo[r][c];
    bool bad = false;
    double logdet = etkf_sweep<G>(x, col, m, br, bc, act, bad);

    // S^T s into LDS, dfs = m - tr(M^-1), srf
    if (lane < MP) {
        double v = 0.0;
        if (lane < m)
            for (int sg = 0; sg < a.nseg; ++sg) {
                if (a.pertype && a.counts[(size_t)node * nt + sg] == 0)
                    continue;
                v += a.sts[((size_t)bnode * a.nseg + sg) * MPG + lane];
            }
        vs[lane] = v;
    }
    double sd = 0.0;
    if (act && br == bc) {
#pragma unroll
        for (int r = 0; r < BS; ++r)
            if (br * BS + r < m)
                sd += x[r][r];
    }
    sd = group_sum<NT>(sd, tmp, lane);   // (its barriers also order the vs stores before the reads)
    if (lane == 0) {
        const float fd = (float)((double)m + sd);
        a.pnlobs[(size_t)nt * a.nnode + node] = p;
        a.pdfs[(size_t)nt * a.nnode + node] = fd;
        a.psrf[(size_t)nt * a.nnode + node] = (float)(sqrt(trg / (double)fd) - 1.0);
    }
    // ---- w = M^-1 S^T s from the blocks' row products, summed in block order
    if (act) {
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            double y = 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c)
                y = fma(-x[r][c], vs[bc * BS + c], y);
            part[br][bc][r] = y;
        }
        if (br != bc) {
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                double y = 0.0;
#pragma unroll
                for (int r = 0; r < BS; ++r)
                    y = fma(-x[r][c], vs[br * BS + r], y);
                part[bc][br][c] = y;
            }
        }
    }
    __syncthreads();
    if (lane < m) {
        const int rb = lane / BS, rr = lane - rb * BS;
        double y = 0.0;
        for (int sl = 0; sl < G; ++sl)
            y += part[rb][sl][rr];
        a.w[(size_t)node * m + lane] = (float)y;
        if (a.w64 != nullptr)
            a.w64[(size_t)node * m + lane] = y;
    }

    // ---- the iteration: mo = M_k, x = -W_k, z = Z_k; M_{k+1} is formed as
    // soon as W_k is known and waits in LDS while the product runs
    double z[BS][BS];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c)
            z[r][c] = (br * BS + r == bc * BS + c) ? 1.0 : 0.0;
    bool last = false;
    for (int it = 0; it < 60; ++it) {
        // the block coordinates, opaque to the compiler in each step: the
        // per-element constants derived from them are formed where they are
        // used rather than held in registers across the steps
        int ib = br, jb = bc;
        asm volatile("" : "+v"(ib), "+v"(jb));
        auto in_mm = [&](int r, int c) { return act && ib * BS + r < m && jb * BS + c < m; };
        auto delta = [&](int r, int c) { return (ib * BS + r == jb * BS + c) ? 1.0 : 0.0; };
        const double mu = last ? 1.0 : exp(-logdet / (2.0 * m));
        const double hm = 0.5 * mu, hi = 0.5 / mu, qm = 0.25 * mu * mu, qi = 0.25 / (mu * mu);
        // M_{k+1} = 1/2 I + mu^2/4 M_k + 1/(4 mu^2) W_k (identity outside m x m)
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const double dlt = delta(r, c);
                mst[r * BS + c][lane] = in_mm(r, c) ? 0.5 * dlt + qm * mo[r][c] - qi * x[r][c] : dlt;
            }
        // pc = Z_k (-W_k): block column k of Z and of x, k = 0 .. G-1
        double pc[BS][BS];
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < BS; ++c)
                pc[r][c] = 0.0;
#pragma unroll 1
        for (int k = 0; k < G; ++k) {
            const int buf = k & 1;
            if (act && bc == k) {          // block (br, k), br >= k
#pragma unroll
                for (int r = 0; r < BS; ++r)
#pragma unroll
                    for (int c = 0; c < BS; ++c) {
                        zc[buf][br][r * BS + c] = z[r][c];
                        wc[buf][br][r * BS + c] = x[r][c];
                    }
            } else if (act && br == k) {   // block (bc, k) = block (k, bc)^T, bc < k
#pragma unroll
                for (int r = 0; r < BS; ++r)
#pragma unroll
                    for (int c = 0; c < BS; ++c) {
                        zc[buf][bc][c * BS + r] = z[r][c];
                        wc[buf][bc][c * BS + r] = x[r][c];
                    }
            }
            __syncthreads();
            if (act) {
                // Z(br, k) (-W)(k, bc), with (-W)(k, bc) = (-W)(bc, k)^T
#pragma unroll
                for (int q = 0; q < BS; ++q) {
                    double zr[BS], wq[BS];
#pragma unroll
                    for (int r = 0; r < BS; ++r)
                        zr[r] = zc[buf][br][r * BS + q];
#pragma unroll
                    for (int c = 0; c < BS; ++c)
                        wq[c] = wc[buf][bc][c * BS + q];
#pragma unroll
                    for (int r = 0; r < BS; ++r)
#pragma unroll
                        for (int c = 0; c < BS; ++c)
                            pc[r][c] = fma(zr[r], wq[c], pc[r][c]);
                }
            }
        }
        // the diagonal blocks of the symmetric product, symmetrised
        if (br == bc) {
#pragma unroll
            for (int r = 0; r < BS; ++r)
#pragma unroll
                for (int c = 0; c < r; ++c) {
                    const double sy = 0.5 * (pc[r][c] + pc[c][r]);
                    pc[r][c] = sy;
                    pc[c][r] = sy;
                }
        }
        // Z_{k+1} = mu/2 Z + 1/(2 mu) Z W
        double dn = 0.0, nn = 0.0;
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const bool in = in_mm(r, c);
                const double zn = in ? hm * z[r][c] - hi * pc[r][c] : delta(r, c);
                const double dz = zn - z[r][c];
                // off-diagonal blocks stand for themselves and their transposes
                const double wgt = (ib == jb) ? 1.0 : 2.0;
                dn += in ? wgt * dz * dz : 0.0;
                nn += in ? wgt * zn * zn : 0.0;
                z[r][c] = zn;
            }
        dn = group_sum<NT>(dn, tmp, lane);   // (its barriers also order the mst stores before the reads)
        nn = group_sum<NT>(nn, tmp, lane);
        if (last)
            break;
        // a relative change of 1e-8: the iteration converges quadratically,
        // so one more unscaled step leaves Z within rounding of its limit
        // (1e-10: one more iteration of the pass, about the same T)
        if (dn <= 1e-16 * nn)
            last = true;
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                mo[r][c] = mst[r * BS + c][lane];
                x[r][c] = mo[r][c];
            }
        logdet = etkf_sweep<G>(x, col, m, br, bc, act, bad);
    }
    if (!last || bad)
        flag = 1;   // a non-positive pivot, or no convergence in 60 steps: reported
    __syncthreads();
    if (lane == 0 && flag)
        atomicExch(a.err, ENKFG_ERR_LINALG);

    // ---- T = Z, relaxation T <- (T - I) alpha + I, FP32: Z one block row at a
    // time through LDS (rows R BS.., from the blocks (R, bc) and, transposed,
    // (br, R)), then written as whole rows of T
    for (int R = 0; R < G && R * BS < md; ++R) {
        __syncthreads();   // the previous slab has been written out
        if (act && br == R) {
#pragma unroll
            for (int r = 0; r < BS; ++r)
#pragma unroll
                for (int c = 0; c < BS; ++c)
                    slab[r][bc * BS + c] = z[r][c];
        }
        if (act && bc == R && br != R) {
#pragma unroll
            for (int c = 0; c < BS; ++c)
#pragma unroll
                for (int r = 0; r < BS; ++r)
                    slab[c][br * BS + r] = z[r][c];
        }
        __syncthreads();
        for (int k = lane; k < BS * m; k += NT) {
            const int rr = k / m, e2 = k - rr * m;
            const int e1 = R * BS + rr;
            if (e1 < md) {
                const double dlt = (e1 == e2) ? 1.0 : 0.0;
                const double t = (slab[rr][e2] - dlt) * a.alpha + dlt;
                const size_t o = ((size_t)node * md + e1) * m + e2;
                a.T[o] = (float)t;
                if (a.T64 != nullptr)
                    a.T64[o] = t;
            }
        }
    }
}

}  // namespace

bool factor_wave_supported(const CalcArgs& a)
{
    const bool etkf = a.scheme == ENKFG_SCHEME_ETKF && a.mode != ENKFG_MODE_ENOI;
    return !etkf && a.m <= 160;
}

hipError_t launch_factor_wave(hipStream_t s, const CalcArgs& a)
{
    if (a.nbatch == 0)
        return hipSuccess;
    const int MT = (a.m + 15) / 16;
    const dim3 grid((unsigned)a.nbatch);
    if (a.adj_g != nullptr) {
        // adjoint: z = M^-1 g and w by the Gauss-Jordan elimination
        switch (MT) {
        case 1: hipLaunchKernelGGL(factor_wave_kernel<1>, grid, dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(factor_wave_kernel<2>, grid, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL(factor_wave_kernel<3>, grid, dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL(factor_wave_kernel<4>, grid, dim3(256), 0, s, a); break;
        case 5: hipLaunchKernelGGL(factor_wave_kernel<5>, grid, dim3(256), 0, s, a); break;
        case 6: hipLaunchKernelGGL(factor_wave_kernel<6>, grid, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (MT) {
        case 1: hipLaunchKernelGGL(factor_sym_kernel<1>, grid, dim3(192), 0, s, a); break;
        case 2: hipLaunchKernelGGL(factor_sym_kernel<2>, grid, dim3(192), 0, s, a); break;
        case 3: hipLaunchKernelGGL(factor_sym_kernel<3>, grid, dim3(192), 0, s, a); break;
        case 4: hipLaunchKernelGGL(factor_sym_kernel<4>, grid, dim3(192), 0, s, a); break;
        case 5: hipLaunchKernelGGL(factor_sym_kernel<5>, grid, dim3(192), 0, s, a); break;
        case 6: hipLaunchKernelGGL(factor_sym_kernel<6>, grid, dim3(192), 0, s, a); break;
        case 7: hipLaunchKernelGGL(factor_sym_kernel<7>, grid, dim3(192), 0, s, a); break;
        case 8: hipLaunchKernelGGL(factor_sym_kernel<8>, grid, dim3(192), 0, s, a); break;
        case 9: hipLaunchKernelGGL(factor_sym_kernel<9>, grid, dim3(192), 0, s, a); break;
        case 10: hipLaunchKernelGGL(factor_sym_kernel<10>, grid, dim3(192), 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.pertype)
        return e;
    return launch_factor_trace(s, a);
}

hipError_t launch_factor_trace(hipStream_t s, const CalcArgs& a)
{
    if (a.nbatch == 0 || a.ntypes <= 0)
        return hipSuccess;
    const int MT = (a.m + 15) / 16;
    // per-obstype diagnostics, one work item per (node, obstype)
    const dim3 tgrid((unsigned)((long long)a.nbatch * a.ntypes)), tblock(192);
    switch (MT) {
    case 1: hipLaunchKernelGGL(factor_trace_kernel<1>, tgrid, tblock, 0, s, a); break;
    case 2: hipLaunchKernelGGL(factor_trace_kernel<2>, tgrid, tblock, 0, s, a); break;
    case 3: hipLaunchKernelGGL(factor_trace_kernel<3>, tgrid, tblock, 0, s, a); break;
    case 4: hipLaunchKernelGGL(factor_trace_kernel<4>, tgrid, tblock, 0, s, a); break;
    case 5: hipLaunchKernelGGL(factor_trace_kernel<5>, tgrid, tblock, 0, s, a); break;
    case 6: hipLaunchKernelGGL(factor_trace_kernel<6>, tgrid, tblock, 0, s, a); break;
    case 7: hipLaunchKernelGGL(factor_trace_kernel<7>, tgrid, tblock, 0, s, a); break;
    case 8: hipLaunchKernelGGL(factor_trace_kernel<8>, tgrid, tblock, 0, s, a); break;
    case 9: hipLaunchKernelGGL(factor_trace_kernel<9>, tgrid, tblock, 0, s, a); break;
    case 10: hipLaunchKernelGGL(factor_trace_kernel<10>, tgrid, tblock, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_factor_etkf(hipStream_t s, const CalcArgs& a)
{
    if (a.nbatch == 0)
        return hipSuccess;
    const dim3 grid((unsigned)a.nbatch);
    const int g = (a.m + kEtkfBS - 1) / kEtkfBS;   // blocks per side
    if (g <= 8)
        hipLaunchKernelGGL(factor_etkf_kernel<8>, grid, dim3(EtkfGeom<8>::NT), 0, s, a);
    else if (g <= 12)
        hipLaunchKernelGGL(factor_etkf_kernel<12>, grid, dim3(EtkfGeom<12>::NT), 0, s, a);
    else if (g <= 16)
        hipLaunchKernelGGL(factor_etkf_kernel<16>, grid, dim3(EtkfGeom<16>::NT), 0, s, a);
    else if (g <= 20)
        hipLaunchKernelGGL(factor_etkf_kernel<20>, grid, dim3(EtkfGeom<20>::NT), 0, s, a);
    else if (g <= 24)
        hipLaunchKernelGGL(factor_etkf_kernel<24>, grid, dim3(EtkfGeom<24>::NT), 0, s, a);
    else
        return hipErrorInvalidValue;
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || !a.pertype)
        return e;
    return launch_factor_trace(s, a);   // per-type diagnostics, after the node's own
}

}  // namespace enkfg
