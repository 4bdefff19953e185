// loc_kernels.hip -- GPU localisation: for every node of the stride grid, the
// ordered list of (obs id, taper coefficient) that obs_findlocal
// (common/observations.c:1837-1927) returns.
//
// Each thread walks the per-type kd-trees built on the host by sequential
// insertion (the trees of obs_createkdtrees).  The walk is the preorder of
// _kd_findnodeswithinrange (common/kdtree.c:516-540) -- node, then the child
// on the query's side (left iff dx <= 0), then the other child iff |dx| <
// range -- done without a stack by following parent links: on the way up, a
// node is re-examined to decide whether the far child is still to be
// visited, which depends only on that node and the query.  Lists are
// therefore identical, element for element, to the reference's, including
// their order.  Squared distances, tapers and the inclusion test use the
// reference's operation order in IEEE double with no contraction (the
// library is compiled with -ffp-contract=off), so the taper bits match too.
//
// Two passes: count (per node and type), exclusive scan, fill.  The obs
// indices in the trees and in the lists are device rows (rows of S in tree
// preorder); the host maps them back to observation indices.
#include "enkfg_internal.h"

namespace enkfg {

namespace {

constexpr int kStatusOK = 0;

__device__ inline double taper_gc(double x)
{
    // common/utils.c:1425-1443
    if (x >= 1.0)
        return 0.0;
    x *= 2.0;
    double x2 = x * x;
    double x3 = x2 * x;
    double f;
    if (x < 1.0)
        f = 1.0 + x2 * (-x3 / 4.0 + x2 / 2.0) + x3 * (5.0 / 8.0) - x2 * (5.0 / 3.0);
    else
        f = x2 * (x3 / 12.0 - x2 / 2.0) + x3 * (5.0 / 8.0) + x2 * (5.0 / 3.0) - x * 5.0 + 4.0 - (2.0 / 3.0) / x;
    return (f >= 0.0) ? f : 0.0;
}

__device__ inline double calclcoeff(const TypeInfo& t, double dist)
{
    // common/obstypes.c:800-810
    double sum = 0.0;
    for (int i = 0; i < t.nlocrad; ++i)
        if (dist <= t.locrad[i])
            sum += t.locweight[i] * taper_gc(dist / t.locrad[i]);
    return sum;
}

// Depth-first range search over one tree; calls emit(node, distsq) for every
// hit in the reference's visiting order.  The split dimension is the depth
// modulo ND, tracked on the way down and up.  Each step selects the next
// node without branching (down to the near child if there is one, else to
// the far child if the range crosses the split, else up; coming up from the
// near child, to the far child under the same condition, else further up),
// so that the lanes of a wave at different stages of their walks stay in
// one instruction stream.
template <int ND, class Emit>
__device__ inline void range_walk_nd(const KdNode* __restrict__ tree, const double* q, double range, Emit&& emit)
{
    const double r2 = range * range;
    const double q0 = q[0], q1 = q[1], q2 = ND == 3 ? q[2] : 0.0;
    int node = 0;
    int from = -1;
    int dir = 0;
    bool down = true;

    while (node >= 0) {
        const KdNode nd = tree[node];
        const double cd = dir == 0 ? nd.x : ((ND == 2 || dir == 1) ? nd.y : nd.z);
        const double qd = dir == 0 ? q0 : ((ND == 2 || dir == 1) ? q1 : q2);
        const double dx = qd - cd;
        if (down) {
            // kd_findnodeswithinrange's sum, in its order (kdtree.c:529-533)
            double distsq = 0.0;
            double d = nd.x - q0;
            distsq += d * d;
            d = nd.y - q1;
            distsq += d * d;
            if (ND == 3) {
                d = nd.z - q2;
                distsq += d * d;
            }
            if (distsq <= r2)
                emit(node, distsq);
        }
        const int left = (nd.link & 0x80000000u) ? node + 1 : -1;
        const bool lside = dx <= 0.0;
        const int near = lside ? left : nd.right;
        const int far = lside ? nd.right : left;
        const int parent = (int)(nd.link & 0x7fffffffu) - 1;
        const bool farok = fabs(dx) < range && far >= 0;
        const bool tonear = down && near >= 0;
        const bool tofar = !tonear && farok && (down || from == near);
        const bool ndown = tonear || tofar;
        from = node;
        node = tonear ? near : (tofar ? far : parent);
        dir = ndown ? (dir + 1 == ND ? 0 : dir + 1) : (dir == 0 ? ND - 1 : dir - 1);
        down = ndown;
    }
}

template <class Emit>
__device__ inline void range_walk(const KdNode* __restrict__ tree, int ndim, const double* q, double range, Emit&& emit)
{
    if (ndim == 3)
        range_walk_nd<3>(tree, q, range, emit);
    else
        range_walk_nd<2>(tree, q, range, emit);
}

// Thread mapping: one thread per (node, obstype); a wave takes an 8 x 8 tile
// of the stride grid and one obstype, so its lanes walk nearly the same
// paths of one tree and their node loads coalesce.  ni is the grid's row
// length (nnode = ni nj); tiles are row-major within a type.
// consecutive walk items per XCD run: four 8 x 8 blocks of tiles (runs of
// one block: localisation 103-105 ms on the tile, of four 101-102 ms, of
// sixteen 135 ms: the XCDs with the densest regions then finish last)
constexpr int kWalkRun = 256;

__device__ inline bool thread_node(int nnode, int ni, int ntypes, int& t, int& node)
{
    const int nj = nnode / ni;
    const int ti = (ni + 7) / 8, tj = (nj + 7) / 8;
    const int sti = (ti + 7) / 8, stj = (tj + 7) / 8;   // 8 x 8 blocks of tiles
    const long long per_type = (long long)sti * stj * 64 * 64;
    // Tiles are ordered in 8 x 8 blocks of tiles (64 x 64 nodes), row-major
    // within and between blocks.  Blocks go to the 8 XCDs round-robin; each
    // XCD takes runs of kWalkRun consecutive (type, tile) items -- four
    // blocks of tiles -- so the walks in flight on one L2 cover a compact patch of
    // the grid, and the runs are dealt round-robin, so dense and sparse
    // regions spread over all XCDs (walk_grid rounds the grid up to a
    // multiple of 8 runs; items past the grid's edge return)
    const unsigned s = blockIdx.x / 8, x = blockIdx.x % 8;
    const long long item = ((long long)(s / kWalkRun) * 8 + x) * kWalkRun + s % kWalkRun;
    const long long gid = item * blockDim.x + threadIdx.x;
    if (gid >= per_type * ntypes)
        return false;
    t = (int)(gid / per_type);
    const long long rem = gid - (long long)t * per_type;
    const int tile = (int)(rem / 64), lane = (int)(rem % 64);
    const int sb = tile / 64, wi = tile % 64;
    const int tx = (sb % sti) * 8 + wi % 8, ty = (sb / sti) * 8 + wi / 8;
    if (tx >= ti || ty >= tj)
        return false;
    const int i = tx * 8 + lane % 8, j = ty * 8 + lane / 8;
    if (i >= ni || j >= nj)
        return false;
    node = j * ni + i;
    return true;
}

__global__ void __launch_bounds__(64) loc_count_kernel(int nnode, int ni, const double* __restrict__ qxyz,
                                                        const unsigned char* __restrict__ isglobal,
                                                        const unsigned char* __restrict__ visible,
                                                        const KdNode* __restrict__ nodes,
                                                        const TypeInfo* __restrict__ types, int ntypes,
                                                        const short* __restrict__ status, int* __restrict__ counts,
                                                        int* __restrict__ raw)
{
    int t, node;
    if (!thread_node(nnode, ni, ntypes, t, node))
        return;
    double q[3] = {qxyz[3 * (size_t)node], qxyz[3 * (size_t)node + 1], qxyz[3 * (size_t)node + 2]};
    const bool glob = isglobal[node] != 0;
    const TypeInfo& ti = types[t];
    int n = 0, nraw = 0;
    if (ti.searchable && (visible == nullptr || visible[t])) {
        if (glob) {
            n = ti.nids;   // obs_getglobal: all obs of the type, no status check
        } else if (ti.nnode > 0 && ti.nlobsmax != ENKFG_NLOBSMAX_NONE) {
            // NLOBSMAX: every hit is a candidate; the sort kernel selects
            range_walk(nodes + ti.node0, ti.ndim, q, ti.range, [&](int, double) { nraw++; });
        } else if (ti.nnode > 0) {
            range_walk(nodes + ti.node0, ti.ndim, q, ti.range, [&](int v, double) {
                if (status[ti.node0 + v] == kStatusOK)
                    n++;
            });
        }
    }
    counts[(size_t)node * ntypes + t] = n;
    if (raw != nullptr)
        raw[(size_t)node * ntypes + t] = nraw;
}

// NLOBSMAX candidates of a (node, type): rows and squared distances in the
// reference's visiting order
__global__ void __launch_bounds__(64) loc_rawfill_kernel(int nnode, int ni, const double* __restrict__ qxyz,
                                                         const unsigned char* __restrict__ isglobal,
                                                         const unsigned char* __restrict__ visible,
                                                         const KdNode* __restrict__ nodes,
                                                         const TypeInfo* __restrict__ types, int ntypes,
                                                         const long long* __restrict__ rawoff, int* __restrict__ rrow,
                                                         double* __restrict__ rd2)
{
    int t, node;
    if (!thread_node(nnode, ni, ntypes, t, node))
        return;
    const TypeInfo& ti = types[t];
    if (isglobal[node] || !ti.searchable || !(visible == nullptr || visible[t]) || ti.nnode == 0 ||
        ti.nlobsmax == ENKFG_NLOBSMAX_NONE)
        return;
    double q[3] = {qxyz[3 * (size_t)node], qxyz[3 * (size_t)node + 1], qxyz[3 * (size_t)node + 2]};
    long long o = rawoff[(size_t)node * ntypes + t];
    range_walk(nodes + ti.node0, ti.ndim, q, ti.range, [&](int v, double distsq) {
        rrow[o] = ti.node0 + v;
        rd2[o] = distsq;
        o++;
    });
}

// The candidates of one (node, type) sorted by distance, ties in visiting
// order (the stable sort of kd_findnodeswithinrange's qsort on this
// platform, common/kdtree.c:569-570), by a bitonic sort of (distsq, index)
// keys in LDS; then obs_findlocal's selection (observations.c:1899-1908):
// the first min(n, NLOBSMAX) candidates, skipping those whose status is not
// OK.  Writes the sorted candidates and the list length.  Lists of more
// than kSortMax candidates are loc_sort_big_kernel's.
constexpr int kSortMax = 8192;

__global__ void __launch_bounds__(256) loc_sort_kernel(int nnode, const TypeInfo* __restrict__ types, int ntypes,
                                                       const short* __restrict__ status,
                                                       const int* __restrict__ raw,
                                                       const long long* __restrict__ rawoff,
                                                       const int* __restrict__ rrow, const double* __restrict__ rd2,
                                                       int* __restrict__ srow, double* __restrict__ sd2,
                                                       int* __restrict__ counts)
{
    __shared__ double kd[kSortMax];
    __shared__ int ki[kSortMax];
    __shared__ int cnt;
    const long long item = blockIdx.x;
    if (item >= (long long)nnode * ntypes)
        return;
    const int n = raw[item];
    if (n == 0 || n > kSortMax)
        return;
    const TypeInfo& ti = types[item % ntypes];
    const long long off = rawoff[item];
    int P = 1;
    while (P < n)
        P <<= 1;
    for (int k = threadIdx.x; k < P; k += 256) {
        kd[k] = (k < n) ? rd2[off + k] : INFINITY;
        ki[k] = k;
    }
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < P / 2; k += 256) {
                const int lo = 2 * stride * (k / stride) + (k % stride);
                const int hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const double a = kd[lo], b = kd[hi];
                const int ia = ki[lo], ib = ki[hi];
                const bool agt = (a > b) || (a == b && ia > ib);
                if (agt == up) {
                    kd[lo] = b;
                    kd[hi] = a;
                    ki[lo] = ib;
                    ki[hi] = ia;
                }
            }
            __syncthreads();
        }
    }
    const int nsel = min(n, ti.nlobsmax);
    int mine = 0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const int row = rrow[off + ki[k]];
        srow[off + k] = row;
        sd2[off + k] = kd[k];
        if (k < nsel && status[row] == kStatusOK)
            mine++;
    }
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0)
        counts[item] = cnt;
}

// Lists of more than kSortMax candidates (the reference has no limit): one
// workgroup sorts a (node, type)'s candidates in place in global memory by a
// bottom-up merge sort on the squared distance, stable -- equal distances
// keep their visiting order, the order loc_sort_kernel's (distance, index)
// keys give -- ping-ponging between (rd2, rrow) and (sd2, srow); each
// element finds its place in the merged run by a binary search in the other
// run.  Then the same selection.
__global__ void __launch_bounds__(256) loc_sort_big_kernel(int nnode, const TypeInfo* __restrict__ types,
                                                           int ntypes, const short* __restrict__ status,
                                                           const int* __restrict__ raw,
                                                           const long long* __restrict__ rawoff, int* __restrict__ rrow,
                                                           double* __restrict__ rd2, int* __restrict__ srow,
                                                           double* __restrict__ sd2, int* __restrict__ counts)
{
    __shared__ int cnt;
    const long long item = blockIdx.x;
    if (item >= (long long)nnode * ntypes)
        return;
    const int n = raw[item];
    if (n <= kSortMax)
        return;
    const TypeInfo& ti = types[item % ntypes];
    const long long off = rawoff[item];
    double* keys[2] = {rd2 + off, sd2 + off};
    int* rows[2] = {rrow + off, srow + off};
    int src = 0;
    for (int width = 1; width < n; width <<= 1) {
        const double* ik = keys[src];
        const int* ir = rows[src];
        double* ok = keys[src ^ 1];
        int* orow = rows[src ^ 1];
        for (int k = threadIdx.x; k < n; k += 256) {
            const int r0 = (k / (2 * width)) * (2 * width);   // the pair of runs [r0, mid) [mid, end)
            const int mid = min(r0 + width, n), end = min(r0 + 2 * width, n);
            const double key = ik[k];
            int pos;
            if (k < mid) {   // before the right run's elements >= key
                int lo = mid, hi = end;
                while (lo < hi) {
                    const int h = (lo + hi) >> 1;
                    if (ik[h] < key)
                        lo = h + 1;
                    else
                        hi = h;
                }
                pos = k + (lo - mid);
            } else {         // after the left run's elements <= key
                int lo = r0, hi = mid;
                while (lo < hi) {
                    const int h = (lo + hi) >> 1;
                    if (ik[h] <= key)
                        lo = h + 1;
                    else
                        hi = h;
                }
                pos = (k - mid) + lo;
            }
            ok[pos] = key;
            orow[pos] = ir[k];
        }
        __syncthreads();   // the pass is complete before the next reads it
        src ^= 1;
    }
    if (src == 0) {   // the sorted candidates end in (sd2, srow)
        for (int k = threadIdx.x; k < n; k += 256) {
            sd2[off + k] = rd2[off + k];
            srow[off + k] = rrow[off + k];
        }
    }
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    const int nsel = min(n, ti.nlobsmax);
    int mine = 0;
    for (int k = threadIdx.x; k < nsel; k += 256)
        if (status[srow[off + k]] == kStatusOK)
            mine++;
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0)
        counts[item] = cnt;
}

__global__ void __launch_bounds__(64) loc_fill_kernel(int nnode, int ni, const double* __restrict__ qxyz,
                                                       const unsigned char* __restrict__ isglobal,
                                                       const unsigned char* __restrict__ visible,
                                                       const KdNode* __restrict__ nodes,
                                                       const TypeInfo* __restrict__ types, int ntypes,
                                                       const short* __restrict__ status,
                                                       const int* __restrict__ obsids,
                                                       const long long* __restrict__ offsets, int* __restrict__ ids,
                                                       double* __restrict__ lc, const long long* __restrict__ rawoff,
                                                       const int* __restrict__ raw, const int* __restrict__ srow,
                                                       const double* __restrict__ sd2)
{
    int t, node;
    if (!thread_node(nnode, ni, ntypes, t, node))
        return;
    double q[3] = {qxyz[3 * (size_t)node], qxyz[3 * (size_t)node + 1], qxyz[3 * (size_t)node + 2]};
    const bool glob = isglobal[node] != 0;
    const TypeInfo& ti = types[t];
    if (!ti.searchable || !(visible == nullptr || visible[t]))
        return;
    long long o = offsets[(size_t)node * ntypes + t];
    if (glob) {
        for (int i = 0; i < ti.nids; ++i) {
            ids[o + i] = obsids[ti.obsid0 + i];
            lc[o + i] = 1.0;
        }
    } else if (ti.nnode > 0 && ti.nlobsmax != ENKFG_NLOBSMAX_NONE) {
        const size_t item = (size_t)node * ntypes + t;
        const long long r0 = rawoff[item];
        const int nsel = min(raw[item], ti.nlobsmax);
        for (int k = 0; k < nsel; ++k) {
            const int row = srow[r0 + k];
            if (status[row] == kStatusOK) {
                ids[o] = row;
                lc[o] = calclcoeff(ti, sqrt(sd2[r0 + k]));
                o++;
            }
        }
    } else if (ti.nnode > 0) {
        range_walk(nodes + ti.node0, ti.ndim, q, ti.range, [&](int v, double distsq) {
            const int row = ti.node0 + v;   // device row of the node's observation
            if (status[row] == kStatusOK) {
                ids[o] = row;
                lc[o] = calclcoeff(ti, sqrt(distsq));
                o++;
            }
        });
    }
}

// ---- exclusive scan of int counts into long long offsets (n + 1 entries)
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ inline long long block_exclusive_scan(long long v, long long* sh, long long* total)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        long long add = (tid >= off) ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    long long incl = sh[tid];
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kScanThreads) scan_partials_kernel(const int* __restrict__ in, long long n,
                                                                    long long* __restrict__ partials)
{
    __shared__ long long sh[kScanThreads];
    const long long base = (long long)blockIdx.x * kScanChunk;
    long long sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        long long i = base + (long long)threadIdx.x * kScanItems + k;
        if (i < n)
            sum += in[i];
    }
    long long total;
    block_exclusive_scan(sum, sh, &total);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) scan_top_kernel(long long* __restrict__ partials, int nb)
{
    __shared__ long long sh[kScanThreads];
    long long carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
        int i = b0 + threadIdx.x;
        long long v = (i < nb) ? partials[i] : 0;
        long long total;
        long long ex = block_exclusive_scan(v, sh, &total);
        if (i < nb)
            partials[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0)
        partials[nb] = carry;
}

__global__ void __launch_bounds__(kScanThreads) scan_final_kernel(const int* __restrict__ in, long long n,
                                                                 const long long* __restrict__ partials,
                                                                 long long* __restrict__ out, int nb)
{
    __shared__ long long sh[kScanThreads];
    const long long base = (long long)blockIdx.x * kScanChunk;
    long long vals[kScanItems];
    long long sum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        long long i = base + (long long)threadIdx.x * kScanItems + k;
        vals[k] = (i < n) ? in[i] : 0;
        sum += vals[k];
    }
    long long total;
    long long ex = block_exclusive_scan(sum, sh, &total) + partials[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        long long i = base + (long long)threadIdx.x * kScanItems + k;
        if (i < n)
            out[i] = ex;
        ex += vals[k];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0)
        out[n] = partials[nb];
}

// ---------------------------------------------------------------- one walk
//
// Counting the lists before filling them walks every tree twice.  With the
// capacity known from an earlier localisation, one walk suffices: a (node,
// type) thread appends its entries to chunks of kLocChunk taken from a pool
// by an atomic counter, linked in order (next[chunk]), and records its count
// and first chunk; after the scan of the counts, loc_compact_kernel copies
// each list into its place in the CSR arrays.  A walk that runs out of pool
// raises the overflow flag; the host then redoes the localisation with the
// two walks.  Types with NLOBSMAX take the two-walk path.
//
// Each thread keeps its last kLocBurst entries in registers and writes them
// as one aligned run (32 bytes of ids, 64 of tapers): single 4- and 8-byte
// stores from 64 threads walking different lists leave their cache lines
// part-written, and such lines are written back to memory several times over.
constexpr int kLocBurst = 8;
static_assert(kLocChunk % kLocBurst == 0, "bursts must not straddle chunks");

__global__ void __launch_bounds__(64) loc_walk_kernel(int nnode, int ni, const double* __restrict__ qxyz,
                                                       const unsigned char* __restrict__ isglobal,
                                                       const unsigned char* __restrict__ visible,
                                                       const KdNode* __restrict__ nodes,
                                                       const TypeInfo* __restrict__ types, int ntypes,
                                                       const short* __restrict__ status, int* __restrict__ counts,
                                                       int* __restrict__ heads, int* __restrict__ cids,
                                                       double* __restrict__ clc, int* __restrict__ cnext,
                                                       int* __restrict__ pool, int pool_cap, int* __restrict__ overflow)
{
    int t, node;
    if (!thread_node(nnode, ni, ntypes, t, node))
        return;
    const size_t item = (size_t)node * ntypes + t;
    double q[3] = {qxyz[3 * (size_t)node], qxyz[3 * (size_t)node + 1], qxyz[3 * (size_t)node + 2]};
    const bool glob = isglobal[node] != 0;
    const TypeInfo& ti = types[t];
    int n = 0;
    int head = -1;
    if (ti.searchable && (visible == nullptr || visible[t])) {
        if (glob) {
            n = ti.nids;   // obs_getglobal: all obs of the type, no status check (written by the compaction)
            head = -2;
        } else if (ti.nnode > 0) {
            int cur = -1, pos = kLocChunk;   // pos: entries of chunk cur, written or in the burst
            int nb = 0;                      // entries in the burst
            int bid[kLocBurst];
            double blc[kLocBurst];
#pragma unroll
            for (int b = 0; b < kLocBurst; ++b) {
                bid[b] = 0;
                blc[b] = 0.0;
            }
            bool full = false;
            range_walk(nodes + ti.node0, ti.ndim, q, ti.range, [&](int v, double distsq) {
                const int row = ti.node0 + v;   // device row of the node's observation
                if (status[row] != kStatusOK)
                    return;
                n++;
                if (pos == kLocChunk) {   // the burst is empty here (chunks hold whole bursts)
                    if (full)
                        return;
                    const int nc = atomicAdd(pool, 1);
                    if (nc >= pool_cap) {
                        full = true;
                        return;
                    }
                    if (cur >= 0)
                        cnext[cur] = nc;
                    else
                        head = nc;
                    cur = nc;
                    pos = 0;
                }
                const double c = calclcoeff(ti, sqrt(distsq));
#pragma unroll
                for (int b = 0; b < kLocBurst; ++b) {
                    bid[b] = (b == nb) ? row : bid[b];
                    blc[b] = (b == nb) ? c : blc[b];
                }
                ++nb;
                ++pos;
                if (nb == kLocBurst) {
                    const size_t o = (size_t)cur * kLocChunk + (pos - kLocBurst);   // a multiple of kLocBurst
                    int4* di = reinterpret_cast<int4*>(cids + o);
                    double2* dl = reinterpret_cast<double2*>(clc + o);
#pragma unroll
                    for (int b = 0; b < kLocBurst; b += 4)
                        di[b / 4] = make_int4(bid[b], bid[b + 1], bid[b + 2], bid[b + 3]);
#pragma unroll
                    for (int b = 0; b < kLocBurst; b += 2)
                        dl[b / 2] = make_double2(blc[b], blc[b + 1]);
                    nb = 0;
                }
            });
            if (nb > 0) {
                const size_t o = (size_t)cur * kLocChunk + (pos - nb);
#pragma unroll
                for (int b = 0; b < kLocBurst; ++b)
                    if (b < nb) {
                        cids[o + b] = bid[b];
                        clc[o + b] = blc[b];
                    }
            }
            if (cur >= 0)
                cnext[cur] = -1;
            if (full)
                atomicExch(overflow, 1);
        }
    }
    counts[item] = n;
    heads[item] = head;
}

// one wavefront per (node, type): the chunks of a list copied to their
// place in the CSR arrays (lanes along the chunk: coalesced)
__global__ void __launch_bounds__(256) loc_compact_kernel(long long nitem, int ntypes,
                                                           const TypeInfo* __restrict__ types,
                                                           const int* __restrict__ obsids,
                                                           const int* __restrict__ counts,
                                                           const int* __restrict__ heads,
                                                           const long long* __restrict__ offsets,
                                                           const int* __restrict__ cids, const double* __restrict__ clc,
                                                           const int* __restrict__ cnext, int* __restrict__ ids,
                                                           double* __restrict__ lc)
{
    const long long item = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= nitem)
        return;
    const int lane = threadIdx.x & 63;
    const int n = counts[item];
    const long long o = offsets[item];
    int cur = heads[item];
    if (cur == -2) {
        const TypeInfo& ti = types[item % ntypes];
        for (int i = lane; i < n; i += 64) {
            ids[o + i] = obsids[ti.obsid0 + i];
            lc[o + i] = 1.0;
        }
        return;
    }
    for (int k0 = 0; k0 < n && cur >= 0; k0 += kLocChunk) {
        const int len = min(kLocChunk, n - k0);
        for (int i = lane; i < len; i += 64) {
            ids[o + k0 + i] = cids[(size_t)cur * kLocChunk + i];
            lc[o + k0 + i] = clc[(size_t)cur * kLocChunk + i];
        }
        cur = cnext[cur];
    }
}

// grid of the walk kernels: one 64-thread block per 8 x 8 tile and obstype
// (thread_node), rounded up to a multiple of 8 runs
inline dim3 walk_grid(int ni, int nj, int ntypes)
{
    const long long n = (long long)ntypes * (((ni + 7) / 8 + 7) / 8) * (((nj + 7) / 8 + 7) / 8) * 64;
    return dim3((unsigned)(8 * kWalkRun * ((n + 8 * kWalkRun - 1) / (8 * kWalkRun))));
}

}  // namespace

hipError_t launch_loc_count(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                            const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                            const short* status, int* counts, int* raw)
{
    const long long nt = (long long)ni * nj * ntypes;
    if (nt <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(loc_count_kernel, walk_grid(ni, nj, ntypes), dim3(64), 0, s, ni * nj, ni, qxyz, isglobal,
                       visible, nodes, types, ntypes, status, counts, raw);
    return hipGetLastError();
}

hipError_t launch_loc_nlobsmax(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                               const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                               const short* status, const int* raw, const long long* rawoff, int* rrow, double* rd2,
                               int* srow, double* sd2, int* counts)
{
    const long long nt = (long long)ni * nj * ntypes;
    if (nt <= 0)
        return hipSuccess;
    if (nt > 0x7fffffffLL)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(loc_rawfill_kernel, walk_grid(ni, nj, ntypes), dim3(64), 0, s, ni * nj, ni, qxyz, isglobal,
                       visible, nodes, types, ntypes, rawoff, rrow, rd2);
    hipLaunchKernelGGL(loc_sort_kernel, dim3((unsigned)nt), dim3(256), 0, s, ni * nj, types, ntypes, status, raw,
                       rawoff, rrow, rd2, srow, sd2, counts);
    hipLaunchKernelGGL(loc_sort_big_kernel, dim3((unsigned)nt), dim3(256), 0, s, ni * nj, types, ntypes, status, raw,
                       rawoff, rrow, rd2, srow, sd2, counts);
    return hipGetLastError();
}

hipError_t launch_loc_fill(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                           const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                           const short* status, const int* obsids, const long long* offsets, int* ids, double* lc,
                           const long long* rawoff, const int* raw, const int* srow, const double* sd2)
{
    const long long nt = (long long)ni * nj * ntypes;
    if (nt <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(loc_fill_kernel, walk_grid(ni, nj, ntypes), dim3(64), 0, s, ni * nj, ni, qxyz, isglobal,
                       visible, nodes, types, ntypes, status, obsids, offsets, ids, lc, rawoff, raw, srow, sd2);
    return hipGetLastError();
}

// tmp must hold (nb + 1) long longs, nb = ceil(n / 2048); with tmp == nullptr
// the required size is returned in *tmp_bytes.
hipError_t launch_scan_counts(hipStream_t s, const int* counts, long long* offsets, long long n, void* tmp,
                              size_t* tmp_bytes)
{
    const long long nb = (n + kScanChunk - 1) / kScanChunk;
    const size_t need = (size_t)(nb + 1) * sizeof(long long);
    if (tmp == nullptr) {
        *tmp_bytes = need;
        return hipSuccess;
    }
    if (*tmp_bytes < need)
        return hipErrorInvalidValue;
    long long* partials = static_cast<long long*>(tmp);
    if (n == 0) {
        // offsets[0] = 0 via the top kernel with nb = 0
        hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanThreads), 0, s, offsets, 0);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(scan_partials_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, counts, n, partials);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanThreads), 0, s, partials, (int)nb);
    hipLaunchKernelGGL(scan_final_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, counts, n, partials, offsets,
                       (int)nb);
    return hipGetLastError();
}

hipError_t launch_loc_walk(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                           const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                           const short* status, int* counts, int* heads, int* cids, double* clc, int* cnext, int* pool,
                           int pool_cap, int* overflow)
{
    const long long nt = (long long)ni * nj * ntypes;
    if (nt <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(loc_walk_kernel, walk_grid(ni, nj, ntypes), dim3(64), 0, s, ni * nj, ni, qxyz, isglobal,
                       visible, nodes, types, ntypes, status, counts, heads, cids, clc, cnext, pool, pool_cap,
                       overflow);
    return hipGetLastError();
}

hipError_t launch_loc_compact(hipStream_t s, long long nitem, int ntypes, const TypeInfo* types, const int* obsids,
                              const int* counts, const int* heads, const long long* offsets, const int* cids,
                              const double* clc, const int* cnext, int* ids, double* lc)
{
    if (nitem <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(loc_compact_kernel, dim3((unsigned)((nitem + 3) / 4)), dim3(256), 0, s, nitem, ntypes, types,
                       obsids, counts, heads, offsets, cids, clc, cnext, ids, lc);
    return hipGetLastError();
}

}  // namespace enkfg
