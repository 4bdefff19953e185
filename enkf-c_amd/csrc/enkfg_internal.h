// enkfg_internal.h -- device data structures and kernel launchers shared by
// the host side (enkfg_abi.cpp) and the HIP kernels (*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "enkfg.h"

namespace enkfg {

// One kd-tree node on the device, 32 bytes (4 per 128-byte line).  The tree
// is the one obs_createkdtrees builds (sequential insertion,
// common/kdtree.c:194-274), stored in preorder: the left child of node v is
// v + 1, the split dimension of a node is its depth modulo ndim (as
// kd_insertnode assigns it), and a node's global index in the node array is
// its observation's device row.  Parent links allow a stackless depth-first
// walk.
struct alignas(16) KdNode {
    double x, y, z;   // coordinates (z unused for 2D trees)
    int right;        // right child within the type's tree, -1 = none
    unsigned link;    // bit 31: has a left child; bits 0-30: parent + 1 (0 at the root)
};
static_assert(sizeof(KdNode) == 32, "KdNode layout");

// Per-obstype search parameters (obstype fields of common/obstypes.h)
struct TypeInfo {
    int node0;        // first node of this type's tree in the node array
    int nnode;        // nodes in the tree (== OK obs at build time)
    int ndim;         // 3 geographic, 2 otherwise
    int searchable;   // tree present and not stats-only
    int nlocrad;
    int nlobsmax;
    int obsid0;       // first entry of this type in the obsids array
    int nids;         // OK obs of this type at tree build (obs_getglobal list)
    double range;     // obstype_getmaxlocrad()
    double locrad[ENKFG_MAXLOCRAD];
    double locweight[ENKFG_MAXLOCRAD];
};

// ---- launchers (defined in the .hip files); all return hipError_t

// localisation (loc_kernels.hip)
hipError_t launch_loc_count(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                            const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                            const short* status, int* counts, int* raw);
// NLOBSMAX types: candidates (raw counts scanned into rawoff) gathered, sorted
// by distance and selected; writes the list lengths into counts (rrow and
// rd2 are overwritten)
hipError_t launch_loc_nlobsmax(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                               const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                               const short* status, const int* raw, const long long* rawoff, int* rrow, double* rd2,
                               int* srow, double* sd2, int* counts);
// one-walk localisation (loc_kernels.hip): lists appended to a pool of
// chunks of kLocChunk entries, then compacted into the CSR arrays
constexpr int kLocChunk = 512;
hipError_t launch_loc_walk(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                           const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                           const short* status, int* counts, int* heads, int* cids, double* clc, int* cnext, int* pool,
                           int pool_cap, int* overflow);
hipError_t launch_loc_compact(hipStream_t s, long long nitem, int ntypes, const TypeInfo* types, const int* obsids,
                              const int* counts, const int* heads, const long long* offsets, const int* cids,
                              const double* clc, const int* cnext, int* ids, double* lc);
hipError_t launch_loc_fill(hipStream_t s, int ni, int nj, const double* qxyz, const unsigned char* isglobal,
                           const unsigned char* visible, const KdNode* nodes, const TypeInfo* types, int ntypes,
                           const short* status, const int* obsids, const long long* offsets, int* ids,
                           double* lc, const long long* rawoff, const int* raw, const int* srow, const double* sd2);
hipError_t launch_scan_counts(hipStream_t s, const int* counts, long long* offsets, long long n, void* tmp,
                              size_t* tmp_bytes);

// standardisation and layout (calc_kernels.hip)
hipError_t launch_standardise(hipStream_t s, int nmem, int nobs, const short* type, const float* estd,
                              const double* sqrt_rfactor, float* S_obsmajor, double* sv);
hipError_t launch_transpose_f32(hipStream_t s, const float* in, float* out, long long rows, long long cols);
// Row permutations between the caller's order (rows of `width` elements)
// and device row order (rows of `stride` >= width elements, zero padded):
// gather writes dst[r] = src[perm[r]] for r < nrow plus one zero row at
// r = nrow; scatter writes dst[perm[r]] = src[r].
hipError_t launch_gather_rows(hipStream_t s, const float* src, const int* perm, float* dst, long long nrow, int width,
                              int stride);
hipError_t launch_gather_rows(hipStream_t s, const double* src, const int* perm, double* dst, long long nrow,
                              int width, int stride);
hipError_t launch_scatter_rows(hipStream_t s, const float* src, const int* perm, float* dst, long long nrow, int width,
                               int stride);
hipError_t launch_scatter_rows(hipStream_t s, const double* src, const int* perm, double* dst, long long nrow,
                               int width, int stride);

// per-column transforms (calc_kernels.hip)
constexpr int kListPad = 256;      // readable entries past the end of the local lists
struct CalcArgs {
    int m, md, mode, scheme, ntypes, pertype;
    double alpha;
    int node0, nbatch;             // batch of nodes [node0, node0 + nbatch)
    int nseg;                      // segments per node in the Gram scratch (ntypes or 1)
    const int* counts;             // [nnode][ntypes]
    const long long* offsets;      // [nnode*ntypes + 1]
    const int* ids;
    const double* lc;
    const float* S;                // [nobs + 1][ms], device row order; row nobs is zero
    const double* sv;              // [nobs + 1], device row order; sv[nobs] = 0
    int ms;                        // row stride of S: m rounded up to a multiple of 4
    int nobs;                      // index of the zero row
    // ids and lc are readable for kListPad entries past the last list
    double* gram;                  // [nbatch][nseg][mp][mp]
    double* sts;                   // [nbatch][nseg][mp]
    float* T;                      // [nnode][md][m]
    float* w;                      // [nnode][m]
    double* T64;                   // optional
    double* w64;                   // optional
    int* pnlobs;                   // [ntypes+1][nnode]
    float* pdfs;
    float* psrf;
    int nnode;
    int* err;                      // device error flag
    // adjoint of the analysis mean: nodes taken from node_list[node0 ..]
    // (when set), and the factor kernel returns z = M^-1 g (adj_g: [node][m])
    // and w = M^-1 S^T s per listed node (adj_z, adj_w: [list index][m])
    const int* node_list;
    // item Gram kernels: the segments in the order their items run ([nseg],
    // obstypes by decreasing total list length), or null for 0, 1, ...
    const int* seg_order;
    const double* adj_g;
    double* adj_z;
    double* adj_w;
    // ... and, when set, M^-1 of each batch node ([bnode][mp][mp], mp = 16
    // ceil(m / 16), zero outside m x m) for the adjoint through T
    double* adj_minv;
    // Gram item kernels (m <= 96): tile groups (wavefronts) per segment, 2 or
    // 3; any other value takes the kernel's default (3 FP64, 2 mixed)
    int tile_groups;
    // ETKF with m > 96: the Denman-Beavers iterates Y, Z and Y^-1 of each batch
    // node in device memory, [bnode][3][B * B][256] doubles (thread-major, so
    // each thread's block elements are coalesced)
    double* etkf_ws;
};
// doubles of etkf_ws per batch node
__host__ __device__ inline size_t etkf_ws_per_node(int m)
{
    const size_t B = (size_t)(m + 15) / 16;
    return 3 * B * B * 256;
}
hipError_t launch_gram(hipStream_t s, const CalcArgs& a);
hipError_t launch_gram_mixed(hipStream_t s, const CalcArgs& a);   // gram_mixed.hip
hipError_t launch_factor(hipStream_t s, const CalcArgs& a);
// factor_wave.hip: DEnKF / EnOI with m <= 160 by the symmetric sweep; the
// per-type diagnostics of every scheme (factor_trace_kernel, after the
// node's own factorisation on the same stream)
bool factor_wave_supported(const CalcArgs& a);
hipError_t launch_factor_wave(hipStream_t s, const CalcArgs& a);
hipError_t launch_factor_trace(hipStream_t s, const CalcArgs& a);
// ETKF with m <= 96 by the symmetric sweep, then the per-type diagnostics
hipError_t launch_factor_etkf(hipStream_t s, const CalcArgs& a);

// update (update_kernels.hip)
struct UpdArgs {
    int m, md, mode;
    int mni, mnj, stride, periodic_i, surfk;
    int ni_s, nj_s;
    int j1, jsize;                 // rows j1 .. j1+jsize-1 of the grid
    int nfield;
    int outputinc;
    const int* nlevels;            // [mnj][mni]
    const float* T;                // [nj_s][ni_s][md][m]
    const float* w;                // [nj_s][ni_s][m]
    const int* flevel;             // [nfield]
    const int* fsurf;
    const int* fapplylog;
    const float* finfl;
    const double* finfratio;
    float* E;                      // [nfield][ncell][m], cell = (j - j1) * mni + i
    long long ncell;               // cells per field in E
    float* infl;                   // optional [nfield][ncell]
};
hipError_t launch_update_mfma(hipStream_t s, const UpdArgs& a, bool mixed);   // update_mfma.hip

// analysed ensemble observations (obs_update.hip): the observations grouped
// by grid cell, HE observation-major [nobs][m]
struct HeArgs {
    int ngroup;                    // cells holding observations
    const int* gcell;              // [ngroup] cell j * mni + i
    const int* goff;               // [ngroup + 1] offsets into order
    const int* order;              // observation indices, grouped by cell
    const float* infl0;            // [nobs] inflation of the obs type's variable
    const double* ratio;           // [nobs] its inflation ratio (NaN = plain)
    float* HE;                     // [nobs][m], updated in place
};
hipError_t launch_update_he(hipStream_t s, const UpdArgs& u, const HeArgs& h);

// the ensemble observations from the device-resident ensemble (he_kernels.hip)
struct HeGetArgs {
    int m, md, mode;
    double gamma;
    int ni, nj, ktop, periodic_i;  // model grid, surface layer id
    long long ncell;
    const int* nlevels;            // [nj][ni]
    const float* E;                // [field][cell][m]
    int nobs;
    const short* type;             // [nobs], observation order
    const short* status;
    const double* fi;
    const double* fj;
    const float* fk;
    const double* value;
    const int* hkind;              // [ntypes] ENKFG_H_SURF / ENKFG_H_SUBSURF
    const int* nk;                 // [ntypes] layers of the type's variable
    const int* field0;             // [ntypes] first entry of the type in fields
    const int* fields;             // field index of each layer
    float* HA;                     // [nobs][m] out: HE minus the member mean
    double* s;                     // [nobs] out: innovations
    double* std;                   // [nobs] out: spread
    int* err;                      // 1: bad forecast, 2: bad innovation
};
hipError_t launch_he(hipStream_t s, const HeGetArgs& a);
// rows of HA and entries of s of the obs that are not OK set to 0
hipError_t launch_he_clear_bad(hipStream_t s, const short* status, int nobs, int m, float* HA, double* sv);
hipError_t launch_moderate(hipStream_t s, int nobs, double kfactor, const short* status, const short* type,
                           const double* std, const double* inn, float* estd, int* nmod);

// adjoint of the analysis mean (adjoint.hip)
// g_nodes[node][m] += sum over cells of beta(cell, node) sum_f dJ/dmean[f][cell] A^f[f][cell][:]
hipError_t launch_adj_cells(hipStream_t s, const UpdArgs& a, const float* dJ_dmean, double* g_nodes);
// adjoint of the analysis ensemble (adjoint.hip): per node, g_n (as above)
// and D_n = dJ/dT_n ([node][md][m]); per field, dJ/dinflation0 (summed)
hipError_t launch_adj_ens_cells(hipStream_t s, const UpdArgs& a, const float* dJ_dXa, double* g_nodes,
                                double* D_nodes, double* g_infl);
// flags[r] |= 1 if row r of the nrow x width matrix has a non-zero entry
hipError_t launch_nonzero_rows(hipStream_t s, const double* g, int nrow, int width, int* flags, bool accumulate);
// B_k = alpha / 2 M_k^-1 D_k M_k^-1 for the batch's nodes (k = node0 ..
// node0 + nbatch - 1 of the list; M^-1 from adj_minv, D from D_nodes)
hipError_t launch_adj_B(hipStream_t s, const CalcArgs& a, const double* D_nodes, double* B);
// per listed node k of the batch (node_list[node0 + k]): dJ/dy and dJ/dR of
// its local obs, by device row; with B, the term of T as well
hipError_t launch_adj_obs(hipStream_t s, const CalcArgs& a, const int* perm, const short* type, const float* estd,
                          const double* sqrt_rfactor, const double* B, double* dJ_dy, double* dJ_dR);
hipError_t launch_update_bg(hipStream_t s, const UpdArgs& a);
// addnoise (update.c:46-87) on one field E [ncell][m]:
// v <- (float)((double)(v * defl) + (double)mult * r[e])
hipError_t launch_addnoise(hipStream_t s, float* E, long long ncell, int m, float defl, float mult, const double* r);

// das_writespread's spread (diag_kernels.hip) of nfield fields: E is
// [field][cell][m] when cells_major, else [field][m][cell]; out [field][cell]
hipError_t launch_spread(hipStream_t s, const float* E, int nfield, long long ncell, int m, bool cells_major,
                         float* out);
hipError_t launch_hybrid_condition(hipStream_t s, const UpdArgs& a, double gamma);
// [nmem][rows][cols] <-> [rows*cols][nmem]
hipError_t launch_members_to_cells(hipStream_t s, const float* in, float* out, int nmem, long long ncell);
hipError_t launch_cells_to_members(hipStream_t s, const float* in, float* out, int nmem, long long ncell);

}  // namespace enkfg
