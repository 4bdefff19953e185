// enkfg_abi.cpp -- the C ABI of include/enkfg.h: context, inputs, the
// host-side kd-tree construction, device memory and kernel sequencing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "enkfg.h"
#include "enkfg_internal.h"
#include "gram_layout.h"

// Scratch of enkfg_update_fields_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never uses its name as an
// address (on the host it is a stand-in); the device address of each GPU's
// copy comes from hipGetSymbolAddress, taken once per device at context
// creation with that device current (scratch_address below).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

using namespace enkfg;

namespace {

// Device buffer that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= bytes && p != nullptr)
            return hipSuccess;
        if (p != nullptr) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess)
                return e;
        }
        if (n == 0)
            n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess)
            bytes = n;
        return e;
    }
    void release()
    {
        if (p != nullptr)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

constexpr double kREarth = 6371.0;            // common/global.h:119
constexpr double kDeg2Rad = M_PI / 180.0;     // common/global.h:120

// common/utils.c:1449-1458, evaluated with the host libm as the reference does.
// The reference is built with gcc -O2, which merges the sin and cos of the
// same argument into one glibc sincos() call; sincos() and separate sin()/cos()
// differ in the last bit for a fraction of inputs, so sincos() is called
// explicitly to obtain the reference's coordinates bit for bit.
void ll2xyz(double lon_deg, double lat_deg, double* out)
{
    const double lon = lon_deg * kDeg2Rad;
    const double lat = lat_deg * kDeg2Rad;
    double sinlat, coslat, sinlon, coslon;
    sincos(lat, &sinlat, &coslat);
    sincos(lon, &sinlon, &coslon);
    out[0] = kREarth * coslon * coslat;
    out[1] = kREarth * sinlon * coslat;
    out[2] = kREarth * sinlat;
}

}  // namespace

// obs-count buckets of the calc batching: lower bounds of the largest
// segment (node, obstype list) of a node
constexpr int kBuckets = 4;
constexpr int kBucketLo[kBuckets] = {0, 256, 1024, 4096};

struct enkfg_ctx {
    enkfg_config cfg{};
    std::string err;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;   // factor kernels, overlapping the next batch's Gram
    hipEvent_t ev[10] = {};
    std::vector<hipEvent_t> bev;     // 4 per calc batch
    double timings[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    // obstypes and obs
    std::vector<enkfg_obstype> types;
    int nobs = 0;
    std::vector<short> otype, ostatus;
    std::vector<float> olon, olat, oestd;
    bool obs_set = false;
    DevBuf d_nodes, d_types, d_obsids, d_type, d_status, d_estd, d_sqrtrf, d_perm;
    std::vector<TypeInfo> tinfo;
    // Device rows of S, s and status are ordered by tree preorder (type by
    // type), so the obs a node's range search returns -- a few subtrees --
    // are a few runs of nearby rows; perm[row] is the observation index.
    // kd nodes and the local lists hold rows; the API speaks obs indices.
    std::vector<int> perm;

    // S and s
    DevBuf d_S, d_s, d_tmp, d_tmp2;
    bool S_set = false;

    // localisation
    int nnode = 0;
    std::vector<double> node_lon, node_lat;   // node grid whose coordinates are on the device
    int node_geo = 0;
    DevBuf d_qxyz, d_glob, d_vis, d_counts, d_offsets, d_scantmp, d_ids, d_lc;
    DevBuf d_raw, d_rawoff, d_rrow, d_rd2, d_srow, d_sd2;   // NLOBSMAX candidates
    // one-walk localisation: chunk pool (capacity from the last two-walk
    // localisation), list heads, pool counter and overflow flag
    DevBuf d_heads, d_cids, d_clc, d_cnext, d_lflag;
    long long pool_cap = 0;
    bool lists_valid = false;

    // calc batching: nodes ordered by obs-count bucket (heaviest first), and
    // the autotuned Gram tile groups per [precision][bucket] (0 = not tuned)
    DevBuf d_order;
    DevBuf d_segord;   // segment order of the item Gram launches
    std::vector<int> h_order;
    int bucket_nodes[kBuckets] = {0, 0, 0, 0};
    int tile_groups[2][kBuckets] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    // captured batch loop (ENKFG_FLAG_GRAPH): the executable graph, the
    // signature of what it was captured for, and its dependency events
    hipGraphExec_t graph_exec = nullptr;
    std::vector<long long> graph_sig;
    std::vector<hipEvent_t> gev;
    int graph_launches = 0;   // replays of the current graph (statistics)

    // transforms and diagnostics
    int ni_s = 0, nj_s = 0, stride = 0;
    DevBuf d_T, d_w, d_T64, d_w64, d_pnlobs, d_pdfs, d_psrf, d_gram, d_sts, d_err, d_etkf;
    bool transforms_valid = false, fp64_valid = false, diag_valid = false;
    enkfg_stats stats{};

    // update
    DevBuf d_E, d_infl, d_nlev, d_flevel, d_fsurf, d_flog, d_finfl, d_fratio, d_spread;
    // RANDOMISE: the context's drand48 stream (48-bit state) and the noise
    // vectors of the randomised fields of the last batch
    unsigned long long rand48 = 0;
    DevBuf d_noise;
    // analysed observations: cell groups and per-obs inflation parameters
    DevBuf d_he_order, d_he_gcell, d_he_goff, d_he_infl, d_he_ratio;
    // this device's copy of the no-allocation scratch array (device address)
    unsigned char* scratch = nullptr;
    // ensemble observations: H operator tables, obs records, spread
    DevBuf d_hg_tab, d_hg_obs, d_hg_std;
    double kfactor = NAN;            // das_moderateobs; NaN = off
    DevBuf d_kmod;
    std::vector<int> kmod;           // per type, from the last moderation
    // adjoint
    DevBuf d_adjg, d_adjflag, d_adjlist, d_adjz, d_adjw, d_gy, d_gR, d_adjD, d_adjinfl, d_minv, d_adjB;
    float* attached_E = nullptr;
    enkfg_grid att_grid{};
    int att_nfield = 0;
    std::vector<enkfg_field> att_fields;
    std::vector<int> att_nlevels;
};

namespace {

int fail(enkfg_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c != nullptr)
        c->err = buf;
    return code;
}

#define HIPCHECK(c, expr)                                                                                  \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return fail((c), (e_ == hipErrorOutOfMemory) ? ENKFG_ERR_NOMEM : ENKFG_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                                 \
    } while (0)

// a kd-tree node while the tree is built on the host
struct BuildNode {
    double x, y, z;
    int left, right, parent, dir, obsid;
};

// obs_createkdtrees for one obstype: sequential insertion of the type's OK obs
// in ascending index order (common/kdtree.c:194-274, observations.c:1695-1720).
void build_tree(const std::vector<double>& xyz, int ndim, const std::vector<int>& ids, std::vector<BuildNode>& out)
{
    const size_t base = out.size();
    for (size_t k = 0; k < ids.size(); ++k) {
        const double* c = &xyz[3 * k];
        if (!std::isfinite(c[0]))
            continue;   // kd_insertnode skips non-finite points
        BuildNode n{};
        n.x = c[0];
        n.y = c[1];
        n.z = (ndim == 3) ? c[2] : 0.0;
        n.left = n.right = -1;
        n.obsid = ids[k];
        const int id = (int)(out.size() - base);
        if (id == 0) {
            n.parent = -1;
            n.dir = 0;
            out.push_back(n);
            continue;
        }
        int cur = 0;
        for (;;) {
            BuildNode& p = out[base + cur];
            const double pc = (p.dir == 0) ? p.x : (p.dir == 1 ? p.y : p.z);
            int& slot = (c[p.dir] < pc) ? p.left : p.right;
            if (slot < 0) {
                slot = id;
                n.parent = cur;
                n.dir = (p.dir + 1) % ndim;
                break;
            }
            cur = slot;
        }
        out.push_back(n);
    }
    // Relabel the nodes in preorder (root first, then the left subtree, then
    // the right one) so that a subtree occupies a contiguous stretch of
    // memory; the tree itself, and so every walk over it, is unchanged.
    const int n = (int)(out.size() - base);
    if (n > 2) {
        std::vector<int> order;
        order.reserve(n);
        std::vector<int> stack(1, 0);
        while (!stack.empty()) {
            const int v = stack.back();
            stack.pop_back();
            order.push_back(v);
            const BuildNode& nd = out[base + v];
            if (nd.right >= 0)
                stack.push_back(nd.right);
            if (nd.left >= 0)
                stack.push_back(nd.left);
        }
        std::vector<int> pos(n);
        for (int k = 0; k < n; ++k)
            pos[order[k]] = k;
        std::vector<BuildNode> tmp(n);
        for (int k = 0; k < n; ++k) {
            BuildNode nd = out[base + order[k]];
            nd.left = (nd.left >= 0) ? pos[nd.left] : -1;
            nd.right = (nd.right >= 0) ? pos[nd.right] : -1;
            nd.parent = (nd.parent >= 0) ? pos[nd.parent] : -1;
            tmp[k] = nd;
        }
        std::copy(tmp.begin(), tmp.end(), out.begin() + base);
    }
}

int check_ctx(enkfg_ctx* c)
{
    return c == nullptr ? ENKFG_ERR_ARG : ENKFG_OK;
}

// search-time status (observations.c:1906) by device row
int upload_status_rows(enkfg_ctx* c)
{
    std::vector<short> st((size_t)std::max(c->nobs, 1));
    for (int r = 0; r < c->nobs; ++r)
        st[r] = c->ostatus[c->perm[r]];
    HIPCHECK(c, c->d_status.reserve(st.size() * sizeof(short)));
    HIPCHECK(c, hipMemcpy(c->d_status.p, st.data(), st.size() * sizeof(short), hipMemcpyHostToDevice));
    return ENKFG_OK;
}

int upload_types(enkfg_ctx* c)
{
    HIPCHECK(c, c->d_types.reserve(sizeof(TypeInfo) * std::max<size_t>(c->tinfo.size(), 1)));
    if (!c->tinfo.empty())
        HIPCHECK(c, hipMemcpy(c->d_types.p, c->tinfo.data(), sizeof(TypeInfo) * c->tinfo.size(), hipMemcpyHostToDevice));
    return ENKFG_OK;
}

int node_inputs(enkfg_ctx* c, const enkfg_nodes* nodes)
{
    if (nodes == nullptr || nodes->ni <= 0 || nodes->nj <= 0 || nodes->lon == nullptr || nodes->lat == nullptr)
        return fail(c, ENKFG_ERR_ARG, "enkfg_nodes: invalid node grid");
    const long long nn = (long long)nodes->ni * nodes->nj;
    if (nn > 0x7fffffffLL)
        return fail(c, ENKFG_ERR_ARG, "too many nodes");
    c->nnode = (int)nn;
    c->ni_s = nodes->ni;
    c->nj_s = nodes->nj;
    c->stride = nodes->stride;
    const int nt = (int)c->types.size();
    if (nodes->visible != nullptr) {
        HIPCHECK(c, c->d_vis.reserve(std::max(nt, 1)));
        HIPCHECK(c, hipMemcpy(c->d_vis.p, nodes->visible, nt, hipMemcpyHostToDevice));
    }
    // the same node grid as the last call (bit for bit): its coordinates are
    // on the device already
    if (c->node_lon.size() == (size_t)nn && c->node_geo == nodes->geographic &&
        std::memcmp(c->node_lon.data(), nodes->lon, (size_t)nn * sizeof(double)) == 0 &&
        std::memcmp(c->node_lat.data(), nodes->lat, (size_t)nn * sizeof(double)) == 0)
        return ENKFG_OK;
    c->node_lon.clear();   // valid again only after the upload below
    std::vector<double> q(3 * (size_t)nn);
    std::vector<unsigned char> g(nn);
    for (long long k = 0; k < nn; ++k) {
        const double lon = nodes->lon[k], lat = nodes->lat[k];
        g[k] = std::isnan(lon + lat) ? 1 : 0;   // obs_getglobal path
        if (g[k]) {
            q[3 * k] = q[3 * k + 1] = q[3 * k + 2] = 0.0;
        } else if (nodes->geographic) {
            ll2xyz(lon, lat, &q[3 * k]);
        } else {
            q[3 * k] = lon;
            q[3 * k + 1] = lat;
            q[3 * k + 2] = std::numeric_limits<double>::quiet_NaN();
        }
    }
    HIPCHECK(c, c->d_qxyz.reserve(q.size() * sizeof(double)));
    HIPCHECK(c, c->d_glob.reserve(nn));
    HIPCHECK(c, hipMemcpy(c->d_qxyz.p, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_glob.p, g.data(), nn, hipMemcpyHostToDevice));
    c->node_lon.assign(nodes->lon, nodes->lon + nn);
    c->node_lat.assign(nodes->lat, nodes->lat + nn);
    c->node_geo = nodes->geographic;
    return ENKFG_OK;
}

double elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess)
        return -1.0;
    return ms;
}

int do_localise(enkfg_ctx* c, const enkfg_nodes* nodes)
{
    if (!c->obs_set)
        return fail(c, ENKFG_ERR_STATE, "observations not set");
    bool nlobsmax = false;
    for (const auto& t : c->types)
        if (t.nlobsmax != ENKFG_NLOBSMAX_NONE)
            nlobsmax = true;
    int r = node_inputs(c, nodes);
    if (r != ENKFG_OK)
        return r;
    const int nt = (int)c->types.size();
    const long long ncnt = (long long)c->nnode * nt;
    HIPCHECK(c, c->d_counts.reserve(std::max<long long>(ncnt, 1) * sizeof(int)));
    HIPCHECK(c, c->d_offsets.reserve((ncnt + 1) * sizeof(long long)));
    const unsigned char* vis = nodes->visible ? c->d_vis.as<unsigned char>() : nullptr;
    if (nlobsmax)
        HIPCHECK(c, c->d_raw.reserve(std::max<long long>(ncnt, 1) * sizeof(int)));
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    size_t tmpb = 0;
    HIPCHECK(c, launch_scan_counts(c->stream, nullptr, nullptr, ncnt, nullptr, &tmpb));
    HIPCHECK(c, c->d_scantmp.reserve(tmpb));
    tmpb = c->d_scantmp.bytes;
    if (!nlobsmax && c->pool_cap > 0) {
        // one walk into the chunk pool, then the compaction; a pool too
        // small for these lists falls back to the two walks below
        HIPCHECK(c, c->d_heads.reserve(std::max<long long>(ncnt, 1) * sizeof(int)));
        HIPCHECK(c, c->d_cids.reserve((size_t)c->pool_cap * kLocChunk * sizeof(int)));
        HIPCHECK(c, c->d_clc.reserve((size_t)c->pool_cap * kLocChunk * sizeof(double)));
        HIPCHECK(c, c->d_cnext.reserve((size_t)c->pool_cap * sizeof(int)));
        HIPCHECK(c, c->d_lflag.reserve(2 * sizeof(int)));
        HIPCHECK(c, hipMemsetAsync(c->d_lflag.p, 0, 2 * sizeof(int), c->stream));
        HIPCHECK(c, launch_loc_walk(c->stream, c->ni_s, c->nj_s, c->d_qxyz.as<double>(), c->d_glob.as<unsigned char>(),
                                    vis, c->d_nodes.as<KdNode>(), c->d_types.as<TypeInfo>(), nt,
                                    c->d_status.as<short>(), c->d_counts.as<int>(), c->d_heads.as<int>(),
                                    c->d_cids.as<int>(), c->d_clc.as<double>(), c->d_cnext.as<int>(),
                                    c->d_lflag.as<int>(), (int)c->pool_cap, c->d_lflag.as<int>() + 1));
        HIPCHECK(c, launch_scan_counts(c->stream, c->d_counts.as<int>(), c->d_offsets.as<long long>(), ncnt,
                                       c->d_scantmp.p, &tmpb));
        long long total = 0;
        int flags[2] = {0, 0};
        HIPCHECK(c, hipMemcpyAsync(&total, c->d_offsets.as<long long>() + ncnt, sizeof total, hipMemcpyDeviceToHost,
                                   c->stream));
        HIPCHECK(c, hipMemcpyAsync(flags, c->d_lflag.p, sizeof flags, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        if (flags[1] == 0) {
            HIPCHECK(c, c->d_ids.reserve((total + kListPad) * sizeof(int)));
            HIPCHECK(c, c->d_lc.reserve((total + kListPad) * sizeof(double)));
            HIPCHECK(c, launch_loc_compact(c->stream, ncnt, nt, c->d_types.as<TypeInfo>(), c->d_obsids.as<int>(),
                                           c->d_counts.as<int>(), c->d_heads.as<int>(),
                                           c->d_offsets.as<long long>(), c->d_cids.as<int>(), c->d_clc.as<double>(),
                                           c->d_cnext.as<int>(), c->d_ids.as<int>(), c->d_lc.as<double>()));
            HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            c->timings[0] = elapsed(c->ev[0], c->ev[1]);
            c->lists_valid = true;
            return ENKFG_OK;
        }
    }
    HIPCHECK(c, launch_loc_count(c->stream, c->ni_s, c->nj_s, c->d_qxyz.as<double>(), c->d_glob.as<unsigned char>(), vis,
                                 c->d_nodes.as<KdNode>(), c->d_types.as<TypeInfo>(), nt, c->d_status.as<short>(),
                                 c->d_counts.as<int>(), nlobsmax ? c->d_raw.as<int>() : nullptr));
    if (nlobsmax) {
        // NLOBSMAX types: candidates in visiting order, sorted by distance,
        // the first NLOBSMAX with status OK kept (observations.c:1899-1908)
        HIPCHECK(c, c->d_rawoff.reserve((ncnt + 1) * sizeof(long long)));
        HIPCHECK(c, launch_scan_counts(c->stream, c->d_raw.as<int>(), c->d_rawoff.as<long long>(), ncnt,
                                       c->d_scantmp.p, &tmpb));
        long long nraw = 0;
        HIPCHECK(c, hipMemcpyAsync(&nraw, c->d_rawoff.as<long long>() + ncnt, sizeof nraw, hipMemcpyDeviceToHost,
                                   c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        HIPCHECK(c, c->d_rrow.reserve(std::max<long long>(nraw, 1) * sizeof(int)));
        HIPCHECK(c, c->d_srow.reserve(std::max<long long>(nraw, 1) * sizeof(int)));
        HIPCHECK(c, c->d_rd2.reserve(std::max<long long>(nraw, 1) * sizeof(double)));
        HIPCHECK(c, c->d_sd2.reserve(std::max<long long>(nraw, 1) * sizeof(double)));
        HIPCHECK(c, launch_loc_nlobsmax(c->stream, c->ni_s, c->nj_s, c->d_qxyz.as<double>(),
                                        c->d_glob.as<unsigned char>(), vis, c->d_nodes.as<KdNode>(),
                                        c->d_types.as<TypeInfo>(), nt, c->d_status.as<short>(), c->d_raw.as<int>(),
                                        c->d_rawoff.as<long long>(), c->d_rrow.as<int>(), c->d_rd2.as<double>(),
                                        c->d_srow.as<int>(), c->d_sd2.as<double>(), c->d_counts.as<int>()));
    }
    HIPCHECK(c, launch_scan_counts(c->stream, c->d_counts.as<int>(), c->d_offsets.as<long long>(), ncnt,
                                   c->d_scantmp.p, &tmpb));
    long long total = 0;
    HIPCHECK(c, hipMemcpyAsync(&total, c->d_offsets.as<long long>() + ncnt, sizeof total, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    // the Gram kernels read up to kListPad entries past the last list
    HIPCHECK(c, c->d_ids.reserve((total + kListPad) * sizeof(int)));
    HIPCHECK(c, c->d_lc.reserve((total + kListPad) * sizeof(double)));
    HIPCHECK(c, launch_loc_fill(c->stream, c->ni_s, c->nj_s, c->d_qxyz.as<double>(), c->d_glob.as<unsigned char>(), vis,
                                c->d_nodes.as<KdNode>(), c->d_types.as<TypeInfo>(), nt, c->d_status.as<short>(),
                                c->d_obsids.as<int>(), c->d_offsets.as<long long>(), c->d_ids.as<int>(),
                                c->d_lc.as<double>(), nlobsmax ? c->d_rawoff.as<long long>() : nullptr,
                                nlobsmax ? c->d_raw.as<int>() : nullptr, nlobsmax ? c->d_srow.as<int>() : nullptr,
                                nlobsmax ? c->d_sd2.as<double>() : nullptr));
    HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->timings[0] = elapsed(c->ev[0], c->ev[1]);
    c->lists_valid = true;
    // pool for the next localisation's single walk: the chunks these lists
    // take, and a margin
    if (!nlobsmax) {
        std::vector<int> cnt((size_t)ncnt);
        if (ncnt > 0)
            HIPCHECK(c, hipMemcpy(cnt.data(), c->d_counts.p, (size_t)ncnt * sizeof(int), hipMemcpyDeviceToHost));
        long long need = 0;
        for (int n : cnt)
            need += (n + kLocChunk - 1) / kLocChunk;
        need += need / 64 + 64;
        c->pool_cap = (need < 0x7fffffffLL) ? need : 0;   // 0: stay with the two walks
    }
    return ENKFG_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// the current device's scratch address (device `dev` must be current)
hipError_t scratch_address(int dev, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

int enkfg_create(enkfg_ctx** out, const enkfg_config* cfg)
{
    if (out == nullptr || cfg == nullptr)
        return ENKFG_ERR_ARG;
    *out = nullptr;
    if (cfg->nmem < 2 || cfg->nmem > 160 || cfg->nmem_dynamic < 1 || cfg->nmem_dynamic > cfg->nmem)
        return ENKFG_ERR_ARG;
    if (cfg->mode != ENKFG_MODE_ENKF && cfg->mode != ENKFG_MODE_ENOI && cfg->mode != ENKFG_MODE_HYBRID)
        return ENKFG_ERR_ARG;
    if (cfg->scheme != ENKFG_SCHEME_DENKF && cfg->scheme != ENKFG_SCHEME_ETKF)
        return ENKFG_ERR_ARG;
    if (cfg->precision != ENKFG_PREC_FP64 && cfg->precision != ENKFG_PREC_MIXED)
        return ENKFG_ERR_ARG;
    if (cfg->mode != ENKFG_MODE_HYBRID && cfg->nmem_dynamic != cfg->nmem)
        return ENKFG_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return ENKFG_ERR_HIP;
    if (cfg->device < 0 || cfg->device >= ndev)
        return ENKFG_ERR_ARG;
    if (hipSetDevice(cfg->device) != hipSuccess)
        return ENKFG_ERR_HIP;
    enkfg_ctx* c = new enkfg_ctx();
    c->cfg = *cfg;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ENKFG_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return ENKFG_ERR_HIP;
    }
    for (auto& e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return ENKFG_ERR_HIP;
        }
    if (c->d_err.reserve(sizeof(int)) != hipSuccess) {
        delete c;
        return ENKFG_ERR_NOMEM;
    }
    if (scratch_address(cfg->device, &c->scratch) != hipSuccess) {
        delete c;
        return ENKFG_ERR_HIP;
    }
    *out = c;
    return ENKFG_OK;
}

void enkfg_destroy(enkfg_ctx* c)
{
    if (c == nullptr)
        return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream != nullptr)
        (void)hipStreamSynchronize(c->stream);
    if (c->stream2 != nullptr)
        (void)hipStreamSynchronize(c->stream2);
    DevBuf* bufs[] = {&c->d_nodes, &c->d_types, &c->d_obsids, &c->d_type, &c->d_status, &c->d_estd, &c->d_sqrtrf,
                      &c->d_perm, &c->d_S, &c->d_s, &c->d_tmp, &c->d_tmp2, &c->d_qxyz, &c->d_glob, &c->d_vis, &c->d_counts, &c->d_offsets,
                      &c->d_scantmp, &c->d_ids, &c->d_lc, &c->d_heads, &c->d_cids, &c->d_clc, &c->d_cnext, &c->d_lflag,
                      &c->d_raw, &c->d_rawoff, &c->d_rrow, &c->d_rd2,
                      &c->d_srow, &c->d_sd2, &c->d_T, &c->d_w, &c->d_T64, &c->d_w64, &c->d_pnlobs,
                      &c->d_pdfs, &c->d_psrf, &c->d_gram, &c->d_order, &c->d_segord, &c->d_sts, &c->d_err, &c->d_E, &c->d_infl, &c->d_nlev,
                      &c->d_flevel, &c->d_fsurf, &c->d_flog, &c->d_finfl, &c->d_fratio, &c->d_adjg, &c->d_adjflag,
                      &c->d_adjlist, &c->d_adjz, &c->d_adjw, &c->d_gy, &c->d_gR, &c->d_etkf, &c->d_noise,
                      &c->d_he_order, &c->d_he_gcell, &c->d_he_goff, &c->d_he_infl, &c->d_he_ratio, &c->d_hg_tab,
                      &c->d_hg_obs, &c->d_hg_std, &c->d_spread, &c->d_adjD, &c->d_adjinfl, &c->d_minv, &c->d_adjB};
    for (DevBuf* b : bufs)
        b->release();
    for (auto& e : c->ev)
        if (e != nullptr)
            (void)hipEventDestroy(e);
    for (auto& e : c->bev)
        (void)hipEventDestroy(e);
    if (c->stream != nullptr)
        (void)hipStreamDestroy(c->stream);
    if (c->graph_exec != nullptr)
        (void)hipGraphExecDestroy(c->graph_exec);
    for (hipEvent_t e : c->gev)
        (void)hipEventDestroy(e);
    if (c->stream2 != nullptr)
        (void)hipStreamDestroy(c->stream2);
    delete c;
}

const char* enkfg_last_error(const enkfg_ctx* c)
{
    return c == nullptr ? "null context" : c->err.c_str();
}

int enkfg_sync(enkfg_ctx* c)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ENKFG_OK;
}

int enkfg_set_obstypes(enkfg_ctx* c, int ntypes, const enkfg_obstype* types)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (ntypes < 0 || (ntypes > 0 && types == nullptr) || ntypes > 32767)
        return fail(c, ENKFG_ERR_ARG, "invalid obstypes");
    for (int t = 0; t < ntypes; ++t)
        if (types[t].nlocrad < 1 || types[t].nlocrad > ENKFG_MAXLOCRAD || !(types[t].rfactor > 0.0))
            return fail(c, ENKFG_ERR_ARG, "obstype %d: invalid nlocrad or rfactor", t);
    c->types.assign(types, types + ntypes);
    c->obs_set = false;
    c->lists_valid = false;
    return ENKFG_OK;
}

int enkfg_set_obs(enkfg_ctx* c, int nobs, const short* type, const short* status, const float* lon, const float* lat,
                  const float* estd)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (nobs < 0 || (nobs > 0 && (!type || !status || !lon || !lat || !estd)))
        return fail(c, ENKFG_ERR_ARG, "invalid obs arrays");
    const int nt = (int)c->types.size();
    for (int i = 0; i < nobs; ++i)
        if (type[i] < 0 || type[i] >= nt)
            return fail(c, ENKFG_ERR_ARG, "obs %d: type %d out of range", i, (int)type[i]);
    c->nobs = nobs;
    c->otype.assign(type, type + nobs);
    c->ostatus.assign(status, status + nobs);
    c->olon.assign(lon, lon + nobs);
    c->olat.assign(lat, lat + nobs);
    c->oestd.assign(estd, estd + nobs);
    c->kmod.clear();

    // trees (obs_createkdtrees)
    std::vector<BuildNode> nodes;
    std::vector<int> obsids;
    c->tinfo.assign(nt, TypeInfo{});
    for (int t = 0; t < nt; ++t) {
        const enkfg_obstype& ot = c->types[t];
        TypeInfo& ti = c->tinfo[t];
        ti.ndim = ot.geographic ? 3 : 2;
        ti.nlocrad = ot.nlocrad;
        ti.nlobsmax = ot.nlobsmax;
        double rmax = 0.0;
        for (int k = 0; k < ot.nlocrad; ++k) {
            ti.locrad[k] = ot.locrad[k];
            ti.locweight[k] = ot.locweight[k];
            if (ot.locrad[k] > rmax)
                rmax = ot.locrad[k];   // obstype_getmaxlocrad
        }
        ti.range = rmax;
        ti.node0 = (int)nodes.size();
        ti.obsid0 = (int)obsids.size();
        if (ot.statsonly)
            continue;
        std::vector<int> ids;
        for (int i = 0; i < nobs; ++i)
            if (type[i] == t && status[i] == ENKFG_STATUS_OK)
                ids.push_back(i);
        std::vector<double> xyz(3 * ids.size());
        for (size_t k = 0; k < ids.size(); ++k) {
            const int i = ids[k];
            if (ot.geographic)
                ll2xyz((double)lon[i], (double)lat[i], &xyz[3 * k]);
            else {
                xyz[3 * k] = (double)lon[i];
                xyz[3 * k + 1] = (double)lat[i];
                xyz[3 * k + 2] = 0.0;
            }
        }
        build_tree(xyz, ti.ndim, ids, nodes);
        ti.nnode = (int)(nodes.size() - ti.node0);
        // obs_getglobal lists obsids in index order, including obs the tree
        // skipped for non-finite coordinates
        obsids.insert(obsids.end(), ids.begin(), ids.end());
        ti.nids = (int)ids.size();
        ti.searchable = ids.empty() ? 0 : 1;
    }
    // row order: each type's tree in preorder, then every obs not in a tree
    {
        std::vector<int> rowof((size_t)nobs, -1);
        c->perm.clear();
        c->perm.reserve(nobs);
        for (const BuildNode& nd : nodes)
            if (rowof[nd.obsid] < 0) {
                rowof[nd.obsid] = (int)c->perm.size();
                c->perm.push_back(nd.obsid);
            }
        for (int i = 0; i < nobs; ++i)
            if (rowof[i] < 0) {
                rowof[i] = (int)c->perm.size();
                c->perm.push_back(i);
            }
        for (int& id : obsids)
            id = rowof[id];
    }
    // device nodes: the row of node k is k (rows follow the node order); in
    // preorder a left child directly follows its parent
    std::vector<KdNode> dnodes(nodes.size());
    for (int t = 0; t < nt; ++t)
        for (int v = 0; v < c->tinfo[t].nnode; ++v) {
            const BuildNode& b = nodes[(size_t)c->tinfo[t].node0 + v];
            if (b.left >= 0 && b.left != v + 1)
                return fail(c, ENKFG_ERR_ARG, "internal: kd-tree of obstype %d is not in preorder", t);
            KdNode& d = dnodes[(size_t)c->tinfo[t].node0 + v];
            d.x = b.x;
            d.y = b.y;
            d.z = b.z;
            d.right = b.right;
            d.link = (unsigned)(b.parent + 1) | (b.left >= 0 ? 0x80000000u : 0u);
        }
    HIPCHECK(c, c->d_perm.reserve(std::max(nobs, 1) * sizeof(int)));
    if (nobs > 0)
        HIPCHECK(c, hipMemcpy(c->d_perm.p, c->perm.data(), nobs * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(c, c->d_nodes.reserve(std::max<size_t>(dnodes.size(), 1) * sizeof(KdNode)));
    if (!dnodes.empty())
        HIPCHECK(c, hipMemcpy(c->d_nodes.p, dnodes.data(), dnodes.size() * sizeof(KdNode), hipMemcpyHostToDevice));
    HIPCHECK(c, c->d_obsids.reserve(std::max<size_t>(obsids.size(), 1) * sizeof(int)));
    if (!obsids.empty())
        HIPCHECK(c, hipMemcpy(c->d_obsids.p, obsids.data(), obsids.size() * sizeof(int), hipMemcpyHostToDevice));
    int r = upload_types(c);
    if (r != ENKFG_OK)
        return r;
    HIPCHECK(c, c->d_type.reserve(std::max(nobs, 1) * sizeof(short)));
    HIPCHECK(c, c->d_status.reserve(std::max(nobs, 1) * sizeof(short)));
    HIPCHECK(c, c->d_estd.reserve(std::max(nobs, 1) * sizeof(float)));
    if (nobs > 0) {
        HIPCHECK(c, hipMemcpy(c->d_type.p, type, nobs * sizeof(short), hipMemcpyHostToDevice));
        HIPCHECK(c, hipMemcpy(c->d_estd.p, estd, nobs * sizeof(float), hipMemcpyHostToDevice));
    }
    {
        int r = upload_status_rows(c);
        if (r != ENKFG_OK)
            return r;
    }
    std::vector<double> srf(std::max(nt, 1));
    for (int t = 0; t < nt; ++t)
        srf[t] = sqrt(c->types[t].rfactor);
    HIPCHECK(c, c->d_sqrtrf.reserve(srf.size() * sizeof(double)));
    HIPCHECK(c, hipMemcpy(c->d_sqrtrf.p, srf.data(), srf.size() * sizeof(double), hipMemcpyHostToDevice));
    c->obs_set = true;
    c->S_set = false;
    c->lists_valid = false;
    return ENKFG_OK;
}

int enkfg_set_obs_status(enkfg_ctx* c, const short* status)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->obs_set || status == nullptr)
        return fail(c, ENKFG_ERR_STATE, "observations not set");
    c->ostatus.assign(status, status + c->nobs);
    int r = upload_status_rows(c);
    if (r != ENKFG_OK)
        return r;
    c->lists_valid = false;
    return ENKFG_OK;
}

}  // extern "C"

namespace {

// S observation-major [nobs][m] in d_S and s [nobs] in d_s, in observation
// order: standardise (unless done), then into device row order with rows
// padded to ms floats, plus the zero row.  The result is copied back rather
// than swapped, so that d_S keeps its address and a captured calc graph that
// reads it stays valid across calls.
int finish_S(enkfg_ctx* c, bool standardised)
{
    const int m = c->cfg.nmem;
    const int ms = (m + 3) & ~3;
    const size_t nrow = (size_t)ms * (c->nobs + 1);
    if (!standardised)
        HIPCHECK(c, launch_standardise(c->stream, m, c->nobs, c->d_type.as<short>(), c->d_estd.as<float>(),
                                       c->d_sqrtrf.as<double>(), c->d_S.as<float>(), c->d_s.as<double>()));
    HIPCHECK(c, launch_gather_rows(c->stream, c->d_S.as<float>(), c->d_perm.as<int>(), c->d_tmp.as<float>(), c->nobs,
                                   m, ms));
    HIPCHECK(c, hipMemcpyAsync(c->d_S.p, c->d_tmp.p, nrow * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(c, launch_gather_rows(c->stream, c->d_s.as<double>(), c->d_perm.as<int>(), c->d_tmp2.as<double>(),
                                   c->nobs, 1, 1));
    HIPCHECK(c, hipMemcpyAsync(c->d_s.p, c->d_tmp2.p, (c->nobs + 1) * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ENKFG_OK;
}

}  // namespace

extern "C" {

int enkfg_set_S(enkfg_ctx* c, const float* S, const double* s, int standardised)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->obs_set)
        return fail(c, ENKFG_ERR_STATE, "observations not set");
    const int m = c->cfg.nmem;
    const int ms = (m + 3) & ~3;
    const size_t n = (size_t)m * c->nobs;
    const size_t nrow = (size_t)ms * (c->nobs + 1);
    if (c->nobs > 0 && (S == nullptr || s == nullptr))
        return fail(c, ENKFG_ERR_ARG, "S or s is NULL");
    for (int i = 0; i < c->nobs; ++i)
        if (!std::isfinite(s[i]))
            return fail(c, ENKFG_ERR_NONFINITE, "obs %d: non-finite innovation", i);
    HIPCHECK(c, c->d_S.reserve(std::max(n, nrow) * sizeof(float)));
    HIPCHECK(c, c->d_tmp.reserve(std::max(n, nrow) * sizeof(float)));
    HIPCHECK(c, c->d_s.reserve((c->nobs + 1) * sizeof(double)));
    HIPCHECK(c, c->d_tmp2.reserve((c->nobs + 1) * sizeof(double)));
    if (c->nobs > 0) {
        HIPCHECK(c, hipMemcpyAsync(c->d_tmp.p, S, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(c->d_s.p, s, c->nobs * sizeof(double), hipMemcpyHostToDevice, c->stream));
        // [m][nobs] -> [nobs][m]
        HIPCHECK(c, launch_members_to_cells(c->stream, c->d_tmp.as<float>(), c->d_S.as<float>(), m, c->nobs));
        const int r = finish_S(c, standardised != 0);
        if (r != ENKFG_OK)
            return r;
    }
    (void)nrow;
    c->S_set = true;
    return ENKFG_OK;
}

int enkfg_get_S(enkfg_ctx* c, float* S, double* s)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->S_set)
        return fail(c, ENKFG_ERR_STATE, "S not set");
    const int m = c->cfg.nmem;
    const size_t n = (size_t)m * c->nobs;
    if (c->nobs == 0)
        return ENKFG_OK;
    const int ms = (m + 3) & ~3;
    HIPCHECK(c, c->d_tmp.reserve(n * sizeof(float)));
    HIPCHECK(c, c->d_tmp2.reserve(n * sizeof(float)));
    if (S != nullptr) {
        HIPCHECK(c, launch_scatter_rows(c->stream, c->d_S.as<float>(), c->d_perm.as<int>(), c->d_tmp.as<float>(),
                                        c->nobs, m, ms));
        HIPCHECK(c, launch_cells_to_members(c->stream, c->d_tmp.as<float>(), c->d_tmp2.as<float>(), m, c->nobs));
        HIPCHECK(c, hipMemcpyAsync(S, c->d_tmp2.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    if (s != nullptr) {
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        HIPCHECK(c, launch_scatter_rows(c->stream, c->d_s.as<double>(), c->d_perm.as<int>(), c->d_tmp.as<double>(),
                                        c->nobs, 1, 1));
        HIPCHECK(c, hipMemcpyAsync(s, c->d_tmp.p, c->nobs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ENKFG_OK;
}

int enkfg_calc_HE(enkfg_ctx* c, const enkfg_grid* g, int ntypes, const enkfg_hop* hops, const double* fi,
                  const double* fj, const float* fk, const double* value, float* HA, double* s, double* std_out)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->obs_set)
        return fail(c, ENKFG_ERR_STATE, "observations not set");
    if (c->attached_E == nullptr)
        return fail(c, ENKFG_ERR_STATE, "no ensemble attached (enkfg_ens_attach)");
    if (c->cfg.mode == ENKFG_MODE_ENOI)
        return fail(c, ENKFG_ERR_UNSUPPORTED, "HE in EnOI mode needs the background; not supported");
    const int nt = (int)c->types.size();
    if (g == nullptr || g->nlevels == nullptr || g->mni != c->att_grid.mni || g->mnj != c->att_grid.mnj)
        return fail(c, ENKFG_ERR_ARG, "the grid must be the attached ensemble's");
    if (ntypes != nt || hops == nullptr)
        return fail(c, ENKFG_ERR_ARG, "one H operator per obs type (%d) is needed", nt);
    const int nobs = c->nobs;
    if (nobs > 0 && (fi == nullptr || fj == nullptr || fk == nullptr || value == nullptr))
        return fail(c, ENKFG_ERR_ARG, "NULL obs array");
    // H operator tables
    std::vector<int> hk(nt), nk(nt), f0(nt), fl;
    for (int t = 0; t < nt; ++t) {
        const enkfg_hop& h = hops[t];
        if (h.hfunction != ENKFG_H_SURF && h.hfunction != ENKFG_H_SUBSURF)
            return fail(c, ENKFG_ERR_UNSUPPORTED, "obs type %d: H function %d is not supported", t, h.hfunction);
        if (h.nk < 1 || h.field == nullptr)
            return fail(c, ENKFG_ERR_ARG, "obs type %d: no fields", t);
        if (g->surfk != 0 && g->surfk != h.nk - 1 && h.nk > 1)
            return fail(c, ENKFG_ERR_ARG, "obs type %d: surface layer %d is neither 0 nor nk - 1", t, g->surfk);
        hk[t] = h.hfunction;
        nk[t] = h.nk;
        f0[t] = (int)fl.size();
        for (int k = 0; k < h.nk; ++k) {
            if (h.field[k] < 0 || h.field[k] >= c->att_nfield)
                return fail(c, ENKFG_ERR_ARG, "obs type %d layer %d: field %d is not attached", t, k, h.field[k]);
            fl.push_back(h.field[k]);
        }
    }
    const int m = c->cfg.nmem;
    const size_t ncell = (size_t)g->mni * g->mnj;
    if (nobs == 0) {
        c->S_set = true;
        return ENKFG_OK;
    }
    HIPCHECK(c, c->d_hg_tab.reserve((3 * (size_t)nt + fl.size()) * sizeof(int)));
    std::vector<int> tab;
    tab.insert(tab.end(), hk.begin(), hk.end());
    tab.insert(tab.end(), nk.begin(), nk.end());
    tab.insert(tab.end(), f0.begin(), f0.end());
    tab.insert(tab.end(), fl.begin(), fl.end());
    HIPCHECK(c, hipMemcpyAsync(c->d_hg_tab.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, c->d_nlev.reserve(ncell * sizeof(int)));
    HIPCHECK(c, hipMemcpyAsync(c->d_nlev.p, g->nlevels, ncell * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, c->d_hg_obs.reserve((size_t)nobs * (3 * sizeof(double) + sizeof(float) + 2 * sizeof(short))));
    unsigned char* ob = c->d_hg_obs.as<unsigned char>();
    double* d_fi = reinterpret_cast<double*>(ob);
    double* d_fj = d_fi + nobs;
    double* d_val = d_fj + nobs;
    float* d_fk = reinterpret_cast<float*>(d_val + nobs);
    short* d_st = reinterpret_cast<short*>(d_fk + nobs);
    short* d_ty = d_st + nobs;
    HIPCHECK(c, hipMemcpyAsync(d_fi, fi, (size_t)nobs * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_fj, fj, (size_t)nobs * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_val, value, (size_t)nobs * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_fk, fk, (size_t)nobs * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_st, c->ostatus.data(), (size_t)nobs * sizeof(short), hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_ty, c->otype.data(), (size_t)nobs * sizeof(short), hipMemcpyHostToDevice,
                               c->stream));
    const int ms = (m + 3) & ~3;
    const size_t nrow = (size_t)ms * (nobs + 1);
    HIPCHECK(c, c->d_S.reserve(std::max((size_t)m * nobs, nrow) * sizeof(float)));
    HIPCHECK(c, c->d_tmp.reserve(std::max((size_t)m * nobs, nrow) * sizeof(float)));
    HIPCHECK(c, c->d_s.reserve((nobs + 1) * sizeof(double)));
    HIPCHECK(c, c->d_tmp2.reserve((nobs + 1) * sizeof(double)));
    HIPCHECK(c, c->d_hg_std.reserve((size_t)nobs * sizeof(double)));
    int zero = 0;
    HIPCHECK(c, hipMemcpyAsync(c->d_err.p, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    HeGetArgs a{};
    a.m = m;
    a.md = c->cfg.nmem_dynamic;
    a.mode = c->cfg.mode;
    a.gamma = c->cfg.gamma;
    a.ni = g->mni;
    a.nj = g->mnj;
    a.ktop = g->surfk;
    a.periodic_i = g->periodic_i;
    a.ncell = (long long)ncell;
    a.nlevels = c->d_nlev.as<int>();
    a.E = c->attached_E;
    a.nobs = nobs;
    a.type = d_ty;
    a.status = d_st;
    a.fi = d_fi;
    a.fj = d_fj;
    a.fk = d_fk;
    a.value = d_val;
    const int* dt = c->d_hg_tab.as<int>();
    a.hkind = dt;
    a.nk = dt + nt;
    a.field0 = dt + 2 * nt;
    a.fields = dt + 3 * nt;
    a.HA = c->d_S.as<float>();          // [nobs][m], observation-major
    a.s = c->d_s.as<double>();
    a.std = c->d_hg_std.as<double>();
    a.err = c->d_err.as<int>();
    HIPCHECK(c, hipEventRecord(c->ev[7], c->stream));
    HIPCHECK(c, launch_he(c->stream, a));
    HIPCHECK(c, hipEventRecord(c->ev[8], c->stream));
    int err = 0;
    HIPCHECK(c, hipMemcpyAsync(&err, c->d_err.p, sizeof err, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->timings[6] = elapsed(c->ev[7], c->ev[8]);
    if (err != 0) {
        c->S_set = false;
        return fail(c, ENKFG_ERR_NONFINITE, err == 1 ? "a forecast observation is not finite or exceeds MAXOBSVAL "
                                                        "(model2obs.c:111-119)"
                                                      : "y - Hx_f is not finite or exceeds MAXOBSVAL (ensobs.c:607-615)");
    }
    // the caller's copies: HA [nmem][nobs], s and std as das_calcinnandspread leaves them
    if (HA != nullptr) {
        HIPCHECK(c, launch_cells_to_members(c->stream, c->d_S.as<float>(), c->d_tmp.as<float>(), m, nobs));
        HIPCHECK(c, hipMemcpyAsync(HA, c->d_tmp.p, (size_t)m * nobs * sizeof(float), hipMemcpyDeviceToHost,
                                   c->stream));
    }
    if (s != nullptr)
        HIPCHECK(c, hipMemcpyAsync(s, c->d_s.p, (size_t)nobs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (std_out != nullptr)
        HIPCHECK(c, hipMemcpyAsync(std_out, c->d_hg_std.p, (size_t)nobs * sizeof(double), hipMemcpyDeviceToHost,
                                   c->stream));
    // das_moderateobs (enkf_calc.c:387 runs it when KFACTOR is finite):
    // the engine's obs errors become the moderated ones before S is
    // standardised with them
    if (std::isfinite(c->kfactor)) {
        HIPCHECK(c, c->d_kmod.reserve(std::max(nt, 1) * sizeof(int)));
        HIPCHECK(c, hipMemsetAsync(c->d_kmod.p, 0, std::max(nt, 1) * sizeof(int), c->stream));
        HIPCHECK(c, launch_moderate(c->stream, nobs, c->kfactor, d_st, d_ty, c->d_hg_std.as<double>(),
                                    c->d_s.as<double>(), c->d_estd.as<float>(), c->d_kmod.as<int>()));
        c->kmod.assign(std::max(nt, 1), 0);
        HIPCHECK(c, hipMemcpyAsync(c->kmod.data(), c->d_kmod.p, std::max(nt, 1) * sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
        HIPCHECK(c, hipMemcpyAsync(c->oestd.data(), c->d_estd.p, (size_t)nobs * sizeof(float), hipMemcpyDeviceToHost,
                                   c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    // the engine's S and s, as enkfg_set_S(HA, s, standardised = 0) leaves them
    HIPCHECK(c, launch_he_clear_bad(c->stream, d_st, nobs, m, c->d_S.as<float>(), c->d_s.as<double>()));
    int r = finish_S(c, false);
    if (r != ENKFG_OK)
        return r;
    c->S_set = true;
    return ENKFG_OK;
}

int enkfg_set_kfactor(enkfg_ctx* c, double kfactor)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!(kfactor > 0.0) && !std::isnan(kfactor))
        return fail(c, ENKFG_ERR_ARG, "KFACTOR must be positive (or NaN: off)");
    c->kfactor = kfactor;
    return ENKFG_OK;
}

int enkfg_get_obs_estd(enkfg_ctx* c, float* estd, int* nmodified)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->obs_set)
        return fail(c, ENKFG_ERR_STATE, "observations not set");
    if (estd != nullptr && c->nobs > 0)
        std::copy(c->oestd.begin(), c->oestd.end(), estd);
    if (nmodified != nullptr)
        for (int t = 0; t < (int)c->types.size(); ++t)
            nmodified[t] = t < (int)c->kmod.size() ? c->kmod[t] : 0;
    return ENKFG_OK;
}

int enkfg_localise(enkfg_ctx* c, const enkfg_nodes* nodes)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    return do_localise(c, nodes);
}

int enkfg_get_local_counts(enkfg_ctx* c, int* counts)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->lists_valid || counts == nullptr)
        return fail(c, ENKFG_ERR_STATE, "no local lists");
    const int nt = (int)c->types.size();
    std::vector<long long> off(c->nnode + 1);
    for (int k = 0; k <= c->nnode; ++k)
        HIPCHECK(c, hipMemcpy(&off[k], c->d_offsets.as<long long>() + (size_t)k * nt, sizeof(long long),
                              hipMemcpyDeviceToHost));
    for (int k = 0; k < c->nnode; ++k)
        counts[k] = (int)(off[k + 1] - off[k]);
    return ENKFG_OK;
}

int enkfg_get_localobs(enkfg_ctx* c, int node, int cap, int* ids, double* lc)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->lists_valid)
        return fail(c, ENKFG_ERR_STATE, "no local lists");
    if (node < 0 || node >= c->nnode)
        return fail(c, ENKFG_ERR_ARG, "node out of range");
    const int nt = (int)c->types.size();
    long long o[2];
    HIPCHECK(c, hipMemcpy(&o[0], c->d_offsets.as<long long>() + (size_t)node * nt, sizeof(long long),
                          hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(&o[1], c->d_offsets.as<long long>() + (size_t)(node + 1) * nt, sizeof(long long),
                          hipMemcpyDeviceToHost));
    const long long p = o[1] - o[0];
    if (p > cap)
        return fail(c, ENKFG_ERR_ARG, "cap %d < %lld local obs", cap, p);
    if (p > 0) {
        if (ids != nullptr) {
            HIPCHECK(c, hipMemcpy(ids, c->d_ids.as<int>() + o[0], p * sizeof(int), hipMemcpyDeviceToHost));
            for (long long k = 0; k < p; ++k)
                ids[k] = c->perm[ids[k]];   // device row -> obs index
        }
        if (lc != nullptr)
            HIPCHECK(c, hipMemcpy(lc, c->d_lc.as<double>() + o[0], p * sizeof(double), hipMemcpyDeviceToHost));
    }
    return (int)ENKFG_OK;
}

int enkfg_calc_transforms(enkfg_ctx* c, const enkfg_nodes* nodes)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->S_set)
        return fail(c, ENKFG_ERR_STATE, "S not set");
    int r = do_localise(c, nodes);
    if (r != ENKFG_OK)
        return r;
    const int m = c->cfg.nmem, md = c->cfg.nmem_dynamic;
    const int nt = (int)c->types.size();
    const int nn = c->nnode;
    const bool dotT = c->cfg.mode != ENKFG_MODE_ENOI;
    const bool keep64 = (c->cfg.flags & ENKFG_FLAG_KEEP_FP64) != 0;
    const bool pertype = (c->cfg.flags & ENKFG_FLAG_PERTYPE_DIAG) != 0 && nt > 0;
    HIPCHECK(c, c->d_T.reserve(std::max<size_t>(dotT ? (size_t)nn * md * m : 1, 1) * sizeof(float)));
    HIPCHECK(c, c->d_w.reserve((size_t)nn * m * sizeof(float)));
    if (keep64) {
        HIPCHECK(c, c->d_T64.reserve(std::max<size_t>(dotT ? (size_t)nn * md * m : 1, 1) * sizeof(double)));
        HIPCHECK(c, c->d_w64.reserve((size_t)nn * m * sizeof(double)));
    }
    HIPCHECK(c, c->d_pnlobs.reserve((size_t)(nt + 1) * nn * sizeof(int)));
    HIPCHECK(c, c->d_pdfs.reserve((size_t)(nt + 1) * nn * sizeof(float)));
    HIPCHECK(c, c->d_psrf.reserve((size_t)(nt + 1) * nn * sizeof(float)));

    const int mp = 16 * ((m + 15) / 16);
    const int nseg = pertype ? nt : 1;
    const size_t per_node = (size_t)nseg * mp * mp * sizeof(double);
    // batches of up to 8 GiB of Gram scratch (two buffers): fewer, larger
    // launches, each ending on a partly empty GPU less often (on the tile 11
    // batches instead of 42 with 2 GiB: FP64 calc 663 -> 658 ms)
    const size_t budget = (size_t)8 << 30;
    int nbatch = (int)std::max<size_t>(1, std::min<size_t>(budget / per_node, (size_t)nn));
    if (nbatch > 65536)
        nbatch = 65536;
    // ETKF with m > 96 keeps its iterates in device memory (etkf_ws): batches
    // of at most 1 GiB of it
    const bool etkf_ws = c->cfg.scheme == ENKFG_SCHEME_ETKF && dotT && m > 96;
    if (etkf_ws) {
        const size_t per = etkf_ws_per_node(m) * sizeof(double);
        nbatch = std::max(1, std::min(nbatch, (int)(((size_t)1 << 30) / per)));
        HIPCHECK(c, c->d_etkf.reserve((size_t)nbatch * per));
    }
    // two scratch buffers: the Gram kernel of batch b + 1 runs while the
    // factor kernel of batch b reads the other one
    const size_t gram_elems = (size_t)nbatch * per_node / sizeof(double);
    const size_t sts_elems = (size_t)nbatch * nseg * mp;
    HIPCHECK(c, c->d_gram.reserve(2 * gram_elems * sizeof(double)));
    HIPCHECK(c, c->d_sts.reserve(2 * sts_elems * sizeof(double)));
    int zero = 0;
    HIPCHECK(c, hipMemcpy(c->d_err.p, &zero, sizeof zero, hipMemcpyHostToDevice));

    CalcArgs a{};
    a.m = m;
    a.md = md;
    a.mode = c->cfg.mode;
    a.scheme = c->cfg.scheme;
    a.ntypes = nt;
    a.pertype = pertype ? 1 : 0;
    a.alpha = c->cfg.alpha;
    a.nseg = nseg;
    a.counts = c->d_counts.as<int>();
    a.offsets = c->d_offsets.as<long long>();
    a.ids = c->d_ids.as<int>();
    a.lc = c->d_lc.as<double>();
    a.S = c->d_S.as<float>();
    a.sv = c->d_s.as<double>();
    a.ms = (m + 3) & ~3;
    a.nobs = c->nobs;
    a.gram = c->d_gram.as<double>();
    a.sts = c->d_sts.as<double>();
    a.T = dotT ? c->d_T.as<float>() : nullptr;
    a.w = c->d_w.as<float>();
    a.T64 = (keep64 && dotT) ? c->d_T64.as<double>() : nullptr;
    a.w64 = keep64 ? c->d_w64.as<double>() : nullptr;
    a.pnlobs = c->d_pnlobs.as<int>();
    a.pdfs = c->d_pdfs.as<float>();
    a.psrf = c->d_psrf.as<float>();
    a.nnode = nn;
    a.err = c->d_err.as<int>();
    a.etkf_ws = etkf_ws ? c->d_etkf.as<double>() : nullptr;
    // Node order and batches.  Buckets by the node's largest segment (the
    // work of its heaviest Gram item), [0, 256), [256, 1024), [1024, 4096)
    // and [4096, ...) local obs: the nodes are processed bucket by bucket,
    // heaviest first, in grid order within a bucket (neighbours, whose lists
    // overlap, stay together for the L2); a batch never spans two buckets,
    // and each bucket's Gram tile shape comes from the autotuner below.
    std::vector<long long> all((size_t)nn * nt + 1);
    HIPCHECK(c, hipMemcpy(all.data(), c->d_offsets.p, all.size() * sizeof(long long), hipMemcpyDeviceToHost));
    std::vector<unsigned char> bucket_of((size_t)nn);
    int bcount[kBuckets] = {0, 0, 0, 0};
    for (int k = 0; k < nn; ++k) {
        long long mx = 0;
        if (pertype) {
            for (int t = 0; t < nt; ++t)
                mx = std::max(mx, all[(size_t)k * nt + t + 1] - all[(size_t)k * nt + t]);
        } else {
            mx = all[(size_t)(k + 1) * nt] - all[(size_t)k * nt];
        }
        int bk = kBuckets - 1;
        while (bk > 0 && mx < kBucketLo[bk])
            --bk;
        bucket_of[k] = (unsigned char)bk;
        bcount[bk]++;
    }
    // the item Gram kernels run the segments (obstypes) by decreasing total
    // list length, so that each launch starts with its longest items
    if (pertype && nt > 1) {
        std::vector<long long> tot((size_t)nt, 0);
        for (int k = 0; k < nn; ++k)
            for (int t = 0; t < nt; ++t)
                tot[(size_t)t] += all[(size_t)k * nt + t + 1] - all[(size_t)k * nt + t];
        std::vector<int> ord((size_t)nt);
        for (int t = 0; t < nt; ++t)
            ord[(size_t)t] = t;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return tot[(size_t)x] > tot[(size_t)y]; });
        HIPCHECK(c, c->d_segord.reserve((size_t)nt * sizeof(int)));
        HIPCHECK(c, hipMemcpy(c->d_segord.p, ord.data(), (size_t)nt * sizeof(int), hipMemcpyHostToDevice));
        a.seg_order = c->d_segord.as<int>();
    }
    int bstart[kBuckets + 1];   // bucket kBuckets-1 first
    bstart[0] = 0;
    for (int q = 0; q < kBuckets; ++q)
        bstart[q + 1] = bstart[q] + bcount[kBuckets - 1 - q];
    c->h_order.resize((size_t)nn);
    {
        int pos[kBuckets];
        for (int q = 0; q < kBuckets; ++q)
            pos[kBuckets - 1 - q] = bstart[q];
        for (int k = 0; k < nn; ++k)
            c->h_order[(size_t)pos[bucket_of[k]]++] = k;
    }
    for (int q = 0; q < kBuckets; ++q)
        c->bucket_nodes[q] = bcount[q];

    // Autotuner: the item Gram kernels run a bucket's segments with one of
    // two tile-group counts (wavefronts) per segment, item_groups_lo(MT) or
    // one more (2 or 3 for m <= 96, 4 or 5 for m = 144).  The first batch of
    // a bucket that is not yet tuned for this precision and holds at least
    // kTuneMin nodes runs both and keeps the faster (both write the same
    // Gram matrices, bit for bit); the choice holds for the engine's life.
    // Once every bucket large enough to tune is tuned and all present
    // buckets share one choice, the nodes run in plain grid order (better L2
    // reuse between neighbours) with that choice.
    const bool mixed = c->cfg.precision == ENKFG_PREC_MIXED;
    const bool items = (m + 15) / 16 <= 10;
    const int kTuneMin = std::min(nbatch, 256);
    int* tg = c->tile_groups[mixed ? 1 : 0];
    const int tg_lo = item_groups_lo((m + 15) / 16);
    const int tg_default = mixed ? tg_lo : tg_lo + 1;
    bool plain = !items;
    if (items) {
        bool tuned = true, same = true;
        int common = 0;
        for (int q = 0; q < kBuckets; ++q) {
            if (bcount[q] == 0)
                continue;
            if (tg[q] == 0 && bcount[q] >= kTuneMin)
                tuned = false;
            const int g = tg[q] ? tg[q] : tg_default;
            if (common == 0)
                common = g;
            else if (g != common)
                same = false;
        }
        plain = tuned && same;
    }
    struct Batch {
        int node0, n, bucket;
    };
    std::vector<Batch> batches;
    if (plain) {
        for (int k = 0; k < nn; k += nbatch)
            batches.push_back({k, std::min(nbatch, nn - k), -1});
        a.node_list = nullptr;
    } else {
        HIPCHECK(c, c->d_order.reserve(std::max<size_t>((size_t)nn, 1) * sizeof(int)));
        if (nn > 0)
            HIPCHECK(c, hipMemcpy(c->d_order.p, c->h_order.data(), (size_t)nn * sizeof(int), hipMemcpyHostToDevice));
        a.node_list = c->d_order.as<int>();
        for (int q = 0; q < kBuckets; ++q)
            for (int k = bstart[q]; k < bstart[q + 1]; k += nbatch)
                batches.push_back({k, std::min(nbatch, bstart[q + 1] - k), kBuckets - 1 - q});
    }
    const int nb = (int)batches.size();
    int plain_groups = 0;
    if (plain && items)
        for (int q = 0; q < kBuckets; ++q)
            if (bcount[q] > 0 && plain_groups == 0)
                plain_groups = tg[q] ? tg[q] : tg_default;

    // Graph mode: once no bucket of this run needs tuning, the batch loop is
    // captured as one HIP graph (kernel nodes and event dependencies only)
    // and replayed while its signature -- the batches, the tile groups and
    // every buffer the kernels read or write -- is unchanged.
    bool need_tune = false;
    for (const Batch& bt : batches)
        if (items && bt.bucket >= 0 && tg[bt.bucket] == 0 && bt.n >= kTuneMin)
            need_tune = true;
    if ((c->cfg.flags & ENKFG_FLAG_GRAPH) && !need_tune && nb > 0) {
        std::vector<long long> sig = {nb, (long long)mixed, plain_groups, (long long)(intptr_t)a.node_list,
                                      (long long)(intptr_t)a.seg_order,
                                      (long long)(intptr_t)a.counts, (long long)(intptr_t)a.offsets,
                                      (long long)(intptr_t)a.ids, (long long)(intptr_t)a.lc,
                                      (long long)(intptr_t)a.S, (long long)(intptr_t)a.sv,
                                      (long long)(intptr_t)c->d_gram.p, (long long)(intptr_t)c->d_sts.p,
                                      (long long)(intptr_t)a.T, (long long)(intptr_t)a.w,
                                      (long long)(intptr_t)a.T64, (long long)(intptr_t)a.w64,
                                      (long long)(intptr_t)a.pnlobs, (long long)(intptr_t)a.pdfs,
                                      (long long)(intptr_t)a.psrf, (long long)(intptr_t)a.err,
                                      (long long)(intptr_t)a.etkf_ws, nn, nbatch, a.m, a.md, a.mode, a.scheme,
                                      a.ntypes, a.pertype, a.nseg, a.ms, a.nobs,
                                      (long long)__builtin_bit_cast(long long, a.alpha)};
        for (const Batch& bt : batches) {
            sig.push_back(bt.node0);
            sig.push_back(bt.n);
            sig.push_back(bt.bucket >= 0 ? tg[bt.bucket] : -1);
        }
        if (c->graph_exec == nullptr || sig != c->graph_sig) {
            if (c->graph_exec != nullptr) {
                HIPCHECK(c, hipGraphExecDestroy(c->graph_exec));
                c->graph_exec = nullptr;
            }
            while ((int)c->gev.size() < 2 * nb + 2) {
                hipEvent_t e;
                HIPCHECK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
                c->gev.push_back(e);
            }
            hipGraph_t graph = nullptr;
            HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            // stream2 joins the capture through the fork event
            hipError_t ce = hipEventRecord(c->gev[2 * nb], c->stream);
            if (ce == hipSuccess)
                ce = hipStreamWaitEvent(c->stream2, c->gev[2 * nb], 0);
            for (int b = 0; b < nb && ce == hipSuccess; ++b) {
                const int buf = b & 1;
                a.node0 = batches[b].node0;
                a.nbatch = batches[b].n;
                a.gram = c->d_gram.as<double>() + buf * gram_elems;
                a.sts = c->d_sts.as<double>() + buf * sts_elems;
                a.tile_groups = batches[b].bucket >= 0 ? tg[batches[b].bucket] : plain_groups;
                if (b >= 2)
                    ce = hipStreamWaitEvent(c->stream, c->gev[2 * (b - 2) + 1], 0);
                if (ce == hipSuccess)
                    ce = mixed ? launch_gram_mixed(c->stream, a) : launch_gram(c->stream, a);
                if (ce == hipSuccess)
                    ce = hipEventRecord(c->gev[2 * b], c->stream);
                if (ce == hipSuccess)
                    ce = hipStreamWaitEvent(c->stream2, c->gev[2 * b], 0);
                if (ce == hipSuccess)
                    ce = launch_factor(c->stream2, a);
                if (ce == hipSuccess)
                    ce = hipEventRecord(c->gev[2 * b + 1], c->stream2);
            }
            // join stream2 back before the capture ends
            if (ce == hipSuccess)
                ce = hipEventRecord(c->gev[2 * nb + 1], c->stream2);
            if (ce == hipSuccess)
                ce = hipStreamWaitEvent(c->stream, c->gev[2 * nb + 1], 0);
            const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
            if (ce != hipSuccess || ee != hipSuccess) {
                if (graph != nullptr)
                    (void)hipGraphDestroy(graph);
                return fail(c, ENKFG_ERR_HIP, "graph capture of the calc batches: %s",
                            hipGetErrorString(ce != hipSuccess ? ce : ee));
            }
            const hipError_t ie = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                c->graph_exec = nullptr;
                return fail(c, ENKFG_ERR_HIP, "graph instantiation: %s", hipGetErrorString(ie));
            }
            c->graph_sig = sig;
            c->graph_launches = 0;
        }
        HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
        HIPCHECK(c, hipGraphLaunch(c->graph_exec, c->stream));
        HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->graph_launches++;
        c->timings[1] = elapsed(c->ev[2], c->ev[3]);
        c->timings[2] = 0.0;
        c->timings[4] = (double)nb;
        c->timings[5] = elapsed(c->ev[0], c->ev[3]);
    } else {
    // batch b: Gram on `stream` into buffer b % 2, factor on `stream2`; the
    // Gram of batch b + 2 waits for the factor of batch b (same buffer)
    while ((int)c->bev.size() < 4 * nb + 2) {
        hipEvent_t e;
        HIPCHECK(c, hipEventCreate(&e));
        c->bev.push_back(e);
    }
    for (int b = 0; b < nb; ++b) {
        const int buf = b & 1;
        hipEvent_t* e = &c->bev[4 * (size_t)b];
        a.node0 = batches[b].node0;
        a.nbatch = batches[b].n;
        a.gram = c->d_gram.as<double>() + buf * gram_elems;
        a.sts = c->d_sts.as<double>() + buf * sts_elems;
        const int bk = batches[b].bucket;
        if (b >= 2)
            HIPCHECK(c, hipStreamWaitEvent(c->stream, c->bev[4 * (size_t)(b - 2) + 3], 0));
        HIPCHECK(c, hipEventRecord(e[0], c->stream));
        if (bk < 0) {
            a.tile_groups = plain_groups;
            HIPCHECK(c, mixed ? launch_gram_mixed(c->stream, a) : launch_gram(c->stream, a));
        } else if (items && tg[bk] == 0 && a.nbatch >= kTuneMin) {
            // the default shape first; the other is kept only when it is
            // clearly faster (3%), so that run-to-run noise does not flip
            // the choice
            const int cands[2] = {tg_default, 2 * tg_lo + 1 - tg_default};
            float ms[2] = {0.0f, 0.0f};
            hipEvent_t t0 = c->bev[4 * (size_t)nb], t1 = c->bev[4 * (size_t)nb + 1];
            for (int k = 0; k < 2; ++k) {
                a.tile_groups = cands[k];
                HIPCHECK(c, hipEventRecord(t0, c->stream));
                HIPCHECK(c, mixed ? launch_gram_mixed(c->stream, a) : launch_gram(c->stream, a));
                HIPCHECK(c, hipEventRecord(t1, c->stream));
                HIPCHECK(c, hipEventSynchronize(t1));
                HIPCHECK(c, hipEventElapsedTime(&ms[k], t0, t1));
            }
            tg[bk] = (ms[1] < 0.97f * ms[0]) ? cands[1] : cands[0];
        } else {
            a.tile_groups = tg[bk];
            HIPCHECK(c, mixed ? launch_gram_mixed(c->stream, a) : launch_gram(c->stream, a));
        }
        HIPCHECK(c, hipEventRecord(e[1], c->stream));
        HIPCHECK(c, hipStreamWaitEvent(c->stream2, e[1], 0));
        HIPCHECK(c, hipEventRecord(e[2], c->stream2));
        HIPCHECK(c, launch_factor(c->stream2, a));
        HIPCHECK(c, hipEventRecord(e[3], c->stream2));
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream2));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    double tgram = 0.0, tf = 0.0;
    for (int b = 0; b < nb; ++b) {
        tgram += elapsed(c->bev[4 * (size_t)b], c->bev[4 * (size_t)b + 1]);
        tf += elapsed(c->bev[4 * (size_t)b + 2], c->bev[4 * (size_t)b + 3]);
    }
    c->timings[1] = tgram;
    c->timings[2] = tf;
    c->timings[4] = (double)nb;
    c->timings[5] = nb > 0 ? elapsed(c->ev[0], c->bev[4 * (size_t)(nb - 1) + 3]) : 0.0;
    }
    int err = 0;
    HIPCHECK(c, hipMemcpy(&err, c->d_err.p, sizeof err, hipMemcpyDeviceToHost));
    if (err != 0)
        return fail(c, ENKFG_ERR_LINALG, "non-positive pivot in the factorisation of I + S^T S, or no "
                                         "convergence of the ETKF square root");

    // calcstats (transforms.c:44-50) from the counts
    std::vector<long long> off((size_t)nn + 1);
    for (int k = 0; k <= nn; ++k)
        off[k] = all[(size_t)k * nt];
    enkfg_stats st{};
    for (int k = 0; k < nn; ++k) {
        const int p = (int)(off[k + 1] - off[k]);
        st.nlobs_sum += p;
        if (p > st.nlobs_max)
            st.nlobs_max = p;
        st.ncell++;
        if (p > 0) {
            if (p >= m)
                st.n_inv_ens++;
            else
                st.n_inv_obs++;
        }
    }
    c->stats = st;
    c->transforms_valid = true;
    c->fp64_valid = keep64;
    c->diag_valid = true;
    return ENKFG_OK;
}

int enkfg_get_transforms(enkfg_ctx* c, float* T, float* w, double* T64, double* w64)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->transforms_valid)
        return fail(c, ENKFG_ERR_STATE, "no transforms");
    const size_t nn = (size_t)c->nnode, m = c->cfg.nmem, md = c->cfg.nmem_dynamic;
    const bool dotT = c->cfg.mode != ENKFG_MODE_ENOI;
    if (T != nullptr && dotT)
        HIPCHECK(c, hipMemcpy(T, c->d_T.p, nn * md * m * sizeof(float), hipMemcpyDeviceToHost));
    if (w != nullptr)
        HIPCHECK(c, hipMemcpy(w, c->d_w.p, nn * m * sizeof(float), hipMemcpyDeviceToHost));
    if ((T64 != nullptr || w64 != nullptr) && !c->fp64_valid)
        return fail(c, ENKFG_ERR_STATE, "FP64 transforms not kept (ENKFG_FLAG_KEEP_FP64)");
    if (T64 != nullptr && dotT)
        HIPCHECK(c, hipMemcpy(T64, c->d_T64.p, nn * md * m * sizeof(double), hipMemcpyDeviceToHost));
    if (w64 != nullptr)
        HIPCHECK(c, hipMemcpy(w64, c->d_w64.p, nn * m * sizeof(double), hipMemcpyDeviceToHost));
    return ENKFG_OK;
}

int enkfg_get_diag(enkfg_ctx* c, int* pnlobs, float* pdfs, float* psrf)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->diag_valid)
        return fail(c, ENKFG_ERR_STATE, "no diagnostics");
    const size_t n = (c->types.size() + 1) * (size_t)c->nnode;
    if (pnlobs != nullptr)
        HIPCHECK(c, hipMemcpy(pnlobs, c->d_pnlobs.p, n * sizeof(int), hipMemcpyDeviceToHost));
    if (pdfs != nullptr)
        HIPCHECK(c, hipMemcpy(pdfs, c->d_pdfs.p, n * sizeof(float), hipMemcpyDeviceToHost));
    if (psrf != nullptr)
        HIPCHECK(c, hipMemcpy(psrf, c->d_psrf.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return ENKFG_OK;
}

int enkfg_get_stats(enkfg_ctx* c, enkfg_stats* st)
{
    if (check_ctx(c) || st == nullptr)
        return ENKFG_ERR_ARG;
    if (!c->diag_valid)
        return fail(c, ENKFG_ERR_STATE, "no statistics");
    *st = c->stats;
    return ENKFG_OK;
}

int enkfg_get_timings(enkfg_ctx* c, double* ms, int n)
{
    if (check_ctx(c) || ms == nullptr)
        return ENKFG_ERR_ARG;
    for (int k = 0; k < n && k < 8; ++k)
        ms[k] = c->timings[k];
    return ENKFG_OK;
}

int enkfg_get_tuning(enkfg_ctx* c, int* out, int n)
{
    if (check_ctx(c) || out == nullptr || n < 1)
        return ENKFG_ERR_ARG;
    std::vector<int> v;
    v.push_back(kBuckets);
    for (int q = 0; q < kBuckets; ++q) {
        v.push_back(kBucketLo[q]);
        v.push_back(c->bucket_nodes[q]);
        v.push_back(c->tile_groups[0][q]);
        v.push_back(c->tile_groups[1][q]);
    }
    for (int k = 0; k < n && k < (int)v.size(); ++k)
        out[k] = v[k];
    return ENKFG_OK;
}

// ---------------------------------------------------------------- update

int enkfg_set_transforms(enkfg_ctx* c, int nj_s, int ni_s, int stride, const float* T, const float* w)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    const bool dotT = c->cfg.mode != ENKFG_MODE_ENOI;
    if (nj_s <= 0 || ni_s <= 0 || stride <= 0 || w == nullptr || (dotT && T == nullptr))
        return fail(c, ENKFG_ERR_ARG, "invalid transforms");
    const size_t nn = (size_t)nj_s * ni_s, m = c->cfg.nmem, md = c->cfg.nmem_dynamic;
    HIPCHECK(c, c->d_w.reserve(nn * m * sizeof(float)));
    HIPCHECK(c, hipMemcpy(c->d_w.p, w, nn * m * sizeof(float), hipMemcpyHostToDevice));
    if (dotT) {
        HIPCHECK(c, c->d_T.reserve(nn * md * m * sizeof(float)));
        HIPCHECK(c, hipMemcpy(c->d_T.p, T, nn * md * m * sizeof(float), hipMemcpyHostToDevice));
    }
    c->nj_s = nj_s;
    c->ni_s = ni_s;
    c->stride = stride;
    c->nnode = (int)nn;
    c->transforms_valid = true;
    c->fp64_valid = false;
    c->diag_valid = false;
    return ENKFG_OK;
}

namespace {

int check_grid(enkfg_ctx* c, const enkfg_grid* g)
{
    if (g == nullptr || g->mni <= 0 || g->mnj <= 0 || g->nlevels == nullptr)
        return fail(c, ENKFG_ERR_ARG, "invalid grid");
    if (!c->transforms_valid)
        return fail(c, ENKFG_ERR_STATE, "no transforms for the update");
    if (g->stride != c->stride)
        return fail(c, ENKFG_ERR_ARG, "grid stride = %d is not equal to the stride of transforms = %d", g->stride,
                    c->stride);
    if ((g->mni - 1) / g->stride + 1 != c->ni_s || (g->mnj - 1) / g->stride + 1 != c->nj_s)
        return fail(c, ENKFG_ERR_ARG, "grid %d x %d does not match the %d x %d transforms grid", g->mnj, g->mni,
                    c->nj_s, c->ni_s);
    return ENKFG_OK;
}

int upload_fields(enkfg_ctx* c, const enkfg_grid* g, int nfield, const enkfg_field* f)
{
    std::vector<int> lev(nfield), sf(nfield), lg(nfield);
    std::vector<float> inf(nfield);
    std::vector<double> rat(nfield);
    for (int k = 0; k < nfield; ++k) {
        lev[k] = f[k].level;
        sf[k] = f[k].issurfacevar;
        lg[k] = f[k].applylog;
        inf[k] = f[k].inflation;
        rat[k] = f[k].inf_ratio;
    }
    HIPCHECK(c, c->d_flevel.reserve(nfield * sizeof(int)));
    HIPCHECK(c, c->d_fsurf.reserve(nfield * sizeof(int)));
    HIPCHECK(c, c->d_flog.reserve(nfield * sizeof(int)));
    HIPCHECK(c, c->d_finfl.reserve(nfield * sizeof(float)));
    HIPCHECK(c, c->d_fratio.reserve(nfield * sizeof(double)));
    HIPCHECK(c, hipMemcpy(c->d_flevel.p, lev.data(), nfield * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_fsurf.p, sf.data(), nfield * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_flog.p, lg.data(), nfield * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_finfl.p, inf.data(), nfield * sizeof(float), hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(c->d_fratio.p, rat.data(), nfield * sizeof(double), hipMemcpyHostToDevice));
    const size_t ncell = (size_t)g->mni * g->mnj;
    HIPCHECK(c, c->d_nlev.reserve(ncell * sizeof(int)));
    HIPCHECK(c, hipMemcpy(c->d_nlev.p, g->nlevels, ncell * sizeof(int), hipMemcpyHostToDevice));
    return ENKFG_OK;
}

UpdArgs upd_args(enkfg_ctx* c, const enkfg_grid* g, int nfield, int j1, int jsize, int flags)
{
    UpdArgs a{};
    a.m = c->cfg.nmem;
    a.md = c->cfg.nmem_dynamic;
    a.mode = c->cfg.mode;
    a.mni = g->mni;
    a.mnj = g->mnj;
    a.stride = g->stride;
    a.periodic_i = g->periodic_i;
    a.surfk = g->surfk;
    a.ni_s = c->ni_s;
    a.nj_s = c->nj_s;
    a.j1 = j1;
    a.jsize = jsize;
    a.nfield = nfield;
    a.outputinc = (flags & ENKFG_UPD_OUTPUTINC) ? 1 : 0;
    a.nlevels = c->d_nlev.as<int>();
    a.T = c->d_T.as<float>();
    a.w = c->d_w.as<float>();
    a.flevel = c->d_flevel.as<int>();
    a.fsurf = c->d_fsurf.as<int>();
    a.fapplylog = c->d_flog.as<int>();
    a.finfl = c->d_finfl.as<float>();
    a.finfratio = c->d_fratio.as<double>();
    a.ncell = (long long)jsize * g->mni;
    return a;
}

// drand48 (POSIX): X <- (0x5DEECE66D X + 0xB) mod 2^48, value X / 2^48
double rand48_next(unsigned long long& X)
{
    X = (0x5DEECE66Dull * X + 0xBull) & ((1ull << 48) - 1);
    return (double)X * (1.0 / 281474976710656.0);
}

// get_normalpair (common/utils.c:1477-1484) on the stream X
void rand48_normalpair(unsigned long long& X, double* x)
{
    const double u = sqrt(-2.0 * log(1.0 - rand48_next(X)));
    const double v = (M_PI * 2.0) * rand48_next(X);
    // gcc -O2 turns the reference's cos(v), sin(v) into one sincos() call,
    // whose results can differ from the separate calls in the last bit
    double sv, cv;
    sincos(v, &sv, &cv);
    x[0] = u * cv;
    x[1] = u * sv;
}

// addnoise (update/update.c:46-87) for the randomised fields of a batch of
// nfield fields in E [field][ncell][m]: the noise vectors are drawn on the
// host from the context's stream, in field order, then applied on the device
int apply_randomise(enkfg_ctx* c, int nfield, const enkfg_field* fields, float* E, long long ncell)
{
    const int m = c->cfg.nmem;
    std::vector<int> fid;
    for (int f = 0; f < nfield; ++f)
        if (std::isfinite(fields[f].deflation))
            fid.push_back(f);
    if (fid.empty())
        return ENKFG_OK;
    const int npair = (m + 1) / 2;
    std::vector<double> r((size_t)fid.size() * 2 * npair);
    for (size_t k = 0; k < fid.size(); ++k) {
        double* rk = &r[k * 2 * npair];
        for (int e = 0; e < npair; ++e)
            rand48_normalpair(c->rand48, &rk[2 * e]);
        double sum = 0.0;
        for (int e = 0; e < m; ++e)
            sum += rk[e];
        sum /= (double)m;
        for (int e = 0; e < m; ++e)
            rk[e] -= sum;
    }
    HIPCHECK(c, c->d_noise.reserve(r.size() * sizeof(double)));
    HIPCHECK(c, hipMemcpyAsync(c->d_noise.p, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    for (size_t k = 0; k < fid.size(); ++k) {
        const enkfg_field& fl = fields[fid[k]];
        const float mult = (float)(sqrt(1.0 - fl.deflation * fl.deflation) * fl.sigma);
        HIPCHECK(c, launch_addnoise(c->stream, E + (size_t)fid[k] * ncell * m, ncell, m, (float)fl.deflation, mult,
                                    c->d_noise.as<double>() + k * 2 * npair));
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return ENKFG_OK;
}

}  // namespace

int enkfg_set_rand48(enkfg_ctx* c, unsigned long long state)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    c->rand48 = state & ((1ull << 48) - 1);
    return ENKFG_OK;
}

int enkfg_get_rand48(const enkfg_ctx* c, unsigned long long* state)
{
    if (c == nullptr || state == nullptr)
        return ENKFG_ERR_ARG;
    *state = c->rand48;
    return ENKFG_OK;
}

int enkfg_rand48_normals(unsigned long long* state, int npair, double* out)
{
    if (state == nullptr || npair < 0 || (npair > 0 && out == nullptr))
        return ENKFG_ERR_ARG;
    unsigned long long X = *state & ((1ull << 48) - 1);
    for (int k = 0; k < npair; ++k)
        rand48_normalpair(X, &out[2 * k]);
    *state = X;
    return ENKFG_OK;
}

int enkfg_update_fields(enkfg_ctx* c, const enkfg_grid* g, int nfield, const enkfg_field* fields, int j1, int j2,
                        float** fb, float* infl, int flags)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (c->cfg.mode == ENKFG_MODE_ENOI)
        return fail(c, ENKFG_ERR_ARG, "das_updatefields needs EnKF or hybrid mode");
    int r = check_grid(c, g);
    if (r != ENKFG_OK)
        return r;
    if (nfield <= 0 || fields == nullptr || fb == nullptr || j1 < 0 || j2 < j1 || j2 >= g->mnj || j1 % g->stride)
        return fail(c, ENKFG_ERR_ARG, "invalid field batch");
    r = upload_fields(c, g, nfield, fields);
    if (r != ENKFG_OK)
        return r;
    const int m = c->cfg.nmem;
    const int jsize = j2 - j1 + 1;
    const size_t ncell = (size_t)jsize * g->mni;
    HIPCHECK(c, c->d_E.reserve((size_t)nfield * ncell * m * sizeof(float)));
    HIPCHECK(c, c->d_tmp.reserve(ncell * m * sizeof(float)));
    for (int f = 0; f < nfield; ++f) {
        if (fb[f] == nullptr)
            return fail(c, ENKFG_ERR_ARG, "field buffer %d is NULL", f);
        HIPCHECK(c, hipMemcpyAsync(c->d_tmp.p, fb[f], ncell * m * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, launch_members_to_cells(c->stream, c->d_tmp.as<float>(), c->d_E.as<float>() + (size_t)f * ncell * m,
                                            m, (long long)ncell));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    UpdArgs a = upd_args(c, g, nfield, j1, jsize, flags);
    a.E = c->d_E.as<float>();
    if (infl != nullptr) {
        HIPCHECK(c, c->d_infl.reserve((size_t)nfield * ncell * sizeof(float)));
        a.infl = c->d_infl.as<float>();
    }
    if ((flags & ENKFG_UPD_HYBRID_CONDITION) && c->cfg.mode == ENKFG_MODE_HYBRID)
        HIPCHECK(c, launch_hybrid_condition(c->stream, a, c->cfg.gamma));
    HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
    HIPCHECK(c, launch_update_mfma(c->stream, a, c->cfg.precision == ENKFG_PREC_MIXED));
    HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev[6]));
    c->timings[3] = elapsed(c->ev[5], c->ev[6]);
    r = apply_randomise(c, nfield, fields, c->d_E.as<float>(), (long long)ncell);
    if (r != ENKFG_OK)
        return r;
    for (int f = 0; f < nfield; ++f) {
        HIPCHECK(c, launch_cells_to_members(c->stream, c->d_E.as<float>() + (size_t)f * ncell * m, c->d_tmp.as<float>(),
                                            m, (long long)ncell));
        HIPCHECK(c, hipMemcpyAsync(fb[f], c->d_tmp.p, ncell * m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    if (infl != nullptr)
        HIPCHECK(c, hipMemcpy(infl, c->d_infl.p, (size_t)nfield * ncell * sizeof(float), hipMemcpyDeviceToHost));
    return ENKFG_OK;
}

int enkfg_update_fields_noalloc(enkfg_ctx* c, const enkfg_grid* g, int nfield, const enkfg_field* fields, int j1,
                                int j2, float** fb, float* infl, int flags)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (c->cfg.mode == ENKFG_MODE_ENOI)
        return fail(c, ENKFG_ERR_ARG, "das_updatefields needs EnKF or hybrid mode");
    int r = check_grid(c, g);
    if (r != ENKFG_OK)
        return r;
    if (nfield <= 0 || fields == nullptr || fb == nullptr || j1 < 0 || j2 < j1 || j2 >= g->mnj || j1 % g->stride)
        return fail(c, ENKFG_ERR_ARG, "invalid field batch");
    for (int f = 0; f < nfield; ++f)
        if (fb[f] == nullptr)
            return fail(c, ENKFG_ERR_ARG, "field buffer %d is NULL", f);
    if (c->scratch == nullptr)
        return fail(c, ENKFG_ERR_STATE, "no scratch address for this device");
    const int m = c->cfg.nmem;
    const int jsize = j2 - j1 + 1;
    const size_t ncell = (size_t)jsize * g->mni;
    // scratch layout: one field's parameters, the batch's rows of nlevels,
    // one field's noise vector, two staging copies of one field (members
    // major, cells major) and its inflation output; 256-byte aligned
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_lev = take(sizeof(int)), o_surf = take(sizeof(int)), o_log = take(sizeof(int));
    const size_t o_inf = take(sizeof(float)), o_rat = take(sizeof(double));
    const size_t o_nlev = take(ncell * sizeof(int));
    const size_t o_noise = take((size_t)m * sizeof(double));
    const size_t o_A = take(ncell * m * sizeof(float)), o_B = take(ncell * m * sizeof(float));
    const size_t o_if = take(ncell * sizeof(float));
    if (off > kNoAllocScratchBytes)
        return fail(c, ENKFG_ERR_ARG,
                    "rows %d..%d of a %d-column grid with %d members need %zu bytes of scratch; the limit of "
                    "enkfg_update_fields_noalloc is %zu bytes (split the rows into smaller tiles)",
                    j1, j2, g->mni, m, off, kNoAllocScratchBytes);
    unsigned char* base = c->scratch;
    int* d_nlev = reinterpret_cast<int*>(base + o_nlev);
    float* A = reinterpret_cast<float*>(base + o_A);
    float* B = reinterpret_cast<float*>(base + o_B);
    double* d_noise = reinterpret_cast<double*>(base + o_noise);
    HIPCHECK(c, hipMemcpyAsync(d_nlev, g->nlevels + (size_t)j1 * g->mni, ncell * sizeof(int), hipMemcpyHostToDevice,
                               c->stream));
    UpdArgs a = upd_args(c, g, 1, j1, jsize, flags);
    // the kernels index nlevels by grid row: the staged rows start at j1
    a.nlevels = d_nlev - (ptrdiff_t)j1 * g->mni;
    a.flevel = reinterpret_cast<const int*>(base + o_lev);
    a.fsurf = reinterpret_cast<const int*>(base + o_surf);
    a.fapplylog = reinterpret_cast<const int*>(base + o_log);
    a.finfl = reinterpret_cast<const float*>(base + o_inf);
    a.finfratio = reinterpret_cast<const double*>(base + o_rat);
    a.E = B;
    a.infl = infl != nullptr ? reinterpret_cast<float*>(base + o_if) : nullptr;
    const bool condition = (flags & ENKFG_UPD_HYBRID_CONDITION) && c->cfg.mode == ENKFG_MODE_HYBRID;
    double tupd = 0.0;
    for (int f = 0; f < nfield; ++f) {
        const enkfg_field& fl = fields[f];
        const int lev = fl.level, surf = fl.issurfacevar, lg = fl.applylog;
        const float inf = fl.inflation;
        const double rat = fl.inf_ratio;
        HIPCHECK(c, hipMemcpyAsync(base + o_lev, &lev, sizeof lev, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(base + o_surf, &surf, sizeof surf, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(base + o_log, &lg, sizeof lg, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(base + o_inf, &inf, sizeof inf, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(base + o_rat, &rat, sizeof rat, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(A, fb[f], ncell * m * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, launch_members_to_cells(c->stream, A, B, m, (long long)ncell));
        if (condition)
            HIPCHECK(c, launch_hybrid_condition(c->stream, a, c->cfg.gamma));
        HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
        HIPCHECK(c, launch_update_mfma(c->stream, a, c->cfg.precision == ENKFG_PREC_MIXED));
        HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
        double r48[160];
        if (std::isfinite(fl.deflation)) {
            // addnoise for this field, drawn in field order as the batch call draws them
            const int npair = (m + 1) / 2;
            for (int e = 0; e < npair; ++e)
                rand48_normalpair(c->rand48, &r48[2 * e]);
            double sum = 0.0;
            for (int e = 0; e < m; ++e)
                sum += r48[e];
            sum /= (double)m;
            for (int e = 0; e < m; ++e)
                r48[e] -= sum;
            const float mult = (float)(sqrt(1.0 - fl.deflation * fl.deflation) * fl.sigma);
            HIPCHECK(c, hipMemcpyAsync(d_noise, r48, (size_t)m * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIPCHECK(c, launch_addnoise(c->stream, B, (long long)ncell, m, (float)fl.deflation, mult, d_noise));
        }
        HIPCHECK(c, launch_cells_to_members(c->stream, B, A, m, (long long)ncell));
        HIPCHECK(c, hipMemcpyAsync(fb[f], A, ncell * m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (infl != nullptr)
            HIPCHECK(c, hipMemcpyAsync(infl + (size_t)f * ncell, base + o_if, ncell * sizeof(float),
                                       hipMemcpyDeviceToHost, c->stream));
        // the stack values above must outlive their copies, and the next
        // field reuses the staging areas
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        tupd += elapsed(c->ev[5], c->ev[6]);
    }
    c->timings[3] = tupd;
    return ENKFG_OK;
}

int enkfg_update_HE(enkfg_ctx* c, const enkfg_grid* g, int nobs, const double* fi, const double* fj,
                    const short* status, const float* inflation, const double* inf_ratio, float* HE)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (!c->transforms_valid)
        return fail(c, ENKFG_ERR_STATE, "no transforms for the update of HE");
    if (g == nullptr || g->mni <= 0 || g->mnj <= 0 || nobs < 0)
        return fail(c, ENKFG_ERR_ARG, "invalid grid or obs count");
    if (nobs > 0 && (fi == nullptr || fj == nullptr || status == nullptr || inflation == nullptr ||
                     inf_ratio == nullptr || HE == nullptr))
        return fail(c, ENKFG_ERR_ARG, "NULL obs array");
    if (g->stride != c->stride || (g->mni - 1) / g->stride + 1 != c->ni_s || (g->mnj - 1) / g->stride + 1 != c->nj_s)
        return fail(c, ENKFG_ERR_ARG, "grid %d x %d, stride %d does not match the transforms grid %d x %d, stride %d",
                    g->mnj, g->mni, g->stride, c->nj_s, c->ni_s, c->stride);
    if (nobs == 0)
        return ENKFG_OK;
    const int m = c->cfg.nmem;
    const long long ncell = (long long)g->mni * g->mnj;
    // the obs the reference updates (ensobs.c:1321-1428): status OK and the
    // nearest row a row of the grid; the column rounding to mni is the last
    // one.  Obs whose column rounds outside the grid, or with non-finite
    // indices, are left unchanged (the reference would read past its row).
    std::vector<int> cellof((size_t)nobs, -1);
    std::vector<int> cnt((size_t)ncell + 1, 0);
    for (int o = 0; o < nobs; ++o) {
        if (status[o] != ENKFG_STATUS_OK || !std::isfinite(fi[o]) || !std::isfinite(fj[o]))
            continue;
        const double ja = fj[o] + 0.5, ia = fi[o] + 0.5;
        if (ja >= (double)g->mnj || ja <= -1.0 || ia <= -1.0 || ia >= (double)g->mni + 1.0)
            continue;
        const int j = (int)ja;
        int i = (int)ia;
        if (i == g->mni)
            i--;
        if (j < 0 || j >= g->mnj || i < 0 || i >= g->mni)
            continue;
        cellof[o] = j * g->mni + i;
        cnt[(size_t)cellof[o]]++;
    }
    std::vector<int> gcell, goff;
    std::vector<long long> start((size_t)ncell, -1);
    long long tot = 0;
    for (long long k = 0; k < ncell; ++k)
        if (cnt[(size_t)k] > 0) {
            gcell.push_back((int)k);
            goff.push_back((int)tot);
            start[(size_t)k] = tot;
            tot += cnt[(size_t)k];
        }
    goff.push_back((int)tot);
    std::vector<int> order((size_t)std::max<long long>(tot, 1));
    for (int o = 0; o < nobs; ++o)
        if (cellof[o] >= 0)
            order[(size_t)start[(size_t)cellof[o]]++] = o;
    const int ng = (int)gcell.size();
    HIPCHECK(c, c->d_he_order.reserve(order.size() * sizeof(int)));
    HIPCHECK(c, c->d_he_gcell.reserve(std::max<size_t>(gcell.size(), 1) * sizeof(int)));
    HIPCHECK(c, c->d_he_goff.reserve(goff.size() * sizeof(int)));
    HIPCHECK(c, c->d_he_infl.reserve((size_t)nobs * sizeof(float)));
    HIPCHECK(c, c->d_he_ratio.reserve((size_t)nobs * sizeof(double)));
    HIPCHECK(c, c->d_E.reserve((size_t)nobs * m * sizeof(float)));
    HIPCHECK(c, c->d_tmp.reserve((size_t)nobs * m * sizeof(float)));
    HIPCHECK(c, hipMemcpyAsync(c->d_he_order.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                               c->stream));
    if (ng > 0)
        HIPCHECK(c, hipMemcpyAsync(c->d_he_gcell.p, gcell.data(), gcell.size() * sizeof(int), hipMemcpyHostToDevice,
                                   c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->d_he_goff.p, goff.data(), goff.size() * sizeof(int), hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->d_he_infl.p, inflation, (size_t)nobs * sizeof(float), hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->d_he_ratio.p, inf_ratio, (size_t)nobs * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->d_tmp.p, HE, (size_t)nobs * m * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, launch_members_to_cells(c->stream, c->d_tmp.as<float>(), c->d_E.as<float>(), m, nobs));
    UpdArgs u{};
    u.m = m;
    u.md = c->cfg.nmem_dynamic;
    u.mode = c->cfg.mode;
    u.mni = g->mni;
    u.mnj = g->mnj;
    u.stride = g->stride;
    u.periodic_i = g->periodic_i;
    u.ni_s = c->ni_s;
    u.nj_s = c->nj_s;
    u.T = c->d_T.as<float>();
    u.w = c->d_w.as<float>();
    HeArgs h{};
    h.ngroup = ng;
    h.gcell = c->d_he_gcell.as<int>();
    h.goff = c->d_he_goff.as<int>();
    h.order = c->d_he_order.as<int>();
    h.infl0 = c->d_he_infl.as<float>();
    h.ratio = c->d_he_ratio.as<double>();
    h.HE = c->d_E.as<float>();
    HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
    HIPCHECK(c, launch_update_he(c->stream, u, h));
    HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
    HIPCHECK(c, launch_cells_to_members(c->stream, c->d_E.as<float>(), c->d_tmp.as<float>(), m, nobs));
    HIPCHECK(c, hipMemcpyAsync(HE, c->d_tmp.p, (size_t)nobs * m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->timings[3] = elapsed(c->ev[5], c->ev[6]);
    return ENKFG_OK;
}

int enkfg_update_bg(enkfg_ctx* c, const enkfg_grid* g, int nfield, const enkfg_field* fields, int j1, int j2,
                    float** fb)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    int r = check_grid(c, g);
    if (r != ENKFG_OK)
        return r;
    if (nfield <= 0 || fields == nullptr || fb == nullptr || j1 < 0 || j2 < j1 || j2 >= g->mnj || j1 % g->stride)
        return fail(c, ENKFG_ERR_ARG, "invalid field batch");
    r = upload_fields(c, g, nfield, fields);
    if (r != ENKFG_OK)
        return r;
    const int m1 = c->cfg.nmem + 1;
    const int jsize = j2 - j1 + 1;
    const size_t ncell = (size_t)jsize * g->mni;
    HIPCHECK(c, c->d_E.reserve((size_t)nfield * ncell * m1 * sizeof(float)));
    HIPCHECK(c, c->d_tmp.reserve(ncell * m1 * sizeof(float)));
    for (int f = 0; f < nfield; ++f) {
        HIPCHECK(c, hipMemcpyAsync(c->d_tmp.p, fb[f], ncell * m1 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, launch_members_to_cells(c->stream, c->d_tmp.as<float>(),
                                            c->d_E.as<float>() + (size_t)f * ncell * m1, m1, (long long)ncell));
    }
    UpdArgs a = upd_args(c, g, nfield, j1, jsize, 0);
    a.E = c->d_E.as<float>();
    HIPCHECK(c, launch_update_bg(c->stream, a));
    for (int f = 0; f < nfield; ++f) {
        HIPCHECK(c, launch_cells_to_members(c->stream, c->d_E.as<float>() + (size_t)f * ncell * m1,
                                            c->d_tmp.as<float>(), m1, (long long)ncell));
        HIPCHECK(c, hipMemcpyAsync(fb[f], c->d_tmp.p, ncell * m1 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    return ENKFG_OK;
}

int enkfg_ens_attach(enkfg_ctx* c, const enkfg_grid* g, int nfield, const enkfg_field* fields, float* dev_E)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    // the ensemble may be attached before any transforms exist (enkfg_calc_HE
    // reads it first); the update checks the grid against the transforms
    if (g == nullptr || g->mni <= 0 || g->mnj <= 0 || g->nlevels == nullptr || g->stride <= 0)
        return fail(c, ENKFG_ERR_ARG, "invalid grid");
    int r = ENKFG_OK;
    if (nfield <= 0 || fields == nullptr || dev_E == nullptr)
        return fail(c, ENKFG_ERR_ARG, "invalid ensemble");
    r = upload_fields(c, g, nfield, fields);
    if (r != ENKFG_OK)
        return r;
    c->attached_E = dev_E;
    c->att_grid = *g;
    c->att_grid.nlevels = nullptr;
    c->att_nfield = nfield;
    c->att_fields.assign(fields, fields + nfield);
    return ENKFG_OK;
}

int enkfg_ens_update(enkfg_ctx* c, int flags, float* dev_infl)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (c->attached_E == nullptr)
        return fail(c, ENKFG_ERR_STATE, "no ensemble attached");
    if (c->cfg.mode == ENKFG_MODE_ENOI)
        return fail(c, ENKFG_ERR_ARG, "das_updatefields needs EnKF or hybrid mode");
    {
        enkfg_grid g = c->att_grid;
        g.nlevels = reinterpret_cast<const int*>(c->d_nlev.p);   // checked for presence only
        const int r = check_grid(c, &g);
        if (r != ENKFG_OK)
            return r;
    }
    UpdArgs a = upd_args(c, &c->att_grid, c->att_nfield, 0, c->att_grid.mnj, flags);
    a.E = c->attached_E;
    a.infl = dev_infl;
    if ((flags & ENKFG_UPD_HYBRID_CONDITION) && c->cfg.mode == ENKFG_MODE_HYBRID)
        HIPCHECK(c, launch_hybrid_condition(c->stream, a, c->cfg.gamma));
    HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
    HIPCHECK(c, launch_update_mfma(c->stream, a, c->cfg.precision == ENKFG_PREC_MIXED));
    HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev[6]));
    c->timings[3] = elapsed(c->ev[5], c->ev[6]);
    return apply_randomise(c, c->att_nfield, c->att_fields.data(), c->attached_E, a.ncell);
}

int enkfg_ens_spread(enkfg_ctx* c, int f0, int nfield, float* spread)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (c->attached_E == nullptr)
        return fail(c, ENKFG_ERR_STATE, "no ensemble attached");
    if (f0 < 0 || nfield <= 0 || f0 + nfield > c->att_nfield || spread == nullptr)
        return fail(c, ENKFG_ERR_ARG, "invalid field range %d + %d of %d, or NULL output", f0, nfield, c->att_nfield);
    const int m = c->cfg.nmem;
    const long long ncell = (long long)c->att_grid.mni * c->att_grid.mnj;
    const size_t bytes = (size_t)nfield * ncell * sizeof(float);
    HIPCHECK(c, c->d_spread.reserve(bytes));
    HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
    HIPCHECK(c, launch_spread(c->stream, c->attached_E + (size_t)f0 * ncell * m, nfield, ncell, m, true,
                              c->d_spread.as<float>()));
    HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
    HIPCHECK(c, hipMemcpyAsync(spread, c->d_spread.p, bytes, hipMemcpyDefault, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->timings[7] = elapsed(c->ev[5], c->ev[6]);
    return ENKFG_OK;
}

int enkfg_spread_fields(enkfg_ctx* c, int nfield, int nij, float* const* fieldbuffer, float* spread)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (nfield <= 0 || nij <= 0 || fieldbuffer == nullptr || spread == nullptr)
        return fail(c, ENKFG_ERR_ARG, "invalid field buffers");
    const int m = c->cfg.nmem;
    const size_t per = (size_t)m * nij;
    HIPCHECK(c, c->d_tmp.reserve(per * sizeof(float)));
    HIPCHECK(c, c->d_spread.reserve((size_t)nfield * nij * sizeof(float)));
    double t = 0.0;
    for (int f = 0; f < nfield; ++f) {
        if (fieldbuffer[f] == nullptr)
            return fail(c, ENKFG_ERR_ARG, "field buffer %d is NULL", f);
        HIPCHECK(c, hipMemcpyAsync(c->d_tmp.p, fieldbuffer[f], per * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipEventRecord(c->ev[5], c->stream));
        HIPCHECK(c, launch_spread(c->stream, c->d_tmp.as<float>(), 1, nij, m, false,
                                  c->d_spread.as<float>() + (size_t)f * nij));
        HIPCHECK(c, hipEventRecord(c->ev[6], c->stream));
        HIPCHECK(c, hipEventSynchronize(c->ev[6]));
        t += elapsed(c->ev[5], c->ev[6]);
    }
    HIPCHECK(c, hipMemcpyAsync(spread, c->d_spread.p, (size_t)nfield * nij * sizeof(float), hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->timings[7] = t;
    return ENKFG_OK;
}

int enkfg_adjoint(enkfg_ctx* c, const float* dev_dJ_dmean, double* dJ_dy, double* dJ_dR, double* dJ_dinfl)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (dev_dJ_dmean == nullptr)
        return fail(c, ENKFG_ERR_ARG, "dJ/dmean is NULL");
    if (!c->S_set || !c->lists_valid || !c->transforms_valid)
        return fail(c, ENKFG_ERR_STATE, "the adjoint needs S, and enkfg_calc_transforms on the same grid");
    if (c->attached_E == nullptr)
        return fail(c, ENKFG_ERR_STATE, "the adjoint needs the forecast ensemble attached (enkfg_ens_attach)");
    const int m = c->cfg.nmem;
    if (m > 96)
        return fail(c, ENKFG_ERR_UNSUPPORTED, "the adjoint supports nmem <= 96 (got %d)", m);
    const enkfg_grid& g = c->att_grid;
    if (g.stride != c->stride || (g.mni - 1) / g.stride + 1 != c->ni_s || (g.mnj - 1) / g.stride + 1 != c->nj_s)
        return fail(c, ENKFG_ERR_ARG, "the attached grid does not match the node grid of the last calc");
    const int nn = c->nnode;
    const int nt = (int)c->types.size();
    // g_n = sum over cells of beta(c, n) sum_f dJ/dmean A^f
    HIPCHECK(c, c->d_adjg.reserve((size_t)nn * m * sizeof(double)));
    HIPCHECK(c, hipMemsetAsync(c->d_adjg.p, 0, (size_t)nn * m * sizeof(double), c->stream));
    UpdArgs ua = upd_args(c, &g, c->att_nfield, 0, g.mnj, 0);
    ua.E = c->attached_E;
    HIPCHECK(c, launch_adj_cells(c->stream, ua, dev_dJ_dmean, c->d_adjg.as<double>()));
    // nodes with a non-zero g
    HIPCHECK(c, c->d_adjflag.reserve((size_t)nn * sizeof(int)));
    HIPCHECK(c, launch_nonzero_rows(c->stream, c->d_adjg.as<double>(), nn, m, c->d_adjflag.as<int>(), false));
    std::vector<int> flag(nn);
    HIPCHECK(c, hipMemcpyAsync(flag.data(), c->d_adjflag.p, (size_t)nn * sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::vector<int> list;
    for (int k = 0; k < nn; ++k)
        if (flag[k])
            list.push_back(k);
    const int nlist = (int)list.size();
    HIPCHECK(c, c->d_gy.reserve((size_t)std::max(c->nobs, 1) * sizeof(double)));
    HIPCHECK(c, c->d_gR.reserve((size_t)std::max(c->nobs, 1) * sizeof(double)));
    HIPCHECK(c, hipMemsetAsync(c->d_gy.p, 0, (size_t)std::max(c->nobs, 1) * sizeof(double), c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_gR.p, 0, (size_t)std::max(c->nobs, 1) * sizeof(double), c->stream));
    if (nlist > 0) {
        HIPCHECK(c, c->d_adjlist.reserve((size_t)nlist * sizeof(int)));
        HIPCHECK(c, hipMemcpy(c->d_adjlist.p, list.data(), (size_t)nlist * sizeof(int), hipMemcpyHostToDevice));
        HIPCHECK(c, c->d_adjz.reserve((size_t)nlist * m * sizeof(double)));
        HIPCHECK(c, c->d_adjw.reserve((size_t)nlist * m * sizeof(double)));
        const int mp = 16 * ((m + 15) / 16);
        const size_t per_node = (size_t)mp * mp * sizeof(double);
        int nbatch = (int)std::max<size_t>(1, std::min<size_t>(((size_t)2 << 30) / per_node, (size_t)nlist));
        nbatch = std::min(nbatch, 65536);
        HIPCHECK(c, c->d_gram.reserve((size_t)nbatch * per_node));
        HIPCHECK(c, c->d_sts.reserve((size_t)nbatch * mp * sizeof(double)));
        int zero = 0;
        HIPCHECK(c, hipMemcpy(c->d_err.p, &zero, sizeof zero, hipMemcpyHostToDevice));
        CalcArgs a{};
        a.m = m;
        a.md = c->cfg.nmem_dynamic;
        a.mode = c->cfg.mode;
        a.scheme = c->cfg.scheme;
        a.ntypes = nt;
        a.pertype = 0;
        a.nseg = 1;
        a.alpha = c->cfg.alpha;
        a.counts = c->d_counts.as<int>();
        a.offsets = c->d_offsets.as<long long>();
        a.ids = c->d_ids.as<int>();
        a.lc = c->d_lc.as<double>();
        a.S = c->d_S.as<float>();
        a.sv = c->d_s.as<double>();
        a.ms = (m + 3) & ~3;
        a.nobs = c->nobs;
        a.gram = c->d_gram.as<double>();
        a.sts = c->d_sts.as<double>();
        a.nnode = nn;
        a.err = c->d_err.as<int>();
        a.node_list = c->d_adjlist.as<int>();
        a.adj_g = c->d_adjg.as<double>();
        a.adj_z = c->d_adjz.as<double>();
        a.adj_w = c->d_adjw.as<double>();
        // M recomputed in FP64 from the local lists, z = M^-1 g, w = M^-1 S^T s
        for (int k0 = 0; k0 < nlist; k0 += nbatch) {
            a.node0 = k0;
            a.nbatch = std::min(nbatch, nlist - k0);
            HIPCHECK(c, launch_gram(c->stream, a));
            HIPCHECK(c, launch_factor_wave(c->stream, a));
            HIPCHECK(c, launch_adj_obs(c->stream, a, c->d_perm.as<int>(), c->d_type.as<short>(), c->d_estd.as<float>(),
                                       c->d_sqrtrf.as<double>(), nullptr, c->d_gy.as<double>(), c->d_gR.as<double>()));
        }
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        int err = 0;
        HIPCHECK(c, hipMemcpy(&err, c->d_err.p, sizeof err, hipMemcpyDeviceToHost));
        if (err != 0)
            return fail(c, ENKFG_ERR_LINALG, "non-positive pivot in the factorisation of I + S^T S, or no "
                                         "convergence of the ETKF square root");
    }
    std::vector<double> gy((size_t)c->nobs), gR((size_t)c->nobs);
    if (c->nobs > 0) {
        HIPCHECK(c, hipMemcpy(gy.data(), c->d_gy.p, (size_t)c->nobs * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHECK(c, hipMemcpy(gR.data(), c->d_gR.p, (size_t)c->nobs * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int r = 0; r < c->nobs; ++r) {
        const int o = c->perm[r];
        if (dJ_dy != nullptr)
            dJ_dy[o] = gy[r];
        if (dJ_dR != nullptr)
            dJ_dR[o] = gR[r];
    }
    if (dJ_dinfl != nullptr)
        *dJ_dinfl = 0.0;   // the analysis mean does not depend on the inflation
    return ENKFG_OK;
}

int enkfg_adjoint_ens(enkfg_ctx* c, const float* dev_dJ_dXa, double* dJ_dy, double* dJ_dR, double* dJ_dinfl)
{
    if (check_ctx(c))
        return ENKFG_ERR_ARG;
    if (dev_dJ_dXa == nullptr)
        return fail(c, ENKFG_ERR_ARG, "dJ/dXa is NULL");
    if (!c->S_set || !c->lists_valid || !c->transforms_valid)
        return fail(c, ENKFG_ERR_STATE, "the adjoint needs S, and enkfg_calc_transforms on the same grid");
    if (c->attached_E == nullptr)
        return fail(c, ENKFG_ERR_STATE, "the adjoint needs the forecast ensemble attached (enkfg_ens_attach)");
    if (c->cfg.mode == ENKFG_MODE_ENOI)
        return fail(c, ENKFG_ERR_ARG, "the analysis ensemble needs EnKF or hybrid mode");
    if (c->cfg.scheme != ENKFG_SCHEME_DENKF)
        return fail(c, ENKFG_ERR_UNSUPPORTED, "the adjoint of the analysis ensemble supports DEnKF only");
    const int m = c->cfg.nmem, md = c->cfg.nmem_dynamic;
    if (m > 96)
        return fail(c, ENKFG_ERR_UNSUPPORTED, "the adjoint supports nmem <= 96 (got %d)", m);
    const enkfg_grid& g = c->att_grid;
    if (g.stride != c->stride || (g.mni - 1) / g.stride + 1 != c->ni_s || (g.mnj - 1) / g.stride + 1 != c->nj_s)
        return fail(c, ENKFG_ERR_ARG, "the attached grid does not match the node grid of the last calc");
    const int nn = c->nnode;
    const int nt = (int)c->types.size();
    const int nf = c->att_nfield;
    const size_t dsz = (size_t)nn * md * m * sizeof(double);
    HIPCHECK(c, c->d_adjg.reserve((size_t)nn * m * sizeof(double)));
    HIPCHECK(c, c->d_adjD.reserve(dsz));
    HIPCHECK(c, c->d_adjinfl.reserve((size_t)nf * sizeof(double)));
    HIPCHECK(c, hipMemsetAsync(c->d_adjg.p, 0, (size_t)nn * m * sizeof(double), c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_adjD.p, 0, dsz, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_adjinfl.p, 0, (size_t)nf * sizeof(double), c->stream));
    UpdArgs ua = upd_args(c, &g, nf, 0, g.mnj, 0);
    ua.E = c->attached_E;
    HIPCHECK(c, launch_adj_ens_cells(c->stream, ua, dev_dJ_dXa, c->d_adjg.as<double>(), c->d_adjD.as<double>(),
                                     c->d_adjinfl.as<double>()));
    // nodes with a non-zero g or D
    HIPCHECK(c, c->d_adjflag.reserve((size_t)nn * sizeof(int)));
    HIPCHECK(c, launch_nonzero_rows(c->stream, c->d_adjg.as<double>(), nn, m, c->d_adjflag.as<int>(), false));
    HIPCHECK(c, launch_nonzero_rows(c->stream, c->d_adjD.as<double>(), nn, md * m, c->d_adjflag.as<int>(), true));
    std::vector<int> flag(nn);
    HIPCHECK(c, hipMemcpyAsync(flag.data(), c->d_adjflag.p, (size_t)nn * sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::vector<int> list;
    for (int k = 0; k < nn; ++k)
        if (flag[k])
            list.push_back(k);
    const int nlist = (int)list.size();
    const size_t nob = (size_t)std::max(c->nobs, 1);
    HIPCHECK(c, c->d_gy.reserve(nob * sizeof(double)));
    HIPCHECK(c, c->d_gR.reserve(nob * sizeof(double)));
    HIPCHECK(c, hipMemsetAsync(c->d_gy.p, 0, nob * sizeof(double), c->stream));
    HIPCHECK(c, hipMemsetAsync(c->d_gR.p, 0, nob * sizeof(double), c->stream));
    if (nlist > 0) {
        HIPCHECK(c, c->d_adjlist.reserve((size_t)nlist * sizeof(int)));
        HIPCHECK(c, hipMemcpy(c->d_adjlist.p, list.data(), (size_t)nlist * sizeof(int), hipMemcpyHostToDevice));
        HIPCHECK(c, c->d_adjz.reserve((size_t)nlist * m * sizeof(double)));
        HIPCHECK(c, c->d_adjw.reserve((size_t)nlist * m * sizeof(double)));
        const int mp = 16 * ((m + 15) / 16);
        const size_t per_node = (size_t)mp * mp * sizeof(double);
        // Gram scratch (reused for D M^-1), M^-1 and B of a batch: 3 per node
        int nbatch = (int)std::max<size_t>(1, std::min<size_t>(((size_t)2 << 30) / (3 * per_node), (size_t)nlist));
        nbatch = std::min(nbatch, 65536);
        HIPCHECK(c, c->d_gram.reserve((size_t)nbatch * per_node));
        HIPCHECK(c, c->d_minv.reserve((size_t)nbatch * per_node));
        HIPCHECK(c, c->d_adjB.reserve((size_t)nbatch * per_node));
        HIPCHECK(c, c->d_sts.reserve((size_t)nbatch * mp * sizeof(double)));
        int zero = 0;
        HIPCHECK(c, hipMemcpy(c->d_err.p, &zero, sizeof zero, hipMemcpyHostToDevice));
        CalcArgs a{};
        a.m = m;
        a.md = md;
        a.mode = c->cfg.mode;
        a.scheme = c->cfg.scheme;
        a.ntypes = nt;
        a.pertype = 0;
        a.nseg = 1;
        a.alpha = c->cfg.alpha;
        a.counts = c->d_counts.as<int>();
        a.offsets = c->d_offsets.as<long long>();
        a.ids = c->d_ids.as<int>();
        a.lc = c->d_lc.as<double>();
        a.S = c->d_S.as<float>();
        a.sv = c->d_s.as<double>();
        a.ms = (m + 3) & ~3;
        a.nobs = c->nobs;
        a.gram = c->d_gram.as<double>();
        a.sts = c->d_sts.as<double>();
        a.nnode = nn;
        a.err = c->d_err.as<int>();
        a.node_list = c->d_adjlist.as<int>();
        a.adj_g = c->d_adjg.as<double>();
        a.adj_z = c->d_adjz.as<double>();
        a.adj_w = c->d_adjw.as<double>();
        a.adj_minv = c->d_minv.as<double>();
        // per batch: M recomputed in FP64 (z = M^-1 g, w, M^-1), B, then the obs
        for (int k0 = 0; k0 < nlist; k0 += nbatch) {
            a.node0 = k0;
            a.nbatch = std::min(nbatch, nlist - k0);
            HIPCHECK(c, launch_gram(c->stream, a));
            HIPCHECK(c, launch_factor_wave(c->stream, a));
            HIPCHECK(c, launch_adj_B(c->stream, a, c->d_adjD.as<double>(), c->d_adjB.as<double>()));
            HIPCHECK(c, launch_adj_obs(c->stream, a, c->d_perm.as<int>(), c->d_type.as<short>(), c->d_estd.as<float>(),
                                       c->d_sqrtrf.as<double>(), c->d_adjB.as<double>(), c->d_gy.as<double>(),
                                       c->d_gR.as<double>()));
        }
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        int err = 0;
        HIPCHECK(c, hipMemcpy(&err, c->d_err.p, sizeof err, hipMemcpyDeviceToHost));
        if (err != 0)
            return fail(c, ENKFG_ERR_LINALG, "non-positive pivot in the factorisation of I + S^T S");
    }
    std::vector<double> gy((size_t)c->nobs), gR((size_t)c->nobs);
    if (c->nobs > 0) {
        HIPCHECK(c, hipMemcpy(gy.data(), c->d_gy.p, (size_t)c->nobs * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHECK(c, hipMemcpy(gR.data(), c->d_gR.p, (size_t)c->nobs * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int r = 0; r < c->nobs; ++r) {
        const int o = c->perm[r];
        if (dJ_dy != nullptr)
            dJ_dy[o] = gy[r];
        if (dJ_dR != nullptr)
            dJ_dR[o] = gR[r];
    }
    if (dJ_dinfl != nullptr)
        HIPCHECK(c, hipMemcpy(dJ_dinfl, c->d_adjinfl.p, (size_t)nf * sizeof(double), hipMemcpyDeviceToHost));
    return ENKFG_OK;
}

}  // extern "C"
