/*
 * ref_harness.c -- driver for the reference's own calc/calcs.c and
 * common/kdtree.c, compiled unmodified from the reference tree into
 * oracle/_ref/libenkfref.so (see oracle/Makefile).
 *
 * TEST INFRASTRUCTURE ONLY: used to write golden vectors (tests/golden/) and
 * as the reference CPU baseline in bench.py; never part of the product.
 *
 * calcs.c needs three helpers that the reference defines in common/utils.c
 * (alloc2d, cast2d, enkf_quit); utils.c itself includes netcdf.h, which this
 * image lacks, so it is not compiled.  The helpers below reproduce the
 * documented layout contract of common/utils.c:829-935 (an array of row
 * pointers followed by one contiguous, zero-filled block) in this file's own
 * code.  kdtree.c needs the global seed_rand48 (common/utils.c:55), which is
 * only read when trees are randomised; it stays 0 here.
 *
 * The driver loop ref_calc_transforms() restates the USE_MPIQUEUE
 * das_calctransforms loop (calc/transforms.c:722-893) around the reference
 * kernels: kd_findnodeswithinrange for the search, calc_wT_* / calc_G /
 * calc_w / traceprod for the algebra.
 */
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <stdlib.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>
#include <math.h>
#include <limits.h>
#include <stdint.h>
#include <setjmp.h>
#include "kdtree.h"
#include "enkf_oracle.h"

/* ---- declarations of the reference routines used (calc/calcs.h:18-22) */
double traceprod(int transposeA, int transposeB, int m, int n, double** A, double** B, int iscompact);
void calc_G(int m, int p, double** Min, double** S, double** G);
void calc_w(int m, int p, double** G, double* s, double* w);
void calc_wT_denkf(int m, int mout, int p, double* s, double** S, double** Sa, double** M, double** G, double* w, double** T);
void calc_wT_etkf(int m, int mout, int p, double* s, double** S, double** Sa, double** M, double** G, double* w, double** T);

long int seed_rand48 = 0;

/* enkf_quit() (common/utils.c:59-99) ends the reference process.  Here it
 * returns control to the harness entry point that is running on this thread:
 * each entry point arms `quit_env` with setjmp, and enkf_quit longjmps back
 * to it, so the caller (a test or the CPU baseline) gets an error status and
 * the message from ref_last_error() instead of an aborted process.  Storage
 * allocated inside the interrupted call is not reclaimed (an error path).
 * Without an armed entry point (a direct call of a reference routine) the
 * message is printed and the process aborts, as the reference does. */
static __thread jmp_buf quit_env;
static __thread int quit_armed = 0;
static __thread char quit_msg[512];

const char* ref_last_error(void)
{
    return quit_msg;
}

void enkf_quit(char* format, ...)
{
    va_list args;

    va_start(args, format);
    vsnprintf(quit_msg, sizeof quit_msg, format, args);
    va_end(args);
    if (quit_armed) {
        quit_armed = 0;
        longjmp(quit_env, 1);
    }
    fflush(stdout);
    fprintf(stderr, "\n  ERROR: reference harness: %s\n", quit_msg);
    abort();
}

/* row-pointer matrix laid over caller storage: nj pointers, then nj * ni
 * zero-filled units */
void* cast2d(void* p, size_t nj, size_t ni, size_t unitsize)
{
    char* base = p;
    char* block = base + nj * sizeof(void*);
    void** rows = p;
    size_t j;

    if (ni == 0 || nj == 0 || unitsize == 0)
        enkf_quit("cast2d(): zero size");
    memset(base, 0, nj * sizeof(void*) + nj * ni * unitsize);
    for (j = 0; j < nj; ++j)
        rows[j] = block + j * ni * unitsize;
    return rows;
}

void* alloc2d(size_t nj, size_t ni, size_t unitsize)
{
    void* p;

    if (ni == 0 || nj == 0 || unitsize == 0)
        enkf_quit("alloc2d(): zero size");
    p = malloc(nj * sizeof(void*) + nj * ni * unitsize);
    if (p == NULL)
        enkf_quit("alloc2d(): out of memory");
    return cast2d(p, nj, ni, unitsize);
}

/* ---------------------------------------------------------------------- */
/* One local analysis with the reference kernels, as in transforms.c:794-879.
 * S: p x m column-major (S[e * p + o]), s: [p].  Outputs T (m x mout
 * column-major, i.e. C [mout][m]), w [m], dfs, srf (double, before the
 * float store of the reference), G (m x p column-major) optional. */
static int local_analysis_body(int mode, int scheme, int m, int mout, int p, const double* S, const double* s, double* T, double* w, double* dfs, double* srf, double* Gout)
{
    double** Sloc = alloc2d(m, p, sizeof(double));
    double** G = alloc2d(p, m, sizeof(double));
    double* sloc = malloc(p * sizeof(double));
    double** Tm = alloc2d(m, m, sizeof(double));
    size_t nM = (size_t) m * (2 * m + 11);
    void* Mstore = malloc(nM * sizeof(double) + (2 * m + 11) * sizeof(void*));
    double** M = Mstore;

    memcpy(Sloc[0], S, (size_t) m * p * sizeof(double));
    memcpy(sloc, s, p * sizeof(double));
    if (mode == OR_MODE_ENOI) {
        calc_G(m, p, M, Sloc, G);
        calc_w(m, p, G, sloc, w);
    } else if (scheme == OR_SCHEME_DENKF)
        calc_wT_denkf(m, mout, p, sloc, Sloc, Sloc, M, G, w, Tm);
    else
        calc_wT_etkf(m, mout, p, sloc, Sloc, Sloc, M, G, w, Tm);
    if (T != NULL && mode != OR_MODE_ENOI)
        memcpy(T, Tm[0], (size_t) m * mout * sizeof(double));
    if (dfs != NULL) {
        *dfs = traceprod(0, 0, p, m, G, Sloc, 1);
        if (srf != NULL)
            *srf = sqrt(traceprod(0, 1, m, p, Sloc, Sloc, 1) / *dfs) - 1.0;
    }
    if (Gout != NULL)
        memcpy(Gout, G[0], (size_t) m * p * sizeof(double));
    free(Sloc);
    free(G);
    free(sloc);
    free(Tm);
    free(Mstore);
    return 0;
}

/* Entry point: returns 0, or -1 when a reference routine called enkf_quit
 * (message in ref_last_error()). */
int ref_local_analysis(int mode, int scheme, int m, int mout, int p, const double* S, const double* s, double* T, double* w, double* dfs, double* srf, double* Gout)
{
    int r;

    if (setjmp(quit_env))
        return -1;
    quit_armed = 1;
    r = local_analysis_body(mode, scheme, m, mout, p, S, s, T, w, dfs, srf, Gout);
    quit_armed = 0;
    return r;
}

/* ---------------------------------------------------------------------- */
/* kd-tree: build by sequential kd_insertnode, query by
 * kd_findnodeswithinrange; returns node data and squared distances. */
void* ref_kd_build(int ndim, int n, const double* coords)
{
    kdtree* tree = kd_create("ref", ndim);
    int i;

    kd_setstorage(tree, n > 0 ? n : 1, NULL, -1);
    for (i = 0; i < n; ++i)
        kd_insertnode(tree, &coords[(size_t) i * ndim], i);
    return tree;
}

void ref_kd_destroy(void* tree)
{
    kd_destroy(tree);
}

int ref_kd_size(void* tree)
{
    return (int) kd_getsize(tree);
}

int ref_kd_range(void* tree, const double* q, double range, int ordered, int cap, int* data, double* d2)
{
    size_t n, i;
    kdresult* res;

    kd_findnodeswithinrange(tree, q, range, ordered, &n, &res);
    for (i = 0; i < n && i < (size_t) cap; ++i) {
        data[i] = (int) kd_getnodedata(tree, res[i].id);
        d2[i] = res[i].distsq;
    }
    return (int) n;
}

/* ---------------------------------------------------------------------- */
/* Per-type trees as obs_createkdtrees (observations.c:1616-1733) builds them,
 * with the reference kd_insertnode. */
typedef struct {
    int nt;
    kdtree** trees;
    int** obsids;
    int* ntobs;
} ref_trees;

void ref_trees_destroy(ref_trees* rt)
{
    int t;

    if (rt == NULL)
        return;
    for (t = 0; t < rt->nt; ++t) {
        if (rt->trees[t] != NULL)
            kd_destroy(rt->trees[t]);
        free(rt->obsids[t]);
    }
    free(rt->trees);
    free(rt->obsids);
    free(rt->ntobs);
    free(rt);
}

ref_trees* ref_trees_create(const or_obs* obs, const short* status_at_build)
{
    int nt = obs->ntypes, t, i;
    ref_trees* rt = calloc(1, sizeof(ref_trees));
    kdtree** trees;
    int** obsids;
    int* ntobs;

    rt->nt = nt;
    rt->trees = trees = calloc(nt, sizeof(kdtree*));
    rt->obsids = obsids = calloc(nt, sizeof(int*));
    rt->ntobs = ntobs = calloc(nt, sizeof(int));
    for (t = 0; t < nt; ++t) {
        const or_obstype* ot = &obs->types[t];
        int n = 0;

        if (ot->statsonly)
            continue;
        for (i = 0; i < obs->nobs; ++i)
            if (obs->type[i] == t && status_at_build[i] == OR_STATUS_OK)
                n++;
        if (n == 0)
            continue;
        obsids[t] = malloc(n * sizeof(int));
        trees[t] = kd_create("t", ot->geographic ? 3 : 2);
        kd_setstorage(trees[t], n, NULL, -1);
        for (i = 0, n = 0; i < obs->nobs; ++i)
            if (obs->type[i] == t && status_at_build[i] == OR_STATUS_OK) {
                double ll[2] = { obs->lon[i], obs->lat[i] };
                double xyz[3];

                obsids[t][n] = i;
                if (ot->geographic) {
                    or_ll2xyz(ll, xyz);
                    kd_insertnode(trees[t], xyz, n);
                } else
                    kd_insertnode(trees[t], ll, n);
                n++;
            }
        ntobs[t] = n;
    }
    return rt;
}

/* The calc loop over nodes node0 .. node0 + nnode_do - 1 with the reference
 * kernels (search, algebra).  Same arguments and outputs as
 * or_calc_transforms (oracle/enkf_oracle.h), the trees from
 * ref_trees_create(). */
static int calc_transforms_body(const or_calcspec* spec, const or_obs* obs, const ref_trees* rt, const float* S, const double* s, int nnode, const double* lon, const double* lat, int geographic, const unsigned char* visible, float* Tout, float* wout, double* T64, double* w64, int* pnlobs, float* pdfs, float* psrf, or_calcstats* stats, int node0, int nnode_do)
{
    int m = spec->nmem, md = spec->nmem_dynamic, nt = obs->ntypes;
    int dotT = (spec->mode == OR_MODE_ENKF || spec->mode == OR_MODE_HYBRID);
    kdtree** trees = rt->trees;
    int** obsids = rt->obsids;
    int* ntobs = rt->ntobs;
    int* lobs = malloc((obs->nobs + 1) * sizeof(int));
    double* lco = malloc((obs->nobs + 1) * sizeof(double));
    double* w = malloc(m * sizeof(double));
    double** T = alloc2d(m, m, sizeof(double));
    size_t Mbytes = (size_t) m * (2 * m + 11) * sizeof(double) + (2 * m + 11) * sizeof(void*);
    void* Mstore = malloc(Mbytes);
    int t, i, node;

    if (stats != NULL)
        memset(stats, 0, sizeof(*stats));
    if (nnode_do < 0)
        nnode_do = nnode - node0;
    for (node = node0; node < node0 + nnode_do; ++node) {
        int p = 0, e, o;
        double xyz[3];
        double** Sloc;
        double** G;
        double** pS;
        double* sloc;

        /* obs_findlocal (observations.c:1837-1927) */
        if (isnan(lon[node] + lat[node])) {
            for (t = 0; t < nt; ++t) {
                if (ntobs[t] == 0 || obs->types[t].statsonly || (visible != NULL && !visible[t]))
                    continue;
                for (i = 0; i < ntobs[t]; ++i, ++p) {
                    lobs[p] = obsids[t][i];
                    lco[p] = 1.0;
                }
            }
        } else {
            if (geographic) {
                double ll[2] = { lon[node], lat[node] };

                or_ll2xyz(ll, xyz);
            } else {
                xyz[0] = lon[node];
                xyz[1] = lat[node];
                xyz[2] = NAN;
            }
            for (t = 0; t < nt; ++t) {
                const or_obstype* ot = &obs->types[t];
                size_t nloc, k;
                kdresult* res;

                if (ntobs[t] == 0 || ot->statsonly || (visible != NULL && !visible[t]))
                    continue;
                kd_findnodeswithinrange(trees[t], xyz, or_getmaxlocrad(ot->nlocrad, ot->locrad), (ot->nlobsmax == INT_MAX) ? 0 : 1, &nloc, &res);
                if (nloc > (size_t) ot->nlobsmax)
                    nloc = ot->nlobsmax;
                for (k = 0; k < nloc; ++k) {
                    int oid = obsids[t][kd_getnodedata(trees[t], res[k].id)];

                    if (obs->status[oid] != OR_STATUS_OK)
                        continue;
                    lobs[p] = oid;
                    lco[p] = or_calclcoeff(ot->nlocrad, ot->locrad, ot->locweight, sqrt(res[k].distsq));
                    p++;
                }
            }
        }
        if (stats != NULL) {
            if (p > stats->nlobs_max)
                stats->nlobs_max = p;
            stats->nlobs_sum += p;
            stats->ncell++;
        }
        if (p == 0) {
            if (dotT)
                for (e = 0; e < md; ++e)
                    for (o = 0; o < m; ++o) {
                        if (Tout)
                            Tout[((size_t) node * md + e) * m + o] = (e == o) ? 1.0f : 0.0f;
                        if (T64)
                            T64[((size_t) node * md + e) * m + o] = (e == o) ? 1.0 : 0.0;
                    }
            for (e = 0; e < m; ++e) {
                if (wout)
                    wout[(size_t) node * m + e] = 0.0f;
                if (w64)
                    w64[(size_t) node * m + e] = 0.0;
            }
            for (t = 0; t <= nt; ++t) {
                if (pnlobs)
                    pnlobs[(size_t) t * nnode + node] = 0;
                if (pdfs)
                    pdfs[(size_t) t * nnode + node] = 0.0f;
                if (psrf)
                    psrf[(size_t) t * nnode + node] = 0.0f;
            }
            continue;
        }
        Sloc = alloc2d(m, p, sizeof(double));
        pS = alloc2d(m, p, sizeof(double));
        G = alloc2d(p, m, sizeof(double));
        sloc = malloc(p * sizeof(double));
        for (e = 0; e < m; ++e) {
            const float* Se = &S[(size_t) e * obs->nobs];

            for (o = 0; o < p; ++o)
                Sloc[e][o] = (double) Se[lobs[o]] * lco[o];
        }
        for (o = 0; o < p; ++o)
            sloc[o] = s[lobs[o]] * lco[o];
        if (spec->mode == OR_MODE_ENOI) {
            calc_G(m, p, Mstore, Sloc, G);
            calc_w(m, p, G, sloc, w);
        } else if (spec->scheme == OR_SCHEME_DENKF)
            calc_wT_denkf(m, md, p, sloc, Sloc, Sloc, Mstore, G, w, T);
        else
            calc_wT_etkf(m, md, p, sloc, Sloc, Sloc, Mstore, G, w, T);
        if (dotT) {
            int e1, e2;

            for (e1 = 0; e1 < md; ++e1) {
                double* Ti = T[e1];

                Ti[e1] -= 1.0;
                for (e2 = 0; e2 < m; ++e2)
                    Ti[e2] *= spec->alpha;
                Ti[e1] += 1.0;
            }
            for (e1 = 0; e1 < md; ++e1)
                for (e2 = 0; e2 < m; ++e2) {
                    if (Tout)
                        Tout[((size_t) node * md + e1) * m + e2] = (float) T[e1][e2];
                    if (T64)
                        T64[((size_t) node * md + e1) * m + e2] = T[e1][e2];
                }
        }
        for (e = 0; e < m; ++e) {
            if (wout)
                wout[(size_t) node * m + e] = (float) w[e];
            if (w64)
                w64[(size_t) node * m + e] = w[e];
        }
        {
            float dfs = (float) traceprod(0, 0, p, m, G, Sloc, 1);
            float srf = (float) (sqrt(traceprod(0, 1, m, p, Sloc, Sloc, 1) / dfs) - 1.0);

            if (pnlobs)
                pnlobs[(size_t) nt * nnode + node] = p;
            if (pdfs)
                pdfs[(size_t) nt * nnode + node] = dfs;
            if (psrf)
                psrf[(size_t) nt * nnode + node] = srf;
        }
        for (t = 0; t < nt; ++t) {
            int pt = 0;
            int* plobs = malloc(p * sizeof(int));

            for (o = 0; o < p; ++o)
                if (obs->type[lobs[o]] == t)
                    plobs[pt++] = o;
            if (pnlobs)
                pnlobs[(size_t) t * nnode + node] = pt;
            if (pt == 0) {
                if (pdfs)
                    pdfs[(size_t) t * nnode + node] = 0.0f;
                if (psrf)
                    psrf[(size_t) t * nnode + node] = 0.0f;
            } else {
                double** pSt = cast2d(pS, m, pt, sizeof(double));
                float pd, ps;

                for (e = 0; e < m; ++e)
                    for (o = 0; o < pt; ++o)
                        pSt[e][o] = Sloc[e][plobs[o]];
                calc_G(m, pt, Mstore, pSt, G);
                pd = (float) traceprod(0, 0, pt, m, G, pSt, 1);
                ps = (pd > 1.0e-6) ? (float) (sqrt(traceprod(0, 1, m, pt, pSt, pSt, 1) / pd) - 1.0) : 0.0f;
                if (pdfs)
                    pdfs[(size_t) t * nnode + node] = pd;
                if (psrf)
                    psrf[(size_t) t * nnode + node] = ps;
            }
            free(plobs);
        }
        if (stats != NULL) {
            if (p >= m)
                stats->n_inv_ens++;
            else
                stats->n_inv_obs++;
        }
        free(Sloc);
        free(pS);
        free(G);
        free(sloc);
    }
    free(lobs);
    free(lco);
    free(w);
    free(T);
    free(Mstore);
    return 0;
}

/* Entry point: returns 0, or -1 when a reference routine called enkf_quit
 * (message in ref_last_error()). */
int ref_calc_transforms(const or_calcspec* spec, const or_obs* obs, const ref_trees* rt, const float* S, const double* s, int nnode, const double* lon, const double* lat, int geographic, const unsigned char* visible, float* Tout, float* wout, double* T64, double* w64, int* pnlobs, float* pdfs, float* psrf, or_calcstats* stats, int node0, int nnode_do)
{
    int r;

    if (setjmp(quit_env))
        return -1;
    quit_armed = 1;
    r = calc_transforms_body(spec, obs, rt, S, s, nnode, lon, lat, geographic, visible, Tout, wout, T64, w64, pnlobs, pdfs, psrf, stats, node0, nnode_do);
    quit_armed = 0;
    return r;
}
