/*
 * enkf_oracle.c -- CPU restatement of the EnKF-C calc/update hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of the
 * product.  Plain C99, no LAPACK: the dense algebra the reference delegates to
 * LAPACK/BLAS is written out here (Cholesky, triangular inverse, cyclic
 * Jacobi for the symmetric eigenproblem that dgesvd solves on an SPD matrix).
 *
 * Build flags: -std=c99 -O2 -ffp-contract=off, matching the reference's
 * "-std=c99 -O2" (enkf/arch/make.inc.*) where ISO C mode keeps GCC from
 * contracting a*b+c; the bit-exact parts (taper, chord distance, range search,
 * interpolation, update arithmetic) depend on it.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   or_ll2xyz, or_taper_gc, or_calclcoeff, or_getmaxlocrad
 *       closed forms (GC(0)=1, GC(1/2)=5/24 from both branches, GC(x>=1)=0,
 *       C1 continuity, chord = 2R sin(dsigma/2)) and the reference kd-tree
 *       golden lists whose tapers are computed from them.
 *   or_kd_build, or_kd_range
 *       golden vectors from the reference common/kdtree.c compiled unmodified
 *       (oracle/_ref), node-for-node; brute-force set equality.
 *   or_calc_G, or_calc_w, or_calc_wT_denkf, or_calc_wT_etkf, or_traceprod
 *       golden vectors from the reference calc/calcs.c compiled unmodified and
 *       linked to MKL (oracle/_ref); closed forms (p = 1, p = 0) and the
 *       invariants T = M^-1/2 (ETKF), T = (I + M^-1)/2 (DEnKF).
 *   or_findlocal, or_calc_transforms
 *       restatement of obs_findlocal / the USE_MPIQUEUE das_calctransforms
 *       loop, checked through the kd-tree and calcs.c golden vectors plus hand
 *       cases (empty column, global node, status filter, NLOBSMAX, per-type
 *       diagnostics == separate per-type analyses).
 *   or_update_fields, or_update_bg, or_sethybridensemble, or_standardise
 *       parity unpinned by reference execution (their sources need NetCDF and
 *       do not compile here); pinned by closed forms: T = I, w = 0 gives
 *       E^a = E^f up to the float mean round trip, exact values at stride
 *       nodes, linearity between nodes, periodic wrap, mean increment = A^f w,
 *       brute-force FP64 evaluation.
 */
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include "enkf_oracle.h"

#define OR_REARTH 6371.0                 /* common/global.h:121 */
#define OR_DEG2RAD (M_PI / 180.0)        /* common/global.h:122 */
#define OR_DFS_MIN 1.0e-6                /* calc/transforms.c:32 */
#define OR_EPSF 1.0e-6f                  /* update/update.c:41 */
#define OR_MAXOBSVAL 999999.0            /* common/global.h:119 */

/* ======================================================================
 * geometry and taper
 * ====================================================================== */

/* common/utils.c:1449-1458 */
void or_ll2xyz(const double ll[2], double xyz[3])
{
    double lon = ll[0] * OR_DEG2RAD;
    double lat = ll[1] * OR_DEG2RAD;
    double coslat = cos(lat);

    xyz[0] = OR_REARTH * cos(lon) * coslat;
    xyz[1] = OR_REARTH * sin(lon) * coslat;
    xyz[2] = OR_REARTH * sin(lat);
}

/* Gaspari-Cohn taper, common/utils.c:1425-1443 (same operation order) */
double or_taper_gc(double x)
{
    double x2, x3, f;

    if (x >= 1.0)
        return 0.0;
    x *= 2.0;
    x2 = x * x;
    x3 = x2 * x;
    if (x < 1.0)
        f = 1.0 + x2 * (-x3 / 4.0 + x2 / 2.0) + x3 * (5.0 / 8.0) - x2 * (5.0 / 3.0);
    else
        f = x2 * (x3 / 12.0 - x2 / 2.0) + x3 * (5.0 / 8.0) + x2 * (5.0 / 3.0) - x * 5.0 + 4.0 - (2.0 / 3.0) / x;
    return (f >= 0.0) ? f : 0.0;
}

/* common/obstypes.c:800-810 */
double or_calclcoeff(int nlocrad, const double* locrad, const double* locweight, double dist)
{
    double sum = 0.0;
    int i;

    for (i = 0; i < nlocrad; ++i)
        if (dist <= locrad[i])
            sum += locweight[i] * or_taper_gc(dist / locrad[i]);
    return sum;
}

/* common/obstypes.c:813-822 */
double or_getmaxlocrad(int nlocrad, const double* locrad)
{
    double max = 0.0;
    int i;

    for (i = 0; i < nlocrad; ++i)
        if (locrad[i] > max)
            max = locrad[i];
    return max;
}

/* ======================================================================
 * kd-tree
 * ====================================================================== */

void or_kd_free(or_kdtree* t)
{
    free(t->coords);
    free(t->left);
    free(t->right);
    free(t->parent);
    free(t->dir);
    free(t->data);
    memset(t, 0, sizeof(*t));
}

/* Sequential, unrandomised insertion as in kd_insertnode / _kd_insertnode
 * (common/kdtree.c:194-215, 253-274): the root splits on dimension 0, a
 * child splits on (parent split + 1) % ndim, and a point goes left iff its
 * coordinate along the split is strictly less than the node's. */
int or_kd_build(or_kdtree* t, int ndim, int n, const double* coords)
{
    int i, k;

    memset(t, 0, sizeof(*t));
    if (ndim < 1 || ndim > 3 || n < 0)
        return OR_ERR_ARG;
    t->ndim = ndim;
    if (n == 0)
        return OR_OK;
    t->coords = malloc((size_t) n * ndim * sizeof(double));
    t->left = malloc((size_t) n * sizeof(int));
    t->right = malloc((size_t) n * sizeof(int));
    t->parent = malloc((size_t) n * sizeof(int));
    t->dir = malloc((size_t) n);
    t->data = malloc((size_t) n * sizeof(int));
    if (!t->coords || !t->left || !t->right || !t->parent || !t->dir || !t->data) {
        or_kd_free(t);
        return OR_ERR_NOMEM;
    }
    for (i = 0; i < n; ++i) {
        const double* c = &coords[(size_t) i * ndim];
        int id, cur, d;

        if (!isfinite(c[0]))
            continue;
        id = t->n++;
        for (k = 0; k < ndim; ++k)
            t->coords[(size_t) id * ndim + k] = c[k];
        t->data[id] = i;
        t->left[id] = t->right[id] = -1;
        if (id == 0) {
            t->dir[0] = 0;
            t->parent[0] = -1;
            t->depth = 1;
            continue;
        }
        for (cur = 0, d = 1;;) {
            int s = t->dir[cur];
            int* slot = (c[s] < t->coords[(size_t) cur * ndim + s]) ? &t->left[cur] : &t->right[cur];

            d++;
            if (*slot < 0) {
                *slot = id;
                t->parent[id] = cur;
                t->dir[id] = (unsigned char) ((t->dir[cur] + 1) % ndim);
                break;
            }
            cur = *slot;
        }
        if (d > t->depth)
            t->depth = d;
    }
    return OR_OK;
}

/* stable merge sort of (data, d2) pairs by d2 -- glibc 2.35 qsort() is a
 * merge sort for such arrays, so equal distances keep their search order */
static void msort_d2(int n, int* data, double* d2, int* tdata, double* td2)
{
    int w, i;

    for (w = 1; w < n; w *= 2) {
        for (i = 0; i < n; i += 2 * w) {
            int a = i, amax = (i + w < n) ? i + w : n;
            int b = amax, bmax = (i + 2 * w < n) ? i + 2 * w : n;
            int o = i;

            while (a < amax && b < bmax) {
                if (d2[b] < d2[a]) {
                    tdata[o] = data[b];
                    td2[o++] = d2[b++];
                } else {
                    tdata[o] = data[a];
                    td2[o++] = d2[a++];
                }
            }
            while (a < amax) {
                tdata[o] = data[a];
                td2[o++] = d2[a++];
            }
            while (b < bmax) {
                tdata[o] = data[b];
                td2[o++] = d2[b++];
            }
        }
        memcpy(data, tdata, (size_t) n * sizeof(int));
        memcpy(d2, td2, (size_t) n * sizeof(double));
    }
}

/* _kd_findnodeswithinrange / kd_findnodeswithinrange (common/kdtree.c:
 * 516-574): preorder recursion visiting a node, then the child on the query's
 * side (left iff dx <= 0), then the other child iff |dx| < range; a node is a
 * hit iff its squared distance (summed over dimensions in order) is <=
 * range^2.  The recursion is replaced by an explicit stack: pushing the far
 * child before the near one gives the same visiting order. */
int or_kd_range(const or_kdtree* t, const double* q, double range, int ordered, int cap, int* data_out, double* d2_out)
{
    int ndim = t->ndim;
    int* stack;
    int top = 0, count = 0, overflow = 0;

    if (t->n == 0)
        return 0;
    stack = malloc((size_t) (t->n + 1) * sizeof(int));
    if (stack == NULL)
        return OR_ERR_NOMEM;
    stack[top++] = 0;
    while (top > 0) {
        int id = stack[--top];
        const double* c = &t->coords[(size_t) id * ndim];
        double distsq = 0.0, dx;
        int i, near, far;

        for (i = 0; i < ndim; i++)
            distsq += (c[i] - q[i]) * (c[i] - q[i]);
        if (distsq <= range * range) {
            if (count < cap) {
                data_out[count] = t->data[id];
                d2_out[count] = distsq;
            } else
                overflow = 1;
            count++;
        }
        dx = q[t->dir[id]] - c[t->dir[id]];
        near = (dx <= 0.0) ? t->left[id] : t->right[id];
        far = (dx <= 0.0) ? t->right[id] : t->left[id];
        if (fabs(dx) < range && far >= 0)
            stack[top++] = far;
        if (near >= 0)
            stack[top++] = near;
    }
    free(stack);
    if (overflow)
        return OR_ERR_CAP;
    if (ordered && count > 1) {
        int* td = malloc((size_t) count * sizeof(int));
        double* tx = malloc((size_t) count * sizeof(double));

        if (td == NULL || tx == NULL) {
            free(td);
            free(tx);
            return OR_ERR_NOMEM;
        }
        msort_d2(count, data_out, d2_out, td, tx);
        free(td);
        free(tx);
    }
    return count;
}

/* ======================================================================
 * observations: trees per type and local search
 * ====================================================================== */

void or_obs_free_trees(or_obs* obs)
{
    int t;

    if (obs->trees != NULL)
        for (t = 0; t < obs->ntypes; ++t)
            or_kd_free(&obs->trees[t]);
    if (obs->obsids != NULL)
        for (t = 0; t < obs->ntypes; ++t)
            free(obs->obsids[t]);
    free(obs->trees);
    free(obs->obsids);
    free(obs->ntobs);
    obs->trees = NULL;
    obs->obsids = NULL;
    obs->ntobs = NULL;
}

/* obs_find_bytype + obs_createkdtrees (observations.c:1543-1565, 1616-1733),
 * default build (no OBS_SHUFFLE): for each non-stats-only type, the OK obs in
 * ascending index order are inserted sequentially; 3D chord coordinates for
 * geographic types, (lon, lat) otherwise. */
int or_obs_build_trees(or_obs* obs, const short* status_at_build)
{
    int t, i;

    obs->ntobs = calloc((size_t) obs->ntypes, sizeof(int));
    obs->obsids = calloc((size_t) obs->ntypes, sizeof(int*));
    obs->trees = calloc((size_t) obs->ntypes, sizeof(or_kdtree));
    if (!obs->ntobs || !obs->obsids || !obs->trees)
        return OR_ERR_NOMEM;
    for (t = 0; t < obs->ntypes; ++t) {
        const or_obstype* ot = &obs->types[t];
        int ndim = ot->geographic ? 3 : 2;
        int n = 0, r;
        double* coords;

        if (ot->statsonly)
            continue;
        for (i = 0; i < obs->nobs; ++i)
            if (obs->type[i] == t && status_at_build[i] == OR_STATUS_OK)
                n++;
        if (n == 0)
            continue;
        obs->obsids[t] = malloc((size_t) n * sizeof(int));
        coords = malloc((size_t) n * ndim * sizeof(double));
        if (obs->obsids[t] == NULL || coords == NULL) {
            free(coords);
            return OR_ERR_NOMEM;
        }
        for (i = 0, n = 0; i < obs->nobs; ++i) {
            if (obs->type[i] == t && status_at_build[i] == OR_STATUS_OK) {
                double ll[2] = { obs->lon[i], obs->lat[i] };

                obs->obsids[t][n] = i;
                if (ot->geographic)
                    or_ll2xyz(ll, &coords[(size_t) n * 3]);
                else {
                    coords[(size_t) n * 2] = ll[0];
                    coords[(size_t) n * 2 + 1] = ll[1];
                }
                n++;
            }
        }
        obs->ntobs[t] = n;
        r = or_kd_build(&obs->trees[t], ndim, n, coords);
        free(coords);
        if (r != OR_OK)
            return r;
    }
    return OR_OK;
}

/* obs_findlocal (observations.c:1837-1927) with obs_getglobal (1773-1817)
 * for nodes with NaN coordinates. */
int or_findlocal(const or_obs* obs, double lon, double lat, int geographic, const unsigned char* visible, int cap, int* ids, double* lc)
{
    double xyz[3];
    int t, n = 0;

    if (obs->ntypes == 0)
        return 0;
    if (isnan(lon + lat)) {
        for (t = 0; t < obs->ntypes; ++t) {
            int i;

            if (obs->ntobs[t] == 0 || obs->types[t].statsonly)
                continue;
            if (visible != NULL && !visible[t])
                continue;
            for (i = 0; i < obs->ntobs[t]; ++i, ++n) {
                if (n >= cap)
                    return OR_ERR_CAP;
                ids[n] = obs->obsids[t][i];
                lc[n] = 1.0;
            }
        }
        return n;
    }
    if (geographic) {
        double ll[2] = { lon, lat };

        or_ll2xyz(ll, xyz);
    } else {
        xyz[0] = lon;
        xyz[1] = lat;
        xyz[2] = NAN;
    }
    for (t = 0; t < obs->ntypes; ++t) {
        const or_obstype* ot = &obs->types[t];
        const or_kdtree* tree = &obs->trees[t];
        int nres, k, ordered;
        int* rdata;
        double* rd2;

        if (obs->ntobs[t] == 0 || ot->statsonly)
            continue;
        if (visible != NULL && !visible[t])
            continue;
        ordered = (ot->nlobsmax == OR_NLOBSMAX_NONE) ? 0 : 1;
        rdata = malloc((size_t) tree->n * sizeof(int));
        rd2 = malloc((size_t) tree->n * sizeof(double));
        if (rdata == NULL || rd2 == NULL) {
            free(rdata);
            free(rd2);
            return OR_ERR_NOMEM;
        }
        nres = or_kd_range(tree, xyz, or_getmaxlocrad(ot->nlocrad, ot->locrad), ordered, tree->n, rdata, rd2);
        if (nres < 0) {
            free(rdata);
            free(rd2);
            return nres;
        }
        if (nres > ot->nlobsmax)
            nres = ot->nlobsmax;
        for (k = 0; k < nres; ++k) {
            int oid = obs->obsids[t][rdata[k]];

            if (obs->status[oid] != OR_STATUS_OK)
                continue;
            if (n >= cap) {
                free(rdata);
                free(rd2);
                return OR_ERR_CAP;
            }
            ids[n] = oid;
            lc[n] = or_calclcoeff(ot->nlocrad, ot->locrad, ot->locweight, sqrt(rd2[k]));
            n++;
        }
        free(rdata);
        free(rd2);
    }
    return n;
}

/* ======================================================================
 * standardisation: calc/ensobs.c:849-898
 * ====================================================================== */

void or_standardise(int nmem, int nobs, const short* type, const float* estd, const or_obstype* types, float* S, double* s)
{
    double mult = sqrt((double) nmem - 1);
    int e, i;

    for (e = 0; e < nmem; ++e) {
        float* Se = &S[(size_t) e * nobs];

        for (i = 0; i < nobs; ++i)
            Se[i] /= estd[i] * sqrt(types[type[i]].rfactor) * mult;
    }
    if (s != NULL)
        for (i = 0; i < nobs; ++i)
            s[i] /= estd[i] * sqrt(types[type[i]].rfactor) * mult;
}

/* ======================================================================
 * dense algebra (column-major, leading dimension = rows)
 * ====================================================================== */

/* C = alpha * op(A) * op(B) + beta * C; op(A) is M x K, op(B) is K x N */
static void gemm(int ta, int tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B, int ldb, double beta, double* C, int ldc)
{
    int i, j, k;

    for (j = 0; j < N; ++j)
        for (i = 0; i < M; ++i) {
            double sum = 0.0;

            for (k = 0; k < K; ++k) {
                double a = ta ? A[(size_t) i * lda + k] : A[(size_t) k * lda + i];
                double b = tb ? B[(size_t) k * ldb + j] : B[(size_t) j * ldb + k];

                sum += a * b;
            }
            C[(size_t) j * ldc + i] = alpha * sum + ((beta == 0.0) ? 0.0 : beta * C[(size_t) j * ldc + i]);
        }
}

/* In-place inverse of an SPD n x n matrix through its Cholesky factor
 * A = U^T U (what dpotrf('U') + dpotri('U') compute), then the lower triangle
 * is filled from the upper one (invm, calcs.c:103-132). */
static int spd_inverse(int n, double* A)
{
    int i, j, k;

    /* Cholesky: upper factor U stored in the upper triangle */
    for (j = 0; j < n; ++j) {
        double d = A[(size_t) j * n + j];

        for (k = 0; k < j; ++k)
            d -= A[(size_t) j * n + k] * A[(size_t) j * n + k];
        if (!(d > 0.0))
            return OR_ERR_LINALG;
        d = sqrt(d);
        A[(size_t) j * n + j] = d;
        for (i = j + 1; i < n; ++i) {
            double v = A[(size_t) i * n + j];

            for (k = 0; k < j; ++k)
                v -= A[(size_t) j * n + k] * A[(size_t) i * n + k];
            A[(size_t) i * n + j] = v / d;  /* U(j, i) */
        }
    }
    /* inverse of U in place (upper triangular) */
    for (j = 0; j < n; ++j) {
        A[(size_t) j * n + j] = 1.0 / A[(size_t) j * n + j];
        for (i = 0; i < j; ++i) {
            double sum = 0.0;

            for (k = i; k < j; ++k)
                sum += A[(size_t) k * n + i] * A[(size_t) j * n + k];
            A[(size_t) j * n + i] = -sum * A[(size_t) j * n + j];
        }
    }
    /* A^-1 = U^-1 U^-T, upper triangle, row i of U^-1 is A(i, i..n-1) */
    for (j = 0; j < n; ++j)
        for (i = 0; i <= j; ++i) {
            double sum = 0.0;

            for (k = j; k < n; ++k)
                sum += A[(size_t) k * n + i] * A[(size_t) k * n + j];
            A[(size_t) j * n + i] = sum;
        }
    for (j = 1; j < n; ++j)
        for (i = 0; i < j; ++i)
            A[(size_t) i * n + j] = A[(size_t) j * n + i];
    return OR_OK;
}

/* Symmetric eigen-decomposition by the cyclic Jacobi method: A (n x n) is
 * destroyed, lam[i] and column i of U are the eigenpairs.  For the SPD
 * matrices of calc_wT_etkf this is the SVD that dgesvd returns. */
static int sym_eig(int n, double* A, double* lam, double* U)
{
    int i, j, k, sweep;

    for (j = 0; j < n; ++j)
        for (i = 0; i < n; ++i)
            U[(size_t) j * n + i] = (i == j) ? 1.0 : 0.0;
    for (sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;

        for (j = 0; j < n; ++j) {
            diag += A[(size_t) j * n + j] * A[(size_t) j * n + j];
            for (i = 0; i < j; ++i)
                off += A[(size_t) j * n + i] * A[(size_t) j * n + i];
        }
        if (off <= 1e-34 * diag || off == 0.0)
            break;
        for (i = 0; i < n - 1; ++i)
            for (j = i + 1; j < n; ++j) {
                double apq = A[(size_t) j * n + i];
                double app = A[(size_t) i * n + i];
                double aqq = A[(size_t) j * n + j];
                double theta, t, c, s;

                if (apq == 0.0)
                    continue;
                theta = (aqq - app) / (2.0 * apq);
                t = ((theta >= 0.0) ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                c = 1.0 / sqrt(t * t + 1.0);
                s = t * c;
                for (k = 0; k < n; ++k) {
                    double akp = A[(size_t) i * n + k];
                    double akq = A[(size_t) j * n + k];

                    A[(size_t) i * n + k] = c * akp - s * akq;
                    A[(size_t) j * n + k] = s * akp + c * akq;
                }
                for (k = 0; k < n; ++k) {
                    double apk = A[(size_t) k * n + i];
                    double aqk = A[(size_t) k * n + j];

                    A[(size_t) k * n + i] = c * apk - s * aqk;
                    A[(size_t) k * n + j] = s * apk + c * aqk;
                }
                for (k = 0; k < n; ++k) {
                    double ukp = U[(size_t) i * n + k];
                    double ukq = U[(size_t) j * n + k];

                    U[(size_t) i * n + k] = c * ukp - s * ukq;
                    U[(size_t) j * n + k] = s * ukp + c * ukq;
                }
            }
    }
    for (i = 0; i < n; ++i)
        lam[i] = A[(size_t) i * n + i];
    return OR_OK;
}

/* ======================================================================
 * per-column algebra: calc/calcs.c
 * ====================================================================== */

/* traceprod, calcs.c:41-97, compact-storage branch */
double or_traceprod(int transposeA, int transposeB, int m, int n, const double* A, const double* B)
{
    double trace = 0.0;
    int i, j;

    if (!transposeA && !transposeB) {
        for (i = 0; i < n; ++i)
            for (j = 0; j < m; ++j)
                trace += A[(size_t) j * n + i] * B[(size_t) i * m + j];
    } else if (transposeA && transposeB) {
        for (i = 0; i < n; ++i)
            for (j = 0; j < m; ++j)
                trace += B[(size_t) j * n + i] * A[(size_t) i * m + j];
    } else {
        for (i = 0; i < n * m; ++i)
            trace += A[i] * B[i];
    }
    return trace;
}

/* calc_Minv (calcs.c:141-178): M = inv(I + S'S) (m x m) if m <= p, else
 * inv(I + S S') (p x p).  S is p x m. */
static int calc_Minv(int p, int m, const double* S, double* M)
{
    int i;

    if (m <= p) {
        gemm(1, 0, m, m, p, 1.0, S, p, S, p, 0.0, M, m);
        for (i = 0; i < m; ++i)
            M[(size_t) i * m + i] += 1.0;
        return spd_inverse(m, M);
    }
    gemm(0, 1, p, p, m, 1.0, S, p, S, p, 0.0, M, p);
    for (i = 0; i < p; ++i)
        M[(size_t) i * p + i] += 1.0;
    return spd_inverse(p, M);
}

/* calc_G (calcs.c:189-221): G = inv(I + S'S) S' = S' inv(I + S S'), m x p */
int or_calc_G(int m, int p, const double* S, double* G)
{
    int n = (p < m) ? p : m;
    double* M = malloc((size_t) n * n * sizeof(double));
    int r;

    if (M == NULL)
        return OR_ERR_NOMEM;
    r = calc_Minv(p, m, S, M);
    if (r == OR_OK) {
        if (p < m)
            gemm(1, 0, m, p, p, 1.0, S, p, M, p, 0.0, G, m);
        else
            gemm(0, 1, m, p, m, 1.0, M, m, S, p, 0.0, G, m);
    }
    free(M);
    return r;
}

/* calc_w (calcs.c:256-263): w = G s */
void or_calc_w(int m, int p, const double* G, const double* s, double* w)
{
    int e, o;

    for (e = 0; e < m; ++e) {
        double sum = 0.0;

        for (o = 0; o < p; ++o)
            sum += G[(size_t) o * m + e] * s[o];
        w[e] = sum;
    }
}

/* calc_wT_denkf (calcs.c:279-284 with calc_T_denkf 232-248): T = I - G S/2
 * (m x mout), w = G s; Sa = S (the reference calls it with Sa = S). */
int or_calc_wT_denkf(int m, int mout, int p, const double* s, const double* S, double* G, double* w, double* T)
{
    int r = or_calc_G(m, p, S, G), i;

    if (r != OR_OK)
        return r;
    gemm(0, 0, m, mout, p, -0.5, G, m, S, p, 0.0, T, m);
    for (i = 0; i < mout; ++i)
        T[(size_t) i * m + i] += 1.0;
    or_calc_w(m, p, G, s, w);
    return OR_OK;
}

/* calc_wT_etkf (calcs.c:300-412): SVD (= eigen-decomposition) of the SPD
 * matrix I + S'S (or I + S S' when p < m) gives (M + M^1/2)^-1, then
 * T = I - (M + M^1/2)^-1 S' Sa and, after rescaling the singular vectors,
 * G = M^-1 S' and w = G s. */
int or_calc_wT_etkf(int m, int mout, int p, const double* s, const double* S, double* G, double* w, double* T)
{
    int n = (p < m) ? p : m;
    double* M = malloc((size_t) n * n * sizeof(double));
    double* U = malloc((size_t) n * n * sizeof(double));
    double* sig = malloc((size_t) n * sizeof(double));
    int i, j, r = OR_OK;

    if (!M || !U || !sig) {
        r = OR_ERR_NOMEM;
        goto done;
    }
    if (p < m)
        gemm(0, 1, p, p, m, 1.0, S, p, S, p, 0.0, M, p);
    else
        gemm(1, 0, m, m, p, 1.0, S, p, S, p, 0.0, M, m);
    for (i = 0; i < n; ++i)
        M[(size_t) i * n + i] += 1.0;
    r = sym_eig(n, M, sig, U);
    if (r != OR_OK)
        goto done;
    for (i = 0; i < n; ++i) {
        double d;

        if (!(sig[i] > 0.0)) {
            r = OR_ERR_LINALG;
            goto done;
        }
        d = sqrt(sig[i] + sqrt(sig[i]));
        for (j = 0; j < n; ++j)
            U[(size_t) i * n + j] /= d;
    }
    gemm(0, 1, n, n, n, 1.0, U, n, U, n, 0.0, M, n);
    if (p < m)
        gemm(1, 0, m, p, p, 1.0, S, p, M, p, 0.0, G, m);
    else
        gemm(0, 1, m, p, m, 1.0, M, m, S, p, 0.0, G, m);
    gemm(0, 0, m, mout, p, -1.0, G, m, S, p, 0.0, T, m);
    for (i = 0; i < mout; ++i)
        T[(size_t) i * m + i] += 1.0;
    for (i = 0; i < n; ++i) {
        double d = sqrt(1.0 + 1.0 / sqrt(sig[i]));

        for (j = 0; j < n; ++j)
            U[(size_t) i * n + j] *= d;
    }
    gemm(0, 1, n, n, n, 1.0, U, n, U, n, 0.0, M, n);
    if (p < m)
        gemm(1, 0, m, p, p, 1.0, S, p, M, p, 0.0, G, m);
    else
        gemm(0, 1, m, p, m, 1.0, M, m, S, p, 0.0, G, m);
    or_calc_w(m, p, G, s, w);
  done:
    free(M);
    free(U);
    free(sig);
    return r;
}

/* ======================================================================
 * das_calctransforms, USE_MPIQUEUE build (calc/transforms.c:722-893)
 * ====================================================================== */

int or_calc_transforms(const or_calcspec* spec, const or_obs* obs, const float* S, const double* s, int nnode, const double* lon, const double* lat, int geographic, const unsigned char* visible, float* Tout, float* wout, double* T64, double* w64, int* pnlobs, float* pdfs, float* psrf, or_calcstats* stats, int node0, int nnode_do)
{
    int m = spec->nmem, md = spec->nmem_dynamic;
    int nt = obs->ntypes;
    int dotT = (spec->mode == OR_MODE_ENKF || spec->mode == OR_MODE_HYBRID);
    int cap = obs->nobs;
    int* lobs = malloc((size_t) (cap > 0 ? cap : 1) * sizeof(int));
    double* lco = malloc((size_t) (cap > 0 ? cap : 1) * sizeof(double));
    double* T = malloc((size_t) m * m * sizeof(double));
    double* w = malloc((size_t) m * sizeof(double));
    int node, r = OR_OK;

    if (!lobs || !lco || !T || !w) {
        r = OR_ERR_NOMEM;
        goto done;
    }
    if (stats != NULL)
        memset(stats, 0, sizeof(*stats));
    if (nnode_do < 0)
        nnode_do = nnode - node0;
    for (node = node0; node < node0 + nnode_do; ++node) {
        int p = or_findlocal(obs, lon[node], lat[node], geographic, visible, cap, lobs, lco);
        double *Sloc, *sloc, *G, *pS;
        int e, o, t;

        if (p < 0) {
            r = p;
            goto done;
        }
        if (stats != NULL) {
            if (p > stats->nlobs_max)
                stats->nlobs_max = p;
            stats->nlobs_sum += p;
            stats->ncell++;
        }
        if (p == 0) {
            if (dotT) {
                for (e = 0; e < md; ++e)
                    for (o = 0; o < m; ++o) {
                        if (Tout)
                            Tout[((size_t) node * md + e) * m + o] = (e == o) ? 1.0f : 0.0f;
                        if (T64)
                            T64[((size_t) node * md + e) * m + o] = (e == o) ? 1.0 : 0.0;
                    }
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
        Sloc = malloc((size_t) m * p * sizeof(double));
        pS = malloc((size_t) m * p * sizeof(double));
        sloc = malloc((size_t) p * sizeof(double));
        G = malloc((size_t) m * p * sizeof(double));
        if (!Sloc || !pS || !sloc || !G) {
            free(Sloc);
            free(pS);
            free(sloc);
            free(G);
            r = OR_ERR_NOMEM;
            goto done;
        }
        /* gather and taper, transforms.c:794-802 */
        for (e = 0; e < m; ++e) {
            const float* Se = &S[(size_t) e * obs->nobs];

            for (o = 0; o < p; ++o)
                Sloc[(size_t) e * p + o] = (double) Se[lobs[o]] * lco[o];
        }
        for (o = 0; o < p; ++o)
            sloc[o] = s[lobs[o]] * lco[o];

        if (spec->mode == OR_MODE_ENOI) {
            r = or_calc_G(m, p, Sloc, G);
            if (r == OR_OK)
                or_calc_w(m, p, G, sloc, w);
        } else if (spec->scheme == OR_SCHEME_DENKF)
            r = or_calc_wT_denkf(m, md, p, sloc, Sloc, G, w, T);
        else if (spec->scheme == OR_SCHEME_ETKF)
            r = or_calc_wT_etkf(m, md, p, sloc, Sloc, G, w, T);
        else
            r = OR_ERR_ARG;
        if (r != OR_OK) {
            free(Sloc);
            free(pS);
            free(sloc);
            free(G);
            goto done;
        }
        if (dotT) {
            int e1, e2;

            /* relaxation to prior, T = I + alpha (T - I), transforms.c:814-828 */
            for (e1 = 0; e1 < md; ++e1) {
                double* Ti = &T[(size_t) e1 * m];

                Ti[e1] -= 1.0;
                for (e2 = 0; e2 < m; ++e2)
                    Ti[e2] *= spec->alpha;
                Ti[e1] += 1.0;
            }
            for (e1 = 0; e1 < md; ++e1)
                for (e2 = 0; e2 < m; ++e2) {
                    if (Tout)
                        Tout[((size_t) node * md + e1) * m + e2] = (float) T[(size_t) e1 * m + e2];
                    if (T64)
                        T64[((size_t) node * md + e1) * m + e2] = T[(size_t) e1 * m + e2];
                }
        }
        for (e = 0; e < m; ++e) {
            if (wout)
                wout[(size_t) node * m + e] = (float) w[e];
            if (w64)
                w64[(size_t) node * m + e] = w[e];
        }
        /* diagnostics, transforms.c:843-879 (queue build: per-type re-solve) */
        {
            float dfs = (float) or_traceprod(0, 0, p, m, G, Sloc);
            float srf = (float) (sqrt(or_traceprod(0, 1, m, p, Sloc, Sloc) / dfs) - 1.0);

            if (pnlobs)
                pnlobs[(size_t) nt * nnode + node] = p;
            if (pdfs)
                pdfs[(size_t) nt * nnode + node] = dfs;
            if (psrf)
                psrf[(size_t) nt * nnode + node] = srf;
        }
        for (t = 0; t < nt; ++t) {
            int pt = 0;
            int* plobs = malloc((size_t) p * sizeof(int));

            if (plobs == NULL) {
                r = OR_ERR_NOMEM;
                break;
            }
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
                float pd;
                float ps;

                for (e = 0; e < m; ++e)
                    for (o = 0; o < pt; ++o)
                        pS[(size_t) e * pt + o] = Sloc[(size_t) e * p + plobs[o]];
                r = or_calc_G(m, pt, pS, G);
                if (r != OR_OK) {
                    free(plobs);
                    break;
                }
                pd = (float) or_traceprod(0, 0, pt, m, G, pS);
                if (pd > OR_DFS_MIN)
                    ps = (float) (sqrt(or_traceprod(0, 1, m, pt, pS, pS) / pd) - 1.0);
                else
                    ps = 0.0f;
                if (pdfs)
                    pdfs[(size_t) t * nnode + node] = pd;
                if (psrf)
                    psrf[(size_t) t * nnode + node] = ps;
            }
            free(plobs);
        }
        free(Sloc);
        free(pS);
        free(sloc);
        free(G);
        if (r != OR_OK)
            goto done;
        if (stats != NULL) {
            if (p >= m)
                stats->n_inv_ens++;
            else
                stats->n_inv_obs++;
        }
    }
  done:
    free(lobs);
    free(lco);
    free(T);
    free(w);
    return r;
}

/* ======================================================================
 * update
 * ====================================================================== */

/* j- and i-interpolation of one model row of T and w from the stride grid,
 * update.c:195-272.  Tjj, Tjj1, Tjj2: [ni_s][md*m]; Tj: [mni][md*m]. */
static void interp_i(int ni, int mni, int stride, int periodic_i, int nT, int nw, const float* Tjj, const float* wjj, float* Tj, float* wj)
{
    int ii, i, stepi, e;

    for (ii = 0, i = 0; ii < ni; ++ii) {
        for (stepi = 0; stepi < stride && i < mni; ++stepi, ++i) {
            if (stepi == 0) {
                if (Tj)
                    memcpy(&Tj[(size_t) i * nT], &Tjj[(size_t) ii * nT], (size_t) nT * sizeof(float));
                memcpy(&wj[(size_t) i * nw], &wjj[(size_t) ii * nw], (size_t) nw * sizeof(float));
            } else {
                float weight2 = (float) stepi / (float) stride;
                float weight1 = 1.0f - weight2;
                int ii2 = (ii < ni - 1) ? ii + 1 : ((periodic_i) ? (ii + 1) % ni : ii);

                if (Tj)
                    for (e = 0; e < nT; ++e)
                        Tj[(size_t) i * nT + e] = Tjj[(size_t) ii * nT + e] * weight1 + Tjj[(size_t) ii2 * nT + e] * weight2;
                for (e = 0; e < nw; ++e)
                    wj[(size_t) i * nw + e] = wjj[(size_t) ii * nw + e] * weight1 + wjj[(size_t) ii2 * nw + e] * weight2;
            }
        }
    }
}

/* das_updatefields for structured grids, update.c:92-564 (without the
 * RANDOMISE noise of addnoise(), which is nondeterministic by design). */
int or_update_fields(const or_updatespec* sp, int nfield, const or_field* fields, int j1, int j2, float** fb, float* infl)
{
    int m = sp->nmem, md = sp->nmem_dynamic;
    int mni = sp->mni, stride = sp->stride;
    int ni = sp->ni_s, nj = sp->nj_s;
    int nT = md * m;
    int jsize = j2 - j1 + 1;
    float *Tj = NULL, *wj = NULL, *Tjj = NULL, *Tjj1 = NULL, *Tjj2 = NULL, *wjj = NULL, *wjj1 = NULL, *wjj2 = NULL;
    float *v_f = NULL, *v_a = NULL;
    int jj, j, stepj, i, e, fid, r = OR_OK;

    if ((sp->mode != OR_MODE_ENKF && sp->mode != OR_MODE_HYBRID) || j1 % stride != 0 || j2 < j1)
        return OR_ERR_ARG;
    Tj = malloc((size_t) mni * nT * sizeof(float));
    wj = malloc((size_t) mni * m * sizeof(float));
    Tjj = malloc((size_t) ni * nT * sizeof(float));
    Tjj1 = malloc((size_t) ni * nT * sizeof(float));
    Tjj2 = malloc((size_t) ni * nT * sizeof(float));
    wjj = malloc((size_t) ni * m * sizeof(float));
    wjj1 = malloc((size_t) ni * m * sizeof(float));
    wjj2 = malloc((size_t) ni * m * sizeof(float));
    v_f = malloc((size_t) m * sizeof(float));
    v_a = malloc((size_t) (md > 0 ? md : 1) * sizeof(float));
    if (!Tj || !wj || !Tjj || !Tjj1 || !Tjj2 || !wjj || !wjj1 || !wjj2 || !v_f || !v_a) {
        r = OR_ERR_NOMEM;
        goto done;
    }
    if (stride > 1) {
        memcpy(Tjj2, &sp->T[(size_t) (j1 / stride) * ni * nT], (size_t) ni * nT * sizeof(float));
        memcpy(wjj2, &sp->w[(size_t) (j1 / stride) * ni * m], (size_t) ni * m * sizeof(float));
    }
    for (jj = j1 / stride, j = 0; jj <= j2 / stride && jj < nj; ++jj) {
        for (stepj = 0; stepj < stride && j <= j2 - j1; ++stepj, ++j) {
            if (stride == 1) {
                memcpy(Tj, &sp->T[(size_t) jj * ni * nT], (size_t) ni * nT * sizeof(float));
                memcpy(wj, &sp->w[(size_t) jj * ni * m], (size_t) ni * m * sizeof(float));
            } else {
                if (stepj == 0) {
                    memcpy(Tjj1, Tjj2, (size_t) ni * nT * sizeof(float));
                    memcpy(Tjj, Tjj2, (size_t) ni * nT * sizeof(float));
                    memcpy(wjj, wjj2, (size_t) ni * m * sizeof(float));
                    memcpy(wjj1, wjj2, (size_t) ni * m * sizeof(float));
                    if (jj < nj - 1) {
                        memcpy(Tjj2, &sp->T[(size_t) ((jj + 1) % nj) * ni * nT], (size_t) ni * nT * sizeof(float));
                        memcpy(wjj2, &sp->w[(size_t) ((jj + 1) % nj) * ni * m], (size_t) ni * m * sizeof(float));
                    }
                } else {
                    float weight2 = (float) stepj / (float) stride;
                    float weight1 = 1.0f - weight2;
                    size_t k;

                    for (k = 0; k < (size_t) ni * nT; ++k)
                        Tjj[k] = Tjj1[k] * weight1 + Tjj2[k] * weight2;
                    for (k = 0; k < (size_t) ni * m; ++k)
                        wjj[k] = wjj1[k] * weight1 + wjj2[k] * weight2;
                }
                interp_i(ni, mni, stride, sp->periodic_i, nT, m, Tjj, wjj, Tj, wj);
            }

            for (fid = 0; fid < nfield; ++fid) {
                const or_field* f = &fields[fid];
                float inflation0 = f->inflation;
                double inf_ratio = f->inf_ratio;
                float* vvv = fb[fid];

                for (i = 0; i < mni; ++i) {
                    float inflation = inflation0;
                    double v1_f, v1_a;
                    int nl = sp->nlevels[(size_t) (j1 + j) * mni + i];
                    const float* Ti = &Tj[(size_t) i * nT];
                    const float* wi = &wj[(size_t) i * m];
#define VV(ee) vvv[((size_t) (ee) * jsize + j) * mni + i]

                    if (sp->surfk == 0) {
                        if (nl <= f->level) {
                            for (e = 0; e < m; ++e)
                                VV(e) = 0.0f;
                            continue;
                        }
                    } else {
                        if (nl <= ((f->issurfacevar) ? 0 : sp->surfk - f->level)) {
                            for (e = 0; e < m; ++e)
                                VV(e) = 0.0f;
                            continue;
                        }
                    }
                    for (e = 0; e < m; ++e)
                        if (!isfinite(VV(e)) || fabsf(VV(e)) > (float) OR_MAXOBSVAL)
                            break;
                    if (e < md) {
                        for (e = 0; e < m; ++e)
                            VV(e) = 0.0f;
                        continue;
                    } else if (e < m) {
                        for (e = md; e < m; ++e)
                            VV(e) = 0.0f;
                        if (md == 0)
                            continue;
                    }
                    for (e = 0; e < m; ++e)
                        v_f[e] = VV(e);

                    for (e = 0, v1_f = 0.0; e < md; ++e)
                        v1_f += v_f[e];
                    v1_f /= (double) md;
                    for (e = 0; e < m; ++e)
                        v_f[e] -= v1_f;
                    /* sgemv('T', m, md, 1, Tj[i], m, v_f, 1, 0, v_a, 1) */
                    {
                        int e1, e2;

                        for (e1 = 0; e1 < md; ++e1) {
                            float sum = 0.0f;

                            for (e2 = 0; e2 < m; ++e2)
                                sum += Ti[(size_t) e1 * m + e2] * v_f[e2];
                            v_a[e1] = sum;
                        }
                    }
                    for (e = 0, v1_a = 0.0; e < m; ++e)
                        v1_a += v_f[e] * wi[e];
                    v1_a += v1_f;
                    if (sp->mode == OR_MODE_HYBRID) {
                        double k_d = (md > 1) ? sqrt((double) (m - 1) / (double) (md - 1)) : 1.0;

                        for (e = 0; e < md; ++e)
                            v_a[e] /= k_d;
                        for (e = 0; e < md; ++e)
                            v_f[e] /= k_d;
                    }
                    if (!isnan(inf_ratio)) {
                        double v2_f = 0.0;
                        double v2_a = 0.0;

                        for (e = 0; e < md; ++e) {
                            v2_f += v_f[e] * v_f[e];
                            v2_a += v_a[e] * v_a[e];
                        }
                        if (v2_a > 0.0) {
                            inflation = (float) (sqrt(v2_f / v2_a) * inf_ratio + 1.0 - inf_ratio);
                            if (inflation >= inflation0)
                                inflation = inflation0;
                        }
                    }
                    if (inflation - 1.0f > OR_EPSF)
                        for (e = 0; e < md; ++e)
                            v_a[e] = v_a[e] * inflation;
                    else
                        inflation = 1.0f;
                    for (e = 0; e < md; ++e)
                        v_a[e] += v1_a;
                    for (e = 0; e < md; ++e)
                        v_f[e] += v1_f;
                    if (!sp->outputinc)
                        for (e = 0; e < md; ++e)
                            VV(e) = v_a[e];
                    else if (!f->applylog)
                        for (e = 0; e < md; ++e)
                            VV(e) = v_a[e] - v_f[e];
                    else
                        for (e = 0; e < md; ++e) {
                            if (!isnormal(v_f[e]))
                                VV(e) = 0.0;
                            else
                                VV(e) = exp10(v_a[e]) - exp10(v_f[e]);
                        }
                    if (infl != NULL)
                        infl[((size_t) fid * jsize + j) * mni + i] = inflation;
#undef VV
                }
            }
        }
    }
  done:
    free(Tj);
    free(wj);
    free(Tjj);
    free(Tjj1);
    free(Tjj2);
    free(wjj);
    free(wjj1);
    free(wjj2);
    free(v_f);
    free(v_a);
    return r;
}

/* das_updatebg for structured grids, update.c:568-814 */
int or_update_bg(const or_updatespec* sp, int nfield, const or_field* fields, int j1, int j2, float** fb)
{
    int m = sp->nmem;
    int mni = sp->mni, stride = sp->stride;
    int ni = sp->ni_s, nj = sp->nj_s;
    int jsize = j2 - j1 + 1;
    float *wj = NULL, *wjj = NULL, *wjj1 = NULL, *wjj2 = NULL;
    int jj, j, stepj, i, e, fid, r = OR_OK;

    if (j1 % stride != 0 || j2 < j1)
        return OR_ERR_ARG;
    wj = malloc((size_t) mni * m * sizeof(float));
    wjj = malloc((size_t) ni * m * sizeof(float));
    wjj1 = malloc((size_t) ni * m * sizeof(float));
    wjj2 = malloc((size_t) ni * m * sizeof(float));
    if (!wj || !wjj || !wjj1 || !wjj2) {
        r = OR_ERR_NOMEM;
        goto done;
    }
    if (stride > 1)
        memcpy(wjj2, &sp->w[(size_t) (j1 / stride) * ni * m], (size_t) ni * m * sizeof(float));
    for (jj = j1 / stride, j = 0; jj <= j2 / stride && jj < nj; ++jj) {
        for (stepj = 0; stepj < stride && j <= j2 - j1; ++stepj, ++j) {
            if (stride == 1)
                memcpy(wj, &sp->w[(size_t) jj * ni * m], (size_t) ni * m * sizeof(float));
            else {
                if (stepj == 0) {
                    memcpy(wjj1, wjj2, (size_t) ni * m * sizeof(float));
                    memcpy(wjj, wjj2, (size_t) ni * m * sizeof(float));
                    if (jj < nj - 1)
                        memcpy(wjj2, &sp->w[(size_t) ((jj + 1) % nj) * ni * m], (size_t) ni * m * sizeof(float));
                } else {
                    float weight2 = (float) stepj / stride;
                    float weight1 = (float) 1.0f - weight2;
                    size_t k;

                    for (k = 0; k < (size_t) ni * m; ++k)
                        wjj[k] = wjj1[k] * weight1 + wjj2[k] * weight2;
                }
                interp_i(ni, mni, stride, sp->periodic_i, 0, m, NULL, wjj, NULL, wj);
            }
            for (fid = 0; fid < nfield; ++fid) {
                const or_field* f = &fields[fid];
                float* vvv = fb[fid];

                for (i = 0; i < mni; ++i) {
                    float xmean = 0.0f;
                    int nl = sp->nlevels[(size_t) (j1 + j) * mni + i];
#define VB(ee) vvv[((size_t) (ee) * jsize + j) * mni + i]

                    if (sp->surfk == 0) {
                        if (nl <= f->level) {
                            VB(m) = 0.0f;
                            continue;
                        }
                    } else {
                        if (nl <= ((f->issurfacevar) ? 1 : sp->surfk - f->level)) {
                            VB(m) = 0.0f;
                            continue;
                        }
                    }
                    if (fabsf(VB(m)) > (float) OR_MAXOBSVAL) {
                        VB(m) = 0.0f;
                        continue;
                    }
                    for (e = 0; e < m; ++e)
                        if (fabsf(VB(e)) > (float) OR_MAXOBSVAL)
                            break;
                    if (e < m) {
                        VB(m) = 0.0f;
                        continue;
                    }
                    for (e = 0; e < m; ++e)
                        xmean += VB(e);
                    xmean /= (float) m;
                    for (e = 0; e < m; ++e)
                        VB(m) += (VB(e) - xmean) * wj[(size_t) i * m + e];
#undef VB
                }
            }
        }
    }
  done:
    free(wj);
    free(wjj);
    free(wjj1);
    free(wjj2);
    return r;
}

/* das_sethybridensemble, common/dasystem.c:775-819 */
void or_sethybridensemble(int nmem, int nmem_d, double gamma, int nij, float* v, double* vmean)
{
    int nmem_s = nmem - nmem_d;
    double k_d = (nmem_d > 1) ? sqrt((double) (nmem - 1) / (double) (nmem_d - 1)) : 0.0;
    double k_s = sqrt(gamma * (double) (nmem - 1) / (double) (nmem_s - 1));
    int i, e;

    memset(vmean, 0, (size_t) nij * sizeof(double));
    for (e = 0; e < nmem_d; ++e)
        for (i = 0; i < nij; ++i)
            vmean[i] += (double) v[(size_t) e * nij + i];
    for (i = 0; i < nij; ++i)
        vmean[i] /= (double) nmem_d;
    for (e = 0; e < nmem_d; ++e)
        for (i = 0; i < nij; ++i)
            v[(size_t) e * nij + i] = (v[(size_t) e * nij + i] - vmean[i]) * k_d + vmean[i];
    for (e = nmem_d; e < nmem; ++e)
        for (i = 0; i < nij; ++i)
            v[(size_t) e * nij + i] = v[(size_t) e * nij + i] * k_s + vmean[i];
}

/* ======================================================================
 * RANDOMISE (update/update.c:46-87)
 * ====================================================================== */

/* drand48: X <- (a X + c) mod 2^48 with a = 0x5DEECE66D, c = 0xB (the
 * POSIX drand48 family); the value is X / 2^48 */
double or_drand48(unsigned long long* state)
{
    *state = (0x5DEECE66DULL * *state + 0xBULL) & 0xFFFFFFFFFFFFULL;
    return (double) *state / 281474976710656.0;
}

/* get_normalpair (common/utils.c:1477-1484) */
void or_get_normalpair(unsigned long long* state, double x[2])
{
    double u = sqrt(-2.0 * log(1.0 - or_drand48(state)));
    double v = (M_PI * 2.0) * or_drand48(state);

    x[0] = u * cos(v);
    x[1] = u * sin(v);
}

/* addnoise: (nmem + 1) / 2 normal pairs, centred over the nmem members, then
 * v = v * (float) deflation + (float) (sqrt(1 - deflation^2) sigma) * r_e */
void or_addnoise(int nmem, double deflation, double sigma, unsigned long long* state, long long n, float** v)
{
    double* random = malloc((size_t) ((nmem + 1) / 2 * 2) * sizeof(double));
    double sum = 0.0;
    float mult;
    int e;
    long long i;

    for (e = 0; e < (nmem + 1) / 2; ++e)
        or_get_normalpair(state, &random[e * 2]);
    for (e = 0; e < nmem; ++e)
        sum += random[e];
    sum /= (double) nmem;
    for (e = 0; e < nmem; ++e)
        random[e] -= sum;
    mult = (float) (sqrt(1.0 - deflation * deflation) * sigma);
    for (e = 0; e < nmem; ++e) {
        float* vv = v[e];

        for (i = 0; i < n; ++i)
            vv[i] = vv[i] * (float) deflation + mult * random[e];
    }
    free(random);
}

/* ======================================================================
 * analysed observations: update_HE_structured (calc/ensobs.c:1228-1553)
 * ====================================================================== */

/* T and w of model row j (update.c:195-272 / ensobs.c:1319-1405): Tj
 * [mni][md m], wj [mni][m]; Tjj, wjj scratch [ni_s][.] */
static void interp_row(const or_updatespec* sp, int j, float* Tjj, float* wjj, float* Tj, float* wj)
{
    int m = sp->nmem, md = sp->nmem_dynamic, ni = sp->ni_s, nj = sp->nj_s, stride = sp->stride;
    int nT = md * m;
    int jj = j / stride, stepj = j % stride;
    size_t k;

    if (Tj == NULL)
        nT = 0;
    if (stride == 1) {
        if (Tj)
            memcpy(Tj, &sp->T[(size_t) j * ni * nT], (size_t) ni * nT * sizeof(float));
        memcpy(wj, &sp->w[(size_t) j * ni * m], (size_t) ni * m * sizeof(float));
        return;
    }
    if (stepj == 0) {
        if (Tj)
            memcpy(Tjj, &sp->T[(size_t) jj * ni * nT], (size_t) ni * nT * sizeof(float));
        memcpy(wjj, &sp->w[(size_t) jj * ni * m], (size_t) ni * m * sizeof(float));
    } else {
        float weight2 = (float) stepj / (float) stride;
        float weight1 = 1.0f - weight2;
        int jj2 = (jj < nj - 1) ? jj + 1 : jj;
        const float* w1 = &sp->w[(size_t) jj * ni * m];
        const float* w2 = &sp->w[(size_t) jj2 * ni * m];

        if (Tj) {
            const float* T1 = &sp->T[(size_t) jj * ni * nT];
            const float* T2 = &sp->T[(size_t) jj2 * ni * nT];

            for (k = 0; k < (size_t) ni * nT; ++k)
                Tjj[k] = T1[k] * weight1 + T2[k] * weight2;
        }
        for (k = 0; k < (size_t) ni * m; ++k)
            wjj[k] = w1[k] * weight1 + w2[k] * weight2;
    }
    interp_i(ni, sp->mni, stride, sp->periodic_i, nT, m, Tjj, wjj, Tj, wj);
}

int or_update_HE(const or_updatespec* sp, int nobs, const double* fi, const double* fj, const short* status, const float* inflation0, const double* inf_ratio, float* HE)
{
    int m = sp->nmem, md = sp->nmem_dynamic, mni = sp->mni, mnj = sp->mnj;
    int enoi = (sp->mode == OR_MODE_ENOI);
    int nT = enoi ? 1 : md * m;
    float* Tj = malloc((size_t) mni * nT * sizeof(float));
    float* wj = malloc((size_t) mni * m * sizeof(float));
    float* Tjj = malloc((size_t) sp->ni_s * nT * sizeof(float));
    float* wjj = malloc((size_t) sp->ni_s * m * sizeof(float));
    float* HEo_f = malloc((size_t) m * sizeof(float));
    float* HEo_a = malloc((size_t) m * sizeof(float));
    int* rowof = malloc((size_t) (nobs > 0 ? nobs : 1) * sizeof(int));
    int j, o, e, e2, r = OR_OK;

    if (!Tj || !wj || !Tjj || !wjj || !HEo_f || !HEo_a || !rowof) {
        r = OR_ERR_NOMEM;
        goto done;
    }
    for (o = 0; o < nobs; ++o)
        rowof[o] = (int) (fj[o] + 0.5);
    for (j = 0; j < mnj; ++j) {
        int have = 0;

        for (o = 0; o < nobs && !have; ++o)
            have = (rowof[o] == j && status[o] == OR_STATUS_OK);
        if (!have)
            continue;
        interp_row(sp, j, Tjj, wjj, enoi ? NULL : Tj, wj);
        for (o = 0; o < nobs; ++o) {
            float inflation = NAN;
            double v1_f = 0.0, v1_a = 0.0;
            int i;
            const float* Ti;
            const float* wi;

            if (rowof[o] != j || status[o] != OR_STATUS_OK)
                continue;
            i = (int) (fi[o] + 0.5);
            if (i == mni)
                i--;
            Ti = &Tj[(size_t) i * nT];
            wi = &wj[(size_t) i * m];
            if (enoi) {
                /* update_Hx_structured (ensobs.c:1933-1962): the mean only */
                double Hx = 0.0, dHx = 0.0;

                for (e = 0; e < m; ++e)
                    Hx += HE[(size_t) e * nobs + o];
                Hx /= (double) m;
                for (e = 0; e < m; ++e)
                    dHx += (HE[(size_t) e * nobs + o] - Hx) * wi[e];
                for (e = 0; e < m; ++e)
                    HE[(size_t) e * nobs + o] += dHx;
                continue;
            }
            for (e = 0; e < m; ++e)
                HEo_f[e] = HE[(size_t) e * nobs + o];
            for (e = 0; e < md; ++e)
                v1_f += HEo_f[e];
            v1_f /= (double) md;
            for (e = 0; e < m; ++e)
                HEo_f[e] -= v1_f;
            /* HA^a = HA^f T: sgemv('T', m, md, T) */
            for (e = 0; e < md; ++e) {
                float sum = 0.0f;

                for (e2 = 0; e2 < m; ++e2)
                    sum += Ti[(size_t) e * m + e2] * HEo_f[e2];
                HEo_a[e] = sum;
            }
            for (e = 0; e < m; ++e)
                v1_a += HEo_f[e] * wi[e];
            v1_a += v1_f;
            for (e = 0; e < md; ++e)
                HEo_a[e] += v1_a;
            for (e = 0; e < m; ++e)
                HEo_f[e] += v1_f;
            if (!isnan(inf_ratio[o])) {
                double v2_f = 0.0, v2_a = 0.0, var_a, var_f;

                for (e = 0; e < md; ++e) {
                    double ve = (double) HEo_f[e];

                    v2_f += ve * ve;
                }
                var_f = v2_f / (double) md - v1_f * v1_f;
                for (e = 0; e < md; ++e) {
                    double ve = (double) HEo_a[e];

                    v2_a += ve * ve;
                }
                var_a = v2_a / (double) md - v1_a * v1_a;
                if (var_a > 0.0) {
                    inflation = (float) (sqrt(var_f / var_a) * inf_ratio[o] + 1.0 - inf_ratio[o]);
                    if (inflation >= inflation0[o])
                        inflation = inflation0[o];
                }
            } else
                inflation = inflation0[o];
            if (fabsf(inflation - 1.0f) > 1.0e-6f)
                for (e = 0; e < md; ++e)
                    HEo_a[e] = (HEo_a[e] - (float) v1_a) * inflation + v1_a;
            for (e = 0; e < md; ++e)
                HE[(size_t) e * nobs + o] = HEo_a[e];
            for (e = md; e < m; ++e)
                HE[(size_t) e * nobs + o] += v1_a - v1_f;
        }
    }
  done:
    free(Tj);
    free(wj);
    free(Tjj);
    free(wjj);
    free(HEo_f);
    free(HEo_a);
    free(rowof);
    return r;
}

/* ======================================================================
 * ensemble observations: das_getHE (calc/ensobs.c:46-508) with the standard
 * H functions (calc/model2obs.c:127-365), das_calcinnandspread (541-690)
 * ====================================================================== */

#define OR_EPS_FLOAT 1.0e-7   /* common/utils.c:51 */

/* interpolate2d_structured, common/utils.c:1140-1213 */
float or_interp2d(const double* fij, int ni, int nj, const float* v, const int* mask, int periodic_i)
{
    int i1, i2, j1, j2;
    double wi1, wi2, wj1, wj2;
    double sum, w, ww;
    double fi = fij[0];
    double fj = fij[1];

    if (fi >= (double) ni)
        fi = (double) ni - OR_EPS_FLOAT;
    if (fj >= (double) nj)
        fj = (double) nj - OR_EPS_FLOAT;
    i1 = (int) floor(fi);
    wi1 = ceil(fi) - fi;
    i2 = (int) ceil(fi);
    wi2 = fi - floor(fi);
    j1 = (int) floor(fj);
    wj1 = ceil(fj) - fj;
    j2 = (int) ceil(fj);
    wj2 = fj - floor(fj);
    if (i1 == i2)
        wi1 = 1.0;
    if (j1 == j2)
        wj1 = 1.0;
    if (i1 == -1)
        i1 = (periodic_i) ? ni - 1 : i2;
    if (i2 == ni)
        i2 = (periodic_i) ? 0 : i1;
    if (j1 == -1)
        j1 = j2;
    if (j2 == nj)
        j2 = j1;
    sum = 0.0;
    w = 0.0;
    if (mask[(size_t) j1 * ni + i1]) {
        ww = wj1 * wi1;
        sum += v[(size_t) j1 * ni + i1] * ww;
        w += ww;
    }
    if (mask[(size_t) j1 * ni + i2]) {
        ww = wj1 * wi2;
        sum += v[(size_t) j1 * ni + i2] * ww;
        w += ww;
    }
    if (mask[(size_t) j2 * ni + i1]) {
        ww = wj2 * wi1;
        sum += v[(size_t) j2 * ni + i1] * ww;
        w += ww;
    }
    if (mask[(size_t) j2 * ni + i2]) {
        ww = wj2 * wi2;
        sum += v[(size_t) j2 * ni + i2] * ww;
        w += ww;
    }
    sum /= w;
    return (float) sum;
}

/* interpolate3d_structured, common/utils.c:1252-1367; v[k] is layer k */
float or_interp3d(const double* fij, double fk, int ni, int nj, int nk, int ktop, const float* const* v, const int* nlevels, int periodic_i)
{
    int i1, i2, j1, j2, k1, k2, k1top, k2top;
    double wi1, wi2, wj1, wj2, wk1, wk2;
    double sum, w, ww;
    double fi = fij[0];
    double fj = fij[1];

    if (fi > (double) ni)
        fi = (double) ni - OR_EPS_FLOAT;
    if (fj >= (double) nj)
        fj = (double) nj - OR_EPS_FLOAT;
    i1 = (int) floor(fi);
    wi1 = ceil(fi) - fi;
    i2 = (int) ceil(fi);
    wi2 = fi - floor(fi);
    j1 = (int) floor(fj);
    wj1 = ceil(fj) - fj;
    j2 = (int) ceil(fj);
    wj2 = fj - floor(fj);
    if (fk < 0.0)
        fk = 0.0;
    if (fk > (double) (nk - 1))
        fk = (double) (nk - 1);
    k1 = (int) floor(fk);
    k1top = abs(ktop - k1);
    wk1 = ceil(fk) - fk;
    k2 = (int) ceil(fk);
    k2top = abs(ktop - k2);
    wk2 = fk - floor(fk);
    if (wk1 == 0.0 && wk2 == 0.0)
        wk1 = 1.0;
    if (i1 == i2)
        wi1 = 1.0;
    if (j1 == j2)
        wj1 = 1.0;
    if (k1 == k2)
        wk1 = 1.0;
    if (i1 == -1)
        i1 = (periodic_i) ? ni - 1 : i2;
    if (i2 == ni)
        i2 = (periodic_i) ? 0 : i1;
    if (j1 == -1)
        j1 = j2;
    if (j2 == nj)
        j2 = j1;
    sum = 0.0;
    w = 0.0;
#define OR_C3(J, I, WJ, WI, K, KT, WK) \
    if (nlevels[(size_t) (J) * ni + (I)] > (KT)) { \
        ww = (WJ) * (WI) * (WK); \
        sum += v[K][(size_t) (J) * ni + (I)] * ww; \
        w += ww; \
    }
    OR_C3(j1, i1, wj1, wi1, k1, k1top, wk1)
    OR_C3(j1, i2, wj1, wi2, k1, k1top, wk1)
    OR_C3(j2, i1, wj2, wi1, k1, k1top, wk1)
    OR_C3(j2, i2, wj2, wi2, k1, k1top, wk1)
    OR_C3(j1, i1, wj1, wi1, k2, k2top, wk2)
    OR_C3(j1, i2, wj1, wi2, k2, k2top, wk2)
    OR_C3(j2, i1, wj2, wi1, k2, k2top, wk2)
    OR_C3(j2, i2, wj2, wi2, k2, k2top, wk2)
#undef OR_C3
    sum = sum / w;
    return (float) sum;
}

/* das_getHE (ensobs.c:46-508) for structured grids with the standard H
 * functions (H_surf_standard, H_subsurf_standard; model2obs.c:127-365,
 * footprint 0), the hybrid scaling (ensobs.c:434-486) and
 * das_calcinnandspread (ensobs.c:541-620).  E[f] is field f of the ensemble,
 * [nmem][nj][ni]; obs type t evaluates hkind[t] (0 surface, 1 subsurface)
 * over the fields fields[t][0 .. nk[t]-1] (layer k; the surface field is
 * fields[t][0]).  Only obs with status OK are evaluated (obs_find_bytype);
 * the others keep NaN.  Outputs: HA [nmem][nobs] (das->S after the call),
 * s [nobs] (s_f) and std [nobs] (std_f).  Returns OR_ERR_NONFINITE where the
 * reference calls enkf_quit (a forecast or innovation that is not finite
 * or exceeds MAXOBSVAL). */
int or_getHE(int mode, int nmem, int nmem_dynamic, double gamma, int ni, int nj, int ktop, int periodic_i, const int* nlevels, const float* const* E, int nobs, const short* type, const short* status, const double* fi, const double* fj, const float* fk, const double* value, const int* hkind, const int* nk, const int* const* fields, float* HA, double* s, double* std)
{
    size_t nij = (size_t) ni * nj;
    int e, o, k;
    const float* v[256];

    for (e = 0; e < nmem; ++e)
        for (o = 0; o < nobs; ++o)
            HA[(size_t) e * nobs + o] = NAN;
    for (o = 0; o < nobs; ++o) {
        int t = type[o];
        double fij[2] = { fi[o], fj[o] };

        if (status[o] != OR_STATUS_OK)
            continue;
        for (e = 0; e < nmem; ++e) {
            float x;

            if (hkind[t] == 0)
                x = or_interp2d(fij, ni, nj, E[fields[t][0]] + (size_t) e * nij, nlevels, periodic_i);
            else {
                for (k = 0; k < nk[t]; ++k)
                    v[k] = E[fields[t][k]] + (size_t) e * nij;
                x = or_interp3d(fij, (double) fk[o], ni, nj, nk[t], ktop, v, nlevels, periodic_i);
            }
            if (!isfinite(x) || fabs(x) > OR_MAXOBSVAL)
                return OR_ERR_NONFINITE;
            HA[(size_t) e * nobs + o] = x;
        }
    }
    if (mode == OR_MODE_HYBRID) {
        int md = nmem_dynamic, ms = nmem - nmem_dynamic;
        double k_d = (md > 1) ? sqrt((double) (nmem - 1) / (double) (md - 1)) : 0.0;
        double k_s = sqrt(gamma * (double) (nmem - 1) / (double) (ms - 1));

        for (o = 0; o < nobs; ++o) {
            double mean_d = 0.0, mean_s = 0.0;

            for (e = 0; e < md; ++e)
                mean_d += HA[(size_t) e * nobs + o];
            mean_d /= (double) md;
            for (e = 0; e < md; ++e) {
                float* x = &HA[(size_t) e * nobs + o];

                *x = (*x - mean_d) * k_d + mean_d;
            }
            for (e = md; e < nmem; ++e)
                mean_s += HA[(size_t) e * nobs + o];
            mean_s /= (double) ms;
            for (e = md; e < nmem; ++e) {
                float* x = &HA[(size_t) e * nobs + o];

                *x = (*x - mean_s) * k_s + mean_d;
            }
        }
    }
    for (o = 0; o < nobs; ++o) {
        s[o] = 0.0;
        std[o] = 0.0;
    }
    for (e = 0; e < nmem; ++e)
        for (o = 0; o < nobs; ++o)
            s[o] += (double) HA[(size_t) e * nobs + o];
    for (o = 0; o < nobs; ++o)
        s[o] /= (double) nmem;
    for (e = 0; e < nmem; ++e)
        for (o = 0; o < nobs; ++o)
            HA[(size_t) e * nobs + o] -= (float) s[o];
    for (e = 0; e < nmem; ++e)
        for (o = 0; o < nobs; ++o) {
            float x = HA[(size_t) e * nobs + o];

            std[o] += (double) (x * x);
        }
    for (o = 0; o < nobs; ++o) {
        if (status[o] != OR_STATUS_OK)
            continue;
        std[o] = sqrt(std[o] / (double) (nmem - 1));
        s[o] = value[o] - s[o];
        if (!isfinite(s[o]) || fabs(s[o]) > OR_MAXOBSVAL)
            return OR_ERR_NONFINITE;
    }
    return OR_OK;
}

/* das_writespread (diag/diags.c:79-168), the reduction at 101-123: sums over
 * all nmem members in member order, in double; the square is a product of
 * two floats (float arithmetic, then promoted); population variance clamped
 * at 0; spread 0 where |mean| > MAXOBSVAL (global.h:119); written as float
 * (ncw_put_var_double of the tile file converts). */
void or_spread(int nmem, long long nij, const float* const* v, float* spread)
{
    long long i;
    int e;

    for (i = 0; i < nij; ++i) {
        double v1 = 0.0, v2 = 0.0;

        for (e = 0; e < nmem; ++e) {
            float x = v[e][i];
            float x2 = x * x;

            v1 += x;
            v2 += x2;
        }
        v1 /= (double) nmem;
        v2 = v2 / (double) nmem - v1 * v1;
        v2 = (v2 < 0.0) ? 0.0 : sqrt(v2);
        if (fabs(v1) > OR_MAXOBSVAL)
            v2 = 0.0;
        spread[i] = (float) v2;
    }
}
