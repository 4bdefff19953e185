/*
 * enkf_oracle.h -- CPU restatement of the EnKF-C calc/update hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is never linked into, loaded by, or called from the
 * product library (libenkfg).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every routine cites the reference function it restates (paths relative to
 * the reference tree, enkf-c v2.47.3).  Array layouts follow the reference:
 * "Fortran column-major via C row pointers" matrices from alloc2d()
 * (common/utils.c:829) are passed here as flat arrays with the same
 * addressing, e.g. S (p x m, column-major) has element (o, e) at S[e * p + o].
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of enkf_oracle.c.
 */
#ifndef ENKF_ORACLE_H
#define ENKF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MODE_ENKF 1
#define OR_MODE_ENOI 2
#define OR_MODE_HYBRID 3
#define OR_SCHEME_DENKF 1
#define OR_SCHEME_ETKF 2
#define OR_STATUS_OK 0
#define OR_MAXLOCRAD 8
#define OR_NLOBSMAX_NONE 2147483647

/* errors (the reference calls enkf_quit() instead) */
#define OR_OK 0
#define OR_ERR_ARG -1
#define OR_ERR_LINALG -2
#define OR_ERR_NOMEM -3
#define OR_ERR_CAP -4
#define OR_ERR_NONFINITE -5

/* ---- geometry and taper (common/utils.c:1425-1458, obstypes.c:800-822) */
void or_ll2xyz(const double ll[2], double xyz[3]);
double or_taper_gc(double x);
double or_calclcoeff(int nlocrad, const double* locrad, const double* locweight, double dist);
double or_getmaxlocrad(int nlocrad, const double* locrad);

/* ---- kd-tree (common/kdtree.c:194-274 insertion, 516-574 range search) */
typedef struct {
    int ndim;
    int n;            /* nodes inserted */
    double* coords;   /* [n][ndim] */
    int* left;        /* -1 = none */
    int* right;
    int* parent;
    unsigned char* dir;
    int* data;        /* user datum of each node (insertion index) */
    int depth;        /* maximal depth (root = 1) */
} or_kdtree;

/* Sequential insertion of n points coords[i*ndim .. ] with data = i; points
 * whose first coordinate is not finite are skipped (kd_insertnode). */
int or_kd_build(or_kdtree* t, int ndim, int n, const double* coords);
void or_kd_free(or_kdtree* t);
/* Range search.  Returns the number of hits (or OR_ERR_CAP if > cap); node
 * data and squared distances are written in depth-first visiting order, or
 * sorted by distance (stable) if ordered != 0. */
int or_kd_range(const or_kdtree* t, const double* q, double range, int ordered, int cap, int* data_out, double* d2_out);

/* ---- observations and localisation (common/observations.c) */
typedef struct {
    int nlocrad;
    double locrad[OR_MAXLOCRAD];
    double locweight[OR_MAXLOCRAD]; /* already normalised to sum 1 */
    double rfactor;
    int geographic;
    int statsonly;
    int nlobsmax;                   /* OR_NLOBSMAX_NONE = unlimited */
} or_obstype;

typedef struct {
    int nobs;
    int ntypes;
    const short* type;     /* [nobs] */
    const short* status;   /* [nobs] status at search time */
    const float* lon;      /* [nobs] */
    const float* lat;
    const float* estd;
    const or_obstype* types;
    /* built by or_obs_build_trees (obs_createkdtrees, observations.c:1616) */
    int* ntobs;            /* [ntypes] OK obs of each type at tree build */
    int** obsids;          /* [ntypes][ntobs] */
    or_kdtree* trees;      /* [ntypes] */
} or_obs;

int or_obs_build_trees(or_obs* obs, const short* status_at_build);
void or_obs_free_trees(or_obs* obs);
/* obs_findlocal (observations.c:1837-1927).  visible[t] != 0 if obstype t is
 * visible from the grid's domain (NULL = all).  Returns p or an error. */
int or_findlocal(const or_obs* obs, double lon, double lat, int geographic, const unsigned char* visible, int cap, int* ids, double* lc);

/* ---- standardisation (calc/ensobs.c:849-898) */
void or_standardise(int nmem, int nobs, const short* type, const float* estd, const or_obstype* types, float* S /*[nmem][nobs]*/, double* s /*[nobs] or NULL*/);

/* ---- per-column algebra (calc/calcs.c) */
double or_traceprod(int transposeA, int transposeB, int m, int n, const double* A, const double* B);
int or_calc_G(int m, int p, const double* S, double* G);
void or_calc_w(int m, int p, const double* G, const double* s, double* w);
int or_calc_wT_denkf(int m, int mout, int p, const double* s, const double* S, double* G, double* w, double* T);
int or_calc_wT_etkf(int m, int mout, int p, const double* s, const double* S, double* G, double* w, double* T);

/* ---- the das_calctransforms loop (calc/transforms.c:605-985) */
typedef struct {
    int mode, scheme;
    int nmem, nmem_dynamic;
    double alpha;
} or_calcspec;

typedef struct {
    long long nlobs_sum;
    int nlobs_max;
    int n_inv_obs;
    int n_inv_ens;
    int ncell;
} or_calcstats;

/* nodes: nnode (lon, lat) pairs of the stride sub-grid in [jj][ii] order.
 * Outputs (any may be NULL):
 *   T    float  [nnode][m_d][m]      (transforms.nc "T")
 *   w    float  [nnode][m]           (transforms.nc "w")
 *   T64  double [nnode][m_d][m]      pre-cast values (after relaxation)
 *   w64  double [nnode][m]
 *   pnlobs int  [ntypes+1][nnode]    (slot ntypes = all types)
 *   pdfs, psrf float [ntypes+1][nnode]
 */
int or_calc_transforms(const or_calcspec* spec, const or_obs* obs, const float* S /*[m][nobs] standardised*/, const double* s /*[nobs]*/, int nnode, const double* lon, const double* lat, int geographic, const unsigned char* visible, float* T, float* w, double* T64, double* w64, int* pnlobs, float* pdfs, float* psrf, or_calcstats* stats, int node0, int nnode_do);

/* ---- update (update/update.c:92-564, 568-814; common/dasystem.c:775-819) */
typedef struct {
    int level;
    int issurfacevar;
    float inflation;
    double inf_ratio;   /* NaN = PLAIN */
    int applylog;
} or_field;

typedef struct {
    int mode;           /* ENKF or HYBRID */
    int nmem, nmem_dynamic;
    int mni, mnj;       /* model grid */
    int stride;
    int periodic_i;
    int surfk;
    const int* nlevels; /* [mnj][mni] */
    int outputinc;
    int ni_s, nj_s;     /* transforms grid */
    const float* T;     /* [nj_s][ni_s][m_d][m] */
    const float* w;     /* [nj_s][ni_s][m] */
} or_updatespec;

/* fields: nfield buffers, each float [nmem][jsize][mni] (jsize = j2 - j1 + 1)
 * updated in place for rows j1..j2 (j1 % stride == 0).  infl (optional):
 * [nfield][jsize][mni]. */
int or_update_fields(const or_updatespec* spec, int nfield, const or_field* fields, int j1, int j2, float** fieldbuffer, float* infl);
/* EnOI background: fieldbuffer[f] is [nmem + 1][jsize][mni], member nmem is
 * the background updated in place. */
int or_update_bg(const or_updatespec* spec, int nfield, const or_field* fields, int j1, int j2, float** fieldbuffer);
/* update_HE_structured (calc/ensobs.c:1228-1553): the analysed ensemble
 * observations.  HE [nmem][nobs] holds HE^f on entry and HE^a on exit for
 * the obs with status OK whose nearest row (int)(fj + 0.5) is a row of the
 * grid; fi, fj are the obs' fractional grid indices; inflation0 and
 * inf_ratio are those of the obs type's model variable (NaN ratio = PLAIN).
 * sp->T, sp->w, stride, periodic_i, mni, mnj as for or_update_fields.  In
 * EnOI mode (sp->T unused) it is update_Hx_structured (ensobs.c:1796-1983):
 * every member moves by the interpolated w applied to the anomalies. */
int or_update_HE(const or_updatespec* sp, int nobs, const double* fi, const double* fj, const short* status, const float* inflation0, const double* inf_ratio, float* HE);
/* ---- ensemble observations (calc/ensobs.c:46-690, common/utils.c:1140-1367) */
/* interpolate2d_structured: v [nj][ni], mask [nj][ni] (non-zero = wet) */
float or_interp2d(const double* fij, int ni, int nj, const float* v, const int* mask, int periodic_i);
/* interpolate3d_structured: v[k] the layer k [nj][ni]; ktop the surface layer
 * index (0 or nk - 1); wet where nlevels > |ktop - k| */
float or_interp3d(const double* fij, double fk, int ni, int nj, int nk, int ktop, const float* const* v, const int* nlevels, int periodic_i);
/* das_getHE with the standard H functions, the hybrid scaling and
 * das_calcinnandspread (see enkf_oracle.c) */
int or_getHE(int mode, int nmem, int nmem_dynamic, double gamma, int ni, int nj, int ktop, int periodic_i, const int* nlevels, const float* const* E, int nobs, const short* type, const short* status, const double* fi, const double* fj, const float* fk, const double* value, const int* hkind, const int* nk, const int* const* fields, float* HA, double* s, double* std);
/* RANDOMISE: drand48 on an explicit 48-bit state (POSIX), get_normalpair
 * (common/utils.c:1477-1484) and addnoise (update/update.c:46-87) applied to
 * v[e][0 .. n-1], e < nmem */
double or_drand48(unsigned long long* state);
void or_get_normalpair(unsigned long long* state, double x[2]);
void or_addnoise(int nmem, double deflation, double sigma, unsigned long long* state, long long n, float** v);
/* das_writespread's arithmetic (diag/diags.c:101-123) for one field layer:
 * v is [nmem][nij]; spread [nij] */
void or_spread(int nmem, long long nij, const float* const* v, float* spread);
/* hybrid conditioning of one layer; v is [nmem][nij], vmean scratch [nij] */
void or_sethybridensemble(int nmem, int nmem_dynamic, double gamma, int nij, float* v, double* vmean);

#ifdef __cplusplus
}
#endif
#endif
