/*
 * enkfg.h -- C ABI of the MI355X (gfx950) EnKF-C analysis engine.
 *
 * The engine replaces two seams of enkf-c v2.47.3:
 *   calc:   void das_calctransforms(dasystem* das)      calc/transforms.c:605
 *           (declared common/dasystem.h:143), which in turn calls
 *           obs_findlocal (common/observations.c:1837), kd_findnodeswithinrange
 *           (common/kdtree.c:558) and calc_wT_denkf / calc_wT_etkf / calc_G /
 *           calc_w / traceprod (calc/calcs.h:18-22);
 *   update: static void das_updatefields(dasystem*, int nfield,
 *           void** fieldbuffer, field fields[])          update/update.c:92
 *           and das_updatebg (update/update.c:568), das_sethybridensemble
 *           (common/dasystem.c:775).
 *
 * Conventions
 *   - Every call returns an int status, ENKFG_OK (0) on success; the engine
 *     never aborts.  enkfg_last_error() describes the last failure.  Where the
 *     reference calls enkf_quit() (LAPACK info != 0 at calc/calcs.c:110-118,
 *     339-343; stride mismatch at update/update.c:152-158) the engine returns
 *     ENKFG_ERR_LINALG / ENKFG_ERR_ARG instead.
 *   - The caller owns host buffers; the engine owns its device buffers, except
 *     buffers passed through an *_attach call, which stay caller-owned.
 *   - One host thread per context; each context has one HIP stream.  All
 *     calls are synchronous unless named *_async.
 *   - Array layouts follow the reference: S is [nmem][nobs] (das->S,
 *     common/dasystem.h:95), T is [node][nmem_dynamic][nmem] and w is
 *     [node][nmem] (the "T" and "w" variables of transforms.nc,
 *     calc/transforms.c:54-94), field buffers are [nmem][jsize][mni]
 *     (update/update.c:1764-1797).
 */
#ifndef ENKFG_H
#define ENKFG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define ENKFG_OK 0
#define ENKFG_ERR_ARG 1          /* invalid argument or inconsistent sizes */
#define ENKFG_ERR_HIP 2          /* HIP runtime error */
#define ENKFG_ERR_NOMEM 3        /* device or host allocation failed */
#define ENKFG_ERR_LINALG 4       /* factorisation failed (non-SPD matrix) */
#define ENKFG_ERR_STATE 5        /* call out of order (e.g. no obs set) */
#define ENKFG_ERR_UNSUPPORTED 6  /* valid input this engine does not handle */
#define ENKFG_ERR_NONFINITE 7    /* non-finite innovation (calc/ensobs.c:609) */

/* ---- modes and schemes: values of common/global.h:64-73 */
#define ENKFG_MODE_ENKF 1
#define ENKFG_MODE_ENOI 2
#define ENKFG_MODE_HYBRID 3
#define ENKFG_SCHEME_DENKF 1
#define ENKFG_SCHEME_ETKF 2

/* ---- arithmetic of the calc and update kernels */
#define ENKFG_PREC_FP64 0        /* FP64 algebra on FP32 S, FP32/FP64 update */
#define ENKFG_PREC_MIXED 1       /* S^T S as BF16 hi/lo MFMA products with FP32
                                    accumulation; factorisation in FP64 */

/* ---- config flags */
#define ENKFG_FLAG_PERTYPE_DIAG 1  /* per-obstype dfs/srf as the default
                                      USE_MPIQUEUE build computes them
                                      (calc/transforms.c:853-879) */
#define ENKFG_FLAG_KEEP_FP64 2     /* keep pre-cast FP64 T and w */
#define ENKFG_FLAG_GRAPH 4         /* run the calc's Gram/factor batches as
                                      one captured HIP graph, replayed while
                                      the batches and buffers are unchanged
                                      (timings then give the Gram and factor
                                      phases as one, in [1]) */

#define ENKFG_NLOBSMAX_NONE 2147483647
#define ENKFG_MAXLOCRAD 8
#define ENKFG_STATUS_OK 0

typedef struct enkfg_ctx enkfg_ctx;

/* dasystem fields used on the path (common/dasystem.h:48-127) */
typedef struct {
    int mode;            /* ENKFG_MODE_* (das->mode) */
    int scheme;          /* ENKFG_SCHEME_* (das->scheme) */
    int precision;       /* ENKFG_PREC_* */
    int nmem;            /* das->nmem */
    int nmem_dynamic;    /* das->nmem_dynamic (== nmem unless hybrid) */
    double alpha;        /* relaxation to prior (das->alpha, ALPHA) */
    double gamma;        /* hybrid static weight (das->gamma, GAMMA) */
    int device;          /* HIP device ordinal */
    int flags;           /* ENKFG_FLAG_* */
} enkfg_config;

/* obstype fields used on the path (common/obstypes.h:29-118) */
typedef struct {
    int nlocrad;
    double locrad[ENKFG_MAXLOCRAD];     /* km */
    double locweight[ENKFG_MAXLOCRAD];  /* normalised to sum 1 (obstypes.c:533-541) */
    double rfactor;                     /* already multiplied by RFACTOR base */
    int geographic;                     /* 3D chord coordinates if non-zero */
    int statsonly;
    int nlobsmax;                       /* ENKFG_NLOBSMAX_NONE = no limit */
} enkfg_obstype;

/* the stride sub-grid of one model grid (calc/transforms.c:665-684) */
typedef struct {
    int ni, nj;          /* ni_s = (mni-1)/stride+1, nj_s likewise */
    int stride;
    int geographic;      /* grid_isgeographic() */
    const double* lon;   /* [nj][ni] node longitudes (grid_ij2xy), NaN = global */
    const double* lat;   /* [nj][ni] node latitudes */
    const unsigned char* visible;  /* [ntypes] 1 if the obstype is visible from
                                      this grid's domain, NULL = all */
} enkfg_nodes;

/* calcstats, calc/transforms.c:44-50 */
typedef struct {
    long long nlobs_sum;
    int nlobs_max;
    int n_inv_obs;
    int n_inv_ens;
    int ncell;
} enkfg_stats;

/* model grid as the update sees it (update/update.c:96-104) */
typedef struct {
    int mni, mnj;        /* model grid size (grid_getsize) */
    int stride;          /* grid_getstride(); must match the transforms */
    int periodic_i;      /* grid_isperiodic_i() */
    int surfk;           /* grid_getsurflayerid() */
    const int* nlevels;  /* [mnj][mni] grid_getnumlevels() */
} enkfg_grid;

/* one horizontal field (common/dasystem.h:36-45 plus model variable data) */
typedef struct {
    int level;           /* field level */
    int issurfacevar;
    float inflation;     /* model_getvarinflation(): inflation0 */
    double inf_ratio;    /* NaN = PLAIN inflation */
    int applylog;        /* model_getvarislog() */
    double deflation;    /* RANDOMISE <deflation> <sigma> (model.c:208-218,
                            model_getvardeflation); NaN = not randomised */
    double sigma;        /* model_getvarsigma() */
} enkfg_field;

/* update flags (common/global.h:92-104) */
#define ENKFG_UPD_OUTPUTINC 32   /* UPDATE_OUTPUTINC */
#define ENKFG_UPD_HYBRID_CONDITION 1024  /* apply das_sethybridensemble first */

/* ---------------------------------------------------------------- context */
int enkfg_create(enkfg_ctx** ctx, const enkfg_config* cfg);
void enkfg_destroy(enkfg_ctx* ctx);
const char* enkfg_last_error(const enkfg_ctx* ctx);
int enkfg_sync(enkfg_ctx* ctx);

/* ---------------------------------------------------------------- calc inputs */
/* obstypes_set() output; replaces the das->obs->obstypes array. */
int enkfg_set_obstypes(enkfg_ctx* ctx, int ntypes, const enkfg_obstype* types);

/* obs_read() output plus obs_createkdtrees() (observations.c:1616): per
 * non-stats-only obstype, obs with status OK are inserted, in ascending index
 * order, into a kd-tree over 3D chord coordinates (geographic types) or
 * (lon, lat).  lon/lat are the float obs coordinates of observations.h:67-68. */
int enkfg_set_obs(enkfg_ctx* ctx, int nobs, const short* type, const short* status, const float* lon, const float* lat, const float* estd);

/* H operators of the obs types (calc/allhs.c:30-44) that enkfg_calc_HE
 * evaluates on the device */
#define ENKFG_H_SURF 0       /* H_surf_standard (model2obs.c:127-181): 2D
                                interpolation of the variable's surface layer */
#define ENKFG_H_SUBSURF 1    /* H_subsurf_standard (model2obs.c:256-355): 3D
                                interpolation between the layers */
typedef struct {
    int hfunction;           /* ENKFG_H_* */
    int nk;                  /* layers of the obs type's model variable (1: 2D) */
    const int* field;        /* [nk] index, in the attached ensemble, of the
                                field holding layer k of the variable */
} enkfg_hop;

/* Search-time obs status (observations.c:1906): obs marked bad after the trees
 * were built (e.g. by obs_markbadbatches) are skipped by the local search. */
int enkfg_set_obs_status(enkfg_ctx* ctx, const short* status);

/* das->S ([nmem][nobs], float) and das->s_f ([nobs], double).  With
 * standardised == 0 they are HA^f and the innovations, and the engine applies
 * das_standardise (calc/ensobs.c:849-898) on the device. */
int enkfg_set_S(enkfg_ctx* ctx, const float* S, const double* s, int standardised);

/* das_getHE (calc/ensobs.c:46-508) from the device-resident ensemble
 * (enkfg_ens_attach), with the hybrid scaling (ensobs.c:434-486) and
 * das_calcinnandspread (ensobs.c:541-620): for every obs of the last
 * enkfg_set_obs with status OK, HE = H(E) for the H operator of its type
 * (structured grids, footprint 0; fi, fj, fk are fij[0], fij[1], fk of the
 * observation record), then the innovations y - mean(HE) and the anomalies.
 * The result becomes the engine's S and s exactly as enkfg_set_S(HA, s, 0)
 * would make them, so S never leaves the device on its way to the calc.
 * grid is the attached ensemble's (nlevels for the land mask, surfk = the
 * surface layer id, 0 or nk - 1).  Optional host copies: HA [nmem][nobs]
 * (das->S in S_MODE_HA_f), s [nobs] (das->s_f) and std [nobs] (das->std_f),
 * with the reference's values for obs that are not OK (NaN).  A forecast or
 * innovation that is not finite or exceeds MAXOBSVAL returns
 * ENKFG_ERR_NONFINITE where the reference calls enkf_quit.  EnOI (the
 * background forecast) is not supported. */
int enkfg_calc_HE(enkfg_ctx* ctx, const enkfg_grid* grid, int ntypes, const enkfg_hop* hops, const double* fi, const double* fj, const float* fk, const double* value, float* HA, double* s, double* std);

/* The standardised S and s_f (the reference standardises in place; callers
 * such as das_updateHE read them back). */
int enkfg_get_S(enkfg_ctx* ctx, float* S, double* s);
/* das_moderateobs (calc/ensobs.c:749-799; KFACTOR, enkfprm.c): with a
 * finite K, enkfg_calc_HE moderates the error of every good observation from
 * its forecast spread and innovation before S is standardised.  NaN (the
 * default) turns it off. */
int enkfg_set_kfactor(enkfg_ctx* ctx, double kfactor);
/* the engine's obs errors (moderated ones after a moderating enkfg_calc_HE)
 * [nobs], and per obs type the count of obs whose error variance more than
 * doubled (obstype nmodified) [ntypes]; either may be NULL */
int enkfg_get_obs_estd(enkfg_ctx* ctx, float* estd, int* nmodified);


/* ---------------------------------------------------------------- calc */
/* das_calctransforms for one grid: local search, gather and taper, per-column
 * transform, relaxation to prior, FP32 cast, diagnostics.  Results stay on
 * the device until fetched with enkfg_get_*. */
int enkfg_calc_transforms(enkfg_ctx* ctx, const enkfg_nodes* nodes);

/* T [node][nmem_dynamic][nmem] float, w [node][nmem] float, and optional
 * pre-cast FP64 copies (need ENKFG_FLAG_KEEP_FP64).  Any pointer may be NULL. */
int enkfg_get_transforms(enkfg_ctx* ctx, float* T, float* w, double* T64, double* w64);

/* enkf_diag.nc contents: arrays [ntypes + 1][node]; slot ntypes holds the
 * totals (nlobs, dfs, srf), slots 0..ntypes-1 the per-type values
 * (calc/transforms.c:843-879).  Any pointer may be NULL. */
int enkfg_get_diag(enkfg_ctx* ctx, int* pnlobs, float* pdfs, float* psrf);

int enkfg_get_stats(enkfg_ctx* ctx, enkfg_stats* stats);

/* Localisation only (obs_findlocal for every node); lists are kept on the
 * device.  counts: [node]; a node's list is read with enkfg_get_localobs. */
int enkfg_localise(enkfg_ctx* ctx, const enkfg_nodes* nodes);
int enkfg_get_local_counts(enkfg_ctx* ctx, int* counts);
int enkfg_get_localobs(enkfg_ctx* ctx, int node, int cap, int* ids, double* lc);

/* ---------------------------------------------------------------- update */
/* The transforms.nc contents for the update (T may be NULL in EnOI mode).
 * Not needed when enkfg_calc_transforms ran in the same context. */
int enkfg_set_transforms(enkfg_ctx* ctx, int nj_s, int ni_s, int stride, const float* T, const float* w);

/* das_updatefields for rows j1..j2 of a structured grid (j1 % stride == 0):
 * fieldbuffer[f] is float [nmem][j2-j1+1][mni], updated in place; infl
 * (optional) receives [nfield][j2-j1+1][mni] inflation factors. */
int enkfg_update_fields(enkfg_ctx* ctx, const enkfg_grid* grid, int nfield, const enkfg_field* fields, int j1, int j2, float** fieldbuffer, float* infl, int flags);

/* RANDOMISE (addnoise, update/update.c:46-87, applied at 554-563): after the
 * update of a field batch, every field with a finite deflation d gets
 *   v[e] <- v[e] * d + sqrt(1 - d^2) sigma (r_e - mean_e r)
 * over all nmem members and all cells of the batch's rows, with r_e from
 * (nmem + 1) / 2 normal pairs of drand48 (get_normalpair, common/utils.c:
 * 1477-1484), drawn field by field in batch order.  The engine keeps its own
 * drand48 stream per context: 48-bit state X, X <- (0x5DEECE66D X + 0xB) mod
 * 2^48, drand48() = X / 2^48.  The initial state is 0, the state of glibc's
 * drand48 when the process never seeds it -- the reference seeds srand()
 * (common/utils.c:102-122), not srand48(), so its stream starts there too. */
int enkfg_set_rand48(enkfg_ctx* ctx, unsigned long long state);
int enkfg_get_rand48(const enkfg_ctx* ctx, unsigned long long* state);
/* The engine's generator on the host, without a context or a device: npair
 * normal pairs (get_normalpair) into out[2 * npair], advancing *state. */
int enkfg_rand48_normals(unsigned long long* state, int npair, double* out);

/* enkfg_update_fields for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and errors as
 * enkfg_update_fields, but it calls no hipMalloc / hipFree: its staging
 * lives in a fixed 64 MiB device array per GPU, held until the process ends,
 * and the fields are processed one at a time through it.  A batch needs
 * about 8 (m + 1) jsize mni bytes (jsize = j2 - j1 + 1); larger batches are
 * refused with ENKFG_ERR_ARG (split the rows).  The transforms (from the
 * calc or enkfg_set_transforms) must already be on the device.  Every call
 * on one GPU shares that array: calls that overlap in time on one GPU (from
 * several threads or contexts) race. */
int enkfg_update_fields_noalloc(enkfg_ctx* ctx, const enkfg_grid* grid, int nfield, const enkfg_field* fields, int j1, int j2, float** fieldbuffer, float* infl, int flags);

/* The analysed ensemble observations of a structured grid: update_HE_structured
 * (calc/ensobs.c:1228-1553; update_Hx_structured, 1796-1983, in EnOI mode),
 * the step of das_updateHE (ensobs.c:2116-2129) between sorting the obs and
 * das_calcinnandspread.  HE is das->S in S_MODE_HE_f, [nmem][nobs], updated
 * in place for the obs with status OK whose nearest cell
 * ((int)(fi + 0.5), (int)(fj + 0.5); a column rounding to mni is the last
 * one) lies in the grid; fi, fj are the obs' fractional indices (fij[0],
 * fij[1]); inflation and inf_ratio per obs are model_getvarinflation() of
 * the obs type's variable (NaN ratio = PLAIN).  Uses the transforms of the
 * last calc or enkfg_set_transforms; grid->nlevels is not used. */
int enkfg_update_HE(enkfg_ctx* ctx, const enkfg_grid* grid, int nobs, const double* fi, const double* fj, const short* status, const float* inflation, const double* inf_ratio, float* HE);

/* das_updatebg (EnOI): fieldbuffer[f] is [nmem + 1][jsize][mni]; member nmem
 * is the background, updated in place. */
int enkfg_update_bg(enkfg_ctx* ctx, const enkfg_grid* grid, int nfield, const enkfg_field* fields, int j1, int j2, float** fieldbuffer);

/* Device-resident ensemble: a caller-owned device buffer laid out
 * [field][cell][member] (cell = j * mni + i, members contiguous) holding all
 * nfield fields of the whole grid.  The update then runs in place with no
 * host traffic. */
int enkfg_ens_attach(enkfg_ctx* ctx, const enkfg_grid* grid, int nfield, const enkfg_field* fields, float* dev_E);
int enkfg_ens_update(enkfg_ctx* ctx, int flags, float* dev_infl);

/* ---------------------------------------------------------------- diagnostics */
/* das_writespread (diag/diags.c:79-168): the spread of each field over all
 * nmem members, sqrt(max(0, sum v^2 / nmem - mean^2)) with the sums in
 * double in member order and v^2 a float product, 0 where |mean| >
 * MAXOBSVAL; the values spread.nc holds.  The reference writes the forecast
 * spread: call before the update.
 *   enkfg_ens_spread: fields f0 .. f0 + nfield - 1 of the attached ensemble;
 *     spread [nfield][mnj][mni], host or device memory.
 *   enkfg_spread_fields: host field buffers as the update takes them,
 *     fieldbuffer[f] [nmem][nij] (nij = jsize * mni); spread [nfield][nij]. */
int enkfg_ens_spread(enkfg_ctx* ctx, int f0, int nfield, float* spread);
int enkfg_spread_fields(enkfg_ctx* ctx, int nfield, int nij, float* const* fieldbuffer, float* spread);

/* ---------------------------------------------------------------- adjoint */
/* Adjoint of the analysis mean.  No counterpart in the reference; it
 * differentiates the reference's own forward path: w = M^-1 S^T s per node
 * (calc/calcs.c:189-263 on the standardised S, s of calc/ensobs.c:849-898),
 * the stride-grid interpolation of w (update/update.c:195-272) and
 * x^a = x^f + A^f w (update.c:381-410).
 *   dev_dJ_dmean  device float [nfield][mnj * mni]: dJ/dx^a of a scalar J for
 *                 the fields of the attached ensemble (zero where J does not
 *                 depend on the cell)
 *   dJ_dy, dJ_dR  host double [nobs]: gradients with respect to every
 *                 observation value and observation-error variance
 *                 (R = estd^2), in observation order; either may be NULL
 *   dJ_dinfl      host double: gradient with respect to the inflation, which
 *                 scales the analysis anomalies only and so is 0 for J of
 *                 the mean
 * Needs enkfg_calc_transforms for the grid and the forecast ensemble attached
 * (enkfg_ens_attach, before enkfg_ens_update); recomputes I + S^T S and its
 * inverse in FP64 for the nodes that J depends on (none is stored); nmem <= 96. */
int enkfg_adjoint(enkfg_ctx* ctx, const float* dev_dJ_dmean, double* dJ_dy, double* dJ_dR, double* dJ_dinfl);

/* Adjoint of the analysis ensemble.  No counterpart in the reference; it
 * differentiates the DEnKF forward path of the reference: M = I + S^T S and
 * w = M^-1 S^T s per node (calc/calcs.c:189-263), T = 1/2 (I + M^-1) with the
 * relaxation to prior of transforms.c:1270-1283, the stride-grid
 * interpolation of T and w (update/update.c:195-272), the analysis
 * v_a = A^f T^T, its hybrid rescaling and the capped inflation, and
 * X^a = infl v_a + x^f + A^f w (update.c:381-458).
 *   dev_dJ_dXa  device float [nfield][mnj * mni][nmem], the attached
 *               ensemble's layout: dJ/dX^a of a scalar J for the dynamic
 *               members (the static members of a hybrid ensemble are not
 *               updated, and their entries are ignored)
 *   dJ_dy, dJ_dR  host double [nobs], as enkfg_adjoint; either may be NULL
 *   dJ_dinfl    host double [nfield]: dJ/d(inflation) of each attached field
 *               (the fields' inflation values; cells where the ratio form is
 *               below the cap, or no inflation is applied, add nothing)
 * The rounding of the forward's FP32 arithmetic (interpolation, products,
 * the inflation factor) is not differentiated.  Same preconditions as
 * enkfg_adjoint; DEnKF only (ENKFG_ERR_UNSUPPORTED for ETKF), nmem <= 96.
 * Memory: nnode * nmem_dynamic * nmem doubles for dJ/dT. */
int enkfg_adjoint_ens(enkfg_ctx* ctx, const float* dev_dJ_dXa, double* dJ_dy, double* dJ_dR, double* dJ_dinfl);

/* hipEvent-measured duration (ms) of the last calc / update kernels, per
 * phase: [0] localisation, [1] Gram, [2] factorisation, [3] update; [4] is
 * the number of Gram (and factor) launches of the last calc and [5] the
 * device time of the whole calc (the factor kernels of one batch run
 * concurrently with the Gram kernel of the next, so [1] + [2] may exceed
 * the time they take together); [6] is the kernel time of the last
 * enkfg_calc_HE and [7] of the last spread (enkfg_ens_spread,
 * enkfg_spread_fields). */
int enkfg_get_timings(enkfg_ctx* ctx, double* ms, int n);

/* Calc batching and the autotuner's choices (no counterpart in the
 * reference, whose calc hands out columns through mpiqueue,
 * transforms.c:700-716).  enkfg_calc_transforms orders the stride nodes by
 * obs-count bucket (the largest (node, obstype) list: [0, 256), [256, 1024),
 * [1024, 4096), [4096, ...)), heaviest bucket first, and runs each bucket's
 * Gram matrices with the tile groups per list (wavefronts: 2 or 3 for
 * nmem <= 96, 3 or 4 up to 112, 4 or 5 up to 144, 5 or 6 up to 160) that the
 * first sufficiently large batch of the bucket measured fastest.  Writes up
 * to n ints: out[0] = number of buckets B, then per bucket the lower bound,
 * the nodes in it at the last calc, and the tile groups chosen for FP64 and
 * for mixed precision (0 = not tuned yet). */
int enkfg_get_tuning(enkfg_ctx* ctx, int* out, int n);

#ifdef __cplusplus
}
#endif
#endif
