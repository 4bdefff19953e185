/*
 * enkfio.h -- file formats of the EnKF-C calc/update path, host side
 * (libenkfio.so; no GPU code).
 *
 * 1. A reader and writer of netCDF classic files (the 64-bit-offset format,
 *    "CDF\2"; the reader also takes "CDF\1"), enough for the files the path
 *    reads and writes: fixed-size dimensions, variables of NC_BYTE ..
 *    NC_DOUBLE, global and variable attributes.  The reference opens its
 *    files through the netCDF library, which reads this format whatever
 *    NCFORMAT it writes with (common/ncw.c), so the files written here are
 *    inputs the reference's enkf_update and tools accept.
 * 2. The path's own files, with the reference's names and layouts:
 *      transforms.nc  T[j][i][m_dyn][m], w[j][i][m], global "stride",
 *                     "grid_name" (calc/transforms.c:54-135)
 *      enkf_diag.nc   nlobs, dfs, srf [j][i]; pnlobs, pdfs, psrf
 *                     [nobstypes][j][i]; global "stride" and one int
 *                     attribute per obs type name (transforms.c:485-520)
 *      spread.nc / inflation.nc  one float [j][i] variable per field
 *                     (diag/diags.c:79-168, 336-390)
 *    and the reader of transforms.nc that enkf_update needs.
 *
 * Every call returns 0 on success or a negative ENKFIO_E* code; files are
 * big-endian on disk as the format requires.
 */
#ifndef ENKFIO_H
#define ENKFIO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ENKFIO_OK 0
#define ENKFIO_EIO -1        /* open, read or write failed */
#define ENKFIO_EFORMAT -2    /* not a netCDF classic file, or a corrupt header */
#define ENKFIO_EARG -3       /* invalid argument, unknown name, wrong mode */
#define ENKFIO_ENOMEM -4
#define ENKFIO_ERANGE -5     /* slab outside the variable, or a size the format cannot hold */

/* netCDF external types */
#define ENKFIO_BYTE 1
#define ENKFIO_CHAR 2
#define ENKFIO_SHORT 3
#define ENKFIO_INT 4
#define ENKFIO_FLOAT 5
#define ENKFIO_DOUBLE 6

#define ENKFIO_GLOBAL -1     /* variable id of the global attributes */
#define ENKFIO_MAXDIMS 8

typedef struct ncfile ncfile;

/* ---- writing: create, define, end the definitions, write, close */
int ncf_create(const char* path, ncfile** out);
int ncf_def_dim(ncfile* f, const char* name, size_t len);             /* returns the dimension id */
int ncf_def_var(ncfile* f, const char* name, int type, int ndims, const int* dimids);   /* variable id */
int ncf_put_att(ncfile* f, int varid, const char* name, int type, size_t n, const void* values);
int ncf_put_att_text(ncfile* f, int varid, const char* name, const char* text);
int ncf_enddef(ncfile* f);
/* values of a variable along its first dimension: records start .. start +
 * count - 1 (the whole variable: start 0, count = first dimension, or 1 for
 * a scalar); data are native-endian of the variable's type */
int ncf_put_slab(ncfile* f, int varid, size_t start, size_t count, const void* data);

/* ---- reading */
int ncf_open(const char* path, ncfile** out);
int ncf_inq(const ncfile* f, int* ndims, int* nvars, int* ngatts);
int ncf_inq_varid(const ncfile* f, const char* name);                  /* variable id, or ENKFIO_EARG */
int ncf_inq_varname(const ncfile* f, int varid, char* name, size_t cap);   /* NUL-terminated, truncated to cap */
int ncf_inq_var(const ncfile* f, int varid, int* type, int* ndims, size_t* dimlens);
int ncf_inq_dimlen(const ncfile* f, const char* name, size_t* len);
/* attribute of varid (ENKFIO_GLOBAL) by name: type and count, and, when
 * values != NULL, up to cap values (text: cap bytes) */
int ncf_get_att(const ncfile* f, int varid, const char* name, int* type, size_t* n, void* values, size_t cap);
int ncf_get_slab(ncfile* f, int varid, size_t start, size_t count, void* data);
/* a hyperslab start[d] .. start[d] + count[d] - 1 of every dimension, in the
 * variable's type (native-endian); record variables are read record by record */
int ncf_get_vara(ncfile* f, int varid, const size_t* start, const size_t* count, void* data);
/* the same converted to `type` (ENKFIO_BYTE .. ENKFIO_DOUBLE, not CHAR), as C
 * casts; a double beyond the float range reads as NaN (common/ncw.c:1078);
 * start == NULL reads the whole variable */
int ncf_get_vara_as(ncfile* f, int varid, const size_t* start, const size_t* count, int type, void* out);
int ncf_inq_natts(const ncfile* f, int varid, int* natts);
int ncf_inq_attname(const ncfile* f, int varid, int k, char* name, size_t cap);

int ncf_close(ncfile* f);

/* ---- the path's files */
/* transforms.nc: T (NULL in EnOI mode) [nj][ni][m_dyn][m] and w [nj][ni][m] */
int enkfio_write_transforms(const char* path, int nj, int ni, int m_dyn, int m, int stride, const char* grid_name, const float* T, const float* w);
/* dimensions and stride of a transforms file (any pointer may be NULL);
 * m_dyn is 0 when the file has no T (EnOI) */
int enkfio_inq_transforms(const char* path, int* nj, int* ni, int* m_dyn, int* m, int* stride);
/* T (may be NULL) and w, rows j0 .. j0 + nrow - 1 */
int enkfio_read_transforms(const char* path, int j0, int nrow, float* T, float* w);
/* enkf_diag.nc: pnlobs, pdfs, psrf [nobstypes + 1][nj][ni] as the engine's
 * enkfg_get_diag returns them (slot nobstypes = totals) */
int enkfio_write_diag(const char* path, int nobstypes, const char* const* typenames, int nj, int ni, int stride, const int* pnlobs, const float* pdfs, const float* psrf);
/* spread.nc / inflation.nc: nfield float [nj][ni] variables named names[f] */
int enkfio_write_fields2d(const char* path, int nfield, const char* const* names, int nj, int ni, const float* v);

/* ---- model fields (ensemble members, backgrounds), structured grids */
/* ncu_getnlevels (common/ncutils.c:54-91): layers of a [(t,)k,]j,i variable */
int enkfio_field_nlevels(const char* path, const char* varname, int* nlevels);
/* ncu_readfield_part (common/ncutils.c:99-531): rows j1 .. j2 (j1 <= 0 and
 * j2 < 0: all) of layer k, record r (-1: the only one), of a [(t,)k,]j,i or
 * j,i variable on an nj x ni grid of nk layers, as float [j2-j1+1][ni];
 * values equal to _FillValue (or the type's default fill), missing_value, or
 * outside valid_min / valid_max / valid_range become NaN; then scale_factor
 * and add_offset apply, as floats.  Dimensions that do not fit the grid
 * return ENKFIO_EARG (the reference quits). */
int enkfio_read_field(const char* path, const char* varname, int r, int k, int j1, int j2, int ni, int nj, int nk, float* v);

/* ---- observations.nc (common/observations.c:452-802) */
#define ENKFIO_MAXTYPES 64
#define ENKFIO_NAMELEN 64
typedef struct {
    long long nobs;
    double da_day;                      /* global DA_DAY; NaN when absent */
    int ntypes;                         /* the obs type names: the int attributes of "type" */
    char typenames[ENKFIO_MAXTYPES][ENKFIO_NAMELEN];
    int *id, *id_orig, *batch, *aux;
    short *status, *type, *product, *instrument, *fid;
    float *value, *estd;                /* estd_orig when present, as obs_read */
    float *footprint;                   /* NULL when absent */
    float *lon, *lat, *depth, *model_depth, *fk, *time;
    double *fi, *fj;                    /* fi, fj; or fi0, fi1 with fi2 (unstructured grids) */
    double *fi2;                        /* NULL for structured grids */
} enkfio_obs;
/* reads every record of the file; STATUS_BADBATCH becomes STATUS_OK as in
 * obs_read.  The times are not clipped to the obs types' windows (that needs
 * the obs type parameters).  Free with enkfio_free_obs. */
int enkfio_read_obs(const char* path, enkfio_obs* obs);
void enkfio_free_obs(enkfio_obs* obs);

#ifdef __cplusplus
}
#endif
#endif
