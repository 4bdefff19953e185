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

#ifdef __cplusplus
}
#endif
#endif
