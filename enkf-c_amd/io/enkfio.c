/*
 * enkfio.c -- netCDF classic files and the EnKF-C path's files (host side).
 *
 * The classic format (the netCDF "File Format Specification", versions 1
 * and 2): a header
 *     "CDF" version numrecs dim_list gatt_list var_list
 * with every integer big-endian, names and values padded to 4 bytes, and an
 * ABSENT list written as two zero words; then each fixed-size variable's
 * values, big-endian, at the offset ("begin") its header entry records,
 * padded to 4 bytes.  Version 2 stores begin as a 64-bit offset, so files
 * may exceed 2 GiB; only the last variable may be larger than 4 GiB (its
 * vsize is then written as 2^32 - 1).  A dimension defined with length 0 is
 * the record (unlimited) dimension, written with no records (numrecs 0):
 * the reference's files have one where a count is zero (a point log without
 * local obs, common/pointlog.c:115).  Its variables, which must have it
 * first, hold no values; their begin is the end of the fixed data.
 *
 * Layouts of the path's files follow the reference writers:
 * calc/transforms.c:54-135 (transforms.nc), 485-520 (enkf_diag.nc),
 * diag/diags.c:79-168 (spread), 336-390 (inflation).
 */
#define _FILE_OFFSET_BITS 64
#define _GNU_SOURCE
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>

#include "enkfio.h"

#define TAG_DIMENSION 0x0Au
#define TAG_VARIABLE 0x0Bu
#define TAG_ATTRIBUTE 0x0Cu
#define MAXNAME 256

typedef struct {
    char* name;
    int type;
    size_t n;
    void* vals;          /* native-endian values */
} ncatt;

typedef struct {
    char* name;
    int type;
    int ndims;
    int dimids[ENKFIO_MAXDIMS];
    ncatt* atts;
    int natt;
    uint64_t begin;
    uint64_t nbytes;     /* unpadded data size (all records) */
    int isrec;           /* along the record dimension */
    uint64_t recbytes;   /* unpadded bytes per record (record variables) */
} ncvar;

typedef struct {
    char* name;
    size_t len;
} ncdim;

struct ncfile {
    FILE* fp;
    int writing;
    int defmode;
    int version;
    ncdim* dims;
    int ndims;
    ncatt* gatts;
    int ngatt;
    ncvar* vars;
    int nvars;
    int recdim;          /* the record (unlimited) dimension, -1 if none */
    uint64_t numrecs;
    uint64_t recsize;    /* bytes of one record of all record variables */
};

static size_t type_size(int t)
{
    switch (t) {
    case ENKFIO_BYTE:
    case ENKFIO_CHAR:
        return 1;
    case ENKFIO_SHORT:
        return 2;
    case ENKFIO_INT:
    case ENKFIO_FLOAT:
        return 4;
    case ENKFIO_DOUBLE:
        return 8;
    default:
        return 0;
    }
}

static size_t pad4(size_t n)
{
    return (n + 3) & ~(size_t) 3;
}

/* native <-> big-endian, in place, for elements of `size` bytes */
static void swap_be(void* p, size_t n, size_t size)
{
    const uint16_t one = 1;
    unsigned char* b = p;
    size_t i, k;

    if (size == 1 || *(const unsigned char*) &one == 0)
        return;                 /* bytes, or a big-endian host */
    for (i = 0; i < n; ++i, b += size)
        for (k = 0; k < size / 2; ++k) {
            unsigned char t = b[k];

            b[k] = b[size - 1 - k];
            b[size - 1 - k] = t;
        }
}

/* ---------------------------------------------------------------- header writing */
typedef struct {
    unsigned char* p;
    size_t n, cap;
} buf;

static int buf_put(buf* b, const void* src, size_t n)
{
    if (b->n + n > b->cap) {
        size_t cap = b->cap ? 2 * b->cap : 4096;
        unsigned char* q;

        while (cap < b->n + n)
            cap *= 2;
        q = realloc(b->p, cap);
        if (q == NULL)
            return ENKFIO_ENOMEM;
        b->p = q;
        b->cap = cap;
    }
    memcpy(b->p + b->n, src, n);
    b->n += n;
    return ENKFIO_OK;
}

static int put_u32(buf* b, uint32_t v)
{
    unsigned char c[4] = { (unsigned char) (v >> 24), (unsigned char) (v >> 16), (unsigned char) (v >> 8), (unsigned char) v };

    return buf_put(b, c, 4);
}

static int put_u64(buf* b, uint64_t v)
{
    int r = put_u32(b, (uint32_t) (v >> 32));

    return r ? r : put_u32(b, (uint32_t) v);
}

static int put_pad(buf* b, size_t n)
{
    static const unsigned char z[4] = { 0, 0, 0, 0 };

    return (pad4(n) > n) ? buf_put(b, z, pad4(n) - n) : ENKFIO_OK;
}

static int put_name(buf* b, const char* s)
{
    size_t n = strlen(s);
    int r = put_u32(b, (uint32_t) n);

    if (!r)
        r = buf_put(b, s, n);
    if (!r)
        r = put_pad(b, n);
    return r;
}

static int put_atts(buf* b, const ncatt* a, int n)
{
    int k, r;

    if (n == 0) {
        r = put_u32(b, 0);
        return r ? r : put_u32(b, 0);
    }
    if ((r = put_u32(b, TAG_ATTRIBUTE)) || (r = put_u32(b, (uint32_t) n)))
        return r;
    for (k = 0; k < n; ++k) {
        size_t ts = type_size(a[k].type), nb = ts * a[k].n;
        unsigned char* tmp = malloc(nb ? nb : 1);

        if (tmp == NULL)
            return ENKFIO_ENOMEM;
        memcpy(tmp, a[k].vals, nb);
        swap_be(tmp, a[k].n, ts);
        r = put_name(b, a[k].name);
        if (!r)
            r = put_u32(b, (uint32_t) a[k].type);
        if (!r)
            r = put_u32(b, (uint32_t) a[k].n);
        if (!r)
            r = buf_put(b, tmp, nb);
        if (!r)
            r = put_pad(b, nb);
        free(tmp);
        if (r)
            return r;
    }
    return ENKFIO_OK;
}

/* the header with the variables' offsets; sets every begin from `data0` */
static int build_header(ncfile* f, buf* b, uint64_t data0)
{
    int k, d, r;
    uint64_t off = data0, recoff = data0;

    b->n = 0;
    if ((r = buf_put(b, "CDF\002", 4)) || (r = put_u32(b, 0)))
        return r;
    if (f->ndims == 0) {
        if ((r = put_u32(b, 0)) || (r = put_u32(b, 0)))
            return r;
    } else {
        if ((r = put_u32(b, TAG_DIMENSION)) || (r = put_u32(b, (uint32_t) f->ndims)))
            return r;
        for (k = 0; k < f->ndims; ++k)
            if ((r = put_name(b, f->dims[k].name)) || (r = put_u32(b, (uint32_t) f->dims[k].len)))
                return r;
    }
    if ((r = put_atts(b, f->gatts, f->ngatt)))
        return r;
    if (f->nvars == 0) {
        if ((r = put_u32(b, 0)) || (r = put_u32(b, 0)))
            return r;
        return ENKFIO_OK;
    }
    if ((r = put_u32(b, TAG_VARIABLE)) || (r = put_u32(b, (uint32_t) f->nvars)))
        return r;
    /* the records follow every fixed variable's values */
    for (k = 0; k < f->nvars; ++k)
        if (!f->vars[k].isrec)
            recoff += pad4(f->vars[k].nbytes);
    for (k = 0; k < f->nvars; ++k) {
        ncvar* v = &f->vars[k];
        uint64_t vsize = pad4(v->nbytes);

        if (vsize > 0xFFFFFFFCull && k != f->nvars - 1)
            return ENKFIO_ERANGE;
        if ((r = put_name(b, v->name)) || (r = put_u32(b, (uint32_t) v->ndims)))
            return r;
        for (d = 0; d < v->ndims; ++d)
            if ((r = put_u32(b, (uint32_t) v->dimids[d])))
                return r;
        if ((r = put_atts(b, v->atts, v->natt)) || (r = put_u32(b, (uint32_t) v->type)))
            return r;
        if ((r = put_u32(b, vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t) vsize)))
            return r;
        if (v->isrec) {
            v->begin = recoff;
            recoff += vsize;
        } else {
            v->begin = off;
            off += vsize;
        }
        if ((r = put_u64(b, v->begin)))
            return r;
    }
    return ENKFIO_OK;
}

/* ---------------------------------------------------------------- writing */
int ncf_create(const char* path, ncfile** out)
{
    ncfile* f;

    if (path == NULL || out == NULL)
        return ENKFIO_EARG;
    *out = NULL;
    f = calloc(1, sizeof(ncfile));
    if (f == NULL)
        return ENKFIO_ENOMEM;
    f->fp = fopen(path, "wb");
    if (f->fp == NULL) {
        free(f);
        return ENKFIO_EIO;
    }
    f->writing = 1;
    f->defmode = 1;
    f->version = 2;
    f->recdim = -1;
    *out = f;
    return ENKFIO_OK;
}

static int valid_name(const char* s)
{
    return s != NULL && s[0] != 0 && strlen(s) < MAXNAME;
}

int ncf_def_dim(ncfile* f, const char* name, size_t len)
{
    ncdim* d;

    if (f == NULL || !f->defmode || !valid_name(name) || len > 0xFFFFFFFFu || (len == 0 && f->recdim >= 0))
        return ENKFIO_EARG;
    d = realloc(f->dims, (f->ndims + 1) * sizeof(ncdim));
    if (d == NULL)
        return ENKFIO_ENOMEM;
    f->dims = d;
    d[f->ndims].name = strdup(name);
    d[f->ndims].len = len;
    if (len == 0)
        f->recdim = f->ndims;
    return f->ndims++;
}

int ncf_def_var(ncfile* f, const char* name, int type, int ndims, const int* dimids)
{
    ncvar* v;
    uint64_t n = 1;
    int k;

    if (f == NULL || !f->defmode || !valid_name(name) || type_size(type) == 0 || ndims < 0 ||
        ndims > ENKFIO_MAXDIMS || (ndims > 0 && dimids == NULL))
        return ENKFIO_EARG;
    for (k = 0; k < ndims; ++k) {
        if (dimids[k] < 0 || dimids[k] >= f->ndims || (k > 0 && dimids[k] == f->recdim))
            return ENKFIO_EARG;
        if (k > 0 || dimids[k] != f->recdim)
            n *= f->dims[dimids[k]].len;
    }
    v = realloc(f->vars, (f->nvars + 1) * sizeof(ncvar));
    if (v == NULL)
        return ENKFIO_ENOMEM;
    f->vars = v;
    v = &f->vars[f->nvars];
    memset(v, 0, sizeof(*v));
    v->name = strdup(name);
    v->type = type;
    v->ndims = ndims;
    for (k = 0; k < ndims; ++k)
        v->dimids[k] = dimids[k];
    v->isrec = ndims > 0 && dimids[0] == f->recdim;
    /* a record variable: the size of one record (the header's vsize) */
    v->nbytes = n * type_size(type);
    return f->nvars++;
}

int ncf_put_att(ncfile* f, int varid, const char* name, int type, size_t n, const void* values)
{
    ncatt** list;
    int* cnt;
    ncatt* a;
    size_t nb = type_size(type) * n;

    if (f == NULL || !f->defmode || !valid_name(name) || type_size(type) == 0 || (n > 0 && values == NULL))
        return ENKFIO_EARG;
    if (varid == ENKFIO_GLOBAL) {
        list = &f->gatts;
        cnt = &f->ngatt;
    } else if (varid >= 0 && varid < f->nvars) {
        list = &f->vars[varid].atts;
        cnt = &f->vars[varid].natt;
    } else
        return ENKFIO_EARG;
    a = realloc(*list, (*cnt + 1) * sizeof(ncatt));
    if (a == NULL)
        return ENKFIO_ENOMEM;
    *list = a;
    a = &a[*cnt];
    a->name = strdup(name);
    a->type = type;
    a->n = n;
    a->vals = malloc(nb ? nb : 1);
    if (a->vals == NULL)
        return ENKFIO_ENOMEM;
    memcpy(a->vals, values, nb);
    (*cnt)++;
    return ENKFIO_OK;
}

int ncf_put_att_text(ncfile* f, int varid, const char* name, const char* text)
{
    return ncf_put_att(f, varid, name, ENKFIO_CHAR, text ? strlen(text) : 0, text ? text : "");
}

int ncf_enddef(ncfile* f)
{
    buf b = { NULL, 0, 0 };
    uint64_t end;
    int r;

    if (f == NULL || !f->defmode)
        return ENKFIO_EARG;
    /* the header's size does not depend on the offsets it records */
    if ((r = build_header(f, &b, 0)) == ENKFIO_OK)
        r = build_header(f, &b, b.n);
    if (r == ENKFIO_OK && fwrite(b.p, 1, b.n, f->fp) != b.n)
        r = ENKFIO_EIO;
    /* extend the file to its full size: every fixed variable's values exist
       (record variables have no records) */
    if (r == ENKFIO_OK && f->nvars > 0) {
        int k;

        end = 0;
        for (k = 0; k < f->nvars; ++k)
            if (!f->vars[k].isrec && f->vars[k].begin + pad4(f->vars[k].nbytes) > end)
                end = f->vars[k].begin + pad4(f->vars[k].nbytes);
        if (end > b.n && (fseeko(f->fp, (off_t) (end - 1), SEEK_SET) != 0 || fputc(0, f->fp) == EOF))
            r = ENKFIO_EIO;
    }
    free(b.p);
    if (r == ENKFIO_OK)
        f->defmode = 0;
    return r;
}

static uint64_t row_elems(const ncfile* f, const ncvar* v)
{
    uint64_t n = 1;
    int k;

    for (k = 1; k < v->ndims; ++k)
        n *= f->dims[v->dimids[k]].len;
    return n;
}

static int slab_range(const ncfile* f, const ncvar* v, size_t start, size_t count, uint64_t* off, uint64_t* nel)
{
    uint64_t first = v->ndims > 0 ? f->dims[v->dimids[0]].len : 1;

    if (start > first || count > first - start)
        return ENKFIO_ERANGE;
    *nel = (uint64_t) count * row_elems(f, v);
    *off = v->begin + (uint64_t) start * row_elems(f, v) * type_size(v->type);
    return ENKFIO_OK;
}

int ncf_put_slab(ncfile* f, int varid, size_t start, size_t count, const void* data)
{
    const ncvar* v;
    uint64_t off, nel, done = 0;
    size_t ts, chunk = (size_t) 1 << 20;
    unsigned char* tmp;
    int r;

    if (f == NULL || !f->writing || f->defmode || varid < 0 || varid >= f->nvars || (count > 0 && data == NULL))
        return ENKFIO_EARG;
    v = &f->vars[varid];
    if ((r = slab_range(f, v, start, count, &off, &nel)))
        return r;
    ts = type_size(v->type);
    if (fseeko(f->fp, (off_t) off, SEEK_SET) != 0)
        return ENKFIO_EIO;
    tmp = malloc(chunk * ts);
    if (tmp == NULL)
        return ENKFIO_ENOMEM;
    while (done < nel) {
        size_t n = (nel - done < chunk) ? (size_t) (nel - done) : chunk;

        memcpy(tmp, (const unsigned char*) data + done * ts, n * ts);
        swap_be(tmp, n, ts);
        if (fwrite(tmp, ts, n, f->fp) != n) {
            free(tmp);
            return ENKFIO_EIO;
        }
        done += n;
    }
    free(tmp);
    return ENKFIO_OK;
}

/* ---------------------------------------------------------------- reading */
typedef struct {
    const unsigned char* p;
    size_t n, pos;
} rd;

static int get_u32(rd* r, uint32_t* v)
{
    if (r->pos + 4 > r->n)
        return ENKFIO_EFORMAT;
    *v = ((uint32_t) r->p[r->pos] << 24) | ((uint32_t) r->p[r->pos + 1] << 16) | ((uint32_t) r->p[r->pos + 2] << 8) | r->p[r->pos + 3];
    r->pos += 4;
    return ENKFIO_OK;
}

static int get_bytes(rd* r, void* dst, size_t n)
{
    if (r->pos + n > r->n)
        return ENKFIO_EFORMAT;
    memcpy(dst, r->p + r->pos, n);
    r->pos += pad4(n);
    if (r->pos > r->n)
        return ENKFIO_EFORMAT;
    return ENKFIO_OK;
}

static int get_name(rd* r, char** out)
{
    uint32_t n;
    int e = get_u32(r, &n);

    if (e)
        return e;
    if (n >= MAXNAME * 16)
        return ENKFIO_EFORMAT;
    *out = calloc(n + 1, 1);
    if (*out == NULL)
        return ENKFIO_ENOMEM;
    return get_bytes(r, *out, n);
}

static int get_atts(rd* r, ncatt** out, int* nout)
{
    uint32_t tag, n, k;
    int e;

    if ((e = get_u32(r, &tag)) || (e = get_u32(r, &n)))
        return e;
    *out = NULL;
    *nout = 0;
    if (tag == 0 && n == 0)
        return ENKFIO_OK;
    if (tag != TAG_ATTRIBUTE || n > 100000)
        return ENKFIO_EFORMAT;
    *out = calloc(n, sizeof(ncatt));
    if (*out == NULL)
        return ENKFIO_ENOMEM;
    for (k = 0; k < n; ++k) {
        ncatt* a = &(*out)[k];
        uint32_t type, nel;
        size_t ts;

        if ((e = get_name(r, &a->name)) || (e = get_u32(r, &type)) || (e = get_u32(r, &nel)))
            return e;
        ts = type_size((int) type);
        if (ts == 0 || (uint64_t) nel * ts > r->n)
            return ENKFIO_EFORMAT;
        a->type = (int) type;
        a->n = nel;
        a->vals = malloc(nel * ts + 1);
        if (a->vals == NULL)
            return ENKFIO_ENOMEM;
        if ((e = get_bytes(r, a->vals, nel * ts)))
            return e;
        swap_be(a->vals, nel, ts);
        (*nout)++;
    }
    return ENKFIO_OK;
}

int ncf_open(const char* path, ncfile** out)
{
    ncfile* f;
    unsigned char* hdr = NULL;
    size_t cap = 1 << 16, got;
    int e = ENKFIO_OK;

    if (path == NULL || out == NULL)
        return ENKFIO_EARG;
    *out = NULL;
    f = calloc(1, sizeof(ncfile));
    if (f == NULL)
        return ENKFIO_ENOMEM;
    f->fp = fopen(path, "rb");
    if (f->fp == NULL) {
        free(f);
        return ENKFIO_EIO;
    }
    /* read a prefix of the file large enough to hold the header: grow it
     * until the header parses (or the whole file is in) */
    for (;;) {
        rd r;
        uint32_t nrec, tag, n, k;

        free(hdr);
        hdr = malloc(cap);
        if (hdr == NULL) {
            e = ENKFIO_ENOMEM;
            break;
        }
        rewind(f->fp);
        got = fread(hdr, 1, cap, f->fp);
        r.p = hdr;
        r.n = got;
        r.pos = 0;
        if (got < 8 || memcmp(hdr, "CDF", 3) != 0 || (hdr[3] != 1 && hdr[3] != 2)) {
            e = ENKFIO_EFORMAT;
            break;
        }
        f->version = hdr[3];
        r.pos = 4;
        e = get_u32(&r, &nrec);
        if (!e && nrec == 0xFFFFFFFFu)
            e = ENKFIO_EFORMAT;        /* a file still being written (streaming) */
        f->numrecs = nrec;
        f->recdim = -1;
        /* dimensions */
        if (!e)
            e = get_u32(&r, &tag);
        if (!e)
            e = get_u32(&r, &n);
        if (!e && !(tag == 0 && n == 0) && (tag != TAG_DIMENSION || n > 65536))
            e = ENKFIO_EFORMAT;
        if (!e && tag == TAG_DIMENSION) {
            f->dims = calloc(n ? n : 1, sizeof(ncdim));
            if (f->dims == NULL)
                e = ENKFIO_ENOMEM;
            for (k = 0; k < n && !e; ++k) {
                uint32_t len = 0;

                e = get_name(&r, &f->dims[k].name);
                if (!e)
                    e = get_u32(&r, &len);
                if (!e && len == 0) {  /* the record dimension */
                    if (f->recdim >= 0)
                        e = ENKFIO_EFORMAT;
                    f->recdim = (int) k;
                    len = nrec;
                }
                f->dims[k].len = len;
                if (!e)
                    f->ndims++;
            }
        }
        if (!e)
            e = get_atts(&r, &f->gatts, &f->ngatt);
        if (!e)
            e = get_u32(&r, &tag);
        if (!e)
            e = get_u32(&r, &n);
        if (!e && !(tag == 0 && n == 0) && (tag != TAG_VARIABLE || n > 65536))
            e = ENKFIO_EFORMAT;
        if (!e && tag == TAG_VARIABLE) {
            f->vars = calloc(n ? n : 1, sizeof(ncvar));
            if (f->vars == NULL)
                e = ENKFIO_ENOMEM;
            for (k = 0; k < n && !e; ++k) {
                ncvar* v = &f->vars[k];
                uint32_t nd = 0, d, type = 0, vsize = 0, hi = 0, lo = 0;
                uint64_t nel = 1;

                e = get_name(&r, &v->name);
                if (!e)
                    e = get_u32(&r, &nd);
                if (!e && nd > ENKFIO_MAXDIMS)
                    e = ENKFIO_EFORMAT;
                for (d = 0; d < nd && !e; ++d) {
                    uint32_t id;

                    e = get_u32(&r, &id);
                    if (!e && id >= (uint32_t) f->ndims)
                        e = ENKFIO_EFORMAT;
                    if (!e) {
                        v->dimids[d] = (int) id;
                        nel *= f->dims[id].len;
                    }
                }
                v->ndims = (int) nd;
                if (!e)
                    e = get_atts(&r, &v->atts, &v->natt);
                if (!e)
                    e = get_u32(&r, &type);
                if (!e)
                    e = get_u32(&r, &vsize);
                if (!e && f->version == 2) {
                    e = get_u32(&r, &hi);
                    if (!e)
                        e = get_u32(&r, &lo);
                    v->begin = ((uint64_t) hi << 32) | lo;
                } else if (!e) {
                    e = get_u32(&r, &lo);
                    v->begin = lo;
                }
                if (!e && type_size((int) type) == 0)
                    e = ENKFIO_EFORMAT;
                v->type = (int) type;
                v->nbytes = nel * type_size((int) type);
                v->isrec = (nd > 0 && f->recdim >= 0 && v->dimids[0] == f->recdim);
                if (v->isrec)
                    v->recbytes = (nrec > 0 ? nel / nrec : 0) * type_size((int) type);
                if (v->isrec && nrec == 0) {   /* no records yet: per-record size from the other dims */
                    uint64_t per = 1;
                    for (d = 1; d < nd; ++d)
                        per *= f->dims[v->dimids[d]].len;
                    v->recbytes = per * type_size((int) type);
                }
                if (!e)
                    f->nvars++;
            }
        }
        if (!e) {
            /* records hold every record variable in order, each padded to
             * 4 bytes, except when there is only one */
            int nrv = 0, k2;

            f->recsize = 0;
            for (k2 = 0; k2 < f->nvars; ++k2)
                nrv += f->vars[k2].isrec;
            for (k2 = 0; k2 < f->nvars; ++k2)
                if (f->vars[k2].isrec)
                    f->recsize += (nrv == 1) ? f->vars[k2].recbytes : pad4(f->vars[k2].recbytes);
        }
        if (e == ENKFIO_EFORMAT && got == cap && cap < ((size_t) 1 << 30)) {
            /* the header may continue past the prefix read: a longer one */
            cap *= 4;
            e = ENKFIO_OK;
            /* reset what was parsed */
            {
                int i, j;

                for (i = 0; i < f->ndims; ++i)
                    free(f->dims[i].name);
                free(f->dims);
                f->dims = NULL;
                f->ndims = 0;
                for (i = 0; i < f->ngatt; ++i) {
                    free(f->gatts[i].name);
                    free(f->gatts[i].vals);
                }
                free(f->gatts);
                f->gatts = NULL;
                f->ngatt = 0;
                for (i = 0; i < f->nvars; ++i) {
                    free(f->vars[i].name);
                    for (j = 0; j < f->vars[i].natt; ++j) {
                        free(f->vars[i].atts[j].name);
                        free(f->vars[i].atts[j].vals);
                    }
                    free(f->vars[i].atts);
                }
                free(f->vars);
                f->vars = NULL;
                f->nvars = 0;
            }
            continue;
        }
        break;
    }
    free(hdr);
    if (e) {
        ncf_close(f);
        return e;
    }
    *out = f;
    return ENKFIO_OK;
}

int ncf_inq(const ncfile* f, int* ndims, int* nvars, int* ngatts)
{
    if (f == NULL)
        return ENKFIO_EARG;
    if (ndims)
        *ndims = f->ndims;
    if (nvars)
        *nvars = f->nvars;
    if (ngatts)
        *ngatts = f->ngatt;
    return ENKFIO_OK;
}

int ncf_inq_varname(const ncfile* f, int varid, char* name, size_t cap)
{
    if (f == NULL || varid < 0 || varid >= f->nvars || name == NULL || cap == 0)
        return ENKFIO_EARG;
    strncpy(name, f->vars[varid].name, cap - 1);
    name[cap - 1] = 0;
    return ENKFIO_OK;
}

int ncf_inq_varid(const ncfile* f, const char* name)
{
    int k;

    if (f == NULL || name == NULL)
        return ENKFIO_EARG;
    for (k = 0; k < f->nvars; ++k)
        if (strcmp(f->vars[k].name, name) == 0)
            return k;
    return ENKFIO_EARG;
}

int ncf_inq_var(const ncfile* f, int varid, int* type, int* ndims, size_t* dimlens)
{
    const ncvar* v;
    int k;

    if (f == NULL || varid < 0 || varid >= f->nvars)
        return ENKFIO_EARG;
    v = &f->vars[varid];
    if (type)
        *type = v->type;
    if (ndims)
        *ndims = v->ndims;
    if (dimlens)
        for (k = 0; k < v->ndims; ++k)
            dimlens[k] = f->dims[v->dimids[k]].len;
    return ENKFIO_OK;
}

int ncf_inq_dimlen(const ncfile* f, const char* name, size_t* len)
{
    int k;

    if (f == NULL || name == NULL)
        return ENKFIO_EARG;
    for (k = 0; k < f->ndims; ++k)
        if (strcmp(f->dims[k].name, name) == 0) {
            if (len)
                *len = f->dims[k].len;
            return ENKFIO_OK;
        }
    return ENKFIO_EARG;
}

int ncf_get_att(const ncfile* f, int varid, const char* name, int* type, size_t* n, void* values, size_t cap)
{
    const ncatt* a;
    int na, k;

    if (f == NULL || name == NULL)
        return ENKFIO_EARG;
    if (varid == ENKFIO_GLOBAL) {
        a = f->gatts;
        na = f->ngatt;
    } else if (varid >= 0 && varid < f->nvars) {
        a = f->vars[varid].atts;
        na = f->vars[varid].natt;
    } else
        return ENKFIO_EARG;
    for (k = 0; k < na; ++k)
        if (strcmp(a[k].name, name) == 0) {
            size_t ts = type_size(a[k].type);
            size_t m = a[k].n < cap ? a[k].n : cap;

            if (type)
                *type = a[k].type;
            if (n)
                *n = a[k].n;
            if (values != NULL)
                memcpy(values, a[k].vals, m * ts);
            return ENKFIO_OK;
        }
    return ENKFIO_EARG;
}

int ncf_get_slab(ncfile* f, int varid, size_t start, size_t count, void* data)
{
    const ncvar* v;
    uint64_t off, nel;
    size_t ts;
    int r;

    if (f == NULL || varid < 0 || varid >= f->nvars || (count > 0 && data == NULL))
        return ENKFIO_EARG;
    v = &f->vars[varid];
    ts = type_size(v->type);
    if (v->isrec) {
        size_t k, per = (size_t) (v->recbytes / ts);

        if (start > f->numrecs || count > f->numrecs - start)
            return ENKFIO_ERANGE;
        for (k = 0; k < count; ++k) {
            unsigned char* dst = (unsigned char*) data + k * v->recbytes;

            if (fseeko(f->fp, (off_t) (v->begin + (start + k) * f->recsize), SEEK_SET) != 0)
                return ENKFIO_EIO;
            if (per > 0 && fread(dst, ts, per, f->fp) != per)
                return ENKFIO_EIO;
            swap_be(dst, per, ts);
        }
        return ENKFIO_OK;
    }
    if ((r = slab_range(f, v, start, count, &off, &nel)))
        return r;
    if (fseeko(f->fp, (off_t) off, SEEK_SET) != 0)
        return ENKFIO_EIO;
    if (nel > 0 && fread(data, ts, (size_t) nel, f->fp) != (size_t) nel)
        return ENKFIO_EIO;
    swap_be(data, (size_t) nel, ts);
    return ENKFIO_OK;
}

int ncf_get_vara(ncfile* f, int varid, const size_t* start, const size_t* count, void* data)
{
    const ncvar* v;
    size_t len[ENKFIO_MAXDIMS], idx[ENKFIO_MAXDIMS], ts, run, nrow = 1;
    uint64_t stride[ENKFIO_MAXDIMS];
    unsigned char* dst = data;
    int nd, d;

    if (f == NULL || varid < 0 || varid >= f->nvars)
        return ENKFIO_EARG;
    v = &f->vars[varid];
    nd = v->ndims;
    ts = type_size(v->type);
    if (nd == 0)
        return ncf_get_slab(f, varid, 0, 1, data);
    if (start == NULL || count == NULL || data == NULL)
        return ENKFIO_EARG;
    if (v->isrec && nd == 1)   /* records are not contiguous */
        return ncf_get_slab(f, varid, start[0], count[0], data);
    for (d = 0; d < nd; ++d) {
        len[d] = f->dims[v->dimids[d]].len;
        if (start[d] > len[d] || count[d] > len[d] - start[d])
            return ENKFIO_ERANGE;
        if (count[d] == 0)
            return ENKFIO_OK;
    }
    /* element strides within a record (record variables) or the variable */
    stride[nd - 1] = 1;
    for (d = nd - 2; d >= 0; --d)
        stride[d] = stride[d + 1] * len[d + 1];
    for (d = 0; d < nd - 1; ++d)
        nrow *= count[d];
    run = count[nd - 1];
    for (d = 0; d < nd; ++d)
        idx[d] = start[d];
    while (nrow-- > 0) {
        uint64_t off;

        if (v->isrec) {
            uint64_t e = 0;

            for (d = 1; d < nd; ++d)
                e += idx[d] * stride[d];
            off = v->begin + idx[0] * f->recsize + e * ts;
        } else {
            uint64_t e = 0;

            for (d = 0; d < nd; ++d)
                e += idx[d] * stride[d];
            off = v->begin + e * ts;
        }
        if (fseeko(f->fp, (off_t) off, SEEK_SET) != 0 || fread(dst, ts, run, f->fp) != run)
            return ENKFIO_EIO;
        swap_be(dst, run, ts);
        dst += run * ts;
        /* next row: odometer over dimensions 0 .. nd - 2 */
        for (d = nd - 2; d >= 0; --d) {
            if (++idx[d] < start[d] + count[d])
                break;
            idx[d] = start[d];
        }
    }
    return ENKFIO_OK;
}

/* value i of a native-endian array of netCDF type t, as double */
static double get_as_double(const void* p, int t, size_t i)
{
    switch (t) {
    case ENKFIO_BYTE:
        return (double) ((const signed char*) p)[i];
    case ENKFIO_CHAR:
        return (double) ((const unsigned char*) p)[i];
    case ENKFIO_SHORT:
        return (double) ((const int16_t*) p)[i];
    case ENKFIO_INT:
        return (double) ((const int32_t*) p)[i];
    case ENKFIO_FLOAT:
        return (double) ((const float*) p)[i];
    default:
        return ((const double*) p)[i];
    }
}

int ncf_get_vara_as(ncfile* f, int varid, const size_t* start, const size_t* count, int type, void* out)
{
    int vt, nd, d, r;
    size_t n = 1, dl[ENKFIO_MAXDIMS], st0[ENKFIO_MAXDIMS], i, ts;
    const size_t* st = start;
    const size_t* ct = count;
    void* raw;

    if ((r = ncf_inq_var(f, varid, &vt, &nd, dl)) != ENKFIO_OK)
        return r;
    if (type_size(type) == 0 || vt == ENKFIO_CHAR || type == ENKFIO_CHAR || out == NULL)
        return ENKFIO_EARG;
    if (st == NULL) {   /* the whole variable */
        for (d = 0; d < nd; ++d)
            st0[d] = 0;
        st = st0;
        ct = dl;
    }
    for (d = 0; d < nd; ++d)
        n *= ct[d];
    ts = type_size(vt);
    raw = malloc(n * ts + 1);
    if (raw == NULL)
        return ENKFIO_ENOMEM;
    if ((r = ncf_get_vara(f, varid, st, ct, raw)) != ENKFIO_OK) {
        free(raw);
        return r;
    }
    for (i = 0; i < n; ++i) {
        const double x = get_as_double(raw, vt, i);

        switch (type) {
        case ENKFIO_BYTE:
            ((signed char*) out)[i] = (signed char) x;
            break;
        case ENKFIO_SHORT:
            ((int16_t*) out)[i] = (int16_t) x;
            break;
        case ENKFIO_INT:
            ((int32_t*) out)[i] = (int32_t) x;
            break;
        case ENKFIO_FLOAT:
            /* a double beyond the float range reads as NaN
             * (ncw_get_vara_float_fixerange, common/ncw.c:1078-1112) */
            ((float*) out)[i] = (!isfinite(x) || x < -FLT_MAX || x > FLT_MAX) && vt == ENKFIO_DOUBLE ? NAN : (float) x;
            break;
        default:
            ((double*) out)[i] = x;
        }
    }
    free(raw);
    return ENKFIO_OK;
}

int ncf_inq_natts(const ncfile* f, int varid, int* natts)
{
    if (f == NULL || natts == NULL)
        return ENKFIO_EARG;
    if (varid == ENKFIO_GLOBAL)
        *natts = f->ngatt;
    else if (varid >= 0 && varid < f->nvars)
        *natts = f->vars[varid].natt;
    else
        return ENKFIO_EARG;
    return ENKFIO_OK;
}

int ncf_inq_attname(const ncfile* f, int varid, int k, char* name, size_t cap)
{
    const ncatt* a;
    int na;

    if (f == NULL || name == NULL || cap == 0)
        return ENKFIO_EARG;
    if (varid == ENKFIO_GLOBAL) {
        a = f->gatts;
        na = f->ngatt;
    } else if (varid >= 0 && varid < f->nvars) {
        a = f->vars[varid].atts;
        na = f->vars[varid].natt;
    } else
        return ENKFIO_EARG;
    if (k < 0 || k >= na)
        return ENKFIO_EARG;
    strncpy(name, a[k].name, cap - 1);
    name[cap - 1] = 0;
    return ENKFIO_OK;
}

int ncf_close(ncfile* f)
{
    int k, j, r = ENKFIO_OK;

    if (f == NULL)
        return ENKFIO_EARG;
    if (f->writing && f->defmode)
        r = ncf_enddef(f);
    if (f->fp != NULL && fclose(f->fp) != 0 && r == ENKFIO_OK)
        r = ENKFIO_EIO;
    for (k = 0; k < f->ndims; ++k)
        free(f->dims[k].name);
    free(f->dims);
    for (k = 0; k < f->ngatt; ++k) {
        free(f->gatts[k].name);
        free(f->gatts[k].vals);
    }
    free(f->gatts);
    for (k = 0; k < f->nvars; ++k) {
        free(f->vars[k].name);
        for (j = 0; j < f->vars[k].natt; ++j) {
            free(f->vars[k].atts[j].name);
            free(f->vars[k].atts[j].vals);
        }
        free(f->vars[k].atts);
    }
    free(f->vars);
    free(f);
    return r;
}

/* ---------------------------------------------------------------- the path's files */
#define TRY(x) do { int r_ = (x); if (r_ < 0) { ncf_close(f); return r_; } } while (0)

int enkfio_write_transforms(const char* path, int nj, int ni, int m_dyn, int m, int stride, const char* grid_name, const float* T, const float* w)
{
    ncfile* f;
    int dj, di, dmd, dm, vT = -1, vw, d4[4], d3[3];

    if (nj <= 0 || ni <= 0 || m <= 0 || w == NULL || (T != NULL && (m_dyn <= 0 || m_dyn > m)))
        return ENKFIO_EARG;
    TRY(ncf_create(path, &f) ? ENKFIO_EIO : 0);
    TRY(dj = ncf_def_dim(f, "j", (size_t) nj));
    TRY(di = ncf_def_dim(f, "i", (size_t) ni));
    if (T != NULL) {
        TRY(dmd = ncf_def_dim(f, "m_dyn", (size_t) m_dyn));
        TRY(dm = ncf_def_dim(f, "m", (size_t) m));
    } else {
        dmd = -1;
        TRY(dm = ncf_def_dim(f, "m", (size_t) m));
    }
    d3[0] = dj;
    d3[1] = di;
    d3[2] = dm;
    /* w first: T, the largest, last (only the last variable may pass 4 GiB) */
    TRY(vw = ncf_def_var(f, "w", ENKFIO_FLOAT, 3, d3));
    if (T != NULL) {
        d4[0] = dj;
        d4[1] = di;
        d4[2] = dmd;
        d4[3] = dm;
        TRY(vT = ncf_def_var(f, "T", ENKFIO_FLOAT, 4, d4));
    }
    TRY(ncf_put_att(f, ENKFIO_GLOBAL, "stride", ENKFIO_INT, 1, &stride));
    TRY(ncf_put_att_text(f, ENKFIO_GLOBAL, "grid_name", grid_name ? grid_name : ""));
    TRY(ncf_enddef(f));
    TRY(ncf_put_slab(f, vw, 0, (size_t) nj, w));
    if (T != NULL)
        TRY(ncf_put_slab(f, vT, 0, (size_t) nj, T));
    return ncf_close(f);
}

int enkfio_inq_transforms(const char* path, int* nj, int* ni, int* m_dyn, int* m, int* stride)
{
    ncfile* f;
    size_t d[ENKFIO_MAXDIMS];
    int vw, vT, nd, st = 0;
    size_t n;

    TRY(ncf_open(path, &f));
    TRY(vw = ncf_inq_varid(f, "w"));
    TRY(ncf_inq_var(f, vw, NULL, &nd, d));
    if (nd != 3) {
        ncf_close(f);
        return ENKFIO_EFORMAT;
    }
    if (nj)
        *nj = (int) d[0];
    if (ni)
        *ni = (int) d[1];
    if (m)
        *m = (int) d[2];
    vT = ncf_inq_varid(f, "T");
    if (m_dyn) {
        *m_dyn = 0;
        if (vT >= 0) {
            TRY(ncf_inq_var(f, vT, NULL, &nd, d));
            *m_dyn = (int) d[2];
        }
    }
    if (stride) {
        TRY(ncf_get_att(f, ENKFIO_GLOBAL, "stride", NULL, &n, &st, 1));
        *stride = st;
    }
    return ncf_close(f);
}

int enkfio_read_transforms(const char* path, int j0, int nrow, float* T, float* w)
{
    ncfile* f;
    int vw, vT;

    if (j0 < 0 || nrow < 0)
        return ENKFIO_EARG;
    TRY(ncf_open(path, &f));
    TRY(vw = ncf_inq_varid(f, "w"));
    if (w != NULL)
        TRY(ncf_get_slab(f, vw, (size_t) j0, (size_t) nrow, w));
    if (T != NULL) {
        TRY(vT = ncf_inq_varid(f, "T"));
        TRY(ncf_get_slab(f, vT, (size_t) j0, (size_t) nrow, T));
    }
    return ncf_close(f);
}

int enkfio_write_diag(const char* path, int nobstypes, const char* const* typenames, int nj, int ni, int stride, const int* pnlobs, const float* pdfs, const float* psrf)
{
    ncfile* f;
    int dims[3], vn, vd, vs, vpn, vpd, vps, t;
    size_t plane = (size_t) nj * ni;

    if (nobstypes < 0 || nj <= 0 || ni <= 0 || pnlobs == NULL || pdfs == NULL || psrf == NULL)
        return ENKFIO_EARG;
    TRY(ncf_create(path, &f) ? ENKFIO_EIO : 0);
    TRY(dims[0] = ncf_def_dim(f, "nobstypes", (size_t) (nobstypes > 0 ? nobstypes : 1)));
    TRY(dims[1] = ncf_def_dim(f, "j", (size_t) nj));
    TRY(dims[2] = ncf_def_dim(f, "i", (size_t) ni));
    TRY(vn = ncf_def_var(f, "nlobs", ENKFIO_INT, 2, &dims[1]));
    TRY(ncf_put_att_text(f, vn, "long_name", "number of local observations"));
    TRY(vd = ncf_def_var(f, "dfs", ENKFIO_FLOAT, 2, &dims[1]));
    TRY(ncf_put_att_text(f, vd, "long_name", "local DFS (Degrees of Freedom for Signal)"));
    TRY(ncf_put_att_text(f, vd, "comment", "DFS = trace(KH) for vanilla KF"));
    TRY(vs = ncf_def_var(f, "srf", ENKFIO_FLOAT, 2, &dims[1]));
    TRY(ncf_put_att_text(f, vs, "long_name", "local SRF (Spread Reduction Factor)"));
    TRY(ncf_put_att_text(f, vs, "comment", "SRF = (forecast spread / analysis spread) - 1 for vanilla ETKF"));
    TRY(vpn = ncf_def_var(f, "pnlobs", ENKFIO_INT, 3, dims));
    TRY(ncf_put_att_text(f, vpn, "long_name", "partial number of local observations"));
    TRY(vpd = ncf_def_var(f, "pdfs", ENKFIO_FLOAT, 3, dims));
    TRY(ncf_put_att_text(f, vpd, "long_name", "partial local DFS"));
    TRY(vps = ncf_def_var(f, "psrf", ENKFIO_FLOAT, 3, dims));
    TRY(ncf_put_att_text(f, vps, "long_name", "partial local SRF"));
    TRY(ncf_put_att(f, ENKFIO_GLOBAL, "stride", ENKFIO_INT, 1, &stride));
    for (t = 0; t < nobstypes; ++t)
        TRY(ncf_put_att(f, ENKFIO_GLOBAL, typenames[t], ENKFIO_INT, 1, &t));
    TRY(ncf_enddef(f));
    TRY(ncf_put_slab(f, vn, 0, (size_t) nj, pnlobs + (size_t) nobstypes * plane));
    TRY(ncf_put_slab(f, vd, 0, (size_t) nj, pdfs + (size_t) nobstypes * plane));
    TRY(ncf_put_slab(f, vs, 0, (size_t) nj, psrf + (size_t) nobstypes * plane));
    if (nobstypes > 0) {
        TRY(ncf_put_slab(f, vpn, 0, (size_t) nobstypes, pnlobs));
        TRY(ncf_put_slab(f, vpd, 0, (size_t) nobstypes, pdfs));
        TRY(ncf_put_slab(f, vps, 0, (size_t) nobstypes, psrf));
    }
    return ncf_close(f);
}

int enkfio_write_fields2d(const char* path, int nfield, const char* const* names, int nj, int ni, const float* v)
{
    ncfile* f;
    int dims[2], k;
    int* vid;

    if (nfield <= 0 || names == NULL || nj <= 0 || ni <= 0 || v == NULL)
        return ENKFIO_EARG;
    vid = malloc((size_t) nfield * sizeof(int));
    if (vid == NULL)
        return ENKFIO_ENOMEM;
    if (ncf_create(path, &f)) {
        free(vid);
        return ENKFIO_EIO;
    }
#undef TRY
#define TRY(x) do { int r_ = (x); if (r_ < 0) { ncf_close(f); free(vid); return r_; } } while (0)
    TRY(dims[0] = ncf_def_dim(f, "j", (size_t) nj));
    TRY(dims[1] = ncf_def_dim(f, "i", (size_t) ni));
    for (k = 0; k < nfield; ++k)
        TRY(vid[k] = ncf_def_var(f, names[k], ENKFIO_FLOAT, 2, dims));
    TRY(ncf_enddef(f));
    for (k = 0; k < nfield; ++k)
        TRY(ncf_put_slab(f, vid[k], 0, (size_t) nj, v + (size_t) k * nj * ni));
    free(vid);
    return ncf_close(f);
}

/* ---------------------------------------------------------------- model fields and observations */
#undef TRY
#define TRY(x) do { int r_ = (x); if (r_ < 0) { ncf_close(f); return r_; } } while (0)

/* the default fill value of a classic file's type, in the type's bytes */
static void default_fill(int t, unsigned char* v)
{
    const signed char b = -127;
    const int16_t sh = -32767;
    const int32_t in = -2147483647;
    const float fl = 9.9692099683868690e+36f;
    const double db = 9.9692099683868690e+36;

    switch (t) {
    case ENKFIO_BYTE:
        memcpy(v, &b, 1);
        break;
    case ENKFIO_CHAR:
        v[0] = 0;
        break;
    case ENKFIO_SHORT:
        memcpy(v, &sh, 2);
        break;
    case ENKFIO_INT:
        memcpy(v, &in, 4);
        break;
    case ENKFIO_FLOAT:
        memcpy(v, &fl, 4);
        break;
    default:
        memcpy(v, &db, 8);
    }
}

int enkfio_field_nlevels(const char* path, const char* varname, int* nlevels)
{
    ncfile* f;
    size_t dl[ENKFIO_MAXDIMS];
    int vid, nd, rec;

    if (nlevels == NULL)
        return ENKFIO_EARG;
    TRY(ncf_open(path, &f));
    TRY(vid = ncf_inq_varid(f, varname));
    TRY(ncf_inq_var(f, vid, NULL, &nd, dl));
    rec = f->vars[vid].isrec || (nd > 0 && dl[0] == 1);
    /* common/ncutils.c:54-91, structured grids */
    if (nd > 4 || (nd == 4 && !rec) || nd < 2) {
        ncf_close(f);
        return ENKFIO_EARG;
    }
    *nlevels = (nd == 4) ? (int) dl[1] : (nd == 3) ? (rec ? 1 : (int) dl[0]) : (rec ? 0 : 1);
    return ncf_close(f);
}

int enkfio_read_field(const char* path, const char* varname, int r, int k, int j1, int j2, int ni, int nj, int nk, float* v)
{
    ncfile* f;
    size_t dl[ENKFIO_MAXDIMS], start[4] = {0, 0, 0, 0}, count[4] = {1, 1, 1, 1};
    int vid, nd, vt, isrec, natt, a;
    size_t n, i, ts;
    unsigned char* raw;

    if (v == NULL || ni <= 0 || nj <= 0)
        return ENKFIO_EARG;
    if (j1 <= 0 && j2 < 0) {
        j1 = 0;
        j2 = nj - 1;
    }
    if (j2 < j1 || j1 < 0 || j2 > nj - 1)
        return ENKFIO_EARG;
    TRY(ncf_open(path, &f));
    TRY(vid = ncf_inq_varid(f, varname));
    TRY(ncf_inq_var(f, vid, &vt, &nd, dl));
    isrec = f->vars[vid].isrec;
    if (vt == ENKFIO_CHAR || (isrec && dl[0] == 0))
        TRY(ENKFIO_EARG);
    if (isrec && r < 0) {
        if (dl[0] > 1)
            TRY(ENKFIO_EARG);   /* several records: which one must be given */
        r = 0;
    }
    /* the dimensions ncu_readfield_part() accepts (common/ncutils.c:99-293) */
    if (nd == 4) {
        if (r < 0) {
            if (dl[0] > 1)
                TRY(ENKFIO_EARG);
            r = 0;
        }
        if ((size_t) r >= dl[0] || dl[3] != (size_t) ni || dl[2] != (size_t) nj)
            TRY(ENKFIO_EARG);
        if (dl[1] != (size_t) nk) {
            if (dl[1] != 1)
                TRY(ENKFIO_EARG);
            k = 0;
        }
        start[0] = (size_t) r;
        start[1] = (size_t) k;
        start[2] = (size_t) j1;
        count[2] = (size_t) (j2 - j1 + 1);
        count[3] = dl[3];
    } else if (nd == 3) {
        if (dl[2] != (size_t) ni || dl[1] != (size_t) nj)
            TRY(ENKFIO_EARG);
        if (!isrec) {
            if (dl[0] != (size_t) nk && !(dl[0] == 1 && (k <= 0 || k == nk - 1)))
                TRY(ENKFIO_EARG);
            start[0] = (dl[0] == 1) ? 0 : (size_t) k;
        } else {
            if ((size_t) r >= dl[0])
                TRY(ENKFIO_EARG);
            start[0] = (size_t) r;   /* a layer of a record variable: k is ignored */
        }
        start[1] = (size_t) j1;
        count[1] = (size_t) (j2 - j1 + 1);
        count[2] = dl[2];
    } else if (nd == 2) {
        if (isrec || dl[1] != (size_t) ni || dl[0] != (size_t) nj)
            TRY(ENKFIO_EARG);
        start[0] = (size_t) j1;
        count[0] = (size_t) (j2 - j1 + 1);
        count[1] = dl[1];
    } else
        TRY(ENKFIO_EARG);
    n = (size_t) (j2 - j1 + 1) * (size_t) ni;
    TRY(ncf_get_vara_as(f, vid, start, count, ENKFIO_FLOAT, v));
    /* fill, missing and valid values by the raw values (ncutils.c:296-510) */
    ts = type_size(vt);
    raw = malloc(n * ts + 1);
    if (raw == NULL)
        TRY(ENKFIO_ENOMEM);
    if (ncf_get_vara(f, vid, start, count, raw) != ENKFIO_OK) {
        free(raw);
        TRY(ENKFIO_EIO);
    }
    {
        unsigned char fill[8], mv[8];
        int tt;
        size_t na;

        if (ncf_get_att(f, vid, "_FillValue", &tt, &na, NULL, 0) == ENKFIO_OK && na == 1 && tt == vt)
            ncf_get_att(f, vid, "_FillValue", NULL, NULL, fill, 1);
        else
            default_fill(vt, fill);
        for (i = 0; i < n; ++i)
            if (memcmp(raw + i * ts, fill, ts) == 0)
                v[i] = NAN;
        if (ncf_get_att(f, vid, "missing_value", &tt, &na, NULL, 0) == ENKFIO_OK && na == 1 && tt == vt) {
            ncf_get_att(f, vid, "missing_value", NULL, NULL, mv, 1);
            for (i = 0; i < n; ++i)
                if (memcmp(raw + i * ts, mv, ts) == 0)
                    v[i] = NAN;
        }
        {
            double lim[2];
            unsigned char lb[16];

            if (ncf_get_att(f, vid, "valid_min", &tt, &na, NULL, 0) == ENKFIO_OK && na == 1 && tt == vt) {
                ncf_get_att(f, vid, "valid_min", NULL, NULL, lb, 1);
                lim[0] = get_as_double(lb, vt, 0);
                for (i = 0; i < n; ++i)
                    if (get_as_double(raw, vt, i) < lim[0])
                        v[i] = NAN;
            }
            if (ncf_get_att(f, vid, "valid_max", &tt, &na, NULL, 0) == ENKFIO_OK && na == 1 && tt == vt) {
                ncf_get_att(f, vid, "valid_max", NULL, NULL, lb, 1);
                lim[1] = get_as_double(lb, vt, 0);
                for (i = 0; i < n; ++i)
                    if (get_as_double(raw, vt, i) > lim[1])
                        v[i] = NAN;
            }
            if (ncf_get_att(f, vid, "valid_range", &tt, &na, NULL, 0) == ENKFIO_OK && na == 2 && tt == vt) {
                ncf_get_att(f, vid, "valid_range", NULL, NULL, lb, 2);
                lim[0] = get_as_double(lb, vt, 0);
                lim[1] = get_as_double(lb, vt, 1);
                for (i = 0; i < n; ++i) {
                    const double x = get_as_double(raw, vt, i);

                    if (x < lim[0] || x > lim[1])
                        v[i] = NAN;
                }
            }
        }
    }
    free(raw);
    /* scale_factor, then add_offset, as floats (ncutils.c:513-531) */
    TRY(ncf_inq_natts(f, vid, &natt));
    for (a = 0; a < 2; ++a) {
        const char* an = a == 0 ? "scale_factor" : "add_offset";
        int tt;
        size_t na;
        double x[1];

        if (ncf_get_att(f, vid, an, &tt, &na, NULL, 0) != ENKFIO_OK || tt == ENKFIO_CHAR)
            continue;
        if (na != 1)
            TRY(ENKFIO_EFORMAT);
        {
            unsigned char b[8];
            float c;

            ncf_get_att(f, vid, an, NULL, NULL, b, 1);
            x[0] = get_as_double(b, tt, 0);
            c = (float) x[0];
            for (i = 0; i < n; ++i)
                v[i] = (a == 0) ? v[i] * c : v[i] + c;
        }
    }
    return ncf_close(f);
}

static int read_var_as(ncfile* f, const char* name, int type, size_t n, void** out)
{
    int vid = ncf_inq_varid(f, name);
    size_t dl[ENKFIO_MAXDIMS];
    int nd;

    if (vid < 0)
        return vid;
    if (ncf_inq_var(f, vid, NULL, &nd, dl) != ENKFIO_OK || nd != 1 || dl[0] != n)
        return ENKFIO_EFORMAT;
    *out = malloc(n * type_size(type) + 1);
    if (*out == NULL)
        return ENKFIO_ENOMEM;
    return ncf_get_vara_as(f, vid, NULL, NULL, type, *out);
}

void enkfio_free_obs(enkfio_obs* o)
{
    if (o == NULL)
        return;
    free(o->id);
    free(o->id_orig);
    free(o->status);
    free(o->type);
    free(o->product);
    free(o->instrument);
    free(o->fid);
    free(o->batch);
    free(o->value);
    free(o->estd);
    free(o->footprint);
    free(o->lon);
    free(o->lat);
    free(o->depth);
    free(o->model_depth);
    free(o->fi);
    free(o->fj);
    free(o->fi2);
    free(o->fk);
    free(o->time);
    free(o->aux);
    memset(o, 0, sizeof(*o));
}

int enkfio_read_obs(const char* path, enkfio_obs* o)
{
    ncfile* f;
    size_t n = 0, i;
    int vid, natt, k, r;

    if (o == NULL)
        return ENKFIO_EARG;
    memset(o, 0, sizeof(*o));
    o->da_day = NAN;
    TRY(ncf_open(path, &f));
    TRY(ncf_inq_dimlen(f, "nobs", &n));
    o->nobs = (long long) n;
    {
        int tt;
        size_t na;

        if (ncf_get_att(f, ENKFIO_GLOBAL, "DA_DAY", &tt, &na, NULL, 0) == ENKFIO_OK && na >= 1 && tt != ENKFIO_CHAR) {
            unsigned char b[8];

            ncf_get_att(f, ENKFIO_GLOBAL, "DA_DAY", NULL, NULL, b, 1);
            o->da_day = get_as_double(b, tt, 0);
        }
    }
    /* obs type names: the int attributes of "type" (observations.c:515-557) */
    TRY(vid = ncf_inq_varid(f, "type"));
    TRY(ncf_inq_natts(f, vid, &natt));
    for (k = 0; k < natt; ++k) {
        char name[ENKFIO_NAMELEN];
        int tt, id;
        size_t na;

        TRY(ncf_inq_attname(f, vid, k, name, sizeof name));
        if (ncf_get_att(f, vid, name, &tt, &na, NULL, 0) != ENKFIO_OK || tt != ENKFIO_INT || na != 1)
            continue;
        ncf_get_att(f, vid, name, NULL, NULL, &id, 1);
        if (id < 0 || id >= ENKFIO_MAXTYPES)
            continue;
        snprintf(o->typenames[id], ENKFIO_NAMELEN, "%s", name);
        if (id + 1 > o->ntypes)
            o->ntypes = id + 1;
    }
#undef TRY
#define TRY(x) do { int r_ = (x); if (r_ < 0) { ncf_close(f); enkfio_free_obs(o); return r_; } } while (0)
    if (n == 0)
        return ncf_close(f);
    TRY(read_var_as(f, "id", ENKFIO_INT, n, (void**) &o->id));
    TRY(read_var_as(f, "id_orig", ENKFIO_INT, n, (void**) &o->id_orig));
    TRY(read_var_as(f, "status", ENKFIO_SHORT, n, (void**) &o->status));
    TRY(read_var_as(f, "type", ENKFIO_SHORT, n, (void**) &o->type));
    TRY(read_var_as(f, "product", ENKFIO_SHORT, n, (void**) &o->product));
    TRY(read_var_as(f, "instrument", ENKFIO_SHORT, n, (void**) &o->instrument));
    TRY(read_var_as(f, "fid", ENKFIO_SHORT, n, (void**) &o->fid));
    TRY(read_var_as(f, "batch", ENKFIO_INT, n, (void**) &o->batch));
    TRY(read_var_as(f, "value", ENKFIO_FLOAT, n, (void**) &o->value));
    TRY(read_var_as(f, ncf_inq_varid(f, "estd_orig") >= 0 ? "estd_orig" : "estd", ENKFIO_FLOAT, n,
                    (void**) &o->estd));
    if (ncf_inq_varid(f, "footprint") >= 0)
        TRY(read_var_as(f, "footprint", ENKFIO_FLOAT, n, (void**) &o->footprint));
    TRY(read_var_as(f, "lon", ENKFIO_FLOAT, n, (void**) &o->lon));
    TRY(read_var_as(f, "lat", ENKFIO_FLOAT, n, (void**) &o->lat));
    TRY(read_var_as(f, "depth", ENKFIO_FLOAT, n, (void**) &o->depth));
    TRY(read_var_as(f, "model_depth", ENKFIO_FLOAT, n, (void**) &o->model_depth));
    if (ncf_inq_varid(f, "fi") >= 0) {
        TRY(read_var_as(f, "fi", ENKFIO_DOUBLE, n, (void**) &o->fi));
        TRY(read_var_as(f, "fj", ENKFIO_DOUBLE, n, (void**) &o->fj));
    } else {
        TRY(read_var_as(f, "fi0", ENKFIO_DOUBLE, n, (void**) &o->fi));
        TRY(read_var_as(f, "fi1", ENKFIO_DOUBLE, n, (void**) &o->fj));
        TRY(read_var_as(f, "fi2", ENKFIO_DOUBLE, n, (void**) &o->fi2));
    }
    TRY(read_var_as(f, "fk", ENKFIO_FLOAT, n, (void**) &o->fk));
    TRY(read_var_as(f, "time", ENKFIO_FLOAT, n, (void**) &o->time));
    TRY(read_var_as(f, "aux", ENKFIO_INT, n, (void**) &o->aux));
    /* a batch rejected by an earlier run is an OK observation again
     * (observations.c:786-796; STATUS_BADBATCH = 5, STATUS_OK = 0) */
    for (i = 0; i < n; ++i)
        if (o->status[i] == 5)
            o->status[i] = 0;
    r = ncf_close(f);
    if (r != ENKFIO_OK)
        enkfio_free_obs(o);
    return r;
}
