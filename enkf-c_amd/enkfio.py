"""ctypes binding of libenkfio.so (include/enkfio.h): netCDF classic files
and the file layouts of the calc/update path.

  NcReader             variables, dimensions and attributes of a classic file
  write_transforms     transforms.nc   (calc/transforms.c:54-135)
  inq_transforms       its dimensions and stride
  read_transforms      rows of T and w, as das_updatefields reads them
                       (update/update.c:195-224)
  write_diag           enkf_diag.nc    (calc/transforms.c:375-520)
  write_fields2d       spread.nc / inflation.nc  (diag/diags.c:79-168, 336-390)

The files are netCDF's 64-bit-offset format (what the reference writes with
NCFORMAT = 64BIT, common/enkfprm.c:490-498), which the netCDF library reads
whatever format its caller would write.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

OK, EIO, EFORMAT, EARG, ENOMEM, ERANGE = 0, -1, -2, -3, -4, -5
BYTE, CHAR, SHORT, INT, FLOAT, DOUBLE = 1, 2, 3, 4, 5, 6
GLOBAL = -1
MAXDIMS = 8
_DTYPE = {BYTE: np.int8, CHAR: np.uint8, SHORT: np.int16, INT: np.int32, FLOAT: np.float32, DOUBLE: np.float64}
_MSG = {EIO: "open, read or write failed", EFORMAT: "not a netCDF classic file, or a corrupt header",
        EARG: "invalid argument or unknown name", ENOMEM: "out of memory",
        ERANGE: "slab outside the variable, or a size the format cannot hold"}

EXPORTED = ["ncf_create", "ncf_def_dim", "ncf_def_var", "ncf_put_att", "ncf_put_att_text", "ncf_enddef",
            "ncf_put_slab", "ncf_open", "ncf_inq", "ncf_inq_varid", "ncf_inq_varname", "ncf_inq_var",
            "ncf_inq_dimlen", "ncf_get_att", "ncf_get_slab", "ncf_get_vara", "ncf_get_vara_as", "ncf_inq_natts",
            "ncf_inq_attname", "ncf_close", "enkfio_write_transforms", "enkfio_inq_transforms",
            "enkfio_read_transforms", "enkfio_write_diag", "enkfio_write_fields2d", "enkfio_field_nlevels",
            "enkfio_read_field", "enkfio_read_obs"]
MAXTYPES, NAMELEN = 64, 64


class ObsStruct(C.Structure):
    """enkfio_obs (include/enkfio.h)"""
    _fields_ = [("nobs", C.c_longlong), ("da_day", C.c_double), ("ntypes", C.c_int),
                ("typenames", (C.c_char * NAMELEN) * MAXTYPES)] + \
        [(n, C.POINTER(C.c_int)) for n in ("id", "id_orig", "batch", "aux")] + \
        [(n, C.POINTER(C.c_short)) for n in ("status", "type", "product", "instrument", "fid")] + \
        [(n, C.POINTER(C.c_float)) for n in ("value", "estd", "footprint", "lon", "lat", "depth", "model_depth",
                                             "fk", "time")] + \
        [(n, C.POINTER(C.c_double)) for n in ("fi", "fj", "fi2")]


class EnkfioError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s: %s (%d)" % (what, _MSG.get(code, "error"), code))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libenkfio.so")
        if not os.path.exists(path):
            raise RuntimeError("libenkfio.so is not built: run make -C enkf-c_amd")
        L = C.CDLL(path)
        vp, cp, sz, ip = C.c_void_p, C.c_char_p, C.c_size_t, C.c_int
        pp = C.POINTER(C.c_void_p)
        L.ncf_create.argtypes = [cp, pp]
        L.ncf_def_dim.argtypes = [vp, cp, sz]
        L.ncf_def_var.argtypes = [vp, cp, ip, ip, C.POINTER(C.c_int)]
        L.ncf_put_att.argtypes = [vp, ip, cp, ip, sz, vp]
        L.ncf_put_att_text.argtypes = [vp, ip, cp, cp]
        L.ncf_enddef.argtypes = [vp]
        L.ncf_put_slab.argtypes = [vp, ip, sz, sz, vp]
        L.ncf_open.argtypes = [cp, pp]
        L.ncf_inq.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.ncf_inq_varid.argtypes = [vp, cp]
        L.ncf_inq_varname.argtypes = [vp, ip, C.c_char_p, sz]
        L.ncf_inq_var.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
        L.ncf_inq_dimlen.argtypes = [vp, cp, C.POINTER(C.c_size_t)]
        L.ncf_get_att.argtypes = [vp, ip, cp, C.POINTER(C.c_int), C.POINTER(C.c_size_t), vp, sz]
        L.ncf_get_slab.argtypes = [vp, ip, sz, sz, vp]
        L.ncf_close.argtypes = [vp]
        L.enkfio_write_transforms.argtypes = [cp, ip, ip, ip, ip, ip, cp, vp, vp]
        L.enkfio_inq_transforms.argtypes = [cp] + [C.POINTER(C.c_int)] * 5
        L.enkfio_read_transforms.argtypes = [cp, ip, ip, vp, vp]
        L.enkfio_write_diag.argtypes = [cp, ip, C.POINTER(C.c_char_p), ip, ip, ip, vp, vp, vp]
        L.enkfio_write_fields2d.argtypes = [cp, ip, C.POINTER(C.c_char_p), ip, ip, vp]
        szp = C.POINTER(C.c_size_t)
        L.ncf_get_vara.argtypes = [vp, ip, szp, szp, vp]
        L.ncf_get_vara_as.argtypes = [vp, ip, szp, szp, ip, vp]
        L.ncf_inq_natts.argtypes = [vp, ip, C.POINTER(C.c_int)]
        L.ncf_inq_attname.argtypes = [vp, ip, ip, C.c_char_p, sz]
        L.enkfio_field_nlevels.argtypes = [cp, cp, C.POINTER(C.c_int)]
        L.enkfio_read_field.argtypes = [cp, cp, ip, ip, ip, ip, ip, ip, ip, vp]
        L.enkfio_read_obs.argtypes = [cp, C.POINTER(ObsStruct)]
        L.enkfio_free_obs.argtypes = [C.POINTER(ObsStruct)]
        L.enkfio_free_obs.restype = None
        for name in EXPORTED:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _chk(r, what):
    if r < 0:
        raise EnkfioError(r, what)
    return r


def _b(s):
    return s.encode() if isinstance(s, str) else s


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class NcReader:
    """A netCDF classic file opened for reading (ncf_open)."""

    def __init__(self, path):
        self.path = path
        self._h = None
        h = C.c_void_p()
        _chk(lib().ncf_open(_b(path), C.byref(h)), path)
        self._h = h

    def close(self):
        if self._h:
            lib().ncf_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def variables(self):
        nv = C.c_int()
        _chk(lib().ncf_inq(self._h, None, C.byref(nv), None), self.path)
        out = []
        buf = C.create_string_buffer(4096)
        for k in range(nv.value):
            _chk(lib().ncf_inq_varname(self._h, k, buf, 4096), self.path)
            out.append(buf.value.decode())
        return out

    def has(self, name):
        return lib().ncf_inq_varid(self._h, _b(name)) >= 0

    def dimlen(self, name):
        n = C.c_size_t()
        _chk(lib().ncf_inq_dimlen(self._h, _b(name), C.byref(n)), "%s: dimension %s" % (self.path, name))
        return n.value

    def shape(self, name):
        vid = _chk(lib().ncf_inq_varid(self._h, _b(name)), "%s: variable %s" % (self.path, name))
        t, nd = C.c_int(), C.c_int()
        dl = (C.c_size_t * MAXDIMS)()
        _chk(lib().ncf_inq_var(self._h, vid, C.byref(t), C.byref(nd), dl), self.path)
        return vid, t.value, tuple(dl[k] for k in range(nd.value))

    def read(self, name, start=0, count=None):
        """Values of a variable, along its first dimension start .. start+count-1."""
        vid, t, shp = self.shape(name)
        if len(shp) == 0:
            shp, first = (), 1
        else:
            first = shp[0]
        if count is None:
            count = first - start
        out = np.empty(((count,) + shp[1:]) if shp else (), _DTYPE[t])
        _chk(lib().ncf_get_slab(self._h, vid, start, count, _ptr(out)), "%s: %s" % (self.path, name))
        return out

    def attr(self, name, var=None):
        vid = GLOBAL if var is None else self.shape(var)[0]
        t, n = C.c_int(), C.c_size_t()
        _chk(lib().ncf_get_att(self._h, vid, _b(name), C.byref(t), C.byref(n), None, 0),
             "%s: attribute %s" % (self.path, name))
        out = np.empty(max(n.value, 1), _DTYPE[t.value])
        _chk(lib().ncf_get_att(self._h, vid, _b(name), None, None, _ptr(out), n.value), self.path)
        if t.value == CHAR:
            return bytes(out[:n.value]).decode()
        return out[:n.value]


def field_nlevels(path, varname):
    """ncu_getnlevels of a structured-grid variable"""
    n = C.c_int()
    _chk(lib().enkfio_field_nlevels(_b(path), _b(varname), C.byref(n)), "%s: %s" % (path, varname))
    return n.value


def read_field(path, varname, k, ni, nj, nk, r=-1, j1=0, j2=-1):
    """ncu_readfield_part: rows j1 .. j2 (all by default) of layer k as float32
    [rows][ni], fill / missing / out-of-range values NaN, scale and offset applied"""
    rows = nj if (j1 <= 0 and j2 < 0) else j2 - j1 + 1
    out = np.empty((rows, ni), np.float32)
    _chk(lib().enkfio_read_field(_b(path), _b(varname), r, k, j1, j2, ni, nj, nk, _ptr(out)),
         "%s: %s" % (path, varname))
    return out


def read_obs(path):
    """observations.nc as obs_read reads it: a dict of numpy arrays (copies),
    plus 'nobs', 'da_day' and 'typenames'"""
    o = ObsStruct()
    _chk(lib().enkfio_read_obs(_b(path), C.byref(o)), path)
    try:
        n = o.nobs
        out = dict(nobs=n, da_day=o.da_day,
                   typenames=[bytes(o.typenames[k].value).decode() for k in range(o.ntypes)])
        for name, typ in ObsStruct._fields_[4:]:
            p = getattr(o, name)
            out[name] = np.ctypeslib.as_array(p, (n,)).copy() if (p and n > 0) else None
        return out
    finally:
        lib().enkfio_free_obs(C.byref(o))


def write_transforms(path, w, T=None, stride=1, grid_name=""):
    """transforms.nc: w [nj][ni][m]; T [nj][ni][m_dyn][m] (None in EnOI mode)."""
    w = np.ascontiguousarray(w, np.float32)
    nj, ni, m = w.shape
    md = 0
    if T is not None:
        T = np.ascontiguousarray(T, np.float32)
        assert T.shape[:2] == (nj, ni) and T.shape[3] == m, (T.shape, w.shape)
        md = T.shape[2]
    _chk(lib().enkfio_write_transforms(_b(path), nj, ni, md, m, stride, _b(grid_name), _ptr(T), _ptr(w)), path)


def inq_transforms(path):
    v = [C.c_int() for _ in range(5)]
    _chk(lib().enkfio_inq_transforms(_b(path), *[C.byref(x) for x in v]), path)
    nj, ni, md, m, stride = (x.value for x in v)
    return dict(nj=nj, ni=ni, m_dyn=md, m=m, stride=stride)


def read_transforms(path, j0=0, nrow=None):
    """(T or None, w) for rows j0 .. j0+nrow-1."""
    d = inq_transforms(path)
    if nrow is None:
        nrow = d["nj"] - j0
    w = np.empty((nrow, d["ni"], d["m"]), np.float32)
    T = np.empty((nrow, d["ni"], d["m_dyn"], d["m"]), np.float32) if d["m_dyn"] else None
    _chk(lib().enkfio_read_transforms(_b(path), j0, nrow, _ptr(T), _ptr(w)), path)
    return T, w


def write_diag(path, typenames, stride, pnlobs, pdfs, psrf):
    """enkf_diag.nc from [nobstypes + 1][nj][ni] arrays (slot nobstypes = totals)."""
    pnlobs = np.ascontiguousarray(pnlobs, np.int32)
    pdfs = np.ascontiguousarray(pdfs, np.float32)
    psrf = np.ascontiguousarray(psrf, np.float32)
    nt = len(typenames)
    assert pnlobs.shape[0] == nt + 1 and pdfs.shape == pnlobs.shape == psrf.shape
    names = (C.c_char_p * max(nt, 1))(*[_b(s) for s in typenames])
    _chk(lib().enkfio_write_diag(_b(path), nt, names, pnlobs.shape[1], pnlobs.shape[2], stride, _ptr(pnlobs),
                                 _ptr(pdfs), _ptr(psrf)), path)


def write_fields2d(path, fields):
    """One float [nj][ni] variable per (name, array) of `fields` (a dict, in order)."""
    names = list(fields)
    v = np.ascontiguousarray(np.stack([np.asarray(fields[k], np.float32) for k in names]))
    arr = (C.c_char_p * len(names))(*[_b(s) for s in names])
    _chk(lib().enkfio_write_fields2d(_b(path), len(names), arr, v.shape[1], v.shape[2], _ptr(v)), path)


class NcWriter:
    """A netCDF classic file written variable by variable (ncf_create)."""

    def __init__(self, path):
        self.path = path
        h = C.c_void_p()
        _chk(lib().ncf_create(_b(path), C.byref(h)), path)
        self._h = h
        self._dims = {}
        self._vars = {}

    def dim(self, name, n):
        self._dims[name] = _chk(lib().ncf_def_dim(self._h, _b(name), n), "%s: dimension %s" % (self.path, name))
        return self._dims[name]

    def var(self, name, dtype, dims):
        t = {np.dtype(v): k for k, v in _DTYPE.items() if k != CHAR}[np.dtype(dtype)]
        ids = (C.c_int * max(len(dims), 1))(*[self._dims[d] for d in dims])
        self._vars[name] = (_chk(lib().ncf_def_var(self._h, _b(name), t, len(dims), ids), self.path), np.dtype(dtype))
        return self._vars[name][0]

    def attr(self, name, value, var=None):
        vid = GLOBAL if var is None else self._vars[var][0]
        if isinstance(value, str):
            _chk(lib().ncf_put_att_text(self._h, vid, _b(name), _b(value)), self.path)
            return
        a = np.atleast_1d(np.asarray(value))
        if a.dtype.kind in "iu":
            a = a.astype(np.int32)
        else:
            a = a.astype(np.float64) if a.dtype != np.float32 else a
        t = {np.dtype(np.int32): INT, np.dtype(np.float32): FLOAT, np.dtype(np.float64): DOUBLE}[a.dtype]
        a = np.ascontiguousarray(a)
        _chk(lib().ncf_put_att(self._h, vid, _b(name), t, a.size, _ptr(a)), self.path)

    def enddef(self):
        _chk(lib().ncf_enddef(self._h), self.path)

    def put(self, name, data, start=0):
        vid, dt = self._vars[name]
        a = np.ascontiguousarray(data, dt)
        n = a.shape[0] if a.ndim else 1
        _chk(lib().ncf_put_slab(self._h, vid, start, n, _ptr(a)), "%s: %s" % (self.path, name))

    def close(self):
        if self._h:
            r = lib().ncf_close(self._h)
            self._h = None
            _chk(r, self.path)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
