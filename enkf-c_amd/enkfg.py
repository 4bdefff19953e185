"""Python binding of the C ABI in include/enkfg.h (libenkfg.so, gfx950).

It mirrors the reference interface of the hot path: calc_transforms() is
das_calctransforms (calc/transforms.c:605) for one grid, update_fields() is
das_updatefields (update/update.c:92), update_bg() is das_updatebg
(update/update.c:568).  Errors raise EnkfgError carrying the engine's status
code and message -- the engine returns codes where the reference calls
enkf_quit().

There is no CPU fallback: if libenkfg.so is missing or no GPU is present the
calls fail loudly.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libenkfg.so")

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_LINALG, ERR_STATE, ERR_UNSUPPORTED, ERR_NONFINITE = range(8)
MODE_ENKF, MODE_ENOI, MODE_HYBRID = 1, 2, 3
SCHEME_DENKF, SCHEME_ETKF = 1, 2
PREC_FP64, PREC_MIXED = 0, 1
FLAG_PERTYPE_DIAG, FLAG_KEEP_FP64, FLAG_GRAPH = 1, 2, 4
UPD_OUTPUTINC, UPD_HYBRID_CONDITION = 32, 1024
NLOBSMAX_NONE = 2147483647
MAXLOCRAD = 8

EXPORTED = [
    "enkfg_create", "enkfg_set_kfactor", "enkfg_get_obs_estd", "enkfg_destroy", "enkfg_last_error", "enkfg_sync", "enkfg_set_obstypes", "enkfg_set_obs",
    "enkfg_set_obs_status", "enkfg_set_S", "enkfg_get_S", "enkfg_calc_transforms", "enkfg_get_transforms",
    "enkfg_get_diag", "enkfg_get_stats", "enkfg_localise", "enkfg_get_local_counts", "enkfg_get_localobs",
    "enkfg_set_transforms", "enkfg_update_fields", "enkfg_update_bg", "enkfg_ens_attach", "enkfg_ens_update",
    "enkfg_get_timings", "enkfg_get_tuning", "enkfg_adjoint", "enkfg_set_rand48", "enkfg_get_rand48",
    "enkfg_rand48_normals", "enkfg_update_HE", "enkfg_update_fields_noalloc", "enkfg_calc_HE",
    "enkfg_ens_spread", "enkfg_spread_fields", "enkfg_adjoint_ens",
]


class EnkfgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("enkfg error %d: %s" % (code, msg))
        self.code = code


class Config(C.Structure):
    _fields_ = [("mode", C.c_int), ("scheme", C.c_int), ("precision", C.c_int), ("nmem", C.c_int),
                ("nmem_dynamic", C.c_int), ("alpha", C.c_double), ("gamma", C.c_double), ("device", C.c_int),
                ("flags", C.c_int)]


class ObsType(C.Structure):
    _fields_ = [("nlocrad", C.c_int), ("locrad", C.c_double * MAXLOCRAD), ("locweight", C.c_double * MAXLOCRAD),
                ("rfactor", C.c_double), ("geographic", C.c_int), ("statsonly", C.c_int), ("nlobsmax", C.c_int)]


class Nodes(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("stride", C.c_int), ("geographic", C.c_int),
                ("lon", C.POINTER(C.c_double)), ("lat", C.POINTER(C.c_double)), ("visible", C.POINTER(C.c_ubyte))]


class Stats(C.Structure):
    _fields_ = [("nlobs_sum", C.c_longlong), ("nlobs_max", C.c_int), ("n_inv_obs", C.c_int), ("n_inv_ens", C.c_int),
                ("ncell", C.c_int)]


class Grid(C.Structure):
    _fields_ = [("mni", C.c_int), ("mnj", C.c_int), ("stride", C.c_int), ("periodic_i", C.c_int), ("surfk", C.c_int),
                ("nlevels", C.POINTER(C.c_int))]


class Hop(C.Structure):
    _fields_ = [("hfunction", C.c_int), ("nk", C.c_int), ("field", C.POINTER(C.c_int))]


H_SURF, H_SUBSURF = 0, 1


class Field(C.Structure):
    _fields_ = [("level", C.c_int), ("issurfacevar", C.c_int), ("inflation", C.c_float), ("inf_ratio", C.c_double),
                ("applylog", C.c_int), ("deflation", C.c_double), ("sigma", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libenkfg.so is not built (run __graft_entry__.build() or make -C enkf-c_amd)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.enkfg_create.argtypes = [C.POINTER(vp), C.POINTER(Config)]
        L.enkfg_destroy.argtypes = [vp]
        L.enkfg_destroy.restype = None
        L.enkfg_last_error.argtypes = [vp]
        L.enkfg_last_error.restype = C.c_char_p
        L.enkfg_sync.argtypes = [vp]
        L.enkfg_set_obstypes.argtypes = [vp, C.c_int, C.POINTER(ObsType)]
        L.enkfg_set_obs.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
        L.enkfg_set_obs_status.argtypes = [vp, vp]
        L.enkfg_set_S.argtypes = [vp, vp, vp, C.c_int]
        L.enkfg_get_S.argtypes = [vp, vp, vp]
        L.enkfg_calc_transforms.argtypes = [vp, C.POINTER(Nodes)]
        L.enkfg_get_transforms.argtypes = [vp, vp, vp, vp, vp]
        L.enkfg_get_diag.argtypes = [vp, vp, vp, vp]
        L.enkfg_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.enkfg_localise.argtypes = [vp, C.POINTER(Nodes)]
        L.enkfg_get_local_counts.argtypes = [vp, vp]
        L.enkfg_get_localobs.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        L.enkfg_set_transforms.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp]
        L.enkfg_update_fields.argtypes = [vp, C.POINTER(Grid), C.c_int, C.POINTER(Field), C.c_int, C.c_int,
                                          C.POINTER(vp), vp, C.c_int]
        L.enkfg_update_bg.argtypes = [vp, C.POINTER(Grid), C.c_int, C.POINTER(Field), C.c_int, C.c_int, C.POINTER(vp)]
        L.enkfg_ens_attach.argtypes = [vp, C.POINTER(Grid), C.c_int, C.POINTER(Field), vp]
        L.enkfg_ens_update.argtypes = [vp, C.c_int, vp]
        L.enkfg_get_timings.argtypes = [vp, C.POINTER(C.c_double), C.c_int]
        L.enkfg_get_tuning.argtypes = [vp, C.POINTER(C.c_int), C.c_int]
        L.enkfg_adjoint.argtypes = [vp, vp, vp, vp, C.POINTER(C.c_double)]
        L.enkfg_adjoint_ens.argtypes = [vp, vp, vp, vp, vp]
        L.enkfg_update_HE.argtypes = [vp, C.POINTER(Grid), C.c_int, vp, vp, vp, vp, vp, vp]
        L.enkfg_calc_HE.argtypes = [vp, C.POINTER(Grid), C.c_int, C.POINTER(Hop), vp, vp, vp, vp, vp, vp, vp]
        L.enkfg_update_fields_noalloc.argtypes = [vp, C.POINTER(Grid), C.c_int, C.POINTER(Field), C.c_int, C.c_int,
                                                  C.POINTER(vp), vp, C.c_int]
        L.enkfg_ens_spread.argtypes = [vp, C.c_int, C.c_int, vp]
        L.enkfg_spread_fields.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp), vp]
        L.enkfg_set_rand48.argtypes = [vp, C.c_ulonglong]
        L.enkfg_get_rand48.argtypes = [vp, C.POINTER(C.c_ulonglong)]
        L.enkfg_rand48_normals.argtypes = [C.POINTER(C.c_ulonglong), C.c_int, vp]
        L.enkfg_set_kfactor.argtypes = [vp, C.c_double]
        L.enkfg_get_obs_estd.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def make_obstypes(specs):
    arr = (ObsType * max(len(specs), 1))()
    for k, sp in enumerate(specs):
        lr = list(sp.get("locrad", [250.0]))
        lw = list(sp.get("locweight", [1.0] * len(lr)))
        s = sum(lw)
        lw = [x / s for x in lw]
        arr[k].nlocrad = len(lr)
        for i in range(len(lr)):
            arr[k].locrad[i] = lr[i]
            arr[k].locweight[i] = lw[i]
        arr[k].rfactor = sp.get("rfactor", 1.0)
        arr[k].geographic = sp.get("geographic", 1)
        arr[k].statsonly = sp.get("statsonly", 0)
        arr[k].nlobsmax = sp.get("nlobsmax", NLOBSMAX_NONE)
    return arr


def make_fields(fields):
    arr = (Field * max(len(fields), 1))()
    for k, f in enumerate(fields):
        arr[k].level = f.get("level", 0)
        arr[k].issurfacevar = f.get("issurfacevar", 0)
        arr[k].inflation = f.get("inflation", 1.0)
        arr[k].inf_ratio = f.get("inf_ratio", 1.0)
        arr[k].applylog = f.get("applylog", 0)
        arr[k].deflation = f.get("deflation", float("nan"))
        arr[k].sigma = f.get("sigma", 0.0)
    return arr


class Engine:
    """One enkfg context (one device, one stream)."""

    def __init__(self, nmem, nmem_dynamic=None, mode=MODE_ENKF, scheme=SCHEME_DENKF, alpha=1.0, gamma=1.0,
                 precision=PREC_FP64, device=0, pertype_diag=True, keep_fp64=False, graph=False):
        self.nmem = nmem
        self.md = nmem if nmem_dynamic is None else nmem_dynamic
        self.mode = mode
        self.scheme, self.alpha, self.gamma = scheme, alpha, gamma
        self.keep_fp64 = keep_fp64
        cfg = Config(mode, scheme, precision, nmem, self.md, alpha, gamma, device,
                     (FLAG_PERTYPE_DIAG if pertype_diag else 0) | (FLAG_KEEP_FP64 if keep_fp64 else 0)
                     | (FLAG_GRAPH if graph else 0))
        h = C.c_void_p()
        r = lib().enkfg_create(C.byref(h), C.byref(cfg))
        if r != OK:
            raise EnkfgError(r, "enkfg_create failed")
        self.h = h
        self.ntypes = 0
        self.nnode = 0
        self._keep = []

    def _check(self, r):
        if r != OK:
            raise EnkfgError(r, lib().enkfg_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None) is not None:
            lib().enkfg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- calc inputs
    def set_obstypes(self, specs):
        self._types = make_obstypes(specs)
        self.ntypes = len(specs)
        self._check(lib().enkfg_set_obstypes(self.h, len(specs), self._types))

    def set_obs(self, type_, status, lon, lat, estd):
        type_ = np.ascontiguousarray(type_, np.int16)
        status = np.ascontiguousarray(status, np.int16)
        lon = np.ascontiguousarray(lon, np.float32)
        lat = np.ascontiguousarray(lat, np.float32)
        estd = np.ascontiguousarray(estd, np.float32)
        self.nobs = len(type_)
        self._check(lib().enkfg_set_obs(self.h, self.nobs, _p(type_), _p(status), _p(lon), _p(lat), _p(estd)))

    def set_obs_status(self, status):
        status = np.ascontiguousarray(status, np.int16)
        self._check(lib().enkfg_set_obs_status(self.h, _p(status)))

    def set_S(self, S, s, standardised=True):
        S = np.ascontiguousarray(S, np.float32)
        s = np.ascontiguousarray(s, np.float64)
        self._check(lib().enkfg_set_S(self.h, _p(S), _p(s), int(standardised)))

    def calc_HE(self, nlevels, surfk, periodic_i, hops, fi, fj, fk, value, want=True):
        """das_getHE + das_calcinnandspread from the attached ensemble; the
        engine's S and s are set from the result.  hops: per obs type
        (hfunction, [field index of layer k]).  Returns (HA, s, std) host
        copies when want."""
        nlevels = np.ascontiguousarray(nlevels, np.int32)
        g = Grid(nlevels.shape[1], nlevels.shape[0], 1, periodic_i, surfk,
                 nlevels.ctypes.data_as(C.POINTER(C.c_int)))
        fl = [np.ascontiguousarray(f, np.int32) for _, f in hops]
        arr = (Hop * max(len(hops), 1))()
        for k, (h, f) in enumerate(hops):
            arr[k].hfunction = h
            arr[k].nk = len(f)
            arr[k].field = fl[k].ctypes.data_as(C.POINTER(C.c_int))
        fi = np.ascontiguousarray(fi, np.float64)
        fj = np.ascontiguousarray(fj, np.float64)
        fk = np.ascontiguousarray(fk, np.float32)
        value = np.ascontiguousarray(value, np.float64)
        HA = np.zeros((self.nmem, self.nobs), np.float32) if want else None
        s = np.zeros(self.nobs) if want else None
        sd = np.zeros(self.nobs) if want else None
        self._check(lib().enkfg_calc_HE(self.h, C.byref(g), len(hops), arr, _p(fi), _p(fj), _p(fk), _p(value),
                                        _p(HA), _p(s), _p(sd)))
        return HA, s, sd

    def set_kfactor(self, kfactor):
        """das_moderateobs in the next calc_HE (NaN: off)"""
        self._check(lib().enkfg_set_kfactor(self.h, C.c_double(kfactor)))

    def obs_estd(self):
        """the engine's obs errors and the per-type counts of obs whose error
        variance the moderation more than doubled"""
        e = np.zeros(max(self.nobs, 1), np.float32)
        n = np.zeros(max(self.ntypes, 1), np.int32)
        self._check(lib().enkfg_get_obs_estd(self.h, _p(e), _p(n)))
        return e[:self.nobs], n[:self.ntypes]

    def get_S(self):
        S = np.zeros((self.nmem, self.nobs), np.float32)
        s = np.zeros(self.nobs, np.float64)
        self._check(lib().enkfg_get_S(self.h, _p(S), _p(s)))
        return S, s

    def _nodes(self, lon, lat, stride=1, geographic=1, visible=None):
        lon = np.ascontiguousarray(lon, np.float64)
        lat = np.ascontiguousarray(lat, np.float64)
        nj, ni = lon.shape
        vis = None if visible is None else np.ascontiguousarray(visible, np.uint8)
        self._keep = [lon, lat, vis]
        self.nnode = ni * nj
        self.nj_s, self.ni_s = nj, ni
        return Nodes(ni, nj, stride, geographic, lon.ctypes.data_as(C.POINTER(C.c_double)),
                     lat.ctypes.data_as(C.POINTER(C.c_double)),
                     None if vis is None else vis.ctypes.data_as(C.POINTER(C.c_ubyte)))

    # ---- calc
    def calc_transforms(self, lon, lat, stride=1, geographic=1, visible=None):
        """lon, lat: [nj_s][ni_s] node coordinates."""
        nd = self._nodes(lon, lat, stride, geographic, visible)
        self._check(lib().enkfg_calc_transforms(self.h, C.byref(nd)))

    def localise(self, lon, lat, stride=1, geographic=1, visible=None):
        nd = self._nodes(lon, lat, stride, geographic, visible)
        self._check(lib().enkfg_localise(self.h, C.byref(nd)))

    def local_counts(self):
        c = np.zeros(self.nnode, np.int32)
        self._check(lib().enkfg_get_local_counts(self.h, _p(c)))
        return c

    def localobs(self, node, cap):
        ids = np.zeros(max(cap, 1), np.int32)
        lc = np.zeros(max(cap, 1), np.float64)
        self._check(lib().enkfg_get_localobs(self.h, int(node), int(cap), _p(ids), _p(lc)))
        return ids, lc

    def transforms(self, fp64=False):
        nn = self.nnode
        T = np.zeros((nn, self.md, self.nmem), np.float32) if self.mode != MODE_ENOI else None
        w = np.zeros((nn, self.nmem), np.float32)
        T64 = np.zeros((nn, self.md, self.nmem)) if (fp64 and T is not None) else None
        w64 = np.zeros((nn, self.nmem)) if fp64 else None
        self._check(lib().enkfg_get_transforms(self.h, _p(T), _p(w), _p(T64), _p(w64)))
        return T, w, T64, w64

    def diag(self):
        n = (self.ntypes + 1, self.nnode)
        a, b, c = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._check(lib().enkfg_get_diag(self.h, _p(a), _p(b), _p(c)))
        return a, b, c

    def stats(self):
        st = Stats()
        self._check(lib().enkfg_get_stats(self.h, C.byref(st)))
        return dict(nlobs_sum=st.nlobs_sum, nlobs_max=st.nlobs_max, n_inv_obs=st.n_inv_obs,
                    n_inv_ens=st.n_inv_ens, ncell=st.ncell)

    def timings(self):
        t = (C.c_double * 4)()
        self._check(lib().enkfg_get_timings(self.h, t, 4))
        return list(t)

    def he_ms(self):
        """kernel time of the last calc_HE"""
        t = (C.c_double * 7)()
        self._check(lib().enkfg_get_timings(self.h, t, 7))
        return t[6]

    def spread_ms(self):
        """kernel time of the last spread"""
        t = (C.c_double * 8)()
        self._check(lib().enkfg_get_timings(self.h, t, 8))
        return t[7]

    def ens_spread(self, f0, nfield, shape, dev_out=None):
        """das_writespread of attached fields f0 .. f0+nfield-1: float32
        [nfield][mnj][mni] (shape = (mnj, mni)); into dev_out (a device
        pointer) instead when given."""
        if dev_out is not None:
            self._check(lib().enkfg_ens_spread(self.h, f0, nfield, C.c_void_p(dev_out)))
            return None
        out = np.empty((nfield,) + tuple(shape), np.float32)
        self._check(lib().enkfg_ens_spread(self.h, f0, nfield, _p(out)))
        return out

    def spread_fields(self, buffers):
        """das_writespread of host field buffers [nmem][...] (float32,
        C-contiguous): float32 [nfield][...]"""
        for b in buffers:
            assert b.dtype == np.float32 and b.flags.c_contiguous and b.shape[0] == self.nmem
        nij = int(np.prod(buffers[0].shape[1:]))
        out = np.empty((len(buffers),) + buffers[0].shape[1:], np.float32)
        bp = (C.c_void_p * len(buffers))(*[b.ctypes.data for b in buffers])
        self._check(lib().enkfg_spread_fields(self.h, len(buffers), nij, bp, _p(out)))
        return out

    def adjoint(self, dev_dJ_dmean, nobs):
        """Gradients of J with respect to the observation values and the
        observation-error variances, given dJ/d(analysis mean) as a device
        pointer to float [nfield][ncell]; returns (dJ/dy, dJ/dR, dJ/dinfl)."""
        gy = np.zeros(nobs, np.float64)
        gR = np.zeros(nobs, np.float64)
        gi = C.c_double()
        self._check(lib().enkfg_adjoint(self.h, C.c_void_p(dev_dJ_dmean), _p(gy), _p(gR), C.byref(gi)))
        return gy, gR, gi.value

    def tuning(self):
        """Obs-count buckets of the calc batching: a list of dicts (lower
        bound of the largest (node, obstype) list, nodes at the last calc,
        Gram tile groups chosen for FP64 and mixed precision, 0 = not tuned)."""
        t = (C.c_int * 64)()
        self._check(lib().enkfg_get_tuning(self.h, t, 64))
        return [dict(lo=t[1 + 4 * q], nodes=t[2 + 4 * q], groups_fp64=t[3 + 4 * q], groups_mixed=t[4 + 4 * q])
                for q in range(t[0])]

    def launches(self):
        """Gram (and factor) launches of the last calc_transforms."""
        t = (C.c_double * 5)()
        self._check(lib().enkfg_get_timings(self.h, t, 5))
        return int(t[4])

    # ---- update
    def set_transforms(self, T, w, stride):
        w = np.ascontiguousarray(w, np.float32)
        T = None if T is None else np.ascontiguousarray(T, np.float32)
        nj_s, ni_s = w.shape[0], w.shape[1]
        self._check(lib().enkfg_set_transforms(self.h, nj_s, ni_s, stride, _p(T), _p(w)))

    @staticmethod
    def _grid(nlevels, stride, periodic_i, surfk):
        nlevels = np.ascontiguousarray(nlevels, np.int32)
        g = Grid(nlevels.shape[1], nlevels.shape[0], stride, periodic_i, surfk,
                 nlevels.ctypes.data_as(C.POINTER(C.c_int)))
        return g, nlevels

    def update_fields(self, nlevels, stride, periodic_i, surfk, fields, buffers, j1, j2, outputinc=False,
                      hybrid_condition=False, want_infl=False, noalloc=False):
        """buffers: float32 arrays [nmem][j2-j1+1][mni], updated in place;
        noalloc: through enkfg_update_fields_noalloc (no device allocation)."""
        g, keep = self._grid(nlevels, stride, periodic_i, surfk)
        fa = make_fields(fields)
        for b in buffers:
            assert b.dtype == np.float32 and b.flags.c_contiguous
        bp = (C.c_void_p * len(buffers))(*[b.ctypes.data for b in buffers])
        infl = np.zeros((len(fields), j2 - j1 + 1, g.mni), np.float32) if want_infl else None
        flags = (UPD_OUTPUTINC if outputinc else 0) | (UPD_HYBRID_CONDITION if hybrid_condition else 0)
        fn = lib().enkfg_update_fields_noalloc if noalloc else lib().enkfg_update_fields
        self._check(fn(self.h, C.byref(g), len(fields), fa, j1, j2, bp, _p(infl), flags))
        return infl

    def update_HE(self, mnj, mni, stride, periodic_i, fi, fj, status, inflation, inf_ratio, HE):
        """das_updateHE's update of the ensemble observations: HE [nmem][nobs]
        float32 (HE^f); returns HE^a (a new array)."""
        g = Grid(mni, mnj, stride, periodic_i, 0, None)
        out = np.ascontiguousarray(HE, np.float32).copy()
        nobs = out.shape[1]
        fi = np.ascontiguousarray(fi, np.float64)
        fj = np.ascontiguousarray(fj, np.float64)
        status = np.ascontiguousarray(status, np.int16)
        inflation = np.ascontiguousarray(np.broadcast_to(inflation, (nobs,)), np.float32)
        inf_ratio = np.ascontiguousarray(np.broadcast_to(inf_ratio, (nobs,)), np.float64)
        self._check(lib().enkfg_update_HE(self.h, C.byref(g), nobs, _p(fi), _p(fj), _p(status), _p(inflation),
                                          _p(inf_ratio), _p(out)))
        return out

    def adjoint_ens(self, dev_dJ_dXa, nobs, nfield):
        """Gradients of J with respect to the observation values, the
        observation-error variances and each field's inflation, given
        dJ/d(analysis ensemble) as a device pointer to float
        [nfield][ncell][nmem]; returns (dJ/dy, dJ/dR, dJ/dinfl[nfield])."""
        gy = np.zeros(nobs, np.float64)
        gR = np.zeros(nobs, np.float64)
        gi = np.zeros(nfield, np.float64)
        self._check(lib().enkfg_adjoint_ens(self.h, C.c_void_p(dev_dJ_dXa), _p(gy), _p(gR), _p(gi)))
        return gy, gR, gi

    def update_bg(self, nlevels, stride, periodic_i, surfk, fields, buffers, j1, j2):
        g, keep = self._grid(nlevels, stride, periodic_i, surfk)
        fa = make_fields(fields)
        bp = (C.c_void_p * len(buffers))(*[b.ctypes.data for b in buffers])
        self._check(lib().enkfg_update_bg(self.h, C.byref(g), len(fields), fa, j1, j2, bp))

    def ens_attach(self, nlevels, stride, periodic_i, surfk, fields, dev_ptr):
        g, keep = self._grid(nlevels, stride, periodic_i, surfk)
        self._gkeep = keep
        self._fa = make_fields(fields)
        self._check(lib().enkfg_ens_attach(self.h, C.byref(g), len(fields), self._fa, C.c_void_p(dev_ptr)))

    def ens_update(self, flags=0, dev_infl=None):
        self._check(lib().enkfg_ens_update(self.h, flags, None if dev_infl is None else C.c_void_p(dev_infl)))

    def sync(self):
        self._check(lib().enkfg_sync(self.h))

    # ---- RANDOMISE stream (drand48 state)
    def set_rand48(self, state):
        self._check(lib().enkfg_set_rand48(self.h, int(state)))

    def rand48(self):
        st = C.c_ulonglong()
        self._check(lib().enkfg_get_rand48(self.h, C.byref(st)))
        return st.value


def rand48_normals(state, npair):
    """The engine's normal pairs (host side, no device): returns (pairs as a
    flat array of 2 npair values, the advanced state)."""
    st = C.c_ulonglong(int(state))
    out = np.zeros(2 * max(npair, 0), np.float64)
    r = lib().enkfg_rand48_normals(C.byref(st), int(npair), _p(out))
    if r != OK:
        raise EnkfgError(r, "enkfg_rand48_normals failed")
    return out, st.value
