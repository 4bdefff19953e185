"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it was
built, for the reference kernels (oracle/_ref/libenkfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker.  The product library never
imports it.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libenkfref.so")

MODE_ENKF, MODE_ENOI, MODE_HYBRID = 1, 2, 3
SCHEME_DENKF, SCHEME_ETKF = 1, 2
NLOBSMAX_NONE = 2147483647
MAXLOCRAD = 8

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_sp = C.POINTER(C.c_short)
_up = C.POINTER(C.c_ubyte)


class ObsType(C.Structure):
    _fields_ = [("nlocrad", C.c_int), ("locrad", C.c_double * MAXLOCRAD),
                ("locweight", C.c_double * MAXLOCRAD), ("rfactor", C.c_double),
                ("geographic", C.c_int), ("statsonly", C.c_int), ("nlobsmax", C.c_int)]


class KdTree(C.Structure):
    _fields_ = [("ndim", C.c_int), ("n", C.c_int), ("coords", _dp), ("left", _ip),
                ("right", _ip), ("parent", _ip), ("dir", _up), ("data", _ip), ("depth", C.c_int)]


class Obs(C.Structure):
    _fields_ = [("nobs", C.c_int), ("ntypes", C.c_int), ("type", _sp), ("status", _sp),
                ("lon", _fp), ("lat", _fp), ("estd", _fp), ("types", C.POINTER(ObsType)),
                ("ntobs", _ip), ("obsids", C.POINTER(_ip)), ("trees", C.POINTER(KdTree))]


class CalcSpec(C.Structure):
    _fields_ = [("mode", C.c_int), ("scheme", C.c_int), ("nmem", C.c_int),
                ("nmem_dynamic", C.c_int), ("alpha", C.c_double)]


class CalcStats(C.Structure):
    _fields_ = [("nlobs_sum", C.c_longlong), ("nlobs_max", C.c_int), ("n_inv_obs", C.c_int),
                ("n_inv_ens", C.c_int), ("ncell", C.c_int)]


class Field(C.Structure):
    _fields_ = [("level", C.c_int), ("issurfacevar", C.c_int), ("inflation", C.c_float),
                ("inf_ratio", C.c_double), ("applylog", C.c_int)]


class UpdateSpec(C.Structure):
    _fields_ = [("mode", C.c_int), ("nmem", C.c_int), ("nmem_dynamic", C.c_int),
                ("mni", C.c_int), ("mnj", C.c_int), ("stride", C.c_int), ("periodic_i", C.c_int),
                ("surfk", C.c_int), ("nlevels", _ip), ("outputinc", C.c_int),
                ("ni_s", C.c_int), ("nj_s", C.c_int), ("T", _fp), ("w", _fp)]


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.or_taper_gc.restype = C.c_double
        L.or_taper_gc.argtypes = [C.c_double]
        L.or_calclcoeff.restype = C.c_double
        L.or_calclcoeff.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.or_getmaxlocrad.restype = C.c_double
        L.or_ll2xyz.argtypes = [_dp, _dp]
        L.or_kd_build.argtypes = [C.POINTER(KdTree), C.c_int, C.c_int, _dp]
        L.or_kd_free.argtypes = [C.POINTER(KdTree)]
        L.or_kd_range.argtypes = [C.POINTER(KdTree), _dp, C.c_double, C.c_int, C.c_int, _ip, _dp]
        L.or_obs_build_trees.argtypes = [C.POINTER(Obs), _sp]
        L.or_obs_free_trees.argtypes = [C.POINTER(Obs)]
        L.or_findlocal.argtypes = [C.POINTER(Obs), C.c_double, C.c_double, C.c_int, _up, C.c_int, _ip, _dp]
        L.or_standardise.argtypes = [C.c_int, C.c_int, _sp, _fp, C.POINTER(ObsType), _fp, _dp]
        L.or_traceprod.restype = C.c_double
        L.or_calc_G.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.or_calc_w.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        for fn in (L.or_calc_wT_denkf, L.or_calc_wT_etkf):
            fn.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_calc_transforms.argtypes = [C.POINTER(CalcSpec), C.POINTER(Obs), _fp, _dp, C.c_int, _dp, _dp,
                                         C.c_int, _up, _fp, _fp, _dp, _dp, _ip, _fp, _fp,
                                         C.POINTER(CalcStats), C.c_int, C.c_int]
        L.or_update_fields.argtypes = [C.POINTER(UpdateSpec), C.c_int, C.POINTER(Field), C.c_int, C.c_int,
                                       C.POINTER(_fp), _fp]
        L.or_update_bg.argtypes = [C.POINTER(UpdateSpec), C.c_int, C.POINTER(Field), C.c_int, C.c_int,
                                   C.POINTER(_fp)]
        L.or_sethybridensemble.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _fp, _dp]
        L.or_update_HE.argtypes = [C.POINTER(UpdateSpec), C.c_int, _dp, _dp, _sp, _fp, _dp, _fp]
        L.or_interp2d.restype = C.c_float
        L.or_interp2d.argtypes = [_dp, C.c_int, C.c_int, _fp, _ip, C.c_int]
        L.or_interp3d.restype = C.c_float
        L.or_interp3d.argtypes = [_dp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_fp), _ip, C.c_int]
        L.or_getHE.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, _ip,
                               C.POINTER(_fp), C.c_int, _sp, _sp, _dp, _dp, _fp, _dp, _ip, _ip, C.POINTER(_ip), _fp,
                               _dp, _dp]
        L.or_spread.argtypes = [C.c_int, C.c_longlong, C.POINTER(_fp), _fp]
        L.or_drand48.restype = C.c_double
        L.or_drand48.argtypes = [C.POINTER(C.c_ulonglong)]
        L.or_get_normalpair.argtypes = [C.POINTER(C.c_ulonglong), _dp]
        L.or_addnoise.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_ulonglong), C.c_longlong,
                                  C.POINTER(_fp)]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference kernels (calc/calcs.c + common/kdtree.c, unmodified)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref not built (needs the reference tree and MKL): make -C oracle ref")
        # one thread, sequential MKL: threaded MKL sharing a process with
        # torch's OpenMP runtime gives wrong products (the non-SPD I + S^T S
        # that once made dpotrf fail with info = 7 in the CPU baseline)
        os.environ["MKL_THREADING_LAYER"] = "SEQUENTIAL"
        os.environ["MKL_NUM_THREADS"] = "1"
        R = C.CDLL(REF_PATH)
        R.ref_last_error.restype = C.c_char_p
        R.ref_last_error.argtypes = []
        R.ref_local_analysis.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                         _dp, _dp, _dp]
        R.ref_kd_build.restype = C.c_void_p
        R.ref_kd_build.argtypes = [C.c_int, C.c_int, _dp]
        R.ref_kd_destroy.argtypes = [C.c_void_p]
        R.ref_kd_size.argtypes = [C.c_void_p]
        R.ref_kd_range.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int, C.c_int, _ip, _dp]
        R.ref_trees_create.restype = C.c_void_p
        R.ref_trees_create.argtypes = [C.POINTER(Obs), _sp]
        R.ref_trees_destroy.argtypes = [C.c_void_p]
        R.ref_calc_transforms.argtypes = [C.POINTER(CalcSpec), C.POINTER(Obs), C.c_void_p, _fp, _dp, C.c_int,
                                          _dp, _dp, C.c_int, _up, _fp, _fp, _dp, _dp, _ip, _fp, _fp,
                                          C.POINTER(CalcStats), C.c_int, C.c_int]
        _ref = R
    return _ref


# ---------------------------------------------------------------- geometry
def ll2xyz(lon, lat):
    out = (C.c_double * 3)()
    ll = (C.c_double * 2)(lon, lat)
    lib().or_ll2xyz(ll, out)
    return np.array(out[:])


def taper_gc(x):
    return lib().or_taper_gc(float(x))


def calclcoeff(locrad, locweight, dist):
    r = np.ascontiguousarray(locrad, np.float64)
    w = np.ascontiguousarray(locweight, np.float64)
    return lib().or_calclcoeff(len(r), _ptr(r, _dp), _ptr(w, _dp), float(dist))


# ---------------------------------------------------------------- kd-tree
class OracleKd:
    def __init__(self, coords):
        coords = np.ascontiguousarray(coords, np.float64)
        self.ndim = coords.shape[1]
        self.t = KdTree()
        r = lib().or_kd_build(C.byref(self.t), self.ndim, coords.shape[0], _ptr(coords, _dp))
        if r != 0:
            raise RuntimeError("or_kd_build: %d" % r)

    def range(self, q, rng, ordered=False):
        n = max(self.t.n, 1)
        data = np.zeros(n, np.int32)
        d2 = np.zeros(n, np.float64)
        q = np.ascontiguousarray(q, np.float64)
        k = lib().or_kd_range(C.byref(self.t), _ptr(q, _dp), float(rng), int(ordered), n,
                              _ptr(data, _ip), _ptr(d2, _dp))
        if k < 0:
            raise RuntimeError("or_kd_range: %d" % k)
        return data[:k].copy(), d2[:k].copy()

    def arrays(self):
        n = self.t.n
        return dict(left=np.ctypeslib.as_array(self.t.left, (n,)).copy(),
                    right=np.ctypeslib.as_array(self.t.right, (n,)).copy(),
                    data=np.ctypeslib.as_array(self.t.data, (n,)).copy(),
                    depth=self.t.depth)

    def __del__(self):
        try:
            lib().or_kd_free(C.byref(self.t))
        except Exception:
            pass


class RefKd:
    def __init__(self, coords):
        coords = np.ascontiguousarray(coords, np.float64)
        self.ndim = coords.shape[1]
        self.n = coords.shape[0]
        self._coords = coords
        self.h = ref().ref_kd_build(self.ndim, self.n, _ptr(coords, _dp))

    def range(self, q, rng, ordered=False):
        n = max(self.n, 1)
        data = np.zeros(n, np.int32)
        d2 = np.zeros(n, np.float64)
        q = np.ascontiguousarray(q, np.float64)
        k = ref().ref_kd_range(self.h, _ptr(q, _dp), float(rng), int(ordered), n, _ptr(data, _ip),
                               _ptr(d2, _dp))
        return data[:k].copy(), d2[:k].copy()

    def __del__(self):
        try:
            ref().ref_kd_destroy(self.h)
        except Exception:
            pass


# ---------------------------------------------------------------- obstypes / obs
def make_obstypes(specs):
    """specs: list of dicts with locrad (list), locweight (list, normalised
    here as obstypes.c:533-541 does), rfactor, geographic, statsonly, nlobsmax."""
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


class OracleObs:
    """Observations plus per-type kd-trees, as obs_createkdtrees builds them."""

    def __init__(self, otypes, type_, lon, lat, estd, status=None, status_at_build=None, build_trees=True):
        self.types = make_obstypes(otypes)
        self.ntypes = len(otypes)
        self.type = np.ascontiguousarray(type_, np.int16)
        self.lon = np.ascontiguousarray(lon, np.float32)
        self.lat = np.ascontiguousarray(lat, np.float32)
        self.estd = np.ascontiguousarray(estd, np.float32)
        n = len(self.type)
        self.status = np.ascontiguousarray(status if status is not None else np.zeros(n), np.int16)
        self.status_build = np.ascontiguousarray(
            status_at_build if status_at_build is not None else self.status, np.int16)
        self.o = Obs(n, self.ntypes, _ptr(self.type, _sp), _ptr(self.status, _sp), _ptr(self.lon, _fp),
                     _ptr(self.lat, _fp), _ptr(self.estd, _fp), self.types, None, None, None)
        if build_trees:
            r = lib().or_obs_build_trees(C.byref(self.o), _ptr(self.status_build, _sp))
            if r != 0:
                raise RuntimeError("or_obs_build_trees: %d" % r)

    def findlocal(self, lon, lat, geographic=1, visible=None):
        cap = max(len(self.type), 1)
        ids = np.zeros(cap, np.int32)
        lc = np.zeros(cap, np.float64)
        vis = None if visible is None else np.ascontiguousarray(visible, np.uint8)
        p = lib().or_findlocal(C.byref(self.o), float(lon), float(lat), int(geographic), _ptr(vis, _up), cap,
                               _ptr(ids, _ip), _ptr(lc, _dp))
        if p < 0:
            raise RuntimeError("or_findlocal: %d" % p)
        return ids[:p].copy(), lc[:p].copy()

    def __del__(self):
        try:
            lib().or_obs_free_trees(C.byref(self.o))
        except Exception:
            pass


def moderateobs(kfactor, ntypes, type_, status, estd, std, inn):
    """das_moderateobs (calc/ensobs.c:749-799): for every obs with status OK,
    estd' = sqrt(sqrt((svar + ovar)^2 + svar inn^2 / K / K) - svar) in double,
    left to right with every operation rounded (numpy does not contract),
    stored as float; counts per type the obs with svar > 0 whose error
    variance more than doubled.  NaN K: nothing changes (ensobs.c:758-759)."""
    e32 = np.array(estd, np.float32)
    cnt = np.zeros(ntypes, np.int64)
    if not np.isfinite(kfactor):
        return e32, cnt
    std = np.asarray(std, np.float64)
    inn = np.asarray(inn, np.float64)
    ov = e32.astype(np.float64)
    svar = std * std
    ovar = ov * ov
    sm = svar + ovar
    with np.errstate(invalid="ignore"):
        new = np.sqrt(np.sqrt(sm * sm + svar * inn * inn / kfactor / kfactor) - svar)
    ok = np.asarray(status) == 0
    mod = ok & (svar > 0.0) & (new * new / ovar > 2.0)
    np.add.at(cnt, np.asarray(type_)[mod], 1)
    e32[ok] = new[ok].astype(np.float32)
    return e32, cnt


def standardise(nmem, obs, S, s):
    S = np.ascontiguousarray(S, np.float32).copy()
    s = np.ascontiguousarray(s, np.float64).copy()
    lib().or_standardise(nmem, len(obs.type), _ptr(obs.type, _sp), _ptr(obs.estd, _fp), obs.types,
                         _ptr(S, _fp), _ptr(s, _dp))
    return S, s


# ---------------------------------------------------------------- per-column algebra
def local_analysis(mode, scheme, m, mout, S, s, use_ref=False):
    """S: (p, m) array (row o, column e); returns T (mout, m) as C [e1][e2],
    w (m), dfs, srf (double, ensemble-space as the reference computes)."""
    S = np.asarray(S, np.float64)
    p = S.shape[0]
    Scm = np.ascontiguousarray(S.T)   # column-major p x m == C [m][p]
    s = np.ascontiguousarray(s, np.float64)
    T = np.zeros((mout, m))
    w = np.zeros(m)
    G = np.zeros((p, m))              # column-major m x p == C [p][m]
    if use_ref:
        dfs = C.c_double()
        srf = C.c_double()
        r = ref().ref_local_analysis(mode, scheme, m, mout, p, _ptr(Scm, _dp), _ptr(s, _dp), _ptr(T, _dp),
                                     _ptr(w, _dp), C.byref(dfs), C.byref(srf), _ptr(G, _dp))
        if r != 0:
            raise RuntimeError("reference local analysis failed: %s" % ref().ref_last_error().decode())
        return T, w, dfs.value, srf.value
    L = lib()
    if mode == MODE_ENOI:
        r = L.or_calc_G(m, p, _ptr(Scm, _dp), _ptr(G, _dp))
        L.or_calc_w(m, p, _ptr(G, _dp), _ptr(s, _dp), _ptr(w, _dp))
        T = None
    elif scheme == SCHEME_DENKF:
        r = L.or_calc_wT_denkf(m, mout, p, _ptr(s, _dp), _ptr(Scm, _dp), _ptr(G, _dp), _ptr(w, _dp),
                               _ptr(T, _dp))
    else:
        r = L.or_calc_wT_etkf(m, mout, p, _ptr(s, _dp), _ptr(Scm, _dp), _ptr(G, _dp), _ptr(w, _dp),
                              _ptr(T, _dp))
    if r != 0:
        raise RuntimeError("oracle local analysis failed: %d" % r)
    L.or_traceprod.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
    dfs = L.or_traceprod(0, 0, p, m, _ptr(G, _dp), _ptr(Scm, _dp))
    srf = math.sqrt(L.or_traceprod(0, 1, m, p, _ptr(Scm, _dp), _ptr(Scm, _dp)) / dfs) - 1.0
    return T, w, dfs, srf


# ---------------------------------------------------------------- calc loop
class RefTrees:
    """The reference's per-type kd-trees (kd_insertnode, sequential)."""

    def __init__(self, obs):
        self.h = ref().ref_trees_create(C.byref(obs.o), _ptr(obs.status_build, _sp))

    def close(self):
        if self.h:
            ref().ref_trees_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calc_transforms(mode, scheme, nmem, nmem_dynamic, alpha, obs, S, s, lon, lat, geographic=1,
                    visible=None, node0=0, nnode_do=-1, use_ref=False, want_T=True, threads=1):
    """Runs the das_calctransforms loop.  S: [nmem][nobs] float32 (already
    standardised), s: [nobs] float64, lon/lat: node coordinates [nnode].

    threads > 1 splits the nodes [node0, node0 + nnode_do) into contiguous
    runs, one per thread (the C loops release the GIL; the oracle is
    reentrant; with use_ref, each thread gets its own reference trees, since
    kd_findnodeswithinrange keeps its results in the tree; use_ref may be a
    list of at least `threads` RefTrees to reuse).  Outputs are per node, so
    the runs write disjoint parts; the stats are summed."""
    import threading

    lon = np.ascontiguousarray(lon, np.float64)
    lat = np.ascontiguousarray(lat, np.float64)
    nnode = len(lon)
    S = np.ascontiguousarray(S, np.float32)
    s = np.ascontiguousarray(s, np.float64)
    md = nmem_dynamic
    nt = obs.ntypes
    dotT = mode in (MODE_ENKF, MODE_HYBRID) and want_T
    T = np.zeros((nnode, md, nmem), np.float32) if dotT else None
    T64 = np.zeros((nnode, md, nmem), np.float64) if dotT else None
    w = np.zeros((nnode, nmem), np.float32)
    w64 = np.zeros((nnode, nmem), np.float64)
    pnlobs = np.zeros((nt + 1, nnode), np.int32)
    pdfs = np.zeros((nt + 1, nnode), np.float32)
    psrf = np.zeros((nt + 1, nnode), np.float32)
    spec = CalcSpec(mode, scheme, nmem, md, alpha)
    vis = None if visible is None else np.ascontiguousarray(visible, np.uint8)
    if nnode_do < 0:
        nnode_do = nnode - node0
    nth = max(1, min(int(threads), nnode_do))
    bounds = [node0 + (nnode_do * k) // nth for k in range(nth + 1)]
    results = [None] * nth
    trees = [None] * nth
    if use_ref:
        if isinstance(use_ref, RefTrees) and nth == 1:
            trees[0] = (use_ref, False)
        elif isinstance(use_ref, (list, tuple)) and len(use_ref) >= nth:
            for k in range(nth):
                trees[k] = (use_ref[k], False)
        else:
            def mk(k):
                trees[k] = (RefTrees(obs), True)
            ths = [threading.Thread(target=mk, args=(k,)) for k in range(nth)]
            for t in ths:
                t.start()
            for t in ths:
                t.join()

    def work(k):
        st = CalcStats()
        a, b = bounds[k], bounds[k + 1]
        if use_ref:
            r = ref().ref_calc_transforms(C.byref(spec), C.byref(obs.o), trees[k][0].h, _ptr(S, _fp),
                                          _ptr(s, _dp), nnode, _ptr(lon, _dp), _ptr(lat, _dp), int(geographic),
                                          _ptr(vis, _up), _ptr(T, _fp), _ptr(w, _fp), _ptr(T64, _dp),
                                          _ptr(w64, _dp), _ptr(pnlobs, _ip), _ptr(pdfs, _fp), _ptr(psrf, _fp),
                                          C.byref(st), a, b - a)
            msg = ref().ref_last_error().decode() if r != 0 else ""
        else:
            r = lib().or_calc_transforms(C.byref(spec), C.byref(obs.o), _ptr(S, _fp), _ptr(s, _dp), nnode,
                                         _ptr(lon, _dp), _ptr(lat, _dp), int(geographic), _ptr(vis, _up),
                                         _ptr(T, _fp), _ptr(w, _fp), _ptr(T64, _dp), _ptr(w64, _dp),
                                         _ptr(pnlobs, _ip), _ptr(pdfs, _fp), _ptr(psrf, _fp), C.byref(st),
                                         a, b - a)
            msg = ""
        results[k] = (r, msg, st)

    if nth == 1:
        work(0)
    else:
        ths = [threading.Thread(target=work, args=(k,)) for k in range(nth)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
    for tr in trees:
        if tr is not None and tr[1]:
            tr[0].close()
    for r, msg, _ in results:
        if r != 0:
            raise RuntimeError("calc_transforms failed: %d %s" % (r, msg))
    stats = dict(nlobs_sum=0, nlobs_max=0, n_inv_obs=0, n_inv_ens=0, ncell=0)
    for _, _, st in results:
        stats["nlobs_sum"] += st.nlobs_sum
        stats["nlobs_max"] = max(stats["nlobs_max"], st.nlobs_max)
        stats["n_inv_obs"] += st.n_inv_obs
        stats["n_inv_ens"] += st.n_inv_ens
        stats["ncell"] += st.ncell
    return dict(T=T, w=w, T64=T64, w64=w64, pnlobs=pnlobs, pdfs=pdfs, psrf=psrf, stats=stats)


# ---------------------------------------------------------------- update
def _fields(fields):
    arr = (Field * max(len(fields), 1))()
    for k, f in enumerate(fields):
        arr[k].level = f.get("level", 0)
        arr[k].issurfacevar = f.get("issurfacevar", 0)
        arr[k].inflation = f.get("inflation", 1.0)
        arr[k].inf_ratio = f.get("inf_ratio", 1.0)
        arr[k].applylog = f.get("applylog", 0)
    return arr


def update_fields(mode, nmem, nmem_dynamic, stride, periodic_i, surfk, nlevels, T, w, fields, buffers,
                  j1, j2, outputinc=0, want_infl=False):
    """buffers: list of float32 arrays [nmem][jsize][mni], updated in place."""
    nlevels = np.ascontiguousarray(nlevels, np.int32)
    mnj, mni = nlevels.shape
    T = np.ascontiguousarray(T, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    nj_s, ni_s = w.shape[0], w.shape[1]
    sp = UpdateSpec(mode, nmem, nmem_dynamic, mni, mnj, stride, periodic_i, surfk, _ptr(nlevels, _ip),
                    outputinc, ni_s, nj_s, _ptr(T, _fp), _ptr(w, _fp))
    fa = _fields(fields)
    bufs = (_fp * len(buffers))(*[_ptr(b, _fp) for b in buffers])
    infl = np.zeros((len(fields), j2 - j1 + 1, mni), np.float32) if want_infl else None
    r = lib().or_update_fields(C.byref(sp), len(fields), fa, j1, j2, bufs, _ptr(infl, _fp))
    if r != 0:
        raise RuntimeError("or_update_fields: %d" % r)
    return infl


def update_bg(nmem, stride, periodic_i, surfk, nlevels, w, fields, buffers, j1, j2):
    nlevels = np.ascontiguousarray(nlevels, np.int32)
    mnj, mni = nlevels.shape
    w = np.ascontiguousarray(w, np.float32)
    nj_s, ni_s = w.shape[0], w.shape[1]
    sp = UpdateSpec(MODE_ENOI, nmem, nmem, mni, mnj, stride, periodic_i, surfk, _ptr(nlevels, _ip), 0,
                    ni_s, nj_s, None, _ptr(w, _fp))
    fa = _fields(fields)
    bufs = (_fp * len(buffers))(*[_ptr(b, _fp) for b in buffers])
    r = lib().or_update_bg(C.byref(sp), len(fields), fa, j1, j2, bufs)
    if r != 0:
        raise RuntimeError("or_update_bg: %d" % r)


def sethybridensemble(nmem, nmem_dynamic, gamma, v):
    """v: float32 [nmem][nij] updated in place."""
    nij = v.shape[1]
    vmean = np.zeros(nij, np.float64)
    lib().or_sethybridensemble(nmem, nmem_dynamic, float(gamma), nij, _ptr(v, _fp), _ptr(vmean, _dp))


def spread(v):
    """das_writespread's spread of v: float32 [nmem][...] -> float32 [...]"""
    v = np.ascontiguousarray(v, np.float32)
    nmem = v.shape[0]
    n = int(np.prod(v.shape[1:]))
    out = np.empty(v.shape[1:], np.float32)
    rows = (_fp * nmem)(*[v[e].ctypes.data_as(_fp) for e in range(nmem)])
    lib().or_spread(nmem, n, rows, _ptr(out, _fp))
    return out


def normalpairs(state, npair):
    """get_normalpair npair times from the drand48 state; returns (values,
    new state)"""
    st = C.c_ulonglong(int(state))
    out = np.zeros(2 * npair)
    for k in range(npair):
        lib().or_get_normalpair(C.byref(st), _ptr(out[2 * k:], _dp))
    return out, st.value


def addnoise(nmem, deflation, sigma, state, v):
    """addnoise on v: float32 [nmem][...] updated in place; returns the new
    drand48 state"""
    st = C.c_ulonglong(int(state))
    n = int(np.prod(v.shape[1:]))
    rows = (_fp * nmem)(*[v[e].ctypes.data_as(_fp) for e in range(nmem)])
    lib().or_addnoise(nmem, float(deflation), float(sigma), C.byref(st), n, rows)
    return st.value


def update_HE(mode, nmem, nmem_dynamic, stride, periodic_i, mnj, mni, T, w, fi, fj, status, inflation0, inf_ratio,
              HE):
    """update_HE_structured: returns HE^a ([nmem][nobs] float32) from HE^f;
    T [nj_s][ni_s][md][m], w [nj_s][ni_s][m]; per-obs inflation0, inf_ratio"""
    T = np.ascontiguousarray(T, np.float32)
    w = np.ascontiguousarray(w, np.float32)
    nj_s, ni_s = w.shape[0], w.shape[1]
    nl = np.ones((mnj, mni), np.int32)
    sp = UpdateSpec(mode, nmem, nmem_dynamic, mni, mnj, stride, periodic_i, 0, _ptr(nl, _ip), 0, ni_s, nj_s,
                    _ptr(T, _fp), _ptr(w, _fp))
    out = np.ascontiguousarray(HE, np.float32).copy()
    nobs = out.shape[1]
    fi = np.ascontiguousarray(fi, np.float64)
    fj = np.ascontiguousarray(fj, np.float64)
    status = np.ascontiguousarray(status, np.int16)
    inflation0 = np.ascontiguousarray(inflation0, np.float32)
    inf_ratio = np.ascontiguousarray(inf_ratio, np.float64)
    r = lib().or_update_HE(C.byref(sp), nobs, _ptr(fi, _dp), _ptr(fj, _dp), _ptr(status, _sp), _ptr(inflation0, _fp),
                           _ptr(inf_ratio, _dp), _ptr(out, _fp))
    if r != 0:
        raise RuntimeError("or_update_HE: %d" % r)
    return out

def getHE(mode, nmem, nmem_dynamic, gamma, nlevels, ktop, periodic_i, fields, otypes_h, type_, status, fi, fj, fk,
          value):
    """das_getHE with the standard H functions + hybrid scaling +
    das_calcinnandspread.  fields: list of float32 [nmem][nj][ni] (one per
    model field); otypes_h: per obs type (hkind, [field index of layer k]).
    Returns (HA [nmem][nobs] float32, s [nobs], std [nobs])."""
    nlevels = np.ascontiguousarray(nlevels, np.int32)
    nj, ni = nlevels.shape
    fields = [np.ascontiguousarray(f, np.float32) for f in fields]
    E = (_fp * len(fields))(*[_ptr(f, _fp) for f in fields])
    nt = len(otypes_h)
    hk = np.array([h for h, _ in otypes_h], np.int32)
    nk = np.array([len(fl) for _, fl in otypes_h], np.int32)
    flarr = [np.ascontiguousarray(fl, np.int32) for _, fl in otypes_h]
    FL = (_ip * nt)(*[_ptr(a, _ip) for a in flarr])
    nobs = len(type_)
    type_ = np.ascontiguousarray(type_, np.int16)
    status = np.ascontiguousarray(status, np.int16)
    fi = np.ascontiguousarray(fi, np.float64)
    fj = np.ascontiguousarray(fj, np.float64)
    fk = np.ascontiguousarray(fk, np.float32)
    value = np.ascontiguousarray(value, np.float64)
    HA = np.zeros((nmem, nobs), np.float32)
    s = np.zeros(nobs)
    sd = np.zeros(nobs)
    r = lib().or_getHE(mode, nmem, nmem_dynamic, float(gamma), ni, nj, ktop, periodic_i, _ptr(nlevels, _ip), E, nobs,
                       _ptr(type_, _sp), _ptr(status, _sp), _ptr(fi, _dp), _ptr(fj, _dp), _ptr(fk, _fp),
                       _ptr(value, _dp), _ptr(hk, _ip), _ptr(nk, _ip), FL, _ptr(HA, _fp), _ptr(s, _dp), _ptr(sd, _dp))
    if r != 0:
        raise RuntimeError("or_getHE: %d" % r)
    return HA, s, sd


def interp2d(fij, v, mask, periodic_i=0):
    nj, ni = v.shape
    fij = np.ascontiguousarray(fij, np.float64)
    v = np.ascontiguousarray(v, np.float32)
    mask = np.ascontiguousarray(mask, np.int32)
    return lib().or_interp2d(_ptr(fij, _dp), ni, nj, _ptr(v, _fp), _ptr(mask, _ip), periodic_i)


def interp3d(fij, fk, v, nlevels, ktop=0, periodic_i=0):
    nk, nj, ni = v.shape
    fij = np.ascontiguousarray(fij, np.float64)
    layers = [np.ascontiguousarray(v[k], np.float32) for k in range(nk)]
    L = (_fp * nk)(*[_ptr(a, _fp) for a in layers])
    nlevels = np.ascontiguousarray(nlevels, np.int32)
    return lib().or_interp3d(_ptr(fij, _dp), float(fk), ni, nj, nk, ktop, L, _ptr(nlevels, _ip), periodic_i)


def ref_trees_parallel(obs, n):
    """n independent sets of the reference's kd-trees, built on n threads"""
    import threading
    out = [None] * n

    def mk(k):
        out[k] = RefTrees(obs)
    ths = [threading.Thread(target=mk, args=(k,)) for k in range(n)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    return out
