"""Full-size synthetic problems generated on the GPU (libenkfsynth.so).

make_tile() builds the one-GPU tile: a curvilinear grid of nj x ni cells at
~0.1 deg, nk layers, an ensemble of nmem members for 3D temp and salt and 2D
eta built as mean + Gaussian random fields (correlation length corr_km),
super-obs of four types (SST on uniform random surface points, SLA along
500 km tracks, T and S profiles over all layers), their ensemble anomalies
HA ([nmem][nobs] float32) and innovations.  A band of the tile (rows row0 ..
row0 + nj - 1 of a tile of nj_map rows) has the tile's geometry and obs
density.

make_clustered_tile() builds the hybrid uneven-batching problem: obs drawn
as clustered SLA tracks and T/S profiles from a log-normal intensity field,
so that the local obs counts per column are roughly log-normal.

The generator is a small HIP library of its own (Philox random numbers,
random-Fourier-feature Gaussian fields); no other GPU library is loaded, so
the process that runs the engine launches only libenkfg's kernels and the
generator's.  Host-side choices (obs positions) use numpy with the given
seed.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libenkfsynth.so")
REARTH = 6371.0

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libenkfsynth.so is not built (run __graft_entry__.build() or make -C enkf-c_amd)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.es_malloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.es_free.argtypes = [vp]
        L.es_h2d.argtypes = [vp, vp, C.c_size_t]
        L.es_d2h.argtypes = [vp, vp, C.c_size_t]
        L.es_zero.argtypes = [vp, C.c_size_t]
        L.es_error_string.restype = C.c_char_p
        L.es_error_string.argtypes = [C.c_int]
        L.es_grf.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint]
        L.es_tile_ensemble.argtypes = [vp, C.c_int, C.c_int, C.c_longlong, vp, vp, vp, vp, vp, vp, vp]
        L.es_tile_ha.argtypes = [vp, vp, C.c_int, C.c_int, C.c_longlong, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                 C.c_uint]
        L.es_uniform.argtypes = [vp, C.c_longlong, C.c_uint, C.c_uint]
        L.es_copy.argtypes = [vp, vp, C.c_size_t]
        _lib = L
    return _lib


def _check(r, what):
    if r != 0:
        raise RuntimeError("%s: %s" % (what, lib().es_error_string(r).decode()))


class DeviceArray:
    """A device buffer of the generator (hipMalloc), freed with the object."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(x) for x in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().es_malloc(C.byref(p), self.nbytes), "device allocation of %d bytes" % self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _check(lib().es_h2d(d.ptr, a.ctypes.data, d.nbytes), "host to device copy")
        return d

    def to_host(self, out=None):
        if out is None:
            out = np.empty(self.shape, self.dtype)
        assert out.nbytes == self.nbytes and out.flags.c_contiguous
        _check(lib().es_d2h(out.ctypes.data, self.ptr, self.nbytes), "device to host copy")
        return out

    def read(self, offset_elems, count, out=None):
        """count elements from element offset_elems, as a flat host array"""
        offset_elems, count = int(offset_elems), int(count)
        assert 0 <= offset_elems and offset_elems + count <= self.nbytes // self.dtype.itemsize
        if out is None:
            out = np.empty(count, self.dtype)
        _check(lib().es_d2h(out.ctypes.data, self.ptr + offset_elems * self.dtype.itemsize,
                            count * self.dtype.itemsize), "device to host copy")
        return out

    def write(self, offset_elems, a):
        a = np.ascontiguousarray(a, self.dtype)
        offset_elems = int(offset_elems)
        assert 0 <= offset_elems and offset_elems + a.size <= self.nbytes // self.dtype.itemsize
        _check(lib().es_h2d(self.ptr + offset_elems * self.dtype.itemsize, a.ctypes.data, a.nbytes),
               "host to device copy")

    def zero(self):
        _check(lib().es_zero(self.ptr, self.nbytes), "device memset")

    def free(self):
        if getattr(self, "ptr", None):
            lib().es_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def grf(n, nj, ni, dx_km, corr_km, seed):
    """n unit-variance Gaussian random fields [n][nj][ni] on the device."""
    d = DeviceArray((n, nj, ni), np.float32)
    _check(lib().es_grf(d.ptr, n, nj, ni, float(dx_km), float(corr_km), int(seed) & 0xFFFFFFFF), "es_grf")
    return d


def grid_lonlat(fi, fj, ni, nj, lon0=-20.0, lat0=-75.0, d=0.1):
    """Curvilinear mapping (fractional indices -> lon, lat)."""
    lon = lon0 + d * fi + 0.5 * np.sin(math.pi * fj / nj)
    lat = lat0 + d * fj + 0.3 * np.sin(2.0 * math.pi * fi / ni)
    return lon, lat


class Tile:
    pass


def _levels(bath, nk, land_frac):
    thr = np.quantile(bath.ravel()[:: max(1, bath.size // 1_000_000)], land_frac)
    lev = np.clip(np.floor((bath - thr) * 40.0), 0, nk).astype(np.int32)
    return np.where(bath > thr, np.maximum(lev, 1), 0).astype(np.int32)


def _finish(X, fi_all, fj_all, k_all, type_all, obs_std, locrad, seed, G, prof, sd, mean0, with_ensemble, G2,
            order):
    nj, ni, nk, nmem = X.nj, X.ni, X.nk, X.nmem
    ncell = nj * ni
    if order == "prep":
        # enkf_prep's super-obs order (prep/enkf_prep.c:192-287): by type,
        # then grid cell i, then j, then level
        ii = np.floor(fi_all).astype(np.int64)
        jj = np.floor(fj_all).astype(np.int64)
        o = np.lexsort((k_all, jj, ii, type_all))
        fi_all, fj_all, k_all, type_all = fi_all[o], fj_all[o], k_all[o], type_all[o]
    lon, lat = grid_lonlat(fi_all, fj_all + X.row0, ni, X.nj_map)
    X.obs_lon = lon.astype(np.float32)
    X.obs_lat = lat.astype(np.float32)
    X.obs_type = type_all.astype(np.int16)
    X.obs_fi, X.obs_fj, X.obs_k = fi_all, fj_all, k_all
    X.nobs = len(type_all)
    X.obs_status = np.zeros(X.nobs, np.int16)
    std_t = np.array([obs_std[0], obs_std[1], obs_std[2], obs_std[2]], np.float32)
    X.obs_estd = std_t[type_all]
    X.obstypes = [dict(locrad=[locrad], rfactor=1.0) for _ in range(4)]

    # HA at the obs (nearest cell) and innovations
    cell = (np.rint(fj_all).astype(np.int64) * ni + np.rint(fi_all).astype(np.int64)).astype(np.int32)
    var = np.array([0, 2, 0, 1], np.int32)[type_all]          # SST, TEM -> temp; SLA -> eta; SAL -> salt
    dcell = DeviceArray.from_host(cell)
    dvar = DeviceArray.from_host(var)
    dlev = DeviceArray.from_host(k_all.astype(np.int32))
    dprof = DeviceArray.from_host(prof.astype(np.float32))
    destd = DeviceArray.from_host(X.obs_estd)
    sdv = np.array([sd["temp"], sd["salt"], sd["eta"]], np.float32)
    dHA = DeviceArray((nmem, X.nobs), np.float32)
    dinn = DeviceArray((X.nobs,), np.float64)
    _check(lib().es_tile_ha(dHA.ptr, dinn.ptr, nmem, X.nobs, ncell, G["temp"].ptr, G["salt"].ptr, G["eta"].ptr,
                            dcell.ptr, dvar.ptr, dlev.ptr, dprof.ptr, destd.ptr, sdv.ctypes.data,
                            (seed * 7919 + 17) & 0xFFFFFFFF), "es_tile_ha")
    X.HA = dHA.to_host()
    X.innov = dinn.to_host()
    del dHA, dinn, dcell, dvar, dlev, destd

    # fields: temp (nk levels), salt (nk levels), eta -> E[field][cell][member]
    X.fields = [dict(level=k, inflation=1.05, inf_ratio=1.0) for k in range(nk)] * 2 + \
        [dict(level=0, issurfacevar=1, inflation=1.05, inf_ratio=1.0)]
    X.nfield = len(X.fields)
    X.E = None
    if with_ensemble:
        X.E = DeviceArray((X.nfield, ncell, nmem), np.float32)
        meanv = np.array([mean0["temp"], mean0["salt"], mean0["eta"]], np.float32)
        _check(lib().es_tile_ensemble(X.E.ptr, nmem, nk, ncell, G["temp"].ptr, G["salt"].ptr, G["eta"].ptr,
                                      G2.ptr, dprof.ptr, sdv.ctypes.data, meanv.ctypes.data), "es_tile_ensemble")
    return X


def make_tile(nj=1500, ni=1800, nk=51, nmem=96, n_sst=4_200_000, n_sla=2_100_000, n_prof=6_900,
              corr_km=300.0, locrad=250.0, stride=3, obs_std=(0.4, 0.04, 0.05), seed=0, with_ensemble=True,
              land_frac=0.05, row0=0, nj_map=None, order="random"):
    """The tile (or a band of rows row0 .. row0 + nj - 1 of a tile nj_map rows
    high); order "prep" sorts the obs as enkf_prep writes super-obs."""
    rng = np.random.default_rng(seed)
    X = Tile()
    X.nj, X.ni, X.nk, X.nmem, X.stride = nj, ni, nk, nmem, stride
    X.row0, X.nj_map = row0, (nj_map if nj_map is not None else nj)
    ncell = nj * ni
    X.ncell = ncell
    dx_km = 0.1 * math.pi / 180.0 * REARTH
    s0 = (seed * 1000003) & 0xFFFFFFF

    G = {name: grf(nmem, nj, ni, dx_km, corr_km, s0 + k) for k, name in enumerate(("temp", "salt", "eta"))}
    G2 = grf(1, nj, ni, dx_km, corr_km, s0 + 11)
    prof = np.exp(-np.arange(nk, dtype=np.float64) / 15.0)
    bath = grf(1, nj, ni, dx_km, 4 * corr_km, s0 + 12).to_host()[0]
    X.nlevels = _levels(bath, nk, land_frac)

    jn = np.arange(0, nj, stride)
    inn = np.arange(0, ni, stride)
    fj, fi = np.meshgrid(jn.astype(np.float64), inn.astype(np.float64), indexing="ij")
    X.node_lon, X.node_lat = grid_lonlat(fi, fj + row0, ni, X.nj_map)
    X.nj_s, X.ni_s = len(jn), len(inn)

    sd = {"temp": 0.5, "salt": 0.05, "eta": 0.05}
    mean0 = {"temp": 15.0, "salt": 35.0, "eta": 0.0}
    # observations: type 0 SST, 1 SLA (tracks), 2 TEM, 3 SAL (profiles)
    sst_i = rng.uniform(0, ni - 1, n_sst)
    sst_j = rng.uniform(0, nj - 1, n_sst)
    ntrack = n_sla // 70 if n_sla >= 70 else (1 if n_sla > 0 else 0)
    per = max(1, n_sla // max(ntrack, 1))
    t_i0 = rng.uniform(0, ni - 1, ntrack)
    t_j0 = rng.uniform(0, nj - 1, ntrack)
    ang = rng.uniform(0, 2 * math.pi, ntrack)
    step = 500.0 / dx_km / per    # 500 km tracks
    s = np.arange(per)
    sla_i = np.clip(t_i0[:, None] + np.cos(ang)[:, None] * s[None, :] * step, 0, ni - 1).ravel()
    sla_j = np.clip(t_j0[:, None] + np.sin(ang)[:, None] * s[None, :] * step, 0, nj - 1).ravel()
    p_i = rng.uniform(0, ni - 1, n_prof)
    p_j = rng.uniform(0, nj - 1, n_prof)
    prof_i = np.repeat(p_i, nk)
    prof_j = np.repeat(p_j, nk)
    prof_k = np.tile(np.arange(nk), n_prof)
    fi_all = np.concatenate([sst_i, sla_i, prof_i, prof_i])
    fj_all = np.concatenate([sst_j, sla_j, prof_j, prof_j])
    k_all = np.concatenate([np.zeros(n_sst + len(sla_i), np.int64), prof_k, prof_k])
    type_all = np.concatenate([np.zeros(n_sst, np.int16), np.ones(len(sla_i), np.int16),
                               np.full(len(prof_i), 2, np.int16), np.full(len(prof_i), 3, np.int16)])
    X = _finish(X, fi_all, fj_all, k_all, type_all, obs_std, locrad, seed, G, prof, sd, mean0, with_ensemble, G2,
                order)
    del G, G2
    return X


def make_clustered_tile(nj=1200, ni=1200, nk=51, nmem=144, nmem_dynamic=96, median=400.0, pmax=20000.0,
                        corr_km=300.0, cluster_km=600.0, locrad=150.0, stride=3, obs_std=(0.4, 0.04, 0.05), seed=0,
                        with_ensemble=True, land_frac=0.05, sigma=None):
    """Hybrid problem with uneven local obs counts: SLA tracks and T/S
    profiles placed by a log-normal intensity exp(mu + sigma g) (g a unit
    random field of correlation cluster_km), mu and sigma chosen so that the
    obs count within locrad of a column is about log-normal with the given
    median and a largest value near pmax.  Members nmem_dynamic .. nmem - 1
    are the static part of the ensemble (their anomalies come from fields of
    their own, as a static ensemble's would)."""
    rng = np.random.default_rng(seed)
    X = Tile()
    X.nj, X.ni, X.nk, X.nmem, X.stride = nj, ni, nk, nmem, stride
    X.nmem_dynamic = nmem_dynamic
    X.row0, X.nj_map = 0, nj
    ncell = nj * ni
    X.ncell = ncell
    dx_km = 0.1 * math.pi / 180.0 * REARTH
    s0 = (seed * 1000003 + 5) & 0xFFFFFFF
    G = {name: grf(nmem, nj, ni, dx_km, corr_km, s0 + k) for k, name in enumerate(("temp", "salt", "eta"))}
    G2 = grf(1, nj, ni, dx_km, corr_km, s0 + 11)
    prof = np.exp(-np.arange(nk, dtype=np.float64) / 15.0)
    bath = grf(1, nj, ni, dx_km, 4 * corr_km, s0 + 12).to_host()[0]
    X.nlevels = _levels(bath, nk, land_frac)
    jn = np.arange(0, nj, stride)
    inn = np.arange(0, ni, stride)
    fj, fi = np.meshgrid(jn.astype(np.float64), inn.astype(np.float64), indexing="ij")
    X.node_lon, X.node_lat = grid_lonlat(fi, fj, ni, nj)
    X.nj_s, X.ni_s = len(jn), len(inn)

    # intensity: per-cell expected obs such that a column's count within
    # locrad, about lambda(cell) * cells in the disc, is log-normal
    g = grf(1, nj, ni, dx_km, cluster_km, s0 + 13).to_host()[0].astype(np.float64)
    if sigma is None:
        zmax = max(float(g.max()), 1.0)
        sigma = math.log(pmax / median) / zmax
    # cells in the disc of radius locrad, with the curvilinear cell area
    lat = grid_lonlat(np.zeros(nj), np.arange(nj, dtype=np.float64), ni, nj)[1]
    cell_km2 = dx_km * dx_km * np.cos(np.radians(lat))[:, None] * np.ones((1, ni))
    lam = median * cell_km2 / (math.pi * locrad * locrad) * np.exp(sigma * g)
    ntot = int(lam.sum())
    # half the obs in 51-level profile pairs (T and S), half in 70-point tracks
    n_prof = max(1, ntot // (4 * nk))
    n_track = max(1, ntot // (2 * 70))
    pcell = lam.ravel() / lam.sum()
    pc = rng.choice(ncell, size=n_prof, p=pcell)
    tc = rng.choice(ncell, size=n_track, p=pcell)
    p_i = np.clip(pc % ni + rng.uniform(-0.5, 0.5, n_prof), 0, ni - 1)
    p_j = np.clip(pc // ni + rng.uniform(-0.5, 0.5, n_prof), 0, nj - 1)
    ang = rng.uniform(0, 2 * math.pi, n_track)
    step = 100.0 / dx_km / 70    # 100 km tracks: a track stays in its cluster
    s = np.arange(70)
    sla_i = np.clip((tc % ni)[:, None] + np.cos(ang)[:, None] * s[None, :] * step, 0, ni - 1).ravel()
    sla_j = np.clip((tc // ni)[:, None] + np.sin(ang)[:, None] * s[None, :] * step, 0, nj - 1).ravel()
    prof_i = np.repeat(p_i, nk)
    prof_j = np.repeat(p_j, nk)
    prof_k = np.tile(np.arange(nk), n_prof)
    fi_all = np.concatenate([sla_i, prof_i, prof_i])
    fj_all = np.concatenate([sla_j, prof_j, prof_j])
    k_all = np.concatenate([np.zeros(len(sla_i), np.int64), prof_k, prof_k])
    type_all = np.concatenate([np.ones(len(sla_i), np.int16), np.full(len(prof_i), 2, np.int16),
                               np.full(len(prof_i), 3, np.int16)])
    sd = {"temp": 0.5, "salt": 0.05, "eta": 0.05}
    mean0 = {"temp": 15.0, "salt": 35.0, "eta": 0.0}
    X.sigma = sigma
    X = _finish(X, fi_all, fj_all, k_all, type_all, obs_std, locrad, seed, G, prof, sd, mean0, with_ensemble, G2,
                "random")
    # four obstypes are declared; type 0 (SST) has no obs here
    del G, G2
    return X


def copy_probe(nbytes):
    """one plain device copy of nbytes (a multiple of 16) with the generator's
    copy kernel; returns its wall time in seconds (after a warm-up copy)"""
    import time
    a = DeviceArray((nbytes // 4,), np.float32)
    b = DeviceArray((nbytes // 4,), np.float32)
    a.zero()
    _check(lib().es_copy(b.ptr, a.ptr, nbytes), "es_copy")
    t = time.perf_counter()
    _check(lib().es_copy(b.ptr, a.ptr, nbytes), "es_copy")
    return time.perf_counter() - t
