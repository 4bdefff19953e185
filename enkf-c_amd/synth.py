"""Synthetic problems for the calc/update path (host, numpy).

A problem has a lat-lon (or gently curvilinear) model grid, a truth and an
ensemble built as truth + Gaussian random fields (spectral, Gaussian
covariance with the given correlation length), observations drawn from the
truth with Gaussian errors, the ensemble observations HE (bilinear in the
grid), and the stride sub-grid node coordinates -- everything
das_calctransforms and das_updatefields consume, in the reference layouts.

Used by the tests and by bench.py; sizes are chosen by the caller.
"""
import math

import numpy as np

REARTH = 6371.0


def grf(rng, nj, ni, dx_km, corr_km, n=1, dtype=np.float32):
    """n unit-variance Gaussian random fields on an nj x ni grid of spacing
    dx_km with Gaussian correlation length corr_km (spectral synthesis)."""
    ky = np.fft.fftfreq(nj, d=dx_km)[:, None]
    kx = np.fft.rfftfreq(ni, d=dx_km)[None, :]
    k2 = kx * kx + ky * ky
    amp = np.exp(-(math.pi ** 2) * corr_km ** 2 * k2)   # sqrt of the Gaussian spectrum
    out = np.empty((n, nj, ni), dtype)
    for k in range(n):
        w = rng.standard_normal((nj, ni))
        f = np.fft.irfft2(np.fft.rfft2(w) * amp, s=(nj, ni))
        f -= f.mean()
        sd = f.std()
        out[k] = f / (sd if sd > 0 else 1.0)
    return out


class Problem:
    pass


def make_problem(nj=60, ni=90, nk=4, dlon=0.1, dlat=0.1, lon0=140.0, lat0=-40.0, nmem=24, nobs=3000,
                 corr_km=150.0, std=(0.5, 0.05, 0.05), obs_std=0.3, locrad=100.0, stride=3, seed=0,
                 curvilinear=False, profiles=0, nmem_static=0, land_frac=0.0):
    """Returns a Problem with fields:
      lon, lat            model grid coordinates [nj][ni] (double)
      nlevels             [nj][ni] int (0 = land)
      ens                 dict var -> float32 [nmem][nk or 1][nj][ni]
      truth               dict var -> float32 [nk or 1][nj][ni]
      obs: type, status, lon, lat, estd, value (float), depth level k
      HE                  float32 [nmem][nobs]
      node_lon, node_lat  [nj_s][ni_s] stride sub-grid coordinates
      obstypes            list of dicts for make_obstypes
    """
    rng = np.random.default_rng(seed)
    P = Problem()
    P.nj, P.ni, P.nk, P.nmem, P.stride = nj, ni, nk, nmem, stride
    jj, ii = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    lon = lon0 + ii * dlon
    lat = lat0 + jj * dlat
    if curvilinear:
        lon = lon + 0.3 * dlon * np.sin(2 * math.pi * jj / max(nj, 1))
        lat = lat + 0.3 * dlat * np.sin(2 * math.pi * ii / max(ni, 1))
    P.lon, P.lat = lon.astype(np.float64), lat.astype(np.float64)
    dx_km = dlat * math.pi / 180.0 * REARTH
    nl = np.full((nj, ni), nk, np.int32)
    if land_frac > 0:
        land = grf(rng, nj, ni, dx_km, 3 * corr_km, 1, np.float64)[0] > np.quantile(
            grf(rng, nj, ni, dx_km, 3 * corr_km, 1, np.float64)[0], 1 - land_frac)
        nl[land] = 0
        shelf = rng.integers(1, nk + 1, size=(nj, ni))
        nl = np.where(nl > 0, np.minimum(nl, np.maximum(shelf, 1)), 0).astype(np.int32)
    P.nlevels = nl
    vars3 = [("temp", std[0], 15.0), ("salt", std[1], 35.0)]
    P.truth, P.ens = {}, {}
    for name, sd, mean in vars3:
        base = grf(rng, nj, ni, dx_km, corr_km, 1)[0]
        P.truth[name] = np.stack([mean - 0.3 * k + sd * base for k in range(nk)]).astype(np.float32)
        anom = grf(rng, nj, ni, dx_km, corr_km, nmem)
        prof = np.exp(-np.arange(nk) / max(nk, 1))
        e = P.truth[name][None] + sd * anom[:, None] * prof[None, :, None, None].astype(np.float32)
        e += (sd * 0.1 * rng.standard_normal(e.shape)).astype(np.float32)
        P.ens[name] = e.astype(np.float32)
    base = grf(rng, nj, ni, dx_km, corr_km, 1)[0]
    P.truth["eta"] = (std[2] * base)[None].astype(np.float32)
    P.ens["eta"] = (P.truth["eta"][None] + std[2] * grf(rng, nj, ni, dx_km, corr_km, nmem)[:, None]).astype(np.float32)

    # observations: SST (type 0) uniform on the surface, optional T profiles (type 1)
    fi = rng.uniform(0, ni - 1, nobs)
    fj = rng.uniform(0, nj - 1, nobs)
    otype = np.zeros(nobs, np.int16)
    ok = np.zeros(nobs, np.int32)
    if profiles > 0:
        pi = rng.uniform(0, ni - 1, profiles)
        pj = rng.uniform(0, nj - 1, profiles)
        fi = np.concatenate([fi, np.repeat(pi, nk)])
        fj = np.concatenate([fj, np.repeat(pj, nk)])
        otype = np.concatenate([otype, np.ones(profiles * nk, np.int16)])
        ok = np.concatenate([ok, np.tile(np.arange(nk, dtype=np.int32), profiles)])
    n = len(fi)

    def bil(f2d):
        i0 = np.clip(np.floor(fi).astype(int), 0, ni - 2)
        j0 = np.clip(np.floor(fj).astype(int), 0, nj - 2)
        a, b = fi - i0, fj - j0
        return ((1 - a) * (1 - b) * f2d[..., j0, i0] + a * (1 - b) * f2d[..., j0, i0 + 1]
                + (1 - a) * b * f2d[..., j0 + 1, i0] + a * b * f2d[..., j0 + 1, i0 + 1])

    i0 = np.clip(np.floor(fi).astype(int), 0, ni - 2)
    j0 = np.clip(np.floor(fj).astype(int), 0, nj - 2)
    a, b = fi - i0, fj - j0
    olon = (1 - a) * (1 - b) * lon[j0, i0] + a * (1 - b) * lon[j0, i0 + 1] + (1 - a) * b * lon[j0 + 1, i0] \
        + a * b * lon[j0 + 1, i0 + 1]
    olat = (1 - a) * (1 - b) * lat[j0, i0] + a * (1 - b) * lat[j0, i0 + 1] + (1 - a) * b * lat[j0 + 1, i0] \
        + a * b * lat[j0 + 1, i0 + 1]
    temp = P.ens["temp"]
    HE = np.empty((nmem, n), np.float32)
    tv = np.empty(n)
    for k in range(nk):
        sel = ok == k
        if not sel.any():
            continue
        fi_s, fj_s = fi[sel], fj[sel]
        ii0 = np.clip(np.floor(fi_s).astype(int), 0, ni - 2)
        jj0 = np.clip(np.floor(fj_s).astype(int), 0, nj - 2)
        aa, bb = fi_s - ii0, fj_s - jj0

        def b2(f):
            return ((1 - aa) * (1 - bb) * f[..., jj0, ii0] + aa * (1 - bb) * f[..., jj0, ii0 + 1]
                    + (1 - aa) * bb * f[..., jj0 + 1, ii0] + aa * bb * f[..., jj0 + 1, ii0 + 1])
        HE[:, sel] = b2(temp[:, k]).astype(np.float32)
        tv[sel] = b2(P.truth["temp"][k])
    estd = np.where(otype == 0, obs_std, 0.5 * obs_std).astype(np.float32)
    value = (tv + estd * rng.standard_normal(n)).astype(np.float32)
    P.obs_type, P.obs_lon, P.obs_lat = otype, olon.astype(np.float32), olat.astype(np.float32)
    P.obs_estd, P.obs_value, P.obs_k = estd, value, ok
    P.obs_status = np.zeros(n, np.int16)
    P.nobs = n
    P.HE = HE
    # innovations and anomalies as das_calcinnandspread forms them (float HA)
    Hx = HE.astype(np.float64).mean(axis=0)
    P.HA = (HE - Hx.astype(np.float32)[None, :]).astype(np.float32)
    P.innov = (value.astype(np.float64) - Hx)
    # stride sub-grid
    jn = np.arange(0, nj, stride)
    inn = np.arange(0, ni, stride)
    P.node_lon = np.ascontiguousarray(P.lon[np.ix_(jn, inn)])
    P.node_lat = np.ascontiguousarray(P.lat[np.ix_(jn, inn)])
    P.nj_s, P.ni_s = len(jn), len(inn)
    P.obstypes = [dict(locrad=[locrad], rfactor=1.0)]
    if profiles > 0:
        P.obstypes.append(dict(locrad=[locrad], rfactor=2.0))
    return P
