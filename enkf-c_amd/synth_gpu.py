"""Full-size synthetic problems generated on the GPU with torch.

make_tile() builds the one-GPU tile: a curvilinear grid of nj x ni cells at
~0.1 deg, nk layers, an ensemble of nmem members for 3D temp and salt and 2D
eta built as mean + Gaussian random fields (spectral, correlation length
corr_km), super-obs of four types (SST on uniform random surface points, SLA
along 500 km tracks, T and S profiles over all layers), their standardisable
ensemble anomalies HA ([nmem][nobs] float32) and innovations.

Torch is used only as device memory and FFT plumbing for the synthetic data;
the analysis itself runs in libenkfg.
"""
import math

import numpy as np
import torch

REARTH = 6371.0


def grid_lonlat(fi, fj, ni, nj, lon0=-20.0, lat0=-75.0, d=0.1):
    """Curvilinear mapping (fractional indices -> lon, lat), numpy or torch."""
    lib = torch if isinstance(fi, torch.Tensor) else np
    lon = lon0 + d * fi + 0.5 * lib.sin(math.pi * fj / nj)
    lat = lat0 + d * fj + 0.3 * lib.sin(2.0 * math.pi * fi / ni)
    return lon, lat


def grf_batch(gen, n, nj, ni, dx_km, corr_km, device):
    """n unit-variance Gaussian random fields [n][nj][ni] (float32)."""
    ky = torch.fft.fftfreq(nj, d=dx_km, device=device)[:, None]
    kx = torch.fft.rfftfreq(ni, d=dx_km, device=device)[None, :]
    amp = torch.exp(-(math.pi ** 2) * corr_km ** 2 * (kx * kx + ky * ky)).to(torch.float32)
    w = torch.randn((n, nj, ni), generator=gen, device=device, dtype=torch.float32)
    f = torch.fft.irfft2(torch.fft.rfft2(w) * amp, s=(nj, ni))
    f = f - f.mean(dim=(1, 2), keepdim=True)
    f = f / f.std(dim=(1, 2), keepdim=True).clamp_min(1e-20)
    return f


class Tile:
    pass


def make_tile(nj=1500, ni=1800, nk=51, nmem=96, n_sst=4_200_000, n_sla=2_100_000, n_prof=6_900,
              corr_km=300.0, locrad=250.0, stride=3, obs_std=(0.4, 0.04, 0.05), seed=0, device="cuda",
              with_ensemble=True, land_frac=0.05):
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    X = Tile()
    X.nj, X.ni, X.nk, X.nmem, X.stride = nj, ni, nk, nmem, stride
    ncell = nj * ni
    X.ncell = ncell
    dx_km = 0.1 * math.pi / 180.0 * REARTH

    # base random fields per variable and member
    G = {}
    for name in ("temp", "salt", "eta"):
        G[name] = torch.empty((nmem, nj, ni), device=dev, dtype=torch.float32)
        for e0 in range(0, nmem, 16):
            e1 = min(e0 + 16, nmem)
            G[name][e0:e1] = grf_batch(gen, e1 - e0, nj, ni, dx_km, corr_km, dev)
    G2 = grf_batch(gen, 2, nj, ni, dx_km, corr_km, dev)   # level decorrelation fields
    prof = torch.exp(-torch.arange(nk, device=dev, dtype=torch.float32) / 15.0)

    # bathymetry: number of wet layers per cell, a fraction land_frac of land
    bath = grf_batch(gen, 1, nj, ni, dx_km, 4 * corr_km, dev)[0]
    thr = torch.quantile(bath.flatten()[:: max(1, ncell // 1_000_000)], land_frac)
    lev = ((bath - thr) * 40.0).clamp(0, nk).floor().to(torch.int32)
    lev = torch.where(bath > thr, lev.clamp_min(1), torch.zeros_like(lev))
    X.nlevels = lev.cpu().numpy()

    # grid and stride nodes
    jn = np.arange(0, nj, stride)
    inn = np.arange(0, ni, stride)
    fj, fi = np.meshgrid(jn.astype(np.float64), inn.astype(np.float64), indexing="ij")
    X.node_lon, X.node_lat = grid_lonlat(fi, fj, ni, nj)
    X.nj_s, X.ni_s = len(jn), len(inn)

    sd = {"temp": 0.5, "salt": 0.05, "eta": 0.05}
    mean0 = {"temp": 15.0, "salt": 35.0, "eta": 0.0}

    # observations: type 0 SST, 1 SLA (tracks), 2 TEM, 3 SAL (profiles)
    sst_i = rng.uniform(0, ni - 1, n_sst)
    sst_j = rng.uniform(0, nj - 1, n_sst)
    ntrack = max(1, n_sla // 70)
    per = n_sla // ntrack
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
    lon, lat = grid_lonlat(fi_all, fj_all, ni, nj)
    X.obs_lon = lon.astype(np.float32)
    X.obs_lat = lat.astype(np.float32)
    X.obs_type = type_all
    X.nobs = len(type_all)
    X.obs_status = np.zeros(X.nobs, np.int16)
    std_t = np.array([obs_std[0], obs_std[1], obs_std[2], obs_std[2]], np.float32)
    X.obs_estd = std_t[type_all]
    X.obstypes = [dict(locrad=[locrad], rfactor=1.0) for _ in range(4)]

    # HA at the obs (nearest cell), [nmem][nobs] float32 anomalies, and innovations
    ci = torch.from_numpy(np.rint(fi_all).astype(np.int64)).to(dev)
    cj = torch.from_numpy(np.rint(fj_all).astype(np.int64)).to(dev)
    ck = torch.from_numpy(k_all).to(dev)
    ty = torch.from_numpy(type_all.astype(np.int64)).to(dev)
    cell = cj * ni + ci
    HA = torch.empty((nmem, X.nobs), device=dev, dtype=torch.float32)
    for name, tsel in (("temp", (0, 2)), ("eta", (1,)), ("salt", (3,))):
        sel = torch.zeros_like(ty, dtype=torch.bool)
        for t in tsel:
            sel |= ty == t
        idx = torch.nonzero(sel).squeeze(1)
        g = G[name].reshape(nmem, ncell)[:, cell[idx]]
        if name != "eta":
            g = g * prof[ck[idx]][None, :]
        HA[:, idx] = sd[name] * g
    HA -= HA.mean(dim=0, keepdim=True)
    spread = HA.std(dim=0)
    innov = torch.randn(X.nobs, generator=gen, device=dev, dtype=torch.float64) * torch.sqrt(
        spread.double() ** 2 + torch.from_numpy(X.obs_estd.astype(np.float64)).to(dev) ** 2)
    X.HA = HA.cpu().numpy()
    X.innov = innov.cpu().numpy()
    del HA, innov

    # fields: temp (nk levels), salt (nk levels), eta -> E[field][cell][member]
    X.fields = [dict(level=k, inflation=1.05, inf_ratio=1.0) for k in range(nk)] * 2 + \
        [dict(level=0, issurfacevar=1, inflation=1.05, inf_ratio=1.0)]
    X.nfield = len(X.fields)
    X.E = None
    if with_ensemble:
        X.E = torch.empty((X.nfield, ncell, nmem), device=dev, dtype=torch.float32)
        f = 0
        for name in ("temp", "salt"):
            base = G[name].reshape(nmem, ncell).t()            # [cell][member]
            for k in range(nk):
                c, s_ = math.cos(0.05 * k), math.sin(0.05 * k)
                m0 = mean0[name] - 0.2 * k + 0.3 * G2[0].reshape(ncell, 1)
                X.E[f] = m0 + sd[name] * float(prof[k]) * (c * base + s_ * base.roll(1, dims=1))
                f += 1
        X.E[f] = mean0["eta"] + 0.2 * G2[0].reshape(ncell, 1) + sd["eta"] * G["eta"].reshape(nmem, ncell).t()
    del G, G2
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)
    return X
