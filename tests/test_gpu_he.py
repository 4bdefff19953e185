"""The ensemble observations on the device (enkfg_calc_HE: das_getHE with the
standard H functions, the hybrid scaling and das_calcinnandspread) against
the oracle, through the C ABI.

HE is a weighted sum in FP64 of the surrounding grid values in the
reference's order, and the member sums run in member order, so HA, the
innovations and the spread match the oracle bit for bit.  The engine's S
and s after the call are those enkfg_set_S would make from the same HA and
innovations, and the calc that follows matches the oracle's.
"""
import numpy as np
import pytest

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu


def problem(seed, m, land_frac):
    P = synth.make_problem(seed=seed, nmem=m, nobs=10, locrad=80.0, nj=30, ni=41, nk=5, stride=3,
                           land_frac=land_frac)
    rng = np.random.default_rng(seed)
    nk = P.nk
    # fields in attached order: temp levels, salt levels, eta
    fields = [P.ens["temp"][:, k] for k in range(nk)] + [P.ens["salt"][:, k] for k in range(nk)] + [P.ens["eta"][:, 0]]
    fl = [dict(level=k) for k in range(nk)] * 2 + [dict(level=0, issurfacevar=1)]
    # obs: type 0 SST (surface temp), 1 T profiles, 2 S profiles, 3 SLA (eta);
    # each obs next to a cell wet at every layer
    deep = np.argwhere(P.nlevels >= nk)
    n = 1800
    cj, ci = deep[rng.integers(0, len(deep), n)].T
    fi = np.clip(ci + rng.uniform(0.0, 0.99, n), 0, P.ni - 1)
    fj = np.clip(cj + rng.uniform(0.0, 0.99, n), 0, P.nj - 1)
    # obs on the last column and on the last row, at cells wet at every layer
    lastc = np.flatnonzero(P.nlevels[:, -1] >= nk)
    lastr = np.flatnonzero(P.nlevels[-1, :] >= nk)
    if len(lastc):
        fi[:5] = P.ni - 1.0
        fj[:5] = lastc[rng.integers(0, len(lastc), 5)]
    if len(lastr):
        fj[5:10] = P.nj - 1.0
        fi[5:10] = lastr[rng.integers(0, len(lastr), 5)]
    typ = rng.integers(0, 4, n).astype(np.int16)
    fk = rng.uniform(-0.5, nk - 0.5, n).astype(np.float32)
    fk[10:20] = 2.0                                  # on a layer centre
    status = np.zeros(n, np.int16)
    status[20:30] = 3
    value = rng.standard_normal(n) + np.where(typ == 3, 0.0, np.where(typ == 2, 35.0, 15.0))
    hops = [(0, list(range(nk))), (1, list(range(nk))), (1, list(range(nk, 2 * nk))), (0, [2 * nk])]
    estd = np.where(typ == 3, 0.05, 0.3).astype(np.float32)
    lon = np.interp(fi, np.arange(P.ni), P.lon[0]).astype(np.float32)
    lat = np.interp(fj, np.arange(P.nj), P.lat[:, 0]).astype(np.float32)
    types = [dict(locrad=[80.0], rfactor=1.0 + t) for t in range(4)]
    return P, fields, fl, dict(fi=fi, fj=fj, fk=fk, typ=typ, status=status, value=value, hops=hops, estd=estd,
                               lon=lon, lat=lat, types=types)


@pytest.mark.parametrize("mode,m,md,land", [(enkfg.MODE_ENKF, 24, 24, 0.0), (enkfg.MODE_ENKF, 96, 96, 0.2),
                                            (enkfg.MODE_HYBRID, 40, 28, 0.1)])
def test_calc_HE_against_oracle(mode, m, md, land):
    print("case calc_HE mode=%d m=%d md=%d land=%.1f" % (mode, m, md, land), flush=True)
    P, fields, fl, o = problem(7 + m, m, land)
    gamma = 0.6
    HA_o, s_o, sd_o = O.getHE(mode, m, md, gamma, P.nlevels, 0, 0, fields, o["hops"], o["typ"], o["status"], o["fi"],
                              o["fj"], o["fk"], o["value"])
    E = enkfg.Engine(m, nmem_dynamic=md, mode=mode, scheme=enkfg.SCHEME_DENKF, gamma=gamma, keep_fp64=True)
    E.set_obstypes(o["types"])
    E.set_obs(o["typ"], o["status"], o["lon"], o["lat"], o["estd"])
    import synth_gpu
    Ecell = np.stack([f.reshape(m, -1).T for f in fields]).astype(np.float32)      # [field][cell][member]
    dE = synth_gpu.DeviceArray.from_host(Ecell)
    E.ens_attach(P.nlevels, P.stride, 0, 0, fl, dE.ptr)
    HA, s, sd = E.calc_HE(P.nlevels, 0, 0, o["hops"], o["fi"], o["fj"], o["fk"], o["value"])
    ok = o["status"] == 0
    assert np.array_equal(HA[:, ok].view(np.uint32), HA_o[:, ok].view(np.uint32))
    assert np.isnan(HA[:, ~ok]).all() and np.isnan(HA_o[:, ~ok]).all()
    assert np.array_equal(s[ok], s_o[ok]) and np.array_equal(sd[ok], sd_o[ok])
    print("  HE of %d obs bit for bit; kernel %.3f ms" % (int(ok.sum()), E.he_ms()), flush=True)
    # the engine's S and s are enkfg_set_S's of the same anomalies and innovations
    ob = O.OracleObs(o["types"], o["typ"], o["lon"], o["lat"], o["estd"])
    HAz = np.where(ok[None, :], HA_o, 0.0).astype(np.float32)
    sz = np.where(ok, s_o, 0.0)
    S_ref, s_ref = O.standardise(m, ob, HAz, sz)
    Sg, sg = E.get_S()
    assert np.array_equal(Sg, S_ref) and np.array_equal(sg, s_ref)
    # and the calc that follows is the oracle's
    ref = O.calc_transforms(mode, enkfg.SCHEME_DENKF, m, md, 1.0, ob, S_ref, s_ref, P.node_lon.ravel(),
                            P.node_lat.ravel())
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    assert np.abs(w64 - ref["w64"]).max() <= 1e-10 * np.abs(ref["w64"]).max()
    assert np.abs(T64 - ref["T64"]).max() <= 1e-10 * np.abs(ref["T64"]).max()
    E.close()


def test_calc_HE_reports_a_bad_forecast():
    """A forecast that is not finite (a NaN member next to the obs) is an
    error, as the reference's enkf_quit at model2obs.c:111-119."""
    print("case calc_HE bad forecast", flush=True)
    m = 24
    P, fields, fl, o = problem(3, m, 0.0)
    j, i = int(o["fj"][100]), int(o["fi"][100])
    f = 0 if o["typ"][100] in (0, 1) else (P.nk if o["typ"][100] == 2 else 2 * P.nk)
    bad = [x.copy() for x in fields]
    bad[f][4, j, i] = np.nan
    if o["typ"][100] in (1, 2):
        for k in range(P.nk):
            bad[f + k][4, j, i] = np.nan
    E = enkfg.Engine(m, mode=enkfg.MODE_ENKF)
    E.set_obstypes(o["types"])
    E.set_obs(o["typ"], o["status"], o["lon"], o["lat"], o["estd"])
    import synth_gpu
    dE = synth_gpu.DeviceArray.from_host(np.stack([x.reshape(m, -1).T for x in bad]).astype(np.float32))
    E.ens_attach(P.nlevels, P.stride, 0, 0, fl, dE.ptr)
    with pytest.raises(enkfg.EnkfgError) as ei:
        E.calc_HE(P.nlevels, 0, 0, o["hops"], o["fi"], o["fj"], o["fk"], o["value"])
    assert ei.value.code == enkfg.ERR_NONFINITE
    E.close()


@pytest.mark.parametrize("kfactor", [0.5, 2.0])
def test_calc_HE_kfactor_moderation(kfactor):
    """das_moderateobs on the device (calc/ensobs.c:749-799): the moderated
    obs errors equal the oracle's bit for bit, the per-type counts of obs whose
    error variance more than doubled are the oracle's, and S and s are
    standardised with the moderated errors."""
    print("case calc_HE kfactor=%g" % kfactor, flush=True)
    m = 24
    P, fields, fl, o = problem(31, m, 0.1)
    HA_o, s_o, sd_o = O.getHE(enkfg.MODE_ENKF, m, m, 1.0, P.nlevels, 0, 0, fields, o["hops"], o["typ"], o["status"],
                              o["fi"], o["fj"], o["fk"], o["value"])
    nt = len(o["types"])
    e_o, cnt_o = O.moderateobs(kfactor, nt, o["typ"], o["status"], o["estd"], sd_o, s_o)
    assert cnt_o.sum() > 0, "the case must moderate some obs substantially"
    E = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, keep_fp64=True)
    E.set_obstypes(o["types"])
    E.set_obs(o["typ"], o["status"], o["lon"], o["lat"], o["estd"])
    E.set_kfactor(kfactor)
    import synth_gpu
    Ecell = np.stack([f.reshape(m, -1).T for f in fields]).astype(np.float32)
    dE = synth_gpu.DeviceArray.from_host(Ecell)
    E.ens_attach(P.nlevels, P.stride, 0, 0, fl, dE.ptr)
    E.calc_HE(P.nlevels, 0, 0, o["hops"], o["fi"], o["fj"], o["fk"], o["value"])
    e, cnt = E.obs_estd()
    assert np.array_equal(e.view(np.uint32), e_o.view(np.uint32))
    assert np.array_equal(cnt, cnt_o)
    print("  %d obs moderated substantially, per type %s" % (int(cnt.sum()), list(cnt)), flush=True)
    ok = o["status"] == 0
    ob = O.OracleObs(o["types"], o["typ"], o["lon"], o["lat"], e_o)
    S_ref, s_ref = O.standardise(m, ob, np.where(ok[None, :], HA_o, 0.0).astype(np.float32), np.where(ok, s_o, 0.0))
    Sg, sg = E.get_S()
    assert np.array_equal(Sg, S_ref) and np.array_equal(sg, s_ref)
    E.close()
