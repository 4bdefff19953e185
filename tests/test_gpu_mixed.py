"""The mixed-precision calc path (S^T S as BF16 hi/lo MFMA products with FP32
accumulation, FP64 factorisation) against the FP64 oracle.

Mixed precision is not bit-compatible with the reference by design; each
output reports its error against the oracle and is held to a stated bound:

  T, w       max |mixed - oracle| / max |oracle|  <= 2e-4 (T), 2e-3 (w)
  analysis   the update applied with the mixed T, w versus the oracle's
             T, w (both through the oracle's update): analysis-mean error
             <= 1e-5 relative to max |mean|, spread error <= 1e-4 relative
  local lists and diagnostic counts stay exact (the search is shared).
"""
import numpy as np
import pytest

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


CASES = [
    # mode, scheme, m, md, profiles
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 24, 24, 0),
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 96, 96, 30),
    (enkfg.MODE_ENKF, enkfg.SCHEME_ETKF, 96, 96, 30),
    (enkfg.MODE_ENOI, enkfg.SCHEME_DENKF, 40, 40, 0),
    (enkfg.MODE_HYBRID, enkfg.SCHEME_DENKF, 48, 32, 20),
]


@pytest.mark.parametrize("mode,scheme,m,md,profiles", CASES)
def test_mixed_transforms_and_analysis_error(mode, scheme, m, md, profiles):
    print("case mixed mode=%d scheme=%d m=%d md=%d" % (mode, scheme, m, md), flush=True)
    P = synth.make_problem(seed=100 + m, nmem=m, profiles=profiles, nobs=3000, locrad=90.0, nj=36, ni=48)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    ref = O.calc_transforms(mode, scheme, m, md, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    E = enkfg.Engine(m, nmem_dynamic=md, mode=mode, scheme=scheme, precision=enkfg.PREC_MIXED, keep_fp64=True)
    E.set_obstypes(P.obstypes)
    E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    E.set_S(P.HA, P.innov, standardised=False)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    pn, pd, ps = E.diag()
    E.close()
    assert np.array_equal(pn, ref["pnlobs"])
    ew = relerr(w64, ref["w64"])
    eT = relerr(T64, ref["T64"]) if mode != enkfg.MODE_ENOI else 0.0
    nz = ref["pdfs"] != 0
    edfs = float(np.abs(pd[nz] - ref["pdfs"][nz]).max() / np.abs(ref["pdfs"][nz]).max())
    print("  max rel error: T %.3g  w %.3g  dfs %.3g" % (eT, ew, edfs), flush=True)
    assert eT <= 2e-4
    assert ew <= 2e-3
    assert edfs <= 1e-3
    if mode == enkfg.MODE_ENOI:
        return
    # analysis with the mixed transforms versus the oracle's, same update
    nj_s, ni_s = P.nj_s, P.ni_s
    Tm = T.reshape(nj_s, ni_s, md, m)
    wm = w.reshape(nj_s, ni_s, m)
    To = ref["T"].reshape(nj_s, ni_s, md, m)
    wo = ref["w"].reshape(nj_s, ni_s, m)
    fl = [dict(level=k) for k in range(P.nk)]
    a = [np.ascontiguousarray(P.ens["temp"][:m, k]) for k in range(P.nk)]
    b = [x.copy() for x in a]
    O.update_fields(mode, m, md, P.stride, 0, 0, P.nlevels, Tm, wm, fl, a, 0, P.nj - 1)
    O.update_fields(mode, m, md, P.stride, 0, 0, P.nlevels, To, wo, fl, b, 0, P.nj - 1)
    A = np.stack(a)[:, :md].astype(np.float64)
    B = np.stack(b)[:, :md].astype(np.float64)
    emean = relerr(A.mean(1), B.mean(1))
    espread = relerr(A.std(1), B.std(1))
    print("  analysis: mean %.3g  spread %.3g" % (emean, espread), flush=True)
    assert emean <= 1e-5
    assert espread <= 1e-4


@pytest.mark.parametrize("mode,m,md,stride", [(enkfg.MODE_ENKF, 96, 96, 3), (enkfg.MODE_HYBRID, 48, 32, 2),
                                              (enkfg.MODE_ENKF, 24, 24, 1)])
def test_mixed_update_error(mode, m, md, stride):
    """The FP16 hi/lo MFMA apply against the oracle's FP32 update with the
    same T and w: every output within 1e-5 of the field's magnitude, masks
    and missing values exact."""
    print("case mixed update mode=%d m=%d md=%d stride=%d" % (mode, m, md, stride), flush=True)
    P = synth.make_problem(seed=200 + m, nmem=m, nobs=1500, locrad=70.0, nj=25, ni=31, stride=stride,
                           land_frac=0.2, nmem_static=m - md)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    r = O.calc_transforms(mode, enkfg.SCHEME_DENKF, m, md, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    T = r["T"].reshape(P.nj_s, P.ni_s, md, m)
    w = r["w"].reshape(P.nj_s, P.ni_s, m)
    bufs, fl = [], []
    for k in range(P.nk):
        bufs.append(np.ascontiguousarray(P.ens["temp"][:m, k]))
        fl.append(dict(level=k, inflation=1.1, inf_ratio=1.0))
        bufs.append(np.ascontiguousarray(P.ens["salt"][:m, k]))
        fl.append(dict(level=k, inflation=1.0, inf_ratio=float("nan")))
    bufs.append(np.ascontiguousarray(P.ens["eta"][:m, 0]))
    fl.append(dict(level=0, issurfacevar=1, inflation=1.05, inf_ratio=0.5))
    bufs[0][min(3, md - 1), 5, 7] = np.nan
    ref = [b.copy() for b in bufs]
    gpu = [b.copy() for b in bufs]
    O.update_fields(mode, m, md, stride, 0, 0, P.nlevels, T, w, fl, ref, 0, P.nj - 1)
    E = enkfg.Engine(m, nmem_dynamic=md, mode=mode, precision=enkfg.PREC_MIXED)
    E.set_transforms(T, w, stride)
    E.update_fields(P.nlevels, stride, 0, 0, fl, gpu, 0, P.nj - 1)
    E.close()
    worst = 0.0
    for a, b in zip(gpu, ref):
        assert np.array_equal(a == 0, b == 0)
        worst = max(worst, float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1.0)))
    A = np.stack(gpu)[:, :md].astype(np.float64)
    B = np.stack(ref)[:, :md].astype(np.float64)
    emean = float(np.abs(A.mean(1) - B.mean(1)).max() / np.abs(B.mean(1)).max())
    espread = float(np.abs(A.std(1) - B.std(1)).max() / np.abs(B.std(1)).max())
    print("  max error %.3g  mean %.3g  spread %.3g" % (worst, emean, espread), flush=True)
    assert worst <= 1e-5
    assert emean <= 1e-5
    assert espread <= 1e-4
