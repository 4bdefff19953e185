"""Parity at the workloads of the benchmark configurations, through the C ABI.

  C1  the whole CPU-runnable case: 100 x 200 grid at 0.1 deg, 10 layers, m =
      24, 20k surface obs, LOCRAD 200 km, stride 3, ETKF, FP64; calc for all
      2,278 nodes and the update of all 21 fields with inflation 1.05.
  C2/C3  a band of the 1500 x 1800 x 51 tile (its geometry and obs density:
      SST, SLA tracks, 51-level T/S profiles, LOCRAD 250 km, m = 96) whose
      local lists fall in the [1024, 4096) and [4096, ...) obs-count buckets;
      DEnKF and ETKF, FP64 (T64, w64 within 1e-10 of the oracle) and mixed
      precision (T, w, dfs, and the analysis mean and spread of a window of
      the grid against the oracle's transforms and update).
  C4  hybrid 96 + 48 members (m = 144) with clustered obs (log-normal local
      counts), FP64 and mixed precision, the m > 96 Gram kernels.

The oracle computes a sample of nodes chosen per obs-count bucket (threads
over nodes); the engine computes every node of the problem, so each node is
run in the batch, bucket and tile shape the full calc gives it.
"""
import math
import os

import numpy as np
import pytest

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
BUCKETS = (0, 256, 1024, 4096)


def relerr(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(float(np.abs(b).max()), 1e-300))


def ulps32(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return int(np.abs(ai - bi).max())


def bucket_of(pnlobs, ntypes):
    """obs-count bucket of each node: its largest (node, obstype) list"""
    mx = pnlobs[:ntypes].max(axis=0)
    return np.searchsorted(np.array(BUCKETS), mx, side="right") - 1


def sample_nodes(bk, per_bucket, rng, extra=()):
    """up to per_bucket nodes of every non-empty bucket, plus `extra`"""
    pick = set(int(k) for k in extra)
    for q in range(len(BUCKETS)):
        idx = np.flatnonzero(bk == q)
        if len(idx):
            pick.update(int(k) for k in rng.choice(idx, size=min(per_bucket, len(idx)), replace=False))
    return np.array(sorted(pick), np.int64)


# ---------------------------------------------------------------------------- C1
def test_c1_whole_etkf_fp64_calc_and_update():
    """C1 in full: every node against the oracle (T64, w64 within 1e-10,
    FP32 T and w within 2 ulp, diagnostics, calcstats), then the update of
    all 21 fields with the engine's transforms against the oracle update."""
    print("case C1 whole: 100x200x10, m=24, 20k obs, ETKF FP64", flush=True)
    P = synth.make_problem(nj=100, ni=200, nk=10, nmem=24, nobs=20000, corr_km=150.0, std=(0.5, 0.05, 0.05),
                           obs_std=0.3, locrad=200.0, stride=3, seed=0)
    assert P.node_lon.size == 2278
    m = P.nmem
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    ref = O.calc_transforms(enkfg.MODE_ENKF, enkfg.SCHEME_ETKF, m, m, 1.0, ob, S, s, P.node_lon.ravel(),
                            P.node_lat.ravel(), threads=THREADS)
    E = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_ETKF, pertype_diag=True, keep_fp64=True)
    E.set_obstypes(P.obstypes)
    E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    E.set_S(P.HA, P.innov, standardised=False)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    pn, pd, ps = E.diag()
    st = E.stats()
    print("  stats", st, "  T64 %.2e  w64 %.2e" % (relerr(T64, ref["T64"]), relerr(w64, ref["w64"])), flush=True)
    assert st == ref["stats"]
    assert st["nlobs_max"] > 1000, "C1 lists are about a thousand obs"
    assert np.array_equal(pn, ref["pnlobs"])
    assert relerr(T64, ref["T64"]) < 1e-10 and relerr(w64, ref["w64"]) < 1e-10
    assert ulps32(T, ref["T"]) <= 2
    assert ulps32(w, ref["w"]) <= 2 or relerr(w, ref["w"]) < 1e-6
    nz = ref["pdfs"] != 0
    assert np.allclose(pd, ref["pdfs"], rtol=1e-6, atol=1e-7)
    assert np.allclose(ps[nz], ref["psrf"][nz], rtol=1e-5, atol=1e-6)
    # update of the 21 fields, inflation 1.05 (capped, ratio 1: the default)
    bufs, fl = [], []
    for name, nk in (("temp", P.nk), ("salt", P.nk), ("eta", 1)):
        for k in range(nk):
            bufs.append(np.ascontiguousarray(P.ens[name][:, k]))
            fl.append(dict(level=k, issurfacevar=int(name == "eta"), inflation=1.05, inf_ratio=1.0))
    assert len(bufs) == 21
    gpu = [b.copy() for b in bufs]
    cpu = [b.copy() for b in bufs]
    E.update_fields(P.nlevels, P.stride, 0, 0, fl, gpu, 0, P.nj - 1)
    E.close()
    O.update_fields(enkfg.MODE_ENKF, m, m, P.stride, 0, 0, P.nlevels, ref["T"].reshape(P.nj_s, P.ni_s, m, m),
                    ref["w"].reshape(P.nj_s, P.ni_s, m), fl, cpu, 0, P.nj - 1)
    worst = max(float(np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()) for a, b in zip(gpu, cpu))
    print("  update: max rel error %.2e over 21 fields" % worst, flush=True)
    assert worst <= 1e-5


# ---------------------------------------------------------------------------- C2 / C3 band
_band = {}


def band():
    """240 rows x 900 columns at the southern edge of the tile geometry
    (lat -75 .. -51): SST lists of about 3k (north) to 9k (south) obs."""
    if "X" not in _band:
        import synth_gpu
        frac = 240 * 900 / (1500 * 1800)
        X = synth_gpu.make_tile(nj=240, ni=900, nk=51, nmem=96, n_sst=int(4_200_000 * frac),
                                n_sla=int(2_100_000 * frac), n_prof=int(6_900 * frac), seed=3, row0=0, nj_map=1500)
        _band["X"] = X
        _band["ob"] = O.OracleObs(X.obstypes, X.obs_type, X.obs_lon, X.obs_lat, X.obs_estd)
        _band["S"] = O.standardise(X.nmem, _band["ob"], X.HA, X.innov)
    return _band["X"]


def band_oracle(scheme, nodes):
    key = ("ref", scheme)
    X = band()
    if key not in _band:
        S, s = _band["S"]
        lon, lat = X.node_lon.ravel()[nodes], X.node_lat.ravel()[nodes]
        _band[key] = O.calc_transforms(enkfg.MODE_ENKF, scheme, X.nmem, X.nmem, 1.0, _band["ob"], S, s, lon, lat,
                                       threads=THREADS)
    return _band[key]


def band_sample():
    """the 2 x 24 block of stride nodes of the update window (rows jb, jb + 1,
    columns ib .. ib + 23); the per-bucket sample is added to it once the
    engine has given every node's bucket"""
    if "block" not in _band:
        X = band()
        jb, ib = X.nj_s // 3, X.ni_s // 2
        _band["block"] = (jb, ib)
        _band["block_nodes"] = np.array([(jb + r) * X.ni_s + ib + c for r in range(2) for c in range(24)])
    return _band["block_nodes"]


CASES_BAND = [(enkfg.SCHEME_DENKF, enkfg.PREC_FP64), (enkfg.SCHEME_ETKF, enkfg.PREC_FP64),
              (enkfg.SCHEME_DENKF, enkfg.PREC_MIXED), (enkfg.SCHEME_ETKF, enkfg.PREC_MIXED)]


@pytest.mark.parametrize("scheme,precision", CASES_BAND)
def test_c2_c3_band_against_oracle(scheme, precision):
    X = band()
    block = band_sample()
    m = X.nmem
    nt = len(X.obstypes)
    print("case C2/C3 band scheme=%d precision=%d: %d obs, %d x %d nodes" % (scheme, precision, X.nobs, X.nj_s,
                                                                         X.ni_s), flush=True)
    E = enkfg.Engine(m, mode=enkfg.MODE_ENKF, scheme=scheme, precision=precision, pertype_diag=True,
                     keep_fp64=True)
    E.set_obstypes(X.obstypes)
    E.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
    E.set_S(X.HA, X.innov, standardised=False)
    E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    pn, pd, ps = E.diag()
    tu = E.tuning()
    E.close()
    bk = bucket_of(pn, nt)
    print("  buckets (nodes):", [int((bk == q).sum()) for q in range(4)], " tuning", tu, flush=True)
    assert (bk == 2).sum() > 0 and (bk == 3).sum() > 0, "the band must fill the >= 1024 and >= 4096 buckets"
    if "nodes" not in _band:
        _band["nodes"] = sample_nodes(bk, 12, np.random.default_rng(5), extra=block)
    nodes = _band["nodes"]
    ref = band_oracle(scheme, nodes)
    print("  oracle on %d nodes; buckets of the sample:" % len(nodes),
          [int((bk[nodes] == q).sum()) for q in range(4)], " max list %d" % ref["pnlobs"][nt].max(), flush=True)
    assert np.array_equal(pn[:, nodes], ref["pnlobs"])
    eT = relerr(T64[nodes], ref["T64"])
    ew = relerr(w64[nodes], ref["w64"])
    nz = ref["pdfs"] != 0
    edfs = float(np.abs(pd[:, nodes][nz] - ref["pdfs"][nz]).max() / np.abs(ref["pdfs"][nz]).max())
    per_bucket = []
    for q in range(4):
        sel = bk[nodes] == q
        if sel.any():
            per_bucket.append((BUCKETS[q], relerr(T64[nodes][sel], ref["T64"][sel]),
                               relerr(w64[nodes][sel], ref["w64"][sel])))
    print("  T64 %.3g  w64 %.3g  dfs %.3g  per bucket (lo, T, w): %s" % (eT, ew, edfs, per_bucket), flush=True)
    if precision == enkfg.PREC_FP64:
        assert eT < 1e-10 and ew < 1e-10
        # the FP32 cast of FP64 values 1e-13 apart: within one float rounding of max |T|
        assert np.abs(T[nodes].astype(np.float64) - ref["T"]).max() <= 2.4e-7 * np.abs(ref["T"]).max()
        assert np.allclose(pd[:, nodes], ref["pdfs"], rtol=1e-6, atol=1e-7)
        assert np.allclose(ps[:, nodes][nz], ref["psrf"][nz], rtol=1e-5, atol=1e-6)
        return
    # mixed precision: transforms, then the analysis of a window of the grid
    assert eT <= 2e-3 and ew <= 2e-2 and edfs <= 1e-3
    jb, ib = _band["block"]
    pos = {int(k): i for i, k in enumerate(nodes)}
    bi = [pos[(jb + r) * X.ni_s + ib + c] for r in range(2) for c in range(24)]
    To = ref["T"][bi].reshape(2, 24, m, m)
    wo = ref["w"][bi].reshape(2, 24, m)
    Tm = T[[(jb + r) * X.ni_s + ib + c for r in range(2) for c in range(24)]].reshape(2, 24, m, m)
    wm = w[[(jb + r) * X.ni_s + ib + c for r in range(2) for c in range(24)]].reshape(2, 24, m)
    r0, c0 = jb * X.stride, ib * X.stride
    rows, cols = X.stride + 1, 23 * X.stride + 1
    nl = np.ascontiguousarray(X.nlevels[r0:r0 + rows, c0:c0 + cols])
    fsel = [0, 10, 30, 51, 70, 102]          # temp 0, 10, 30; salt 0, 19; eta
    bufs, fl = [], []
    for f in fsel:
        cells = ((r0 + np.arange(rows))[:, None] * X.ni + (c0 + np.arange(cols))[None, :]).ravel()
        a = np.stack([X.E.read((f * X.ncell + c) * m, m) for c in cells])       # [cell][m]
        bufs.append(np.ascontiguousarray(a.T).reshape(m, rows, cols))
        fl.append(X.fields[f])
    A = [b.copy() for b in bufs]
    B = [b.copy() for b in bufs]
    Em = enkfg.Engine(m, mode=enkfg.MODE_ENKF, precision=enkfg.PREC_MIXED)
    Em.set_transforms(Tm, wm, X.stride)
    Em.update_fields(nl, X.stride, 0, 0, fl, A, 0, rows - 1)
    Em.close()
    O.update_fields(enkfg.MODE_ENKF, m, m, X.stride, 0, 0, nl, To, wo, fl, B, 0, rows - 1)
    a = np.stack(A).astype(np.float64)
    b = np.stack(B).astype(np.float64)
    f0 = np.stack(bufs).astype(np.float64)
    ma, mb, mf = a.mean(1), b.mean(1), f0.mean(1)
    emean = float(np.abs(ma - mb).max() / np.abs(mb).max())
    einc = float(np.abs(ma - mb).max() / np.abs(mb - mf).max())
    espread = float(np.abs(a.std(1) - b.std(1)).max() / np.abs(b.std(1)).max())
    print("  analysis (mixed transforms + mixed update vs oracle): mean %.3g (of max |mean|), %.3g (of max "
          "|increment|), spread %.3g" % (emean, einc, espread), flush=True)
    assert emean <= 1e-5
    assert espread <= 1e-3
    assert einc <= 1e-2


# ---------------------------------------------------------------------------- C4
_c4 = {}


def c4_problem():
    if "X" not in _c4:
        import synth_gpu
        X = synth_gpu.make_clustered_tile(nj=360, ni=360, nk=8, nmem=144, nmem_dynamic=96, median=400.0,
                                          pmax=6000.0, locrad=150.0, seed=4)
        _c4["X"] = X
        _c4["ob"] = O.OracleObs(X.obstypes, X.obs_type, X.obs_lon, X.obs_lat, X.obs_estd)
        _c4["S"] = O.standardise(X.nmem, _c4["ob"], X.HA, X.innov)
    return _c4["X"]


@pytest.mark.parametrize("precision", [enkfg.PREC_FP64, enkfg.PREC_MIXED])
def test_c4_hybrid_m144_clustered_against_oracle(precision):
    X = c4_problem()
    m, md = X.nmem, X.nmem_dynamic
    nt = len(X.obstypes)
    print("case C4 hybrid %d+%d precision=%d: %d obs, %d nodes" % (md, m - md, precision, X.nobs,
                                                               X.node_lon.size), flush=True)
    E = enkfg.Engine(m, nmem_dynamic=md, mode=enkfg.MODE_HYBRID, scheme=enkfg.SCHEME_DENKF, precision=precision,
                     pertype_diag=True, keep_fp64=True)
    E.set_obstypes(X.obstypes)
    E.set_obs(X.obs_type, X.obs_status, X.obs_lon, X.obs_lat, X.obs_estd)
    E.set_S(X.HA, X.innov, standardised=False)
    E.calc_transforms(X.node_lon, X.node_lat, stride=X.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    pn, pd, ps = E.diag()
    E.close()
    tot = pn[nt]
    bk = bucket_of(pn, nt)
    nzl = tot[tot > 0]
    print("  local obs per column: median %.0f  max %d  (log-sd %.2f); buckets %s" % (
        np.median(nzl), tot.max(), np.log(nzl).std(), [int((bk == q).sum()) for q in range(4)]), flush=True)
    assert (bk >= 2).sum() > 0, "clustered obs must give lists of >= 1024"
    if "nodes" not in _c4:
        _c4["nodes"] = sample_nodes(bk, 10, np.random.default_rng(7))
    nodes = _c4["nodes"]
    if "ref" not in _c4:
        S, s = _c4["S"]
        _c4["ref"] = O.calc_transforms(enkfg.MODE_HYBRID, enkfg.SCHEME_DENKF, m, md, 1.0, _c4["ob"], S, s,
                                       X.node_lon.ravel()[nodes], X.node_lat.ravel()[nodes], threads=THREADS)
    ref = _c4["ref"]
    assert np.array_equal(pn[:, nodes], ref["pnlobs"])
    eT = relerr(T64[nodes], ref["T64"])
    ew = relerr(w64[nodes], ref["w64"])
    nz = ref["pdfs"] != 0
    edfs = float(np.abs(pd[:, nodes][nz] - ref["pdfs"][nz]).max() / np.abs(ref["pdfs"][nz]).max())
    print("  T64 %.3g  w64 %.3g  dfs %.3g on %d sampled nodes" % (eT, ew, edfs, len(nodes)), flush=True)
    if precision == enkfg.PREC_FP64:
        assert eT < 1e-10 and ew < 1e-10
        assert np.allclose(pd[:, nodes], ref["pdfs"], rtol=1e-6, atol=1e-7)
    else:
        # dfs of the per-type rows with few obs carries the relative error of
        # the BF16 hi/lo Gram (measured 1.9e-3 at m = 144)
        assert eT <= 2e-3 and ew <= 2e-2 and edfs <= 5e-3
