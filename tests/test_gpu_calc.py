"""GPU parity of the calc path (localisation, per-column transforms,
diagnostics) against the CPU oracle, through the C ABI.

Tolerances: local lists (ids and taper coefficients) are compared bit for
bit; pre-cast FP64 T and w within 1e-10 relative to max|T|, max|w| (the
bar set for the plain FP64 path); the FP32 outputs within 2 ulp of the oracle's cast;
dfs/srf (stored as float by the reference) within 1e-6 relative.
"""
import numpy as np
import pytest

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu


def problem(**kw):
    P = synth.make_problem(**kw)
    return P


def oracle_obs(P, status=None, status_build=None):
    return O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd,
                       status=status if status is not None else P.obs_status, status_at_build=status_build)


def engine_for(P, mode, scheme, nmem_dynamic=None, alpha=1.0, pertype=True):
    E = enkfg.Engine(P.nmem, nmem_dynamic=nmem_dynamic, mode=mode, scheme=scheme, alpha=alpha,
                     pertype_diag=pertype, keep_fp64=True)
    E.set_obstypes(P.obstypes)
    E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    return E


@pytest.mark.parametrize("seed,curv,profiles", [(0, False, 0), (1, True, 40)])
def test_localisation_bitwise(seed, curv, profiles):
    print("case localisation seed=%d curv=%s profiles=%d" % (seed, curv, profiles), flush=True)
    P = problem(seed=seed, curvilinear=curv, profiles=profiles, nobs=2500, locrad=80.0)
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF)
    E.localise(P.node_lon, P.node_lat, stride=P.stride)
    counts = E.local_counts()
    ob = oracle_obs(P)
    for k in range(P.node_lon.size):
        ids_o, lc_o = ob.findlocal(P.node_lon.flat[k], P.node_lat.flat[k])
        assert counts[k] == len(ids_o), "node %d count" % k
        ids_g, lc_g = E.localobs(k, counts[k])
        assert np.array_equal(ids_g[:counts[k]], ids_o), "node %d ids" % k
        assert np.array_equal(lc_g[:counts[k]].view(np.uint64), lc_o.view(np.uint64)), "node %d taper bits" % k
    E.close()


def test_localisation_status_visibility_global():
    print("case localisation status/visibility/global", flush=True)
    P = problem(seed=3, profiles=30, nobs=2000, locrad=70.0)
    rng = np.random.default_rng(5)
    status = P.obs_status.copy()
    status[rng.random(P.nobs) < 0.1] = 7          # bad after the trees were built
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF)
    E.set_obs_status(status)
    lon = P.node_lon.copy()
    lat = P.node_lat.copy()
    lon[0, 0] = np.nan                              # a node without coordinates -> all obs
    vis = np.array([1, 0], np.uint8)
    E.localise(lon, lat, stride=P.stride, visible=vis)
    counts = E.local_counts()
    ob = oracle_obs(P, status=status, status_build=P.obs_status)
    for k in range(lon.size):
        ids_o, lc_o = ob.findlocal(lon.flat[k], lat.flat[k], visible=vis)
        assert counts[k] == len(ids_o)
        ids_g, lc_g = E.localobs(k, counts[k])
        assert np.array_equal(ids_g[:counts[k]], ids_o)
        assert np.array_equal(lc_g[:counts[k]].view(np.uint64), lc_o.view(np.uint64))
    E.close()


def test_localisation_one_walk_and_pool_overflow():
    """Localisations after the first walk the trees once, into a pool of
    chunks sized by the previous lists, and compact them; a pool too small
    for the new lists falls back to counting first.  Three localisations of
    one engine: type 1 invisible (two walks), all visible with a global node
    (the pool overflows: two walks again), and the same again (one walk,
    with the global node's list written by the compaction) -- each against
    the oracle, bit for bit."""
    print("case localisation one walk / pool overflow", flush=True)
    P = problem(seed=7, profiles=30, nobs=3000, locrad=90.0)
    rng = np.random.default_rng(11)
    status = P.obs_status.copy()
    status[rng.random(P.nobs) < 0.05] = 7
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF)
    E.set_obs_status(status)
    ob = oracle_obs(P, status=status, status_build=P.obs_status)
    lon = P.node_lon.copy()
    lat = P.node_lat.copy()
    lon[1, 2] = np.nan                              # a node without coordinates -> all obs of each type
    for vis in (np.array([1, 0], np.uint8), np.array([1, 1], np.uint8), np.array([1, 1], np.uint8)):
        E.localise(lon, lat, stride=P.stride, visible=vis)
        counts = E.local_counts()
        assert counts.sum() > 0
        for k in range(lon.size):
            ids_o, lc_o = ob.findlocal(lon.flat[k], lat.flat[k], visible=vis)
            assert counts[k] == len(ids_o), "node %d count" % k
            ids_g, lc_g = E.localobs(k, counts[k])
            assert np.array_equal(ids_g[:counts[k]], ids_o), "node %d ids" % k
            assert np.array_equal(lc_g[:counts[k]].view(np.uint64), lc_o.view(np.uint64))
    E.close()


@pytest.mark.parametrize("precision,nmem", [(enkfg.PREC_FP64, None), (enkfg.PREC_MIXED, None),
                                            (enkfg.PREC_FP64, 144), (enkfg.PREC_MIXED, 144)])
def test_calc_buckets_and_autotune_bitwise(precision, nmem):
    """The calc orders the nodes by obs-count bucket and tunes each bucket's
    Gram tile groups (wavefronts per list: 2 or 3 for m <= 96, 4 or 5 for
    m = 144) on its first large batch, running both.  A node's Gram matrix
    does not depend on the order or the tile groups, so a second calc on the
    same engine (tuned choices) gives T, w and the diagnostics of the first
    (tuning) calc bit for bit; the buckets account for every node."""
    print("case calc buckets / autotune precision=%d nmem=%s" % (precision, nmem), flush=True)
    kw = {} if nmem is None else dict(nmem=nmem)
    P = problem(seed=4, profiles=30, nobs=8000, locrad=150.0, **kw)
    lo = 2 if P.nmem <= 96 else 4
    E = enkfg.Engine(P.nmem, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, precision=precision,
                     pertype_diag=True, keep_fp64=True)
    E.set_obstypes(P.obstypes)
    E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    E.set_S(P.HA, P.innov, standardised=False)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T1, w1, T641, w641 = E.transforms(fp64=True)
    d1 = E.diag()
    tu = E.tuning()
    print("  buckets", tu, flush=True)
    assert sum(b["nodes"] for b in tu) == P.node_lon.size
    key = "groups_fp64" if precision == enkfg.PREC_FP64 else "groups_mixed"
    assert all(b[key] in (0, lo, lo + 1) for b in tu)
    assert any(b[key] in (lo, lo + 1) for b in tu), "no bucket was tuned"
    assert len([b for b in tu if b["nodes"] > 0]) >= 2, "the problem should span several buckets"
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T2, w2, T642, w642 = E.transforms(fp64=True)
    d2 = E.diag()
    assert E.tuning() == tu
    assert np.array_equal(T641.view(np.uint64), T642.view(np.uint64))
    assert np.array_equal(w641.view(np.uint64), w642.view(np.uint64))
    assert np.array_equal(T1.view(np.uint32), T2.view(np.uint32))
    for x, y in zip(d1, d2):
        assert np.array_equal(x, y)
    E.close()


@pytest.mark.parametrize("precision", [enkfg.PREC_FP64, enkfg.PREC_MIXED])
def test_calc_graph_replay_bitwise(precision):
    """ENKFG_FLAG_GRAPH: the calc's Gram/factor batches are captured as one
    HIP graph and replayed.  Every replay -- with eager work (localisation,
    result reads, a new S) between them -- must equal an eager engine's
    result bit for bit."""
    print("case calc graph replay precision=%d" % precision, flush=True)
    P = problem(seed=4, profiles=30, nobs=8000, locrad=150.0)

    def make(graph):
        E = enkfg.Engine(P.nmem, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, precision=precision,
                         pertype_diag=True, keep_fp64=True, graph=graph)
        E.set_obstypes(P.obstypes)
        E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
        E.set_S(P.HA, P.innov, standardised=False)
        return E

    def result(E):
        E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
        T, w, T64, w64 = E.transforms(fp64=True)
        return [T64.view(np.uint64), w64.view(np.uint64), T.view(np.uint32)] + list(E.diag())

    E0, E1 = make(False), make(True)
    ref = result(E0)
    for k in range(4):
        got = result(E1)
        for x, y in zip(got, ref):
            assert np.array_equal(x, y), "replay %d" % k
    # new S (same buffers): the replayed graph reads the current data
    E0.set_S(P.HA * 0.75, P.innov, standardised=False)
    E1.set_S(P.HA * 0.75, P.innov, standardised=False)
    ref = result(E0)
    for k in range(2):
        got = result(E1)
        for x, y in zip(got, ref):
            assert np.array_equal(x, y), "replay after set_S %d" % k
    E0.close()
    E1.close()


@pytest.mark.parametrize("nmax", [1, 17, 60])
def test_localisation_nlobsmax(nmax):
    """NLOBSMAX (observations.c:1899-1908): candidates sorted by distance,
    ties (profile levels at one position) in visiting order, the first
    NLOBSMAX kept, then the search-time status filter."""
    print("case localisation nlobsmax=%d" % nmax, flush=True)
    P = problem(seed=5, profiles=40, nobs=2500, locrad=90.0)
    P.obstypes[0]["nlobsmax"] = nmax
    P.obstypes[1]["nlobsmax"] = nmax // 2 + 1
    rng = np.random.default_rng(9)
    status = P.obs_status.copy()
    status[rng.random(P.nobs) < 0.1] = 7          # bad after the trees were built
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF)
    E.set_obs_status(status)
    E.localise(P.node_lon, P.node_lat, stride=P.stride)
    counts = E.local_counts()
    ob = oracle_obs(P, status=status, status_build=P.obs_status)
    ties = 0
    for k in range(P.node_lon.size):
        ids_o, lc_o = ob.findlocal(P.node_lon.flat[k], P.node_lat.flat[k])
        assert counts[k] == len(ids_o), "node %d count" % k
        ids_g, lc_g = E.localobs(k, counts[k])
        assert np.array_equal(ids_g[:counts[k]], ids_o), "node %d ids" % k
        assert np.array_equal(lc_g[:counts[k]].view(np.uint64), lc_o.view(np.uint64))
        ties += len(lc_o) - len(np.unique(lc_o))
    if nmax > 1:
        assert ties > 0, "the case must contain equal distances"
    # and the transforms with NLOBSMAX lists
    S, s = O.standardise(P.nmem, ob, P.HA, P.innov)
    ref = O.calc_transforms(enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, P.nmem, P.nmem, 1.0, ob, S, s,
                            P.node_lon.ravel(), P.node_lat.ravel())
    E.set_S(P.HA, P.innov, standardised=False)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    assert np.array_equal(E.diag()[0], ref["pnlobs"])
    assert relerr(w64, ref["w64"]) < 1e-10 and relerr(T64, ref["T64"]) < 1e-10
    E.close()


def test_localisation_nlobsmax_long_candidate_lists():
    """NLOBSMAX with more candidates within the radius than the LDS sort
    holds (8192): those lists are sorted in global memory, stably by
    distance, and must still equal the reference's ordered search, with the
    selection and the status filter, entry for entry."""
    print("case localisation nlobsmax, candidate lists above 8192", flush=True)
    P = problem(seed=12, profiles=20, nobs=40000, locrad=220.0, nj=30, ni=45)
    P.obstypes[0]["nlobsmax"] = 10000
    P.obstypes[1]["nlobsmax"] = 1000000
    rng = np.random.default_rng(2)
    status = P.obs_status.copy()
    status[rng.random(P.nobs) < 0.1] = 7
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF)
    E.set_obs_status(status)
    E.localise(P.node_lon, P.node_lat, stride=P.stride)
    counts = E.local_counts()
    ob = oracle_obs(P, status=status, status_build=P.obs_status)
    big = 0
    for k in range(P.node_lon.size):
        ids_o, lc_o = ob.findlocal(P.node_lon.flat[k], P.node_lat.flat[k])
        assert counts[k] == len(ids_o), "node %d count" % k
        ids_g, lc_g = E.localobs(k, counts[k])
        assert np.array_equal(ids_g[:counts[k]], ids_o), "node %d ids" % k
        assert np.array_equal(lc_g[:counts[k]].view(np.uint64), lc_o.view(np.uint64))
        big += int(np.count_nonzero(P.obs_type[ids_o] == 0) > 8192)   # so more candidates than that
    print("  %d nodes, %d with a type-0 list of more than 8192" % (P.node_lon.size, big), flush=True)
    assert big > 0, "the case must hold lists beyond the LDS sort"
    E.close()


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def ulps32(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi).max()


CASES = [
    # mode, scheme, m, md, alpha, profiles
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 24, 24, 1.0, 0),
    (enkfg.MODE_ENKF, enkfg.SCHEME_ETKF, 24, 24, 1.0, 0),
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 96, 96, 0.5, 30),
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 37, 37, 1.0, 10),      # odd m: the last pivot pairs with padding
    (enkfg.MODE_ENKF, enkfg.SCHEME_ETKF, 96, 96, 1.0, 30),
    (enkfg.MODE_ENOI, enkfg.SCHEME_DENKF, 40, 40, 1.0, 0),
    (enkfg.MODE_HYBRID, enkfg.SCHEME_DENKF, 48, 32, 1.0, 20),
    (enkfg.MODE_HYBRID, enkfg.SCHEME_DENKF, 144, 96, 1.0, 20),    # 96 dynamic + 48 static members
    (enkfg.MODE_HYBRID, enkfg.SCHEME_ETKF, 144, 96, 1.0, 20),     # ETKF, m > 96: iterates in device memory
    (enkfg.MODE_ENKF, enkfg.SCHEME_ETKF, 160, 160, 0.8, 10),      # the largest supported m
    # the symmetric sweep past m = 96: 7, 8 and 10 blocks of 16 per side
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 100, 100, 1.0, 10),
    (enkfg.MODE_HYBRID, enkfg.SCHEME_DENKF, 128, 80, 0.7, 10),
    (enkfg.MODE_ENOI, enkfg.SCHEME_DENKF, 150, 150, 1.0, 10),
    (enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 160, 160, 1.0, 10),
]


@pytest.mark.parametrize("mode,scheme,m,md,alpha,profiles", CASES)
def test_calc_transforms_parity(mode, scheme, m, md, alpha, profiles):
    print("case calc mode=%d scheme=%d m=%d md=%d alpha=%g" % (mode, scheme, m, md, alpha), flush=True)
    P = problem(seed=m + md, nmem=m, profiles=profiles, nobs=2500, locrad=75.0, nj=45, ni=60)
    ob = oracle_obs(P)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    ref = O.calc_transforms(mode, scheme, m, md, alpha, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    E = engine_for(P, mode, scheme, nmem_dynamic=md, alpha=alpha)
    E.set_S(P.HA, P.innov, standardised=False)
    Sg, sg = E.get_S()
    assert np.array_equal(Sg, S), "standardised S"
    assert np.array_equal(sg, s), "standardised s"
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    pn, pd, ps = E.diag()
    st = E.stats()
    assert st == ref["stats"]
    assert np.array_equal(pn, ref["pnlobs"])
    assert relerr(w64, ref["w64"]) < 1e-10
    assert ulps32(w, ref["w"]) <= 2 or relerr(w, ref["w"]) < 1e-6
    if mode != enkfg.MODE_ENOI:
        assert relerr(T64, ref["T64"]) < 1e-10
        assert np.abs(T.astype(np.float64) - ref["T"]).max() <= 2.4e-7 * np.abs(ref["T"]).max()
    nz = ref["pdfs"] != 0
    assert np.allclose(pd, ref["pdfs"], rtol=1e-6, atol=1e-7)
    assert np.allclose(ps[nz], ref["psrf"][nz], rtol=1e-5, atol=1e-6)
    E.close()



def test_calc_without_pertype_diagnostics():
    """ENKFG_FLAG_PERTYPE_DIAG off: the per-type rows of the diagnostics hold
    the local obs counts with zero dfs and srf; T, w and the all-types row
    match the oracle as with the flag on."""
    m = 96
    print("case calc DEnKF m=96 per-type diagnostics off", flush=True)
    P = problem(seed=7, nmem=m, profiles=30, nobs=2500, locrad=75.0, nj=45, ni=60)
    ob = oracle_obs(P)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    ref = O.calc_transforms(enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, m, m, 1.0, ob, S, s, P.node_lon.ravel(),
                            P.node_lat.ravel())
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, pertype=False)
    E.set_S(P.HA, P.innov, standardised=False)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    pn, pd, ps = E.diag()
    nt = len(P.obstypes)
    assert relerr(w64, ref["w64"]) < 1e-10 and relerr(T64, ref["T64"]) < 1e-10
    assert np.array_equal(pn, ref["pnlobs"])
    assert not np.any(pd[:nt]) and not np.any(ps[:nt])
    assert np.allclose(pd[nt], ref["pdfs"][nt], rtol=1e-6, atol=1e-7)
    nz = ref["pdfs"][nt] != 0
    assert np.allclose(ps[nt][nz], ref["psrf"][nt][nz], rtol=1e-5, atol=1e-6)
    E.close()

def test_calc_empty_columns_and_stats():
    print("case calc empty columns", flush=True)
    P = problem(seed=11, nobs=40, locrad=20.0, nj=30, ni=30)
    ob = oracle_obs(P)
    S, s = O.standardise(P.nmem, ob, P.HA, P.innov)
    ref = O.calc_transforms(1, 1, P.nmem, P.nmem, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    assert (ref["pnlobs"][-1] == 0).any(), "fixture must contain empty columns"
    E = engine_for(P, enkfg.MODE_ENKF, enkfg.SCHEME_DENKF)
    E.set_S(S, s, standardised=True)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    T, w, T64, w64 = E.transforms(fp64=True)
    empty = ref["pnlobs"][-1] == 0
    eye = np.eye(P.nmem, dtype=np.float32)
    assert all(np.array_equal(T[k], eye) for k in np.where(empty)[0])
    assert not w[empty].any()
    assert E.stats() == ref["stats"]
    E.close()
