"""The CPU oracle against the golden vectors written by the reference's own
calc/calcs.c and common/kdtree.c (tests/golden/make_golden.py).

Tolerances: kd-tree lists bit for bit; FP64 transforms within 1e-12 of
max|T| (the oracle's Cholesky and Jacobi replace MKL's dpotrf/dpotri and
dgesvd); FP32 outputs within 1 ulp; the float-stored dfs/srf within 2e-7
relative.
"""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_local_analyses_match_reference_calcs():
    g = load("calcs.npz")
    n = int(g["n"])
    assert n >= 30
    seen = set()
    for k in range(n):
        c = {key: g["%d_%s" % (k, key)] for key in ("m", "mout", "p", "mode", "scheme", "S", "lc", "s", "T", "w",
                                                   "dfs", "srf")}
        m, mout, p, mode, scheme = (int(c[x]) for x in ("m", "mout", "p", "mode", "scheme"))
        seen.add((p < m, p == m, p > m))
        Sloc = c["S"].astype(np.float64) * c["lc"][:, None]
        sloc = c["s"] * c["lc"]
        T, w, dfs, srf = O.local_analysis(mode, scheme, m, mout, Sloc, sloc)
        assert relmax(w, c["w"]) < 1e-12, (k, m, p, mode, scheme)
        if mode != O.MODE_ENOI:
            assert relmax(T, c["T"]) < 1e-12, (k, m, p, mode, scheme)
        assert abs(dfs - float(c["dfs"])) <= 1e-12 * max(1.0, abs(float(c["dfs"])))
        assert abs(srf - float(c["srf"])) <= 1e-10 * max(1.0, abs(float(c["srf"])))
    assert len(seen) == 3, "fixtures must cover obs-space, square and ensemble-space analyses"


def test_kdtree_lists_match_reference_kdtree():
    g = load("kdtree.npz")
    for name in ("geo", "flat"):
        t = O.OracleKd(g[name + "_pts"])
        rng = float(g[name + "_range"])
        for k, q in enumerate(g[name + "_q"]):
            for ordered in (0, 1):
                d, d2 = t.range(q, rng, ordered=bool(ordered))
                assert np.array_equal(d, g["%s_%d_%d_data" % (name, k, ordered)]), (name, k, ordered)
                assert np.array_equal(d2.view(np.uint64), g["%s_%d_%d_d2" % (name, k, ordered)].view(np.uint64))


@pytest.mark.parametrize("tag", ["m1_s1", "m1_s2", "m2_s1"])
def test_calc_loop_matches_reference(tag):
    g = load("calcloop.npz")
    types = [dict(locrad=[float(r)], rfactor=float(f)) for r, f in zip(g["locrad"], g["rfactor"])]
    ob = O.OracleObs(types, g["obs_type"], g["obs_lon"], g["obs_lat"], g["obs_estd"], status=g["obs_status"])
    S, s = O.standardise(24, ob, g["HA"], g["innov"])
    mode, scheme = int(tag[1]), int(tag[4])
    alpha = float(g[tag + "_alpha"])
    r = O.calc_transforms(mode, scheme, 24, 24, alpha, ob, S, s, g["node_lon"].ravel(), g["node_lat"].ravel())
    assert np.array_equal(r["pnlobs"], g[tag + "_pnlobs"])
    for key in ("nlobs_sum", "nlobs_max", "n_inv_obs", "n_inv_ens", "ncell"):
        assert r["stats"][key] == int(g[tag + "_stat_" + key])
    assert relmax(r["w64"], g[tag + "_w64"]) < 1e-12
    assert np.abs(r["w"].view(np.int32).astype(np.int64) - g[tag + "_w"].view(np.int32)).max() <= 1
    if mode != O.MODE_ENOI:
        assert relmax(r["T64"], g[tag + "_T64"]) < 1e-12
        assert np.abs(r["T"].view(np.int32).astype(np.int64) - g[tag + "_T"].view(np.int32)).max() <= 1
    assert np.allclose(r["pdfs"], g[tag + "_pdfs"], rtol=2e-7, atol=0)
    assert np.allclose(r["psrf"], g[tag + "_psrf"], rtol=2e-7, atol=1e-7)
    # the fixture exercises empty columns, obs- and ensemble-space analyses
    # and both obstypes
    assert (g[tag + "_pnlobs"][-1] == 0).any()
    assert int(g[tag + "_stat_n_inv_obs"]) > 0 and int(g[tag + "_stat_n_inv_ens"]) > 0
    assert (g[tag + "_pnlobs"][1] > 0).any()


def test_standardise_fixture_inputs():
    """das_standardise (calc/ensobs.c:849-898) on the fixture: S and s divided
    by estd * sqrt(rfactor) * sqrt(m - 1), S rounded to float."""
    g = load("calcloop.npz")
    types = [dict(locrad=[float(r)], rfactor=float(f)) for r, f in zip(g["locrad"], g["rfactor"])]
    ob = O.OracleObs(types, g["obs_type"], g["obs_lon"], g["obs_lat"], g["obs_estd"], build_trees=False)
    S, s = O.standardise(24, ob, g["HA"], g["innov"])
    div = g["obs_estd"].astype(np.float64) * np.sqrt(g["rfactor"][g["obs_type"]]) * np.sqrt(23.0)
    assert np.array_equal(S, (g["HA"].astype(np.float64) / div[None, :]).astype(np.float32))
    assert np.array_equal(s, g["innov"] / div)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree and MKL)")
def test_reference_harness_reports_lapack_failure():
    """calcs.c calls enkf_quit when dpotrf fails (calcs.c:109-118).  The
    harness turns that into an error status with the reference's message,
    and the process carries on: the next analysis on the same thread works."""
    rng = np.random.default_rng(3)
    S = rng.standard_normal((40, 12)) * 0.3
    bad = S.copy()
    bad[5, 3] = np.nan
    with pytest.raises(RuntimeError, match="dpotrf"):
        O.local_analysis(O.MODE_ENKF, O.SCHEME_DENKF, 12, 12, bad, np.ones(40), use_ref=True)
    T, w, _, _ = O.local_analysis(O.MODE_ENKF, O.SCHEME_DENKF, 12, 12, S, np.ones(40), use_ref=True)
    To, wo, _, _ = O.local_analysis(O.MODE_ENKF, O.SCHEME_DENKF, 12, 12, S, np.ones(40))
    assert relmax(T, To) < 1e-12 and relmax(w, wo) < 1e-12


@pytest.mark.parametrize("use_ref", [False, True])
def test_threaded_calc_loop_equals_serial(use_ref):
    """The calc loop split over threads (disjoint node runs, per-thread
    reference trees) writes what the serial loop writes, bit for bit."""
    if use_ref and not O.ref_available():
        pytest.skip("oracle/_ref not built")
    import synth
    P = synth.make_problem(seed=8, nj=24, ni=30, nmem=16, nobs=1500, locrad=90.0, profiles=20)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(P.nmem, ob, P.HA, P.innov)
    args = (O.MODE_ENKF, O.SCHEME_ETKF, P.nmem, P.nmem, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    a = O.calc_transforms(*args, use_ref=use_ref)
    b = O.calc_transforms(*args, use_ref=use_ref, threads=5)
    for k in ("T64", "w64", "pnlobs", "pdfs", "psrf"):
        assert np.array_equal(a[k], b[k]), k
    assert a["stats"] == b["stats"]
