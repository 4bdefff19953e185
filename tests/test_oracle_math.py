"""The CPU oracle against what the mathematics fixes: published closed forms,
textbook library routines (numpy.linalg), brute force on small inputs and
invariants of the update.  Each check is chosen so that a dropped term, a
wrong sign, a transposed operand or an off-by-one index fails it.
"""
import math

import numpy as np
import pytest

import oracle as O
import synth

R_EARTH = 6371.0


# ---------------------------------------------------------------- taper
def gc_textbook(r, c):
    """Gaspari & Cohn (1999), QJRMS 125, eq. (4.10), half-width c."""
    z = abs(r) / c
    if z <= 1.0:
        return -0.25 * z ** 5 + 0.5 * z ** 4 + 0.625 * z ** 3 - 5.0 / 3.0 * z ** 2 + 1.0
    if z <= 2.0:
        return z ** 5 / 12.0 - 0.5 * z ** 4 + 0.625 * z ** 3 + 5.0 / 3.0 * z ** 2 - 5.0 * z + 4.0 - 2.0 / (3.0 * z)
    return 0.0


def test_taper_closed_forms():
    assert O.taper_gc(0.0) == 1.0
    assert O.taper_gc(1.0) == 0.0 and O.taper_gc(1.5) == 0.0
    assert abs(O.taper_gc(0.5) - 5.0 / 24.0) < 1e-15
    # both branches meet at x = 1/2 in value and slope; the taper and its
    # slope vanish at x = 1
    h = 1e-7
    assert abs(O.taper_gc(0.5 - h) - O.taper_gc(0.5 + h)) < 1e-6
    d_lo = (O.taper_gc(0.5 - h) - O.taper_gc(0.5 - 2 * h)) / h
    d_hi = (O.taper_gc(0.5 + 2 * h) - O.taper_gc(0.5 + h)) / h
    assert abs(d_lo - d_hi) < 1e-5
    assert O.taper_gc(1.0 - h) < 1e-12
    for x in np.linspace(0.0, 0.999, 301):
        assert abs(O.taper_gc(x) - gc_textbook(x, 0.5)) < 2e-15, x


def test_calclcoeff_weighted_sum():
    """obstypes.c:800-822: sum over radii of weight * GC(dist / radius)."""
    for d in (0.0, 10.0, 49.0, 120.0, 199.0, 260.0):
        assert O.calclcoeff([200.0], [1.0], d) == O.taper_gc(d / 200.0)
        two = O.calclcoeff([100.0, 250.0], [0.3, 0.7], d)
        ref = 0.3 * gc_textbook(d, 50.0) + 0.7 * gc_textbook(d, 125.0)
        assert abs(two - ref) < 1e-14


# ---------------------------------------------------------------- geometry
def test_ll2xyz_chord():
    rng = np.random.default_rng(5)
    for _ in range(200):
        lo1, lo2 = rng.uniform(-180, 360, 2)
        la1, la2 = rng.uniform(-89, 89, 2)
        a = np.array(O.ll2xyz(lo1, la1))
        b = np.array(O.ll2xyz(lo2, la2))
        assert abs(np.linalg.norm(a) - R_EARTH) < 1e-9
        p1, p2, dl = math.radians(la1), math.radians(la2), math.radians(lo2 - lo1)
        hav = math.sin((p2 - p1) / 2) ** 2 + math.cos(p1) * math.cos(p2) * math.sin(dl / 2) ** 2
        dsig = 2 * math.asin(math.sqrt(hav))
        assert abs(np.linalg.norm(a - b) - 2 * R_EARTH * math.sin(dsig / 2)) < 1e-8


# ---------------------------------------------------------------- kd-tree
@pytest.mark.parametrize("ndim", [2, 3])
def test_kd_range_brute_force(ndim):
    rng = np.random.default_rng(11 + ndim)
    pts = np.round(rng.uniform(0, 100, (3000, ndim)), 2)
    pts[500:520] = pts[499]                      # duplicates
    t = O.OracleKd(pts)
    for q in rng.uniform(-5, 105, (40, ndim)):
        for r in (0.5, 7.0, 30.0):
            d, d2 = t.range(q, r)
            dd = ((pts - q) ** 2).sum(1)
            want = np.flatnonzero(dd <= r * r)
            assert np.array_equal(np.sort(d), want)
            assert np.allclose(d2, dd[d], rtol=1e-15, atol=0)
            do, d2o = t.range(q, r, ordered=True)
            assert np.array_equal(np.sort(do), want)
            assert np.all(np.diff(d2o) >= 0)
    # common/kdtree.c:538 prunes the far child on |dx| < range (strict), so a
    # zero-radius search at a duplicated point stops at the first copy: the
    # later copies were inserted to its right (kd_insertnode goes left only
    # on coord < node) and |dx| = 0 is not < 0
    d, _ = t.range(pts[499], 0.0)
    assert d.tolist() == [499]
    d, _ = t.range(pts[499], 1e-9)
    assert set(d.tolist()) == set(range(499, 520))


# ---------------------------------------------------------------- local search
def test_findlocal_brute_force():
    """obs_findlocal (calc/obs.c:1064-1133): per type, every good observation
    within the type's largest radius, tapered by calclcoeff."""
    P = synth.make_problem(seed=2, nj=10, ni=12, nmem=12, nobs=1500, locrad=150.0, profiles=40)
    P.obstypes[1]["locrad"] = [80.0, 220.0]
    P.obstypes[1]["locweight"] = [2.0, 1.0]
    status = P.obs_status.copy()
    status[::13] = 2
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd, status=status)
    xyz = np.array([O.ll2xyz(a, b) for a, b in zip(P.obs_lon.astype(np.float64), P.obs_lat.astype(np.float64))])
    for lon, lat in zip(P.node_lon.ravel()[::7], P.node_lat.ravel()[::7]):
        ids, lc = ob.findlocal(lon, lat)
        q = np.array(O.ll2xyz(lon, lat))
        want_ids, want_lc = [], []
        for t, ot in enumerate(P.obstypes):
            rmax = max(ot["locrad"])
            lw = np.array(ot.get("locweight", [1.0] * len(ot["locrad"])), float)
            lw /= lw.sum()
            d2 = ((xyz - q) ** 2).sum(1)
            sel = np.flatnonzero((P.obs_type == t) & (status == 0) & (d2 <= rmax * rmax))
            want_ids += sel.tolist()
            for o in sel:
                d = math.sqrt(d2[o])
                want_lc.append(sum(w * gc_textbook(d, r / 2.0) for r, w in zip(ot["locrad"], lw)))
        order = np.argsort(ids)
        assert np.array_equal(ids[order], np.sort(want_ids))
        wl = np.array(want_lc)[np.argsort(want_ids)]
        assert np.allclose(lc[order], wl, rtol=0, atol=1e-12)
        # types are listed in order
        assert np.all(np.diff(P.obs_type[ids]) >= 0)


@pytest.mark.parametrize("nmax", [1, 9, 40])
def test_findlocal_nlobsmax_brute_force(nmax):
    """NLOBSMAX (observations.c:1899-1908, kdtree.c:569-570): of the
    candidates in the type's tree within range, the nmax nearest by squared
    distance, equal distances in visiting order (a stable sort), and only then
    the search-time status filter."""
    P = synth.make_problem(seed=4, nj=10, ni=12, nmem=12, nobs=1200, locrad=150.0, profiles=30)
    status = P.obs_status.copy()
    status[::11] = 3                  # bad at search time, in the trees
    lim = [dict(o, nlobsmax=nmax) for o in P.obstypes]
    ob = O.OracleObs(lim, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd, status=status,
                     status_at_build=P.obs_status)
    full = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)   # visiting order, all candidates
    xyz = np.array([O.ll2xyz(a, b) for a, b in zip(P.obs_lon.astype(np.float64), P.obs_lat.astype(np.float64))])
    ties = 0
    for lon, lat in zip(P.node_lon.ravel()[::5], P.node_lat.ravel()[::5]):
        ids, lc = ob.findlocal(lon, lat)
        cand, clc = full.findlocal(lon, lat)
        q = np.array(O.ll2xyz(lon, lat))
        want, want_lc = [], []
        for t in range(len(P.obstypes)):
            sel = P.obs_type[cand] == t
            ct, lt = cand[sel], clc[sel]
            d2 = ((xyz[ct] - q) ** 2).sum(1)
            keep = np.argsort(d2, kind="stable")[:nmax]
            ties += len(d2[keep]) - len(np.unique(d2[keep]))
            good = status[ct[keep]] == 0
            want += ct[keep][good].tolist()
            want_lc += lt[keep][good].tolist()
        assert np.array_equal(ids, np.array(want, np.int32))
        assert np.array_equal(lc, np.array(want_lc))
    if nmax > 1:
        assert ties > 0, "the case must contain equal distances"


def test_findlocal_global_node_and_visibility_brute_force():
    """The two branches of obs_findlocal (observations.c:1837-1927) that do
    not search a tree, derived from the obs arrays alone: a node with NaN
    coordinates gets, type by type in ascending type order, every observation
    of the type that was good when the trees were built (obs_getglobal,
    observations.c:1773-1817; ascending index, taper 1, no search-time status
    filter); a type that is not visible from the grid's domain, or is
    stats-only, contributes nothing to any node (observations.c:1877-1897)."""
    P = synth.make_problem(seed=6, nj=10, ni=12, nmem=12, nobs=900, locrad=150.0, profiles=30)
    types = [dict(o) for o in P.obstypes] + [dict(locrad=[150.0], rfactor=1.0, statsonly=1)]
    otype = P.obs_type.copy()
    otype[::17] = 2                       # a stats-only type
    build = P.obs_status.copy()
    build[::9] = 5                        # bad at tree build: never listed
    status = build.copy()
    status[1::7] = 4                      # bad after the build: searched lists drop them, global lists keep them
    ob = O.OracleObs(types, otype, P.obs_lon, P.obs_lat, P.obs_estd, status=status, status_at_build=build)
    full = O.OracleObs(types, otype, P.obs_lon, P.obs_lat, P.obs_estd, status=status, status_at_build=build)
    for vis in (None, np.array([1, 1, 1], np.uint8), np.array([0, 1, 1], np.uint8), np.array([1, 0, 0], np.uint8)):
        visible = np.ones(3, bool) if vis is None else vis.astype(bool)
        ids, lc = ob.findlocal(float("nan"), 10.0, visible=vis)
        want = np.concatenate([np.flatnonzero((otype == t) & (build == 0))
                               for t in range(2) if visible[t]] + [np.zeros(0, np.int64)])
        assert np.array_equal(ids, want.astype(np.int32))
        assert np.all(lc == 1.0)
        assert (status[ids] != 0).any(), "the global list must keep obs bad at search time"
        # a geographic node: the visible types' lists, in type order, are the
        # all-visible lists restricted to those types
        for lon, lat in zip(P.node_lon.ravel()[::11], P.node_lat.ravel()[::11]):
            ids_v, lc_v = ob.findlocal(lon, lat, visible=vis)
            ids_a, lc_a = full.findlocal(lon, lat)
            keep = visible[otype[ids_a]]
            assert np.array_equal(ids_v, ids_a[keep]) and np.array_equal(lc_v, lc_a[keep])
            assert not np.isin(otype[ids_a], [2]).any(), "stats-only types are never searched"
            assert np.all(status[ids_a] == 0)


# ---------------------------------------------------------------- per-column algebra
def spd_inv_sqrt(M):
    lam, U = np.linalg.eigh(M)
    return (U / np.sqrt(lam)) @ U.T


@pytest.mark.parametrize("m,mout,p", [(10, 10, 1), (10, 10, 4), (10, 10, 10), (10, 7, 25), (40, 40, 17),
                                      (40, 32, 90)])
def test_local_analysis_closed_forms(m, mout, p):
    rng = np.random.default_rng(m * 1000 + p)
    S = rng.standard_normal((p, m)) * 0.2
    s = rng.standard_normal(p) * 0.3
    M = np.eye(m) + S.T @ S
    Minv = np.linalg.inv(M)
    w_ref = np.linalg.solve(M, S.T @ s)
    dfs_ref = np.trace(Minv @ S.T @ S)
    srf_ref = math.sqrt(np.trace(S.T @ S) / dfs_ref) - 1.0
    T, w, dfs, srf = O.local_analysis(O.MODE_ENKF, O.SCHEME_DENKF, m, mout, S, s)
    assert np.allclose(w, w_ref, rtol=0, atol=1e-12)
    assert np.allclose(T, (0.5 * (np.eye(m) + Minv))[:mout], rtol=0, atol=1e-12)
    assert abs(dfs - dfs_ref) < 1e-11 and abs(srf - srf_ref) < 1e-11
    T, w, _, _ = O.local_analysis(O.MODE_ENKF, O.SCHEME_ETKF, m, mout, S, s)
    assert np.allclose(w, w_ref, rtol=0, atol=1e-12)
    assert np.allclose(T, spd_inv_sqrt(M)[:mout], rtol=0, atol=1e-11)
    _, w, _, _ = O.local_analysis(O.MODE_ENOI, O.SCHEME_DENKF, m, mout, S, s)
    assert np.allclose(w, w_ref, rtol=0, atol=1e-12)


def test_single_observation():
    """p = 1: M = I + a a^T, so M^-1 = I - a a^T / (1 + |a|^2) (Sherman-
    Morrison), dfs = |a|^2 / (1 + |a|^2), srf = sqrt(1 + |a|^2) - 1."""
    m = 16
    a = np.linspace(-0.4, 0.5, m)
    s0 = 0.7
    n2 = a @ a
    T, w, dfs, srf = O.local_analysis(O.MODE_ENKF, O.SCHEME_DENKF, m, m, a[None, :], np.array([s0]))
    assert np.allclose(w, a * s0 / (1 + n2), atol=1e-15)
    assert np.allclose(T, np.eye(m) - 0.5 * np.outer(a, a) / (1 + n2), atol=1e-15)
    assert abs(dfs - n2 / (1 + n2)) < 1e-15
    assert abs(srf - (math.sqrt(1 + n2) - 1)) < 1e-14
    T, _, _, _ = O.local_analysis(O.MODE_ENKF, O.SCHEME_ETKF, m, m, a[None, :], np.array([s0]))
    assert np.allclose(T, np.eye(m) + (1 / math.sqrt(1 + n2) - 1) * np.outer(a, a) / n2, atol=1e-14)


def test_calc_loop_relaxation_and_pertype_diagnostics():
    """das_calctransforms (calc/transforms.c:605-985): T relaxed to the prior
    as I + alpha (T - I), per-type diagnostics equal separate analyses of each
    type's observations, empty columns give T = I, w = 0."""
    P = synth.make_problem(seed=4, nj=8, ni=9, nmem=16, nobs=700, locrad=90.0, profiles=30)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(16, ob, P.HA, P.innov)
    lon = np.concatenate([P.node_lon.ravel(), [P.node_lon.max() + 30.0]])
    lat = np.concatenate([P.node_lat.ravel(), [P.node_lat.mean()]])
    alpha = 0.6
    r = O.calc_transforms(O.MODE_ENKF, O.SCHEME_ETKF, 16, 12, alpha, ob, S, s, lon, lat)
    nt = len(P.obstypes)
    for n in range(0, len(lon) - 1, 5):
        ids, lc = ob.findlocal(lon[n], lat[n])
        Sl = S[:, ids].T.astype(np.float64) * lc[:, None]
        sl = s[ids] * lc
        T, w, dfs, srf = O.local_analysis(O.MODE_ENKF, O.SCHEME_ETKF, 16, 12, Sl, sl)
        Trel = np.eye(12, 16) + alpha * (T - np.eye(12, 16))
        assert np.allclose(r["T64"][n], Trel, rtol=0, atol=1e-14)
        assert np.array_equal(r["w64"][n], w)
        assert r["pnlobs"][nt, n] == len(ids)
        assert r["pdfs"][nt, n] == np.float32(dfs)
        for t in range(nt):
            sel = P.obs_type[ids] == t
            assert r["pnlobs"][t, n] == sel.sum()
            if sel.any():
                _, _, dt, st = O.local_analysis(O.MODE_ENOI, O.SCHEME_DENKF, 16, 16, Sl[sel], sl[sel])
                assert abs(r["pdfs"][t, n] - dt) <= 2e-7 * abs(dt)
                assert abs(r["psrf"][t, n] - st) <= 2e-7 * max(abs(st), 1.0)
    last = len(lon) - 1
    assert np.array_equal(r["T64"][last], np.eye(12, 16)) and not r["w64"][last].any()
    assert not r["pnlobs"][:, last].any()


# ---------------------------------------------------------------- update
def update_problem(nj=10, ni=11, m=8, md=8, stride=3, nk=2, seed=0):
    rng = np.random.default_rng(seed)
    nj_s, ni_s = (nj - 1) // stride + 1, (ni - 1) // stride + 1
    T = (np.eye(md, m)[None, None] + 0.1 * rng.standard_normal((nj_s, ni_s, md, m))).astype(np.float32)
    w = (0.1 * rng.standard_normal((nj_s, ni_s, m))).astype(np.float32)
    nlev = np.full((nj, ni), nk, np.int32)
    E = rng.standard_normal((m, nj, ni)).astype(np.float32) * 2.0 + 5.0
    return T, w, nlev, E


def brute_cell(v, T, w, md, mode=O.MODE_ENKF, inflation=1.0):
    m = len(v)
    v = v.astype(np.float64)
    mean = v[:md].mean()
    a = v - mean
    va = T.astype(np.float64) @ a
    mean_a = mean + a @ w.astype(np.float64)
    if mode == O.MODE_HYBRID and md > 1:
        va /= math.sqrt((m - 1) / (md - 1))
    return va * inflation + mean_a


def test_update_identity_transform_is_noop():
    T, w, nlev, E = update_problem(md=8)
    T[:] = np.eye(8, dtype=np.float32)
    w[:] = 0.0
    buf = E.copy()
    O.update_fields(O.MODE_ENKF, 8, 8, 3, 0, 0, nlev, T, w, [dict(level=0)], [buf], 0, 9)
    assert np.allclose(buf, E, rtol=2e-6, atol=1e-6)


@pytest.mark.parametrize("mode,md", [(O.MODE_ENKF, 8), (O.MODE_HYBRID, 5)])
def test_update_brute_force_at_nodes_and_between(mode, md):
    """At stride nodes the transforms are used as stored; between nodes they
    are the bilinear blend (update.c:195-272), j first, then i."""
    m, stride = 8, 3
    T, w, nlev, E = update_problem(m=m, md=md, stride=stride)
    buf = E.copy()
    O.update_fields(mode, m, md, stride, 0, 0, nlev, T, w, [dict(level=0)], [buf], 0, 9)
    for j in range(10):
        for i in range(11):
            jj, ii = j // stride, i // stride
            fj, fi = (j % stride) / stride, (i % stride) / stride
            jj2 = min(jj + 1, T.shape[0] - 1)
            ii2 = min(ii + 1, T.shape[1] - 1)
            Tc = ((1 - fj) * ((1 - fi) * T[jj, ii] + fi * T[jj, ii2]) + fj * ((1 - fi) * T[jj2, ii] + fi * T[jj2, ii2]))
            wc = ((1 - fj) * ((1 - fi) * w[jj, ii] + fi * w[jj, ii2]) + fj * ((1 - fi) * w[jj2, ii] + fi * w[jj2, ii2]))
            want = brute_cell(E[:, j, i], Tc, wc, md, mode)
            assert np.allclose(buf[:md, j, i], want, rtol=1e-5, atol=1e-5), (j, i)
            assert np.array_equal(buf[md:, j, i], E[md:, j, i])


def test_update_periodic_wrap():
    m = 6
    T, w, nlev, E = update_problem(ni=12, m=m, md=m, stride=3)   # ni_s = 4, last column 11 wraps
    a = E.copy()
    b = E.copy()
    O.update_fields(O.MODE_ENKF, m, m, 3, 0, 0, nlev, T, w, [dict(level=0)], [a], 0, 9)
    O.update_fields(O.MODE_ENKF, m, m, 3, 1, 0, nlev, T, w, [dict(level=0)], [b], 0, 9)
    assert np.array_equal(a[:, :, :10], b[:, :, :10])
    Tc = (1 / 3) * T[0, 3] + (2 / 3) * T[0, 0]
    wc = (1 / 3) * w[0, 3] + (2 / 3) * w[0, 0]
    assert np.allclose(b[:, 0, 11], brute_cell(E[:, 0, 11], Tc, wc, m), rtol=1e-5, atol=1e-5)
    assert not np.allclose(a[:, 0, 11], b[:, 0, 11])


def test_update_masks_missing_and_increments():
    m, md = 8, 6
    T, w, nlev, E = update_problem(m=m, md=md, stride=1, nj=4, ni=4)
    nlev[1, 1] = 0
    E[2, 2, 2] = np.nan            # dynamic member missing: whole cell zeroed
    E[7, 3, 3] = 1e7               # static member out of range: statics zeroed
    buf = E.copy()
    O.update_fields(O.MODE_ENKF, m, md, 1, 0, 0, nlev, T, w, [dict(level=0)], [buf], 0, 3)
    assert not buf[:, 1, 1].any() and not buf[:, 2, 2].any()
    assert not buf[md:, 3, 3].any()
    inc = E.copy()
    O.update_fields(O.MODE_ENKF, m, md, 1, 0, 0, nlev, T, w, [dict(level=0)], [inc], 0, 3, outputinc=1)
    ok = np.ones((4, 4), bool)
    ok[1, 1] = ok[2, 2] = ok[3, 3] = False
    assert np.allclose(inc[:md][:, ok], buf[:md][:, ok] - E[:md][:, ok], rtol=1e-5, atol=1e-5)


def test_update_inflation_fixed_and_capped():
    m = 8
    T, w, nlev, E = update_problem(m=m, md=m, stride=1, nj=4, ni=4)
    T = (0.7 * T).astype(np.float32)   # shrinks the spread, so the capped factor engages
    fixed, capped = E.copy(), E.copy()
    infl = O.update_fields(O.MODE_ENKF, m, m, 1, 0, 0, nlev, T, w,
                           [dict(level=0, inflation=1.1, inf_ratio=float("nan"))], [fixed], 0, 3, want_infl=True)
    assert np.all(infl == np.float32(1.1))
    infl2 = O.update_fields(O.MODE_ENKF, m, m, 1, 0, 0, nlev, T, w,
                            [dict(level=0, inflation=2.0, inf_ratio=0.5)], [capped], 0, 3, want_infl=True)
    for j in range(4):
        for i in range(4):
            v = E[:, j, i].astype(np.float64)
            a = v - v.mean()
            va = T[j, i].astype(np.float64) @ a
            want = min(math.sqrt((a @ a) / (va @ va)) * 0.5 + 0.5, 2.0)
            assert abs(infl2[0, j, i] - want) < 1e-5
            assert np.allclose(fixed[:, j, i], brute_cell(E[:, j, i], T[j, i], w[j, i], m, inflation=1.1),
                               rtol=1e-5, atol=1e-5)
            assert np.allclose(capped[:, j, i], brute_cell(E[:, j, i], T[j, i], w[j, i], m, inflation=want),
                               rtol=1e-5, atol=1e-5)


def test_update_bg_increment():
    """das_updatebg (update.c:568-814): x^a = x^f + (E - mean(E)) w."""
    m = 8
    _, w, nlev, E = update_problem(m=m, md=m, stride=2, nj=7, ni=7)
    bg = np.random.default_rng(1).standard_normal((1, 7, 7)).astype(np.float32)
    buf = np.concatenate([E, bg])
    O.update_bg(m, 2, 0, 0, nlev, w, [dict(level=0)], [buf], 0, 6)
    for j in (0, 2, 4, 6):
        for i in (0, 2, 4, 6):
            v = E[:, j, i].astype(np.float64)
            want = bg[0, j, i] + (v - v.mean()) @ w[j // 2, i // 2].astype(np.float64)
            assert abs(buf[m, j, i] - want) < 1e-5 * max(1.0, abs(want))


def test_sethybridensemble_closed_form():
    """das_sethybridensemble (common/dasystem.c:775-819): dynamic anomalies
    scaled by sqrt((m-1)/(md-1)) about the dynamic mean; static members
    scaled by sqrt(gamma (m-1)/(ms-1)) and shifted by the dynamic mean."""
    m, md, gamma = 10, 6, 0.3
    v = np.random.default_rng(3).standard_normal((m, 50)).astype(np.float32)
    out = v.copy()
    O.sethybridensemble(m, md, gamma, out)
    mean = v[:md].astype(np.float64).mean(0)
    kd = math.sqrt((m - 1) / (md - 1))
    ks = math.sqrt(gamma * (m - 1) / (m - md - 1))
    assert np.allclose(out[:md], (v[:md] - mean) * kd + mean, rtol=1e-6, atol=1e-6)
    assert np.allclose(out[md:], v[md:] * ks + mean, rtol=1e-6, atol=1e-6)
    assert np.allclose(out[:md].astype(np.float64).mean(0), mean, atol=1e-6)


# ---------------------------------------------------------------- RANDOMISE
def test_drand48_matches_glibc_and_default_state():
    """or_drand48 is the POSIX drand48 recurrence: it matches glibc's
    erand48 on the same 48-bit state draw for draw, and a process that never
    seeds drand48 (the reference seeds srand(), common/utils.c:102-122)
    starts from state 0."""
    import ctypes as C
    import subprocess
    import sys
    libc = C.CDLL("libc.so.6")
    libc.erand48.restype = C.c_double
    libc.erand48.argtypes = [C.POINTER(C.c_ushort)]
    for seed in (0, 1, 0x1234ABCD330E, (1 << 48) - 1):
        xs = (C.c_ushort * 3)(seed & 0xFFFF, (seed >> 16) & 0xFFFF, (seed >> 32) & 0xFFFF)
        st = C.c_ulonglong(seed)
        for _ in range(200):
            assert O.lib().or_drand48(C.byref(st)) == libc.erand48(xs)
    code = ("import ctypes; L = ctypes.CDLL('libc.so.6'); L.drand48.restype = ctypes.c_double; "
            "print(repr(L.drand48())); print(repr(L.drand48()))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout.split()
    st = C.c_ulonglong(0)
    assert [float(x) for x in out] == [O.lib().or_drand48(C.byref(st)), O.lib().or_drand48(C.byref(st))]


def test_normalpair_and_addnoise_closed_forms():
    """get_normalpair is Box-Muller on two draws; addnoise adds the centred
    draws of the first nmem normals, scaled by sqrt(1 - d^2) sigma, to
    v * d, with the reference's float / double rounding."""
    import ctypes as C
    st = C.c_ulonglong(12345)
    d1 = O.lib().or_drand48(C.byref(st))
    d2 = O.lib().or_drand48(C.byref(st))
    x, st2 = O.normalpairs(12345, 1)
    u = math.sqrt(-2.0 * math.log(1.0 - d1))
    assert abs(x[0] - u * math.cos(2 * math.pi * d2)) <= 4e-16 * u
    assert abs(x[1] - u * math.sin(2 * math.pi * d2)) <= 4e-16 * u
    assert st2 == st.value
    for nmem in (7, 8):
        rng = np.random.default_rng(nmem)
        v = rng.standard_normal((nmem, 5, 6)).astype(np.float32)
        d, sigma = 0.8, 0.3
        r, _ = O.normalpairs(99, (nmem + 1) // 2)
        r = r[:nmem] - r[:nmem].sum() / nmem
        mult = np.float32(math.sqrt(1.0 - d * d) * sigma)
        want = ((v * np.float32(d)).astype(np.float64) + np.float64(mult) * r[:, None, None]).astype(np.float32)
        got = v.copy()
        st_after = O.addnoise(nmem, d, sigma, 99, got)
        assert np.array_equal(got, want)
        assert st_after == O.normalpairs(99, (nmem + 1) // 2)[1]
        # zero-mean noise: the member mean moves by d only
        assert np.allclose(got.astype(np.float64).mean(0), d * v.astype(np.float64).mean(0), atol=1e-6)


# ---------------------------------------------------------------- analysed observations
@pytest.mark.parametrize("mode,md", [(1, 24)])
def test_update_HE_is_the_field_update_at_the_obs_cell(mode, md):
    """update_HE_structured (ensobs.c:1228-1553) applies the update of
    update.c to the ensemble observation vector at the obs' nearest cell
    ((int)(f + 0.5), the last column clamped): for obs whose HE^f is the
    field at that cell, HE^a is the updated field there (capped and plain
    inflation), and obs that are not OK or whose row rounds past the grid are
    left unchanged."""
    m = 24
    P = synth.make_problem(seed=21 + md, nj=20, ni=26, nk=1, nmem=m, nobs=500, locrad=80.0, stride=3)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    r = O.calc_transforms(mode, 1, m, md, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    T = r["T"].reshape(P.nj_s, P.ni_s, md, m)
    w = r["w"].reshape(P.nj_s, P.ni_s, m)
    rng = np.random.default_rng(md)
    nobs = 400
    fi = rng.uniform(-0.49, P.ni - 0.51, nobs)
    fj = rng.uniform(-0.49, P.nj - 0.51, nobs)
    fi[:5] = P.ni - 0.5           # rounds to mni: clamped to the last column
    fj[5:8] = P.nj - 0.4          # rounds past the last row: not processed
    status = np.zeros(nobs, np.int16)
    status[10:15] = 3
    field = P.ens["temp"][:, 0]                      # [m][nj][ni]
    ii = np.minimum((fi + 0.5).astype(int), P.ni - 1)
    jj = np.minimum((fj + 0.5).astype(int), P.nj - 1)
    HE = np.ascontiguousarray(field[:, jj, ii])
    for ratio in (1.0, float("nan")):
        infl0 = np.full(nobs, 1.08, np.float32)
        rat = np.full(nobs, ratio)
        HEa = O.update_HE(mode, m, md, 3, 0, P.nj, P.ni, T, w, fi, fj, status, infl0, rat, HE)
        buf = np.ascontiguousarray(field.copy())
        nl = np.ones((P.nj, P.ni), np.int32)
        O.update_fields(mode, m, md, 3, 0, 0, nl, T, w, [dict(level=0, inflation=1.08, inf_ratio=ratio)], [buf], 0,
                        P.nj - 1)
        want = buf[:, jj, ii]
        ok = (status == 0) & ((fj + 0.5).astype(int) < P.nj)
        assert np.allclose(HEa[:, ok], want[:, ok], rtol=0, atol=2e-6 * np.abs(want).max())
        assert np.array_equal(HEa[:, ~ok], HE[:, ~ok])
        assert not np.array_equal(HEa[:, ok], HE[:, ok])


def test_update_HE_hybrid_at_stride_nodes_brute_force():
    """Hybrid mode at obs on stride-node cells (T there is the node's T):
    dynamic members HE^a = (HA^f T + x^a) with x^a = x^f + HA^f w and plain
    inflation about x^a, no k_d rescaling (unlike update.c:413-427); static
    members shift by x^a - x^f (ensobs.c:1524-1533)."""
    m, md = 24, 18
    P = synth.make_problem(seed=5, nj=19, ni=25, nk=1, nmem=m, nobs=500, locrad=80.0, stride=3)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    r = O.calc_transforms(3, 1, m, md, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    T = r["T"].reshape(P.nj_s, P.ni_s, md, m).astype(np.float64)
    w = r["w"].reshape(P.nj_s, P.ni_s, m).astype(np.float64)
    rng = np.random.default_rng(2)
    nobs = 60
    jn = rng.integers(0, P.nj_s, nobs)
    inn = rng.integers(0, P.ni_s, nobs)
    fi = inn * 3 + rng.uniform(-0.4, 0.4, nobs)
    fj = jn * 3 + rng.uniform(-0.4, 0.4, nobs)
    HE = (15.0 + rng.standard_normal((m, nobs))).astype(np.float32)
    infl = 1.1
    HEa = O.update_HE(3, m, md, 3, 0, P.nj, P.ni, T.astype(np.float32), w.astype(np.float32), fi, fj,
                      np.zeros(nobs, np.int16), np.full(nobs, infl, np.float32), np.full(nobs, np.nan), HE)
    X = HE.astype(np.float64)
    for o in range(nobs):
        Tn, wn = T[jn[o], inn[o]], w[jn[o], inn[o]]
        xf = X[:md, o].mean()
        A = X[:, o] - xf
        xa = xf + A @ wn
        dyn = (Tn @ A) * infl + xa
        assert np.allclose(HEa[:md, o], dyn, rtol=0, atol=3e-5 * np.abs(dyn).max())
        assert np.allclose(HEa[md:, o], X[md:, o] + xa - xf, rtol=0, atol=3e-5 * np.abs(dyn).max())


def test_update_Hx_enoi_at_stride_nodes_brute_force():
    """EnOI (update_Hx_structured, ensobs.c:1933-1962): every member of the
    ensemble observation moves by (HE - mean(HE)) . w, with w of the obs'
    nearest cell (here stride-node cells, where w is the node's w)."""
    m = 20
    P = synth.make_problem(seed=9, nj=19, ni=25, nk=1, nmem=m, nobs=400, locrad=80.0, stride=3)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    r = O.calc_transforms(2, 1, m, m, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    w = r["w"].reshape(P.nj_s, P.ni_s, m)
    rng = np.random.default_rng(4)
    nobs = 50
    jn = rng.integers(0, P.nj_s, nobs)
    inn = rng.integers(0, P.ni_s, nobs)
    fi = inn * 3 + rng.uniform(-0.4, 0.4, nobs)
    fj = jn * 3 + rng.uniform(-0.4, 0.4, nobs)
    HE = (3.0 + rng.standard_normal((m, nobs))).astype(np.float32)
    HEa = O.update_HE(2, m, m, 3, 0, P.nj, P.ni, np.zeros((P.nj_s, P.ni_s, 1, 1), np.float32), w, fi, fj,
                      np.zeros(nobs, np.int16), np.ones(nobs, np.float32), np.full(nobs, 1.0), HE)
    X = HE.astype(np.float64)
    d = ((X - X.mean(0)) * w[jn, inn].T.astype(np.float64)).sum(0)
    assert np.allclose(HEa, X + d[None, :], rtol=0, atol=1e-5)
    assert np.abs(d).max() > 1e-3


# ---------------------------------------------------------------- ensemble observations
def test_interp2d_linear_masked_edges_periodic():
    """interpolate2d_structured (utils.c:1140-1213): a field linear in (i, j)
    is reproduced where the four corners are wet; a dry corner drops out and
    the others are renormalised; columns past the last wrap on periodic grids
    and clamp otherwise; indices at or past the end are pulled back by
    EPS_FLOAT."""
    nj, ni = 7, 9
    jj, ii = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    v = (2.0 + 0.5 * ii - 0.25 * jj).astype(np.float32)
    wet = np.ones((nj, ni), np.int32)
    rng = np.random.default_rng(0)
    for _ in range(50):
        fi, fj = rng.uniform(0, ni - 1), rng.uniform(0, nj - 1)
        assert abs(O.interp2d([fi, fj], v, wet) - (2.0 + 0.5 * fi - 0.25 * fj)) < 1e-5
    # a dry corner: weights of the wet three, renormalised
    mask = wet.copy()
    mask[3, 4] = 0
    fi, fj = 4.3, 2.6
    wts = {(2, 4): 0.4 * 0.7, (2, 5): 0.4 * 0.3, (3, 5): 0.6 * 0.3}
    want = sum(float(v[j, i]) * w for (j, i), w in wts.items()) / sum(wts.values())
    assert abs(O.interp2d([fi, fj], v, mask) - want) < 1e-6
    # past the last column: periodic wrap to column 0, else the last column
    u = np.zeros((nj, ni), np.float32)
    u[:, 0] = 1.0
    assert abs(O.interp2d([ni - 0.75, 3.0], u, wet, periodic_i=1) - 0.25) < 1e-7
    assert O.interp2d([ni - 0.75, 3.0], u, wet, periodic_i=0) == 0.0
    # an index at the end is pulled back by 1e-7
    assert abs(O.interp2d([ni - 1.0, nj * 1.0], v, wet) - O.interp2d([ni - 1.0, nj - 1e-7], v, wet)) == 0.0


def test_interp3d_levels_and_depth_mask():
    """interpolate3d_structured (utils.c:1252-1367): linear in k between layer
    centres, clamped to the first and last layers, and layers deeper than a
    column's level count drop out (|ktop - k| >= nlevels)."""
    nk, nj, ni = 5, 4, 4
    v = np.stack([np.full((nj, ni), 10.0 + 3.0 * k, np.float32) for k in range(nk)])
    nl = np.full((nj, ni), nk, np.int32)
    assert abs(O.interp3d([1.2, 1.7], 2.25, v, nl) - (10.0 + 3.0 * 2.25)) < 1e-5
    assert O.interp3d([1.2, 1.7], -0.4, v, nl) == 10.0
    assert O.interp3d([1.2, 1.7], nk - 0.6, v, nl) == 10.0 + 3.0 * (nk - 1)
    # ktop = nk - 1: layer k is |nk - 1 - k| from the surface
    assert abs(O.interp3d([1.0, 1.0], 1.5, v, nl, ktop=nk - 1) - (10.0 + 3.0 * 1.5)) < 1e-5
    # the deeper layer is dry in every column: the shallower one alone
    shallow = np.full((nj, ni), 3, np.int32)
    assert O.interp3d([2.0, 2.0], 2.5, v, shallow) == v[2, 0, 0]


@pytest.mark.parametrize("mode,md", [(1, 12), (3, 8)])
def test_getHE_cell_centres_innovations_spread_hybrid(mode, md):
    """das_getHE + das_calcinnandspread: at cell centres (integer indices) HE
    is the field value; the innovation is y minus the member mean and HA the
    anomaly from it (FP32), the spread the unbiased sample deviation; in
    hybrid mode the dynamic members' anomalies are scaled by k_d and the
    static ones by k_s about the dynamic mean (ensobs.c:434-486); obs not OK
    stay NaN."""
    m = 12
    rng = np.random.default_rng(md)
    nk, nj, ni = 3, 6, 7
    fields = [rng.standard_normal((m, nj, ni)).astype(np.float32) + 20.0 + k for k in range(nk)]
    fields.append(rng.standard_normal((m, nj, ni)).astype(np.float32))          # a 2D variable
    nl = np.full((nj, ni), nk, np.int32)
    nobs = 40
    type_ = rng.integers(0, 2, nobs).astype(np.int16)
    fi = rng.integers(0, ni, nobs).astype(np.float64)
    fj = rng.integers(0, nj, nobs).astype(np.float64)
    fk = rng.integers(0, nk, nobs).astype(np.float32)
    status = np.zeros(nobs, np.int16)
    status[5] = 2
    value = rng.standard_normal(nobs) + 20.0
    hops = [(1, [0, 1, 2]), (0, [3])]
    gamma = 0.7
    HA, s, sd = O.getHE(mode, m, md, gamma, nl, 0, 0, fields, hops, type_, status, fi, fj, fk, value)
    HE = np.full((m, nobs), np.nan, np.float64)
    for o in range(nobs):
        if status[o]:
            continue
        f = 3 if type_[o] == 1 else int(fk[o])
        HE[:, o] = fields[f][:, int(fj[o]), int(fi[o])]
    if mode == 3:
        ms = m - md
        kd = math.sqrt((m - 1) / (md - 1))
        ks = math.sqrt(gamma * (m - 1) / (ms - 1))
        mu_d = HE[:md].mean(0)
        mu_s = HE[md:].mean(0)
        HE = np.concatenate([(HE[:md] - mu_d) * kd + mu_d, (HE[md:] - mu_s) * ks + mu_d])
    mean = HE.mean(0)
    ok = status == 0
    assert np.allclose(HA[:, ok], (HE - mean)[:, ok], atol=2e-5)
    assert np.allclose(s[ok], (value - mean)[ok], atol=1e-5)
    assert np.allclose(sd[ok], (HE - mean)[:, ok].std(0, ddof=1), rtol=1e-5)
    assert np.isnan(HA[:, ~ok]).all()


def test_spread_population_std_float_square_and_limits():
    """or_spread (diag/diags.c:101-123) against numpy's population standard
    deviation; the members' squares are float products (a constant member
    value whose float square is above its exact square gives a spread of
    sqrt(float(c*c) - c^2), not 0); |mean| > MAXOBSVAL gives 0; a negative
    variance from rounding is clamped to 0."""
    rng = np.random.default_rng(11)
    v = (rng.standard_normal((96, 40, 30)) * 0.7).astype(np.float32)
    sp = O.spread(v)
    ref = v.astype(np.float64).std(axis=0)          # ddof 0: / nmem, not / (nmem - 1)
    assert np.abs(sp - ref).max() <= 2e-6 * ref.max()
    assert np.abs(sp - v.astype(np.float64).std(axis=0, ddof=1)).max() > 1e-3 * ref.max()
    # constant members: the float square decides (above the exact square: a
    # positive spread; below: a negative variance, clamped to 0)
    cs = [np.float32(1.0 + 0.01 * k) for k in range(1, 100)]
    up = next(c for c in cs if float(np.float32(c * c)) > np.float64(c) ** 2)
    down = next(c for c in cs if float(np.float32(c * c)) < np.float64(c) ** 2)
    gap = float(np.float32(up * up)) - np.float64(up) ** 2
    assert np.all(O.spread(np.full((24, 5), up, np.float32)) == np.float32(np.sqrt(gap)))
    assert np.all(O.spread(np.full((24, 3), down, np.float32)) == 0.0)
    # |mean| > MAXOBSVAL: 0, whatever the variance
    z = rng.standard_normal((24, 7))
    assert np.all(O.spread((z * 1e5 + 2e6).astype(np.float32)) == 0.0)
    assert np.all(O.spread((z * 1e5 - 2e6).astype(np.float32)) == 0.0)
    assert np.all(O.spread((z * 1e5 + 4e5).astype(np.float32)) > 1e4)


def test_moderateobs_closed_forms():
    """das_moderateobs (calc/ensobs.c:749-799) pinned by what the formula
    fixes: no innovation or no spread leaves the error unchanged (exactly,
    for dyadic values), the known answer std 3, estd 4, inn = K sqrt(59) gives
    5, large innovations give estd'^2 -> std |inn| / K, bad obs and a NaN K
    change nothing, and only obs whose error variance more than doubles are
    counted, per type."""
    K = 2.0
    typ = np.array([0, 0, 1, 1, 1, 0, 1])
    st = np.array([0, 0, 0, 0, 0, 7, 0])
    estd = np.array([0.25, 0.5, 4.0, 1.0, 0.75, 0.5, 0.125], np.float32)
    std = np.array([0.5, 0.0, 3.0, 1e-3, 1.5, 2.0, 0.5])
    inn = np.array([0.0, 5.0, K * np.sqrt(59.0), 1e8, 40.0, 9.0, 0.0])
    e, cnt = O.moderateobs(K, 2, typ, st, estd, std, inn)
    assert e[0] == estd[0] and e[1] == estd[1] and e[6] == estd[6]        # no innovation / no spread
    assert abs(e[2] - 5.0) < 1e-6                                        # sqrt(34 - 9)
    assert abs(e[3] ** 2 / (1e-3 * 1e8 / K) - 1.0) < 1e-2                # the large-innovation limit
    assert e[5] == estd[5]                                               # bad status
    big = [k for k in range(7) if st[k] == 0 and std[k] > 0 and float(e[k]) ** 2 / float(estd[k]) ** 2 > 2.0]
    assert sorted(big) == [3, 4]
    assert list(cnt) == [0, 2]
    # a larger K moderates less; NaN turns it off
    e4, _ = O.moderateobs(4.0, 2, typ, st, estd, std, inn)
    assert (e4 <= e).all() and (e4 >= estd - 1e-7).all()
    en, cn = O.moderateobs(float("nan"), 2, typ, st, estd, std, inn)
    assert np.array_equal(en, estd) and not cn.any()
