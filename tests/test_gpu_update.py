"""GPU parity of the update path (das_updatefields, das_updatebg,
das_sethybridensemble) against the CPU oracle, through the C ABI.

The oracle's FP32 transform product is a sequential sum (the reference calls
MKL sgemv, whose summation order is its own), so outputs are compared within
1e-5 of the field's magnitude; the interpolation, masks and missing-value
handling are exact (zeros compared exactly).
"""
import numpy as np
import pytest

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu


def transforms(P, mode, scheme, md, alpha=1.0):
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(P.nmem, ob, P.HA, P.innov)
    r = O.calc_transforms(mode, scheme, P.nmem, md, alpha, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    T = r["T"].reshape(P.nj_s, P.ni_s, md, P.nmem) if r["T"] is not None else None
    w = r["w"].reshape(P.nj_s, P.ni_s, P.nmem)
    return T, w


def fields_of(P, nmem):
    bufs, fl = [], []
    for k in range(P.nk):
        bufs.append(np.ascontiguousarray(P.ens["temp"][:nmem, k]))
        fl.append(dict(level=k, inflation=1.1, inf_ratio=1.0))
        bufs.append(np.ascontiguousarray(P.ens["salt"][:nmem, k]))
        fl.append(dict(level=k, inflation=1.0, inf_ratio=float("nan")))
    bufs.append(np.ascontiguousarray(P.ens["eta"][:nmem, 0]))
    fl.append(dict(level=0, issurfacevar=1, inflation=1.05, inf_ratio=0.5))
    return bufs, fl


def check(a, b, tol=1e-5):
    scale = max(np.abs(b).max(), 1.0)
    assert np.array_equal(a == 0, b == 0), "zero pattern (masks / missing values)"
    assert np.abs(a.astype(np.float64) - b).max() <= tol * scale


@pytest.mark.parametrize("stride,periodic,outputinc,surfk", [(3, 0, 0, 0), (3, 1, 1, 0), (1, 0, 0, 0), (2, 0, 0, 3)])
def test_update_fields_parity(stride, periodic, outputinc, surfk):
    print("case update stride=%d periodic=%d outputinc=%d surfk=%d" % (stride, periodic, outputinc, surfk),
          flush=True)
    P = synth.make_problem(seed=stride + 7 * periodic, nmem=24, nobs=1200, locrad=60.0, nj=31, ni=44,
                           stride=stride, land_frac=0.2)
    T, w = transforms(P, 1, 1, 24, alpha=0.7)
    bufs, fl = fields_of(P, 24)
    # plant missing values: a dynamic member NaN at one cell, a huge value at another
    bufs[0][3, 5, 7] = np.nan
    bufs[1][0, 9, 2] = 2e6
    ref = [b.copy() for b in bufs]
    gpu = [b.copy() for b in bufs]
    infl_o = O.update_fields(1, 24, 24, stride, periodic, surfk, P.nlevels, T, w, fl, ref, 0, P.nj - 1,
                             outputinc=outputinc, want_infl=True)
    E = enkfg.Engine(24, mode=enkfg.MODE_ENKF)
    E.set_transforms(T, w, stride)
    infl_g = E.update_fields(P.nlevels, stride, periodic, surfk, fl, gpu, 0, P.nj - 1, outputinc=bool(outputinc),
                             want_infl=True)
    for a, b in zip(gpu, ref):
        check(a, b)
    assert np.abs(infl_g - infl_o).max() <= 1e-5
    E.close()


@pytest.mark.parametrize("nmem", [30, 17])
def test_update_member_count_not_multiple_of_four(nmem):
    """m % 4 != 0: the scalar interpolation of T_cell (the quad path needs
    whole float4 rows) and a partial last member tile."""
    print("case update nmem=%d" % nmem, flush=True)
    P = synth.make_problem(seed=31 + nmem, nmem=nmem, nobs=1000, locrad=60.0, nj=28, ni=35, stride=3, land_frac=0.1)
    T, w = transforms(P, 1, 1, nmem, alpha=0.9)
    bufs, fl = fields_of(P, nmem)
    bufs[0][2, 4, 6] = np.nan
    ref = [b.copy() for b in bufs]
    gpu = [b.copy() for b in bufs]
    infl_o = O.update_fields(1, nmem, nmem, 3, 0, 0, P.nlevels, T, w, fl, ref, 0, P.nj - 1, want_infl=True)
    E = enkfg.Engine(nmem, mode=enkfg.MODE_ENKF)
    E.set_transforms(T, w, 3)
    infl_g = E.update_fields(P.nlevels, 3, 0, 0, fl, gpu, 0, P.nj - 1, want_infl=True)
    for a, b in zip(gpu, ref):
        check(a, b)
    assert np.abs(infl_g - infl_o).max() <= 1e-5
    E.close()


@pytest.mark.parametrize("nmem,precision", [(48, enkfg.PREC_FP64), (80, enkfg.PREC_FP64), (64, enkfg.PREC_FP64),
                                             (80, enkfg.PREC_MIXED), (112, enkfg.PREC_MIXED)])
def test_update_tile_layouts(nmem, precision):
    """The update kernel's T_cell layouts and workgroup shapes: FP32 with an
    odd number of 16-member output tiles (48, 80: rows padded, 8-wave
    workgroups) and an even one (64: 16-column blocks swapped, 4-wave
    workgroups); mixed up to 96 members in 4-wave workgroups and beyond (112)
    in 8-wave ones."""
    print("case update layout nmem=%d precision=%d" % (nmem, precision), flush=True)
    P = synth.make_problem(seed=60 + nmem, nmem=nmem, nobs=1000, locrad=60.0, nj=25, ni=31, stride=3, land_frac=0.1)
    T, w = transforms(P, 1, 1, nmem)
    bufs, fl = fields_of(P, nmem)
    ref = [b.copy() for b in bufs]
    gpu = [b.copy() for b in bufs]
    infl_o = O.update_fields(1, nmem, nmem, 3, 0, 0, P.nlevels, T, w, fl, ref, 0, P.nj - 1, want_infl=True)
    E = enkfg.Engine(nmem, mode=enkfg.MODE_ENKF, precision=precision)
    E.set_transforms(T, w, 3)
    infl_g = E.update_fields(P.nlevels, 3, 0, 0, fl, gpu, 0, P.nj - 1, want_infl=True)
    tol, itol = (1e-5, 1e-5) if precision == enkfg.PREC_FP64 else (2e-5, 1e-4)
    for a, b in zip(gpu, ref):
        check(a, b, tol)
    assert np.abs(infl_g - infl_o).max() <= itol
    E.close()


def test_update_tile_rows():
    print("case update j-tile", flush=True)
    P = synth.make_problem(seed=21, nmem=24, nobs=1000, locrad=60.0, nj=33, ni=40, stride=3)
    T, w = transforms(P, 1, 2, 24)
    bufs, fl = fields_of(P, 24)
    j1, j2 = 12, 23
    tile = [np.ascontiguousarray(b[:, j1:j2 + 1]) for b in bufs]
    ref = [b.copy() for b in tile]
    gpu = [b.copy() for b in tile]
    O.update_fields(1, 24, 24, 3, 0, 0, P.nlevels, T, w, fl, ref, j1, j2)
    E = enkfg.Engine(24, mode=enkfg.MODE_ENKF)
    E.set_transforms(T, w, 3)
    E.update_fields(P.nlevels, 3, 0, 0, fl, gpu, j1, j2)
    for a, b in zip(gpu, ref):
        check(a, b)
    E.close()


@pytest.mark.parametrize("m,md", [(40, 28), (144, 96), (37, 25)])
def test_update_hybrid_with_conditioning(m, md):
    print("case update hybrid m=%d md=%d" % (m, md), flush=True)
    P = synth.make_problem(seed=5, nmem=m, nobs=1200, locrad=60.0, nj=30, ni=36, stride=3)
    T, w = transforms(P, 3, 1, md)
    bufs, fl = fields_of(P, m)
    ref = [b.copy() for b in bufs]
    for b in ref:
        O.sethybridensemble(m, md, 0.7, b.reshape(m, -1))
    gpu = [b.copy() for b in bufs]
    O.update_fields(3, m, md, 3, 0, 0, P.nlevels, T, w, fl, ref, 0, P.nj - 1)
    E = enkfg.Engine(m, nmem_dynamic=md, mode=enkfg.MODE_HYBRID, gamma=0.7)
    E.set_transforms(T, w, 3)
    E.update_fields(P.nlevels, 3, 0, 0, fl, gpu, 0, P.nj - 1, hybrid_condition=True)
    for a, b in zip(gpu, ref):
        check(a, b, tol=2e-5)
    E.close()


def test_update_bg_enoi():
    print("case update bg", flush=True)
    P = synth.make_problem(seed=9, nmem=24, nobs=900, locrad=60.0, nj=27, ni=30, stride=3)
    _, w = transforms(P, 2, 1, 24)
    bufs = []
    fl = []
    for k in range(P.nk):
        b = np.concatenate([P.ens["temp"][:, k], P.truth["temp"][k][None]], axis=0)
        bufs.append(np.ascontiguousarray(b))
        fl.append(dict(level=k))
    ref = [b.copy() for b in bufs]
    gpu = [b.copy() for b in bufs]
    O.update_bg(24, 3, 0, 0, P.nlevels, w, fl, ref, 0, P.nj - 1)
    E = enkfg.Engine(24, mode=enkfg.MODE_ENOI)
    E.set_transforms(None, w, 3)
    E.update_bg(P.nlevels, 3, 0, 0, fl, gpu, 0, P.nj - 1)
    for k, (a, b) in enumerate(zip(gpu, ref)):
        d = np.argwhere(a != b)
        msg = "field %d: %d differences, first at %s: gpu %r oracle %r" % (
            k, len(d), d[:3].tolist(), [float(a[tuple(x)]) for x in d[:3]], [float(b[tuple(x)]) for x in d[:3]])
        assert len(d) == 0, "EnOI background update is float-sequential and must match bit for bit; " + msg
    E.close()


@pytest.mark.parametrize("precision", [enkfg.PREC_FP64, enkfg.PREC_MIXED])
def test_update_randomise(precision):
    """RANDOMISE fields (addnoise, update.c:46-87, applied after the update
    at 554-563): the engine's drand48 stream from a set state against the
    oracle's update followed by addnoise on the same state, field by field in
    batch order, for the host-buffer call and for the device-resident
    ensemble; the streams end in the same state."""
    print("case update randomise precision=%d" % precision, flush=True)
    m = 24
    P = synth.make_problem(seed=41, nmem=m, nobs=1000, locrad=60.0, nj=25, ni=31, stride=3, land_frac=0.1)
    T, w = transforms(P, 1, 1, m)
    bufs, fl = fields_of(P, m)
    for k in (0, 3, len(fl) - 1):           # three randomised fields, one of them 2D
        fl[k] = dict(fl[k], deflation=0.9 - 0.1 * (k % 2), sigma=0.2 + 0.01 * k)
    state0 = 0x5EED1234
    ref = [b.copy() for b in bufs]
    O.update_fields(1, m, m, 3, 0, 0, P.nlevels, T, w, fl, ref, 0, P.nj - 1)
    st = state0
    for k, f in enumerate(fl):
        if "deflation" in f:
            st = O.addnoise(m, f["deflation"], f["sigma"], st, ref[k])
    tol = 1e-5 if precision == enkfg.PREC_FP64 else 2e-5
    E = enkfg.Engine(m, mode=enkfg.MODE_ENKF, precision=precision)
    E.set_transforms(T, w, 3)
    assert E.rand48() == 0, "the stream starts where an unseeded drand48 starts"
    E.set_rand48(state0)
    gpu = [b.copy() for b in bufs]
    E.update_fields(P.nlevels, 3, 0, 0, fl, gpu, 0, P.nj - 1)
    assert E.rand48() == st
    for k, (a, b) in enumerate(zip(gpu, ref)):
        scale = max(np.abs(b).max(), 1.0)
        assert np.abs(a.astype(np.float64) - b).max() <= tol * scale, "field %d" % k
    # device-resident ensemble: the same fields in [field][cell][member]
    import synth_gpu
    Ecell = np.stack([b.reshape(m, -1).T for b in bufs]).astype(np.float32)
    dE = synth_gpu.DeviceArray.from_host(Ecell)
    E.set_rand48(state0)
    E.ens_attach(P.nlevels, 3, 0, 0, fl, dE.ptr)
    E.ens_update()
    assert E.rand48() == st
    out = dE.to_host()
    for k, b in enumerate(ref):
        scale = max(np.abs(b).max(), 1.0)
        assert np.abs(out[k].T.reshape(b.shape).astype(np.float64) - b).max() <= tol * scale, "field %d" % k
    E.close()


@pytest.mark.parametrize("mode,m,md,stride,periodic", [(enkfg.MODE_ENKF, 24, 24, 3, 0), (enkfg.MODE_ENKF, 96, 96, 3, 1),
                                                      (enkfg.MODE_HYBRID, 40, 28, 2, 0),
                                                      (enkfg.MODE_ENOI, 30, 30, 3, 0)])
def test_update_HE_parity(mode, m, md, stride, periodic):
    """The analysed ensemble observations (update_HE_structured /
    update_Hx_structured) against the oracle: obs anywhere in the grid,
    several per cell (profiles), capped and plain inflation, obs not OK or
    rounding past the last row left unchanged; within 1e-5 of max |HE|."""
    print("case update_HE mode=%d m=%d md=%d stride=%d periodic=%d" % (mode, m, md, stride, periodic), flush=True)
    P = synth.make_problem(seed=50 + m, nmem=m, nobs=1500, locrad=70.0, nj=31, ni=40, stride=stride)
    scheme = enkfg.SCHEME_DENKF
    T, w = transforms(P, mode, scheme, md)
    rng = np.random.default_rng(m + stride)
    nobs = 3000
    fi = rng.uniform(-0.49, P.ni - 0.51, nobs)
    fj = rng.uniform(-0.49, P.nj - 0.51, nobs)
    fi[:200] = np.repeat(fi[:10], 20)            # 20 obs per cell
    fj[:200] = np.repeat(fj[:10], 20)
    fi[200:210] = P.ni - 0.5                     # rounds to mni: the last column
    fj[210:220] = P.nj - 0.4                     # rounds past the last row: unchanged
    status = np.zeros(nobs, np.int16)
    status[300:340] = 2
    HE = (10.0 + rng.standard_normal((m, nobs))).astype(np.float32)
    infl = rng.uniform(1.0, 1.2, nobs).astype(np.float32)
    ratio = np.where(rng.random(nobs) < 0.5, 1.0, np.nan)
    Tt = np.zeros((P.nj_s, P.ni_s, 1, 1), np.float32) if T is None else T
    ref = O.update_HE(mode, m, md, stride, periodic, P.nj, P.ni, Tt, w, fi, fj, status, infl, ratio, HE)
    E = enkfg.Engine(m, nmem_dynamic=md, mode=mode)
    E.set_transforms(T, w, stride)
    got = E.update_HE(P.nj, P.ni, stride, periodic, fi, fj, status, infl, ratio, HE)
    E.close()
    same = np.all(ref == HE, axis=0)
    assert same[210:220].all() and same[300:340].all()
    assert np.array_equal(got[:, same], HE[:, same])
    err = float(np.abs(got.astype(np.float64) - ref).max() / np.abs(ref).max())
    print("  max rel error %.2e, %d obs updated" % (err, int((~same).sum())), flush=True)
    assert err <= 1e-5
