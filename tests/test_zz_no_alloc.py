"""enkfg_update_fields_noalloc (no device allocation: staging in a fixed
per-GPU scratch array) against enkfg_update_fields, through the C ABI.

Both run the same kernels on the same inputs, so the outputs -- fields,
inflation factors and the RANDOMISE stream state -- must be identical bit for
bit.  Two calls in a row on one context, the second with fewer fields and
fewer rows, show that nothing of the first call leaks into the second; a
batch beyond the scratch size is refused with ENKFG_ERR_ARG before any
device work.
"""
import numpy as np
import pytest

import enkfg
import oracle as O
import synth

pytestmark = pytest.mark.gpu


def transforms(P, md):
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(P.nmem, ob, P.HA, P.innov)
    r = O.calc_transforms(enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, P.nmem, md, 1.0, ob, S, s, P.node_lon.ravel(),
                          P.node_lat.ravel())
    return r["T"].reshape(P.nj_s, P.ni_s, md, P.nmem), r["w"].reshape(P.nj_s, P.ni_s, P.nmem)


def test_noalloc_update_equals_update_fields():
    print("case update_fields_noalloc vs update_fields", flush=True)
    m = 24
    P = synth.make_problem(seed=61, nmem=m, nobs=1200, locrad=60.0, nj=31, ni=44, stride=3, land_frac=0.15)
    T, w = transforms(P, m)
    bufs, fl = [], []
    for k in range(P.nk):
        bufs.append(np.ascontiguousarray(P.ens["temp"][:m, k]))
        fl.append(dict(level=k, inflation=1.1, inf_ratio=1.0))
        bufs.append(np.ascontiguousarray(P.ens["salt"][:m, k]))
        fl.append(dict(level=k, inflation=1.0, inf_ratio=float("nan")))
    bufs.append(np.ascontiguousarray(P.ens["eta"][:m, 0]))
    fl.append(dict(level=0, issurfacevar=1, inflation=1.05, inf_ratio=0.5, deflation=0.9, sigma=0.05))
    bufs[0][3, 5, 7] = np.nan                     # a missing dynamic member
    E1 = enkfg.Engine(m, mode=enkfg.MODE_ENKF)
    E2 = enkfg.Engine(m, mode=enkfg.MODE_ENKF)
    for E in (E1, E2):
        E.set_transforms(T, w, 3)
        E.set_rand48(12345)
    calls = [(list(range(len(fl))), 0, P.nj - 1), ([len(fl) - 1, 2], 3, 14)]   # the second call is smaller
    for c, (fsel, j1, j2) in enumerate(calls):
        f = [fl[k] for k in fsel]
        orig = [bufs[k][:, j1:j2 + 1].copy() for k in fsel]
        a = [x.copy() for x in orig]
        b = [x.copy() for x in orig]
        ia = E1.update_fields(P.nlevels, 3, 0, 0, f, a, j1, j2, want_infl=True)
        ib = E2.update_fields(P.nlevels, 3, 0, 0, f, b, j1, j2, want_infl=True, noalloc=True)
        for k, (x, y) in enumerate(zip(a, b)):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), "call %d field %d" % (c, k)
        assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32)), "call %d inflation" % c
        assert E1.rand48() == E2.rand48(), "call %d RANDOMISE stream" % c
        assert not any(np.array_equal(x, y, equal_nan=True) for x, y in zip(orig, b)), "a field was not updated"
    E1.close()
    E2.close()


def test_noalloc_refuses_batches_beyond_the_scratch():
    print("case update_fields_noalloc size limit", flush=True)
    m, mnj, mni = 2, 1500, 2000
    E = enkfg.Engine(m, mode=enkfg.MODE_ENKF)
    T = np.tile(np.eye(m, dtype=np.float32), (mnj, mni, 1, 1))
    w = np.zeros((mnj, mni, m), np.float32)
    E.set_transforms(T, w, 1)
    nl = np.ones((mnj, mni), np.int32)
    small = [np.zeros((m, 1, mni), np.float32)]
    with pytest.raises(enkfg.EnkfgError) as ei:
        E.update_fields(nl, 1, 0, 0, [dict(level=0)], [np.zeros((m, mnj, mni), np.float32)], 0, mnj - 1,
                        want_infl=True, noalloc=True)
    assert ei.value.code == enkfg.ERR_ARG and "scratch" in str(ei.value)
    # a tile of rows within the limit works, and equals the batch call
    a = [np.random.default_rng(1).standard_normal((m, 1, mni)).astype(np.float32)]
    b = [a[0].copy()]
    E.update_fields(nl, 1, 0, 0, [dict(level=0)], a, 7, 7)
    E.update_fields(nl, 1, 0, 0, [dict(level=0)], b, 7, 7, noalloc=True)
    assert np.array_equal(a[0], b[0])
    del small
    E.close()
