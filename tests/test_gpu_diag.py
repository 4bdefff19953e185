"""Update-side diagnostics on the device (diag/diags.c): the ensemble spread
of das_writespread (enkfg_ens_spread on the attached ensemble,
enkfg_spread_fields on host field buffers) against the oracle bit for bit,
and spread.nc / inflation.nc written with libenkfio and read back with an
independent netCDF reader (scipy)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import enkfg
import enkfio
import oracle as O
import synth
import synth_gpu

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    """equal bit for bit, NaN where the other is NaN (its payload aside)"""
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def ensemble(seed, m, nf, nj, ni):
    rng = np.random.default_rng(seed)
    E = (rng.standard_normal((nf, m, nj, ni)) * 0.6 + rng.uniform(-20, 30, (nf, 1, nj, ni))).astype(np.float32)
    E[0, :, 0, :5] = np.float32(1.1)                        # constant cells: the float square decides
    E[0, :, 1, :5] += np.float32(2e6)                       # |mean| > MAXOBSVAL
    E[1 % nf, :, 2, :5] = -np.float32(3e6)
    E[-1, 3, 3, 3] = np.nan                                 # a missing member: NaN as in the reference
    return E


@pytest.mark.parametrize("m", [96, 37, 144])
def test_ens_spread_matches_oracle(m):
    print("case ens_spread m=%d" % m, flush=True)
    nf, nj, ni = 3, 41, 53
    E = ensemble(m, m, nf, nj, ni)
    ref = np.stack([O.spread(E[f]) for f in range(nf)])
    eng = enkfg.Engine(m, mode=enkfg.MODE_ENKF)
    nlev = np.ones((nj, ni), np.int32)
    dE = synth_gpu.DeviceArray.from_host(np.ascontiguousarray(E.reshape(nf, m, -1).transpose(0, 2, 1)))
    eng.ens_attach(nlev, 1, 0, 0, [dict(level=0)] * nf, dE.ptr)
    sp = eng.ens_spread(0, nf, (nj, ni))
    assert same_bits(sp, ref) and np.isnan(ref[-1, 3, 3])
    sp1 = eng.ens_spread(1, 2, (nj, ni))
    assert same_bits(sp1, ref[1:])
    # into device memory
    dout = synth_gpu.DeviceArray((nf, nj, ni), np.float32)
    eng.ens_spread(0, nf, (nj, ni), dev_out=dout.ptr)
    assert same_bits(dout.to_host(), ref)
    print("  spread bit for bit; kernel %.3f ms" % eng.spread_ms(), flush=True)
    with pytest.raises(enkfg.EnkfgError):
        eng.ens_spread(2, 2, (nj, ni))
    eng.close()


def test_spread_fields_host_buffers_and_spread_file(tmp_path):
    print("case spread_fields", flush=True)
    m, nf, nj, ni = 24, 4, 30, 36
    E = ensemble(5, m, nf, nj, ni)
    eng = enkfg.Engine(m, mode=enkfg.MODE_ENKF)
    bufs = [np.ascontiguousarray(E[f]) for f in range(nf)]
    sp = eng.spread_fields(bufs)
    ref = np.stack([O.spread(E[f]) for f in range(nf)])
    assert same_bits(sp, ref)
    names = ["temp", "salt", "u", "v"]
    p = str(tmp_path / "spread.nc")
    enkfio.write_fields2d(p, dict(zip(names, sp)))
    with netcdf_file(p, "r", mmap=False) as f:
        for k, n in enumerate(names):
            assert same_bits(f.variables[n][...].astype(np.float32), sp[k])
    eng.close()


def test_inflation_file_from_the_update(tmp_path):
    """The update's inflation factors (update/update.c:518-526) written as
    inflation.nc and read back, equal to the oracle's."""
    print("case inflation file", flush=True)
    P = synth.make_problem(seed=2, nmem=24, nobs=1200, locrad=70.0, nj=24, ni=30, stride=3)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd)
    S, s = O.standardise(P.nmem, ob, P.HA, P.innov)
    ref = O.calc_transforms(1, 1, P.nmem, P.nmem, 1.0, ob, S, s, P.node_lon.ravel(), P.node_lat.ravel())
    E = enkfg.Engine(P.nmem, mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF)
    E.set_obstypes(P.obstypes)
    E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    E.set_S(P.HA, P.innov, standardised=False)
    E.calc_transforms(P.node_lon, P.node_lat, stride=P.stride)
    fl = [dict(level=0, inflation=1.08, inf_ratio=1.0), dict(level=1, inflation=1.05, inf_ratio=float("nan"))]
    bufs = [np.ascontiguousarray(P.ens["temp"][:, 0]), np.ascontiguousarray(P.ens["temp"][:, 1])]
    rb = [b.copy() for b in bufs]
    infl = E.update_fields(P.nlevels, P.stride, 0, 0, fl, bufs, 0, P.nj - 1, want_infl=True)
    rinfl = O.update_fields(1, P.nmem, P.nmem, P.stride, 0, 0, P.nlevels,
                            ref["T"].reshape(P.nj_s, P.ni_s, P.nmem, P.nmem), ref["w"].reshape(P.nj_s, P.ni_s, P.nmem),
                            fl, rb, 0, P.nj - 1, want_infl=True)
    assert (rinfl > 1.0).any()
    assert np.abs(infl - rinfl).max() <= 1e-5
    p = str(tmp_path / "inflation.nc")
    enkfio.write_fields2d(p, {"temp_0": infl[0], "temp_1": infl[1]})
    with netcdf_file(p, "r", mmap=False) as f:
        assert np.array_equal(f.variables["temp_0"][...].astype(np.float32), infl[0])
        assert np.array_equal(f.variables["temp_1"][...].astype(np.float32), infl[1])
    E.close()
