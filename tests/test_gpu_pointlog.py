"""GPU parity of the point logs (das_calcpointlogtransforms,
calc/transforms.c:1640-1770): the log points' transforms come from the
engine's calc over a row of points and are written to one file per point;
the files are read back and compared with the oracle's calc at the same
points.  Local obs ids equal the oracle's search and the taper coefficients
its float cast; s and S are the float casts of the oracle's standardised
values times the taper, exactly; w and T (double) within 1e-10 relative to
max|w|, max|T| (the bar of the FP64 calc).  A point far from every obs gets
the empty-point file: w = 0, T = I."""
import numpy as np
import pytest
from scipy.io import netcdf_file

import enkfg
import oracle as O
import pointlog
import synth

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = [(enkfg.MODE_ENKF, enkfg.SCHEME_DENKF, 24, 24, 0.8),
         (enkfg.MODE_HYBRID, enkfg.SCHEME_ETKF, 40, 24, 1.0),
         (enkfg.MODE_ENOI, enkfg.SCHEME_DENKF, 32, 32, 1.0)]


@pytest.mark.parametrize("mode,scheme,m,md,alpha", CASES)
def test_pointlog_transforms_parity(tmp_path, mode, scheme, m, md, alpha):
    print("case pointlog mode=%d scheme=%d m=%d md=%d" % (mode, scheme, m, md), flush=True)
    P = synth.make_problem(seed=7 + m, nmem=m, profiles=20, nobs=2500, locrad=75.0, nj=45, ni=60)
    ob = O.OracleObs(P.obstypes, P.obs_type, P.obs_lon, P.obs_lat, P.obs_estd, status=P.obs_status)
    S, s = O.standardise(m, ob, P.HA, P.innov)
    # points between the stride nodes, one on a node, one far from every obs
    lon0, lat0 = P.node_lon.ravel(), P.node_lat.ravel()
    pts = [(lon0[7] + 0.037, lat0[7] + 0.021), (lon0[40] - 0.113, lat0[40] + 0.07), (lon0[100], lat0[100]),
           (lon0[5] + 40.0, lat0[5] - 30.0)]
    plon = np.array([p[0] for p in pts])
    plat = np.array([p[1] for p in pts])
    ref = O.calc_transforms(mode, scheme, m, md, alpha, ob, S, s, plon, plat)

    E = enkfg.Engine(m, nmem_dynamic=md, mode=mode, scheme=scheme, alpha=alpha, keep_fp64=True)
    E.set_obstypes(P.obstypes)
    E.set_obs(P.obs_type, P.obs_status, P.obs_lon, P.obs_lat, P.obs_estd)
    E.set_S(P.HA, P.innov, standardised=False)
    obs = dict(lon=P.obs_lon, lat=P.obs_lat, estd=P.obs_estd, value=P.obs_value, type=P.obs_type)
    paths = pointlog.calc_pointlogs(E, pts, obs, outdir=str(tmp_path))
    E.close()
    assert len(paths) == len(pts)
    nonempty = 0
    for k, path in enumerate(paths):
        ids_o, lc_o = ob.findlocal(plon[k], plat[k])
        with netcdf_file(path, "r", mmap=False) as f:
            v = f.variables
            assert f.lon == plon[k] and f.lat == plat[k]
            if len(ids_o) == 0:
                assert "s" not in v and v["obs_ids"][:].shape == (0,)
                assert np.array_equal(v["w"][:], np.zeros(m))
                if mode != enkfg.MODE_ENOI:
                    assert np.array_equal(v["T"][:], np.eye(md, m))
                continue
            nonempty += 1
            assert np.array_equal(v["obs_ids"][:], ids_o), "point %d ids" % k
            assert np.array_equal(v["obs_lcoeffs"][:], lc_o.astype(np.float32))
            assert np.array_equal(v["obs_val"][:], P.obs_value[ids_o].astype(np.float32))
            assert np.array_equal(v["s"][:], (s[ids_o] * lc_o).astype(np.float32))
            assert np.array_equal(v["S"][:], (S[:, ids_o].astype(np.float64) * lc_o).astype(np.float32))
            assert relerr(v["w"][:], ref["w64"][k]) < 1e-10, "point %d w" % k
            if mode != enkfg.MODE_ENOI:
                assert v["T"][:].shape == (md, m)
                assert relerr(v["T"][:], ref["T64"][k]) < 1e-10, "point %d T" % k
            else:
                assert "T" not in v
    assert nonempty == 3
