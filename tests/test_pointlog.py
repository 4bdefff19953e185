"""Point-log files (common/pointlog.c:83-426 das_createplog and
das_writeplogtransform, file names as common/dasystem.c:585-588), written by
pointlog.write_pointlog and read back with scipy's netCDF reader: dimensions,
variables, their values (floats as the reference casts them) and the
attributes, for an EnKF point with local obs, a hybrid point on two grids
(the per-grid p-<gid> dimensions), and a point with none (p is then the
record dimension with no records, as the reference's ncw_def_dim(..., 0)
makes it)."""
import numpy as np
from scipy.io import netcdf_file

import enkfg
import pointlog


def obs_dict(n, rng):
    return dict(lon=rng.uniform(140, 150, n).astype(np.float32), lat=rng.uniform(-45, -35, n).astype(np.float32),
                depth=rng.uniform(0, 100, n).astype(np.float32), value=rng.normal(size=n).astype(np.float32),
                estd=rng.uniform(0.1, 1, n).astype(np.float32), fi=rng.uniform(0, 90, n), fj=rng.uniform(0, 60, n),
                fk=rng.uniform(0, 3, n).astype(np.float32), type=rng.integers(0, 2, n).astype(np.int16),
                instrument=rng.integers(0, 3, n).astype(np.int16), time=rng.uniform(-1, 1, n).astype(np.float32))


TYPES = [dict(name="SST", rfactor=2.0, locrad=[100.0, 300.0], locweight=[3.0, 1.0]),
         dict(name="TEM", rfactor=1.0, locrad=[150.0], domains=["Default", "Deep"])]


def test_pointlog_enkf_values_and_attributes(tmp_path):
    rng = np.random.default_rng(0)
    n, m, p = 50, 8, 7
    obs = obs_dict(n, rng)
    ids = np.sort(rng.choice(n, p, replace=False)).astype(np.int32)
    lc = rng.uniform(0, 1, p)
    tr = dict(s=rng.normal(size=p), S=rng.normal(size=(m, p)), w=rng.normal(size=m), T=rng.normal(size=(m, m)))
    path = str(tmp_path / pointlog.plog_fname(145.25, -40.5))
    assert path.endswith("pointlog_145.250,-40.500.nc")
    pointlog.write_pointlog(path, enkfg.MODE_ENKF, m, m, 145.25, -40.5, ids, lc, obs,
                            [dict(name="t-grid", fi=12.5, fj=7.25, nk=4, model_depth=55.0)], [tr],
                            scheme=enkfg.SCHEME_DENKF, alpha=0.8, obstypes=TYPES, instruments=["A", "B", "C"],
                            datestr="2024-01-01")
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions["m"] == m and f.dimensions["p"] == p
        v = f.variables
        assert np.array_equal(v["obs_ids"][:], ids)
        assert np.array_equal(v["obs_lcoeffs"][:], lc.astype(np.float32))
        for name, key in (("obs_lon", "lon"), ("obs_lat", "lat"), ("obs_val", "value"), ("obs_estd", "estd"),
                          ("obs_fi", "fi"), ("obs_fj", "fj"), ("obs_fk", "fk"), ("obs_time", "time")):
            assert np.array_equal(v[name][:], obs[key][ids].astype(np.float32)), name
        assert np.array_equal(v["obs_type"][:], obs["type"][ids])
        assert np.array_equal(v["obs_inst"][:], obs["instrument"][ids])
        assert np.array_equal(v["s"][:], tr["s"].astype(np.float32))
        assert v["S"].dimensions == ("m", "p") and np.array_equal(v["S"][:], tr["S"].astype(np.float32))
        assert v["w"][:].dtype.kind == "f" and v["w"][:].dtype.itemsize == 8 and np.array_equal(v["w"][:], tr["w"])
        assert v["T"].dimensions == ("m", "m") and np.array_equal(v["T"][:], tr["T"])
        assert v["obs_time"].units == b"days from 2024-01-01"
        ot = v["obs_type"]
        assert ot.SST == 0 and ot.TEM == 1 and ot.RFACTOR_SST == 2.0
        assert np.allclose(ot.LOCRAD_SST, [100, 300]) and np.allclose(ot.WEIGHT_SST, [0.75, 0.25])
        assert ot.DOMAINS_TEM == b"Default Deep"
        assert v["obs_inst"].C == 2
        g = v["grid"]
        assert g.id == 0 and g.name == b"t-grid" and g.fi == 12.5 and g.fj == 7.25 and g.nk == 4
        assert f.lon == 145.25 and f.lat == -40.5 and f.MODE == b"EnKF" and f.SCHEME == b"DEnKF"
        assert f.ALPHA == 0.8 and f.ngrids == 1
        assert v["w"].long_name == b'increment coefficients on grid 0 ("t-grid")'


def test_pointlog_hybrid_two_grids_and_empty_point(tmp_path):
    rng = np.random.default_rng(1)
    n, m, md = 30, 12, 8
    obs = obs_dict(n, rng)
    ids = np.arange(5, dtype=np.int32)
    lc = np.linspace(1, 0.2, 5)
    trs = [dict(s=rng.normal(size=5), S=rng.normal(size=(m, 5)), w=rng.normal(size=m), T=rng.normal(size=(md, m))),
           dict(s=rng.normal(size=3), S=rng.normal(size=(m, 3)), w=rng.normal(size=m), T=rng.normal(size=(md, m)))]
    path = str(tmp_path / "hyb.nc")
    pointlog.write_pointlog(path, enkfg.MODE_HYBRID, m, md, 10.0, 20.0, ids, lc, obs,
                            [dict(name="a"), dict(name="b")], trs, scheme=enkfg.SCHEME_ETKF, alpha=1.0, gamma=0.5)
    with netcdf_file(path, "r", mmap=False) as f:
        d = f.dimensions
        assert d["m1"] == md and d["m2"] == m and d["p"] == 5 and d["p-0"] == 5 and d["p-1"] == 3
        v = f.variables
        assert v["T-1"].dimensions == ("m1", "m2") and np.array_equal(v["T-1"][:], trs[1]["T"])
        assert v["S-0"].dimensions == ("m2", "p-0")
        assert np.array_equal(v["S-1"][:], trs[1]["S"].astype(np.float32))
        assert f.MODE == b"Hybrid" and f.SCHEME == b"ETKF" and f.GAMMA == 0.5 and f.ngrids == 2
        assert getattr(v["dimension_reference"], "p-1") == b"number of observations on grid 1"

    # no local obs: p is the record dimension with no records; w = 0, T = I
    path0 = str(tmp_path / "empty.nc")
    pointlog.write_pointlog(path0, enkfg.MODE_ENKF, m, m, 0.0, 0.0, np.zeros(0, np.int32), np.zeros(0), obs,
                            [dict(name="a")], [dict(s=np.zeros(0), S=np.zeros((m, 0)), w=np.zeros(m), T=np.eye(m))])
    with netcdf_file(path0, "r", mmap=False) as f:
        assert f.dimensions["p"] is None or f.dimensions["p"] == 0
        v = f.variables
        assert v["obs_ids"][:].shape == (0,) and v["obs_lon"][:].shape == (0,)
        assert "s" not in v and "S" not in v
        assert np.array_equal(v["w"][:], np.zeros(m)) and np.array_equal(v["T"][:], np.eye(m))
