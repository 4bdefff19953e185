"""libenkfio.so (include/enkfio.h) on the CPU: the netCDF classic files it
writes are read back by an independent implementation (scipy.io.netcdf_file),
files scipy writes are read by it, and the path's files carry the reference's
names, dimensions and attributes (calc/transforms.c:54-135, 375-520;
diag/diags.c:79-168)."""
import os
import re
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import enkfio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "enkfio.h")
LIB = os.path.join(ROOT, "enkf-c_amd", "libenkfio.so")


def test_header_declarations_are_exported():
    txt = open(HEADER).read()
    names = sorted(set(re.findall(r"^\s*int\s+((?:ncf|enkfio)_\w+)\s*\(", txt, re.M)))
    assert names == sorted(enkfio.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in names:
        assert n in syms, n


def scipy_read(path):
    with netcdf_file(path, "r", mmap=False) as f:
        return ({k: v for k, v in f.dimensions.items()},
                {k: (v[...].copy(), dict(v._attributes)) for k, v in f.variables.items()},
                dict(f._attributes), f.version_byte)


@pytest.mark.parametrize("enoi", [False, True])
def test_transforms_file_round_trip(tmp_path, enoi):
    rng = np.random.default_rng(3)
    nj, ni, md, m = 5, 7, 6, 9
    w = rng.standard_normal((nj, ni, m)).astype(np.float32)
    T = None if enoi else rng.standard_normal((nj, ni, md, m)).astype(np.float32)
    p = str(tmp_path / "transforms.nc")
    enkfio.write_transforms(p, w, T, stride=3, grid_name="t-grid")
    dims, vars_, gatts, ver = scipy_read(p)
    assert ver == 2
    if enoi:
        assert dims == {"j": nj, "i": ni, "m": m} and set(vars_) == {"w"}
    else:
        assert dims == {"j": nj, "i": ni, "m_dyn": md, "m": m} and set(vars_) == {"w", "T"}
        assert vars_["T"][0].dtype == np.dtype(">f4")
        assert np.array_equal(vars_["T"][0].astype(np.float32).view(np.uint32), T.view(np.uint32))
    assert np.array_equal(vars_["w"][0].astype(np.float32).view(np.uint32), w.view(np.uint32))
    assert int(gatts["stride"]) == 3 and gatts["grid_name"] == b"t-grid"
    d = enkfio.inq_transforms(p)
    assert d == dict(nj=nj, ni=ni, m_dyn=0 if enoi else md, m=m, stride=3)
    T2, w2 = enkfio.read_transforms(p, 2, 2)
    assert np.array_equal(w2, w[2:4])
    if not enoi:
        assert np.array_equal(T2, T[2:4])
    else:
        assert T2 is None


def test_diag_file_layout(tmp_path):
    rng = np.random.default_rng(4)
    names = ["SST", "SLA", "TEM"]
    nj, ni = 4, 6
    pn = rng.integers(0, 500, (4, nj, ni)).astype(np.int32)
    pd = rng.uniform(0, 20, (4, nj, ni)).astype(np.float32)
    ps = rng.uniform(0, 2, (4, nj, ni)).astype(np.float32)
    ps[3, 0, 0] = np.nan
    p = str(tmp_path / "enkf_diag.nc")
    enkfio.write_diag(p, names, 3, pn, pd, ps)
    dims, v, g, _ = scipy_read(p)
    assert dims == {"nobstypes": 3, "j": nj, "i": ni}
    assert np.array_equal(v["nlobs"][0], pn[3]) and np.array_equal(v["pnlobs"][0], pn[:3])
    assert np.array_equal(v["dfs"][0], pd[3]) and np.array_equal(v["pdfs"][0], pd[:3])
    assert np.array_equal(v["srf"][0], ps[3], equal_nan=True) and np.array_equal(v["psrf"][0], ps[:3])
    assert v["nlobs"][0].dtype == np.dtype(">i4")
    assert v["dfs"][1]["long_name"] == b"local DFS (Degrees of Freedom for Signal)"
    assert int(g["stride"]) == 3
    assert [int(g[n]) for n in names] == [0, 1, 2]


def test_fields2d_file(tmp_path):
    rng = np.random.default_rng(5)
    f = {"temp": rng.standard_normal((3, 4)), "salt": rng.standard_normal((3, 4)), "eta": np.ones((3, 4))}
    p = str(tmp_path / "spread.nc")
    enkfio.write_fields2d(p, f)
    dims, v, _, _ = scipy_read(p)
    assert dims == {"j": 3, "i": 4} and list(v) == ["temp", "salt", "eta"]
    for k in f:
        assert np.array_equal(v[k][0], f[k].astype(np.float32))


@pytest.mark.parametrize("version", [1, 2])
def test_reads_what_scipy_writes(tmp_path, version):
    rng = np.random.default_rng(6)
    p = str(tmp_path / "obs.nc")
    data = dict(type=rng.integers(0, 5, 11).astype(np.int16), status=rng.integers(0, 3, 11).astype(np.int32),
                lon=rng.uniform(0, 360, 11).astype(np.float32), value=rng.standard_normal(11),
                flag=rng.integers(-100, 100, 11).astype(np.int8), grid=rng.standard_normal((3, 5, 2)))
    with netcdf_file(p, "w", version=version) as f:
        f.createDimension("nobs", 11)
        f.createDimension("a", 3)
        f.createDimension("b", 5)
        f.createDimension("c", 2)
        f.DA_DAY = 21245.5
        f.title = "synthetic observations"
        f.counts = np.array([1, 2, 3], np.int32)
        for k, a in data.items():
            var = f.createVariable(k, a.dtype, ("a", "b", "c") if a.ndim == 3 else ("nobs",))
            var[...] = a
        f.variables["value"].units = "degC"
        f.variables["value"].scale = np.float32(0.5)
    with enkfio.NcReader(p) as r:
        assert r.variables() == list(data)
        assert r.dimlen("nobs") == 11 and r.dimlen("b") == 5
        for k, a in data.items():
            got = r.read(k)
            assert got.dtype == a.dtype.newbyteorder("=") and np.array_equal(got, a), k
        assert np.array_equal(r.read("grid", 1, 2), data["grid"][1:3])
        assert r.attr("DA_DAY")[0] == 21245.5
        assert r.attr("title") == "synthetic observations"
        assert list(r.attr("counts")) == [1, 2, 3]
        assert r.attr("units", "value") == "degC" and r.attr("scale", "value")[0] == np.float32(0.5)
        with pytest.raises(enkfio.EnkfioError) as e:
            r.read("grid", 2, 2)
        assert e.value.code == enkfio.ERANGE
        with pytest.raises(enkfio.EnkfioError):
            r.read("missing")


def test_header_longer_than_the_first_read(tmp_path):
    """A header beyond the 64 KiB read first (many variables, long
    attributes) is read by growing the prefix."""
    p = str(tmp_path / "big.nc")
    with netcdf_file(p, "w", version=2) as f:
        f.createDimension("n", 4)
        f.history = "x" * 150000
        for k in range(300):
            v = f.createVariable("v%03d" % k, np.float32, ("n",))
            v[...] = np.arange(4, dtype=np.float32) + k
            v.long_name = "variable number %d" % k
    with enkfio.NcReader(p) as r:
        assert len(r.variables()) == 300
        assert np.array_equal(r.read("v299"), np.arange(4, dtype=np.float32) + 299)
        assert r.attr("history") == "x" * 150000


def test_writer_api_and_errors(tmp_path):
    p = str(tmp_path / "w.nc")
    with enkfio.NcWriter(p) as w:
        w.dim("j", 3)
        w.dim("i", 2)
        w.var("a", np.float64, ("j", "i"))
        w.var("k", np.int16, ("i",))
        w.var("s", np.int32, ())
        w.attr("note", "two rows then one", var="a")
        w.attr("ratio", 0.25)
        w.enddef()
        a = np.arange(6, dtype=np.float64).reshape(3, 2) - 2.5
        w.put("a", a[:2])
        w.put("a", a[2:], start=2)
        w.put("k", np.array([-3, 7], np.int16))
        w.put("s", np.array(42, np.int32))
        with pytest.raises(enkfio.EnkfioError) as e:
            w.put("a", a, start=1)
        assert e.value.code == enkfio.ERANGE
    dims, v, g, _ = scipy_read(p)
    assert np.array_equal(v["a"][0], a) and list(v["k"][0]) == [-3, 7] and int(v["s"][0]) == 42
    assert v["a"][1]["note"] == b"two rows then one" and float(g["ratio"]) == 0.25

    bad = str(tmp_path / "not.nc")
    open(bad, "wb").write(b"HDF\x89 not a classic file at all")
    with pytest.raises(enkfio.EnkfioError) as e:
        enkfio.NcReader(bad)
    assert e.value.code == enkfio.EFORMAT
    with pytest.raises(enkfio.EnkfioError) as e:
        enkfio.NcReader(str(tmp_path / "absent.nc"))
    assert e.value.code == enkfio.EIO
    # a header cut short is a format error, values cut short a read error
    raw = open(p, "rb").read()
    open(bad, "wb").write(raw[:40])
    with pytest.raises(enkfio.EnkfioError) as e:
        enkfio.NcReader(bad)
    assert e.value.code == enkfio.EFORMAT
    open(bad, "wb").write(raw[:-8])
    with enkfio.NcReader(bad) as r:
        with pytest.raises(enkfio.EnkfioError) as e:
            r.read("s")
        assert e.value.code == enkfio.EIO


def test_record_files_are_refused(tmp_path):
    p = str(tmp_path / "rec.nc")
    with netcdf_file(p, "w") as f:
        f.createDimension("t", None)
        v = f.createVariable("x", np.float32, ("t",))
        v[:3] = [1, 2, 3]
    with pytest.raises(enkfio.EnkfioError) as e:
        enkfio.NcReader(p)
    assert e.value.code == enkfio.EFORMAT
