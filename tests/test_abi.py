"""The C ABI of libenkfg.so (include/enkfg.h), checked without a GPU: every
declared entry point is exported, the Python binding's structs have the C
layout, argument checks fire before any device call, and the product library
carries none of the oracle's code."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

import pytest

import enkfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "enkfg.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s*(enkfg_\w+)\s*\(", txt, re.M)))


def test_header_declarations_match_binding():
    assert declared() == sorted(enkfg.EXPORTED)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(enkfg.LIB_PATH), "libenkfg.so is not built"
    out = subprocess.run(["nm", "-D", "--defined-only", enkfg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared():
        assert name in syms, name
    L = enkfg.lib()
    for name in declared():
        assert getattr(L, name) is not None


def test_product_library_has_no_oracle_code():
    out = subprocess.run(["nm", "-D", enkfg.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert not re.search(r"\bor_\w+", out)
    assert not re.search(r"\bref_\w+", out)
    ldd = subprocess.run(["ldd", enkfg.LIB_PATH], capture_output=True, text=True).stdout
    assert "enkforacle" not in ldd and "enkfref" not in ldd
    src = open(os.path.join(ROOT, "enkf-c_amd", "enkfg.py")).read()
    assert "import oracle" not in src and "oracle." not in src


def c_layout(work):
    """sizeof/offsetof of the header's structs as the C compiler lays them out."""
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "enkfg.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("enkfg_config %zu\nenkfg_obstype %zu\nenkfg_nodes %zu\nenkfg_stats %zu\nenkfg_grid %zu\nenkfg_field %zu\n",
         sizeof(enkfg_config), sizeof(enkfg_obstype), sizeof(enkfg_nodes), sizeof(enkfg_stats), sizeof(enkfg_grid),
         sizeof(enkfg_field));
  P(enkfg_config, alpha) P(enkfg_config, gamma) P(enkfg_config, flags)
  P(enkfg_obstype, locweight) P(enkfg_obstype, rfactor) P(enkfg_obstype, nlobsmax)
  P(enkfg_nodes, lon) P(enkfg_nodes, visible)
  P(enkfg_stats, ncell) P(enkfg_grid, nlevels) P(enkfg_field, inf_ratio) P(enkfg_field, applylog)
  return 0;
}
"""
    src = os.path.join(work, "layout.c")
    exe = os.path.join(work, "layout")
    with open(src, "w") as f:
        f.write(prog)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_struct_layout_matches_c(tmp_path):
    lay = c_layout(str(tmp_path))
    pairs = {"enkfg_config": enkfg.Config, "enkfg_obstype": enkfg.ObsType, "enkfg_nodes": enkfg.Nodes,
             "enkfg_stats": enkfg.Stats, "enkfg_grid": enkfg.Grid, "enkfg_field": enkfg.Field}
    for cname, py in pairs.items():
        assert C.sizeof(py) == lay[cname], cname
        for key, off in lay.items():
            if key.startswith(cname + "."):
                assert getattr(py, key.split(".")[1]).offset == off, key


def test_argument_errors_before_device_calls():
    L = enkfg.lib()
    h = C.c_void_p()
    assert L.enkfg_create(None, None) == enkfg.ERR_ARG
    bad = [dict(nmem=1), dict(nmem=200), dict(nmem_dynamic=0), dict(mode=7), dict(scheme=9),
           dict(mode=enkfg.MODE_ENKF, nmem_dynamic=10), dict(precision=5)]
    for b in bad:
        cfg = dict(mode=enkfg.MODE_ENKF, scheme=enkfg.SCHEME_DENKF, precision=0, nmem=24, nmem_dynamic=24,
                   alpha=1.0, gamma=0.0, device=0, flags=0)
        cfg.update(b)
        assert L.enkfg_create(C.byref(h), C.byref(enkfg.Config(**cfg))) == enkfg.ERR_ARG, b
        assert not h.value


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(enkfg.EnkfgError) as ei:
        enkfg.Engine(24)
    assert ei.value.code == enkfg.ERR_HIP


def test_engine_noise_generator_matches_oracle():
    """The engine's RANDOMISE stream (host side, callable without a device)
    draws the oracle's get_normalpair values bit for bit and advances the
    state as the reference's drand48 would."""
    import oracle as O
    for state in (0, 77, (1 << 48) - 5):
        got, st = enkfg.rand48_normals(state, 37)
        want, st_o = O.normalpairs(state, 37)
        assert np.array_equal(got, want)
        assert st == st_o
