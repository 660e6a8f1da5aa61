"""C ABI and host-side logic checks that need no GPU.

The library must load and export every entry point include/gaga_amd.h
declares; option defaults and argument validation follow the reference's
driver (GAGA/gaga/gaga_gen.cu:176-197, gaga_cs.m).  No device work is issued.
"""
import ctypes
import os
import re
import warnings

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gaga_amd.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xcs_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(lib):
    from gaga_amd import _capi
    L = lib.load()
    declared = _declared()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_capi.EXPORTS) == declared


def test_version_and_last_error(lib):
    L = lib.load()
    assert b"gfx950" in L.xcs_version()
    o = lib._capi.Options()
    rc = L.xcs_options_default(b"NOPE", b"gen", 100, ctypes.byref(o))
    assert rc == -1 and b"unknown algorithm" in L.xcs_last_error()


@pytest.mark.parametrize("alg,ens,n,maxiter,num_bins,mat", [
    ("NIHT", "gen", 1024, 5000, 1000, 1),
    ("HTP", "gen", 2 ** 18, 300, 13107, 1),
    ("CSMPSP", "smv", 2 ** 18, 300, 13107, 3),
    ("CGIHT", "dct", 2 ** 20, 5000, 52428, 1),
])
def test_option_defaults(alg, ens, n, maxiter, num_bins, mat, lib):
    o = lib._capi.default_options(alg, ens, n)
    assert o.maxiter == maxiter and o.num_bins == num_bins and o.matrix_ensemble == mat
    assert o.vec_distribution == 1 and o.restart == 0 and o.mode == 0
    assert o.tol == np.float32(1e-4) and o.tol_reference_default == 0


def test_options_mirror_gagaOptions(lib):
    import gaga_amd as ga
    opts = ga.options("tol", 1e-6, "maxiter", 40, restartFlag="on", vecDistribution="gaussian",
                      matrixEnsemble="binary", numBins=2000)
    o = ga._apply_options("CGIHT", "smv", 4096, opts)
    assert o.tol == np.float32(1e-6) and o.maxiter == 40 and o.restart == 1
    assert o.vec_distribution == 2 and o.matrix_ensemble == 2 and o.num_bins == 2000
    with pytest.raises(ValueError):
        ga._apply_options("CGIHT", "gen", 64, ga.options(matrixEnsemble="ones"))
    with pytest.raises(ValueError):
        ga.options("tol")
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        ga._apply_options("NIHT", "gen", 64, [("bogusOption", 1)])
        assert any("not recognised" in str(x.message) for x in w)


def test_cs_argument_validation(lib):
    import gaga_amd as ga
    with pytest.raises(ValueError, match="ens must be either dct, gen, or smv"):
        ga.cs("NIHT", "foo", 10, 100, 200)
    with pytest.raises(ValueError, match="algorithms"):
        ga.cs("IHTX", "gen", 10, 100, 200)
    with pytest.raises(ValueError, match="six or seven"):
        ga.cs("NIHT", "smv", 10, 100, 200)
    with pytest.raises(ValueError, match="five or six"):
        ga.cs("NIHT", "gen", 10, 100)
