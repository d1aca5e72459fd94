"""The Julia binding (julia/HDGAccel.jl) calls only entry points that
include/hdg.h declares and libhdg exports, and imports what it names
(Julia is not installed here, so the shim is checked as text)."""
import os
import re

from conftest import REPO, load_package

SHIM = os.path.join(REPO, "hdiscontinuousgalerkin.jl_amd", "julia", "HDGAccel.jl")


def test_shim_ccalls_are_exported():
    H = load_package()
    src = open(SHIM).read()
    names = set(re.findall(r"ccall\(\(:(\w+),\s*libhdg\)", src))
    assert names, "no ccall found"
    header = open(os.path.join(REPO, "include", "hdg.h")).read()
    lib = H.lib()
    for n in sorted(names):
        assert n in H._lib.EXPORTS, n
        assert re.search(r"\b" + n + r"\s*\(", header), n
        assert hasattr(lib, n), n


def test_shim_imports_what_it_uses():
    src = open(SHIM).read()
    if "LinearAlgebra." in src:
        assert re.search(r"^\s*(import|using) LinearAlgebra", src, re.M)
    if "SparseMatrixCSC" in src:
        assert re.search(r"^\s*using .*SparseArrays", src, re.M)
