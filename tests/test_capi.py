"""CPU checks of the C ABI: the HIP library loads and exports every function
``include/grim_hip.h`` declares, the host I/O library every function of
``include/grim_io.h``, and the Python bindings cover them all.  No compute
entry point is called (there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="grim_hip.h"):
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn != header:
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(grim_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    from grim_amd import capi

    if not os.path.exists(capi.LIB_PATH):
        import __graft_entry__

        __graft_entry__.build()
    return ctypes.CDLL(capi.LIB_PATH)


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("grim_batched_newton", "grim_residual", "grim_jacobian", "grim_batch_linear_solve",
                 "grim_conserved", "grim_default_params", "grim_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from grim_amd import capi

    assert set(declared_functions()) == set(capi.SIGNATURES)


def test_headers_are_the_two_libraries():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["grim_hip.h", "grim_io.h"]


def test_io_library_exports_and_binding():
    from grim_amd import checkpoint

    io = checkpoint.lib()
    names = declared_functions("grim_io.h")
    assert names == ["grim_io_dump", "grim_io_dump_vts", "grim_io_last_error", "grim_io_load"]
    assert all(hasattr(io, n) and getattr(io, n).argtypes is not None for n in names)


def test_default_params_match_grim_linear_modes(lib):
    from grim_amd import capi

    p = capi.default_params(emhd=True)
    # src/problem/linear_modes/params.cpp:40-69
    assert (p.conduction, p.viscosity, p.maxNonLinearIter, p.maxLineSearchIters) == (1, 1, 3, 3)
    assert p.nonlinearsolve_atol == 1e-20 and p.JacobianAssembleEpsilon == 4e-8
    assert p.linesearchfloor == 1e-24 and p.rhoFloorInFluidElement == 1e-20
    assert p.nf == 7 and p.dof == 10


def test_struct_layouts_match_oracle():
    """The product and oracle params structs carry the same fields in order."""
    from grim_amd import capi
    from oracle import oracle as O

    assert [f[0] for f in capi.Params._fields_] == [f[0] for f in O.Params._fields_]
