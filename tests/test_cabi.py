"""CPU checks of the C ABI boundary: the HIP library builds, loads, exports
every entry point declared in include/fenics_hip.h, and fails loudly (no CPU
fallback) when no GPU is visible."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fenics_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(fh_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("fh_solve_csc_cg", "fh_solve_csr_cg", "fh_solve_bsr3_pcg", "fh_assemble_stiffness",
              "fh_assemble_elasticity", "fh_apply_dirichlet", "fh_solve", "fh_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import fenics_hip
    L = C.CDLL(fenics_hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_no_device_fails_loudly():
    import fenics_hip as fh
    n = C.c_int(-1)
    assert fh.lib().fh_device_count(C.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(fh.FenicsError) as e:
        fh.unit_square(4, 4)
    assert e.value.code == fh.FH_ERR_NO_DEVICE
    assert fh.discover_gpus() == [] and fh.best_gpu_index() is None


def test_oracle_is_not_linked_into_the_product():
    import fenics_hip
    with open(fenics_hip.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"liboracle" not in blob and b"or_solve_cg" not in blob


def test_null_arguments_are_rejected_not_dereferenced():
    """Every entry point answers null pointers and zero sizes with a status and
    a message (the reference's Err), never a crash; destroy calls accept null.
    The checks run before any device work, so this needs no GPU."""
    import fenics_hip as fh
    L = fh.lib()
    for name in declared_symbols():
        if name == "fh_last_error":
            continue
        f = getattr(L, name)
        args = []
        for t in f.argtypes or []:
            if isinstance(t, type) and issubclass(t, C._CFuncPtr):
                args.append(C.cast(None, t))
            elif t in (C.c_double, C.c_float):
                args.append(0.0)
            elif t in (C.c_int, C.c_uint64, C.c_uint32, C.c_size_t, C.c_int64):
                args.append(0)
            else:
                args.append(None)
        rc = f(*args)
        if name.endswith("_destroy"):
            assert rc == fh.FH_OK, name
        elif name == "fh_synchronize":
            assert rc in (fh.FH_OK, fh.FH_ERR_NO_DEVICE), name
        else:
            assert rc != fh.FH_OK and fh.last_error(), name


def test_rust_binding_stub_declares_the_header():
    """The Rust extern block of the integration guide (the binding a
    maintainer copies) names exactly the entry points the header declares."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = text[text.index('#[link(name = "fenics_hip")]'):]
    block = block[:block.index("\n}\n")]
    rust = sorted(set(re.findall(r"pub fn (fh_\w+)\(", block)))
    assert rust == declared_symbols()
