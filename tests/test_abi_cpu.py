"""CPU checks of the native boundary: the shared library loads without a GPU
and exports every entry point include/frank_amd.h declares."""
import ctypes

from frank_amd import _lib as L


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.lib()
    declared = L.header_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []


def test_every_declared_symbol_has_a_binding_signature():
    assert sorted(L._SIGS) == L.header_symbols()


def test_version_and_error_string_without_gpu():
    lib = L.lib()
    assert lib.frk_version() >= 1
    # an argument error is reported, not an abort
    rc = lib.frk_ctx_synchronize(None)
    assert rc == L.FRK_ERR_ARG
    assert b"null" in lib.frk_last_error()
