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


def test_no_host_pointer_to_temporaries():
    """Host arrays handed to the C ABI must outlive the call: an address taken
    from a temporary numpy array (``_hptr(np.array(...))``) dangles."""
    import re
    from pathlib import Path
    pkg = Path(L.__file__).resolve().parent
    bad = []
    for f in pkg.glob("*.py"):
        for i, line in enumerate(f.read_text().splitlines(), 1):
            if re.search(r"_hptr\((np\.|_i32\(|_i64\(|\[)", line):
                bad.append(f"{f.name}:{i}")
    assert bad == []
