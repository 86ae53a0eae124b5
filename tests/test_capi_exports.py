"""CPU checks of the C-ABI library: it loads and exports every symbol include/*.h declares
(no compute calls: there is no GPU here)."""
import ctypes as C
import os

import pytest

from tlfea_amd import capi


def test_library_built_and_exports_all_declared_symbols():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libtlfea_hip.so not built (run __graft_entry__.build())")
    L = C.CDLL(capi.LIB_PATH)
    declared = capi.declared_symbols()
    assert len(declared) > 30
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_table_matches_header():
    declared = set(capi.declared_symbols())
    bound = set(capi._SIGS)
    assert bound <= declared, f"bound but not declared: {sorted(bound - declared)}"


def test_host_only_calls():
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libtlfea_hip.so not built")
    L = capi.lib()
    assert b"gfx950" in L.tlfea_version()
    assert capi.padded_elems(1) == 64 and capi.padded_elems(64) == 64 and capi.padded_elems(65) == 128
    q = capi.quadrature(10)
    assert q.n == 5 and abs(sum(q.w) - 1.0 / 6.0) < 1e-16
    q4 = capi.quadrature(4)
    assert q4.n == 1 and q4.w[0] == 1.0 / 6.0
    with pytest.raises(capi.TlfeaError):
        capi.quadrature(8)
