"""C-ABI checks that need no GPU: the library loads, exports every symbol that
include/globe.h declares, reports its ABI version, and fails cleanly when no
device is visible (no CPU fallback exists)."""
import ctypes

import pytest

import globe_amd


def test_library_exports_every_header_symbol():
    L = globe_amd.load_library()
    syms = globe_amd.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert L.globe_abi_version() == 1


def test_no_device_means_error_not_fallback():
    L = globe_amd.load_library()
    if L.globe_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert L.globe_create(0, ctypes.byref(h)) == 3  # GLOBE_E_DEVICE
    with pytest.raises(globe_amd.GlobeError):
        globe_amd.Globe(0)


def test_null_context_is_rejected():
    L = globe_amd.load_library()
    assert L.globe_set_varcomp(None, None) == 1
    assert L.globe_vx(None, None, 0, 1, None, 0, 0, 0) == 1
    assert L.globe_last_error(None) == b"null context"
