"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

import pytest

from ticit_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "ticit_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ticit_hip_[a-z_]+)\s*\(", text)))


def test_header_declares_what_the_binding_knows():
    assert declared_symbols() == sorted(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.library_path())
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_create_without_device_fails_with_device_error():
    import ticit_amd
    if ticit_amd.device_properties()["devices"]:
        pytest.skip("a device is visible")
    with pytest.raises(ticit_amd.TicitError) as e:
        ticit_amd.Circuit("H 0\nM 0\n").compile()
    assert e.value.kind in ("DeviceError", "OutOfDeviceMemory")


def test_non_hip_backend_is_rejected():
    import ticit_amd
    with pytest.raises(ValueError):
        ticit_amd.Circuit("M 0\n").compile(backend="cpu")
