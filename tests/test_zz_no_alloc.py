"""The sampling entry point that allocates nothing
(ticit_hip_sample_device_postselected_noalloc; gpu-marked).

It stages a postselected launch in a file-scope device array instead of a
buffer the program allocates. Bars: on two calls in a row, the second with
fewer shots (so the array holds the first call's contents), its survivor
count, survivors' rows, expectation values and observable counts equal
ticit_hip_sample_device_postselected's bit for bit, on the lane and group
tiers and for a program without postselection; a call whose staging
exceeds the array, and a wide-tier program, are refused before any launch.
"""
import ctypes

import numpy as np
import pytest

from ticit_amd import _capi
from ticit_amd import workloads as w

pytestmark = pytest.mark.gpu

NOISY = ("H 0\nT 0\nH 0\nX_ERROR(0.2) 1\nM 0 1\nDETECTOR rec[-1]\nH 2\nT 2\nM 2\n"
         "DETECTOR rec[-1] rec[-2]\nDISCARD rec[-2]\nOBSERVABLE_INCLUDE(0) rec[-3]\n"
         "EXP_VAL Z0\n")
SURFACE = w.surface_code_t(3, 3, 2e-2, 1, 0)
STATIC_SCRATCH_BYTES = 32 << 20


def buffers(s, n, fill):
    import torch
    row = lambda c: max(1, (c + 7) // 8)
    return (torch.full((n, row(s.num_measurements)), fill, dtype=torch.uint8, device="cuda"),
            torch.full((n, row(s.num_detectors)), fill, dtype=torch.uint8, device="cuda"),
            torch.full((n, row(s.num_observables)), fill, dtype=torch.uint8, device="cuda"),
            torch.zeros(max(1, s.num_observables), dtype=torch.int64, device="cuda"),
            torch.full((n, max(1, s.num_expectation_values)), 7.0, dtype=torch.float64, device="cuda"),
            torch.full((1,), -1, dtype=torch.int64, device="cuda"))


def call(fn, s, shots, seed, b, first_shot=0):
    p = lambda t: ctypes.c_void_p(int(t.data_ptr()))
    ev = p(b[4]) if s.num_expectation_values else None
    return fn(s.handle, shots, seed, first_shot, 1, 1, p(b[0]), p(b[1]), p(b[2]), p(b[3]), ev, p(b[5]), None)


def same(s, shots, seed, first_shot=0):
    import torch
    lib = _capi.load()
    a = buffers(s, shots, 0xAB)
    z = buffers(s, shots, 0x5C)
    print("noalloc", s.tier, "shots", shots, "seed", seed, flush=True)
    _capi.check(call(lib.ticit_hip_sample_device_postselected, s, shots, seed, a, first_shot))
    _capi.check(call(lib.ticit_hip_sample_device_postselected_noalloc, s, shots, seed, z, first_shot))
    torch.cuda.synchronize()
    acc = int(a[5].item())
    assert int(z[5].item()) == acc
    for x, y in zip(a[:3], z[:3]):
        assert torch.equal(x[:acc], y[:acc])
    if s.num_expectation_values:
        assert torch.equal(a[4][:acc], z[4][:acc])
    assert torch.equal(a[3], z[3])
    return acc


@pytest.mark.parametrize("name,mask,tier", [("noisy", [0, 1], "lane"), ("surface", [1] * 12, "group"),
                                            ("surface_group64", [1] * 12, "group64")])
def test_noalloc_matches_the_allocating_entry_point(hip, name, mask, tier):
    text = NOISY if name == "noisy" else SURFACE
    kw = {"group_threads": 64} if tier == "group64" else {}
    s = hip.Circuit(text).compile(postselection_mask=mask, tier=tier[:5], **kw)
    assert s.has_postselection
    first = same(s, 50000, 11)
    second = same(s, 3001, 12, first_shot=77)
    assert 0 < second < 3001 and 0 < first < 50000


def test_noalloc_without_postselection(hip):
    s = hip.Circuit(SURFACE).compile()
    assert not s.has_postselection
    assert same(s, 4000, 3) == 4000
    assert same(s, 1000, 4) == 1000


def test_noalloc_refuses_oversized_calls_and_the_wide_tier(hip):
    import torch
    lib = _capi.load()
    s = hip.Circuit(SURFACE).compile(postselection_mask=[1] * 12)
    b = buffers(s, 1, 0)
    too_many = STATIC_SCRATCH_BYTES + 1  # one keep byte per shot alone exceeds the array
    p = lambda t: ctypes.c_void_p(int(t.data_ptr()))
    rc = lib.ticit_hip_sample_device_postselected_noalloc(s.handle, too_many, 1, 0, 1, 0, None, None, None,
                                                          p(b[3]), None, p(b[5]), None)
    assert _capi.STATUS_NAMES.get(rc) == "InvalidInput"
    assert b"static scratch" in lib.ticit_hip_last_error()
    wide = hip.Circuit(SURFACE).compile(postselection_mask=[1] * 12, tier="wide")
    rc = lib.ticit_hip_sample_device_postselected_noalloc(wide.handle, 10, 1, 0, 1, 0, None, None, None,
                                                          p(b[3]), None, p(b[5]), None)
    assert _capi.STATUS_NAMES.get(rc) == "Unsupported"
    torch.cuda.synchronize()
    assert int(b[5].item()) == -1  # nothing ran
