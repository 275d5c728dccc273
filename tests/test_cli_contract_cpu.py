"""The command-line contract through the C ABI, no GPU: every flag list of tests/golden/cli_contract.json gives the same
status and message as the fixture (the flag grammar, values, checks and error strings of reference src/cli.cpp, which
dl_run keeps). Lists the fixture saw accepted stop at the device ("no HIP device"); where a GPU is visible they are
left out."""
import ctypes as C
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cli_contract.json")


@pytest.fixture(scope="module")
def handle():
    import daylight_amd as d
    L = d.lib()
    m = L.dl_mesh_create(b"a", np.array([0, 0, 0, 1, 0, 0, 1, 1, 0], np.float64).ctypes.data, 3, np.array([3], np.int32).ctypes.data, 1,
                         np.array([0, 1, 2], np.int32).ctypes.data, 3, None)
    assert m
    yield L, m
    L.dl_mesh_free(m)


def gpu_visible():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.parametrize("tool", ["daylight", "views"])
def test_flag_contract(handle, tool):
    L, m = handle
    rows = json.load(open(GOLDEN))[tool]
    skip_device = gpu_visible()
    hs = (C.c_void_p * 1)(m)
    checked = 0
    for row in rows:
        if skip_device and row["message"] == "no HIP device":
            continue
        argv = row["argv"]
        av = (C.c_char_p * max(1, len(argv)))(*[a.encode() for a in argv])
        res = C.c_void_p()
        rc = L.dl_run(0 if tool == "daylight" else 1, len(argv), av, hs, 1, None, 0, None, C.byref(res))
        msg = (L.dl_last_error() or b"").decode()
        if res.value:
            L.dl_result_free(res)
        assert (rc, msg) == (row["status"], row["message"]), argv
        checked += 1
    assert checked >= 50
