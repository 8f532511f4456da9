"""Host-side argument checks of the torch wrappers: a tensor handed to the
C-ABI as a raw pointer must be contiguous float32 (int64/int32 for CSR
arrays) of the expected shape on the context's device, else LightNEError is
raised before anything is launched (CPU only: the checks need no device)."""
import numpy as np
import pytest
import torch

import lightne_amd as L
from lightne_amd import api


def test_need_rejects_wrong_dtype_layout_shape_device():
    x = torch.zeros(8, 16)
    with pytest.raises(L.LightNEError, match="cuda"):
        api._need(x, "X", torch.float32, (8, 16))  # host tensor
    x64 = torch.zeros(8, 16, dtype=torch.float64, device="meta")
    with pytest.raises(L.LightNEError):
        api._need(x64, "X", torch.float32, (8, 16))
    with pytest.raises(L.LightNEError, match="expected a torch tensor"):
        api._need(np.zeros((8, 16), np.float32), "X", torch.float32)


def test_out_ptr_host_arrays():
    good = np.zeros((5, 3), np.float32)
    assert api._out_ptr(good, (5, 3), "out", None, True) == good.ctypes.data
    with pytest.raises(L.LightNEError, match="device tensor"):
        api._out_ptr(good, (5, 3), "out", None, False)
    for bad in [np.zeros((5, 3), np.float64), np.zeros((3, 5), np.float32).T, np.zeros((5, 4), np.float32)]:
        with pytest.raises(L.LightNEError):
            api._out_ptr(bad, (5, 3), "out", None, True)
    ro = np.zeros((5, 3), np.float32)
    ro.flags.writeable = False
    with pytest.raises(L.LightNEError):
        api._out_ptr(ro, (5, 3), "out", None, True)


def test_device_graph_desc_checks_arrays():
    g = api.DeviceGraph(torch.zeros(4, dtype=torch.int64), torch.zeros(6, dtype=torch.int32), 3, 6)
    with pytest.raises(L.LightNEError, match="graph offsets"):
        g.desc()  # host tensors
