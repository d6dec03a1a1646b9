"""Helpers for GPU tests: device planes as torch tensors handed to the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from quasar_amd import _lib


def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def qgrid(nx, ny, ng, lx=1.0, ly=1.0, ox=0.0, oy=0.0):
    return _lib.QGrid(nx, ny, ng, 0, lx, ly, ox, oy, 0.0, 0.0)


def to_dev(arrs):
    torch = torch_cuda()
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda() for a in arrs]


def zeros_dev(n, count):
    torch = torch_cuda()
    return [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(count)]


def ptrs(tensors):
    return (C.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def to_host(tensors):
    torch = torch_cuda()
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in tensors]


def flags(f):
    return (C.c_int32 * 4)(*f)


def scratch(g):
    torch = torch_cuda()
    L = _lib.load()
    return torch.zeros(L.qmhd_scratch_doubles(C.byref(g)), dtype=torch.float64, device="cuda")


def call(fn, *args):
    torch = torch_cuda()
    torch.cuda.synchronize()
    _lib.check(fn(*args))
    torch.cuda.synchronize()


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    denom = np.abs(b).sum()
    diff = np.abs(a - b).sum()
    if denom == 0:
        return diff
    return diff / denom


GROUPS = {0: (0,), 1: (1, 2, 3), 2: (1, 2, 3), 3: (1, 2, 3), 4: (4,), 5: (5, 6, 7), 6: (5, 6, 7),
          7: (5, 6, 7)}


def rel_l1_state(got, ref, c, offset=0):
    """Relative L1 of component c.  Momentum and magnetic components are
    normalised by the L1 norm of their vector group, so an identically zero
    component (mz, bz in 2-D problems), which both implementations carry only as
    round-off residue, is measured against the physical scale of its vector."""
    diff = np.abs(np.asarray(got[offset + c]) - np.asarray(ref[offset + c])).sum()
    denom = max(np.abs(np.asarray(ref[offset + k])).sum() for k in GROUPS[c])
    return diff / denom if denom > 0 else diff


def interior(a, nx, ny, ng):
    return np.asarray(a).reshape(ny + 2 * ng, nx + 2 * ng)[ng:ng + ny, ng:ng + nx]
