"""Device-buffer plumbing shared by the drop-in solver classes.

Buffers are torch tensors (ROCm device memory on the GPU); kernels receive
their raw addresses through the C ABI.  `FieldView` gives each buffer the
reference's Taichi-field surface: `to_numpy()` / `from_numpy()` in the
reference shapes ((nx,ny,nz[,k]) for lattice fields, (P[,3[,3]]) for
particles) while the device layout stays SoA.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise _native.FsiError("no ROCm GPU visible: the FSI solver runs only on the GPU (libfsi_hip.so)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle(device: torch.device):
    """hipStream_t of torch's current stream on `device` (None for host buffers)."""
    if device.type != "cuda":
        return None
    return torch.cuda.current_stream(device).cuda_stream


class FieldView:
    """to_numpy / from_numpy over a SoA tensor of shape (comps, *space) or (*space)."""

    def __init__(self, get_tensor, space_shape, comps=None, before_read=None, after_write=None, np_dtype=np.float32,
                 before_write=None):
        self._get = get_tensor
        self.space_shape = tuple(space_shape)
        self.comps = comps  # None (scalar), int, or (3, 3)
        self._before_read = before_read
        self._before_write = before_write
        self._after_write = after_write
        self.np_dtype = np_dtype

    @property
    def shape(self):
        if self.comps is None:
            return self.space_shape
        extra = self.comps if isinstance(self.comps, tuple) else (self.comps,)
        return self.space_shape + extra

    def tensor(self):
        return self._get()

    def to_numpy(self):
        if self._before_read is not None:
            self._before_read()
        t = self._get()
        if self.comps is None:
            out = t.reshape(self.space_shape)
        else:
            k = int(np.prod(self.comps)) if isinstance(self.comps, tuple) else self.comps
            out = t.reshape((k,) + self.space_shape)
            out = out.permute(*range(1, len(self.space_shape) + 1), 0)
            if isinstance(self.comps, tuple):
                out = out.reshape(self.space_shape + self.comps)
        # a fresh array, as Taichi's to_numpy gives: host tensors would otherwise share their memory
        return np.array(out.contiguous().cpu().numpy(), dtype=self.np_dtype, copy=True)

    def from_numpy(self, arr):
        a = np.asarray(arr)
        if a.shape != self.shape:
            raise ValueError(f"expected array of shape {self.shape}, got {a.shape}")
        if self._before_write is not None:
            self._before_write()
        t = self._get()
        src = torch.from_numpy(np.ascontiguousarray(a.astype(self.np_dtype, copy=False)))
        if self.comps is not None:
            k = int(np.prod(self.comps)) if isinstance(self.comps, tuple) else self.comps
            src = src.reshape(self.space_shape + (k,)).permute(len(self.space_shape), *range(len(self.space_shape)))
        t.copy_(src.reshape(t.shape).to(t.device))
        if self._after_write is not None:
            self._after_write()

    def fill(self, value):
        if self._before_write is not None:
            self._before_write()
        self._get().fill_(value)
        if self._after_write is not None:
            self._after_write()


class ScalarView:
    """`field[None]` access for reference code that reads scalar fields."""

    def __init__(self, getter):
        self._getter = getter

    def __getitem__(self, key):
        if key is not None:
            raise KeyError("scalar fields are indexed with None")
        return self._getter()


def read_struct(buf: torch.Tensor, cls):
    """Copy a device byte buffer holding one ctypes struct to the host."""
    host = buf.detach().cpu().numpy().tobytes()
    return cls.from_buffer_copy(host)


def struct_buffer(cls, device, count: int = 1) -> torch.Tensor:
    import ctypes

    return torch.zeros(ctypes.sizeof(cls) * count, dtype=torch.uint8, device=device)


def bits_to_float(u: int) -> float:
    return float(np.array([u], dtype=np.uint32).view(np.float32)[0])


class BinBuffers:
    """Device buffers of one particle binning (the FsiBins struct) for
    `n_particles` particles over a node grid of `shape`."""

    def __init__(self, n_particles: int, shape, device):
        from . import _native

        nb = _native.bin_count(*shape)
        i32 = dict(dtype=torch.int32, device=device)
        self.shape = tuple(int(v) for v in shape)
        self.key = torch.zeros(max(1, int(n_particles)), **i32)
        self.start = torch.zeros(nb + 2, **i32)
        self.order = torch.zeros(max(1, int(n_particles)), **i32)
        self.work = torch.zeros(_native.bin_work_ints(nb), **i32)
        self.struct = _native.Bins(self.key.data_ptr(), self.start.data_ptr(), self.order.data_ptr(),
                                   self.work.data_ptr())

