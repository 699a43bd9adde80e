"""Particle-binning reference built from the oracle's own stencil bases
(orc_bspline, fsi_oracle.cpp weights()): each particle's key is its stencil
base per axis laid out over the (n + 3)^3 bin box of include/fsi_hip.h
(bases -2 .. n, offset by 2; anything else the overflow bin), start is the
exclusive prefix of the bin counts, and order lists each bin's particles in
increasing index (a stable sort by key)."""
import ctypes

import numpy as np
import torch

from mpm_lbm import _native
from mpm_lbm._device import BinBuffers
from oracle import oracle as orc


def oracle_binning(x, inv_dx, shape):
    """(key, start, order) of particles x (P, 3) on spacing 1/inv_dx over a node grid of `shape`."""
    x = np.ascontiguousarray(x, np.float32)
    P = x.shape[0]
    base = np.zeros((P, 3), np.int64)
    for d in range(3):
        X = np.ascontiguousarray(x[:, d] * np.float32(inv_dx), np.float32)
        b = np.zeros(P, np.int32)
        fx = np.zeros(P, np.float32)
        w = np.zeros((P, 3), np.float32)
        orc.lib().orc_bspline_f32(P, orc._p(X), orc._p(b), orc._p(fx), orc._p(w))
        base[:, d] = b
    nx, ny, nz = shape
    nb = (nx + 3) * (ny + 3) * (nz + 3)
    inside = np.all((base >= -2) & (base <= np.array(shape)), axis=1)
    key = np.where(inside, ((base[:, 0] + 2) * (ny + 3) + (base[:, 1] + 2)) * (nz + 3) + (base[:, 2] + 2), nb)
    counts = np.bincount(key, minlength=nb + 1)
    start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    order = np.argsort(key, kind="stable").astype(np.int32)
    return key.astype(np.int32), start, order


def native_binning(native, x, inv_dx, shape, device):
    """fsi_bin_particles through the C ABI on `native` (x as (P, 3))."""
    P = x.shape[0]
    xs = torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32).T)).to(device)
    b = BinBuffers(P, shape, device)
    stream = torch.cuda.current_stream(device).cuda_stream if torch.device(device).type == "cuda" else None
    native.call("fsi_bin_particles", _native.ptr(xs), int(P), ctypes.c_float(inv_dx), *[int(v) for v in shape],
                ctypes.byref(b.struct), stream)
    if torch.device(device).type == "cuda":
        torch.cuda.synchronize(device)
    return b.key[:P].cpu().numpy(), b.start.cpu().numpy(), b.order[:P].cpu().numpy()


def cloud_for_binning(seed, shape, P):
    """Positions over and around a node grid: many particles per bin,
    positions on the half-integer key boundaries, and a few outside."""
    rng = np.random.default_rng(seed)
    n = np.array(shape, np.float64)
    x = rng.uniform(-0.3, 1.3, (P, 3)) * n / n.max()
    k = rng.integers(0, int(n.min()), (P // 8, 3))
    x[: P // 8] = (k + 0.5) / n.max()  # X - 0.5 exactly an integer
    x[P // 8: P // 4] = (rng.integers(0, 4, (P // 8, 3)) + rng.uniform(0.4, 0.6, (P // 8, 3))) / n.max()  # crowded bins
    return x.astype(np.float32)
