"""Launch-shape autotuner of the fused LBM substep (fsi_lbm_substep).

The substep's launch shape changes no result, only the time:
  * the tile: cells per workgroup, a run of 128, 256 or 512 cells of one
    i-plane's (j,k) plane (at 64^3: 2, 4 or 8 lattice rows);
  * the register budget: no cap (a grid that fits the chip in one wave
    generation), or 5 waves per SIMD with the populations loaded together
    with the cell flags;
  * the workgroup order: one contiguous slab of i-planes per XCD, or plain
    dispatch order.
`tune_substep` times every candidate on the lattice it will run - its own
populations, flags and walls, with the substep flags of the run - and keeps
the fastest in the lattice's launch parameters (FsiLbmParams.launch_*).  The
substeps it times overwrite only what a substep writes, which is saved first
and restored after.  Choices are cached per signature (device, lattice,
storage, collision, open boundary, substep flags, wall count) for the process
and, given a cache file, across processes.

No reference counterpart: the reference's Taichi kernels take Taichi's
launch shape.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Optional

import numpy as np

from . import _native

BLOCKS = (128, 256, 512)
MINWS = (1, 5)
ORDERS = (0, 1)  # 0: one slab of i-planes per XCD, 1: dispatch order
_CACHE: dict = {}


def candidates(storage: int):
    """(block, minw, linear) launch shapes to try; 16-bit storage is built for 256-cell tiles only."""
    blocks = BLOCKS if storage == _native.STORE_F32 else (256,)
    return [(b, w, o) for b in blocks for w in MINWS for o in ORDERS]


def signature(lbm, flags: int) -> str:
    import torch

    p = lbm._params
    dev = torch.cuda.get_device_name(lbm.device) if lbm.device.type == "cuda" else "cpu"
    walls = int((lbm._cellinfo < 0).sum().item())  # solid cells (bit 31): the wall work varies with them
    return json.dumps([dev, p.nx, p.ny, p.nz, p.storage, p.collision, p.open_bc, int(flags), walls])


def _apply(lbm, shape):
    p = lbm._params
    p.launch_block, p.launch_minw, p.launch_linear = (int(v) for v in shape)


def _load(path):
    try:
        with open(path, "r", encoding="utf-8") as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune_substep(lbm, flags: int, launches: int = 20, repeats: int = 3, cache_path: Optional[str] = None) -> dict:
    """Time the substep of `lbm` (LBMFluid3D on a GPU) under every candidate
    launch shape with `flags` (the FSI_SUB_* flags of the run's steady
    substep) and keep the fastest.  Returns {"choice", "us" per shape,
    "default_us", "cached"}."""
    import torch

    if lbm.device.type != "cuda":
        raise RuntimeError("the substep autotuner times GPU launches")
    lbm._ensure_ready()
    if flags & _native.SUB_MOVING_WALL:
        lbm._ensure_link_buffers()
    key = signature(lbm, flags)
    if key not in _CACHE and cache_path:
        _CACHE.update(_load(cache_path))
    if key in _CACHE:
        rec = dict(_CACHE[key], cached=True)
        _apply(lbm, rec["choice"])
        return rec
    stream = torch.cuda.current_stream(lbm.device)
    cur = lbm.cur
    parity = 0 if lbm._pending is None else 1 - lbm._pending
    # what a substep writes: the other population buffer, the link buffers,
    # the accumulator and the counters
    saved = [lbm._F[1 - cur].clone(), lbm._counters.clone()]
    if lbm._link is not None:
        saved += [lbm._link[0].clone(), lbm._link[1].clone(), lbm._link_corr.clone(), lbm._accum.clone()]
    targets = [lbm._F[1 - cur], lbm._counters]
    if lbm._link is not None:
        targets += [lbm._link[0], lbm._link[1], lbm._link_corr, lbm._accum]
    old = (lbm._params.launch_block, lbm._params.launch_minw, lbm._params.launch_linear)
    if lbm._cell_force_active:
        flags |= _native.SUB_CELL_FORCE
    args = (ctypes.byref(lbm._params), ctypes.byref(lbm._state), cur, parity, int(flags), lbm._stream())

    def time_shape(shape):
        _apply(lbm, shape)
        for _ in range(2):
            lbm.native.call("fsi_lbm_substep", *args)
        per = []
        for _ in range(repeats):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(launches):
                lbm.native.call("fsi_lbm_substep", *args)
            b.record(stream)
            b.synchronize()
            per.append(a.elapsed_time(b) * 1000.0 / launches)
        return float(np.median(per))

    try:
        us = {",".join(str(v) for v in s): time_shape(s) for s in candidates(lbm._params.storage)}
        default_us = time_shape((0, 0, 0))
    finally:
        torch.cuda.synchronize(lbm.device)
        for t, s in zip(targets, saved):
            t.copy_(s)
        _apply(lbm, old)
    best = min(us, key=us.get)
    choice = [int(v) for v in best.split(",")]
    rec = {"choice": choice, "us": {k: round(v, 3) for k, v in us.items()}, "default_us": round(default_us, 3)}
    # keep the built-in shape unless a candidate beats it clearly (timing noise)
    if us[best] > 0.98 * default_us:
        rec["choice"] = [0, 0, 0]
    _CACHE[key] = rec
    if cache_path:
        data = _load(cache_path)
        data[key] = rec
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        tmp = cache_path + ".tmp"
        with open(tmp, "w", encoding="utf-8") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, cache_path)
    _apply(lbm, rec["choice"])
    return dict(rec, cached=False)


def driver_substep_flags(driver) -> int:
    """The FSI_SUB_* flags of a driver's steady substep (the one its FSI
    steps repeat): moving-wall with the previous substep's gather for a
    subcycled moving-boundary run, the stored-moments collision otherwise."""
    c = driver.config
    if c.coupling_mode == "moving_boundary":
        f = _native.SUB_MOVING_WALL
        if c.fsi_exchange_mode == "lbm_subcycled_per_fsi_step":
            f |= _native.SUB_GATHER_PREV
        return f
    return _native.SUB_STORED_MOMENTS
