"""Per-cell bounce-back link classification against the oracle's.

The oracle classifies a fluid cell's link in direction s as bounce-back when
its streaming target (the periodic neighbour c + e_s) is solid
(orc_link_masks, the helper its streaming passes use; fluid.py:432-440,
:489-519).  The product keeps that classification per cell in `cellinfo`
(bit 31 solid; bits 1..18 links of a fluid cell, incoming links of a solid
cell; bit 30 a link to, or for a solid cell itself, a nonzero wall velocity):
every bit is compared exactly."""
import numpy as np

from oracle import oracle as orc

CI_SOLID = np.uint32(1 << 31)
CI_MOVING = np.uint32(1 << 30)
CI_MASK = np.uint32(0x7FFFE)


def oracle_link_masks(solid):
    s = np.ascontiguousarray(np.asarray(solid, np.int8))
    out = np.zeros(s.shape, np.uint32)
    orc.lib().orc_link_masks(*[int(v) for v in s.shape], orc._p(s), orc._p(out))
    return out


def expected_cellinfo(solid, solid_vel):
    """cellinfo that the oracle's classification implies for flags `solid`
    and wall velocities `solid_vel` (n,n,n,3)."""
    solid = np.asarray(solid) != 0
    mask = oracle_link_masks(solid.astype(np.int8))
    moving_cell = np.any(np.asarray(solid_vel) != 0, axis=-1)
    incoming = np.zeros(solid.shape, np.uint32)
    to_moving = np.zeros(solid.shape, bool)
    for s in range(1, 19):
        e = orc.E[s]
        bit = (mask >> np.uint32(s)) & np.uint32(1)
        # the fluid cell x = y - e_s streams s into y: y's incoming bit s
        incoming |= np.roll(bit, shift=tuple(int(v) for v in e), axis=(0, 1, 2)) << np.uint32(s)
        # does fluid cell x link to a moving wall in direction s?
        tgt_moving = np.roll(moving_cell, shift=tuple(int(-v) for v in e), axis=(0, 1, 2))
        to_moving |= (bit != 0) & tgt_moving
    ci = np.where(solid, CI_SOLID | (incoming & CI_MASK), mask & CI_MASK).astype(np.uint32)
    ci |= np.where(solid & moving_cell, CI_MOVING, np.uint32(0))
    ci |= np.where(~solid & to_moving, CI_MOVING, np.uint32(0))
    return ci


def cellinfo_problems(d, solid_oracle):
    """Differences between the driver's cellinfo and the oracle's links."""
    lbm = d.lbm
    got = lbm._cellinfo.cpu().numpy().view(np.uint32).reshape(lbm.nx, lbm.ny, lbm.nz)
    want = expected_cellinfo(solid_oracle, lbm.solid_vel.to_numpy())
    bad = got != want
    if not bad.any():
        return []
    idx = np.argwhere(bad)[:5]
    return [f"{int(bad.sum())} cells differ, e.g. {tuple(int(v) for v in c)}: {got[tuple(c)]:#010x} vs "
            f"{want[tuple(c)]:#010x}" for c in idx[:1]]
