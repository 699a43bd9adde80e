"""The resident MPM subcycle on the GPU (k_mpm_resident: the MPM substeps of
an FSI step in one launch, particles on chip, one grid barrier per substep)
against n launches of the per-substep kernels (fsi_mpm_substep_coupled).

Bit for bit: a cloud of separated clusters of 64 particles (one particle
workgroup each, in lattice order), whose stencil boxes share no grid node, run
on the library built without fma fusion.  Every node then takes its sum from
one workgroup's tile in a fixed order, so both forms are deterministic and
must agree exactly after every FSI step: particles, Jp, the grid, the
reaction counters, and the lattice they drive.  Any race in the buffer
rotation, the clears or the barriers shows up as a difference.

Within the atomic bar: the default build on the 64^3 flap, where workgroups'
boxes overlap and the float adds run in schedule order."""
import numpy as np
import pytest

from mpm_lbm import _native, workloads
from mpm_lbm.sim.drivers.fsi_config import FSIDriverConfig
from mpm_lbm.sim.drivers.fsi_driver import FSIDriver3D
from tests.parity import driver_fields

pytestmark = pytest.mark.gpu


def _cluster_cloud(n, rho_s, seed, origins):
    """Clusters of 2 x 2 x 2 cells x 8 jittered particles (64, lattice order:
    cell-major, then 4 ox + 2 oy + oz) at the given cell origins."""
    dx = 1.0 / n
    rng = np.random.default_rng(seed)
    sub = np.stack(np.meshgrid([0.25, 0.75], [0.25, 0.75], [0.25, 0.75], indexing="ij"), axis=-1).reshape(-1, 3)
    cell = np.stack(np.meshgrid([0, 1], [0, 1], [0, 1], indexing="ij"), axis=-1).reshape(-1, 3)
    xs, cells = [], []
    for o in origins:
        c = np.asarray(o)[None, :] + cell
        xs.append((c[:, None, :] + sub[None, :, :]).reshape(-1, 3))
        cells.append(np.repeat(c, 8, axis=0))
    base = np.concatenate(xs)
    cell_of = np.concatenate(cells)
    x = ((base + rng.uniform(-0.2, 0.2, size=base.shape)) * dx).astype(np.float32)
    vol0 = np.full(len(x), dx ** 3 / 8.0, dtype=np.float32)
    j0 = min(o[1] for o in origins)
    j1 = max(o[1] for o in origins) + 1
    return {"x": x, "vol0": vol0, "mass": (vol0 * np.float32(rho_s)).astype(np.float32),
            "fixed_base_mask": (cell_of[:, 1] == j0).astype(np.int8),
            "free_tip_proxy_mask": (cell_of[:, 1] == j1).astype(np.int8)}


def _node_owners(x, inv_dx):
    """How many particle workgroups' stencil boxes hold each node."""
    base = np.floor(x.astype(np.float64) * inv_dx - 0.5).astype(np.int64)
    seen = {}
    for b in range(len(x) // 64):
        lo, hi = base[64 * b:64 * b + 64].min(0), base[64 * b:64 * b + 64].max(0) + 2
        for node in np.stack(np.meshgrid(*[np.arange(l, h + 1) for l, h in zip(lo, hi)], indexing="ij"),
                             -1).reshape(-1, 3):
            seen[tuple(node)] = seen.get(tuple(node), 0) + 1
    return max(seen.values())


def _driver(w, out, native, resident, cfg=None):
    d = FSIDriver3D(cfg or w.config, str(out), particle_cloud=w.cloud, native=native, device="cuda",
                    write_geometry_file=False)
    d.initialize()
    d.solid.resident_subcycle = resident
    d.solid.resident_min_substeps = 1
    return d


def _state(d):
    f = driver_fields(d)
    s = d.solid
    for k in ("_grid_v", "_grid_m", "_grid_f", "_Jp"):
        f[k] = getattr(s, k).cpu().numpy()
    f["counters"] = d.mb_coupler._counters_buf.cpu().numpy()
    return f


def test_gpu_mpm_resident_bit_exact_separated_clusters(tmp_path):
    n = 64
    origins = [(30, j, k) for j in (1, 8, 15, 22) for k in (18, 25, 32, 39)]
    w = workloads.moving_boundary_flap(n=n, steps=2, substeps=120, out_dir=str(tmp_path))
    w.cloud = _cluster_cloud(n, w.material[2], 3, origins)
    P = len(w.cloud["x"])
    geo = workloads._write_geometry(str(tmp_path), "clusters", workloads.duct_geometry_dict((n, n, n), P, w.material))
    cfg = FSIDriverConfig(**{**w.config.to_dict(), "n_particles": P, "geometry_config_path": geo})
    assert _node_owners(w.cloud["x"], n) == 1
    exact = _native.load_hip(exact=True)
    a = _driver(w, tmp_path / "launch", exact, False, cfg)
    b = _driver(w, tmp_path / "res", exact, True, cfg)
    print(f"[mpm resident] {P} particles in {P // 64} separated clusters, 64^3, exact build", flush=True)
    for step in range(2):
        a.step_once()
        b.step_once()
        sa, sb = _state(a), _state(b)
        # the clusters stay apart: every node still has one owner
        assert _node_owners(sb["x"], n) == 1
        diff = {k: int(np.count_nonzero(sa[k] != sb[k])) for k in sa
                if not np.array_equal(np.asarray(sa[k]).view(np.uint8), np.asarray(sb[k]).view(np.uint8))}
        print(f"[mpm resident] FSI step {step}: differing fields {diff}", flush=True)
        assert not diff, f"FSI step {step}: {diff}"
    b.solid.check_resident_waits()
    assert b.solid._res_sync is not None and not b.solid._res_declined
    sync = b.solid._res_sync.cpu().numpy()
    assert not np.any(sync[:_native.MPM_RES_TIMEOUT_WORD]), "barrier words are reset at the end of every launch"
    assert a.solid._res_sync is None


def test_gpu_mpm_resident_default_build(tmp_path):
    """The default build takes the resident subcycle on the 64^3 flap (24,576
    particles, overlapping workgroup boxes): no barrier wait timed out, and
    the state stays within the float-add reordering bar of the launch path."""
    w = workloads.moving_boundary_flap(n=64, steps=1, substeps=120, out_dir=str(tmp_path))
    a = _driver(w, tmp_path / "launch", None, False)
    b = _driver(w, tmp_path / "res", None, True)
    a.step_once()
    b.step_once()
    b.solid.check_resident_waits()
    assert b.solid._res_sync is not None and not b.solid._res_declined
    fa, fb = driver_fields(a), driver_fields(b)
    scale_v = float(np.max(np.abs(fa["pv"])))
    errs = {
        "x": float(np.max(np.abs(fa["x"] - fb["x"]))) / float(np.max(np.abs(fa["x"]))),
        "pv": float(np.max(np.abs(fa["pv"] - fb["pv"]))) / scale_v,
        "Fp": float(np.max(np.abs(fa["Fp"] - fb["Fp"]))),
        "C": float(np.max(np.abs(fa["C"] - fb["C"]))) / (scale_v * 64.0),
    }
    print(f"[mpm resident] default build, resident vs launch: {errs}", flush=True)
    assert all(np.all(np.isfinite(fb[k])) for k in ("x", "pv", "Fp", "C"))
    assert errs["x"] < 1e-6 and errs["Fp"] < 1e-5 and errs["pv"] < 1e-3 and errs["C"] < 1e-3
