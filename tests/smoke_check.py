"""One small moving-boundary FSI step on cuda:0, checked against the oracle."""
from mpm_lbm import workloads


def run_smoke():
    import tempfile

    import torch

    from mpm_lbm import _native
    from tests.parity import make_pair, run_and_compare

    assert torch.cuda.is_available(), "smoke needs cuda:0"
    hip = _native.load_hip()
    with tempfile.TemporaryDirectory() as tmp:
        w = workloads.moving_boundary_flap(n=24, steps=1, substeps=12, flap=(2, 10, 8), mpm_substeps=2, out_dir=tmp)
        d, o = make_pair(w, tmp + "/run", hip, None)
        errs = run_and_compare(d, o, 1)[-1]
    assert errs["solid_equal"], "solid flags differ from the oracle"
    bad = {k: v for k, v in errs.items() if k != "solid_equal" and v > 1.0e-5}
    assert not bad, f"smoke parity failed: {bad}"
    print("smoke ok", {k: v for k, v in errs.items()}, flush=True)
