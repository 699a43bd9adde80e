"""Closed-form pins of the oracle's regularised collision and quadratic
B-spline weights, the same checks applied to the kernels' own bodies (the
emulation build's test-only entries), and proof that every check catches a
planted mistake: oracle variants compiled with -DORACLE_MUTANT=n.

The regularised collision (Latt & Chopard 2006, with Guo's BGK forcing) has no
reference counterpart; the reference only has a regularised *boundary*
(src/mpm_lbm/sim/lbm/config.py:15-26).  What fixes it is the D3Q19 moment
algebra: with Pi_neq = sum_s e_s e_s (F_s - feq_s) and the lattice's isotropy
sum_s w_s e_a e_b e_c e_d = (d_ab d_cd + d_ac d_bd + d_ad d_bc) / 9,
  - an equilibrium input is a fixed point;
  - mass and momentum are conserved without forcing;
  - a pure second-order Hermite input feq + w (9/2)(e e - I/3) : A comes back
    as feq + (1 - s_v) w (9/2)(e e - I/3) : A, i.e. Pi_neq = A is relaxed by
    exactly (1 - 1/tau);
  - every component orthogonal to {1, e, e e} is removed;
  - the Guo term adds no mass, momentum (1 - s_v/2) G and second moment
    (1 - s_v/2)(u G + G u).
The B-spline weights (src/mpm_lbm/sim/mpm/solid.py:222-230) are fixed by the
quadratic B-spline: base = trunc(X - 0.5), and the three weights sum to one
with zero first moment and second moment 1/4 about the particle."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = Path(__file__).resolve().parents[1]
E = orc.E.astype(np.float64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
# the ten second-order-and-below monomials 1, e_a, e_a e_b (a <= b) per direction
_LOW = np.stack([np.ones(19)] + [E[:, a] for a in range(3)] +
                [E[:, a] * E[:, b] for a in range(3) for b in range(a, 3)], axis=1)


def feq(rho, u):
    eu = u @ E.T
    return W * rho[:, None] * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (u * u).sum(1)[:, None])


def hermite2(A):
    """w_s (9/2) (e_s e_s - I/3) : A for each symmetric A (n,3,3)."""
    ee = np.einsum("sa,sb,nab->ns", E, E, A)
    return W * 4.5 * (ee - np.trace(A, axis1=1, axis2=2)[:, None] / 3)


def high_order(rng, n):
    """Random populations with zero sum, first and second moments."""
    h = rng.standard_normal((n, 19))
    coef, *_ = np.linalg.lstsq(_LOW, h.T, rcond=None)
    return (h.T - _LOW @ coef).T * 1e-3


def _sym(rng, n, scale):
    A = rng.standard_normal((n, 3, 3)) * scale
    return A + A.transpose(0, 2, 1)


# --- collision callables -----------------------------------------------------

def oracle_collide(lib, T=np.float64):
    ct = ctypes.c_double if T == np.float64 else ctypes.c_float
    fn = getattr(lib, "orc_colission_regularized_" + ("f64" if T == np.float64 else "f32"))
    LbmArgs = orc._structs(ct)[0]

    def collide(F, rho, u, G, s_v):
        n = F.shape[0]
        a = LbmArgs()
        a.nx, a.ny, a.nz = n, 1, 1
        a.s_v, a.s_other, a.rho0 = s_v, 1.0, 1.0
        Fc, r, v, g = (np.ascontiguousarray(x, dtype=T) for x in (F, rho, u, G))
        f = np.zeros_like(Fc)
        solid = np.zeros(n, np.int8)
        fn(ctypes.byref(a), orc._p(Fc), orc._p(f), orc._p(r), orc._p(v), orc._p(solid), orc._p(g))
        return f.astype(np.float64)
    return collide


def kernel_collide(emu_lib):
    def collide(F, rho, u, G, s_v):
        n = F.shape[0]
        Fc, r, v, g = (np.ascontiguousarray(x, dtype=np.float32) for x in (F, rho, u, G))
        f = np.zeros_like(Fc)
        rc = emu_lib.fsi_emu_regularized_collide(n, orc._p(Fc), orc._p(r), orc._p(v), orc._p(g),
                                                 ctypes.c_float(s_v), orc._p(f))
        assert rc == 0
        return f.astype(np.float64)
    return collide


# --- the checks (each returns the largest violation found) ----------------------

def _inputs(seed, n=24):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.05 * rng.standard_normal(n)
    u = 0.04 * rng.standard_normal((n, 3))
    return rng, rho, u


def check_equilibrium_fixed_point(collide, s_v=1 / 0.55):
    _, rho, u = _inputs(1)
    F = feq(rho, u)
    return np.abs(collide(F, rho, u, np.zeros_like(u), s_v) - F).max()


def check_conservation(collide, s_v=1 / 0.55):
    rng, rho, u = _inputs(2)
    F = feq(rho, u) + hermite2(_sym(rng, len(rho), 2e-3)) + high_order(rng, len(rho))
    rho_in = F.sum(1)
    mom_in = F @ E
    f = collide(F, rho_in, mom_in / rho_in[:, None], np.zeros_like(u), s_v)
    return max(np.abs(f.sum(1) - rho_in).max(), np.abs(f @ E - mom_in).max())


def check_second_order_relaxation(collide, s_v=1 / 0.55):
    rng, rho, u = _inputs(3)
    H2 = hermite2(_sym(rng, len(rho), 3e-3))
    F = feq(rho, u) + H2
    return np.abs(collide(F, rho, u, np.zeros_like(u), s_v) - (feq(rho, u) + (1 - s_v) * H2)).max()


def check_high_order_removed(collide, s_v=1 / 0.55):
    rng, rho, u = _inputs(4)
    H2 = hermite2(_sym(rng, len(rho), 3e-3))
    F = feq(rho, u) + H2 + high_order(rng, len(rho))
    return np.abs(collide(F, rho, u, np.zeros_like(u), s_v) - (feq(rho, u) + (1 - s_v) * H2)).max()


def check_guo_moments(collide, s_v=1 / 0.55):
    rng, rho, u = _inputs(5)
    F = feq(rho, u) + hermite2(_sym(rng, len(rho), 1e-3))
    G = 1e-3 * rng.standard_normal(u.shape)
    dg = collide(F, rho, u, G, s_v) - collide(F, rho, u, np.zeros_like(G), s_v)
    k = 1 - 0.5 * s_v
    second = np.einsum("ns,sa,sb->nab", dg, E, E)
    want2 = k * (u[:, :, None] * G[:, None, :] + G[:, :, None] * u[:, None, :])
    return max(np.abs(dg.sum(1)).max(), np.abs(dg @ E - k * G).max(), np.abs(second - want2).max())


COLLISION_CHECKS = (check_equilibrium_fixed_point, check_conservation, check_second_order_relaxation,
                    check_high_order_removed, check_guo_moments)

BSPLINE_X = np.concatenate([np.linspace(0.3, 30.7, 203), np.arange(1, 31) + 0.5, np.arange(1, 31),
                            np.arange(1, 31) + 0.5 - 1e-6, np.arange(1, 31) + 0.5 + 1e-6])


def check_bspline(bs, X, tol):
    """Largest violation of base = trunc(X - 0.5), fx = X - base, partition of
    unity, zero first moment and second moment 1/4 of the weights."""
    base, fx, w = bs(X)
    off = np.arange(3)[None, :] - fx[:, None]
    bad_base = np.any(base != np.trunc(X - 0.5).astype(int))
    return max(float(bad_base), np.abs(fx - (np.asarray(X, np.float64) - base)).max() - tol,
               np.abs(w.sum(1) - 1).max(), np.abs((w * off).sum(1)).max(), np.abs((w * off ** 2).sum(1) - 0.25).max())


def oracle_bspline(lib, T=np.float64):
    fn = getattr(lib, "orc_bspline_" + ("f64" if T == np.float64 else "f32"))

    def bs(X):
        X = np.ascontiguousarray(X, dtype=T)
        base = np.zeros(len(X), np.int32)
        fx = np.zeros(len(X), T)
        w = np.zeros((len(X), 3), T)
        fn(len(X), orc._p(X), orc._p(base), orc._p(fx), orc._p(w))
        return base, fx.astype(np.float64), w.astype(np.float64)
    return bs


def kernel_bspline(emu_lib):
    def bs(X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        base = np.zeros(len(X), np.int32)
        fx = np.zeros(len(X), np.float32)
        w = np.zeros((len(X), 3), np.float32)
        assert emu_lib.fsi_emu_bspline(len(X), orc._p(X), orc._p(base), orc._p(fx), orc._p(w)) == 0
        return base, fx, w
    return bs


# --- the tests ------------------------------------------------------------------

@pytest.mark.parametrize("check", COLLISION_CHECKS, ids=lambda c: c.__name__)
def test_oracle_regularized_collision_fp64(check):
    assert check(oracle_collide(orc.lib())) < 1e-14


@pytest.mark.parametrize("check", COLLISION_CHECKS, ids=lambda c: c.__name__)
def test_oracle_regularized_collision_fp32(check):
    assert check(oracle_collide(orc.lib(), np.float32)) < 2e-7


@pytest.mark.parametrize("check", COLLISION_CHECKS, ids=lambda c: c.__name__)
def test_kernel_regularized_collision(emu, check):
    """The same closed forms on the kernels' regularized_collide (fsi_core.h)."""
    assert check(kernel_collide(emu.lib)) < 2e-7


def test_oracle_bspline_weights():
    assert check_bspline(oracle_bspline(orc.lib()), BSPLINE_X, 1e-15) < 1e-14
    assert check_bspline(oracle_bspline(orc.lib(), np.float32), BSPLINE_X.astype(np.float32), 0.0) < 1e-6


def test_kernel_bspline_weights_and_binning_keys_match_oracle(emu):
    """The kernels' bspline() against the FP32 oracle: base (the particle's
    binning key along the axis), fx and the weights bit for bit, on particles
    across every cell including the half-integer key boundaries."""
    rng = np.random.default_rng(7)
    X = np.concatenate([BSPLINE_X, rng.uniform(0.2, 255.8, 20000)]).astype(np.float32)
    kb, kf, kw = kernel_bspline(emu.lib)(X)
    ob, of, ow = oracle_bspline(orc.lib(), np.float32)(X)
    assert np.array_equal(kb, ob)
    assert np.array_equal(kf.view(np.uint32), of.astype(np.float32).view(np.uint32))
    assert np.array_equal(kw.view(np.uint32), ow.astype(np.float32).view(np.uint32))
    assert check_bspline(kernel_bspline(emu.lib), BSPLINE_X.astype(np.float32), 0.0) < 1e-6


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    libs = {}
    for n in range(1, 7):
        out = d / f"liboracle_mutant{n}.so"
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", f"-DORACLE_MUTANT={n}",
                        str(ROOT / "oracle" / "fsi_oracle.cpp"), "-o", str(out)], check=True)
        libs[n] = ctypes.CDLL(str(out))
        assert libs[n].orc_mutant() == n
    return libs


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_planted_collision_mistakes_are_caught(mutants, n):
    """Each planted mistake in the regularised collision (sign of the trace
    term, the 9/2 factor, the relaxation factor, the Guo prefactor) fails at
    least one closed-form check, in FP64 and in FP32."""
    for T, tol in ((np.float64, 1e-14), (np.float32, 2e-7)):
        worst = max(c(oracle_collide(mutants[n], T)) for c in COLLISION_CHECKS)
        assert worst > 100 * tol, f"mutant {n} ({T.__name__}) passes every check: {worst:.2e}"


@pytest.mark.parametrize("n", [5, 6])
def test_planted_bspline_mistakes_are_caught(mutants, n):
    assert check_bspline(oracle_bspline(mutants[n]), BSPLINE_X, 1e-15) > 1e-3
