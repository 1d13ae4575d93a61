"""The FP64 discrete adjoint of the oracle against central finite differences.

No reference counterpart exists (the reference has no adjoint); the gradient
dJ/drho of the point objective J = sum_f sum_c |DFT_f E_c(r0)|^2 is pinned by
central differences of J in FP64, which a wrong transpose (a sign, a dropped
CPML or pole term, an off-by-one in the curl) fails.
"""
import numpy as np
import pytest

import oracle as O
from conftest import small_params

C0 = 2.99792458e8
LAMS = np.array([633e-9, 785e-9, 850e-9])


def design_problem(seed=0, shape=(3, 3, 2), steps=260):
    p = small_params(O, n=18, ncpml=5, sphere=0, steps=steps)
    p.inc_phi = 90.0
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.0, 1.0, size=shape)
    o = O.Oracle(p)
    c = o.info.NX // 2
    lo = (c - shape[0] // 2, c - shape[1] // 2, c - shape[2] // 2)
    r0 = (c, c, lo[2] + shape[2] + 1)
    return p, rho, lo, r0


def run_J(p, rho, lo, r0, steps):
    o = O.Oracle(p)
    o.set_design(lo, rho)
    o.set_objective(r0, C0 / LAMS)
    o.step(steps)
    return o.objective()


def test_adjoint_matches_central_differences():
    steps = 260
    p, rho, lo, r0 = design_problem(steps=steps)
    o = O.Oracle(p)
    o.set_design(lo, rho)
    o.set_objective(r0, C0 / LAMS)
    J, grad = o.adjoint(steps)
    assert J > 0
    assert J == pytest.approx(run_J(p, rho, lo, r0, steps), rel=1e-12)
    h = 1e-4
    rng = np.random.default_rng(1)
    idx = [tuple(rng.integers(0, s) for s in rho.shape) for _ in range(6)]
    for v in idx:
        rp, rm = rho.copy(), rho.copy()
        rp[v] += h
        rm[v] -= h
        fd = (run_J(p, rp, lo, r0, steps) - run_J(p, rm, lo, r0, steps)) / (2 * h)
        scale = np.abs(grad).max()
        assert abs(grad[v] - fd) < 1e-5 * scale, (v, grad[v], fd, scale)


def test_adjoint_vacuum_limit_has_no_gradient_from_outside():
    """With every density at zero the design block is vacuum; the gradient is
    still defined (metal growth direction) and finite, and J equals the plain
    vacuum run."""
    steps = 200
    p, rho, lo, r0 = design_problem(steps=steps)
    rho[:] = 0.0
    o = O.Oracle(p)
    o.set_design(lo, rho)
    o.set_objective(r0, C0 / LAMS)
    J, grad = o.adjoint(steps)
    ov = O.Oracle(p)
    ov.set_objective(r0, C0 / LAMS)
    ov.step(steps)
    assert J == pytest.approx(ov.objective(), rel=1e-12)
    assert np.all(np.isfinite(grad))
