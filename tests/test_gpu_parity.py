"""HIP solver against the FP64 oracle on small grids, through the C ABI.

Tolerances: fields and DFT spectra are FP32 on the device against FP64 in the
oracle; after ~160 steps the relative error (L-inf normalised by the field
maximum) stays near 1e-6..1e-5, so 2e-4 is the bar here.
"""
import numpy as np
import pytest

import fdtd_amd
import oracle as O
from conftest import small_params

pytestmark = pytest.mark.gpu

TOL = 2e-4


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def run_pair(p, steps, fused):
    o = O.Oracle(p)
    s = fdtd_amd.Simulation(p)
    o.step(steps)
    if fused:
        s.step(steps)
    else:
        s.split_step(steps)
    return o, s


@pytest.mark.parametrize("fused", [False, True])
def test_fields_gold_sphere(fused):
    p = small_params(O, steps=160)
    o, s = run_pair(p, 160, fused)
    assert s.get_info().n_dispersive[0] == o.info.n_dispersive[0] > 0
    for c in range(6):
        err = rel(s.get_field(c), o.get_field(c))
        print(f"fused={fused} comp={c} rel={err:.3e}", flush=True)
        assert err < TOL, (c, err)


@pytest.mark.parametrize("fused", [False, True])
def test_fields_vacuum_x_polarised(fused):
    p = small_params(O, sphere=0, phi=90.0, steps=140)
    o, s = run_pair(p, 140, fused)
    for c in range(6):
        err = rel(s.get_field(c), o.get_field(c))
        print(f"vacuum fused={fused} comp={c} rel={err:.3e}", flush=True)
        if np.max(np.abs(o.get_field(c))) > 1e-6:
            assert err < TOL, (c, err)


def test_line_matches_oracle():
    p = small_params(O, steps=100)
    o, s = run_pair(p, 100, True)
    eo, ho = o.get_line()
    es, hs = s.get_line()
    assert rel(es, eo) < 1e-12
    assert rel(hs, ho) < 1e-12


@pytest.mark.parametrize("fused", [False, True])
def test_dft_and_flux(fused):
    p = small_params(O, steps=160, block=8)
    o, s = run_pair(p, 161, fused)   # ends mid-block: exercises the partial contraction
    worst = 0.0
    for face in range(12):
        for comp in range(4):
            a, b = s.get_dft(face, comp), o.get_dft(face, comp)
            worst = max(worst, rel(a, b))
    print(f"dft worst rel={worst:.3e}", flush=True)
    assert worst < 1e-3
    fo, fs = o.flux(), s.flux()
    for k in ("raw", "sca", "abs", "ext"):
        err = rel(fs[k], fo[k])
        print(f"flux {k} rel={err:.3e}", flush=True)
        assert err < 1e-3, (k, err)


def test_fused_equals_split():
    p = small_params(O, steps=120)
    a = fdtd_amd.Simulation(p)
    b = fdtd_amd.Simulation(p)
    a.step(120)
    b.split_step(120)
    for c in range(6):
        fa, fb = a.get_field(c), b.get_field(c)
        err = rel(fa, fb)
        print(f"fused vs split comp={c} rel={err:.3e} bitwise={np.array_equal(fa, fb)}", flush=True)
        assert err < 1e-6


@pytest.mark.parametrize("tj", [4, 8, 16])
def test_tile_shapes_agree(tj, monkeypatch):
    p = small_params(O, steps=60)
    ref = fdtd_amd.Simulation(p)
    ref.split_step(60)
    monkeypatch.setenv("FDTD_TJ", str(tj))
    monkeypatch.setenv("FDTD_ICHUNK", "7")
    s = fdtd_amd.Simulation(p)
    s.step(60)
    for c in range(6):
        assert rel(s.get_field(c), ref.get_field(c)) < 1e-6
