"""The Lorentz-Drude material (material 7, setup.cu:1256-1275, 1384-1402)
and the grad-div nonlocal Drude pole (NONLOCAL = 1, updatefields.cu:876-891,
1424-1604) on the GPU against the FP64 oracle, split and fused paths, and
through the spectra.  Fields are normalised by the largest E (or H) value;
the bar is that of the other parity tests (2e-4), spectra 1e-4."""
import numpy as np
import pytest

import fdtd_amd
import oracle as O
from conftest import small_params
from fdtd_amd.abi import FDTD_ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu

TOL = 2e-4


def field_errors(s, o):
    fo = [o.get_field(c) for c in range(6)]
    fs = [s.get_field(c) for c in range(6)]
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    return [float(np.abs(fs[c] - fo[c]).max() / (me if c < 3 else mh)) for c in range(6)]


def spec_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("material,nonlocal_,beta", [(7, 0, 0.0), (3, 1, 1.2038e6), (2, 1, 4e6), (7, 1, 2e6)])
@pytest.mark.parametrize("fused", [False, True])
def test_material_against_oracle(material, nonlocal_, beta, fused):
    p = small_params(O, steps=500, phi=90.0)
    p.material = material
    p.NONLOCAL = nonlocal_
    p.hydro_beta = beta
    o = O.Oracle(p)
    s = fdtd_amd.Simulation(p)
    assert s.get_info().n_dispersive[0] == o.info.n_dispersive[0] > 0
    o.step(500)
    if fused:
        s.step(500)
    else:
        s.split_step(500)
    errs = field_errors(s, o)
    fs, fo = s.flux(), o.flux()
    e = spec_err(fs["raw"], fo["raw"])
    print(f"material {material} NONLOCAL {nonlocal_} beta {beta:.3g} fused={fused}: fields "
          f"{['%.1e' % x for x in errs]} face fluxes {e:.1e}", flush=True)
    assert max(errs) < TOL
    assert e < 1e-4


def test_nonlocal_grad_div_changes_the_fields():
    """The grad-div term is live: beta = 4e6 m/s moves the fields away from
    the local run by far more than the parity bar."""
    out = []
    for nl in (0, 1):
        p = small_params(fdtd_amd, steps=500, phi=90.0)
        p.NONLOCAL = nl
        p.hydro_beta = 4e6
        s = fdtd_amd.Simulation(p)
        s.step(500)
        out.append(s.get_field(0))
    d = float(np.abs(out[0] - out[1]).max() / np.abs(out[0]).max())
    print(f"local vs nonlocal at beta 4e6: {d:.2e}")
    assert d > 1e-3


def test_nonlocal_models_rejected_where_unsupported():
    p = small_params(fdtd_amd, steps=100)
    p.NONLOCAL = 1
    s = fdtd_amd.Simulation(p)
    with pytest.raises(fdtd_amd.FDTDError) as e:
        fdtd_amd.Batch([s])
    assert e.value.code == FDTD_ERR_UNSUPPORTED
    with pytest.raises(fdtd_amd.FDTDError) as e:
        s.set_design((15, 15, 15), np.full((2, 2, 2), 0.5))
    assert e.value.code == FDTD_ERR_UNSUPPORTED
