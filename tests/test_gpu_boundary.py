"""Failure handling and checkpoint/resume through the C ABI.

* Divergence: a NaN or Inf in the fields is reported as FDTD_ERR_NONFINITE
  (the reference probes ex for NaN every step and stops, main.cu:2728-2731).
  The condition is forced from the host with fdtd_set_field; every kernel
  still runs to completion.
* Checkpoint/resume: 2N steps straight against N steps, an image saved,
  restored into a fresh handle, N more steps; fields, DFT spectra and the
  point monitor must agree bit for bit.
"""
import numpy as np
import pytest

import fdtd_amd
from fdtd_amd.abi import FDTD_ERR_INVALID, FDTD_ERR_NONFINITE
from conftest import small_params
import oracle as O

pytestmark = pytest.mark.gpu


def probe_cell(info):
    # the reference's NaN probe: ex at (NX/2, NY/2, NZ/2 - NZ/4) of the augmented grid
    return (info.NX // 2, info.NY // 2, info.NZ // 2 - info.NZ // 4)


def test_inf_at_the_probe_cell_is_reported():
    p = small_params(fdtd_amd)
    s = fdtd_amd.Simulation(p)
    s.step(10)
    s.sync()
    ex = s.get_field(0)
    ex[probe_cell(s.info)] = np.inf
    s.set_field(0, ex)
    s.step(1)                      # the flag is raised by this step's kernels
    with pytest.raises(fdtd_amd.FDTDError) as e:
        s.sync()
    assert e.value.code == FDTD_ERR_NONFINITE
    t = s.get_info().t
    with pytest.raises(fdtd_amd.FDTDError) as e:
        s.step(5)
    assert e.value.code == FDTD_ERR_NONFINITE
    assert s.get_info().t == t     # no step ran after the flag was seen
    with pytest.raises(fdtd_amd.FDTDError) as e:
        s.flux()
    assert e.value.code == FDTD_ERR_NONFINITE


def test_nan_off_the_probe_reaches_the_monitors():
    p = small_params(fdtd_amd, sphere=0)
    s = fdtd_amd.Simulation(p)
    s.step(3)
    info = s.info
    c = info.NY // 2
    hz = s.get_field(5)
    hz[info.sca_start[0] - 3, c + 2, c + 5] = np.nan
    s.set_field(5, hz)
    s.step(16)
    with pytest.raises(fdtd_amd.FDTDError) as e:
        s.sync()
    assert e.value.code == FDTD_ERR_NONFINITE


def test_finite_run_reports_no_error():
    p = small_params(fdtd_amd)
    s = fdtd_amd.Simulation(p)
    s.step(200)
    s.sync()
    assert np.all(np.isfinite(s.flux()["ext"]))


def _run_pair(p, n, monitor=False, alter_target=None):
    a = fdtd_amd.Simulation(p)
    c = a.info.NX // 2
    if monitor:
        a.set_point_monitor((c, c, c), 4 * n)
    a.step(n)
    img = a.checkpoint()
    a.step(n)
    b = fdtd_amd.Simulation(alter_target(p) if alter_target else p)
    b.step(3)                      # a different time level and state before the load
    b.restore(img)
    assert b.get_info().t == n
    b.step(n)
    for comp in range(6):
        fa, fb = a.get_field(comp), b.get_field(comp)
        assert np.array_equal(fa, fb), comp
        assert np.abs(fa).max() > 0
    for k in ("sca", "abs", "ext", "raw"):
        assert np.array_equal(a.flux()[k], b.flux()[k]), k
    assert a.get_info().sample_count == b.get_info().sample_count
    if monitor:
        f = a.spectrum_setup()[0]
        ea, _ = a.point_spectrum(f)
        eb, _ = b.point_spectrum(f)
        assert np.array_equal(ea, eb)
    return a, b


@pytest.mark.parametrize("hydro", [0, 3])
def test_checkpoint_resume_is_bit_exact(hydro):
    p = small_params(fdtd_amd, steps=300)
    p.NONLOCAL = hydro
    _run_pair(p, 150, monitor=True)


def test_checkpoint_restores_materials_into_a_plain_handle():
    """The target was created without the sphere: the load brings the saved masks."""
    p = small_params(fdtd_amd, steps=300)

    def plain(q):
        r = fdtd_amd.default_params()
        for name, _ in type(q)._fields_:
            setattr(r, name, getattr(q, name))
        r.nano_sphere = 0
        return r

    a, b = _run_pair(p, 70, alter_target=plain)
    assert list(b.get_info().n_dispersive) == list(a.get_info().n_dispersive)
    assert np.array_equal(a.get_materials(1), b.get_materials(1))


def test_checkpoint_rejects_other_parameters():
    a = fdtd_amd.Simulation(small_params(fdtd_amd))
    a.step(5)
    img = a.checkpoint()
    b = fdtd_amd.Simulation(small_params(fdtd_amd, n=26))
    with pytest.raises(fdtd_amd.FDTDError) as e:
        b.restore(img)
    assert e.value.code == FDTD_ERR_INVALID
    with pytest.raises(fdtd_amd.FDTDError):
        b.restore(img[:100])


def test_checkpoint_resume_matches_the_oracle():
    """Resumed run against the FP64 oracle run straight through."""
    p = small_params(fdtd_amd, steps=320)
    a = fdtd_amd.Simulation(p)
    a.step(160)
    img = a.checkpoint()
    a.close()
    b = fdtd_amd.Simulation(p)
    b.restore(img)
    b.step(160)
    o = O.Oracle(small_params(O, steps=320))
    o.step(320)
    fo = [o.get_field(c) for c in range(6)]
    me = max(np.abs(f).max() for f in fo[:3])
    mh = max(np.abs(f).max() for f in fo[3:])
    worst = max(float(np.abs(b.get_field(c) - fo[c]).max() / (me if c < 3 else mh)) for c in range(6))
    assert worst < 2e-4
