"""fdtd_flux_noalloc (the allocation-free flux entry point) against fdtd_flux,
through the C ABI.

Both reduce the same accumulated monitor DFTs with the same kernel, one into a
buffer of the handle and one into the library's fixed device array, so the
spectra must agree bit for bit.  Two calls in a row on different handles, the
second with fewer frequencies (a smaller reduction), then the first handle
again: what the larger call left in the shared array must not leak into the
smaller one, nor the smaller into the next.
"""
import numpy as np
import pytest

import fdtd_amd
from conftest import small_params

pytestmark = pytest.mark.gpu


def same(a, b):
    for k in ("sca", "abs", "ext", "raw"):
        assert np.array_equal(a[k], b[k]), k
        assert np.all(np.isfinite(a[k])), k


def test_noalloc_flux_equals_flux_on_two_calls_in_a_row():
    big = fdtd_amd.Simulation(small_params(fdtd_amd, n=28, steps=300, nfreq=48, block=16))
    small = fdtd_amd.Simulation(small_params(fdtd_amd, n=20, ncpml=5, radius=6e-9, steps=200, nfreq=5, block=8))
    big.step(300)
    small.step(200)
    # first call: the allocation-free path contracts the pending samples itself
    a = big.flux_noalloc()
    b = big.flux()
    same(a, b)
    assert np.abs(b["ext"]).max() > 0
    # second call, smaller input: 5 frequencies, flux first this time
    d = small.flux()
    c = small.flux_noalloc()
    same(c, d)
    assert np.abs(d["raw"]).max() > 0
    assert c["raw"].shape == (12, 5)
    # the larger handle again after the smaller call left its sums in the array
    big.step(40)
    e = big.flux_noalloc()
    same(e, big.flux())
    print("ext[0] big %.6e small %.6e" % (a["ext"][0], c["ext"][0]), flush=True)
    big.close()
    small.close()


def test_noalloc_flux_rejects_more_frequencies_than_its_array_holds():
    n = fdtd_amd.NOALLOC_MAX_FREQ + 1
    s = fdtd_amd.Simulation(small_params(fdtd_amd, n=20, ncpml=5, radius=4e-9, steps=8, nfreq=n, block=8))
    s.step(8)
    with pytest.raises(RuntimeError, match="FDTD_NOALLOC_MAX_FREQ"):
        s.flux_noalloc()
    # the handle stays usable through the allocating entry point
    assert np.all(np.isfinite(s.flux()["raw"]))
    s.close()
