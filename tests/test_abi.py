"""The C ABI boundary: the library loads, exports every entry point that
include/fdtd.h declares, and the ctypes structs match the C layout.  No
compute calls (CPU only)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import fdtd_amd
from fdtd_amd.abi import GridInfo, Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdtd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s*(fdtd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(fdtd_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(fdtd_amd.EXPORTS) | set(names)
    assert set(fdtd_amd.EXPORTS) <= set(names)


@pytest.mark.parametrize("struct,cname", [(Params, "fdtd_params"), (GridInfo, "fdtd_grid_info")])
def test_struct_layout_matches_header(tmp_path, struct, cname):
    src = tmp_path / "layout.c"
    fields = [f for f, _ in struct._fields_]
    lines = [f'#include "{HEADER}"', "#include <stdio.h>", "#include <stddef.h>", "int main(void){",
             f'printf("%zu\\n", sizeof({cname}));']
    lines += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f in fields]
    lines += ["return 0;}"]
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(struct)
    for f, off in zip(fields, out[1:]):
        assert getattr(struct, f).offset == off, f


def test_default_params_match_oracle():
    import oracle as O
    a, b = fdtd_amd.default_params(), O.default_params()
    assert a.as_dict() == b.as_dict()


def test_create_without_device_fails_cleanly():
    """On a machine without a GPU, create reports an error code instead of exiting."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a device is present")
    lib = fdtd_amd.load_library()
    p = fdtd_amd.default_params()
    h = C.c_void_p()
    rc = lib.fdtd_create(C.byref(p), C.byref(h))
    assert rc == fdtd_amd.abi.FDTD_ERR_DEVICE
    assert b"device" in lib.fdtd_last_error(None)


def _create_rc(**over):
    lib = fdtd_amd.load_library()
    p = fdtd_amd.default_params()
    for k, v in over.items():
        setattr(p, k, v)
    h = C.c_void_p()
    rc = lib.fdtd_create(C.byref(p), C.byref(h))
    if rc == 0:
        lib.fdtd_destroy(h)
    return rc, lib.fdtd_last_error(None).decode()


@pytest.mark.parametrize("over,code,word", [
    ({"TimeFactor": 0.8}, fdtd_amd.abi.FDTD_ERR_INVALID, "Courant"),
    ({"NONLOCAL": 2}, fdtd_amd.abi.FDTD_ERR_UNSUPPORTED, "NONLOCAL"),
    ({"NONLOCAL": 3, "material": 7}, fdtd_amd.abi.FDTD_ERR_UNSUPPORTED, "hydrodynamic"),
    ({"periodic_xy": 1}, fdtd_amd.abi.FDTD_ERR_INVALID, "NcpmlX"),
    ({"periodic_xy": 1, "NcpmlX": 0, "NcpmlY": 0, "inc_theta": 30.0}, fdtd_amd.abi.FDTD_ERR_UNSUPPORTED, "normal"),
    ({"material": 9}, fdtd_amd.abi.FDTD_ERR_INVALID, "material"),
    ({"dx": 1e-9}, fdtd_amd.abi.FDTD_ERR_INVALID, "cubic"),
])
def test_invalid_and_unsupported_parameters_are_rejected_before_the_device(over, code, word):
    """Parameters are checked before any device call, so these codes hold with
    or without a GPU: a time step beyond the 3D Courant limit (TimeFactor <
    sqrt(3)/2) and the reference's inputs this build does not run."""
    rc, msg = _create_rc(**over)
    assert rc == code, (rc, msg)
    assert word in msg


def test_time_factor_at_the_courant_limit_is_accepted_by_validation():
    import torch
    rc, msg = _create_rc(TimeFactor=0.5 * 3 ** 0.5)
    if torch.cuda.device_count() == 0:
        assert rc == fdtd_amd.abi.FDTD_ERR_DEVICE, msg
    else:
        assert rc == 0, msg
