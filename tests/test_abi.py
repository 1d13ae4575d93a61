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
