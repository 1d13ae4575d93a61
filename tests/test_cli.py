"""fdtd_run's Input.txt reader (READ_DATA_FILE, setup.cu:17-58), on the CPU:
--parse-only prints the parameters without touching the device."""
import os
import subprocess

import pytest

from cli_input import ORDER, write_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fdtd-cuda_amd", "fdtd_run")

USED = {"Tend", "dx", "dy", "dz", "NCELLX", "NCELLY", "NCELLZ", "NcpmlX", "NcpmlY", "NcpmlZ", "nano_sphere",
        "nano_sphere_radius", "material", "first_medium", "NUM_freq", "WL_or_freq", "BandWidth", "NONLOCAL",
        "MAX_AMP", "TimeFactor", "lam_0", "StaticField"}


def parse(path):
    r = subprocess.run([EXE, str(path), "--parse-only"], capture_output=True, text=True, timeout=60)
    return r


def test_parse_only_reads_every_token(tmp_path):
    assert os.path.exists(EXE), "fdtd_run is not built (run __graft_entry__.build())"
    vals = write_input(tmp_path / "Input.txt", Tend=777, NCELLY=48, nano_sphere_radius=1.5e-8, lam_0=5.5e-7,
                       t_skip=13, NONLOCAL=3)
    r = parse(tmp_path / "Input.txt")
    assert r.returncode == 0, r.stderr
    got, unused = {}, {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if parts[0] == "unused":
            unused[parts[1]] = parts[2]
        else:
            got[parts[0]] = parts[1]
    assert set(got) == USED
    assert set(unused) == {n for n, _ in ORDER} - USED
    for name in USED:
        v = vals[name]
        if isinstance(v, float):
            assert float(got[name]) == v, name
        else:
            assert int(got[name]) == v, name
    assert int(unused["t_skip"]) == 13


def test_truncated_and_malformed_inputs_fail_cleanly(tmp_path):
    p = tmp_path / "Input.txt"
    write_input(p)
    lines = p.read_text().splitlines()
    p.write_text("\n".join(lines[:20]) + "\n")
    r = parse(p)
    assert r.returncode != 0 and "input ends before" in r.stderr
    write_input(p, Tend=10)
    text = p.read_text().replace("10 Tend", "ten Tend")
    p.write_text(text)
    r = parse(p)
    assert r.returncode != 0 and "bad integer for Tend" in r.stderr
    r = parse(tmp_path / "missing.txt")
    assert r.returncode != 0 and "cannot open" in r.stderr


def test_dispersive_slab_inputs_are_refused(tmp_path):
    """inf_disp_slab 1 or 2 asks for a dispersive half-space or slab
    (setup.cu:156-200), which fdtd_run does not build: it exits with an error
    before touching the device instead of running another geometry."""
    p = tmp_path / "Input.txt"
    write_input(p, inf_disp_slab=1)
    r = subprocess.run([EXE, str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "inf_disp_slab" in r.stderr
