"""The seeded nanorod sweep workload (fdtd_amd/sweep.py), on the CPU:
splitmix64 against the published test vectors of its reference
implementation (Steele, Lea and Flood 2014; seed 0 and seed 1234567), and
the 32 draws of seed 20240601 against the workload's stated ranges."""
import sys
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "fdtd-cuda_amd"))

from fdtd_amd import sweep  # noqa: E402


def test_splitmix64_known_answers():
    r = sweep.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r = sweep.SplitMix64(1234567)
    assert [r.next_u64() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                                  4593380528125082431, 16408922859458223821]
    r = sweep.SplitMix64(42)
    u = [r.uniform() for _ in range(1000)]
    assert min(u) >= 0.0 and max(u) < 1.0


def test_draws_follow_the_workload():
    cfgs, redraws = sweep.draw(seed=20240601)
    assert len(cfgs) == 32 and redraws == 3
    geom = {"tfsf_inset": 10, "abs_offset": 5}
    for c in cfgs:
        assert all(e in sweep.EDGES for e in c["edges"])
        assert c["axis"] in (0, 1, 2)
        assert 40e-9 <= c["length"] <= 120e-9 and 10e-9 <= c["diameter"] <= 30e-9
        assert sweep.rod_fits(c["edges"], c["axis"], c["length"], c["diameter"], geom)
    again, _ = sweep.draw(seed=20240601)
    assert again == cfgs
    assert sweep.total_cells(cfgs) == 146675344
