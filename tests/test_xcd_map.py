"""Host check of the workgroup-to-XCD maps (cdocker_amd/csrc/xcd_map.hpp)
that the kz-plane and line-group kernels call: tests/cpp/xcd_map_test.cpp,
built from the same header by tests/cpp/Makefile, runs both maps' grids for
the item counts of every grid edge (N/2 + 1 kz planes, the plane form's line
groups) and unit counts 1-80 plus large ones, and checks that every (item,
unit) pair is covered exactly once, nothing falls out of range, and the
plane-major map's XCD loads differ by at most one item.  A wrong map would
skip or repeat a plane of some pose, which the GPU parity tests would see
only for the pose counts they happen to use."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "xcd_map_test")


def test_xcd_maps_cover_every_pair_once():
    subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "_build/xcd_map_test"], check=True,
                   capture_output=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
