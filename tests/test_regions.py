"""The memory-region helpers that need no device, on the host: the two modes
of a set of pair-walk times (neighbours, straddlers and outliers moving
neither; one region, too few times or modes closer than the ratio giving no
split) and the order in which a spread range takes its pieces."""
import subprocess

from paths import PKG

TOOL = PKG / "lib" / "ck_region_tests"


def test_region_helpers():
    out = subprocess.run([str(TOOL)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("0 failed"), out.stdout
