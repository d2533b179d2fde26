"""Packed residual-entry tables (44-bit and 48-bit) against a bit-level model,
on the host: reads, writes after reads, writes before reads, emptied entries,
and the whole table image (slack included) after every write."""
import subprocess

from paths import PKG

TOOL = PKG / "lib" / "ck_entry_tests"


def test_packed_entries_match_bit_model():
    out = subprocess.run([str(TOOL)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("0 failed"), out.stdout
