"""Import helpers: the product binding (package dir has hyphens) and the oracle."""
import importlib.util
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "pbc-based-dmpf-monte-carlo-test_amd"
HEADER = ROOT / "include" / "cuckoo_mc.h"


def load_product():
    if "cuckoo_mc" in sys.modules:
        return sys.modules["cuckoo_mc"]
    spec = importlib.util.spec_from_file_location("cuckoo_mc", PKG / "cuckoo_mc.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cuckoo_mc"] = mod
    spec.loader.exec_module(mod)
    return mod


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False
