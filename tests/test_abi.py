"""The C-ABI library loads and exports every symbol include/madsdp_hip.h
declares (no device calls)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "madsdp_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msdp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "msdp_solver_solve" in syms and "msdp_k_syevd" in syms
    assert len(syms) > 30


def test_library_exports_every_symbol():
    from madsdp_amd import lib as L
    lib = ctypes.CDLL(str(L.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from madsdp_amd import lib as L
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_options_struct_defaults():
    from madsdp_amd import lib as L
    lib = L.load()
    opt = L.Options()
    lib.msdp_default_options(ctypes.byref(opt))
    assert opt.tol == 1e-7 and opt.max_iter == 100 and opt.regularize_schur == 1e-12
    assert opt.max_schur_reg == 8 and opt.tau == 0.95 and opt.dc_A == 2.0
