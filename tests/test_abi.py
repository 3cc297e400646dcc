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
    assert opt.precision == 0 and opt.shared_device == 0


def test_tuning_parameters():
    """Kernel-selection parameters: known names read and set within range,
    unknown names and out-of-range values rejected; the autotuner's table, if
    present, names only known parameters."""
    import json
    from madsdp_amd import lib as L
    lib = L.load()
    v = ctypes.c_double()
    assert lib.msdp_get_tuning(b"streamk_occupancy", ctypes.byref(v)) == 0 and v.value in (1, 2)
    old = v.value
    assert lib.msdp_set_tuning(b"streamk_occupancy", 1.0) == 0
    assert lib.msdp_get_tuning(b"streamk_occupancy", ctypes.byref(v)) == 0 and v.value == 1
    assert lib.msdp_set_tuning(b"streamk_occupancy", old) == 0
    assert lib.msdp_set_tuning(b"streamk_occupancy", 3.0) < 0
    assert lib.msdp_set_tuning(b"no_such_parameter", 1.0) < 0
    if L.TUNING_PATH.exists():
        params = json.loads(L.TUNING_PATH.read_text())["params"]
        for name, value in params.items():
            assert lib.msdp_get_tuning(name.encode(), ctypes.byref(v)) == 0
