"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU: no compute call is made)."""
import ctypes
import glob
import os
import re

from rchem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("rchem_ctx_create", "rchem_jk", "rchem_eri_tensor", "rchem_one_electron",
                 "rchem_scf", "rchem_basis_from_atoms", "rchem_rhf_gradient",
                 "coulomb_repulsion", "contr_coulomb", "overlap", "kinetic",
                 "nuclear_attraction", "Fgamma"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    assert declared_functions() <= set(_lib.SIGNATURES)


def test_opts_defaults_without_gpu():
    o = _lib.default_opts()
    assert o.boys_mode == _lib.BOYS_EXACT and o.precision == _lib.PREC_FP64
    assert o.schwarz_thresh == 1e-12 and o.mixed_thresh == 1e-7


def test_context_without_device_fails_cleanly():
    """Where no gfx950 device is visible (this CPU-only run), creating a
    context returns RCHEM_ERR_NO_DEVICE with a message and no handle -- there
    is no CPU fallback to fall into."""
    import pytest
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    L = _lib.lib()
    h = ctypes.c_void_p(1)
    assert L.rchem_ctx_create(0, ctypes.byref(h)) == 3
    assert h.value is None
    assert L.rchem_last_error(None)
    assert L.rchem_opts_default(None) == 1
    assert L.rchem_jk(None, None, None, None, None) == 1
