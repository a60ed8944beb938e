"""The C ABI: the library loads and exports exactly what include/peps.h declares.

No GPU is touched: only symbol lookup and the host-only model entry point.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from lgt_amd import peps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "peps.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(peps_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ["peps_ctx_create", "peps_bond_env", "peps_compress_mps", "peps_absorb_row",
                 "peps_als_truncate", "peps_charge_svd", "peps_su_bond", "peps_fu_bond_v4",
                 "peps_trotter_step", "peps_energy_v4", "peps_norm2", "peps_dl_traced"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(peps.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    assert sorted(peps.EXPORTED) == _declared()


def test_bond_operator_matches_oracle_model():
    """H_merged_*_site and exp(-tau H) from the host C++ model vs the oracle (no GPU)."""
    from oracle import peps_oracle as O
    for gd in (2, 3):
        nx, ny = 3, 3
        for (orient, ix, iy) in [(0, 1, 1), (0, 2, 3), (1, 1, 1), (1, 3, 2), (1, 2, 2)]:
            if orient == 0:
                H = O.H_merged_h_site(ix, iy, nx, ny, gd, 1.0, 1.0, 0.5).real
            else:
                H = O.H_merged_v_site(ix, iy, nx, ny, gd, 1.0, 1.0, 0.5).real
            Hl = peps.bond_operator(orient, ix, iy, nx, ny, gd, 1.0, 1.0, 0.5)
            np.testing.assert_allclose(Hl, H, atol=1e-14)
            Gl = peps.bond_operator(orient, ix, iy, nx, ny, gd, 1.0, 1.0, 0.5, tau=0.05)
            G = O.gate(H, 0.05).real
            np.testing.assert_allclose(Gl, G, atol=2e-14, rtol=1e-13)
