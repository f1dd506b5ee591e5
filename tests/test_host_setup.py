"""CPU checks of the product's host layer (no GPU): the element operators, metrics
and interface tables it hands to the device ABI equal the oracle's, and the C ABI
library exports every entry point the headers declare."""
import ctypes
import os
import re

import numpy as np
import pytest

import cases
import hifr
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(cases.GOLDEN))


@pytest.mark.parametrize("name", cases.golden_names())
def test_host_setup_matches_oracle(name):
    """Quads and hexes: bit-identical (both sides evaluate the same Lagrange products
    in the same order).  Tets: the host builds its operators independently of the
    oracle (iterative orthonormal Jacobi recurrence, Gauss-Jordan inverse with
    partial pivoting, against the oracle's recursive Jacobi and full-pivoting
    inverse), so they agree to rounding: 1e-14 of each operator's max magnitude."""
    g = cases.load_golden(name)
    s = cases.solver_of(g, host_only=True)
    o = cases.oracle_of(g)
    tet = int(g["ele_type"]) == 2
    for what in ("opp_0", "opp_3", "opp_4", "detjac_upts", "JGinv_upts", "tdA_fpts", "norm_fpts", "pos_upts"):
        a, b = s.get(what), o.get(what)
        assert a.shape == b.shape, what
        if tet and what.startswith("opp"):
            err = np.abs(a - b).max() / np.abs(b).max()
            assert err < 1e-14, (what, err)
        else:
            assert np.array_equal(a, b), what
    l, r = s.interfaces()
    lo, ro = o.face_pairs()
    assert set(map(tuple, map(sorted, zip(l, r)))) == set(map(tuple, map(sorted, zip(lo, ro))))


def test_host_rejects_unsupported_decks():
    g = cases.load_golden("hex_p3_ns_tgv_2x2x2")
    ele_type, xv, c2v = cases.mesh_of(g)
    # LES is built for the Smagorinsky and WALE models without wall model
    with pytest.raises(RuntimeError, match="SGS models"):
        hifr.Solver.from_arrays(cases.product_deck(g, host_only=True, LES=1, SGS_model=2, filter_ratio=1.0,
                                                   wall_model=0), ele_type, xv, c2v)
    with pytest.raises(RuntimeError, match="wall model"):
        hifr.Solver.from_arrays(cases.product_deck(g, host_only=True, LES=1, SGS_model=1, filter_ratio=1.0,
                                                   wall_model=1), ele_type, xv, c2v)
    with pytest.raises(RuntimeError, match="no neighbour"):
        hifr.Solver.from_arrays(cases.product_deck(g, host_only=True, dx_cyclic=float("inf")), ele_type, xv, c2v)


def _declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hifr_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["hifr.h", "hifr_solver.h"])
def test_library_exports_declared_symbols(header):
    lib = ctypes.CDLL(hifr.LIB_PATH)
    names = _declared(os.path.join(ROOT, "include", header))
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
