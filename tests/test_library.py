"""CPU checks of the C-ABI library: it loads, exports every entry point
include/smearfem.h declares, and its host utilities reproduce the reference's
known answers (test/runtests.jl:15-35).  No device calls here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import smearfem
from smearfem import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "smearfem.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sfem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.lib_path()))
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) == set(names)


def test_quadrature_known_answers():
    assert smearfem.gaussian_quadrature(-1, 1, 2) == ([-1 / np.sqrt(3), 1 / np.sqrt(3)], [1.0, 1.0])
    assert smearfem.gaussian_quadrature(-1, 1, 3) == ([-np.sqrt(3 / 5), 0.0, np.sqrt(3 / 5)], [5 / 9, 8 / 9, 5 / 9])
    with pytest.raises(smearfem.UndefVarError):
        smearfem.gaussian_quadrature(-1, 1, 5)


def test_basis_known_answers():
    assert list(smearfem.basis_function(-1)[0]) == [1.0, 0.0]
    assert list(smearfem.basis_function(1)[0]) == [0.0, 1.0]
    corners2 = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    for k, c in enumerate(corners2):
        assert list(smearfem.basis_function(*c)[0]) == list(np.eye(4)[k])
    for k, c in enumerate([(x, y, z) for z in (-1, 1) for (x, y) in corners2]):
        assert list(smearfem.basis_function(*c)[0]) == list(np.eye(8)[k])
    N, dN = smearfem.basis_function(0.2)
    assert dN.shape == (1, 2) and list(dN[0]) == [-0.5, 0.5]  # the reference's 1x2 row
    with pytest.raises(smearfem.UndefVarError):
        smearfem.basis_function(0.1, FunctionClass="Q2")


@pytest.mark.parametrize("args,cls", [((0.3,), "Q1"), ((0.3, -0.2), "Q1"), ((0.3, -0.2, 0.7), "Q1"),
                                      ((0.3, -0.2), "Q2"), ((0.3, -0.2, 0.5), "Q2")])
def test_basis_matches_oracle(args, cls):
    import oracle as O

    N1, d1 = smearfem.basis_function(*args, FunctionClass=cls)
    N2, d2 = O.basis_function(*args, FunctionClass=cls)
    assert np.array_equal(N1, N2) and np.array_equal(d1, d2)
