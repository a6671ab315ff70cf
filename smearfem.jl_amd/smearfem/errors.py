"""Exceptions named after the Julia exceptions the reference raises, so code
ported from smearFEM can catch the same conditions.

Status codes are the SFEM_ERR_* values of include/smearfem.h.
"""
from __future__ import annotations


class SmearFEMError(Exception):
    code = -8


class ArgumentError(SmearFEMError, ValueError):
    """Julia ArgumentError, e.g. sparse() on an index < 1 (src/fem.jl:253)."""
    code = -1


class BoundsError(SmearFEMError, IndexError):
    """Julia BoundsError: IEN/ID out of range (src/fem.jl:180,243)."""
    code = -2


class DimensionMismatch(SmearFEMError, ValueError):
    """Julia DimensionMismatch: ndim == 1 or nDof != ndim (src/fem.jl:192,212)."""
    code = -3


class SingularException(SmearFEMError, ArithmeticError):
    """LinearAlgebra.SingularException from inv(Jac) (src/fem.jl:195)."""
    code = -4


class UndefVarError(SmearFEMError, NameError):
    """Julia UndefVarError: unknown FunctionClass, nGaussPoints not 2/3."""
    code = -5


class NotConvergedError(SmearFEMError, RuntimeError):
    code = -6


class DeviceMemoryError(SmearFEMError, MemoryError):
    code = -7


class DeviceError(SmearFEMError, RuntimeError):
    code = -8


class UnsupportedError(SmearFEMError, ValueError):
    code = -9


class StateError(SmearFEMError, RuntimeError):
    code = -10


_BY_CODE = {c.code: c for c in (ArgumentError, BoundsError, DimensionMismatch, SingularException,
                                UndefVarError, NotConvergedError, DeviceMemoryError, DeviceError,
                                UnsupportedError, StateError)}


def raise_for(code: int, msg: str = "") -> None:
    if code == 0:
        return
    raise _BY_CODE.get(code, DeviceError)(msg or f"status {code}")
