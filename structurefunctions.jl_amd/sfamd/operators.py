"""Structure-function operators (src/StructureFunctionTypes.jl).

Each operator is a pairwise functor of the velocity increment ``du`` and the longitudinal unit
vector ``rhat``.  On the device they are evaluated by the kernels; this module carries the
operator identity (kind and exponents) across the C ABI and evaluates single pairs in numpy for
documentation and tests.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .errors import ArgumentError

# operator kind codes of the C ABI (include/sfmi355x.h)
OP_PROJECTED, OP_S2, OP_S3, OP_T2COMP, OP_L1T2COMP, OP_FULLVEC = range(6)


class AbstractStructureFunctionType:
    """src/StructureFunctionTypes.jl:6."""

    def __call__(self, du=None, rhat=None):
        if du is None and rhat is None:  # identity call (StructureFunctionTypes.jl:11)
            return self
        return self.value(np.asarray(du, dtype=float), np.asarray(rhat, dtype=float))


class AbstractPairwiseStructureFunctionType(AbstractStructureFunctionType):
    """src/StructureFunctionTypes.jl:7."""

    code: tuple  # (kind, nl, nt) for the C ABI

    def order(self) -> int:
        raise NotImplementedError


class AbstractDerivedStructureFunctionType(AbstractStructureFunctionType):
    """Not pairwise (src/StructureFunctionTypes.jl:8); rejected by calculate_structure_function."""


def _mdu_l(du, rh):
    return float(np.dot(du, rh))


def _nhat(rh):
    # HelperFunctions.jl:450-471: 2D quarter turn, 3D normalize(zhat x rhat)
    if len(rh) == 2:
        return np.array([-rh[1], rh[0]])
    c = np.cross([0.0, 0.0, 1.0], rh)
    return c / np.sqrt(np.dot(c, c))


def _mdu_t(du, rh):
    return float(np.dot(du, _nhat(rh)))


def _transverse_norm2(du, rh):
    L = _mdu_l(du, rh)
    return float(np.dot(du, du)) - L * L


@dataclass(frozen=True)
class ProjectedStructureFunctionType(AbstractPairwiseStructureFunctionType):
    """du_L^NL * (transverse term)^NT  (src/StructureFunctionTypes.jl:43-89)."""

    NL: int
    NT: int

    def __post_init__(self):
        if self.NL < 0 or self.NT < 0:
            raise ArgumentError("NL and NT must be non-negative")

    @property
    def code(self):
        return (OP_PROJECTED, self.NL, self.NT)

    def order(self):
        return self.NL + self.NT

    def value(self, du, rh):
        ex = 1.0
        if self.NL:
            ex *= _mdu_l(du, rh) ** self.NL
        if self.NT == 2:
            ex *= _transverse_norm2(du, rh)
        elif self.NT:
            ex *= _mdu_t(du, rh) ** self.NT
        return ex


@dataclass(frozen=True)
class SecondOrderStructureFunctionType(AbstractPairwiseStructureFunctionType):
    """||du||^2 (src/StructureFunctionTypes.jl:97,120)."""

    code = (OP_S2, 0, 0)

    def order(self):
        return 2

    def value(self, du, rh):
        return float(np.dot(du, du))


@dataclass(frozen=True)
class ThirdOrderStructureFunctionType(AbstractPairwiseStructureFunctionType):
    """du_L ||du||^2 = L3 + L1T2 (src/StructureFunctionTypes.jl:106,122)."""

    code = (OP_S3, 0, 0)

    def order(self):
        return 3

    def value(self, du, rh):
        return _mdu_l(du, rh) * float(np.dot(du, du))


@dataclass(frozen=True)
class TransverseComponentSecondOrderStructureFunctionType(AbstractPairwiseStructureFunctionType):
    """||du_t||^2 / (D-1) (src/StructureFunctionTypes.jl:137,147)."""

    code = (OP_T2COMP, 0, 0)

    def order(self):
        return 2

    def value(self, du, rh):
        return _transverse_norm2(du, rh) / (len(rh) - 1)


@dataclass(frozen=True)
class LongitudinalTransverseComponentThirdOrderStructureFunctionType(AbstractPairwiseStructureFunctionType):
    """du_L ||du_t||^2 / (D-1) (src/StructureFunctionTypes.jl:145,150)."""

    code = (OP_L1T2COMP, 0, 0)

    def order(self):
        return 3

    def value(self, du, rh):
        return _mdu_l(du, rh) * _transverse_norm2(du, rh) / (len(rh) - 1)


@dataclass(frozen=True)
class FullVectorStructureFunctionType(AbstractPairwiseStructureFunctionType):
    """||du||^NF (src/StructureFunctionTypes.jl:114,125-128)."""

    NF: int

    @property
    def code(self):
        return (OP_FULLVEC, self.NF, 0)

    def order(self):
        return self.NF

    def value(self, du, rh):
        if self.NF == 2:
            return float(np.dot(du, du))
        return float(np.sqrt(np.dot(du, du))) ** self.NF


@dataclass(frozen=True)
class RotationalSecondOrderStructureFunctionType(AbstractDerivedStructureFunctionType):
    """Helmholtz-derived (src/StructureFunctionTypes.jl:253)."""


@dataclass(frozen=True)
class DivergentSecondOrderStructureFunctionType(AbstractDerivedStructureFunctionType):
    """Helmholtz-derived (src/StructureFunctionTypes.jl:262)."""


@dataclass(frozen=True)
class HelmholtzDecomposition2DType(AbstractDerivedStructureFunctionType):
    """Helmholtz-derived (src/StructureFunctionTypes.jl:270)."""


# named aliases (src/StructureFunctionTypes.jl:157-212)
LongitudinalSecondOrderStructureFunction = L2SF = ProjectedStructureFunctionType(2, 0)
TransverseSecondOrderStructureFunction = T2SF = ProjectedStructureFunctionType(0, 2)
DiagonalConsistentThirdOrderStructureFunction = L3SF = ProjectedStructureFunctionType(3, 0)
DiagonalInconsistentThirdOrderStructureFunction = L2T1SF = ProjectedStructureFunctionType(2, 1)
OffDiagonalInconsistentThirdOrderStructureFunction = L1T2SF = ProjectedStructureFunctionType(1, 2)
OffDiagonalConsistentThirdOrderStructureFunction = T3SF = ProjectedStructureFunctionType(0, 3)
SecondOrderStructureFunction = S2SF = SecondOrderStructureFunctionType()
ThirdOrderStructureFunction = S3SF = ThirdOrderStructureFunctionType()
T2ComponentSF = TransverseComponentSecondOrderStructureFunctionType()
L1T2ComponentSF = LongitudinalTransverseComponentThirdOrderStructureFunctionType()
L1SF = ProjectedStructureFunctionType(1, 0)
T1SF = ProjectedStructureFunctionType(0, 1)
RotationalSecondOrderStructureFunction = RotationalSecondOrderStructureFunctionType()
DivergentSecondOrderStructureFunction = DivergentSecondOrderStructureFunctionType()
HelmholtzDecomposition2DOperator = HelmholtzDecomposition2DType()

SF_TYPE_MAP = {
    "SecondOrderStructureFunction": S2SF,
    "LongitudinalSecondOrderStructureFunction": L2SF,
    "TransverseSecondOrderStructureFunction": T2SF,
    "T2ComponentSF": T2ComponentSF,
    "RotationalSecondOrderStructureFunction": RotationalSecondOrderStructureFunction,
    "DivergentSecondOrderStructureFunction": DivergentSecondOrderStructureFunction,
    "HelmholtzDecomposition2D": HelmholtzDecomposition2DOperator,
    "ThirdOrderStructureFunction": S3SF,
    "DiagonalConsistentThirdOrderStructureFunction": L3SF,
    "DiagonalInconsistentThirdOrderStructureFunction": L2T1SF,
    "OffDiagonalInconsistentThirdOrderStructureFunction": L1T2SF,
    "OffDiagonalConsistentThirdOrderStructureFunction": T3SF,
    "L2SF": L2SF, "T2SF": T2SF, "RotationalSF": RotationalSecondOrderStructureFunction,
    "DivergentSF": DivergentSecondOrderStructureFunction, "L3SF": L3SF, "S2SF": S2SF,
    "S3SF": S3SF, "T3SF": T3SF, "L2T1SF": L2T1SF, "L1T2SF": L1T2SF,
    "L1T2ComponentSF": L1T2ComponentSF,
}


def get_structure_function_type(x, mode=None):
    """Symbol lookup or (order, mode) lookup (src/StructureFunctionTypes.jl:348-409)."""
    if mode is None:
        key = str(x).lstrip(":")
        if key not in SF_TYPE_MAP:
            raise KeyError(f"Unknown structure function type symbol: {x}")
        return SF_TYPE_MAP[key]
    order, mode = int(x), str(mode).lstrip(":")
    if order == 2:
        if mode in ("longitudinal", "long", "L"):
            return L2SF
        if mode in ("transverse", "trans", "T"):
            return T2SF
        if mode in ("transverse_component", "trans_component", "Tcomponent"):
            return T2ComponentSF
        if mode in ("scalar", "total", "S", "full"):
            return S2SF
        if mode in ("rotational", "rot"):
            return RotationalSecondOrderStructureFunction
        if mode in ("divergent", "div"):
            return DivergentSecondOrderStructureFunction
    elif order == 3:
        if mode in ("longitudinal", "long", "L", "diagonal_consistent"):
            return L3SF
        if mode in ("transverse", "trans", "T", "off_diagonal_consistent"):
            return T3SF
        if mode in ("scalar", "total", "S", "full"):
            return S3SF
        if mode == "diagonal_inconsistent":
            return L2T1SF
        if mode == "off_diagonal_inconsistent":
            return L1T2SF
        if mode in ("off_diagonal_inconsistent_component", "L1T2_component"):
            return L1T2ComponentSF
    raise KeyError(f"No mapping for order {order} and mode {mode}")


def order(sf) -> int:
    """src/StructureFunctionTypes.jl:416-424."""
    if isinstance(sf, AbstractDerivedStructureFunctionType):
        return 2
    return sf.order()
