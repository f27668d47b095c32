"""Host-side mirror of the reference API (no GPU): operators, bin specifications, result
containers, and the validation calculate_structure_function performs before any device work."""
import math

import numpy as np
import pytest

import sfamd
from sfamd import dispatch
from sfamd.errors import ArgumentError, DimensionMismatch


def test_operator_values_match_reference_kats():
    # test/test_core_correctness.jl:142-177
    du, rh = [2.0, 3.0, 4.0], [1.0, 0.0, 0.0]
    assert sfamd.L2SF(du, rh) == pytest.approx(4.0)
    assert sfamd.T2SF(du, rh) == pytest.approx(25.0)
    assert sfamd.T2ComponentSF(du, rh) == pytest.approx(12.5)
    assert sfamd.L3SF(du, rh) == pytest.approx(8.0)
    assert sfamd.L1T2SF(du, rh) == pytest.approx(50.0)
    assert sfamd.L1T2ComponentSF(du, rh) == pytest.approx(25.0)
    assert sfamd.S3SF(du, rh) == pytest.approx(sfamd.L3SF(du, rh) + sfamd.L1T2SF(du, rh))
    assert sfamd.FullVectorStructureFunctionType(3)(du, rh) == pytest.approx(math.sqrt(29) ** 3)
    assert sfamd.L2T1SF([2.0, 3.0], [1.0, 0.0]) == pytest.approx(12.0)
    assert sfamd.T3SF([2.0, 3.0], [1.0, 0.0]) == pytest.approx(27.0)


def test_operator_lookup_and_order():
    # src/StructureFunctionTypes.jl:348-424
    assert sfamd.get_structure_function_type("L2SF") is sfamd.L2SF
    assert sfamd.get_structure_function_type(":T3SF") is sfamd.T3SF
    assert sfamd.get_structure_function_type(2, "longitudinal") == sfamd.L2SF
    assert sfamd.get_structure_function_type(3, "off_diagonal_inconsistent") == sfamd.L1T2SF
    assert sfamd.order(sfamd.L1T2SF) == 3 and sfamd.order(sfamd.S2SF) == 2
    assert sfamd.ProjectedStructureFunctionType(2, 0) == sfamd.L2SF
    assert sfamd.L2SF() is sfamd.L2SF
    with pytest.raises(KeyError):
        sfamd.get_structure_function_type("nope")


def test_bin_specifications():
    e = np.geomspace(1e-3, 1.0, 9)
    lb = sfamd.LogBinEdges(e)
    np.testing.assert_allclose(lb.edges(), e, rtol=1e-14)
    np.testing.assert_array_equal(lb.edges(), np.exp(np.linspace(np.log(e[0]), np.log(e[-1]), 9)))
    with pytest.raises(ArgumentError):
        sfamd.LogBinEdges([0.0, 1.0])
    ip = sfamd.InfPaddedBinEdges([-np.inf, 1.0, 2.0, np.inf])
    assert ip.edges().tolist() == [-np.inf, 1.0, 2.0, np.inf]
    assert sfamd.n_histogram_bins(ip) == 3
    assert sfamd.LinearBinEdges(0.0, 1.0, 5).edges().tolist() == [0.0, 0.25, 0.5, 0.75, 1.0]
    assert sfamd.midpoints([0.0, 1.0, 3.0]).tolist() == [0.5, 2.0]
    with pytest.raises(ArgumentError):
        sfamd.as_edges([1.0])


def test_result_containers():
    e = [0.0, 1.0, 2.0]
    with pytest.raises(DimensionMismatch):
        sfamd.StructureFunction(sfamd.L2SF, e, [1.0])
    a = sfamd.StructureFunctionSumsAndCounts(sfamd.L2SF, e, np.array([1.0, 2.0]), np.array([1, 2]))
    b = sfamd.StructureFunctionSumsAndCounts(sfamd.L2SF, e, np.array([3.0, 4.0]), np.array([3, 0]))
    c = a + b
    assert c.sums.tolist() == [4.0, 6.0] and c.counts.tolist() == [4, 2]
    with pytest.raises(ArgumentError):
        a + sfamd.StructureFunctionSumsAndCounts(sfamd.T2SF, e, np.zeros(2), np.zeros(2, int))
    avg = dispatch._finalize(b, None)
    assert avg.values[0] == 1.0 and math.isnan(avg.values[1])  # empty bin -> NaN
    j = sfamd.StructureFunction2DSumsAndCounts(sfamd.L1SF, e, [0.0, 1.0, 2.0, 3.0], np.ones((2, 3)),
                                               np.ones((2, 3), int))
    m = sfamd.marginalize(j)
    assert m.sums.tolist() == [3.0, 3.0] and m.counts.tolist() == [3, 3]
    with pytest.raises(ArgumentError):
        dispatch._finalize(j, sfamd.StructureFunction)
    with pytest.raises(DimensionMismatch):
        sfamd.StructureFunction2DSumsAndCounts(sfamd.L1SF, e, [0.0, 1.0], np.ones((2, 3)), np.ones((2, 3)))


def test_validation_before_device():
    x = np.zeros((2, 10))
    u = np.zeros((2, 10))
    e = [0.0, 1.0]
    with pytest.raises(ArgumentError):
        sfamd.calculate_structure_function(sfamd.RotationalSecondOrderStructureFunction, x, u, e)
    with pytest.raises(ArgumentError):
        sfamd.calculate_structure_function(sfamd.L2SF, (x[0], x[1]), (u[0], u[1]), e)
    with pytest.raises(DimensionMismatch):
        sfamd.calculate_structure_function(sfamd.L2SF, np.zeros((4, 10)), np.zeros((4, 10)), e)
    with pytest.raises(DimensionMismatch):
        sfamd.calculate_structure_function(sfamd.L2SF, np.zeros((3, 10)), u, e)
    with pytest.raises(DimensionMismatch):
        sfamd.calculate_structure_function(sfamd.L2SF, x, np.zeros((2, 9)), e)
    with pytest.raises(DimensionMismatch):  # sphere takes 2 coordinates even for 3D velocity
        sfamd.calculate_structure_function(sfamd.L2SF, np.zeros((3, 10)), np.zeros((3, 10)), e,
                                           distance_metric=sfamd.Haversine(1.0))
    with pytest.raises(ArgumentError):  # UInt32 counts cannot hold N(N-1)/2 for N = 92683
        sfamd.calculate_structure_function(sfamd.L2SF, np.zeros((2, 92683)), np.zeros((2, 92683)), e)

    class Cityblock:
        pass
    with pytest.raises(ArgumentError):
        sfamd.calculate_structure_function(sfamd.L2SF, x, u, e, distance_metric=Cityblock())
    with pytest.raises(ArgumentError):
        sfamd.MI355XBackend(precision="fp16")
