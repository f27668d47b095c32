"""Argument checks of calculate_structure_functions_fused that run before any device call."""
import numpy as np
import pytest

import sfamd


def _xu():
    rng = np.random.default_rng(0)
    return rng.uniform(0, 1, (2, 20)), rng.normal(size=(2, 20))


def test_rejects_two_joint_outputs():
    x, u = _xu()
    vb = sfamd.LinearBinEdges(-1.0, 1.0, 4)
    with pytest.raises(sfamd.ArgumentError, match="at most one joint"):
        sfamd.calculate_structure_functions_fused([(sfamd.L1SF, vb), (sfamd.T1SF, vb)], x, u,
                                                  np.geomspace(1e-2, 1.5, 5))


def test_rejects_more_than_eight_1d_outputs():
    x, u = _xu()
    with pytest.raises(sfamd.ArgumentError, match="at most 8"):
        sfamd.calculate_structure_functions_fused([sfamd.L2SF] * 9, x, u, np.geomspace(1e-2, 1.5, 5))


def test_rejects_derived_and_single_pass_and_empty():
    x, u = _xu()
    with pytest.raises(sfamd.ArgumentError):
        sfamd.calculate_structure_functions_fused([sfamd.RotationalSecondOrderStructureFunction], x, u,
                                                  np.geomspace(1e-2, 1.5, 5))
    with pytest.raises(sfamd.ArgumentError):
        sfamd.calculate_structure_functions_fused([], x, u, np.geomspace(1e-2, 1.5, 5))
    with pytest.raises(sfamd.ArgumentError):
        sfamd.calculate_structure_functions_fused(["L2"], x, u, np.geomspace(1e-2, 1.5, 5))
