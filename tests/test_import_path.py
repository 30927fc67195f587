"""The reference's import path (carcassonne.system, carcassonne.tensors._2d.{dense,sparse},
carcassonne.utils, carcassonne.compression, carcassonne.data, carcassonne.policies, carcassonne.sparse)
resolves to the MI355X implementation.  The CPU test checks the module mapping and the reference's
public names; the GPU test is the reference's own 2D-in-1D transverse-Ising known answer
(carcassonne/tests/test_simulator_2d_in_1d.py:36-47) written with the reference's imports."""
import importlib

import pytest

REFERENCE_NAMES = {
    "carcassonne.system": ["System"],
    "carcassonne.system._2d": ["System"],
    "carcassonne.system.base": ["BaseSystem"],
    "carcassonne.tensors._2d.dense": ["absorbDenseSideIntoCornerFromLeft", "absorbDenseSideIntoCornerFromRight",
                                      "absorbDenseCenterSSIntoSide", "absorbDenseCenterSOSIntoSide",
                                      "formNormalizationStage1", "formNormalizationStage2", "formNormalizationStage3",
                                      "formDenseStage3", "formNormalizationMultiplier", "formNormalizationSubmatrix"],
    "carcassonne.tensors._2d.sparse": ["absorbSparseSideIntoCornerFromLeft", "absorbSparseSideIntoCornerFromRight",
                                       "absorbSparseCenterSOSIntoSide", "formExpectationAndNormalizationMultipliers",
                                       "formExpectationStage1", "formExpectationStage2", "formExpectationStage3"],
    "carcassonne.utils": ["relaxOver", "Multiplier", "Join", "formDataContractor", "prependDataContractor",
                          "computeCompressor", "unitize", "O", "L", "R", "RelaxFailed", "DimensionMismatchError",
                          "UnexpectedTensorRankError", "InvariantViolatedError"],
    "carcassonne.compression": ["computeProductCompressor"],
    "carcassonne.data": ["NDArrayData"],
    "carcassonne.data.cost_tracker": ["CostTracker", "computeCostOfContracting"],
    "carcassonne.policies": ["ConstantStateCompressionPolicy", "RepeatPatternContractionPolicy",
                             "OneDirectionIncrementBandwidthIncreasePolicy",
                             "RelativeStateDifferenceThresholdConvergencePolicy",
                             "RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy"],
    "carcassonne.sparse": ["Identity", "Complete", "OneSiteOperator", "TwoSiteOperator", "TwoSiteOperatorCompressed",
                           "makeSparseOperator", "makeSimpleSparseOperator", "contractSparseTensors"],
}


@pytest.mark.parametrize("module", sorted(REFERENCE_NAMES))
def test_reference_module_resolves_to_the_implementation(module):
    m = importlib.import_module(module)
    impl = importlib.import_module(module.replace("carcassonne", "carcassonne_amd", 1))
    assert m is impl
    for name in REFERENCE_NAMES[module]:
        assert hasattr(m, name), (module, name)
    import carcassonne.data
    from carcassonne_amd.data import HipData
    assert carcassonne.data.NDArrayData is HipData


@pytest.mark.gpu
@pytest.mark.parametrize("direction", [0, 1])
def test_reference_imports_run_the_2d_in_1d_transverse_ising_known_answer(gpu, direction):
    from carcassonne.data import NDArrayData
    from carcassonne.policies import (OneDirectionIncrementBandwidthIncreasePolicy, RepeatPatternContractionPolicy,
                                      RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy,
                                      RelativeStateDifferenceThresholdConvergencePolicy)
    from carcassonne.system import System
    if direction == 0:
        system = System.newTrivialWithSimpleSparseOperator(O=-NDArrayData.Z, OO_LR=[NDArrayData.X, -0.01 * NDArrayData.X])
    else:
        system = System.newTrivialWithSimpleSparseOperator(O=-NDArrayData.Z, OO_UD=[NDArrayData.X, -0.01 * NDArrayData.X])
    system.setPolicy("sweep convergence", RelativeStateDifferenceThresholdConvergencePolicy(1e-5))
    system.setPolicy("run convergence", RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy(1e-7))
    system.setPolicy("bandwidth increase", OneDirectionIncrementBandwidthIncreasePolicy(direction, 2))
    system.setPolicy("contraction", RepeatPatternContractionPolicy([0 + direction, 2 + direction]))
    system.runUntilConverged()
    assert round(system.computeOneSiteExpectation() - (-1.0000250001562545), 7) == 0
    from carcassonne_amd import _lib
    assert _lib._lib is not None       # the HIP library did the work
