"""GPU parity of the System layer (absorption, observables, solves, compression, gauge moves,
driver) against the oracle on identical seeded tensors, and the reference's known answers
(carcassonne/tests/test_system.py, test_simulator_2d_in_1d.py) computed on the GPU."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(seed, maximum_dimension=2, O=None):
    from oracle.system import System as OSystem
    from tests.oracle_bridge import system_to_product
    osys = OSystem.newRandom(np.random.default_rng(seed), maximum_dimension=maximum_dimension, O=O)
    return system_to_product(osys), osys


@pytest.mark.parametrize("seed", range(4))
def test_random_walk_absorption_matches_oracle(gpu, seed):
    """contractTowards (normalizeAxis + three absorptions) on a random system; the polar factor
    is unique for full-rank matricizations, so raw tensors must agree, not only energies."""
    from tests.oracle_bridge import compare_systems
    psys, osys = _pair(seed, maximum_dimension=3)
    rng = np.random.default_rng(100 + seed)
    for direction in rng.integers(0, 4, 4):
        psys.contractTowards(int(direction))
        osys.contractTowards(int(direction))
        assert compare_systems(psys, osys) < 1e-8
    e_p, n_p = psys.computeExpectationAndNormalization()
    e_o, n_o = osys.computeExpectationAndNormalization()
    assert abs(e_p - e_o) <= 1e-9 * abs(e_o) and abs(n_p - n_o) <= 1e-9 * abs(n_o)


def test_observables_match_oracle(gpu):
    psys, osys = _pair(11, maximum_dimension=2)
    for d in (0, 1, 2, 3):
        psys.contractTowards(d)
        osys.contractTowards(d)
    assert abs(psys.computeCenterSiteExpectation() - osys.computeCenterSiteExpectation()) < 1e-9
    assert abs(psys.computeOneSiteExpectation() - osys.computeOneSiteExpectation()) < 1e-9
    assert abs(psys.computeEstimatedOneSiteExpectation(0) - osys.computeEstimatedOneSiteExpectation(0)) < 1e-9


def test_minimize_expectation_relaxover_matches_oracle(gpu):
    """The reference's solve (relaxOver, k=3, tol 1e-7, 100 multiplications) on a grown random
    environment with a D=2 center on both axes: the same tensors on both sides (the oracle's random
    system, uploaded), five contractions, then the solve takes the same number of restarts and
    multiplications as the oracle's (more than one restart) and lands on the same energy.  (A bandwidth
    increase from D=1 is not used to grow the center: its new isometries come from SVDs of rank-one
    matricizations, whose null-space columns are a free choice of each SVD.)"""
    psys, osys = _pair(5, maximum_dimension=2)
    for d in (0, 1, 2, 3, 0):
        osys.contractTowards(d)
        psys.contractTowards(d)
    assert tuple(psys.state_center_data.shape) == (2, 2, 2, 2, 2) == tuple(osys.state.shape)
    e_p0, e_o0 = psys.computeExpectation(), osys.computeExpectation()
    assert abs(e_p0 - e_o0) <= 1e-9 * abs(e_o0)
    osys.minimizeExpectation()
    psys.minimizeExpectation()
    print("oracle", osys.last_solve, "device", psys.last_solve, flush=True)
    assert osys.last_solve["restarts"] > 1
    assert psys.last_solve["restarts"] == osys.last_solve["restarts"]
    assert psys.last_solve["mults"] == osys.last_solve["mults"]
    e_p, e_o = psys.computeExpectation(), osys.computeExpectation()
    assert abs(e_p - e_o.real) <= 1e-9 * abs(e_o)
    assert e_p < e_p0


def test_gauge_moves_leave_expectations_invariant(gpu):
    """normalizeCorner/Side/Center... (reference test_system.py:276-305)."""
    psys, _ = _pair(5, maximum_dimension=3)
    psys.contractTowards(0)
    psys.contractTowards(1)
    e1, n1 = psys.computeExpectationAndNormalization()
    psys.normalizeCornerAndDenormalizeSide(2, 1)
    psys.normalizeSideAndDenormalizeCenter(3)
    psys.normalizeCenterAndDenormalizeSide(0)
    e2, n2 = psys.computeExpectationAndNormalization()
    assert abs(e2 / e1 - 1) < 1e-9 and abs(n2 / n1 - 1) < 1e-9


def test_compression_exact_at_full_rank(gpu):
    """compressCornerStateTowards to the current dimension is exact (reference test_system.py:64-81)."""
    from carcassonne_amd.sparse import Identity
    psys, _ = _pair(8, maximum_dimension=2)
    psys.contractTowards(0)
    psys.contractTowards(1)
    e1, n1 = psys.computeExpectationAndNormalization()
    for corner_id in range(4):
        for direction in range(2):
            dim = psys.corners[corner_id][Identity()].shape[3 * direction]
            psys.compressCornerStateTowards(corner_id, direction, dim)
    e2, n2 = psys.computeExpectationAndNormalization()
    assert abs(e2 / e1 - 1) < 1e-7 and abs(n2 / n1 - 1) < 1e-7


def test_compression_recovers_enlarged_bond(gpu):
    """Enlarge a corner-side bond by an isometry, then compress back (reference test_system.py:40-63)."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.sparse import Identity, mapOverSparseData
    from carcassonne_amd.system._2d import sideFromCorner
    psys, _ = _pair(9, maximum_dimension=2)
    for i in range(4):
        psys.contractTowards(i)
    e1, n1 = psys.computeExpectationAndNormalization()
    corner_id, direction = 1, 0
    axis = 3 * direction
    old = psys.corners[corner_id][Identity()].shape[axis]
    new = 2 * (old + 1)
    enl, enl_c = HipData.newEnlargener(old, new)
    psys.corners[corner_id] = mapOverSparseData(lambda d: d.absorbMatrixAt(axis, enl).absorbMatrixAt(axis + 1, enl_c),
                                                psys.corners[corner_id])
    side_id = sideFromCorner(corner_id, direction)
    saxis = 3 - axis
    psys.sides[side_id] = mapOverSparseData(lambda d: d.absorbMatrixAt(saxis, enl_c).absorbMatrixAt(saxis + 1, enl),
                                            psys.sides[side_id])
    e2, n2 = psys.computeExpectationAndNormalization()
    assert abs(n2 / n1 - 1) < 1e-9 and abs(e2 / e1 - 1) < 1e-9
    psys.compressCornerStateTowards(corner_id, direction, old)
    e3, n3 = psys.computeExpectationAndNormalization()
    assert abs(n3 / n1 - 1) < 1e-6 and abs(e3 / e1 - 1) < 1e-6


@pytest.mark.parametrize("walk", [[0], [1, 2], [0, 0, 3], [2, 3, 1, 0, 1]])
def test_identity_operator_walk_is_width_times_height(gpu, walk):
    """Silly field walk (reference test_system.py:308-358): E = width*height exactly, N = 1."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    system = System.newTrivialWithSimpleSparseOperator(O=HipData.newIdentity(1))
    width = height = 1
    for d in walk:
        system.contractTowards(d)
        if d in (0, 2):
            width += 1
        else:
            height += 1
    assert system.computeExpectation() == pytest.approx(width * height, abs=1e-12)
    assert system.computeNormalization() == pytest.approx(1.0, abs=1e-12)


def test_magnetic_field_product_state_walk(gpu):
    """Z field with up/down product states (reference test_system.py:360-439)."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    system = System.newTrivialWithSimpleSparseOperator(O=HipData.Z)
    up = HipData.fromArray(np.array([[[[[1.0, 0.0]]]]]))
    down = HipData.fromArray(np.array([[[[[0.0, 1.0]]]]]))
    states, spins = (up, down), (1, -1)
    rng = np.random.default_rng(3)
    UL = U = UR = L_ = R_ = DL = D = DR = 0
    for _ in range(6):
        direction, spin = int(rng.integers(0, 4)), int(rng.integers(0, 2))
        system.setStateCenter(states[spin])
        system.contractTowards(direction)
        if direction == 0:
            R_ += spins[spin]; UR += U; DR += D
        elif direction == 1:
            U += spins[spin]; UL += L_; UR += R_
        elif direction == 2:
            L_ += spins[spin]; UL += U; DL += D
        else:
            D += spins[spin]; DL += L_; DR += R_
    final = 1
    system.setStateCenter(states[final])
    total = UL + U + UR + L_ + R_ + spins[final] + DL + D + DR
    assert system.computeExpectation() == pytest.approx(total, abs=1e-10)
    assert system.computeNormalization() == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("direction", [0, 1])
def test_2d_in_1d_transverse_ising_known_answer(gpu, direction):
    """-1.0000250001562545 to 7 places (reference test_simulator_2d_in_1d.py:36-47), the whole
    sweep (bandwidth increases, contractions, relaxOver solves, one-site observables) on the GPU."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    from carcassonne_amd import policies as P
    X, Z = HipData.X, HipData.Z
    if direction == 0:
        system = System.newTrivialWithSimpleSparseOperator(O=-Z, OO_LR=[X, X * -0.01])
    else:
        system = System.newTrivialWithSimpleSparseOperator(O=-Z, OO_UD=[X, X * -0.01])
    system.setPolicy("sweep convergence", P.RelativeStateDifferenceThresholdConvergencePolicy(1e-5))
    system.setPolicy("run convergence", P.RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy(1e-7))
    system.setPolicy("bandwidth increase", P.OneDirectionIncrementBandwidthIncreasePolicy(direction, 2))
    system.setPolicy("contraction", P.RepeatPatternContractionPolicy([0 + direction, 2 + direction]))
    system.runUntilConverged(maximum_number_of_runs=30, maximum_number_of_steps=200)
    assert round(system.computeOneSiteExpectation() - (-1.0000250001562545), 7) == 0


def test_2d_in_1d_ferromagnet_known_answer(gpu):
    """-1 (reference test_simulator_2d_in_1d.py:23-34)."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    from carcassonne_amd import policies as P
    Z = HipData.Z
    system = System.newTrivialWithSimpleSparseOperator(OO_LR=[Z, -Z])
    system.setPolicy("sweep convergence", P.RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy(1e-7))
    system.setPolicy("run convergence", P.RelativeOneSiteExpectationDifferenceThresholdConvergencePolicy(1e-7))
    system.setPolicy("bandwidth increase", P.OneDirectionIncrementBandwidthIncreasePolicy(0))
    system.setPolicy("contraction", P.RepeatPatternContractionPolicy([0, 2]))
    system.runUntilConverged(maximum_number_of_runs=20, maximum_number_of_steps=100)
    assert round(system.computeOneSiteExpectation() - (-1.0), 7) == 0


def _operator_systems(corner_id, direction, seed):
    """The same two-site-term system in the oracle and on the GPU (reference test_system.py:82-177)."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    from carcassonne_amd.system._2d import sideFromCorner
    from carcassonne_amd.utils import L, R
    from oracle.system import System as OSystem
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(2):
        a = rng.standard_normal((2, 2))
        ops.append(a + a.T)
    side_id = sideFromCorner(corner_id, direction)
    key = "OO_LR" if side_id in (1, 3) else "OO_UD"
    gs = System.newTrivialWithSimpleSparseOperator(**{key: [HipData.fromArray(o) for o in ops]})
    os_ = OSystem.newTrivialWithSimpleSparseOperator(**{key: ops})
    first = R(side_id) if direction == 0 else L(side_id)
    for s in (gs, os_):
        s.contractTowards(first)
    return gs, os_, side_id


@pytest.mark.parametrize("corner_id,direction", [(0, 0), (1, 1), (2, 0), (3, 1)])
def test_operator_compression_exact_and_matches_oracle(gpu, corner_id, direction):
    from carcassonne_amd.sparse import TwoSiteOperatorCompressed
    gs, os_, side_id = _operator_systems(corner_id, direction, 20 + corner_id)
    for _ in range(4):
        gs.contractTowards(side_id)
        os_.contractTowards(side_id)
    e1, n1 = gs.computeExpectationAndNormalization()
    gs.compressCornerTwoSiteOperatorTowards(corner_id, direction, 4)
    os_.compressCornerTwoSiteOperatorTowards(corner_id, direction, 4)
    assert TwoSiteOperatorCompressed(direction) in gs.corners[corner_id]
    assert TwoSiteOperatorCompressed(1 - direction) in gs.sides[side_id]
    e2, n2 = gs.computeExpectationAndNormalization()
    oe, on = os_.computeExpectationAndNormalization()
    assert abs(n2 / n1 - 1) < 1e-10 and abs(e2 - e1) < 1e-10 * max(1, abs(e1))
    assert abs(e2 - oe) < 1e-10 * max(1, abs(oe))
    # the compressed channels keep working in further growth steps
    gs.contractTowards(side_id)
    os_.contractTowards(side_id)
    assert abs(gs.computeExpectation() - os_.computeExpectation()) < 1e-10 * max(1, abs(os_.computeExpectation()))


def test_operator_compression_below_the_term_count(gpu):
    """More two-site terms than the channel rank: compressing to the rank is exact (reference
    test_system.py:82-115)."""
    gs, os_, side_id = _operator_systems(0, 1, 7)
    for _ in range(2):
        gs.contractTowards(side_id)
    state = gs.state_center_data
    for _ in range(5):
        gs.contractTowards(side_id)
    gs.setStateCenter(state)
    e1, n1 = gs.computeExpectationAndNormalization()
    gs.compressCornerTwoSiteOperatorTowards(0, 1, 2)
    e2, n2 = gs.computeExpectationAndNormalization()
    assert abs(n2 / n1 - 1) < 1e-10 and abs(e2 - e1) < 1e-9 * max(1, abs(e1))


@pytest.mark.parametrize("walk,axes", [((0, 2, 1, 3), (0,)), ((1, 0, 3, 2, 1), (1,)), ((0, 1, 2), (0, 1))])
def test_full_eigensolver_matches_oracle_eigh(gpu, walk, axes):
    """minimizeExpectationUsingFullEigensolver (system/_2d.py:498-502) on transverse-field Ising systems
    grown by bandwidth increases and contractions on the oracle's side, then uploaded: centers
    (2,1,2,1,2) and (1,2,1,2,2) have no fused kernel (the dense generalized eigensolve runs), (2,2,2,2,2)
    has one (Lanczos runs).  Both give the oracle's eigh(H, N) energy; so does a trivial D=1 system."""
    from carcassonne_amd.data import HipData
    from carcassonne_amd.system import System
    from oracle.system import System as OSystem
    from tests.oracle_bridge import system_to_product
    Z = np.diag([1.0, -1.0])
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    osys = OSystem.newTrivialWithSimpleSparseOperator(O=-0.7 * X + 0.2 * Z, OO_LR=[Z, -Z], OO_UD=[Z, -Z],
                                                      rng=np.random.default_rng(3))
    for a in axes:
        osys.increaseBandwidth(a, by=1, do_as_much_as_possible=True)
        osys.setStateCenter(osys.state)
    for d in walk:
        osys.contractTowards(d)
    psys = system_to_product(osys)
    H, _ = psys.formExpectationAndNormalizationMultipliers()
    print(walk, axes, "center", tuple(psys.state_center_data.shape), "fused", H.operator.use_fused, flush=True)
    assert H.operator.use_fused == (len(axes) == 2)
    for s in (psys, osys):
        s.minimizeExpectationUsingFullEigensolver()
    e_p, e_o = psys.computeExpectation(), osys.computeExpectation()
    assert abs(e_p - e_o) <= 1e-10 * abs(e_o), (e_p, e_o)
    osys = OSystem.newTrivialWithSimpleSparseOperator(O=-0.7 * X + 0.2 * Z, OO_LR=[Z, -Z], OO_UD=[Z, -Z])
    psys = System.newTrivialWithSimpleSparseOperator(O=HipData.fromArray(-0.7 * X + 0.2 * Z),
                                                     OO_LR=[HipData.Z, -HipData.Z], OO_UD=[HipData.Z, -HipData.Z])
    for s in (psys, osys):
        s.minimizeExpectationUsingFullEigensolver()
    e_p, e_o = psys.computeExpectation(), osys.computeExpectation()
    assert abs(e_p - e_o) <= 1e-12 * max(1.0, abs(e_o)), (e_p, e_o)


def test_gauge_moves_match_oracle(gpu):
    """normalizeCornerAndDenormalizeSide / normalizeSideAndDenormalizeCenter / normalizeCenterAndDenormalizeSide
    and the full normalize() (system/_2d.py:503-549) on the GPU and in the oracle from the same tensors.
    The normalizer is diag(S^-1/2) V* of the SVD M = U S V (data/__init__.py:263-301), so each move is fixed up
    to the sign of every right singular vector; a sign flips index a of the ket leg and of the bra leg
    together in the normalized tensor and in its neighbour, which leaves every entry's magnitude unchanged.
    The magnitudes of every corner, side and center entry are compared, besides the invariant energy and
    normalization."""
    from tests.oracle_bridge import product_to_oracle_tag, rel_err

    def compare_systems(psys, osys):
        worst = 0.0
        for pl, ol in ((psys.corners, osys.corners), (psys.sides, osys.sides)):
            for pd, od in zip(pl, ol):
                for t, v in pd.items():
                    worst = max(worst, rel_err(np.abs(v.toArray()), np.abs(od[product_to_oracle_tag(t)])))
        return max(worst, rel_err(np.abs(psys.state_center_data.toArray()), np.abs(osys.state)))

    psys, osys = _pair(5, maximum_dimension=3)
    for s in (psys, osys):
        s.contractTowards(0)
        s.contractTowards(1)
    for corner_id, direction in ((2, 1), (0, 0), (3, 1)):
        psys.normalizeCornerAndDenormalizeSide(corner_id, direction)
        osys.normalizeCornerAndDenormalizeSide(corner_id, direction)
        assert compare_systems(psys, osys) < 1e-9, ("corner", corner_id, direction)
    for side_id in (3, 1):
        psys.normalizeSideAndDenormalizeCenter(side_id)
        osys.normalizeSideAndDenormalizeCenter(side_id)
        assert compare_systems(psys, osys) < 1e-9, ("side", side_id)
    psys.normalizeCenterAndDenormalizeSide(0)
    osys.normalizeCenterAndDenormalizeSide(0)
    assert compare_systems(psys, osys) < 1e-9
    psys.normalize()
    osys.normalize()
    assert compare_systems(psys, osys) < 1e-8
    e_p, n_p = psys.computeExpectationAndNormalization()
    e_o, n_o = osys.computeExpectationAndNormalization()
    assert abs(e_p - e_o) <= 1e-9 * abs(e_o) and abs(n_p - n_o) <= 1e-9 * abs(n_o)


@pytest.mark.parametrize("seed,new", [(8, 1), (9, 2), (12, 2)])
def test_state_compression_matches_oracle(gpu, seed, new):
    """compressCornerStateTowards (system/_2d.py:191-228, compression.py:26-45) with the deterministic start on
    both sides: the compressors agree with the oracle's, and so do the compressed corners and sides, the
    gauge-invariant product of each compressed corner with its side over the bond (L P R), the energy and
    the normalization."""
    import numpy as np
    from carcassonne_amd.sparse import Identity
    from oracle.tags import Identity as OIdentity
    from tests.oracle_bridge import compare_systems
    psys, osys = _pair(seed, maximum_dimension=2)
    for s in (psys, osys):
        s.compressor_start = "svd"
        for d in (0, 1, 2, 3):
            s.contractTowards(d)
    for corner_id in range(4):
        for direction in range(2):
            dim = osys.corners[corner_id][OIdentity()].shape[3 * direction]
            target = max(1, min(new, dim - 1))
            psys.compressCornerStateTowards(corner_id, direction, target)
            osys.compressCornerStateTowards(corner_id, direction, target)
    assert compare_systems(psys, osys) < 1e-7
    for corner_id in range(4):
        pc, oc = psys.corners[corner_id][Identity()].toArray(), osys.corners[corner_id][OIdentity()]
        ps, os_ = psys.sides[corner_id][Identity()].toArray(), osys.sides[corner_id][OIdentity()]
        lpr_p = np.tensordot(pc, ps, ((3, 4, 5), (0, 1, 2)))
        lpr_o = np.tensordot(oc, os_, ((3, 4, 5), (0, 1, 2)))
        assert np.linalg.norm(lpr_p - lpr_o) <= 1e-7 * np.linalg.norm(lpr_o)
    e_p, n_p = psys.computeExpectationAndNormalization()
    e_o, n_o = osys.computeExpectationAndNormalization()
    assert abs(e_p - e_o) <= 1e-8 * abs(e_o) and abs(n_p - n_o) <= 1e-8 * abs(n_o)
