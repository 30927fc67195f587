"""The batched form: one HIP graph holding the eigensolves of several systems of different shapes.
The graph is replayed several times with eager work in between; every replay reproduces the eager
fixed-work solves exactly and matches the plain FP64 energies."""
import pytest

pytestmark = pytest.mark.gpu


def test_graph_batch_matches_plain_fp64(gpu):
    from carcassonne_amd.batched import BatchedSolve
    from carcassonne_amd.device import synchronize
    specs = [(2, 6, 2.5, 1), (3, 5, 3.1, 2), (2, 8, 2.2, 3), (3, 6, 3.7, 4), (4, 4, 2.9, 5)]
    b = BatchedSolve(specs, dtype="f16")
    b.prepare()
    b.tune_work()
    ref = b.reference_energies()
    eager = [float(s["solver"].solve_fixed(s["x0"], lp_iters=b.tuning[s["key"]]["lp_iters"],
                                           outer=b.tuning[s["key"]]["outer"])[1].item()) for s in b.systems]
    b.capture()
    for rep in range(3):
        b.replay()
        got = b.energies()
        assert got == eager, (rep, got, eager)
        assert len(b.convergence()) == len(specs)
        for g, r in zip(got, ref):
            assert abs(g - r) <= 1e-10 * abs(r), (g, r)
        b.reference_energies()          # eager work between replays
        synchronize()
