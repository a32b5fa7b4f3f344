"""Placement tuning of the Lanczos vectors (ftlm_set_vector_placement): the
engine times the batched hop-list apply from several buffers and keeps the
fastest as the gather sources. Which buffers hold the vectors must not change
any result: the tree-order reductions are fixed by the launch shape, not by
addresses, so a tuned and an untuned run give the same bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M44 = dict(lx=4, ly=4, u=7.0)


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pkg():
    import ftlm_extended_hubbard_development_amd as p

    p.load_library()
    return p


def run(pkg, candidates, case=(4, 4, 0, 0), R=64, steps=30):
    eng = pkg.Engine(**M44)
    try:
        eng.set_vector_placement(candidates, 1)
        eng.set_start_source(pkg.START_DEVICE)
        eng.sector(case[0], case[1])
        eng.block(case[2], case[3])
        e, w = eng.ftlm_block(R, steps, pkg.block_seed(12345, *case), reduce=pkg.REDUCE_TREE)
        # a second block run on the same (already tuned) buffers
        e2, w2 = eng.ftlm_block(R, steps, pkg.block_seed(777, *case), reduce=pkg.REDUCE_TREE)
        return e, w, e2, w2, eng.vector_placement_report()
    finally:
        eng.close()


def test_placement_does_not_change_results(pkg):
    say("case 4x4 (4,4,0,0) R=64 steps=30: placement off / 6 candidates / 3 candidates")
    base = run(pkg, 0)
    assert len(base[4][0]) == 0  # off: no tuning ran
    for cand in (6, 3):
        got = run(pkg, cand)
        ms, pick = got[4]
        say(f"  candidates {cand}: apply ms {np.round(ms, 3).tolist()} -> {pick}")
        assert 3 <= len(ms) <= cand and np.all(ms > 0)
        assert sorted(set(pick)) == sorted(pick) and len(pick) == 3
        assert ms[pick[0]] <= ms[pick[1]] <= min(np.delete(ms, pick[:2]))
        for a, b in zip(base[:4], got[:4]):
            assert np.array_equal(a, b)


def test_placement_arguments(pkg):
    eng = pkg.Engine(**M44)
    try:
        for bad in (1, 2, 17, -1):
            with pytest.raises(pkg.FtlmError):
                eng.set_vector_placement(bad, 0)
        eng.set_vector_placement(0, 0)
        eng.set_vector_placement(16, 1 << 30)
    finally:
        eng.close()
