"""Queued exact blocks (ftlm_exact_queue_*): blocks of several sectors and
twists solved in one batched launch give, bit for bit, the eigenvalues of the
per-sector batches (ftlm_exact_blocks) -- the same dense matrices and the same
per-block eigensolver -- and a budget that fits one block at a time forces
one run per block with the same results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODEL = dict(lx=3, ly=3, tp=-0.25, u=6.0, v=1.0)
TWISTS = [(0.3, 0.7), (0.0, 0.5)]
CASES = [(2, 1), (3, 3), (4, 3), (1, 1), (5, 4)]


def say(*a):
    print(*a, flush=True)


def enumerate_blocks(eng):
    out = []
    for tw in TWISTS:
        eng.set_twist(*tw)
        for nup, ndn in CASES:
            eng.sector(nup, ndn)
            for my in range(3):
                for mx in range(3):
                    d = eng.block(mx, my)
                    if 0 < d <= 256:
                        out.append((tw, nup, ndn, mx, my, d))
    return out


def test_queue_equals_per_sector_batches():
    import ftlm_extended_hubbard_development_amd as pkg

    eng = pkg.Engine(**MODEL)
    blocks = enumerate_blocks(eng)
    say(f"{len(blocks)} exact blocks, dims {min(b[5] for b in blocks)}..{max(b[5] for b in blocks)}")
    ref = {}
    for tw in TWISTS:
        eng.set_twist(*tw)
        for nup, ndn in CASES:
            eng.sector(nup, ndn)
            ks = [(b[3], b[4]) for b in blocks if b[0] == tw and b[1:3] == (nup, ndn)]
            for k, e in zip(ks, eng.exact_blocks(ks)):
                ref[(tw, nup, ndn) + k] = e
    eng.exact_queue_begin()
    idx = {}
    for tw, nup, ndn, mx, my, d in blocks:
        eng.set_twist(*tw)
        if (eng.nup, eng.ndn) != (nup, ndn):
            eng.sector(nup, ndn)
        idx[(tw, nup, ndn, mx, my)] = eng.exact_queue_add(mx, my)
    eng.exact_queue_run()
    for key, i in idx.items():
        assert np.array_equal(eng.exact_queue_result(i), ref[key]), key
    # a budget of one matrix: every add beyond the first reports a full queue
    eng.exact_queue_begin(budget_bytes=16 * 256 * 256)
    n_runs = 0
    for tw, nup, ndn, mx, my, d in blocks[:12]:
        eng.set_twist(*tw)
        eng.sector(nup, ndn)
        i = eng.exact_queue_add(mx, my)
        if i is None:
            eng.exact_queue_run()
            n_runs += 1
            eng.exact_queue_begin(budget_bytes=16 * 256 * 256)
            i = eng.exact_queue_add(mx, my)
        assert i is not None
        eng.exact_queue_run()
        assert np.array_equal(eng.exact_queue_result(i), ref[(tw, nup, ndn, mx, my)])
    eng.close()
