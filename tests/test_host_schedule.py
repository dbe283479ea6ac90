"""Host logic of the fill's LDS schedule (lcg_pair_schedule, the function
lcg_batch_create uses): every group of a row is placed once, a round of a
workgroup member holds at most `waves` groups on distinct waves, their LDS
regions are disjoint and inside the pool, and the row's round count covers
them.  Runs on the CPU."""
import collections

import pytest

from locarna_amd import api, synth


def _pairs():
    pps, pairs, span = synth.workload("allpairs64", limit_pairs=6)
    out = [api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span)) for a, b in pairs]
    pps, pairs, span = synth.workload("long500", limit_pairs=1)
    out += [api.Pair(pps[a], pps[b], api.locarna_params(max_bp_span=span, sequ_local=1)) for a, b in pairs]
    return out


@pytest.mark.parametrize("waves,pool,members", [(16, 40000, 1), (8, 20000, 1), (16, 40000, 4), (16, 40000, 64)])
def test_schedule_invariants(waves, pool, members):
    for P in _pairs():
        sched, rounds = P.schedule(waves, pool, members)
        assert len(sched) == P.num_groups
        by_slot = collections.defaultdict(list)
        for (rnd, wave, member, off, size, al, bl) in sched:
            assert 0 <= wave < waves and 0 <= member < members
            assert rnd < rounds[al]
            assert 0 <= off and off + size <= pool
            by_slot[(al, member, rnd)].append((wave, off, size))
        for key, items in by_slot.items():
            assert len(items) <= waves
            assert len({w for w, _, _ in items}) == len(items), key
            spans = sorted((o, o + s) for _, o, s in items)
            for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
                assert a1 <= b0, (key, spans)
        rows = {al for (_, _, _, _, _, al, _) in sched}
        for al in range(1, P.lenA + 1):
            if al not in rows:
                assert rounds[al] == 0


def test_schedule_rejects_small_pool():
    P = _pairs()[0]
    with pytest.raises(api.LcgError):
        P.schedule(16, 64, 1)
