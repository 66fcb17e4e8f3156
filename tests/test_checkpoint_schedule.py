"""Binomial checkpointing schedule of the batched gradient (revolve): every
segment is reversed once, last first, from the state it started from, with
no more slots than given, and the number of segment advances equals the
closed form of binomial checkpointing (Griewank and Walther, ACM TOMS 26,
2000): with S slots, t segments and r the least repetition number with
C(S + r, S) >= t, r t - C(S + r, S + 1)."""
from math import comb

import pytest

from gpuocean_amd.adjoint import binomial_schedule, schedule_split


def _closed_form(t, S):
    r = 0
    while comb(S + r, S) < t:
        r += 1
    return r * t - comb(S + r, S + 1)


def _simulate(ops, n_seg, n_slots):
    """Replays the actions on symbolic states: slots hold segment indices,
    the forward run a position; returns the turn order and advance count."""
    slots = {0: 0}
    at, turns, advances = 0, [], 0
    for op in ops:
        if op[0] == "load":
            assert slots.get(op[1]) == op[2], op
            at = op[2]
        elif op[0] == "advance":
            assert at == op[1] and op[2] > op[1], op
            advances += op[2] - op[1]
            at = op[2]
        elif op[0] == "save":
            assert at == op[2] and 0 < op[1] < n_slots, op
            slots[op[1]] = op[2]
        else:
            assert op[0] == "turn" and at == op[1], op
            turns.append(op[1])
    return turns, advances


@pytest.mark.parametrize("n_seg", [1, 2, 3, 5, 8, 15, 21, 40])
@pytest.mark.parametrize("n_slots", [1, 2, 3, 4, 6])
def test_schedule_reverses_every_segment_with_the_fewest_advances(n_seg, n_slots):
    ops = binomial_schedule(n_seg, n_slots)
    turns, advances = _simulate(ops, n_seg, n_slots)
    assert turns == list(reversed(range(n_seg)))
    assert advances == _closed_form(n_seg, min(n_slots, n_seg))


def test_enough_slots_advance_each_segment_once():
    ops = binomial_schedule(15, 15)
    assert _simulate(ops, 15, 15)[1] == 14
    assert [op for op in ops if op[0] == "save"] == [("save", i, i) for i in range(1, 15)]


def test_a_planted_mistake_fails_the_replay():
    ops = binomial_schedule(8, 3)
    i = next(i for i, op in enumerate(ops) if op[0] == "save")
    bad = ops[:i] + [("save", ops[i][1], ops[i][2] + 1)] + ops[i + 1:]
    with pytest.raises(AssertionError):
        _simulate(bad, 8, 3)


@pytest.mark.parametrize("n_seg,n_slots", [(2, 1), (6, 3), (15, 4), (15, 2), (40, 5)])
def test_forward_share_then_sweep(n_seg, n_slots):
    """The forward run executes the schedule up to schedule_split, then steps
    to the end; the sweep's first action must bring it back by a load."""
    ops = binomial_schedule(n_seg, n_slots)
    j = schedule_split(ops)
    assert all(op[0] != "turn" for op in ops[:j]) and ops[j][0] == "load"
    _simulate(ops[:j], n_seg, n_slots)
    turns, _ = _simulate(ops, n_seg, n_slots)
    assert turns == list(reversed(range(n_seg)))
