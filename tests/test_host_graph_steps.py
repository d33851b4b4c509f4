"""Host logic of RBCBatch.step: n steps split into replays of the longest
captured graphs that fit, then single launches (no GPU: the graphs and the
step are stand-ins that record what ran)."""
from pypde_amd.engine import RBCBatch


class _Graph:
    def __init__(self, log, length):
        self.log, self.length = log, length

    def replay(self):
        self.log.append(self.length)


class _Engine:
    def __init__(self, lengths):
        self.log = []
        self.graphs = {L: _Graph(self.log, L) for L in lengths}

    def step_eager(self):
        self.log.append(0)


def run(lengths, n):
    e = _Engine(lengths)
    RBCBatch.step(e, n)
    return e.log


def test_longest_graphs_first_then_single_steps():
    assert run((1, 10), 23) == [10, 10, 1, 1, 1]
    assert run((10,), 23) == [10, 10, 0, 0, 0]
    assert run((1, 10), 7) == [1] * 7
    assert run((1, 5, 10), 19) == [10, 5, 1, 1, 1, 1]


def test_no_graphs_runs_launches():
    assert run((), 3) == [0, 0, 0]
    assert run((1, 10), 0) == []


def test_step_count_is_kept():
    for n in range(0, 40):
        log = run((1, 10), n)
        assert sum(L if L else 1 for L in log) == n
