"""Host logic of the autotuner, on a stand-in graph whose kernel times are a
function of the launch parameters (no GPU)."""
import pytest

from everest_amd import abi, autotune


class FakeGraph:
    """count() returns a fixed count per query; last_kernel_ms() a cost
    model: base time per query scaled by per-field factors."""

    def __init__(self, factors, wrong=None):
        self.factors = factors
        self.wrong = wrong
        self.runs = 0

    def count(self, m, d, tune=None):
        t = abi.tuning_dict(tune)
        self.runs += 1
        ms = float(d)
        for f, v in t.items():
            ms *= self.factors.get((f, v), 1.0)
        lim = t.get("time_limit_ms", 0)
        if lim and ms > lim:
            raise abi.TMError(abi.TM_ERR_TIMEOUT, "time limit")
        self._ms = ms
        if self.wrong and self.wrong in t.items():
            return d + 1
        return d * 10

    def last_kernel_ms(self):
        return self._ms


QS = [([(0, 1, 0), (1, 2, 1), (2, 0, 2)], 100), ([(0, 1, 0), (0, 2, 1), (2, 3, 2)], 40),
      ([(0, 1, 0), (1, 2, 1), (2, 3, 2), (3, 0, 3)], 300), ([(0, 1, 0), (0, 2, 1), (0, 3, 2), (0, 1, 3)], 20)]


def test_descent_keeps_winning_values_and_stops_losers():
    fg = FakeGraph({("tail_swap", 4): 0.5, ("waves_per_simd_hint", 6): 3.0, ("claim_blocks", 8): 0.9,
                    ("claim_blocks", 2): 0.99})
    table, meas, screen = autotune.autotune(fg, QS, per_class=2)
    assert len(screen) == len(QS)
    for cls, t in table.items():
        assert t == {"tail_swap": 4, "claim_blocks": 8}, (cls, t)
        # the 3x slower register cap was abandoned after its heaviest query
        rows = dict((tuple(sorted(c.items())), ms) for c, ms in meas[cls])
        assert rows[(("waves_per_simd_hint", 6),)] is None
        # a 1% gain is within the 2% tie band: not taken
        assert rows[(("claim_blocks", 2), ("tail_swap", 4))] is not None


def test_count_change_is_an_error():
    fg = FakeGraph({("tail_swap", 2): 0.5}, wrong=("tail_swap", 2))
    with pytest.raises(RuntimeError, match="differs"):
        autotune.autotune(fg, QS, per_class=2)


def test_start_table_and_second_pass():
    # a start entry better than the defaults is where the descent begins,
    # and the descent still adds the values that pay on top of it
    fg = FakeGraph({("tail_swap", 4): 0.5, ("claim_blocks", 2): 0.7})
    start = {autotune.abi.shape_class(m): {"claim_blocks": 2} for m, _ in QS}
    table, meas, _ = autotune.autotune(fg, QS, per_class=2, start=start)
    for cls, t in table.items():
        assert t == {"tail_swap": 4, "claim_blocks": 2}, (cls, t)
        assert meas[cls][1][0] == {"claim_blocks": 2}
