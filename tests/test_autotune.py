"""The autotuner's coordinate descent (fcl_amd.autotune), on a stand-in
context whose batch time is a known function of its settings, read off a
fake clock (no GPU)."""
import pytest

from fcl_amd import autotune


class FakeCtx:
    def __init__(self, params):
        self.params = dict(params)

    def get_engine_param(self, name):
        return float(self.params[name])

    def set_engine_param(self, name, value):
        self.params[name] = float(value)


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


@pytest.fixture
def clock(monkeypatch):
    c = FakeClock()
    monkeypatch.setattr(autotune.time, "perf_counter", c)
    return c


SPACE = (("a", (1, 2, 3, 4)), ("b", (10, 20, 30)))


def test_descent_finds_the_minimum_of_a_separable_cost(clock):
    ctx = FakeCtx({"a": 1, "b": 10})

    def run():  # cost in seconds, minimal at a=3, b=20
        clock.t += 1.0 + (ctx.params["a"] - 3) ** 2 + 0.5 * abs(ctx.params["b"] - 20) / 10
        return 42

    best, best_ms, trials = autotune.tune(ctx, run, SPACE, repeats=2)
    assert best == {"a": 3.0, "b": 20.0}
    assert ctx.params == {"a": 3.0, "b": 20.0}  # the best stays applied
    assert best_ms == pytest.approx(1000.0)
    # the start, 3 other values of a, 2 other values of b
    assert len(trials) == 1 + 3 + 2


def test_a_change_within_the_margin_is_not_taken(clock):
    ctx = FakeCtx({"a": 1, "b": 10})

    def run():  # a=2 is 1 % faster: below the 2 % margin
        clock.t += 1.0 - (0.01 if ctx.params["a"] == 2 else 0.0)

    best, _, _ = autotune.tune(ctx, run, SPACE, repeats=1)
    assert best == {"a": 1.0, "b": 10.0}


def test_a_setting_that_changes_the_result_is_an_error(clock):
    ctx = FakeCtx({"a": 1, "b": 10})

    def run():
        clock.t += 1.0
        return 7 if ctx.params["a"] != 4 else 8

    with pytest.raises(autotune.ResultChanged):
        autotune.tune(ctx, run, SPACE)


def test_spaces_name_engine_tunables():
    spaces = autotune.FAST_SPACE + autotune.BOOL_SPACE + autotune.DIST_SPACE
    names = [n for n, _ in spaces]
    assert len(names) == len(set(names))
    assert all(len(v) >= 2 for _, v in spaces)
