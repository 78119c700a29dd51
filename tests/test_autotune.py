"""The autotuner's selection (autotune.tune) against a synthetic cost model
whose step time drifts upwards while it runs, as the real step cost does
(omega_max only grows): the tuner must still find the fastest setting, where
timing the candidates one after another would prefer the early ones."""
import autotune


def _model(costs, drift=0.02):
    """apply/time_block over a table of costs per (knob, value); every block
    costs a little more than the one before."""
    state = {"s": None, "n": 0}

    def apply(s):
        state["s"] = dict(s)

    def time_block():
        state["n"] += 1
        c = 1.0
        for k, v in state["s"].items():
            c *= costs[k][v]
        return c * (1.0 + drift * state["n"])
    return apply, time_block


def test_tuner_finds_the_fastest_setting_under_drift():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30]}
    costs = {"a": {1: 1.10, 2: 1.00, 3: 0.95, 4: 1.30}, "b": {10: 1.0, 20: 0.90, 30: 0.97}}
    apply, time_block = _model(costs)
    best, trials = autotune.tune(space, {"a": 1, "b": 10}, apply, time_block)
    assert best == {"a": 3, "b": 20}
    assert len(trials) == 5


def test_tuner_keeps_the_start_within_the_margin():
    space = {"a": [1, 2]}
    costs = {"a": {1: 1.0, 2: 0.995}}
    apply, time_block = _model(costs, drift=0.0)
    best, _ = autotune.tune(space, {"a": 1}, apply, time_block, margin=0.01)
    assert best == {"a": 1}


def test_sequential_timing_would_be_fooled_by_the_drift():
    """Why the alternating blocks: timed one after another, a slower early
    candidate beats a faster late one."""
    costs = {"a": {1: 1.00, 2: 0.97}}
    apply, time_block = _model(costs, drift=0.05)
    apply({"a": 1})
    t1 = time_block()
    apply({"a": 2})
    t2 = time_block()
    assert t2 > t1  # the faster setting looks slower
    best, _ = autotune.tune({"a": [1, 2]}, {"a": 1}, *_model(costs, drift=0.05))
    assert best == {"a": 2}
