"""The rollout controller's schedule (lbm_amd/rollout.py) on the CPU, with a
stand-in grid that records the calls and returns scripted check scores: which
cycles are checks, which jump and which fall back, the interval's back-off
(doubled after a check passed with at most half the threshold, kept after one
passed closer to it, reset after a fallback) and the step count.  No
reference counterpart: the reference has no surrogate."""
import pytest

from lbm_amd.rollout import Rollout


class ScriptedGrid:
    """Records what Rollout.run asks of a grid; velocity_skill returns the next scripted score."""

    def __init__(self, scores):
        self.scores = list(scores)
        self.log = []

    def velocity_mark(self, slot=0):
        self.log.append(("mark", slot))

    def snapshot(self):
        self.log.append(("snapshot",))

    def restore(self):
        self.log.append(("restore",))

    def surrogate_jump(self, net, jump, *scales):
        self.log.append(("jump", jump))

    def step(self, n):
        self.log.append(("step", n))

    def velocity_skill(self):
        return self.scores.pop(0), 0.004


def cycle_kinds(g):
    """Per cycle: 'check' (snapshot .. restore, then the plain cycle), 'jump' or 'plain'."""
    kinds, i, log = [], 0, g.log
    while i < len(log):
        if log[i] == ("mark", 0):
            kinds.append("check")
            i += 7  # mark, snapshot, jump, correctors, mark, restore, plain cycle
        elif log[i][0] == "jump":
            kinds.append("jump")
            i += 2
        else:
            kinds.append("plain")
            i += 1
    return kinds


def test_fixed_interval_checks_every_kth_cycle():
    g = ScriptedGrid([0.1, 0.1, 0.1])
    st = Rollout(g, None, check_every=4).run(72 * 10 + 5)
    assert cycle_kinds(g)[:10] == ["check", "jump", "jump", "jump"] * 2 + ["check", "jump"]
    assert (st.checks, st.jumps, st.fallbacks, st.cycles) == (3, 7, 0, 10)
    assert st.steps == 72 * 10 + 5 and g.log[-1] == ("step", 5)  # the rest runs plain


def test_backoff_doubles_after_clear_passes_and_keeps_near_the_threshold():
    # checks at 0 (0.1: 2 -> 4), 4 (0.2: -> 8), 12 (0.3, within the threshold but above half: stays 8),
    # 20 (0.1: -> capped at 8), 28
    g = ScriptedGrid([0.1, 0.2, 0.3, 0.1, 0.1])
    st = Rollout(g, None, check_every=2, max_check_every=8).run(72 * 30)
    kinds = cycle_kinds(g)
    assert [c for c, k in enumerate(kinds) if k == "check"] == [0, 4, 12, 20, 28]
    assert (st.checks, st.jumps, st.fallbacks) == (5, 25, 0)


def test_fallback_resets_the_interval_and_runs_plain_until_the_next_check():
    # 0: pass (2 -> 4); 4: fail (-> 2, cycles 5 plain); 6: pass near the threshold (stays 2); 8: pass (-> 4); 12
    g = ScriptedGrid([0.1, 0.9, 0.4, 0.2, 0.1])
    st = Rollout(g, None, check_every=2, max_check_every=16).run(72 * 14)
    kinds = cycle_kinds(g)
    assert kinds == ["check", "jump", "jump", "jump", "check", "plain", "check", "jump", "check", "jump", "jump",
                     "jump", "check", "jump"]
    assert (st.checks, st.fallbacks, st.jumps) == (5, 2, 8)
    assert st.residuals == [0.1, 0.9, 0.4, 0.2, 0.1]


def test_no_checks_and_bad_intervals():
    g = ScriptedGrid([])
    st = Rollout(g, None, check_every=0).run(72 * 3)
    assert cycle_kinds(g) == ["jump"] * 3 and st.checks == 0
    with pytest.raises(ValueError):
        Rollout(g, None, check_every=-1)
    with pytest.raises(ValueError):
        Rollout(g, None, check_every=8, max_check_every=4)
