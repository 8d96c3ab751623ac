"""Host-side logic of the autotuner (no GPU): the coordinate descent over a fake context
whose step time depends on the variant and on how many steps ran since the last reset,
as a train step's does (the first steps from the initialisation get cheaper)."""
from msplat import autotune


class FakeCtx:
    def __init__(self):
        self.tuning = {key: values[0] for key, values in autotune.SPACE.values()}
        self.clock = 0.0
        self.since_reset = 0

    def get_tuning(self, key):
        return self.tuning[key]

    def set_tuning(self, key, value):
        self.tuning[key] = value

    def sync(self):
        pass


# per-variant cost: 4 pixels per lane is 3% faster than 2, which is 10% faster than 1;
# every other tunable costs the same at each value
COST = {4: 1.00, 2: 1.03, 1: 1.13}


def _step(ctx):
    ppt = ctx.tuning[autotune.SPACE["bwd_ppt"][0]]
    # a fresh scene: the first steps are slower (50% at step 0, fading over 20 steps)
    drift = 1.0 + 0.5 * max(0.0, 1.0 - ctx.since_reset / 20.0)
    ctx.clock += COST[ppt] * drift * 1e-3
    ctx.since_reset += 1


def _patched(monkeypatch, ctx):
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: ctx.clock)


def test_reset_gives_every_trial_the_same_start(monkeypatch):
    ctx = FakeCtx()
    _patched(monkeypatch, ctx)
    resets = []

    def reset():
        resets.append(ctx.since_reset)
        ctx.since_reset = 0

    res = autotune.autotune(ctx, lambda: _step(ctx), warmup=2, repeats=8, reset=reset)
    # the discarded first trial, the baseline, then every other value of every tunable
    n_trials = 1 + sum(len(values) - 1 for _, values in autotune.SPACE.values())
    assert len(res["trials"]) == n_trials and len(resets) == n_trials + 1
    assert res["bwd_ppt"] == 4
    times = [ms for choice, ms in res["trials"] if choice["bwd_ppt"] == 4]
    assert max(times) - min(times) < 1e-9  # equal variants time equal from equal starts


def test_without_reset_the_drift_favours_later_trials(monkeypatch):
    """What the reset is for: timed back to back, the scene's own speed-up outweighs the
    3% between the variants, and the tuner keeps the slower one tried after the default."""
    ctx = FakeCtx()
    _patched(monkeypatch, ctx)
    res = autotune.autotune(ctx, lambda: _step(ctx), warmup=0, repeats=4)
    assert res["bwd_ppt"] == 2
