"""bench.py's warmup on the CPU, with a stand-in simulation that records the
calls the bench makes: the warmup steps run in blocks of --warmup steps
until --settle-ms of device time have passed, the initial state is then made
again and sorted, and the timed run of exactly --steps steps follows."""
import importlib.util
import os
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_warmup_under_test",
                                                  os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)  # main() runs only as a script
    return mod


class _Ctx:
    def __init__(self, log):
        self.log = log

    def sync(self):
        self.log.append(("sync",))


class FakeSim:
    """Records calls; run() takes step_ms of wall time per step, as pic_run
    takes device time and synchronises before it returns."""

    def __init__(self, step_ms):
        self.log = []
        self.ctx = _Ctx(self.log)
        self.step_ms = step_ms

    def __getattr__(self, name):  # the set_* calls the bench makes
        if name.startswith("set_"):
            return lambda *a, **k: self.log.append((name,) + a)
        raise AttributeError(name)

    def sort(self, lanes=False):
        self.log.append(("sort", lanes))

    def run(self, dt, nsteps, sort_every=0):
        self.log.append(("run", nsteps))
        time.sleep(nsteps * self.step_ms * 1e-3)
        return np.zeros((nsteps, 2))

    def sample_landau(self, n, alpha, kx, state):
        self.log.append(("reload",))

    def slow_chunks(self):
        return 0

    def timing(self):
        return 1.0, 1, 0.5

    def step_pass(self, mode, dt):
        self.log.append(("step_pass", mode))

    def grid(self, i):
        return np.zeros((4, 4))

    def close(self):
        self.log.append(("close",))


def _run(bench, monkeypatch, warmup, settle_ms, step_ms=1.0, steps=3):
    sim = FakeSim(step_ms)
    monkeypatch.setattr(bench, "_new_sim", lambda *a, **k: sim)
    p0 = bench.DeviceLandau(1000, 0.01, 0.5)
    out = bench.run_path(None, None, p0, None, 1, 0.05, steps, warmup, 12, settle_ms=settle_ms)
    return sim.log, out


def _warm_and_timed(log):
    reload = log.index(("reload",))
    warm = [e[1] for e in log[:reload] if e[0] == "run"]
    after = log[reload:]
    timed = next(e[1] for e in after if e[0] == "run")
    return warm, after, timed


@pytest.mark.parametrize("warmup,settle_ms,step_ms", [(5, 50.0, 1.0), (2, 20.0, 1.5)])
def test_warmup_runs_blocks_until_settled(bench, monkeypatch, warmup, settle_ms, step_ms):
    log, out = _run(bench, monkeypatch, warmup, settle_ms, step_ms)
    warm, after, timed = _warm_and_timed(log)
    assert warm and all(k == warmup for k in warm)
    # blocks until settle_ms had passed: at least that much work ran, and the
    # last block was needed (the work before it was shorter than settle_ms)
    assert sum(warm) * step_ms >= settle_ms
    assert (sum(warm) - warmup) * step_ms < settle_ms * 1.5 + 10
    assert out["warmup_steps"] == sum(warm)
    # then: the initial state made again, lane-sorted, and the timed run of
    # exactly `steps` steps
    assert after[1] == ("sort", True)
    assert timed == 3


def test_settle_zero_runs_the_warmup_steps_only(bench, monkeypatch):
    log, out = _run(bench, monkeypatch, 5, 0.0)
    warm, _, timed = _warm_and_timed(log)
    assert warm == [5] and out["warmup_steps"] == 5 and timed == 3


def test_no_warmup_runs_no_settle(bench, monkeypatch):
    log, out = _run(bench, monkeypatch, 0, 50.0)
    warm, _, timed = _warm_and_timed(log)
    assert warm == [] and out["warmup_steps"] == 0 and timed == 3


def test_settle_default_and_flag(bench):
    assert bench.SETTLE_MS == 50.0
    import inspect
    assert inspect.signature(bench.run_path).parameters["settle_ms"].default == bench.SETTLE_MS
