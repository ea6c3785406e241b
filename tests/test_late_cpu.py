"""The Sim-side late path on the CPU: the oracle's TraceGate pinned by the
reference's rules (crates/omg-scene/src/sim.rs:38-76) on hand cases, the
library's host gate (omg_trace_gate_step) against it, and the oracle's
late step (sim.rs:374-397) on cases whose answers are fixed by hand."""
import numpy as np

import oracle as O
import omg_hip as H


def test_gate_fires_first_then_on_motion_energy_or_age():
    g = O.TraceGate(phase=0)
    s, l, e = [1.0, 2.0, 1.5], [4.0, 2.0, 1.5], [1.0, 1.0, 1.0]
    assert g.should_trace(s, l, e, 8)                 # last positions start at f32::MAX
    assert not g.should_trace(s, l, e, 8)
    assert not g.should_trace([1.24, 2.0, 1.5], l, e, 8)   # 0.24 m: stale is fine
    assert g.should_trace([1.26, 2.0, 1.5], l, e, 8)       # 0.26 m: re-trace
    assert not g.should_trace([1.26, 2.0, 1.5], [4.0, 2.2, 1.5], e, 8)
    assert g.should_trace([1.26, 2.0, 1.5], [4.0, 2.3, 1.5], e, 8)
    s2, l2 = [1.26, 2.0, 1.5], [4.0, 2.3, 1.5]
    assert not g.should_trace(s2, l2, [1.09, 1.0, 1.0], 8)  # 9 % of max(e0, e1)
    assert g.should_trace(s2, l2, [1.0, 1.0, 1.12], 8)     # 12 %: moved
    # idle: ages to max_age, then re-traces every max_age ticks
    fires = [g.should_trace(s2, l2, [1.0, 1.0, 1.12], 8) for _ in range(24)]
    assert [i for i, f in enumerate(fires) if f] == [7, 15, 23]


def test_gate_phase_staggers_idle_refresh():
    s, l, e = [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0] * 3
    first = []
    for phase in range(4):
        g = O.TraceGate(phase=phase)
        g.should_trace(s, l, e, 8)                    # the motion-triggered first trace
        fires = [g.should_trace(s, l, e, 8) for _ in range(10)]
        first.append(fires.index(True))
    assert first == [7, 7, 7, 7]                      # after a trace the age restarts at 0
    g = O.TraceGate(phase=7)
    g.g.last_src[0] = 0.0                              # as if positions were known: only age fires
    g.g.last_lis[0] = 1.0
    for b in range(3):
        g.g.last_energy[b] = 1.0
    assert g.should_trace(s, l, e, 8)                 # phase 7 + 1 reaches the age limit at once


def test_library_gate_matches_oracle_gate():
    rng = np.random.default_rng(11)
    for phase, max_age in [(0, 8), (3, 12), (5, 16), (1, 1)]:
        go, gl = O.TraceGate(phase), H.TraceGate(phase)
        s = rng.uniform(-5, 5, 3).astype(np.float32)
        l = rng.uniform(-5, 5, 3).astype(np.float32)
        e = np.ones(3, np.float32)
        for _ in range(400):
            s = s + rng.normal(0, 0.08, 3).astype(np.float32)
            l = l + rng.normal(0, 0.05, 3).astype(np.float32)
            if rng.random() < 0.1:
                e = (e * rng.uniform(0.85, 1.15, 3)).astype(np.float32)
            assert go.should_trace(s, l, e, max_age) == gl.should_trace(s, l, e, max_age)


def test_late_step_hand_cases():
    nbins, nb, dt = 400, 3, 0.005
    slot = O.LateSlot(nbins, nb, dt)
    # untraced before any trace: zero echogram, default RT60, no compensation
    r = slot.step(False, None, None, [0, 0, 0], [3, 0, 0])
    assert r["version"] == 1 and np.all(r["rt60"] == 0.5) and np.all(r["level"] == 0.0)
    assert np.array_equal(r["dir"], [1.0, 0.0, 0.0]) and r["aniso"] == 0.0
    # first trace: alpha 1.0 copies the echogram; an exponential decay with
    # T60 = 0.6 s gives RT60 0.6 within one bin quantum
    t = (np.arange(nbins) + 0.5) * dt
    decay = (0.2 * 10.0 ** (-6.0 * t / 0.6)).astype(np.float32)
    bins = np.repeat(decay[:, None], nb, axis=1)
    dirs = np.zeros((nbins, 3), np.float32)
    dirs[:, 0] = -decay
    # version counts every tick (sim.rs:380): this slot's first trace comes at
    # version 1 and blends with alpha 0.3 into the zero average
    slot.step(True, bins, dirs, [0, 0, 0], [4, 0, 0])
    assert np.array_equal(slot.bins, (np.float32(0.3) * bins).astype(np.float32))
    # a fresh slot traced on its first tick takes alpha 1.0: 0 + 1 * (o - 0) = o exactly
    slot = O.LateSlot(nbins, nb, dt)
    r = slot.step(True, bins, dirs, [0, 0, 0], [4, 0, 0])
    assert slot.s.version == 1
    assert np.array_equal(slot.bins, bins)
    assert np.all(np.abs(r["rt60"] - 0.6) <= 60.0 * dt / 20.0 + 1e-6)
    assert np.allclose(r["dir"], [-1.0, 0.0, 0.0]) and r["aniso"] == np.float32(1.0)
    late = np.sqrt(np.float32(decay[int(np.float32(0.08) / np.float32(dt)):].sum(dtype=np.float32)))
    assert np.allclose(r["level"], min(late, 0.7), rtol=1e-6)
    lvl_traced = r["level"].copy()
    # a second, untraced tick 6 m further away: the level scales by (4+1)/(10+1)
    r = slot.step(False, None, None, [0, 0, 0], [10, 0, 0])
    k = np.float32(np.float32(5.0) / np.float32(11.0))
    assert np.array_equal(r["level"], (lvl_traced * k).astype(np.float32))
    # closer than at the trace: never scaled up
    r = slot.step(False, None, None, [0, 0, 0], [1, 0, 0])
    assert np.array_equal(r["level"], lvl_traced)
    # later traces blend with alpha 0.3
    before = slot.bins.copy()
    slot.step(True, np.zeros_like(bins), np.zeros_like(dirs), [0, 0, 0], [1, 0, 0])
    assert np.array_equal(slot.bins, (before + np.float32(0.3) * (0 - before)).astype(np.float32))
