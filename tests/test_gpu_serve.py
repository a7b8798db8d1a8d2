"""The continuous-batching server: an episode's results are the same alone
and in a mix (bit for bit: every member's arithmetic is independent of the
others), and admission respects the device-memory budget."""
import sys

import numpy as np
import pytest

from freeflow_amd import serve

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_episodes(seed, n=5, steps=40):
    eps = serve.make_episodes(rate=1000.0, duration=1.0, seed=seed, steps=steps, action_every=10,
                              sizes2d=((160, 240), (80, 120)), sizes3d=((32, 48), (32, 32), (24, 32)))[:n]
    assert {e.kind for e in eps} == {2, 3}, "the seed must give both kinds"
    return eps


def test_alone_equals_mix():
    eps = small_episodes(7)
    note(f"case serve mix {[(e.kind, e.nx, e.ny, e.nz) for e in eps]}")
    srv = serve.Server(storage="f16", keep_final=True)
    summary = srv.run(eps, realtime=False)
    srv.close()
    assert summary["episodes_done"] == len(eps)
    assert summary["peak_active"] == len(eps)
    for target in (eps[0], next(e for e in eps if e.kind == 3)):
        solo = serve.make_episodes(rate=1000.0, duration=1.0, seed=7, steps=40, action_every=10,
                                   sizes2d=((160, 240), (80, 120)), sizes3d=((32, 48), (32, 32), (24, 32)))
        alone = solo[target.eid]
        s2 = serve.Server(storage="f16", keep_final=True)
        s2.run([alone], realtime=False)
        s2.close()
        assert np.array_equal(alone.final_moments, target.final_moments), f"episode {target.eid}"
        assert len(alone.wrench) == len(target.wrench) == 4
        for a, b in zip(alone.wrench, target.wrench):
            assert np.array_equal(a, b)
        assert np.abs(target.wrench[-1]).max() > 0


def test_admission_by_footprint():
    eps = small_episodes(9, n=4, steps=20)
    srv = serve.Server(storage="f16", keep_final=True)
    fps = sorted(srv.footprint(e) for e in eps)
    srv.budget = fps[-1] + fps[-2]  # room for two of the largest at a time
    note(f"case admission budget {srv.budget}")
    summary = srv.run(eps, realtime=False)
    srv.close()
    assert summary["episodes_done"] == len(eps)
    assert summary["peak_bytes"] <= srv.budget
    assert summary["peak_active"] < len(eps)
    assert summary["p99_step_ms"] >= summary["p50_step_ms"] > 0


def test_episode_larger_than_budget_is_rejected():
    # it would block the queue's head for good; the others are still served
    eps = small_episodes(9, n=4, steps=20)
    srv = serve.Server(storage="f16")
    fps = [srv.footprint(e) for e in eps]
    srv.budget = max(fps) - 1
    big = [e for e, f in zip(eps, fps) if f == max(fps)]
    note(f"case budget {srv.budget}, {len(big)} episode(s) above it")
    summary = srv.run(eps, realtime=False)
    srv.close()
    assert summary["episodes_rejected"] == len(big) >= 1
    assert all(e.rejected and e.finished_at is None for e in big)
    assert summary["episodes_done"] == len(eps) - len(big)
    assert summary["peak_bytes"] <= srv.budget


def test_batched_equals_plain_form():
    """Served in one batch per kind or every episode on its own launches: the
    same moments and wrench histories, bit for bit (actions re-detect links in
    one batched launch in the first form, per domain in the second)."""
    res = {}
    for batched in (True, False):
        eps = small_episodes(11, n=4, steps=30)
        note(f"case serve batched={batched} {[(e.kind, e.nx, e.ny, e.nz) for e in eps]}")
        srv = serve.Server(storage="f16", keep_final=True, batched=batched)
        summary = srv.run(eps, realtime=False)
        srv.close()
        assert summary["episodes_done"] == len(eps)
        res[batched] = eps
    for a, b in zip(res[True], res[False]):
        assert np.array_equal(a.final_moments, b.final_moments), f"episode {a.eid} kind {a.kind}"
        assert len(a.wrench) == len(b.wrench) == 3
        for wa, wb in zip(a.wrench, b.wrench):
            assert np.array_equal(wa, wb)
