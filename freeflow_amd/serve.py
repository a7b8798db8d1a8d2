"""Continuous-batching server for RL-rollout episodes on one GPU.

The reference drives one simulator per environment from Python
(rl/env.py:163-187: reset, then `step(action)` = 10 fluid steps with new wave
parameters, reading back forces).  Here many episodes of different grid
sizes share the GPU: each episode's fluid and boundary state lives on the
device in its own domain; domains are members of one batch per kind (2D and
3D, FP16 storage and FP32 fast-math arithmetic by default: FMA-reassociated
collision and reconstruction, FP32 wave kinematics, FP64 only in the
per-body force sums).  A server iteration advances every active
episode to its next action boundary: the due actions (new wave amplitude and
phase, uploaded on the batch stream), one multi-step graph replay per batch,
one wait, and the per-step FP64 wrench of the interval read back.  Episodes
that arrive while others run are admitted at the next action boundary as
soon as their device footprint (ff_domain_footprint) fits the memory budget
and a slot is free, so all running episodes share one action phase (an
episode larger than the whole budget is rejected at submission); finished
episodes leave the batch and free their memory.

Per-step latency of an episode is the wall time of the iteration that
advanced it divided by the steps it advanced; the server reports p50/p99
over all episode-steps, the aggregate MLUPS and the p50/p99 wait from an
episode's arrival to its admission.
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi, workloads as wl


@dataclass
class Episode:
    """One rollout request.  `kind` 2 or 3; `actions[k]` = (a, phi) applied
    before step k*action_every: the wave amplitude becomes a times the
    episode's base amplitude and the phase becomes phi."""
    eid: int
    kind: int
    nx: int
    ny: int
    nz: int = 1
    steps: int = 500
    action_every: int = 10
    seed: int = 0
    arrival: float = 0.0
    actions: np.ndarray = None
    # filled by the server
    dom: object = None
    scene: object = None
    done_steps: int = 0
    admitted_at: float = None
    finished_at: float = None
    footprint: int = 0
    rejected: bool = False  # its footprint alone exceeds the memory budget
    wrench: list = field(default_factory=list)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz


class Scene3D:
    """3D episode: rest fluid in a clamped-halo box with a sphere that
    oscillates along y with the device-side wave kinematics (s = 1 on every
    vertex and no travelling term: y = y_rest + A sin(phi - 2 pi f t))."""

    def __init__(self, nx, ny, nz, rng, viscosity=0.02):
        self.nx, self.ny, self.nz, self.viscosity = nx, ny, nz, viscosity
        self.center = np.array([0.35 * nx, 0.5 * ny, 0.5 * nz])
        self.r = 0.12 * min(ny, nz)
        self.v, self.f = wl.icosphere(self.center, self.r, subdiv=2)
        self.s = np.ones(len(self.v), np.float32)
        self.params = np.array([float(rng.uniform(0.5, 2.0)), 1.0 / 200.0, float(rng.uniform(0, 2 * np.pi)), 0.0])

    def domain_args(self):
        return dict(dim=3, nx=self.nx, ny=self.ny, nz=self.nz, viscosity=self.viscosity)

    def setup(self, dom, history=0):
        dom.init_state([0.0] * 6)  # rest fluid, computed on the device
        dom.set_boundary(self.v, np.zeros_like(self.v), self.f)
        dom.set_wave(self.v, self.s, self.params[None, :])
        if history:
            dom.set_wrench_history(history)


def make_episodes(rate, duration, seed, frac2d=0.7, steps=500, action_every=10,
                  sizes2d=((600, 1200), (200, 400)), sizes3d=((96, 192), (96, 96), (96, 96))):
    """Poisson arrivals at `rate` episodes/s over `duration` s (seeded)."""
    rng = np.random.default_rng(seed)
    t, eps = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        k = 2 if rng.random() < frac2d else 3
        n_act = (steps + action_every - 1) // action_every
        if k == 2:
            nx = int(rng.integers(sizes2d[0][0], sizes2d[0][1] + 1))
            ny = int(rng.integers(sizes2d[1][0], sizes2d[1][1] + 1))
            nz = 1
        else:
            nx = int(rng.integers(sizes3d[0][0], sizes3d[0][1] + 1))
            ny = int(rng.integers(sizes3d[1][0], sizes3d[1][1] + 1))
            nz = int(rng.integers(sizes3d[2][0], sizes3d[2][1] + 1))
        acts = np.stack([rng.uniform(0.5, 1.0, n_act), rng.uniform(0, 2 * np.pi, n_act)], 1)
        eps.append(Episode(len(eps), k, nx, ny, nz, steps, action_every, int(rng.integers(1 << 30)), t, acts))
    return eps


class Server:
    def __init__(self, storage="f16", budget_bytes=None, max_active=512, device=0, keep_final=False, batched=True,
                 fast_math=True):
        self.storage = storage
        # fast_math=False: the reference's operation order (2D with its FP16 tile staging)
        self.fast_math = fast_math
        self.keep_final = keep_final  # read back each episode's final moments (ep.final_moments)
        # batched=False: the plain per-domain form, every episode stepped by
        # its own launches on its own stream (the same admission and actions)
        self.batched = batched
        self.budget = budget_bytes if budget_bytes is not None else 288 * 10 ** 9
        self.max_active = max_active
        self.batches = {2: capi.Batch(2, storage, stage_fp16=not fast_math, fast_math=fast_math, device=device),
                        3: capi.Batch(3, storage, fast_math=fast_math, device=device)}
        self.device = device
        self.active = []
        self.used = 0
        self.lat = []          # one entry per episode-step: iteration wall time (s)
        self.cell_steps = 0
        self.iterations = 0
        # wall time by phase (s): actions, replays + wait, wrench readback, admissions, retirements
        self.phase = {"act": 0.0, "replay": 0.0, "read": 0.0, "admit": 0.0, "retire": 0.0}

    # ------------------------------------------------------------ episodes
    def _scene(self, ep):
        rng = np.random.default_rng(ep.seed)
        if ep.kind == 2:
            return wl.SwimmerScene2D(ep.nx, ep.ny, int(rng.integers(200, 801)), rng)
        return Scene3D(ep.nx, ep.ny, ep.nz, rng)

    def footprint(self, ep):
        if ep.scene is None:
            ep.scene = self._scene(ep)
        sc = ep.scene
        if ep.kind == 2:
            v, e = sc.swimmer.rest, sc.swimmer.elems
            return capi.footprint(2, ep.nx, ep.ny, storage=self.storage, verts=v, elems=e, dx=sc.dx)
        return capi.footprint(3, ep.nx, ep.ny, ep.nz, storage=self.storage, verts=sc.v, elems=sc.f)

    def admit(self, ep, now):
        sc = ep.scene
        d = capi.Domain(**sc.domain_args(), storage=self.storage, stage_fp16=ep.kind == 2 and not self.fast_math,
                        fast_math=self.fast_math, device=self.device)
        sc.setup(d, history=ep.action_every)
        ep.dom = d
        ep.admitted_at = now
        if self.batched:
            self.batches[ep.kind].add(d)
        self.active.append(ep)
        self.used += ep.footprint

    def _act(self, ep):
        k = ep.done_steps // ep.action_every
        A, phi = ep.actions[min(k, len(ep.actions) - 1)]
        sc = ep.scene
        p = (sc.swimmer.params if ep.kind == 2 else sc.params).copy()
        p[0] = p[0] * A
        p[2] = phi
        ep.dom.set_wave_params(p[None, :])

    def _retire(self, ep, now):
        ep.finished_at = now
        if self.batched:
            self.batches[ep.kind].remove(ep.dom)
        if self.keep_final:
            ep.final_moments = ep.dom.moments(0)
        ep.dom.close()
        ep.dom = None
        self.active.remove(ep)
        self.used -= ep.footprint

    # ------------------------------------------------------------ loop
    def at_boundary(self):
        """Every active episode is at an action boundary (admissions wait for
        one, so the episodes' action phases stay aligned)."""
        return all(ep.done_steps % ep.action_every == 0 for ep in self.active)

    def step_all(self):
        """Advance every active episode to its next action boundary (or its
        end): due actions, one multi-step graph replay per batch, one wait,
        the wrench readings of the interval.  Returns the iteration's wall
        time; each episode-step of the iteration is charged the iteration's
        time divided by its steps."""
        t0 = time.perf_counter()
        n = min(min(ep.action_every - ep.done_steps % ep.action_every, ep.steps - ep.done_steps)
                for ep in self.active)
        for ep in self.active:
            if ep.done_steps % ep.action_every == 0:
                self._act(ep)
        t1 = time.perf_counter()
        if self.batched:
            for k in (2, 3):
                if len(self.batches[k].members):
                    self.batches[k].step(n)
            for k in (2, 3):
                if len(self.batches[k].members):
                    self.batches[k].sync()
        else:
            for ep in self.active:
                ep.dom.step(n)
            for ep in self.active:
                ep.dom.sync()
        t2 = time.perf_counter()
        for ep in self.active:
            ep.done_steps += n
            self.cell_steps += ep.cells * n
            if ep.done_steps % ep.action_every == 0 or ep.done_steps == ep.steps:
                ep.wrench.append(ep.dom.wrench_history(ep.action_every))
        t3 = time.perf_counter()
        self.phase["act"] += t1 - t0
        self.phase["replay"] += t2 - t1
        self.phase["read"] += t3 - t2
        dt = t3 - t0
        self.lat.extend([dt / n] * (len(self.active) * n))
        self.iterations += 1
        return dt

    def run(self, episodes, time_limit=None, realtime=True):
        """Serve `episodes` (sorted by arrival).  With `realtime`, arrivals
        are released at their arrival time on the wall clock; otherwise all
        are queued at once.  Returns a summary dict."""
        pending = sorted(episodes, key=lambda e: e.arrival)
        for ep in pending:
            ep.footprint = self.footprint(ep)
            ep.rejected = ep.footprint > self.budget
        # an episode that can never fit would block the queue's head for good
        pending = [ep for ep in pending if not ep.rejected]
        queue = []
        t_start = time.perf_counter()
        i = 0
        busy = 0.0
        peak = peak_used = 0
        while i < len(pending) or queue or self.active:
            now = time.perf_counter() - t_start
            if time_limit is not None and now > time_limit:
                break
            while i < len(pending) and (not realtime or pending[i].arrival <= now):
                queue.append(pending[i])
                i += 1
            # admission in arrival order while memory and slots allow, at an
            # action boundary of the running episodes
            ta = time.perf_counter()
            while (queue and len(self.active) < self.max_active and self.used + queue[0].footprint <= self.budget
                   and self.at_boundary()):
                self.admit(queue.pop(0), now)
            self.phase["admit"] += time.perf_counter() - ta
            if not self.active:
                if i < len(pending) and realtime:
                    time.sleep(max(0.0, min(0.01, pending[i].arrival - now)))
                continue
            peak = max(peak, len(self.active))
            peak_used = max(peak_used, self.used)
            busy += self.step_all()
            now = time.perf_counter() - t_start
            tr = time.perf_counter()
            for ep in [e for e in self.active if e.done_steps >= e.steps]:
                self._retire(ep, now)
            self.phase["retire"] += time.perf_counter() - tr
        wall = time.perf_counter() - t_start
        lat = np.array(self.lat) * 1e3
        done = [e for e in episodes if e.finished_at is not None]
        # arrival to first step (queueing for memory, a slot and an action boundary)
        wait = np.array([(e.admitted_at - e.arrival) * 1e3 for e in done]) if realtime else np.array([])
        return {
            "episodes_done": len(done),
            "episodes_submitted": len(episodes),
            "episodes_rejected": sum(1 for e in episodes if e.rejected),
            "wall_s": wall,
            "gpu_busy_s": busy,
            "iterations": self.iterations,
            "aggregate_mlups": self.cell_steps / wall / 1e6 if wall > 0 else 0.0,
            "busy_mlups": self.cell_steps / busy / 1e6 if busy > 0 else 0.0,
            "p50_step_ms": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_step_ms": float(np.percentile(lat, 99)) if lat.size else None,
            "p50_admission_wait_ms": float(np.percentile(wait, 50)) if wait.size else None,
            "p99_admission_wait_ms": float(np.percentile(wait, 99)) if wait.size else None,
            "peak_active": peak,
            "peak_bytes": peak_used,
            "budget_bytes": self.budget,
            "phase_s": {k: round(v, 4) for k, v in self.phase.items()},
        }

    def close(self):
        for ep in list(self.active):
            self._retire(ep, 0.0)
        for b in self.batches.values():
            b.close()
