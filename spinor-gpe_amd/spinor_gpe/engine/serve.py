"""Continuous-batching solve server.

Solve requests arrive over time: a ground state (imaginary time until the GP
energy changes by at most ``conv_tol`` per full step, or ``max_steps``), or a
quench (that ground state, then ``quench_steps`` real-time steps after the
detuning is raised by ``quench_shift_er`` recoil energies). The reference
solves one problem per ``TensorPropagator`` (tensor_propagator.py:67-212);
here every grid size has one resident plan with a fixed number of slots, and:

* a request is admitted at the next block boundary once a slot of its grid
  is free and its device memory (state, tables) fits the HBM budget; its
  state then stays on the device until it finishes;
* all admitted requests advance together, ``block`` full steps at a time, in
  one captured hipGraph per set of busy plans (concurrent branches), with the
  energy evaluation and convergence test on the device after every block;
* a finished request leaves at a block boundary, freeing its slot and memory
  for the queue, while the others continue.

Slots only share launches: each slot's arithmetic depends on its own data
alone, so a request's result does not depend on what else is in flight.
"""

import dataclasses
import time

import numpy as np
import torch

from spinor_gpe import _native as nat
from spinor_gpe.engine import workloads as wl
from spinor_gpe.engine.plan import InstanceSpec, Plan
from spinor_gpe.engine.sweep import ENERGY_KEYS, Group

DEFAULT_SLOTS = {128: 32, 256: 32, 512: 16, 1024: 8}


@dataclasses.dataclass
class Request:
    """One solve: ``draw`` fixes the physics (see :mod:`workloads`)."""
    rid: int
    draw: wl.Draw
    kind: str = 'ground'          # 'ground' or 'quench'
    arrival: float = 0.0          # seconds after the server starts


@dataclasses.dataclass
class Outcome:
    """What a finished request reports (times in seconds from the start)."""
    rid: int
    grid: int
    kind: str
    arrival: float
    admitted: float
    finished: float
    steps_imag: int
    steps_real: int
    converged: bool
    energy: dict
    psik: object = None

    @property
    def latency(self):
        return self.finished - self.arrival


class _Job:
    # pylint: disable=too-few-public-methods
    def __init__(self, req, plan, slot, state, nbytes, admitted):
        self.req, self.plan, self.slot, self.state = req, plan, slot, state
        self.nbytes, self.admitted = nbytes, admitted
        self.phase = 'imag'
        self.steps_imag = 0
        self.converged = False


def state_bytes(grid, precision):
    """Device bytes a request holds while admitted: its [2, n, n] state plus
    its slot tables (1-D phase tables, norm partials, energy partials). The
    energy evaluation works in place on the state, so this is all."""
    b = 16 if nat.precision_of(precision) == nat.C128 else 8
    return 2 * grid * grid * b + 64 * grid * b


class Server:
    """Continuous-batching server on one GPU (see the module docstring).

    ``slots``: slots per grid size; ``hbm_budget``: bytes of device memory the
    server may hold for requests (default: 80 % of the free memory when the
    server starts); a request is charged its state and slot tables while it
    is admitted.
    """

    # pylint: disable=too-many-instance-attributes,too-many-arguments
    def __init__(self, device='cuda', precision='c64', slots=None, hbm_budget=None,
                 block=100, conv_tol=1e-8, max_steps=8000, dt_imag=0.02,
                 quench_steps=5000, dt_real=1e-3, quench_shift_er=0.5, keep_states=False):
        self.device = nat.require_device(device)
        self.precision = precision
        self.slots = dict(DEFAULT_SLOTS if slots is None else slots)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.budget = int(hbm_budget)
        self.used = 0
        self.block = int(block)
        self.conv_tol = float(conv_tol)
        self.max_steps = int(max_steps)
        self.dt_imag, self.dt_real = float(dt_imag), float(dt_real)
        self.quench_steps = int(quench_steps)
        self.quench_shift_er = float(quench_shift_er)
        self.keep_states = keep_states
        self.plans = {}
        self.free = {}
        self.groups = {}
        if self.max_steps % self.block or self.quench_steps % self.block:
            raise ValueError("max_steps and quench_steps must be multiples of block")

    # ------------------------------------------------------------ resources
    def _plan(self, grid):
        if grid not in self.plans:
            if grid not in self.slots:
                raise ValueError(f"no slots configured for grid {grid}")
            n = self.slots[grid]
            self.plans[grid] = Plan(self.device, self.precision, grid, grid, n)
            self.free[grid] = list(range(n))
        return self.plans[grid]

    def _group(self, grids):
        key = tuple(sorted(grids))
        if key not in self.groups:
            self.groups[key] = Group([self.plans[g] for g in key])
        return self.groups[key]

    def _admit(self, req, now):
        """Start ``req`` if a slot and memory are free; returns the job or None."""
        grid = req.draw.grid
        need = state_bytes(grid, self.precision)
        plan = self._plan(grid)
        if plan is None or not self.free[grid] or self.used + need > self.budget:
            return None
        slot = self.free[grid].pop(0)
        ps = wl.raman_spinor(f'serve{req.rid}', req.draw)
        state = plan.new_state(ps.psik)
        plan.set_slot(slot, InstanceSpec.from_spin(ps, 'imag', self.dt_imag), state)
        plan.set_run(slot, None, conv_tol=self.conv_tol)
        self.used += need
        return _Job(req, plan, slot, state, need, now)

    def _finish(self, job, steps, converged, energy, now):
        job.plan.set_active(job.slot, False)
        self.free[job.req.draw.grid].append(job.slot)
        self.used -= job.nbytes
        psik = job.state.cpu().numpy() if self.keep_states else None
        steps_real = steps if job.phase == 'real' else 0
        steps_imag = job.steps_imag if job.phase == 'real' else steps
        return Outcome(job.req.rid, job.req.draw.grid, job.req.kind, job.req.arrival, job.admitted,
                       now, steps_imag, steps_real, converged,
                       dict(zip(ENERGY_KEYS, (float(v) for v in energy))), psik)

    # ----------------------------------------------------------------- loop
    def run(self, requests, clock=None):
        """Serve ``requests`` (arrival times relative to the call); returns
        the outcomes in completion order."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        waiting, jobs, done = [], [], []
        clock = clock or time.perf_counter
        t0 = clock()
        while pending or waiting or jobs:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            still = []
            for req in waiting:
                job = self._admit(req, now)
                if job is None:
                    still.append(req)
                else:
                    jobs.append(job)
            waiting = still
            if not jobs:
                if waiting and not pending:
                    raise MemoryError("requests cannot be admitted: HBM budget or slots too small")
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            busy = {j.req.draw.grid for j in jobs}
            self._group(busy).run(self.block, True, 1)
            status = {g: self.plans[g].status() for g in busy}
            now = clock() - t0
            remaining = []
            for job in jobs:
                st, en = status[job.req.draw.grid]
                active, converged, steps = (int(v) for v in st[job.slot])
                if job.phase == 'imag':
                    if active and steps < self.max_steps:
                        remaining.append(job)
                        continue
                    job.converged = bool(converged)
                    if job.req.kind == 'quench':
                        self._start_quench(job, steps)
                        remaining.append(job)
                        continue
                    done.append(self._finish(job, steps, job.converged, en[job.slot], now))
                else:
                    if steps < self.quench_steps:
                        remaining.append(job)
                        continue
                    done.append(self._finish(job, steps, job.converged, en[job.slot], now))
            jobs = remaining
        return done

    def _start_quench(self, job, steps):
        """Ground state reached: raise the detuning and run in real time from
        the normalized k-space state the energy check left in place."""
        job.steps_imag = steps
        job.phase = 'real'
        ps = wl.raman_spinor(f'serve{job.req.rid}q', job.req.draw, delta_shift_er=self.quench_shift_er)
        job.plan.set_slot(job.slot, InstanceSpec.from_spin(ps, 'real', self.dt_real), job.state)
        job.plan.set_run(job.slot, None, conv_tol=0.0)


def request_stream(n, rate, seed=3, grids=((128, 0.3), (256, 0.4), (512, 0.2), (1024, 0.1)),
                   quench_frac=0.3):
    """``n`` seeded requests with Poisson arrivals at ``rate`` per second
    (rate None: all at t = 0); grid sizes and kinds drawn with the given
    probabilities, physics drawn as in :func:`workloads.draw_params`."""
    rng = np.random.default_rng(seed)
    sizes = np.array([g for g, _ in grids])
    probs = np.array([p for _, p in grids], dtype=np.float64)
    t = 0.0
    out = []
    for i in range(n):
        if rate:
            t += float(rng.exponential(1.0 / rate))
        grid = int(rng.choice(sizes, p=probs / probs.sum()))
        kind = 'quench' if rng.uniform() < quench_frac else 'ground'
        out.append(Request(i, wl.draw_params(rng, grid), kind, t if rate else 0.0))
    return out


def summarize(outcomes, wall):
    """Throughput and latency figures of a served run. ``solves_per_s``
    counts the results whose ground-state phase met the stop test;
    ``results_per_s`` counts every result, the unconverged ones (stopped at
    the step cap) included; latencies are over all results."""
    lat = np.array([o.latency for o in outcomes])
    wait = np.array([o.admitted - o.arrival for o in outcomes])
    conv = int(sum(o.converged for o in outcomes))
    return {'requests': len(outcomes), 'wall_s': round(wall, 3),
            'converged': conv, 'unconverged': len(outcomes) - conv,
            'solves_per_s': round(conv / wall, 3) if wall > 0 else None,
            'results_per_s': round(len(outcomes) / wall, 3) if wall > 0 else None,
            'latency_p50_s': round(float(np.percentile(lat, 50)), 4),
            'latency_p99_s': round(float(np.percentile(lat, 99)), 4),
            'queue_wait_p50_s': round(float(np.percentile(wait, 50)), 4),
            'queue_wait_p99_s': round(float(np.percentile(wait, 99)), 4)}
