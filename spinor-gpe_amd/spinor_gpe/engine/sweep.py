"""Batched ground-state parameter sweeps.

The reference solves one instance per ``TensorPropagator`` with a fixed step
count (spinor_gpe/pspinor/tensor_propagator.py:151-212). A :class:`Sweep`
solves many instances of possibly different grid sizes at once: instances
are grouped by grid into plans (one batched launch per pass per plan), all
plans advance together through one captured hipGraph per check interval
(concurrent branches, k full steps plus an on-device energy evaluation and
convergence test each), and converged instances freeze on the device while
the others continue. The host only polls a small status array; when a
plan's active instances fall to half of the slots it launches, the plan is
compacted (``Plan.compact``: active slots first, kernels launched for those
only, run state moved with each slot), so finished instances stop costing
workgroups.
"""

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import InstanceSpec, Plan

ENERGY_KEYS = ('E_tot', 'E_kin', 'E_pot', 'E_int', 'E_cpl', 'mu', 'N', 'rel_change')


@dataclass
class SolveResult:
    """Outcome of one instance of a sweep."""
    index: int
    grid: tuple
    steps: int
    converged: bool
    energy: dict
    psik: object = None          # host complex array (2, ny, nx) if requested
    extra: dict = field(default_factory=dict)


class Group:
    """Python handle of an ``sgpe_group`` (plans advanced by one graph)."""

    def __init__(self, plans):
        self.plans = list(plans)
        ctx = self.plans[0].ctx
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        handle = ctypes.c_void_p()
        ctx.bind_stream()
        nat.check(nat.lib().sgpe_group_create(ctx.handle, len(self.plans), arr, ctypes.byref(handle)))
        self.handle = handle
        self.ctx = ctx

    def run(self, k, check=True, n_graphs=1):
        self.ctx.bind_stream()
        nat.check(nat.lib().sgpe_group_run(self.handle, int(k), int(bool(check)), int(n_graphs)))

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h:
            try:
                nat.lib().sgpe_group_destroy(h)
            except Exception:  # pylint: disable=broad-except
                pass
            self.handle = None


class Sweep:
    """Solve a set of instances to convergence in one batched, graph-driven loop.

    ``precision``: 'c64' (float32 arithmetic, float64 reductions) or 'c128'.
    ``conv_tol``: relative energy change per full step that counts as
    converged; ``check_every``: full steps between energy checks;
    ``max_steps``: cap per instance.
    """

    def __init__(self, device='cuda', precision='c64', conv_tol=1e-8, check_every=50,
                 max_steps=8000, poll_every=4):
        self.device = nat.require_device(device)
        self.precision = precision
        self.conv_tol = float(conv_tol)
        self.check_every = int(check_every)
        self.max_steps = int(max_steps)
        self.poll_every = int(poll_every)
        self.items = []
        self.trace = []  # (seconds, full steps, active instances) at each poll

    def add(self, spec, psik):
        """Queue an instance (:class:`InstanceSpec`, k-space state (2, ny, nx))."""
        if spec.time != 'imag':
            raise ValueError("sweeps solve ground states: use imaginary time")
        self.items.append((spec, np.asarray(psik)))
        return len(self.items) - 1

    def add_spin(self, spin, t_step):
        """Queue a PSpinor's ground-state problem (imaginary time, ``t_step``)."""
        return self.add(InstanceSpec.from_spin(spin, 'imag', t_step), np.stack(spin.psik))

    def run(self, keep_states=True):
        """Solve everything; returns (list of SolveResult, wall seconds)."""
        groups = {}
        for idx, (spec, _) in enumerate(self.items):
            groups.setdefault((spec.ny, spec.nx), []).append(idx)
        plans, states, slots = [], [], {}
        for (ny, nx), idxs in sorted(groups.items()):
            plan = Plan(self.device, self.precision, ny, nx, len(idxs))
            st = torch.empty((len(idxs), 2, ny, nx), dtype=plan.dtype, device=self.device)
            host = np.stack([self.items[i][1] for i in idxs])
            st.copy_(torch.as_tensor(host, dtype=plan.dtype))
            for slot, i in enumerate(idxs):
                plan.set_slot(slot, self.items[i][0], st[slot])
                plan.set_run(slot, None, conv_tol=self.conv_tol)
                slots[i] = (len(plans), slot)
            plans.append(plan)
            states.append(st)
        live = list(range(len(plans)))
        group = Group(plans)
        group.run(self.check_every, True, 0)  # allocations and sync outside the timed loop
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        steps = 0
        self.trace = []
        while steps < self.max_steps:
            n = min(self.poll_every, -(-(self.max_steps - steps) // self.check_every))
            group.run(self.check_every, True, n)
            steps += n * self.check_every
            flags = {g: plans[g].status()[0][:, 0] for g in live}
            active = {g: int(f.sum()) for g, f in flags.items()}
            self.trace.append((time.perf_counter() - t0, steps, sum(active.values())))
            if not any(active.values()):
                break
            rebuild = not all(active.values())
            # a plan whose active slots fell to half of its launched ones
            # launches those only from now on
            for g in live:
                plan = plans[g]
                if 0 < active[g] <= plan.n_launch // 2:
                    order = [s for s in range(plan.n_slots) if flags[g][s]]
                    order += [s for s in range(plan.n_slots) if not flags[g][s]]
                    plan.compact(order, active[g])
                    where = {old: new for new, old in enumerate(order)}
                    for i, (gi, si) in slots.items():
                        if gi == g:
                            slots[i] = (g, where[si])
                    rebuild = True
            # plans whose instances have all converged leave the graph
            if rebuild:
                live = [g for g in live if active[g]]
                group = Group([plans[g] for g in live])
        torch.cuda.synchronize(self.device)
        wall = time.perf_counter() - t0
        results = []
        stats = [p.status() for p in plans]
        for i in range(len(self.items)):
            g, slot = slots[i]
            st, en = stats[g]
            spec = self.items[i][0]
            res = SolveResult(i, (spec.ny, spec.nx), int(st[slot, 2]), bool(st[slot, 1]),
                              dict(zip(ENERGY_KEYS, (float(v) for v in en[slot]))))
            if keep_states:
                res.psik = plans[g].states[slot].cpu().numpy()
            results.append(res)
        del group
        return results, wall
