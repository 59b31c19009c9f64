"""Profiling workload: split steps of one grid, launched eagerly.

Runs a Raman-coupled real-time instance (the tuner's synthetic instance) on
the shipped or cached FFT shapes and issues ``--steps`` full steps through
``sgpe_profile_full_steps`` (eager launches, per-kernel execution times).
Meant to run under ``rocprofv3`` for kernel traces and hardware counters:

    rocprofv3 --pmc FETCH_SIZE -- python3 -m spinor_gpe.engine.profile_run --grid 512

(with ``spinor-gpe_amd`` on ``PYTHONPATH``). ``--calibrate-copy MIB`` first
copies a buffer of that size once, a known-size reference for HBM counters.
``--others N`` then runs N of each of the other kernels of the library on
the same state: the sync (exit column pass and normalize kernel), the GP
energy evaluation (energy column, row and final kernels), the standalone
fft_2d / ifft_2d (row and column FFT kernels), norm() and calc_pops()
(density sums, norm apply, pops final).
"""

import argparse
import json
import sys

import numpy as np
import torch

from spinor_gpe.engine.autotune import synthetic_spec
from spinor_gpe.engine.plan import Plan


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--grid', type=int, default=512)
    ap.add_argument('--precision', default='c128', choices=['c128', 'c64'])
    ap.add_argument('--slots', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB')
    ap.add_argument('--others', type=int, default=0, metavar='N',
                    help='also run N of each of the other kernels (sync, energy, fft2, norm, pops)')
    args = ap.parse_args(argv)
    dev = torch.device('cuda', 0)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy * 2 ** 20 // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    n = args.grid
    spec, psik = synthetic_spec(n, n)
    plan = Plan(dev, args.precision, n, n, args.slots)
    states = torch.empty((args.slots, 2, n, n), dtype=plan.dtype, device=dev)
    states.copy_(torch.as_tensor(np.broadcast_to(psik, (args.slots,) + psik.shape).copy(), dtype=plan.dtype))
    for s in range(args.slots):
        plan.set_slot(s, spec, states[s])
    plan.profile_steps(2)
    col_ms, row_ms, n_col, n_row = plan.profile_steps(args.steps)
    if args.others:
        from spinor_gpe import _native as nat  # pylint: disable=import-outside-toplevel
        dr = spec.dr
        plan.sync()
        plan.energy()  # allocates the energy work buffer outside the counted calls
        torch.cuda.synchronize()
        for _ in range(args.others):
            plan.full_steps(1)
            plan.sync()
        for _ in range(args.others):
            plan.energy()
        st = states[0]
        for _ in range(args.others):
            out = nat.fft2([st[0], st[1]], True, dr[0], dr[1])
            nat.fft2(out, False, dr[0], dr[1], out=out)
        for _ in range(args.others):
            nat.norm([st[0], st[1]], float(dr[0] * dr[1]), 1e4)
            nat.pops([st[0], st[1]], float(dr[0] * dr[1]))
        torch.cuda.synchronize()
    print(json.dumps({'grid': n, 'precision': args.precision, 'slots': args.slots,
                      'shape': plan.tuning(), 'streaming': plan.streaming(),
                      'col_pass_us': col_ms / n_col * 1e3,
                      'row_pass_us': row_ms / n_row * 1e3}), flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
