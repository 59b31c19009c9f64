"""FFT-shape autotuner for the split-step passes.

For a grid length the library compiles several shapes of each pass: E
elements per thread (the FFT's radix) and the number of rows (row pass) or
columns (column pass) per workgroup; each pass can also stream its state
accesses past the caches (non-temporal) or not. Which choice is fastest
depends on the grid, the precision and how many instances share a launch
(small single instances are latency-bound and want many small waves and
streaming; large or batched ones want fewer, fatter threads). The tuner times
every compiled shape in both cache modes with ``sgpe_profile_full_steps`` on
a scratch plan (the caller's states are never touched), picks the fastest
row and column choices independently and caches them in a JSON file keyed by
precision, grid and batch-size bucket as
``[row E, rows, column E, columns, row streaming, column streaming]``.

Usage: ``python -m spinor_gpe.engine.autotune --grids 128 256 512 1024 2048``
writes the cache (``$SGPE_TUNE_FILE``, default
``~/.cache/spinor_gpe/tune.json``).
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np
import torch

from spinor_gpe import _native as nat
from spinor_gpe.engine.plan import InstanceSpec, Plan

PACKAGED = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuned_mi355x.json')


def cache_path():
    return os.environ.get('SGPE_TUNE_FILE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'spinor_gpe', 'tune.json'))


def batch_bucket(n_slots):
    for b in (1, 4, 16, 64):
        if n_slots <= b:
            return b
    return 256


def key(precision, ny, nx, n_slots):
    return f"{nat.precision_of(precision)}:{ny}x{nx}:{batch_bucket(n_slots)}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _lookup_exact(precision, ny, nx, n_slots):
    k = key(precision, ny, nx, n_slots)
    for path in (cache_path(), PACKAGED):
        entry = _load(path).get(k)
        if entry:
            return tuple(entry['shape'])
    return None


def lookup(precision, ny, nx, n_slots):
    """Cached (row_e, row_rows, col_e, col_w[, row_stream, col_stream]) or
    None (user cache, then packaged). A grid without an entry of its own
    takes its row-pass choice from the nx x nx entry and its column-pass
    choice from the ny x ny entry (the row transforms have length nx, the
    column transforms length ny)."""
    exact = _lookup_exact(precision, ny, nx, n_slots)
    if exact or ny == nx:
        return exact
    rows = _lookup_exact(precision, nx, nx, n_slots)
    cols = _lookup_exact(precision, ny, ny, n_slots)
    if not rows or not cols:
        return None
    shape = [rows[0], rows[1], cols[2], cols[3]]
    if len(rows) >= 6 and len(cols) >= 6:
        shape += [rows[4], cols[5]]
    return tuple(shape)


def store(precision, ny, nx, n_slots, shape, timings):
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    data = _load(path)
    data[key(precision, ny, nx, n_slots)] = {'shape': list(shape), 'timings_us': timings,
                                             'when': time.strftime('%Y-%m-%d %H:%M:%S')}
    tmp = path + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def shapes(pass_, precision, n):
    """Compiled (E, rows-or-columns) pairs of one pass (0 rows, 1 columns)."""
    prec = nat.precision_of(precision)
    cnt = nat.lib().sgpe_list_shapes(pass_, prec, n, None, 0)
    if cnt < 0:
        nat.check(cnt)
    buf = (ctypes.c_int * (2 * cnt))()
    nat.lib().sgpe_list_shapes(pass_, prec, n, buf, cnt)
    return [(buf[2 * i], buf[2 * i + 1]) for i in range(cnt)]


def persistent_shapes(precision, n):
    """Shapes (row E, rows per block, column E, columns per block) with a
    compiled persistent step kernel for n x n grids."""
    prec = nat.precision_of(precision)
    cnt = nat.lib().sgpe_list_persistent(prec, n, None, 0)
    if cnt < 0:
        nat.check(cnt)
    buf = (ctypes.c_int * (4 * cnt))()
    nat.lib().sgpe_list_persistent(prec, n, buf, cnt)
    return [tuple(buf[4 * i:4 * i + 4]) for i in range(cnt)]


def synthetic_spec(ny, nx, time_mode='real'):
    """A Raman-coupled Thomas-Fermi-like instance of the given grid (lab frame)."""
    r = 16.0
    x = np.linspace(-r, r, nx, endpoint=False)
    y = np.linspace(-r, r, ny, endpoint=False)
    kx = np.linspace(-np.pi * nx / (2 * r), np.pi * nx / (2 * r), nx, endpoint=False)
    ky = np.linspace(-np.pi * ny / (2 * r), np.pi * ny / (2 * r), ny, endpoint=False)
    xm, ym = np.meshgrid(x, y)
    kxm, kym = np.meshgrid(kx, ky)
    pot = (xm ** 2 + ym ** 2) / 2
    kin = (kxm ** 2 + kym ** 2) / 2
    psi = np.sqrt(np.maximum(30.0 - pot, 0.0)) + 0j
    c = (2 * r / nx) * (2 * r / ny) / (2 * np.pi)
    psik = np.fft.fftshift(np.fft.fftn(psi)) * c
    spec = InstanceSpec([kin, kin], [pot + 1.0, pot - 1.0], [2 * r / nx, 2 * r / ny], 1e4,
                        {'uu': 0.11, 'dd': 0.11, 'ud': 0.105}, time_mode, 1e-3,
                        coupling=np.full((ny, nx), 9.0), theta_x=6.0 * x)
    return spec, np.stack([psik, psik])


def tune(device, precision, ny, nx, n_slots=1, steps=30, reps=3, verbose=False, time_mode=None):
    """Time every compiled shape; returns (best shape, timings dict in us).

    ``time_mode`` of the synthetic instance: by default real time for one
    slot (the single-instance real-time runs) and imaginary time for batched
    plans (ground-state sweeps and served solves); the phase arithmetic of
    the row pass differs (sin/cos vs exp), and with it the fastest shape."""
    if time_mode is None:
        time_mode = 'real' if n_slots == 1 else 'imag'
    spec, psik = synthetic_spec(ny, nx, time_mode)
    plan = Plan(device, precision, ny, nx, n_slots)
    states = torch.empty((n_slots, 2, ny, nx), dtype=plan.dtype, device=plan.device)
    states.copy_(torch.as_tensor(np.broadcast_to(psik, (n_slots,) + psik.shape).copy(),
                                 dtype=plan.dtype))
    for s in range(n_slots):
        plan.set_slot(s, spec, states[s])
    base = plan.tuning()
    base_st = plan.streaming()
    timings = {'row': {}, 'col': {}}
    best = list(base) + list(base_st)
    for pass_, name in ((0, 'row'), (1, 'col')):
        best_t = float('inf')
        for st in (0, 1):
            for e, p in shapes(pass_, precision, nx if pass_ == 0 else ny):
                try:
                    if pass_ == 0:
                        plan.tune(row_e=e, row_rows=p, col_e=base[2], col_w=base[3])
                        plan.set_streaming(st, base_st[1])
                    else:
                        plan.tune(row_e=base[0], row_rows=base[1], col_e=e, col_w=p)
                        plan.set_streaming(base_st[0], st)
                except nat.SgpeError:
                    continue  # exceeds LDS or does not tile this grid
                plan.profile_steps(2)
                idx = 1 if pass_ == 0 else 0  # profile_steps -> (col ms, row ms, ...)
                t = min(plan.profile_steps(steps)[idx] for _ in range(reps)) / (3 * steps) * 1e3
                timings[name][f"{e}x{p}" + ("s" if st else "")] = round(t, 3)
                if verbose:
                    print(f"  {name} E={e} {p}{' streaming' if st else ''}: {t:.2f} us", flush=True)
                if t < best_t:
                    best_t = t
                    if pass_ == 0:
                        best[0], best[1], best[4] = e, p, st
                    else:
                        best[2], best[3], best[5] = e, p, st
    apply_shape(plan, best)
    return tuple(best), timings


def apply_shape(plan, shape):
    """Set a cached choice: 4 shape entries, optionally 2 streaming flags."""
    plan.tune(*shape[:4])
    if len(shape) >= 6:
        plan.set_streaming(shape[4], shape[5])


def apply(plan):
    """Set a plan's shapes from the cache if an entry exists; returns it."""
    shape = lookup(plan.prec, plan.ny, plan.nx, plan.n_slots)
    if shape:
        try:
            apply_shape(plan, shape)
        except nat.SgpeError:
            return None
    return shape


def main(argv=None):
    ap = argparse.ArgumentParser(description="Tune FFT shapes of the split-step passes.")
    ap.add_argument('--grids', type=int, nargs='+', default=[128, 256, 512, 1024, 2048])
    ap.add_argument('--precisions', nargs='+', default=['c128', 'c64'])
    ap.add_argument('--slots', type=int, nargs='+', default=[1])
    ap.add_argument('--steps', type=int, default=30)
    ap.add_argument('--out', default=None, help='cache file (default: $SGPE_TUNE_FILE)')
    args = ap.parse_args(argv)
    if args.out:
        os.environ['SGPE_TUNE_FILE'] = args.out
    dev = torch.device('cuda', 0)
    for prec in args.precisions:
        for n in args.grids:
            for slots in args.slots:
                print(f"[tune] {prec} {n}x{n} slots={slots}", flush=True)
                shape, t = tune(dev, prec, n, n, slots, steps=args.steps, verbose=True)
                store(prec, n, n, slots, shape, t)
                print(f"[tune] best {shape}", flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
