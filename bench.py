"""Benchmark: split-steps/s of the 512^2 real-time Raman-coupled propagation.

Workload (one instance on one GPU): 512 x 512 grid over +/-24 oscillator
lengths, isotropic 50 Hz trap, 1e4 atoms, g = (1, 1, 0.995) x g_scale, Raman
coupling Omega = 2 E_R, detuning 0.5 E_R, kL = 3 along x (lab frame),
dt = 1e-3. The initial state is the imaginary-time ground state of the
256^2 / +/-16 configuration (4000 steps, dt = 0.02, Thomas-Fermi start with
seeded +/-0.1 rad phase noise) resampled to the 512^2 grid, times seeded
Gaussian phase noise (sigma 0.05). A step is one full split step (three
sub-steps) of the whole state; inputs are resident in HBM before timing.

Prints one JSON line. ``value`` is the complex128 rate (the reference's
precision); the complex64 rate is reported under ``"c64"``. ``roofline``
prices the dominant kernel's algorithmic bytes (read + write of the state,
2 * 2 * Nx * Ny * 16 B) against the 8 TB/s HBM peak, with per-kernel average
durations from HIP events on the launch stream; ``traffic`` is that kernel's
measured HBM bytes per dispatch (rocprofv3 counters, profiles folder).
``cpu_baseline`` times the CPU oracle (an op-for-op torch restatement of the
reference's path) on a bounded sample of the same workload (and of the 256^2
ground-state workload). Further legs, reported beside the metric: ``large``
(the 2048^2 vortex workload, where the passes are bandwidth-bound),
``sweep`` (96 batched ground states) and ``serve`` (a continuous-batching
request stream); on request, ``sizes`` and ``fft`` (the reference's own
propagation and fft_2d benchmarks on its 13 grids).
"""

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, 'spinor-gpe_amd'), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
BLOCK_REPS = 20  # graph blocks timed one by one for the median and MAD
TRAFFIC_FILE = 'profiles/hbm_traffic_pass_kernels.json'
TITAN_V_512 = 149.1  # split-steps/s, bench_data/TitanV_cuda_step.npz (512^2, complex128)
METRIC = ("split-steps/s at 512² (1 GPU); sweep solves/s + p50/p99 latency at "
          "ΔE≤1e-6")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def c1_spinor(device):
    """The 512^2 Raman-coupled instance; its initial state comes from the
    256^2 imaginary-time ground state (4000 steps of dt = 0.02)."""
    from spinor_gpe.engine import workloads as wl
    ps0 = wl.c0_spinor('c0')
    res, _ = ps0.imaginary(0.02, 4000, device, progress=False)
    return wl.c1_spinor(res.psi, ps0.space, 'c1')


def time_engine(ps, device, precision, steps, warmup, block, prof_steps, tune, dt=1e-3):
    from spinor_gpe.engine import autotune
    from spinor_gpe.pspinor.tensor_propagator import TensorPropagator
    ny, nx = ps.psi[0].shape
    if tune:
        shape, timings = autotune.tune(torch.device(device), precision, ny, nx, 1)
        autotune.store(precision, ny, nx, 1, shape, timings)
        log(f"tuned {precision}: {shape} {timings}")
    prop = TensorPropagator(ps, dt, steps, device, time='real', precision=precision,
                            block_steps=block, progress=False)
    plan = prop._plan  # pylint: disable=protected-access
    # The timed region is whole captured graphs: blocks of min(block, steps)
    # steps and one graph of the remainder, all captured during the warm-up.
    blk = min(block, steps)
    plan.run(warmup, block=blk)
    plan.graph_steps(blk, 1)
    if steps % blk:
        plan.graph_steps(steps % blk, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.run(steps, block=blk)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # Per-block spread, as the reference reports median and MAD of repeated
    # timings (benchmark_prop.py:92-93): BLOCK_REPS more graph blocks, each
    # bracketed by events on the launch stream.
    events = [torch.cuda.Event(enable_timing=True) for _ in range(BLOCK_REPS + 1)]
    events[0].record()
    for i in range(BLOCK_REPS):
        plan.run(block, block=block)
        events[i + 1].record()
    torch.cuda.synchronize()
    per_step = np.array([events[i].elapsed_time(events[i + 1]) / block for i in range(BLOCK_REPS)])
    med = float(np.median(per_step))
    blocks = {'blocks': BLOCK_REPS, 'steps_per_block': block, 'median_ms_per_step': round(med, 5),
              'mad_ms_per_step': round(float(np.median(np.abs(per_step - med))), 6)}
    col_ms, row_ms, n_col, n_row = plan.profile_steps(prof_steps)
    plan.sync()
    st = plan.states[0]
    atoms = float(torch.sum(torch.abs(st) ** 2).item()) * ps.space['dv_k']
    energy = plan.energy()[0].cpu().numpy()
    return {'elapsed': elapsed, 'timed_graph_block': blk, 'col_ms': col_ms / n_col, 'row_ms': row_ms / n_row,
            'tuning': plan.tuning(), 'streaming': plan.streaming(), 'atoms': atoms, 'blocks': blocks,
            'state': st.cpu().numpy().astype(np.complex128), 'energy': energy,
            'steps_total': warmup + blk + (steps % blk) + steps + BLOCK_REPS * block + prof_steps}


def c64_vs_c128(r64, r128, dv_k):
    """Errors of the complex64 run against the complex128 run of the same
    steps from the same state: state fidelity |<a|b>|^2 / (<a|a><b|b>),
    relative GP energy difference, and population differences (of N_at)."""
    a = np.asarray(r64['state'], dtype=np.complex128)  # sums in float64 (a complex64 norm is off by ~1e-7)
    b = np.asarray(r128['state'], dtype=np.complex128)
    fid = abs(np.vdot(b, a)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
    pops_a = [float(np.sum(np.abs(c) ** 2)) * dv_k for c in a]
    pops_b = [float(np.sum(np.abs(c) ** 2)) * dv_k for c in b]
    n_at = sum(pops_b)
    e64, e128 = float(r64['energy'][0]), float(r128['energy'][0])
    return {'steps': r64['steps_total'], 'state_fidelity': float(fid),
            'energy_rel_err': abs(e64 - e128) / abs(e128),
            'pops_err_of_atoms': max(abs(x - y) for x, y in zip(pops_a, pops_b)) / n_at}


def sweep_leg(device, check_every, fp64_stride):
    """Batched ground-state sweep: 96 seeded Raman instances on 128^2, 256^2
    and 512^2 grids solved together in complex64 (FP64 reductions) to a
    relative energy change of 1e-8 per step or 8000 steps, against the same
    set solved in complex128 (the accuracy reference) and against the plain
    FP64 form: one instance at a time in complex128, on a stratified subset
    (every fp64_stride-th instance of each grid, so the subset has the set's
    grid mix). Both rates count the same thing: solves that met the stop
    test and whose energy is within 1e-6 of the complex128 solve, per second
    of wall time."""
    from spinor_gpe.engine import workloads as wl
    from spinor_gpe.engine.sweep import Sweep
    draws = wl.sweep_draws(seed=2)
    spins = [wl.raman_spinor(f'sweep{i}', d) for i, d in enumerate(draws)]
    conv, cap, dt = 1e-8, 8000, 0.02

    def solve(prec, items, keep=False):
        sw = Sweep(device, precision=prec, conv_tol=conv, check_every=check_every, max_steps=cap)
        for ps in items:
            sw.add_spin(ps, dt)
        res, wall = sw.run(keep_states=keep)  # states are copied out after the timed loop
        return res, wall, sw.trace

    log("sweep: 96 instances, complex64 batched")
    res32, wall32, trace32 = solve('c64', spins, keep=True)
    log(f"sweep: c64 batched {wall32:.2f} s; complex128 batched (accuracy reference)")
    res64, wall64, _ = solve('c128', spins, keep=True)
    rel = np.array([abs(a.energy['E_tot'] - b.energy['E_tot']) / abs(b.energy['E_tot'])
                    for a, b in zip(res32, res64)])
    ok = np.array([a.converged and r <= 1e-6 for a, r in zip(res32, rel)])
    errs = [c64_vs_c128({'state': a.psik, 'energy': [a.energy['E_tot']], 'steps_total': a.steps},
                        {'state': b.psik, 'energy': [b.energy['E_tot']], 'steps_total': b.steps},
                        ps.space['dv_k']) for a, b, ps in zip(res32, res64, spins)]
    fid = np.array([e['state_fidelity'] for e in errs])
    pops = np.array([e['pops_err_of_atoms'] for e in errs])
    for r in res32 + res64:
        r.psik = None
    sample = []
    for grid in sorted({d.grid for d in draws}):
        sample += [i for i, d in enumerate(draws) if d.grid == grid][::fp64_stride]
    t_single, ok_single, by_grid = 0.0, 0, {}
    for i in sample:
        (one,), w, _ = solve('c128', [spins[i]])
        good = one.converged and abs(one.energy['E_tot'] / res64[i].energy['E_tot'] - 1) <= 1e-6
        t_single += w
        ok_single += int(good)
        g = by_grid.setdefault(str(draws[i].grid), {'instances': 0, 'solves': 0, 'wall_s': 0.0})
        g['instances'] += 1
        g['solves'] += int(good)
        g['wall_s'] = round(g['wall_s'] + w, 4)
    log(f"sweep: FP64 single-instance: {ok_single} solves of {len(sample)} sampled in {t_single:.2f} s")
    rate32 = int(ok.sum()) / wall32
    rate_single = ok_single / t_single
    steps = np.array([r.steps for r in res32])
    return {
        'workload': '96 Raman ground states (32x128^2, 48x256^2, 16x512^2), r=16, kL=3 rotated frame, '
                    'Omega~U[0,4]E_R, delta~U[-1,1]E_R, N~10^U[3,5], g_ud/g_uu~U[0.9,1.1], seed 2, '
                    'imaginary dt=0.02',
        'precision': 'c64 (f32 arithmetic, f64 reductions)', 'conv_tol_per_step': conv,
        'max_steps': cap, 'check_every': check_every,
        'wall_s': round(wall32, 3), 'converged': int(sum(r.converged for r in res32)),
        'solves_within_1e-6': int(ok.sum()), 'solves_per_s': round(rate32, 3),
        'steps_mean': float(steps.mean()), 'steps_max': int(steps.max()),
        'active_trace': [[round(t, 3), n, a] for t, n, a in trace32],
        'energy_rel_err_vs_c128': {'max': float(rel.max()), 'median': float(np.median(rel))},
        'state_fidelity_vs_c128': {'min_counted': float(fid[ok].min()), 'median': float(np.median(fid)),
                                   'min': float(fid.min()),
                                   'counted_above_0.999': int(np.sum(fid[ok] >= 0.999))},
        'pops_err_of_atoms_vs_c128': {'max_counted': float(pops[ok].max()), 'median': float(np.median(pops)),
                                      'max': float(pops.max())},
        'c128_batched': {'wall_s': round(wall64, 3), 'converged': int(sum(r.converged for r in res64)),
                         'solves_per_s': round(len(draws) / wall64, 3)},
        'fp64_single': {'sample': f'every {fp64_stride}th instance of each grid ({len(sample)} of '
                                  f'{len(draws)}, the set\'s grid mix)',
                        'wall_s': round(t_single, 3), 'solves_within_1e-6': ok_single,
                        'by_grid': by_grid, 'solves_per_s': round(rate_single, 3),
                        'note': 'one instance at a time, complex128, same stop test and cap; '
                                'solves counted as on the batched side'},
        'ratio_vs_fp64_single': round(rate32 / rate_single, 2),
    }


def serve_leg(device, n_requests, load):
    """Served stream: seeded requests on 128^2-1024^2 grids (p = .3/.4/.2/.1),
    70 % ground states to convergence, 30 % quenches (ground state, then
    delta += 0.5 E_R and 5000 real-time steps of dt = 1e-3), complex64 with
    FP64 reductions. A burst (all requests at t = 0) measures the capacity;
    then the same requests arrive as a Poisson stream at ``load`` x capacity
    and their arrival-to-result latencies are reported."""
    from spinor_gpe.engine import serve
    srv = serve.Server(device, 'c64')
    log("serve: warm-up (one request per grid size)")
    warm = [serve.Request(1000 + i, r.draw, 'ground', 0.0)
            for i, r in enumerate(serve.request_stream(4, None, seed=99, grids=((128, 1), (256, 1), (512, 1), (1024, 1))))]
    for i, g in enumerate((128, 256, 512, 1024)):
        warm[i].draw.grid = g
    srv.run(warm)
    reqs = serve.request_stream(n_requests, None, seed=3)
    log(f"serve: burst of {n_requests} requests")
    burst = srv.run(reqs)
    makespan = max(o.finished for o in burst)
    capacity = n_requests / makespan
    rate = load * capacity
    log(f"serve: capacity {capacity:.2f} solves/s; Poisson stream at {rate:.2f}/s")
    stream = serve.request_stream(n_requests, rate, seed=3)
    out = srv.run(stream)
    wall = max(o.finished for o in out)
    res = serve.summarize(out, wall)
    res.update({
        'workload': f'{n_requests} requests, grids 128/256/512/1024 at p=.3/.4/.2/.1, 70% ground state '
                    '(imaginary dt=0.02 to |dE/E|<=1e-8 per step or 8000 steps), 30% quench (+0.5 E_R '
                    'detuning, 5000 real-time steps dt=1e-3), Raman kL=3 rotated frame, seed 3',
        'precision': 'c64 (f32 arithmetic, f64 reductions)',
        'slots': {str(k): v for k, v in srv.slots.items()}, 'block_steps': srv.block,
        'capacity_solves_per_s': round(capacity, 3), 'arrival_rate_per_s': round(rate, 3),
        'load': load, 'burst': serve.summarize(burst, makespan)})
    acc = serve_accuracy(device, n_requests)
    res['ground_state_errors_vs_c128'] = acc
    # solves: ground states that met the stop test within 1e-6 of the
    # complex128 solve, and quenches whose ground-state phase met the stop
    # test (their real-time phase is not re-solved in complex128)
    good = set(acc['rids_within_1e-6']) | {o.rid for o in out if o.kind == 'quench' and o.converged}
    res['solves_within_1e-6'] = len(good)
    res['solves_within_1e-6_per_s'] = round(len(good) / wall, 3)
    return res


def serve_accuracy(device, n_requests):
    """Errors of the served complex64 ground states against complex128 solves
    of the same draws (same step, stop test and cap): the requests are served
    once more, untimed, keeping their states, and the ground-state ones are
    solved as a complex128 sweep."""
    from spinor_gpe.engine import serve
    from spinor_gpe.engine import workloads as wl
    from spinor_gpe.engine.sweep import Sweep
    log("serve: accuracy of the served ground states against complex128 solves")
    srv = serve.Server(device, 'c64', keep_states=True)
    reqs = serve.request_stream(n_requests, None, seed=3)
    draws = {r.rid: r.draw for r in reqs}
    ground = [o for o in srv.run(reqs) if o.kind == 'ground']
    sw = Sweep(device, precision='c128', conv_tol=srv.conv_tol, check_every=srv.block, max_steps=srv.max_steps)
    spins = [wl.raman_spinor(f'serve_ref{o.rid}', draws[o.rid]) for o in ground]
    for ps in spins:
        sw.add_spin(ps, srv.dt_imag)
    ref, _ = sw.run(keep_states=True)
    errs = [c64_vs_c128({'state': o.psik, 'energy': [o.energy['E_tot']], 'steps_total': o.steps_imag},
                        {'state': r.psik, 'energy': [r.energy['E_tot']], 'steps_total': r.steps},
                        ps.space['dv_k']) for o, r, ps in zip(ground, ref, spins)]
    both = [o.converged and r.converged for o, r in zip(ground, ref)]
    e = np.array([x['energy_rel_err'] for x in errs])
    within = [o.rid for o, x in zip(ground, errs) if o.converged and x['energy_rel_err'] <= 1e-6]
    f = np.array([x['state_fidelity'] for x in errs])
    p = np.array([x['pops_err_of_atoms'] for x in errs])
    return {'ground_requests': len(ground), 'both_converged': int(sum(both)), 'rids_within_1e-6': within,
            'energy_rel_err': {'median': float(np.median(e)), 'max': float(e.max()),
                               'within_1e-6': int(np.sum(e <= 1e-6))},
            'state_fidelity': {'median': float(np.median(f)), 'min': float(f.min())},
            'pops_err_of_atoms': {'median': float(np.median(p)), 'max': float(p.max())}}


def measured_traffic(kernel):
    """HBM bytes per dispatch of one kernel instance, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes summarized in the profiles folder (each
    counter scaled by what it reads for a 256 MiB copy); None if that kernel
    instance was not profiled."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            entry = json.load(f)['kernels'].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return entry['traffic_bytes'] if entry else None


def cpu_info():
    """The host CPU as lscpu reports it: model name, sockets, physical cores
    and hardware threads (logical CPUs), and the CPUs this process may use."""
    info = {}
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=20, check=False).stdout
        fields = dict(line.split(':', 1) for line in out.splitlines() if ':' in line)
        fields = {k.strip(): v.strip() for k, v in fields.items()}
        info['model'] = fields.get('Model name', platform.processor() or 'unknown')
        sockets = int(fields.get('Socket(s)', '1'))
        per_socket = int(fields.get('Core(s) per socket', '0'))
        info['sockets'] = sockets
        info['physical_cores'] = sockets * per_socket
        info['threads_per_core'] = int(fields.get('Thread(s) per core', '1'))
        info['logical_cpus'] = int(fields.get('CPU(s)', '0'))
    except (OSError, ValueError, subprocess.SubprocessError):
        info['model'] = platform.processor() or 'unknown'
    try:
        info['cpus_available_to_process'] = len(os.sched_getaffinity(0))
    except OSError:
        pass
    return info


def time_oracle(ps, dt, time_mode, budget_s):
    """Median seconds of the CPU oracle's full_step() over ~budget_s."""
    from oracle.reference_path import OraclePropagator
    prop = OraclePropagator(ps, dt, time=time_mode)
    prop.full_step()
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s or len(times) < 3:
        t0 = time.perf_counter()
        prop.full_step()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times)


def cpu_baseline(ps, budget_s):
    """Time the CPU oracle's full_step() on the same workload (bounded
    sample), and on the 256^2 ground-state workload for a third of that."""
    from spinor_gpe.engine import workloads as wl
    med, n = time_oracle(ps, 1e-3, 'real', budget_s)
    med256, n256 = time_oracle(wl.c0_spinor('c0cpu'), 0.02, 'imag', budget_s / 3)
    return {'value': 1.0 / med, 'unit': 'split-steps/s', 'cores': torch.get_num_threads(),
            'cores_note': 'torch intra-op threads used by the oracle (one per core given to the process)',
            'kind': 'port', 'cpu': cpu_info(),
            'sample': (f"512^2 Raman workload, complex128, {n} timed full_step() calls "
                       f"of the torch-CPU oracle (median {med * 1e3:.1f} ms)"),
            'at_256': {'value': 1.0 / med256, 'unit': 'split-steps/s',
                       'sample': (f"256^2 imaginary-time ground-state workload, complex128, {n256} "
                                  f"timed full_step() calls (median {med256 * 1e3:.1f} ms)")}}


def roofline(res, prec, n):
    """HBM roofline of the slower pass of an n x n single-instance run."""
    dom = 'row_pass' if res['row_ms'] >= res['col_ms'] else 'col_pass'
    ms = max(res['row_ms'], res['col_ms'])
    alg = 2 * 2 * n * n * {'c128': 16, 'c64': 8}[prec]
    achieved = alg / (ms * 1e-3) / 1e9
    row_e, row_rows, col_e, col_w = res['tuning']
    real = 'double' if prec == 'c128' else 'float'
    if dom == 'col_pass':
        kname = f"col_pass_kernel<{real}, {n}, {col_e}, {col_w}>"
    elif row_rows < 0:  # split row layout, -rows rows per workgroup
        kname = f"row_split_kernel<{real}, {n}, {row_e}, {-row_rows}>"
    else:
        kname = f"row_pass_kernel<{real}, {n}, {row_e}, {row_rows}>"
    if res['streaming'][0 if dom == 'row_pass' else 1]:
        kname += ' streaming'
    traffic = measured_traffic(kname)
    return {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4),
            'traffic': traffic, 'traffic_source': TRAFFIC_FILE if traffic else None,
            'traffic_note': ('rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch: bytes between the L2 and '
                             'the data fabric (Infinity Cache hits included, so not HBM-only), each '
                             'scaled by its reading for a copy of known size'),
            'kernel_instance': kname, 'avg_us': round(ms * 1e3, 3), 'algorithmic_bytes': alg,
            'col_pass_us': round(res['col_ms'] * 1e3, 3), 'row_pass_us': round(res['row_ms'] * 1e3, 3)}


def large_leg(args, device):
    """The 2048^2 workload (real time, dt = 5e-4, Raman kL = 3 with Omega =
    1.5 E_R in the lab frame, 1e6 atoms, 64 seeded vortices on component 0;
    workloads.c4_spinor) in complex128 and complex64: one instance large
    enough that the passes are bound by bandwidth rather than latency."""
    from spinor_gpe.engine import workloads as wl
    log("building the 2048^2 vortex workload")
    ps = wl.c4_spinor('c4bench')
    n = 2048
    out = {'workload': '2048x2048 real-time Raman-coupled GPE (kL=3, Omega=1.5 E_R, lab frame), '
                       '64 vortices on component 0, 1e6 atoms, dt=5e-4, 1 instance'}
    runs = {}
    for prec in args.precisions:
        log(f"timing 2048^2 {prec}: warmup {args.large_warmup}, steps {args.large_steps}")
        r = runs[prec] = time_engine(ps, device, prec, args.large_steps, args.large_warmup, args.block,
                                     args.profile_steps // 4, False, dt=5e-4)
        sps = args.large_steps / r['elapsed']
        out[prec] = {'value': round(sps, 2), 'unit': 'split-steps/s',
                     'ms_per_step': round(r['elapsed'] / args.large_steps * 1e3, 5),
                     'grid_point_steps_per_s': round(sps * n * n, 1), 'block_timing': r['blocks'],
                     'tuning_row_e_rows_col_e_w': list(r['tuning']), 'streaming_row_col': list(r['streaming']),
                     'roofline': roofline(r, prec, n), 'atoms_after_run': r['atoms']}
    if 'c64' in runs and 'c128' in runs:
        out['c64']['vs_c128'] = c64_vs_c128(runs['c64'], runs['c128'], ps.space['dv_k'])
    return out


# The reference's propagation benchmark (spinor_gpe/benchmarks/benchmark_prop.py:
# 28-40 grids; bench_data/TitanV_cuda_step.npz medians, ms per full_step).
SIZES = [(64, 64), (64, 128), (128, 128), (128, 256), (256, 256), (256, 512), (512, 512),
         (512, 1024), (1024, 1024), (1024, 2048), (2048, 2048), (2048, 4096), (4096, 4096)]
TITAN_V_MS = {12: 3.774, 13: 3.785, 14: 3.842, 15: 3.951, 16: 4.062, 17: 5.285, 18: 6.708,
              19: 12.375, 20: 22.370, 21: 41.160, 22: 79.821, 23: 158.277}


def sizes_leg(args, device):
    """The reference's own benchmark workload (benchmark_prop.py:48-67: 1e2
    atoms, 50/50/2000 Hz trap, g = (1, 1, 1.04), +/-8 oscillator lengths,
    Raman coupling set up with Omega = 0, imaginary time dt = 1/50,
    complex128) on its 13 grids from 64x64 to 4096x4096. Per grid: the
    reference's method (prop.full_step() timed call by call, here with a
    device synchronize after each call; median and MAD) and the engine's
    captured-graph rate (blocks of full steps timed with events; median and
    MAD of the per-step time)."""
    from spinor_gpe.pspinor import pspinor as spin
    w = 2 * np.pi * 50
    out = []
    for grid in SIZES:
        n_pts = grid[0] * grid[1]
        log(f"sizes: {grid[0]}x{grid[1]}")
        ps = spin.PSpinor(f'bp{grid[0]}x{grid[1]}', overwrite=True, atom_num=1e2,
                          omeg={'x': w, 'y': w, 'z': 40 * w}, g_sc={'uu': 1, 'dd': 1, 'ud': 1.04},
                          pop_frac=(0.5, 0.5), r_sizes=(8, 8), mesh_points=grid)
        ps.coupling_setup(wavel=790.1e-9, kin_shift=False)
        _, prop = ps.imaginary(1 / 50, 1, device, is_sampling=False, progress=False)
        calls = []
        for _ in range(args.sizes_calls):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            prop.full_step()
            torch.cuda.synchronize()
            calls.append(time.perf_counter() - t0)
        calls = np.array(calls[3:]) * 1e3
        plan = prop._plan  # pylint: disable=protected-access
        block = max(1, min(args.block, int(2e6 // n_pts)))
        plan.run(block, block=block)
        events = [torch.cuda.Event(enable_timing=True) for _ in range(args.sizes_blocks + 1)]
        events[0].record()
        for i in range(args.sizes_blocks):
            plan.run(block, block=block)
            events[i + 1].record()
        torch.cuda.synchronize()
        per = np.array([events[i].elapsed_time(events[i + 1]) / block for i in range(args.sizes_blocks)])
        size = int(round(np.log2(n_pts)))
        med_c, med_g = float(np.median(calls)), float(np.median(per))
        row = {'grid': list(grid), 'log2_points': size,
               'full_step_call_ms': {'median': round(med_c, 5),
                                     'mad': round(float(np.median(np.abs(calls - med_c))), 6)},
               'graph_ms_per_step': {'median': round(med_g, 5),
                                     'mad': round(float(np.median(np.abs(per - med_g))), 6),
                                     'steps_per_block': block},
               'split_steps_per_s': round(1e3 / med_g, 1)}
        if size in TITAN_V_MS:
            row['titan_v_ms'] = TITAN_V_MS[size]
            row['speedup_vs_titan_v'] = {'full_step_call': round(TITAN_V_MS[size] / med_c, 2),
                                         'graph': round(TITAN_V_MS[size] / med_g, 2)}
        out.append(row)
        del prop, plan
        torch.cuda.empty_cache()
    return {'workload': 'reference benchmark_prop.py: 1e2 atoms, r=8, g_ud=1.04, Raman set up with '
                        'Omega=0, imaginary dt=1/50, complex128', 'grids': out}


# The reference's fft_2d medians on a Titan V in ms by log2(points)
# (benchmark_fft.py, bench_data/TitanV_cuda_fft.npz; 4096^2 did not run).
TITAN_V_FFT_MS = {12: 0.1433, 13: 0.1177, 14: 0.1188, 15: 0.1438, 16: 0.1421, 17: 0.1171, 18: 0.1221,
                  19: 0.1259, 20: 0.1853, 21: 1.6537, 22: 2.8285, 23: 4.9641}


def fft_leg(args, device):
    """The reference's FFT benchmark (benchmark_fft.py:21-84): the set-up of
    benchmark_prop on its 13 grids, one imaginary step, then
    ``ttools.fft_2d(prop.psik, prop.space['dr'])`` timed call by call (median
    and MAD). The reference's timer brackets the asynchronous launches
    without a device synchronize; here every call is followed by one, so the
    figures include the transforms' completion. Also the device time per call
    from events around a batch of back-to-back calls."""
    from spinor_gpe.pspinor import pspinor as spin
    from spinor_gpe.pspinor import tensor_tools as ttools
    w = 2 * np.pi * 50
    out = []
    for grid in SIZES:
        n_pts = grid[0] * grid[1]
        log(f"fft: {grid[0]}x{grid[1]}")
        ps = spin.PSpinor(f'bf{grid[0]}x{grid[1]}', overwrite=True, atom_num=1e2,
                          omeg={'x': w, 'y': w, 'z': 40 * w}, g_sc={'uu': 1, 'dd': 1, 'ud': 1.04},
                          pop_frac=(0.5, 0.5), r_sizes=(8, 8), mesh_points=grid)
        ps.coupling_setup(wavel=790.1e-9, kin_shift=False)
        _, prop = ps.imaginary(1 / 50, 1, device, is_sampling=False, progress=False)
        psik, dr = prop.psik, prop.space['dr']
        for _ in range(3):
            ttools.fft_2d(psik, dr)
        calls = []
        for _ in range(args.fft_calls):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ttools.fft_2d(psik, dr)
            torch.cuda.synchronize()
            calls.append(time.perf_counter() - t0)
        calls = np.array(calls) * 1e3
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.fft_batch):
            ttools.fft_2d(psik, dr)
        e1.record()
        torch.cuda.synchronize()
        dev_ms = e0.elapsed_time(e1) / args.fft_batch
        size = int(round(np.log2(n_pts)))
        med = float(np.median(calls))
        row = {'grid': list(grid), 'log2_points': size,
               'fft_2d_call_ms': {'median': round(med, 5), 'mad': round(float(np.median(np.abs(calls - med))), 6)},
               'device_ms_per_call': round(dev_ms, 5)}
        if size in TITAN_V_FFT_MS:
            row['titan_v_ms'] = TITAN_V_FFT_MS[size]
            row['speedup_vs_titan_v'] = round(TITAN_V_FFT_MS[size] / med, 2)
        out.append(row)
        del prop, psik
        torch.cuda.empty_cache()
    return {'workload': 'reference benchmark_fft.py: benchmark_prop set-up, fft_2d of the two-component '
                        'k-space state, complex128', 'grids': out}


def step_leg(args, device):
    """The metric value: 512^2 split-steps/s in complex128 (and complex64)."""
    log("building the 512^2 initial state (256^2 ground state, 4000 imaginary steps)")
    ps = c1_spinor(device)
    n_pts = 512 * 512
    results = {}
    for prec in args.precisions:
        log(f"timing {prec}: warmup {args.warmup}, steps {args.steps}")
        results[prec] = time_engine(ps, device, prec, args.steps, args.warmup, args.block,
                                    args.profile_steps, args.tune)
    r = results['c128']
    sps = args.steps / r['elapsed']
    r64 = results.get('c64')
    line = {
        'metric': METRIC, 'value': round(sps, 2), 'unit': 'split-steps/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(r['elapsed'] / args.steps * 1e3, 5),
        'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': round(sps / TITAN_V_512, 2),
        'dtype': 'c128 (f64)', 'data': 'synthetic (seeded Thomas-Fermi/ground-state inputs)',
        'config': {'workload': '512x512 real-time Raman-coupled GPE, 1 instance, c128',
                   'grid': [512, 512], 'r_sizes': [24, 24], 'dt': 1e-3, 'omega_ER': 2.0,
                   'delta_ER': 0.5, 'kL': 3.0, 'frame': 'lab', 'atoms': 1e4,
                   'graph_block_steps': args.block, 'tuning_row_e_rows_col_e_w': list(r['tuning']),
                   'streaming_row_col': list(r['streaming']),
                   'baseline': 'Titan V 149.1 split-steps/s at 512^2, complex128 '
                               '(reference bench_data/TitanV_cuda_step.npz)'},
        'grid_point_steps_per_s': round(sps * n_pts, 1),
        'block_timing': r['blocks'],
        'roofline': roofline(r, 'c128', 512),
        'atoms_after_run': {p: res['atoms'] for p, res in results.items()},
    }
    if r64:
        sps64 = args.steps / r64['elapsed']
        line['c64'] = {'value': round(sps64, 2), 'ms_per_step': round(r64['elapsed'] / args.steps * 1e3, 5),
                       'grid_point_steps_per_s': round(sps64 * n_pts, 1), 'block_timing': r64['blocks'],
                       'roofline': roofline(r64, 'c64', 512),
                       'vs_c128': c64_vs_c128(r64, r, ps.space['dv_k']),
                       'dtype': 'c64 (f32 arithmetic, f64 reductions)'}
    if not args.no_cpu_baseline:
        log(f"CPU baseline: oracle full_step() for ~{args.cpu_budget:.0f} s")
        line['cpu_baseline'] = cpu_baseline(ps, args.cpu_budget)
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20000)
    ap.add_argument('--warmup', type=int, default=300)
    ap.add_argument('--block', type=int, default=100, help='full steps per captured graph')
    ap.add_argument('--profile-steps', type=int, default=200)
    ap.add_argument('--cpu-budget', type=float, default=15.0, help='seconds of CPU baseline sampling')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--tune', action='store_true',
                    help='re-measure the FFT shapes of the workload instead of using the shipped table')
    ap.add_argument('--precisions', nargs='+', default=['c128', 'c64'], choices=['c128', 'c64'])
    ap.add_argument('--legs', nargs='+', default=['step', 'large', 'sweep', 'serve'],
                    choices=['step', 'large', 'sweep', 'serve', 'sizes', 'fft'],
                    help='step: 512^2 split-step rate (the metric value); large: the 2048^2 '
                         'workload; sweep: batched ground states; serve: continuous-batching '
                         'request stream; sizes and fft (not run by default): the reference\'s '
                         'own propagation and fft_2d benchmarks on its 13 grids, beside its '
                         'Titan V times')
    ap.add_argument('--large-steps', type=int, default=1000)
    ap.add_argument('--sizes-calls', type=int, default=53, help='timed full_step() calls per grid (sizes leg)')
    ap.add_argument('--sizes-blocks', type=int, default=10, help='timed graph blocks per grid (sizes leg)')
    ap.add_argument('--fft-calls', type=int, default=50, help='timed fft_2d calls per grid (fft leg)')
    ap.add_argument('--fft-batch', type=int, default=20, help='back-to-back fft_2d calls per event pair (fft leg)')
    ap.add_argument('--large-warmup', type=int, default=100)
    ap.add_argument('--requests', type=int, default=48, help='requests of the served stream')
    ap.add_argument('--load', type=float, default=0.8, help='arrival rate / measured capacity')
    ap.add_argument('--sweep-check', type=int, default=100, help='full steps between convergence checks')
    ap.add_argument('--fp64-stride', type=int, default=4,
                    help='every n-th instance of each grid is timed alone in complex128 (FP64 baseline)')
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                    help='before timing, copy a MIB-sized device buffer once (a known-size '
                         'copy for scaling HBM counters in profiles)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine measures one GPU; --gpus {args.gpus} is not supported "
              "(instances shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py: no HIP device available", file=sys.stderr)
        sys.exit(1)
    os.environ.setdefault('SPINOR_GPE_DATA', tempfile.mkdtemp(prefix='sgpe_bench_'))
    if args.tune:
        os.environ['SGPE_TUNE_FILE'] = os.path.join(tempfile.mkdtemp(prefix='sgpe_tune_'), 'tune.json')
    device = 'cuda:0'
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy * 2 ** 20 // 4, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst
    line = step_leg(args, device) if 'step' in args.legs else {'metric': METRIC, 'value': None}
    if 'large' in args.legs:
        line['large'] = large_leg(args, device)
    if 'sizes' in args.legs:
        line['sizes'] = sizes_leg(args, device)
    if 'fft' in args.legs:
        line['fft'] = fft_leg(args, device)
    if 'sweep' in args.legs:
        line['sweep'] = sweep_leg(device, args.sweep_check, args.fp64_stride)
    if 'serve' in args.legs:
        line['serve'] = serve_leg(device, args.requests, args.load)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
