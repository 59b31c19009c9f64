"""Every compiled FFT shape of both split-step passes gives the oracle's
result (GPU): rows E x rows-per-block and columns E x columns-per-block for
each grid length, complex128 and complex64; the kernel variants for dense
operators (non-separable kinetic energy and potential, a coupling array) on
a subset of grid lengths. The 4096-point shapes (the largest grid) run one
step against the oracle, the others three.

Each separable shape also runs on a two-slot plan (the slot-table path:
state address read per slot, second slot holding a different state) with
normal and with non-temporal state accesses, both slots against their own
oracle runs."""

import numpy as np
import pytest

from cases import raman_lab
from oracle.reference_path import OraclePropagator
from spinor_gpe import _native as nat
from spinor_gpe.engine import autotune
from spinor_gpe.engine.plan import InstanceSpec, Plan

pytestmark = pytest.mark.gpu

GRIDS = [64, 128, 256, 512, 1024, 2048, 4096]
_REF = {}


def steps(n):
    return 1 if n >= 4096 else 3


GENERAL_GRIDS = [64, 256, 1024]


def make_case(n, general):
    ps, time, dt = raman_lab(name=f'shape{n}{"g" if general else ""}', mesh=(n, n), r=(24, 24))
    if general:
        x, y = ps.space['x_mesh'], ps.space['y_mesh']
        kx, ky = ps.space['kx_mesh'], ps.space['ky_mesh']
        ps.pot_eng = ps.pot_eng + 2.0 * np.exp(-((x - 1.0) ** 2 + y ** 2) / 4)
        ps.detuning_uniform(0.5 * ps.EL_recoil)
        ps.coupling_grad(0.2, 2.0, axis=0)
        ps.coupling = np.abs(ps.coupling)
        ps.kin_eng_spin = [k + 0.02 * kx * ky for k in ps.kin_eng_spin]
        ps.kin_eng_spin = [k - np.min(k) for k in ps.kin_eng_spin]
    return ps, time, dt


def reference(n, general=False, seed=1):
    if (n, general, seed) not in _REF:
        if seed == 1:
            ps, time, dt = make_case(n, general)
        else:  # the second slot: other phase noise, imaginary time
            ps, _, _ = raman_lab(name=f'shape{n}s{seed}', mesh=(n, n), r=(24, 24), seed=seed)
            time, dt = 'imag', 0.01
        o = OraclePropagator(ps, dt, time=time)
        o.run(steps(n))
        _REF[(n, general, seed)] = (ps, time, dt, np.stack(o.psik_numpy()))
    return _REF[(n, general, seed)]


def cases():
    """(n, precision, pass, shape, general, slots, streaming); streaming None
    keeps the plan's default."""
    out = []
    for n in GRIDS:
        for prec in ('c128', 'c64'):
            for pass_ in (0, 1):
                for shape in autotune.shapes(pass_, prec, n):
                    out.append((n, prec, pass_, shape, False, 1, None))
                    if n in GENERAL_GRIDS:
                        out.append((n, prec, pass_, shape, True, 1, None))
                    for stream in ((0,) if n >= 4096 else (0, 1)):
                        out.append((n, prec, pass_, shape, False, 2, stream))
    return out


@pytest.mark.parametrize('n,prec,pass_,shape,general,slots,stream', cases())
def test_shape_matches_oracle(gpu, n, prec, pass_, shape, general, slots, stream):
    print(f"[shape] n={n} {prec} pass={'row' if pass_ == 0 else 'col'} shape={shape} general={general} "
          f"slots={slots} streaming={stream}", flush=True)
    runs = [reference(n, general)] + ([reference(n, False, seed=8)] if slots == 2 else [])
    plan = Plan(gpu, prec, n, n, slots)
    cur = plan.tuning()
    try:
        if pass_ == 0:
            plan.tune(shape[0], shape[1], cur[2], cur[3])
        else:
            plan.tune(cur[0], cur[1], shape[0], shape[1])
    except nat.SgpeError as err:
        assert 'LDS' in str(err) or 'divide' in str(err)
        pytest.skip(f"shape not usable here: {err}")
    if stream is not None:
        plan.set_streaming(stream, stream)
    states = []
    for slot, (ps, time, dt, _) in enumerate(runs):
        st = plan.new_state(ps.psik)
        plan.set_slot(slot, InstanceSpec.from_spin(ps, time, dt), st)
        states.append(st)
    plan.full_steps(steps(n))
    plan.sync()
    for slot, (st, run) in enumerate(zip(states, runs)):
        want = run[3]
        got = st.cpu().numpy()
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < (1e-12 if prec == 'c128' else 2e-5), (slot, err)
