"""Every compiled FFT shape of both split-step passes gives the oracle's
result (GPU): rows E x rows-per-block and columns E x columns-per-block for
each grid length, complex128 and complex64; the kernel variants for dense
operators (non-separable kinetic energy and potential, a coupling array) on
a subset of grid lengths. The 4096-point shapes (the largest grid) run one
step against the oracle, the others three."""

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


def reference(n, general=False):
    if (n, general) not in _REF:
        ps, time, dt = make_case(n, general)
        o = OraclePropagator(ps, dt, time=time)
        o.run(steps(n))
        _REF[(n, general)] = (ps, time, dt, np.stack(o.psik_numpy()))
    return _REF[(n, general)]


def cases():
    out = []
    for n in GRIDS:
        for prec in ('c128', 'c64'):
            for pass_ in (0, 1):
                for shape in autotune.shapes(pass_, prec, n):
                    out.append((n, prec, pass_, shape, False))
                    if n in GENERAL_GRIDS:
                        out.append((n, prec, pass_, shape, True))
    return out


@pytest.mark.parametrize('n,prec,pass_,shape,general', cases())
def test_shape_matches_oracle(gpu, n, prec, pass_, shape, general):
    print(f"[shape] n={n} {prec} pass={'row' if pass_ == 0 else 'col'} shape={shape} general={general}",
          flush=True)
    ps, time, dt, want = reference(n, general)
    spec = InstanceSpec.from_spin(ps, time, dt)
    plan = Plan(gpu, prec, n, n, 1)
    cur = plan.tuning()
    try:
        if pass_ == 0:
            plan.tune(shape[0], shape[1], cur[2], cur[3])
        else:
            plan.tune(cur[0], cur[1], shape[0], shape[1])
    except nat.SgpeError as err:
        assert 'LDS' in str(err) or 'divide' in str(err)
        pytest.skip(f"shape not usable here: {err}")
    st = plan.new_state(ps.psik)
    plan.set_slot(0, spec, st)
    plan.full_steps(steps(n))
    plan.sync()
    got = st.cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < (1e-12 if prec == 'c128' else 2e-5), err
