"""Every compiled FFT shape of both split-step passes gives the oracle's
result (GPU): rows E x rows-per-block and columns E x columns-per-block for
each grid length, complex128 and complex64."""

import numpy as np
import pytest

from cases import raman_lab
from oracle.reference_path import OraclePropagator
from spinor_gpe import _native as nat
from spinor_gpe.engine import autotune
from spinor_gpe.engine.plan import InstanceSpec, Plan

pytestmark = pytest.mark.gpu

GRIDS = [64, 128, 256, 512, 1024, 2048]
STEPS = 3
_REF = {}


def reference(n):
    if n not in _REF:
        ps, time, dt = raman_lab(name=f'shape{n}', mesh=(n, n), r=(24, 24))
        o = OraclePropagator(ps, dt, time=time)
        o.run(STEPS)
        _REF[n] = (ps, time, dt, np.stack(o.psik_numpy()))
    return _REF[n]


def cases():
    out = []
    for n in GRIDS:
        for prec in ('c128', 'c64'):
            for pass_ in (0, 1):
                for shape in autotune.shapes(pass_, prec, n):
                    out.append((n, prec, pass_, shape))
    return out


@pytest.mark.parametrize('n,prec,pass_,shape', cases())
def test_shape_matches_oracle(gpu, n, prec, pass_, shape):
    print(f"[shape] n={n} {prec} pass={'row' if pass_ == 0 else 'col'} shape={shape}", flush=True)
    ps, time, dt, want = reference(n)
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
    plan.full_steps(STEPS)
    plan.sync()
    got = st.cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < (1e-12 if prec == 'c128' else 2e-5), err
