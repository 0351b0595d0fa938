"""check_temp (out.f90:1487-1557, echo.f90:189-190): the run stops at the
first step whose starting state has its maximum temperature (or energy
density) below temp_end; the stopped state does not move."""
import numpy as np
import pytest

from libs import oracle_get_stop, oracle_set_stop
from scenarios import gubser2d, rel_linf


def _tmax(oracle, v, freeze_type):
    mx = 0.0
    for p in v[4].ravel():
        e, out = oracle.eos_derived(1.0, p)
        mx = max(mx, out[1] if freeze_type == 0 else out[0])
    return mx


@pytest.mark.parametrize("freeze_type", [0, 1])
def test_oracle_stops_where_the_reference_loop_exits(oracle, freeze_type):
    ic, rc, v0 = gubser2d(24)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(3)
    v3, _ = oracle.state()
    m3 = _tmax(oracle, v3, freeze_type)
    oracle.steps(1)
    v4, _ = oracle.state()
    m4 = _tmax(oracle, v4, freeze_type)
    assert m4 < m3                       # the Gubser flow cools
    # threshold between the maxima after steps 3 and 4: the run makes steps
    # 1..4 and stops at the start of step 5
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle_set_stop(oracle, 0.5 * (m3 + m4), freeze_type)
    assert oracle.steps(10) == 0
    st, mx = oracle_get_stop(oracle)
    assert st == 1 and mx == m4
    assert oracle.scalars()["iter"] == 4
    v, _ = oracle.state()
    assert np.array_equal(v, v4)
    oracle.steps(5)                       # stays stopped
    v, _ = oracle.state()
    assert np.array_equal(v, v4) and oracle.scalars()["iter"] == 4


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
def test_gpu_stop_matches_oracle(oracle, fused):
    from echoamd import EchoGrid
    ic, rc, v0 = gubser2d(32)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle.steps(5)
    v5, _ = oracle.state()
    m5 = _tmax(oracle, v5, 0)
    oracle.steps(1)
    v6, _ = oracle.state()
    m6 = _tmax(oracle, v6, 0)
    thr = 0.5 * (m5 + m6)
    oracle.setup(ic, rc)
    oracle.init_from_prim(v0)
    oracle_set_stop(oracle, thr, 0)
    oracle.steps(12)
    g = EchoGrid(ic, rc)
    g.init_from_prim(v0)
    g.set_stop(thr, 0)
    if fused:
        g.steps(12)
    else:
        for _ in range(12):
            for irk in (1, 2, 3):
                g.rk_stage(irk)
            g.end_step()
    st, mx = g.stop_state()
    assert st == 1 and abs(mx - m6) <= 1e-12 * m6
    assert g.scalars()["iter"] == oracle.scalars()["iter"] == 6
    vg, ug = g.state()
    vo, uo = oracle.state()
    assert rel_linf(vg, vo) <= 1e-10 and rel_linf(ug, uo) <= 1e-10
    g.close()
