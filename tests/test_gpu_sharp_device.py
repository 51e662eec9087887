"""The device-controlled SharpClaw SSP33 loop (clawhip_sharp2_ssp33_run): whole
runs equal the host-driven loop bit for bit -- dt sequence, accept/reject,
final state, status, the indicator's flag cadence, positivity fallbacks and
audits -- and the FP64 run equals the oracle's."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle.driver import OracleRun
from tests._cases import seeded_state

pytestmark = pytest.mark.gpu


def _swe_run(device, precision="fp64", indicator=None, n=(48, 40), tfinal=0.06, nout=3,
             dt_initial=0.004, q=None, char_decomp=2, bc=(2, 2)):
    import pyclaw_amd as pyclaw
    from pyclaw_amd import riemann
    if q is None:
        q = seeded_state(np.random.default_rng(71), O.SHALLOW, n)
    solver = pyclaw.SharpClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.time_integrator = "SSP33"
    solver.char_decomp = char_decomp
    solver.cfl_desired, solver.cfl_max = 0.45, 0.5
    solver.dt_initial = dt_initial
    solver.precision = precision
    solver.indicator = indicator
    solver.device_loop = device
    solver.graph_chunk = 4
    for k in range(2):
        solver.bc_lower[k] = bc[k]
        solver.bc_upper[k] = bc[k]
    dom = pyclaw.Domain([0., 0.], [1., 1.], list(q.shape[1:]))
    sol = pyclaw.Solution(3, dom)
    sol.problem_data["grav"] = 1.0
    sol.q = q
    claw = pyclaw.Controller()
    claw.solution = sol
    claw.solver = solver
    claw.tfinal = tfinal
    claw.num_output_times = nout
    claw.keep_copy = True
    claw.output_format = None
    claw.run()
    return claw


def _same(a, b):
    for fa, fb in zip(a.frames, b.frames):
        assert fa.t == fb.t
        assert np.array_equal(fa.state.q, fb.state.q), np.abs(fa.state.q - fb.state.q).max()
    sa, sb = a.solver, b.solver
    assert sa.status["numsteps"] == sb.status["numsteps"]
    assert sa.status["cflmax"] == sb.status["cflmax"]
    assert sa.dt == sb.dt
    assert sa.cfl.get_cached_max() == sb.cfl.get_cached_max()


@pytest.mark.parametrize("precision", ["fp64", "fp64fast", "fp32k"])
def test_device_loop_equals_host_loop(precision):
    """dt_initial 0.05 makes the first attempt exceed cfl_max: the rejected
    attempt, the retry's get_dt without stage 1 and the graph-replayed chunks
    all take the host loop's decisions."""
    print("sharp device loop", precision, flush=True)
    dev = _swe_run(True, precision, dt_initial=0.05)
    host = _swe_run(False, precision, dt_initial=0.05)
    _same(dev, host)
    assert dev.solver.status["numsteps"] > 8


def test_device_loop_fp64_equals_oracle():
    print("sharp device loop vs oracle", flush=True)
    q = seeded_state(np.random.default_rng(71), O.SHALLOW, (48, 40))
    dev = _swe_run(True, q=q, dt_initial=0.05)
    run = OracleRun(O.make_rp(O.SHALLOW, efix=1, grav=1.0), q, [1.0 / 48, 1.0 / 40], [2, 2],
                    [2, 2], solver="sharp", time_integrator="SSP33", char_decomp=2,
                    cfl_desired=0.45, cfl_max=0.5, dt_initial=0.05)
    run.run(0.06, 3)
    assert dev.solver.status["numsteps"] == run.numsteps
    assert dev.solution.t == run.t
    assert np.array_equal(dev.frames[-1].state.q, run.q)


def test_device_loop_indicator_equals_host_loop():
    """The learned indicator's path: flags refreshed on the device every second
    attempt, the full-WENO5 pass over the flagged interfaces, audits every 5
    judged attempts on the host path between device runs."""
    from pyclaw_amd import indicator as I
    print("sharp device loop, indicator", flush=True)
    runs = []
    for device in (True, False):
        ind = I.TroubledInterfaceIndicator()
        ind.audit_every = 5
        runs.append((_swe_run(device, "fp32k", ind, n=(64, 56), tfinal=0.08), ind))
    (dev, di), (host, hi) = runs
    _same(dev, host)
    assert di.judge_stats == hi.judge_stats
    assert di.judge_stats["audits"] >= 2
    assert di._age == hi._age


def test_device_loop_positivity_fallback_equals_host_loop():
    """An indicator that flags nothing on a strong dam break next to a nearly dry
    region: the linear weights overshoot to a negative depth, the judge recomputes
    those steps with full WENO5 -- on the device as on the host."""
    from pyclaw_amd import indicator as I
    mx, my = 64, 32
    q = np.zeros((3, mx, my), order="F")
    q[0] = 0.02
    q[0, : mx // 3] = 1.0
    print("sharp device loop, positivity fallback", flush=True)
    runs = []
    for device in (True, False):
        ind = I.TroubledInterfaceIndicator(threshold=1e30)
        ind.audit_every = 0
        runs.append((_swe_run(device, "fp64", ind, q=q, tfinal=0.04, dt_initial=0.002,
                              char_decomp=0, bc=(1, 1)), ind))
    (dev, di), (host, hi) = runs
    _same(dev, host)
    assert di.judge_stats["positivity_fallbacks"] == hi.judge_stats["positivity_fallbacks"] > 0
    assert np.isfinite(dev.frames[-1].state.q).all()


def test_run_steps_graph_replay_matches_eager():
    """run_steps with graph-captured chunks equals the same attempts launched
    one by one."""
    import pyclaw_amd as pyclaw
    from pyclaw_amd import indicator as I
    out = []
    for graph in (True, False):
        ind = I.TroubledInterfaceIndicator()
        ind.audit_every = 0
        solver = I.swe_solver("fp32k", ind)
        solver.use_graph = graph
        solver.graph_chunk = 4
        sol = I.swe_solution(3, 96, 80)
        solver.dt = 0.45 / 96 / 1.6
        solver.setup(sol)
        print("run_steps graph", graph, flush=True)
        solver.run_steps(sol, 13)
        out.append((sol.state.q.copy(), sol.t, solver.status["numsteps"]))
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2] == 13
    assert np.array_equal(out[0][0], out[1][0])
