"""Graph mode of the resident step (cube_p3m_graph_enable): a full kick's
compute passes -- PM1+PM2, then PM3+PP+buffer_v, whose lists are built on a
second stream meanwhile -- are replayed from two HIP graphs captured once per
launch signature, with the kick coefficient read from device memory.

Two contexts start from the same seeded Zel'dovich input; one steps with the
graph off, the other with it on.  The step length changes from step to step,
so a replay that kept the captured coefficient would differ.  After every
step, with downloads and synchronisations of the program's own between the
replays, the resident states (coordinates, velocities, IDs, counts) and the
drift and kick outputs agree bit for bit, and the graph context reports its
graph-run kicks and captures (fewer captures than replays)."""
import numpy as np
import pytest

from cube_nu_amd import P3M

pytestmark = pytest.mark.gpu

NG = 64


@pytest.mark.parametrize("izipx,prec,ws", [(2, 32, 0), (1, 32, 0), (2, 32, 1)],
                         ids=["int16-fp32", "int8-fp32", "per-tile-batches"])
def test_graph_steps_bit_identical_to_uncaptured(izipx, prec, ws):
    """ws = 1: a one-byte workspace splits every tile level into batches of
    one mesh, which a captured kick does not hold; the compute pass then runs
    uncaptured and the results are the same."""
    ps = [P3M(NG, izipx=izipx, precision=prec, box=600.0 * NG / 512, workspace_bytes=ws) for _ in range(2)]
    try:
        ps[1].graph(True)
        for p in ps:
            p.timing(True)  # the graph path holds its timing scopes out of the capture
        ics = [p.make_ic(seed=7, z=1.0, nic=2) for p in ps]
        assert ics[0] == ics[1]
        a = 0.5
        k0 = [p.kick_resident(a_mid=a, dt=0.0) for p in ps]
        assert k0[0] == k0[1]
        dt0 = min(k0[0]["dt_pm1"], k0[0]["dt_pm2"], k0[0]["dt_pm3"], k0[0]["dt_pp"], 1.0) * 2
        vz = [ics[0]["vz_max"]] * 2
        dt_old = dt0
        nsteps = 5
        for step in range(nsteps):
            dt = dt0 * (1.0 + 0.1 * step)
            r = [p.step_resident(dt_old=dt_old, dt=dt, vz_max=vz[i], a_mid=a + 0.01 * step) for i, p in enumerate(ps)]
            assert r[0][0] == r[1][0], f"step {step}: drift outputs"
            assert r[0][1] == r[1][1], f"step {step}: kick outputs"
            s0, s1 = ps[0].download(), ps[1].download()
            for name, x, y in zip(("xp", "vp", "pid", "rhoc"), s0, s1):
                assert np.array_equal(x, y), f"step {step}: {name}"
            vz = [r[0][1]["vz_max"], r[1][1]["vz_max"]]
            dt_old = dt
        rep = ps[1].timing_report()
        launches, captures = rep["graph_launches"]["calls"], rep["graph_captures"]["calls"]
        print(f"izipx {izipx} fp{prec} ws {ws}: {launches} graph launches, {captures} captures")
        assert ps[0].timing_report()["graph_launches"]["calls"] == 0
        if ws:
            assert launches == 0 and captures == 0
        else:
            # the kick's two parts (PM1+PM2, PM3+PP+buffer_v) are captured at
            # their first uses -- the second part once its autotuner trials
            # have run uncaptured, and again when buffer_v joins it -- then
            # replayed
            assert launches >= nsteps - 1
            assert 2 <= captures <= 4
    finally:
        for p in ps:
            p.close()


def test_graph_kicks_fp64_match_uncaptured():
    """The FP64 form accumulates its tile deposits in FP64 LDS atomics, whose
    order follows the scheduling, so two runs agree to rounding rather than
    bit for bit: replayed FP64 kicks (coefficient changing between them) match
    uncaptured ones to 1e-12 relative."""
    from inputs import buffered_set, rel_rms
    from cube_nu_amd.p3m import Geometry

    geo = Geometry(NG)
    xp, vp, pid, rhoc = buffered_set(geo, seed=13, mean=6.0)
    ps = [P3M(NG, precision=64) for _ in range(2)]
    try:
        ps[1].graph(True)
        for p in ps:
            p.upload(xp, vp, pid, rhoc)
        for i, dt in enumerate((0.2, 0.3, 0.25)):
            k = [p.kick_resident(a_mid=0.5 + 0.01 * i, dt=dt) for p in ps]
            v = [p.download()[1] for p in ps]
            err = rel_rms(v[1] - vp, v[0] - vp)
            print(f"fp64 kick {i}: rms rel difference {err:.3e}")
            assert err <= 1e-12
            assert k[0]["napm"] == k[1]["napm"]
            for key in ("dt_pm1", "dt_pm2", "dt_pm3", "dt_pp", "vz_max"):
                assert abs(k[0][key] - k[1][key]) <= 1e-6 * abs(k[0][key]), key
        rep = ps[1].timing_report()
        # the FP64 form tunes nothing: both parts are captured at the first
        # kick and replayed by the next two
        assert rep["graph_launches"]["calls"] == 3 and rep["graph_captures"]["calls"] == 2
    finally:
        for p in ps:
            p.close()
