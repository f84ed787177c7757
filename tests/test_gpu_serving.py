"""Serving on the GPU: sliced runs equal uninterrupted ones, and a request
gives the same result alone and inside a mixed stream."""
from __future__ import annotations

import time

import numpy as np
import pytest

import parity_cases
from cmtj_amd import Junction, SolverMode, workloads
from cmtj_amd.batch import JunctionBatch
from cmtj_amd.serving import Server

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sliced_run_equals_uninterrupted(precision):
    js, kws = [], []
    for n in ("bilayer_iec_rk4", "trilayer_heun", "sot_dynamic_heun", "stt_drivers", "five_layer_chain",
              "rh_strip"):
        j, kw = parity_cases.CASES[n](Junction)
        js.append(j)
        kws.append(kw)
    def mk():
        for j in js:  # each packing takes the next noise stream: start both batches from the same one
            j.setNoiseSeed(11)
        return JunctionBatch.from_junctions(js, [k["totalTime"] for k in kws], [k["timeStep"] for k in kws],
                                            [k["writeFrequency"] for k in kws],
                                            [k.get("solverMode", SolverMode.RK4) for k in kws],
                                            precision=precision, log="full").set_lanes("per_layer")

    a = mk().run()
    b = mk().run_sliced(slice_steps=777, n_streams=3)
    la, lb = a.log.cpu().numpy(), b.log.cpu().numpy()
    print(f"[slices] {precision}: max log difference {np.max(np.abs(la - lb)):.3e}", flush=True)
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(a.final_m.cpu().numpy(), b.final_m.cpu().numpy())
    sa, sb = a.status(), b.status()
    assert (sa["steps_done"] == sb["steps_done"]).all() and (sa["entries_done"] == sb["entries_done"]).all()


def _alone(spec, precision):
    kind, j, T, dt, wf, sm, log = spec
    b = JunctionBatch.from_junctions([j], T, dt, wf, sm, precision=precision, log=log).set_lanes("per_component")
    r = b.run()
    return r.instance_log(0), r.final_m.cpu().numpy()[0, : j.layerNo]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_results_are_the_same_alone_and_in_the_mix(precision):
    mix = workloads.serving_mix(24, seed=5, t_min=0.05e-9, t_max=0.4e-9)
    with Server(precision=precision, slice_steps=1500, n_streams=4, log_elements=1 << 22) as srv:
        futs = []
        for q, (kind, j, T, dt, wf, sm, log) in enumerate(mix):
            j.setNoiseSeed(1000 + q)  # the same (seed, stream) as the run alone below
            futs.append(srv.submit(j, T, dt, wf, sm, log=log))
            if q % 6 == 5:
                time.sleep(0.05)  # later requests join while earlier ones run
        res = [f.result(timeout=300) for f in futs]
        ticks = srv.ticks
    worst = 0.0
    for q, (spec, r) in enumerate(zip(mix, res)):
        spec[1].setNoiseSeed(1000 + q)
        lg, m = _alone(spec, precision)
        assert set(lg) == set(r.log)
        for k in lg:
            worst = max(worst, float(np.max(np.abs(lg[k] - r.log[k]))) if lg[k].size else 0.0)
        worst = max(worst, float(np.max(np.abs(m - r.final_m))))
    print(f"[serving] {precision}: {len(mix)} requests in {ticks} ticks, max difference alone vs mix "
          f"{worst:.3e}", flush=True)
    assert worst <= 1e-12


def test_a_failing_request_does_not_disturb_the_others():
    from cmtj_amd import AxialDriver, CVector, Layer, constantDriver, stepDriver

    bad = Junction([Layer("free", mag=CVector(1, 0, 0), anis=CVector(1, 0, 0), Ms=1.0, thickness=1e-9,
                          cellSurface=1e-16, demagTensor=parity_cases.DEMAG, damping=0.01)])
    bad.setLayerExternalFieldDriver("all", AxialDriver(stepDriver(0, float("inf"), 5e-12, 1e-9),
                                                       constantDriver(0), constantDriver(1e5)))
    good = [workloads.pimm_junction() for _ in range(3)]
    with Server(precision="fp64", slice_steps=500, log_elements=1 << 20) as srv:
        fb = srv.submit(bad, 1e-10, 1e-13, 1e-13)
        fg = [srv.submit(j, 2e-10, 1e-13, 1e-13, log="mx") for j in good]
        with pytest.raises(RuntimeError, match="NAN magnetisation") as ei:
            fb.result(timeout=120)
        assert 0 < len(ei.value.partial_log["time"]) < 1000
        rs = [f.result(timeout=120) for f in fg]
    for r in rs:
        assert r.steps == 2000 and np.all(np.isfinite(r.log["free_mx"]))


def test_admission_waits_for_log_memory():
    """An arena that holds one request at a time serialises admission; every
    request still completes."""
    js = [workloads.pimm_junction() for _ in range(4)]
    need = 3 * 1000  # mx of 2 layers + R, 1000 entries
    with Server(precision="fp32", slice_steps=400, log_elements=need + 10) as srv:
        futs = [srv.submit(j, 1e-10, 1e-13, 1e-13, log="mx") for j in js]
        rs = [f.result(timeout=120) for f in futs]
    for a, b in zip(rs, rs[1:]):
        assert b.t_admit >= a.t_done - 1e-3  # admitted only after the previous one left
    assert all(r.steps == 1000 for r in rs)
