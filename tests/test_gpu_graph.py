"""A plan captured as a HIP graph (cmtj_plan_graph_create / _launch): every
replay equals a direct execute of the plan bit for bit, with the program's
own device work and synchronisations between the replays, and with the
output buffers overwritten before each replay (so a replay that skipped a
group would show)."""
from __future__ import annotations

import numpy as np
import pytest

import parity_cases
from cmtj_amd import Junction, SolverMode, workloads
from cmtj_amd.batch import JunctionBatch

pytestmark = pytest.mark.gpu


def _mixed(precision):
    js, kws = [], []
    for n in ("bilayer_iec_rk4", "trilayer_heun", "sot_dynamic_heun", "stt_drivers", "five_layer_chain",
              "rh_strip", "exotic_eulerheun"):
        j, kw = parity_cases.CASES[n](Junction)
        j.setNoiseSeed(31)
        js.append(j)
        kws.append(kw)
    for q in range(6):  # a few PIMM points as well
        js.append(workloads.pimm_junction())
        kws.append(dict(totalTime=(1 + q) * 0.05e-9, timeStep=1e-13, writeFrequency=1e-13))
    return JunctionBatch.from_junctions(js, [k["totalTime"] for k in kws], [k["timeStep"] for k in kws],
                                        [k["writeFrequency"] for k in kws],
                                        [k.get("solverMode", SolverMode.RK4) for k in kws],
                                        precision=precision, log="full")


def _snapshot(r, b):
    logs = [r.instance_log(i) for i in range(len(b.js))]
    nl = b.js["n_layers"]
    mag = r.final_m.cpu().numpy()
    return logs, [mag[i, : int(nl[i])].copy() for i in range(len(b.js))], r.status().copy()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_equal_direct_execution(precision):
    import torch

    b = _mixed(precision)
    groups = b.plan().info.n_kernels
    print(f"[graph] {precision}: {len(b.js)} instances in {groups} groups", flush=True)
    assert groups >= 4
    want = _snapshot(b.run(), b)
    b.capture()
    x = torch.randn(512, 512, device="cuda")
    for rep in range(3):
        log, mag, st = b._bufs
        log.fill_(float("nan"))
        mag.fill_(float("nan"))
        st.fill_(0xFF)
        y = (x @ x).sum()  # the program's own device work between replays
        torch.cuda.synchronize()
        r = b.replay()
        torch.cuda.synchronize()
        got = _snapshot(r, b)
        for lw, lg in zip(want[0], got[0]):
            assert lw.keys() == lg.keys()
            for k in lw:
                np.testing.assert_array_equal(lw[k], lg[k])
        for mw, mg in zip(want[1], got[1]):
            np.testing.assert_array_equal(mw, mg)
        assert want[2].tobytes() == got[2].tobytes()
        print(f"[graph] replay {rep}: bitwise equal ({float(y):.3e})", flush=True)
