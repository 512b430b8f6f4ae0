"""A mirgecom-style driver end to end on the GPU (examples/vortex.py's
structure): advance_state with RK4 and the prescribed-vortex boundary,
restart written from a post-step callback (mirgecom/restart.py:62-78) and
read back, the march continued from the restart bit for bit equal to an
uninterrupted march, and the VTK output of the device state
(mirgecom/simutil.py:309-386)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def test_restart_continues_march_exactly(tmp_path):
    from mirgecom_amd.euler import euler_operator
    from mirgecom_amd.fluid import ConservedVars
    from mirgecom_amd.gas_model import make_fluid_state
    from mirgecom_amd.integrators import rk4_step
    from mirgecom_amd.inviscid import inviscid_facial_flux_rusanov
    from mirgecom_amd.restart import read_restart_data, write_restart_file
    from mirgecom_amd.simutil import write_visfile
    from mirgecom_amd.steppers import advance_state
    from mirgecom_amd.visualization import make_visualizer
    from mirgecom_amd.workloads import vortex_2d

    wl, _ = vortex_2d(n=8, order=3)
    dcoll, gm, bnds = wl.dcoll, wl.gas_model, wl.boundaries
    dt = wl.dt

    def rhs(t, state):
        fs = make_fluid_state(state, gm)
        return euler_operator(dcoll, fs, gm, bnds, time=t,
                              inviscid_numerical_flux_func=inviscid_facial_flux_rusanov)

    rst = str(tmp_path / "restart_data" / "vortex-{step:06d}.pkl")
    vis = make_visualizer(dcoll)

    def post(state, step, t, dt):
        if step == 5:
            write_restart_file(None, {"cv": state, "t": t, "step": step,
                                      "order": 3}, rst.format(step=step))
            write_visfile(dcoll, [("cv", state)], vis,
                          vizname=str(tmp_path / "viz" / "vortex"), step=step,
                          t=t)
        return state, dt

    print("driver: 10 RK4 steps with a restart at step 5", flush=True)
    _, t_full, full = advance_state(rhs, rk4_step, wl.cv, t_final=1.0, dt=dt,
                                    max_steps=10, post_step_callback=post)
    data = read_restart_data(None, rst.format(step=5))
    cv5 = data["cv"]
    assert isinstance(cv5, ConservedVars) and cv5.q.device.type == "cuda"
    assert data["step"] == 5
    istep, t_b, cont = advance_state(rhs, rk4_step, cv5, t_final=1.0,
                                     t=data["t"], istep=data["step"], dt=dt,
                                     max_steps=10)
    assert istep == 10 and t_b == t_full
    assert torch.equal(cont.q, full.q)
    assert (tmp_path / "viz" / "vortex-000000005.pvtu").exists()
    assert (tmp_path / "viz" / "vortex-000000005-0000.vtu").exists()
