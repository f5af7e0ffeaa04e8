"""pmotion_RK3 on the GPU (qv_rk3_step) against the reference build's
trajectories (runge_kutta T): positions within 1e-15 cm, the Runge-Kutta
storage u1 within 1e-11 relative per point (the stages' velocities agree with
the reference's to 1e-12)."""
import numpy as np
import pytest

from conftest import gpu_params, load_golden, rel_err, rk3_names

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", rk3_names())
def test_rk3_trajectory_vs_golden(name):
    from qvort_amd import gpu
    g = load_golden(name)
    meta = g["meta"]
    print("case", name, meta["nsteps"], flush=True)
    with gpu.Context(gpu_params(meta["params"])) as c:
        c.upload(g["x0"], g["infront"], g["behind"])
        for _ in range(meta["nsteps"]):
            c.rk3_step()
            c.enforce_periodic()
        x, u, u1, _, _ = c.download()
    act = g["infront"] != 0
    assert np.max(np.abs(x - g["x_final"])[act]) <= 1e-15
    assert np.max(rel_err(u1, g["u1_final"], act)) <= 1e-11
