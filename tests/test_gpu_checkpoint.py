"""Resume from a checkpoint on the GPU path: RK4 steps with the full padding
refresh on the refined-octree sample, a checkpoint written from the device
state midway (dgr.checkpoint, the reference's file set and index rule), and
the run restored from it into a new plan continues with the same bits as the
run that was not interrupted (the padded state is saved whole, so the restored
state needs no refresh before its first right-hand side)."""
import numpy as np
import pytest

import dgr
from dgr import checkpoint as CK
from dgr import workloads as W
from gpu_util import log, to_dev
from test_oracle_halo import smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_restore_continues_bit_identically(tmp_path, prec):
    import torch
    t = W.two_level_lattice()
    st = smooth_state(t, amp=1e-2)
    p = dgr.default_params(ko_sigma=0.4)
    dt = 0.02
    prefix = str(tmp_path / "bssn")

    def plan_for(blocks, n):
        plan = dgr.Plan(blocks, n)
        plan.set_precision(prec)
        plan.set_halo_refresh("full")
        return plan

    plan = plan_for(t, st.shape[1])
    u = to_dev(st)
    plan.rk4_steps(u, dt, p, nsteps=2)
    torch.cuda.synchronize()
    idx, _ = CK.checkpoint_index(2, 1, ((3.0, 0.0, 0.0), (-3.0, 0.0, 0.0)), False)
    CK.write_checkpoint(prefix, idx, u, t, time_begin=0.0, time_end=1.0, time=2 * dt, step=2, dt=dt,
                        bh=((3.0, 0.0, 0.0), (-3.0, 0.0, 0.0)))
    plan.rk4_steps(u, dt, p, nsteps=3)
    torch.cuda.synchronize()
    straight = u.cpu().numpy()
    plan.close()

    ridx, cp, blocks, state = CK.restore_checkpoint(prefix)
    assert ridx == idx and int(cp["DENDRO_TS_STEP_CURRENT"]) == 2
    assert np.array_equal(blocks, t)
    plan = plan_for(blocks, state.shape[1])
    u = to_dev(state)
    plan.rk4_steps(u, float(cp["DENDRO_TS_TIME_STEP_SIZE"]), p, nsteps=3)
    torch.cuda.synchronize()
    resumed = u.cpu().numpy()
    plan.close()
    log(f"resume ({prec}): largest difference {np.abs(resumed - straight).max():.3e}")
    assert np.array_equal(resumed, straight)
