"""The reference's own GPU test case (cuda/src/rhs_test.cu:21-266): 13^3
blocks at (-0.5, -0.5, -0.5) with spacing 1e-2/13/(b % 3 + 1), its input
recipe fake_initial_data (grUtils.cpp:1368-1620), KO sigma 0.4, one face flag
word for every block (argv[3] there). That harness prints the largest
pointwise relative error per variable and sets no threshold; here the GPU
right-hand side is held to the FP64 bound, 1e-12 normwise per variable
against the oracle, and the pointwise form is logged for information. The
spacing is tiny (7.7e-4 at most) against O(1-30) fields, so the second
derivatives are differences of nearly equal numbers: the stencils round alike
on both sides (same terms, same order, no contraction)."""
import numpy as np
import pytest

import oracle as O
import dgr
from gpu_util import fake_initial_data, gpu_rhs, harness_blocks, interior_values, log, normwise, oracle_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dgr.load()
    yield


@pytest.mark.parametrize("bflag", [0, 1, 63])
def test_harness_recipe_rhs_matches_oracle(bflag):
    log(f"reference harness recipe: 96 blocks of 13^3, bflag {bflag}")
    t = harness_blocks(96, bflag=bflag)
    st = fake_initial_data(t)
    p = dgr.default_params(ko_sigma=0.4)
    got = gpu_rhs(t, st, p)
    ref = O.bssnrhs_batch(st, t, oracle_params(p), nthreads=8)
    g, r = interior_values(got, t), interior_values(ref, t)
    e = normwise(g, r)
    with np.errstate(divide="ignore", invalid="ignore"):
        pw = np.nanmax(np.abs(g - r) / np.abs(r), axis=1)  # rhs_test.cu:245, for information
    worst = np.argsort(e)[::-1][:3]
    log(f"  normwise max {e.max():.3e} ({[(dgr.api.VAR_NAMES[v], f'{e[v]:.2e}') for v in worst]}); "
        f"pointwise relative max {np.nanmax(pw):.3e}")
    assert np.isfinite(g).all()
    assert e.max() <= 1e-12
