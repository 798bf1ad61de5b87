"""cs_calc_time_step_noalloc (time step with a static per-GPU scratch, no
allocation, no synchronisation) against cs_calc_time_step on the same
inputs: two calls in a row, the second on a smaller grid, so the second
reuses the scratch the first left behind."""
import numpy as np
import pytest

from gpu_util import case, dev


def _fields(dims, seed):
    n = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(seed)
    xim, etm, zem = (rng.uniform(0, 1, 3 * n) for _ in range(3))
    de = rng.uniform(0, 1, 5 * n)
    yaco, ssk = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    return [dev(a) for a in (xim, etm, zem, de, yaco, ssk)]


@pytest.mark.gpu
def test_time_step_noalloc_matches_time_step():
    import torch
    import compact_schemes_amd as cs
    umf = (0.3, 0.1, -0.2)
    out3 = torch.full((3,), float("nan"), dtype=torch.float64, device="cuda")
    # 256x128x64 needs all 2048 partial pairs; 40x24x16 needs 60
    for dims, seed in (((256, 128, 64), 21), ((40, 24, 16), 22)):
        f = _fields(dims, seed)
        case(f"time step noalloc {dims}")
        ph = cs.Physics(dims, True, umf)
        dte, m2 = ph.calc_time_step(*f, 0.95)
        cs.calc_time_step_noalloc(dims, umf, *f, 0.95, out3)
        got = out3.cpu().numpy()
        assert got[1] == m2[0] and got[2] == m2[1]       # a max is exact
        assert got[0] == dte
        out3.fill_(float("nan"))


@pytest.mark.gpu
def test_time_step_noalloc_rejects_bad_arguments():
    import torch
    import compact_schemes_amd as cs
    dims = (8, 8, 8)
    f = _fields(dims, 23)
    out3 = torch.zeros(3, dtype=torch.float64, device="cuda")
    with pytest.raises(cs.CsError, match="CS_ERR_ARG"):
        cs.calc_time_step_noalloc(dims, (0, 0, 0), *f, 0.95, None)
    with pytest.raises(cs.CsError, match="CS_ERR_SHAPE"):
        cs.calc_time_step_noalloc((0, 8, 8), (0, 0, 0), *f, 0.95, out3)
