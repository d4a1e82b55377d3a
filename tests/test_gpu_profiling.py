"""Per-kernel event timing of DEER solves (pmcmc_ctx_set_profiling /
pmcmc_ctx_set_profiling_kinds): a kind mask records only its kinds, every
kind by default, nothing when off; a bad mask is an argument error.  The
profiled solve returns the same trajectory as the unprofiled one."""
import numpy as np
import pytest

from oracle import deer_oracle as O

pytestmark = pytest.mark.gpu


def test_profiling_kinds_select_the_recorded_kernels():
    import torch

    from ParallelMCMC import DEER, DensityModel, LogisticTarget, _lib

    rng = np.random.default_rng(5)
    N, D, T = 800, 12, 300
    Xd, y, _ = O.make_logreg_data(rng, N, D)
    model = DensityModel(dim=D, target=LogisticTarget(Xd, y))
    Xi, U = O.make_tape(rng, D, T, np.float32)
    x0 = (0.1 * rng.standard_normal(D)).astype(np.float32)
    rec = DEER.TapedRecursion(model, 0.02, Xi, U, precision="f32")
    ws = DEER.DEERWorkspace(D, T, _lib.PREC_F32)
    ctx = ws.ctx
    kw = dict(tol_abs=1e-6, tol_rel=1e-5, maxiter=100, jacobian="stoch_diag", return_info=True, workspace=ws,
              seed=3)

    def run(**prof):
        ctx.reset_kernel_times()
        if prof:
            ctx.set_profiling(**prof)
        S, info = DEER.solve(rec, x0, **kw)
        torch.cuda.synchronize()
        ctx.set_profiling(False)
        return S.cpu().numpy(), info, ctx.kernel_times()

    S0, info0, kt0 = run()
    assert all(n == 0 for n, _ in kt0.values())
    S1, info1, kt1 = run(on=True, kinds=["glm_pass"])
    print({k: n for k, (n, _) in kt1.items()}, flush=True)
    assert kt1["glm_pass"][0] == 2 * info1.iters and kt1["glm_pass"][1] > 0
    assert all(kt1[k][0] == 0 for k in kt1 if k != "glm_pass")
    S2, info2, kt2 = run(on=True)
    assert kt2["glm_pass"][0] == 2 * info2.iters
    assert kt2["iteration"][0] == info2.iters and kt2["tile_apply"][0] > 0
    assert info0.iters == info1.iters == info2.iters
    assert np.array_equal(S0, S1) and np.array_equal(S0, S2)
    with pytest.raises(Exception):
        _lib.check(_lib.load().pmcmc_ctx_set_profiling_kinds(ctx.h, 1 << 20), ctx.h)
