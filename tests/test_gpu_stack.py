"""The slice-stack configuration at its own size: 512 slices of 512^2, parallel beam,
720 views over 180 degrees x 729 bins, E = 20, Poisson counts at I0 = 1e5.  Fifty PNPG
iterations (npg/PNPG.m:50-167, non-negativity prox), each line-search trial replayed from
one captured HIP graph.  At every fifth kept iterate the device's f and grad_alpha of
four fixed slices are compared with the FP64 restatement of bhc/poissLAlpha.m over the
oracle's stored A (<= 1e-5 relative each; the errors are printed per iterate); the
accepted costs never increase."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_stack_pnpg_50_iterations(oracle_mod):
    import torch
    import bench
    dev = torch.device("cuda", 0)
    out, cd = bench.pnpg_leg(512, 50, dev, torch.cuda.current_stream(), compare=False)
    print({k: out[k] for k in ("iterations", "trials", "ms_per_trial", "cost_first", "cost_last",
                               "rejects", "restarts")})
    assert out["iterations"] == 50
    assert out["cost_monotone"] and out["cost_last"] < out["cost_first"]
    keep = list(range(0, len(cd["alphas"]), 5))
    for key in ("alphas", "f", "g"):
        cd[key] = [cd[key][i] for i in keep]
    res = bench.stack_fp64_check(cd)
    for i, ef, eg in zip(keep, res["f_rel_err"], res["grad_rel_l2"]):
        print(f"iterate {i}: f rel err {ef:.2e}, grad rel L2 {eg:.2e}")
    assert res["max_f_rel_err"] <= 1e-5 and res["max_grad_rel_l2"] <= 1e-5
    assert np.isfinite(out["t"]) and out["t"] > 0
