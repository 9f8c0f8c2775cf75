"""Helper process of tests/test_gpu_adjoint.py (not collected by pytest):
torch is imported and its GPU started before the HIP library is loaded, so
both run on torch's HIP runtime; the collapse adjoint then runs once on CPU
tensors and once on cuda:0 tensors, and one JSON line reports the comparison."""
import importlib
import json
import os
import sys

import numpy as np
import torch

torch.zeros(1, device="cuda:0")  # start torch's HIP runtime first

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
wr = importlib.import_module("warp-reactor_amd")
ag = importlib.import_module("warp-reactor_amd.autograd")


def main():
    n, nsteps = 4 ** 3, 20
    seed, sigma, t0, t1 = wr.RECIPES["uniform16"]
    cells = wr.make_recipe_states(n, 2, sigma, t0, t1)
    base = ag.state_of(cells)
    out, kept = {}, True
    for dev in ("cpu", "cuda:0"):
        s0 = torch.tensor(base, dtype=torch.float64, device=dev, requires_grad=True)
        theta = torch.zeros(wr.adjoint_ntheta(), dtype=torch.float64, device=dev,
                            requires_grad=True)
        with wr.Context(n) as ctx:
            sN = ag.collapse(theta, s0, ctx, nsteps=nsteps)
            ag.h2_log_t_loss(sN).backward()
        kept = kept and sN.device == s0.device and theta.grad.device == theta.device \
            and s0.grad.device == s0.device
        out[dev] = (sN.detach().cpu().numpy(), s0.grad.cpu().numpy(), theta.grad.cpu().numpy())

    def rel(a, b):
        d = np.abs(a - b)
        m = np.abs(b) > 0
        return float((d[m] / np.abs(b[m])).max()) if m.any() else float(d.max())

    print(json.dumps({
        "devices_kept": bool(kept),
        "one_runtime": len(ag._hip_runtimes()) == 1,
        "states_bit_identical": bool(np.array_equal(out["cpu"][0].view(np.uint64),
                                                    out["cuda:0"][0].view(np.uint64))),
        "s0_grad_max_rel": rel(out["cuda:0"][1], out["cpu"][1]),
        "theta_grad_max_rel": rel(out["cuda:0"][2], out["cpu"][2]),
    }))


if __name__ == "__main__":
    main()
