"""The server with one 2D lane per PN order (four 2D lanes and the 1D lane).

The number of 2D lanes follows the runtime's hardware queue count
(GPU_MAX_HW_QUEUES), which the runtime reads when it initialises, so the check
runs in a child process started with six queues.  Every served result must
equal the same request solved alone."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, math, sys
sys.path.insert(0, {root!r})
import numpy as np
from pnepma_amd import serving
from pnepma_amd.problem import solve_inputs
from pnepma_amd.solver import DeviceProblem

specs = serving.serving_requests(16, seed=21, sizes_1d=(30, 70), sizes_2d=(10, 20), max_steps=36)
preps = [serving.prepare(s, want_rho_bar=True) for s in specs]
srv = serving.Server(hbm_budget_bytes=1 << 30, max_active=64)
tickets = [srv.submit(pr, arrival_s=0.001 * n) for n, pr in enumerate(preps)]
for t in tickets:
    assert srv.wait(t).status == 0
srv.close()
worst = 0.0
orders = set()
for spec, pr in zip(specs, preps):
    p = spec.problem
    orders.add((p.N, int(p.gspec.axes[1].M > 1)))
    K = max(1, math.ceil(math.sqrt(p.n_steps)))
    dp = DeviceProblem(p, solve_inputs(p, spec.material, spec.detector), ckpt_every=K)
    I, rb, gen = dp.forward_adjoint(spec.I_bar[None, :])
    dp.close()
    def rel(a, b):
        s = np.max(np.abs(b))
        return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a)))
    worst = max(worst, rel(pr.I_out, I[0]), rel(pr.gen, gen))
    for c in p.gspec.active_classes():
        worst = max(worst, rel(pr.rho_bar[c], rb[c]))
print(json.dumps({{"worst": worst, "orders": sorted(orders)}}), flush=True)
"""


def test_served_equals_alone_four_lanes():
    env = dict(os.environ, GPU_MAX_HW_QUEUES="6")
    res = subprocess.run([sys.executable, "-c", CHILD.format(root=ROOT)], env=env, capture_output=True, text=True,
                         timeout=240)
    assert res.returncode == 0, res.stderr[-3000:]
    out = json.loads(res.stdout.strip().splitlines()[-1])
    print("four-lane server vs alone:", out, flush=True)
    # every PN order with a 2D grid present, so each of the four 2D lanes ran
    assert {n for n, is2d in out["orders"] if is2d} == {3, 5, 7, 9}
    assert out["worst"] < 1e-12
