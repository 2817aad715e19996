"""The uneven batch (config-4 table at full scale, seeded breast phantoms and
Poisson counts, I0 5e3) against the CPU oracle: three of its cases, the
smallest (800x400x30, 9 views, 3 subsets), the 25-view one (971x486x45, 5
subsets) and the largest (1400x700x80, 15 views, 5 subsets), run as one
batched launch per kernel; after each case's first full iteration its image
is compared with the oracle's OSTR iteration on the same inputs. FP32 subset
sinograms: the parity bar (1e-4 relative RMSE); BF16 subset sinograms (h):
their error budget against the same FP32 oracle (5e-3)."""
import numpy as np
import pytest
import torch

from newrecon_amd.batch import BatchOSTR
from newrecon_amd.geometry import OSTRConfig, scan_batch
from newrecon_amd.ostr import OSTR
from newrecon_amd.workload import make_measurement

pytestmark = pytest.mark.gpu

PICK = (0, 2, 7)


def rel_rmse(a, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="module")
def batch_cases(dev, oracle_lib):
    scans = [scan_batch()[k] for k in PICK]
    data = []
    for s in scans:
        _, y, b = make_measurement(s, device=dev)
        data.append((s, y.cpu().numpy(), b.cpu().numpy()))
        print(f"[batch-oracle] {s.name}: data", flush=True)
    refs = []
    for s, y, b in data:
        ref, _ = oracle_lib.run(s.geom, s.angles, y, b, n_subsets=s.recon.n_subsets, n_iters=1,
                                beta=1e3, delta=5e-4, potential=2, gamma_mode=1, guard=True,
                                nthreads=16)
        refs.append(ref)
        print(f"[batch-oracle] {s.name}: oracle iteration", flush=True)
    return data, refs


@pytest.mark.parametrize("prj_dtype,bound", [(torch.float32, 1e-4), (torch.bfloat16, 5e-3)])
def test_batched_iteration_matches_oracle(batch_cases, dev, prj_dtype, bound):
    data, refs = batch_cases
    recs = [OSTR(s.geom, s.angles, y, b, cfg=OSTRConfig(n_subsets=s.recon.n_subsets, beta=1e3,
                                                         delta=5e-4), prj_dtype=prj_dtype,
                 device=dev) for s, y, b in data]
    bt = BatchOSTR(recs)
    errs = {}
    for step in range(1, max(r.M for r in recs) + 1):
        bt.step(1)
        for k, r in enumerate(recs):
            if step == r.M:  # the case's first full iteration is done
                torch.cuda.synchronize()
                errs[data[k][0].name] = rel_rmse(r.x.cpu().numpy(), refs[k])
    bt.close()
    print(f"[batch-oracle] {prj_dtype}: rel RMSE after one iteration {errs}", flush=True)
    assert len(errs) == len(recs) and max(errs.values()) <= bound, errs
