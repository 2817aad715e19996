"""OSTR iterations replayed from a HIP graph equal the eager iterations bit for
bit: fused SBP with 3 subsets, matched with 3 subsets, and the accelerated
matched iteration with the device-side momentum restart. Each replay chunk is
separated by eager work and synchronisation and checked against the eager
run."""
import numpy as np
import pytest
import torch

from newrecon_amd.geometry import OSTRConfig
from newrecon_amd.ostr import OSTR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("backprojector,M,momentum", [("sbp", 3, False), ("matched", 3, False),
                                                      ("matched", 1, True)])
def test_graph_replay_equals_eager(small_case, dev, backprojector, M, momentum):
    s, x, y, b = small_case
    cfg = OSTRConfig(n_subsets=M, beta=s.recon.beta, delta=s.recon.delta,
                     backprojector=backprojector, momentum=momentum)
    eager = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev)
    graph = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev)
    done = 0
    for chunk in (7, 1, 12, 5):
        graph.iterate(chunk, graph=True)
        # eager work and a synchronisation between replays
        junk = torch.randn(1 << 20, device=dev).sum()
        torch.cuda.synchronize()
        eager.iterate(chunk)
        done += chunk
        assert graph.iteration == eager.iteration == done
        assert torch.equal(graph.estimate, eager.estimate), chunk
        assert torch.isfinite(junk)
    if momentum:
        assert graph.restarts == eager.restarts
    assert graph._graph is not None and 1 <= graph._graph_iters <= 12


def test_device_momentum_matches_host_formula(small_case, dev):
    """The device coefficient sequence equals t' = (1 + sqrt(1 + 4t^2))/2,
    c = (t - 1)/t' evaluated on the host, while no restart occurs."""
    import math
    from newrecon_amd.ostr import DBTSystem
    s = small_case[0]
    S = DBTSystem(s.geom, s.angles, device=dev)
    state = torch.empty(4, dtype=torch.float64, device=dev)
    S.momentum_init(state)
    z = torch.rand(s.geom.vol_shape, device=dev)
    zp = z * 0.5                      # z - zp > 0
    xk = z.clone()                    # <xk - z, z - zp> = 0: no restart
    xn = torch.empty_like(z)
    t = 1.0
    for _ in range(6):
        S.momentum_step(xn, z, zp, xk, state)
        tn = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
        c, t = (t - 1.0) / tn, tn
        st = state.cpu().numpy()
        assert st[0] == t and st[1] == c and st[2] == 0.0
        want = torch.clamp(z + np.float32(c) * (z - zp), min=0.0)
        assert torch.allclose(xn, want, rtol=1e-6, atol=0)
    xk2 = z * 2.0                     # <xk - z, z - zp> > 0: restart
    S.momentum_step(xn, z, zp, xk2, state)
    st = state.cpu().numpy()
    assert st[0] == 1.0 and st[1] == 0.0 and st[2] == 1.0 and st[3] > 0
    assert torch.equal(xn, z)


@pytest.mark.parametrize("backprojector,M,momentum,fused", [("matched", 1, True, True),
                                                            ("sbp", 3, False, False),
                                                            ("matched", 3, False, False)])
def test_graph_replay_after_work_buffers_grow(small_case, dev, backprojector, M, momentum, fused):
    """Calls that grow the handle's work buffers between replays (the FP64
    objective and norms, the BP image) must not free memory a recorded graph
    still writes: the momentum step's partials and the unfused
    backprojectors' transposes live there. Replays equal the eager run."""
    s, x, y, b = small_case
    cfg = OSTRConfig(n_subsets=M, beta=s.recon.beta, delta=s.recon.delta,
                     backprojector=backprojector, momentum=momentum, fused=fused)
    eager = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev)
    graph = OSTR(s.geom, s.angles, y, b, cfg=cfg, x0="bp", device=dev)
    yd = torch.as_tensor(y, device=dev)
    bd = torch.as_tensor(b, device=dev)
    done = 0
    for chunk in (6, 6, 7):
        graph.iterate(chunk, graph=True)
        # grow dwork (objective > norms > momentum partials) and work (BP image)
        f = graph.objective()
        n = graph.sys.diff_norms(graph.estimate, graph.estimate)
        bp = graph.sys.bp_init(y=yd, b=bd)
        torch.cuda.synchronize()
        eager.iterate(chunk)
        done += chunk
        assert np.isfinite(f) and n[0] == 0.0 and torch.isfinite(bp).all()
        assert graph.iteration == eager.iteration == done
        assert torch.equal(graph.estimate, eager.estimate), chunk
    assert graph._graph is not None
