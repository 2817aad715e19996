"""The FBP row filter (ostr_fbp_filter: cosine weight, then the convolution
along detector x as an FFT convolution in LDS) against the FP64 direct sum of
oracle/fbp.py, over detector widths that take each of its launch forms: two
column pairs per work-group with 16-byte rows (nby % 4 == 0) and with scalar
rows and a partial last chunk, the widest FFT the LDS holds, a one-bin
detector; windowed ramp taps, short taps and taps longer than 2 nbx - 1 (the
excess never meets data); a slot range that does not start at 0; misaligned
buffers (the scalar path gives the same bits); and the size limit as an
error. Bar: 1e-5 of the peak (an FP32 FFT of L points errs by about
1e-7 sqrt(log2 L) of the signal's norm; observed below 1e-6)."""
import numpy as np
import pytest
import torch

from newrecon_amd.fbp import filter_taps
from newrecon_amd.geometry import Geometry, arc_angles
from newrecon_amd.ostr import DBTSystem, OSTRError

pytestmark = pytest.mark.gpu


def _geom(nbx, nby, dbx=0.14):
    return Geometry.half_cone(16, 16, 4, 1.0, 1.0, 1.0, nbx, nby, dbx, dbx)


def _views(g, n, seed):
    """Projection-like rows: a plateau with edges plus noise, positive."""
    rng = np.random.default_rng(seed)
    i = np.arange(g.nbx)[None, :, None]
    body = 2.0 * (np.abs(i - g.nbx / 2) < max(g.nbx / 3, 1))
    return (body + 0.05 * rng.standard_normal((n, g.nbx, g.nby)) + 0.5).astype(np.float32)


def _check(dev, g, taps, n_views=3, a0=1, a1=3, seed=0):
    from oracle import fbp as F
    angles = arc_angles(n_views, 7.5)
    S = DBTSystem(g, angles, device=dev)
    p = _views(g, a1 - a0, seed)
    tt = torch.from_numpy(taps.astype(np.float32)).to(dev)
    pg = torch.from_numpy(p).to(dev)
    q = S.fbp_filter(pg, tt, a0, a1)
    ref = F.fbp_filter(g, angles[a0:a1], p, taps.astype(np.float32).astype(np.float64))
    err = float(np.abs(q.cpu().numpy() - ref).max() / np.abs(ref).max())
    assert err <= 1e-5, (g.nbx, g.nby, len(taps), err)
    assert torch.equal(S.fbp_filter(pg, tt, a0, a1), q)
    return S, pg, tt, q, err


@pytest.mark.parametrize("nbx,nby", [(2048, 64), (2048, 66), (37, 5), (1, 3), (4096, 8)])
def test_fbp_filter_sizes(dev, nbx, nby):
    g = _geom(nbx, nby)
    _, _, _, _, err = _check(dev, g, filter_taps(nbx, g.dbx, "hann"))
    print(f"nbx={nbx} nby={nby}: max error / peak {err:.2e}")


@pytest.mark.parametrize("ntaps", [1, 5, 2 * 300 + 9])
def test_fbp_filter_tap_lengths(dev, ntaps):
    g = _geom(300, 12)
    taps = np.random.default_rng(ntaps).standard_normal(ntaps)
    _check(dev, g, taps)


def test_fbp_filter_misaligned_buffers_same_bits(dev):
    """p and q 4 bytes off a 16-byte boundary take the scalar row path, which
    computes the same values as the 16-byte one."""
    g = _geom(512, 16)
    S, pg, tt, q, _ = _check(dev, g, filter_taps(512, g.dbx, "ramp"))
    n = pg.numel()
    pbuf = torch.zeros(n + 1, device=dev)
    pbuf[1:] = pg.reshape(-1)
    qbuf = torch.zeros(n + 1, device=dev)
    S.fbp_filter(pbuf[1:].view(pg.shape), tt, 1, 3, q=qbuf[1:].view(pg.shape))
    assert torch.equal(qbuf[1:].view(pg.shape), q)


def test_fbp_filter_size_limit(dev):
    """A row whose FFT does not fit one work-group's LDS is refused."""
    g = _geom(8192, 4)
    S = DBTSystem(g, arc_angles(1, 7.5), device=dev)
    p = S.sinogram(1)
    taps = torch.ones(3, device=dev)
    with pytest.raises(OSTRError, match="LDS"):
        S.fbp_filter(p, taps, 0, 1)
