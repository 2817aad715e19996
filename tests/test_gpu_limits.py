"""Size limits of the backprojecting calls: one launch reads at most 64
sinogram slots through 32-bit byte offsets, so a call whose slots span 2^32
bytes or more is refused, and one just below the limit reads its highest
slot correctly (byte offsets above 2^31)."""
import numpy as np
import pytest
import torch

from newrecon_amd.geometry import Geometry
from newrecon_amd.ostr import DBTSystem, OSTRError

pytestmark = pytest.mark.gpu


def test_sinogram_span_limit(dev):
    nb = 4096  # 64 views of 4096 x 4096 FP32 bins = 2^32 bytes
    g = Geometry.half_cone(16, 16, 4, 0.1, 0.1, 1.0, nb, nb, 0.1, 0.1)
    angles = np.linspace(-0.13, 0.13, 64).astype(np.float32)
    S = DBTSystem(g, angles, device=dev)
    print(f"[limits] 64 x {nb} x {nb} float32 rows", flush=True)
    prj = torch.zeros((64, nb, nb), dtype=torch.float32, device=dev)
    img = S.volume()
    with pytest.raises(OSTRError) as ei:
        S.bproj_sbp_rows(prj, 0, 64, img=img)
    assert ei.value.code == -3 and "2^32" in str(ei.value)
    with pytest.raises(OSTRError) as ei:
        S.update_fused(S.volume(), S.volume(), S.volume(), prj, 0, 64, 1.0, 1.0, "hyperbola",
                       5e-4, 10.0, True)
    assert ei.value.code == -3
    # 63 views fit: put data in the last slot only and compare with that
    # view backprojected from its own one-slot sinogram
    del prj
    prj = torch.zeros((63, nb, nb), dtype=torch.float32, device=dev)
    prj[62].uniform_(0.5, 1.5, generator=torch.Generator(device=dev).manual_seed(7))
    assert 62 * nb * nb * 4 > 2 ** 31
    big = S.bproj_sbp_rows(prj, 0, 63, img=S.volume())
    one = S.bproj_sbp_rows(prj[62:63].clone(), 62, 63, img=S.volume())
    torch.cuda.synchronize()
    assert one.abs().max() > 0
    assert torch.equal(big, one)
