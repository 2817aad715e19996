"""ostr_objective_noalloc (the allocation-free FP64 objective, static
per-device scratch) through the C ABI: equal to ostr_objective bit for bit
on two calls in a row, the second on a smaller case, and a case needing more
partials than the scratch holds is refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from newrecon_amd.geometry import Geometry, scan_small, scan_tiny
from newrecon_amd.ostr import POTENTIALS, DBTSystem, _dt, _p, lib
from newrecon_amd.phantom import breast_phantom, poisson_counts

pytestmark = pytest.mark.gpu

NOALLOC_DOUBLES = 1 << 20


def _case(scan, dev, seed):
    g = scan.geom
    S = DBTSystem(g, scan.angles, device=dev)
    x = torch.from_numpy(breast_phantom(g, scan.phantom_seed, scan.n_fg, scan.n_calc)).to(dev)
    l = S.fproj(x, 0, scan.n_views).cpu().numpy()
    y = torch.from_numpy(poisson_counts(l, scan.i0, seed)).to(dev)
    b = torch.full_like(y, scan.i0)
    r = torch.full_like(y, 3.0)
    return S, x, y, b, r


def _both(S, img, y, b, r, potential, delta):
    L = lib()
    st = ctypes.c_void_p(torch.cuda.current_stream(S.device).cuda_stream)
    rho = S.geom.dz / S.geom.dx
    ref = (ctypes.c_double * 2)()
    got = (ctypes.c_double * 2)()
    args = (S._h, _p(img), _dt(img), _p(y), _p(b), _p(r), POTENTIALS[potential], delta, rho)
    assert L.ostr_objective(*args, ref, st) == 0
    rc = L.ostr_objective_noalloc(*args, got, st)
    assert rc == 0, L.ostr_last_error(S._h).decode()
    return tuple(ref), tuple(got)


def test_noalloc_objective_equals_objective(small_case, dev):
    big = _case(scan_small(), dev, 1001)
    small = _case(scan_tiny(nx=24, ny=16, nz=8, nbx=40, nby=24, n_views=3), dev, 5)
    print("[noalloc] 200x100x20 then 24x16x8", flush=True)
    for (S, x, y, b, r), pot in ((big, "hyperbola"), (small, "lange")):
        for img in (x, x.to(torch.float16)):
            ref, got = _both(S, img, y, b, r, pot, 5e-4)
            assert np.isfinite(ref).all() and ref[0] != 0.0 and ref[1] != 0.0
            assert got == ref, (S.geom.nx, img.dtype, ref, got)
    # and back to the large case: the scratch keeps the small case's partials
    S, x, y, b, r = big
    ref, got = _both(S, x, y, b, None, "q2", 1.0)
    assert got == ref


def test_noalloc_objective_refuses_too_many_partials(dev):
    nb = 4096
    g = Geometry.half_cone(16, 16, 4, 0.1, 0.1, 1.0, nb, nb, 0.1, 0.1)
    nv = 17
    need = (nb // 64) * (nb // 4) * nv + 1 * 4 * 4 + 2
    assert need > NOALLOC_DOUBLES
    S = DBTSystem(g, np.linspace(-0.1, 0.1, nv).astype(np.float32), device=dev)
    img = S.volume()
    y = torch.ones((nv, nb, nb), device=dev)
    L = lib()
    out = (ctypes.c_double * 2)()
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = L.ostr_objective_noalloc(S._h, _p(img), _dt(img), _p(y), _p(y), None, 2, 5e-4, 10.0,
                                  out, st)
    assert rc == -3
    assert b"static scratch" in L.ostr_last_error(S._h)
