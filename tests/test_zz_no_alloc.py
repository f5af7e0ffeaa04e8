"""qv_vjp_device_noalloc (the VJP with its scratch in a static device array,
for callers that cannot allocate) against qv_vjp_device on the same state,
through the C ABI: bit-identical on two calls in a row, the second on a
smaller system uploaded to the same context (the scratch keeps the first
call's contents), FP64 and FP32C pair terms; more slots than the scratch
holds is refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import ROOT  # noqa: F401  (puts the repository on sys.path)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _vjp_pair(ctx, torch, x, w):
    xd = torch.from_numpy(x).cuda()
    wd = torch.from_numpy(w).cuda()
    a = torch.empty_like(wd)
    b = torch.empty_like(wd)
    ctx.vjp_device(xd.data_ptr(), wd.data_ptr(), a.data_ptr())
    torch.cuda.synchronize()
    ctx.vjp_device_noalloc(xd.data_ptr(), wd.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_noalloc_vjp_equals_vjp_device(precision):
    import torch

    from qvort_amd import geometry as G
    from qvort_amd import gpu
    big = G.random_loops(0.1, 5e-4, target_points=3000, seed=11)
    small = G.random_loops(0.1, 5e-4, target_points=700, seed=12)
    rng = np.random.default_rng(5)
    p = gpu.params(velocity="BS", boundary="periodic", box_size=0.1, corea=1.3e-8,
                   precision=precision)
    with gpu.Context(p) as c:
        for fil in (big, small):
            _say("case", precision, fil.pcount, "points")
            c.upload_filaments(fil)
            w = rng.standard_normal((fil.pcount, 3))
            a, b = _vjp_pair(c, torch, np.ascontiguousarray(fil.x), w)
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b)


def test_noalloc_vjp_refuses_more_slots_than_scratch():
    import torch

    from qvort_amd import gpu
    n = 65537
    ang = 2 * np.pi * np.arange(n) / n
    x = np.stack([0.04 * np.cos(ang), 0.04 * np.sin(ang), np.zeros(n)], axis=1)
    inf = (np.arange(n) + 1) % n + 1
    beh = (np.arange(n) - 1) % n + 1
    p = gpu.params(velocity="BS", boundary="open", box_size=0.1, corea=1.3e-8)
    with gpu.Context(p) as c:
        c.upload(x, inf.astype(np.int32), beh.astype(np.int32))
        xd = torch.from_numpy(x).cuda()
        out = torch.empty_like(xd)
        rc = gpu.lib().qv_vjp_device_noalloc(c._h, C.c_void_p(xd.data_ptr()),
                                              C.c_void_p(xd.data_ptr()),
                                              C.c_void_p(out.data_ptr()))
        assert rc == gpu.QV_ERR_SCRATCH
