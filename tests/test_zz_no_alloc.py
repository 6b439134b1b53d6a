"""optk_collect_noalloc, the binning entry point that allocates nothing,
against optk_collect with a caller-owned workspace (both through the C ABI).

Two calls in a row on the same GPU share the library's scratch array; the
second uses fewer rays, so stale entries of the first call sit past its
end.  Counts (integer weights) must agree exactly, weighted sums to
float64 rounding (the tiles are summed with float64 atomics in either
entry point, so their order is not fixed).
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rays(n, f32, seed):
    import torch

    dt = torch.float32 if f32 else torch.float64
    g = torch.Generator(device="cuda").manual_seed(seed)
    half = 15.36
    # a margin outside the sensor so some rays miss every pixel
    px = (torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 2.2 - 1.1) * half
    py = (torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 2.2 - 1.1) * half
    dz = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 0.5 + 0.5
    wl = torch.rand(n, generator=g, device="cuda", dtype=torch.float64) * 300 + 400
    # integer weights: every pixel sum is exact in either summation order
    inten = torch.randint(0, 4, (n,), generator=g, device="cuda").to(torch.float64)
    mask = (torch.rand(n, generator=g, device="cuda") < 0.9).to(torch.uint8)
    return [t.to(dt).contiguous() for t in (px, py, dz, wl, inten)] + [mask]


def _args(rays, edges, image, moment, nx, ny, f32):
    from optika_amd import _lib

    px, py, dz, wl, inten, mask = rays
    a = _lib.CollectArgs()
    a.position[0], a.position[1] = px.data_ptr(), py.data_ptr()
    a.direction_z = dz.data_ptr()
    a.wavelength = wl.data_ptr()
    a.intensity = inten.data_ptr()
    a.unvignetted = mask.data_ptr()
    a.wavelength_edges = edges.data_ptr()
    a.n_wavelength, a.nx, a.ny = edges.numel() - 1, nx, ny
    a.precision = _lib.F32 if f32 else _lib.F64
    a.x_lo, a.x_hi, a.y_lo, a.y_hi = -15.36, 15.36, -15.36, 15.36
    a.image, a.moment = image.data_ptr(), moment.data_ptr()
    return a


def _bin(rays, n, nx, ny, f32, noalloc, edges):
    import torch

    from optika_amd import _lib

    L = _lib.lib()
    edges = torch.tensor(edges, dtype=torch.float64, device="cuda")
    image = torch.zeros((edges.numel() - 1, nx, ny), dtype=torch.float64, device="cuda")
    moment = torch.zeros_like(image)
    a = _args(rays, edges, image, moment, nx, ny, f32)
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = L.optk_collect_noalloc(ctypes.byref(a), n, stream)
    else:
        nbytes = int(L.optk_collect_workspace(ctypes.byref(a), n))
        assert nbytes > 0
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        a.workspace, a.workspace_bytes = ws.data_ptr(), nbytes
        rc = L.optk_collect(ctypes.byref(a), n, stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return image.cpu().numpy(), moment.cpu().numpy()


@pytest.mark.parametrize("f32", [True, False], ids=["f32", "f64"])
def test_noalloc_collect_matches_workspace_collect(f32):
    from optika_amd import _lib

    import torch

    nx = ny = 2048
    three = [400.0, 500.0, 600.0, 700.0]  # 12288 tiles: the unstaged scatter
    one = [400.0, 700.0]                  # 4096 tiles: the LDS-staged scatter
    # the largest call the scratch holds for three wavelength bins (found
    # with the host-side workspace query), then a small call with one bin
    probe = _lib.CollectArgs()
    probe.n_wavelength, probe.nx, probe.ny = 3, nx, ny
    probe.precision = _lib.F32 if f32 else _lib.F64
    lo, hi = 1, 1 << 23
    while lo < hi:
        mid = (lo + hi + 1) // 2
        need = _lib.lib().optk_collect_workspace(ctypes.byref(probe), mid)
        lo, hi = (mid, hi) if need <= _lib.COLLECT_SCRATCH_BYTES else (lo, mid - 1)
    big = lo
    assert 2_500_000 < big < 3_300_000
    for n, seed, edges in ((big, 11, three), (20_480, 12, one)):
        print(f"noalloc collect: f32={f32} n={n} seed={seed} bins={len(edges) - 1}", flush=True)
        rays = _rays(n, f32, seed)
        ref_img, ref_mom = _bin(rays, n, nx, ny, f32, False, edges)
        img, mom = _bin(rays, n, nx, ny, f32, True, edges)
        assert np.array_equal(img, ref_img), "pixel sums differ"
        assert ref_img.sum() > 0.5 * n  # most rays land on the image
        np.testing.assert_allclose(mom, ref_mom, rtol=1e-12, atol=1e-9)
    # a call needing more than the scratch is refused, before any device
    # work (the arrays hold all its rays all the same)
    L = _lib.lib()
    n = 4_000_000
    rays = _rays(n, f32, 13)
    edges = torch.tensor([400.0, 700.0], dtype=torch.float64, device="cuda")
    image = torch.zeros((1, nx, ny), dtype=torch.float64, device="cuda")
    a = _args(rays, edges, image, torch.zeros_like(image), nx, ny, f32)
    assert L.optk_collect_workspace(ctypes.byref(a), n) > _lib.COLLECT_SCRATCH_BYTES
    assert L.optk_collect_noalloc(ctypes.byref(a), n, None) == -1
    assert b"scratch" in L.optk_last_error()
    torch.cuda.synchronize()
    assert float(image.abs().sum()) == 0.0
