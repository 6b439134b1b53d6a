"""The surrogate's kernels on the GPU: the bf16 MFMA GEMM against float64
products of the same bf16 values, the gradients of the training kernels
against the CPU network definition (oracle/mlp.py), and the trained
surrogate against the float64 trace -- including what the network gains
over the polynomial alone (the method without it)."""

import numpy as np
import pytest
import torch

from oracle import mlp as ref_mlp
from optika_amd import surrogate, workloads
from optika_amd.vectors import Cartesian2dVectorArray

pytestmark = pytest.mark.gpu


def _bf16(rng, shape, scale=1.0):
    x = ref_mlp.bf16_round(rng.normal(0, scale, shape))
    bits = (x.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    return x, torch.as_tensor(bits.view(np.int16), device="cuda")


@pytest.mark.parametrize("ta,tb,split", [(0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1), (1, 0, 4)])
def test_gemm_bf16_matches_float64(ta, tb, split):
    rng = np.random.default_rng(ta * 2 + tb + 10 * split)
    m, n, k = 128, 192, 256
    A, a = _bf16(rng, (k, m) if ta else (m, k))
    B, b = _bf16(rng, (n, k) if tb else (k, n))
    Am = A.T if ta else A
    Bm = B.T if tb else B
    want = Am @ Bm
    c = torch.empty(m, n, dtype=torch.float32, device="cuda")
    ws = torch.empty(split * m * n, dtype=torch.float32, device="cuda") if split > 1 else None
    print(f"case: gemm ta={ta} tb={tb} split={split}", flush=True)
    surrogate._gemm(ta, tb, m, n, k, a, m if ta else k, b, k if tb else n, c, n, split, ws)
    got = c.cpu().numpy().astype(np.float64)
    bound = 1e-5 * (np.abs(Am) @ np.abs(Bm)) + 1e-6
    assert np.all(np.abs(got - want) <= bound), np.max(np.abs(got - want) / bound)


@pytest.mark.parametrize("back", [False, True])
def test_row_panel_gemm_matches_tiled_path(back):
    """The hidden layers' row-panel product (ta = 0, tb = 1, n = k = 256)
    gives the same bits as the 128 x 128 tiled kernel on the same operands
    (b stored k-major through the transpose op, tb = 0), and the float64
    product of the same bf16 values within accumulation rounding, for both
    fused epilogues."""
    rng = np.random.default_rng(70 + back)
    m, n, k = 384, 256, 256
    A, a = _bf16(rng, (m, k), 0.5)
    Bt, bt = _bf16(rng, (n, k), 0.1)  # row j: output column j's weights
    b = torch.empty(k, n, dtype=torch.int16, device="cuda")
    surrogate._op(surrogate._lib.MLP_TRANSPOSE16, u16=[bt, b], rows=n, cols=k)
    assert np.array_equal(b.cpu().numpy(), bt.cpu().numpy().T)
    bias64 = rng.normal(0, 0.1, n).astype(np.float32).astype(np.float64)
    pre64 = rng.normal(0, 2, (m, n)).astype(np.float32).astype(np.float64)
    bias = torch.as_tensor(bias64, dtype=torch.float32, device="cuda")
    pre = torch.as_tensor(pre64, dtype=torch.float32, device="cuda")
    epi = surrogate._lib.GEMM_SILU_BACK if back else surrogate._lib.GEMM_BIAS_SILU
    print(f"case: row-panel gemm epilogue {epi}", flush=True)
    outs = []
    for tb, bm in ((1, bt), (0, b)):
        c = torch.empty(m, n, dtype=torch.float32, device="cuda")
        o16 = torch.empty(m, n, dtype=torch.int16, device="cuda")
        surrogate._gemm_fused(0, tb, m, n, k, a, k, bm, k if tb else n, c, n, epi, o16,
                              bias=None if back else bias, pre=pre if back else None)
        outs.append((c.cpu().numpy(), o16.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    if back:
        # the same product with the column sums of its float32 result (the
        # previous layer's bias gradient), with and without that result
        ws = torch.empty(m // 64 * 256, dtype=torch.float32, device="cuda")
        for keep in (True, False):
            c = torch.full((m, n), 7.0, dtype=torch.float32, device="cuda")
            o16 = torch.empty(m, n, dtype=torch.int16, device="cuda")
            sums = torch.empty(n, dtype=torch.float32, device="cuda")
            surrogate._gemm_silu_back_sums(m, a, k, bt, k, pre, n, c if keep else None, o16, sums, ws)
            assert np.array_equal(o16.cpu().numpy(), outs[0][1])
            if keep:
                assert np.array_equal(c.cpu().numpy(), outs[0][0])
            else:
                assert torch.all(c == 7.0).item()
            col = outs[0][0].astype(np.float64)
            err = np.abs(sums.cpu().numpy() - col.sum(0))
            assert np.all(err <= 1e-6 * np.abs(col).sum(0) + 1e-30), err.max()
    got = outs[0][0].astype(np.float64)
    mag = np.abs(A) @ np.abs(Bt).T
    if back:
        # silu' through the hardware exp and reciprocal: ~1e-7 absolute
        # (it crosses zero at x = -1.28, so no relative bound holds there)
        g = ref_mlp.silu_grad(pre64)
        want, bound = (A @ Bt.T) * g, (1e-5 * mag + 1e-6) * np.abs(g) + 1e-6 * np.abs(A @ Bt.T)
        act = got
    else:
        want, bound = A @ Bt.T + bias64, 1e-5 * mag + 1e-6
        act = ref_mlp.silu(got)
    assert np.all(np.abs(got - want) <= bound), np.max(np.abs(got - want) / bound)
    o = (outs[0][1].astype(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    assert np.all(np.abs(o - act) <= 2.0 ** -8 * np.abs(act) + 1e-30)


def test_training_gradients_match_cpu_network():
    rng = np.random.default_rng(3)
    dev = torch.device("cuda")
    net = surrogate.MLP(dev, seed=5)
    flat = ref_mlp.init_params(rng) * 0.7 + rng.normal(0, 0.02, surrogate.NUM_PARAMS)
    net.set_params(torch.as_tensor(flat, dtype=torch.float32, device=dev))
    p32 = net.p.cpu().numpy().astype(np.float64)  # the float32 master values
    rows = 256
    x = rng.uniform(-1, 1, (rows, 5)).astype(np.float32)
    t = rng.normal(0, 1, (rows, 3)).astype(np.float32)
    w = np.array([1.0, 1.0, 0.25], dtype=np.float32)
    print("case: mlp gradients vs cpu", flush=True)
    loss = net.loss_grad(torch.as_tensor(x, device=dev), torch.as_tensor(t, device=dev),
                         torch.as_tensor(w, device=dev))
    g = net.g.cpu().numpy().astype(np.float64)
    l_bf, g_bf = ref_mlp.loss_and_grad(p32, x, t, w, bf16=True)
    l_ex, g_ex = ref_mlp.loss_and_grad(p32, x, t, w, bf16=False)
    assert abs(loss.item() - l_bf) <= 1e-4 * abs(l_bf)
    sizes = [int(np.prod(s)) for s in ref_mlp.param_shapes()]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    for i in range(len(sizes)):
        sl = slice(starts[i], starts[i + 1])
        rel_bf = np.linalg.norm(g[sl] - g_bf[sl]) / np.linalg.norm(g_bf[sl])
        rel_ex = np.linalg.norm(g[sl] - g_ex[sl]) / np.linalg.norm(g_ex[sl])
        print(f"  block {i}: rel err vs bf16-emulating {rel_bf:.2e}, vs exact float64 {rel_ex:.2e}")
        assert rel_bf < 2e-2, (i, rel_bf)
        assert rel_ex < 5e-2, (i, rel_ex)


def test_adam_step_matches_formula():
    dev = torch.device("cuda")
    net = surrogate.MLP(dev, seed=1)
    p0 = net.p.clone()
    net.g.copy_(torch.linspace(-1, 1, surrogate.NUM_PARAMS, device=dev))
    net.adam_step(1e-3)
    g = net.g.double()
    m = 0.1 * g
    v = 0.001 * g * g
    want = p0.double() - 1e-3 * (m / 0.1) / (torch.sqrt(v / 0.001) + 1e-8)
    assert torch.max(torch.abs(net.p.double() - want)).item() < 1e-6


@pytest.fixture(scope="module")
def trained():
    s = workloads.spectrograph()
    lo = [-0.05, -0.05, -40.0, -40.0, 58.0]
    hi = [0.05, 0.05, 40.0, 40.0, 63.0]
    sg = surrogate.Surrogate(s, lo, hi, path=0, seed=0)
    print("case: surrogate train (2M rays)", flush=True)
    n = sg.prepare(lambda m, seed: workloads.spectrograph_grid(m, seed=seed), 1 << 21,
                   chunk=1 << 20, poly_rays=1 << 20)
    losses = sg.train(epochs=6, batch=65536, lr=2e-3, log_every=20)
    print(f"  {n} training rays; losses {losses}", flush=True)
    return sg


def _eval_grid(n, seed):
    g = workloads.spectrograph_grid(n, seed=seed)
    t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device="cuda").double()  # noqa: E731
    return t(g.field_x), t(g.field_y), t(g.pupil_x), t(g.pupil_y), t(g.wavelength)


def test_network_beats_polynomial_alone(trained):
    """Ablation: the same prediction with the network's correction removed."""
    sg = trained
    ins = _eval_grid(1 << 18, seed=77)
    ref = sg.trace64(*ins)
    p = sg.predict(*ins)
    u = sg.normalised(*ins)
    poly = sg.poly(u)
    ok = ref["unvignetted"] & ~p["fallback"]
    e_net = torch.sqrt((p["x"] - ref["x"]) ** 2 + (p["y"] - ref["y"]) ** 2)[ok] / sg.pixel
    e_poly = torch.sqrt((poly[:, 0] - ref["x"]) ** 2 + (poly[:, 1] - ref["y"]) ** 2)[ok] / sg.pixel
    rms_net = torch.sqrt((e_net ** 2).mean()).item()
    rms_poly = torch.sqrt((e_poly ** 2).mean()).item()
    print(f"ablation: RMS spot error polynomial alone {rms_poly:.3e} px, with network "
          f"{rms_net:.3e} px ({rms_poly / rms_net:.1f}x)", flush=True)
    assert rms_net < 0.5 * rms_poly


def test_surrogate_with_fallback_meets_bound(trained):
    sg = trained
    sg.calibrate(*_eval_grid(1 << 18, seed=78))
    ins = _eval_grid(1 << 18, seed=79)
    ref = sg.trace64(*ins)
    p = sg.run(*ins, precision="f32")
    both = ref["unvignetted"] & p["unvignetted"]
    err = torch.sqrt((p["x"] - ref["x"]) ** 2 + (p["y"] - ref["y"]) ** 2)[both] / sg.pixel
    rms = torch.sqrt((err ** 2).mean()).item()
    agree = (ref["unvignetted"] == p["unvignetted"]).double().mean().item()
    frac = sg.predict(*ins)["fallback"].double().mean().item()
    print(f"surrogate + fallback: RMS {rms:.3e} px, mask agreement {agree:.5f}, "
          f"fallback fraction {frac:.3f}", flush=True)
    assert rms <= 0.05
    assert agree >= 0.999


def test_fused_inference_matches_layered_path(trained):
    """The one-kernel inference against the layer-by-layer kernels and the
    float64 polynomial: positions within 1e-3 of the residual scale plus the
    float32 polynomial's rounding, identical fallback decisions except at
    the margin."""
    sg = trained
    ins = _eval_grid(100_000, seed=80)
    ref = sg.predict(*ins)
    f32 = [t.float().contiguous() for t in ins]
    out = sg.infer(*f32)
    torch.cuda.synchronize()
    ok = ~ref["fallback"] & ~out["fallback"]
    tol = 1e-3 * float(sg.scale[:2].max()) + 2e-5
    dx = torch.abs(out["x"].double() - ref["x"])[ok]
    dy = torch.abs(out["y"].double() - ref["y"])[ok]
    print(f"fused vs layered: max |dx| {dx.max().item():.2e} mm, |dy| {dy.max().item():.2e} mm, "
          f"tol {tol:.2e}", flush=True)
    assert dx.max().item() <= tol and dy.max().item() <= tol
    agree = (out["fallback"] == ref["fallback"]).double().mean().item()
    assert agree > 0.999
    count = int(out["count"].item())
    assert count == int(out["fallback"].sum().item())
    lst = out["list"][:count].long()
    assert torch.all(out["fallback"][lst]) and torch.unique(lst).numel() == count


def test_fused_inference_with_fallback_traced(trained):
    """infer_traced: the rays the fused kernel flags carry the float32
    trace of their own order (the same as order -1 of the two-order trace),
    the others the fused prediction unchanged; twice in a row (the cached
    fallback plan), the second time with fewer rays."""
    sg = trained
    for n, seed in ((100_000, 81), (60_001, 82)):
        ins = [t.float().contiguous() for t in _eval_grid(n, seed=seed)]
        print(f"case: infer_traced n={n}", flush=True)
        plain = {k: v.clone() for k, v in sg.infer(*ins).items()}
        out = sg.infer_traced(*ins, precision="f32")
        torch.cuda.synchronize()
        fb = plain["fallback"]
        assert int(fb.sum().item()) > 0, "the case should flag some rays"
        for k in ("x", "y", "intensity", "unvignetted"):
            assert torch.equal(out[k][~fb], plain[k][~fb]), k
        idx = torch.nonzero(fb).reshape(-1)
        sel = [t.index_select(0, idx).contiguous() for t in ins]
        r = sg.system.rayfunction(
            wavelength=sel[4], field=Cartesian2dVectorArray(sel[0], sel[1]),
            pupil=Cartesian2dVectorArray(sel[2], sel[3]), normalized_field=False,
            normalized_pupil=False, precision="f32").outputs
        r = r[sg.path] if r.position.x.ndim == 2 else r
        assert torch.equal(out["unvignetted"][idx], r.unvignetted)
        m = r.unvignetted
        for k, e in (("x", r.position.x), ("y", r.position.y), ("intensity", r.intensity)):
            a = out[k][idx][m].double()
            e = e[m].double()
            # the same float32 arithmetic; the bound only absorbs a different
            # kernel variant's contraction (a wrong ray mapping is off by mm)
            assert torch.all(torch.abs(a - e) <= 1e-5 + 1e-6 * torch.abs(e)), k
    # a fallback list too short for the flagged rays: the mask is used instead
    n = ins[0].numel()
    small = {"x": torch.empty(n, dtype=torch.float32, device="cuda"),
             "y": torch.empty(n, dtype=torch.float32, device="cuda"),
             "intensity": torch.empty(n, dtype=torch.float32, device="cuda"),
             "unvignetted": torch.empty(n, dtype=torch.bool, device="cuda"),
             "fallback": torch.empty(n, dtype=torch.bool, device="cuda"),
             "list": torch.empty(16, dtype=torch.int32, device="cuda"),
             "count": torch.zeros(1, dtype=torch.int32, device="cuda")}
    got = sg.infer_traced(*ins, out=small, precision="f32")
    assert int(got["count"].item()) > 16
    for k in ("x", "y", "intensity", "unvignetted"):
        assert torch.equal(got[k], out[k]), f"overflowed list: {k}"


@pytest.mark.parametrize("n", [1, 6400, 100_037])
def test_polynomial_normal_equations_and_eval(n):
    """optk_poly_normal / optk_poly_eval against numpy float64 on the same
    monomials: rays flagged invalid (and their non-finite targets) add
    nothing; the sums are deterministic."""
    rng = np.random.default_rng(n)
    u = rng.uniform(-1, 1, (n, 5))
    nt = 7
    tg = rng.normal(0, 1, (n, nt))
    valid = rng.uniform(0, 1, n) > 0.2
    tg[~valid, 0] = np.nan
    dev = torch.device("cuda")
    sg = surrogate.Surrogate.__new__(surrogate.Surrogate)
    sg.dev = dev
    print(f"case: poly normal/eval n={n}", flush=True)
    ut = torch.as_tensor(u, device=dev)
    F = ref_mlp.poly_features(u, surrogate.EXPONENTS)
    G = np.concatenate([F, np.nan_to_num(tg)], 1)
    A = F[valid].T @ G[valid]
    if valid.sum() >= 4 * 56:  # a well-posed fit
        sg.fit_polynomial(ut, torch.as_tensor(tg, device=dev), torch.as_tensor(valid, device=dev))
        want = np.linalg.solve(A[:, :56], A[:, 56:])
        np.testing.assert_allclose(sg.coef.cpu().numpy(), want, rtol=1e-8, atol=1e-10)
    # the accumulated matrix itself, twice (deterministic)
    from optika_amd import _lib

    outs = []
    for _ in range(2):
        out = torch.empty(56, 56 + nt, dtype=torch.float64, device=dev)
        nb = int(_lib.lib().optk_poly_workspace(n))
        ws = torch.empty(nb // 8 + 1, dtype=torch.float64, device=dev)
        tgt = torch.as_tensor(tg, device=dev)
        ok = torch.as_tensor(valid.astype(np.uint8), device=dev)
        _lib.check(_lib.lib().optk_poly_normal(
            ut.data_ptr(), tgt.data_ptr(), ok.data_ptr(), n, nt,
            surrogate._EXP_CODES.ctypes.data, out.data_ptr(), ws.data_ptr(), nb, None))
        outs.append(out.cpu().numpy())
    scale = np.abs(F[valid]).T @ np.abs(G[valid]) + 1e-300
    assert np.all(np.abs(outs[0] - A) <= 1e-13 * scale)
    assert np.array_equal(outs[0], outs[1])
    coef = rng.normal(0, 1, (56, nt))
    sg.coef = torch.as_tensor(coef, device=dev)
    got = sg.poly(ut).cpu().numpy()
    np.testing.assert_allclose(got, F @ coef, rtol=0, atol=1e-13 * (np.abs(F) @ np.abs(coef)).max())
