"""A learned surrogate of one system's ``rayfunction``.

A small MLP (5 -> 256 -> 256 -> 256 -> 256 -> 3, SiLU, bf16 weights and
activations, float32 accumulation on MFMA) maps (field x, y; pupil x, y;
wavelength) to (sensor x, y; log intensity) for one diffraction-order path.
It is trained on rays the float64 trace produces, by the library's own
kernels (``optk_gemm_bf16``, ``optk_mlp_op``; Adam), and evaluated against
the float64 trace.

The network does not predict positions from scratch: a least-squares cubic
polynomial in the five normalised inputs (the method without the network)
carries the smooth part of the map, and the network learns the residual,
scaled to unit variance.  The same polynomial predicts where each ray
crosses every aperture; rays predicted within ``edge_margin`` (0.5 mm) of
an aperture edge, or in an input-space cell where validation showed the
surrogate above its error bound, fall back to the trace.

No counterpart in the reference.
"""

from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np
import torch

from . import _lib, _trace

__all__ = ["Surrogate", "NUM_POLY", "EXPONENTS"]

WIDTHS = (5, 256, 256, 256, 256, 3)
NUM_POLY = 56  # monomials of total degree <= 3 in 5 variables


def _exponents(degree=3, nvar=5):
    out = []

    def rec(prefix, remaining, var):
        if var == nvar:
            out.append(tuple(prefix))
            return
        for e in range(remaining + 1):
            rec(prefix + [e], remaining - e, var + 1)

    rec([], degree, 0)
    out.sort(key=lambda e: (sum(e), e[::-1]))
    return out


EXPONENTS = _exponents()
assert len(EXPONENTS) == NUM_POLY
# the monomials' exponents packed 2 bits per input, as the kernels take them
_EXP_CODES = np.array([sum(int(k) << (2 * v) for v, k in enumerate(e)) for e in EXPONENTS],
                      dtype=np.uint32)


def _param_layout(widths=WIDTHS):
    """(offset, shape) of W and b of each layer in the flat parameter vector."""
    lay, o = [], 0
    for a, b in zip(widths[:-1], widths[1:]):
        lay.append(((o, (b, a)), (o + a * b, (b,))))
        o += a * b + b
    return lay, o


LAYOUT, NUM_PARAMS = _param_layout()


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return _trace.stream_handle()


def _bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), as float32."""
    return x.to(torch.bfloat16).to(torch.float32)


def _op(op, f32=(), u16=(), rows=0, cols=0, inner=0, count=0, flag=0, scalars=()):
    a = _lib.MlpArgs()
    for i, t in enumerate(f32):
        a.f32[i] = None if t is None else t.data_ptr()
    for i, t in enumerate(u16):
        a.u16[i] = None if t is None else t.data_ptr()
    a.rows, a.cols, a.inner, a.count, a.flag = rows, cols, inner, count, flag
    for i, v in enumerate(scalars):
        a.scalars[i] = v
    _lib.check(_lib.lib().optk_mlp_op(op, ctypes.byref(a), _stream()))


def _gemm(ta, tb, m, n, k, a, lda, b, ldb, c, ldc, split=1, ws=None):
    _lib.check(_lib.lib().optk_gemm_bf16(ta, tb, m, n, k, _ptr(a), lda, _ptr(b), ldb, _ptr(c), ldc,
                                         split, _ptr(ws), _stream()))


def _gemm_fused(ta, tb, m, n, k, a, lda, b, ldb, c, ldc, epilogue, out16, bias=None, pre=None):
    """GEMM with the layer's bias + SiLU (forward) or SiLU derivative
    (backward) in its epilogue (optk_gemm_bf16_fused)."""
    _lib.check(_lib.lib().optk_gemm_bf16_fused(ta, tb, m, n, k, _ptr(a), lda, _ptr(b), ldb, _ptr(c),
                                               ldc, epilogue, _ptr(bias), _ptr(pre), _ptr(out16),
                                               _stream()))


def _gemm_silu_back_sums(m, a, lda, b, ldb, pre, ldc, c, out16, sums, ws):
    """Backward product of a 256-wide layer with the SiLU derivative, the
    bf16 (and optionally float32) result and its column sums -- the
    previous layer's bias gradient (optk_gemm_bf16_silu_back_sums)."""
    _lib.check(_lib.lib().optk_gemm_bf16_silu_back_sums(m, _ptr(a), lda, _ptr(b), ldb, _ptr(pre), ldc,
                                                        _ptr(c), _ptr(out16), _ptr(sums), _ptr(ws),
                                                        _stream()))


class MLP:
    """Parameters (float32 master + bf16 copy) and Adam state of the
    network, with forward/backward passes on the library's kernels."""

    def __init__(self, dev, seed: int = 0):
        self.dev = dev
        g = torch.Generator(device="cpu").manual_seed(seed)
        parts = []
        for (wo, ws), (bo, bs) in LAYOUT:
            bound = math.sqrt(6.0 / ws[1])
            parts.append((torch.rand(ws, generator=g, dtype=torch.float64) * 2 - 1) * bound)
            parts.append(torch.zeros(bs, dtype=torch.float64))
        flat = torch.cat([p.reshape(-1) for p in parts]).to(torch.float32)
        self.set_params(flat.to(dev))

    def set_params(self, flat32: torch.Tensor):
        assert flat32.numel() == NUM_PARAMS
        self.p = flat32.contiguous().clone()
        self.p16 = torch.empty(NUM_PARAMS, dtype=torch.int16, device=self.dev)
        _op(_lib.MLP_TO_BF16, f32=[self.p], u16=[self.p16], rows=1, cols=NUM_PARAMS)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.g = torch.zeros_like(self.p)
        self.t = 0

    def view(self, flat, layer, which):
        off, shape = LAYOUT[layer][which]
        return flat[off:off + int(np.prod(shape))]

    # -- forward ---------------------------------------------------------
    def _buffers(self, rows, train):
        key = (rows, train)
        if getattr(self, "_buf_key", None) != key:
            f = lambda *s: torch.empty(s, dtype=torch.float32, device=self.dev)  # noqa: E731
            h = lambda *s: torch.empty(s, dtype=torch.int16, device=self.dev)  # noqa: E731
            H = WIDTHS[1]
            self.pre = [f(rows, H) for _ in range(4)]
            self.act = [h(rows, H) for _ in range(4)]
            self.y = f(rows, 3)
            if train:
                self.dy = f(rows, 3)
                # two bf16 gradient buffers: the backward GEMM of a layer
                # reads one while its epilogue writes the previous layer's
                # gradient into the other
                self.dpre16 = h(rows, H)
                self.dpre16_b = h(rows, H)
                self.dpre32 = f(rows, H)
                # the hidden layers' bf16 weights transposed ([in][out]): the
                # backward product's B operand with k contiguous
                self.wt16 = [h(H, H) for _ in range(3)]
                self.ones = torch.ones(rows, 1, dtype=torch.float32, device=self.dev)
                # split the weight-gradient GEMM's k = rows so that its four
                # 128 x 128 output tiles make ~2 blocks per CU
                split = max(1, min(128, rows // 512))
                while rows % (32 * split):
                    split //= 2
                self.split = split
                # also the per-block partials of the small outer products
                # (at most min(512, rows / 32) blocks of H x 8)
                self.ws = f(max(split * H * H, min(512, (rows + 31) // 32) * H * 8))
                self.loss_part = f((rows * 3 + 255) // 256)
            self._buf_key = key

    def forward(self, x: torch.Tensor, train: bool = False) -> torch.Tensor:
        """x: [rows][5] float32 (normalised inputs), rows a multiple of 64."""
        rows = x.shape[0]
        assert rows % 64 == 0 and x.dtype == torch.float32 and x.is_contiguous()
        self._buffers(rows, train)
        H = WIDTHS[1]
        _op(_lib.MLP_DENSE_IN, f32=[x, self.view(self.p, 0, 0), self.view(self.p, 0, 1),
                                    self.pre[0]],
            u16=[self.act[0]], rows=rows, cols=H, inner=5)
        for layer in (1, 2, 3):
            w16 = self.view(self.p16, layer, 0)
            if rows % 128 == 0:
                _gemm_fused(0, 1, rows, H, H, self.act[layer - 1], H, w16, H, self.pre[layer], H,
                            _lib.GEMM_BIAS_SILU, self.act[layer], bias=self.view(self.p, layer, 1))
            else:
                _gemm(0, 1, rows, H, H, self.act[layer - 1], H, w16, H, self.pre[layer], H)
                _op(_lib.MLP_BIAS_SILU, f32=[self.pre[layer], self.view(self.p, layer, 1)],
                    u16=[self.act[layer]], rows=rows, cols=H)
        _op(_lib.MLP_DENSE_OUT, f32=[self.view(self.p, 4, 0), self.view(self.p, 4, 1), self.y],
            u16=[self.act[3]], rows=rows, cols=H, inner=3)
        return self.y

    # -- backward --------------------------------------------------------
    def loss_grad(self, x, target, weight) -> torch.Tensor:
        """Forward, weighted-MSE loss and gradients into ``self.g``;
        returns the loss (a device scalar).  rows: a multiple of 128."""
        rows = x.shape[0]
        if rows % 128:
            raise ValueError("training batches must hold a multiple of 128 rows")
        H = WIDTHS[1]
        self.forward(x, train=True)
        _op(_lib.MLP_MSE_GRAD, f32=[self.y, target, weight, self.dy, self.loss_part], rows=rows,
            cols=3)
        g = self.g
        # output layer
        _op(_lib.MLP_OUTER_GRAD, f32=[None, self.dy, self.view(g, 4, 0), self.ws],
            u16=[self.act[3]], rows=rows, cols=H, inner=3, count=self.ws.numel(), flag=0)
        _op(_lib.MLP_OUTER_GRAD, f32=[self.dy, self.ones, self.view(g, 4, 1), self.ws],
            rows=rows, cols=3, inner=1, count=self.ws.numel(), flag=1)
        # the last hidden layer's dpre (bf16) and its column sums, that
        # layer's bias gradient; the other bias gradients come with the
        # backward products below
        _op(_lib.MLP_DENSE_OUT_BACK, f32=[self.dy, self.view(self.p, 4, 0), self.pre[3], None,
                                          self.view(g, 3, 1), self.ws], u16=[self.dpre16],
            rows=rows, cols=H, inner=3, count=self.ws.numel())
        for layer in (1, 2, 3):
            _op(_lib.MLP_TRANSPOSE16, u16=[self.view(self.p16, layer, 0), self.wt16[layer - 1]],
                rows=H, cols=H)
        d16, d16_next = self.dpre16, self.dpre16_b
        for layer in (3, 2, 1):
            # weight gradient: gW[o][i] = sum_r dpre[r][o] act[r][i]
            _gemm(1, 0, H, H, rows, d16, H, self.act[layer - 1], H, self.view(g, layer, 0),
                  H, self.split, self.ws)
            # back to the previous layer: dpre = (dpre W) * silu'(pre), the
            # derivative in the GEMM's epilogue with the column sums (that
            # layer's bias gradient); the float32 dpre is kept only for the
            # first layer's weight gradient
            _gemm_silu_back_sums(rows, d16, H, self.wt16[layer - 1], H, self.pre[layer - 1], H,
                                 self.dpre32 if layer == 1 else None, d16_next,
                                 self.view(g, layer - 1, 1), self.ws)
            d16, d16_next = d16_next, d16
        # first layer (float32)
        _op(_lib.MLP_OUTER_GRAD, f32=[self.dpre32, x, self.view(g, 0, 0), self.ws], rows=rows,
            cols=H, inner=5, count=self.ws.numel(), flag=1)
        return self.loss_part.sum()

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        c1 = 1.0 - beta1 ** self.t
        c2 = 1.0 - beta2 ** self.t
        _op(_lib.MLP_ADAM, f32=[self.p, self.g, self.m, self.v], u16=[self.p16],
            count=NUM_PARAMS, scalars=[lr, beta1, beta2, eps, c1, c2])


@dataclasses.dataclass
class Normalisation:
    lo: np.ndarray  # input ranges (field x, y deg; pupil x, y mm; wavelength nm)
    hi: np.ndarray

    def apply(self, cols):
        mid = (self.lo + self.hi) / 2
        half = (self.hi - self.lo) / 2
        return [(c - float(mid[i])) / float(half[i]) for i, c in enumerate(cols)]


# ---------------------------------------------------------------------------
# apertures: signed distance to the edge (positive inside)
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class ApertureEdge:
    """One active aperture of the system, in the frame where it clips."""

    surface: int        # index into surfaces_all
    kind: int           # OPTK_APERTURE_*
    params: np.ndarray  # aperture_p
    vertices: np.ndarray | None  # polygon [nv][2]
    inverted: bool
    m: np.ndarray       # system frame -> aperture frame: p_ap = m @ p + t
    t: np.ndarray

    def distance(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        if self.kind == 1:  # circle
            d = float(self.params[0]) - torch.sqrt(x * x + y * y)
        elif self.kind == 4:  # rectangle
            d = torch.minimum(float(self.params[0]) - torch.abs(x), float(self.params[1]) - torch.abs(y))
        elif self.kind == 5:  # polygon: distance to the boundary, signed by the crossing rule
            v = self.vertices
            dmin = None
            inside = torch.zeros_like(x, dtype=torch.bool)
            nv = len(v)
            for i in range(nv):
                x1, y1 = float(v[i - 1][0]), float(v[i - 1][1])
                x2, y2 = float(v[i][0]), float(v[i][1])
                ex, ey = x2 - x1, y2 - y1
                L2 = ex * ex + ey * ey
                s = torch.clamp(((x - x1) * ex + (y - y1) * ey) / L2, 0.0, 1.0)
                dx, dy = x - (x1 + s * ex), y - (y1 + s * ey)
                dd = torch.sqrt(dx * dx + dy * dy)
                dmin = dd if dmin is None else torch.minimum(dmin, dd)
                cond = (y1 > y) != (y2 > y)
                xc = (x2 - x1) * (y - y1) / (y2 - y1 if y2 != y1 else 1e-300) + x1
                inside = inside ^ (cond & (x < xc))
            d = torch.where(inside, dmin, -dmin)
        else:  # sectors and ellipses: no distance model, always near an edge
            d = torch.zeros_like(x)
        return -d if self.inverted else d


def _aperture_edges(system) -> list[ApertureEdge]:
    from . import _compile

    rec = _compile.build_records(system.surfaces_all)
    edges = []
    for k in range(rec.num_surfaces):
        s = rec.surfaces[k]
        if s.aperture_kind == 0 or not (s.aperture_flags & 1) or (s.aperture_flags & 4):
            continue

        def mt(a):
            if a.kind == 0:
                return np.eye(3), np.zeros(3)
            m = np.array(a.m[:]).reshape(3, 3) if a.kind == 2 else np.eye(3)
            return m, np.array(a.t[:])

        mT, tT = mt(s.transform)
        mA, tA = mt(s.aperture_transform)
        # system -> surface: mT^T (p - tT); surface -> aperture: mA^T (q - tA)
        m = mA.T @ mT.T
        t = -mA.T @ (mT.T @ tT) - mA.T @ tA
        verts = None
        if s.aperture_kind == 5:
            tab = rec.tables
            o, c = s.polygon_offset, s.polygon_count
            verts = np.stack([tab[o:o + c], tab[o + c:o + 2 * c]], 1)
        edges.append(ApertureEdge(k, int(s.aperture_kind), np.array(s.aperture_p[:]), verts,
                                  bool(s.aperture_flags & 2), m, t))
    return edges


# ---------------------------------------------------------------------------
# the surrogate
# ---------------------------------------------------------------------------

class Surrogate:
    """Learned surrogate of ``system.rayfunction`` for diffraction-order
    path ``path``, over the input box [lo, hi] (field x, y in degrees;
    pupil x, y in mm; wavelength in nm)."""

    def __init__(self, system, lo, hi, path: int = 0, pixel: float = 0.015,
                 edge_margin: float = 0.5, bound_px: float = 0.05, seed: int = 0):
        self.system = system
        self.path = path
        self.norm = Normalisation(np.asarray(lo, float), np.asarray(hi, float))
        self.pixel = pixel
        self.edge_margin = edge_margin
        self.bound_px = bound_px
        self.dev = _trace.device()
        self.mlp = MLP(self.dev, seed)
        self.edges = _aperture_edges(system)
        self.coef = None       # polynomial coefficients [NUM_POLY][targets]
        self.scale = None      # residual scales of the three network targets
        self.err_map = None    # [16][8][8] flags of input cells above the bound
        self.relevant = []     # apertures that come within reach of the rays

    # -- data ------------------------------------------------------------
    def normalised(self, fx, fy, px, py, wl) -> torch.Tensor:
        u = self.norm.apply([fx.double(), fy.double(), px.double(), py.double(), wl.double()])
        return torch.stack(u, 1)

    def trace64(self, fx, fy, px, py, wl) -> dict:
        """The float64 trace: sensor-frame position and intensity of path
        ``path``, plus the aperture-frame position at every aperture."""
        from .vectors import Cartesian2dVectorArray

        kw = dict(normalized_field=False, normalized_pupil=False, precision="f64")
        rf = self.system.rayfunction(wavelength=wl, field=Cartesian2dVectorArray(fx, fy),
                                     pupil=Cartesian2dVectorArray(px, py), **kw).outputs
        acc = self.system.raytrace(wavelength=wl, field=Cartesian2dVectorArray(fx, fy),
                                   pupil=Cartesian2dVectorArray(px, py), accumulate=True,
                                   **kw).outputs
        r = rf if rf.position.x.ndim == 1 else rf[self.path]
        out = {"x": r.position.x, "y": r.position.y, "intensity": r.intensity,
               "unvignetted": r.unvignetted, "ap": []}
        for e in self.edges:
            # accumulate axis 0 = surface, then path
            s = acc[e.surface]
            s = s if s.position.x.ndim == 1 else s[self.path]
            p = torch.stack([s.position.x, s.position.y, s.position.z], 0)
            m = torch.as_tensor(e.m, device=p.device)
            t = torch.as_tensor(e.t, device=p.device)
            q = m @ p + t[:, None]
            out["ap"].append((q[0], q[1]))
        return out

    def _targets(self, d: dict) -> torch.Tensor:
        cols = [d["x"], d["y"], torch.log(torch.clamp(d["intensity"], min=1e-300))]
        for ax, ay in d["ap"]:
            cols += [ax, ay]
        return torch.stack(cols, 1)

    # -- fitting -----------------------------------------------------------
    def fit_polynomial(self, u: torch.Tensor, targets: torch.Tensor, valid: torch.Tensor):
        """Least squares of every target on the cubic monomials: the normal
        equations accumulated in float64 on the device (optk_poly_normal),
        solved on the host."""
        n, nt = u.shape[0], targets.shape[1]
        u = u.double().contiguous()
        tg = targets.double().contiguous()
        ok = valid.to(torch.uint8).contiguous()
        out = torch.empty(NUM_POLY, NUM_POLY + nt, dtype=torch.float64, device=self.dev)
        nbytes = int(_lib.lib().optk_poly_workspace(n))
        ws = torch.empty(nbytes // 8 + 1, dtype=torch.float64, device=self.dev)
        _lib.check(_lib.lib().optk_poly_normal(
            _ptr(u), _ptr(tg), _ptr(ok), n, nt, _EXP_CODES.ctypes.data_as(ctypes.c_void_p),
            _ptr(out), _ptr(ws), nbytes, _stream()))
        m = out.cpu().numpy()
        coef = np.linalg.solve(m[:, :NUM_POLY], m[:, NUM_POLY:])
        self.coef = torch.as_tensor(coef, dtype=torch.float64, device=self.dev)

    def poly(self, u: torch.Tensor) -> torch.Tensor:
        """The fitted polynomial at normalised inputs u [n][5]: [n][targets]
        (optk_poly_eval, float64)."""
        u = u.double().contiguous()
        n, nt = u.shape[0], self.coef.shape[1]
        out = torch.empty(n, nt, dtype=torch.float64, device=self.dev)
        coef = self.coef.contiguous()
        _lib.check(_lib.lib().optk_poly_eval(_ptr(u), _ptr(coef), n, nt,
                                             _EXP_CODES.ctypes.data_as(ctypes.c_void_p),
                                             _ptr(out), _stream()))
        return out

    # -- training ----------------------------------------------------------
    def prepare(self, grid_fn, n_train: int, chunk: int = 1 << 22, poly_rays: int = 1 << 22):
        """Trace ``n_train`` rays from ``grid_fn(count, seed)`` with the
        float64 trace, fit the polynomial, and keep the network's inputs
        and normalised residual targets on the device.  ``grid_fn`` returns
        a grid of host arrays, or the five input columns (field x, y, pupil
        x, y, wavelength) as device tensors already."""
        dev = self.dev
        us, ts, oks = [], [], []
        done, seed = 0, 1000
        while done < n_train:
            m = min(chunk, n_train - done)
            g = grid_fn(m, seed)
            seed += 1
            if isinstance(g, (list, tuple)):
                fx, fy, px, py, wl = [torch.as_tensor(a, device=dev).double() for a in g]
            else:
                t = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev).double()  # noqa: E731
                fx, fy, px, py, wl = t(g.field_x), t(g.field_y), t(g.pupil_x), t(g.pupil_y), t(g.wavelength)
            d = self.trace64(fx, fy, px, py, wl)
            u = self.normalised(fx, fy, px, py, wl)
            tg = self._targets(d)
            ok = torch.isfinite(tg).all(1)
            us.append(u)
            ts.append(tg)
            oks.append(ok)
            done += m
        u = torch.cat(us)
        tg = torch.cat(ts)
        ok = torch.cat(oks)
        k = min(poly_rays, u.shape[0])
        self.fit_polynomial(u[:k], tg[:k], ok[:k])
        # apertures the rays come near (anything within 4 margins of an edge)
        self.relevant = []
        for i, e in enumerate(self.edges):
            ax, ay = tg[ok, 3 + 2 * i], tg[ok, 4 + 2 * i]
            if torch.abs(e.distance(ax, ay)).min().item() < 4 * self.edge_margin:
                self.relevant.append(i)
        res = tg[:, :3] - self.poly(u)[:, :3]
        self.scale = res[ok].std(0).clamp(min=1e-12)
        self.train_u = u[ok].float().contiguous()
        self.train_t = (res[ok] / self.scale).float().contiguous()
        return int(ok.sum().item())

    def train(self, epochs: float = 1.0, batch: int = 65536, lr: float = 2e-3,
              weights=(1.0, 1.0, 0.25), log_every: int = 0) -> list:
        """Adam over the prepared rays (cosine-decayed learning rate)."""
        self.invalidate()
        n = self.train_u.shape[0]
        steps = max(1, int(epochs * n / batch))
        w = torch.tensor(weights, dtype=torch.float32, device=self.dev)
        losses = []
        g = torch.Generator(device=self.dev).manual_seed(1234)
        perm = None
        pos = n
        for step in range(steps):
            if pos + batch > n:
                perm = torch.randperm(n, device=self.dev, generator=g)
                pos = 0
            idx = perm[pos:pos + batch]
            pos += batch
            x = self.train_u.index_select(0, idx).contiguous()
            t = self.train_t.index_select(0, idx).contiguous()
            loss = self.mlp.loss_grad(x, t, w)
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / steps))
            self.mlp.adam_step(cur)
            if log_every and (step % log_every == 0 or step == steps - 1):
                losses.append((step, float(loss.item())))
        return losses

    # -- prediction --------------------------------------------------------
    def predict(self, fx, fy, px, py, wl):
        """Surrogate outputs for float32/float64 device tensors: dict of
        x, y, intensity (float64), unvignetted, fallback (bool)."""
        n = fx.numel()
        u = self.normalised(fx, fy, px, py, wl)
        P = self.poly(u)
        pad = (-n) % 64
        u32 = u.float()
        if pad:
            u32 = torch.cat([u32, torch.zeros(pad, 5, dtype=torch.float32, device=self.dev)])
        y = self.mlp.forward(u32.contiguous())[:n].double()
        out = P[:, :3] + y * self.scale
        near = torch.zeros(n, dtype=torch.bool, device=self.dev)
        inside = torch.ones(n, dtype=torch.bool, device=self.dev)
        for i in self.relevant:
            e = self.edges[i]
            d = e.distance(P[:, 3 + 2 * i], P[:, 4 + 2 * i])
            near |= torch.abs(d) < self.edge_margin
            inside &= d > 0
        fallback = near | ~torch.isfinite(out).all(1)
        if self.err_map is not None:
            fallback |= self._cell_flag(u)
        return {"x": out[:, 0], "y": out[:, 1], "intensity": torch.exp(out[:, 2]),
                "unvignetted": inside, "fallback": fallback}

    def _cell(self, u):
        b = torch.clamp(((u[:, [4, 0, 1]] + 1) * 0.5 * torch.tensor([16, 8, 8], device=u.device,
                                                                     dtype=u.dtype)).long(), 0)
        b = torch.minimum(b, torch.tensor([15, 7, 7], device=u.device))
        return (b[:, 0] * 8 + b[:, 1]) * 8 + b[:, 2]

    def _cell_flag(self, u):
        return self.err_map[self._cell(u)]

    def calibrate(self, fx, fy, px, py, wl, ref=None):
        """Flag input cells (wavelength x field x x field y, 16 x 8 x 8)
        whose worst validation error exceeds the bound."""
        if ref is None:
            ref = self.trace64(fx, fy, px, py, wl)
        self.invalidate()
        self.err_map = None
        p = self.predict(fx, fy, px, py, wl)
        ok = ref["unvignetted"] & p["unvignetted"] & ~p["fallback"]
        err = torch.sqrt((p["x"] - ref["x"]) ** 2 + (p["y"] - ref["y"]) ** 2) / self.pixel
        cell = self._cell(self.normalised(fx, fy, px, py, wl))
        worst = torch.zeros(16 * 8 * 8, dtype=torch.float64, device=self.dev)
        worst.scatter_reduce_(0, cell[ok], err[ok], reduce="amax")
        self.err_map = worst > self.bound_px
        return worst

    def run(self, fx, fy, px, py, wl, precision: str = "f32"):
        """Surrogate outputs with the flagged rays traced: returns the
        prediction dict (x, y, intensity, unvignetted) and the fallback
        mask."""
        p = self.predict(fx, fy, px, py, wl)
        fb = torch.nonzero(p["fallback"]).reshape(-1)
        if fb.numel():
            from .vectors import Cartesian2dVectorArray

            dt = torch.float64 if precision == "f64" else torch.float32
            sel = lambda a: a.index_select(0, fb).to(dt).contiguous()  # noqa: E731
            r = self.system.rayfunction(
                wavelength=sel(wl), field=Cartesian2dVectorArray(sel(fx), sel(fy)),
                pupil=Cartesian2dVectorArray(sel(px), sel(py)), normalized_field=False,
                normalized_pupil=False, precision=precision).outputs
            r = r if r.position.x.ndim == 1 else r[self.path]
            p["x"] = p["x"].index_copy(0, fb, r.position.x.double())
            p["y"] = p["y"].index_copy(0, fb, r.position.y.double())
            p["intensity"] = p["intensity"].index_copy(0, fb, r.intensity.double())
            p["unvignetted"] = p["unvignetted"].index_copy(0, fb, r.unvignetted)
        return p

    # -- fused inference -----------------------------------------------------
    def _fused_desc(self):
        """The optk_surrogate_desc of the trained model (weights are read in
        place from the parameter tensors)."""
        L = _lib
        d = L.SurrogateDesc()
        m = self.mlp
        d.w1 = m.view(m.p, 0, 0).data_ptr()
        d.b1 = m.view(m.p, 0, 1).data_ptr()
        for k, layer in enumerate((1, 2, 3)):
            d.w16[k] = m.view(m.p16, layer, 0).data_ptr()
            d.b[k] = m.view(m.p, layer, 1).data_ptr()
        d.w5 = m.view(m.p, 4, 0).data_ptr()
        d.b5 = m.view(m.p, 4, 1).data_ptr()
        # the first layer on MFMA: row o = [hi | hi | lo | 0] of w1[o]
        w1 = m.view(m.p, 0, 0).reshape(WIDTHS[1], WIDTHS[0])
        hi1 = _bf16_round(w1)
        rows = torch.zeros(WIDTHS[1], 32, dtype=torch.float32, device=self.dev)
        rows[:, 0:5] = hi1
        rows[:, 5:10] = hi1
        rows[:, 10:15] = w1 - hi1
        self._w1_16 = torch.empty(WIDTHS[1] * 32, dtype=torch.int16, device=self.dev)
        _op(_lib.MLP_TO_BF16, f32=[rows.reshape(-1)], u16=[self._w1_16], rows=1, cols=WIDTHS[1] * 32)
        d.w1_16 = self._w1_16.data_ptr()
        # the output layer runs on MFMA in the fused kernel: bf16 high and
        # low parts of the rows, padded to one 16-row tile
        w5 = m.view(m.p, 4, 0).reshape(3, WIDTHS[4])
        hi = _bf16_round(w5)
        tile = torch.zeros(16, WIDTHS[4], dtype=torch.float32, device=self.dev)
        tile[:3] = hi
        tile[3:6] = w5 - hi
        self._w5_16 = torch.empty(16 * WIDTHS[4], dtype=torch.int16, device=self.dev)
        _op(_lib.MLP_TO_BF16, f32=[tile.reshape(-1)], u16=[self._w5_16], rows=1, cols=16 * WIDTHS[4])
        d.w5_16 = self._w5_16.data_ptr()
        T = 3 + 2 * len(self.relevant)
        if len(self.relevant) > L.SURROGATE_MAX_APERTURES:
            raise ValueError("too many apertures within reach for the fused kernel")
        cols = [0, 1, 2]
        for i in self.relevant:
            cols += [3 + 2 * i, 4 + 2 * i]
        coef = torch.zeros(NUM_POLY, L.SURROGATE_MAX_TARGETS, dtype=torch.float32, device=self.dev)
        coef[:, :T] = self.coef[:, cols].float()
        self._coef32 = coef.contiguous()
        d.poly = self._coef32.data_ptr()
        if self.err_map is not None:
            self._emap = self.err_map.to(torch.uint8).contiguous()
            d.error_map = self._emap.data_ptr()
        mid = (self.norm.lo + self.norm.hi) / 2
        half = (self.norm.hi - self.norm.lo) / 2
        for k in range(5):
            d.in_mid[k] = float(mid[k])
            d.in_inv_half[k] = float(1.0 / half[k])
        for k in range(3):
            d.scale[k] = float(self.scale[k])
        d.edge_margin = self.edge_margin
        for t, e in enumerate(EXPONENTS):
            d.poly_exp[t] = int(_EXP_CODES[t])
        d.num_targets = T
        d.num_apertures = len(self.relevant)
        for a, i in enumerate(self.relevant):
            e = self.edges[i]
            ap = d.apertures[a]
            ap.kind = e.kind if e.kind in (1, 4, 5) else 0
            ap.inverted = 1 if e.inverted else 0
            for k in range(min(4, len(e.params))):
                ap.p[k] = float(e.params[k])
            if e.vertices is not None:
                if len(e.vertices) > L.SURROGATE_MAX_VERTICES:
                    raise ValueError("polygon with too many vertices for the fused kernel")
                ap.num_vertices = len(e.vertices)
                for v, (vx, vy) in enumerate(e.vertices):
                    ap.vx[v], ap.vy[v] = float(vx), float(vy)
        return d

    def infer(self, fx, fy, px, py, wl, out=None, want_list=True):
        """Fused inference (one kernel): float32 device inputs; returns a
        dict of x, y, intensity (float32), unvignetted, fallback (bool) and
        the fallback ray indices (when ``want_list``)."""
        n = fx.numel()
        dev = self.dev
        if out is None:
            out = {"x": torch.empty(n, dtype=torch.float32, device=dev),
                   "y": torch.empty(n, dtype=torch.float32, device=dev),
                   "intensity": torch.empty(n, dtype=torch.float32, device=dev),
                   "unvignetted": torch.empty(n, dtype=torch.bool, device=dev),
                   "fallback": torch.empty(n, dtype=torch.bool, device=dev),
                   "list": torch.empty(n if want_list else 0, dtype=torch.int32, device=dev),
                   "count": torch.zeros(1, dtype=torch.int32, device=dev)}
        else:
            out["count"].zero_()
        if not hasattr(self, "_desc_cache"):
            self._desc_cache = self._fused_desc()
        io = _lib.SurrogateIO()
        for name, t in (("field_x", fx), ("field_y", fy), ("pupil_x", px), ("pupil_y", py),
                        ("wavelength", wl)):
            if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n:
                raise ValueError(f"{name}: expected a contiguous float32 tensor of {n} rays")
            setattr(io, name, t.data_ptr())
        io.x, io.y, io.intensity = out["x"].data_ptr(), out["y"].data_ptr(), out["intensity"].data_ptr()
        io.unvignetted, io.fallback = out["unvignetted"].data_ptr(), out["fallback"].data_ptr()
        if want_list:
            io.fallback_list = out["list"].data_ptr()
            io.fallback_count = out["count"].data_ptr()
            io.fallback_capacity = out["list"].numel()
        io.n = n
        _lib.check(_lib.lib().optk_surrogate_infer(ctypes.byref(self._desc_cache), ctypes.byref(io),
                                                   _stream()))
        return out

    def _fallback_plan(self, n: int, precision: str):
        """A prepared sensor-frame trace of ``n`` rays through the system
        with every multi-order grating restricted to the surrogate's order
        path (one path: the other order is not part of the emulated rays).
        The last plan is kept, so repeated calls with the same fallback
        count allocate nothing."""
        key = (n, precision)
        if getattr(self, "_fb_key", None) != key:
            from .stops import single_path

            one = dataclasses.replace(self.system, surfaces=single_path(list(self.system.surfaces),
                                                                         self.path))
            self._fb_plan = None  # release the old buffers first
            self._fb_plan = _trace.GridTracePlan(one, n, sensor_frame=True, precision=precision)
            self._fb_key = key
        return self._fb_plan

    def infer_traced(self, fx, fy, px, py, wl, out=None, precision: str = "f32"):
        """Fused inference with the fallback rays traced: the surrogate's
        outputs for every ray, where the rays it flags carry the trace's
        position, intensity and mask (``precision`` "f32" or "f64") instead
        of the prediction.  One host read of the fallback count per call;
        the flagged rays' inputs are gathered on the device, traced in one
        launch and scattered back."""
        out = self.infer(fx, fy, px, py, wl, out=out)
        count = int(out["count"].item())
        if count > out["list"].numel():  # the list overflowed: take the mask
            idx = torch.nonzero(out["fallback"]).reshape(-1)
        else:
            idx = out["list"][:count].long()
        if idx.numel() == 0:
            return out
        plan = self._fallback_plan(idx.numel(), precision)
        sel = [t.index_select(0, idx).to(plan.dtype) for t in (fx, fy, px, py, wl)]
        r = plan(*sel)
        out["x"].index_copy_(0, idx, r.position.x.to(out["x"].dtype))
        out["y"].index_copy_(0, idx, r.position.y.to(out["y"].dtype))
        out["intensity"].index_copy_(0, idx, r.intensity.to(out["intensity"].dtype))
        out["unvignetted"].index_copy_(0, idx, r.unvignetted)
        return out

    def release_fallback_plan(self):
        """Free the buffers of the cached fallback trace plan."""
        self._fb_plan = None
        self._fb_key = None

    def invalidate(self):
        """Forget the fused descriptor (after retraining or recalibrating)."""
        if hasattr(self, "_desc_cache"):
            del self._desc_cache
