"""CPU reference of the learned-proposal U-Net (test infrastructure).

A plain numpy restatement of the network the library runs on the GPU
(csrc/fs_unet.hip): NHWC tensors, 3x3 convs with zero padding + ReLU, 2x2
average pooling, nearest 2x upsampling concatenated with the skip, a 1x1
head, loss mean((y - t)^2), and the hand-derived backward pass.  The
parameter vector has the library's layout: per conv layer W[tap][out][in]
then bias[out]; then the head's weights and bias.

``bf16=True`` rounds to BF16 where the device stores BF16 (input, weights,
every activation and every activation gradient), so the GPU can be
compared to it closely; ``bf16=False`` with float64 is the exact reference
that the finite-difference tests pin.
"""
import numpy as np

IN_CHANNELS = 8
WIDTHS = (32, 64, 128, 160)


def bf16_round(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def layer_shapes(widths=WIDTHS):
    """[(cin, cout, level)] in parameter order: encoder levels 0..3 (two convs
    each), then decoder levels 2, 1, 0 (two convs each)."""
    L = len(widths)
    out = []
    for l in range(L):
        out.append((IN_CHANNELS if l == 0 else widths[l - 1], widths[l], l))
        out.append((widths[l], widths[l], l))
    for l in range(L - 2, -1, -1):
        below = widths[l + 1]
        out.append((below + widths[l], widths[l], l))
        out.append((widths[l], widths[l], l))
    return out


def param_count(widths=WIDTHS):
    n = sum(9 * ci * co + co for ci, co, _ in layer_shapes(widths))
    return n + widths[0] + 1


def unpack(p, widths=WIDTHS):
    convs, off = [], 0
    for ci, co, _ in layer_shapes(widths):
        w = p[off:off + 9 * co * ci].reshape(9, co, ci)
        off += 9 * co * ci
        b = p[off:off + co]
        off += co
        convs.append((w, b))
    hw = p[off:off + widths[0]]
    hb = p[off + widths[0]]
    return convs, hw, hb


def init_params(seed, widths=WIDTHS):
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co, _ in layer_shapes(widths):
        parts.append(rng.normal(0, np.sqrt(2.0 / (9 * ci)), 9 * co * ci))
        parts.append(np.zeros(co))
    parts.append(rng.normal(0, np.sqrt(1.0 / widths[0]), widths[0]))
    parts.append(np.zeros(1))
    return np.concatenate(parts)


def _shift(x, t):
    """x[b, y + dy, x + dx, c] with zero padding, (dy, dx) = (t//3 - 1, t%3 - 1)."""
    B, H, W, C = x.shape
    xp = np.zeros((B, H + 2, W + 2, C), dtype=x.dtype)
    xp[:, 1:-1, 1:-1] = x
    dy, dx = t // 3, t % 3
    return xp[:, dy:dy + H, dx:dx + W]


def _unshift_add(g, t, out):
    """Adjoint of _shift: out += the gradient g of _shift(x, t) w.r.t. x."""
    B, H, W, C = g.shape
    gp = np.zeros((B, H + 2, W + 2, C), dtype=g.dtype)
    dy, dx = t // 3, t % 3
    gp[:, dy:dy + H, dx:dx + W] = g
    out += gp[:, 1:-1, 1:-1]


class Net:
    def __init__(self, params, widths=WIDTHS, bf16=False, dtype=np.float64):
        self.widths = widths
        self.bf16 = bf16
        self.dtype = dtype
        p = np.asarray(params, dtype=dtype)
        if bf16:
            convs, hw, hb = unpack(p, widths)
            self.convs = [(bf16_round(w).astype(dtype), b.astype(dtype)) for w, b in convs]
            self.hw, self.hb = hw.astype(dtype), dtype(hb)
        else:
            self.convs, self.hw, self.hb = unpack(p, widths)
        self.shapes = layer_shapes(widths)

    def _r(self, a):
        return bf16_round(a).astype(self.dtype) if self.bf16 else a

    def _conv(self, i, x):
        w, b = self.convs[i]
        z = np.zeros(x.shape[:3] + (w.shape[1],), dtype=self.dtype)
        for t in range(9):
            z += np.einsum("bhwc,oc->bhwo", _shift(x, t), w[t], optimize=True)
        return self._r(np.maximum(z + b, 0))

    def forward(self, terrain, params):
        B, H, W = terrain.shape
        x = np.zeros((B, H, W, IN_CHANNELS), dtype=self.dtype)
        x[..., 0] = terrain
        x[..., 1:4] = np.asarray(params, dtype=self.dtype)[:, None, None, :]
        x = self._r(x)
        L = len(self.widths)
        c = self.cache = {"x0": x}
        inp = x
        for l in range(L):
            c["ea%d" % l] = self._conv(2 * l, inp)
            c["e%d" % l] = self._conv(2 * l + 1, c["ea%d" % l])
            if l < L - 1:
                e = c["e%d" % l]
                inp = c["p%d" % l] = self._r(
                    0.25 * (e[:, 0::2, 0::2] + e[:, 0::2, 1::2] + e[:, 1::2, 0::2] + e[:, 1::2, 1::2]))
        lo = c["e%d" % (L - 1)]
        for l in range(L - 2, -1, -1):
            up = lo.repeat(2, axis=1).repeat(2, axis=2)
            c["u%d" % l] = np.concatenate([up, c["e%d" % l]], axis=-1)
            k = 2 * L + 2 * (L - 2 - l)
            c["da%d" % l] = self._conv(k, c["u%d" % l])
            c["d%d" % l] = self._conv(k + 1, c["da%d" % l])
            lo = c["d%d" % l]
        y = np.einsum("bhwc,c->bhw", c["d0"], self.hw) + self.hb
        c["y"] = y
        return y

    def loss_grad(self, terrain, params, target):
        """(loss, gradient vector in the library's parameter layout)."""
        y = self.forward(terrain, params)
        c = self.cache
        M = y.size
        e = y - target
        loss = float(np.sum(e * e) / M)
        gy = 2.0 * e / M
        L = len(self.widths)
        grads = [None] * len(self.convs)
        ghw = np.einsum("bhw,bhwc->c", gy, c["d0"])
        ghb = gy.sum()
        g = {"d0": self._r(gy[..., None] * self.hw)}

        def conv_back(i, x, yout, gout, need_input):
            w, _ = self.convs[i]
            gz = self._r(np.where(yout > 0, gout, 0))
            gw = np.stack([np.einsum("bhwo,bhwc->oc", gz, _shift(x, t), optimize=True)
                           for t in range(9)])
            gb = gz.sum(axis=(0, 1, 2))
            grads[i] = (gw, gb)
            if not need_input:
                return None
            gx = np.zeros(x.shape, dtype=self.dtype)
            for t in range(9):
                _unshift_add(np.einsum("bhwo,oc->bhwc", gz, w[t], optimize=True), t, gx)
            return self._r(gx)

        for l in range(0, L - 1):
            k = 2 * L + 2 * (L - 2 - l)
            g["da%d" % l] = conv_back(k + 1, c["da%d" % l], c["d%d" % l], g["d%d" % l], True)
            gu = conv_back(k, c["u%d" % l], c["da%d" % l], g["da%d" % l], True)
            below = c["u%d" % l].shape[-1] - self.widths[l]
            gup = gu[..., :below]
            glo = self._r(gup[:, 0::2, 0::2] + gup[:, 0::2, 1::2] + gup[:, 1::2, 0::2] +
                          gup[:, 1::2, 1::2])
            g["d%d" % (l + 1) if l + 1 <= L - 2 else "e%d" % (L - 1)] = glo
            g["skip%d" % l] = gu[..., below:]
        for l in range(L - 1, -1, -1):
            ge = g["e%d" % l] if l == L - 1 else None
            if l < L - 1:
                gp = g["p%d" % l]
                ge = self._r(g["skip%d" % l] + 0.25 * gp.repeat(2, axis=1).repeat(2, axis=2))
            g["ea%d" % l] = conv_back(2 * l + 1, c["ea%d" % l], c["e%d" % l], ge, True)
            inp = c["x0"] if l == 0 else c["p%d" % (l - 1)]
            r = conv_back(2 * l, inp, c["ea%d" % l], g["ea%d" % l], l > 0)
            if l > 0:
                g["p%d" % (l - 1)] = r
        parts = []
        for gw, gb in grads:
            parts.append(gw.ravel())
            parts.append(gb.ravel())
        parts.append(ghw.ravel())
        parts.append(np.array([ghb]))
        return loss, np.concatenate(parts)
