"""CPU ORACLE of the warm-start U-Net (test infrastructure only): the same architecture and
flat parameter layout as the library's rt_unet (forward and backward in numpy), used to check
the GPU gradients.  Itself pinned by central finite differences in float64 on a tiny net."""
from __future__ import annotations

import numpy as np


def layout(levels: int, base: int, in_ch: int):
    """[(name, cin, cout, w_offset, b_offset)] in parameter order, and the total count."""
    convs, off = [], 0

    def add(name, cin, cout):
        nonlocal off
        convs.append((name, cin, cout, off, off + 9 * cin * cout))
        off += 9 * cin * cout + cout

    for l in range(levels):
        add(f"enc{l}a", base << (l - 1) if l else in_ch, base << l)
        add(f"enc{l}b", base << l, base << l)
    for l in range(levels - 2, -1, -1):
        add(f"dec{l}up", base << (l + 1), base << l)
        add(f"dec{l}a", base << l, base << l)
    add("out", base, 1)
    return convs, off


def im2col(x):
    B, H, W, C = x.shape
    p = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    taps = [p[:, dy:dy + H, dx:dx + W, :] for dy in range(3) for dx in range(3)]
    return np.stack(taps, axis=3).reshape(B * H * W, 9 * C)


def col2im(dcol, B, H, W, C):
    d = dcol.reshape(B, H, W, 9, C)
    out = np.zeros((B, H + 2, W + 2, C), dcol.dtype)
    for t in range(9):
        dy, dx = divmod(t, 3)
        out[:, dy:dy + H, dx:dx + W, :] += d[:, :, :, t, :]
    return out[:, 1:H + 1, 1:W + 1, :]


class UNetRef:
    def __init__(self, levels, base, in_ch, params, dtype=np.float32):
        self.L, self.C, self.cin = levels, base, in_ch
        self.convs, self.n = layout(levels, base, in_ch)
        self.p = np.asarray(params, dtype=dtype)
        self.by = {c[0]: c for c in self.convs}

    def _wb(self, name):
        _, cin, cout, wo, bo = self.by[name]
        return self.p[wo:wo + 9 * cin * cout].reshape(9 * cin, cout), self.p[bo:bo + cout]

    def _conv(self, name, x, relu):
        W, b = self._wb(name)
        B, H, Wd, _ = x.shape
        y = (im2col(x) @ W + b).reshape(B, H, Wd, -1)
        return np.maximum(y, 0) if relu else y

    def forward(self, x):
        L, cache = self.L, {}
        e, xin = [], x
        for l in range(L):
            if l:
                p = e[-1]
                B, H, W, C = p.shape
                xin = p.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))
            t1 = self._conv(f"enc{l}a", xin, True)
            el = self._conv(f"enc{l}b", t1, True)
            cache[f"x{l}"], cache[f"t{l}"] = xin, t1
            e.append(el)
        d = e[-1]
        for l in range(L - 2, -1, -1):
            u = d.repeat(2, axis=1).repeat(2, axis=2)
            v = self._conv(f"dec{l}up", u, True)
            s = v + e[l]
            d = self._conv(f"dec{l}a", s, True)
            cache[f"u{l}"], cache[f"v{l}"], cache[f"s{l}"], cache[f"d{l}"] = u, v, s, d
        cache["e"] = e
        cache["d0"] = d
        return self._conv("out", d, False)[..., 0], cache

    def _conv_back(self, name, x, dy, grad):
        _, cin, cout, wo, bo = self.by[name]
        W, _ = self._wb(name)
        B, H, Wd, _ = x.shape
        col = im2col(x)
        g = dy.reshape(-1, cout)
        grad[wo:wo + 9 * cin * cout] = (col.T @ g).ravel()
        grad[bo:bo + cout] = g.sum(axis=0)
        return col2im(g @ W.T, B, H, Wd, cin)

    def loss_grad(self, x, target, mask):
        y, c = self.forward(x)
        m = (mask != 0).astype(self.p.dtype)
        cnt = m.sum()
        r = (y - target) * m
        loss = float((r * r).sum() / cnt)
        grad = np.zeros_like(self.p)
        dy = (2 * r * m / cnt)[..., None]
        dd = self._conv_back("out", c["d0"], dy, grad)
        e = c["e"]
        de = [np.zeros_like(t) for t in e]
        L = self.L
        for l in range(L - 1):
            dd = dd * (c[f"d{l}"] > 0)
            ds = self._conv_back(f"dec{l}a", c[f"s{l}"], dd, grad)
            de[l] += ds
            dv = ds * (c[f"v{l}"] > 0)
            du = self._conv_back(f"dec{l}up", c[f"u{l}"], dv, grad)
            B, H, W, C = du.shape
            dd = du.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))
            if l + 1 == L - 1:
                de[L - 1] += dd
        for l in range(L - 1, -1, -1):
            g = de[l] * (e[l] > 0)
            dt = self._conv_back(f"enc{l}b", c[f"t{l}"], g, grad) * (c[f"t{l}"] > 0)
            dx = self._conv_back(f"enc{l}a", c[f"x{l}"], dt, grad)
            if l:
                de[l - 1] += 0.25 * dx.repeat(2, axis=1).repeat(2, axis=2)
        return loss, grad
