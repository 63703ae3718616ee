"""CPU restatement of the learned preconditioner (3-level 3-D U-Net) -- TEST INFRASTRUCTURE ONLY: the checker of the
HIP network kernels, used by the tests; never by the product library.

The network has no counterpart in the reference (its preconditioner is the multigrid V-cycle,
SloshPressureMG.usf); it is this build's own component, so the restatement below is its definition:

  input (level 0, 4 channels): r / s, fluid, air, solid       (s = RMS of r over fluid cells)
  E0a conv3 4->16 + ReLU, E0b conv3 16->16 + ReLU = s0        (level 0: the grid)
  pool (2^3 mean) -> E1a conv3 16->32 + ReLU, E1b 32->32 + ReLU = s1   (level 1: half resolution)
  pool -> B2a conv3 32->64 + ReLU, B2b 64->64 + ReLU, D1p conv3 64->32 (no activation)   (level 2)
  t1 = ReLU(up(D1p) + s1); D1b conv3 32->32 + ReLU; D0p conv3 32->16 (no activation)     (level 1)
  t0 = ReLU(up(D0p) + s0); D0b conv3 16->16 + ReLU; head 1x1 16->1 on fluid cells = zhat  (level 0)
  preconditioned residual z = s * zhat

Convolutions are correlations over the 3x3x3 neighbourhood (tap t = (dx+1) + 3(dy+1) + 9(dz+1)) with zero
padding: cells of unallocated bricks and outside the grid read 0, and every level's output is kept only on cells
of allocated bricks. Weights and activations are BF16 (round to nearest even) with FP32 (here FP64) sums; the
parameter vector is, per layer, W[27][cout][cin] then bias[cout] (head: W[cin], bias).

Training loss on one system: L = sum_fluid (A zhat - rhat)^2 / sum_fluid rhat^2, rhat = r / s, with A the
pressure operator (diag + 1e-4) x - sum_{fluid n} x_n on fluid cells.
"""
import numpy as np

# (name, cin, cout, level, kernel taps, relu)
LAYERS = [
    ("E0a", 4, 16, 0, 27, True),
    ("E0b", 16, 16, 0, 27, True),
    ("E1a", 16, 32, 1, 27, True),
    ("E1b", 32, 32, 1, 27, True),
    ("B2a", 32, 64, 2, 27, True),
    ("B2b", 64, 64, 2, 27, True),
    ("D1p", 64, 32, 2, 27, False),
    ("D1b", 32, 32, 1, 27, True),
    ("D0p", 32, 16, 1, 27, False),
    ("D0b", 16, 16, 0, 27, True),
    ("H", 16, 1, 0, 1, False),
]
DIAG_EPS = np.float32(1e-4)


def layer_offsets():
    """(weight offset, bias offset) of each layer in the flat parameter vector, and the total count."""
    offs, o = [], 0
    for _, cin, cout, _, taps, _ in LAYERS:
        offs.append((o, o + taps * cout * cin))
        o += taps * cout * cin + cout
    return offs, o


def init_params(seed=0):
    """He-normal weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    offs, n = layer_offsets()
    p = np.zeros(n, np.float32)
    for (wo, bo), (_, cin, cout, _, taps, relu) in zip(offs, LAYERS):
        std = np.sqrt((2.0 if relu else 1.0) / (taps * cin))
        p[wo:bo] = (rng.standard_normal(taps * cout * cin) * std).astype(np.float32)
    return p


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _shift(a, dz, dy, dx):
    """out[..., z, y, x] = a[..., z + dz, y + dy, x + dx] with zero outside."""
    out = np.zeros_like(a)
    Z, Y, X = a.shape[-3:]
    zs, ys, xs = slice(max(0, -dz), Z - max(0, dz)), slice(max(0, -dy), Y - max(0, dy)), slice(max(0, -dx), X - max(0, dx))
    zt, yt, xt = slice(max(0, dz), Z - max(0, -dz)), slice(max(0, dy), Y - max(0, -dy)), slice(max(0, dx), X - max(0, -dx))
    out[..., zs, ys, xs] = a[..., zt, yt, xt]
    return out


TAPS = [((t % 3) - 1, ((t // 3) % 3) - 1, (t // 9) - 1) for t in range(27)]  # (dx, dy, dz)


def conv3(x, w, b, mask):
    """x (cin, Z, Y, X) float64, w (27, cout, cin), b (cout,) -> (cout, Z, Y, X) on `mask` cells."""
    out = np.zeros((w.shape[1],) + x.shape[1:])
    for t, (dx, dy, dz) in enumerate(TAPS):
        out += np.einsum("oc,czyx->ozyx", w[t], _shift(x, dz, dy, dx))
    out += b[:, None, None, None]
    return out * mask


def conv3_t(g, w, mask_in):
    """Adjoint of conv3 in its input: g (cout, ...) -> (cin, ...), kept on the input level's cells."""
    out = np.zeros((w.shape[2],) + g.shape[1:])
    for t, (dx, dy, dz) in enumerate(TAPS):
        out += _shift(np.einsum("oc,ozyx->czyx", w[t], g), -dz, -dy, -dx)
    return out * mask_in


def conv3_dw(x, g):
    """Gradient of conv3 in its weights: (27, cout, cin), and in its bias."""
    dw = np.stack([np.einsum("ozyx,czyx->oc", g, _shift(x, dz, dy, dx)) for (dx, dy, dz) in TAPS])
    return dw, g.sum(axis=(1, 2, 3))


def pool(x):
    """Mean over 2x2x2 children, summed in child order (x fastest) then divided by 8."""
    C, Z, Y, X = x.shape
    s = np.zeros((C, Z // 2, Y // 2, X // 2))
    for oz in range(2):
        for oy in range(2):
            for ox in range(2):
                s = s + x[:, oz::2, oy::2, ox::2]
    return s / 8.0


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def down_sum(x):
    """Adjoint of up: the sum over the 2x2x2 children."""
    C, Z, Y, X = x.shape
    return x.reshape(C, Z // 2, 2, Y // 2, 2, X // 2, 2).sum(axis=(2, 4, 6))


def params_to_layers(p, bf=True):
    offs, _ = layer_offsets()
    out = []
    for (wo, bo), (_, cin, cout, _, taps, _) in zip(offs, LAYERS):
        w = bf16(p[wo:bo].astype(np.float32)) if bf else p[wo:bo]
        out.append((w.reshape(taps, cout, cin).astype(np.float64), p[bo:bo + cout].astype(np.float64)))
    return out


class Domain:
    """Masks and operator of one pressure system on a grid whose dims are multiples of 8.

    ctype: (Z, Y, X) uint32 cell flags (bits 0-1 type); alloc: (Z, Y, X) bool, cells of allocated bricks;
    diag: (Z, Y, X) float32 ghost-fluid diagonal (without the 1e-4)."""

    def __init__(self, ctype, alloc, diag):
        t = np.asarray(ctype) & 3
        self.alloc = [alloc.astype(np.float64)]
        for _ in range(2):
            a = self.alloc[-1]
            self.alloc.append(a[::2, ::2, ::2] * 1.0)  # bricks are 8^3: a coarse cell is allocated iff its brick is
        self.fluid = ((t == 1) & alloc).astype(np.float64)
        self.air = ((t == 0) & alloc).astype(np.float64)
        self.solid = ((t == 2) & alloc).astype(np.float64)
        self.diag = np.asarray(diag, np.float32)

    def apply_a(self, x):
        """(diag + 1e-4) x - sum over fluid neighbours, on fluid cells (FP64)."""
        f = self.fluid
        s = np.zeros_like(x)
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            s += _shift(x * f, *d)
        dg = (self.diag + DIAG_EPS).astype(np.float64)
        return (dg * x - s) * f


def forward(p, dom, r, keep=False, exact=False):
    """z = s * zhat for residual r (Z, Y, X); with keep=True also returns the intermediate tensors. exact=True
    drops every BF16 rounding (the smooth FP64 function that finite differences check)."""
    L = params_to_layers(p, bf=not exact)
    rnd = (lambda v: np.asarray(v, np.float64)) if exact else (lambda v: bf16(np.asarray(v, np.float32)).astype(np.float64))
    f = dom.fluid
    nf = max(f.sum(), 1.0)
    rr = float((np.asarray(r, np.float64) ** 2 * f).sum())
    s_inv = np.float32(1.0 / np.sqrt(rr / nf)) if rr > 0 else np.float32(0.0)
    s = np.float32(np.sqrt(rr / nf))
    rh = (np.asarray(r, np.float64) * float(s_inv) if exact else bf16(np.asarray(r, np.float32) * s_inv)) * f
    x0 = np.stack([rh, f, dom.air, dom.solid])
    m0, m1, m2 = dom.alloc
    act = lambda v, m: rnd(np.maximum(v, 0.0)) * m
    lin = lambda v, m: rnd(v) * m
    a0 = act(conv3(x0, *L[0], m0), m0)
    s0 = act(conv3(a0, *L[1], m0), m0)
    p0 = lin(pool(s0), m1)
    a1 = act(conv3(p0, *L[2], m1), m1)
    s1 = act(conv3(a1, *L[3], m1), m1)
    p1 = lin(pool(s1), m2)
    a2 = act(conv3(p1, *L[4], m2), m2)
    b2 = act(conv3(a2, *L[5], m2), m2)
    c2 = lin(conv3(b2, *L[6], m2), m2)
    t1 = act(up(c2) + s1, m1)
    d1 = act(conv3(t1, *L[7], m1), m1)
    c1 = lin(conv3(d1, *L[8], m1), m1)
    t0 = act(up(c1) + s0, m0)
    d0 = act(conv3(t0, *L[9], m0), m0)
    wh, bh = L[10]
    zh = (np.einsum("c,czyx->zyx", wh[0, 0], d0) + bh[0]) * f
    z = zh * float(s) if exact else zh.astype(np.float32) * s
    if keep:
        return z, dict(x0=x0, a0=a0, s0=s0, p0=p0, a1=a1, s1=s1, p1=p1, a2=a2, b2=b2, c2=c2, t1=t1, d1=d1, c1=c1,
                       t0=t0, d0=d0, zh=zh, rh=rh, s=s, L=L)
    return z


def loss_and_grad(p, dom, r, exact=False):
    """Loss L and dL/dp (FP64) for one system, the BF16 rounding treated as identity in the backward pass."""
    _, T = forward(p, dom, r, keep=True, exact=exact)
    L = T["L"]
    m0, m1, m2 = dom.alloc
    e = dom.apply_a(T["zh"]) - T["rh"]
    den = float((T["rh"] ** 2).sum())
    loss = float((e ** 2).sum()) / den
    gz = 2.0 * dom.apply_a(e) / den  # A is symmetric on fluid cells
    offs, n = layer_offsets()
    grad = np.zeros(n)

    def put(i, dw, db):
        wo, bo = offs[i]
        grad[wo:bo] = dw.ravel()
        grad[bo:bo + db.size] = db

    # head
    put(10, np.einsum("zyx,czyx->c", gz, T["d0"])[None, None, :], np.array([gz.sum()]))
    g = L[10][0][0, 0][:, None, None, None] * gz[None] * (T["d0"] > 0)
    # D0b (input t0)
    put(9, *conv3_dw(T["t0"], g))
    g = conv3_t(g, L[9][0], m0) * (T["t0"] > 0)
    g_s0 = g.copy()
    g_c1 = down_sum(g) * m1
    # D0p (input d1, no activation)
    put(8, *conv3_dw(T["d1"], g_c1))
    g = conv3_t(g_c1, L[8][0], m1) * (T["d1"] > 0)
    # D1b (input t1)
    put(7, *conv3_dw(T["t1"], g))
    g = conv3_t(g, L[7][0], m1) * (T["t1"] > 0)
    g_s1 = g.copy()
    g_c2 = down_sum(g) * m2
    # D1p (input b2)
    put(6, *conv3_dw(T["b2"], g_c2))
    g = conv3_t(g_c2, L[6][0], m2) * (T["b2"] > 0)
    # B2b (input a2)
    put(5, *conv3_dw(T["a2"], g))
    g = conv3_t(g, L[5][0], m2) * (T["a2"] > 0)
    # B2a (input p1)
    put(4, *conv3_dw(T["p1"], g))
    g = conv3_t(g, L[4][0], m2)
    g_s1 = (g_s1 + up(g) / 8.0 * m1) * (T["s1"] > 0)
    # E1b (input a1)
    put(3, *conv3_dw(T["a1"], g_s1))
    g = conv3_t(g_s1, L[3][0], m1) * (T["a1"] > 0)
    # E1a (input p0)
    put(2, *conv3_dw(T["p0"], g))
    g = conv3_t(g, L[2][0], m1)
    g_s0 = (g_s0 + up(g) / 8.0 * m0) * (T["s0"] > 0)
    # E0b (input a0)
    put(1, *conv3_dw(T["a0"], g_s0))
    g = conv3_t(g_s0, L[1][0], m0) * (T["a0"] > 0)
    # E0a (input x0)
    put(0, *conv3_dw(T["x0"], g))
    return loss, grad
