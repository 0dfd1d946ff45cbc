"""CPU restatement of the U-Net preconditioner (numpy), for checking the HIP
path in carpios_amd/csrc/unet.hip and unet_kernels.hpp.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product package.  The network has no counterpart in the reference (whose
Poisson solvers are the CSR Krylov methods of lib/algebra/solver_matrix.hpp),
so parity is pinned here by
  * the primitives against textbook routines: conv3 against
    scipy.signal.correlate2d, the stride-4 transposed convolution against a
    Kronecker-product construction, pooling against a reshape mean;
  * the backward pass against central finite differences of the float64
    network (tests/test_unet_oracle.py);
  * the initialisation against an independent xoshiro256** + Box-Muller.

Two modes: ``bf16=False`` is the exact float64 network (what the finite
differences check); ``bf16=True`` rounds where the device rounds (BF16
weights of the 3x3 and transposed convolutions, BF16 activations, BF16
copies of the masked gradients fed to the matrix cores) so that the device
result agrees to accumulation-order effects.

Arrays are NHWC: (H, W, C), H rows of W contiguous pixels.
"""
import numpy as np

# (kind, cin, cout); kinds: first 3x3 2->16 in FP32 (residual, wall code),
# conv 3x3, up 4x4/4, out 1x1
LAYERS = [("first", 2, 16), ("conv", 16, 16), ("conv", 16, 32), ("conv", 32, 32),
          ("conv", 32, 64), ("conv", 64, 64), ("conv", 64, 128),
          ("conv", 128, 128), ("up", 128, 64), ("conv", 128, 64),
          ("conv", 64, 64), ("up", 64, 32), ("conv", 64, 32), ("conv", 32, 32),
          ("up", 32, 16), ("conv", 32, 16), ("conv", 16, 16), ("out", 16, 1)]


def ntaps(kind):
    return {"up": 16, "out": 1}.get(kind, 9)


def layout():
    """[(w_offset, b_offset, shape_of_w)] and the total parameter count;
    weights [tap][cout][cin] then biases [cout], layer after layer."""
    out, off = [], 0
    for kind, ci, co in LAYERS:
        shape = (ntaps(kind), co, ci)
        nw = int(np.prod(shape))
        out.append((off, off + nw, shape))
        off += nw + co
    return out, off


def unpack(params):
    lay, n = layout()
    assert params.size == n
    ws, bs = [], []
    for (wo, bo, shape), (_, _, co) in zip(lay, LAYERS):
        ws.append(params[wo:wo + int(np.prod(shape))].reshape(shape))
        bs.append(params[bo:bo + co])
    return ws, bs


def pack(ws, bs):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()])
                           for w, b in zip(ws, bs)])


def bf16(a):
    """Round to bfloat16 (nearest even) via float32; NaN stays NaN."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x7FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


# ----------------------------------------------------------------- primitives
def _shift(x, dy, dx):
    """s[i, j] = x[i + dy, j + dx] with zeros outside."""
    H, W = x.shape[:2]
    p = np.pad(x, ((1, 1), (1, 1)) + ((0, 0),) * (x.ndim - 2))
    return p[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]


def conv3(x, w, b):
    """y[i,j,co] = b[co] + sum_tap sum_ci w[tap,co,ci] x[i+dy, j+dx, ci],
    dy = tap // 3 - 1, dx = tap % 3 - 1 (zero padding)."""
    H, W, _ = x.shape
    y = np.broadcast_to(b, (H, W, b.size)).astype(np.float64).copy()
    for t in range(9):
        y += _shift(x, t // 3 - 1, t % 3 - 1) @ w[t].T
    return y


def conv3_dx(g, w):
    """Input gradient of conv3 for output gradient g."""
    dx = 0.0
    for t in range(9):
        dx = dx + _shift(g, -(t // 3 - 1), -(t % 3 - 1)) @ w[t]
    return dx


def conv3_dw(g, x):
    H, W, co = g.shape
    ci = x.shape[2]
    G = g.reshape(-1, co)
    return np.stack([G.T @ _shift(x, t // 3 - 1, t % 3 - 1).reshape(-1, ci)
                     for t in range(9)])


def up4(x, w, b):
    """y[4I+a, 4J+c, co] = b[co] + sum_ci w[4a+c, co, ci] x[I, J, ci]."""
    Hc, Wc, _ = x.shape
    co = w.shape[1]
    y = np.empty((4 * Hc, 4 * Wc, co))
    for ab in range(16):
        a, c = divmod(ab, 4)
        y[a::4, c::4] = x @ w[ab].T + b
    return y


def up4_dx(g, w):
    dx = 0.0
    for ab in range(16):
        a, c = divmod(ab, 4)
        dx = dx + g[a::4, c::4] @ w[ab]
    return dx


def up4_dw(g, x):
    ci = x.shape[2]
    X = x.reshape(-1, ci)
    out = []
    for ab in range(16):
        a, c = divmod(ab, 4)
        ga = g[a::4, c::4]
        out.append(ga.reshape(-1, ga.shape[2]).T @ X)
    return np.stack(out)


def pool4(x):
    H, W, C = x.shape
    return x.reshape(H // 4, 4, W // 4, 4, C).mean(axis=(1, 3))


def pool4_dx(g):
    return np.repeat(np.repeat(g, 4, axis=0), 4, axis=1) / 16.0


# ------------------------------------------------------------------- network
def wall_map(H, W, walls=None):
    """The second input channel: per cell the sum over the walls it touches
    of +1 (Dirichlet, type 1) or -1 (Neumann, type 2); walls = the types of
    the x-, x+, y-, y+ segments (rows i = 0 and H-1, columns j = 0 and W-1),
    None for no wall information (all zero)."""
    m = np.zeros((H, W))
    if walls is None:
        return m
    code = {0: 0.0, 1: 1.0, 2: -1.0}
    m[0, :] += code[int(walls[0])]
    m[H - 1, :] += code[int(walls[1])]
    m[:, 0] += code[int(walls[2])]
    m[:, W - 1] += code[int(walls[3])]
    return m


def forward(params, x, bf=False, walls=None):
    """Network output y (H, W) for input x (H, W) on a domain whose walls
    have the given types (wall_map); returns (y, cache)."""
    ws, bs = unpack(np.asarray(params, dtype=np.float64))
    R = bf16 if bf else (lambda a: a)
    wq = [R(w) if LAYERS[k][0] in ("conv", "up") else w for k, w in enumerate(ws)]
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    H, W = x.shape
    x0 = np.asarray(x, dtype=np.float64).reshape(H, W, 1)
    if bf:
        x0 = x0.astype(np.float32).astype(np.float64)
    x0 = np.concatenate([x0, wall_map(H, W, walls)[:, :, None]], axis=2)
    c = {"x0": x0, "wq": wq}
    c["a0"] = R(relu(conv3(x0, ws[0], bs[0])))
    e0 = R(relu(conv3(c["a0"], wq[1], bs[1])))
    c["p0"] = R(pool4(e0))
    c["a1"] = R(relu(conv3(c["p0"], wq[2], bs[2])))
    e1 = R(relu(conv3(c["a1"], wq[3], bs[3])))
    c["p1"] = R(pool4(e1))
    c["a2"] = R(relu(conv3(c["p1"], wq[4], bs[4])))
    e2 = R(relu(conv3(c["a2"], wq[5], bs[5])))
    c["p2"] = R(pool4(e2))
    c["a3"] = R(relu(conv3(c["p2"], wq[6], bs[6])))
    c["b3"] = R(relu(conv3(c["a3"], wq[7], bs[7])))
    u2 = R(up4(c["b3"], wq[8], bs[8]))
    c["C2"] = np.concatenate([u2, e2], axis=2)
    c["d2a"] = R(relu(conv3(c["C2"], wq[9], bs[9])))
    c["d2"] = R(relu(conv3(c["d2a"], wq[10], bs[10])))
    u1 = R(up4(c["d2"], wq[11], bs[11]))
    c["C1"] = np.concatenate([u1, e1], axis=2)
    c["d1a"] = R(relu(conv3(c["C1"], wq[12], bs[12])))
    c["d1"] = R(relu(conv3(c["d1a"], wq[13], bs[13])))
    u0 = R(up4(c["d1"], wq[14], bs[14]))
    c["C0"] = np.concatenate([u0, e0], axis=2)
    c["d0a"] = R(relu(conv3(c["C0"], wq[15], bs[15])))
    c["d0"] = R(relu(conv3(c["d0a"], wq[16], bs[16])))
    y = c["d0"] @ ws[17][0, 0] + bs[17][0]
    c["ws"], c["bs"], c["bf"] = ws, bs, bf
    return y, c


def backward(c, dy):
    """Gradient of sum(dy * y) with respect to the flat parameters."""
    ws, wq, bf = c["ws"], c["wq"], c["bf"]
    R = bf16 if bf else (lambda a: a)
    gw = [None] * len(LAYERS)
    gb = [None] * len(LAYERS)
    dy = np.asarray(dy, dtype=np.float64)[:, :, None]

    def conv_bwd(k, x, y, gout, need_dx=True):
        g = gout * (y > 0)
        gq = R(g)
        gw[k] = conv3_dw(gq, x)
        gb[k] = g.sum(axis=(0, 1))
        return conv3_dx(gq, wq[k]) if need_dx else None

    def up_bwd(k, x, gout):
        gq = R(gout)
        gw[k] = up4_dw(gq, x)
        gb[k] = gout.sum(axis=(0, 1))
        return up4_dx(gq, wq[k])

    # output 1x1
    d0 = c["d0"]
    gw[17] = (dy * d0).sum(axis=(0, 1)).reshape(1, 1, 16)
    gb[17] = dy.sum(axis=(0, 1))
    g_d0 = dy * ws[17][0, 0]
    g_d0a = conv_bwd(16, c["d0a"], d0, g_d0)
    g_C0 = conv_bwd(15, c["C0"], c["d0a"], g_d0a)
    g_d1 = up_bwd(14, c["d1"], g_C0[:, :, :16])
    g_d1a = conv_bwd(13, c["d1a"], c["d1"], g_d1)
    g_C1 = conv_bwd(12, c["C1"], c["d1a"], g_d1a)
    g_d2 = up_bwd(11, c["d2"], g_C1[:, :, :32])
    g_d2a = conv_bwd(10, c["d2a"], c["d2"], g_d2)
    g_C2 = conv_bwd(9, c["C2"], c["d2a"], g_d2a)
    g_b3 = up_bwd(8, c["b3"], g_C2[:, :, :64])
    g_a3 = conv_bwd(7, c["a3"], c["b3"], g_b3)
    g_p2 = conv_bwd(6, c["p2"], c["a3"], g_a3)
    g_e2 = g_C2[:, :, 64:] + pool4_dx(g_p2)
    g_a2 = conv_bwd(5, c["a2"], c["C2"][:, :, 64:], g_e2)
    g_p1 = conv_bwd(4, c["p1"], c["a2"], g_a2)
    g_e1 = g_C1[:, :, 32:] + pool4_dx(g_p1)
    g_a1 = conv_bwd(3, c["a1"], c["C1"][:, :, 32:], g_e1)
    g_p0 = conv_bwd(2, c["p0"], c["a1"], g_a1)
    g_e0 = g_C0[:, :, 16:] + pool4_dx(g_p0)
    g_a0 = conv_bwd(1, c["a0"], c["C0"][:, :, 16:], g_e0)
    # first layer, FP32 weights on the FP32 input, no rounding
    g = g_a0 * (c["a0"] > 0)
    gw[0] = conv3_dw(g, c["x0"])
    gb[0] = g.sum(axis=(0, 1))
    return pack(gw, gb)


# ------------------------------------------------------------- init and Adam
class Xoshiro256ss:
    """xoshiro256** (Blackman & Vigna, 2018) seeded by SplitMix64."""

    M = (1 << 64) - 1

    def __init__(self, seed):
        x = seed & self.M
        self.s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & self.M
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
            self.s.append(z ^ (z >> 31))

    def next(self):
        s, M = self.s, self.M
        rotl = lambda v, k: ((v << k) | (v >> (64 - k))) & M  # noqa: E731
        r = (rotl((s[1] * 5) & M, 7) * 9) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return r

    def uniform(self):
        return (self.next() >> 11) * 2.0 ** -53


def kaiming_init(seed, layers=None):
    """Kaiming-normal weights (fan-in, gain sqrt 2; 1 for the output layer)
    by Box-Muller on xoshiro256** uniforms, zero biases; float32 vector.
    ``layers`` limits the draw to the first layers (the rest stay zero)."""
    import math
    lay, n = layout()
    out = np.zeros(n, dtype=np.float32)
    g = Xoshiro256ss(seed)
    nl = len(LAYERS) if layers is None else layers
    for k in range(nl):
        kind, ci, co = LAYERS[k]
        fan = ci if kind == "up" else ntaps(kind) * ci
        sd = (1.0 if kind == "out" else math.sqrt(2.0)) / math.sqrt(fan)
        wo, _, shape = lay[k]
        for i in range(int(np.prod(shape))):
            u1 = 1.0 - g.uniform()
            u2 = g.uniform()
            out[wo + i] = sd * math.sqrt(-2.0 * math.log(u1)) * \
                math.cos(6.283185307179586 * u2)
    return out


def adam(p, g, m1, m2, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba) with bias correction, in float32."""
    f = np.float32
    p, g, m1, m2 = (np.asarray(a, dtype=f) for a in (p, g, m1, m2))
    b1, b2 = f(b1), f(b2)
    # 1 - beta evaluated in float32, as the device does
    m1 = b1 * m1 + (f(1) - b1) * g
    m2 = b2 * m2 + (f(1) - b2) * g * g
    c1 = f(1) - b1 ** f(t)
    c2 = f(1) - b2 ** f(t)
    p = p - f(lr) * (m1 / c1) / (np.sqrt(m2 / c2) + f(eps))
    return p, m1, m2


# ------------------------------------------------------------ training loss
def energy_loss(params, A, r, e, bf=False):
    """L = (s y - e)' A (s y - e) / e' A e, y = net(r / s), s = rms(r), and
    dL/dy.  A is the oracle's CSR matrix over the interior unknowns in the
    x-fastest order; r and e are (H, W) arrays with W = ny contiguous."""
    H, W = r.shape
    s = float(np.sqrt(np.mean(np.asarray(r, dtype=np.float64) ** 2))) or 1.0
    y, c = forward(params, r / s, bf=bf)
    d = s * y - e
    # interior arrays are (nx, ny) with j contiguous; the CSR is i-fastest
    dx = d.T.ravel()
    Ad = A.mult(dx).reshape(W, H).T
    ex = np.asarray(e, dtype=np.float64).T.ravel()
    E = float(ex @ A.mult(ex))
    L = float(dx @ A.mult(dx)) / E
    dLdy = 2.0 * s * Ad / E
    return L, dLdy, c
