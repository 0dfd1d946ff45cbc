"""Pins of the U-Net oracle (oracle/unet.py) on the CPU: primitives against
textbook routines, the backward pass against finite differences, BF16
rounding against torch, the generator against the library's own host-side
xoshiro256**, Adam and the loss against closed forms."""
import numpy as np
import pytest
import scipy.signal

import oracle.unet as U


def test_conv3_matches_correlate2d():
    rng = np.random.default_rng(0)
    H, W, ci, co = 9, 13, 3, 2
    x = rng.standard_normal((H, W, ci))
    w = rng.standard_normal((9, co, ci))
    b = rng.standard_normal(co)
    y = U.conv3(x, w, b)
    for o in range(co):
        ref = np.full((H, W), b[o])
        for c in range(ci):
            K = w[:, o, c].reshape(3, 3)  # K[dy+1, dx+1]
            ref += scipy.signal.correlate2d(x[:, :, c], K, mode="same")
        assert np.allclose(y[:, :, o], ref, atol=1e-12)


def test_up4_is_a_kronecker_product_and_pool4_a_block_mean():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 5, 1))
    w = rng.standard_normal((16, 1, 1))
    b = np.array([0.25])
    y = U.up4(x, w, b)
    assert np.allclose(y[:, :, 0], np.kron(x[:, :, 0], w[:, 0, 0].reshape(4, 4)) + 0.25)
    z = rng.standard_normal((8, 12, 2))
    p = U.pool4(z)
    for I in range(2):
        for J in range(3):
            assert np.isclose(p[I, J, 1], z[4 * I:4 * I + 4, 4 * J:4 * J + 4, 1].mean())


def test_adjoints_of_the_primitives():
    # <conv3(x), g> = <x, conv3_dx(g)>, <conv3(x), g> = <w, conv3_dw(g, x)>
    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 12, 3))
    w = rng.standard_normal((9, 4, 3))
    g = rng.standard_normal((8, 12, 4))
    z = np.zeros(4)
    lhs = np.sum(U.conv3(x, w, z) * g)
    assert np.isclose(lhs, np.sum(x * U.conv3_dx(g, w)))
    assert np.isclose(lhs, np.sum(w * U.conv3_dw(g, x)))
    xc = rng.standard_normal((2, 3, 3))
    wu = rng.standard_normal((16, 4, 3))
    gu = rng.standard_normal((8, 12, 4))
    lhs = np.sum(U.up4(xc, wu, z) * gu)
    assert np.isclose(lhs, np.sum(xc * U.up4_dx(gu, wu)))
    assert np.isclose(lhs, np.sum(wu * U.up4_dw(gu, xc)))


def _random_params(seed):
    rng = np.random.default_rng(seed)
    lay, n = U.layout()
    p = np.zeros(n)
    for (wo, bo, shape), (kind, ci, co) in zip(lay, U.LAYERS):
        fan = ci if kind == "up" else U.ntaps(kind) * ci
        p[wo:wo + int(np.prod(shape))] = rng.standard_normal(int(np.prod(shape))) * np.sqrt(2.0 / fan)
        p[bo:bo + co] = 0.05 * rng.standard_normal(co)
    return p


def test_parameter_count():
    assert U.layout()[1] == 610913  # the first layer reads 2 channels


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    H, W = 64, 128
    p = _random_params(4)
    x = rng.standard_normal((H, W))
    dy = rng.standard_normal((H, W))
    walls = (1, 2, 2, 1)  # the wall channel's weights get gradients too
    y, c = U.forward(p, x, walls=walls)
    g = U.backward(c, dy)
    lay, _ = U.layout()
    # The network is piecewise smooth (about 10^6 ReLUs at this size): a
    # step can cross a kink, so each directional derivative is accepted when
    # one of two small steps reproduces it.
    for k, (wo, bo, shape) in enumerate(lay):
        nw = int(np.prod(shape))
        co = U.LAYERS[k][2]
        for lo, hi in ((wo, wo + nw), (bo, bo + co)):
            v = np.zeros_like(p)
            v[lo:hi] = rng.standard_normal(hi - lo)
            an = g @ v
            errs = []
            for eps in (1e-7, 3e-8):
                fp = np.sum(dy * U.forward(p + eps * v, x, walls=walls)[0])
                fm = np.sum(dy * U.forward(p - eps * v, x, walls=walls)[0])
                errs.append(abs((fp - fm) / (2 * eps) - an))
            assert min(errs) <= 1e-6 * max(1.0, abs(an)), (k, lo, errs, an)
    # a plausible mistake fails: the gradient of a transposed first conv
    gbad = g.copy()
    wo, bo, shape = lay[1]
    gbad[wo:bo] = np.transpose(g[wo:bo].reshape(shape), (0, 2, 1)).ravel()
    v = np.zeros_like(p)
    v[wo:bo] = rng.standard_normal(bo - wo)
    eps = 1e-7
    fd = (np.sum(dy * U.forward(p + eps * v, x, walls=walls)[0]) -
          np.sum(dy * U.forward(p - eps * v, x, walls=walls)[0])) / (2 * eps)
    assert abs(fd - gbad @ v) > 1e-3 * abs(fd)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-30, 30, 5000),
                        [0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 65504.0,
                         3.0e38, -3.0e38, 1e-40]]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(U.bf16(a).astype(np.float32), ref)
    assert U.bf16(np.float32(1 + 2 ** -8)) == 1.0          # tie to even
    assert U.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6


def test_generator_matches_the_library_host_xoshiro():
    # cpgpu_blob_params draws with the same xoshiro256** / SplitMix64 /
    # Box-Muller as the U-Net initialisation; it runs on the host only
    import carpios_amd as ca
    out = ca.blob_params(seed=7, nblobs=5, wmin=0.01, wmax=0.1)
    g = U.Xoshiro256ss(7)
    import math
    for b in range(5):
        cx, cy = g.uniform(), g.uniform()
        w = 0.01 + (0.1 - 0.01) * g.uniform()
        u1, u2 = 1.0 - g.uniform(), g.uniform()
        a = math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586 * u2)
        assert np.array_equal(out[b], [cx, cy, w, a]), (b, out[b], (cx, cy, w, a))


def test_adam_constant_gradient_moves_lr_per_step():
    p = np.zeros(4, dtype=np.float32)
    g = np.array([3.0, -2.0, 0.5, 1e3], dtype=np.float32)
    m1 = np.zeros(4, dtype=np.float32)
    m2 = np.zeros(4, dtype=np.float32)
    for t in range(1, 6):
        p, m1, m2 = U.adam(p, g, m1, m2, t, lr=0.01)
    assert np.allclose(p, -0.05 * np.sign(g), rtol=1e-4)


def test_energy_loss_is_one_for_a_zero_network_and_zero_at_the_solution():
    import oracle as O
    n = 64
    og = O.StructGrid(n, n, 1)
    A = O.Matrix(og, O.make_bc(og, [1, 1, 1, 1]), 1.0)
    rng = np.random.default_rng(6)
    e = rng.standard_normal((n, n))
    r = A.mult(e.T.ravel()).reshape(n, n).T
    p = _random_params(7)
    lay, _ = U.layout()
    wo, bo, _ = lay[-1]
    p[wo:bo + 1] = 0.0  # output layer zero: y = 0
    L, dLdy, _ = U.energy_loss(p, A, r, e)
    assert L == pytest.approx(1.0, rel=1e-14)
    # dL/dy at y = 0 is -2 s r / e'Ae
    s = np.sqrt(np.mean(r ** 2))
    E = e.T.ravel() @ A.mult(e.T.ravel())
    assert np.allclose(dLdy, -2 * s * r / E, rtol=1e-12, atol=1e-300)
    # output bias b alone: y = b, L(b) = (s b 1 - e)'A(s b 1 - e)/E is a
    # quadratic in b whose derivative is sum(dL/dy)
    p[bo] = 0.3
    L1, dLdy1, _ = U.energy_loss(p, A, r, e)
    one = np.ones(n * n)
    ex = e.T.ravel()
    Lq = ((0.3 * s * one - ex) @ A.mult(0.3 * s * one - ex)) / E
    assert L1 == pytest.approx(Lq, rel=1e-12)
    h = 1e-6
    p[bo] = 0.3 + h
    Lp = U.energy_loss(p, A, r, e)[0]
    p[bo] = 0.3 - h
    Lm = U.energy_loss(p, A, r, e)[0]
    assert (Lp - Lm) / (2 * h) == pytest.approx(dLdy1.sum(), rel=1e-6)


def test_wall_map():
    m = U.wall_map(4, 5, (1, 2, 2, 1))
    # rows: x- Dirichlet (+1), x+ Neumann (-1); columns: y- Neumann, y+ Dirichlet
    assert m[0, 2] == 1 and m[3, 2] == -1 and m[2, 0] == -1 and m[2, 4] == 1
    assert m[0, 0] == 0 and m[0, 4] == 2 and m[3, 0] == -2 and m[3, 4] == 0
    assert (m[1:3, 1:4] == 0).all() and (U.wall_map(4, 5) == 0).all()
