"""The fused level-0 kernels of an applied U-Net (first layer + convolution
+ pooling in one pass; up-convolution + both decoder convolutions + output
layer in another) against the layer-per-kernel path they replace (grids of
at least 128^2 cells, where that path takes the tiled convolutions): the same
products in the same order, so the applied network's output, the level-0
skip channels and the pooled level-1 input agree bit for bit, and a
flexible-CG solve preconditioned by either takes the same iterates."""
import numpy as np
import pytest

import carpios_amd as ca
import oracle as O
import oracle.unet as OU
from test_gpu_unet import random_params

pytestmark = pytest.mark.gpu


def problem(ctx, nx, ny, types, seed):
    og = O.StructGrid(nx, ny, 1)
    rng = np.random.default_rng(seed)
    src = og.field()
    og.interior(src)[:] = rng.standard_normal((nx, ny))
    P = ca.Poisson(ctx, ca.uniform_grid_desc(nx, ny, 1, types, O.make_bc(og, types)._keep, 1.0))
    P.set_source(src)
    return og, P


def warm_start(P, og, net, fused):
    # the warm start phi = s net(b / s) is one application of the network;
    # the fused encoder stores the skip channels densely in a0's buffer, the
    # layer-per-kernel path as the second half of C0
    net.set_fusion(fused)
    P.set_unet(net, ca.UNET_WARM)
    P.fill_phi(0.0)
    P.solve("JPCG", max_iter=0, tol=1e-300, norm=ca.NORM_TRUE)
    phi = og.interior(P.get_phi()).copy()
    skip = (net.tensor("a0") if fused else net.tensor("C0")[:, :, 16:]).copy()
    p0 = net.tensor("p0").copy()
    d1 = net.tensor("d1").copy()
    P.set_unet(None, 0)
    return phi, skip, p0, d1


@pytest.mark.parametrize("nx,ny,types,seed", [
    (256, 256, [1, 2, 2, 1], 3),   # mixed walls: every edge tile takes the wall channel
    (128, 320, [1, 1, 1, 1], 4),
    (256, 64, [2, 2, 2, 2], 5),    # all-Neumann; the smallest level 0 that fuses
    (1024, 1024, [1, 1, 1, 2], 6),
])
def test_fused_application_is_bitwise_the_layer_per_kernel_one(ctx, nx, ny, types, seed):
    print(f"fused level 0: {nx}x{ny} walls {types}", flush=True)
    og, P = problem(ctx, nx, ny, types, seed)
    net = ca.UNet(ctx, seed=1)
    net.set_params(random_params(seed + 40).astype(np.float32))
    ref = warm_start(P, og, net, False)
    for fam in (8, 9):  # the fused encoder and decoder launches
        ctx.timer_enable(fam)
        ctx.timer_read(fam)
    got = warm_start(P, og, net, True)
    launches = [ctx.timer_read(fam)[0] for fam in (8, 9)]
    for fam in (8, 9):
        ctx.timer_enable(fam, False)
    assert launches == [1, 1], launches
    for name, a, b in zip(("phi", "skip", "p0", "d1"), got, ref):
        nd = int(np.count_nonzero(a != b))
        print(f"  {name}: {nd} of {a.size} differ, max |a| {np.abs(b).max():.3e}", flush=True)
        assert np.array_equal(a, b), name
    assert np.abs(ref[0]).max() > 0 and np.abs(ref[1]).max() > 0
    net.close()
    P.close()


def test_fused_preconditioner_takes_the_same_iterates(ctx):
    n = 256
    og, P = problem(ctx, n, n, [1, 1, 2, 1], 8)
    net = ca.UNet(ctx, seed=1)
    net.set_params(random_params(17).astype(np.float32))
    out = []
    for fused in (False, True):
        net.set_fusion(fused)
        P.set_unet(net, ca.UNET_PRECOND)
        P.fill_phi(0.0)
        code, it, res, _ = P.solve("FCG_UNET", max_iter=5000, tol=1e-10, norm=ca.NORM_TRUE)
        print(f"  FCG fused={fused}: code {code} iterations {it} res {res:.3e}", flush=True)
        out.append((code, it, res, og.interior(P.get_phi()).copy()))
        P.set_unet(None, 0)
    assert out[0][:3] == out[1][:3]
    assert np.array_equal(out[0][3], out[1][3])
    assert out[1][0] == 0 and out[1][2] <= 1e-10
    net.close()
    P.close()


def rne_bf16(x):
    """FP32 -> BF16 -> FP32, round to nearest even, on the host (finite
    non-negative inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def test_device_bf16_rounding_is_round_to_nearest_even(ctx):
    # Every U-Net kernel rounds its FP32 results to BF16 with one device
    # conversion.  With the first layer's centre-tap input weight 1 and all
    # other parameters 0 the layer stores bf16(relu(x)) exactly
    # (fma(1, x, 0) = x and the zero products change nothing), so the stored
    # activation checks that conversion against round-to-nearest-even on the
    # host: ties to even and odd, the values either side of a tie,
    # subnormals, the largest finite value (rounds to infinity) and random
    # bit patterns over every exponent.
    H = W = 128
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 0x7F800000, size=H * W, dtype=np.uint64).astype(np.uint32)
    k = H * W // 8
    bits[:k] = (bits[:k] & 0xFFFF0000) | 0x8000          # exact ties
    bits[k:2 * k] = (bits[k:2 * k] & 0xFFFF0000) | 0x7FFF
    bits[2 * k:3 * k] = (bits[2 * k:3 * k] & 0xFFFF0000) | 0x8001
    bits[3 * k:4 * k] &= 0x007FFFFF                      # subnormals
    bits[4 * k:4 * k + 8] = [0, 1, 0x8000, 0x18000, 0x007FFFFF, 0x7F7FFFFF, 0x7F7F7FFF,
                             0x7F7F8000]
    x = bits.view(np.float32).reshape(H, W)
    lay, n = OU.layout()
    p = np.zeros(n, dtype=np.float32)
    wo = lay[0][0]
    for co in range(16):
        p[wo + (4 * 16 + co) * 2] = 1.0   # tap 4 (centre), input channel 0
    net = ca.UNet(ctx, seed=1)
    net.set_params(p)
    net.set_walls(None)
    net.forward(x)
    a0 = net.tensor("a0")
    want = rne_bf16(x)
    for co in range(16):
        got = a0[:, :, co]
        bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        assert bad.size == 0, (co, [hex(int(b)) for b in bits[bad[:5]]])
    assert np.isinf(want.ravel()[4 * k + 5]) and np.count_nonzero(want.ravel()[:k] == 0) < k
    net.close()


def test_replayed_iteration_graphs_take_the_same_iterates(ctx):
    # with use_graphs = 2 the flexible-CG iterations after the first replay
    # one captured graph per direction parity; otherwise the same kernels
    # launch directly
    n = 256
    og, P = problem(ctx, n, n, [2, 2, 1, 2], 9)
    net = ca.UNet(ctx, seed=1)
    net.set_params(random_params(23).astype(np.float32))
    out = []
    for graphs in (True, 2, 2):  # the second graphed solve replays cached graphs
        P.set_batch(16, use_graphs=graphs)
        P.set_unet(net, ca.UNET_PRECOND)
        P.fill_phi(0.0)
        code, it, res, _ = P.solve("FCG_UNET", max_iter=5000, tol=1e-10, norm=ca.NORM_TRUE)
        print(f"  FCG graphs={graphs}: code {code} iterations {it} res {res:.3e}", flush=True)
        out.append((code, it, res, og.interior(P.get_phi()).copy()))
        P.set_unet(None, 0)
    for o in out[1:]:
        assert o[:3] == out[0][:3]
        assert np.array_equal(o[3], out[0][3])
    assert out[0][0] == 0 and out[0][1] > 2
    net.close()
    P.close()
