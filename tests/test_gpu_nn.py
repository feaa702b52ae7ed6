"""The surrogate's HIP kernels (libswe_nn.so, through its C-ABI) against
references of the same ops computed on the CPU in float64 (convolutions and
their gradients by PyTorch autograd, the rest in closed form): 3x3 periodic
convolution forward on BF16 MFMA, its data and weight gradients, pooling,
upsampling, the loss, Adam and the fixed-point mass projection.  Inputs are
bf16 values, exact in float64, so the tolerances only cover the kernels'
float32 accumulation (and bf16 rounding of bf16 outputs)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
F = torch.nn.functional

pytestmark = pytest.mark.gpu

from simplewinderosion_amd import nn  # noqa: E402

DEV = "cuda"


def ref_conv(x16, w16, bias=None):
    """float64 CPU periodic 3x3 convolution of NHWC x with [Cout,9,Cin] w,
    returned on the device as float32 for comparison."""
    x = x16.cpu().double().permute(0, 3, 1, 2)
    Cout, _, Cin = w16.shape
    w = w16.cpu().double().view(Cout, 3, 3, Cin).permute(0, 3, 1, 2).contiguous()
    b = None if bias is None else bias.cpu().double()
    y = F.conv2d(F.pad(x, (1, 1, 1, 1), mode="circular"), w, b)
    return y.permute(0, 2, 3, 1).contiguous().float().to(DEV)


def rnd(shape, dtype=torch.bfloat16, seed=0, scale=1.0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return (torch.randn(shape, generator=g) * scale).to(DEV).to(dtype).contiguous()


@pytest.mark.parametrize("N,H,W,Cin,Cout,relu", [
    (2, 12, 80, 16, 48, True),    # partial 64-pixel segment, two accumulators per wave
    (1, 8, 64, 8, 16, False),     # one accumulator, K = 72 (a partial 32-step)
    (1, 6, 128, 32, 80, True),    # four accumulators, a partial output-channel block
    (3, 4, 16, 64, 128, False),   # narrow rows
    (1, 128, 520, 32, 136, True),  # 128-channel workgroups (many pixel blocks), a partial channel block
    (2, 9, 13, 40, 24, False),     # pixel blocks across rows and images, a partial 32-channel step
    # row strips (W % 32 == 0, Cin <= 64): weights in LDS, input rows streamed through a ring
    (2, 9, 64, 32, 32, True),      # two strips a row, several chunks a strip
    (1, 7, 32, 40, 72, False),     # 64-channel ring, partial output-channel block
    (1, 1, 32, 16, 8, False),      # one row
    (1, 2, 64, 64, 24, True),      # a chunk of two rows
    (1, 11, 96, 24, 40, True),     # three strips a row, odd rows per chunk
    (8, 64, 64, 64, 64, True),     # a training layer's shape
])
def test_conv3x3_forward(N, H, W, Cin, Cout, relu):
    print("case conv", N, H, W, Cin, Cout, relu, flush=True)
    x = rnd((N, H, W, Cin), seed=1)
    w = rnd((Cout, 9, Cin), seed=2, scale=0.2)
    b = rnd((Cout,), torch.float32, seed=3)
    y32 = torch.empty((N, H, W, Cout), device=DEV)
    y16 = torch.empty((N, H, W, Cout), device=DEV, dtype=torch.bfloat16)
    nn.conv3x3(x, w, b, relu=relu, out16=y16, out32=y32)
    torch.cuda.synchronize()
    ref = ref_conv(x, w, b)
    if relu:
        ref = ref.clamp_min(0)
    assert torch.allclose(y32, ref, atol=1e-4 * (9 * Cin) ** 0.5, rtol=1e-4)
    assert torch.equal(y16, y32.to(torch.bfloat16))


@pytest.mark.parametrize("Cin,Cout,N,H,W", [(16, 32, 2, 10, 70), (24, 384, 2, 10, 70), (64, 72, 2, 60, 70),
                                             (128, 8, 1, 33, 40),
                                             (48, 40, 3, 5, 12),      # a 32-pixel step spans rows and images
                                             (64, 64, 2, 16, 16)])    # the 16^2 bottom level's shape
def test_conv3x3_gradients(Cin, Cout, N, H, W):
    """Data gradient (forward on flipped weights) and the weight and bias
    gradients, over several pixel chunks with a partial last one and
    partial 64-channel blocks."""
    print("case conv grads", Cin, Cout, N, H, W, flush=True)
    x = rnd((N, H, W, Cin), seed=4)
    w = rnd((Cout, 9, Cin), seed=5, scale=0.2)
    dy = rnd((N, H, W, Cout), seed=6)
    # data gradient = forward convolution of dy with the flipped weights
    dx = torch.empty((N, H, W, Cin), device=DEV)
    nn.conv3x3(dy, nn.flip(w), None, out32=dx)
    dw = torch.empty((Cout, 9, Cin), device=DEV)
    db = torch.empty((Cout,), device=DEV)
    nn.wgrad(x, dy, dw, db)
    torch.cuda.synchronize()
    xr = x.cpu().double().requires_grad_(True)
    wr = w.cpu().double().requires_grad_(True)
    br = torch.zeros(Cout, dtype=torch.float64, requires_grad=True)
    y = ref_conv_grad(xr, wr, br)
    y.backward(dy.cpu().double())
    assert torch.allclose(dx.cpu().double(), xr.grad, atol=1e-3, rtol=1e-4)
    assert torch.allclose(dw.cpu().double(), wr.grad, atol=2e-3, rtol=1e-4)
    assert torch.allclose(db.cpu().double(), br.grad, atol=2e-3, rtol=1e-4)


@pytest.mark.parametrize("N,H,W,Cin,Cout", [
    (1, 3, 32, 8, 16),      # 32 x 32 blocks, one chunk, channels padded to the block
    (2, 9, 64, 32, 32),     # two strips a row, several chunks a strip, a short last chunk
    (1, 20, 32, 64, 72),    # 64 x 64 blocks, a partial output-channel block
    (1, 1, 32, 16, 8),      # one row: the ring's three rows are the same row
    (1, 17, 96, 40, 64),    # 64 (co) x 32 (ci) blocks, three strips a row
    (2, 2, 32, 72, 24),     # 32 (co) x 64 (ci) blocks, a partial input-channel block
    (8, 64, 64, 128, 128),  # a training layer's shape (four 64 x 64 blocks)
])
def test_wgrad_row_strips(N, H, W, Cin, Cout):
    """The weight and bias gradients of maps whose rows are whole 32-pixel
    strips (all nine taps of a channel block per workgroup, rows streamed
    through an LDS ring): against float64 autograd, and the same bits on a
    second call."""
    print("case wgrad rows", N, H, W, Cin, Cout, flush=True)
    x = rnd((N, H, W, Cin), seed=7)
    dy = rnd((N, H, W, Cout), seed=8)
    w = rnd((Cout, 9, Cin), seed=9, scale=0.2)
    dw = torch.empty((Cout, 9, Cin), device=DEV)
    db = torch.empty((Cout,), device=DEV)
    work = nn.wgrad(x, dy, dw, db)
    dw2 = torch.empty_like(dw)
    db2 = torch.empty_like(db)
    nn.wgrad(x, dy, dw2, db2, work)
    torch.cuda.synchronize()
    assert torch.equal(dw, dw2) and torch.equal(db, db2)
    xr = x.cpu().double().requires_grad_(True)
    wr = w.cpu().double().requires_grad_(True)
    br = torch.zeros(Cout, dtype=torch.float64, requires_grad=True)
    ref_conv_grad(xr, wr, br).backward(dy.cpu().double())
    tol = 1e-5 * (N * H * W) ** 0.5 + 1e-3
    assert torch.allclose(dw.cpu().double(), wr.grad, atol=tol, rtol=1e-4)
    assert torch.allclose(db.cpu().double(), br.grad, atol=tol, rtol=1e-4)


@pytest.mark.parametrize("Cin,Cout,N,H,W", [(16, 32, 2, 10, 70), (64, 8, 1, 33, 40), (32, 64, 2, 16, 32)])
def test_masked_conv_matches_unfused(Cin, Cout, N, H, W):
    """The data gradient through a ReLU in one pass: the same bits as the
    float convolution followed by the ReLU gradient's conversion."""
    print("case masked conv", Cin, Cout, N, H, W, flush=True)
    x = rnd((N, H, W, Cin), seed=10)
    w = rnd((Cout, 9, Cin), seed=11, scale=0.2)
    mask = rnd((N, H, W, Cout), seed=12)
    y32 = torch.empty((N, H, W, Cout), device=DEV)
    nn.conv3x3(x, w, None, out32=y32)
    for m in (mask, None):
        want = nn.relu_grad(y32, m, torch.empty((N, H, W, Cout), device=DEV, dtype=torch.bfloat16))
        got = nn.conv3x3_masked(x, w, m, torch.empty((N, H, W, Cout), device=DEV, dtype=torch.bfloat16))
        torch.cuda.synchronize()
        assert torch.equal(got, want)


@pytest.mark.parametrize("N,H,W,Cin,Cout", [(8, 16, 16, 256, 192), (8, 32, 32, 192, 192), (1, 4, 4, 96, 40)])
def test_conv3x3_split_k(N, H, W, Cin, Cout):
    """Layers with few pixel blocks split their reduction over several
    workgroups (partials added in a fixed order): against float64, the same
    bits on a second call, the masked data-gradient form equal to the float
    output passed through the ReLU gradient, and unsplit without workspace."""
    print("case split-k", N, H, W, Cin, Cout, flush=True)
    need = nn.conv_work(N, H, W, Cin, Cout)
    assert need > 0
    work = torch.empty(need, device=DEV)
    x = rnd((N, H, W, Cin), seed=21)
    w = rnd((Cout, 9, Cin), seed=22, scale=0.1)
    b = rnd((Cout,), torch.float32, seed=23)
    mask = rnd((N, H, W, Cout), seed=24)
    y32 = torch.empty((N, H, W, Cout), device=DEV)
    y32b = torch.empty_like(y32)
    nn.conv3x3(x, w, b, out32=y32, work=work)
    nn.conv3x3(x, w, b, out32=y32b, work=work)
    plain = torch.empty_like(y32)
    nn.conv3x3(x, w, b, out32=plain)
    g16 = nn.conv3x3_masked(x, w, mask, torch.empty((N, H, W, Cout), device=DEV, dtype=torch.bfloat16), bias=b,
                            work=work)
    torch.cuda.synchronize()
    assert torch.equal(y32, y32b)
    ref = ref_conv(x, w, b)
    assert torch.allclose(y32, ref, atol=1e-4 * (9 * Cin) ** 0.5, rtol=1e-4)
    assert torch.allclose(y32, plain, atol=1e-4 * (9 * Cin) ** 0.5, rtol=1e-4)
    want = nn.relu_grad(y32, mask, torch.empty((N, H, W, Cout), device=DEV, dtype=torch.bfloat16))
    torch.cuda.synchronize()
    assert torch.equal(g16, want)


def test_conv3x3_refuses_tensors_of_2_gib():
    """The convolution addresses x and w with 32-bit byte offsets: shapes
    whose input or weights reach 2 GiB are refused (-1) before any launch."""
    x = rnd((1, 4, 4, 32), seed=25)
    w = rnd((32, 9, 32), seed=26, scale=0.1)
    y = torch.empty((1, 4, 4, 32), device=DEV, dtype=torch.bfloat16)
    lib = nn.load()
    # 2^14 x 2^8 x 2^8 pixels x 32 channels x 2 bytes = 2^36 bytes of input
    assert lib.swe_nn_conv3x3(x.data_ptr(), 1 << 14, 256, 256, 32, w.data_ptr(), None, 32, 0, y.data_ptr(), None,
                              None) == -1
    # 2^18 output channels x 9 x 512 x 2 bytes of weights > 2 GiB
    assert lib.swe_nn_conv3x3(x.data_ptr(), 1, 4, 4, 512, w.data_ptr(), None, 1 << 18, 0, y.data_ptr(), None,
                              None) == -1
    assert lib.swe_nn_conv3x3(x.data_ptr(), 1, 4, 4, 32, w.data_ptr(), None, 32, 0, y.data_ptr(), None, None) == 0
    torch.cuda.synchronize()
    assert torch.allclose(y.float(), ref_conv(x, w), atol=1e-2, rtol=1e-2)


def test_fused_skip_passes_match_unfused():
    """up2 + add, pool2_grad + add + ReLU mask and up2_grad + ReLU mask in
    one pass each: the same bits as the separate kernels."""
    N, h, w, C = 2, 6, 10, 24
    bf = torch.bfloat16

    def empty(*shape):
        return torch.empty(shape, device=DEV, dtype=bf)

    lo = rnd((N, h, w, C), seed=13)
    hi = rnd((N, 2 * h, 2 * w, C), seed=14)
    hi2 = rnd((N, 2 * h, 2 * w, C), seed=15)
    mlo = rnd((N, h, w, C), seed=16)
    want = nn.add(nn.up2(lo, empty(N, 2 * h, 2 * w, C)), hi, empty(N, 2 * h, 2 * w, C))
    assert torch.equal(nn.up2_add(lo, hi, empty(N, 2 * h, 2 * w, C)), want)
    for m in (hi2, None):
        want = nn.add(nn.pool2_grad(lo, empty(N, 2 * h, 2 * w, C)), hi, empty(N, 2 * h, 2 * w, C))
        if m is not None:
            nn.relu_mask(want, m)
        assert torch.equal(nn.pool2_grad_add(lo, hi, m, empty(N, 2 * h, 2 * w, C)), want)
    for m in (mlo, None):
        want = nn.up2_grad(hi, empty(N, h, w, C))
        if m is not None:
            nn.relu_mask(want, m)
        assert torch.equal(nn.up2_grad_mask(hi, m, empty(N, h, w, C)), want)
    torch.cuda.synchronize()


def ref_conv_grad(x, w, b):
    Cout, _, Cin = w.shape
    wc = w.view(Cout, 3, 3, Cin).permute(0, 3, 1, 2)
    y = F.conv2d(F.pad(x.permute(0, 3, 1, 2), (1, 1, 1, 1), mode="circular"), wc, b)
    return y.permute(0, 2, 3, 1)


def test_pool_upsample_add_relu_grad():
    print("case pool/up", flush=True)
    x = rnd((2, 8, 12, 16), seed=7)
    y = torch.empty((2, 4, 6, 16), device=DEV, dtype=torch.bfloat16)
    nn.pool2(x, y)
    ref = F.avg_pool2d(x.float().permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
    assert torch.equal(y, ref.to(torch.bfloat16))
    dx = torch.empty_like(x)
    nn.pool2_grad(y, dx)
    ref = (0.25 * y.float()).repeat_interleave(2, 1).repeat_interleave(2, 2)
    assert torch.equal(dx, ref.to(torch.bfloat16))
    u = torch.empty((2, 16, 24, 16), device=DEV, dtype=torch.bfloat16)
    nn.up2(x, u)
    assert torch.equal(u, x.repeat_interleave(2, 1).repeat_interleave(2, 2))
    du = torch.empty_like(x)
    nn.up2_grad(u, du)
    assert torch.equal(du, (4 * x.float()).to(torch.bfloat16))
    s = torch.empty_like(x)
    nn.add(x, x, s)
    assert torch.equal(s, (2 * x.float()).to(torch.bfloat16))
    g32 = rnd((2, 8, 12, 16), torch.float32, seed=8)
    g16 = torch.empty_like(x)
    nn.relu_grad(g32, x, g16)
    assert torch.equal(g16, torch.where(x.float() > 0, g32, torch.zeros_like(g32)).to(torch.bfloat16))
    m16 = g32.to(torch.bfloat16)
    ref = torch.where(x.float() > 0, m16.float(), torch.zeros_like(g32)).to(torch.bfloat16)
    nn.relu_mask(m16, x)
    assert torch.equal(m16, ref)


def test_mse_and_adam():
    print("case mse/adam", flush=True)
    pred = rnd((500, 16), torch.float32, seed=9)
    target = rnd((500, 2), torch.float32, seed=10)
    grad = torch.empty_like(pred)
    grad16 = torch.empty(pred.shape, device=DEV, dtype=torch.bfloat16)
    part = torch.empty(((500 * 16 + 255) // 256,), device=DEV)
    part2 = torch.empty_like(part)
    nn.mse(pred, target, grad, part)
    nn.mse(pred, target, None, part2, grad16=grad16)
    torch.cuda.synchronize()
    d = pred[:, :2] - target
    assert abs(float(part.sum()) - float((d * d).mean())) < 1e-5
    # the block sums against a serial binary tree over each block's 256 terms
    ld = torch.zeros(part.numel() * 256, dtype=torch.float64)
    pad = torch.zeros_like(pred)
    pad[:, :2] = (d * d) * (1.0 / d.numel())
    ld[:pred.numel()] = pad.reshape(-1).cpu().double()
    tree = ld.float().view(-1, 256)
    k = 128
    while k:
        tree = tree[:, :k] + tree[:, k:2 * k]
        k //= 2
    assert torch.equal(part.cpu(), tree[:, 0])
    ref = torch.zeros_like(pred)
    ref[:, :2] = 2 * d / d.numel()
    assert torch.allclose(grad, ref, atol=1e-8, rtol=1e-5)
    assert torch.equal(grad16, grad.to(torch.bfloat16)) and torch.equal(part2, part)
    w = rnd((1000,), torch.float32, seed=11)
    g = rnd((1000,), torch.float32, seed=12)
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    w16 = torch.empty(1000, device=DEV, dtype=torch.bfloat16)
    w0 = w.clone()
    nn.adam(w, w16, g, m, v, step=1, lr=1e-2)
    torch.cuda.synchronize()
    mr = 0.1 * g
    vr = 0.001 * g * g
    wr = w0 - 1e-2 * (mr / 0.1) / (torch.sqrt(vr / 0.001) + 1e-8)
    assert torch.allclose(w, wr, atol=1e-6, rtol=1e-5)
    assert torch.equal(w16, w.to(torch.bfloat16))


def test_adam_multi_matches_per_tensor_adam_and_flip():
    """One launch over several tensors: the same bits as swe_nn_adam per
    tensor, and the flipped copies as swe_nn_flip of the new bf16 weights."""
    shapes = [(24, 9, 16), (40,), (8, 9, 32), (3,)]
    ts = []
    for k, sh in enumerate(shapes):
        w = rnd(sh, torch.float32, seed=20 + k)
        g = rnd(sh, torch.float32, seed=30 + k, scale=0.1)
        m = rnd(sh, torch.float32, seed=40 + k, scale=0.01)
        v = rnd(sh, torch.float32, seed=50 + k, scale=0.01).abs()
        ts.append((w, g, m, v))
    ref = [tuple(t.clone() for t in x) for x in ts]
    segs = []
    outs = []
    for (w, g, m, v), sh in zip(ts, shapes):
        conv = len(sh) == 3
        w16 = torch.empty(sh, device=DEV, dtype=torch.bfloat16) if conv else None
        wt16 = torch.empty((sh[2], 9, sh[0]), device=DEV, dtype=torch.bfloat16) if conv else None
        outs.append((w16, wt16))
        segs.append((w, w16, wt16, g, m, v, sh[2] if conv else 0, sh[0] if conv else 0))
    table = nn.adam_table(segs, DEV)
    for step in (1, 2):
        nn.adam_multi(table, step, lr=1e-2)
        for (w, g, m, v), sh in zip(ref, shapes):
            w16 = torch.empty(sh, device=DEV, dtype=torch.bfloat16) if len(sh) == 3 else None
            nn.adam(w, w16, g, m, v, step=step, lr=1e-2)
    torch.cuda.synchronize()
    for (w, g, m, v), (rw, rg, rm, rv), (w16, wt16) in zip(ts, ref, outs):
        assert torch.equal(w, rw) and torch.equal(m, rm) and torch.equal(v, rv)
        if w16 is not None:
            assert torch.equal(w16, w.to(torch.bfloat16))
            assert torch.equal(wt16, nn.flip(w16))


@pytest.mark.parametrize("target", [0, 12345, -987654321])
def test_mass_projection_exact(target):
    print("case projection", target, flush=True)
    n = 10007
    d = rnd((n,), torch.float32, seed=13, scale=1e-3)
    master = torch.zeros(n, device=DEV, dtype=torch.int64)
    work = torch.empty(256, device=DEV, dtype=torch.int64)
    nn.project_mass(d, 1.0, target, master, work)
    torch.cuda.synchronize()
    got = master.cpu().numpy()
    assert sum(int(v) for v in got) == target
    q = np.rint(d.cpu().numpy().astype(np.float32) * np.float32(2.0 ** 32)).astype(np.int64)
    assert np.abs(got - q).max() <= abs(target - int(q.sum())) // n + 1


def test_features_match_float64():
    print("case features", flush=True)
    g = torch.Generator(device="cpu").manual_seed(14)
    h = torch.rand((2, 16, 24), generator=g, dtype=torch.float64)
    s = torch.rand((2, 16, 24), generator=g, dtype=torch.float64) * 0.1
    az = torch.tensor([0.3, 2.5], dtype=torch.float64)
    mean = h.mean(dim=(1, 2))
    x = nn.features(h.float().to(DEV), s.float().to(DEV), mean.float().to(DEV), az.float().to(DEV), 0.2, 0.05, 0.01,
                    0.02).cpu().double()
    r = lambda a, k, d: torch.roll(a, k, dims=d)
    ref = torch.zeros((2, 16, 24, 8), dtype=torch.float64)
    ref[..., 0] = (h - mean[:, None, None]) / 0.2
    ref[..., 1] = s / 0.05
    ref[..., 2] = torch.cos(az)[:, None, None]
    ref[..., 3] = torch.sin(az)[:, None, None]
    ref[..., 4] = 0.5 * (r(h, -1, 1) - r(h, 1, 1)) / 0.01
    ref[..., 5] = 0.5 * (r(h, -1, 2) - r(h, 1, 2)) / 0.01
    ref[..., 6] = (r(h, -1, 1) + r(h, 1, 1) + r(h, -1, 2) + r(h, 1, 2) - 4 * h) / 0.02
    # bf16 keeps 8 significant bits; inputs were rounded to float32 first
    assert torch.allclose(x, ref, rtol=1e-2, atol=1e-2)
