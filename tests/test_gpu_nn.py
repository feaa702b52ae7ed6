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
                                             (128, 8, 1, 33, 40)])
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
    part = torch.empty(((500 * 16 + 255) // 256,), device=DEV)
    nn.mse(pred, target, grad, part)
    torch.cuda.synchronize()
    d = pred[:, :2] - target
    assert abs(float(part.sum()) - float((d * d).mean())) < 1e-5
    ref = torch.zeros_like(pred)
    ref[:, :2] = 2 * d / d.numel()
    assert torch.allclose(grad, ref, atol=1e-8, rtol=1e-5)
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
