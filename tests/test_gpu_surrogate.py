"""The surrogate (simplewinderosion_amd/surrogate.py on libswe_nn.so): its
hand-written backward pass against float32 PyTorch autograd of the same
network, training that fits a fixed batch, and a rollout whose fixed-point
mass projection keeps the map's total exact while the judge runs the plain
method."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
F = torch.nn.functional

pytestmark = pytest.mark.gpu

from simplewinderosion_amd import surrogate as S  # noqa: E402
from simplewinderosion_amd import workloads  # noqa: E402


def conv_ref(x, w, b):
    Cout, _, Cin = w.shape
    wc = w.view(Cout, 3, 3, Cin).permute(0, 3, 1, 2)
    return F.conv2d(F.pad(x, (1, 1, 1, 1), mode="circular"), wc, b)


def forward_ref(model, x0):
    """The same network in float32 (NCHW), on the model's bf16 weights."""
    W = {n: model.w16[n].float().requires_grad_(True) for n in model.w16}
    B = {n: model.b32[n].clone().requires_grad_(True) for n in model.b32}

    def c(n, x, relu=True):
        y = conv_ref(x, W[n], B[n])
        return F.relu(y) if relu else y

    x = x0.float().permute(0, 3, 1, 2)
    s0 = c("e0b", c("e0a", x))
    s1 = c("e1b", c("e1a", F.avg_pool2d(s0, 2)))
    b2 = c("m2b", c("m2a", F.avg_pool2d(s1, 2)))
    c1 = F.interpolate(b2, scale_factor=2, mode="nearest") + s1
    b3 = c("d1b", c("d1a", c1))
    c0 = F.interpolate(b3, scale_factor=2, mode="nearest") + s0
    out = c("d0b", c("d0a", c0), relu=False)
    return out.permute(0, 2, 3, 1), W, B


def test_backward_matches_autograd():
    print("case surrogate backward", flush=True)
    m = S.Surrogate(seed=3)
    g = torch.Generator(device="cpu").manual_seed(0)
    h = (torch.randn((2, 64, 128), generator=g) * 0.1).cuda()
    dh = (torch.randn((2, 64, 128), generator=g) * 0.01).cuda()
    m.h_scale, m.d_scale = 0.1, 0.01
    x0 = m.features(h)
    out, A = m.forward(x0)
    ref, W, B = forward_ref(m, x0)
    assert torch.allclose(out, ref, atol=0.05 * float(ref.abs().max()) + 1e-3)
    N, H, Wd, C = out.shape
    target = (dh / m.d_scale).reshape(-1, 1).contiguous()
    grad = torch.empty_like(out)
    part = torch.empty(((N * H * Wd * C + 255) // 256,), device="cuda")
    S.nn.mse(out.view(-1, C), target, grad.view(-1, C), part)
    m.backward(A, grad)
    torch.cuda.synchronize()
    loss = ((ref[..., 0].reshape(-1) - target[:, 0]) ** 2).mean()
    loss.backward()
    for name in m.gw:
        a, b = m.gw[name].flatten(), W[name].grad.flatten()
        if name == "e0a":
            a, b = m.gw[name][:, :, 0].flatten(), W[name].grad[:, :, 0].flatten()
        cos = float(torch.dot(a, b) / (a.norm() * b.norm() + 1e-30))
        assert cos > 0.98, (name, cos)
        cb = float(torch.dot(m.gb[name], B[name].grad) / (m.gb[name].norm() * B[name].grad.norm() + 1e-30))
        assert cb > 0.98, (name, cb)


def test_training_fits_a_batch():
    print("case surrogate fit", flush=True)
    m = S.Surrogate(seed=4)
    g = torch.Generator(device="cpu").manual_seed(1)
    h = (torch.randn((2, 64, 64), generator=g) * 0.1).cuda()
    dh = 0.01 * torch.roll(h, 1, dims=2) - 0.01 * h  # a smooth function of h
    m.h_scale, m.d_scale = float(h.std()), float(dh.std())
    losses = [float(m.train_step(h, dh, 3e-3)) for _ in range(150)]
    assert losses[-1] < 0.3 * losses[0], losses[::30]


def test_rollout_mass_exact_and_judged(statics):
    print("case surrogate rollout", flush=True)
    m = S.Surrogate(seed=5)
    m.K = 2
    m.h_scale, m.d_scale = 0.1, 1e-4
    world = workloads.make_world("surr256", seed=3, deposit_on_death=True)
    plain = workloads.make_world("surr256", seed=3, deposit_on_death=True)
    rec = S.rollout(m, world, plain, 20000, blocks=4, judge_every=2, tol_rel=1e9, tol_w1=1e9)
    assert rec["mass_exact"]
    assert rec["surrogate_blocks"] == 2 and rec["plain_blocks"] == 2 and len(rec["judged"]) == 2
    assert world.cycle == 8
    # a judge that rejects hands the rest to the plain method
    rec = S.rollout(m, world, plain, 20000, blocks=4, judge_every=2, tol_rel=-1.0)
    assert rec["fallback_at"] == 1 and rec["plain_blocks"] == 3 and rec["mass_exact"]
    world.close()
    plain.close()
