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

    L = len(S.CH)
    x = x0.float().permute(0, 3, 1, 2)
    skips = []
    for l in range(L - 1):
        s = c("enc%db" % l, c("enc%da" % l, x))
        skips.append(s)
        x = F.avg_pool2d(s, 2)
    b = c("midb", c("mida", x))
    for l in range(L - 2, -1, -1):
        cat = F.interpolate(b, scale_factor=2, mode="nearest") + skips[l]
        b = c("dec%db" % l, c("dec%da" % l, cat), relu=(l > 0))
    return b.permute(0, 2, 3, 1), W, B


def test_backward_matches_autograd():
    print("case surrogate backward", flush=True)
    m = S.Surrogate(seed=3)
    g = torch.Generator(device="cpu").manual_seed(0)
    h = (torch.randn((2, 64, 128), generator=g) * 0.1).cuda()
    s = (torch.rand((2, 64, 128), generator=g) * 0.05).cuda()
    dh = (torch.randn((2, 64, 128), generator=g) * 0.01).cuda()
    ds = (torch.randn((2, 64, 128), generator=g) * 0.01).cuda()
    m.h_scale, m.s_scale, m.d_scale, m.ds_scale = 0.1, 0.05, 0.01, 0.01
    x0 = m.features(h, s, [0.3, 1.2])
    out, A = m.forward(x0)
    ref, W, B = forward_ref(m, x0)
    assert torch.allclose(out, ref, atol=0.05 * float(ref.abs().max()) + 1e-3)
    N, H, Wd, C = out.shape
    target = torch.stack((dh / m.d_scale, ds / m.ds_scale), dim=-1).reshape(-1, 2).contiguous()
    grad = torch.empty_like(out)
    part = torch.empty(((N * H * Wd * C + 255) // 256,), device="cuda")
    S.nn.mse(out.view(-1, C), target, grad.view(-1, C), part)
    m.backward(A, grad)
    torch.cuda.synchronize()
    loss = ((ref[..., :2].reshape(-1, 2) - target) ** 2).mean()
    loss.backward()
    for name in m.gw:
        a, b = m.gw[name].flatten(), W[name].grad.flatten()
        if name == "enc0a":
            a, b = m.gw[name][:, :, :S.N_IN].flatten(), W[name].grad[:, :, :S.N_IN].flatten()
        cos = float(torch.dot(a, b) / (a.norm() * b.norm() + 1e-30))
        assert cos > 0.98, (name, cos)
        cb = float(torch.dot(m.gb[name], B[name].grad) / (m.gb[name].norm() * B[name].grad.norm() + 1e-30))
        assert cb > 0.98, (name, cb)


def test_graph_training_steps_match_eager():
    """Training steps replayed from one captured HIP graph (TrainGraph) give
    the eager steps' bits -- weights, bf16 and flipped copies, Adam moments,
    loss -- on every replay, with other eager GPU work between replays."""
    print("case train graph", flush=True)
    a = S.Surrogate(seed=3)
    b = S.Surrogate(seed=3)
    tg = S.TrainGraph(b, 2, 64, 64)
    g = torch.Generator(device="cpu").manual_seed(0)
    az = np.array([0.3, 1.9], np.float32)
    for step in range(5):
        h = (torch.rand((2, 64, 64), generator=g) * 0.5).to("cuda")
        s = (torch.rand((2, 64, 64), generator=g) * 0.02).to("cuda")
        dh = (torch.randn((2, 64, 64), generator=g) * 0.01).to("cuda")
        ds = (torch.randn((2, 64, 64), generator=g) * 0.01).to("cuda")
        lr = 1e-3 * (1.0 - 0.1 * step)
        la = a.train_step(h, s, az, dh, ds, lr)
        lb = tg.step(h, s, az, dh, ds, lr)
        b.predict(h, s, az)  # eager work between replays
        torch.cuda.synchronize()
        assert float(la) == float(lb), step
        assert a.step == b.step
        for n in a.w32:
            assert torch.equal(a.w32[n], b.w32[n]) and torch.equal(a.b32[n], b.b32[n]), (step, n)
            assert torch.equal(a.w16[n], b.w16[n]) and torch.equal(a.wt16[n], b.wt16[n]), (step, n)
            for k in (n, n + ".b"):
                assert torch.equal(a.m[k], b.m[k]) and torch.equal(a.v[k], b.v[k]), (step, k)
    assert tg.graph is not None


def test_crop_training_masks_the_border():
    """Training on crops of a larger periodic map: with the border band
    covering the whole crop the loss gradient is zero and Adam leaves every
    weight bit for bit; with a narrow band the weights move."""
    print("case crop training", flush=True)
    g = torch.Generator(device="cpu").manual_seed(3)
    states = (torch.rand((6, 2, 2, 48, 48), generator=g) * 0.1).to("cuda")
    az = np.array([0.3, 1.7])

    def run(border):
        m = S.Surrogate(seed=0)
        S.fit_scales(m, states, 2)
        before = {k: v.clone() for k, v in m.w32.items()}
        curve, _ = S.train(m, states, az, 2, 2, batch=4, crop=32, border=border, seed=1)
        torch.cuda.synchronize()
        assert all(np.isfinite(l) for _, l in curve)
        return sum(int(not torch.equal(before[k], m.w32[k])) for k in before), len(before)

    changed, total = run(16)  # bands of 16 on both sides of 32 cover the crop
    assert changed == 0
    changed, total = run(4)
    assert changed == total


def test_training_fits_a_batch():
    print("case surrogate fit", flush=True)
    m = S.Surrogate(seed=4)
    g = torch.Generator(device="cpu").manual_seed(1)
    h = (torch.randn((2, 64, 64), generator=g) * 0.1).cuda()
    s = (torch.rand((2, 64, 64), generator=g) * 0.05).cuda()
    dh = 0.01 * torch.roll(h, 1, dims=2) - 0.01 * h  # smooth functions of the state
    ds = 0.5 * dh + 0.1 * s
    m.h_scale, m.s_scale, m.d_scale, m.ds_scale = float(h.std()), float(s.std()), float(dh.std()), float(ds.std())
    losses = [float(m.train_step(h, s, [0.5, 2.0], dh, ds, 2e-3)) for _ in range(150)]
    assert losses[-1] < 0.3 * losses[0], losses[::30]


def test_rollout_mass_exact_and_judged(statics):
    print("case surrogate rollout", flush=True)
    m = S.Surrogate(seed=5)
    m.K = 2
    m.h_scale, m.s_scale, m.d_scale, m.ds_scale = 0.1, 0.05, 1e-4, 1e-4
    world = workloads.make_world("surr256", seed=3, deposit_on_death=True)
    plain = workloads.make_world("surr256", seed=3, deposit_on_death=True)
    rec = S.rollout(m, world, plain, 20000, blocks=4, azimuth=0.785, judge_every=2, tol_rms=1e9)
    assert rec["mass_exact"] and rec["mass_drift_fx"] == 0
    assert rec["surrogate_blocks"] == 2 and rec["judge_blocks"] == 2 and rec["plain_blocks"] == 0
    assert len(rec["judged"]) == 2 and rec["noise_floor_rms"] is not None
    assert world.cycle == 8
    _, sm = world.master()
    assert sm.min() >= 0
    # a judge that rejects hands the next blocks to the plain method
    rec = S.rollout(m, world, plain, 20000, blocks=4, azimuth=0.785, judge_every=2, tol_rms=-1.0)
    assert rec["surrogate_blocks"] == 1 and rec["plain_blocks"] == 1 and rec["judge_blocks"] == 2
    assert not any(j["accepted"] for j in rec["judged"]) and rec["mass_exact"]
    world.close()
    plain.close()


def test_trained_network_beats_no_change(statics):
    """What the network gains over the method without it: trained briefly on
    pairs the plain method generates (64x64 maps of the surrogate workload's
    dynamics, K = 20), it predicts a held-out map's K-cycle height change
    with a smaller RMS error than the no-change prediction."""
    print("case surrogate gain", flush=True)
    import math

    from simplewinderosion_amd.world import Batch, glm_normalize

    K, T, n = 20, 100, 5
    workloads.apply("surr256", 11)
    b = Batch([(64, 64)] * n, periodic=True, layers=2, abrasion=1e-4, deposit_on_death=True)
    az = [0.3 + 1.1 * k for k in range(n)]
    for k in range(n):
        h, s = workloads.make_inputs("surr256", 100 + k, nx=64, ny=64)
        b.upload(k, h, s)
        b.set_terrain(k, 100 + k, glm_normalize((math.cos(az[k]), 0.0, math.sin(az[k]))))
    states = torch.empty((T + 1, n, 2, 64, 64), device="cuda")
    for t in range(T + 1):
        if t:
            b.erode([4096] * n)
        for k in range(n):
            d = b.download(k)
            states[t, k, 0] = torch.from_numpy(d["height"]).cuda()
            states[t, k, 1] = torch.from_numpy(d["sediment"]).cuda()
    b.close()
    train_states, test_states = states[:, : n - 1], states[:, n - 1:]
    m = S.Surrogate(seed=2)
    S.fit_scales(m, train_states, K)
    S.train(m, train_states, az[: n - 1], K, steps=1500, batch=8, lr=2e-3)
    h0, h1 = test_states[: T - K, 0], test_states[K:T, 0]
    dh, _ = m.predict(h0[:, 0].contiguous(), h0[:, 1].contiguous(), [az[-1]] * (T - K))
    true = h1[:, 0] - h0[:, 0]
    err = float(((dh - true) ** 2).mean().sqrt())
    base = float((true ** 2).mean().sqrt())
    print("held-out RMS error %.3g, no-change %.3g" % (err, base), flush=True)
    assert err < 0.9 * base


def test_config_d_pairs_training_and_gradients_at_256(statics):
    """The training configuration's shapes: pairs (state, state after K = 50
    cycles) generated by the plain method on 256x256 maps of the training
    workload (one particle per cell, random winds, obstacles on every other
    map); Adam steps on batches of them at 256x256; and the hand-written
    backward against float32 autograd at 256x256, every layer at its real
    shape."""
    print("case surrogate config D", flush=True)
    states, az, _ = S.generate_states(2, 56, seed=5)
    assert states.shape == (57, 2, 2, 256, 256)
    K = 50
    m = S.Surrogate(seed=6)
    S.fit_scales(m, states, K)
    assert m.K == K and m.d_scale > 0
    curve, _ = S.train(m, states, az, K, steps=12, batch=4, lr=1e-3, log_every=1)
    losses = [v for _, v in curve]
    assert all(np.isfinite(losses)) and min(losses[-4:]) < losses[0], losses
    # gradient check at 256x256
    x0, x1 = states[:1, 0], states[K:K + 1, 0]
    h, s = x0[:, 0].contiguous(), x0[:, 1].contiguous()
    dh, ds = (x1[:, 0] - x0[:, 0]).contiguous(), (x1[:, 1] - x0[:, 1]).contiguous()
    feats = m.features(h, s, [float(az[0])])
    out, A = m.forward(feats)
    ref, W, B = forward_ref(m, feats)
    assert torch.allclose(out, ref, atol=0.05 * float(ref.abs().max()) + 1e-3)
    N, H, Wd, C = out.shape
    target = torch.stack((dh / m.d_scale, ds / m.ds_scale), dim=-1).reshape(-1, 2).contiguous()
    grad = torch.empty_like(out)
    part = torch.empty(((N * H * Wd * C + 255) // 256,), device="cuda")
    S.nn.mse(out.view(-1, C), target, grad.view(-1, C), part)
    m.backward(A, grad)
    torch.cuda.synchronize()
    ((ref[..., :2].reshape(-1, 2) - target) ** 2).mean().backward()
    for name in m.gw:
        a, b = m.gw[name], W[name].grad
        if name == "enc0a":
            a, b = a[:, :, :S.N_IN], b[:, :, :S.N_IN]
        a, b = a.flatten(), b.flatten()
        cos = float(torch.dot(a, b) / (a.norm() * b.norm() + 1e-30))
        assert cos > 0.98, (name, cos)


def test_config_e_rollout_blocks_at_1024(statics):
    """The rollout configuration's shapes: K = 50 blocks on the 1024x1024 map
    with its 8 obstacles, the plain method judging every 2nd block (with a
    second plain run for the noise floor) and taking over after a rejected
    judgement; the map's mass stays exact (drift 0 in fixed point)."""
    print("case surrogate config E", flush=True)
    m = S.Surrogate(seed=7)
    m.K = 50
    m.h_scale, m.s_scale, m.g_scale, m.l_scale, m.d_scale, m.ds_scale = 0.2, 0.05, 1e-3, 1e-3, 0.02, 0.02
    world = workloads.make_world("surr1024", seed=7, deposit_on_death=True)
    plain = workloads.make_world("surr1024", seed=7, deposit_on_death=True)
    rec = S.rollout(m, world, plain, 1_000_000, blocks=3, azimuth=0.785, judge_every=2, tol_rms=-1.0)
    assert rec["K"] == 50 and rec["judge_blocks"] == 1
    assert rec["surrogate_blocks"] == 1 and rec["plain_blocks"] == 1  # rejected judge: fallback
    j = rec["judged"][0]
    assert not j["accepted"] and j["rms_change_plain"] > 0 and rec["noise_floor_rms"] > 0
    assert rec["mass_exact"] and rec["mass_drift_fx"] == 0
    assert world.cycle == 150
    _, sm = world.master()
    assert sm.min() >= 0
    world.close()
    plain.close()
