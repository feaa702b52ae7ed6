"""Learned surrogate of K map-cycles of the erosion (World::erode,
world.hpp:36-77, run K times) with a judge and a fallback to the plain
method.

* Model: a periodic convolutional encoder-decoder (five resolutions, 3x3
  convolutions, ReLU, average pooling, nearest upsampling, additive skips;
  2.6M parameters) mapping (height anomaly, sediment, cos and sin of the
  wind azimuth broadcast) to the height and sediment changes over K cycles.
  All of its arithmetic runs in libswe_nn.so (BF16 MFMA convolutions with
  FP32 accumulation, FP32 master weights, Adam); torch only holds buffers.
* Data: pairs (state, state after K cycles) generated on the GPU by the plain
  method, many terrains with random wind azimuths at once (a Batch).
* Mass: the predicted height change is projected in int64 fixed point so that
  the map's total changes by exactly 0 (deposit on death: the plain method
  conserves it exactly too); the sediment change is applied in the same
  fixed point and the layer clamped at 0.
* Rollout: blocks of K cycles.  Every `judge_every`-th block the plain method
  re-simulates the block from the same state; the plain result becomes the
  state, and the block's RMS height error decides the next blocks: the
  surrogate while it passes (<= tol_rms), the plain method (fallback) until
  a later judged block passes again.
"""
import json
import math
import time

import numpy as np

from . import nn

# name, input channels, output channels, ReLU
# channels per resolution level (level l = 1/2^l of the map); the encoder
# pools after levels 0 .. L-2, the bottom runs at level L-1
CH = (32, 64, 128, 192, 256)
C_IN = 8    # height anomaly, sediment, cos and sin of the azimuth, two slopes, Laplacian, 0
C_OUT = 16  # height change, sediment change, unused
N_IN, N_OUT = 7, 2


def _layers():
    """(name, input channels, output channels, ReLU, level) of every convolution."""
    L = len(CH)
    out = []
    for l in range(L - 1):
        out.append(("enc%da" % l, C_IN if l == 0 else CH[l - 1], CH[l], 1, l))
        out.append(("enc%db" % l, CH[l], CH[l], 1, l))
    out.append(("mida", CH[L - 2], CH[L - 1], 1, L - 1))
    out.append(("midb", CH[L - 1], CH[L - 2], 1, L - 1))
    for l in range(L - 2, -1, -1):
        out.append(("dec%da" % l, CH[l], CH[l], 1, l))
        out.append(("dec%db" % l, CH[l], CH[l - 1] if l > 0 else C_OUT, 1 if l > 0 else 0, l))
    return out


LAYERS = [(n, ci, co, r) for n, ci, co, r, _ in _layers()]
LEVEL = {n: lv for n, _, _, _, lv in _layers()}


def _torch():
    import torch

    return torch


class Surrogate:
    def __init__(self, seed=0, device="cuda"):
        torch = _torch()
        self.dev = torch.device(device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.w32, self.w16, self.wt16, self.b32 = {}, {}, {}, {}
        self.gw, self.gb, self.m, self.v = {}, {}, {}, {}
        for name, cin, cout, relu in LAYERS:
            std = math.sqrt(2.0 / (9 * cin)) if relu else 0.1 * math.sqrt(1.0 / (9 * cin))
            w = (torch.randn((cout, 9, cin), generator=g) * std).to(self.dev)
            if name == "enc0a":
                w[:, :, N_IN:] = 0.0  # the padding channels carry no data
            self.w32[name] = w.contiguous()
            self.w16[name] = w.to(torch.bfloat16).contiguous()
            self.wt16[name] = nn.flip(self.w16[name])  # the data gradient's weights, kept by Adam
            self.b32[name] = torch.zeros(cout, device=self.dev)
            self.gw[name] = torch.zeros_like(self.w32[name])
            self.gb[name] = torch.zeros_like(self.b32[name])
            for key in (name, name + ".b"):
                shape = self.w32[name].shape if key == name else self.b32[name].shape
                self.m[key] = torch.zeros(shape, device=self.dev)
                self.v[key] = torch.zeros(shape, device=self.dev)
        self.h_scale = 1.0   # height anomaly scale of the inputs
        self.s_scale = 1.0   # sediment scale of the inputs
        self.g_scale = 1.0   # scale of the height's central differences
        self.l_scale = 1.0   # scale of the height's Laplacian
        self.d_scale = 1.0   # scale of the K-cycle height change
        self.ds_scale = 1.0  # scale of the K-cycle sediment change
        self.K = 50
        self.step = 0
        self._work = None
        self._adam = None  # swe_nn_adam_multi table over every parameter tensor
        self._cw = None    # split-K partials of the low-resolution convolutions

    # -- bookkeeping ----------------------------------------------------------
    def n_params(self):
        return sum(w.numel() + self.b32[n].numel() for n, w in self.w32.items())

    @staticmethod
    def flops_forward(N, H, W):
        """2 * Cin * Cout * 9 * pixels summed over the convolutions."""
        return sum(2 * cin * cout * 9 * N * (H >> LEVEL[n]) * (W >> LEVEL[n]) for n, cin, cout, _ in LAYERS)

    def state_dict(self):
        d = {"w." + n: self.w32[n].cpu().numpy() for n in self.w32}
        d.update({"b." + n: self.b32[n].cpu().numpy() for n in self.b32})
        d["scales"] = np.array([self.h_scale, self.s_scale, self.g_scale, self.l_scale, self.d_scale, self.ds_scale,
                                self.K], np.float64)
        return d

    def load_state_dict(self, d):
        torch = _torch()
        for n in self.w32:
            self.w32[n] = torch.from_numpy(np.asarray(d["w." + n], np.float32)).to(self.dev).contiguous()
            self.w16[n] = self.w32[n].to(torch.bfloat16).contiguous()
            self.wt16[n] = nn.flip(self.w16[n])
            self.b32[n] = torch.from_numpy(np.asarray(d["b." + n], np.float32)).to(self.dev).contiguous()
        self._adam = None
        (self.h_scale, self.s_scale, self.g_scale, self.l_scale, self.d_scale, self.ds_scale,
         k) = (float(x) for x in d["scales"])
        self.K = int(k)

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path, device="cuda"):
        s = cls(device=device)
        with np.load(path, allow_pickle=False) as f:
            s.load_state_dict({k: f[k] for k in f.files})
        return s

    # -- network ----------------------------------------------------------------
    def features(self, h, s, az):
        """h, s [N,H,W] float (device), az [N] wind azimuths -> x0 [N,H,W,8]
        bf16 (libswe_nn's feature kernel: anomaly, sediment, cos, sin,
        slopes and Laplacian computed in float)."""
        torch = _torch()
        N = h.shape[0]
        h = h.contiguous().float()
        s = s.contiguous().float()
        mean = h.mean(dim=(1, 2)).contiguous()
        azt = torch.as_tensor(np.asarray(az, np.float32).reshape(N)).to(self.dev).contiguous()
        return nn.features(h, s, mean, azt, self.h_scale, self.s_scale, self.g_scale, self.l_scale)

    def _conv_work(self, shape, cout):
        """The workspace a convolution of input `shape` (NHWC) to `cout`
        channels splits its reduction into, or None."""
        torch = _torch()
        need = nn.conv_work(shape[0], shape[1], shape[2], shape[3], cout)
        if need == 0:
            return None
        if self._cw is None or self._cw.numel() < need:
            self._cw = torch.empty(need, device=self.dev, dtype=torch.float32)
        return self._cw

    def _buf(self, shape, dtype=None):
        torch = _torch()
        return torch.empty(shape, device=self.dev, dtype=dtype or torch.bfloat16)

    def forward(self, x0):
        """x0 [N,H,W,8] bf16 (H, W multiples of 2^(levels-1)) -> (out
        [N,H,W,16] float, activations)."""
        torch = _torch()
        N, H, W, _ = x0.shape
        L = len(CH)
        A = {"x0": x0}

        def conv(name, x, out32=False):
            cout = self.w16[name].shape[0]
            y = self._buf(x.shape[:3] + (cout,), torch.float32 if out32 else None)
            ws = self._conv_work(x.shape, cout)
            if out32:
                nn.conv3x3(x, self.w16[name], self.b32[name], relu=False, out32=y, work=ws)
            else:
                nn.conv3x3(x, self.w16[name], self.b32[name], relu=True, out16=y, work=ws)
            A[name] = y
            return y

        x = x0
        for l in range(L - 1):
            s = conv("enc%db" % l, conv("enc%da" % l, x))
            x = nn.pool2(s, self._buf((N, H >> (l + 1), W >> (l + 1), CH[l])))
            A["pool%d" % l] = x
        b = conv("midb", conv("mida", x))
        for l in range(L - 2, -1, -1):
            hl, wl = H >> l, W >> l
            c = nn.up2_add(b, A["enc%db" % l], self._buf((N, hl, wl, CH[l])))
            A["cat%d" % l] = c
            b = conv("dec%db" % l, conv("dec%da" % l, c), out32=(l == 0))
        return b, A

    def backward(self, A, dout):
        """dout [N,H,W,16] float or bf16: gradient of the loss by the output;
        fills self.gw / self.gb."""
        torch = _torch()
        N, H, W, _ = dout.shape
        L = len(CH)

        def layer(name, x_in, g16, mask=None, need_dx=True):
            """g16: bf16 gradient by this layer's pre-activation output;
            returns the bf16 gradient by its input, zero where mask (the
            ReLU output that is this layer's input) is not positive."""
            self._work = nn.wgrad(x_in, g16, self.gw[name], self.gb[name], self._work)
            if not need_dx:
                return None
            dx = self._buf(x_in.shape[:3] + (self.w16[name].shape[2],))
            return nn.conv3x3_masked(g16, self.wt16[name], mask, dx,
                                     work=self._conv_work(g16.shape, self.w16[name].shape[2]))

        skip = {}
        # the last layer is linear: its gradient is dout in bf16
        g = dout if dout.dtype == torch.bfloat16 else nn.relu_grad(dout, None, self._buf(dout.shape))
        for l in range(L - 1):
            # dec_l_b (its input is dec_l_a's ReLU output), dec_l_a (input cat_l)
            g = layer("dec%db" % l, A["dec%da" % l], g, mask=A["dec%da" % l])
            dcat = layer("dec%da" % l, A["cat%d" % l], g)
            skip[l] = dcat
            hl, wl = H >> (l + 1), W >> (l + 1)
            below = "midb" if l == L - 2 else "dec%db" % (l + 1)
            g = nn.up2_grad_mask(dcat, A[below], self._buf((N, hl, wl, CH[l])))
        # bottom
        g = layer("midb", A["mida"], g, mask=A["mida"])
        dpool = layer("mida", A["pool%d" % (L - 2)], g)
        for l in range(L - 2, -1, -1):
            hl, wl = H >> l, W >> l
            ds = nn.pool2_grad_add(dpool, skip[l], A["enc%db" % l], self._buf((N, hl, wl, CH[l])))
            g = layer("enc%db" % l, A["enc%da" % l], ds, mask=A["enc%da" % l])
            x_in = A["x0"] if l == 0 else A["pool%d" % (l - 1)]
            if l == 0:
                layer("enc0a", x_in, g, need_dx=False)
            else:
                dpool = layer("enc%da" % l, x_in, g)

    def adam(self, lr):
        """One Adam step of every weight and bias in one launch (which also
        keeps the bf16 weights and their flipped copies)."""
        self.step += 1
        nn.adam_multi(self._adam_table(), self.step, lr=lr)

    def train_step(self, h, s, az, dh, ds, lr, border=0):
        """One Adam step on a batch: heights h, sediment s [N,H,W], wind
        azimuths az [N], and their K-cycle changes dh, ds (device float).
        With `border` > 0 the loss gradient is zero on that band along the
        edges (crops of a larger map, whose own edges are not periodic).
        Returns the loss tensor (device)."""
        torch = _torch()
        out, A = self.forward(self.features(h, s, az))
        N, H, W, C = out.shape
        target = torch.stack((dh / self.d_scale, ds / self.ds_scale), dim=-1).reshape(N * H * W, N_OUT).contiguous()
        grad = self._buf(out.shape)
        part = torch.empty(((N * H * W * C + 255) // 256,), device=self.dev)
        nn.mse(out.view(N * H * W, C), target, None, part, grad16=grad.view(N * H * W, C))
        if border > 0:
            grad[:, :border] = 0
            grad[:, H - border:] = 0
            grad[:, :, :border] = 0
            grad[:, :, W - border:] = 0
        self.backward(A, grad)
        self.adam(lr)
        return part.sum()

    def _adam_table(self):
        if self._adam is None:
            segs = []
            for name, cin, cout, _ in LAYERS:
                segs.append((self.w32[name], self.w16[name], self.wt16[name], self.gw[name], self.m[name],
                             self.v[name], cin, cout))
                segs.append((self.b32[name], None, None, self.gb[name], self.m[name + ".b"], self.v[name + ".b"],
                             0, 0))
            self._adam = nn.adam_table(segs, self.dev)
        return self._adam

    def predict(self, h, s, az):
        """K-cycle height and sediment changes [N,H,W] (device float)."""
        out, _ = self.forward(self.features(h, s, az))
        return out[..., 0] * self.d_scale, out[..., 1] * self.ds_scale


class TrainGraph:
    """Surrogate.train_step for a fixed batch shape, captured once as a HIP
    graph and replayed: the batch goes into static device buffers, the mean
    heights are computed before the replay (a torch reduction stays outside
    the graph), and Adam reads its learning rate and bias corrections from
    device memory.  The first step runs eagerly (it allocates the work
    buffers); every step computes the same bits as train_step."""

    def __init__(self, model, N, H, W):
        torch = _torch()
        dev = model.dev
        self.model = model
        self.shape = (N, H, W)
        f = dict(device=dev, dtype=torch.float32)
        self.h, self.s = torch.empty((N, H, W), **f), torch.empty((N, H, W), **f)
        self.dh, self.ds = torch.empty((N, H, W), **f), torch.empty((N, H, W), **f)
        self.mean, self.az = torch.empty((N,), **f), torch.empty((N,), **f)
        self.hyper = torch.empty((3,), **f)
        self.hyper_host = torch.empty((3,), dtype=torch.float32, pin_memory=True)
        self.graph = None
        self.part = None

    def _kernels(self):
        """The captured work: features, forward, loss gradient, backward, Adam."""
        torch = _torch()
        m = self.model
        x0 = nn.features(self.h, self.s, self.mean, self.az, m.h_scale, m.s_scale, m.g_scale, m.l_scale)
        out, A = m.forward(x0)
        N, H, W, C = out.shape
        target = torch.stack((self.dh / m.d_scale, self.ds / m.ds_scale), dim=-1).reshape(N * H * W, N_OUT)
        target = target.contiguous()
        grad = m._buf(out.shape)
        part = torch.empty(((N * H * W * C + 255) // 256,), device=m.dev)
        nn.mse(out.view(N * H * W, C), target, None, part, grad16=grad.view(N * H * W, C))
        m.backward(A, grad)
        nn.adam_multi_dev(m._adam_table(), self.hyper)
        return part

    def step(self, h, s, az, dh, ds, lr):
        """One Adam step on a batch (as Surrogate.train_step); returns the
        loss tensor (device)."""
        torch = _torch()
        m = self.model
        if tuple(h.shape) != self.shape:
            raise ValueError("batch shape %s, graph captured for %s" % (tuple(h.shape), self.shape))
        if self.graph is None:
            loss = m.train_step(h, s, az, dh, ds, lr)  # eager: allocates what the graph will reuse
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.part = self._kernels()
            return loss
        self.h.copy_(h)
        self.s.copy_(s)
        self.dh.copy_(dh)
        self.ds.copy_(ds)
        self.mean.copy_(self.h.mean(dim=(1, 2)))
        self.az.copy_(torch.as_tensor(np.asarray(az, np.float32).reshape(-1)))
        m.step += 1
        self.hyper_host.copy_(torch.tensor(nn.adam_hyper(lr, m.step), dtype=torch.float32))
        self.hyper.copy_(self.hyper_host, non_blocking=True)
        self.graph.replay()
        return self.part.sum()


# -----------------------------------------------------------------------------
# data
# -----------------------------------------------------------------------------
def generate_states(n_terrains, cycles, seed=1, workload="surr256", device=0, log=None):
    """Height and sediment of `n_terrains` maps of a surrogate workload, each
    with its own seed and a random wind azimuth (horizontal wind), after 0 ..
    `cycles` cycles of the plain method, eroded together in one Batch.
    Returns (states float32 device tensor [cycles + 1, n, 2, nx, ny],
    azimuths [n], seconds of erosion)."""
    torch = _torch()
    from .world import Batch, glm_normalize
    from . import workloads

    w = workloads.WORKLOADS[workload]
    workloads.apply(workload, seed)
    nx, ny = w["nx"], w["ny"]
    rng = np.random.default_rng(seed)
    az = rng.uniform(0.0, 2.0 * math.pi, n_terrains)
    b = Batch([(nx, ny)] * n_terrains, device=device, periodic=w["periodic"], layers=w["layers"],
              abrasion=w["params"].get("abrasion", 0.0), deposit_on_death=True)
    for k in range(n_terrains):
        npyr = w["terrain"]["n_pyramids"] * (k % 2)  # obstacles on every other map
        h, sed = workloads.make_inputs(workload, seed * 1000 + k, n_pyramids=npyr)
        b.upload(k, h, sed)
        b.set_terrain(k, seed * 1000 + k, glm_normalize((math.cos(az[k]), 0.0, math.sin(az[k]))))
    out = torch.empty((cycles + 1, n_terrains, 2, nx, ny), device="cuda")

    def grab(t):
        for k in range(n_terrains):
            d = b.download(k)
            out[t, k, 0] = torch.from_numpy(d["height"]).to("cuda")
            out[t, k, 1] = torch.from_numpy(d["sediment"]).to("cuda")

    grab(0)
    t_erode = 0.0
    for t in range(1, cycles + 1):
        t0 = time.perf_counter()
        b.erode([w["particles"]] * n_terrains)
        b.synchronize()
        t_erode += time.perf_counter() - t0
        grab(t)
        if log and t % 100 == 0:
            log("generated %d / %d cycles" % (t, cycles))
    b.close()
    return out, az, t_erode


def fit_scales(model, states, K):
    """Normalisation scales from (a subsample of at most ~100 times of) the data."""
    T = states.shape[0] - K
    step = max(1, T // 100)
    h, s = states[:T:step, :, 0], states[:T:step, :, 1]
    torch = _torch()
    model.h_scale = float((h - h.mean(dim=(2, 3), keepdim=True)).std())
    model.s_scale = float(s.std()) or 1.0
    gx = 0.5 * (torch.roll(h, -1, dims=2) - torch.roll(h, 1, dims=2))
    lap = torch.roll(h, -1, dims=2) + torch.roll(h, 1, dims=2) + torch.roll(h, -1, dims=3) + torch.roll(h, 1, dims=3) - 4 * h
    model.g_scale = float(gx.std()) or 1.0
    model.l_scale = float(lap.std()) or 1.0
    model.d_scale = float((states[K:K + T:step, :, 0] - h).std())
    model.ds_scale = float((states[K:K + T:step, :, 1] - s).std()) or 1.0
    model.K = K


def symmetry(x, az, k):
    """One of the eight symmetries of the square periodic map applied to maps
    x [..., X, Y] and wind azimuths az (wind (cos az, sin az) along (X, Y)):
    bit 0 swaps the axes (az -> pi/2 - az), bit 1 reverses X (az -> pi -
    az), bit 2 reverses Y (az -> -az).  The erosion is isotropic on a
    periodic map -- spawn, the wind's effect and the cascade treat the two
    axes alike -- so a transformed pair is a pair of the same process."""
    torch = _torch()
    az = np.asarray(az, np.float64)
    if k & 1:
        x = x.transpose(-1, -2)
        az = 0.5 * math.pi - az
    if k & 2:
        x = torch.flip(x, dims=(-2,))
        az = math.pi - az
    if k & 4:
        x = torch.flip(x, dims=(-1,))
        az = -az
    return x.contiguous(), az


def train(model, states, az, K, steps, batch=8, lr=1e-3, seed=0, log=None, log_every=100, augment=True,
          pushforward=0.0, pushforward_from=0.5, pushforward_blocks=1, graph=False, crop=0, border=0):
    """Adam on random (terrain, t) pairs of `states` with a cosine-decayed
    learning rate; with `augment`, each batch under a random one of the
    map's eight symmetries (square maps).  With `pushforward` p > 0, after
    the first `pushforward_from` fraction of the steps a batch is, with
    probability p, a pushforward pair: its input is the state the network
    reaches by itself in b blocks of K cycles from a plain state (b uniform
    in 1 .. pushforward_blocks, as many as the data allows) -- the state a
    rollout reaches after b surrogate blocks -- and its target the plain
    state (b + 1) K cycles on, so that the network learns to correct its own
    drift.  With `graph`, the steps replay one captured HIP graph
    (TrainGraph: the same bits, less host time).  With `crop` > 0 each
    pair is a crop x crop window at a random place of its (periodic, larger)
    map, the loss gradient zero on a `border` band along the crop's edges,
    which are not periodic.  Returns the loss curve and the time of the
    steps."""
    torch = _torch()
    T, n = states.shape[0] - K, states.shape[1]
    T2 = states.shape[0] - 2 * K  # pushforward needs two blocks of data after t
    MX, MY = states.shape[-2], states.shape[-1]
    crop = int(crop) if crop and crop < min(MX, MY) else 0
    if crop and (pushforward > 0 or graph):
        raise ValueError("crops are not combined with pushforward or graph training")
    dev = states.device

    def crops(t, k):
        """[batch, 2, crop, crop] windows of states[t, k] at random places (wrapping)."""
        ox = torch.randint(0, MX, (batch,), generator=g)
        oy = torch.randint(0, MY, (batch,), generator=g)
        ar = torch.arange(crop)
        ix = ((ox[:, None] + ar[None, :]) % MX).to(dev)
        iy = ((oy[:, None] + ar[None, :]) % MY).to(dev)
        ch = torch.arange(2, device=dev)[None, :, None, None]

        def take(tt):
            return states[tt.to(dev)[:, None, None, None], k.to(dev)[:, None, None, None], ch,
                          ix[:, None, :, None], iy[:, None, None, :]]
        return take(t), take(t + K)

    g = torch.Generator(device="cpu").manual_seed(seed)
    az = np.asarray(az, np.float64)
    augment = augment and states.shape[-1] == states.shape[-2]
    curve = []
    tg = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for st in range(steps):
        push = (pushforward > 0 and T2 > 0 and st >= pushforward_from * steps
                and float(torch.rand((1,), generator=g)) < pushforward)
        nb = 1
        if push and pushforward_blocks > 1:
            nbmax = min(int(pushforward_blocks), (states.shape[0] - 1) // K - 1)
            nb = int(torch.randint(1, nbmax + 1, (1,), generator=g))
        t = torch.randint(0, (states.shape[0] - (nb + 1) * K) if push else T, (batch,), generator=g)
        k = torch.randint(0, n, (batch,), generator=g)
        a = az[k.numpy()]
        if push:
            xs = states[t, k]
            h, sd = xs[:, 0].contiguous(), xs[:, 1].contiguous()
            for _ in range(nb):
                dh, ds = model.predict(h, sd, a.astype(np.float32))
                h, sd = (h + dh).contiguous(), (sd + ds).clamp_min(0.0).contiguous()
            x0 = torch.stack((h, sd), dim=1)
            x1 = states[t + (nb + 1) * K, k]
        elif crop:
            x0, x1 = crops(t, k)
        else:
            x0, x1 = states[t, k], states[t + K, k]
        if augment:
            sym = int(torch.randint(0, 8, (1,), generator=g))
            pair, a = symmetry(torch.stack((x0, x1)), a, sym)
            x0, x1 = pair[0], pair[1]
        rate = lr * 0.5 * (1.0 + math.cos(math.pi * st / max(1, steps)))
        args = (x0[:, 0].contiguous(), x0[:, 1].contiguous(), a.astype(np.float32),
                (x1[:, 0] - x0[:, 0]).contiguous(), (x1[:, 1] - x0[:, 1]).contiguous(), rate)
        if graph:
            if tg is None:
                tg = TrainGraph(model, *args[0].shape)
            loss = tg.step(*args)
        else:
            loss = model.train_step(*args, border=border if crop else 0)
        if st % log_every == 0 or st == steps - 1:
            curve.append((st, float(loss)))
            if log:
                log("step %d loss %.4f" % (st, curve[-1][1]))
    torch.cuda.synchronize()
    return curve, time.perf_counter() - t0


# -----------------------------------------------------------------------------
# rollout with judge and fallback
# -----------------------------------------------------------------------------
def rms(a):
    return float(np.sqrt(np.mean(np.square(np.asarray(a, np.float64)))))


def surrogate_block(model, zm, sm, h, s, az, work):
    """The masters after one surrogate block: int64 height master with the
    predicted change projected to sum exactly 0, sediment master with its
    change applied and clamped at 0."""
    torch = _torch()
    nx, ny = zm.shape
    dh, ds = model.predict(torch.from_numpy(h).to("cuda")[None], torch.from_numpy(s).to("cuda")[None], [az])
    zt = torch.from_numpy(zm.reshape(-1)).to("cuda")
    st = torch.from_numpy(sm.reshape(-1)).to("cuda")
    nn.project_mass(dh[0].contiguous().view(-1), 1.0, 0, zt, work)
    target = int(round(float(ds.double().sum()) * 2.0 ** 32))
    nn.project_mass(ds[0].contiguous().view(-1), 1.0, target, st, work)
    st.clamp_(min=0)
    return zt.cpu().numpy().reshape(nx, ny), st.cpu().numpy().reshape(nx, ny)


def rollout(model, world, plain, particles, blocks, azimuth, judge_every=10, tol_rms=1e-3, noise_floor=True,
            tol_noise=None, noise_every_judge=False, log=None):
    """Advance `world` (two layers, deposit on death) by `blocks` blocks of
    model.K cycles.  A block is the surrogate's (mass-projected) prediction
    while the last judge passed; every `judge_every`-th block the plain
    method (on the twin context `plain`) re-simulates it from the same state
    and its result becomes the state; the judged block's RMS height error
    (surrogate vs plain) against tol_rms decides whether the next blocks use
    the surrogate or the plain method.  With noise_floor, the first judged
    block also runs a second plain realisation (other random streams) to give
    the RMS difference of two plain runs; with tol_noise, the tolerance
    becomes max(tol_rms, tol_noise x that noise floor).  With
    noise_every_judge, every judged block measures the floor at its own
    state (one more plain block per judge) instead of reusing the first."""
    torch = _torch()
    K = model.K
    work = torch.empty(256, device="cuda", dtype=torch.int64)
    log = log or (lambda *_: None)
    tol_rms0 = tol_rms
    rec = {"blocks": blocks, "K": K, "tol_rms": tol_rms, "judged": [], "surrogate_blocks": 0, "plain_blocks": 0,
           "judge_blocks": 0, "t_surrogate_s": 0.0, "t_plain_s": 0.0, "noise_floor_rms": None}
    zm0, _ = world.master()
    mass0 = int(zm0.astype(object).sum())
    use_sur = True
    scale = 2.0 ** -32
    n_plain_runs = [0]

    def plain_block(zm, sm, cyc):
        plain.upload_master(zm, sm)
        plain.cycle = cyc
        t0 = time.perf_counter()
        plain.erode(particles, n_cycles=K)
        plain.synchronize()
        rec["t_plain_s"] += time.perf_counter() - t0
        n_plain_runs[0] += 1
        return plain.master()

    t_start = time.perf_counter()
    for blk in range(blocks):
        zm, sm = world.master()
        cyc = world.cycle
        judged = (blk + 1) % judge_every == 0
        if judged or use_sur:
            d = world.download()
            t0 = time.perf_counter()
            zs, ss = surrogate_block(model, zm, sm, d["height"], d["sediment"], azimuth, work)
            rec["t_surrogate_s"] += time.perf_counter() - t0
        if judged:
            zp, sp = plain_block(zm, sm, cyc)
            rec["judge_blocks"] += 1
            m = {"block": blk, "rms_err": rms((zs - zp) * scale), "rms_change_plain": rms((zp - zm) * scale),
                 "rms_change_surrogate": rms((zs - zm) * scale)}
            if (noise_floor or tol_noise) and (rec["noise_floor_rms"] is None or noise_every_judge):
                zp2, _ = plain_block(zm, sm, cyc + 1_000_003)  # other random streams
                rec["noise_floor_rms"] = rms((zp2 - zp) * scale)
                m["noise_floor_rms"] = rec["noise_floor_rms"]
                if tol_noise:
                    tol_rms = max(tol_rms0, tol_noise * rec["noise_floor_rms"])
                rec["tol_rms_effective"] = tol_rms
            m["tol_rms"] = tol_rms
            m["accepted"] = m["rms_err"] <= tol_rms
            use_sur = m["accepted"]
            rec["judged"].append(m)
            log("block %d judge %s" % (blk, json.dumps(m)))
            world.upload_master(zp, sp)
        elif use_sur:
            world.upload_master(zs, ss)
            rec["surrogate_blocks"] += 1
        else:
            zp, sp = plain_block(zm, sm, cyc)
            world.upload_master(zp, sp)
            rec["plain_blocks"] += 1
        world.cycle = cyc + K
    rec["t_total_s"] = time.perf_counter() - t_start
    rec["plain_seconds_per_block"] = rec["t_plain_s"] / max(1, n_plain_runs[0])
    zm, _ = world.master()
    rec["mass_drift_fx"] = int(zm.astype(object).sum()) - mass0
    rec["mass_exact"] = rec["mass_drift_fx"] == 0
    return rec


def main():
    """Train on 256^2 maps with random winds and roll out on the 1024^2 map
    with obstacles; prints one JSON summary (losses, MFMA throughput of
    training and inference, judge results per block, speedup over the plain
    method, mass drift)."""
    import argparse

    torch = _torch()
    ap = argparse.ArgumentParser()
    ap.add_argument("--terrains", type=int, default=16)
    ap.add_argument("--cycles", type=int, default=1300, help="cycles of training data per terrain")
    ap.add_argument("--K", type=int, default=50)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--blocks", type=int, default=400)
    ap.add_argument("--judge-every", type=int, default=10)
    ap.add_argument("--tol-rms", type=float, default=1e-3)
    ap.add_argument("--tol-noise", type=float, default=None,
                    help="accept blocks within this multiple of the plain method's own run-to-run RMS")
    ap.add_argument("--rollout", default="surr1024", help="comma-separated workloads to roll out on")
    ap.add_argument("--no-augment", action="store_true", help="train without the map's symmetries")
    ap.add_argument("--pushforward", type=float, default=0.0,
                    help="fraction of the later batches that start from the network's own prediction")
    ap.add_argument("--save", default=None)
    ap.add_argument("--load", default=None, help="skip data and training: use a saved model")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    from . import workloads

    res = {"config": vars(args)}
    if args.load:
        model = Surrogate.load(args.load)
    else:
        states, az, t_gen = generate_states(args.terrains, args.cycles, log=log)
        pairs = (args.cycles + 1 - args.K) * args.terrains
        gen_steps = args.terrains * args.cycles * workloads.WORKLOADS["surr256"]["particles"] * 513
        res["data"] = {"pairs": pairs, "erode_seconds": t_gen, "particle_steps": gen_steps}
        log("data: %d pairs, %.1f s of erosion" % (pairs, t_gen))
        model = Surrogate(seed=1)
        fit_scales(model, states, args.K)
        curve, t_train = train(model, states, az, args.K, args.steps, batch=args.batch, lr=args.lr, log=log,
                               augment=not args.no_augment, pushforward=args.pushforward)
        fps = 3 * Surrogate.flops_forward(args.batch, 256, 256)
        res["train"] = {"params": model.n_params(), "loss_curve": curve, "seconds": t_train,
                        "flops_per_step": fps, "tflops": fps * args.steps / t_train / 1e12,
                        "mfma_bf16_peak_tflops": 2500.0}
        del states
        torch.cuda.empty_cache()
        if args.save:
            model.save(args.save)
    # inference throughput at the rollout size
    h = torch.zeros((1, 1024, 1024), device="cuda")
    model.predict(h, h, [0.0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        model.predict(h, h, [0.0])
    torch.cuda.synchronize()
    ti = (time.perf_counter() - t0) / 10
    res["inference_1024"] = {"seconds": ti, "tflops": Surrogate.flops_forward(1, 1024, 1024) / ti / 1e12}
    # rollouts: the 1024^2 map with obstacles (and optionally a 256^2 map of
    # the training kind with an unseen seed), each with its plain twin
    for wl in args.rollout.split(","):
        world = workloads.make_world(wl, seed=7, deposit_on_death=True)
        plain = workloads.make_world(wl, seed=7, deposit_on_death=True)
        ps = workloads.WORKLOADS[wl]["params"]["pspeed"]
        azimuth = math.atan2(ps[2], ps[0])
        n = workloads.WORKLOADS[wl]["particles"]
        rec = rollout(model, world, plain, n, args.blocks, azimuth, judge_every=args.judge_every,
                      tol_rms=args.tol_rms, tol_noise=args.tol_noise, log=log)
        rec["map_cycles_per_s"] = args.blocks * model.K / rec["t_total_s"]
        rec["plain_map_cycles_per_s"] = model.K / rec["plain_seconds_per_block"]
        rec["speedup_vs_plain"] = rec["map_cycles_per_s"] / rec["plain_map_cycles_per_s"]
        res["rollout" if wl == "surr1024" else "rollout_" + wl] = rec
        world.close()
        plain.close()
    line = json.dumps(res)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)


if __name__ == "__main__":
    main()
