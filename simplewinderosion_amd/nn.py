"""ctypes binding of libswe_nn.so (include/swe_nn.h), the surrogate's HIP
kernels, over torch device tensors (torch supplies device memory and the
stream; the arithmetic is the library's).  Every function launches on
torch's current stream and raises on a bad argument or launch error.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswe_nn.so")

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F = ctypes.c_float
SIGNATURES = {
    "swe_nn_conv3x3": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _P, _P, _I32, _I32, _P, _P, _P]),
    "swe_nn_conv3x3_masked": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _P, _P, _I32, _P, _P, _P]),
    "swe_nn_conv3x3_work": (_I64, [_I32, _I32, _I32, _I32, _I32]),
    "swe_nn_conv3x3_ws": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _P, _P, _I32, _I32, _P, _P, _P, _P, _I64, _P]),
    "swe_nn_flip": (ctypes.c_int, [_P, _I32, _I32, _P, _P]),
    "swe_nn_wgrad_work": (_I64, [_I32, _I32, _I32, _I32, _I32]),
    "swe_nn_wgrad": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P]),
    "swe_nn_relu_grad": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "swe_nn_relu_mask": (ctypes.c_int, [_P, _P, _I64, _P]),
    "swe_nn_add": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "swe_nn_pool2": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_pool2_grad": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_up2": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_up2_grad": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_up2_add": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_pool2_grad_add": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_up2_grad_mask": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _P]),
    "swe_nn_mse": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I32, _I32, _P]),
    "swe_nn_adam": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _F, _F, _F, _F, _I32, _P]),
    "swe_nn_adam_multi": (ctypes.c_int, [_P, _I32, _I64, _F, _F, _F, _F, _I32, _P]),
    "swe_nn_adam_multi_dev": (ctypes.c_int, [_P, _I32, _I64, _P, _F, _F, _F, _P]),
    "swe_nn_adam_hyper": (None, [_F, _F, _F, _I32, _P]),
    "swe_nn_features": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _I32, _F, _F, _F, _F, _P, _P]),
    "swe_nn_project_mass": (ctypes.c_int, [_P, _I64, _F, _I64, _P, _P, _P]),
}

_lib = None


class AdamSeg(ctypes.Structure):
    """swe_nn_adam_seg (include/swe_nn.h)."""
    _fields_ = [("w", _P), ("w16", _P), ("wt16", _P), ("g", _P), ("m", _P), ("v", _P), ("off", _I64), ("n", _I64),
                ("cin", _I32), ("cout", _I32)]


def load():
    """Load libswe_nn.so (built in-tree by simplewinderosion_amd/build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libswe_nn.so is not built (%s); run `python -c 'import __graft_entry__ as g; g.build()'`"
                               % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d)" % (what, rc))


def _dense(t, dtype, what):
    import torch

    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError("%s must be a contiguous %s device tensor" % (what, dtype))
    if t.data_ptr() % 16:
        raise ValueError("%s must be 16-byte aligned" % what)
    return t


def conv_work(N, H, W, Cin, Cout):
    """Floats of workspace a split convolution of this shape uses (0: unsplit)."""
    return int(load().swe_nn_conv3x3_work(N, H, W, Cin, Cout))


def conv3x3(x, w, bias=None, relu=False, out16=None, out32=None, work=None):
    """x [N,H,W,Cin] bf16, w [Cout,9,Cin] bf16, bias [Cout] float -> out16
    [N,H,W,Cout] bf16 and/or out32 float (allocated when neither is given).
    With `work` (float device tensor of conv_work(...) floats), layers with
    few pixel blocks split their reduction over several workgroups."""
    import torch

    _dense(x, torch.bfloat16, "x")
    _dense(w, torch.bfloat16, "w")
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    if w.shape[1] != 9 or w.shape[2] != Cin:
        raise ValueError("weights must be [Cout, 9, Cin]")
    if out16 is None and out32 is None:
        out16 = torch.empty((N, H, W, Cout), device=x.device, dtype=torch.bfloat16)
    if work is not None:
        _ck(load().swe_nn_conv3x3_ws(_p(x), N, H, W, Cin, _p(w), _p(bias), Cout, int(bool(relu)), None, _p(out16),
                                     _p(out32), _p(work), work.numel(), _stream()), "swe_nn_conv3x3_ws")
    else:
        _ck(load().swe_nn_conv3x3(_p(x), N, H, W, Cin, _p(w), _p(bias), Cout, int(bool(relu)), _p(out16),
                                  _p(out32), _stream()), "swe_nn_conv3x3")
    return out16 if out16 is not None else out32


def conv3x3_masked(x, w, mask, out16, bias=None, work=None):
    """out16 [N,H,W,Cout] bf16 = bf16(conv3x3(x, w) + bias), zero where mask
    [N,H,W,Cout] bf16 <= 0 (mask optional): a data gradient through a ReLU."""
    import torch

    _dense(x, torch.bfloat16, "x")
    _dense(w, torch.bfloat16, "w")
    _dense(out16, torch.bfloat16, "out16")
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    if w.shape[1] != 9 or w.shape[2] != Cin or tuple(out16.shape) != (N, H, W, Cout):
        raise ValueError("weights must be [Cout, 9, Cin] and out16 [N, H, W, Cout]")
    if mask is not None and (mask.shape != out16.shape or mask.dtype != torch.bfloat16):
        raise ValueError("mask must be a bf16 tensor shaped like out16")
    if work is not None:
        _ck(load().swe_nn_conv3x3_ws(_p(x), N, H, W, Cin, _p(w), _p(bias), Cout, 2 if mask is not None else 0,
                                     _p(mask), _p(out16), None, _p(work), work.numel(), _stream()),
            "swe_nn_conv3x3_ws")
    else:
        _ck(load().swe_nn_conv3x3_masked(_p(x), N, H, W, Cin, _p(w), _p(bias), Cout, _p(mask), _p(out16),
                                         _stream()), "swe_nn_conv3x3_masked")
    return out16


def flip(w):
    """[Cout,9,Cin] -> [Cin,9,Cout] weights of the data gradient."""
    import torch

    Cout, _, Cin = w.shape
    wt = torch.empty((Cin, 9, Cout), device=w.device, dtype=torch.bfloat16)
    _ck(load().swe_nn_flip(_p(w), Cin, Cout, _p(wt), _stream()), "swe_nn_flip")
    return wt


def wgrad(x, dy, dw, db=None, work=None):
    """dw [Cout,9,Cin] float = sum_p dy x^T; db [Cout] float (optional)."""
    import torch

    N, H, W, Cin = x.shape
    Cout = dy.shape[3]
    need = load().swe_nn_wgrad_work(N, H, W, Cin, Cout)
    if work is None or work.numel() < need:
        work = torch.empty(need, device=x.device, dtype=torch.float32)
    _ck(load().swe_nn_wgrad(_p(x), _p(dy), N, H, W, Cin, Cout, _p(dw), _p(db), _p(work), work.numel(), _stream()),
        "swe_nn_wgrad")
    return work


def relu_grad(g32, y, g16):
    _ck(load().swe_nn_relu_grad(_p(g32), _p(y), _p(g16), g32.numel(), _stream()), "swe_nn_relu_grad")
    return g16


def relu_mask(g16, y):
    _ck(load().swe_nn_relu_mask(_p(g16), _p(y), g16.numel(), _stream()), "swe_nn_relu_mask")
    return g16


def add(a, b, out):
    _ck(load().swe_nn_add(_p(a), _p(b), _p(out), a.numel(), _stream()), "swe_nn_add")
    return out


def pool2(x, y):
    N, H, W, C = x.shape
    _ck(load().swe_nn_pool2(_p(x), _p(y), N, H, W, C, _stream()), "swe_nn_pool2")
    return y


def pool2_grad(dy, dx):
    N, H, W, C = dx.shape
    _ck(load().swe_nn_pool2_grad(_p(dy), _p(dx), N, H, W, C, _stream()), "swe_nn_pool2_grad")
    return dx


def up2(x, y):
    N, h, w, C = x.shape
    _ck(load().swe_nn_up2(_p(x), _p(y), N, h, w, C, _stream()), "swe_nn_up2")
    return y


def up2_grad(dy, dx):
    N, h, w, C = dx.shape
    _ck(load().swe_nn_up2_grad(_p(dy), _p(dx), N, h, w, C, _stream()), "swe_nn_up2_grad")
    return dx


def up2_add(x, skip, out):
    """out [N,2h,2w,C] = up2(x [N,h,w,C]) + skip (bf16)."""
    N, h, w, C = x.shape
    _ck(load().swe_nn_up2_add(_p(x), _p(skip), _p(out), N, h, w, C, _stream()), "swe_nn_up2_add")
    return out


def pool2_grad_add(dy, skip, mask, dx):
    """dx [N,H,W,C] = pool2_grad(dy) + skip, zero where mask <= 0 (mask optional)."""
    N, H, W, C = dx.shape
    _ck(load().swe_nn_pool2_grad_add(_p(dy), _p(skip), _p(mask), _p(dx), N, H, W, C, _stream()),
        "swe_nn_pool2_grad_add")
    return dx


def up2_grad_mask(dy, mask, dx):
    """dx [N,h,w,C] = up2_grad(dy), zero where mask <= 0 (mask optional)."""
    N, h, w, C = dx.shape
    _ck(load().swe_nn_up2_grad_mask(_p(dy), _p(mask), _p(dx), N, h, w, C, _stream()), "swe_nn_up2_grad_mask")
    return dx


def mse(pred, target, grad, part, grad16=None):
    """pred [npix, C] float, target [npix, used] float -> loss partials; the
    gradient by pred into grad [npix, C] float and/or grad16 (bf16)."""
    npix, C = pred.shape[0], pred.shape[-1]
    used = target.shape[-1]
    _ck(load().swe_nn_mse(_p(pred), _p(target), _p(grad), _p(grad16), _p(part), npix, C, used, _stream()),
        "swe_nn_mse")
    return part


def adam(w, w16, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    _ck(load().swe_nn_adam(_p(w), _p(w16), _p(g), _p(m), _p(v), w.numel(), lr, b1, b2, eps, int(step), _stream()),
        "swe_nn_adam")


def adam_table(segs, device):
    """Device table of swe_nn_adam_multi segments from (w, w16, wt16, g, m, v,
    cin, cout) tuples of device tensors (w16, wt16 may be None); valid while
    those tensors live.  Returns (table, nseg, total)."""
    import torch

    arr = (AdamSeg * len(segs))()
    off = 0
    for k, (w, w16, wt16, g, m, v, cin, cout) in enumerate(segs):
        n = w.numel()
        if g.numel() != n or m.numel() != n or v.numel() != n:
            raise ValueError("segment %d: w, g, m, v must have the same size" % k)
        if wt16 is not None and (wt16.numel() != n or cin * 9 * cout != n):
            raise ValueError("segment %d: wt16 needs [cout, 9, cin] weights" % k)
        arr[k] = AdamSeg(w.data_ptr(), None if w16 is None else w16.data_ptr(),
                         None if wt16 is None else wt16.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(),
                         off, n, int(cin), int(cout))
        off += n
    raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
    return raw.to(device), len(segs), off


def adam_multi(table, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    buf, nseg, total = table
    _ck(load().swe_nn_adam_multi(_p(buf), nseg, total, lr, b1, b2, eps, int(step), _stream()), "swe_nn_adam_multi")


def adam_multi_dev(table, hyper, b1=0.9, b2=0.999, eps=1e-8):
    """adam_multi with (lr, c1, c2) read from the device tensor `hyper`."""
    buf, nseg, total = table
    _ck(load().swe_nn_adam_multi_dev(_p(buf), nseg, total, _p(hyper), b1, b2, eps, _stream()),
        "swe_nn_adam_multi_dev")


def adam_hyper(lr, step, b1=0.9, b2=0.999):
    """(lr, 1 - b1^step, 1 - b2^step) as the library computes them (float32)."""
    out = (ctypes.c_float * 3)()
    load().swe_nn_adam_hyper(lr, b1, b2, int(step), ctypes.cast(out, ctypes.c_void_p))
    return [out[0], out[1], out[2]]


def project_mass(d, scale, target, master, work):
    """master (int64, n) += round(d * scale * 2^32) corrected to sum exactly `target`."""
    _ck(load().swe_nn_project_mass(_p(d), d.numel(), float(scale), int(target), _p(master), _p(work), _stream()),
        "swe_nn_project_mass")


def features(h, s, mean, az, h_scale, s_scale, g_scale, l_scale):
    """h, s [N,H,W] float (s may be None), mean, az [N] float -> x [N,H,W,8] bf16."""
    import torch

    N, H, W = h.shape
    x = torch.empty((N, H, W, 8), device=h.device, dtype=torch.bfloat16)
    _ck(load().swe_nn_features(_p(h), _p(s), _p(mean), _p(az), N, H, W, float(h_scale), float(s_scale),
                               float(g_scale), float(l_scale), _p(x), _stream()), "swe_nn_features")
    return x
