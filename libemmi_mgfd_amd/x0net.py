"""Initial-guess network: a 3D U-Net whose convolutions, pooling, upsampling,
loss and optimiser run as the HIP kernels of libemmi_mgfd.so
(csrc/x0net_kernels.hip); device memory comes from the same library, and all
kernels run on the default stream of its HIP runtime.

Architecture (all convolutions 3x3x3, zero padding 1, LeakyReLU 0.1):

    level 0  16 ch  (full multigrid resolution n3 x n2 x n1)
    level 1  32 ch  (1/2)
    level 2  48 ch  (1/4)
    level 3  64 ch  (1/8, bottleneck)
    encoder: two convolutions per level, 2x average pooling between levels
    decoder: 2x nearest upsampling, concatenation with the encoder tensor of
             the same level, two convolutions
    output:  one linear convolution to 9 channels: per component of E, a
             change of log-amplitude and a relative phasor that refine the
             field after one multigrid cycle (zero output = that field)

Inputs (32 channels, 18 used): log10 sigma11, log10 sigma33, the offset to
the source in skin depths of the source medium (3), log10 of the distance,
log10 of the cell size, a source blob, a constant, and the log-amplitude and
unit phasor of each component of the field x1 after one multigrid cycle (9).
The network learns the correction from x1 to the converged field; with the
output convolution initialised to zero it starts as the identity on x1.  Amplitude and phase
are learnt separately: the log-amplitude spans the many decades of the field
with a relative accuracy, and the phasor varies smoothly where the real and
imaginary parts change sign.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native

SLOPE = 0.1
EREF = 1e-15
LEVELS = (16, 32, 48, 64)
CIN = 32
COUT_STORE = 9
PHASOR_WEIGHT = 4.0


def _lib():
    L = native.lib()
    if getattr(L, "_x0_bound", False):
        return L
    vp, i, f, d, l = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_long
    sig = {
        "emmi_x0_kpad": (i, [i]),
        "emmi_x0_conv_fwd": (i, [vp, i, i, i, i, vp, i, vp, i, f, vp, vp, vp, i, vp]),
        "emmi_x0_conv_dw": (i, [vp, vp, i, i, i, i, i, vp, i, vp, vp, vp]),
        "emmi_x0_prep_weights": (i, [vp, i, i, vp, vp, vp]),
        "emmi_x0_pool2": (i, [vp, i, i, i, i, vp, vp]),
        "emmi_x0_pool2_bwd": (i, [vp, i, i, i, i, vp, i, vp]),
        "emmi_x0_up2_cat": (i, [vp, vp, i, i, i, i, i, vp, vp]),
        "emmi_x0_up2_cat_bwd": (i, [vp, i, i, i, i, i, vp, vp, i, vp]),
        "emmi_x0_dact": (i, [vp, vp, l, f, vp]),
        "emmi_x0_mse": (i, [vp, vp, vp, l, i, f, vp, vp, vp]),
        "emmi_x0_adam": (i, [vp, vp, vp, vp, l, f, f, f, f, i, f, vp]),
        "emmi_x0_pack_bf16": (i, [vp, l, i, i, vp, vp]),
        "emmi_x0_out_to_edges": (i, [vp, vp, i, i, i, vp, vp]),
        "emmi_x0_out_to_edges_avg": (i, [vp, vp, vp, f, i, i, i, vp, vp]),
        "emmi_x0_edges_to_target": (i, [vp, vp, i, i, i, d, f, vp, vp, vp]),
        "emmi_x0_features": (i, [vp] * 8 + [i, i, i, f, f, f, f, f, vp, d, vp, vp]),
        "emmi_x0_scale": (i, [vp, l, d, vp]),
        "emmi_x0_fill": (i, [vp, l, C.c_uint, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L._x0_bound = True
    return L


def _ok(rc, what):
    if rc != 0:
        raise native.NativeError(f"{what} failed")


class Buf:
    """Device memory of the product library (never another HIP runtime)."""

    def __init__(self, nbytes, zero=True):
        self.nbytes = int(nbytes)
        self.d = native.DeviceBuffer(max(self.nbytes, 16))
        if zero:
            self.zero()

    def zero(self):
        _ok(_lib().emmi_x0_fill(self.d.ptr, (self.nbytes + 3) // 4, 0, None), "fill")

    @property
    def ptr(self):
        return self.d.ptr

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, zero=False)
        if a.nbytes:
            b.d.upload(a) if a.nbytes == b.d.nbytes else b._upload(a)
        return b

    def _upload(self, a):
        if native.lib().emmi_memcpy_h2d(self.d.ptr, a.ctypes.data, a.nbytes):
            raise native.NativeError("host to device copy failed")

    def get(self, dtype, shape=None):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        native.lib().emmi_device_sync()
        if native.lib().emmi_memcpy_d2h(out.ctypes.data, self.d.ptr, out.nbytes):
            raise native.NativeError("device to host copy failed")
        return out.reshape(shape) if shape is not None else out


def bf16_bits(a):
    """float array -> BF16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_values(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def kpad(cin):
    return (27 * cin + 31) // 32 * 32


class Conv:
    """One 3x3x3 convolution: FP32 master weights/bias, their gradients and
    Adam moments, BF16 forward and input-gradient weight copies."""

    def __init__(self, cin, cout, rng, linear=False, cout_store=None):
        self.cin, self.cout, self.linear = cin, cout, linear
        self.cout_store = cout_store if cout_store is not None else cout
        fan_in = 27 * cin
        std = np.sqrt(2.0 / (fan_in * (1 + SLOPE ** 2))) if not linear else np.sqrt(1.0 / fan_in)
        w = rng.standard_normal((cout, 27 * cin)).astype(np.float32) * np.float32(std)
        if cout_store is not None:
            w[cout_store:] = 0.0   # channels beyond cout_store stay zero (zero gradients)
        if linear:
            w[:] = 0.0             # the output layer starts as "no correction"
        self.nw = w.size
        self.w = Buf.of(w)
        self.b = Buf(4 * cout)
        self.gw = Buf(4 * self.nw)
        self.gb = Buf(4 * cout)
        self.m = [Buf(4 * self.nw), Buf(4 * self.nw), Buf(4 * cout), Buf(4 * cout)]
        self.wf = Buf(2 * cout * kpad(cin))
        self.wt = Buf(2 * cin * kpad(cout))
        self.prep()

    def prep(self):
        _ok(_lib().emmi_x0_prep_weights(self.w.ptr, self.cin, self.cout, self.wf.ptr, self.wt.ptr, None),
            "prep_weights")

    def get(self):
        return (self.w.get(np.float32, (self.cout, 27 * self.cin)), self.b.get(np.float32))

    def set(self, w, b):
        self.w._upload(np.ascontiguousarray(w, np.float32))
        self.b._upload(np.ascontiguousarray(b, np.float32))
        self.prep()


class UNet:
    def __init__(self, shape, seed=0, levels=LEVELS):
        """shape = (Z, Y, X) of the level-0 grid; each must be divisible by
        2^(len(levels)-1)."""
        self.shape = tuple(int(v) for v in shape)
        self.levels = tuple(levels)
        nl = len(levels)
        for v in self.shape:
            if v % (2 ** (nl - 1)):
                raise ValueError("grid not divisible by the pooling depth")
        rng = np.random.default_rng(seed)
        self.enc = []
        cin = CIN
        for c in levels:
            self.enc.append((Conv(cin, c, rng), Conv(c, c, rng)))
            cin = c
        self.dec = []
        for lev in range(nl - 2, -1, -1):
            c_up, c_skip = levels[lev + 1], levels[lev]
            self.dec.append((Conv(c_up + c_skip, c_skip, rng), Conv(c_skip, c_skip, rng)))
        self.out = Conv(levels[0], 16, rng, linear=True, cout_store=COUT_STORE)
        self.convs = [c for pair in self.enc for c in pair] + [c for pair in self.dec for c in pair] + [self.out]
        self.step_count = 0
        self._alloc()

    # -- buffers -------------------------------------------------------------
    def dims(self, lev):
        Z, Y, X = self.shape
        return Z >> lev, Y >> lev, X >> lev

    def _alloc(self):
        nl = len(self.levels)

        def bf(lev, c):
            Z, Y, X = self.dims(lev)
            return Buf(2 * Z * Y * X * c)

        self.a_enc = [(bf(l, c), bf(l, c)) for l, c in enumerate(self.levels)]
        self.pooled = [bf(l + 1, self.levels[l]) for l in range(nl - 1)]
        self.cat = []
        self.a_dec = []
        for i, lev in enumerate(range(nl - 2, -1, -1)):
            c_up, c_skip = self.levels[lev + 1], self.levels[lev]
            self.cat.append(bf(lev, c_up + c_skip))
            self.a_dec.append((bf(lev, c_skip), bf(lev, c_skip)))
        Z, Y, X = self.dims(0)
        self.nv = Z * Y * X
        self.y = Buf(4 * self.nv * COUT_STORE)
        # gradients (w.r.t. post-activation values): three buffers per level,
        # each wide enough for the widest tensor of that level
        self.cap = []
        for l in range(nl):
            c = self.levels[l] + (self.levels[l + 1] if l + 1 < nl else 0)
            self.cap.append(max(c, self.levels[l], 16))
        self.gA = [bf(l, self.cap[l]) for l in range(nl)]
        self.gB = [bf(l, self.cap[l]) for l in range(nl)]
        self.gC = [bf(l, self.cap[l]) for l in range(nl)]
        self.g_skip = [bf(l, self.levels[l]) for l in range(nl)]
        self.g_out = bf(0, 16)
        maxp = max(27 * cv.cin * cv.cout for cv in self.convs)
        self.part = Buf(4 * 32 * maxp)
        self.nlp = (self.nv + 255) // 256
        self.loss_part = Buf(4 * self.nlp)

    # -- kernels -------------------------------------------------------------
    def _conv(self, cv, lev, x, out_bf=None, out_f=None, epi=0, act=None):
        Z, Y, X = self.dims(lev)
        s = None
        _ok(_lib().emmi_x0_conv_fwd(x.ptr, Z, Y, X, cv.cin, cv.wf.ptr, cv.cout, cv.b.ptr, epi,
                                    SLOPE, act.ptr if act is not None else None,
                                    out_bf.ptr if out_bf is not None else None,
                                    out_f.ptr if out_f is not None else None, cv.cout_store, s), "conv_fwd")

    def _conv_dx(self, cv, lev, dz, dx):
        """dx [V][cin] (BF16) = input gradient of conv cv from dz [V][cout]."""
        Z, Y, X = self.dims(lev)
        _ok(_lib().emmi_x0_conv_fwd(dz.ptr, Z, Y, X, cv.cout, cv.wt.ptr, cv.cin, None, 3, SLOPE, None,
                                    dx.ptr, None, 0, None), "conv_dx")

    def _conv_dw(self, cv, lev, x, dz):
        Z, Y, X = self.dims(lev)
        zslab = max(1, Z // 16)
        _ok(_lib().emmi_x0_conv_dw(x.ptr, dz.ptr, Z, Y, X, cv.cin, cv.cout, self.part.ptr, zslab,
                                   cv.gw.ptr, cv.gb.ptr, None), "conv_dw")

    def _dact(self, g, a, n):
        _ok(_lib().emmi_x0_dact(g.ptr, a.ptr, n, SLOPE, None), "dact")

    # -- forward / backward --------------------------------------------------
    def forward(self, feat):
        """feat: BF16 [V][16] tensor; returns the FP32 output [V][6]."""
        L = _lib()
        s = None
        nl = len(self.levels)
        self.feat = feat
        x = feat
        for lev in range(nl):
            c1, c2 = self.enc[lev]
            a1, a2 = self.a_enc[lev]
            self._conv(c1, lev, x, out_bf=a1)
            self._conv(c2, lev, a1, out_bf=a2)
            if lev < nl - 1:
                Z, Y, X = self.dims(lev)
                _ok(L.emmi_x0_pool2(a2.ptr, Z, Y, X, self.levels[lev], self.pooled[lev].ptr, s), "pool")
                x = self.pooled[lev]
        up = self.a_enc[nl - 1][1]
        for i, lev in enumerate(range(nl - 2, -1, -1)):
            Z, Y, X = self.dims(lev)
            c_up, c_skip = self.levels[lev + 1], self.levels[lev]
            _ok(L.emmi_x0_up2_cat(up.ptr, self.a_enc[lev][1].ptr, Z, Y, X, c_up, c_skip,
                                  self.cat[i].ptr, s), "up2_cat")
            c1, c2 = self.dec[i]
            a1, a2 = self.a_dec[i]
            self._conv(c1, lev, self.cat[i], out_bf=a1)
            self._conv(c2, lev, a1, out_bf=a2)
            up = a2
        self._conv(self.out, 0, up, out_f=self.y, epi=1)
        return self.y

    def loss_and_grad(self, tgt, mask, count):
        """Weighted mean-square error of the last forward against tgt [V][9]
        (mask with `count` ones); writes the output gradient and returns the
        loss (synchronises)."""
        _ok(_lib().emmi_x0_mse(self.y.ptr, tgt.ptr, mask.ptr, self.nv, COUT_STORE, 1.0 / count,
                               self.g_out.ptr, self.loss_part.ptr, None), "mse")
        return float(self.loss_part.get(np.float32).astype(np.float64).sum()) / count

    def backward(self):
        """Parameter gradients of the loss of the last loss_and_grad."""
        L = _lib()
        s = None
        nl = len(self.levels)
        top = self.a_dec[-1][1] if nl > 1 else self.a_enc[0][1]
        self._conv_dw(self.out, 0, top, self.g_out)
        g = self.gA[0]
        self._conv_dx(self.out, 0, self.g_out, g)
        # decoder, finest level first; g = gradient of the level's output
        for i in range(nl - 2, -1, -1):
            lev = nl - 2 - i
            c1, c2 = self.dec[i]
            a1, a2 = self.a_dec[i]
            Z, Y, X = self.dims(lev)
            nvl = Z * Y * X
            self._dact(g, a2, nvl * c2.cout)
            self._conv_dw(c2, lev, a1, g)
            g1 = self.gB[lev]
            self._conv_dx(c2, lev, g, g1)
            self._dact(g1, a1, nvl * c1.cout)
            self._conv_dw(c1, lev, self.cat[i], g1)
            gcat = self.gC[lev]
            self._conv_dx(c1, lev, g1, gcat)
            c_up, c_skip = self.levels[lev + 1], self.levels[lev]
            gup = self.gA[lev + 1]
            _ok(L.emmi_x0_up2_cat_bwd(gcat.ptr, Z, Y, X, c_up, c_skip, gup.ptr,
                                      self.g_skip[lev].ptr, 0, s), "up2_cat_bwd")
            g = gup
        # encoder, coarsest level first; g = gradient of a_enc[lev][1]
        for lev in range(nl - 1, -1, -1):
            c1, c2 = self.enc[lev]
            a1, a2 = self.a_enc[lev]
            Z, Y, X = self.dims(lev)
            nvl = Z * Y * X
            self._dact(g, a2, nvl * c2.cout)
            self._conv_dw(c2, lev, a1, g)
            g1 = self.gB[lev]
            self._conv_dx(c2, lev, g, g1)
            self._dact(g1, a1, nvl * c1.cout)
            xin = self.feat if lev == 0 else self.pooled[lev - 1]
            self._conv_dw(c1, lev, xin, g1)
            if lev > 0:
                gin = self.gC[lev]
                self._conv_dx(c1, lev, g1, gin)
                # a_enc[lev-1][1] feeds both the pooling and the skip path
                Zf, Yf, Xf = self.dims(lev - 1)
                gprev = self.g_skip[lev - 1]
                _ok(L.emmi_x0_pool2_bwd(gin.ptr, Zf, Yf, Xf, self.levels[lev - 1], gprev.ptr, 1, s),
                    "pool_bwd")
                g = gprev

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, clip=1e3):
        L = _lib()
        s = None
        self.step_count += 1
        for cv in self.convs:
            _ok(L.emmi_x0_adam(cv.w.ptr, cv.gw.ptr, cv.m[0].ptr, cv.m[1].ptr,
                               cv.nw, lr, b1, b2, eps, self.step_count, clip, s), "adam")
            _ok(L.emmi_x0_adam(cv.b.ptr, cv.gb.ptr, cv.m[2].ptr, cv.m[3].ptr,
                               cv.cout, lr, b1, b2, eps, self.step_count, clip, s), "adam")
            cv.prep()

    # -- persistence ---------------------------------------------------------
    def state(self):
        d = {"levels": np.array(self.levels), "shape": np.array(self.shape)}
        for i, cv in enumerate(self.convs):
            d[f"w{i}"], d[f"b{i}"] = cv.get()
        return d

    def load(self, d):
        if tuple(int(v) for v in d["levels"]) != self.levels:
            raise ValueError("weights are for another architecture")
        for i, cv in enumerate(self.convs):
            cv.set(d[f"w{i}"], d[f"b{i}"])


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak


def conv_rates(shape=(96, 96, 96), steps=20, seed=0):
    """Training-step rates of the convolution kernels on random inputs of the
    given grid: per class (conv: forward and input-gradient launches,
    conv_dw: weight-gradient launches) the event-timed milliseconds per step,
    the algorithmic TFLOP/s (2 * 27 * Cin * Cout per voxel) and the fraction
    of the dense BF16 peak."""
    net = UNet(shape, seed=seed)
    rng = np.random.default_rng(seed)
    nv = int(np.prod(shape))
    feat = Buf.of(bf16_bits(rng.standard_normal((nv, CIN)).astype(np.float32) * 0.5))
    tgt = Buf.of(rng.standard_normal((nv, COUT_STORE)).astype(np.float32))
    mask = Buf.of(np.ones((nv, COUT_STORE), np.float32))

    def step():
        net.forward(feat)
        net.loss_and_grad(tgt, mask, float(nv * COUT_STORE))
        net.backward()
        net.adam(lr=1e-4)

    step()
    native.lib().emmi_device_sync()
    native.timing_enable(True)
    for _ in range(steps):
        step()
    native.lib().emmi_device_sync()
    kt = native.timing_read()
    native.timing_enable(False)
    out = {"shape": list(shape), "steps": steps}
    for k in ("conv", "conv_dw"):
        ms, n, flops, _ = kt[k]
        tf = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        out[k] = {"ms_per_step": ms / steps, "launches_per_step": n / steps, "tflops": tf,
                  "frac_of_bf16_peak": tf / MFMA_BF16_PEAK_TFLOPS}
    return out


if __name__ == "__main__":
    import json
    import sys
    shp = tuple(int(v) for v in sys.argv[1].split("x")) if len(sys.argv) > 1 else (96, 96, 96)
    print(json.dumps(conv_rates(shp)), flush=True)
