"""The initial-guess network's HIP kernels against the NumPy oracle
(oracle/x0net_ref.py): convolution forward, input and weight gradients on the
matrix cores, pooling, upsampling with concatenation, loss and optimiser, and
the parameter gradients of a whole (small) U-Net.

Inputs are BF16-representable, so the FP32-accumulated outputs agree with
float64 to ~1e-5; outputs stored in BF16 agree to its resolution (~4e-3)."""
import numpy as np
import pytest

import x0net_ref as R

pytestmark = pytest.mark.gpu


def B():
    from libemmi_mgfd_amd import x0net
    return x0net.Buf


def dev_bf16(a):
    return B().of(R.to_bf16_bits(a))


def host_bf16(buf, shape):
    return R.from_bf16_bits(buf.get(np.uint16)).astype(np.float64).reshape(shape)


def dev_f32(a):
    return B().of(np.ascontiguousarray(a, np.float32))


def zeros(nbytes):
    return B()(nbytes)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


@pytest.fixture(scope="module")
def L():
    from libemmi_mgfd_amd import x0net
    return x0net._lib()


# the last two have rows whose length is not a multiple of the 16-voxel tile
SHAPES = [((4, 4, 16), 16, 16), ((4, 8, 32), 32, 48), ((2, 4, 64), 16, 112), ((3, 5, 16), 48, 16),
          ((3, 4, 12), 16, 32), ((2, 3, 24), 32, 16)]


@pytest.mark.parametrize("shape,cin,cout", SHAPES)
def test_conv_forward(L, shape, cin, cout):
    from libemmi_mgfd_amd import x0net
    rng = np.random.default_rng(cin + cout)
    x = R.bf16_round(rng.standard_normal(shape + (cin,)))
    w = R.bf16_round(rng.standard_normal((cout, 27 * cin)) * 0.1)
    b = rng.standard_normal(cout).astype(np.float32)
    dx, dw = dev_bf16(x), dev_f32(w)
    wf = zeros(2 * cout * x0net.kpad(cin))
    wt = zeros(2 * cin * x0net.kpad(cout))
    assert L.emmi_x0_prep_weights(dw.ptr, cin, cout, wf.ptr, wt.ptr, None) == 0
    y = zeros(4 * int(np.prod(shape)) * cout)
    db = dev_f32(b)
    Z, Y, X = shape
    print(f"conv forward {shape} {cin}->{cout}", flush=True)
    assert L.emmi_x0_conv_fwd(dx.ptr, Z, Y, X, cin, wf.ptr, cout, db.ptr, 1, R.SLOPE, None, None,
                              y.ptr, cout, None) == 0
    ref = R.conv_fwd(x, w, b)
    assert rel(y.get(np.float32, ref.shape), ref) < 1e-5
    # fused bias + LeakyReLU, BF16 out
    yb = zeros(2 * int(np.prod(shape)) * cout)
    assert L.emmi_x0_conv_fwd(dx.ptr, Z, Y, X, cin, wf.ptr, cout, db.ptr, 0, R.SLOPE, None,
                              yb.ptr, None, 0, None) == 0
    assert rel(host_bf16(yb, ref.shape), R.leaky(ref)) < 8e-3


@pytest.mark.parametrize("shape,cin,cout", SHAPES)
def test_conv_input_and_weight_gradients(L, shape, cin, cout):
    from libemmi_mgfd_amd import x0net
    rng = np.random.default_rng(7 + cin)
    x = R.bf16_round(rng.standard_normal(shape + (cin,)))
    g = R.bf16_round(rng.standard_normal(shape + (cout,)))
    w = R.bf16_round(rng.standard_normal((cout, 27 * cin)) * 0.1)
    Z, Y, X = shape
    dw_ = dev_f32(w)
    wf = zeros(2 * cout * x0net.kpad(cin))
    wt = zeros(2 * cin * x0net.kpad(cout))
    assert L.emmi_x0_prep_weights(dw_.ptr, cin, cout, wf.ptr, wt.ptr, None) == 0
    dg = dev_bf16(g)
    gx = zeros(2 * int(np.prod(shape)) * cin)
    print(f"conv gradients {shape} {cin}->{cout}", flush=True)
    assert L.emmi_x0_conv_fwd(dg.ptr, Z, Y, X, cout, wt.ptr, cin, None, 3, R.SLOPE, None, gx.ptr,
                              None, 0, None) == 0
    assert rel(host_bf16(gx, shape + (cin,)), R.conv_dx(g, w, cin)) < 8e-3
    zslab = max(1, Z // 2)
    nslab = -(-Z // zslab)
    part = zeros(4 * nslab * 27 * cin * cout)
    gw = zeros(4 * cout * 27 * cin)
    gb = zeros(4 * cout)
    assert L.emmi_x0_conv_dw(dev_bf16(x).ptr, dg.ptr, Z, Y, X, cin, cout, part.ptr, zslab,
                             gw.ptr, gb.ptr, None) == 0
    assert rel(gw.get(np.float32, (cout, 27 * cin)), R.conv_dw(x, g)) < 1e-5
    assert rel(gb.get(np.float32), g.sum(axis=(0, 1, 2))) < 1e-5


def test_pool_upsample_concat(L):
    rng = np.random.default_rng(3)
    Z, Y, X, C1, C2 = 4, 4, 32, 32, 16
    a = R.bf16_round(rng.standard_normal((Z, Y, X, C2)))
    p = zeros(2 * (Z // 2) * (Y // 2) * (X // 2) * C2)
    assert L.emmi_x0_pool2(dev_bf16(a).ptr, Z, Y, X, C2, p.ptr, None) == 0
    assert rel(host_bf16(p, (Z // 2, Y // 2, X // 2, C2)), R.pool2(a)) < 8e-3
    c = R.bf16_round(rng.standard_normal((Z // 2, Y // 2, X // 2, C1)))
    out = zeros(2 * Z * Y * X * (C1 + C2))
    assert L.emmi_x0_up2_cat(dev_bf16(c).ptr, dev_bf16(a).ptr, Z, Y, X, C1, C2, out.ptr, None) == 0
    ref = np.concatenate([R.up2(c), a], axis=3)
    assert rel(host_bf16(out, ref.shape), ref) == 0.0
    g = R.bf16_round(rng.standard_normal(ref.shape))
    din = zeros(2 * (Z // 2) * (Y // 2) * (X // 2) * C1)
    prior = R.bf16_round(rng.standard_normal((Z, Y, X, C2)))
    dskip = dev_bf16(prior)
    assert L.emmi_x0_up2_cat_bwd(dev_bf16(g).ptr, Z, Y, X, C1, C2, din.ptr, dskip.ptr, 1, None) == 0
    assert rel(host_bf16(din, (Z // 2, Y // 2, X // 2, C1)), R.up2_bwd(g[..., :C1])) < 8e-3
    assert rel(host_bf16(dskip, (Z, Y, X, C2)), prior + g[..., C1:]) < 8e-3
    gp = R.bf16_round(rng.standard_normal((Z // 2, Y // 2, X // 2, C2)))
    acc = dev_bf16(prior)
    assert L.emmi_x0_pool2_bwd(dev_bf16(gp).ptr, Z, Y, X, C2, acc.ptr, 1, None) == 0
    assert rel(host_bf16(acc, (Z, Y, X, C2)), prior + R.pool2_bwd(gp)) < 8e-3


def test_mse_and_adam(L):
    rng = np.random.default_rng(4)
    nv = 1000
    out = rng.standard_normal((nv, 6)).astype(np.float32)
    tgt = rng.standard_normal((nv, 6)).astype(np.float32)
    mask = (rng.random((nv, 6)) > 0.3).astype(np.float32)
    grad = zeros(2 * nv * 16)
    part = zeros(4 * ((nv + 255) // 256))
    cnt = mask.sum()
    assert L.emmi_x0_mse(dev_f32(out).ptr, dev_f32(tgt).ptr, dev_f32(mask).ptr, nv, 6,
                         1.0 / cnt, grad.ptr, part.ptr, None) == 0
    d = (out - tgt) * mask
    got = float(part.get(np.float32).astype(np.float64).sum())
    assert abs(got - float((d ** 2).sum())) <= 1e-4 * float((d ** 2).sum())
    gref = np.zeros((nv, 16))
    gref[:, :6] = 2 * d / cnt
    assert rel(host_bf16(grad, (nv, 16)), gref) < 8e-3
    p = rng.standard_normal(50).astype(np.float32)
    g = rng.standard_normal(50).astype(np.float32)
    m1 = np.zeros(50, np.float32)
    m2 = np.zeros(50, np.float32)
    tp, tg, tm1, tm2 = dev_f32(p), dev_f32(g), dev_f32(m1), dev_f32(m2)
    assert L.emmi_x0_adam(tp.ptr, tg.ptr, tm1.ptr, tm2.ptr, 50, 1e-2, 0.9, 0.999, 1e-8, 1,
                          1e3, None) == 0
    a = 0.1 * g
    b = 0.001 * g * g
    ref = p - 1e-2 * (a / 0.1) / (np.sqrt(b / 0.001) + 1e-8)
    assert rel(tp.get(np.float32), ref) < 1e-6


@pytest.mark.parametrize("shape", [(8, 8, 64), (8, 8, 24)])
def test_unet_gradients_match_oracle(shape):
    """Forward output and every parameter gradient of a 3-level U-Net
    (16/32/48 channels) on an 8x8x64 grid, and on 8x8x24 whose rows end in
    partial tiles, against the float64 oracle run on
    the same BF16 weights and inputs.  With the oracle rounding to BF16 where
    the kernels store BF16, the remaining differences are FP32 accumulation
    order and the rounding flips it causes, which accumulate along the
    backward chain: at most 3% relative L2 per tensor (the first layer's
    weight gradient, last in the chain), against O(1) for an indexing error.
    Without that emulation the plain float64 network differs by up to 6%."""
    from libemmi_mgfd_amd import x0net
    levels = (16, 32, 48)
    net = x0net.UNet(shape, seed=5, levels=levels)
    rng = np.random.default_rng(6)
    # the output layer starts at zero in training; give it weights here so
    # that every upstream gradient is exercised
    wo = (rng.standard_normal((16, 27 * levels[0])) * 0.05).astype(np.float32)
    wo[x0net.COUT_STORE:] = 0
    net.out.set(wo, np.zeros(16, np.float32))
    cs = x0net.COUT_STORE
    feat = R.bf16_round(rng.standard_normal(shape + (x0net.CIN,)))
    tgt = rng.standard_normal(shape + (cs,)).astype(np.float32)
    mask = np.ones(shape + (cs,), np.float32)
    print("unet forward", flush=True)
    y = net.forward(dev_bf16(feat))
    loss = net.loss_and_grad(dev_f32(tgt), dev_f32(mask), mask.sum())
    net.backward()
    params = []
    for cv in net.convs:
        w, b = cv.get()
        params.append((R.bf16_round(w), b.astype(np.float64)))
    yref, tape = R.unet_forward(feat, params, levels, rnd=R.bf16_round)
    yref6 = yref[..., :cs]
    assert rel(y.get(np.float32, yref6.shape), yref6) < 0.015
    d = yref6 - tgt
    assert abs(loss - float((d ** 2).mean())) <= 0.015 * float((d ** 2).mean())
    gy = np.zeros(yref.shape)
    gy[..., :cs] = R.bf16_round(2 * d / d.size)
    grads = R.unet_backward(gy, params, levels, tape, rnd=R.bf16_round)
    plain, ptape = R.unet_forward(feat, params, levels)
    pgrads = R.unet_backward(gy, params, levels, ptape)
    worst = 0.0
    for i, (cv, (gw, gb)) in enumerate(zip(net.convs, grads)):
        ew = rel(cv.gw.get(np.float32, gw.shape), gw)
        eb = rel(cv.gb.get(np.float32), gb)
        ep = rel(cv.gw.get(np.float32, gw.shape), pgrads[i][0])
        print(f"  conv {i}: {cv.cin}->{cv.cout} dw rel {ew:.2e} db rel {eb:.2e} (float64 network: {ep:.2e})")
        worst = max(worst, ew, eb)
        assert ep < 0.06
        assert np.linalg.norm(gw) > 0
    assert worst < 0.03


def test_unet_training_reduces_loss():
    """Adam steps on one sample lower the loss (the whole training loop runs
    on the project's kernels).  The rate is 1e-3: at 3e-3 this single-sample
    loss oscillates (0.40 -> 0.45 -> 0.38 over 60 steps), and whether step 60
    lands below 0.8 of the start depends on the rounding order of the
    gradient sums."""
    from libemmi_mgfd_amd import x0net
    shape = (8, 8, 64)
    net = x0net.UNet(shape, seed=1, levels=(16, 32, 48))
    rng = np.random.default_rng(2)
    feat = dev_bf16(R.bf16_round(rng.standard_normal(shape + (x0net.CIN,))))
    cs = x0net.COUT_STORE
    tgt = dev_f32(np.tanh(rng.standard_normal(shape + (cs,))))
    m = np.ones(shape + (cs,), np.float32)
    mask = dev_f32(m)
    losses = []
    for it in range(60):
        net.forward(feat)
        losses.append(net.loss_and_grad(tgt, mask, m.sum()))
        net.backward()
        net.adam(lr=1e-3)
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.8 * losses[0]


def test_config4_learned_guess_gain_reproduces_report():
    """What the trained network gains, on the config-4 test earths: the first
    16 test seeds (2000-2015, each a random earth on its own 96^3 multigrid)
    re-evaluated with the committed weights (libemmi_mgfd_amd/
    x0net_weights_96.npz) give the committed report's cycle counts exactly
    (zero guess; one cycle then a solve from x1; one cycle, the network, then
    a solve from its guess).  The report's summary over all 200 test
    seeds is printed: it is what the network saves (a negative number means
    it costs cycles)."""
    import json
    import os
    from libemmi_mgfd_amd import configs, x0eval, x0net
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    rep = json.load(open(os.path.join(here, "libemmi_mgfd_amd", "x0net_eval_96.json")))
    d = np.load(os.path.join(here, "libemmi_mgfd_amd", "x0net_weights_96.npz"))
    net = x0net.UNet((96, 96, 96), levels=tuple(int(v) for v in d["levels"]))
    net.load(d)
    seeds = list(configs.CONFIG4_TEST)[:16]
    print("config4 test seeds 2000-2015 on 96^3, reconstructions", rep["recons"], flush=True)
    now = x0eval.evaluate(net, seeds, chunk=16, recons=tuple(rep["recons"]), first_cycle=rep["first_cycle"],
                          log=lambda m: print(m, flush=True))
    keys = ["cycles_zero", "cycles_x1", "status"] + [f"{k}_{rc}" for rc in rep["recons"]
                                                     for k in ("cycles_learned", "rejected")]
    for a, b in zip(now["rows"], rep["rows"][:16]):
        assert a["seed"] == b["seed"]
        for k in keys:
            assert a[k] == b[k], (a["seed"], k, a[k], b[k])
        if rep["first_cycle"] == 1:
            # continuing the pattern sequence after x1's cycle reproduces the
            # zero-guess solve exactly
            assert a["cycles_x1"] == a["cycles_zero"]
    s = rep["summary"]
    print(json.dumps({k: s[k] for k in s if k != "wall_s"}), flush=True)
    assert s["models"] == 200 and s["all_converged"]
