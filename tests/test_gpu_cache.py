"""The learned radiance cache: BF16 MFMA inference and training kernels against a
NumPy FP32 reference of the same network (weights and inter-layer activations
rounded to BF16 as the kernels store them, FP32 accumulation), and the
progressive renderer built on them."""
import inspect

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16(x):
    """float32 values rounded to BF16 (round to nearest even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bits(x):
    return (bf16(x).view(np.uint32) >> 16).astype(np.uint16)


def mm(a, b):
    """a @ b for BF16-valued a, b as the MFMA sums them: the products of each group
    of 8 consecutive K values are summed exactly (BF16 products are exact in
    FP32), rounded to FP32 and added to the FP32 accumulator in K order.  Another
    FP32 order (BLAS's) moves a pre-activation that sits within rounding of a
    BF16 rounding boundary or of zero; the BF16-rounded activations and the ReLU
    mask then differ and the gradient differences grow layer by layer (measured
    up to 9e-3 in the first layer's gradient on such inputs), while this order
    reproduces the kernels to ~1e-7."""
    acc = np.zeros((a.shape[0], b.shape[1]), np.float32)
    for k0 in range(0, a.shape[1], 8):
        part = (a[:, k0:k0 + 8].astype(np.float64) @ b[k0:k0 + 8].astype(np.float64)).astype(np.float32)
        acc = (acc + part).astype(np.float32)
    return acc


def forward(W, x):
    Wb = bf16(W)
    acts = [x]
    h = x
    for ly in range(4):
        h = bf16(np.maximum(mm(h, Wb[ly]), 0.0))
        acts.append(h)
    return mm(h, Wb[4][:, :3]), acts


def backward(W, x, t):
    """Loss and weight gradient of the mean relative-L2 loss, with the output and
    layer gradients rounded to BF16 as the kernels store them."""
    out, acts = forward(W, x)
    n = len(x)
    den = out * out + 0.01
    loss = float(np.sum((out - t) ** 2 / den) / n)
    Wb = bf16(W)
    g = bf16(2.0 * (out - t) / den / n)
    dW = np.zeros_like(W)
    dW[4][:, :3] = mm(acts[4].T, g)
    gin = bf16(mm(g, Wb[4][:, :3].T) * (acts[4] > 0))
    for ly in range(3, -1, -1):
        dW[ly] = mm(acts[ly].T, gin)
        if ly > 0:
            gin = bf16(mm(gin, Wb[ly].T) * (acts[ly] > 0))
    return loss, dW


def he_weights(rng):
    a = np.sqrt(6.0 / 64.0)
    W = rng.uniform(-a, a, (5, 64, 64)).astype(np.float32)
    W[4][:, 3:] = 0.0
    return W


@pytest.fixture(scope="module")
def dev(hip):
    from tinypt_amd import scenes

    return scenes.triangle_room(64, 48, seed=2, tris=5000).to("hip")


def test_mlp_inference_matches_reference(hip, dev):
    rng = np.random.default_rng(1)
    W = he_weights(rng)
    hip.cache_weights(dev, set=W)
    for n in (1, 63, 1000, 4099):  # partial tiles and several workgroups
        x = bf16(rng.uniform(-1, 1, (n, 64)).astype(np.float32))
        got = hip.cache_eval(dev, bits(x))
        ref, _ = forward(W, x)
        scale = np.abs(ref).max() + 1e-6
        rel = np.abs(got - ref) / scale
        within = float(np.mean(rel <= 1e-5))
        print(f"n={n}: max |gpu - ref| / max |ref| = {rel.max():.2e}, within 1e-5: {within:.4f}")
        assert np.all(np.isfinite(got))
        # the inference kernel groups the K values of each 32-wide step in a
        # permuted order (its activations stay in registers), so a few
        # activations within rounding of a BF16 boundary may still round the
        # other way (one BF16 ulp, 2^-8, in one input of the next layer); an
        # operand-layout slip moves almost every output by > 1e-3
        assert within >= 0.99
        assert rel.max() < 1e-2


def test_mlp_inference_persistent_tiles(hip, dev):
    # more records than one pass of the persistent inference wavefronts takes
    # (2048 wavefronts x 32 records per CU-filling launch): each wavefront walks
    # several tiles, loading the next tile's rows while one computes; a sample of
    # the outputs, spread over the whole range, against the reference
    rng = np.random.default_rng(4)
    W = he_weights(rng)
    hip.cache_weights(dev, set=W)
    n = 300_001  # a partial last tile
    x = bf16(rng.uniform(-1, 1, (n, 64)).astype(np.float32))
    got = hip.cache_eval(dev, bits(x))
    pick = np.unique(np.concatenate([rng.integers(0, n, 1500), np.arange(n - 40, n), np.arange(0, 40)]))
    ref, _ = forward(W, x[pick])
    scale = np.abs(ref).max() + 1e-6
    rel = np.abs(got[pick] - ref) / scale
    within = float(np.mean(rel <= 1e-5))
    print(f"n={n}: {len(pick)} sampled outputs, max rel {rel.max():.2e}, within 1e-5: {within:.4f}")
    assert np.all(np.isfinite(got))
    assert within >= 0.99
    assert rel.max() < 1e-2


def test_mlp_gradient_matches_reference(hip, dev):
    rng = np.random.default_rng(2)
    W = he_weights(rng)
    hip.cache_weights(dev, set=W)
    n = 700
    x = bf16(rng.uniform(-1, 1, (n, 64)).astype(np.float32))
    t = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    loss, grad = hip.cache_train(dev, bits(x), t, 0.0)  # lr 0: the step leaves the weights unchanged
    rloss, rgrad = backward(W, x, t)
    print(f"loss gpu {loss:.6f} ref {rloss:.6f}")
    assert abs(loss - rloss) <= 1e-5 * abs(rloss)
    for ly in range(5):
        a, b = grad[ly], rgrad[ly]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        print(f"layer {ly}: |dW - ref| / |ref| = {rel:.2e}")
        # the reference sums in the MFMA's order: ~1e-7 (FP32 reassociation of the
        # per-tile gradient atomics), measured 1.5e-7 on these inputs
        assert rel < 1e-5
    assert np.all(grad[4][:, 3:] == 0.0)
    assert np.array_equal(hip.cache_weights(dev), W)  # step size 0


def test_mlp_training_fits_a_function(hip, dev):
    rng = np.random.default_rng(3)
    hip.cache_weights(dev, set=he_weights(rng))
    n = 4096
    x = bf16(rng.uniform(-1, 1, (n, 64)).astype(np.float32))
    t = np.stack([0.6 + 0.4 * np.sin(2.0 * x[:, 0] + c) * x[:, 1 + c] for c in range(3)], 1).astype(np.float32)
    losses = [hip.cache_train(dev, bits(x), t, 3e-3)[0] for _ in range(300)]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.2 * losses[0]


def test_progressive_without_cache_equals_plain_render(hip, dev):
    # frame f is sample f of the plain sampler with spp = frames (b = 1): same
    # streams, same per-pixel FP32 summation order
    img, st = hip.render_cached(dev, 8, cache_depth=-1, max_depth=6, rr_depth=3, seed=5)
    ref = hip.PathTracer("hip").render(dev, 8, max_depth=6, rr_depth=3, seed=5, hdr=True).raw()
    assert np.array_equal(img, ref)
    assert st["queries"] == 0 and st["records"] == 0


@pytest.mark.parametrize("fp64", [False, True])
def test_frame_groups_leave_plain_frames_unchanged(hip, dev, monkeypatch, fp64):
    """Plain frames traced several per pipeline wave (TPT_FRAME_GROUP) are the
    same samples, added to the sums in frame order: the image equals one frame
    per wave bit for bit, also for a frame count the group does not divide and
    for a session split over calls."""
    kw = dict(cache_depth=-1, max_depth=6, rr_depth=3, seed=5, fp64=fp64)
    out = {}
    for g in ("1", "3"):
        monkeypatch.setenv("TPT_FRAME_GROUP", g)
        img, _ = hip.render_cached(dev, 7, **kw)
        sess = hip.ProgressiveRender(dev, max_depth=6, rr_depth=3, seed=5, fp64=fp64)
        sess.add(2).add(5)
        split, n = sess.read()
        sess.close()
        assert n == 7
        out[g] = (img, split)
    assert np.array_equal(out["1"][0], out["3"][0])
    assert np.array_equal(out["1"][0], out["1"][1]) and np.array_equal(out["1"][0], out["3"][1])


@pytest.mark.parametrize("suffix_len", [0, 3])
def test_cached_render_trains_online(hip, suffix_len):
    """suffix_len 0: training paths run to max_depth; 3: they end in the cache
    three vertices past their cache vertex (self-training)."""
    from tinypt_amd import scenes

    dev = scenes.triangle_room(128, 96).to("hip")
    kw = dict(max_depth=8, rr_depth=5)
    ref = hip.PathTracer("hip").render(dev, 2048, seed=99, hdr=True, fp64=True, **kw).raw()
    plain, _ = hip.render_cached(dev, 64, cache_depth=-1, seed=1, **kw)
    # 12,288 pixels, 2048 training records per frame (about one path in six), 4 Adam steps per frame
    img, st = hip.render_cached(dev, 64, cache_depth=2, train_records=2048, learning_rate=1e-3, reset=True,
                                weight_seed=7, seed=1, suffix_len=suffix_len, train_steps=4, **kw)
    print(st)
    assert np.all(np.isfinite(img)) and img.min() >= 0.0
    assert st["queries"] > 5000 and 0 < st["records"] <= 2048
    assert st["steps"] == 64 * 4 and np.isfinite(st["last_loss"])

    def rel(x):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    print(f"relMSE vs 2048 spp: cached {rel(img):.4f}, plain {rel(plain):.4f}; "
          f"mean cached {img.mean():.4f} plain {plain.mean():.4f} ref {ref.mean():.4f}")
    # trained runs are stochastic in the record order; over weight seeds 1..10 this
    # briefly trained cache read -1.1% .. -1.9% (suffix 0) and +3.2% .. +4.6%
    # (suffix 3: the targets include the cache's own early predictions) against the
    # reference, with relMSE 0.062 .. 0.078 against 0.154 for the plain frames
    assert abs(img.mean() - ref.mean()) < (0.05 if suffix_len == 0 else 0.08) * ref.mean()
    assert rel(img) < rel(plain)  # lower error than the plain estimator at the same frames
    # the network keeps training across calls
    _, st2 = hip.render_cached(dev, 2, cache_depth=2, train_records=2048, seed=2, suffix_len=suffix_len,
                               train_steps=4, **kw)
    assert st2["steps"] == 64 * 4 + 2 * 4


def test_progressive_session_splits_equal_one_render(hip):
    # frames added over several calls continue the same streams: 3 + 5 frames
    # equal one 8-frame progressive render (and the plain 8-spp HDR render)
    from tinypt_amd import scenes

    dev = scenes.triangle_room(64, 48).to("hip")
    kw = dict(max_depth=6, rr_depth=3, seed=9)
    one, _ = hip.render_cached(dev, 8, cache_depth=-1, **kw)
    sess = hip.ProgressiveRender(dev, **kw)
    sess.add(3)  # enqueued, returns at once
    mid, n = sess.read()
    assert n == 3 and np.all(np.isfinite(mid))
    sess.add(2).add(3)
    img, n = sess.read()
    assert n == 8
    assert np.array_equal(img, one)
    plain = hip.PathTracer("hip").render(dev, 8, hdr=True, **kw).raw()
    assert np.array_equal(img, plain)
    sess.close()


def test_progressive_session_with_cache(hip):
    from tinypt_amd import scenes

    dev = scenes.triangle_room(128, 96).to("hip")
    kw = dict(max_depth=8, rr_depth=5, seed=1)
    ref = hip.PathTracer("hip").render(dev, 1024, hdr=True, fp64=True, max_depth=8, rr_depth=5, seed=99).raw()
    # 2048 records per frame in 4 Adam steps (mini-batches of 512; the default 8
    # steps suit 2^16 records, mini-batches of 8192)
    sess = hip.ProgressiveRender(dev, cache_depth=2, train_records=2048, reset=True, weight_seed=7, train_steps=4, **kw)
    for _ in range(4):
        sess.add(16)
    sess.wait()
    img, n = sess.read()
    assert n == 64 and np.all(np.isfinite(img)) and img.min() >= 0
    # training is stochastic in the record order; over weight seeds 1..10 the
    # briefly trained cache read +2.1% .. +4.3% high (mean +3.4%) with relMSE
    # 0.061 .. 0.068 against 0.152 for the plain 64 frames
    plain = hip.PathTracer("hip").render(dev, 64, hdr=True, **kw).raw()
    rel = lambda x: float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
    assert abs(img.mean() - ref.mean()) < 0.08 * ref.mean()
    assert rel(img) < 0.6 * rel(plain)
    with pytest.raises(Exception):
        sess.add(0)


@pytest.mark.parametrize("lanes", ["2", "3", "4"])
def test_cached_frames_on_several_streams(hip, monkeypatch, lanes):
    """Frames alternate between path pools on separate streams (TPT_STREAMS lanes;
    3 by default): a frame's inference waits for the previous frame's training,
    and the frames add into the shared sums in frame order.  Without the cache,
    and with the cache at learning rate 0 (fixed weights), the image equals the
    one-stream image bit for bit.  Training runs differ through the order in
    which paths append their records (it splits the mini-batches), also between
    two runs on one stream."""
    from tinypt_amd import scenes

    kw = dict(max_depth=8, rr_depth=5, seed=3)
    ck = dict(cache_depth=2, train_records=1024, reset=True, weight_seed=5, **kw)
    out = {}
    for streams in ("1", lanes):
        monkeypatch.setenv("TPT_STREAMS", streams)
        dev = scenes.triangle_room(96, 64, seed=2, tris=5000).to("hip")
        plain, _ = hip.render_cached(dev, 12, cache_depth=-1, **kw)
        sess = hip.ProgressiveRender(dev, **kw)
        sess.add(5).add(7)
        split, n = sess.read()
        sess.close()
        fixed, _ = hip.render_cached(dev, 12, learning_rate=0.0, **ck)
        cached, st = hip.render_cached(dev, 24, **ck)
        out[streams] = dict(plain=plain, split=split, fixed=fixed, cached=cached, st=st)
    one, many = out["1"], out[lanes]
    assert np.array_equal(one["plain"], many["plain"]) and np.array_equal(one["split"], many["plain"])
    assert np.array_equal(one["fixed"], many["fixed"])
    k = inspect.signature(hip.render_cached).parameters["train_steps"].default
    assert one["st"]["steps"] == many["st"]["steps"] == 24 * k
    assert one["st"]["queries"] == many["st"]["queries"]
    # trained runs are stochastic in the record order (mean |diff| / mean between
    # runs measured 0.9e-2 .. 2.5e-2), so each is held to a bias bar against a
    # converged plain render instead of to the other run's mean; briefly trained
    # on this small image the cache reads +2.8% .. +4.7% high (measured means
    # 0.1667 .. 0.1709 against 0.1633), hence 8%
    ref = hip.PathTracer("hip").render(dev, 1024, hdr=True, fp64=True, **kw).raw()
    a, b = one["cached"], many["cached"]
    diff = float(np.mean(np.abs(a - b)) / a.mean())
    print(f"cached, one vs {lanes} streams: mean |diff| / mean = {diff:.2e}; "
          f"means {a.mean():.5f} {b.mean():.5f} reference {ref.mean():.5f}")
    assert abs(a.mean() - ref.mean()) < 0.08 * ref.mean()
    assert abs(b.mean() - ref.mean()) < 0.08 * ref.mean()
    assert diff < 0.05


def test_training_batches_stay_within_the_stored_records(hip, dev):
    # 1000 records over 3 Adam steps: mini-batches of 334, 334 and the last 332
    # (shade counts candidates past the 1000 it stores; training reads only those)
    img, st = hip.render_cached(dev, 6, cache_depth=2, train_records=1000, train_steps=3, reset=True,
                                weight_seed=3, max_depth=6, seed=4)
    w = hip.cache_weights(dev)
    print(st)
    assert np.all(np.isfinite(w)) and np.all(np.isfinite(img))
    assert 0 < st["records"] <= 1000 and st["steps"] == 6 * 3 and np.isfinite(st["last_loss"])


def test_weights_read_after_session_frames(hip):
    # a session's frames train on every lane's stream; reading the weights waits
    # for all of them, so it sees the same weights as a read after wait()
    from tinypt_amd import scenes

    dev = scenes.triangle_room(96, 64, seed=2, tris=5000).to("hip")
    sess = hip.ProgressiveRender(dev, cache_depth=2, train_records=1024, reset=True, weight_seed=7, max_depth=6,
                                 seed=2)
    sess.add(16)
    w1 = hip.cache_weights(dev)
    sess.wait()
    w2 = hip.cache_weights(dev)
    assert np.array_equal(w1, w2)
    # a cached render on the same scene trains after the session's steps
    _, st = hip.render_cached(dev, 4, cache_depth=2, train_records=1024, max_depth=6, seed=3)
    k = inspect.signature(hip.render_cached).parameters["train_steps"].default
    assert k == inspect.signature(hip.ProgressiveRender).parameters["train_steps"].default
    assert st["steps"] == 16 * k + 4 * k
    sess.close()
