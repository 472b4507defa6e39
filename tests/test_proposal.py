"""The learned proposal network: the FP64 CPU model's backward pass is pinned
by finite differences; the GPU's BF16-MFMA forward and backward kernels are
checked against that model (with the same BF16 rounding points)."""
import numpy as np
import pytest

from oracle import proposal as P


def random_params(rng, scale=1.0):
    p = P.init_params(7).astype(np.float64)
    # non-zero output layer so every path through the network is exercised
    o = 3 * (64 * 64 + 64)
    p[o:o + 64 * 33] = rng.normal(scale=0.2 * scale, size=64 * 33)
    p[o + 64 * 33:] = rng.normal(scale=0.1, size=33)
    for l in range(3):
        b0 = l * (64 * 64 + 64) + 64 * 64
        p[b0:b0 + 64] = rng.normal(scale=0.05, size=64)
    return p.astype(np.float32)


def inputs(rng, k):
    return rng.uniform(0, 1, (k, 7)).astype(np.float32)


def test_encoding_known_values():
    s = np.array([[0.5, 0.25, 0, 1, 0.125, 0.75, 0.3]])
    f = P.encode(s)
    assert f.shape == (1, 64) and f[0, 63] == 1.0
    assert f[0, 0] == 0.5
    assert abs(f[0, 1] - 1.0) < 1e-15          # sin(pi/2)
    assert abs(f[0, 5] - 0.0) < 1e-15          # cos(pi/2)
    assert abs(f[0, 9 + 2] - 1.0) < 1e-15      # sin(2 pi 0.25)


def test_cpu_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = random_params(rng).astype(np.float64)
    s = inputs(rng, 6)
    dl = rng.normal(size=(6, 33))
    L = 7
    dl[:, L:32] = 0.0
    g = P.backward(p, s, dl, L, emulate_bf16=False)

    def loss(pp):
        z = P.forward(pp, s, L, emulate_bf16=False)["z"]
        return float((z * dl).sum())

    idx = np.concatenate([rng.choice(P.N_PARAMS, 300, replace=False),
                          np.arange(3 * (64 * 64 + 64) + 64 * 33, P.N_PARAMS)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss(p + e) - loss(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_init_is_uniform_pmf_and_half_mixing():
    p = P.init_params(3)
    fw = P.forward(p, inputs(np.random.default_rng(1), 5), 8)
    np.testing.assert_allclose(fw["qp"], 1 / 8, rtol=1e-12)
    np.testing.assert_allclose(fw["c"], 0.5, rtol=1e-12)


# ------------------------------------------------------------------ GPU
gpu = pytest.mark.gpu


@gpu
def test_gpu_eval_rejects_misaligned_probs(ctx):
    """Rows of q' are written with 16-byte stores: a probs pointer off that
    alignment is refused on the host, before any launch."""
    import ctypes as C

    from chakumrender_amd import abi, host

    pr = host.Proposal(ctx, 4, seed=1)
    k = 16
    s = abi.DevBuf.from_array(np.zeros((k, 7), np.float32))
    probs, mix = abi.DevBuf(128 * k + 16), abi.DevBuf(4 * k)
    off = C.c_void_p(probs.ptr.value + 4)
    st = pr.L.chk_proposal_eval(pr.h, s.ptr, k, off, mix.ptr, None, None)
    assert st == 1 and b"16-byte aligned" in pr.L.chk_last_error()
    assert pr.L.chk_proposal_eval(pr.h, s.ptr, k, probs.ptr, mix.ptr, None, None) == 0


@gpu
def test_gpu_init_params_match_cpu(ctx):
    from chakumrender_amd import host

    pr = host.Proposal(ctx, 12, seed=3)
    np.testing.assert_array_equal(pr.get_params(), P.init_params(3))


@gpu
@pytest.mark.parametrize("n_lights", [1, 8, 32])
def test_gpu_forward_matches_cpu_model(ctx, n_lights):
    from chakumrender_amd import host

    rng = np.random.default_rng(n_lights)
    pr = host.Proposal(ctx, n_lights, seed=5)
    p = random_params(rng)
    pr.set_params(p)
    s = inputs(rng, 1000)  # not a multiple of 16
    probs, mix, logits = pr.eval(s)
    fw = P.forward(p, s, n_lights)
    scale = np.abs(fw["z"]).max()
    err = np.abs(logits - fw["z"])
    # A hidden activation whose FP32 sum lies within FP32 rounding of a bf16
    # rounding midpoint may round the other way on the GPU (MFMA summation
    # order) than in the FP64 model: that point's 33 logits then move by one
    # bf16 step times the downstream weights.  Expected ~2e-5 per activation,
    # 192 activations per point: a few of 1000 points.
    row = err.max(1)
    exact = row < 1e-4 * scale
    print(f"L={n_lights}: max logit err {err.max():.2e} (scale {scale:.2f}), "
          f"points off by a bf16 flip {np.sum(~exact)}")
    assert exact.mean() > 0.98
    assert err.max() < 2e-2 * scale
    np.testing.assert_allclose(probs[exact, :n_lights], fw["qp"][exact], rtol=2e-3, atol=1e-6)
    assert np.all(probs[:, n_lights:] == 0)
    np.testing.assert_allclose(probs.sum(1), 1.0, rtol=1e-5)
    assert probs[:, :n_lights].min() >= P.EPS_Q / n_lights * (1 - 1e-6)
    np.testing.assert_allclose(mix[exact], fw["c"][exact], rtol=2e-3)
    assert np.all((mix >= P.EPS_C - 1e-7) & (mix <= 1 - P.EPS_C + 1e-7))


@gpu
@pytest.mark.parametrize("k", [16, 1000, 4096])
def test_gpu_gradient_matches_cpu_model(ctx, k):
    from chakumrender_amd import host

    rng = np.random.default_rng(k)
    L = 20
    pr = host.Proposal(ctx, L, seed=9)
    p = random_params(rng)
    pr.set_params(p)
    s = inputs(rng, k)
    dl = rng.normal(size=(k, 33)).astype(np.float32)
    g_gpu, _ = pr.grad(s, dl)
    g_cpu = P.backward(p, s, dl, L)
    rel = np.linalg.norm(g_gpu - g_cpu) / np.linalg.norm(g_cpu)
    print(f"k={k}: |g_gpu - g_cpu| / |g_cpu| = {rel:.2e}")
    assert rel < 2e-3
    # per-block agreement (every layer's W and b)
    o = 0
    for l in range(4):
        O = 64 if l < 3 else 33
        for n in (64 * O, O):
            a, b = g_gpu[o:o + n], g_cpu[o:o + n]
            assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b) + 1e-6, (l, n)
            o += n


@gpu
def test_gpu_adam_step(ctx):
    from chakumrender_amd import host

    rng = np.random.default_rng(3)
    pr = host.Proposal(ctx, 16, seed=2)
    p0 = random_params(rng)
    pr.set_params(p0)
    s = inputs(rng, 512)
    dl = rng.normal(size=(512, 33)).astype(np.float32)
    g, gbuf = pr.grad(s, dl)
    pr.adam(gbuf, lr=1e-3)
    p1 = pr.get_params()
    # first Adam step: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
    expect = p0 - 1e-3 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=1e-5, atol=1e-7)
