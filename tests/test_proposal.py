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


# ------------------------------------------------------------ online-training loss
def random_records(rng, k, n_lights, n_slots=None):
    """Training records of the learned render's shape: a light drawn per
    record (some without one), positive values and pdfs, occlusion slots."""
    from chakumrender_amd import abi

    n_slots = n_slots or k
    rec = np.zeros(k, abi.TRAIN_RECORD_DTYPE)
    rec["light"] = rng.integers(0, n_lights, k)
    rec["light"][rng.random(k) < 0.1] = P.NO_INDEX          # no light sample
    rec["gl"] = rng.exponential(1.0, k) * (rng.random(k) < 0.8)
    rec["pl"] = rng.uniform(0.05, 3.0, k)
    rec["pbl"] = rng.uniform(0.05, 3.0, k)
    rec["fb"] = rng.exponential(0.5, k)
    rec["le"] = rng.exponential(2.0, k) * (rng.random(k) < 0.5)
    rec["plb"] = rng.uniform(0.05, 3.0, k)
    rec["pbb"] = rng.uniform(0.05, 3.0, k)
    rec["sh_slot"] = rng.permutation(n_slots)[:k]
    rec["sh_slot"][rng.random(k) < 0.1] = P.NO_INDEX
    rec["light_b"] = P.NO_INDEX
    occ = (rng.random(n_slots) < 0.3).astype(np.uint8)
    return rec, occ


def defensive_pmf(rng, k, n_lights):
    q = rng.dirichlet(np.ones(n_lights), k)
    return (1 - P.EPS_Q) * q + P.EPS_Q / n_lights


@pytest.mark.parametrize("n_lights", [32, 5])
@pytest.mark.parametrize("chi2", [False, True])
def test_training_rows_are_gradients_of_their_objectives(n_lights, chi2):
    """Each record's row is the logit gradient of its objective: the light
    head of a (-log q'_j) (KL) or a^2 q'_s,j / q'_j (second moment), the
    mixing head of F(c)^2 -- central differences of the objectives."""
    rng = np.random.default_rng(11 + n_lights + chi2)
    k = 48
    rec, occ = random_records(rng, k, n_lights)
    z = rng.normal(scale=1.5, size=(k, 33))
    qs = defensive_pmf(rng, k, n_lights)
    qc = P.qprime(z, n_lights)
    c = P.mixing(z[:, 32])
    rows, wa, wh = P.train_rows(rec, qs, qc, c, occ, n_lights,
                                flags=P.TRAIN_LIGHTS_CHI2 if chi2 else 0)
    h = 1e-6
    for col in list(range(n_lights)) + [32]:
        zp, zm = z.copy(), z.copy()
        zp[:, col] += h
        zm[:, col] -= h
        if col == 32:
            fd = (P.mix_objective(zp, rec, occ, n_lights) - P.mix_objective(zm, rec, occ, n_lights)) / (2 * h)
        else:
            fd = (P.light_objective(zp, rec, qs, occ, n_lights, chi2)
                  - P.light_objective(zm, rec, qs, occ, n_lights, chi2)) / (2 * h)
        np.testing.assert_allclose(rows[:, col], fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max() + 1e-12,
                                   err_msg=f"logit {col}")
    # rows beyond the scene's lights stay zero; weights match the rows' heads
    assert not rows[:, n_lights:32].any()
    assert np.all((wa > 0) == np.any(rows[:, :n_lights] != 0, axis=1))
    np.testing.assert_array_equal(wh, np.abs(rows[:, 32]))


@pytest.mark.parametrize("chi2", [False, True])
def test_light_rows_estimate_the_objective_over_all_lights(chi2):
    """One shading configuration with known per-light values g_j and area pdfs
    p_j, lights drawn from q'_s: the expectation over j ~ q'_s of the row is
    the logit gradient of Z KL(p* || q') with p*_j = (g_j / p_j) / Z (KL
    form), or of the light estimate's second moment sum_j g_j^2 / (q'_j
    p_j^2) (second-moment form), both by central differences."""
    from chakumrender_amd import abi

    rng = np.random.default_rng(5 + chi2)
    L = 7
    g = rng.exponential(1.0, L)
    parea = rng.uniform(0.2, 2.0, L)
    qs = defensive_pmf(rng, 1, L)[0]
    z0 = rng.normal(size=33)
    rec = np.zeros(L, abi.TRAIN_RECORD_DTYPE)
    rec["light"] = np.arange(L)
    rec["gl"] = g
    rec["pl"] = qs * parea                     # the record's light pdf q'_s,j p_j
    rec["pbl"] = rec["plb"] = rec["pbb"] = 1.0
    rec["sh_slot"] = 0
    rec["light_b"] = P.NO_INDEX
    occ = np.zeros(1, np.uint8)
    z = np.repeat(z0[None], L, 0)
    qc = P.qprime(z, L)
    rows, _, _ = P.train_rows(rec, np.repeat(qs[None], L, 0), qc, P.mixing(z[:, 32]), occ, L,
                              flags=P.TRAIN_FREEZE_MIX | (P.TRAIN_LIGHTS_CHI2 if chi2 else 0))
    expect = (qs[:, None] * rows[:, :L]).sum(0)

    def objective(zz):
        q = P.qprime(zz[None], L)[0]
        if chi2:
            return float(np.sum(g ** 2 / (q * parea ** 2)))
        v = g / parea
        Z = v.sum()
        p_star = v / Z
        return float(Z * np.sum(p_star * np.log(p_star / q)))

    h = 1e-6
    fd = np.array([(objective(z0 + h * np.eye(33)[i]) - objective(z0 - h * np.eye(33)[i])) / (2 * h)
                   for i in range(L)])
    np.testing.assert_allclose(expect, fd, rtol=1e-6, atol=1e-9)


def test_training_dlogits_normalisation():
    """Each head is divided by its batch sum of row weights: the dlogits are
    unchanged when every value (g_l and f L_e: the light values g_l and the
    emission L_e) is scaled by one constant, the
    light head's weights then sum to 1, and a batch without any light value
    gives a zero light head."""
    rng = np.random.default_rng(21)
    k, L = 64, 16
    rec, occ = random_records(rng, k, L)
    z = rng.normal(size=(k, 33))
    qs = defensive_pmf(rng, k, L)
    qc, c = P.qprime(z, L), P.mixing(z[:, 32])
    d1 = P.train_dlogits(rec, qs, qc, c, occ, L)
    r2 = rec.copy()
    for f in ("gl", "le"):
        r2[f] = r2[f] * 37.0
    d2 = P.train_dlogits(r2, qs, qc, c, occ, L)
    np.testing.assert_allclose(d2, d1, rtol=1e-6, atol=1e-12)
    _, wa, _ = P.train_rows(rec, qs, qc, c, occ, L)
    assert abs((wa / wa.sum()).sum() - 1) < 1e-12
    r3 = rec.copy()
    r3["gl"] = 0
    d3 = P.train_dlogits(r3, qs, qc, c, occ, L)
    assert not d3[:, :L].any() and d3[:, 32].any()


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, 4, 1, 2])
def test_gpu_training_loss_matches_oracle(ctx, flags):
    """chk_proposal_loss_grad (the learned render's training rows) against the
    oracle's restatement on the same records, at the q'_c and c the GPU's own
    inference gives for the records' inputs."""
    from chakumrender_amd import host

    rng = np.random.default_rng(31 + flags)
    L, k = 32, 4096
    pr = host.Proposal(ctx, L, seed=3)
    pr.set_params(random_params(rng))
    rec, occ = random_records(rng, k, L, n_slots=2 * k)
    s = inputs(rng, k)
    qs = defensive_pmf(rng, k, L).astype(np.float32)
    print(f"loss grad: {k} records, flags {flags}", flush=True)
    got = pr.loss_grad(rec, s, qs, occ, flags)
    qc, c, _ = pr.eval(s)
    want = P.train_dlogits(rec, qs, qc, c, occ, L, flags)
    scale = np.abs(want).max(0) + 1e-30
    err = np.abs(got - want) / scale
    print(f"max error relative to each logit's largest row: {err.max():.2e}")
    assert err.max() < 1e-5
    np.testing.assert_array_equal(got[:, L:32], 0)
    pr.close()
