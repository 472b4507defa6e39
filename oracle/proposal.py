"""ORACLE -- test infrastructure only.

FP64 CPU model of the learned light-selection proposal (the network of
chakumrender_amd/csrc/chk_mlp.hip): frequency encoding, 64-64-64-64-33 ReLU
MLP, defensive softmax and mixing weight, and its backward pass.  With
emulate_bf16 the model rounds exactly where the GPU kernels hold BF16 values
(encoded inputs, weights, hidden activations, output deltas, hidden deltas),
so GPU gradients agree with it up to FP32 accumulation order.  The backward
formulas themselves are pinned by finite differences of the pure FP64 model
(tests/test_proposal.py).
"""
from __future__ import annotations

import numpy as np

N_PARAMS = 3 * (64 * 64 + 64) + 64 * 33 + 33
EPS_Q = 0.1
EPS_C = 0.1


def bf16(x):
    """Round to bfloat16 (nearest even) via float32, as the kernels' f2bf."""
    a = np.asarray(x, np.float64).astype(np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params):
    p = np.asarray(params, np.float64)
    out, o = [], 0
    for l in range(4):
        O = 64 if l < 3 else 33
        W = p[o:o + 64 * O].reshape(64, O)
        o += 64 * O
        b = p[o:o + O]
        o += O
        out.append((W, b))
    assert o == N_PARAMS
    return out


def encode(s):
    """7 inputs in [0,1] -> 64 features: per input s, sin(2^k pi s) and
    cos(2^k pi s) for k = 0..3 after s itself; feature 63 is 1."""
    s = np.asarray(s, np.float64)
    k = s.shape[0]
    f = np.zeros((k, 64))
    for i in range(7):
        x = s[:, i]
        f[:, 9 * i] = x
        for j in range(4):
            f[:, 9 * i + 1 + j] = np.sin(2 ** j * np.pi * x)
            f[:, 9 * i + 5 + j] = np.cos(2 ** j * np.pi * x)
    f[:, 63] = 1.0
    return f


def forward(params, s, n_lights, emulate_bf16=True):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    # the kernels add FP32 biases; the pure FP64 model keeps them exact
    rb = (lambda b: b.astype(np.float32)) if emulate_bf16 else (lambda b: b)
    layers = split(params)
    x = r(encode(s))
    acts = [x]
    h = x
    for l in range(3):
        W, b = layers[l]
        h = r(np.maximum(h @ r(W) + rb(b), 0.0))
        acts.append(h)
    W, b = layers[3]
    z = h @ r(W) + rb(b)
    L = n_lights
    zl = z[:, :L]
    e = np.exp(zl - zl.max(1, keepdims=True))
    q = e / e.sum(1, keepdims=True)
    qp = (1 - EPS_Q) * q + EPS_Q / L
    c = EPS_C + (1 - 2 * EPS_C) / (1 + np.exp(-z[:, 32]))
    return dict(z=z, q=q, qp=qp, c=c, acts=acts)


def backward(params, s, dlogits, n_lights, emulate_bf16=True):
    """Parameter gradient (flat, N_PARAMS) for upstream logit gradients."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = split(params)
    fw = forward(params, s, n_lights, emulate_bf16)
    acts = fw["acts"]
    d = np.array(dlogits, np.float64)
    d[:, n_lights:32] = 0.0
    d = r(d)
    grads = [None] * 4
    for l in (3, 2, 1, 0):
        W, _ = layers[l]
        grads[l] = (acts[l].T @ d, d.sum(0))
        if l > 0:
            dh = d @ r(W).T
            d = r(dh * (acts[l] > 0))
    flat = []
    for l in range(4):
        flat += [grads[l][0].ravel(), grads[l][1]]
    return np.concatenate(flat)


def init_params(seed):
    """The library's initial parameters for a given seed (chk_proposal_create):
    He-uniform hidden layers from splitmix64, zero output layer."""
    st = (seed ^ 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF

    def nxt():
        nonlocal st
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11) * (1.0 / 9007199254740992.0)

    p = np.zeros(N_PARAMS, np.float32)
    lim = np.sqrt(6.0 / 64.0)
    for l in range(3):
        base = l * (64 * 64 + 64)
        for i in range(64 * 64):
            p[base + i] = np.float32((2.0 * nxt() - 1.0) * lim)
    return p
