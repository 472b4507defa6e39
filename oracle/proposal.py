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


# ------------------------------------------------------------ online-training loss
# Restatement of the learned render's training rows (chk_mlp.hip k_train_rows,
# k_train_norm, k_train_scale; no reference counterpart: the reference gives
# every light the same sample count, mis.h:9).  Records are arrays of the
# library's chk_train_record (chakumrender_amd.abi.TRAIN_RECORD_DTYPE).
NO_INDEX = 0xFFFFFFFF
TRAIN_FREEZE_MIX, TRAIN_FREEZE_LIGHTS, TRAIN_LIGHTS_CHI2 = 1, 2, 4


def qprime(z, n_lights):
    """Defensive pmf q' = (1 - eps) softmax(z[:L]) + eps / L of logits z (k, 33)."""
    zl = np.asarray(z, np.float64)[:, :n_lights]
    e = np.exp(zl - zl.max(1, keepdims=True))
    return (1 - EPS_Q) * e / e.sum(1, keepdims=True) + EPS_Q / n_lights


def mixing(z32):
    """Mixing weight c = eps + (1 - 2 eps) sigmoid(z_32)."""
    return EPS_C + (1 - 2 * EPS_C) / (1 + np.exp(-np.asarray(z32, np.float64)))


def record_terms(rec, occluded):
    """Per record: visible light-sample value g_l V, its pdfs (p_l, p_b),
    the BSDF sample's value g_b = f cos * L_e and its pdfs (p_l, p_b)."""
    rec = np.asarray(rec)
    occ = np.asarray(occluded, np.uint8)
    slot = rec["sh_slot"].astype(np.int64)
    has = slot != NO_INDEX
    vis = np.zeros(len(rec))
    vis[has] = 1.0 - occ[slot[has]]
    gl = rec["gl"].astype(np.float64) * vis
    gb = rec["fb"].astype(np.float64) * rec["le"].astype(np.float64)
    return (gl, rec["pl"].astype(np.float64), rec["pbl"].astype(np.float64), gb,
            rec["plb"].astype(np.float64), rec["pbb"].astype(np.float64))


def train_rows(rec, qs, qc, c, occluded, n_lights, flags=0):
    """Unnormalised loss gradients (k, 33) with respect to the logits at the
    current proposal (pmf qc (k, >= L), mixing weight c (k,)), and the two
    heads' per-row weights (light: the importance weight w; mixing: |row 32|).
    qs: the pmf each record's light was drawn with."""
    k = len(rec)
    L = n_lights
    rows = np.zeros((k, 33))
    wa, wh = np.zeros(k), np.zeros(k)
    qs = np.asarray(qs, np.float64)
    qc = np.asarray(qc, np.float64)
    c = np.asarray(c, np.float64)
    gl, pl, pbl, gb, plb, pbb = record_terms(rec, occluded)
    light = np.asarray(rec)["light"].astype(np.int64)
    for b in range(k):
        j = light[b]
        if j >= L:
            continue
        if not flags & TRAIN_FREEZE_LIGHTS:
            a = gl[b] / pl[b] if pl[b] > 0 else 0.0
            w = a * a * qs[b, j] / qc[b, j] if flags & TRAIN_LIGHTS_CHI2 else a
            if w > 0 and np.isfinite(w):
                q = (qc[b, :L] - EPS_Q / L) / (1 - EPS_Q)  # the softmax part
                onehot = np.zeros(L)
                onehot[j] = 1.0
                rows[b, :L] = -w * (1 - EPS_Q) * q[j] / qc[b, j] * (onehot - q)
                wa[b] = w
        if not flags & TRAIN_FREEZE_MIX:
            cb = c[b]
            pcl = cb * pl[b] + (1 - cb) * pbl[b]
            pcb = cb * plb[b] + (1 - cb) * pbb[b]
            al = gl[b] / pcl if (gl[b] > 0 and pcl > 0) else 0.0
            ab = gb[b] / pcb if (gb[b] > 0 and pcb > 0) else 0.0
            F = cb * al + (1 - cb) * ab
            dF = al - ab
            if al > 0:
                dF -= cb * al * (pl[b] - pbl[b]) / pcl
            if ab > 0:
                dF -= (1 - cb) * ab * (plb[b] - pbb[b]) / pcb
            H = 2 * F * dF
            if np.isfinite(H) and H != 0:
                sg = (cb - EPS_C) / (1 - 2 * EPS_C)
                rows[b, 32] = H * (1 - 2 * EPS_C) * sg * (1 - sg)
                wh[b] = abs(rows[b, 32])
    return rows, wa, wh


def train_dlogits(rec, qs, qc, c, occluded, n_lights, flags=0):
    """The dlogits a training step feeds to the backward pass: each head's
    rows divided by the batch's sum of that head's row weights."""
    rows, wa, wh = train_rows(rec, qs, qc, c, occluded, n_lights, flags)
    out = rows.copy()
    sa, sh = wa.sum(), wh.sum()
    out[:, :n_lights] *= (1.0 / sa) if sa > 0 else 0.0
    out[:, 32] *= (1.0 / sh) if sh > 0 else 0.0
    return out


# The objectives whose logit gradients the rows are (for the finite-difference
# checks): per record, as functions of the current logits z (k, 33).
def light_objective(z, rec, qs, occluded, n_lights, chi2=False):
    """Per record: a (-log q'_j(z)) (KL form) or a^2 q'_s,j / q'_j(z)
    (second-moment form), a = g_l V / p_l; 0 for records without a light."""
    qp = qprime(z, n_lights)
    gl, pl, _, _, _, _ = record_terms(rec, occluded)
    j = np.asarray(rec)["light"].astype(np.int64)
    ok = j < n_lights
    a = np.where(pl > 0, gl / np.where(pl > 0, pl, 1), 0.0)
    jj = np.where(ok, j, 0)
    qj = qp[np.arange(len(j)), jj]
    if chi2:
        v = a * a * np.asarray(qs, np.float64)[np.arange(len(j)), jj] / qj
    else:
        v = -a * np.log(qj)
    return np.where(ok, v, 0.0)


def mix_objective(z, rec, occluded, n_lights):
    """Per record: F(c(z_32))^2, F the vertex's two-sample estimate
    c g_l / p_c(l) + (1 - c) g_b / p_c(b); 0 for records without a light."""
    gl, pl, pbl, gb, plb, pbb = record_terms(rec, occluded)
    c = mixing(np.asarray(z, np.float64)[:, 32])
    pcl, pcb = c * pl + (1 - c) * pbl, c * plb + (1 - c) * pbb
    al = np.where((gl > 0) & (pcl > 0), gl / np.where(pcl > 0, pcl, 1), 0.0)
    ab = np.where((gb > 0) & (pcb > 0), gb / np.where(pcb > 0, pcb, 1), 0.0)
    ok = np.asarray(rec)["light"].astype(np.int64) < n_lights
    F = c * al + (1 - c) * ab
    return np.where(ok, F * F, 0.0)
