"""CPU oracle of the MLP surrogate (test infrastructure only).

Loaded by tests/ as the checker of the GPU surrogate kernels
(ftlm_extended_hubbard_development_amd/csrc/ftlm_mlp.hip); nothing in the
product imports it. The surrogate has no counterpart in the reference, so
this file restates the network the engine implements, in float64 numpy:

  features X (n x 32; 16 used) -> 4 x [Z = H W^T + b, H = GELU(Z)] (256 wide)
  -> Y = H W_4^T + b_4 (32 wide; 2 used),
  loss L = sum_{rows < n, c < 2} (Y - T)^2 / (2 n),

with its backward pass written out by hand and torch.optim.Adam's update.
``emulate=True`` rounds to bfloat16 (round to nearest even) exactly where
the device kernels store bf16: the input features, the weights, each
hidden activation after GELU, the output gradient dL/dY and each dL/dZ; all
sums are float64 (the device accumulates in fp32). ``emulate=False`` is the
plain float64 network, pinned by finite differences and torch autograd in
the tests.
"""
import math

import numpy as np

FEAT = 32
HID = 256
NHID = 4
OUTPAD = 32
SHAPES = [(HID, FEAT)] + [(HID, HID)] * (NHID - 1) + [(OUTPAD, HID)]


def offsets():
    """(weight offsets, bias offsets, total) of the flat parameter vector
    W_0, b_0, ..., W_4, b_4 (row-major out x in)."""
    wo, bo, o = [], [], 0
    for n_out, n_in in SHAPES:
        wo.append(o)
        o += n_out * n_in
        bo.append(o)
        o += n_out
    return wo, bo, o


W_OFF, B_OFF, N_PARAMS = offsets()


def bf16(x):
    """float -> float32 -> bfloat16 (round to nearest even) -> float64."""
    x32 = np.ascontiguousarray(np.asarray(x, np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(x):
    """Feature rows of input rows (U, V, t', phix, phiy, beta, mu), float32."""
    x = np.asarray(x, np.float64).reshape(-1, 7)
    U, V, tp, px, py, beta, mu = (x[:, i] for i in range(7))
    f = np.zeros((x.shape[0], FEAT), np.float64)
    f[:, 0] = (U - 7.0) / 3.0
    f[:, 1] = V - 1.0
    f[:, 2] = tp / 0.3
    f[:, 3] = np.sin(2 * math.pi * px)
    f[:, 4] = np.cos(2 * math.pi * px)
    f[:, 5] = np.sin(2 * math.pi * py)
    f[:, 6] = np.cos(2 * math.pi * py)
    f[:, 7] = (beta - 11.0) / 9.0
    f[:, 8] = (1.0 / beta - 0.275) / 0.225
    f[:, 9] = (mu - 5.0) / 10.0
    f[:, 10] = (mu - 0.5 * U) / 5.0
    f[:, 11] = beta * (mu - 0.5 * U) / 100.0
    f[:, 12] = beta * mu / 100.0
    f[:, 13] = beta * (mu - 0.5 * U - 4.0 * V) / 100.0
    f[:, 14] = (mu - 0.5 * U - 4.0 * V) / 5.0
    f[:, 15] = beta * U / 100.0
    return f.astype(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return [(p[W_OFF[l]:W_OFF[l] + n * k].reshape(n, k), p[B_OFF[l]:B_OFF[l] + n]) for l, (n, k) in enumerate(SHAPES)]


def pack(layers):
    p = np.zeros(N_PARAMS)
    for l, (W, b) in enumerate(layers):
        p[W_OFF[l]:W_OFF[l] + W.size] = W.ravel()
        p[B_OFF[l]:B_OFF[l] + b.size] = b
    return p


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, feat, emulate=True):
    """Outputs (n x 32) and the cache of the backward pass."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    layers = unpack(p)
    Wb = [r(W) for W, _ in layers]
    H = r(feat)
    Hs, Zs = [H], []
    for j in range(NHID):
        Z = H @ Wb[j].T + layers[j][1]
        Zs.append(Z)
        H = r(gelu(Z))
        Hs.append(H)
    Y = H @ Wb[NHID].T + layers[NHID][1]
    return Y, (Hs, Zs, Wb)


def loss_and_grad(p, feat, target, emulate=True):
    """Loss and flat gradient for feature rows `feat` and normalised targets
    `target` (n x 2)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Y, (Hs, Zs, Wb) = forward(p, feat, emulate)
    n = Y.shape[0]
    E = Y[:, :2] - np.asarray(target, np.float64)
    loss = 0.5 * float((E * E).sum()) / n
    dY = np.zeros_like(Y)
    dY[:, :2] = E / n
    dY = r(dY)
    grads = [None] * (NHID + 1)
    grads[NHID] = (dY.T @ Hs[NHID], dY.sum(axis=0))
    dH = dY @ Wb[NHID]
    for j in range(NHID - 1, -1, -1):
        dZ = r(dH * dgelu(Zs[j]))
        grads[j] = (dZ.T @ Hs[j], dZ.sum(axis=0))
        dH = dZ @ Wb[j]
    return loss, pack(grads)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One step of torch.optim.Adam (no weight decay, no amsgrad); t is the
    1-based step. Returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    bc1 = 1.0 - b1 ** t
    bc2 = 1.0 - b2 ** t
    denom = np.sqrt(v) / math.sqrt(bc2) + eps
    return p - (lr / bc1) * (m / denom), m, v
