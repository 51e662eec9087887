"""CPU restatement of the learned troubled-interface indicator
(pyclaw_amd/csrc/indicator.hip): features, labels, the 6->32->32->1 MLP
forward/backward with the kernel's BF16 rounding points, and Adam.

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker).

The indicator has no counterpart in the reference: the reference always runs the
full WENO5 (sharpclaw/reconstruct.f90:201-299).  The label is defined from that
routine's nonlinear weights (reconstruct.f90:264-284, applied per component to
the cell differences).  Parity pins: the backward pass equals torch autograd of
the same network (float64, rounding off); linear data has exactly linear WENO
weights (label 0); an isolated jump is labelled 1 on the interfaces whose
stencils straddle it.
"""
import numpy as np

from . import oracle as O

NF, NH = 16, 32
P_W1 = 0
P_W2 = P_W1 + NF * NH
P_B2 = P_W2 + NH * NH
P_W3 = P_B2 + NH
P_B3 = P_W3 + NH
P_N = P_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(params):
    p = np.asarray(params, dtype=np.float32)
    return (p[P_W1:P_W2].reshape(NF, NH), p[P_W2:P_B2].reshape(NH, NH), p[P_B2:P_W3],
            p[P_W3:P_B3], p[P_B3])


def init_params(seed=0):
    """He-style initialisation; W1 rows 7..15 (padding) are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P_N, np.float32)
    w1 = rng.normal(0, np.sqrt(2.0 / 7.0), (NF, NH)).astype(np.float32)
    w1[6] = 0.0
    w1[7:] = 0.0
    p[P_W1:P_W2] = w1.ravel()
    p[P_W2:P_B2] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    p[P_W3:P_B3] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    return p


# ---------------------------------------------------------------- features
def stencils(q, bc_lower, bc_upper, reflect=(1, 2)):
    """(interface, component, 6) float64 stencils of a planar state q (meqn, my, mx):
    x-interfaces (i = 1..mx+1, j = 1..my) then y-interfaces (i = 1..mx,
    j = 1..my+1), row-major; boundary cells from the reference BC fill."""
    meqn, my, mx = q.shape
    qbc = O.with_ghosts(np.transpose(q, (0, 2, 1)), 3)   # (meqn, mx+6, my+6)
    O.bc2(qbc, 3, bc_lower, bc_upper, reflect)
    out = []
    # x-interface (c, r): cells c-3..c+2 along x; qbc index of cell c is c+2
    for r in range(1, my + 1):
        for c in range(1, mx + 2):
            out.append(qbc[:, c - 1:c + 5, r + 2])
    for r in range(1, my + 2):
        for c in range(1, mx + 1):
            out.append(qbc[:, c + 2, r - 1:r + 5])
    return np.stack(out)                                  # (nint, meqn, 6)


def departure(v):
    """Largest |w_k - d_k| of the WENO5 weights of both edges (reconstruct.f90:
    264-284 on the cell differences of one component), v (..., 6) float64."""
    hh = np.diff(v, axis=-1)                               # (..., 5)
    worst = np.zeros(v.shape[:-1])
    eps = float(np.float32(1.0e-36))
    for im in (1, -1):
        a2, a1 = (hh[..., 0], hh[..., 1]) if im == 1 else (hh[..., 4], hh[..., 3])
        h0 = hh[..., 2]
        b1 = hh[..., 3] if im == 1 else hh[..., 1]
        t1, t2, t3 = im * (a2 - a1), im * (a1 - h0), im * (h0 - b1)
        e1, e2, e3 = a2 - 3.0 * a1, a1 + h0, 3.0 * h0 - b1
        tt1 = 13.0 * (t1 * t1) + 3.0 * (e1 * e1)
        tt2 = 13.0 * (t2 * t2) + 3.0 * (e2 * e2)
        tt3 = 13.0 * (t3 * t3) + 3.0 * (e3 * e3)
        tt1 = (eps + tt1) * (eps + tt1)
        tt2 = (eps + tt2) * (eps + tt2)
        tt3 = (eps + tt3) * (eps + tt3)
        s1, s2, s3 = tt2 * tt3, 6.0 * tt1 * tt3, 3.0 * tt1 * tt2
        tot = s1 + s2 + s3
        d = np.maximum(np.maximum(np.abs(s1 / tot - 0.1), np.abs(s2 / tot - 0.6)),
                       np.abs(s3 / tot - 0.3))
        worst = np.maximum(worst, d)
    return worst


def features(q, bc_lower, bc_upper, flat, wtol, reflect=(1, 2)):
    """Features (n, 8) float32 and labels (n,) uint8, sample = (interface, component)."""
    st = stencils(q, bc_lower, bc_upper, reflect)           # (nint, meqn, 6)
    v = st.reshape(-1, 6)
    lo, hi = v.min(axis=1), v.max(axis=1)
    rng = hi - lo
    smooth = rng <= np.float64(np.float32(flat))
    inv = 2.0 / np.where(smooth, 1.0, rng)
    f = ((v - lo[:, None]) * inv[:, None] - 1.0).astype(np.float32)
    f[smooth] = 0.0
    feat = np.zeros((v.shape[0], 8), np.float32)
    feat[:, :6] = f
    feat[:, 6] = 1.0
    rngf = rng.astype(np.float32)
    label = ((rngf > np.float32(flat)) & (departure(v) > np.float64(np.float32(wtol))))
    return feat, label.astype(np.uint8)


# ---------------------------------------------------------------- network
def forward(params, feat, quantize=True):
    """Logits and the intermediates; quantize=True rounds operands to BF16 where
    the kernel does (parameters, inputs, hidden activations)."""
    qz = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    W1, W2, b2, w3, b3 = [qz(a) for a in unpack(params)]
    X = np.zeros((feat.shape[0], NF), np.float32)
    X[:, :8] = feat[:, :8]
    X = qz(X).astype(np.float64)
    Z1 = X @ W1.astype(np.float64)
    H1 = qz(np.maximum(Z1, 0.0).astype(np.float32 if quantize else np.float64)).astype(np.float64)
    Z2 = H1 @ W2.astype(np.float64) + b2.astype(np.float64)
    z3 = np.maximum(Z2, 0.0) @ w3.astype(np.float64) + float(b3)
    return z3, (X, Z1, H1, Z2, (W1, W2, b2, w3, b3))


def loss_grad(params, feat, label, idx, pos_weight, quantize=True):
    """Weighted binary cross-entropy (mean over the batch idx) and its gradient
    in the parameter layout, with the kernel's rounding points when quantize."""
    qz = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    f = feat[idx]
    y = label[idx].astype(np.float64)
    z3, (X, Z1, H1, Z2, (W1, W2, b2, w3, b3)) = forward(params, f, quantize)
    B = len(idx)
    p = 1.0 / (1.0 + np.exp(-z3))
    pc = np.clip(p, 1e-7, 1 - 1e-7)
    loss = float(np.sum(-(pos_weight * y * np.log(pc) + (1 - y) * np.log(1 - pc))) / B)
    w = pos_weight * y + (1 - y)
    dz3 = (p * w - pos_weight * y) / B
    A2 = np.maximum(Z2, 0.0)
    g = np.zeros(P_N)
    g[P_B3] = dz3.sum()
    g[P_W3:P_B3] = A2.T @ dz3
    dZ2 = (Z2 > 0) * (dz3[:, None] * w3.astype(np.float64)[None, :])
    g[P_B2:P_W3] = dZ2.sum(axis=0)
    dZ2q = qz(dZ2.astype(np.float32) if quantize else dZ2).astype(np.float64)
    g[P_W2:P_B2] = (H1.T @ dZ2q).ravel()
    dA1 = dZ2q @ W2.astype(np.float64).T
    dZ1 = (Z1 > 0) * dA1
    dZ1q = qz(dZ1.astype(np.float32) if quantize else dZ1).astype(np.float64)
    g[P_W1:P_W2] = (X.T @ dZ1q).ravel()
    return loss, g


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in place; W1 padding rows stay 0."""
    m[:] = b1 * m + (1 - b1) * grad
    v[:] = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    params -= lr * mh / (np.sqrt(vh) + eps)
    params[P_W1 + 7 * NH:P_W2] = 0.0
    return params


def flags(q, bc_lower, bc_upper, flat, params, threshold, reflect=(1, 2)):
    """Per-interface flags: any component whose stencil range exceeds flat has a
    (quantized) logit above threshold.  Returns (flags uint8, margin = min over
    those components of |logit - threshold|)."""
    meqn = q.shape[0]
    feat, _ = features(q, bc_lower, bc_upper, flat, 1.0, reflect)
    v = stencils(q, bc_lower, bc_upper, reflect).reshape(-1, 6)
    rough = ((v.max(axis=1) - v.min(axis=1)).astype(np.float32) > np.float32(flat))
    z, _ = forward(params, feat, True)
    z = z.reshape(-1, meqn)
    rough = rough.reshape(-1, meqn)
    hit = (z > threshold) & rough
    margin = np.where(rough, np.abs(z - threshold), np.inf).min(axis=1)
    return hit.any(axis=1).astype(np.uint8), margin
