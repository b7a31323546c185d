"""numpy restatement of the learned SGS closure (test infrastructure only):
BF16 rounding, the 8 -> 64 -> 64 -> 1 network (tanh or ReLU hidden layers) with BF16 operands and
FP32 accumulation, its mean-squared-error gradient in the same rounding
steps as the device kernel, the per-cell features and the plane profiles.
The parameter layout is the one include/fsl.h documents."""
import numpy as np

NIN, NH = 8, 64
OW1, OB1 = 0, 512
OW2, OB2 = 576, 4672
OW3, OB3 = 4736, 4800
NPARAM = 4801


def bf16(x):
    """Round FP32 to BF16 (nearest even), returned as FP32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return (p[OW1:OB1].reshape(NIN, NH), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def normalise(X, norm):
    X = np.asarray(X, dtype=np.float32)
    return bf16((X - norm[:8]) * norm[8:])


TANH_K = np.float32(2.8853900817779268)   # 2 log2(e): tanh(z) = 1 - 2 / (2^(K z) + 1)

def forward(X, params, norm, exact=False, inference=False, act="tanh"):
    """Network output for raw features X (n x 8).  exact=True skips every
    BF16 rounding (FP32 network).  inference=True follows the inference
    kernels instead: BF16 operands, hidden values r = 1 / (2^u + 1) of the
    scaled arguments u (tanh = 1 - 2 r): layer 1 BF16(x) BF16(K W1) + K b1,
    layer 2 BF16(r1) BF16(-2 K W2) + K (b2 + sum_k W2[k]), output
    BF16(r2) BF16(-2 w3) + b3 + sum_k w3[k]; training uses BF16(x), BF16(W)
    and tanh.  act="relu": h = max(z, 0) rounded to BF16, the same steps in
    training and inference."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    xn = (np.asarray(X, np.float32) - norm[:8]) * norm[8:]
    if act == "relu":
        relu = lambda z: np.maximum(z, 0.0)
        r = (lambda a: np.asarray(a, np.float32)) if exact else bf16
        x = r(xn)
        h1 = r(relu(x.astype(np.float64) @ r(W1) + b1).astype(np.float32))
        h2 = r(relu(h1.astype(np.float64) @ r(W2) + b2).astype(np.float32))
        y = h2.astype(np.float64) @ r(w3) + b3
        return y.astype(np.float32), (x, h1, h2)
    if inference and not exact:
        x = bf16(xn)
        u1 = x.astype(np.float64) @ bf16(TANH_K * W1) + TANH_K * b1
        r1 = bf16((1.0 / (np.exp2(u1) + 1.0)).astype(np.float32))
        c2 = TANH_K * (b2 + W2.astype(np.float32).sum(0, dtype=np.float32))
        u2 = r1.astype(np.float64) @ bf16(np.float32(-2.0) * TANH_K * W2) + c2
        r2 = bf16((1.0 / (np.exp2(u2) + 1.0)).astype(np.float32))
        y = r2.astype(np.float64) @ bf16(np.float32(-2.0) * w3) + (b3 + w3.sum(dtype=np.float32))
        return y.astype(np.float32), (x, 1.0 - 2.0 * r1, 1.0 - 2.0 * r2)
    r = (lambda a: np.asarray(a, np.float32)) if exact else bf16
    x = r(xn)
    h1 = r(np.tanh(x.astype(np.float64) @ r(W1) + b1).astype(np.float32))
    h2 = r(np.tanh(h1.astype(np.float64) @ r(W2) + b2).astype(np.float32))
    y = h2.astype(np.float64) @ r(w3) + b3
    return y.astype(np.float32), (x, h1, h2)


def gradient(X, t, params, norm, weight=None, act="tanh"):
    """Gradient of mean(w (y - t)^2) with the device kernel's rounding steps:
    BF16 activations and back-propagated errors, FP32 master w3."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    y, (x, h1, h2) = forward(X, params, norm, act=act)
    if act == "relu":
        d2, d1 = (h2 > 0).astype(np.float64), (h1 > 0).astype(np.float64)
    else:
        d2, d1 = 1.0 - h2.astype(np.float64) ** 2, 1.0 - h1.astype(np.float64) ** 2
    n = len(t)
    err = (y - t).astype(np.float64)
    w = np.ones(n) if weight is None else np.asarray(weight, np.float64)
    dy = w * err * (2.0 / n)
    g = np.zeros(NPARAM)
    g[OW3:OB3] = h2.T.astype(np.float64) @ dy
    g[OB3] = dy.sum()
    dz2 = bf16((dy[:, None] * w3[None, :] * d2).astype(np.float32))
    g[OB2:OW3] = dz2.sum(0)
    g[OW2:OB2] = (h1.T.astype(np.float64) @ dz2).reshape(-1)
    dz1 = bf16(((dz2.astype(np.float64) @ bf16(W2).T.astype(np.float64)) * d1).astype(np.float32))
    g[OB1:OW2] = dz1.sum(0)
    g[OW1:OB1] = (x.T.astype(np.float64) @ dz1).reshape(-1)
    return g, float((w * err ** 2).sum())


def velocity_gradient(U, case):
    """du_r/dx_e by central differences over the interior, shape (L, H, D, 3, 3),
    and the second-difference Laplacian of u_r, shape (L, H, D, 3)."""
    d = (case.deltaX, case.deltaY, case.deltaZ)
    r2 = [2.0 * (1.0 / (4.0 * de)) for de in d]
    vel = U[1:4] / U[0]
    sl = (slice(1, -1),) * 3
    g = np.zeros((case.LEN, case.HIG, case.DEP, 3, 3))
    lap = np.zeros((case.LEN, case.HIG, case.DEP, 3))
    for e in range(3):
        p = list(sl); m = list(sl)
        p[e] = slice(2, None); m[e] = slice(0, -2)
        for r in range(3):
            g[..., r, e] = (vel[r][tuple(p)] - vel[r][tuple(m)]) * r2[e]
            lap[..., r] += (vel[r][tuple(p)] - 2 * vel[r][sl] + vel[r][tuple(m)]) / d[e] ** 2
    return g, lap


def features(U, case, dU, dw):
    """The 8 features of every interior cell (k fastest), float64, and |S|."""
    gr, lap = velocity_gradient(U, case)
    gr = gr.reshape(-1, 3, 3)
    lap = lap.reshape(-1, 3)
    S = 0.5 * (gr + gr.transpose(0, 2, 1))
    W = 0.5 * (gr - gr.transpose(0, 2, 1))
    s2 = (S * S).sum((1, 2))
    w2 = (W * W).sum((1, 2))
    SS = S @ S
    WW = W @ W
    s3 = np.einsum("nij,nji->n", SS, S)
    sww = np.einsum("nij,nji->n", S, WW)
    ssww = np.einsum("nij,nji->n", SS, WW)
    smag = np.sqrt(2 * s2)
    sn = np.sqrt(s2)
    om = np.stack([gr[:, 2, 1] - gr[:, 1, 2], gr[:, 0, 2] - gr[:, 2, 0], gr[:, 1, 0] - gr[:, 0, 1]], 1)
    ws = np.einsum("ni,nij,nj->n", om, S, om)
    oo = (om * om).sum(1)
    L, H, D = case.LEN, case.HIG, case.DEP
    j = np.broadcast_to(np.arange(1, H + 1)[None, :, None], (L, H, D)).reshape(-1)
    i = np.broadcast_to(np.arange(1, L + 1)[:, None, None], (L, H, D)).reshape(-1)
    y = (j - H // 2 - 0.5) * case.deltaY
    DD = (case.deltaX * case.deltaY * case.deltaZ) ** 0.666666666666667
    tiny = 1e-30
    f = np.stack([(w2 - s2) / (w2 + s2 + tiny), s3 / (s2 * sn + tiny), sww / (sn * w2 + tiny),
                  ssww / (s2 * w2 + tiny), np.sqrt((lap ** 2).sum(1)) * np.sqrt(DD) / (smag + tiny),
                  ws / (oo * smag + tiny), smag * np.sqrt(DD) / dU, y / np.asarray(dw, np.float64)[i]], 1)
    return f, smag


def plane_profiles(U):
    """z sums over the interior per (i, j): rho, rho u, rho v, rho w, (rho u)^2/rho, ..."""
    I = U[:, 1:-1, 1:-1, 1:-1]
    r = I[0]
    out = [r.sum(2), I[1].sum(2), I[2].sum(2), I[3].sum(2),
           (I[1] ** 2 / r).sum(2), (I[2] ** 2 / r).sum(2), (I[3] ** 2 / r).sum(2)]
    return np.stack(out)


def thickness(prof, dU, dy, fallback):
    """delta_w(x) per x index 0..LEN+1 from plane profiles."""
    ut = prof[1] / prof[0]
    g = np.abs(ut[:, 2:] - ut[:, :-2]) / (2 * dy)
    m = g.max(1)
    dw = np.where(m > 0, dU / np.where(m > 0, m, 1), fallback)
    return np.concatenate([[dw[0]], dw, [dw[-1]]]).astype(np.float32)
