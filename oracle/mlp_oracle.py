"""numpy reference of the preconditioner network -- TEST INFRASTRUCTURE ONLY.

Restates the network, its loss and gradient in float64, with the same BF16 rounding of the
operands the GPU feeds to its matrix cores (inputs, weights, hidden activations and the
back-propagated gradients), so GPU gradients can be compared entry by entry. ``exact=True``
skips the rounding: that variant is pinned against central finite differences.

Parameter layout (f32 vector): W1 [128][64], b1 [128], W2 [128][128], b2 [128], W3 [48][128],
b3 [48]; weights are [out][in]; input columns 47..63 and output rows 45..47 are padding.
"""
import numpy as np

N_IN, IN_PAD, H, N_OUT, OUT_PAD = 47, 64, 128, 45, 48
OUT_SCALE = 0.125
SIZES = [("W1", H * IN_PAD), ("b1", H), ("W2", H * H), ("b2", H), ("W3", OUT_PAD * H), ("b3", OUT_PAD)]
N_PARAMS = sum(n for _, n in SIZES)


def bf16(x):
    """Round to nearest even BF16, returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, off = {}, 0
    for name, n in SIZES:
        out[name] = np.asarray(p[off:off + n], dtype=np.float64)
        off += n
    out["W1"] = out["W1"].reshape(H, IN_PAD)
    out["W2"] = out["W2"].reshape(H, H)
    out["W3"] = out["W3"].reshape(OUT_PAD, H)
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def gelu(x):
    c = 0.7978845608028654
    return 0.5 * x * (1 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    c = 0.7978845608028654
    t = np.tanh(c * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * c * (1 + 3 * 0.044715 * x * x)


def tril_index():
    """Packed lower column-major index -> (row, col) for 9x9."""
    idx = []
    for j in range(9):
        for i in range(j, 9):
            idx.append((i, j))
    return idx


TRIL = tril_index()


def unpack_sym(a):
    A = np.zeros(a.shape[:-1] + (9, 9))
    for k, (i, j) in enumerate(TRIL):
        A[..., i, j] = a[..., k]
        A[..., j, i] = a[..., k]
    return A


def g_matrix(y):
    G = np.zeros(y.shape[:-1] + (9, 9))
    for k, (i, j) in enumerate(TRIL):
        G[..., i, j] = OUT_SCALE * y[..., k]
    G += np.eye(9)
    return G


def forward(p, x, exact=False):
    r = (lambda v: v) if exact else bf16
    w = unpack(p)
    X = np.zeros((x.shape[0], IN_PAD))
    X[:, :N_IN] = x
    X = r(X)
    z1 = X @ r(w["W1"]).T + w["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ r(w["W2"]).T + w["b2"]
    h2 = r(gelu(z2))
    y = h2 @ r(w["W3"]).T + w["b3"]
    return {"X": X, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "y": y[:, :N_OUT]}


def loss(p, x, targets, exact=False):
    f = forward(p, x, exact)
    G = g_matrix(f["y"])
    A = unpack_sym(targets)
    E = np.swapaxes(G, -1, -2) @ A @ G - np.eye(9)
    return float(np.sum(E * E) / x.shape[0])


def loss_grad(p, x, targets, exact=False):
    """Mean loss and its gradient with respect to the parameter vector."""
    r = (lambda v: v) if exact else bf16
    n = x.shape[0]
    w = unpack(p)
    f = forward(p, x, exact)
    G = g_matrix(f["y"])
    A = unpack_sym(targets)
    AG = A @ G
    E = np.swapaxes(G, -1, -2) @ AG - np.eye(9)
    L = float(np.sum(E * E) / n)
    dG = 4.0 * AG @ E / n
    dy = np.zeros((n, OUT_PAD))
    for k, (i, j) in enumerate(TRIL):
        dy[:, k] = OUT_SCALE * dG[:, i, j]
    dy = r(dy)
    g = {}
    g["W3"] = dy.T @ f["h2"]
    g["b3"] = dy.sum(0)
    dz2 = r((dy @ r(w["W3"])) * gelu_grad(f["z2"]))
    g["W2"] = dz2.T @ f["h1"]
    g["b2"] = dz2.sum(0)
    dz1 = r((dz2 @ r(w["W2"])) * gelu_grad(f["z1"]))
    g["W1"] = dz1.T @ f["X"]
    g["b1"] = dz1.sum(0)
    return L, pack(g)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    a1 = np.sqrt(6.0 / (N_IN + H))
    d["W1"] = np.zeros((H, IN_PAD))
    d["W1"][:, :N_IN] = rng.uniform(-a1, a1, (H, N_IN))
    d["b1"] = rng.normal(0, 0.1, H)
    a2 = np.sqrt(6.0 / (2 * H))
    d["W2"] = rng.uniform(-a2, a2, (H, H))
    d["b2"] = rng.normal(0, 0.1, H)
    d["W3"] = np.zeros((OUT_PAD, H))
    d["W3"][:N_OUT] = rng.normal(0, 0.05, (N_OUT, H))
    d["b3"] = np.zeros(OUT_PAD)
    d["b3"][:N_OUT] = rng.normal(0, 0.05, N_OUT)
    return pack(d).astype(np.float32)
