"""Float64 numpy oracle of the warm-start MLP (TEST INFRASTRUCTURE ONLY).

Forward pass, analytic backward pass and Adam, for the network
22 -> 64 -> 64 -> 64 -> 3 (SiLU), unit output u = y/|y|, loss mean(1 - u.t).
Parameter layout W1[22][64] b1 W2[64][64] b2 W3[64][64] b3 W4[64][3] b4.
The backward pass is pinned by central finite differences in
tests/test_oracle_mlp.py; the features by direct geometric evaluation.
"""
import numpy as np

NIN, NH, NOUT = 22, 64, 3
SHAPES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
          ("W4", (NH, NOUT)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    w = np.asarray(w, np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = w[o:o + k].reshape(shape)
        o += k
    return out


def pack(p):
    return np.concatenate([p[name].reshape(-1) for name, _ in SHAPES])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, X):
    p = unpack(w)
    X = np.asarray(X, np.float64)
    z1 = X @ p["W1"] + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = silu(z2)
    z3 = h2 @ p["W3"] + p["b3"]
    h3 = silu(z3)
    y = h3 @ p["W4"] + p["b4"]
    return y, (X, z1, h1, z2, h2, z3, h3)


def loss(w, X, T):
    y, _ = forward(w, X)
    u = y / np.linalg.norm(y, axis=1, keepdims=True)
    return float(np.mean(1.0 - (u * np.asarray(T, np.float64)).sum(1)))


def loss_and_grad(w, X, T):
    p = unpack(w)
    y, (X, z1, h1, z2, h2, z3, h3) = forward(w, X)
    T = np.asarray(T, np.float64)
    B = len(X)
    ny = np.linalg.norm(y, axis=1, keepdims=True)
    u = y / ny
    ut = (u * T).sum(1, keepdims=True)
    L = float(np.mean(1.0 - ut))
    dy = -(T - ut * u) / ny / B
    g = {"W4": h3.T @ dy, "b4": dy.sum(0)}
    dz3 = (dy @ p["W4"].T) * dsilu(z3)
    g["W3"], g["b3"] = h2.T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ p["W3"].T) * dsilu(z2)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * dsilu(z1)
    g["W1"], g["b1"] = X.T @ dz1, dz1.sum(0)
    return L, pack(g)


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward_bf16(w, X):
    """Inference with BF16 weights and BF16 activations, FP32 accumulation."""
    p = {k: bf16(v.astype(np.float32)).astype(np.float64) for k, v in unpack(w).items()}
    for k in ("b1", "b2", "b3", "b4"):
        p[k] = unpack(w)[k]
    h = bf16(np.asarray(X, np.float32)).astype(np.float64)
    for i in (1, 2, 3):
        z = h @ p["W%d" % i] + p["b%d" % i]
        h = bf16(silu(z).astype(np.float32)).astype(np.float64)
    y = h @ p["W4"] + p["b4"]
    return y / np.linalg.norm(y, axis=1, keepdims=True)


def quat_mat(q):
    w, x, y, z = (np.float64(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def quat_mul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


def descriptor(h, i):
    v = h["verts4"][h["vstart"][i]:h["vstart"][i] + h["nv"][i], :3].astype(np.float64)
    r = np.linalg.norm(v, axis=1)
    return np.array([*np.abs(v).max(0), r.mean(), r.max(), r.min(), len(v) / 32.0])


def features(h, pairs):
    out = np.zeros((len(pairs), NIN))
    for k, (a, b) in enumerate(pairs):
        da, db = descriptor(h, a), descriptor(h, b)
        s = da[4] + db[4]
        dc = h["pos4"][b, :3].astype(np.float64) - h["pos4"][a, :3]
        pl = quat_mat(h["quat4"][a]).T @ dc
        qa = h["quat4"][a].astype(np.float64)
        q = quat_mul(np.array([qa[0], -qa[1], -qa[2], -qa[3]]), h["quat4"][b].astype(np.float64))
        if q[0] < 0:
            q = -q
        out[k] = np.concatenate([pl / s, q, da[:6] / s, [da[6]], db[:6] / s, [db[6]], [np.linalg.norm(dc) / s]])
    return out
