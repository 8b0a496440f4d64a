"""numpy reference of the learned preconditioner's MLP (test code).

Same network and parameter layout as sbd-pub_amd/csrc/mlp.hip:
160 -> 128 -> 128 -> 1, GELU (tanh form), BF16 weights and BF16 rounding of
X, Z1, H1, Z2, H2, dZ1, dZ2 at the same points as the kernels (rounding
treated as identity in the derivative); accumulation in float64 here
(FP32 on the GPU).  With bf16=False everything stays float64, which the
finite-difference test uses to pin the backward formulas themselves.
"""
import numpy as np

DIN, HID = 160, 128
W1, B1 = 0, DIN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID
NPARAM = B3 + 4


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


C_GELU = np.sqrt(2.0 / np.pi)


def gelu(z):
    """Tanh-form GELU, 0.5 z (1 + tanh(u)) = z * sigmoid(2u),
    u = sqrt(2/pi) (z + 0.044715 z^3)."""
    z = np.asarray(z, dtype=np.float64)
    u = C_GELU * (z + 0.044715 * z ** 3)
    return z / (1.0 + np.exp(-2.0 * u))


def gelu_d(z):
    z = np.asarray(z, dtype=np.float64)
    u = C_GELU * (z + 0.044715 * z ** 3)
    s = 1.0 / (1.0 + np.exp(-2.0 * u))
    return s + z * s * (1.0 - s) * 2.0 * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(DIN, HID), p[B1:W2], p[W2:B2].reshape(HID, HID), p[B2:W3],
            p[W3:B3], p[B3])


def forward(p, X, rb=True):
    r = bf16 if rb else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    Xr = r(X)
    Z1 = r(Xr @ r(w1) + b1)
    H1 = r(gelu(Z1))
    Z2 = r(H1 @ r(w2) + b2)
    H2 = r(gelu(Z2))
    y = H2 @ r(w3) + b3
    return y, (Xr, Z1, H1, Z2, H2)


def loss_and_grad(p, X, t, rb=True):
    r = bf16 if rb else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    y, (Xr, Z1, H1, Z2, H2) = forward(p, X, rb)
    n = X.shape[0]
    d = y - t
    dy = 2.0 * d / n
    dZ2 = r(dy[:, None] * r(w3)[None, :] * gelu_d(Z2))
    dH1 = dZ2 @ r(w2).T
    dZ1 = r(dH1 * gelu_d(Z1))
    g = np.zeros(NPARAM)
    g[W1:B1] = (Xr.T @ dZ1).ravel()
    g[B1:W2] = dZ1.sum(axis=0)
    g[W2:B2] = (H1.T @ dZ2).ravel()
    g[B2:W3] = dZ2.sum(axis=0)
    g[W3:B3] = H2.T @ dy
    g[B3] = dy.sum()
    return float(np.mean(d * d)), g


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[W1:B1] = rng.normal(0, scale * np.sqrt(2.0 / DIN), DIN * HID)
    p[W2:B2] = rng.normal(0, scale * np.sqrt(2.0 / HID), HID * HID)
    p[W3:B3] = rng.normal(0, scale * np.sqrt(1.0 / HID), HID)
    return p.astype(np.float32)
