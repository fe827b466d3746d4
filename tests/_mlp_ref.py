"""CPU reference of the learned predictor (16 -> 64 -> 64 -> 2 MLP, ReLU, mean
binary cross-entropy with logits over masked outputs) in float64 -- test
infrastructure for the GPU kernels' gradient check.

`bf16=True` rounds to bfloat16 where the GPU kernels do (inputs, weights,
hidden activations and the backpropagated signals that feed the matrix
products), so that only accumulation order separates it from the GPU; with
`bf16=False` it is the exact float64 model."""
import numpy as np

NF, NH, NO = 16, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NF * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NO
NPAR = OFF_B3 + NO


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NF, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(NH, NO), p[OFF_B3:NPAR])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def forward(p, x, bf=False):
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = q(x)
    Z1 = X @ q(W1) + b1
    H1 = q(np.maximum(Z1, 0.0))
    Z2 = H1 @ q(W2) + b2
    H2 = q(np.maximum(Z2, 0.0))
    Z3 = H2 @ q(W3) + b3
    return X, H1, H2, Z3


def loss_and_grad(p, x, t, bf=False):
    """Mean (over the n samples) of sum_o m_o * BCE(z_o, y_o); t[:, 2o] = y_o, t[:, 2o+1] = m_o."""
    q = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    n = len(x)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X, H1, H2, Z3 = forward(p, x, bf)
    y = t[:, 0::2].astype(np.float64)
    m = t[:, 1::2].astype(np.float64)
    sp = np.where(Z3 > 0, Z3 + np.log1p(np.exp(-np.abs(Z3))), np.log1p(np.exp(-np.abs(Z3))))
    loss = float(np.sum(m * (sp - y * Z3)) / n)
    d3 = m * (1.0 / (1.0 + np.exp(-Z3)) - y) / n
    G3 = q(d3.astype(np.float32)) if bf else d3
    gW3 = H2.T @ G3
    gb3 = d3.sum(axis=0)
    dz2 = (G3 @ q(W3).T) * (H2 > 0)
    G2 = q(dz2.astype(np.float32)) if bf else dz2
    gW2 = H1.T @ G2
    gb2 = dz2.sum(axis=0)
    dz1 = (G2 @ q(W2).T) * (H1 > 0)
    G1 = q(dz1.astype(np.float32)) if bf else dz1
    gW1 = X.T @ G1
    gb1 = dz1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def random_case(n, seed):
    """Parameters at the model's initialisation scale, features in the voxel ranges, and
    targets: half the samples voxel samples (output 0), half draw samples (output 1)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPAR)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2 / NF), NF * NH)
    p[OFF_B1:OFF_W2] = rng.normal(0, 0.1, NH)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2 / NH), NH * NH)
    p[OFF_B2:OFF_W3] = rng.normal(0, 0.1, NH)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1 / NH), NH * NO)
    p[OFF_B3:] = rng.normal(0, 0.1, NO)
    x = rng.uniform(-1, 1, (n, NF)).astype(np.float32)
    t = np.zeros((n, 4), dtype=np.uint8)
    t[:, 0] = rng.integers(0, 2, n)
    t[:, 2] = rng.integers(0, 2, n)
    half = rng.integers(0, 2, n).astype(bool)
    t[half, 1] = 1
    t[~half, 3] = 1
    return p.astype(np.float32), x, t
