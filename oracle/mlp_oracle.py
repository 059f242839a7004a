"""NumPy restatement of the surrogate MLP's forward and backward passes --
TEST INFRASTRUCTURE ONLY (tests/ import it; nothing in the engine does).

The surrogate has no counterpart in the reference; this restates the
engine's training kernels (fjsph_amd/csrc/surrogate_train.hip) in FP64:

    H_0 = x (normalised, dim values), Z_l = W_l H_{l-1} + b_l,
    H_l = gelu(Z_l) = Z_l Phi(Z_l) for l = 1..L, Y = W_o H_L + b_o,
    loss = mean over the batch and the dim + 2 outputs of (Y - T)^2,
    dY = 2 (Y - T) / (B (dim + 2)), dZ_l = dH_l gelu'(Z_l),
    dW_l = dZ_l H_{l-1}^T, db_l = sum_points dZ_l, dH_{l-1} = W_l^T dZ_l.

With bf16=True the operands are rounded to BF16 (round to nearest even)
where the kernels round them: the weights, the activations H_l (H_0
included) and dY, dZ_l as operands of the products; the bias gradients
sum the unrounded FP32 values.  With bf16=False it is the exact FP64
gradient of the FP64 network, which the finite-difference test pins.

Parameter layout (as fjsph_set_surrogate takes it): weights W_1 [64][dim],
W_2..W_L [64][64], W_o [dim+2][64] row-major, flattened in that order;
biases b_1..b_L [64], b_o [dim+2].
"""
import numpy as np
from scipy.special import erf

WIDTH = 64


def bf16_round(a):
    """FP32 values rounded to the nearest BF16 (ties to even), as FP64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    finite = np.isfinite(f)
    r = np.where(finite, r, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(weights, biases, n_hidden, dim):
    """Per-layer (W, b) list: L hidden layers then the output layer."""
    sizes = [(WIDTH, dim)] + [(WIDTH, WIDTH)] * (n_hidden - 1) + [(dim + 2, WIDTH)]
    layers, ow, ob = [], 0, 0
    w = np.asarray(weights, dtype=np.float64)
    b = np.asarray(biases, dtype=np.float64)
    for o, i in sizes:
        layers.append((w[ow:ow + o * i].reshape(o, i), b[ob:ob + o]))
        ow += o * i
        ob += o
    return layers


def forward_backward(weights, biases, n_hidden, dim, xn, tn, bf16=True):
    """(dweights, dbiases, loss) over the batch xn [B][dim] -> tn [B][dim+2]."""
    layers = unpack(weights, biases, n_hidden, dim)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    f32 = (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)) if bf16 else (lambda a: a)
    B = xn.shape[0]
    no = dim + 2
    H = [rnd(xn)]
    Z = []
    for W, b in layers[:-1]:
        z = f32(H[-1] @ rnd(W).T + b)
        Z.append(z)
        H.append(rnd(f32(gelu(z))))
    Wo, bo = layers[-1]
    Y = f32(H[-1] @ rnd(Wo).T + bo)
    diff = Y - tn
    loss = float(np.mean(diff * diff))
    dY = f32(2.0 * diff / (B * no))
    grads_w, grads_b = [None] * len(layers), [None] * len(layers)
    g = rnd(dY)
    grads_w[-1] = g.T @ H[-1]
    grads_b[-1] = dY.sum(axis=0)
    dH = f32(g @ rnd(Wo))
    for l in range(n_hidden - 1, -1, -1):
        dZ = f32(dH * gelu_grad(Z[l]))
        g = rnd(dZ)
        grads_w[l] = g.T @ H[l]
        grads_b[l] = dZ.sum(axis=0)
        if l > 0:
            dH = f32(g @ rnd(layers[l][0]))
    dw = np.concatenate([gw.ravel() for gw in grads_w])
    db = np.concatenate(grads_b)
    return dw, db, loss


def loss_only(weights, biases, n_hidden, dim, xn, tn):
    """FP64 loss of the FP64 network (for finite differences)."""
    layers = unpack(weights, biases, n_hidden, dim)
    h = np.asarray(xn, dtype=np.float64)
    for W, b in layers[:-1]:
        h = gelu(h @ W.T + b)
    Wo, bo = layers[-1]
    d = h @ Wo.T + bo - tn
    return float(np.mean(d * d))
