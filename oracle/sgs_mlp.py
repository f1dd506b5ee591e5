"""CPU restatement of the learned SGS closure's network (test infrastructure only:
imported by tests/, never by the product).

The network is the eddy-viscosity model of the learned LES closure: 10 inputs,
three hidden layers of 64 SiLU units, one softplus output,

    a0 = bf16(x)                      (x padded with zeros to 16 inputs)
    z_l = bf16(W_l) a_{l-1} + b_l,    h_l = silu(z_l),  a_l = bf16(h_l),  l = 1..3
    z_4 = bf16(w_4) . a_3 + b_4,      y = softplus(z_4)

with BF16 weights and activations as the operands of every product and FP32 (here
FP64) accumulation.  The loss is the batch mean of (y - t)^2.  Gradients treat the
BF16 roundings of the activations as the identity (straight-through) and round
each layer's output gradient dz_l to BF16 before it enters a product, as the
device kernels do:

    dz_4 = 2 (y - t) / B * sigmoid(z_4)
    dW_4 = sum_s dz_4 a_3,            da_3 = bf16(w_4) dz_4
    dz_l = da_l * silu'(z_l),         dW_l = bf16(dz_l) a_{l-1}^T,
    db_l = sum_s dz_l,                da_{l-1} = bf16(W_l)^T bf16(dz_l)

Parameter vector (9,089 FP32 values): W1 (64x10, row-major [out][in]), b1 (64),
W2 (64x64), b2, W3 (64x64), b3, w4 (64), b4 (1).

The pieces are pinned in tests/test_sgs_mlp_oracle.py: with the roundings off the
loss gradient equals central finite differences of the loss and PyTorch's
autograd in float64; silu, softplus and the BF16 rounding are checked against
their closed forms.
"""
import numpy as np

N_IN = 10
N_PAD = 16
H = 64
SHAPES = [("W1", (H, N_IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("w4", (H,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float32).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), through float32 as on the device."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def d_silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def softplus(z):
    return np.where(z > 20.0, z, np.log1p(np.exp(np.minimum(z, 20.0))))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def _r(x, rounding):
    return bf16(x) if rounding else np.asarray(x, dtype=np.float64)


def forward(p, x, rounding=True, cache=False):
    """y (n,) for inputs x (n, 10)."""
    w = unpack(p)
    a = _r(np.asarray(x, dtype=np.float32).T if rounding else np.asarray(x, dtype=np.float64).T, rounding)
    acts, zs = [a], []
    for l in (1, 2, 3):
        z = _r(w["W%d" % l], rounding) @ a + w["b%d" % l][:, None]
        zs.append(z)
        if rounding:
            z = z.astype(np.float32).astype(np.float64)
        a = _r(silu(z), rounding)
        acts.append(a)
    z4 = _r(w["w4"], rounding) @ a + w["b4"][0]
    y = softplus(z4)
    if cache:
        return y, (w, acts, zs, z4)
    return y


def loss_grad(p, x, t, rounding=True):
    """(loss, gradient as an N_PARAMS float64 vector) of mean (y - t)^2."""
    t = np.asarray(t, dtype=np.float64)
    n = t.size
    y, (w, acts, zs, z4) = forward(p, x, rounding, cache=True)
    loss = float(np.mean((y - t) ** 2))
    g = {}
    dz = 2.0 * (y - t) / n * sigmoid(z4)
    g["w4"] = acts[3] @ dz
    g["b4"] = np.array([dz.sum()])
    da = np.outer(_r(w["w4"], rounding), dz)
    for l in (3, 2, 1):
        dzl = da * d_silu(zs[l - 1])
        g["b%d" % l] = dzl.sum(axis=1)
        dzr = _r(dzl, rounding)
        g["W%d" % l] = dzr @ acts[l - 1].T
        if l > 1:
            da = _r(w["W%d" % l], rounding).T @ dzr
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])


def adam_step(p, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) of FP32 parameters; step counts from 1."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """He-uniform weights for the SiLU layers, zero biases (the device's own
    initialisation uses the same rule with its own generator)."""
    rng = np.random.default_rng(seed)
    d = {}
    for l, fan_in in ((1, N_IN), (2, H), (3, H)):
        lim = np.sqrt(6.0 / fan_in)
        d["W%d" % l] = rng.uniform(-lim, lim, size=(H, fan_in))
        d["b%d" % l] = np.zeros(H)
    d["w4"] = rng.uniform(-np.sqrt(3.0 / H), np.sqrt(3.0 / H), size=H)
    d["b4"] = np.zeros(1)
    return pack(d)
