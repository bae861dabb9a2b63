"""FP64 CPU restatement of the learned T2 guess (test infrastructure only).

The reference has no learned guess; what it replaces is the MP2 doubles
guess of CCWfn::CCWfn (ccwfn.cc:81-87), t2_ijab = <ij|ab> / D_ijab with
D_ijab = f_ii + f_jj - f_aa - f_bb.  This module restates, in numpy FP64:

- the per-amplitude features (`features`),
- the network f: 8 -> 128 -> 128 -> 128 -> 1 with GELU (tanh form) after the
  hidden layers, on standardised inputs (x - shift) / scale (`forward`),
- the loss (1/n) sum (f(x) - y)^2 and its gradient by backpropagation
  (`loss_grad`),
- one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) (`adam_step`),
- the guess t2 = t2(MP2) + y_scale (f(x_ijab) + f(x_jiba)) / 2 (`guess`).

Parameter order (FP32 on the device): W1[128][8], b1, W2[128][128], b2,
W3[128][128], b3, W4[128], b4; each layer computes z = W h + b.

Pins (tests/test_mlp_oracle.py): the gradient against central finite
differences in FP64; GELU against its closed form through erf at large |z|
and its known values; the features inverted back to <ij|ab>, <ij|ba> and
t(MP2) reproduce the MP2 energy of the reference's saved outputs.
"""
import numpy as np

NFEAT, WIDTH = 8, 128
C_GELU = np.sqrt(2.0 / np.pi)


def layout(width=WIDTH, nfeat=NFEAT):
    """(name, shape) of each parameter block in vector order."""
    return [("W1", (width, nfeat)), ("b1", (width,)), ("W2", (width, width)), ("b2", (width,)),
            ("W3", (width, width)), ("b3", (width,)), ("W4", (width,)), ("b4", (1,))]


def unpack(p, width=WIDTH, nfeat=NFEAT):
    out, off = {}, 0
    for name, shape in layout(width, nfeat):
        n = int(np.prod(shape))
        out[name] = np.asarray(p[off:off + n], dtype=np.float64).reshape(shape)
        off += n
    assert off == len(p)
    return out


def pack(d, width=WIDTH, nfeat=NFEAT):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in layout(width, nfeat)])


def gelu(z):
    u = C_GELU * (z + 0.044715 * z ** 3)
    return 0.5 * z * (1.0 + np.tanh(u))


def gelu_grad(z):
    u = C_GELU * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * z * z)


def features(no, nv, eps, oovv):
    """Features of every amplitude, [o*o*v*v][8] in ijab order (b fastest).

    eps: Fock diagonal (nact); oovv[i][j][a][b] = <ij|ab>."""
    e_o, e_v = eps[:no], eps[no:no + nv]
    D = (e_o[:, None, None, None] + e_o[None, :, None, None] - e_v[None, None, :, None] - e_v[None, None, None, :])
    I = oovv
    K = oovv.transpose(0, 1, 3, 2)
    shp = D.shape
    X = np.stack([np.arcsinh(1e3 * I / D), np.log(np.maximum(-D, 1e-6)), np.arcsinh(1e2 * I), np.arcsinh(1e2 * K),
                  np.broadcast_to(e_o[:, None, None, None], shp), np.broadcast_to(e_o[None, :, None, None], shp),
                  np.broadcast_to(e_v[None, None, :, None], shp), np.broadcast_to(e_v[None, None, None, :], shp)],
                 axis=-1)
    return X.reshape(-1, NFEAT)


def forward(p, X, shift, scale, keep=False):
    P = unpack(p)
    h = (np.asarray(X, np.float64) - shift) / scale
    cache = [h]
    for l in (1, 2, 3):
        z = h @ P["W%d" % l].T + P["b%d" % l]
        h = gelu(z)
        cache += [z, h]
    y = h @ P["W4"] + P["b4"][0]
    return (y, cache) if keep else y


def loss_grad(p, X, y, shift, scale):
    P = unpack(p)
    f, cache = forward(p, X, shift, scale, keep=True)
    n = len(y)
    d = f - y
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    g = {"W4": cache[6].T @ dy, "b4": np.array([dy.sum()])}
    dh = np.outer(dy, P["W4"])
    for l in (3, 2, 1):
        z, hprev = cache[2 * l - 1], cache[2 * l - 2]
        dz = dh * gelu_grad(z)
        g["W%d" % l] = dz.T @ hprev
        g["b%d" % l] = dz.sum(axis=0)
        dh = dz @ P["W%d" % l]
    return loss, pack(g)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam update number t (1-based); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def guess(p, no, nv, eps, oovv, shift, scale, y_scale):
    """t2 = t2(MP2) + y_scale (f(x_ijab) + f(x_jiba)) / 2, [i][j][a][b]."""
    e_o, e_v = eps[:no], eps[no:no + nv]
    D = (e_o[:, None, None, None] + e_o[None, :, None, None] - e_v[None, None, :, None] - e_v[None, None, None, :])
    X = features(no, nv, eps, oovv)
    f = forward(p, X, shift, scale).reshape(no, no, nv, nv)
    return oovv / D + y_scale * 0.5 * (f + f.transpose(1, 0, 3, 2))


def invert_features(X):
    """(t_MP2, <ij|ab>, <ij|ba>, D) recovered from feature rows."""
    X = np.asarray(X, np.float64)
    return np.sinh(X[:, 0]) / 1e3, np.sinh(X[:, 2]) / 1e2, np.sinh(X[:, 3]) / 1e2, -np.exp(X[:, 1])
