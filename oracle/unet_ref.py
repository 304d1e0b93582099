"""Test infrastructure: float64 numpy reference of the pressure warm-start
U-Net (forward, backward, loss), the CPU oracle its HIP kernels are checked
against.  It has no counterpart in ARTSS; the network is this project's own.

Layout: channels-last arrays of shape (Z, Y, X, C), i.e. linear voxel index
v = i + X*(j + Y*k) with the channel innermost (offset v*C + c), the layout
of the device tensors.  Convolutions are 3x3x3 with zero padding and tap
t = 9*(dz+1) + 3*(dy+1) + (dx+1); weights W[t, c_in, c_out], bias b[c_out].

Network (levels 0..3, channels 16, 32, 64, 64):
    h0 = relu(conv_in(x))                 2 -> 16
    e0 = relu(conv(h0))                  16 -> 16
    e1 = relu(conv(down(e0)))            16 -> 32
    e2 = relu(conv(down(e1)))            32 -> 64
    e3 = relu(conv(down(e2)))            64 -> 64
    d2 = relu(conv(up(e3))) + e2         64 -> 64
    d1 = relu(conv(up(d2))) + e1         64 -> 32
    d0 = relu(conv(up(d1))) + e0         32 -> 16
    y  = conv_out(d0)                    16 -> 1
down = 2x2x2 mean, up = nearest-neighbour doubling.
"""
import numpy as np

LAYERS = [  # name, c_in, c_out, level
    ("in", 2, 16, 0), ("e0", 16, 16, 0), ("e1", 16, 32, 1), ("e2", 32, 64, 2), ("e3", 64, 64, 3),
    ("d2", 64, 64, 2), ("d1", 64, 32, 1), ("d0", 32, 16, 0), ("out", 16, 1, 0),
]


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co, _ in LAYERS:
        std = scale * np.sqrt(2.0 / (27 * ci))
        p[name + ".W"] = rng.standard_normal((27, ci, co)) * std
        p[name + ".b"] = np.zeros(co)
    return p


def conv(x, W, b):
    Z, Y, X, _ = x.shape
    co = W.shape[2]
    xp = np.zeros((Z + 2, Y + 2, X + 2, x.shape[3]))
    xp[1:-1, 1:-1, 1:-1] = x
    out = np.broadcast_to(b, (Z, Y, X, co)).copy()
    for t in range(27):
        dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        out += xp[1 + dz:Z + 1 + dz, 1 + dy:Y + 1 + dy, 1 + dx:X + 1 + dx] @ W[t]
    return out


def conv_backward(x, W, gout):
    """gradients of conv(x, W, b) given d loss / d out"""
    Z, Y, X, ci = x.shape
    xp = np.zeros((Z + 2, Y + 2, X + 2, ci))
    xp[1:-1, 1:-1, 1:-1] = x
    gxp = np.zeros_like(xp)
    gW = np.zeros_like(W)
    for t in range(27):
        dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        sl = (slice(1 + dz, Z + 1 + dz), slice(1 + dy, Y + 1 + dy), slice(1 + dx, X + 1 + dx))
        gW[t] = xp[sl].reshape(-1, ci).T @ gout.reshape(-1, gout.shape[3])
        gxp[sl] += gout @ W[t].T
    gb = gout.reshape(-1, gout.shape[3]).sum(axis=0)
    return gxp[1:-1, 1:-1, 1:-1], gW, gb


def down(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def down_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, axis=0), 2, axis=1), 2, axis=2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, axis=0), 2, axis=1), 2, axis=2)


def up_backward(g):
    Z, Y, X, C = g.shape
    return g.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


def forward(p, x):
    """returns y and the activations the backward pass needs"""
    a = {}
    a["x"] = x
    a["h0"] = np.maximum(conv(x, p["in.W"], p["in.b"]), 0)
    a["e0"] = np.maximum(conv(a["h0"], p["e0.W"], p["e0.b"]), 0)
    a["p1"] = down(a["e0"])
    a["e1"] = np.maximum(conv(a["p1"], p["e1.W"], p["e1.b"]), 0)
    a["p2"] = down(a["e1"])
    a["e2"] = np.maximum(conv(a["p2"], p["e2.W"], p["e2.b"]), 0)
    a["p3"] = down(a["e2"])
    a["e3"] = np.maximum(conv(a["p3"], p["e3.W"], p["e3.b"]), 0)
    a["u2"] = up(a["e3"])
    a["c2"] = conv(a["u2"], p["d2.W"], p["d2.b"])
    a["d2"] = np.maximum(a["c2"], 0) + a["e2"]
    a["u1"] = up(a["d2"])
    a["c1"] = conv(a["u1"], p["d1.W"], p["d1.b"])
    a["d1"] = np.maximum(a["c1"], 0) + a["e1"]
    a["u0"] = up(a["d1"])
    a["c0"] = conv(a["u0"], p["d0.W"], p["d0.b"])
    a["d0"] = np.maximum(a["c0"], 0) + a["e0"]
    y = conv(a["d0"], p["out.W"], p["out.b"])
    return y, a


def loss_and_grad(p, x, target):
    """mean squared error over all voxels and its gradient with respect to every parameter"""
    y, a = forward(p, x)
    n = y.size
    diff = y - target
    loss = float((diff * diff).sum() / n)
    g = {}
    gy = 2.0 * diff / n
    gd0, g["out.W"], g["out.b"] = conv_backward(a["d0"], p["out.W"], gy)
    ge0 = gd0.copy()
    gc0 = gd0 * (a["c0"] > 0)
    gu0, g["d0.W"], g["d0.b"] = conv_backward(a["u0"], p["d0.W"], gc0)
    gd1 = up_backward(gu0)
    ge1 = gd1.copy()
    gc1 = gd1 * (a["c1"] > 0)
    gu1, g["d1.W"], g["d1.b"] = conv_backward(a["u1"], p["d1.W"], gc1)
    gd2 = up_backward(gu1)
    ge2 = gd2.copy()
    gc2 = gd2 * (a["c2"] > 0)
    gu2, g["d2.W"], g["d2.b"] = conv_backward(a["u2"], p["d2.W"], gc2)
    ge3 = up_backward(gu2)
    gp3, g["e3.W"], g["e3.b"] = conv_backward(a["p3"], p["e3.W"], ge3 * (a["e3"] > 0))
    ge2 += down_backward(gp3)
    gp2, g["e2.W"], g["e2.b"] = conv_backward(a["p2"], p["e2.W"], ge2 * (a["e2"] > 0))
    ge1 += down_backward(gp2)
    gp1, g["e1.W"], g["e1.b"] = conv_backward(a["p1"], p["e1.W"], ge1 * (a["e1"] > 0))
    ge0 += down_backward(gp1)
    gh0, g["e0.W"], g["e0.b"] = conv_backward(a["h0"], p["e0.W"], ge0 * (a["e0"] > 0))
    _, g["in.W"], g["in.b"] = conv_backward(a["x"], p["in.W"], gh0 * (a["h0"] > 0))
    return loss, g
