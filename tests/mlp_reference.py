"""Reference of the learned warm-start MLP (16-64-64-64-3, ReLU) for the tests.

`forward`/`grad` follow the device arithmetic: bf16 (round to nearest even)
weights, inputs, hidden activations and back-propagated pre-activation
gradients, fp32 biases, accumulation in float64 here (fp32 on the device).
`exact_loss`/`exact_grad` are the float64 model without any rounding, used to
pin the back-propagation formulas against finite differences.
"""
import numpy as np

IN, HID, OUT = 16, 64, 3
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, HID)), ("b3", (HID,)),
         ("W4", (OUT, HID)), ("b4", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shp in SIZES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float32).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """float32 -> bfloat16 (round to nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def init_params(rng, last_scale=0.1):
    d = {}
    for name, shp in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shp[1])
            d[name] = rng.uniform(-lim, lim, shp)
        else:
            d[name] = rng.normal(0, 0.1, shp)
    d["W4"] = d["W4"] * last_scale
    return pack(d)


def random_norm(rng):
    mean_in = rng.normal(0, 1, IN)
    inv_std = rng.uniform(0.5, 2.0, IN)
    std_out = rng.uniform(0.5, 3.0, OUT)
    mean_out = rng.normal(0, 1, OUT)
    return np.concatenate([mean_in, inv_std, std_out, mean_out]).astype(np.float32)


def normalise_inputs(x, norm):
    x = np.asarray(x, np.float32)
    n = np.asarray(norm, np.float32)
    return (x - n[:16]) * n[16:32]  # fp32, as on the device


def forward(p, norm, x, keep=False):
    P = unpack(p)
    n = np.asarray(norm, np.float32)
    h0 = bf16(normalise_inputs(x, norm))
    z1 = h0 @ bf16(P["W1"]).T + P["b1"]
    h1 = bf16(np.maximum(z1, 0))
    z2 = h1 @ bf16(P["W2"]).T + P["b2"]
    h2 = bf16(np.maximum(z2, 0))
    z3 = h2 @ bf16(P["W3"]).T + P["b3"]
    h3 = bf16(np.maximum(z3, 0))
    yn = h3 @ bf16(P["W4"]).T + P["b4"]
    y = yn * n[32:35].astype(np.float64) + n[35:38].astype(np.float64)
    if keep:
        return y, dict(h0=h0, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, yn=yn)
    return y


def grad(p, norm, x, t):
    """Gradient of mean((y' - t')^2) over the n x 3 normalised outputs, device rounding points."""
    P = unpack(p)
    n = np.asarray(norm, np.float32)
    _, a = forward(p, norm, x, keep=True)
    tn = (np.asarray(t, np.float32) - n[35:38]) / n[32:35]
    d = a["yn"] - tn.astype(np.float64)
    loss = float(np.mean(d * d))
    gscale = np.float32(2.0) / np.float32(3.0 * len(x))
    g4 = bf16(np.float32(gscale) * d.astype(np.float32))
    g3 = bf16(np.where(a["z3"] > 0, g4 @ bf16(P["W4"]), 0.0))
    g2 = bf16(np.where(a["z2"] > 0, g3 @ bf16(P["W3"]), 0.0))
    g1 = bf16(np.where(a["z1"] > 0, g2 @ bf16(P["W2"]), 0.0))
    G = dict(W1=g1.T @ a["h0"], b1=g1.sum(0), W2=g2.T @ a["h1"], b2=g2.sum(0), W3=g3.T @ a["h2"], b3=g3.sum(0),
             W4=g4.T @ a["h3"], b4=g4.sum(0))
    return pack(G).astype(np.float64), loss


def _exact_forward(P, xn):
    z1 = xn @ P["W1"].T + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ P["W3"].T + P["b3"]
    h3 = np.maximum(z3, 0)
    return (z1, h1, z2, h2, z3, h3), h3 @ P["W4"].T + P["b4"]


def exact_loss(p, xn, tn):
    _, yn = _exact_forward(unpack(p), xn)
    return float(np.mean((yn - tn) ** 2))


def exact_grad(p, xn, tn):
    P = unpack(p)
    (z1, h1, z2, h2, z3, h3), yn = _exact_forward(P, xn)
    g4 = 2.0 * (yn - tn) / yn.size
    g3 = np.where(z3 > 0, g4 @ P["W4"], 0.0)
    g2 = np.where(z2 > 0, g3 @ P["W3"], 0.0)
    g1 = np.where(z1 > 0, g2 @ P["W2"], 0.0)
    G = dict(W1=g1.T @ xn, b1=g1.sum(0), W2=g2.T @ h1, b2=g2.sum(0), W3=g3.T @ h2, b3=g3.sum(0), W4=g4.T @ h3,
             b4=g4.sum(0))
    return np.concatenate([np.asarray(G[name], np.float64).reshape(-1) for name, _ in SIZES])
