"""CPU oracle of the surrogate network -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the network the HIP kernels in vortexlab_amd/csrc/nn.hip
implement (4 x [conv 3x3 pad 1 + bias + ReLU + 2x2 max-pool], channels 1-32-32-64-64,
flatten + scaled log Re, FC 1025-128 + ReLU, FC 128-3, MSE loss), with the same flat parameter
layout.  It is pinned by central finite differences (tests/test_surrogate_oracle.py) and is
the reference the GPU gradients are checked against.
"""
import numpy as np

CH = [1, 32, 32, 64, 64]
SP = [64, 32, 16, 8, 4]
FLAT, FIN, HID, OUT = 1024, 1025, 128, 3


def layout():
    off, o = {}, 0
    for l in range(4):
        off[f"cw{l}"] = (o, (CH[l + 1], CH[l], 3, 3)); o += CH[l + 1] * CH[l] * 9
        off[f"cb{l}"] = (o, (CH[l + 1],)); o += CH[l + 1]
    off["w1"] = (o, (HID, FIN)); o += HID * FIN
    off["b1"] = (o, (HID,)); o += HID
    off["w2"] = (o, (OUT, HID)); o += OUT * HID
    off["b2"] = (o, (OUT,)); o += OUT
    return off, o


def unpack(p):
    off, _ = layout()
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in off.items()}


def conv(x, w, b):
    n, ci, h, wd = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    out = np.zeros((n, w.shape[0], h, wd)) + b[None, :, None, None]
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("oc,nchw->nohw", w[:, :, ky, kx], xp[:, :, ky:ky + h, kx:kx + wd])
    return out


def pool(a):
    n, c, h, w = a.shape
    v = a.reshape(n, c, h // 2, 2, w // 2, 2).transpose(0, 1, 2, 4, 3, 5).reshape(n, c, h // 2, w // 2, 4)
    return v.max(-1), v.argmax(-1)  # argmax: first maximum in raster order of the 2x2 window


def forward(p, masks, s):
    P = unpack(p)
    x = masks.reshape(-1, 1, 64, 64).astype(np.float64)
    cache = []
    for l in range(4):
        z = conv(x, P[f"cw{l}"], P[f"cb{l}"])
        a = np.maximum(z, 0.0)
        pl, am = pool(a)
        cache.append((x, a, am))
        x = pl
    f = np.concatenate([x.reshape(len(x), -1), s.reshape(-1, 1)], axis=1)
    h = np.maximum(f @ P["w1"].T + P["b1"], 0.0)
    y = h @ P["w2"].T + P["b2"]
    return y, (cache, f, h)


def loss_and_grad(p, masks, s, t):
    P = unpack(p)
    y, (cache, f, h) = forward(p, masks, s)
    n = len(y)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    G = {}
    G["w2"] = dy.T @ h
    G["b2"] = dy.sum(0)
    dh = (dy @ P["w2"]) * (h > 0)
    G["w1"] = dh.T @ f
    G["b1"] = dh.sum(0)
    dpool = (dh @ P["w1"])[:, :FLAT].reshape(n, 64, 4, 4)
    for l in range(3, -1, -1):
        x, a, am = cache[l]
        nn_, c, hh, ww = a.shape
        route = np.zeros((nn_, c, hh // 2, ww // 2, 4))
        np.put_along_axis(route, am[..., None], dpool[..., None], axis=-1)
        dz = route.reshape(nn_, c, hh // 2, ww // 2, 2, 2).transpose(0, 1, 2, 4, 3, 5).reshape(nn_, c, hh, ww)
        dz = dz * (a > 0)
        xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
        gw = np.zeros_like(P[f"cw{l}"])
        for ky in range(3):
            for kx in range(3):
                gw[:, :, ky, kx] = np.einsum("nohw,nchw->oc", dz, xp[:, :, ky:ky + hh, kx:kx + ww])
        G[f"cw{l}"] = gw
        G[f"cb{l}"] = dz.sum((0, 2, 3))
        if l > 0:
            dzp = np.pad(dz, ((0, 0), (0, 0), (1, 1), (1, 1)))
            dx = np.zeros_like(x)
            w = P[f"cw{l}"]
            for ky in range(3):
                for kx in range(3):
                    # x[y,x] contributes to z[y-ky+1, x-kx+1]
                    dx += np.einsum("oc,nohw->nchw", w[:, :, ky, kx], dzp[:, :, 2 - ky:2 - ky + hh, 2 - kx:2 - kx + ww])
            dpool = dx
    off, total = layout()
    g = np.zeros(total)
    for k, (o, shp) in off.items():
        g[o:o + int(np.prod(shp))] = G[k].reshape(-1)
    return loss, g
