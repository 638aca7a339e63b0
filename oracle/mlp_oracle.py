"""CPU oracle of the learned limiter's network (test infrastructure only: the
product path is the HIP library; nothing here is imported by it).

The network, its loss and its features are defined by this project (the
reference has no learned limiter), so this module restates them in numpy
from their definition in include/mahyco_gpu.h:
  * forward: H1 = relu(X W1^T + b1), H2 = relu(H1 W2^T + b2), Y = H2 W3^T + b3,
    with X, W1..W3 and the activations H1, H2 rounded to BF16 (round to
    nearest even) and FP32 accumulation when bf16=True, all in float64 when
    bf16=False;
  * loss: for direction d and channel k < 4, e = y_d sc - st; L = sum e^2 / m;
  * backward: the chain rule of the two lines above, with the ReLU-masked
    gradients dL/dH2 and dL/dH1 rounded to BF16 when bf16=True;
  * features: 5 x 5 stencil of (density, specific internal energy, cell-mean
    x and y velocity, largest environment volume fraction), edge cells
    repeated, each channel centred on the cell and divided by its largest
    deviation plus 1e-6 |centre| + 1e-30, channel-major, padded to 128.
The float64 mode is pinned by finite differences (tests/test_mlp_oracle.py).
"""
import numpy as np

IN, HID, OUT = 128, 64, 2
SIZES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (OUT, HID)),
         ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n]).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).ravel() for name, _ in SIZES])


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def forward(p, X, bf=True):
    w = unpack(p)
    if bf:
        f = np.float32
        Xb = bf16(X)
        W1, W2, W3 = bf16(w["W1"]), bf16(w["W2"]), bf16(w["W3"])
        H1 = bf16(np.maximum(Xb @ W1.T + w["b1"].astype(f), 0))
        H2 = bf16(np.maximum(H1 @ W2.T + w["b2"].astype(f), 0))
        Y = H2 @ W3.T + w["b3"].astype(f)
        return Y.astype(np.float32), (Xb, H1, H2, W1, W2, W3)
    X = np.asarray(X, dtype=np.float64)
    H1 = np.maximum(X @ w["W1"].T + w["b1"], 0)
    H2 = np.maximum(H1 @ w["W2"].T + w["b2"], 0)
    Y = H2 @ w["W3"].T + w["b3"]
    return Y, (X, H1, H2, w["W1"], w["W2"], w["W3"])


def loss_grad(p, X, T, bf=True):
    """Loss and gradient (flat, in the parameter layout) over the rows of X."""
    m = X.shape[0]
    Y, (Xa, H1, H2, W1, W2, W3) = forward(p, X, bf)
    Tt = np.asarray(T, dtype=Y.dtype).reshape(m, 2, 4, 2)
    sc, st = Tt[..., 0], Tt[..., 1]
    e = Y[:, :, None] * sc - st
    loss = float(np.sum(e * e) / m)
    dY = 2.0 * np.sum(e * sc, axis=2) / m
    r = bf16 if bf else (lambda a: a)
    g = {"W3": dY.T @ H2, "b3": dY.sum(0)}
    D2 = r((dY @ W3) * (H2 > 0))
    g["W2"], g["b2"] = D2.T @ H1, D2.sum(0)
    D1 = r((D2 @ W2) * (H1 > 0))
    g["W1"], g["b1"] = D1.T @ Xa, D1.sum(0)
    return loss, pack(g)


def features(nx, ny, rho, e, ux, uy, vf=None):
    """Network inputs [nx*ny][128] and channel scales [nx*ny][5] of a 2D state
    (cells i fastest; node velocities on (nx+1) x (ny+1) nodes; vf the largest
    environment volume fraction per cell, 1 if None)."""
    ux = np.asarray(ux, dtype=np.float64).reshape(ny + 1, nx + 1)
    uy = np.asarray(uy, dtype=np.float64).reshape(ny + 1, nx + 1)
    uc = 0.25 * (((ux[:-1, :-1] + ux[:-1, 1:]) + ux[1:, 1:]) + ux[1:, :-1])
    vc = 0.25 * (((uy[:-1, :-1] + uy[:-1, 1:]) + uy[1:, 1:]) + uy[1:, :-1])
    f = np.ones((ny, nx)) if vf is None else np.asarray(vf).reshape(ny, nx)
    chans = [np.asarray(rho).reshape(ny, nx), np.asarray(e).reshape(ny, nx), uc, vc, f]
    X = np.zeros((nx * ny, IN), dtype=np.float32)
    S = np.zeros((nx * ny, 5), dtype=np.float32)
    for k, ch in enumerate(chans):
        pad = np.pad(ch.astype(np.float32), 2, mode="edge")
        st = np.stack([pad[dj:dj + ny, di:di + nx] for dj in range(5) for di in range(5)], axis=-1)
        st = st.reshape(nx * ny, 25)
        ctr = st[:, 12:13]
        sc = np.max(np.abs(st - ctr), axis=1, keepdims=True)
        s = sc + (np.float32(1e-6) * np.abs(ctr) + np.float32(1e-30))
        X[:, k * 25:(k + 1) * 25] = (st - ctr) / s
        S[:, k] = s[:, 0]
    return X, S
