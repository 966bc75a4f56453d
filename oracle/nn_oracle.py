"""CPU oracle of the learned preconditioner (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import
this module; the library never calls it.

A numpy restatement of the network the GPU kernels implement (libpps:
pps_nn.hip), forward and backward, with the training loss.  The reference
has no learned preconditioner (its preconditioners are NoneSolver,
ChebyshevIteration, BaseCG and BiCGSTAB chained through inputParam.hpp:22-31),
so this oracle is pinned by mathematics, not by reference vectors:
  * conv3d against torch.nn.functional.conv3d (a library routine), both
    orientations of the filter (forward and flipped/transposed);
  * the backward pass against central finite differences of the loss in
    float64 (rounding off);
  * pool / up against explicit loops, and the loss against its closed form
    (loss = 0 for y = e, invariance to the scale of y).

With rounding=True every value the GPU stores in BF16 is rounded at the same
point (layer outputs, staged pool / up+skip inputs, the BF16 copy of the loss
gradient, data gradients), so GPU and oracle agree to FP32 accumulation order
and occasional one-ulp BF16 flips.

Layout: activations [B][D][H][W][16] (channels last); parameters as the
library stores them: 14 layers x (W[27][16 ci][16 co] + bias[16]); tap =
9*dz + 3*dy + dx reads the input at offset (dz-1, dy-1, dx-1).
"""
import numpy as np

TAPS = 27
C = 16
LAYER_PARAMS = TAPS * C * C + C
# (cin, cout, level, mode, in_act, skip_act, relu): mode 'field'|'plain'|'pool'|'upadd'
NET = [
    (1, 16, 0, "field", -1, -1, 1),   # lift
    (16, 16, 1, "pool", 0, -1, 1),    # b0
    (16, 16, 1, "plain", 1, -1, 1),   # b1
    (16, 16, 2, "pool", 2, -1, 1),    # c0
    (16, 16, 2, "plain", 3, -1, 1),   # c1
    (16, 16, 2, "plain", 4, -1, 1),   # c2
    (16, 16, 2, "plain", 5, -1, 1),   # c3
    (16, 16, 2, "plain", 6, -1, 1),   # c4
    (16, 16, 2, "plain", 7, -1, 1),   # c5
    (16, 16, 2, "plain", 8, -1, 1),   # c6
    (16, 16, 2, "plain", 9, -1, 1),   # c7
    (16, 16, 1, "upadd", 10, 2, 1),   # e0
    (16, 16, 1, "plain", 11, -1, 1),  # e1
    (16, 1, 0, "upadd", 12, 0, 0),    # out
]
N_LAYERS = len(NET)
N_PARAMS = N_LAYERS * LAYER_PARAMS


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).reshape(x.shape)
    return np.where(np.isnan(x), x, out)


def _rnd(x, on):
    return bf16(x) if on else x


def split(params):
    """Per-layer (W[27][16][16], b[16]) views of the flat parameter vector."""
    p = np.asarray(params)
    out = []
    for l in range(N_LAYERS):
        blk = p[l * LAYER_PARAMS:(l + 1) * LAYER_PARAMS]
        out.append((blk[:TAPS * C * C].reshape(TAPS, C, C), blk[TAPS * C * C:]))
    return out


def shift_view(xp, tap, D, H, W):
    """Input at v + offset(tap) for all output voxels v, from the zero-padded xp."""
    dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
    return xp[:, dz:dz + D, dy:dy + H, dx:dx + W, :]


def conv(x, W, acc_dtype=np.float64):
    """out[v][co] = sum_tap sum_ci W[tap][ci][co] * x[v + off(tap)][ci] (zero padding)."""
    B, D, H, Wd, _ = x.shape
    xp = np.pad(x.astype(acc_dtype), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, D, H, Wd, W.shape[2]), acc_dtype)
    for t in range(TAPS):
        out += shift_view(xp, t, D, H, Wd) @ W[t].astype(acc_dtype)
    return out


def flip_transpose(W):
    """Filter of the backward-data pass: W'[tap][co][ci] = W[26 - tap][ci][co]."""
    return np.ascontiguousarray(W[::-1].transpose(0, 2, 1))


def conv_dw(x, dz):
    """dW[tap][ci][co] = sum_v x[v + off(tap)][ci] dz[v][co]."""
    B, D, H, Wd, _ = dz.shape
    xp = np.pad(x.astype(np.float64), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    z2 = dz.astype(np.float64).reshape(-1, dz.shape[-1])
    dW = np.zeros((TAPS, x.shape[-1], dz.shape[-1]))
    for t in range(TAPS):
        dW[t] = shift_view(xp, t, D, H, Wd).reshape(-1, x.shape[-1]).T @ z2
    return dW


def coarse_dims(d):
    return tuple((v + 1) // 2 for v in d)


def pool(x):
    """2x2x2 average over the children inside the finer grid (missing children count as 0)."""
    B, D, H, W, Cc = x.shape
    Dc, Hc, Wc = coarse_dims((D, H, W))
    xp = np.zeros((B, 2 * Dc, 2 * Hc, 2 * Wc, Cc), x.dtype)
    xp[:, :D, :H, :W] = x
    return xp.reshape(B, Dc, 2, Hc, 2, Wc, 2, Cc).sum(axis=(2, 4, 6)) * np.asarray(0.125, x.dtype)


def up(xc, dims):
    """Nearest-neighbour copy of the coarser level onto a grid of extent dims."""
    D, H, W = dims
    return xc.repeat(2, 1).repeat(2, 2).repeat(2, 3)[:, :D, :H, :W]


def pool_back(dc, dims):
    """Adjoint of pool: every child receives dcoarse / 8."""
    return up(dc, dims) * np.asarray(0.125, dc.dtype)


def up_back(dx, cdims):
    """Adjoint of up: the coarse voxel collects the sum over its children."""
    B, D, H, W, Cc = dx.shape
    Dc, Hc, Wc = cdims
    xp = np.zeros((B, 2 * Dc, 2 * Hc, 2 * Wc, Cc), dx.dtype)
    xp[:, :D, :H, :W] = dx
    return xp.reshape(B, Dc, 2, Hc, 2, Wc, 2, Cc).sum(axis=(2, 4, 6))


def _layer_input(l, r, acts, rounding, exact_skip=False):
    cin, cout, lev, mode, ia, sk, relu = NET[l]
    dt = np.float32 if rounding else np.float64
    if mode == "field":
        x = np.zeros(r.shape + (C,), dt)
        x[..., 0] = _rnd(r.astype(dt), rounding)
        return x
    if mode == "plain":
        return acts[ia]
    if mode == "pool":
        return _rnd(pool(acts[ia].astype(dt)), rounding)
    skip = acts[sk]
    return _rnd(up(acts[ia], skip.shape[1:4]).astype(dt) + skip, rounding and not exact_skip)


def _weights(params, rounding):
    layers = split(params)
    out = []
    dt = np.float32 if rounding else np.float64
    for l, (W, b) in enumerate(layers):
        cin, cout = NET[l][0], NET[l][1]
        Wm = np.zeros((TAPS, C, C), dt)
        Wm[:, :cin, :cout] = W[:, :cin, :cout]
        bm = np.zeros(C, dt)
        bm[:cout] = b[:cout]
        out.append((_rnd(Wm, rounding), bm))
    return out


def forward(params, r, rounding=True, exact_out_skip=False):
    """Network output y [B][n2][n1][n0] (float32/64) and activations act0..act12.

    exact_out_skip: the output layer's input up(act12) + act0 stays unrounded
    (the fused inference kernel adds the two inside its MFMA, libpps
    pps_nn.hip k_outf); every other rounding as with rounding=True."""
    r = np.asarray(r)
    dt = np.float32 if rounding else np.float64
    Ws = _weights(params, rounding)
    acts = []
    y = None
    for l in range(N_LAYERS):
        x = _layer_input(l, r, acts, rounding, exact_out_skip and l == N_LAYERS - 1).astype(dt)
        W, b = Ws[l]
        z = conv(x, W.astype(dt), np.float64) + b.astype(np.float64)
        if l == N_LAYERS - 1:
            y = z[..., 0].astype(dt)
        else:
            acts.append(_rnd(np.maximum(z, 0.0).astype(dt), rounding))
    return y, acts


def output_layer(params, act0, act12, exact_skip=False):
    """The output convolution alone, from given level-0 and level-1 activations
    (float32, as stored): conv(up(act12) + act0) + bias, channel 0.  With
    exact_skip the sum is not rounded to BF16 (the fused inference kernel)."""
    W, b = _weights(params, True)[N_LAYERS - 1]
    x = up(np.asarray(act12, np.float32), act0.shape[1:4]).astype(np.float32) + np.asarray(act0, np.float32)
    x = x if exact_skip else bf16(x)
    return (conv(x, W.astype(np.float32), np.float64) + b.astype(np.float64))[..., 0]


def k_apply(y):
    """K y with K = 6 I - (index-space neighbours), zero outside (Dirichlet)."""
    yp = np.pad(y, ((0, 0), (1, 1), (1, 1), (1, 1)))
    c = yp[:, 1:-1, 1:-1, 1:-1]
    return (6 * c - yp[:, :-2, 1:-1, 1:-1] - yp[:, 2:, 1:-1, 1:-1] - yp[:, 1:-1, :-2, 1:-1] - yp[:, 1:-1, 2:, 1:-1]
            - yp[:, 1:-1, 1:-1, :-2] - yp[:, 1:-1, 1:-1, 2:])


def loss_and_dy(y, r, G):
    """Mean over samples of 1 - (y.r)^2 / ((y.Ky) G), and its gradient in y."""
    B = y.shape[0]
    ky = k_apply(y)
    a = np.einsum("bzyx,bzyx->b", y.astype(np.float64), r.astype(np.float64))
    c = np.einsum("bzyx,bzyx->b", y.astype(np.float64), ky.astype(np.float64))
    G = np.asarray(G, np.float64)
    loss = np.mean(1.0 - a * a / (c * G))
    c1 = 2 * a * a / (c * c * G * B)
    c2 = -2 * a / (c * G * B)
    dy = c1[:, None, None, None] * ky + c2[:, None, None, None] * r
    return loss, dy


def loss_grad(params, r, G, rounding=True):
    """Loss and the gradient of every stored parameter (flat, like the library's)."""
    dt = np.float32 if rounding else np.float64
    y, acts = forward(params, r, rounding)
    loss, dy = loss_and_dy(y, r.astype(dt), G)
    dy = dy.astype(dt)
    Ws = _weights(params, rounding)
    grad = np.zeros(N_PARAMS)
    dacts = [None] * len(acts)
    for l in range(N_LAYERS - 1, -1, -1):
        cin, cout, lev, mode, ia, sk, relu = NET[l]
        x = _layer_input(l, r, acts, rounding).astype(dt)
        if l == N_LAYERS - 1:
            dz = np.zeros(dy.shape + (C,), dt)
            dz[..., 0] = _rnd(dy, rounding)
        else:
            dz = np.where(acts[l] > 0, dacts[l], 0).astype(dt)
        dW = conv_dw(x, dz)
        db = dz.astype(np.float64).sum(axis=(0, 1, 2, 3))
        blk = grad[l * LAYER_PARAMS:(l + 1) * LAYER_PARAMS]
        blk[:TAPS * C * C] = dW.reshape(-1)
        blk[TAPS * C * C:] = db
        if l == 0:
            break
        dx = _rnd(conv(dz, flip_transpose(Ws[l][0]).astype(dt), np.float64).astype(dt), rounding)
        if mode == "plain":
            dacts[ia] = dx
        elif mode == "pool":
            fine = acts[ia]
            dacts[ia] = _rnd(dacts[ia] + pool_back(dx, fine.shape[1:4]), rounding)
        else:
            dacts[sk] = dx
            dacts[ia] = _rnd(up_back(dx, acts[ia].shape[1:4]), rounding)
    # parameters outside a layer's channel counts are not parameters
    for l in range(N_LAYERS):
        cin, cout = NET[l][0], NET[l][1]
        W = grad[l * LAYER_PARAMS:l * LAYER_PARAMS + TAPS * C * C].reshape(TAPS, C, C)
        W[:, cin:, :] = 0
        W[:, :, cout:] = 0
        grad[l * LAYER_PARAMS + TAPS * C * C + cout:(l + 1) * LAYER_PARAMS] = 0
    return loss, grad


def init_params(seed):
    """He-normal weights (std sqrt(2/fan_in), sqrt(1/fan_in) for the linear output), zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for l, (cin, cout, lev, mode, ia, sk, relu) in enumerate(NET):
        sd = np.sqrt((2.0 if relu else 1.0) / (TAPS * cin))
        W = p[l * LAYER_PARAMS:l * LAYER_PARAMS + TAPS * C * C].reshape(TAPS, C, C)
        W[:, :cin, :cout] = rng.normal(0.0, sd, (TAPS, cin, cout))
    return p
