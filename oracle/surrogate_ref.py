"""CPU restatement of the learned barotropic surrogate's network.

TEST INFRASTRUCTURE ONLY: the checker the HIP kernels of
pcom_amd/csrc/surrogate.hip are compared against; only tests/ load it.

The network has six 3x3 convolutions with their taps DIL[l] pixels apart
(1, 2, 4, 8, 4, 1), zonally periodic with zero padding at the walls, on
NHWC tensors.  Its channels are 32 (16 used) -> 64 -> 64
-> 64 -> 64 -> 64 -> 32 (4 used), with ReLU after all but the last layer.
The loss is the masked mean-squared error over the 4 used outputs: channel
c < 2 is masked by input channel 5 (T-point sea) and the others by input
channel 6 (U-point sea), divided by pixels x 4.  Parameters are laid out as
pcom_nn_get_params gives them: per layer W [cout][3][3][cin] then b
[cout], then 31 normalisation values.

With rnd=True the same BF16 roundings as the device are applied at the
same points:
- the input, weights, hidden activations and the loss gradient are
  rounded to BF16;
- the output of each backward-data pass is rounded to BF16;
- products are accumulated in float64 here and in FP32 on the device.
With rnd=False the network runs in plain float64: that version is pinned
by central finite differences and a brute-force convolution in the tests.
The reference has no learned part, so nothing here cites it.
"""
import numpy as np

IN_C, HID, OUT_C, IN_USED, OUT_USED = 32, 64, 32, 16, 4
DIMS = [(IN_C, HID), (HID, HID), (HID, HID), (HID, HID), (HID, HID), (HID, OUT_C)]
DIL = [1, 2, 4, 8, 4, 1]
N_NORM = 31


def param_count():
    return sum(9 * ci * co + co for ci, co in DIMS) + N_NORM


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    layers, off = [], 0
    for ci, co in DIMS:
        w = np.asarray(params[off:off + 9 * ci * co], np.float64).reshape(co, 3, 3, ci)
        off += 9 * ci * co
        b = np.asarray(params[off:off + co], np.float64)
        off += co
        layers.append((w, b))
    return layers


def shift(x, oy, ox):
    """r[:, y, x] = x[:, y + oy, (x + ox) mod W], zero where y + oy is outside."""
    _, h, _, _ = x.shape
    r = np.roll(x, -ox, axis=2)
    out = np.zeros_like(r)
    if abs(oy) >= h:
        return out
    if oy >= 0:
        out[:, :h - oy] = r[:, oy:]
    else:
        out[:, -oy:] = r[:, :h + oy]
    return out


def conv(x, w, b, dil=1):
    """out[p][co] = b[co] + sum_{dy,dx,ci} w[co,dy,dx,ci] x[p + dil (dy-1, dx-1)][ci]"""
    out = np.zeros(x.shape[:3] + (w.shape[0],)) + b
    for dy in range(3):
        for dx in range(3):
            out += shift(x, (dy - 1) * dil, (dx - 1) * dil) @ w[:, dy, dx, :].T
    return out


def conv_data_grad(d, w, dil=1):
    """din[p][ci] = sum_{dy,dx,co} w[co,dy,dx,ci] d[p - dil (dy-1, dx-1)][co]"""
    out = np.zeros(d.shape[:3] + (w.shape[3],))
    for dy in range(3):
        for dx in range(3):
            out += shift(d, (1 - dy) * dil, (1 - dx) * dil) @ w[:, dy, dx, :]
    return out


def conv_weight_grad(x, d, dil=1):
    """gw[co,dy,dx,ci] = sum_p d[p][co] x[p + dil (dy-1, dx-1)][ci]"""
    co, ci = d.shape[3], x.shape[3]
    g = np.zeros((co, 3, 3, ci))
    dm = d.reshape(-1, co)
    for dy in range(3):
        for dx in range(3):
            g[:, dy, dx, :] = dm.T @ shift(x, (dy - 1) * dil, (dx - 1) * dil).reshape(-1, ci)
    return g


def _r(a, rnd):
    return bf16(a).astype(np.float64) if rnd else np.asarray(a, np.float64)


def forward(params, x8, rnd=True):
    """x8: [B, H, W, 16] -> (activations incl. the padded input, output [B,H,W,32])."""
    layers = unpack(params)
    x = np.zeros(x8.shape[:3] + (IN_C,))
    x[..., :IN_USED] = x8
    acts = [_r(x, rnd)]
    out = None
    for li, (w, b) in enumerate(layers):
        z = conv(acts[-1], _r(w, rnd), np.asarray(b, np.float32 if rnd else np.float64), DIL[li])
        if li < len(layers) - 1:
            acts.append(_r(np.maximum(z, 0.0), rnd))
        else:
            out = z.astype(np.float32).astype(np.float64) if rnd else z
    return acts, out


def loss_and_grads(params, x8, target, rnd=True):
    """Masked MSE and its gradient in parameter order (normalisation zero)."""
    layers = unpack(params)
    acts, out = forward(params, x8, rnd)
    npix = np.prod(x8.shape[:3])
    inv = np.float32(1.0 / (npix * OUT_USED)) if rnd else 1.0 / (npix * OUT_USED)
    mask = np.concatenate([np.repeat(acts[0][..., 5:6], 2, axis=-1), np.repeat(acts[0][..., 6:7], 2, axis=-1)],
                          axis=-1)
    if rnd:
        d = (mask.astype(np.float32) * (out[..., :OUT_USED].astype(np.float32) - target.astype(np.float32)))
        loss = float(np.sum(d.astype(np.float64) ** 2)) * (1.0 / (npix * OUT_USED))
        dpre = np.zeros(out.shape)
        dpre[..., :OUT_USED] = bf16(np.float32(2.0) * d * inv)
    else:
        d = mask * (out[..., :OUT_USED] - target)
        loss = float(np.sum(d ** 2)) * inv
        dpre = np.zeros(out.shape)
        dpre[..., :OUT_USED] = 2.0 * d * inv
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        w, _ = layers[li]
        grads[li] = (conv_weight_grad(acts[li], dpre, DIL[li]), dpre.reshape(-1, dpre.shape[-1]).sum(axis=0))
        if li == 0:
            break
        din = _r(conv_data_grad(dpre, _r(w, rnd), DIL[li]), rnd)
        dpre = din * (acts[li] > 0)
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads] + [np.zeros(N_NORM)])
    return loss, out[..., :OUT_USED], flat
