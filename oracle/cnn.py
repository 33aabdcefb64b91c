"""Oracle (test infrastructure only): numpy FP64 restatement of the learned
proposal network, its losses and gradients.

The network has no counterpart in the reference; its definition is the one
stated in include/spectrapack_hip.h (sp_hip_proposal_*) and
cpp-max-packer_amd/csrc/learned.hip.  With emulate=True the oracle rounds to
BF16 (round to nearest even) at the points where the device kernels store
BF16: the input fractions, the conv weights, every conv activation after
ReLU, and every gated activation gradient.  Sums run in FP64 here and in FP32
on the device, so the two agree to BF16 tolerance, not bit for bit.

Pinning (tests/test_oracle_cnn.py): conv taps against scipy.ndimage.correlate
(stride 1) and its subsampling (stride 2); the downsample against closed forms
and an independent per-bin loop; the analytic gradients against central
finite differences of the FP64 (emulate=False) loss.
"""
from __future__ import annotations

import numpy as np

IN = 64
CH = 32
DESC = 32
HIDDEN = 64
ORIENT = 96
REGIONS = 16

LAYOUT = [
    ("w1", (CH, 27)), ("b1", (CH,)),
    ("w2", (CH, 27, CH)), ("b2", (CH,)),
    ("w3", (CH, 27, CH)), ("b3", (CH,)),
    ("w4", (CH, 27, CH)), ("b4", (CH,)),
    ("wf", (CH, DESC)),
    ("wh", (CH,)), ("bh", (1,)),
    ("wa", (HIDDEN, 2 * CH + DESC)), ("ba", (HIDDEN,)),
    ("wo", (ORIENT, HIDDEN)), ("bo", (ORIENT,)),
]
PARAMETER_COUNT = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(flat):
    """Flat parameter vector -> dict of views (the order of LAYOUT)."""
    flat = np.asarray(flat)
    out, off = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = flat[off:off + n].reshape(shape)
        off += n
    assert off == flat.size, (off, flat.size)
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def bf16(x):
    """Round to BF16 (nearest even) through float32, returned as float64."""
    a = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    a = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def downsample(env):
    """Occupancy [ez, ey, ex] (0/1) -> uint8 [64, 64, 64] (z, y, x): bin b of an
    axis of extent e covers cells [b e // 64, (b + 1) e // 64); the value is
    (255 count + cells // 2) // cells."""
    env = (np.asarray(env) != 0).astype(np.int64)
    counts = env
    cells = np.ones(1, np.int64)
    sizes = []
    for axis in range(3):
        e = env.shape[axis]
        if e < IN:
            raise ValueError("extent below 64")
        edges = (np.arange(IN + 1) * e) // IN
        counts = np.add.reduceat(counts, edges[:-1], axis=axis)
        sizes.append(np.diff(edges))
    cells = sizes[0][:, None, None] * sizes[1][None, :, None] * sizes[2][None, None, :]
    return ((255 * counts + cells // 2) // cells).astype(np.uint8)


def input_values(q, emulate=True):
    x = np.float32(np.asarray(q, dtype=np.float32)) * np.float32(np.float32(1.0) / np.float32(255.0))
    return bf16(x) if emulate else np.asarray(q, np.float64) / 255.0


def _im2col(x, stride):
    """x [B, D, D, D, C] -> columns [B, Do, Do, Do, 27, C]; tap t = 9 dz + 3 dy +
    dx reads source voxel stride * v + d - 1 (zero outside)."""
    B, D = x.shape[0], x.shape[1]
    do = D // stride
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    cols = []
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        cols.append(xp[:, dz:dz + stride * do:stride, dy:dy + stride * do:stride, dx:dx + stride * do:stride, :])
    return np.stack(cols, axis=-2)


def _col2im(cols, stride, D):
    """Adjoint of _im2col: columns [B, Do, Do, Do, 27, C] -> [B, D, D, D, C]."""
    B, do, C = cols.shape[0], cols.shape[1], cols.shape[-1]
    xp = np.zeros((B, D + 2, D + 2, D + 2, C))
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        xp[:, dz:dz + stride * do:stride, dy:dy + stride * do:stride, dx:dx + stride * do:stride, :] += cols[..., t, :]
    return xp[:, 1:-1, 1:-1, 1:-1, :]


def conv(x, w, b, stride):
    """Pre-activation of one conv layer: x [B, D, D, D, Cin], w [Cout, 27, Cin]."""
    cols = _im2col(x, stride)
    shp = cols.shape
    y = cols.reshape(-1, 27 * shp[-1]) @ w.reshape(w.shape[0], -1).T + b
    return y.reshape(shp[0], shp[1], shp[2], shp[3], w.shape[0]), cols


def forward(flat, q, desc, emulate=True):
    """q: uint8 [B, 64, 64, 64] (z, y, x); desc [B, 32].  Returns a dict with
    logits [B, 96], heat [B, 4096] and the intermediates backward needs."""
    P = unpack(np.asarray(flat, np.float64))
    rw = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    w1 = rw(P["w1"]).reshape(CH, 27, 1)
    w2, w3, w4 = rw(P["w2"]), rw(P["w3"]), rw(P["w4"])
    q = np.asarray(q)
    desc = np.asarray(desc, np.float64)
    B = q.shape[0]
    x0 = input_values(q, emulate)[..., None]
    film = desc @ P["wf"].T                                   # [B, 32]
    z1, c1 = conv(x0, w1, P["b1"], 2)
    a1 = rw(np.maximum(z1, 0))
    z2, c2 = conv(a1, w2, P["b2"], 2)
    z2 = z2 + film[:, None, None, None, :]
    a2 = rw(np.maximum(z2, 0))
    z3, c3 = conv(a2, w3, P["b3"], 1)
    a3 = rw(np.maximum(z3, 0))
    z4, c4 = conv(a3, w4, P["b4"], 1)
    a4 = rw(np.maximum(z4, 0))
    a4f = a4.reshape(B, -1, CH)                               # [B, 4096, 32]
    heat = a4f @ P["wh"] + P["bh"][0]
    g = a4f.mean(axis=1)
    pi = np.exp(heat - heat.max(axis=1, keepdims=True))
    pi /= pi.sum(axis=1, keepdims=True)                       # softmax of the heat logits
    fatt = np.einsum("bv,bvc->bc", pi, a4f)                   # heat-weighted pooling
    u = np.concatenate([g, fatt, desc], axis=1)
    hid = np.maximum(u @ P["wa"].T + P["ba"], 0)
    logits = hid @ P["wo"].T + P["bo"]
    return dict(P=P, w=(w1, w2, w3, w4), x0=x0, cols=(c1, c2, c3, c4), a=(a1, a2, a3, a4), a4f=a4f, heat=heat, g=g,
                pi=pi, fatt=fatt, u=u, hid=hid, logits=logits, desc=desc, emulate=emulate)


def _log_softmax(z, valid):
    z = np.where(valid, z, -np.inf)
    m = z.max(axis=-1, keepdims=True)
    s = np.log(np.exp(z - m).sum(axis=-1, keepdims=True))
    return z - m - s


def loss_and_grads(flat, q, desc, target, mask, region, heat_weight=1.0, emulate=True):
    """Mean orientation and region losses (as sp_hip_proposal_train) and the
    flat gradient of loss_o + heat_weight * loss_h."""
    f = forward(flat, q, desc, emulate)
    P, B = f["P"], f["logits"].shape[0]
    rw = bf16 if emulate else (lambda a: a)
    valid = np.asarray(mask) != 0
    target = np.where(valid, np.asarray(target, np.float64), 0.0)
    region = np.asarray(region)
    lp = _log_softmax(f["logits"], valid)
    loss_o = -(target * np.where(valid, lp, 0.0)).sum() / B
    dlog = np.where(valid, np.exp(lp) - target, 0.0) / B
    lh = _log_softmax(f["heat"], np.ones_like(f["heat"], bool))
    has = region >= 0
    loss_h = -sum(lh[b, region[b]] for b in range(B) if has[b]) / B
    dheat = np.exp(lh)
    for b in range(B):
        if has[b]:
            dheat[b, region[b]] -= 1.0
        else:
            dheat[b] = 0.0
    dheat *= heat_weight / B
    G = {name: np.zeros(shape) for name, shape in LAYOUT}
    G["wo"] = dlog.T @ f["hid"]
    G["bo"] = dlog.sum(0)
    dpre = (dlog @ P["wo"]) * (f["hid"] > 0)
    G["wa"] = dpre.T @ f["u"]
    G["ba"] = dpre.sum(0)
    du = dpre @ P["wa"]
    dg = du[:, :CH] / f["a4f"].shape[1]
    df = du[:, CH:2 * CH]
    # attention pooling: f = sum_v pi_v a_v with pi = softmax(heat)
    sv = np.einsum("bvc,bc->bv", f["a4f"], df)
    dheat = dheat + f["pi"] * (sv - (f["pi"] * sv).sum(axis=1, keepdims=True))
    G["wh"] = np.einsum("bv,bvc->c", dheat, f["a4f"])
    G["bh"] = np.array([dheat.sum()])
    a1, a2, a3, a4 = f["a"]
    da4 = dheat[:, :, None] * P["wh"][None, None, :] + dg[:, None, :] + f["pi"][:, :, None] * df[:, None, :]
    dz4 = rw(np.where(a4.reshape(da4.shape) > 0, da4, 0.0)).reshape(a4.shape)
    w1, w2, w3, w4 = f["w"]
    c1, c2, c3, c4 = f["cols"]

    def layer(dz, cols, w, stride, D, gate):
        co = dz.shape[-1]
        dzf = dz.reshape(-1, co)
        dw = dzf.T @ cols.reshape(dzf.shape[0], -1)
        db = dzf.sum(0)
        if gate is None:
            return dw.reshape(w.shape), db, None
        dcols = (dzf @ w.reshape(co, -1)).reshape(cols.shape)
        dx = _col2im(dcols, stride, D)
        return dw.reshape(w.shape), db, rw(np.where(gate > 0, dx, 0.0))

    G["w4"], G["b4"], dz3 = layer(dz4, c4, w4, 1, 16, a3)
    G["w3"], G["b3"], dz2 = layer(dz3, c3, w3, 1, 16, a2)
    G["w2"], G["b2"], dz1 = layer(dz2, c2, w2, 2, 32, a1)
    dw1, G["b1"], _ = layer(dz1, c1, w1, 2, 64, None)
    G["w1"] = dw1.reshape(CH, 27)
    G["wf"] = dz2.reshape(B, -1, CH).sum(1).T @ f["desc"]
    return loss_o, loss_h, pack(G), f


def init_parameters(seed):
    """He-style random initial parameters (FP32 values as float64)."""
    rs = np.random.default_rng(seed)
    d = {}
    for name, shape in LAYOUT:
        if name.startswith("b"):
            d[name] = np.zeros(shape)
            continue
        fan_in = {"w1": 27, "w2": 864, "w3": 864, "w4": 864, "wf": DESC, "wh": CH, "wa": 2 * CH + DESC,
                  "wo": HIDDEN}[name]
        d[name] = rs.normal(0.0, np.sqrt(2.0 / fan_in), shape)
    d["wf"] *= 0.1
    return pack(d).astype(np.float32).astype(np.float64)
