"""FP64 restatement of the warm-start network (test infrastructure only).

Same architecture, parameter layout and rounding points as the HIP
implementation in lbm_amd/csrc/lbm_surrogate.hip:
  * network input, every hidden activation and pooled tensor, the bf16
    working copies of the weights and every gradient fed to a matrix product
    are rounded to bf16 (round to nearest even of the FP32 value);
  * biases, the output layer and all accumulations stay wide (FP64 here,
    FP32 on the GPU).
With exact=True no rounding is applied at all; that mode is checked against
central finite differences in tests/test_surrogate.py, which pins the
backward pass.  There is no reference counterpart (SFCMM/LBM has no learned
component); the layer semantics are the textbook ones: 3x3 "same"
convolution with zero padding, ReLU, 2x2 average pooling, nearest
upsampling, channel concatenation [upsampled, skip], mean squared error over
fluid cells and the 3 outputs.
"""
import numpy as np

# widths of the six resolution levels
WIDTHS = (32, 32, 64, 96, 128, 160)


def build():
    """Layer list and pooling steps, identical to build_net() in the HIP code."""
    tensors = [(0, 16)]  # (level, channels)
    layers = []
    pools = []

    def conv(a, s, co, relu, last=False):
        ca = tensors[a][1] if a is not None else 0
        cs = tensors[s][1]
        out = None
        if not last:
            tensors.append((tensors[s][0], co))
            out = len(tensors) - 1
        layers.append(dict(a=a, s=s, ca=ca, cs=cs, co=co, relu=relu, last=last, out=out))
        return out

    def pool(src):
        tensors.append((tensors[src][0] + 1, tensors[src][1]))
        pools.append((src, len(tensors) - 1))
        return len(tensors) - 1

    enc = []
    cur = 0
    for lv, wdt in enumerate(WIDTHS):
        if lv > 0:
            cur = pool(cur)
        cur = conv(None, conv(None, cur, wdt, True), wdt, True)
        enc.append(cur)
    for lv in range(len(WIDTHS) - 2, -1, -1):
        cur = conv(None, conv(cur, enc[lv], WIDTHS[lv], True), WIDTHS[lv], True)
    conv(None, cur, 32, False, last=True)
    off = 0
    for l in layers:
        ci = l["ca"] + l["cs"]
        l["woff"] = off
        off += l["co"] * 9 * ci
        l["boff"] = off
        off += l["co"]
    return tensors, layers, pools, off


def bf16(x):
    """Round to nearest even bf16 of the FP32 value of x; returned as float64."""
    u = np.ascontiguousarray(np.asarray(x, dtype=np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(x, dy, dx):
    """out[b, y, x] = x[b, y + dy, x + dx] (zero outside)."""
    out = np.zeros_like(x)
    H, W = x.shape[1], x.shape[2]
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def pool2(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


class Net:
    def __init__(self, params, exact=False):
        self.tensors, self.layers, self.pools, self.n = build()
        self.p = np.asarray(params, np.float64)
        assert self.p.size == self.n
        self.exact = exact

    def r(self, x):
        return x if self.exact else bf16(x)

    def weights(self, l):
        ci = l["ca"] + l["cs"]
        w = self.p[l["woff"]:l["woff"] + l["co"] * 9 * ci].reshape(l["co"], 9, ci)
        b = self.p[l["boff"]:l["boff"] + l["co"]]
        return self.r(w), (b if self.exact else b.astype(np.float32).astype(np.float64))

    def _input(self, acts, l):
        xs = acts[l["s"]]
        if l["a"] is None:
            return xs
        return np.concatenate([up(acts[l["a"]]), xs], axis=-1)

    def forward(self, x):
        """x[n][H][W][cin] -> outputs [n][H][W][3]; keeps activations for backward."""
        x = np.asarray(x, np.float64)
        n, H, W, cin = x.shape
        x0 = np.zeros((n, H, W, 16))
        x0[..., :cin] = x
        acts = {0: self.r(x0)}
        pools = list(self.pools)
        for l in self.layers:
            while pools and pools[0][1] == l["s"]:
                src, dst = pools.pop(0)
                acts[dst] = self.r(pool2(acts[src]))
            X = self._input(acts, l)
            w, b = self.weights(l)
            acc = np.zeros(X.shape[:3] + (l["co"],))
            for tap in range(9):
                acc += shift(X, tap // 3 - 1, tap % 3 - 1) @ w[:, tap, :].T
            y = acc + b
            if l["last"]:
                out = y[..., :3] if self.exact else y[..., :3].astype(np.float32).astype(np.float64)
            else:
                if l["relu"]:
                    y = np.maximum(y, 0.0)
                acts[l["out"]] = self.r(y)
        self.acts = acts
        return out

    def loss_grad(self, x, t, mask):
        """Mean squared error over fluid cells and outputs, and its gradient
        with respect to the flat parameter vector."""
        o = self.forward(x)
        t = np.asarray(t, np.float64)
        m = np.asarray(mask, bool)[..., None]
        nfl = max(1, int(m.sum()))
        d = np.where(m, o - t, 0.0)
        loss = float((d * d).sum() / (3 * nfl))
        g = np.zeros(self.n)
        gout = np.zeros(o.shape[:3] + (32,))
        gout[..., :3] = 2.0 * d / (3 * nfl)
        gout = self.r(gout)
        acts = self.acts
        grads = {k: np.zeros_like(v) for k, v in acts.items()}
        pools = list(self.pools)
        for li in range(len(self.layers) - 1, -1, -1):
            l = self.layers[li]
            if l["last"]:
                gy = gout
            else:
                gy = grads[l["out"]]
                if l["relu"]:
                    gy = np.where(acts[l["out"]] > 0, gy, 0.0)
                gy = self.r(gy)
            X = self._input(acts, l)
            w, _ = self.weights(l)
            ci = l["ca"] + l["cs"]
            gw = np.zeros((l["co"], 9, ci))
            gyf = gy.reshape(-1, l["co"])
            for tap in range(9):
                gw[:, tap, :] = gyf.T @ shift(X, tap // 3 - 1, tap % 3 - 1).reshape(-1, ci)
            g[l["woff"]:l["woff"] + gw.size] = gw.reshape(-1)
            g[l["boff"]:l["boff"] + l["co"]] = gyf.sum(axis=0)
            if li > 0:
                dX = np.zeros(X.shape)
                for tap in range(9):
                    dX += shift(gy, -(tap // 3 - 1), -(tap % 3 - 1)) @ w[:, tap, :]
                grads[l["s"]] += dX[..., l["ca"]:]
                if l["a"] is not None:
                    dA = dX[..., :l["ca"]]
                    grads[l["a"]] += dA[:, 0::2, 0::2] + dA[:, 0::2, 1::2] + dA[:, 1::2, 0::2] + dA[:, 1::2, 1::2]
            while pools and pools[-1][1] == l["s"]:
                src, dst = pools.pop()
                grads[src] += 0.25 * up(grads[dst])
        return loss, g
