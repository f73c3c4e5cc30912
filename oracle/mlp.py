"""CPU oracle of the closure surrogate network (numpy).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the library's BF16-MFMA
forward/backward kernels; the product never imports it.

The network: 14 inputs -> 4 hidden layers of 64 SiLU units -> 4 outputs,
with the library's parameter layout (row-major, out x in):
  W1 64x32 (inputs 14..31 are zero padding), b1 64,
  W2, W3, W4 64x64, b2..b4 64,
  W5 32x64 (output rows 4..31 are zero padding), b5 32.
Loss: mean over the batch and the 4 outputs of (pred - target)^2.

forward/backward() evaluate in float64. With bf16=True the operands are
rounded to bfloat16 (round to nearest even) exactly where the GPU kernels
round them (the normalised inputs, every weight matrix, every hidden
activation and every upstream gradient fed to a GEMM), and sums are taken
in float64, so the GPU result differs only by its float32 accumulation.
"""
import numpy as np

NIN, NIN_PAD, H, NOUT, NOUT_PAD = 14, 32, 64, 4, 32
SHAPES = [("W1", (H, NIN_PAD)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
          ("W3", (H, H)), ("b3", (H,)), ("W4", (H, H)), ("b4", (H,)),
          ("W5", (NOUT_PAD, H)), ("b5", (NOUT_PAD,))]


def param_count():
    return sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, k = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = np.asarray(flat[k:k + n], np.float64).reshape(s)
        k += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, bf=False):
    """x: [B, 14] normalised inputs. Returns (pred [B, 4], cache)."""
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    B = x.shape[0]
    a = np.zeros((B, NIN_PAD))
    a[:, :NIN] = x
    a = r(a)
    cache = {"a0": a}
    for l in range(1, 5):
        z = a @ r(params["W%d" % l]).T + params["b%d" % l]
        cache["z%d" % l] = z
        a = r(silu(z))
        cache["a%d" % l] = a
    z5 = a @ r(params["W5"]).T + params["b5"]
    return z5[:, :NOUT], cache


def loss_and_grads(params, x, y, bf=False):
    """Mean squared error and its gradients (dict like params)."""
    r = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    B = x.shape[0]
    pred, c = forward(params, x, bf)
    diff = pred - y
    loss = float((diff ** 2).mean())
    dz = np.zeros((B, NOUT_PAD))
    dz[:, :NOUT] = 2.0 * diff / (B * NOUT)
    g = {}
    for l in range(5, 0, -1):
        g["b%d" % l] = dz.sum(0)
        dzr = r(dz)
        g["W%d" % l] = dzr.T @ c["a%d" % (l - 1)]
        if l > 1:
            da = dzr @ r(params["W%d" % l])
            dz = da * dsilu(c["z%d" % (l - 1)])
    g["W1"][:, NIN:] = 0.0
    return loss, g


def init_params(seed=0):
    """He-style uniform initialisation (fixed seed), zero padding."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, s in SHAPES:
        if name.startswith("W"):
            fan_in = NIN if name == "W1" else H
            lim = np.sqrt(6.0 / fan_in) * 0.5
            w = rng.uniform(-lim, lim, size=s)
            if name == "W1":
                w[:, NIN:] = 0.0
            if name == "W5":
                w[NOUT:, :] = 0.0
            p[name] = w
        else:
            p[name] = np.zeros(s)
    return p
