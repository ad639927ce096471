"""CPU oracle of the learned interface closure's network arithmetic.

Test infrastructure only: the product path runs the HIP kernels of
``libicyclosure.so``; tests use these FP64 NumPy restatements as the checker.
The closure has no counterpart in the reference (which uses the coarse
exchange formula of iceflow3d/thermal.py:308-392 alone), so nothing here
follows a reference file; each routine is pinned in the tests against a
library routine (``torch.nn.functional.conv3d`` and autograd on the CPU) or
against finite differences.

Conventions (those of include/icyclosure.h): activations are channels-last
arrays of shape (nx, ny, nz, 32); taps t = 9 (di + 1) + 3 (dj + 1) + (dk + 1)
with zero padding outside the grid; forward weights wf[t, co, ci].
"""

import numpy as np

C = 32
TAPS = 27
LAYER_PARAMS = TAPS * C * C + C


def bf16_round(x):
    """Round FP32 values to the nearest BF16 (ties to even), as FP32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


def bf16_bits(x):
    """BF16 bit patterns (uint16) of FP32 values rounded to nearest even."""
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def _shifted(x, di, dj, dk):
    """y[i, j, k] = x[i + di, j + dj, k + dk] (zero outside)."""
    nx, ny, nz = x.shape[:3]
    y = np.zeros_like(x)
    si = slice(max(0, -di), min(nx, nx - di))
    sj = slice(max(0, -dj), min(ny, ny - dj))
    sk = slice(max(0, -dk), min(nz, nz - dk))
    y[si, sj, sk] = x[si.start + di:si.stop + di, sj.start + dj:sj.stop + dj, sk.start + dk:sk.stop + dk]
    return y


def conv3d(x, wf, bias=None):
    """out[n, co] = bias[co] + sum_t sum_ci x[n + off(t), ci] wf[t, co, ci] (FP64)."""
    x = np.asarray(x, dtype=np.float64)
    wf = np.asarray(wf, dtype=np.float64).reshape(TAPS, C, C)
    out = np.zeros(x.shape[:3] + (C,), dtype=np.float64)
    for t, (di, dj, dk) in enumerate(taps()):
        out += _shifted(x, di, dj, dk) @ wf[t].T
    if bias is not None:
        out += np.asarray(bias, dtype=np.float64)
    return out


def conv3d_backward_data(dout, wf):
    """dx[m, ci] = sum_t sum_co dout[m - off(t), co] wf[t, co, ci] (the adjoint of conv3d)."""
    dout = np.asarray(dout, dtype=np.float64)
    wf = np.asarray(wf, dtype=np.float64).reshape(TAPS, C, C)
    dx = np.zeros_like(dout)
    for t, (di, dj, dk) in enumerate(taps()):
        dx += _shifted(dout, -di, -dj, -dk) @ wf[t]
    return dx


def conv3d_weight_grad(x, dout):
    """dW[t, co, ci] = sum_n dout[n, co] x[n + off(t), ci]; db[co] = sum_n dout[n, co]."""
    x = np.asarray(x, dtype=np.float64)
    dout = np.asarray(dout, dtype=np.float64)
    d2 = dout.reshape(-1, C)
    dw = np.zeros((TAPS, C, C), dtype=np.float64)
    for t, (di, dj, dk) in enumerate(taps()):
        dw[t] = d2.T @ _shifted(x, di, dj, dk).reshape(-1, C)
    return dw, d2.sum(axis=0)


def weighted_mse(pred, target, weight):
    """loss = sum w (pred - target)^2 and its gradient 2 w (pred - target)."""
    d = np.asarray(pred, np.float64) - np.asarray(target, np.float64)
    w = np.asarray(weight, np.float64)
    return float((w * d * d).sum()), 2.0 * w * d


def adam(master, m, v, grad, lr, beta1, beta2, eps, step):
    """One Adam step with bias correction (FP64)."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    c1, c2 = 1.0 - beta1 ** step, 1.0 - beta2 ** step
    return master - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def split(params):
    """Layer parameter block -> (wf[27, 32, 32], bias[32])."""
    p = np.asarray(params, dtype=np.float64)
    return p[:TAPS * C * C].reshape(TAPS, C, C), p[TAPS * C * C:]


def network(features, params, *, emulate_bf16=True):
    """Forward of the four-layer network and the activations it keeps.

    With ``emulate_bf16`` the weights, the input and every hidden activation
    are rounded to BF16 as the kernels store them (accumulation stays exact
    here).  Returns (prediction[nx, ny, nz], activations [x, a1, a2, a3])."""
    layers = len(params) // LAYER_PARAMS
    rnd = bf16_round if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    a = rnd(features).astype(np.float64)
    acts = [a]
    for layer in range(layers):
        wf, b = split(params[layer * LAYER_PARAMS:(layer + 1) * LAYER_PARAMS])
        pre = conv3d(a, rnd(wf), b)
        if layer + 1 < layers:
            a = rnd(np.maximum(pre, 0.0)).astype(np.float64)
            acts.append(a)
        else:
            return pre[..., 0], acts
    raise ValueError("no layers")


def network_gradient(features, params, target, weight, *, emulate_bf16=True):
    """Loss and parameter gradient of the weighted squared error, by the
    same backward pass the kernels run (BF16 storage of every gradient
    tensor when ``emulate_bf16``)."""
    layers = len(params) // LAYER_PARAMS
    rnd = bf16_round if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    pred, acts = network(features, params, emulate_bf16=emulate_bf16)
    loss, dpred = weighted_mse(pred, target, weight)
    dpre = np.zeros(pred.shape + (C,), dtype=np.float64)
    dpre[..., 0] = rnd(dpred)
    grad = np.zeros(len(params), dtype=np.float64)
    for layer in reversed(range(layers)):
        dw, db = conv3d_weight_grad(acts[layer], dpre)
        grad[layer * LAYER_PARAMS:layer * LAYER_PARAMS + TAPS * C * C] = dw.ravel()
        grad[layer * LAYER_PARAMS + TAPS * C * C:(layer + 1) * LAYER_PARAMS] = db
        if layer > 0:
            wf, _ = split(params[layer * LAYER_PARAMS:(layer + 1) * LAYER_PARAMS])
            da = conv3d_backward_data(dpre, rnd(wf))
            dpre = rnd(np.where(acts[layer] > 0.0, da, 0.0)).astype(np.float64)
    return loss, grad, pred
