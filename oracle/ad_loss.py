"""CPU restatement of ADLossFunc.backward (pydtrt/auto_grad.py:10-74), in numpy.

Test infrastructure only: it is the checker of the adjoint's loss-side weighting
and is never called by the product path.
"""
import math

import numpy as np


def downsample(x):
    """auto_grad.py:10-16."""
    if x.shape[0] % 2 == 1:
        x = np.concatenate((x, x[-1:]), axis=0)
    if x.shape[1] % 2 == 1:
        x = np.concatenate((x, x[:, -1:]), axis=1)
    return (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2]) * 0.25


def backward(derivs, grad_input, pyramid_level=1, pyramid_scale=4.0, out_of_range=None, penalty_scale=None,
             clamping=0):
    """Per-parameter gradient of auto_grad.py:46-74: derivs [nder][H][W][3] (image
    channels 1..nder), grad_input [H][W][3] = dL/d(channel 0)."""
    nder = derivs.shape[0]
    oor = np.zeros(nder) if out_of_range is None else np.asarray(out_of_range, np.float64)
    pen = np.ones(nder) if penalty_scale is None else np.asarray(penalty_scale, np.float64)
    lvl = int(min(math.log(derivs.shape[1], 2) + 1, math.log(derivs.shape[2], 2) + 1, pyramid_level))
    ret = np.zeros(nder)
    cur = [np.asarray(derivs[j], np.float64) for j in range(nder)]
    g = np.asarray(grad_input, np.float64)
    for i in range(lvl):
        for j in range(nder):
            if abs(oor[j]) > 1e-4:
                ret[j] = oor[j] * pen[j]
            else:
                gi = g
                if clamping:
                    gi = g.copy()
                    gi[g > 2.0] = 0.0
                    gi[g < -2.0] = 0.0
                ret[j] += pyramid_scale ** i * (cur[j] * gi).sum()
            if i < lvl - 1:
                cur[j] = downsample(cur[j])
        if i < lvl - 1:
            g = downsample(g)
    return ret
