"""FP64 closed form of the reference's Adam kernels (test infrastructure only).

cuda/src/adam_step.cu:21-158: m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
unbiased: p -= (lr / (1-b1^t)) * m / (sqrt(v) / sqrt(1-b2^t) + eps);
biased:   p -= lr * m / (sqrt(v) + eps);
sparse: elements with g == 0 are left untouched (parameter and moments).

Pinning (tests/test_oracle_misc.py): the unbiased dense form equals torch.optim.Adam (a library
routine) step for step; the biased form matches hand-computed values for two steps, and with
lr' = lr sqrt(1-b2^t)/(1-b1^t), eps' = eps sqrt(1-b2^t) it equals torch.optim.Adam step for step;
the sparse form leaves zero-gradient elements untouched.
"""
from __future__ import annotations

import numpy as np


def adam_step(p, g, m, v, step, lr, b1, b2, eps, biased=False, sparse=False):
    p, g, m, v = (np.asarray(x, dtype=np.float64).copy() for x in (p, g, m, v))
    upd = np.ones_like(g, dtype=bool) if not sparse else (g != 0)
    m_new = b1 * m + (1 - b1) * g
    v_new = b2 * v + (1 - b2) * g * g
    if biased:
        p_new = p - lr * m_new / (np.sqrt(v_new) + eps)
    else:
        bc1 = 1 - b1 ** step
        bc2 = 1 - b2 ** step
        p_new = p - (lr / bc1) * m_new / (np.sqrt(v_new) / np.sqrt(bc2) + eps)
    return (np.where(upd, p_new, p), np.where(upd, m_new, m), np.where(upd, v_new, v))
