"""Fidelity of the atom fits: l2_squared_loss (GaussDictCode/transport_loss.py:567-600).

The Wasserstein "Transport" loss of the same module is not part of this
package; problem.getFid raises for it.
"""
from __future__ import annotations

import torch


def _arr(v):
    return v.asarray() if hasattr(v, "asarray") else v


class l2_squared_loss:
    """f(x, y) = 1/2 ||x - y||^2, summed in FP64 on the arrays' device
    (complex arrays: 1/2 sum |x - y|^2, transport_loss.py:569-583);
    grad(x, y, axis=0) = x - y, axis != 0: y - x (:585-597)."""

    is_l2_squared = True

    def __init__(self, dim=None):
        self.dim = dim

    def __call__(self, x, y):
        x, y = _arr(x), _arr(y)
        d = torch.as_tensor(x) - torch.as_tensor(y)
        if d.is_complex():
            d = torch.view_as_real(d.to(torch.complex128))
        d = d.double()
        return float((d * d).sum()) / 2

    def grad(self, x, y, axis=0):
        x, y = _arr(x), _arr(y)
        return x - y if axis == 0 else y - x
