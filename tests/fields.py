"""Closed-form fields for the box FGT tests.

For the unnormalised kernel ``G(x) = exp(-|x|^2/delta)`` on the periodic unit
cell, a Fourier mode ``f = cos(k.x + phi)`` with ``k`` in ``2 pi Z^2`` maps to
``pi delta exp(-delta |k|^2 / 4) cos(k.x + phi)``; its gradient and Hessian
follow by differentiation.
"""
import numpy as np


class FourierMode:
    def __init__(self, a: int, b: int, phi: float = 0.3, amp: float = 1.0):
        self.k = 2 * np.pi * np.array([a, b], dtype=np.float64)
        self.phi = phi
        self.amp = amp

    def __call__(self, p):
        return self.amp * np.cos(p[..., 0] * self.k[0] + p[..., 1] * self.k[1] + self.phi)

    def exact(self, p, delta):
        """(pot, grad (..., 2), hess (..., 3) ordered xx, xy, yy)."""
        kx, ky = self.k
        s = self.amp * np.pi * delta * np.exp(-delta * (kx * kx + ky * ky) / 4)
        arg = p[..., 0] * kx + p[..., 1] * ky + self.phi
        c, sn = np.cos(arg), np.sin(arg)
        pot = s * c
        grad = np.stack([-s * sn * kx, -s * sn * ky], axis=-1)
        hess = np.stack([-s * c * kx * kx, -s * c * kx * ky, -s * c * ky * ky], axis=-1)
        return pot, grad, hess


def leaf_rel_err(tree, x, y):
    """Relative L2 error over leaf boxes of arrays shaped (nboxes, ...)."""
    lv = tree.nchild == 0
    num = np.linalg.norm((x - y)[lv])
    den = np.linalg.norm(y[lv])
    return num / den if den > 0 else num
