"""Image error metrics used to compare estimators at equal quality."""
from __future__ import annotations

import numpy as np


def rel_mse(img: np.ndarray, ref: np.ndarray, eps: float = 1e-2) -> float:
    """Relative MSE: mean over pixels and RGB of (x - r)^2 / (r^2 + eps)."""
    x = np.asarray(img, np.float64)[..., :3]
    r = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def mad(img: np.ndarray, ref: np.ndarray) -> float:
    """Mean absolute difference over pixels and RGB."""
    return float(np.mean(np.abs(np.asarray(img, np.float64)[..., :3] - np.asarray(ref, np.float64)[..., :3])))
