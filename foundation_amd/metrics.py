"""Image error metric of the path: relative mean squared error.

relMSE(x, y) = mean over pixels and channels of (x - y)^2 / (y^2 + eps),
eps = 0.01, the form used to compare renders with a converged reference
(Rousselle et al., "Robust Denoising using Feature and Color Information",
2013; also the loss of the radiance cache).  Evaluated in float64.
"""
import numpy as np

REL_MSE_EPS = 0.01


def rel_mse(img, ref, eps=REL_MSE_EPS):
    x = np.asarray(img, np.float64)
    y = np.asarray(ref, np.float64)
    if x.shape != y.shape:
        raise ValueError(f"shape mismatch {x.shape} vs {y.shape}")
    return float(np.mean((x - y) ** 2 / (y * y + eps)))
