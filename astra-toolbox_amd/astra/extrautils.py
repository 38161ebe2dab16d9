"""Small array utilities (python/astra/extrautils.pyx)."""
import numpy as np


def clipCircle(img):
    """Zero the pixels of square image `img` (in place) outside the disk of
    diameter img.shape[0] centred on the image; returns how many were
    zeroed."""
    n = img.shape[0]
    mid = (n - 1.0) / 2.0
    i = (np.arange(n) - mid) ** 2
    outside = (i[:, None] + i[None, :]) > n * n / 4.0
    img[:n, :n][outside] = 0
    return int(outside.sum())
