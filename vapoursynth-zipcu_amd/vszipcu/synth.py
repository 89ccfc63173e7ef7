"""Seeded synthetic frames for benchmarks and parity tests.

Clean luma: two linear gradients (random direction, amplitude U[0, 0.5]) as the base, then
40 axis-aligned rectangles painted in order (size U[4, W/4] x U[4, H/4], value U[0, 1]), then
three sinusoidal gratings added (period U[4, 32] px, random orientation and phase, amplitude
U[0, 0.25]), clipped to [0, 1]. Chroma: gradients plus painted rectangles, box-blurred, scaled
to [-0.5, 0.5]. Noise: iid N(0, (sigma/255)^2) from Box-Muller in float64, cast to float32.

Every draw comes from numpy's PCG64 stream for the given seed, so any machine with the same
numpy produces the same bytes.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _gradients(rng, h, w, count=2):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.zeros((h, w), np.float64)
    for _ in range(count):
        theta = rng.random() * 2.0 * np.pi
        amp = rng.random() * 0.5
        proj = (x / max(w - 1, 1)) * np.cos(theta) + (y / max(h - 1, 1)) * np.sin(theta)
        span = proj.max() - proj.min()
        out += amp * ((proj - proj.min()) / span if span > 0 else 0.0)
    return out


def _rectangles(rng, img, count=40):
    h, w = img.shape
    for _ in range(count):
        rw = 4 + int(rng.random() * max(w // 4 - 4, 1))
        rh = 4 + int(rng.random() * max(h // 4 - 4, 1))
        x0 = int(rng.random() * max(w - rw, 1))
        y0 = int(rng.random() * max(h - rh, 1))
        img[y0:y0 + rh, x0:x0 + rw] = rng.random()
    return img


def _gratings(rng, h, w, count=3):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.zeros((h, w), np.float64)
    for _ in range(count):
        period = 4.0 + rng.random() * 28.0
        phi = rng.random() * np.pi
        phase = rng.random() * 2.0 * np.pi
        amp = rng.random() * 0.25
        out += amp * np.sin(2.0 * np.pi * (x * np.cos(phi) + y * np.sin(phi)) / period + phase)
    return out


def clean_luma(width: int, height: int, seed: int) -> np.ndarray:
    rng = _rng(seed)
    img = _gradients(rng, height, width)
    img = _rectangles(rng, img)
    img = img + _gratings(rng, height, width)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def _box_blur(img, r):
    k = 2 * r + 1
    p = np.pad(img, r, mode="edge")
    c = np.cumsum(np.cumsum(p, axis=0), axis=1)
    c = np.pad(c, ((1, 0), (1, 0)))
    h, w = img.shape
    s = c[k:k + h, k:k + w] - c[0:h, k:k + w] - c[k:k + h, 0:w] + c[0:h, 0:w]
    return s / (k * k)


def clean_chroma(width: int, height: int, seed: int) -> np.ndarray:
    rng = _rng(seed)
    img = _gradients(rng, height, width)
    img = _rectangles(rng, img, count=20)
    img = _box_blur(img, max(2, min(width, height) // 64))
    lo, hi = img.min(), img.max()
    img = (img - lo) / (hi - lo) - 0.5 if hi > lo else img * 0.0
    return img.astype(np.float32)


def gaussian_noise(shape, sigma255: float, seed: int) -> np.ndarray:
    """iid N(0, (sigma/255)^2) by Box-Muller in float64, cast to float32."""
    rng = _rng(seed)
    n = int(np.prod(shape))
    m = (n + 1) // 2
    u1 = 1.0 - rng.random(m)  # (0, 1]
    u2 = rng.random(m)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)])[:n]
    return (z * (sigma255 / 255.0)).reshape(shape).astype(np.float32)


def noisy(clean: np.ndarray, sigma255: float, seed: int) -> np.ndarray:
    return (clean + gaussian_noise(clean.shape, sigma255, seed)).astype(np.float32)


def gray_frame(width, height, sigma255, seed=1):
    """(clean, noisy) float32 luma planes."""
    c = clean_luma(width, height, seed)
    return c, noisy(c, sigma255, seed + 1000)


def yuv444_frame(width, height, sigmas255=(5.0, 3.0, 3.0), seed=2):
    """(clean [3,H,W], noisy [3,H,W]) float32; Y from the luma generator, U/V low-frequency."""
    clean = np.stack([clean_luma(width, height, seed),
                      clean_chroma(width, height, seed + 1),
                      clean_chroma(width, height, seed + 2)])
    nz = np.stack([noisy(clean[p], sigmas255[p], seed + 1000 + p) for p in range(3)])
    return clean, nz


def moving_sequence(width, height, frames, sigma255, seed=3, n_rects=20, max_shift=3):
    """(clean [F,H,W], noisy [F,H,W]): a luma frame under a global integer translation of
    U{-max_shift..max_shift} px per frame, plus rectangles moving at their own integer
    velocities."""
    rng = _rng(seed)
    pad = 4 * max_shift * frames + 8
    base = clean_luma(width + 2 * pad, height + 2 * pad, seed)
    rects = []
    for _ in range(n_rects):
        rects.append((int(rng.random() * width), int(rng.random() * height),
                      4 + int(rng.random() * max(width // 8, 1)), 4 + int(rng.random() * max(height // 8, 1)),
                      int(rng.integers(-3, 4)), int(rng.integers(-3, 4)), float(rng.random())))
    ox = oy = 0
    clean = np.empty((frames, height, width), np.float32)
    for f in range(frames):
        img = base[pad + oy:pad + oy + height, pad + ox:pad + ox + width].copy()
        for (x0, y0, rw, rh, vx, vy, val) in rects:
            x = (x0 + vx * f) % width
            y = (y0 + vy * f) % height
            img[y:y + rh, x:x + rw] = val
        clean[f] = img
        ox += int(rng.integers(-max_shift, max_shift + 1))
        oy += int(rng.integers(-max_shift, max_shift + 1))
        ox = max(-pad + 1, min(pad - 1, ox))
        oy = max(-pad + 1, min(pad - 1, oy))
    nz = np.stack([noisy(clean[f], sigma255, seed + 100 + f) for f in range(frames)])
    return clean, nz


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)
