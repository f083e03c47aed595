"""Seeded synthetic clips for tests and benchmarks.

Clean plane: clip to [0, 1] of
    0.5 + sum_{k=1..3} 0.1 sin(2 pi (x cos t_k + y sin t_k) / l_k + p_k)
with t ~ U[0, pi), l ~ U[8, 64], p ~ U[0, 2 pi), then 20 axis-aligned
rectangles of constant value v ~ U[0.2, 0.8] painted over it (top-left
uniform in the plane, sides U[W/16, W/4] x U[H/16, H/4]), then clipped again.
Chroma (U, V) planes of a YUV clip use clean - 0.5.

Video: frame t is frame t-1 shifted by an integer (dx, dy) ~ U{-3..3}^2
with edge replication. Noise: iid N(0, (sd)^2) per pixel and frame, not
clipped, rounded to float32.

Random streams (numpy PCG64 via SeedSequence):
    content of plane p     : SeedSequence([seed, p, 0])
    motion                 : SeedSequence([seed, 0, 1])
    noise of plane p, frame t : SeedSequence([seed, p, 2, t])
"""
from __future__ import annotations

import hashlib

import numpy as np


def _rng(*key):
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence(list(key))))


def clean_plane(height: int, width: int, seed: int, plane: int = 0) -> np.ndarray:
    rng = _rng(seed, plane, 0)
    y, x = np.mgrid[0:height, 0:width].astype(np.float64)
    img = np.full((height, width), 0.5)
    for _ in range(3):
        theta = rng.uniform(0.0, np.pi)
        lam = rng.uniform(8.0, 64.0)
        phi = rng.uniform(0.0, 2.0 * np.pi)
        img += 0.1 * np.sin(2.0 * np.pi * (x * np.cos(theta) + y * np.sin(theta)) / lam + phi)
    img = np.clip(img, 0.0, 1.0)
    for _ in range(20):
        w = int(rng.integers(max(1, width // 16), max(2, width // 4) + 1))
        h = int(rng.integers(max(1, height // 16), max(2, height // 4) + 1))
        x0 = int(rng.integers(0, width))
        y0 = int(rng.integers(0, height))
        img[y0:y0 + h, x0:x0 + w] = rng.uniform(0.2, 0.8)
    return np.clip(img, 0.0, 1.0).astype(np.float32)


def _shift(img: np.ndarray, dx: int, dy: int) -> np.ndarray:
    """out[y][x] = img[clamp(y - dy)][clamp(x - dx)] (edge replication)."""
    h, w = img.shape
    ys = np.clip(np.arange(h) - dy, 0, h - 1)
    xs = np.clip(np.arange(w) - dx, 0, w - 1)
    return img[ys][:, xs]


def make_clip(height: int, width: int, frames: int = 1, planes: int = 1, noise_sd: float = 10.0,
              seed: int = 0, motion: bool = False, yuv: bool = False):
    """Returns (clean, noisy) float32 arrays [frames, planes, height, width].
    noise_sd is in 8-bit units (sd = noise_sd / 255)."""
    clean = np.empty((frames, planes, height, width), np.float32)
    base = [clean_plane(height, width, seed, p) for p in range(planes)]
    if yuv:
        for p in range(1, planes):
            base[p] = base[p] - np.float32(0.5)
    mrng = _rng(seed, 0, 1)
    cur = base
    for t in range(frames):
        if t > 0 and motion:
            dx, dy = (int(v) for v in mrng.integers(-3, 4, size=2))
            cur = [_shift(c, dx, dy) for c in cur]
        for p in range(planes):
            clean[t, p] = cur[p]
    noisy = np.empty_like(clean)
    sd = noise_sd / 255.0
    for p in range(planes):
        for t in range(frames):
            nrng = _rng(seed, p, 2, t)
            noisy[t, p] = (clean[t, p].astype(np.float64) +
                           nrng.normal(0.0, sd, size=(height, width))).astype(np.float32)
    return clean, noisy


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


# The 64-clip serving mix: clip k draws its resolution, temporal radius and
# sigma with seed k; content and integer motion as make_clip, noise sd =
# sigma / 255. Every clip uses block_step 4, bm_range 9, ps_num 2, ps_range 4.
BATCH_SIZES = [(360, 640), (720, 1280), (1080, 1920), (2160, 3840)]
BATCH_PARAMS = dict(block_step=4, bm_range=9, ps_num=2, ps_range=4)


def batch_clip_params(k: int):
    """(height, width, radius, sigma) of clip k of the mix."""
    rng = np.random.default_rng(k)
    h, w = BATCH_SIZES[int(rng.integers(0, 4))]
    radius = int(rng.integers(0, 3))
    sigma = float(rng.uniform(5.0, 25.0))
    return h, w, radius, sigma


def batch_clips(n_clips: int = 64, frames: int = 8):
    """The mix as dicts (noisy [frames, 1, H, W] float32, radius, sigma)."""
    clips = []
    for k in range(n_clips):
        h, w, radius, sigma = batch_clip_params(k)
        _, noisy = make_clip(h, w, frames, 1, noise_sd=sigma, seed=k, motion=True)
        clips.append(dict(noisy=noisy, radius=radius, sigma=sigma))
    return clips
