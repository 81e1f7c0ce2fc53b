"""Seeded synthetic YUV420 frames for the benchmark and parity workloads.

Texture (every workload): a plane of 8 random 2-D sinusoids (periods 8-64 px,
amplitude 16, random orientation and phase) around 128, plus N(0, 4^2) noise,
clipped to [0, 255] and rounded to uint8.  U and V use the same recipe at
half resolution with seeds ``seed+1`` and ``seed+2``.

Motion (chunk workloads): frame t is the crop of a larger canvas at the
cumulative offset of per-frame integer motion vectors drawn from U{-8..8}^2,
and 10 % of its pixels are then replaced by U{0..255} ("occlusion").

Conversion to codec input follows the reference harness: 420 -> 444 by
nearest-neighbour upsampling (src/utils/transforms.py:69-80), /255 and -0.5
(test_video.py:115-122), channels concatenated frame-major for a chunk
(test_video.py:112), a short chunk padded to 8 frames by repeating the last
frame (test_video.py:104-110).
"""
from __future__ import annotations

import numpy as np


def _texture_plane(h, w, seed, n_waves=8, amp=16.0, sigma=4.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    plane = np.full((h, w), 128.0)
    for _ in range(n_waves):
        p = rng.uniform(8.0, 64.0)
        th = rng.uniform(0.0, np.pi)
        ph = rng.uniform(0.0, 2 * np.pi)
        plane += amp * np.sin(2 * np.pi * (xx * np.cos(th) + yy * np.sin(th)) / p + ph)
    plane += rng.normal(0.0, sigma, size=(h, w))
    return np.clip(np.rint(plane), 0, 255).astype(np.uint8)


def texture_yuv420(h, w, seed=0):
    """One uint8 YUV420 frame (y [h,w], u [h/2,w/2], v [h/2,w/2])."""
    y = _texture_plane(h, w, seed)
    u = _texture_plane(h // 2, w // 2, seed + 1)
    v = _texture_plane(h // 2, w // 2, seed + 2)
    return y, u, v


def motion_sequence_yuv420(h, w, n_frames, seed=0, max_mv=8, occlusion=0.10):
    """``n_frames`` YUV420 frames: frame 0 plus ``n_frames-1`` moved frames.

    The canvas is (h+2R) x (w+2R) with R = max_mv * n_frames (even), textured
    once; frame t is a crop at the cumulative motion offset.  Chroma crops use
    half the (even-rounded) luma offset so luma and chroma move together.
    """
    R = max_mv * max(n_frames, 1)
    R += R % 2
    cy, cu, cv = texture_yuv420(h + 2 * R, w + 2 * R, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    oy, ox = 0, 0
    frames = []
    for t in range(n_frames):
        if t > 0:
            dy, dx = rng.integers(-max_mv, max_mv + 1, size=2)
            oy = int(np.clip(oy + dy, -R, R))
            ox = int(np.clip(ox + dx, -R, R))
        ey, ex = (oy // 2) * 2, (ox // 2) * 2
        y = cy[R + ey:R + ey + h, R + ex:R + ex + w].copy()
        u = cu[(R + ey) // 2:(R + ey) // 2 + h // 2, (R + ex) // 2:(R + ex) // 2 + w // 2].copy()
        v = cv[(R + ey) // 2:(R + ey) // 2 + h // 2, (R + ex) // 2:(R + ex) // 2 + w // 2].copy()
        if t > 0 and occlusion > 0:
            m = rng.random((h, w)) < occlusion
            y[m] = rng.integers(0, 256, size=int(m.sum()), dtype=np.uint8)
            mc = rng.random((h // 2, w // 2)) < occlusion
            u[mc] = rng.integers(0, 256, size=int(mc.sum()), dtype=np.uint8)
            mc = rng.random((h // 2, w // 2)) < occlusion
            v[mc] = rng.integers(0, 256, size=int(mc.sum()), dtype=np.uint8)
        frames.append((y, u, v))
    return frames


def yuv420_to_444(y, u, v):
    """Nearest-neighbour chroma upsampling; returns float32 [3, h, w] in [0, 255]."""
    u4 = np.repeat(np.repeat(u, 2, axis=0), 2, axis=1)
    v4 = np.repeat(np.repeat(v, 2, axis=0), 2, axis=1)
    return np.stack([y, u4, v4]).astype(np.float32)


def frame_to_input(yuv):
    """[3, h, w] float32 codec input in [-0.5, 0.5]."""
    return yuv420_to_444(*yuv) / 255.0 - 0.5


def chunk_to_input(frames, frame_delay=8):
    """Concatenate up to ``frame_delay`` frames on channels, padding by repetition."""
    frames = list(frames)
    while len(frames) < frame_delay:
        frames.append(frames[-1])
    return np.concatenate([frame_to_input(f) for f in frames[:frame_delay]], axis=0)


def get_padding_size(height, width, p=16):
    """(padding_right, padding_bottom) to a multiple of p (common_model.py:102-108)."""
    new_h = (height + p - 1) // p * p
    new_w = (width + p - 1) // p * p
    return new_w - width, new_h - height
