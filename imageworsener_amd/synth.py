"""Deterministic synthetic inputs for the benchmark and parity configurations.

Every value comes from IEEE-754 double arithmetic done element-wise by numpy
(no libm, no fused operations) driven by SplitMix64, so the same seed gives
the same bytes on any host.  Draw order is part of each recipe.

  c0(): 640x480 RGB8 sRGB, seed 1 (left: bilinear gradient between 4 random
        corner colours; right: uniform noise with a 16x16 random checkerboard
        of hard 0/255 blocks)
  c1(): 4096x4096 RGBA16, seed 2 (RGB: 32768 + three random sinusoid gratings
        of amplitude 10922 + uniform noise in [-2048, 2048], clamped; alpha:
        64x64 tiles of uniform values, 10% of tiles exactly 0)
  c2(): 2048x2048 Gray16 linear, seed 3 (radial ramp 0..65535 + uniform noise
        in [-512, 512], clamped)
  c3_list()/c3_image(): 4096 RGBA8 images, seed 4, sides log-uniform in
        [64, 1024], per-axis scale U[0.25, 2], filter from {triangle, catrom,
        lanczos2, lanczos3}
  c4_list()/c4_image(): serving requests, seed 5: RGB 8- or 16-bit (p=0.5),
        sides log-uniform in [256, 8192], C1's RGB recipe, one scale U[0.1,
        1.5], Lanczos3 or Mitchell (p=0.5), Poisson arrival gaps
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
TWO_PI = 6.283185307179586
PI = 3.141592653589793


class SplitMix64:
    """SplitMix64; draw n is mix(seed + (n+1)*golden), so blocks of draws are
    produced vectorised."""

    def __init__(self, seed):
        self.state = np.uint64(seed)
        self.n = 0

    def u64(self, count):
        with np.errstate(over="ignore"):
            idx = np.arange(self.n + 1, self.n + 1 + count, dtype=np.uint64)
            z = self.state + idx * GOLDEN
            z = (z ^ (z >> np.uint64(30))) * M1
            z = (z ^ (z >> np.uint64(27))) * M2
            z = z ^ (z >> np.uint64(31))
        self.n += count
        return z

    def uniform(self, count):
        """Doubles in [0, 1) from the top 53 bits."""
        return (self.u64(count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def ints(self, count, lo, hi):
        """Integers in [lo, hi] (inclusive)."""
        return np.floor(self.uniform(count) * (hi - lo + 1)).astype(np.int64) + lo


def sin_poly(x):
    """sin(x) from an IEEE-only recipe: reduce to [-pi, pi] with a two-part
    2*pi, fold to [-pi/2, pi/2], odd Taylor polynomial to x^15."""
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x * (1.0 / TWO_PI))
    r = (x - k * 6.28125) - k * 0.0019353071795864769
    r = np.where(r > PI / 2, PI - r, r)
    r = np.where(r < -PI / 2, -PI - r, r)
    r2 = r * r
    p = -1.0 / 1307674368000.0
    for c in (1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0, 1.0 / 120.0, -1.0 / 6.0, 1.0):
        p = p * r2 + c
    return r * p


def _round_clip(v, hi):
    return np.clip(np.floor(v + 0.5), 0, hi)


def c0(width=640, height=480, seed=1):
    """RGB8 interleaved, shape (height, width, 3)."""
    rng = SplitMix64(seed)
    corners = rng.ints(12, 0, 255).astype(np.float64).reshape(4, 3)
    half = width // 2
    out = np.zeros((height, width, 3), np.float64)
    fx = np.arange(half, dtype=np.float64) / max(half - 1, 1)
    fy = np.arange(height, dtype=np.float64) / max(height - 1, 1)
    FX, FY = fx[None, :, None], fy[:, None, None]
    grad = (corners[0] * ((1.0 - FX) * (1.0 - FY)) + corners[1] * (FX * (1.0 - FY)) +
            corners[2] * ((1.0 - FX) * FY) + corners[3] * (FX * FY))
    out[:, :half] = _round_clip(grad, 255)
    out[:, half:] = rng.ints(height * (width - half) * 3, 0, 255).reshape(height, width - half, 3)
    bh, bw = (height + 15) // 16, (width - half + 15) // 16
    coins = rng.uniform(bh * bw).reshape(bh, bw) < 0.5
    by, bx = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
    val = np.where((by + bx) % 2 == 0, 255.0, 0.0)
    blk = np.repeat(np.repeat(np.where(coins, val, -1.0), 16, 0), 16, 1)[:height, :width - half]
    right = out[:, half:]
    right[blk >= 0] = blk[blk >= 0][:, None]
    return out.astype(np.uint8)


def c1(size=4096, seed=2, tile=64):
    """RGBA16 samples (uint16, shape (size, size, 4)); gratings are separable:
    sin(a x + b y + phi) = sin(a x + phi) cos(b y) + cos(a x + phi) sin(b y)."""
    rng = SplitMix64(seed)
    out = np.empty((size, size, 4), np.uint16)
    coords = np.arange(size, dtype=np.float64)
    acc = []
    for ch in range(3):
        params = rng.uniform(9).reshape(3, 3)
        v = np.full((size, size), 32768.0)
        for f_u, th_u, ph_u in params:
            f = 1.0 + f_u * 255.0
            th = th_u * PI
            ph = ph_u * TWO_PI
            cth, sth = sin_poly(th + PI / 2), sin_poly(th)
            ax = (TWO_PI * f / size) * cth * coords + ph
            by = (TWO_PI * f / size) * sth * coords
            sx, cx = sin_poly(ax), sin_poly(ax + PI / 2)
            sy, cy = sin_poly(by), sin_poly(by + PI / 2)
            v += 10922.0 * (sy[:, None] * cx[None, :] + cy[:, None] * sx[None, :])
        acc.append(v)
    noise = rng.ints(size * size * 3, -2048, 2048).reshape(size, size, 3)
    for ch in range(3):
        out[..., ch] = _round_clip(acc[ch] + noise[..., ch], 65535).astype(np.uint16)
    nt = (size + tile - 1) // tile
    u = rng.uniform(nt * nt * 2).reshape(nt * nt, 2)
    vals = np.floor(u[:, 0] * 65536.0)
    vals[u[:, 1] < 0.1] = 0.0
    alpha = np.repeat(np.repeat(vals.reshape(nt, nt), tile, 0), tile, 1)[:size, :size]
    out[..., 3] = alpha.astype(np.uint16)
    return out


def c2(size=2048, seed=3):
    """Gray16 samples (uint16, shape (size, size))."""
    rng = SplitMix64(seed)
    c = (size - 1) / 2.0
    y, x = np.meshgrid(np.arange(size, dtype=np.float64) - c, np.arange(size, dtype=np.float64) - c,
                       indexing="ij")
    r = np.sqrt(x * x + y * y)
    rmax = np.sqrt(2.0 * c * c)
    v = 65535.0 * r / rmax + rng.ints(size * size, -512, 512).reshape(size, size)
    return _round_clip(v, 65535).astype(np.uint16)


def be16(samples):
    """uint16 samples -> big-endian bytes, the library's 16-bit layout."""
    return np.ascontiguousarray(samples, dtype=">u2").view(np.uint8)


C3_FILTERS = ["triangle", "catrom", "lanczos2", "lanczos3"]


def c3_list(count=4096, seed=4):
    """[(w, h, out_w, out_h, filter, image_seed)] for the batched config."""
    rng = SplitMix64(seed)
    u = rng.uniform(count * 5).reshape(count, 5)
    lo, hi = np.log(64.0), np.log(1024.0)
    out = []
    for i in range(count):
        w = int(np.floor(np.exp(lo + u[i, 0] * (hi - lo)) + 0.5))
        h = int(np.floor(np.exp(lo + u[i, 1] * (hi - lo)) + 0.5))
        sx = 0.25 + u[i, 2] * 1.75
        sy = 0.25 + u[i, 3] * 1.75
        ow = max(1, int(np.floor(w * sx + 0.5)))
        oh = max(1, int(np.floor(h * sy + 0.5)))
        filt = C3_FILTERS[min(3, int(u[i, 4] * 4))]
        out.append((w, h, ow, oh, filt, 1000 + i))
    return out


def c3_image(w, h, seed):
    """RGBA8 image for C3: C0's recipe for RGB at this size, alpha from 16x16
    tiles of uniform values with 10% exactly 0."""
    rgb = c0(w, h, seed)
    rng = SplitMix64(seed ^ 0x5A5A)
    th, tw = (h + 15) // 16, (w + 15) // 16
    u = rng.uniform(th * tw * 2).reshape(th * tw, 2)
    vals = np.floor(u[:, 0] * 256.0)
    vals[u[:, 1] < 0.1] = 0.0
    a = np.repeat(np.repeat(vals.reshape(th, tw), 16, 0), 16, 1)[:h, :w]
    return np.concatenate([rgb, a[..., None].astype(np.uint8)], axis=2)


def gratings_rgb(w, h, depth, seed):
    """C1's RGB recipe at w x h: three sinusoid gratings of amplitude 1/6 of
    full scale around mid-grey, plus uniform noise of 1/32 of full scale,
    clamped; uint8 or uint16 samples, shape (h, w, 3)."""
    rng = SplitMix64(seed)
    full = 255.0 if depth == 8 else 65535.0
    mid, amp = (full + 1.0) / 2.0, (full + 1.0) / 6.0
    nz = int(full + 1.0) // 32
    xs, ys = np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64)
    size = float(max(w, h))
    out = np.empty((h, w, 3), np.uint8 if depth == 8 else np.uint16)
    for ch in range(3):
        params = rng.uniform(9).reshape(3, 3)
        v = np.full((h, w), mid)
        for f_u, th_u, ph_u in params:
            f = 1.0 + f_u * 255.0
            th, ph = th_u * PI, ph_u * TWO_PI
            cth, sth = sin_poly(th + PI / 2), sin_poly(th)
            ax = (TWO_PI * f / size) * cth * xs + ph
            by = (TWO_PI * f / size) * sth * ys
            sx, cx = sin_poly(ax), sin_poly(ax + PI / 2)
            sy, cy = sin_poly(by), sin_poly(by + PI / 2)
            v += amp * (sy[:, None] * cx[None, :] + cy[:, None] * sx[None, :])
        v += rng.ints(w * h, -nz, nz).reshape(h, w)
        out[..., ch] = _round_clip(v, full).astype(out.dtype)
    return out


C4_FILTERS = ["lanczos3", "mitchell"]


def c4_list(count=1000, seed=5, lo_side=256, hi_side=8192):
    """[(w, h, depth, out_w, out_h, filter, image_seed, gap)] for the serving
    config; gap is the exponential inter-arrival time in units of 1/rate."""
    rng = SplitMix64(seed)
    u = rng.uniform(count * 6).reshape(count, 6)
    lo, hi = np.log(float(lo_side)), np.log(float(hi_side))
    out = []
    for i in range(count):
        w = int(np.floor(np.exp(lo + u[i, 0] * (hi - lo)) + 0.5))
        h = int(np.floor(np.exp(lo + u[i, 1] * (hi - lo)) + 0.5))
        depth = 8 if u[i, 2] < 0.5 else 16
        sc = 0.1 + u[i, 3] * 1.4
        ow, oh = max(1, int(np.floor(w * sc + 0.5))), max(1, int(np.floor(h * sc + 0.5)))
        filt = C4_FILTERS[0 if u[i, 4] < 0.5 else 1]
        gap = -np.log(1.0 - u[i, 5])
        out.append((w, h, depth, ow, oh, filt, 2000 + i, gap))
    return out
