"""Deterministic synthetic dynamic-MRI inputs for the benchmark workloads.

Everything is generated from a seed with numpy's PCG64 so that the HIP path,
the oracle and every box see identical inputs:

* phantom: ellipses whose centres and semi-axes move sinusoidally over time
  (3 px amplitude), optionally over a static background with a periodic
  component of a given cycle length, scaled to a maximum magnitude of 255
  (the range the reference normalises its data to, matlab/prepare_data.m:40-53);
* coil maps: Gaussian magnitude bumps (sigma = 0.4 FOV) centred on a circle,
  with a linear phase, normalised to sum-of-squares 1;
* masks: variable-density random phase-encode lines (rows) per frame with a
  fully sampled centre, optionally a per-frame shifted pattern;
* k-space: d = K rho + M n with complex Gaussian noise n, sigma = 1% of
  max |K rho|, in the reference's uncentred FFT order
  (CUDA/src/main.cpp:755-756 builds the operator with centered = false);
* initial image u0 = sum_c conj(b1_c) F_ortho(sum_t d_{t,c} / max(1, sum_t M_t))
  (the time-averaged zero-filled image of matlab/get_sos_crec_opt.m:7-15).

Layout: image [T, H, W]; k-space [T, C, H, W]; mask [T, H, W]; coils [C, H, W].
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np


@dataclasses.dataclass
class Workload:
    name: str
    method: str
    n: int
    frames: int
    coils: int
    accel: float
    center_lines: int
    ellipses: int
    seed: int = 0
    shifted: bool = False
    cycle: int | None = None
    static_background: bool = False
    iterations: int = 500
    params: dict = dataclasses.field(default_factory=dict)


# The benchmark configurations (matrix size, frames, coils, acceleration, ...).
CONFIGS = {
    "tv_128x12x8": Workload("tv_128x12x8", "tv", 128, 12, 8, 8, 12, 6, seed=0, iterations=500,
                            params={"lambda": 0.01}),
    "tgv2_256x25x12": Workload("tgv2_256x25x12", "tgv2", 256, 25, 12, 10, 16, 10, seed=1,
                               iterations=1000, params={"alpha0": 2.0, "alpha1": 1.0}),
    "ictgv2_256x30x16": Workload("ictgv2_256x30x16", "ictgv2", 256, 30, 16, 12, 16, 10, seed=2,
                                 shifted=True, cycle=15, static_background=True,
                                 iterations=1000),
}


def _ellipse(yy, xx, cy, cx, a, b, th):
    c, s = np.cos(th), np.sin(th)
    u = (xx - cx) * c + (yy - cy) * s
    v = -(xx - cx) * s + (yy - cy) * c
    return (u / a) ** 2 + (v / b) ** 2 <= 1.0


def phantom(n, T, n_ellipses, seed, amp=3.0, cycle=None, static_background=False):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    yy -= n / 2.0
    xx -= n / 2.0
    period = float(cycle if cycle else T)
    img = np.zeros((T, n, n), np.complex128)
    if static_background:
        bg = 0.4 * _ellipse(yy, xx, 0, 0, 0.42 * n, 0.36 * n, 0.0)
        bg = bg + 0.2 * _ellipse(yy, xx, 0.05 * n, -0.05 * n, 0.30 * n, 0.22 * n, 0.3)
        img += bg[None]
    for _ in range(n_ellipses):
        cy, cx = rng.uniform(-0.25 * n, 0.25 * n, 2)
        a, b = rng.uniform(0.04 * n, 0.16 * n, 2)
        th = rng.uniform(0, np.pi)
        val = rng.uniform(0.2, 1.0) * np.exp(1j * rng.uniform(-0.3, 0.3))
        ph = rng.uniform(0, 2 * np.pi, 3)
        for t in range(T):
            w = 2 * np.pi * t / period
            ccy = cy + amp * np.sin(w + ph[0])
            ccx = cx + amp * np.cos(w + ph[1])
            da = amp * np.sin(w + ph[2])
            img[t] += val * _ellipse(yy, xx, ccy, ccx, max(a + da, 1.0), max(b + da, 1.0), th)
    img *= 255.0 / max(np.abs(img).max(), 1e-12)
    return img


def coil_maps(n, C, seed, sigma=0.4):
    rng = np.random.default_rng(seed + 1000)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64) / n - 0.5
    maps = np.zeros((C, n, n), np.complex128)
    for c in range(C):
        ang = 2 * np.pi * c / C
        cy, cx = 0.5 * np.sin(ang), 0.5 * np.cos(ang)
        mag = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sigma ** 2))
        ky, kx = rng.uniform(-np.pi, np.pi, 2)
        maps[c] = mag * np.exp(1j * (ky * yy + kx * xx + rng.uniform(0, 2 * np.pi)))
    maps /= np.sqrt((np.abs(maps) ** 2).sum(0))[None]
    return maps


def vd_line_mask(n, T, accel, center_lines, seed, shifted=False):
    """Variable-density phase-encode line mask [T, n, n] in centred
    coordinates (row n//2 is k_y = 0)."""
    rng = np.random.default_rng(seed + 2000)
    per_frame = int(round(n / accel))
    c0 = n // 2 - center_lines // 2
    centre = np.arange(c0, c0 + center_lines)
    outer = np.setdiff1d(np.arange(n), centre)
    dist = np.abs(outer - n / 2.0) / (n / 2.0)
    w = (1.0 - dist) ** 2 + 1e-3
    w /= w.sum()
    n_out = max(per_frame - center_lines, 0)
    mask = np.zeros((T, n, n), np.float64)
    base = rng.choice(outer, size=n_out, replace=False, p=w) if n_out else np.array([], int)
    for t in range(T):
        if shifted:
            rows = (base + t) % n
            rows = np.where(np.isin(rows, centre), (rows + center_lines) % n, rows)
        else:
            rows = rng.choice(outer, size=n_out, replace=False, p=w) if n_out else base
        mask[t, centre, :] = 1.0
        mask[t, rows, :] = 1.0
    return mask


def _ifft2(a):
    try:
        import scipy.fft as sf
        return sf.ifft2(a, norm="ortho", workers=min(16, os.cpu_count() or 1))
    except ImportError:
        return np.fft.ifft2(a, norm="ortho")


def kspace(img, b1, mask_unc):
    """K rho: uncentred, unitary inverse DFT (+i exponent), masked."""
    T, H, W = img.shape
    C = b1.shape[0]
    out = np.empty((T, C, H, W), np.complex128)
    for t in range(T):
        out[t] = _ifft2(b1 * img[t][None]) * mask_unc[t][None]
    return out


def initial_image(data, b1, mask_unc):
    T, C, H, W = data.shape
    num = data.sum(0)
    den = np.maximum(1.0, mask_unc.sum(0))
    kavg = num / den[None]
    return (np.conj(b1) * np.fft.fft2(kavg, norm="ortho")).sum(0)


def make(wl: Workload | str, with_truth=False):
    """Returns dict(data, b1, mask, x0[, truth]) in uncentred k-space order,
    plus the workload description."""
    if isinstance(wl, str):
        wl = CONFIGS[wl]
    n, T, C = wl.n, wl.frames, wl.coils
    rho = phantom(n, T, wl.ellipses, wl.seed, cycle=wl.cycle,
                  static_background=wl.static_background)
    b1 = coil_maps(n, C, wl.seed)
    mask_c = vd_line_mask(n, T, wl.accel, wl.center_lines, wl.seed, wl.shifted)
    mask = np.fft.ifftshift(mask_c, axes=(1, 2))
    data = kspace(rho, b1, mask)
    rng = np.random.default_rng(wl.seed + 3000)
    sig = 0.01 * np.abs(data).max()
    # complex Gaussian noise on the sampled phase-encode rows only
    rows = mask.any(axis=2)  # [T, H]
    tt, yy = np.nonzero(rows)
    shape = (len(tt), C, n)
    noise = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * sig
    data[tt, :, yy, :] += noise * mask[tt, yy, :][:, None, :]
    u0 = initial_image(data, b1, mask)
    out = dict(data=data, b1=b1, mask=mask, x0=np.repeat(u0[None], T, axis=0), workload=wl)
    if with_truth:
        out["truth"] = rho
    return out


# ------------------------------------------------------------- serving stream
SERVE_SIZES = (128, 160, 192, 224, 256)
# acceleration / centre lines / shifted pattern per method, as in the three
# single-reconstruction workloads above
SERVE_SAMPLING = {"tv": (8, 12, False), "tgv2": (10, 16, False), "ictgv2": (12, 16, True)}


def serving_stream(rate=50.0, seconds=2.0, seed=4):
    """Poisson arrivals at `rate` per second over `seconds`: a list of
    (arrival_s, method, n, frames, coils), method uniform over TV / TGV2 /
    ICTGV2, n uniform over SERVE_SIZES, frames uniform 8..30, coils 8..20."""
    rng = np.random.default_rng(seed + 5000)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return out
        out.append((t, str(rng.choice(["tv", "tgv2", "ictgv2"])), int(rng.choice(SERVE_SIZES)),
                    int(rng.integers(8, 31)), int(rng.integers(8, 21))))


def serving_pool(seed=4, frames=30, coils=20):
    """Per matrix size: phantom [frames, n, n] (10 ellipses), coil maps
    [coils, n, n] and, per method, the line mask [frames, n, n] (uncentred).
    Requests slice the first T frames and C coils."""
    pool = {}
    for n in SERVE_SIZES:
        rho = phantom(n, frames, 10, seed + n)
        b1 = coil_maps(n, coils, seed + n)
        masks = {m: np.fft.ifftshift(vd_line_mask(n, frames, a, c, seed + n, s), axes=(1, 2))
                 for m, (a, c, s) in SERVE_SAMPLING.items()}
        pool[n] = dict(rho=rho, b1=b1, masks=masks)
    return pool


# ------------------------------------------------------------ radial TGV2
@dataclasses.dataclass
class RadialWorkload:
    name: str = "tgv2_radial_256x30x12"
    method: str = "tgv2"
    n: int = 256
    frames: int = 30
    coils: int = 12
    n_fe: int = 512          # samples per spoke (2x readout oversampling)
    spokes: int = 21         # spokes per frame, golden-angle increments
    kernel_width: int = 3
    osf: float = 1.5
    ellipses: int = 10
    seed: int = 1
    iterations: int = 1000


RADIAL = RadialWorkload()


def radial_trajectory(n_fe, spokes_per_frame, frames):
    """Golden-angle (111.246 deg) radial spokes, [frames, 2, spokes*n_fe] in
    cycles per pixel, k in [-0.5, 0.5)."""
    ga = np.pi * (np.sqrt(5.0) - 1.0) / 2.0
    r = (np.arange(n_fe) - n_fe / 2.0) / n_fe
    out = np.zeros((frames, 2, spokes_per_frame * n_fe))
    for t in range(frames):
        a = (t * spokes_per_frame + np.arange(spokes_per_frame)) * ga
        out[t, 0] = (r[None, :] * np.cos(a)[:, None]).ravel()
        out[t, 1] = (r[None, :] * np.sin(a)[:, None]).ravel()
    return out


def make_radial_host(wl: RadialWorkload = RADIAL):
    """Host inputs of the radial workload: phantom, coil maps, trajectory and
    density |k|.  The k-space data is K rho + noise, formed with the
    operator itself (on the device) by the caller."""
    rho = phantom(wl.n, wl.frames, wl.ellipses, wl.seed)
    b1 = coil_maps(wl.n, wl.coils, wl.seed)
    traj = radial_trajectory(wl.n_fe, wl.spokes, wl.frames)
    dcf = np.sqrt(traj[:, 0] ** 2 + traj[:, 1] ** 2)
    return dict(rho=rho, b1=b1, traj=traj, dcf=dcf)
