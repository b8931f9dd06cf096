"""Seeded synthetic inputs for the codec's benchmark workloads and tests.

Two families:

* ``exact_field``: a smooth-plus-noise float field built from numpy's PCG64
  stream and IEEE-754 add/multiply/divide only, so the same seed gives the
  same bits on every machine.  Golden fixtures are generated from it.
* ``sinusoid_field``: the benchmark recipe -- a sum of random 3-D sinusoids on
  the unit grid, an optional linear ramp and Gaussian or lognormal noise,
  computed in float64 (on a torch device when given) and rounded to the
  target dtype.  ``inject_specials`` replaces a seeded fraction of positions
  by NaN (with payload), +Inf, -Inf, a denormal and -0.0, round-robin.

The parameters come from a numpy ``default_rng(seed)`` stream, so the recipe
is reproducible; the transcendental functions make the low bits of large
fields platform-dependent, which is why parity tests always compare against
the oracle on the very array they generated.
"""
import math

import numpy as np

__all__ = ["exact_field", "sinusoid_field", "inject_specials", "config_field", "batch_sizes", "array_field"]


def exact_field(n, seed, dtype=np.float32, noise=1e-3, scale=10.0, offset=0.0):
    """Smooth polynomial/triangle field plus Irwin-Hall noise, bit-reproducible."""
    rng = np.random.default_rng(seed)
    c = rng.random(6) * 2.0 - 1.0
    t = np.arange(n, dtype=np.float64) / float(max(n, 1))
    f = c[0] + t * (c[1] + t * (c[2] + t * c[3]))
    # triangle wave with a random period: |2 * frac(t * p) - 1|
    p = 3.0 + 20.0 * rng.random()
    tp = t * p
    tri = np.abs((tp - np.floor(tp)) * 2.0 - 1.0)
    f = (f + c[4] * tri) * scale + offset
    u = rng.random((4, n)).sum(axis=0) - 2.0
    f = f + u * noise
    return f.astype(dtype)


def _torch():
    import torch

    return torch


def sinusoid_field(shape, seed, dtype=np.float32, n_terms=8, noise="normal", noise_sigma=1e-3,
                   noise_mu=-10.0, ramp=None, device=None):
    """Sum of ``n_terms`` random 3-D sinusoids on the grid i/N.

    amplitude ~ U[0.1, 10], wavenumbers ~ U[0.5, 8], phase ~ U[0, 2*pi).
    ``noise``: "normal" adds N(0, noise_sigma); "lognormal" adds
    +-exp(N(noise_mu, noise_sigma)).  ``ramp=(r0, r1)`` adds a linear ramp
    from r0 to r1 along the main diagonal.  Returns a torch tensor on
    ``device`` when given, else a numpy array.
    """
    torch = _torch()
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    dev = torch.device(device) if device is not None else torch.device("cpu")
    amp = rng.uniform(0.1, 10.0, n_terms)
    kx, ky, kz = (rng.uniform(0.5, 8.0, n_terms) for _ in range(3))
    ph = rng.uniform(0.0, 2.0 * math.pi, n_terms)
    x = torch.arange(nx, dtype=torch.float64, device=dev) / nx
    y = torch.arange(ny, dtype=torch.float64, device=dev) / ny
    z = torch.arange(nz, dtype=torch.float64, device=dev) / nz
    tdt = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
    out = torch.empty((nz, ny, nx), dtype=tdt, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
    # slab by slab along z keeps the float64 temporaries small
    slab = max(1, (1 << 24) // max(1, nx * ny))
    for z0 in range(0, nz, slab):
        zz = z[z0:z0 + slab]
        f = torch.zeros((zz.numel(), ny, nx), dtype=torch.float64, device=dev)
        for k in range(n_terms):
            a = 2.0 * math.pi * (kx[k] * x)[None, None, :]
            b = 2.0 * math.pi * (ky[k] * y)[None, :, None]
            c = 2.0 * math.pi * (kz[k] * zz)[:, None, None] + ph[k]
            f += amp[k] * torch.sin(a + b + c)
        if ramp is not None:
            r0, r1 = ramp
            s = (x[None, None, :] + y[None, :, None] + zz[:, None, None]) / 3.0
            f += r0 + (r1 - r0) * s
        if noise == "normal":
            f += noise_sigma * torch.randn(f.shape, dtype=torch.float64, device=dev, generator=gen)
        elif noise == "lognormal":
            mag = torch.exp(noise_mu + noise_sigma * torch.randn(f.shape, dtype=torch.float64, device=dev, generator=gen))
            sgn = torch.where(torch.rand(f.shape, device=dev, generator=gen) < 0.5, -1.0, 1.0).to(torch.float64)
            f += sgn * mag
        out[z0:z0 + zz.numel()] = f.to(tdt)
        del f
    return out if device is not None else out.numpy()


def _special_bits(dtype, rng, count):
    """Bit patterns, round-robin: qNaN with payload, +Inf, -Inf, denormal, -0.0."""
    if np.dtype(dtype) == np.float32:
        it = np.uint32
        nan = 0x7FC00000 | int(rng.integers(1, 1 << 22))
        pos_inf, neg_inf, neg0 = 0x7F800000, 0xFF800000, 0x80000000
        den = rng.integers(1, 1 << 23, size=count, dtype=np.uint64).astype(it)
    else:
        it = np.uint64
        nan = 0x7FF8000000000000 | int(rng.integers(1, 1 << 51))
        pos_inf, neg_inf, neg0 = 0x7FF0000000000000, 0xFFF0000000000000, 0x8000000000000000
        den = rng.integers(1, 1 << 52, size=count, dtype=np.uint64).astype(it)
    out = np.empty(count, dtype=it)
    kinds = np.arange(count) % 5
    out[kinds == 0] = nan
    out[kinds == 1] = pos_inf
    out[kinds == 2] = neg_inf
    out[kinds == 3] = den[kinds == 3]
    out[kinds == 4] = neg0
    return out


def inject_specials(arr, frac, seed):
    """Replace round(frac * n) seeded positions by special values, in place.

    ``arr`` is a numpy array or a torch tensor; returns the positions."""
    n = arr.size if isinstance(arr, np.ndarray) else arr.numel()
    rng = np.random.default_rng(seed ^ 0x5EC1A1)
    count = int(round(frac * n))
    if count == 0:
        return np.zeros(0, dtype=np.int64)
    pos = rng.choice(n, size=count, replace=False)
    if isinstance(arr, np.ndarray):
        dt = arr.dtype
        bits = _special_bits(dt, rng, count)
        flat = arr.reshape(-1).view(bits.dtype)
        flat[pos] = bits
    else:
        torch = _torch()
        dt = np.float32 if arr.dtype == torch.float32 else np.float64
        bits = _special_bits(dt, rng, count)
        ibits = torch.from_numpy(bits.view(np.int32 if dt == np.float32 else np.int64)).to(arr.device)
        iv = arr.reshape(-1).view(torch.int32 if dt == np.float32 else torch.int64)
        iv[torch.from_numpy(pos).to(arr.device)] = ibits
    return pos


def finite_range(arr):
    """max - min over finite values, in the data's precision."""
    if isinstance(arr, np.ndarray):
        f = arr[np.isfinite(arr)]
        if f.size == 0:
            return arr.dtype.type(0)
        return arr.dtype.type(f.max() - f.min())
    torch = _torch()
    f = arr[torch.isfinite(arr)]
    if f.numel() == 0:
        return 0.0
    return float((f.max() - f.min()).item())


def config_field(config, seed=0, device=None, scale=1.0):
    """The input array of a benchmark workload.

    config 0: 256^3 float32, 8 sinusoids + N(0, 1e-3)
    config 1: 512^3 float32, as config 0 + diagonal ramp U[-1e3, 1e3] + 0.01 % specials
    config 2: 384^3 float64, 12 sinusoids + lognormal(-10, 2) noise + 0.01 % specials
    ``scale`` < 1 shrinks every axis (for tests).
    """
    rng = np.random.default_rng(seed + 1000 * config)
    if config == 0:
        n = max(1, int(256 * scale))
        return sinusoid_field((n, n, n), seed, np.float32, 8, device=device)
    if config == 1:
        n = max(1, int(512 * scale))
        ramp = tuple(rng.uniform(-1e3, 1e3, 2))
        f = sinusoid_field((n, n, n), seed, np.float32, 8, ramp=ramp, device=device)
        inject_specials(f, 1e-4, seed)
        return f
    if config == 2:
        n = max(1, int(384 * scale))
        f = sinusoid_field((n, n, n), seed, np.float64, 12, noise="lognormal", noise_sigma=2.0,
                           noise_mu=-10.0, device=device)
        inject_specials(f, 1e-4, seed)
        return f
    raise ValueError(f"unknown config {config}")


def batch_sizes(n_arrays, seed, lo=4096, hi=64 << 20, f32_frac=0.7):
    """Log-uniform array sizes in bytes and dtypes for the batched workload.

    Returns a list of (nbytes, dtype) with nbytes a multiple of the element size."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), n_arrays))
    is32 = rng.random(n_arrays) < f32_frac
    out = []
    for s, f in zip(sizes, is32):
        es = 4 if f else 8
        out.append((max(es, int(s) // es * es), np.float32 if f else np.float64))
    return out


def array_field(nbytes, dtype, seed, device=None):
    """One array of the batched and serving workloads: a float32 array is
    built like the 256^3 field (8 sinusoids + N(0, 1e-3)), a float64 one like
    the 384^3 field (12 sinusoids + lognormal noise, 0.01 % specials), on the
    smallest cube holding nbytes / itemsize values, flattened and cut."""
    it = np.dtype(dtype).itemsize
    n = nbytes // it
    side = max(1, int(math.ceil(n ** (1.0 / 3.0) - 1e-9)))
    while side ** 3 < n:
        side += 1
    if np.dtype(dtype) == np.float32:
        f = sinusoid_field((side, side, side), seed, np.float32, 8, device=device)
    else:
        f = sinusoid_field((side, side, side), seed, np.float64, 12, noise="lognormal", noise_sigma=2.0,
                           noise_mu=-10.0, device=device)
    f = f.reshape(-1)[:n]
    f = f.clone() if device is not None else f.copy()
    if np.dtype(dtype) == np.float64:
        inject_specials(f, 1e-4, seed)
    return f
