"""Input recipes of the golden fixtures.

Every input is rebuilt from integer PCG64 draws and IEEE-754 add/multiply/
divide only (``sleek_amd.datagen.exact_field``), so it is bit-identical on
every machine; ``golden.json`` records its SHA-256 next to the stream the
reference CPU codec produced for it.
"""
import numpy as np

from sleek_amd.datagen import exact_field, inject_specials

CS = 16384


def _bits_as(dtype, bits):
    return bits.view(np.dtype(dtype))


def random_bits(n, seed, dtype):
    """Uniformly random bit patterns (all NaN payloads, infinities, denormals)."""
    rng = np.random.default_rng(seed)
    if np.dtype(dtype) == np.float32:
        b = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    else:
        b = rng.integers(0, np.iinfo(np.uint64).max, size=n, dtype=np.uint64, endpoint=True)
    return _bits_as(dtype, b)


def finite_range(a):
    f = a[np.isfinite(a)]
    if f.size == 0:
        return a.dtype.type(0)
    return a.dtype.type(f.max() - f.min())


def build_input(spec):
    """Rebuild a case input from its recipe dict."""
    dt = np.float32 if spec["dtype"] == "f32" else np.float64
    n = int(spec["n"])
    kind = spec["kind"]
    if kind == "zeros":
        return np.zeros(n, dtype=dt)
    if kind == "const":
        return np.full(n, spec["value"], dtype=dt)
    if kind == "field":
        a = exact_field(n, spec["seed"], dt, noise=spec.get("noise", 1e-3), scale=spec.get("scale", 10.0),
                        offset=spec.get("offset", 0.0))
        if spec.get("specials", 0.0) > 0:
            inject_specials(a, spec["specials"], spec["seed"])
        return a
    if kind == "bits":
        return random_bits(n, spec["seed"], dt)
    if kind == "huge":
        # magnitudes near the top of the range, both signs, to exercise the
        # wrap of the sign-folded code for tiny bounds
        rng = np.random.default_rng(spec["seed"])
        if dt == np.float32:
            b = (rng.integers(0x7E000000, 0x7F7FFFFF, size=n, dtype=np.uint64).astype(np.uint32)
                 | (rng.integers(0, 2, size=n, dtype=np.uint64).astype(np.uint32) << np.uint32(31)))
        else:
            b = (rng.integers(0x7FD0000000000000, 0x7FEFFFFFFFFFFFFF, size=n, dtype=np.uint64)
                 | (rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63)))
        a = _bits_as(dt, b)
        if spec.get("sparse", 0.0) > 0:
            keep = rng.random(n) < spec["sparse"]
            a = np.where(keep, a, dt(0)).astype(dt)
        return a
    raise ValueError(kind)


def eb_value(spec, arr):
    """Absolute error bound of a lossy case (None for lossless)."""
    if "eb_rel" in spec:
        r = finite_range(arr)
        return arr.dtype.type(r * arr.dtype.type(spec["eb_rel"]))
    if "eb_abs" in spec:
        return np.dtype(arr.dtype).type(spec["eb_abs"])
    return None


def case_list():
    """The fixture matrix: sizes x dtypes x modes x value mixes."""
    cases = []

    def add(name, **spec):
        spec["name"] = name
        cases.append(spec)

    # stream known answers (zero chunks, one value, a constant chunk)
    add("zeros_1chunk_f32", dtype="f32", n=4096, kind="zeros")
    add("zeros_1chunk_f64", dtype="f64", n=2048, kind="zeros")
    add("zeros_1chunk_f32_lossy", dtype="f32", n=4096, kind="zeros", eb_abs=0.01)
    add("one_value_f32", dtype="f32", n=1, kind="const", value=1.0)
    add("ones_1chunk_f32", dtype="f32", n=4096, kind="const", value=1.0)
    add("ones_1chunk_f64", dtype="f64", n=2048, kind="const", value=1.0)
    add("const_3chunks_f32_lossy", dtype="f32", n=3 * 4096 + 5, kind="const", value=-3.25, eb_abs=1e-3)

    for dt, per in (("f32", 4096), ("f64", 2048)):
        ebs = (1e-2, 1e-3, 1e-4) if dt == "f32" else (1e-6, 1e-9, 1e-3)
        sizes = {"1v": 1, "chunk-1": per - 1, "chunk": per, "3.5chunks": per * 7 // 2, "64chunks": per * 64}
        for sname, n in sizes.items():
            seed = 100 + n + (0 if dt == "f32" else 7)
            add(f"field_{sname}_{dt}", dtype=dt, n=n, kind="field", seed=seed)
            for eb in ebs:
                add(f"field_{sname}_{dt}_eb{eb:g}", dtype=dt, n=n, kind="field", seed=seed, eb_rel=eb)
        # ramped field with many special values
        add(f"special_{dt}", dtype=dt, n=per * 9 + 3, kind="field", seed=777, offset=500.0, specials=0.01)
        for eb in ebs:
            add(f"special_{dt}_eb{eb:g}", dtype=dt, n=per * 9 + 3, kind="field", seed=777, offset=500.0,
                specials=0.01, eb_rel=eb)
        # incompressible bit patterns
        add(f"bits_{dt}", dtype=dt, n=per * 5 + 1, kind="bits", seed=31)
        add(f"bits_{dt}_eb", dtype=dt, n=per * 5 + 1, kind="bits", seed=31, eb_abs=1.0)
        # huge magnitudes with tiny bounds (wrap of the folded code)
        add(f"huge_{dt}_tinyeb", dtype=dt, n=per * 2, kind="huge", seed=5,
            eb_abs=(1e-30 if dt == "f32" else 1e-300))
        # mostly zeros with a few huge magnitudes: compressed chunks holding
        # full-width codes
        add(f"sparse_huge_{dt}", dtype=dt, n=per * 6, kind="huge", seed=6, sparse=0.01)
        add(f"sparse_huge_{dt}_tinyeb", dtype=dt, n=per * 6, kind="huge", seed=6, sparse=0.01,
            eb_abs=(1e-30 if dt == "f32" else 1e-300))
        # a denormal bound: zero quantizes to a non-zero code
        add(f"field_{dt}_denormal_eb", dtype=dt, n=per * 2 + 9, kind="field", seed=9,
            eb_abs=(1e-45 if dt == "f32" else 5e-324))
        # a bound just below the limit where the reference exits
        add(f"field_{dt}_huge_eb", dtype=dt, n=per + 17, kind="field", seed=11, scale=1e20,
            eb_abs=(2.0 ** 100 if dt == "f32" else 2.0 ** 900))
        # noisier, less compressible field
        add(f"noisy_{dt}", dtype=dt, n=per * 16, kind="field", seed=21, noise=1.0)
        add(f"noisy_{dt}_eb", dtype=dt, n=per * 16, kind="field", seed=21, noise=1.0, eb_rel=1e-3)
    return cases
