"""Autotuner of the codec's launch shape.

The stream format fixes the chunk (16 KiB) and everything the bytes depend
on; the tuner searches only what changes speed:

* encoder workgroup shape among the ticket encoders: two 64-thread waves per
  chunk (128-thread workgroups) or one wave per chunk, and its persistent
  grid (a fraction of the resident grid: a larger grid only adds look-back
  waits, and the library caps it);
* look-back geometry: chunks per look-back group (8..64) and the window of
  groups the first group-level poll watches (1..64);
* decoder: offset scan + one workgroup per chunk, or the persistent decoder
  that derives offsets from the size table, and its grid.

Coordinate descent: encoder shape x grid at the default look-back, then
look-back group x window, then the decoder shape.  Every candidate's stream
and decoded bytes are compared with the default configuration's: a
configuration that changed a byte would be a bug, and the tuner raises.
Results are cached in a JSON file keyed by device, dtype, mode and size
bucket (``SLEEK_TUNE_CACHE``, default ~/.cache/sleek_amd/tune.json) and
applied with ``apply``.
"""
import json
import math
import os

DEFAULT = {"encode_variant": 0, "encode_grid_mult": 0.0, "lb_group": 32, "lb_window": 0, "decoder": 0,
           "decode_grid_mult": 0.0}   # 0: the library's own choice (lb_window: by encoder)

# ticket encoders only: they hand chunks to running workgroups, so their
# look-backs never wait on a workgroup that has not started (two waves per
# chunk = 0/6, one wave per chunk = 7; grid multiples of the resident grid)
ENCODE_SHAPES = [(0, 0.0), (7, 0.0), (6, 0.75), (7, 0.75)]
MIN_GAIN = 0.01   # a candidate must beat the best so far by 1 % (timing noise is about that)
CONFIRM = 3       # a winner is timed again this often, alternately with the default, before it is kept
# knob groups confirmed one by one: (keys, 0 = timed by the encoder / 1 = by the decoder)
KNOB_GROUPS = ((("encode_variant", "encode_grid_mult"), 0), (("lb_group", "lb_window"), 0),
               (("decoder", "decode_grid_mult"), 1))
LOOKBACK = [(32, 8), (32, 16), (32, 32), (16, 8), (64, 8), (64, 16)]
DECODE_SHAPES = [(0, 1.0), (0, 2.0), (1, 0.0)]   # (decoder, grid multiple; 0.0 = automatic)


def cache_path():
    return os.environ.get("SLEEK_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "sleek_amd", "tune.json"))


def size_bucket(nbytes):
    """Power-of-4 size bucket (16 KiB, 64 KiB, ... ) of an input size."""
    return int(4 ** max(7, math.ceil(math.log(max(nbytes, 1), 4))))


def key(device_name, elem_bytes, lossy, nbytes):
    """Cache key; lossy is a bool, or the number of lossy modes next to a lossless one
    when a mix of modes was tuned together."""
    if isinstance(lossy, bool) or lossy in (0, 1):
        mode = "lossy" if lossy else "lossless"
    else:
        mode = f"mixed{int(lossy)}"
    return f"{device_name}|f{8 * elem_bytes}|{mode}|{size_bucket(nbytes)}"


def load_cache(path=None):
    p = path or cache_path()
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=None):
    p = path or cache_path()
    os.makedirs(os.path.dirname(p) or ".", exist_ok=True)
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def search(measure, encode_shapes=ENCODE_SHAPES, lookback=LOOKBACK, decode_shapes=DECODE_SHAPES):
    """Coordinate descent over the knobs.  measure(config) -> (encode_s,
    decode_s, digest); digest identifies the bytes produced.  Returns (best
    config, trace of (config, encode_s, decode_s))."""
    trace = []
    base_enc, base_dec, base_digest = measure(dict(DEFAULT))
    trace.append((dict(DEFAULT), base_enc, base_dec))

    def run(cfg):
        e, d, dig = measure(cfg)
        if dig != base_digest:
            raise RuntimeError(f"tuning configuration {cfg} changed the output bytes")
        trace.append((dict(cfg), e, d))
        return e, d

    best = dict(DEFAULT)
    best_enc, best_dec = base_enc, base_dec
    for variant, mult in encode_shapes:
        cfg = dict(best, encode_variant=variant, encode_grid_mult=mult)
        if cfg == best:
            continue
        e, _ = run(cfg)
        if e < best_enc * (1 - MIN_GAIN):
            best, best_enc = cfg, e
    for group, window in lookback:
        cfg = dict(best, lb_group=group, lb_window=window)
        if cfg == best:
            continue
        e, _ = run(cfg)
        if e < best_enc * (1 - MIN_GAIN):
            best, best_enc = cfg, e
    for dec, mult in decode_shapes:
        cfg = dict(best, decoder=dec, decode_grid_mult=mult)
        _, d = run(cfg)
        if d < best_dec * (1 - MIN_GAIN):
            best, best_dec = cfg, d
    # A single timing can win by noise: each knob group's winner must beat
    # the group's default again, in CONFIRM rounds timed alternately with it
    # (medians compared, the other groups as chosen), or the default stays.
    for keys, side in KNOB_GROUPS:
        if all(best[k] == DEFAULT[k] for k in keys):
            continue
        base = dict(best, **{k: DEFAULT[k] for k in keys})
        tb, tc = [], []
        for _ in range(CONFIRM):
            tb.append(run(base)[side])
            tc.append(run(best)[side])
        if sorted(tc)[CONFIRM // 2] >= sorted(tb)[CONFIRM // 2] * (1 - MIN_GAIN):
            best = base
    return best, trace


def apply(cfg):
    """Sets the library's knobs to a configuration.  Grids are multiples of
    the variant's resident encoder grid (decoder: twice that)."""
    from . import _lib

    lib = _lib.load()
    variant = int(cfg["encode_variant"])
    _lib.check(lib.sleek_gpu_set_tuning(variant, 0, 0), "tuning")
    res = max(1, int(lib.sleek_gpu_thread_blocks()))
    eg = int(round(cfg["encode_grid_mult"] * res)) if cfg["encode_grid_mult"] > 0 else 0
    dg = int(round(cfg["decode_grid_mult"] * res * 2)) if cfg["decode_grid_mult"] > 0 else 0
    _lib.check(lib.sleek_gpu_set_tuning(variant, eg, dg), "tuning")
    _lib.check(lib.sleek_gpu_set_lookback(int(cfg["lb_group"]), int(cfg["lb_window"])), "lookback")
    _lib.check(lib.sleek_gpu_set_decoder(int(cfg.get("decoder", 0))), "decoder")


def tune(x, eb=None, reps=5, cache=True, path=None):
    """Tunes the single-array codec on the CUDA tensor x (a representative
    input) and applies the result; returns (config, trace).  eb: one bound
    (None = lossless) or a list of them: a configuration is then timed over
    all those modes together, so one choice serves the mix."""
    import torch

    from . import codec

    ebs = list(eb) if isinstance(eb, (list, tuple)) else [eb]
    dev = x.device
    nbytes = x.numel() * x.element_size()
    nlossy = sum(e is not None for e in ebs)
    k = key(torch.cuda.get_device_name(dev), x.element_size(),
            (nlossy > 0) if len(ebs) == 1 else nlossy + 1, nbytes)
    if cache:
        c = load_cache(path)
        if k in c:
            apply(c[k])
            return c[k], []
    out = torch.empty(codec.upper_bound_size(nbytes), dtype=torch.uint8, device=dev)
    enc = torch.zeros(1, dtype=torch.int64, device=dev)
    dec = torch.empty_like(x)
    dsz = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = codec.Workspace(nbytes, dev)
    stream = torch.cuda.current_stream()

    ref = {}

    def measure(cfg):
        apply(cfg)
        te_tot, td_tot, same = 0.0, 0.0, True
        for i, b in enumerate(ebs):
            te, td = [], []
            for _ in range(reps + 1):
                a0, a1, a2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
                a0.record(stream)
                codec.compress(x, b, out=out, encsize=enc, workspace=ws)
                a1.record(stream)
                codec.decompress(out, x.dtype, b, out=dec, decsize=dsz, workspace=ws)
                a2.record(stream)
                torch.cuda.synchronize()
                te.append(a0.elapsed_time(a1) / 1e3)
                td.append(a1.elapsed_time(a2) / 1e3)
            te_tot += float(sorted(te[1:])[reps // 2])
            td_tot += float(sorted(td[1:])[reps // 2])
            n = int(enc.item())
            if i not in ref:
                ref[i] = (n, out[:n].clone(), dec.clone())
            else:
                rn, rout, rdec = ref[i]
                same = same and n == rn and torch.equal(out[:n], rout) and \
                    torch.equal(dec.view(torch.uint8), rdec.view(torch.uint8))
        return te_tot, td_tot, "reference" if same else "changed"

    try:
        best, trace = search(measure)
    except Exception:
        apply(DEFAULT)
        raise
    apply(best)
    if cache:
        c = load_cache(path)
        c[k] = best
        save_cache(c, path)
    return best, trace


def main():
    import argparse

    import numpy as np
    import torch

    from .datagen import array_field

    p = argparse.ArgumentParser(description="tune the SLEEK codec's launch shape for an input size")
    p.add_argument("--bytes", type=int, default=512 << 20)
    p.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    p.add_argument("--eb", type=float, default=None, help="absolute error bound (default lossless)")
    p.add_argument("--no-cache", action="store_true")
    a = p.parse_args()
    dt = np.float32 if a.dtype == "f32" else np.float64
    x = array_field(a.bytes, dt, seed=1, device=torch.device("cuda"))
    best, trace = tune(x, a.eb, cache=not a.no_cache)
    for cfg, e, d in trace:
        print(json.dumps({"config": cfg, "encode_ms": round(e * 1e3, 4), "decode_ms": round(d * 1e3, 4)}))
    print(json.dumps({"best": best}))


if __name__ == "__main__":
    main()
