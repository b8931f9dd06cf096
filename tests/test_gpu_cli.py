"""The command-line drivers (sleek_amd/cli) against the oracle.

They take the reference CLIs' arguments and print their lines
(lossless_compressors_gpu/sleek_single_compressor.cu:59-150 and the lossy and
double-precision twins, memcpy.cu:68-139); the reference's harness reads them
by position (compress_decompress_inputs_single.py:176-184, :233-235), and its
verify script diffs the stream files against the CPU codec's
(verify_cpu_gpu_single.py:91-100).  Here the files are compared with the
oracle's streams and decodes, and the lines are parsed the harness's way.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sleek_amd", "cli", "bin")


def run(*args):
    name = os.path.basename(args[0])
    print(f"cli: {name} {' '.join(os.path.basename(a) for a in args[1:])}", flush=True)
    r = subprocess.run([os.path.join(BIN, args[0])] + list(args[1:]), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, f"{name} failed ({r.returncode}): {r.stdout[-500:]} {r.stderr[-500:]}"
    return r.stdout.splitlines()


def harness_compress_fields(lines):
    """compress_decompress_inputs_single.py:179-181"""
    return float(lines[-3].split()[-1][0:-1]), float(lines[-1].split()[-2])


def field(dt, n, specials):
    rng = np.random.default_rng(7)
    t = np.arange(n, dtype=np.float64) / n
    a = (50.0 * np.sin(9.0 * t) + 20.0 * t + rng.normal(0, 1e-3, n)).astype(dt)
    if specials:
        a[::9973] = np.nan
        a[5::20011] = np.inf
        a[7::30011] = -np.inf
        a[11::40009] = np.finfo(dt).tiny / 4
    else:
        a[::9973] = np.nan   # NaNs only: the value range stays finite
    return a


def range_bound(a, rel):
    """The drivers' relative bound: (max - min over non-NaN values) * rel, in the data's precision."""
    dt = a.dtype.type
    f = a[~np.isnan(a)]
    return dt(dt(f.max()) - dt(f.min())) * dt(rel)


@pytest.mark.parametrize("prec,dt", [("single", np.float32), ("double", np.float64)])
def test_cli_lossless(tmp_path, prec, dt):
    a = field(dt, 300_001, specials=True)
    src, enc, dec = tmp_path / "in", tmp_path / "TEMPC", tmp_path / "TEMPD"
    a.tofile(src)
    lines = run(f"sleek_{prec}_compressor_lossless", str(src), str(enc))
    ratio, thp = harness_compress_fields(lines)
    want = po.compress(a)
    got = np.fromfile(enc, dtype=np.uint8)
    assert got.tobytes() == want.tobytes()
    assert abs(ratio - a.nbytes / want.size) < 2e-3 * ratio and thp > 0
    assert lines[-4] == f"encoded size: {want.size} bytes"
    lines = run(f"sleek_{prec}_decompressor_lossless", str(enc), str(dec))
    assert float(lines[-1].split()[-2]) > 0 and lines[-1].startswith("decoding throughput:")
    assert np.fromfile(dec, dtype=np.uint8).tobytes() == a.tobytes()


@pytest.mark.parametrize("prec,dt,rel", [("single", np.float32, "1e-3"), ("double", np.float64, "1e-6")])
def test_cli_lossy_relative_bound(tmp_path, prec, dt, rel):
    a = field(dt, 250_007, specials=False)
    src, enc, dec = tmp_path / "in", tmp_path / "TEMPC", tmp_path / "TEMPD"
    a.tofile(src)
    lines = run(f"sleek_{prec}_compressor_lossy", str(src), str(enc), rel)
    ratio, thp = harness_compress_fields(lines)
    eb = range_bound(a, np.float32(rel) if dt == np.float32 else np.float64(rel))
    want = po.compress(a, eb)
    assert np.fromfile(enc, dtype=np.uint8).tobytes() == want.tobytes()
    assert thp > 0 and ratio > 1
    lines = run(f"sleek_{prec}_decompressor_lossy", str(enc), str(dec), rel, str(src))
    assert lines[-1].startswith("decoding throughput:") and float(lines[-1].split()[-2]) > 0
    tag = "f32" if dt == np.float32 else "d64"
    assert f"MAXABS_{tag} verification passed" in lines
    od, rc = po.decompress(want, dt, eb)
    assert rc == 0 and np.fromfile(dec, dtype=np.uint8).tobytes() == od.tobytes()


def test_cli_standalone_absolute_bound(tmp_path):
    a = field(np.float32, 70_000, specials=True)
    src, enc, dec = tmp_path / "in", tmp_path / "c", tmp_path / "d"
    a.tofile(src)
    lines = run("standalone/sleek_single_compressor_lossy", str(src), str(enc), "0.01")
    assert lines[-1].startswith("ratio:")
    want = po.compress(a, np.float32("0.01"))
    assert np.fromfile(enc, dtype=np.uint8).tobytes() == want.tobytes()
    run("standalone/sleek_single_decompressor_lossy", str(enc), str(dec), "0.01")
    od, rc = po.decompress(want, np.float32, np.float32("0.01"))
    assert rc == 0 and np.fromfile(dec, dtype=np.uint8).tobytes() == od.tobytes()


def test_gpu_memcpy_lines(tmp_path):
    src = tmp_path / "in"
    np.arange(1 << 20, dtype=np.float32).tofile(src)
    lines = run("GPUmemcpy", str(src))
    assert lines[7].startswith("DtD throughput:") and float(lines[7].split()[-2]) > 0   # compress_decompress_inputs_single.py:235
    assert lines[3] == f"original size: {4 << 20} bytes"


def test_verify_harness_on_an_input_tree(tmp_path):
    """The reference-style verify run over a small tree of SDRBench-like raw files
    (tests/verify_inputs.py): every run bit-for-bit equal to the CPU codec."""
    import verify_inputs

    (tmp_path / "setA").mkdir()
    field(np.float32, 40_000, specials=False).tofile(tmp_path / "setA" / "a.f32")
    field(np.float32, 16_384 * 3, specials=False).tofile(tmp_path / "setA" / "b.f32")
    (tmp_path / "setB").mkdir()
    np.linspace(-1, 1, 5_001, dtype=np.float32).tofile(tmp_path / "setB" / "ramp.dat")
    v, t = verify_inputs.main([str(tmp_path)])
    assert (v, t) == (21, 21)
    (tmp_path / "d").mkdir()
    field(np.float64, 20_000, specials=False).tofile(tmp_path / "d" / "x.d64")
    v, t = verify_inputs.main([str(tmp_path / "d"), "--precision", "double"])
    assert (v, t) == (7, 7)
