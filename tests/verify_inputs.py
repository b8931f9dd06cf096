"""Bit-for-bit check of the GPU command-line drivers against the CPU codec
(the oracle) on every file of an input tree, as the reference's
verify_cpu_gpu_single.py / _double.py do with its GPU and CPU programs: for
each file, lossless and six range-relative bounds; a run is verified when the
compressed and the decompressed files both equal the CPU codec's.  Lossy runs
also get the point-wise bound check (MAXABS).

    python tests/verify_inputs.py <input_directory> [--precision single|double] [-v]
"""
import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import pyoracle as po  # noqa: E402
from sleek_amd import sdr  # noqa: E402

BIN = os.path.join(ROOT, "sleek_amd", "cli", "bin")
ERRBND_LIST = ["0.1", "0.01", "0.001", "0.0001", "0.00001", "0.000001"]   # verify_cpu_gpu_single.py


def relative_bound(a, rel):
    """The drivers' bound: (max - min over non-NaN values) * rel, in the data's precision."""
    dt = a.dtype.type
    f = a[~np.isnan(a)]
    if f.size == 0:
        return None
    return dt(dt(f.max()) - dt(f.min())) * dt(rel)


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    return r.returncode == 0, r.stdout


def verify_file(path, precision, tmp, verbose=False):
    """Number of verified runs (of 7) for one input file."""
    dt = np.float32 if precision == "single" else np.float64
    a = np.ascontiguousarray(sdr.load(path, dt))
    c, d = os.path.join(tmp, "TEMPC"), os.path.join(tmp, "TEMPD")
    ok_runs = 0
    ok, _ = _run([os.path.join(BIN, f"sleek_{precision}_compressor_lossless"), path, c])
    ok = ok and _run([os.path.join(BIN, f"sleek_{precision}_decompressor_lossless"), c, d])[0]
    if ok and np.fromfile(c, np.uint8).tobytes() == po.compress(a).tobytes() and \
            np.fromfile(d, np.uint8).tobytes() == a.tobytes():
        ok_runs += 1
    elif verbose:
        print(f"{path}: lossless runs differ")
    for rel in ERRBND_LIST:
        eb = relative_bound(a, np.float32(rel) if dt == np.float32 else np.float64(rel))
        ok, _ = _run([os.path.join(BIN, f"sleek_{precision}_compressor_lossy"), path, c, rel])
        ok2, out = _run([os.path.join(BIN, f"sleek_{precision}_decompressor_lossy"), c, d, rel, path]) if ok else \
            (False, "")
        same = False
        if ok and ok2 and eb is not None:
            try:
                want = po.compress(a, eb)
            except ValueError:   # a bound too large for the format: the drivers refuse it too
                want = None
            if want is not None:
                wd, rc = po.decompress(want, dt, eb)
                same = rc == 0 and np.fromfile(c, np.uint8).tobytes() == want.tobytes() and \
                    np.fromfile(d, np.uint8).tobytes() == wd.tobytes() and \
                    f"MAXABS_{'f32' if dt == np.float32 else 'd64'} verification passed" in out
        if same:
            ok_runs += 1
        elif verbose:
            print(f"{path}: errbnd {rel} runs differ")
    return ok_runs


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("input_directory")
    p.add_argument("--precision", choices=["single", "double"], default="single")
    p.add_argument("-v", "--verbose", action="store_true")
    args = p.parse_args(argv)
    files = sdr.list_inputs(args.input_directory)
    verified = 0
    with tempfile.TemporaryDirectory() as tmp:
        for f in files:
            n = verify_file(f, args.precision, tmp, args.verbose)
            if args.verbose:
                print(f"{f}: {n} / {1 + len(ERRBND_LIST)}")
            verified += n
    print(f"Bit-for-bit verified runs over total runs: {verified} / {len(files) * (1 + len(ERRBND_LIST))}")
    return verified, len(files) * (1 + len(ERRBND_LIST))


if __name__ == "__main__":
    v, t = main()
    sys.exit(0 if v == t else 1)
