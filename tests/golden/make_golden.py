"""Regenerate tests/golden/golden.json from the reference CPU codec.

Run in a tree where the reference is present and oracle/_ref is built
(``make -C oracle``):  python tests/golden/make_golden.py

For every case of cases.case_list() it records the SHA-256 of the input,
of the stream the reference compressor wrote and of the array the reference
decompressor returned, and inlines streams of up to 2 KiB as hex.  The
oracle restatement is checked against the reference on the way.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from cases import build_input, case_list, eb_value  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def sha(b):
    return hashlib.sha256(np.ascontiguousarray(b).view(np.uint8).tobytes()).hexdigest()


def main():
    if po.ref() is None:
        sys.exit("oracle/_ref/libsleek_ref.so is not built (needs the reference tree)")
    out = []
    for spec in case_list():
        a = build_input(spec)
        eb = eb_value(spec, a)
        ebf = None if eb is None else float(eb)
        s = po.ref_compress(a, ebf)
        d = po.ref_decompress(s, a.dtype, ebf)
        so = po.compress(a, ebf)
        do, rc = po.decompress(so, a.dtype, ebf)
        ok = (so.tobytes() == s.tobytes()) and rc == 0 and do.tobytes() == d.tobytes()
        if not ok:
            print("ORACLE MISMATCH", spec["name"])
        ent = dict(spec)
        ent.update(
            eb_bits=None if eb is None else int(np.array(eb).view(np.uint32 if a.dtype == np.float32 else np.uint64)),
            input_sha256=sha(a), stream_len=int(s.size), stream_sha256=sha(s),
            decoded_sha256=sha(d), stream_hex=s.tobytes().hex() if s.size <= 2048 else None,
        )
        out.append(ent)
        print(f"{spec['name']:40s} n={a.size:8d} stream={s.size:8d} {'ok' if ok else 'MISMATCH'}")
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump({"generator": "reference CPU codec (include/sleek_*_cpu.h) via oracle/ref_shim.cpp",
                   "cases": out}, f, indent=1)


if __name__ == "__main__":
    main()
