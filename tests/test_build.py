"""Structure of the compiled kernels.

Every workgroup barrier of every kernel must sit in the kernel's single
outer loop (never inside a nested, lane-masked loop), and each kernel has the
barrier count its design needs.  A barrier inside a divergent inner loop can
leave lanes spinning on stale shared state; this check reads the gfx950
assembly hipcc emits for the codec."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRCS = [os.path.join(ROOT, "sleek_amd", "csrc", f) for f in ("sleek_kernels.hip", "sleek_batch.hip")]


def kernel_barriers(asm):
    fn, depth, out = None, 0, {}
    lines = asm.split("\n")
    for i, l in enumerate(lines):
        m = re.match(r"^(_Z\w+):", l)
        if m:
            fn, depth = m.group(1), 0
            continue
        if re.match(r"^(\.LBB\d+_\d+|; %bb\.\d+):", l):
            depth = 0
            j = i
            while j < len(lines):
                mm = re.search(r"Depth=(\d+)", lines[j])
                if mm:
                    depth = max(depth, int(mm.group(1)))
                j += 1
                if j >= len(lines) or not lines[j].strip().startswith(";"):
                    break
        if "s_barrier" in l and fn:
            out.setdefault(fn, []).append(depth)
    return out


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("asm")
    text = ""
    for i, src in enumerate(SRCS):
        out = d / f"k{i}.s"
        subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", src, "-o", str(out)],
                       check=True, capture_output=True)
        text += out.read_text()
    return text


def test_barriers_in_single_outer_loop(asm):
    b = kernel_barriers(asm)
    enc = [k for k in b if "encode_kernel" in k]
    dec = [k for k in b if "decode_kernel" in k and "batch" not in k]
    assert len(enc) == 8 and len(dec) == 4
    for k in enc:
        assert b[k] == [1, 1, 1], (k, b[k])      # bit lengths, payload ready, stage reuse
    for k in dec:
        assert b[k] == [1, 1], (k, b[k])          # payload cover loaded, stage reuse
    # the batched decoder inlines the chunk decoder once per (dtype, mode)
    bdec = [k for k in b if "batch_decode_kernel" in k]
    assert len(bdec) == 1 and b[bdec[0]] == [1] * 8, b.get(bdec[0] if bdec else None)
    # the fused decoder: the first chunk's offset sum before the loop, then per
    # chunk the payload cover, the stage reuse and the next offset window
    fused = [k for k in b if "decode_fused_kernel" in k]
    assert len(fused) == 4
    for k in fused:
        assert b[k] == [0, 1, 1, 1], (k, b[k])
    # the two-wave encoder: the first chunk's bit-length exchange before the loop,
    # then per chunk the exchange, the stream offset and the staging buffer's reuse;
    # with chunk tickets, first the exchange of the first chunk's ticket
    pair = [k for k in b if "encode_pair_kernel" in k and "batch" not in k]
    assert len(pair) == 8
    for k in pair:
        ticket = k.endswith("Lb1EEEvNS_10EncodeArgsE")
        assert b[k] == ([0] if ticket else []) + [0, 1, 1, 1], (k, b[k])
    fused += pair
    # the two-wave batched encoder: the same, and the last tile's payload and
    # stream offset after the loop (all waves leave the loop together)
    bpair = [k for k in b if "batch_encode_pair_kernel" in k]
    assert len(bpair) == 1   # with tile tickets: first the exchange of the first tile's ticket
    for k in bpair:
        assert b[k] == [0, 0, 1, 1, 1, 0, 0], (k, b[k])
    fused += bpair
    # the offset scan: the sum of the table entries before the workgroup's rounds
    # (uniform: a launch constant), then per round the waves' sums and their reuse
    offs = [k for k in b if "offsets_kernel" in k and "batch" not in k]
    assert len(offs) == 1 and b[offs[0]] == [0, 0, 1, 1], b.get(offs[0] if offs else None)
    fused += offs
    # single-wave workgroups (one-wave encoders, one-wave batched encoder) need no barrier instruction
    assert not [k for k in b if "wave" in k or ("batch_encode" in k and "pair" not in k)]
    assert all(d == 1 for k, v in b.items() if k not in fused for d in v), b


CLI_NAMES = [f"sleek_{p}_{d}_{m}" for p in ("single", "double") for d in ("compressor", "decompressor")
             for m in ("lossless", "lossy")]


def test_cli_drivers_build_and_print_usage():
    """The command-line drivers compile against the drop-in headers with the host
    compiler, and answer a wrong argument count with the reference's usage line and
    exit status (-1) before touching a device."""
    if shutil.which("g++") is None or not os.path.exists(os.path.join(ROOT, "sleek_amd", "libsleek_amd.so")):
        pytest.skip("needs g++ and the built library")
    subprocess.run(["make", "-j8", "-C", os.path.join(ROOT, "sleek_amd", "cli")], check=True, capture_output=True)
    for sub in ("", "standalone"):
        for name in CLI_NAMES:
            exe = os.path.join(ROOT, "sleek_amd", "cli", "bin", sub, name)
            r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
            assert r.returncode == 255, (name, r.returncode, r.stderr)
            want = "compressed_file_name decompressed_file_name" if "decompressor" in name else \
                "input_file_name compressed_file_name"
            assert f"USAGE: {exe} {want}" in r.stdout
    r = subprocess.run([os.path.join(ROOT, "sleek_amd", "cli", "bin", "GPUmemcpy")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 255 and "USAGE:" in r.stdout
