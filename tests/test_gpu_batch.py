"""Batched launches (many arrays per launch, mixed float32/float64 and
lossless/lossy): every array's stream and decoded bytes equal the oracle's,
and so the single-array calls' and the reference codec's, at the automatic
grid and at small grids, eagerly and replayed from a captured HIP graph.
All GPU calls go through the C ABI of libsleek_amd.so."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mixed_arrays(seed, extra, max_n=600_000):
    """Edge sizes (empty, 1 value, chunk +-1, partial chunks) plus `extra` random sizes;
    70 % float32; a quarter each lossless and eb in {1e-2, 1e-4, 1e-7}; specials injected."""
    from sleek_amd.datagen import exact_field, inject_specials

    rng = np.random.default_rng(seed)
    sizes = [0, 1, 5, 2047, 2048, 2049, 4095, 4096, 4097, 12295] + [int(v) for v in rng.integers(1, max_n, extra)]
    arrays, ebs = [], []
    for i, n in enumerate(sizes):
        dt = np.float32 if rng.random() < 0.7 else np.float64
        a = exact_field(n, seed * 1000 + i, dt) if n > 0 else np.zeros(0, dt)
        if n >= 1000:
            inject_specials(a, 1e-3, i)
        arrays.append(a)
        ebs.append([None, 1e-2, 1e-4, 1e-7][int(rng.integers(0, 4))])
    return arrays, ebs


def run_batch(arrays, ebs, grid=0):
    """Batched compress then batched decompress; returns (streams, decoded, statuses) as numpy."""
    import torch
    from sleek_amd.batch import compress_batch, decompress_batch

    xs = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]
    outs, sizes, bc = compress_batch(xs, ebs, grid=grid)
    st_c = bc.status()
    n = sizes.cpu().numpy()[: len(arrays)]
    streams = [outs[i][: int(n[i])] for i in range(len(arrays))]
    dec, dsz, bd = decompress_batch(streams, [a.dtype for a in arrays], ebs, [a.size for a in arrays], grid=grid)
    st_d = bd.status()
    torch.cuda.synchronize()
    return ([s.cpu().numpy() for s in streams], [d.cpu().numpy() for d in dec], dsz.cpu().numpy()[: len(arrays)],
            st_c, st_d)


@pytest.mark.parametrize("grid", [0, 5, 1])
def test_batch_matches_oracle(gpu, grid):
    arrays, ebs = mixed_arrays(11, 30 if grid else 60)
    say(f"batch of {len(arrays)} arrays, grid={grid}")
    streams, dec, dsz, st_c, st_d = run_batch(arrays, ebs, grid)
    assert not st_c.any() and not st_d.any()
    for i, (a, eb) in enumerate(zip(arrays, ebs)):
        ref = po.compress(a, eb)
        assert streams[i].tobytes() == ref.tobytes(), (i, a.size, a.dtype, eb)
        od, rc = po.decompress(ref, a.dtype, eb)
        assert rc == 0 and dec[i].tobytes() == od.tobytes(), (i, a.size, a.dtype, eb)
        assert dsz[i] == a.nbytes
        if eb is None:
            assert dec[i].tobytes() == a.tobytes()


def help_events():
    from sleek_amd import _lib

    n = _lib.load().sleek_gpu_help_events()
    assert n >= 0, n
    return n


@pytest.mark.parametrize("variant", [4, 5, 6])
def test_batch_encoder_forms_match_oracle(gpu, variant):
    """The batched encoder's one-wave form (encoder variant 4, tiles in launch order)
    and its two-wave form with tile tickets (every other variant; 5 and 6 here)
    write the oracle's streams, at the automatic grid and at 1, 3 and 5 workgroups.
    With tile tickets every tile goes to a running workgroup: no look-back help."""
    from sleek_amd import _lib

    lib = _lib.load()
    arrays, ebs = mixed_arrays(17, 25)
    try:
        for grid in (0, 1, 3, 5):
            _lib.check(lib.sleek_gpu_set_tuning(variant, 0, 0), "tuning")
            say(f"batch encoder variant={variant}, {len(arrays)} arrays, grid={grid}")
            h0 = help_events()
            streams, dec, _, st_c, st_d = run_batch(arrays, ebs, grid)
            assert not st_c.any() and not st_d.any()
            for i, (a, eb) in enumerate(zip(arrays, ebs)):
                assert streams[i].tobytes() == po.compress(a, eb).tobytes(), (variant, grid, i, a.size, a.dtype, eb)
            if variant != 4:
                assert help_events() == h0, f"tile tickets needed look-back help at grid {grid}"
    finally:
        lib.sleek_gpu_set_tuning(0, 0, 0)


@pytest.mark.parametrize("variant", [4, 6])
def test_batch_lookback_help_always_on(gpu, variant):
    """The batched encoders with the help threshold at 0: look-backs size missing
    tiles and groups from the inputs themselves, and the bytes stay the oracle's."""
    from sleek_amd import _lib

    lib = _lib.load()
    arrays, ebs = mixed_arrays(19, 12)
    try:
        _lib.check(lib.sleek_gpu_set_help_after(0), "help")
        _lib.check(lib.sleek_gpu_set_tuning(variant, 0, 0), "tuning")
        for grid in (0, 3):
            say(f"batch help always on, variant={variant}, grid={grid}")
            streams, _, _, st_c, st_d = run_batch(arrays, ebs, grid)
            assert not st_c.any() and not st_d.any()
            for i, (a, eb) in enumerate(zip(arrays, ebs)):
                assert streams[i].tobytes() == po.compress(a, eb).tobytes(), (variant, grid, i, a.size, a.dtype, eb)
    finally:
        lib.sleek_gpu_set_help_after(100)
        lib.sleek_gpu_set_tuning(0, 0, 0)


def test_batch_equals_single_array_calls(gpu):
    """A batch of larger arrays against the plain single-array entry points."""
    import torch
    from sleek_amd import codec

    arrays, ebs = mixed_arrays(23, 40, max_n=3_000_000)
    say(f"batch vs plain, {len(arrays)} arrays, {sum(a.nbytes for a in arrays) >> 20} MiB")
    streams, dec, _, st_c, st_d = run_batch(arrays, ebs)
    assert not st_c.any() and not st_d.any()
    for i, (a, eb) in enumerate(zip(arrays, ebs)):
        x = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        out, n = codec.compress(x, eb)
        d, _ = codec.decompress(out, x.dtype, eb, numel=a.size)
        torch.cuda.synchronize()
        assert out[: int(n.item())].cpu().numpy().tobytes() == streams[i].tobytes(), i
        assert d.cpu().numpy().tobytes() == dec[i].tobytes(), i


def test_batch_graph_replay(gpu):
    """Captured compress and decompress graphs, replayed three times with eager
    work and synchronisations between replays, reproduce the eager results."""
    import torch
    from sleek_amd.batch import plan_compress, plan_decompress

    arrays, ebs = mixed_arrays(5, 20)
    say(f"graph replay, {len(arrays)} arrays")
    xs = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]
    bc = plan_compress(xs, ebs).launch()
    torch.cuda.synchronize()
    sizes0 = bc.sizes.clone()
    outs0 = [o.clone() for o in bc.outs]
    n = sizes0.cpu().numpy()
    streams = [outs0[i][: int(n[i])] for i in range(len(arrays))]
    bd = plan_decompress(streams, [a.dtype for a in arrays], ebs, [a.size for a in arrays]).launch()
    torch.cuda.synchronize()
    dec0 = [d.clone() for d in bd.outs]
    gc, gd = bc.capture(), bd.capture()
    try:
        for rep in range(3):
            for o in bc.outs:
                o.fill_(0xA5)
            bc.sizes.zero_()
            torch.cuda.synchronize()
            gc.launch()
            torch.cuda.synchronize()
            assert torch.equal(bc.sizes, sizes0), rep
            for i in range(len(arrays)):
                assert torch.equal(bc.outs[i][: int(n[i])], outs0[i][: int(n[i])]), (rep, i)
            for d in bd.outs:
                if d.numel():
                    d.view(torch.uint8).fill_(0x5A)
            torch.cuda.synchronize()
            gd.launch()
            torch.cuda.synchronize()
            for i in range(len(arrays)):
                assert torch.equal(bd.outs[i].view(torch.uint8), dec0[i].view(torch.uint8)), (rep, i)
            assert not bc.status().any() and not bd.status().any()
    finally:
        gc.close()
        gd.close()


def test_batch_reports_damage_per_array(gpu):
    import torch
    from sleek_amd import _lib
    from sleek_amd.batch import decompress_batch

    say("batch corrupt / capacity")
    a = np.linspace(0, 1, 4096 * 4, dtype=np.float32)
    ref = po.compress(a)
    sizes = ref[8:16].view(np.uint16)
    bad = ref.copy()
    bad[8 + 8 + int(sizes[0]) - 2] ^= 0x5A          # trailer of chunk 0
    streams = [torch.from_numpy(s).cuda() for s in (ref, bad, ref, ref)]
    numels = [a.size, a.size, a.size - 1, a.size]    # the third is one value short
    dec, dsz, b = decompress_batch(streams, [np.float32] * 4, [None] * 4, numels)
    st = b.status()
    assert list(st) == [0, _lib.ECORRUPT, _lib.ECAPACITY, 0]
    assert dec[0].cpu().numpy().tobytes() == a.tobytes()
    assert dec[3].cpu().numpy().tobytes() == a.tobytes()
