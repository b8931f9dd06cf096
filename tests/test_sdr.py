"""Input-file handling for the reference-style harnesses (host only)."""
import numpy as np
import pytest

from sleek_amd import sdr


def test_list_and_load(tmp_path):
    (tmp_path / "x" / "y").mkdir(parents=True)
    a = np.arange(10, dtype=np.float32)
    b = np.arange(6, dtype=np.float64)
    a.tofile(tmp_path / "x" / "a.f32")
    b.tofile(tmp_path / "x" / "y" / "b.d64")
    files = sdr.list_inputs(str(tmp_path))
    assert [f.split("/")[-1] for f in files] == ["a.f32", "b.d64"]
    assert np.array_equal(sdr.load(files[0]), a) and sdr.load(files[0]).dtype == np.float32
    assert np.array_equal(sdr.load(files[1]), b) and sdr.load(files[1]).dtype == np.float64
    assert sdr.load(files[1], np.float32).size == 12           # an explicit type wins
    assert sdr.load(files[0], shape=(2, 5)).shape == (2, 5)


def test_load_rejects_bad_sizes(tmp_path):
    p = tmp_path / "odd.f32"
    p.write_bytes(b"\0" * 6)
    with pytest.raises(ValueError):
        sdr.load(str(p))
    np.zeros(8, np.float32).tofile(tmp_path / "e.f32")
    with pytest.raises(ValueError):
        sdr.load(str(tmp_path / "e.f32"), shape=(3, 3))


def test_empty_file(tmp_path):
    p = tmp_path / "empty.f32"
    p.write_bytes(b"")
    assert sdr.load(str(p)).size == 0
