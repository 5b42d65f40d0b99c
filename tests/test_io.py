"""TIFF stack output (the examples' dxchange.write_tiff_stack calls,
examples/foam/foam_example.py:103-104): file names, exact float32 round trip,
overwrite rules, and the per-slice loop of the examples."""
import os

import numpy as np
import pytest

from rectv_gpu_amd import io


def test_tiff_stack_round_trip_and_names(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((3, 7, 5)).astype(np.float32)
    base = str(tmp_path / "rec_tv" / "rec_0")
    names = io.write_tiff_stack(a, base)
    assert [os.path.basename(x) for x in names] == ["rec_0_00000.tiff", "rec_0_00001.tiff", "rec_0_00002.tiff"]
    np.testing.assert_array_equal(io.read_tiff_stack(base, 3), a)
    with pytest.raises(FileExistsError):
        io.write_tiff_stack(a, base)
    io.write_tiff_stack(a * 2, base, overwrite=True)
    np.testing.assert_array_equal(io.read_tiff_stack(base, 3), a * 2)
    # another axis and a start index
    names = io.write_tiff_stack(a, str(tmp_path / "x.tiff"), axis=2, start=10, digit=3)
    assert os.path.basename(names[0]) == "x_010.tiff" and len(names) == 5
    np.testing.assert_array_equal(io.read_tiff_stack(str(tmp_path / "x"), 5, digit=3, start=10),
                                  np.moveaxis(a, 2, 0))


def test_save_reconstruction_writes_every_slice_and_frame(tmp_path):
    rtv = np.arange(2 * 3 * 4 * 4, dtype=np.float32).reshape(2, 3, 4, 4)
    names = io.save_reconstruction(rtv, str(tmp_path / "rec_tv" / "rec"))
    assert len(names) == 6 and os.path.basename(names[-1]) == "rec_1_00002.tiff"
    np.testing.assert_array_equal(io.read_tiff_stack(str(tmp_path / "rec_tv" / "rec_1"), 3), rtv[1])
