"""The native output writers (csrc/io.hip) on the CPU: byte-identical to the
format module's reference writers, and the part stitching of
stitch_temperatures (175706/Run-04/thermal.cpp:104-244)."""
import filecmp
import os

import numpy as np
import pytest

from binarytpm_amd import _lib, formats, inputs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libbtpm.so not built")


def test_native_text_writer_is_byte_identical(tmp_path):
    from binarytpm_amd import thermal as th
    V, F = inputs.ellipsoid(2, (0.6, 0.5, 0.45))
    rng = np.random.default_rng(0)
    T = (120.0 + 280.0 * rng.random((len(F), 333))).astype(np.float32)
    T[0, :6] = [1e-5, 123456789.0, 0.0, 1e30, 3.4e-39, -7.25]
    a, b = str(tmp_path / "a.txt"), str(tmp_path / "b.txt")
    formats.write_temperature_txt(a, V, F, T)
    for threads in (1, 3, 0):
        th.write_temperature_txt(b, V, F, T, threads=threads)
        assert filecmp.cmp(a, b, shallow=False)


def test_native_crater_writer_and_stitching(tmp_path):
    from binarytpm_amd import thermal as th
    c = np.load(os.path.join(GOLD, "new_crater2.npz"))
    rng = np.random.default_rng(1)
    temps = rng.random((10, 48, 100)).astype(np.float32)
    full_py, full = str(tmp_path / "py.dat"), str(tmp_path / "full.dat")
    formats.write_roughness_dat(full_py, c["V"], c["F"], temps)
    th.write_crater_dat(full, c["V"], c["F"], temps)
    assert filecmp.cmp(full_py, full, shallow=False)
    parts = []
    for k, (a, b) in enumerate(((0, 4), (4, 7), (7, 10)), start=1):
        p = str(tmp_path / f"s_roughness_temperature_part_{k}(50).dat")
        th.write_crater_dat(p, c["V"], c["F"], temps[a:b])
        parts.append(p)
    from binarytpm_amd import stitch
    assert stitch.main(["--shape", str(tmp_path / "s"), "--parts", "3", "--ntau", "48",
                        "--ti", "50"]) == 0
    assert filecmp.cmp(full, str(tmp_path / "s_roughness_temperature(50).dat"), shallow=False)
    # a part whose size is not a header plus whole facets is refused
    with open(parts[1], "ab") as fh:
        fh.write(b"\0" * 12)
    with pytest.raises(_lib.BtpmError):
        th.stitch_crater_parts(str(tmp_path / "bad.dat"), parts, 61, 100, 48)
