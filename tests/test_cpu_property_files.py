"""CPU: the driver's reading of get_property's numgrid files
(jfield.f90:612-657): what is missing skips, what disagrees is an error."""
import numpy as np
import pytest

from gimic_amd import driver


def test_missing_files_skip(tmp_path):
    assert driver.read_property_files(str(tmp_path)) is None
    np.savetxt(tmp_path / "coord.au", np.zeros((1, 3)))
    assert driver.read_property_files(str(tmp_path)) is None


def test_files_read_and_counts_checked(tmp_path):
    np.savetxt(tmp_path / "coord.au", np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]]))
    np.savetxt(tmp_path / "gridfile.grd", np.arange(15.0).reshape(5, 3))
    np.savetxt(tmp_path / "grid_w.grd", np.full(5, 0.5))
    np.savetxt(tmp_path / "nelpts.info", np.array([[1, 2], [2, 3]]), fmt="%d")
    coords, pts, w, sizes = driver.read_property_files(str(tmp_path))
    assert coords.shape == (2, 3) and pts.shape == (5, 3) and list(sizes) == [2, 3]
    np.savetxt(tmp_path / "nelpts.info", np.array([[1, 2], [2, 2]]), fmt="%d")
    with pytest.raises(ValueError):
        driver.read_property_files(str(tmp_path))
