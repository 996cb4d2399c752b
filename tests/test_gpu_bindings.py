"""Host bindings on the GPU: the Fortran iso_c_binding caller and the
pygimic-compatible Gimic class, against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err
from oracle import molio, oracle

pytestmark = pytest.mark.gpu

DEMO = os.path.join(ROOT, "gimic_amd", "fortran", "fgimic_hip_demo")


@pytest.mark.skipif(not os.path.exists(DEMO), reason="Fortran demo not built (no flang)")
def test_fortran_caller_batched_and_per_point(c4h4_files, tmp_path):
    mol, xd = c4h4_files
    pts = np.random.default_rng(8).uniform(-5, 5, (40, 3))
    pf = tmp_path / "pts.txt"
    with open(pf, "w") as f:
        f.write(f"{len(pts)}\n")
        np.savetxt(f, pts, fmt="%.17e")
    out = tmp_path / "out.txt"
    print("case fortran demo", flush=True)
    subprocess.run([DEMO, mol, xd, str(pf), str(out)], check=True, timeout=120)
    res = np.loadtxt(out)
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    Jo = oracle.jtensor(b.arrays(), dens, pts, screening=1e-6)
    assert rel_err(res[:, :9], Jo).max() <= 1e-10
    assert rel_err(res[:, 9:], Jo).max() <= 1e-10


def test_pygimic_class_interface(c4h4_files):
    from gimic_amd.gimic import Gimic
    mol, xd = c4h4_files
    g = Gimic(mol, xd)
    g.set_property("magnet", [0.0, 0.0, 1.0])
    g.set_property("spin", "total")
    with pytest.raises(TypeError):
        g.set_screening(1)
    pts = np.random.default_rng(6).uniform(-4, 4, (12, 3))
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    Jo = oracle.jtensor(b.arrays(), dens, pts, screening=1e-6)
    for k in range(len(pts)):
        assert rel_err(g.jtensor(pts[k])[None], Jo[k:k + 1], floor_frac=1.0).max() <= 1e-10
        assert np.allclose(g.jvector(pts[k]), Jo[k, 6:9], rtol=1e-10, atol=1e-16)
    assert rel_err(g.jtensors(pts), Jo).max() <= 1e-10
    assert np.allclose(g.jvectors(pts), Jo[:, 6:9], rtol=1e-10, atol=1e-16)
    g.close()
