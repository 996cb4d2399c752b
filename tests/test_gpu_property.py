"""GPU: magnetic-property quadrature (gimic_ctx_properties, jfield.f90:584-929)
through the C ABI.

The FP64 path against the oracle restatement fed the oracle's tensors, on
the c4h4 fixture molecule with random weights, partitions and nuclei
(<= 1e-10 relative), and against the closed forms for one s Gaussian
(Lamb shielding, Langevin magnetizability, to quadrature accuracy).
"""
import numpy as np
import pytest

from oracle import molio, oracle, property_ref

pytestmark = pytest.mark.gpu


def test_properties_match_oracle_on_c4h4(c4h4_files):
    from gimic_amd.lib import Context
    mol, xd = c4h4_files
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    rng = np.random.default_rng(21)
    coords = np.asarray(b.coords).reshape(-1, 3)
    # per-atom clouds of points, as an external numgrid grid would give
    sizes = rng.integers(40, 90, len(coords))
    pts = np.concatenate([c + rng.normal(0.0, 1.2, (n, 3)) for c, n in zip(coords, sizes)])
    w = rng.uniform(0.01, 0.2, len(pts))
    with Context.from_files(mol, xd, uhf=False, screening=1e-8) as ctx:
        print("case properties c4h4 fp64", flush=True)
        res = ctx.properties(pts, w, sizes, coords, precision="fp64", return_jt=True)
    jt = oracle.jtensor(b.arrays(), dens, pts, screening=1e-8)
    assert np.abs(res["jt"] - jt).max() <= 1e-10 * np.abs(jt).max()
    shield, magnet = property_ref.properties(jt, pts, w, sizes, coords)
    for got, ref in ((res["shield"], shield), (res["magnet"], magnet)):
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max(), np.abs(got - ref).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_properties_closed_forms_for_one_s_gaussian(tmp_path, precision):
    from gimic_amd.lib import Context
    alpha, dval, R = 0.9, 2.0, np.array([0.2, 0.1, -0.3])
    pts, w = property_ref.spherical_grid(R, rmax=8.0 / np.sqrt(alpha))
    # one-centre basis and density through the array constructor
    ctx = Context.from_arrays(R.reshape(1, 3), np.array([0], np.int32), np.array([0], np.int32),
                              np.array([1], np.int32), np.array([0], np.int32), np.array([alpha]),
                              np.array([1.0]), screening=None)
    d = np.zeros((1, 4, 1, 1))
    d[0, 0] = dval
    ctx.set_density(d)
    print("case properties closed form", precision, flush=True)
    res = ctx.properties(pts, w, [len(w)], R.reshape(1, 3), precision=precision)
    ctx.close()
    sigma, chi = property_ref.s_shell_closed_forms(alpha, dval)
    tol = 1e-9 if precision == "fp64" else 1e-6
    assert np.allclose(res["shield"][0, 0, :3], sigma, rtol=tol), (res["shield"], sigma)
    assert np.allclose(res["magnet"][0, :3], chi, rtol=tol), (res["magnet"], chi)


def test_driver_reads_numgrid_files_and_prints_reference_blocks(c4h4_files, tmp_path):
    """the fgimic-compatible driver's property mode: coord.au, gridfile.grd,
    grid_w.grd, nelpts.info in, the reference's shielding and
    magnetizability blocks out, with the oracle's numbers"""
    import io
    from gimic_amd import driver
    from gimic_amd.lib import Context
    mol, xd = c4h4_files
    b = molio.read_mol(mol)
    dens = molio.read_xdens(xd, b)
    coords = np.asarray(b.coords).reshape(-1, 3)
    rng = np.random.default_rng(8)
    sizes = rng.integers(30, 60, len(coords))
    pts = np.concatenate([c + rng.normal(0.0, 1.0, (n, 3)) for c, n in zip(coords, sizes)])
    w = rng.uniform(0.01, 0.1, len(pts))
    np.savetxt(tmp_path / "coord.au", coords)
    np.savetxt(tmp_path / "gridfile.grd", pts, fmt="%.17e")
    np.savetxt(tmp_path / "grid_w.grd", w, fmt="%.17e")
    np.savetxt(tmp_path / "nelpts.info", np.stack([np.arange(1, len(sizes) + 1), sizes], 1), fmt="%d")
    buf = io.StringIO()
    with Context.from_files(mol, xd, uhf=False, screening=1e-8) as ctx:
        print("case property driver", flush=True)
        driver.run_property(ctx, str(tmp_path), out=buf)
    text = buf.getvalue()
    jt = oracle.jtensor(b.arrays(), dens, pts, screening=1e-8)
    shield, magnet = property_ref.properties(jt, pts, w, sizes, coords)
    iso = [float(line.split()[-1]) for line in text.splitlines() if "shielding constant" in line]
    assert len(iso) == len(coords)
    ref = shield.sum(axis=1)[:, :3].sum(axis=1) / 3.0
    assert np.allclose(iso, ref, atol=2e-6)
    chi = [float(line.split()[-1]) for line in text.splitlines()
           if "isotropic magnetizability chi" in line]
    assert len(chi) == 1 and abs(chi[0] - magnet.sum(axis=0)[:3].sum() / 3.0) <= 2e-6
