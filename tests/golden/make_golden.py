"""Write the golden fixtures under tests/golden/ (run in a checkout that has
the reference tree; the GPU box only reads the files written here).

  c4h4.npz        test/c4h4/{MOL,XDENS} (basis text + density values as
                  stored), read-grid points and reference jvec.vtu vectors,
                  integration/reference/stdout flux numbers
  openshell.npz   test/open-shell/{MOL,XDENS}, integration and 3d inputs and
                  the integration reference stdout
  writers.npz     the reference's own output files that pin the field
                  writers, as the test data holds them (xz-compressed
                  bytes): open-shell/3d/reference/{jvec,jmod}[tag].vti,
                  c4h4/read-grid/reference/jvec.vtu, and the TetGen cells
                  c4h4/read-grid/grid.1.ele that run reads
  benzene.npz     what the benzene tests and example hold that needs no
                  density (their XDENS is not in the tree): test/benzene/MOL,
                  each test's input, reference stdout and VTI headers (up to
                  the DataArray line), int-cdens/reference/jmod.txt, every
                  29th point of the 3d and 3d-keyword-magnet jvec.vti /
                  jmod.vti values and of the 2d pair's jvec.vti values (as
                  the files print them), and examples/benzene/{int,3D}
                  inputs with the mol.xyz and grid.xyz they produced
  ring12_ref.npz  flang-built reference (oracle/_ref) tensors, all nine
                  components, for the synthetic 12-atom ring at 256 seeded
                  points, the cartesian-equivalent density it was fed, and
                  composite Gauss/Lobatto rules from the reference's
                  setup_gauss_data

Usage: python tests/golden/make_golden.py [/root/reference] [--writers-only|--benzene-only]
"""
import lzma
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?")


def numbers(text):
    return [float(x.replace("D", "E").replace("d", "e")) for x in NUM.findall(text)]


def xdens_values(path):
    vals = []
    with open(path) as f:
        for line in f:
            s = line.split()
            if s:
                vals.append(float(s[0].replace("D", "E")))
    return np.array(vals)


def vtu_vectors(path, npts):
    lines = open(path).read().splitlines()
    i = next(k for k, l in enumerate(lines) if 'Name="vectors"' in l)
    return np.array([[float(x) for x in lines[i + 1 + k].split()] for k in range(npts)])


def vti_data(path):
    txt = open(path).read()
    head = txt[:txt.index("<PointData")]
    extent = [int(x) for x in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    origin = [float(x) for x in re.search(r'Origin="([^"]+)"', head).group(1).split()]
    spacing = [float(x) for x in re.search(r'Spacing="([^"]+)"', head).group(1).split()]
    body = txt[txt.index("<DataArray"):]
    body = body[body.index(">") + 1:body.index("</DataArray>")]
    return np.array(extent), np.array(origin), np.array(spacing), np.array(numbers(body))


def flux_block(stdout, marker):
    """numbers of the 10-line block starting at `marker` (runtest filter)."""
    lines = stdout.splitlines()
    i = next(k for k, l in enumerate(lines) if marker in l)
    return numbers("\n".join(lines[i:i + 10]))


def writer_files(ref):
    """writers.npz: the reference's output files named in the module
    docstring, each as xz-compressed bytes under its path in test/"""
    t = os.path.join(ref, "test")
    names = ["open-shell/3d/reference/%s%s.vti" % (k, tag) for k in ("jvec", "jmod")
             for tag in ("", "alpha", "beta", "spindens")]
    names += ["c4h4/read-grid/reference/jvec.vtu", "c4h4/read-grid/grid.1.ele"]
    out = {}
    for n in names:
        raw = open(os.path.join(t, n), "rb").read()
        out[n] = np.frombuffer(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME), dtype=np.uint8)
    np.savez(os.path.join(HERE, "writers.npz"), **out)


BENZENE_TESTS = ["2d", "2d-keyword-magnet", "3d", "3d-keyword-magnet", "diamag-off", "giao-test",
                 "int-cdens", "int-grid-bond-even", "integration-gauss", "integration-lobatto",
                 "keyword-magnet", "keyword-radius", "keyword-rotation", "keyword-rotation_origin",
                 "keyword-spacing", "paramag-off", "skip-jmod-integration", "vectors"]
SAMPLE_STRIDE = 29


def data_tokens(text, name):
    """the printed values of DataArray `name`, as text tokens"""
    i = text.index(f'Name="{name}"')
    i = text.index(">", i) + 1
    return text[i:text.index("</DataArray>", i)].split()


def benzene_files(ref):
    """benzene.npz, named in the module docstring"""
    t = os.path.join(ref, "test", "benzene")
    out = {"mol": np.array(open(os.path.join(t, "MOL")).read())}
    for name in BENZENE_TESTS:
        d = os.path.join(t, name)
        inp = "vectors.inp" if name == "vectors" else "gimic.inp"
        out[f"{name}/input"] = np.array(open(os.path.join(d, inp)).read())
        rd = os.path.join(d, "reference")
        for f in sorted(os.listdir(rd)):
            text = open(os.path.join(rd, f)).read()
            if f == "stdout":
                out[f"{name}/stdout"] = np.array(text)
            elif f.endswith(".vti"):
                k = text.index("<DataArray")
                out[f"{name}/{f}:head"] = np.array(text[:text.index("\n", k) + 1])
    out["int-cdens/jmod.txt"] = np.array(open(os.path.join(t, "int-cdens", "reference",
                                                         "jmod.txt")).read())
    for name, files in (("3d", ("jvec", "jmod")), ("3d-keyword-magnet", ("jvec", "jmod")),
                        ("2d", ("jvec",)), ("2d-keyword-magnet", ("jvec",))):
        for f in files:
            text = open(os.path.join(t, name, "reference", f + ".vti")).read()
            tok = data_tokens(text, "vectors" if f == "jvec" else "scalars")
            ncomp = 3 if f == "jvec" else 1
            pts = range(0, len(tok) // ncomp, SAMPLE_STRIDE)
            out[f"{name}/{f}.vti:sample"] = np.array(
                " ".join(" ".join(tok[p * ncomp:(p + 1) * ncomp]) for p in pts))
    out["sample_stride"] = np.array(SAMPLE_STRIDE)
    ex = os.path.join(ref, "examples", "benzene")
    for d in ("int", "3D"):
        for f in ("gimic.inp", "mol.xyz", "grid.xyz"):
            out[f"examples/{d}/{f}"] = np.array(open(os.path.join(ex, d, f)).read())
    np.savez_compressed(os.path.join(HERE, "benzene.npz"), **out)


def main(ref):
    t = os.path.join(ref, "test")
    # ---- c4h4 ----
    d = os.path.join(t, "c4h4")
    grid = np.loadtxt(os.path.join(d, "read-grid", "gridfile.grd"))
    out = dict(
        mol=np.array(open(os.path.join(d, "MOL")).read()),
        xdens=xdens_values(os.path.join(d, "XDENS")),
        readgrid_points=grid,
        readgrid_jvec=vtu_vectors(os.path.join(d, "read-grid", "reference", "jvec.vtu"), len(grid)),
        integration_inp=np.array(open(os.path.join(d, "integration", "gimic.inp")).read()),
        integration_stdout=np.array(open(os.path.join(d, "integration", "reference",
                                                      "stdout")).read()),
    )
    np.savez_compressed(os.path.join(HERE, "c4h4.npz"), **out)
    # ---- open-shell ----
    d = os.path.join(t, "open-shell")
    out = dict(
        mol=np.array(open(os.path.join(d, "MOL")).read()),
        xdens=xdens_values(os.path.join(d, "XDENS")),
        integration_inp=np.array(open(os.path.join(d, "integration", "gimic.inp")).read()),
        integration_stdout=np.array(open(os.path.join(d, "integration", "reference",
                                                      "stdout")).read()),
        threed_inp=np.array(open(os.path.join(d, "3d", "gimic.inp")).read()),
    )
    np.savez_compressed(os.path.join(HERE, "openshell.npz"), **out)
    writer_files(ref)
    benzene_files(ref)
    # ---- synthetic ring through the flang-built reference ----
    from gimic_amd import synth
    from oracle import molio, refrun, sph
    import tempfile
    s = synth.ring12()
    rng = np.random.default_rng(7)
    pts = rng.uniform(-6.0, 6.0, (256, 3))
    pts[:, 2] *= 0.5
    dcart = sph.to_cartesian(s.dens, s.shell_l)
    b = molio.Basis()
    b.symbols, b.charges, b.coords = s.symbols, s.charges, s.coords
    b.shell_atom, b.shell_l, b.shell_nprim, b.shell_poff = (s.shell_atom, s.shell_l,
                                                             s.shell_nprim, s.shell_poff)
    b.exps, b.coefs = s.exps, s.coefs
    with tempfile.TemporaryDirectory() as td:
        mol = os.path.join(td, "MOL")
        molio.write_mol(mol, b)
        res = {}
        for flags in [(1, 1, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]:
            key = "jt_%d%d%d" % flags
            res[key] = refrun.ref_jtensor(mol, pts, dens=molio.pack_dens(dcart), screening=1e-8,
                                          giao=flags[0], diamag=flags[1], paramag=flags[2])
        res["jt_noscreen"] = refrun.ref_jtensor(mol, pts, dens=molio.pack_dens(dcart),
                                                screening=-1)
        molread = molio.read_mol(mol)
        assert molread.nbf == s.nbf_cart
    rules = {}
    for kind in ("gauss", "lobatto"):
        for order in (3, 5, 7, 9, 11):
            p, w = refrun.ref_gauss_points(0.0, 7.3, order, 4 * order, kind)
            rules[f"{kind}_{order}_pts"] = p
            rules[f"{kind}_{order}_wgt"] = w
    np.savez_compressed(os.path.join(HERE, "ring12_ref.npz"), points=pts, dens_cart=dcart,
                        **res, **rules)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    ref = args[0] if args else "/root/reference"
    if "--writers-only" in sys.argv:
        writer_files(ref)
    elif "--benzene-only" in sys.argv:
        benzene_files(ref)
    else:
        main(ref)
