import lzma
import os
import re
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def have_reference_build():
    from oracle import refrun
    return refrun.available() and os.path.isdir(REFERENCE)


def rel_err(a, b, floor_frac=1e-3):
    """per-point ||a-b|| / max(||b||, floor_frac * max ||b||) over (N,9) tensors"""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    nb = np.linalg.norm(b, axis=1)
    den = np.maximum(nb, floor_frac * nb.max())
    return np.linalg.norm(a - b, axis=1) / den


def write_fixture(tmpdir, npz):
    """MOL and XDENS text files from a golden npz (values as the reference stored them)"""
    d = golden(npz)
    mol = os.path.join(tmpdir, "MOL")
    xd = os.path.join(tmpdir, "XDENS")
    with open(mol, "w") as f:
        f.write(str(d["mol"]))
    np.savetxt(xd, d["xdens"], fmt="%.17e")
    return mol, xd


@pytest.fixture(scope="session")
def c4h4_files(tmp_path_factory):
    return write_fixture(str(tmp_path_factory.mktemp("c4h4")), "c4h4.npz")


@pytest.fixture(scope="session")
def openshell_files(tmp_path_factory):
    return write_fixture(str(tmp_path_factory.mktemp("openshell")), "openshell.npz")


def writer_fixture(name):
    """text of one of the reference's output files held in writers.npz
    (name: its path under the reference's test/ directory)"""
    d = np.load(os.path.join(GOLDEN, "writers.npz"))
    return lzma.decompress(d[name].tobytes()).decode()


_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?")


def runtest_mismatches(out_text, ref_text, rel_tolerance):
    """Whole-file comparison with runtest's numeric-filter semantics: every
    number in both files in order, integers exactly, reals to a relative
    tolerance (a zero reference value must be matched exactly).  Returns the
    (written, reference) token pairs that fail; raises when the files hold
    different counts of numbers."""
    a, r = _NUM.findall(out_text), _NUM.findall(ref_text)
    if len(a) != len(r):
        raise AssertionError(f"{len(a)} numbers written, {len(r)} in the reference file")
    bad = []
    for x, y in zip(a, r):
        if x == y:
            continue
        if not re.search(r"[.eEdD]", y):
            bad.append((x, y))
            continue
        fx, fy = float(x.replace("D", "E")), float(y.replace("D", "E"))
        if fy == 0.0 or abs(fx - fy) > rel_tolerance * abs(fy):
            bad.append((x, y))
    return bad


def stdout_setup_lines(text):
    """the lines the reference prints between reading the density and naming
    the shell type (gimic.F90:159-171): the grid's set-up, the grid plot
    note and the field check"""
    lines = text.splitlines()
    end = next(i for i, l in enumerate(lines) if "-shell calculation" in l)
    # the first line new_grid prints: the empty record before a bond grid's
    # data block, the mode, or a file grid's point count
    start = min(i - (l == " Integration grid data") for i, l in enumerate(lines[:end])
                if l == " Integration grid data" or l.startswith(" Grid mode = ")
                or l.startswith("   Total number of grid points"))
    return lines[start:end]


def stdout_integral_lines(text):
    """the printout of run_integral (gimic.F90:222-261): from 'Integrating
    current density' up to the deallocation notes (or the end)"""
    lines = text.splitlines()
    start = next(i for i, l in enumerate(lines) if l == " Integrating current density")
    end = next((i for i, l in enumerate(lines) if "Deallocated grid data" in l), len(lines))
    return lines[start:end]


def printout_mismatches(ours, ref, atol):
    """line-by-line comparison of two printouts: the text around the numbers
    exactly, every number to `atol`.  Returns the differing line pairs."""
    bad = []
    if len(ours) != len(ref):
        return [(f"{len(ours)} lines", f"{len(ref)} lines")]
    for a, r in zip(ours, ref):
        if a == r:
            continue
        na, nr = _NUM.findall(a), _NUM.findall(r)
        if _NUM.sub("#", a) != _NUM.sub("#", r) or len(na) != len(nr) or any(
                abs(float(x) - float(y)) > atol for x, y in zip(na, nr)):
            bad.append((a, r))
    return bad
