"""Reference file formats for the oracle side (test infrastructure only).

Independent restatement of the reference readers, used to feed the C oracle
and to write synthetic inputs for the flang-built reference:
  * MOL/INTGRL reader  -- intgrl.f90:20-264 (general contractions split into
    segmented ones, TURBOMOLE flag from line 2)
  * XDENS reader       -- dens.f90:56-135 (first value of each non-blank
    record, UHF perturbed densities halved, TURBOMOLE l-sorted reorder
    reorder.f90:54-96 + dens.f90:210-234)
  * MOL and XDENS writers (this project's harness; non-TURBOMOLE header).
"""
import numpy as np


class Basis:
    """Segmented contractions in reference function order."""

    def __init__(self):
        self.symbols = []
        self.charges = []
        self.coords = []          # bohr
        self.shell_atom = []
        self.shell_l = []
        self.shell_nprim = []
        self.shell_poff = []
        self.exps = []
        self.coefs = []
        self.turbomole = False

    @property
    def natoms(self):
        return len(self.coords)

    @property
    def nshell(self):
        return len(self.shell_l)

    @property
    def nbf(self):
        return sum((l + 1) * (l + 2) // 2 for l in self.shell_l)

    def arrays(self):
        return dict(
            coords=np.ascontiguousarray(np.asarray(self.coords, dtype=np.float64).reshape(-1, 3)),
            shell_atom=np.asarray(self.shell_atom, dtype=np.int32),
            shell_l=np.asarray(self.shell_l, dtype=np.int32),
            shell_nprim=np.asarray(self.shell_nprim, dtype=np.int32),
            shell_poff=np.asarray(self.shell_poff, dtype=np.int32),
            exps=np.asarray(self.exps, dtype=np.float64),
            coefs=np.asarray(self.coefs, dtype=np.float64),
            turbomole=int(self.turbomole))

    def add_shell(self, atom, l, exps, coefs):
        self.shell_atom.append(atom)
        self.shell_l.append(l)
        self.shell_nprim.append(len(exps))
        self.shell_poff.append(len(self.exps))
        self.exps.extend(float(e) for e in exps)
        self.coefs.extend(float(c) for c in coefs)

    def bf_l(self):
        """angular momentum of every function, reference order"""
        out = []
        for l in self.shell_l:
            out += [l] * ((l + 1) * (l + 2) // 2)
        return np.array(out, dtype=np.int32)


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


class _Records:
    """Fortran list-directed reads: each read starts on a fresh record and
    consumes records until it has its values."""

    def __init__(self, lines):
        self.lines = lines
        self.i = 0

    def line(self):
        s = self.lines[self.i]
        self.i += 1
        return s

    def values(self, n):
        vals = []
        while len(vals) < n:
            toks = self.line().replace(",", " ").split()
            vals.extend(toks)
        return vals[:n]


def read_mol(path):
    with open(path) as f:
        lines = f.read().splitlines()
    rec = _Records(lines)
    if rec.line()[:6] != "INTGRL":
        raise ValueError(f"{path}: not an INTGRL file")
    b = Basis()
    b.turbomole = rec.line()[:9] == "TURBOMOLE"
    rec.line()
    natoms = int(rec.values(1)[0])
    rec.line()
    for a in range(natoms):
        hdr = rec.line().replace(",", " ").split()
        charge, nsh = _num(hdr[0]), int(hdr[2])
        while len(hdr) < 3 + nsh:          # list-directed: may continue
            hdr += rec.line().replace(",", " ").split()
        nctrps = [int(x) for x in hdr[3:3 + nsh]]
        sline = rec.line()
        b.symbols.append(sline[0:2].ljust(2))   # the two columns as read (intgrl.f90:112)
        b.charges.append(charge)
        b.coords.append([_num(x) for x in sline[4:].split()[:3]])
        for ish in range(nsh):
            for _ in range(nctrps[ish]):
                npf, ncf = (int(x) for x in rec.values(2))
                rows = [[_num(x) for x in rec.values(1 + ncf)] for _ in range(npf)]
                xp = [r[0] for r in rows]
                for k in range(ncf):
                    b.add_shell(a, ish, xp, [r[1 + k] for r in rows])
    return b


def turbomole_perm(basis):
    """sv of reorder.f90:54-96: TM-file position q -> atom-major index"""
    lvec = basis.bf_l()
    sv = []
    for l in range(int(lvec.max()) + 1 if len(lvec) else 0):
        sv += [i for i in range(len(lvec)) if lvec[i] == l]
    return np.array(sv, dtype=np.int64)


def read_xdens(path, basis, uhf=False):
    """-> array (nset, 4, n, n) with [.., mu, nu] = D(mu, nu)"""
    n = basis.nbf
    nset = 2 if uhf else 1
    need = nset * 4 * n * n
    vals = np.empty(need, dtype=np.float64)
    k = 0
    with open(path) as f:
        for line in f:
            s = line.split()
            if not s:
                continue
            vals[k] = _num(s[0].rstrip(","))
            k += 1
            if k == need:
                break
    if k < need:
        raise ValueError(f"{path}: {k} values, need {need}")
    # column-major blocks: element (mu, nu) at mu + n*nu
    d = vals.reshape(nset, 4, n, n).transpose(0, 1, 3, 2).copy()
    if uhf:
        d[:, 1:4] *= 0.5
    if basis.turbomole:
        sv = turbomole_perm(basis)
        out = np.empty_like(d)
        out[:, :, sv[:, None], sv[None, :]] = d
        d = out
    return d


def pack_dens(d):
    """(nset,4,n,n) [mu,nu] -> flat column-major blocks as the oracle takes"""
    d = np.asarray(d, dtype=np.float64)
    return np.ascontiguousarray(d.transpose(0, 1, 3, 2)).reshape(-1)


def write_mol(path, basis, title="CFOUR"):
    """non-TURBOMOLE MOL/INTGRL in the layout read by intgrl.f90:20-264"""
    out = ["INTGRL        1    0    1    0    0    0    0    0    0", title,
           "              written by gimic_amd harness",
           f"{basis.natoms}    0            0.10E-08              0    0",
           "9999.00      3.00"]
    order = sorted(range(basis.nshell), key=lambda s: (basis.shell_atom[s], basis.shell_l[s]))
    if order != list(range(basis.nshell)):
        raise ValueError("write_mol: shells must be atom-major and l-sorted per atom")
    for a in range(basis.natoms):
        sh = [s for s in range(basis.nshell) if basis.shell_atom[s] == a]
        lmax = max(basis.shell_l[s] for s in sh)
        counts = [sum(1 for s in sh if basis.shell_l[s] == l) for l in range(lmax + 1)]
        out.append(f"{basis.charges[a]:.1f}    1 {lmax + 1} " + " ".join(str(c) for c in counts))
        x, y, z = basis.coords[a]
        sym = (basis.symbols[a] + "  ")[:2]
        out.append(f"{sym}{(a + 1) % 100:2d} {x:22.15f} {y:22.15f} {z:22.15f}")
        for l in range(lmax + 1):
            for s in sh:
                if basis.shell_l[s] != l:
                    continue
                o, npf = basis.shell_poff[s], basis.shell_nprim[s]
                out.append(f"{npf:6d}  1")
                for p in range(npf):
                    out.append(f"{basis.exps[o + p]:25.17e} {basis.coefs[o + p]:25.17e}")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_xdens(path, dens):
    """(nset,4,n,n) [mu,nu] -> one value per line, column-major blocks"""
    flat = pack_dens(dens)
    np.savetxt(path, flat, fmt="%.17e")
