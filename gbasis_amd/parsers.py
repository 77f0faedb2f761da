"""Basis-set file reader (NWChem format) and shell construction.

Behaviour follows ``gbasis/parsers.py:9-72`` (``parse_nwchem``) and
``gbasis/parsers.py:169-247`` (``make_contractions``): an element block header
``<El> <SHELLTYPES>`` is followed by rows ``exponent coeff [coeff ...]``; a
multi-letter shell type such as ``SP`` yields one shell per letter that shares
the exponents and takes the matching coefficient column.
"""

import re

import numpy as np

from .contractions import GeneralizedContractionShell

_ANGMOM = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4, "h": 5, "i": 6, "k": 7}
_HEADER = re.compile(r"^\s*([A-Z][a-z]?)\s+([A-Za-z]+)\s*$")
_NUMBER = r"[0-9.DdEe+\-]+"
_ROW = re.compile(rf"^\s*({_NUMBER})((?:\s+{_NUMBER})+)\s*$")


def _to_float(tok):
    return float(tok.lower().replace("d", "e"))


def parse_nwchem(nwchem_basis_file):
    """Return ``{element: [(angmom, exps (K,), coeffs (K,) or (K, M)), ...]}``."""
    with open(nwchem_basis_file, "r") as fh:
        lines = fh.read().splitlines()
    out = {}
    block = None

    def flush():
        if block is None or not block["exps"]:
            return
        exps = np.array(block["exps"])
        coeffs = np.array(block["coeffs"])
        types = block["types"]
        if len(types) == 1:
            out.setdefault(block["atom"], []).append((types[0], exps, coeffs))
        else:
            for col, l in enumerate(types):
                out.setdefault(block["atom"], []).append((l, exps, coeffs[:, col]))

    for line in lines:
        head = _HEADER.match(line)
        if head and head.group(2).lower() != "end" and all(
                ch in _ANGMOM for ch in head.group(2).lower()):
            flush()
            block = {"atom": head.group(1), "types": [_ANGMOM[c] for c in head.group(2).lower()],
                     "exps": [], "coeffs": []}
            continue
        row = _ROW.match(line)
        if row and block is not None:
            block["exps"].append(_to_float(row.group(1)))
            block["coeffs"].append([_to_float(t) for t in row.group(2).split()])
            continue
        if line.strip().upper() == "END" or line.strip().upper().startswith("BASIS"):
            flush()
            block = None
    flush()
    return out


def parse_gbs(gbs_basis_file):
    """Return ``{element: [(angmom, exps (K,), coeffs (K, M)), ...]}`` of a Gaussian94
    basis file, the contract of gbasis/parsers.py:75-166: an "SP" shell gives one
    s and one p entry, and a contraction directly following one of the same
    angular momentum and exponents joins it as a further column (a generalised
    contraction)."""
    with open(gbs_basis_file, "r") as fh:
        lines = fh.read().splitlines()
    out = {}
    atom = None
    i = 0
    while i < len(lines):
        tok = lines[i].split()
        i += 1
        if not tok or tok[0].startswith("!"):
            continue
        if tok[0] == "****":
            atom = None
            continue
        if atom is None:
            if len(tok) == 2 and tok[0][0].isalpha():
                atom = tok[0]
                out.setdefault(atom, [])
            continue
        label = tok[0].lower()
        if len(tok) < 2 or not all(c in _ANGMOM for c in label):
            continue
        nprim = int(tok[1])
        rows = [[_to_float(x) for x in lines[i + r].split()] for r in range(nprim)]
        i += nprim
        exps = np.array([r[0] for r in rows])
        coeffs = np.array([r[1:] for r in rows])
        for col, ch in enumerate(label):
            l = _ANGMOM[ch]
            c = coeffs[:, col:col + 1]
            prev = out[atom][-1] if out[atom] else None
            if (prev is not None and prev[0] == l and prev[1].shape == exps.shape
                    and np.allclose(prev[1], exps)):
                out[atom][-1] = (l, exps, np.hstack([prev[2], c]))
            else:
                out[atom].append((l, exps, c))
    return out


def make_contractions(basis_dict, atoms, coords, coord_types, tol=1e-20, ovr=False):
    """Build shells for ``atoms`` at ``coords`` (``gbasis/parsers.py:169-247``)."""
    if not (isinstance(atoms, (list, tuple)) and all(isinstance(a, str) for a in atoms)):
        raise TypeError("Atoms must be provided as a list or tuple.")
    if not (isinstance(coords, np.ndarray) and coords.ndim == 2 and coords.shape[1] == 3):
        raise TypeError("Coordinates must be provided as a two-dimensional `numpy` array with "
                        "three columns.")
    if len(atoms) != coords.shape[0]:
        raise ValueError("Number of atoms must be equal to the number of rows in the coordinates.")
    n_shells = sum(len(basis_dict[a]) for a in atoms)
    if isinstance(coord_types, str):
        if coord_types not in ("c", "cartesian", "p", "spherical"):
            raise ValueError("If coord_types is a string, it must be either 'spherical'/'p' or "
                             "'cartesian'/'c'.")
        coord_types = [coord_types] * n_shells
    coord_types = list(coord_types)
    if len(coord_types) != n_shells:
        raise ValueError("If coord_types is a list, it must be the same length as the total "
                         "number of contractions.")
    basis = []
    for icenter, (atom, coord) in enumerate(zip(atoms, coords)):
        for l, exps, coeffs in basis_dict[atom]:
            basis.append(GeneralizedContractionShell(l, np.array(coord, dtype=float), coeffs,
                                                     exps, coord_types.pop(0), icenter=icenter))
    return tuple(basis)


# ---------------------------------------------------------------------------
# Gaussian formatted checkpoint (fchk) files.
#
# The reference reads fchk files through iodata (gbasis/wrappers.py:89-257,
# ``from_iodata``): segmented shells whose contraction coefficients multiply
# L2-normalised primitives and are used as given (``norm_cont`` = 1), with the
# component orders the file defines.  The reader below restates that for the
# fchk fields the ERI path needs.

_FCHK_ARRAY = re.compile(r"^(.{40})\s+([IRCL])\s+N=\s*(\d+)\s*$")
_FCHK_SCALAR = re.compile(r"^(.{40})\s+([IRCL])\s+(\S+)\s*$")

# Gaussian's Cartesian component order per l <= 3, as exponent triples
_GAUSSIAN_CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1), (1, 0, 2), (0, 1, 2),
        (0, 2, 1), (1, 1, 1)],
}


def _gaussian_cart(l):
    if l in _GAUSSIAN_CART:
        return np.array(_GAUSSIAN_CART[l])
    # l >= 4: reverse alphabetical order (zzzz, yzzz, yyzz, ..., xxxx)
    return np.array([(a, b, l - a - b) for a in range(l + 1) for b in range(l - a + 1)])


def _gaussian_sph(l):
    labels = ["c0"]
    for m in range(1, l + 1):
        labels += [f"c{m}", f"s{m}"]
    return tuple(labels)


def read_fchk(path):
    """Fields of a Gaussian fchk file: ``{name: int | float | str | np.ndarray}``."""
    out = {}
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    i = 2  # title and job lines
    while i < len(lines):
        line = lines[i]
        m = _FCHK_ARRAY.match(line)
        if m:
            name, kind, n = m.group(1).strip(), m.group(2), int(m.group(3))
            vals = []
            i += 1
            while len(vals) < n and i < len(lines) and not _FCHK_ARRAY.match(lines[i]) \
                    and not _FCHK_SCALAR.match(lines[i]):
                vals.extend(lines[i].split())
                i += 1
            if kind == "I":
                out[name] = np.array([int(v) for v in vals[:n]], dtype=np.int64)
            elif kind == "R":
                out[name] = np.array([_to_float(v) for v in vals[:n]], dtype=np.float64)
            else:
                out[name] = " ".join(vals)
            continue
        m = _FCHK_SCALAR.match(line)
        if m:
            name, kind, v = m.group(1).strip(), m.group(2), m.group(3)
            out[name] = int(v) if kind == "I" else (_to_float(v) if kind == "R" else v)
        i += 1
    return out


class FchkShell(GeneralizedContractionShell):
    """A shell read from an fchk file: Gaussian's component orders, coefficients used
    as given (``norm_cont`` = 1), as the reference's iodata shells
    (gbasis/wrappers.py:133-227)."""

    def assign_norm_cont(self):
        ncart = (self.angmom + 1) * (self.angmom + 2) // 2
        self.norm_cont = np.ones((self.coeffs.shape[1], ncart))

    @property
    def angmom_components_cart(self):
        return _gaussian_cart(self.angmom)

    @property
    def angmom_components_sph(self):
        return _gaussian_sph(self.angmom)


def fchk_basis(fields):
    """Shells of an fchk file (``read_fchk`` output) in the file's function order:
    SP shells become an s and a p shell; negative shell types are pure."""
    types = fields["Shell types"]
    nprim = fields["Number of primitives per shell"]
    atom = fields["Shell to atom map"] - 1
    exps_all = fields["Primitive exponents"]
    coef_all = fields["Contraction coefficients"]
    coef_p = fields.get("P(S=P) Contraction coefficients")
    xyz = fields["Coordinates of each shell"].reshape(-1, 3)
    basis = []
    p0 = 0
    for s, (t, k) in enumerate(zip(types.tolist(), nprim.tolist())):
        exps = exps_all[p0:p0 + k].copy()
        coeffs = coef_all[p0:p0 + k].copy()
        coord = xyz[s].astype(np.float64).copy()
        if t == -1:
            basis.append(FchkShell(0, coord, coeffs, exps, "cartesian", icenter=int(atom[s])))
            basis.append(FchkShell(1, coord, coef_p[p0:p0 + k].copy(), exps.copy(), "cartesian",
                                   icenter=int(atom[s])))
        else:
            l = abs(int(t))
            kind = "spherical" if (t < 0 and l >= 2) else "cartesian"
            basis.append(FchkShell(l, coord, coeffs, exps, kind, icenter=int(atom[s])))
        p0 += k
    return basis


def fchk_rhf_orbitals(fields):
    """(MO coefficients (nbasis, nmo), occupations (nmo,)) of a closed-shell fchk."""
    c, occ = fchk_orbitals(fields, "alpha")
    return c, 2.0 * occ


def fchk_orbitals(fields, spin="alpha"):
    """(MO coefficients (nbasis, nmo), occupations 0/1 (nmo,)) of one spin of an fchk;
    a restricted file has no beta set, and its beta orbitals are the alpha ones."""
    if spin not in ("alpha", "beta"):
        raise ValueError("spin must be 'alpha' or 'beta'")
    nb = int(fields["Number of basis functions"])
    nmo = int(fields["Number of independent functions"])
    key = "Beta MO coefficients" if spin == "beta" and "Beta MO coefficients" in fields \
        else "Alpha MO coefficients"
    c = fields[key].reshape(nmo, nb).T
    occ = np.zeros(nmo)
    occ[:int(fields[f"Number of {spin} electrons"])] = 1.0
    return c, occ
