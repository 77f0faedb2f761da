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
