"""Basis-file parsers (CPU): the Gaussian94 reader against the NWChem reader on the
same basis (6-31G, tests/data_631g.{gbs,nwchem} of the reference), plus spot values
of tests/test_parsers.py:459-577 of the reference."""
import os

import numpy as np

from gbasis_amd.parsers import parse_gbs, parse_nwchem

HERE = os.path.dirname(os.path.abspath(__file__))
GBS = os.path.join(HERE, "golden", "data_631g.gbs")
NW = os.path.join(HERE, "golden", "data_631g.nwchem")


def test_gbs_matches_nwchem_631g():
    g, n = parse_gbs(GBS), parse_nwchem(NW)
    assert sorted(g) == sorted(n) and len(g) == 36
    for atom in n:
        assert len(g[atom]) == len(n[atom]), atom
        for (lg, eg, cg), (ln, en, cn) in zip(g[atom], n[atom]):
            assert lg == ln and np.allclose(eg, en, rtol=1e-12, atol=0)
            assert cg.shape == np.asarray(cn).reshape(len(en), -1).shape
            assert np.allclose(cg, np.asarray(cn).reshape(cg.shape), rtol=1e-12, atol=0)


def test_gbs_631g_values():
    g = parse_gbs(GBS)
    assert [x[0] for x in g["H"]] == [0, 0]
    assert np.allclose(g["H"][0][1], [18.73113696, 2.825394365, 0.6401216923])
    assert np.allclose(g["H"][0][2], np.array([0.03349460434, 0.2347269535, 0.8137573261])[:, None])
    # Li: s, then the SP shells give s / p pairs
    assert [x[0] for x in g["Li"]] == [0, 0, 1, 0, 1]
    assert np.allclose(g["Li"][0][1], [642.4189150, 96.79851530, 22.09112120, 6.201070250,
                                       1.935117680, 0.636735789])
    assert all(c.ndim == 2 and c.shape[1] == 1 for _, _, c in g["Li"])
