"""The table cache in the reference's layout and text format
(GMRES/G_FMM2DTree.hpp:1252-1313 M tables, :1806-1950 neighbour tables):
lsg_write_precomputations and lsg_inputs.precomputations_dir. CPU only."""
import os
import re

import numpy as np
import pytest

import lippmann_schwinger_solver_amd as L
import oracle as O
from tests import helpers as H

ENTRY = re.compile(r"^\(([^,()]+),([^,()]+)\)$")


@pytest.fixture(scope="module")
def cache(tmp_path_factory):
    d = tmp_path_factory.mktemp("cache")
    L.write_precomputations(H.inputs(depth=3, kappa=10.0), d)
    return d


def test_layout_matches_the_reference_reader(cache):
    md = cache / "precomputations" / "M" / "M_6_10_6"
    nd = cache / "precomputations" / "Neighbor" / "Neighbor_6_10_6"
    assert sorted(os.listdir(md)) == [f"M_6_{l}" for l in range(4)]          # readMFromFile: levels 0..nLevels
    names = set(os.listdir(nd))
    for kind, count in (("colleague", 9), ("fine", 12), ("separatedFine", 20), ("coarse", 12)):
        for level in range(2, 11):                                          # getNeighborInteractions: 2..10
            for s in range(count):
                assert f"{kind}NeighborInteraction_{level}_{s}" in names
    assert len(names) == 53 * 9


@pytest.mark.parametrize("name,rows", [("M/M_6_10_6/M_6_3", 31), ("Neighbor/Neighbor_6_10_6/colleagueNeighborInteraction_3_8", 36)])
def test_text_format(cache, name, rows):
    """Eigen's matrix printing of complex entries at stream precision 6: one row
    per line, entries right-aligned to the widest, one trailing newline
    (ReadFromTextFile counts rows as newlines and columns as words / rows)."""
    text = (cache / "precomputations" / name).read_text()
    assert text.endswith("\n") and not text.endswith("\n\n")
    lines = text[:-1].split("\n")
    assert len(lines) == rows and text.count("\n") == rows
    width = None
    for line in lines:
        toks = line.split()
        assert len(toks) == 36
        for t in toks:
            m = ENTRY.match(t)
            assert m, t
            for v in m.groups():
                assert "%.6g" % float(v) == v                               # six significant digits, %g style
        # fixed-width columns: entry k ends at (k + 1) * (width + 1) - 1
        w = (len(line) + 1) // 36 - 1
        assert (len(line) + 1) % 36 == 0
        width = w if width is None else width
        assert w == width


def test_cached_tables_reproduce_the_build(cache):
    inp = H.inputs(depth=3, kappa=10.0)
    h = L.HostBuild(inp)
    inp2 = H.inputs(depth=3, kappa=10.0, q=0)
    inp2.precomputations_dir = str(cache)
    h2 = L.HostBuild(inp2)
    o = O.Oracle(inp.to_c())
    assert np.array_equal(h2.far_table(), h.far_table())
    assert np.array_equal(h2.near_tables(), h.near_tables())
    assert np.array_equal(h2.far_table(), o.far_table())
    assert np.array_equal(h2.skeleton(), h.skeleton())


def test_reader_accepts_any_whitespace_and_rejects_bad_files(cache, tmp_path):
    """The reference reads with operator>> (any whitespace); a missing or
    ragged file fails the build instead of the reference's exit(2)."""
    src = cache / "precomputations"
    dst = tmp_path / "precomputations"
    for sub in ("M/M_6_10_6", "Neighbor/Neighbor_6_10_6"):
        os.makedirs(dst / sub)
        for n in os.listdir(src / sub):
            t = (src / sub / n).read_text()
            (dst / sub / n).write_text(re.sub(r" +", "\t", t.strip()) + "\n")   # tabs, no padding
    inp = H.inputs(depth=3, kappa=10.0)
    inp.precomputations_dir = str(tmp_path)
    ref = L.HostBuild(H.inputs(depth=3, kappa=10.0))
    assert np.array_equal(L.HostBuild(inp).near_tables(), ref.near_tables())
    os.remove(dst / "M/M_6_10_6/M_6_3")
    with pytest.raises(L.LsgError):
        L.HostBuild(inp)
    (dst / "M/M_6_10_6/M_6_3").write_text("(1,2) (3,4)\n(5,6)\n")
    with pytest.raises(L.LsgError):
        L.HostBuild(inp)
