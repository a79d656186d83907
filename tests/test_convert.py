"""Format converters against the reference's own data files and against the
semantics of its SNAP converter."""
import json
import os

import pytest

import cuts_amd
from cuts_amd import convert

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "ref_gsi.json")))["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_cuts_to_gsi_reproduces_reference_files(case):
    assert convert.cuts_to_gsi(case["ours"]) == case["gsi"]


def test_snap_to_cuts(tmp_path):
    # SNAP files: '#' comment lines, TAB-separated arcs (every kept line must
    # hold a TAB, as the reference converter splits each of them)
    snap = "# Directed graph\n# Nodes: 4 Edges: 4\n0\t1\n1\t3\n3\t0\n3\t0\n"
    out = convert.snap_to_cuts(snap)
    assert out == "4\n0\t1\n1\t3\n3\t0\n3\t0\n"
    p = tmp_path / "g.g"
    p.write_text(out)
    v = cuts_amd.Graph.load(str(p)).view()  # the duplicate arc collapses
    assert v["V"] == 4 and v["arcs"] == 3
