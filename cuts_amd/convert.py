"""Edge-list format converters.

snap_to_cuts: a SNAP edge list to this format, with the semantics of
data_set/data/cuts_graph_converter.py:9-25 -- lines containing '#' are
dropped, every other line is kept verbatim, and the header is
V = (largest id on a TAB-separated line) + 1.

cuts_to_gsi: this format to GSI's, as data_set/query/convert_ours_to_gsi.py:3-19
writes it -- "t # 0", "V E 1 1", one "v i 1" per vertex, one "e s t 1" per
line after the header (taken from the first two TAB fields), "t # -1".
"""
from __future__ import annotations


def snap_to_cuts(text: str) -> str:
    kept = [ln for ln in text.splitlines(keepends=True) if "#" not in ln]
    max_id = 0
    for ln in kept:
        parts = ln.split("\t")
        s, e = int(parts[0]), int(parts[1])
        max_id = max(max_id, s, e)
    return f"{max_id + 1}\n" + "".join(kept)


def cuts_to_gsi(text: str) -> str:
    lines = text.splitlines(keepends=True)
    V = int(lines[0])
    out = ["t # 0\n", f"{V} {len(lines) - 1} 1 1\n"]
    out += [f"v {i} 1\n" for i in range(V)]
    for ln in lines[1:]:
        parts = ln.split("\t")
        out.append(f"e {int(parts[0])} {int(parts[1])} 1\n")
    out.append("t # -1\n")
    return "".join(out)


def convert_file(fn, src: str, dst: str):
    with open(src, encoding="utf-8") as f:
        text = f.read()
    with open(dst, "w", encoding="utf-8") as f:
        f.write(fn(text))
