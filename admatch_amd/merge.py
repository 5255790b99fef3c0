"""Merge per-rule ``.nfa`` files into one rule set.

Restates script/script.py:16-133 (the reference's Python-2 merge tool):
input files are ordered by natural sort of their names (:81-103); each
file's state ids are offset by the running state total (:68-78); every
accepting line is rewritten to carry the file's 1-based position as its
rule id (:57-60); the output starts with ``# Merge N NFA files`` and the
total state count (:129-130), followed by per-file ``# NFA file`` blocks.
Empty, comment and unparsable lines are dropped, as there.
"""
from __future__ import annotations

import re
from typing import Iterable, List, Sequence


def natural_key(text: str):
    return [int(c) if c.isdigit() else c for c in re.split(r"(\d+)", text)]


def _merge_one(fid: int, lines: Iterable[str], offset: int, out: List[str]) -> int:
    state_count = 0
    for raw in lines:
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        if state_count == 0:
            state_count = int(line)
            continue
        parts = line.split(":")
        if len(parts) != 2:
            continue
        lhs, rhs = parts[0].strip(), parts[1].strip()
        src_dst = lhs.split("->")
        if len(src_dst) == 1:
            if "initial" in rhs:
                out.append("%d : initial" % (int(lhs) + offset))
            elif "accepting" in rhs:
                out.append("%d : accepting %d" % (int(lhs) + offset, fid))
        elif len(src_dst) == 2:
            s, d = int(src_dst[0].strip()) + offset, int(src_dst[1].strip()) + offset
            out.append("%d -> %d : %s" % (s, d, rhs))
    return offset + state_count


def merge_files(paths: Sequence[str]) -> str:
    """Merged .nfa text of ``paths`` (rule id = position after natural sort)."""
    names = sorted(paths, key=natural_key)
    body: List[str] = []
    offset = 0
    for fid, name in enumerate(names, start=1):
        body.append("# NFA file %s " % name)
        with open(name, "r", newline="") as f:
            offset = _merge_one(fid, f.read().split("\n"), offset, body)
    return "# Merge %d NFA files\n%d\n" % (len(names), offset) + "\n".join(body) + "\n"
