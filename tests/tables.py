"""Reads the marching-cubes case table from torchhull_amd/csrc/mc_cases.h (test helper)."""

from __future__ import annotations

import pathlib
import re

HEADER = pathlib.Path(__file__).resolve().parents[1] / "torchhull_amd" / "csrc" / "mc_cases.h"


def case_table() -> tuple[list[list[int]], list[int]]:
    text = HEADER.read_text()
    body = text[text.index("kMcCaseText[256] = {"):]
    body = body[: body.index("};")]
    cases = re.findall(r'"([0-9a-f]*)"', body)
    assert len(cases) == 256
    edges = [[int(ch, 16) for ch in c] for c in cases]
    return edges, [len(e) for e in edges]
