"""Write the reference-held golden vectors this repo's tests use as fixtures.

Reads the reference's test data files as text (nothing is compiled, imported
or run) and writes small JSON fixtures next to this script.  Each table of
``tests/test_safe_math_golden.inc`` holds IEEE binary64 bit rows produced by
the contraction-off CPU builds of the reference's position-independent math
(``rdb_sm_*``):

* ``cbrt_golden.json`` -- ``G_CBRT`` (:557-606), rows [x, cbrt(x)];
* ``safe_math_golden.json`` -- ``G_EXP`` (:7-56), ``G_LOG`` (:57-106),
  ``G_POW`` (:507-556) and ``G_POWBIG`` (:657-706), rows [x, y, pow(x, y)],
  and ``G_ERFC`` (:607-656).

Usage: python tests/golden/extract_reference_vectors.py [/path/to/reference]
"""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
TABLES = {"G_EXP": 2, "G_LOG": 2, "G_POW": 3, "G_POWBIG": 3, "G_ERFC": 2}


def extract_table(text: str, name: str, width: int) -> list[list[str]]:
    m = re.search(name + r"\(" + str(width) + r", NG_\w+\) = reshape\(\[character\(len=16\) :: &(.*?)\]", text, re.S)
    if not m:
        raise SystemExit(f"table {name} not found")
    hexes = re.findall(r'"([0-9A-F]{16})"', m.group(1))
    if len(hexes) % width:
        raise SystemExit(f"table {name}: {len(hexes)} entries not a multiple of {width}")
    return [hexes[i:i + width] for i in range(0, len(hexes), width)]


def line_span(text: str, name: str) -> str:
    start = text.index(f"{name}(") 
    first = text.count("\n", 0, start) + 1
    end = text.index("])", start)
    return f"{first - 1}-{text.count(chr(10), 0, end) + 1}"


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = ref / "tests" / "test_safe_math_golden.inc"
    text = src.read_text()
    rows = extract_table(text, "G_CBRT", 2)
    out = {
        "source": "tests/test_safe_math_golden.inc:557-606 (G_CBRT, NG_CBRT=48)",
        "format": "[argument_bits_hex, result_bits_hex] IEEE binary64",
        "rows": rows,
    }
    (HERE / "cbrt_golden.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"cbrt rows: {len(rows)}")
    tables = {}
    for name, width in TABLES.items():
        t = extract_table(text, name, width)
        tables[name] = {"source": f"tests/test_safe_math_golden.inc:{line_span(text, name)}",
                        "format": "[x, y, result]" if width == 3 else "[x, result]", "rows": t}
        print(f"{name} rows: {len(t)}")
    (HERE / "safe_math_golden.json").write_text(json.dumps(tables, indent=1) + "\n")


if __name__ == "__main__":
    main()
