"""Write the reference-held golden vectors this repo's tests use as fixtures.

Reads the reference's test data files as text (nothing is compiled, imported
or run) and writes small JSON fixtures next to this script:

* ``cbrt_golden.json`` -- the 48 (argument, result) IEEE binary64 bit rows of
  ``G_CBRT`` in ``tests/test_safe_math_golden.inc:557-606``, produced there by
  the contraction-off CPU builds of ``rdb_sm_cbrt``.

Usage: python tests/golden/extract_reference_vectors.py [/path/to/reference]
"""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def extract_table(text: str, name: str, width: int) -> list[list[str]]:
    m = re.search(name + r"\(" + str(width) + r", NG_\w+\) = reshape\(\[character\(len=16\) :: &(.*?)\]", text, re.S)
    if not m:
        raise SystemExit(f"table {name} not found")
    hexes = re.findall(r'"([0-9A-F]{16})"', m.group(1))
    if len(hexes) % width:
        raise SystemExit(f"table {name}: {len(hexes)} entries not a multiple of {width}")
    return [hexes[i:i + width] for i in range(0, len(hexes), width)]


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


if __name__ == "__main__":
    main()
