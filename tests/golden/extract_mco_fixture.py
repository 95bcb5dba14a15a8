"""Write tests/golden/mcml_fixture.mco: the MCML output file that the reference's own test of its
.mco parser holds as a string literal (tests/python/test_mcml_compare.py, the MCO constant).
The fixture is data (an MCML output file); the reference's expected values for it are restated
in tests/test_mco.py.

    python tests/golden/extract_mco_fixture.py
"""

import ast
from pathlib import Path

SRC = Path("/root/reference/tests/python/test_mcml_compare.py")
OUT = Path(__file__).resolve().parent / "mcml_fixture.mco"


def main() -> int:
    tree = ast.parse(SRC.read_text())
    for node in tree.body:
        if isinstance(node, ast.Assign) and [t.id for t in node.targets] == ["MCO"]:
            OUT.write_text(ast.literal_eval(node.value))
            print(f"wrote {OUT}")
            return 0
    raise SystemExit("no MCO fixture found")


if __name__ == "__main__":
    raise SystemExit(main())
