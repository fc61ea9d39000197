"""Writes tests/golden/ewald_100.json from the reference's Ewald known-answer test.

Reads test/ewald_test.cpp of the reference (as text, in this container only):
the 100 explicit +-1 charges and positions in a periodic cube L = 100
(test/ewald_test.cpp:11-114) and the EwaldSummations.jl answer
E/(4 pi) = -0.07897291545871976 (test/ewald_test.cpp:143-144).  The JSON holds
only that data.
"""
import json
import pathlib
import re
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "test" / "ewald_test.cpp"
OUT = pathlib.Path(__file__).with_name("ewald_100.json")

text = REF.read_text()
q_block = re.search(r"double q_array\[\]\s*=\s*\{([^}]*)\}", text).group(1)
q = [float(v) for v in q_block.replace("\n", " ").split(",") if v.strip()]
r_block = re.search(r"double r_array\[\]\[3\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
rows = re.findall(r"\{([^}]*)\}", r_block)
r = [[float(v) for v in row.split(",")] for row in rows]
expected = float(re.search(r"EXPECT_NEAR\(E / \(4 \* M_PI\), ([-0-9.e]+), ([0-9.e-]+)\)", text).group(1))
assert len(q) == 100 and len(r) == 100, (len(q), len(r))
OUT.write_text(json.dumps({
    "source": "test/ewald_test.cpp:11-156",
    "L": 100.0, "s": 5.0, "alphas": [0.2, 0.3, 0.4, 0.5], "eps": 1.0,
    "q": q, "r": r,
    "E_over_4pi": expected, "tol": 1e-10,
    "potential_point": [10.0, 10.0, 10.0],
}, indent=1))
print("wrote", OUT, len(q), "charges")
