"""Writes abd_orthogonality.json: the expected orthogonality gradient B and Hessian H that the
reference states in the comment of UnitTest/test_abd_energy.cpp (kappa 1e5, unit volume) for its
q. Run in a container that has the reference tree; the output is committed."""
import json
import re
import sys

src = open(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/UnitTest/test_abd_energy.cpp").read()
i = src.index("// Curr q = [[")
block = src[i:src.index("for (uint ii = 0; ii < 3; ii++)", i)]
txt = " ".join(l.strip().lstrip("/").strip() for l in block.splitlines())
q = [float(v) for v in re.search(r"Curr q = \[\[(.*?)\]\]", txt).group(1).split()]
B = [float(v) for v in re.findall(r"\[\s*(-?[\d.]+(?:e[-+]?\d+)?)\s*\]", txt.split("Output ortho B =")[1].split("], H =")[0] + "]")]
H = [[float(v) for v in r.split()] for r in re.findall(r"\[?\s*([-\d.e+ ]+?)\s*\]", txt.split("H = [[")[1])]
H = [r for r in H if len(r) == 12]
assert len(q) == 12 and len(B) == 12 and len(H) == 12
json.dump({"source": "UnitTest/test_abd_energy.cpp:55-84 (comment: expected orthogonality B and H)",
           "kappa": 1e5, "volume": 1.0, "q": q, "B": B, "H": H}, open("abd_orthogonality.json", "w"), indent=1)
