"""Write tests/golden/a3m_golden.json: inputs and outputs of the reference's
own `sw_affine_backtrack` (sseqs/sw.py:27-219, Triton DP + traceback) and
`to_a3m` (sseqs/to_a3m.py:5-72), run from the reference checkout in this
container on the CPU with Triton's interpreter (TRITON_INTERPRET=1).

sseqs/sw.py builds a CUDA extension at import (sseqs/sw.py:9) that neither
function uses; torch.utils.cpp_extension.load is replaced by a stub for the
import so that nothing is compiled, and the package's __init__ (which loads
a database) is not executed.  Run once, by hand:

    python tests/golden/make_a3m_golden.py [/path/to/reference]

Only the JSON it writes is data for the tests; this script is not imported
by them.
"""
import functools
import json
import os
import sys
import tempfile
import types
from pathlib import Path

import numpy as np

os.environ["TRITON_INTERPRET"] = "1"
REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "a3m_golden.json"

AA = "ARNDCQEGHILKMFPSTWYV"


def mutate(rng, s, sub, indel):
    out = []
    for c in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append("".join(rng.choice(list(AA), int(rng.integers(1, 4)))))
        out.append(str(rng.choice(list(AA))) if rng.random() < sub else c)
    return "".join(out) or "A"


def cases():
    rng = np.random.default_rng(20261202)
    query = "".join(rng.choice(list(AA), 45)) + "X" + "".join(rng.choice(list(AA), 14))
    targets = []
    for i in range(10):
        targets.append(mutate(rng, query, sub=0.1 + 0.06 * i, indel=0.04))
    # a fragment, an extension on both sides, unrelated sequences, letters
    # outside the 20 residues, exact duplicates (equal scores and rows)
    targets.append(query[10:40])
    targets.append("".join(rng.choice(list(AA), 8)) + query + "".join(rng.choice(list(AA), 6)))
    targets += ["".join(rng.choice(list(AA), int(rng.integers(5, 70)))) for _ in range(5)]
    targets.append(mutate(rng, query, 0.2, 0.0).replace("L", "B").replace("K", "Z"))
    targets.append("WWWW")
    targets += [targets[0], targets[3], targets[3]]
    return [(query, targets, 11, 1), (query[:33], targets[:12], 5, 2)]


def main():
    import torch.utils.cpp_extension as ce

    ce.load = lambda *a, **k: types.SimpleNamespace()
    pkg = types.ModuleType("sseqs")
    pkg.__path__ = [str(REF / "sseqs")]
    sys.modules["sseqs"] = pkg
    import sseqs.sw as rsw
    import sseqs.to_a3m as ra3m

    ra3m.sw_affine_backtrack = functools.partial(rsw.sw_affine_backtrack, device="cpu")
    out = []
    for query, targets, go, ge in cases():
        _, best, aligned = rsw.sw_affine_backtrack(query, targets, go, ge, device="cpu")
        with tempfile.TemporaryDirectory() as d:
            fn = os.path.join(d, "out.a3m")
            ra3m.to_a3m(query, targets, fn, gap_open=go, gap_extend=ge)
            text = open(fn).read()
        out.append({"query": query, "targets": targets, "gap_open": go, "gap_extend": ge,
                    "best": [int(x) for x in best.tolist()], "aligned": aligned, "a3m": text})
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}: {len(out)} cases")


if __name__ == "__main__":
    main()
