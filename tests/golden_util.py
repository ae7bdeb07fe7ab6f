"""Loading the golden fixtures of tests/golden (arrays only, no pickle)."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
# deck fixtures (make_golden.py); other fixtures (make_xxz512_golden.py)
# hold no deck
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz") if not p.stem.endswith("_e0"))


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["deck"] = json.loads(str(d["deck"]))
    return d


def matrices(deck):
    return [(np.array([[complex(re, im) for re, im in row] for row in t["matrix"]]), t["sites"])
            for t in deck["terms"]]
