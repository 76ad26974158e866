"""Writes transports-h6.npz: the T/O/I FCC transports at h = 6 (CPU float64) from the
reference's tests/fixtures/browser-reference.json (which also holds h = 2 and 12, already
in reference-small.npz and transports-h8-h12.npz).  Run in the container that holds the
reference checkout:  python tests/golden/make_transports_h6.py
"""
import json
from pathlib import Path

import numpy as np

SOURCE = Path("/root/reference/tests/fixtures/browser-reference.json")
OUT = Path(__file__).resolve().parent / "transports-h6.npz"

if __name__ == "__main__":
    data = json.loads(SOURCE.read_text(encoding="utf-8"))
    np.savez_compressed(OUT, **{k: np.asarray(data[k], np.float64) for k in ("T_6", "O_6", "I_6")})
    print(OUT)
