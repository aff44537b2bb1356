"""Convert the reference's committed FP32 snapshot files into compact fixtures.

The reference (Float/main.cu:1187-1211, writer wfile3d at Float/main.cu:1563-1588)
wrote eleven 192x192 y-slices of txx/vux/vuz/vwx/vwz as "%e " text with a blank
line after every row.  This script parses them, checks that re-printing the
parsed float32 values with "%e" reproduces every token of the source text, and
stores them as float32 arrays in ``float_snapshots.npz`` plus a manifest with
the sha256 of every source file.

Run from the repository root:  python tests/golden/make_fixtures.py
(it reads /root/reference/Float/*.dat; the outputs are committed so nothing
else ever needs the reference tree).
"""
import hashlib
import json
import os
import sys

import numpy as np

SRC = "/root/reference/Float"
HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ["txx500", "txx1000", "txx1500", "txx2000", "txx2500", "txx3000", "txx",
         "vux500", "vuz500", "vwx500", "vwz500"]


def parse(path):
    with open(path, "rb") as fh:
        raw = fh.read()
    text = raw.decode("ascii")
    rows = [ln.split() for ln in text.split("\n") if ln.strip()]
    arr = np.array([[float(t) for t in r] for r in rows], dtype=np.float32)
    toks = [t for r in rows for t in r]
    back = ["%e" % float(v) for v in arr.ravel()]
    if toks != back:
        bad = sum(a != b for a, b in zip(toks, back))
        raise SystemExit(f"{path}: {bad} tokens do not round-trip through float32")
    return arr, hashlib.sha256(raw).hexdigest()


def main():
    arrays, manifest = {}, {}
    for n in NAMES:
        arr, sha = parse(os.path.join(SRC, n + ".dat"))
        arrays[n] = arr
        manifest[n] = {"file": f"Float/{n}.dat", "shape": list(arr.shape), "sha256": sha,
                       "l2": float(np.sqrt(np.sum(arr.astype(np.float64) ** 2))),
                       "max_abs": float(np.max(np.abs(arr)))}
        print(n, arr.shape, manifest[n]["l2"], manifest[n]["max_abs"], sha[:16])
    np.savez_compressed(os.path.join(HERE, "float_snapshots.npz"), **arrays)
    with open(os.path.join(HERE, "float_snapshots.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
