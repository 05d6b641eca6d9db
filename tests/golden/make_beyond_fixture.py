"""Writes tests/golden/beyond_net{A,B,C,D}.npz from the reference's beyond nets
(dataset/beyond/networks/{tsv_weights,tsv_biases}/<net>/), read as data:

  A  n128-l18-acc94.94  l1-dense (784 x 128 input layer), l2..l19 sparse, l20-dense output
  B  n256-l18-acc96.88  l2..l19 sparse, l20-dense output (l1-dense weights are not in the
                        reference tree; its biases are)
  C  n256-l12-acc95.61  l2..l13 sparse, l14-dense output (l1-dense weights likewise absent)
  D  n256-l12-acc75.86  no input layer (CIFAR-10 inputs are 256 features fed to the hidden
                        layers, src/beyond/SNICIT/SNICIT.hpp:96,514-516), l1..l12 sparse,
                        l13-dense output

File formats (src/beyond/SNICIT/SNICIT.hpp:175-365):
  l<k>-dense.tsv   one value per line, input-major (w[i*out + j] feeds input i to output j)
  l<k>-sparse.tsv  "in<TAB>out<TAB>value" lines grouped by input row (CSR over inputs)
  tsv_biases/...   one bias per line for each of the same files
Values are parsed with the C library's strtof, as the reference's std::stof does.

Run here (the reference tree is not on the GPU machines); the .npz files are committed.
Column indices are stored as uint8 (N <= 256).
"""
import ctypes
import os
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/dataset/beyond/networks"
NETS = {
    "A": dict(dir="n128-l18-acc94.94", N=128, L=18, first=2, nin=784),
    "B": dict(dir="n256-l18-acc96.88", N=256, L=18, first=2, nin=784),
    "C": dict(dir="n256-l12-acc95.61", N=256, L=12, first=2, nin=784),
    "D": dict(dir="n256-l12-acc75.86", N=256, L=12, first=1, nin=0),
}
NOUT = 10
_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def strtof(s):
    return _libc.strtof(s.encode(), None)


def lines(path):
    with open(path) as f:
        return [ln.strip() for ln in f if ln.strip()]


def values(path):
    return np.array([strtof(v) for v in lines(path)], np.float32)


def make(key):
    d = NETS[key]
    N, L, first = d["N"], d["L"], d["first"]
    wdir = os.path.join(SRC, "tsv_weights", d["dir"])
    bdir = os.path.join(SRC, "tsv_biases", d["dir"])
    out = {}
    if d["nin"]:
        win = os.path.join(wdir, "l1-dense.tsv")
        if os.path.exists(win):
            out["w_in"] = values(win).reshape(d["nin"], N)
        out["b_in"] = values(os.path.join(bdir, "l1-dense.tsv"))
    row_ptr = np.zeros((L, N + 1), np.int64)
    cols, vals, bias = [], [], []
    base = 0
    for h in range(L):
        name = f"l{h + first}-sparse.tsv"
        ent = [ln.split("\t") for ln in lines(os.path.join(wdir, name))]
        rows = np.array([int(e[0]) for e in ent], np.int64)
        c = np.array([int(e[1]) for e in ent], np.int64)
        assert np.all(np.diff(rows) >= 0), "entries must be grouped by input row"
        assert rows.min() >= 0 and rows.max() < N and c.min() >= 0 and c.max() < N
        row_ptr[h, 1:] = np.cumsum(np.bincount(rows, minlength=N))
        row_ptr[h] += base
        base += len(ent)
        cols.append(c.astype(np.uint8))
        vals.append(np.array([strtof(e[2]) for e in ent], np.float32))
        bias.append(values(os.path.join(bdir, name)))
    oname = f"l{L + first}-dense.tsv"
    out.update(row_ptr=row_ptr, cols=np.concatenate(cols), vals=np.concatenate(vals),
               bias=np.stack(bias), w_out=values(os.path.join(wdir, oname)).reshape(N, NOUT),
               b_out=values(os.path.join(bdir, oname)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"beyond_net{key}.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    for k in (sys.argv[2:] or NETS):
        make(k)
