"""Writes tests/golden/beyond_netA.npz from the reference's beyond net A files
(dataset/beyond/networks/{tsv_weights,tsv_biases}/n128-l18-acc94.94/), read as data:

  l1-dense.tsv        784 x 128 input-layer weights, one value per line, input-major
                      (weight[i*128 + k] feeds input i to neuron k, src/beyond/general.hpp:5-33)
  l{2..19}-sparse.tsv hidden layers, "in<TAB>out<TAB>value" lines grouped by input row
                      (CSR over input neurons, src/beyond/SNICIT/SNICIT.hpp:261-283)
  l20-dense.tsv       128 x 10 output weights, input-major (general.hpp:74-102)
  tsv_biases/...      one bias per line for each of the same files

Run here (the reference tree is not on the GPU machines); the .npz is committed.
"""
import os
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/dataset/beyond/networks"
NET = "n128-l18-acc94.94"
N, L, NIN, NOUT = 128, 18, 784, 10


def lines(path):
    with open(path) as f:
        return [ln.strip() for ln in f if ln.strip()]


def main():
    wdir = os.path.join(SRC, "tsv_weights", NET)
    bdir = os.path.join(SRC, "tsv_biases", NET)
    w_in = np.array(lines(os.path.join(wdir, "l1-dense.tsv")), np.float32).reshape(NIN, N)
    b_in = np.array(lines(os.path.join(bdir, "l1-dense.tsv")), np.float32)
    row_ptr = np.zeros((L, N + 1), np.int64)
    cols, vals, bias = [], [], []
    base = 0
    for h in range(L):
        name = f"l{h + 2}-sparse.tsv"
        ent = [ln.split("\t") for ln in lines(os.path.join(wdir, name))]
        rows = np.array([int(e[0]) for e in ent], np.int64)
        assert np.all(np.diff(rows) >= 0), "entries must be grouped by input row"
        counts = np.bincount(rows, minlength=N)
        row_ptr[h, 1:] = np.cumsum(counts)
        row_ptr[h] += base
        base += len(ent)
        cols.append(np.array([int(e[1]) for e in ent], np.int32))
        vals.append(np.array([float(e[2]) for e in ent], np.float32))
        bias.append(np.array(lines(os.path.join(bdir, name)), np.float32))
    w_out = np.array(lines(os.path.join(wdir, f"l{L + 2}-dense.tsv")), np.float32).reshape(N, NOUT)
    b_out = np.array(lines(os.path.join(bdir, f"l{L + 2}-dense.tsv")), np.float32)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "beyond_netA.npz")
    np.savez_compressed(out, w_in=w_in, b_in=b_in, row_ptr=row_ptr,
                        cols=np.concatenate(cols), vals=np.concatenate(vals),
                        bias=np.stack(bias), w_out=w_out, b_out=b_out)
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
