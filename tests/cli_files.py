"""Datasets in the reference's file formats for the drop-in executables: the Graph Challenge
TSVs read by snicit_sdgc (main/SDGC.cu:58-62, utility.hpp:109-113,243-293) and the beyond
TSVs, idx label file and CIFAR-10 files read by snicit_beyond (main/beyond.cu:57-95,
src/beyond/SNICIT/SNICIT.hpp:175-365,402-415,462-478)."""
import os
import struct

import numpy as np

import oracle as orc
import snicit_amd as sa
from tests import nets


def _values(path, arr):
    with open(path, "w") as f:
        f.write("\n".join(repr(float(v)) for v in arr.ravel()) + "\n")


def write_sdgc_dataset(root, N, L, seed, B):
    cols = sa.gen_sdgc_cols(N, L, seed)
    wdir = os.path.join(root, "tsv_weights", f"neuron{N}")
    os.makedirs(wdir)
    for l in range(L):
        with open(os.path.join(wdir, f"n{N}-l{l + 1}.tsv"), "w") as f:
            for o in range(N):
                for i in cols[l, o]:
                    f.write(f"{i + 1}\t{o + 1}\t0.0625\n")
    y0 = sa.gen_images(N, 5, 0, B)
    mdir = os.path.join(root, "MNIST")
    os.makedirs(mdir)
    with open(os.path.join(mdir, f"sparse-images-{N}.tsv"), "w") as f:
        for b, n in zip(*np.nonzero(y0)):
            f.write(f"{b + 1}\t{n + 1}\t1\n")
    _, cat, _ = orc.plain_f32(cols, np.full(cols.shape, 0.0625, np.float32), sa.SDGC_BIAS[N],
                              32.0, y0, want_y=False)
    with open(os.path.join(mdir, f"neuron{N}-l{L}-categories.tsv"), "w") as f:
        for b in np.nonzero(cat)[0]:
            f.write(f"{b + 1}\n")
    return cat


def write_beyond_net_a(root, labels):
    a = nets.beyond_net_a()
    name = "n128-l18-acc94.94"
    wdir = os.path.join(root, "networks", "tsv_weights", name)
    bdir = os.path.join(root, "networks", "tsv_biases", name)
    os.makedirs(wdir)
    os.makedirs(bdir)

    _values(os.path.join(wdir, "l1-dense.tsv"), a["w_in"])
    _values(os.path.join(bdir, "l1-dense.tsv"), a["b_in"])
    for h in range(18):
        rp = a["row_ptr"][h]
        with open(os.path.join(wdir, f"l{h + 2}-sparse.tsv"), "w") as f:
            for i in range(128):
                for e in range(rp[i], rp[i + 1]):
                    f.write(f"{i}\t{a['cols'][e]}\t{float(a['vals'][e])!r}\n")
        _values(os.path.join(bdir, f"l{h + 2}-sparse.tsv"), a["bias"][h])
    _values(os.path.join(wdir, "l20-dense.tsv"), a["w_out"])
    _values(os.path.join(bdir, "l20-dense.tsv"), a["b_out"])
    mdir = os.path.join(root, "MNIST")
    os.makedirs(mdir)
    with open(os.path.join(mdir, "t10k-labels-idx1-ubyte"), "wb") as f:
        f.write(struct.pack(">ii", 2049, len(labels)) + bytes(int(v) for v in labels))
    return a


def write_beyond_net_d(root):
    """Net D's files in the reference layout (no input layer: hidden layers l1..l12, output
    layer l13, src/beyond/SNICIT/SNICIT.hpp:175-365)."""
    a = nets.beyond_net("D")
    name = "n256-l12-acc75.86"
    wdir = os.path.join(root, "networks", "tsv_weights", name)
    bdir = os.path.join(root, "networks", "tsv_biases", name)
    os.makedirs(wdir)
    os.makedirs(bdir)

    for h in range(12):
        rp = a["row_ptr"][h]
        with open(os.path.join(wdir, f"l{h + 1}-sparse.tsv"), "w") as f:
            for i in range(256):
                for e in range(rp[i], rp[i + 1]):
                    f.write(f"{i}\t{a['cols'][e]}\t{float(a['vals'][e])!r}\n")
        _values(os.path.join(bdir, f"l{h + 1}-sparse.tsv"), a["bias"][h])
    _values(os.path.join(wdir, "l13-dense.tsv"), a["w_out"])
    _values(os.path.join(bdir, "l13-dense.tsv"), a["b_out"])
    return a


def write_cifar(root, x, labels, seed=2):
    """CIFAR-10 inputs as text, one feature per line, and labels as 3073-byte binary records
    (the label byte, then 3072 pixel bytes, here random)."""
    cdir = os.path.join(root, "CIFAR-10")
    os.makedirs(cdir)
    _values(os.path.join(cdir, "cifar-input.txt"), x)
    rec = np.random.default_rng(seed).integers(0, 256, (len(labels), 3073)).astype(np.uint8)
    rec[:, 0] = labels
    with open(os.path.join(cdir, "cifar-label.bin"), "wb") as f:
        f.write(rec.tobytes())
