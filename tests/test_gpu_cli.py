"""The drop-in executables, run on datasets written in the reference's file formats: the Graph
Challenge TSVs for snicit_sdgc (main/SDGC.cu:58-62) and the beyond TSVs and MNIST label file
for snicit_beyond (main/beyond.cu:57-95).  Expected categories come from the oracle."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu

PKG = os.path.dirname(sa.LIB_PATH)


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout, flush=True)
    return r


def _write_sdgc_dataset(root, N, L, seed, B):
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


@pytest.mark.parametrize("mode", ["SNICIT", "BF", "XY"])
def test_sdgc_cli_on_graph_challenge_files(tmp_path, mode):
    N, L, B = 1024, 24, 2000
    data = tmp_path / "dataset" / "SDGC"
    cat = _write_sdgc_dataset(str(data), N, L, 3, B)
    assert 0 < cat.sum() < B
    r = _run([os.path.join(PKG, "snicit_sdgc"), "-m", mode, "-n", str(N), "-l", str(L), "-t", "6",
              "-b", "1000", "--inputs", str(B), "--dataset", str(data)], cwd=str(tmp_path))
    assert r.returncode == 0
    assert "CHALLENGE PASSED" in r.stdout
    assert re.search(rf"{mode} info: runtime [0-9.e+-]+ ?ms", r.stdout)


@pytest.mark.parametrize("mode", ["SNICIT", "BF"])
def test_sdgc_cli_hash_reorder_keeps_categories(tmp_path, mode):
    """--hash-reorder renumbers every neuron of the weight and input files by the reference's
    bucket hash (src/SDGC/reorder/hash.h:14-21) before the run; the golden categories file is
    written for the original numbering, so passing shows the renumbered run classifies alike."""
    N, L, B = 1024, 24, 2000
    data = tmp_path / "dataset" / "SDGC"
    _write_sdgc_dataset(str(data), N, L, 4, B)
    base = [os.path.join(PKG, "snicit_sdgc"), "-m", mode, "-n", str(N), "-l", str(L), "-t", "6",
            "-b", "1000", "--inputs", str(B), "--dataset", str(data)]
    r = _run(base + ["--hash-reorder"], cwd=str(tmp_path))
    assert r.returncode == 0
    assert "renumbered by the bucket hash" in r.stdout
    assert "Number of different categories: 0" in r.stdout and "CHALLENGE PASSED" in r.stdout
    r = _run([os.path.join(PKG, "snicit_sdgc"), "--synthetic", "--hash-reorder", "-m", mode,
              "-n", "4096", "-l", "30", "-t", "8", "-b", "2000", "--inputs", "4000"],
             cwd=str(tmp_path))
    assert r.returncode == 0 and "CHALLENGE PASSED" in r.stdout


def test_sdgc_cli_synthetic(tmp_path):
    r = _run([os.path.join(PKG, "snicit_sdgc"), "--synthetic", "-n", "4096", "-l", "40", "-t", "10",
              "-b", "3000", "--inputs", "6000"], cwd=str(tmp_path))
    assert r.returncode == 0 and "CHALLENGE PASSED" in r.stdout
    assert re.search(r"SNICIT info: runtime [0-9.e+-]+ms avgpost [0-9.e+-]+ms", r.stdout)


def _write_beyond_net_a(root, labels):
    a = nets.beyond_net_a()
    name = "n128-l18-acc94.94"
    wdir = os.path.join(root, "networks", "tsv_weights", name)
    bdir = os.path.join(root, "networks", "tsv_biases", name)
    os.makedirs(wdir)
    os.makedirs(bdir)

    def values(path, arr):
        with open(path, "w") as f:
            f.write("\n".join(repr(float(v)) for v in arr.ravel()) + "\n")

    values(os.path.join(wdir, "l1-dense.tsv"), a["w_in"])
    values(os.path.join(bdir, "l1-dense.tsv"), a["b_in"])
    for h in range(18):
        rp = a["row_ptr"][h]
        with open(os.path.join(wdir, f"l{h + 2}-sparse.tsv"), "w") as f:
            for i in range(128):
                for e in range(rp[i], rp[i + 1]):
                    f.write(f"{i}\t{a['cols'][e]}\t{float(a['vals'][e])!r}\n")
        values(os.path.join(bdir, f"l{h + 2}-sparse.tsv"), a["bias"][h])
    values(os.path.join(wdir, "l20-dense.tsv"), a["w_out"])
    values(os.path.join(bdir, "l20-dense.tsv"), a["b_out"])
    mdir = os.path.join(root, "MNIST")
    os.makedirs(mdir)
    with open(os.path.join(mdir, "t10k-labels-idx1-ubyte"), "wb") as f:
        f.write(struct.pack(">ii", 2049, len(labels)) + bytes(int(v) for v in labels))
    return a


@pytest.mark.parametrize("mode", ["BF", "SNICIT"])
def test_beyond_cli_on_reference_files(tmp_path, mode):
    """Net A's files in the reference layout, idx images and labels written from a seeded
    sparse input set; labels are the oracle's plain categories, so BF scores 100%."""
    B = 2000
    r0 = np.random.default_rng(1)
    img = ((r0.random((B, 784)) < 0.2) * r0.integers(1, 256, (B, 784))).astype(np.uint8)
    x = img.astype(np.float32) / 255.0
    a = nets.beyond_net_a()
    cols, vals = nets.csr_inputs_to_ell(a["row_ptr"], a["cols"], a["vals"], 128)
    _, _, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, w_in=a["w_in"].T.copy(),
                                b_in=a["b_in"], in_clip=1.0, w_out=a["w_out"].T.copy(),
                                b_out=a["b_out"])
    data = tmp_path / "dataset" / "beyond"
    _write_beyond_net_a(str(data), cat)
    with open(os.path.join(str(data), "MNIST", "t10k-images-idx3-ubyte"), "wb") as f:
        f.write(struct.pack(">iiii", 2051, B, 28, 28) + img.tobytes())
    r = _run([os.path.join(PKG, "snicit_beyond"), "-m", mode, "-k", "A", "-n", str(B), "-b", "1000",
              "--dataset", str(data)], cwd=str(tmp_path))
    assert r.returncode == 0
    m = re.search(rf"{mode} info: accuracy ([0-9.]+)% runtime [0-9.e+-]+ms avgpost [0-9.e+-]+ms",
                  r.stdout)
    assert m
    acc = float(m.group(1))
    if mode == "BF":
        assert acc == 100.0
    else:
        want = 0.0
        for off in range(0, B, 1000):
            rr = orc.model_snicit_lossy(cols, vals, a["bias"], 1.0, x[off:off + 1000], 8,
                                        w_in=a["w_in"].T.copy(), b_in=a["b_in"],
                                        w_out=a["w_out"].T.copy(), b_out=a["b_out"])
            want += (rr["cat"] == cat[off:off + 1000]).sum()
        assert abs(acc - 100.0 * want / B) < 1e-3


def _write_beyond_net_d(root):
    """Net D's files in the reference layout (no input layer: hidden layers l1..l12, output
    layer l13, src/beyond/SNICIT/SNICIT.hpp:175-365)."""
    a = nets.beyond_net("D")
    name = "n256-l12-acc75.86"
    wdir = os.path.join(root, "networks", "tsv_weights", name)
    bdir = os.path.join(root, "networks", "tsv_biases", name)
    os.makedirs(wdir)
    os.makedirs(bdir)

    def values(path, arr):
        with open(path, "w") as f:
            f.write("\n".join(repr(float(v)) for v in arr.ravel()) + "\n")

    for h in range(12):
        rp = a["row_ptr"][h]
        with open(os.path.join(wdir, f"l{h + 1}-sparse.tsv"), "w") as f:
            for i in range(256):
                for e in range(rp[i], rp[i + 1]):
                    f.write(f"{i}\t{a['cols'][e]}\t{float(a['vals'][e])!r}\n")
        values(os.path.join(bdir, f"l{h + 1}-sparse.tsv"), a["bias"][h])
    values(os.path.join(wdir, "l13-dense.tsv"), a["w_out"])
    values(os.path.join(bdir, "l13-dense.tsv"), a["b_out"])
    return a


@pytest.mark.parametrize("mode", ["BF", "SNICIT"])
def test_beyond_cli_cifar_net_d(tmp_path, mode):
    """Net D's files, CIFAR-10 inputs as text (one of the 256 features per line) and labels as
    3073-byte binary records (label byte first), the formats the reference reads
    (src/beyond/SNICIT/SNICIT.hpp:402-415,462-478).  Labels are the oracle's plain
    categories, so BF scores 100% and lossy SNICIT scores the oracle's agreement."""
    B, N = 1200, 256
    x = nets.prototype_features(N, B, 41)
    data = tmp_path / "dataset" / "beyond"
    a = _write_beyond_net_d(str(data))
    cols, vals = nets.csr_inputs_to_ell(a["row_ptr"], a["cols"], a["vals"], N)
    kw = dict(w_out=a["w_out"].T.copy(), b_out=a["b_out"])
    _, _, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, **kw)
    cdir = os.path.join(str(data), "CIFAR-10")
    os.makedirs(cdir)
    with open(os.path.join(cdir, "cifar-input.txt"), "w") as f:
        f.write("\n".join(repr(float(v)) for v in x.ravel()) + "\n")
    rec = np.random.default_rng(2).integers(0, 256, (B, 3073)).astype(np.uint8)
    rec[:, 0] = cat
    with open(os.path.join(cdir, "cifar-label.bin"), "wb") as f:
        f.write(rec.tobytes())
    r = _run([os.path.join(PKG, "snicit_beyond"), "-m", mode, "-k", "D", "-n", str(B), "-b", "600",
              "--dataset", str(data)], cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout
    m = re.search(rf"{mode} info: accuracy ([0-9.]+)% runtime [0-9.e+-]+ms avgpost [0-9.e+-]+ms",
                  r.stdout)
    assert m, r.stdout
    acc = float(m.group(1))
    if mode == "BF":
        assert acc == 100.0
    else:
        want = 0
        for off in range(0, B, 600):
            rr = orc.model_snicit_lossy(cols, vals, a["bias"], 1.0, x[off:off + 600], 6, **kw)
            want += int((rr["cat"] == cat[off:off + 600]).sum())
        assert abs(acc - 100.0 * want / B) < 1e-3
