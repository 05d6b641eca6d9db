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
from tests import cli_files, nets

pytestmark = pytest.mark.gpu

PKG = os.path.dirname(sa.LIB_PATH)


def _run(cmd, cwd):
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout, flush=True)
    return r


@pytest.mark.parametrize("mode", ["SNICIT", "BF", "XY"])
def test_sdgc_cli_on_graph_challenge_files(tmp_path, mode):
    N, L, B = 1024, 24, 2000
    data = tmp_path / "dataset" / "SDGC"
    cat = cli_files.write_sdgc_dataset(str(data), N, L, 3, B)
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
    cli_files.write_sdgc_dataset(str(data), N, L, 4, B)
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
    cli_files.write_beyond_net_a(str(data), cat)
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


@pytest.mark.parametrize("mode", ["BF", "SNICIT"])
def test_beyond_cli_cifar_net_d(tmp_path, mode):
    """Net D's files, CIFAR-10 inputs as text (one of the 256 features per line) and labels as
    3073-byte binary records (label byte first), the formats the reference reads
    (src/beyond/SNICIT/SNICIT.hpp:402-415,462-478).  Labels are the oracle's plain
    categories, so BF scores 100% and lossy SNICIT scores the oracle's agreement."""
    B, N = 1200, 256
    x = nets.prototype_features(N, B, 41)
    data = tmp_path / "dataset" / "beyond"
    a = cli_files.write_beyond_net_d(str(data))
    cols, vals = nets.csr_inputs_to_ell(a["row_ptr"], a["cols"], a["vals"], N)
    kw = dict(w_out=a["w_out"].T.copy(), b_out=a["b_out"])
    _, _, cat = orc.model_plain(cols, vals, a["bias"], 1.0, x, **kw)
    cli_files.write_cifar(str(data), x, cat)
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
