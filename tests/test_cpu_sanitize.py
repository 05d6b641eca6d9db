"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host-only code (tests/sanitize/):
the oracle's entry points, the library's seeded generators and panel packer, and the two
drop-in executables' file readers on well-formed and malformed datasets in the reference's
formats.  Without a GPU the executables stop, with an error, where they first need the
device, after every file has been read and checked."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import snicit_amd as sa
from tests import cli_files

HERE = os.path.dirname(os.path.abspath(__file__))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
SANITIZER_MARKS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs the host compilers")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("sanitize"))
    targets = ["oracle_harness", "host_harness"]
    have_lib = os.path.exists(sa.LIB_PATH)
    if have_lib:
        targets += ["snicit_sdgc", "snicit_beyond"]
    r = subprocess.run(["make", "-s", "-j4", "-C", os.path.join(HERE, "sanitize"), f"OUT={out}"]
                       + [os.path.join(out, t) for t in targets],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    return out, have_lib


def run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, env=ENV, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=600)
    for m in SANITIZER_MARKS:
        assert m not in r.stderr, r.stderr[-4000:]
    return r


def test_oracle_under_sanitizers(built):
    r = run([os.path.join(built[0], "oracle_harness")])
    assert r.returncode == 0 and "oracle harness ok" in r.stdout, r.stderr


def test_generators_and_packer_under_sanitizers(built):
    r = run([os.path.join(built[0], "host_harness")])
    assert r.returncode == 0 and "host harness ok" in r.stdout, r.stderr


def _stopped_at_device_or_ran(r):
    """Without a GPU the run ends at the context; with one it runs to its result line."""
    if r.returncode == 0:
        return True
    assert r.returncode == 1 and "context:" in r.stderr, (r.returncode, r.stdout, r.stderr)
    return False


def _cli(built, name):
    out, have_lib = built
    if not have_lib:
        pytest.skip("libsnicit.so is not built")
    return os.path.join(out, name)


def test_sdgc_readers(built, tmp_path):
    exe = _cli(built, "snicit_sdgc")
    N, L, B = 1024, 3, 40
    data = tmp_path / "SDGC"
    cli_files.write_sdgc_dataset(str(data), N, L, 3, B)
    base = [exe, "-n", str(N), "-l", str(L), "-t", "1", "-b", "20", "--inputs", str(B),
            "--dataset", str(data)]
    _stopped_at_device_or_ran(run(base))
    _stopped_at_device_or_ran(run(base + ["--hash-reorder"]))
    _stopped_at_device_or_ran(run([exe, "--synthetic", "-n", "1024", "-l", "4", "-b", "8",
                                   "--inputs", "16"]))
    # malformed weight files
    layer = data / "tsv_weights" / f"neuron{N}" / f"n{N}-l2.tsv"
    good = layer.read_text()
    layer.write_text(good + f"{N + 1}\t1\t0.0625\n")
    r = run(base)
    assert r.returncode == 1 and "index out of range" in r.stderr
    layer.write_text(good + "".join(f"{i}\t5\t0.0625\n" for i in range(1, 40)))
    r = run(base)
    assert r.returncode == 1 and "more than 32 nonzeros" in r.stderr
    layer.unlink()
    r = run(base)
    assert r.returncode == 1 and "cannot open" in r.stderr
    # arguments
    r = run([exe, "-n", "1000"])
    assert r.returncode == 1 and "N must be one of" in r.stderr
    r = run([exe, "-b", "7", "--inputs", "20", "--synthetic"])
    assert r.returncode == 1 and "factor" in r.stderr


def test_beyond_readers(built, tmp_path):
    exe = _cli(built, "snicit_beyond")
    data = tmp_path / "beyond"
    cli_files.write_beyond_net_a(str(data), np.arange(30) % 10)
    base = [exe, "-k", "A", "-n", "30", "-b", "10", "--dataset", str(data)]
    r = run(base)  # labels but no images
    assert r.returncode == 1 and "input images not found" in r.stderr
    _stopped_at_device_or_ran(run(base + ["--synthetic"]))
    wdir = data / "networks" / "tsv_weights" / "n128-l18-acc94.94"
    bdir = data / "networks" / "tsv_biases" / "n128-l18-acc94.94"
    sp = wdir / "l5-sparse.tsv"
    good = sp.read_text()
    sp.write_text(good + "128\t0\t0.5\n")
    r = run(base + ["--synthetic"])
    assert r.returncode == 1 and "index out of range" in r.stderr
    sp.write_text(good + "0\t0\t0.5\n")
    r = run(base + ["--synthetic"])
    assert r.returncode == 1 and "not grouped by input row" in r.stderr
    sp.write_text(good)
    bias = bdir / "l7-sparse.tsv"
    bgood = bias.read_text()
    bias.write_text(bgood + "0.25\n")
    r = run(base + ["--synthetic"])
    assert r.returncode == 1 and "bias size mismatch" in r.stderr
    bias.write_text(bgood.replace("\n", "\nx\n", 1))
    r = run(base + ["--synthetic"])
    assert r.returncode == 1 and "not a number at line 2" in r.stderr
    bias.write_text(bgood)
    dense = wdir / "l1-dense.tsv"
    dgood = dense.read_text()
    dense.write_text("0.5\n" * 100)  # a short file is refused before it is transposed
    r = run(base + ["--synthetic"])
    assert r.returncode == 1 and "dense layer size mismatch" in r.stderr
    dense.write_text(dgood)
    (wdir / "l20-dense.tsv").write_text("0.5\n" * 1281)
    r = run(base + ["--synthetic"])
    assert r.returncode == 1 and "dense layer size mismatch" in r.stderr


def test_beyond_cifar_readers(built, tmp_path):
    exe = _cli(built, "snicit_beyond")
    data = tmp_path / "beyond"
    cli_files.write_beyond_net_d(str(data))
    x = np.random.default_rng(3).random((20, 256)).astype(np.float32)
    cli_files.write_cifar(str(data), x, np.arange(20) % 10)
    _stopped_at_device_or_ran(run([exe, "-k", "D", "-n", "20", "-b", "10",
                                   "--dataset", str(data)]))
    # fewer label records than samples: the label file is not used, the inputs are
    lab = data / "CIFAR-10" / "cifar-label.bin"
    lab.write_bytes(lab.read_bytes()[:3073 * 19])
    _stopped_at_device_or_ran(run([exe, "-k", "D", "-n", "20", "-b", "10",
                                   "--dataset", str(data)]))
    lab.write_bytes(lab.read_bytes()[:3073 * 18 + 1])  # the 19th record's label byte alone
    _stopped_at_device_or_ran(run([exe, "-k", "D", "-n", "19", "-b", "19",
                                   "--dataset", str(data)]))
    r = run([exe, "-k", "D", "-n", "21", "-b", "7", "--dataset", str(data)])
    assert r.returncode == 1 and "input images not found" in r.stderr
    r = run([exe, "-k", "E"])
    assert r.returncode == 1 and "Error benchmark" in r.stderr
