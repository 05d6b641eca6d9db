"""The benchmark script's own logic, without a GPU: it refuses more than one GPU, its
algorithmic byte counts are the stated closed forms, the committed counter summaries are
found for the kernels the benchmark line can name, and tuned shapes reach every setter."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_layer_bytes_closed_forms():
    # plain layer: read and write each live activation, 32 ids and 32 weights per panel,
    # P + 1 offsets, N output ids
    assert bench.spmm_layer_bytes(16, 1, 3) == 8 * 16 * 3 + 32 * 4 * 2 + 2 * 4 + 16 * 4
    assert bench.spmm_layer_bytes(65536, 2048, 0) == 2048 * 256 + 2049 * 4 + 65536 * 4
    # delta layer: adds the centroid index per column and a 64-value chain record per panel
    assert bench.post_layer_bytes(16, 1, 3) == (8 * 16 * 3 + 4 * 3 + (32 * 4 * 2 + 64 * 4)
                                                + 2 * 4 + 16 * 4)


@pytest.mark.parametrize("cfg", ["c2", "c3"])
@pytest.mark.parametrize("kernel", ["panel_spmm_tiles_kernel<1>", "panel_spmm_tiles_kernel<2>",
                                    "panel_spmm_tiles_kernel<4>", "panel_post_wave_kernel"])
def test_committed_traffic_is_found(cfg, kernel):
    wl = bench.WORKLOADS[cfg]
    traffic, source = bench.measured_traffic(wl, kernel)
    assert source is not None and source.split("<")[0] == kernel.split("<")[0]
    rows = wl["B"]
    # between the live bytes of a nearly empty launch and 1.5x a full one
    assert 0 < traffic < 1.5 * 8 * wl["N"] * rows + 1e9
    assert bench.measured_traffic(dict(wl, t=wl["t"] + 1), kernel) == (None, None)


def test_tuned_shapes_reach_every_setter():
    calls = {}

    class Plan:
        def __getattr__(self, name):
            return lambda *a: calls.__setitem__(name, a)

    e = {"grid_target": 8192, "narrow": 0, "ell_lds_kb": 80, "ell_chunked": 2,
         "post_grid_target": 4096, "post_narrow": 128, "post_form": 2, "compact_slack": 16,
         "plain_tiles": 4}
    bench.apply_shape_dict(Plan(), e)
    assert calls == {"set_shape": (8192, 0, 80, 2), "set_post_shape": (4096, 128),
                     "set_post_form": (2,), "set_compact_slack": (16,), "set_plain_tiles": (4,)}
