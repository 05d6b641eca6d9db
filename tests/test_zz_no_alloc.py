"""GPU test of snicit_infer_noalloc (every device buffer from the library's static scratch)
against snicit_infer, through the C-ABI: two calls in a row, the second smaller, then a run
too large for the scratch, then a normal call again."""
import numpy as np
import pytest

import oracle as orc
import snicit_amd as sa
from tests import nets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["SNICIT", "BF"])
def test_noalloc_equals_infer(ctx, mode):
    N, L = 1024, 40
    cols, vals = nets.sdgc(N, L, seed=7)
    net = sa.Net.from_ell(ctx, cols, vals, -0.3)
    Method = sa.SNICIT if mode == "SNICIT" else sa.BF
    m = Method(net, threshold=8) if mode == "SNICIT" else Method(net)
    for B, batch, seed in ((1800, 900, 11), (300, 300, 12)):
        y0 = nets.images(N, B, seed=seed)
        ref = m.infer(y0, batch=batch)
        got = m.infer(y0, batch=batch, noalloc=True)
        np.testing.assert_array_equal(got, ref)
        _, cat, _ = orc.plain_f32(cols, vals, -0.3, 32.0, y0, want_y=False)
        np.testing.assert_array_equal(got, cat)
        assert m.stats.batch == B and m.stats.ms_total == 0.0


def test_noalloc_lossy_model(ctx):
    """A net with dense input and output layers and lossy clustering: equal to snicit_infer."""
    Nin, N, L, K, Nout, B = 784, 256, 6, 40, 10, 400
    c, v, w_in, w_out = sa.gen_unstructured(Nin, N, L, K, Nout, 3)
    rng = np.random.default_rng(3)
    net = sa.Net.from_ell(ctx, c, v, 0.0, clip=1.0)
    net.set_io(w_in, (rng.standard_normal(N) * 0.05).astype(np.float32), 1.0, w_out,
               (rng.standard_normal(Nout) * 0.1).astype(np.float32))
    x = nets.prototype_features(Nin, B, 5)
    m = sa.SNICIT(net, threshold=2, cluster=sa.CLUSTER_LOSSY)
    np.testing.assert_array_equal(m.infer(x, batch=200, noalloc=True), m.infer(x, batch=200))


def test_noalloc_refuses_runs_beyond_the_scratch(ctx):
    N, L = 4096, 4
    net = sa.Net.generate_sdgc(ctx, N, L, 1, bias=-0.35)
    y0 = nets.images(N, 4000, seed=3)
    m = sa.SNICIT(net, threshold=2)
    with pytest.raises(sa.SnicitError, match="static scratch"):
        m.infer(y0, batch=4000, noalloc=True)
    small = y0[:256]
    np.testing.assert_array_equal(m.infer(small, noalloc=True), m.infer(small))
