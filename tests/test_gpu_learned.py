"""Learned search-radius model on the GPU, through the C ABI.

- the 16x16x32 bf16 MFMA lane layout the kernels assume (exact products);
- occupancy features bit-exact against the CPU reference (oracle/mlp.py);
- forward pass and training gradient against the CPU reference with the same
  bf16 rounding points (fp32 accumulation order differs: tolerance stated);
- GPU training lowers the loss and the predictions track the plain search's
  own counters;
- the learned schedule and any schedule forced from the host (adversarial
  predictions) give results bitwise equal to the plain path; fallbacks are
  flagged and counted.
"""
import numpy as np
import pytest

import tess
from oracle import mlp

pytestmark = pytest.mark.gpu
LO, HI = np.zeros(3), np.ones(3)


def test_mfma_lane_layout(ctx):
    rng = np.random.default_rng(0)
    a = rng.integers(-8, 9, (16, 32)).astype(np.float32)
    b = rng.integers(-8, 9, (32, 16)).astype(np.float32)
    d = ctx.mfma_probe(mlp.bf16_bits(a), mlp.bf16_bits(b))
    assert np.array_equal(d, (a.astype(np.float64) @ b).astype(np.float32))


def test_features_bit_exact(ctx):
    pts = tess.clustered(20000, 11)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    d.compute()
    f = d.features()
    cell, _, delim = d.copy_grid()
    _, _, cpd = d.grid()
    assert np.array_equal(f, mlp.features(cell, delim, cpd))


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, mlp.IN), np.float32)
    x[:, :mlp.FEATURES] = np.log1p(rng.poisson(rng.uniform(1, 60, (n, 1)), (n, mlp.FEATURES)))
    t = rng.normal(0, 1, (n, 2)).astype(np.float32)
    return mlp.bf16_bits(x), t


def test_forward_matches_cpu(ctx):
    m = tess.Model(ctx, seed=3)
    p = mlp.init_params(3)
    m.set(p)
    x, _ = _batch(1000, 4)  # not a multiple of 16: ragged last tile
    y = m.forward(x)
    ref = mlp.forward(p, x, rounded=True)
    err = np.abs(y - ref)
    assert err.max() <= 2e-2 * (1 + np.abs(ref).max()), err.max()
    assert err.mean() <= 1e-3


@pytest.mark.parametrize("n", [32, 200, 4096])
def test_gradient_matches_cpu(ctx, n):
    m = tess.Model(ctx, seed=5)
    p = mlp.init_params(5)
    m.set(p)
    x, t = _batch(n, 6 + n)
    loss, g = m.gradient(x, t)
    ref_loss, ref = mlp.gradient(p, x, t, rounded=True)
    assert loss == pytest.approx(ref_loss, rel=1e-3)
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    assert rel < 5e-3, rel
    # blocks separately: a transposed operand in one layer would fail here
    o = 0
    for size in (mlp.N_W1, mlp.N_W2, mlp.N_W3, mlp.H, mlp.H, mlp.OUT_PAD):
        gb, rb = g[o:o + size], ref[o:o + size]
        assert np.linalg.norm(gb - rb) <= 1e-2 * np.linalg.norm(rb) + 1e-7
        o += size
    # pads carry no gradient
    W1g = g[:mlp.N_W1].reshape(mlp.IN, mlp.H)
    W3g = g[mlp.N_W1 + mlp.N_W2:mlp.WEIGHTS].reshape(mlp.H, mlp.OUT_PAD)
    assert np.all(W1g[mlp.FEATURES:] == 0) and np.all(W3g[:, mlp.OUT:] == 0)
    assert np.all(g[-mlp.OUT_PAD + mlp.OUT:] == 0)


@pytest.fixture(scope="module")
def trained(ctx):
    """A model trained on the plain search's own counters of two clustered sets."""
    m = tess.Model(ctx, seed=7)
    for seed in (20, 21):
        d = tess.Diagram(ctx, tess.clustered(200000, seed), LO, HI, density=5.0, collect_stats=True)
        d.compute()
        m.add_samples(d)
        d.close()
    assert m.sample_count() == 400000
    loss = m.train(epochs=4, batch=4096, lr=2e-3, seed=1)
    return m, loss


def test_training_lowers_loss(trained):
    _, loss = trained
    assert np.all(np.isfinite(loss))
    assert loss[-1] < loss[0], loss
    # standardised targets: predicting the mean gives 0.5 * (1 + 1) = 1.0
    assert loss[-1] < 0.5, loss


def test_predictions_track_counters(ctx, trained):
    m, _ = trained
    d = tess.Diagram(ctx, tess.clustered(200000, 25), LO, HI, density=5.0, collect_stats=True)
    d.compute()
    st = d.stats()
    pred = d.predict(m)
    r = np.sqrt(st["r2"])
    c = st["classified"].astype(np.float64)
    cr = np.corrcoef(np.log(pred[:, 0]), np.log(r))[0, 1]
    cc = np.corrcoef(np.log1p(pred[:, 1]), np.log1p(c))[0, 1]
    assert cr > 0.8 and cc > 0.6, (cr, cc)


def _same(a, b):
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(a[k], b[k]), k


def test_learned_schedule_bitwise_equal_to_plain(ctx, trained):
    m, _ = trained
    pts = tess.clustered(200000, 26)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0)
    d.compute()
    plain = d.cells()
    tp = d.timings()
    d.compute_learned(m)
    learned = d.cells()
    tl = d.timings()
    _same(learned, plain)
    assert tl["fallback_cells"] <= tp["capacity_reruns"] + 1000
    assert tl["plan_ms"] > 0


@pytest.mark.parametrize("kind", ["zeros", "reversed", "random", "nan", "all_large"])
def test_forced_schedules_bitwise_equal(ctx, kind):
    """Any prediction, however wrong, only reorders and re-sizes work."""
    pts = tess.clustered(60000, 27)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, collect_stats=True)
    d.compute()
    plain = d.cells()
    st = d.stats()
    c = np.log1p(st["classified"].astype(np.float64)).astype(np.float32)
    rng = np.random.default_rng(9)
    v = {"zeros": np.zeros_like(c), "reversed": -c, "random": rng.uniform(-5, 20, c.size).astype(np.float32),
         "nan": np.full_like(c, np.nan), "all_large": np.full_like(c, 30.0)}[kind]
    d.compute_scheduled(v, t1=4.0, t2=6.0)
    _same(d.cells(), plain)
    t = d.timings()
    status = d.stats()["status"]
    nfb = int(np.count_nonzero(status & tess.CELL_FALLBACK))
    assert nfb == t["fallback_cells"]
    if kind == "all_large":  # every cell starts at capacity 256: nothing falls back
        assert nfb == 0


def test_fallback_cells_flagged(ctx):
    """A schedule that starts every cell at 64 vertices falls back exactly
    where the plain pass reran: cells that outgrew 64 vertices (or ran out
    of free slots during a cut)."""
    pts = tess.clustered(60000, 28)
    d = tess.Diagram(ctx, pts, LO, HI, density=5.0, collect_stats=True)
    d.compute()
    plain = d.cells()
    reruns = d.timings()["capacity_reruns"]
    st0 = d.stats()
    over = st0["max_vertices"] > 64
    assert over.sum() > 0
    rerun = (st0["status"] & tess.CELL_RERUN) != 0  # plain mode flags its reruns too
    assert int(rerun.sum()) == reruns and np.all(rerun[over])
    assert not np.any(st0["status"] & tess.CELL_FALLBACK)
    d.compute_scheduled(np.zeros(len(pts), np.float32), t1=np.inf, t2=np.inf)
    _same(d.cells(), plain)
    flagged = (d.stats()["status"] & tess.CELL_FALLBACK) != 0
    assert np.all(flagged[over]) and np.array_equal(flagged, rerun)
    assert int(flagged.sum()) == d.timings()["fallback_cells"] == reruns


def test_model_save_load(ctx, trained, tmp_path):
    m, _ = trained
    p, norm, cal = m.get()
    path = tmp_path / "model.bin"
    m.save(str(path))
    m2 = tess.Model(ctx, seed=99)
    m2.load(str(path))
    p2, norm2, cal2 = m2.get()
    assert np.array_equal(p, p2) and np.array_equal(norm, norm2) and np.array_equal(cal, cal2)
    with pytest.raises(tess.TessError):
        m2.load(str(tmp_path / "missing.bin"))
