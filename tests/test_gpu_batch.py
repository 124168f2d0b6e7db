"""Batched tessellation of many independent sets (tess_batch) on the GPU.

Per set, results must be bitwise equal to a single-set Diagram of the same
points and box, for every configuration the autotuner may pick (first-pass
capacity, XCD-contiguous dispatch, rerun chunk, learned order), on replays of
the captured graphs, and with capacity reruns in play.
"""
import numpy as np
import pytest

import tess

pytestmark = pytest.mark.gpu


def _same(a, b):
    for k in ("volume", "face_count", "face_offset", "neighbor", "area"):
        assert np.array_equal(a[k], b[k]), k


def _sets(nsets, seed0, nmax=20000):
    out = []
    for i in range(nsets):
        pts, side = tess.batch_set(seed0 + i, 1000.0, float(nmax))
        out.append((pts, np.zeros(3), side))
    return out


def _clustered_sets(sizes, seed0):
    out = []
    for i, n in enumerate(sizes):
        out.append((tess.clustered(n, seed0 + i) * 2.0, np.zeros(3), np.full(3, 2.0)))
    return out


@pytest.fixture(scope="module")
def uneven(ctx):
    sets = _sets(12, 3000)
    ref = []
    for pts, lo, hi in sets:
        d = tess.Diagram(ctx, pts, lo, hi, density=5.0)
        d.compute()
        ref.append(d.cells())
        d.close()
    return sets, ref


def test_batch_matches_single_sets(ctx, uneven):
    sets, ref = uneven
    b = tess.Batch(ctx, sets, density=5.0)
    b.compute()
    for s in range(len(sets)):
        _same(b.cells(s), ref[s])
    t = b.timings()
    assert t["total_ms"] > 0 and t["cells_ms"] > 0


@pytest.mark.parametrize("cfg", [dict(first_capacity=128), dict(first_capacity=256), dict(xcd_chunk=256), dict(xcd_chunk=7),
                                 dict(redo_chunk=1), dict(redo_chunk=16)])
def test_batch_configs_bitwise(ctx, uneven, cfg):
    sets, ref = uneven
    b = tess.Batch(ctx, sets, density=5.0)
    b.set_config(**cfg)
    b.compute()
    b.compute()  # graph replay
    for s in range(len(sets)):
        _same(b.cells(s), ref[s])


def test_batch_reruns_and_learned_order(ctx):
    """Clustered sets overflow the 64-vertex pass; the learned order and the
    rerun pass must not change a bit."""
    sets = _clustered_sets([30000, 5000, 60000], 40)
    ref = []
    reruns = 0
    for pts, lo, hi in sets:
        d = tess.Diagram(ctx, pts, lo, hi, density=5.0, collect_stats=True)
        d.compute()
        ref.append(d.cells())
        reruns += d.timings()["capacity_reruns"]
    assert reruns > 0
    m = tess.Model(ctx, seed=1)
    d = tess.Diagram(ctx, tess.clustered(100000, 41), np.zeros(3), np.ones(3), density=5.0, collect_stats=True)
    d.compute()
    m.add_samples(d)
    m.train(epochs=2, batch=2048, lr=2e-3, seed=2)
    b = tess.Batch(ctx, sets, density=5.0)
    b.compute()
    assert b.timings()["capacity_reruns"] == reruns
    for s in range(len(sets)):
        _same(b.cells(s), ref[s])
    b.set_config(learned=1)
    for _ in range(2):
        b.compute(m)
        for s in range(len(sets)):
            _same(b.cells(s), ref[s])
    with pytest.raises(tess.TessError):
        b.compute(None)  # the learned order needs a model


def test_batch_tune_and_cache(ctx, uneven, tmp_path):
    sets, ref = uneven
    b = tess.Batch(ctx, sets, density=5.0)
    cache = tmp_path / "tune.txt"
    best = b.tune(cache_path=str(cache))
    assert best["first_capacity"] in (64, 128) and best["learned"] == 0
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].split()[0].startswith(("gfx", "unknown"))
    again = b.tune(cache_path=str(cache))  # served from the cache
    assert again == best and len(cache.read_text().splitlines()) == 1
    b.compute()
    for s in range(len(sets)):
        _same(b.cells(s), ref[s])


def test_batch_rejects_bad_input(ctx):
    pts = tess.uniform(100, 1)
    with pytest.raises(tess.TessError):
        tess.Batch(ctx, [(pts, np.ones(3), np.zeros(3))])
    b = tess.Batch(ctx, [(pts, np.zeros(3), np.ones(3))])
    with pytest.raises(tess.TessError):
        b.set_config(first_capacity=100)
    with pytest.raises(tess.TessError):
        b.cells(0)  # before compute


def test_fcc_lattices_learned_batched(ctx):
    """Near-degenerate input (jittered FCC, many near-co-spherical vertices),
    learned order and batch together: bitwise equal to single diagrams."""
    sets = []
    for i in range(3):
        pts = tess.fcc(4000 + i, ncell=16, a=1.0, sigma=1e-3)
        sets.append((pts, np.zeros(3), np.full(3, 16.0)))
    ref = []
    for pts, lo, hi in sets:
        d = tess.Diagram(ctx, pts, lo, hi, density=5.0)
        d.compute()
        ref.append(d.cells())
    m = tess.Model(ctx, seed=3)
    d = tess.Diagram(ctx, sets[0][0], sets[0][1], sets[0][2], density=5.0, collect_stats=True)
    d.compute()
    m.add_samples(d)
    m.train(epochs=2, batch=2048, lr=2e-3, seed=4)
    b = tess.Batch(ctx, sets, density=5.0)
    b.set_config(learned=1)
    b.compute(m)
    for s in range(len(sets)):
        _same(b.cells(s), ref[s])
    assert b.timings()["fallback_cells"] <= b.timings()["capacity_reruns"]
