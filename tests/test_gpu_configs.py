"""The benchmark workloads at full size on the GPU (Cell::compute_voronoi_cell
over every point, interface.rs:257-313):

* 4M clustered points (80% in 64 Gaussian clusters, sigma 0.02; seed 25),
  plain and learned-radius paths;
* 512 uneven sets (N log-uniform in [1e3, 2e5], box sides U[0.5, 2]; seeds
  3000 + i) in one batch;
* 16 jittered FCC lattices of 63^3 cells (1,000,188 points, cubic constant
  a = 1, jitter sigma 1e-3; seeds 4000 + i), learned order + batch.

Every cell: the volumes sum to the box volume within 1e-10 (relative), the
neighbour relation is symmetric and the two sides of a shared face have the
same area to rounding (|a_ij - a_ji| <= 1e-10 v^(2/3), v the mean cell
volume).  Against the oracle, bitwise (volumes, areas, neighbours and the
search counters): seeded samples of cells, including the heaviest ones.
Learned and batched paths are bitwise equal to the plain kernel."""
import numpy as np
import pytest

import tess
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
LO, HI = np.zeros(3), np.ones(3)
DENSITY = 5.0
COUNTERS = ("streamed", "classified", "cuts", "entries", "flops", "incident_cuts")


def log(msg):
    print(msg, flush=True)


def check_invariants(cells, n, box_volume):
    fc = cells["face_count"].astype(np.int64)
    vol = cells["volume"]
    assert np.all(fc >= 4) and np.all(vol > 0)
    assert abs(vol.sum() - box_volume) <= 1e-10 * box_volume, (vol.sum(), box_volume)
    owner = np.repeat(np.arange(n, dtype=np.int64), fc)
    nb = cells["neighbor"].astype(np.int64)
    inner = nb >= 0
    a, b, ar = owner[inner], nb[inner], cells["area"][inner]
    k1, k2 = a * n + b, b * n + a
    o1, o2 = np.argsort(k1, kind="stable"), np.argsort(k2, kind="stable")
    assert np.array_equal(k1[o1], k2[o2]), "neighbour relation not symmetric"
    scale = (box_volume / n) ** (2.0 / 3.0)
    d = np.abs(ar[o1] - ar[o2])
    assert d.max(initial=0.0) <= 1e-10 * scale, d.max()
    # every cell sits in its own neighbour lists' cells and is not its own neighbour
    assert not np.any(a == b)


def pick(cells, idx):
    """The listed cells' outputs, in list order, as one CSR."""
    fo = cells["face_offset"].astype(np.int64)
    parts = [np.arange(fo[i], fo[i + 1]) for i in idx]
    sel = np.concatenate(parts) if parts else np.zeros(0, np.int64)
    return dict(volume=cells["volume"][idx], face_count=cells["face_count"][idx], neighbor=cells["neighbor"][sel],
                area=cells["area"][sel])


def check_sample(pts, lo, hi, cells, stats, idx):
    ref = orc.compute(pts, lo, hi, density=DENSITY, cells=idx, threads=16)
    got = pick(cells, idx)
    for k in ("face_count", "neighbor", "volume", "area"):
        assert np.array_equal(got[k], ref[k]), k
    if stats is not None:
        for k in COUNTERS:
            assert np.array_equal(stats[k][idx].astype(np.uint64), ref["stats"][k].astype(np.uint64)), k


@pytest.fixture(scope="module")
def cfg2(ctx):
    pts = tess.clustered(4_000_000, 25)
    d = tess.Diagram(ctx, pts, LO, HI, density=DENSITY, collect_stats=True)
    d.compute()
    out = (pts, d, d.cells(), d.stats())
    log(f"clustered 4M: {d.timings()['total_ms']:.1f} ms")
    yield out
    d.close()


def test_clustered_4m_invariants(cfg2):
    pts, _, cells, _ = cfg2
    check_invariants(cells, len(pts), 1.0)


def test_clustered_4m_oracle_sample(cfg2):
    pts, _, cells, st = cfg2
    rng = np.random.default_rng(2025)
    heavy = np.argsort(st["classified"], kind="stable")[-500:]
    idx = np.unique(np.concatenate([rng.choice(len(pts), 3000, replace=False), heavy])).astype(np.uint32)
    log(f"oracle sample: {len(idx)} cells, heaviest classified {int(st['classified'].max())}")
    check_sample(pts, LO, HI, cells, st, idx)


def test_clustered_4m_learned_equals_plain(ctx, cfg2):
    """The model trained on five other seeded sets (20-24) by the GPU kernels;
    the learned path must not change a bit."""
    pts, _, cells, _ = cfg2
    m = tess.Model(ctx, seed=7)
    for s in (20, 21, 22, 23, 24):
        t = tess.Diagram(ctx, tess.clustered(4_000_000, s), LO, HI, density=DENSITY, collect_stats=True)
        t.compute()
        m.add_samples(t)
        t.close()
    loss = m.train(epochs=3, batch=8192, lr=2e-3, seed=1)
    log(f"model loss {np.round(loss, 4)}")
    assert loss[-1] < loss[0]
    d = tess.Diagram(ctx, pts, LO, HI, density=DENSITY)
    d.compute_learned(m)
    got = d.cells()
    for k in cells:
        assert np.array_equal(got[k], cells[k]), k
    t = d.timings()
    # what the trained network gains over the method without it, on this set
    plain = tess.Diagram(ctx, pts, LO, HI, density=DENSITY)
    plain.compute()
    pm = min(plain.compute() or plain.timings()["total_ms"] for _ in range(3))
    lm = min(d.compute_learned(m) or d.timings()["total_ms"] for _ in range(3))
    rate = t["fallback_cells"] / len(pts)
    log(f"plain {len(pts) / pm * 1e3 / 1e6:.2f}M cells/s, learned {len(pts) / lm * 1e3 / 1e6:.2f}M cells/s "
        f"(gain {pm / lm:.3f}x), fallback rate {rate:.5f}, routed long {t['planned_long']}")
    assert 0.0 <= rate < 0.05
    plain.close()
    b = tess.Batch(ctx, [(pts, LO, HI)], density=DENSITY)
    b.set_config(learned=1)
    b.compute(m)
    got = b.cells(0)
    for k in cells:
        assert np.array_equal(got[k], cells[k]), k
    d.close()
    b.close()
    m.close()


def test_batch_512_uneven_sets(ctx):
    sets = []
    for i in range(512):
        p, side = tess.batch_set(3000 + i)
        sets.append((p, np.zeros(3), side))
    n = np.array([len(p) for p, _, _ in sets])
    log(f"batch: 512 sets, {n.sum()} points")
    b = tess.Batch(ctx, sets, density=DENSITY)
    b.compute()
    log(f"batch compute: {b.timings()['total_ms']:.1f} ms")
    sides = np.array([s for _, _, s in sets])
    aniso = sides.max(axis=1) / sides.min(axis=1)
    check = set(np.argsort(n)[:8]) | set(np.argsort(n)[-8:]) | set(np.argsort(aniso)[-8:])
    rng = np.random.default_rng(3)
    for s, (p, lo, hi) in enumerate(sets):
        cells = b.cells(s)
        check_invariants(cells, len(p), float(np.prod(hi - lo)))
        if s in check:
            idx = np.unique(rng.choice(len(p), min(len(p), 300), replace=False)).astype(np.uint32)
            check_sample(p, lo, hi, cells, None, idx)
        if s % 64 == 0:
            log(f"set {s}: {len(p)} points checked")


def test_fcc_16_lattices_learned_batched(ctx):
    """cfg4: 16 perturbed FCC lattices (spacing 1 = the cubic lattice constant
    a, nearest-neighbour distance a / sqrt 2), learned order in one batch,
    against the plain batch and the oracle; the fallback rate is reported."""
    lattices = [tess.fcc(4000 + i, 63) for i in range(16)]
    assert all(len(p) == 1_000_188 for p in lattices)
    box = np.full(3, 63.0)
    m = tess.Model(ctx, seed=3)
    for s in (4100, 4101):
        t = tess.Diagram(ctx, tess.fcc(s, 63), np.zeros(3), box, density=DENSITY, collect_stats=True)
        t.compute()
        m.add_samples(t)
        t.close()
    m.train(epochs=2, batch=8192, lr=2e-3, seed=1)
    sets = [(p, np.zeros(3), box) for p in lattices]
    plain = tess.Batch(ctx, sets, density=DENSITY)
    plain.compute()
    learned = tess.Batch(ctx, sets, density=DENSITY)
    learned.set_config(learned=1)
    learned.compute(m)
    t = learned.timings()
    rate = t["fallback_cells"] / (16 * 1_000_188)
    log(f"fcc 16 x 1M: plain {plain.timings()['total_ms']:.1f} ms, learned {t['total_ms']:.1f} ms, "
        f"fallback rate {rate:.6f}")
    assert 0.0 <= rate <= 1.0
    rng = np.random.default_rng(4)
    for s, (p, lo, hi) in enumerate(sets):
        got, ref = learned.cells(s), plain.cells(s)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (s, k)
        check_invariants(got, len(p), 63.0 ** 3)
        idx = np.unique(rng.choice(len(p), 200, replace=False)).astype(np.uint32)
        check_sample(p, lo, hi, got, None, idx)
    plain.close()
    learned.close()
    m.close()
