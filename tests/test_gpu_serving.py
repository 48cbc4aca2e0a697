"""Continuous-batching server: every result is the same alone and in the mix.

Requests of three camera streams (different sizes and schedules, one also
feeding an ensemble) are submitted while earlier batches are on the GPU; the
server batches them per stream into power-of-two buckets, replays captured
HIP graphs (many replays, with eager work and synchronisations between) and
hands results back as batches finish.  Each pair is then run alone through a
fresh eager MultipassPIV: the fields must be bit-identical.
"""

import numpy as np
import pytest
import torch

from pivtools_amd import synthetic as S
from pivtools_amd.ensemble import EnsemblePIV
from pivtools_amd.multipass import MultipassPIV
from pivtools_amd.serving import PIVServer, StreamSpec

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")

SPECS = [
    StreamSpec("cam0", (256, 256), [(64, 64), (32, 32), (16, 16)], [50, 50, 50]),
    StreamSpec("cam1", (192, 320), [(32, 32), (16, 16)], [50, 50], ensemble=True),
    StreamSpec("cam2", (256, 256), [(128, 128), (64, 64), (32, 32), (16, 16)], [50, 50, 50, 50]),
]


def _pair(spec, seed):
    a, b = S.multipass_pair(seed=seed, shape=spec.shape, device=dev)
    return a, b


def _same(x, y):
    return np.array_equal(x.cpu().numpy(), y.cpu().numpy(), equal_nan=True)


@pytest.mark.parametrize("use_graphs", [True, False])
def test_results_alone_equal_results_in_mix(use_graphs):
    rng = np.random.default_rng(5)
    srv = PIVServer(SPECS, max_batch=4, max_slots_per_stream=2, use_graphs=use_graphs)
    reqs = []
    black = None
    for k in range(30):
        i = int(rng.integers(0, 3))
        a, b = _pair(SPECS[i], 7000 + k)
        if k == 11:                         # a black frame pair in the mix
            a.zero_()
            b.zero_()
        r = srv.submit(i, a, b)
        if k == 11:
            black = r
        reqs.append(r)
        if k % 3 == 0:
            srv.poll()                      # earlier batches launched / retired
    print("serving: drain", flush=True)
    srv.drain()
    torch.cuda.synchronize()
    assert len(srv.completed) == 30
    n_graphs = sum(len(s.graphs) for slots in srv.slots for s in slots)
    if use_graphs:
        assert n_graphs >= 3
    assert srv.stats["pairs"] == 30 and srv.stats["padded_pairs"] == 0
    assert srv.stats["peak_bytes"] <= srv.budget
    alone = [MultipassPIV(s.shape, s.window_sizes, s.overlaps, peak_finder=s.peak_finder) for s in SPECS]
    for r in reqs:
        ux, uy, nm = r.result()
        res = alone[r.stream].run(r.a[None], r.b[None])[-1]
        torch.cuda.synchronize()
        assert _same(ux, res.ux[0]) and _same(uy, res.uy[0]), (r.rid, r.stream)
        assert _same(nm, res.nan_mask[0]), (r.rid, r.stream)
    assert bool(black.nan_mask.bool().all())
    assert all(r.latency > 0 for r in reqs)
    # the ensemble stream accumulated exactly its own pairs
    ens = srv.finalize_ensembles()["cam1"]
    mine = [r for r in reqs if r.stream == 1]
    ref = EnsemblePIV(SPECS[1].shape, (32, 32), 75)
    ref.accumulate(torch.stack([r.a for r in mine]), torch.stack([r.b for r in mine]))
    rres = ref.finalize()
    assert ens.n_images == len(mine)
    for p, q in zip(ens.planes, rres.planes):
        p, q = p.cpu().numpy(), q.cpu().numpy()
        assert np.linalg.norm(p - q) <= 1e-6 * np.linalg.norm(q)


def test_graph_replays_match_eager_repeatedly():
    """One slot's graph replayed for several different inputs, with eager
    runs of other inputs in between, against eager results of the same input."""
    spec = SPECS[0]
    srv = PIVServer([spec], max_batch=2, max_slots_per_stream=1)
    eager = MultipassPIV(spec.shape, spec.window_sizes, spec.overlaps)
    for k in range(6):
        a, b = _pair(spec, 8100 + k)
        r = srv.submit(0, a, b)
        srv.drain()
        ea, eb = _pair(spec, 8200 + k)
        eager.run(ea[None], eb[None])        # unrelated eager work between replays
        res = eager.run(a[None], b[None])[-1]
        ux, uy, _ = r.result()
        torch.cuda.synchronize()
        assert _same(ux, res.ux[0]) and _same(uy, res.uy[0])
    assert 1 in srv.slots[0][0].graphs


def test_stream_with_shipped_options_replays_its_graph_exactly():
    """A stream configured like the reference's shipped settings (peak
    magnitude and median tests, biharmonic infill, a pixel mask, several
    peaks with substitution) served through captured graphs equals eager runs
    of the same options, pair for pair, over several replays."""
    shape = (256, 256)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    opts = {"outlier_methods": [{"type": "peak_mag", "threshold": 0.25},
                                {"type": "median_2d", "epsilon": 0.1, "threshold": 2}],
            "mid_infill": "biharmonic", "final_infill": "biharmonic",
            "pixel_mask": (yy - 120) ** 2 + (xx - 100) ** 2 < 30 ** 2,
            "num_peaks": 2, "secondary_peak": True}
    spec = StreamSpec("shipped", shape, [(64, 64), (32, 32), (16, 16)], [50, 50, 50], piv_options=opts)
    srv = PIVServer([spec], max_batch=2, max_slots_per_stream=1)
    eager = MultipassPIV(shape, spec.window_sizes, spec.overlaps, **opts)
    for k in range(5):
        a, b = _pair(spec, 8300 + k)
        r = srv.submit(0, a, b)
        srv.drain()
        res = eager.run(a[None], b[None])[-1]
        ux, uy, nm = r.result()
        torch.cuda.synchronize()
        assert _same(ux, res.ux[0]) and _same(uy, res.uy[0]) and _same(nm, res.nan_mask[0]), k
    eager.check_infill()
    assert 1 in srv.slots[0][0].graphs


def _poisson_serve(srv, plan, rate, seed):
    """Submit (stream, pair) in plan order at Poisson arrival times of the
    given mean rate, polling the server in between."""
    import time
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(plan))
    reqs = []
    t_next = time.perf_counter()
    for (i, a, b), gap in zip(plan, gaps):
        t_next += gap
        while time.perf_counter() < t_next:
            srv.poll()
        reqs.append(srv.submit(i, a, b))
        srv.poll()
    srv.drain()
    torch.cuda.synchronize()
    return reqs


def test_full_size_streams_alone_equal_mix_under_poisson_arrivals():
    """The serving workload's stream sizes: a 2160 x 2560 stream with the
    4-pass 128 -> 16 schedule and a 2048 x 2048 stream with 64 -> 16, pairs
    submitted at Poisson arrival times (mean 400 pairs/s) while earlier
    batches run.  Every result is bit-identical to the pair run alone."""
    specs = [StreamSpec("cam_2560", (2160, 2560), [(128, 128), (64, 64), (32, 32), (16, 16)], [50] * 4),
             StreamSpec("cam_2048", (2048, 2048), [(64, 64), (32, 32), (16, 16)], [50] * 3)]
    srv = PIVServer(specs, max_batch=4, max_slots_per_stream=2)
    plan = []
    for k in range(10):
        i = k % 2
        a, b = S.multipass_pair(seed=9400 + k, shape=specs[i].shape, device=dev)
        plan.append((i, a, b))
    print("serving: 10 full-size pairs", flush=True)
    reqs = _poisson_serve(srv, plan, 400.0, 3)
    assert len(srv.completed) == 10 and srv.stats["padded_pairs"] == 0
    alone = [MultipassPIV(s.shape, s.window_sizes, s.overlaps) for s in specs]
    for r in reqs:
        ux, uy, nm = r.result()
        res = alone[r.stream].run(r.a[None], r.b[None])[-1]
        torch.cuda.synchronize()
        assert _same(ux, res.ux[0]) and _same(uy, res.uy[0]) and _same(nm, res.nan_mask[0]), (r.rid, r.stream)
        assert r.latency > 0
    assert srv.stats["peak_bytes"] <= srv.budget


def test_hbm_budget_admission_and_eviction():
    """A slot is charged its staging plus pass-loop buffers at max_batch;
    a budget below one slot is refused up front; with room for one slot at a
    time, a stream that waits with no slot gets the idle slot of the other
    stream released, and both streams' results stay exact."""
    specs = [StreamSpec("a", (256, 256), [(64, 64), (32, 32), (16, 16)], [50] * 3),
             StreamSpec("b", (192, 320), [(32, 32), (16, 16)], [50] * 2)]
    probe = PIVServer(specs, max_batch=2, max_slots_per_stream=1)
    sizes = [probe._slot_bytes(i, probe._slot_piv(i)) for i in range(2)]
    mp = probe._slot_piv(0)
    assert sizes[0] == 2 * 2 * 256 * 256 * 2 + mp.footprint_bytes(2)
    with pytest.raises(ValueError):
        PIVServer(specs, max_batch=2, hbm_budget_bytes=min(sizes) - 1)
    srv = PIVServer(specs, max_batch=2, max_slots_per_stream=2, hbm_budget_bytes=max(sizes) + min(sizes) // 2)
    alone = [MultipassPIV(s.shape, s.window_sizes, s.overlaps) for s in specs]
    for rnd in range(3):
        for i in (0, 1):
            reqs = [srv.submit(i, *_pair(specs[i], 9600 + 10 * rnd + 3 * i + k)) for k in range(3)]
            srv.drain()
            assert len(srv.slots[i]) == 1 and len(srv.slots[1 - i]) == 0
            assert srv.bytes_in_use <= srv.budget
            for r in reqs:
                ux, uy, nm = r.result()
                res = alone[i].run(r.a[None], r.b[None])[-1]
                torch.cuda.synchronize()
                assert _same(ux, res.ux[0]) and _same(nm, res.nan_mask[0])
    assert srv.stats["evicted_slots"] >= 5
    assert srv.stats["peak_bytes"] <= srv.budget
