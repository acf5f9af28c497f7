"""GPU training parity: device gather + search + per-cluster sums driven by
the host recipe, against the CPU oracle's train (src/lib.rs:348-491)."""
import numpy as np
import pytest

import emeans_hip as E
import oracle as O

pytestmark = pytest.mark.gpu


def _i8_blobs(n, d, blobs, seed):
    rng = np.random.default_rng(seed)
    centres = rng.integers(-90, 90, size=(blobs, d))
    x = centres[rng.integers(0, blobs, n)] + rng.integers(-12, 13, size=(n, d))
    return np.clip(x, -128, 127).astype(np.int8)


@pytest.mark.parametrize("bits", ["i8", "i16"])
@pytest.mark.parametrize("balance", [0.0, 0.1])
def test_i8_training_is_bit_exact(hip, balance, bits):
    """int8 rows with the exact integer search (CentroidBits::I8 or I16):
    every label, distance and per-cluster sum is exact, so the whole
    trajectory (centroids, last-pass labels, objective, revivals) must equal
    the CPU oracle's."""
    print(f"{bits} training balance={balance}", flush=True)
    x = _i8_blobs(6000, 48, 60, 1)
    cfg_o = O.Config(nepochs=3, balance=balance, seed=99)
    cfg_e = E.Config(nepochs=3, balance=balance, seed=99)
    ref = O.train(x, 40, cfg_o, assigner=O.ASSIGNER_I8 if bits == "i8" else O.ASSIGNER_I16)
    got = E.train(x, 40, cfg_e, E.HipAssigner(E.I8 if bits == "i8" else E.I16))
    assert got.recipe.batch_size == ref.recipe.batch_size
    assert np.array_equal(got.labels, ref.labels)
    assert got.objective == ref.objective
    assert got.n_splits == ref.n_splits
    assert np.array_equal(got.centroids, ref.centroids)


def test_i8_training_past_63_epochs(hip):
    """100 epochs (past the old 64-entry tables) of int8 training equal the
    oracle's bit for bit, with one objective entry per epoch."""
    x = _i8_blobs(1500, 16, 12, 5)
    ref = O.train(x, 10, O.Config(nepochs=100.0, seed=3), assigner=O.ASSIGNER_I8)
    got = E.train(x, 10, E.Config(nepochs=100.0, seed=3), E.HipAssigner(E.I8))
    assert len(got.objective) == 100 and got.objective == ref.objective
    assert np.array_equal(got.centroids, ref.centroids) and np.array_equal(got.labels, ref.labels)


def _blobs_6d():
    n, d = 20_000, 6
    r = O.OracleRng(7)
    u = r.open01_many(20 * d)
    centers = (np.float32(8.0) * (u.astype(np.float32) - np.float32(0.5))).astype(np.float32)
    noise_u = r.open01_many(n * d * 4).astype(np.float32).reshape(n, d, 4) - np.float32(0.5)
    noise = np.zeros((n, d), np.float32)
    for t in range(4):
        noise = (noise + noise_u[:, :, t]).astype(np.float32)
    return (centers.reshape(20, d)[np.arange(n) % 20] + noise).astype(np.float32)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_closing_mean_and_loss(hip, prec):
    # src/lib.rs:575-614 on the GPU assigner
    x = _blobs_6d()
    n, d, k = x.shape[0], x.shape[1], 32
    out = E.train(x, k, E.Config(), E.HipAssigner(prec))
    counts = np.bincount(out.labels, minlength=k)
    assert counts.sum() == n
    sums = np.zeros((k, d))
    np.add.at(sums, out.labels, x.astype(np.float64))
    live = counts > 0
    assert np.all(np.abs(out.centroids[live] - sums[live] / counts[live, None]) < 1e-3)
    trained = E.assign(x, out.centroids)[1].astype(np.float64).mean()
    start = E.assign(x, x[O.permutation(1234, n)[:k]])[1].astype(np.float64).mean()
    assert trained < 0.75 * start and trained < 2.0
    assert all(b < a for a, b in zip(out.objective, out.objective[1:]))


@pytest.mark.parametrize("prec,m", [("bf16", 4), ("fp8", 8)])
def test_rescored_training_matches_fp32(hip, prec, m):
    """Training on the candidate search (top-m + FP32 rescore) ends within
    0.5% of the FP32 search's exact re-assignment MSE, and its closing mean
    is exact on its own labels."""
    rng = np.random.default_rng(4)
    n, d, k = 60_000, 64, 1024
    centres = rng.normal(scale=4, size=(2048, d)).astype(np.float32)
    x = (centres[rng.integers(0, 2048, n)] + rng.normal(size=(n, d))).astype(np.float32)
    ref = E.train(x, k, E.Config(nepochs=2), E.HipAssigner(E.FP32))
    got = E.train(x, k, E.Config(nepochs=2), E.HipAssigner(prec, rescore_m=m))
    mse_ref = E.assign(x, ref.centroids)[1].astype(np.float64).mean()
    mse_got = E.assign(x, got.centroids)[1].astype(np.float64).mean()
    assert abs(mse_got / mse_ref - 1) < 5e-3, (mse_got, mse_ref)
    counts = np.bincount(got.labels, minlength=k)
    sums = np.zeros((k, d))
    np.add.at(sums, got.labels, x.astype(np.float64))
    live = counts > 0
    assert np.all(np.abs(got.centroids[live] - sums[live] / counts[live, None]) < 1e-3 * (1 + np.abs(got.centroids[live])))


def test_f32_training_tracks_oracle(hip):
    """f32 rows: near-tie labels may differ from the CPU's (FP32 sums in a
    different order), so the trajectories are compared statistically."""
    rng = np.random.default_rng(2)
    centres = rng.normal(scale=4, size=(64, 24)).astype(np.float32)
    x = (centres[rng.integers(0, 64, 8000)] + rng.normal(size=(8000, 24))).astype(np.float32)
    ref = O.train(x, 50, O.Config(nepochs=2))
    got = E.train(x, 50, E.Config(nepochs=2))
    mse_ref = O.assign(x, ref.centroids)[1].astype(np.float64).mean()
    mse_got = O.assign(x, got.centroids)[1].astype(np.float64).mean()
    assert abs(mse_got / mse_ref - 1) < 0.02
    assert abs(got.objective[-1] / ref.objective[-1] - 1) < 0.02


@pytest.mark.parametrize("n,m,k,d", [(5000, 3000, 50, 37), (20000, 7001, 300, 37), (70000, 66000, 70000, 37),
                                     (9000, 1, 1, 37), (20000, 15000, 7, 37), (6000, 5000, 40, 300),
                                     (20000, 17001, 2048, 37), (9000, 4097, 257, 96), (70000, 66001, 3000, 37),
                                     (70000, 66001, 1500, 37)])
def test_device_accumulate_is_batch_order_f32(hip, n, m, k, d):
    """The stable sort + streamed fold equals the sequential host loop bit
    for bit: one, two and three 8-bit sort passes, one 11-bit pass (with and
    without the scan launch) and one 12-bit pass, blocks with a ragged end,
    a single row, empty clusters, clusters of thousands of rows (folded one
    wave per 64 dimensions) and rows wider than 256 (streamed 256 dimensions
    at a time)."""
    import torch

    print(f"accumulate n={n} m={m} k={k} d={d}", flush=True)
    rng = np.random.default_rng(3)
    x = rng.normal(scale=100, size=(n, d)).astype(np.float32)
    ids = rng.permutation(n)[:m].astype(np.uint32)
    labels = rng.integers(0, k, m).astype(np.uint32)
    labels[labels == 7] = 8  # an empty cluster (when k > 8)
    a = E.HipAssigner(E.FP32)
    a.set_centroids(np.zeros((k, d), np.float32))
    xd = torch.from_numpy(x).cuda()
    idd = torch.from_numpy(ids.view(np.int32)).cuda()
    ld = torch.from_numpy(labels.view(np.int32)).cuda()
    sums = torch.full((k, d), 7.0, device="cuda")
    counts = torch.full((k,), 9, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    a.accumulate_device(xd, ld, sums, counts, ids=idd)
    a.synchronize()
    # the host loop of src/lib.rs:431-441: sequential f32 adds in batch order
    want = np.zeros((k, d), np.float32)
    want_c = np.zeros(k, np.int64)
    for r in range(m):
        want[labels[r]] += x[ids[r]]
        want_c[labels[r]] += 1
    assert np.array_equal(counts.cpu().numpy(), want_c)
    assert np.array_equal(sums.cpu().numpy(), want)


@pytest.mark.parametrize("n,m,k,d,dt", [(20000, 15000, 4096, 512, "i8"), (9000, 8000, 3000, 16, "u8"),
                                        (9000, 8000, 2500, 1024, "i8"), (9000, 8000, 2500, 96, "i8"),
                                        (9000, 8000, 2500, 37, "u8"), (9000, 8000, 300, 64, "i8"),
                                        (12000, 11000, 2600, 256, "u8")])
def test_device_accumulate_is_batch_order_int(hip, n, m, k, d, dt):
    """Integer rows: the same bit-identity with the sequential f32 host loop
    (src/lib.rs:431-441), over the whole-row integer fold (d = 16 x 2^L),
    the 256-dimension stream (other d), the per-cluster fold (k <= 2048),
    full-range values, a cluster of 600 members (folded as a big one) and
    an empty cluster."""
    import torch

    print(f"accumulate int n={n} m={m} k={k} d={d} {dt}", flush=True)
    rng = np.random.default_rng(4)
    if dt == "i8":
        x = rng.integers(-128, 128, size=(n, d)).astype(np.int8)
    else:
        x = rng.integers(0, 256, size=(n, d)).astype(np.uint8)
    ids = rng.permutation(n)[:m].astype(np.uint32)
    labels = rng.integers(0, k, m).astype(np.uint32)
    labels[labels == 7] = 8  # an empty cluster
    labels[rng.permutation(m)[:600]] = 3  # a big cluster, members spread over the batch
    a = E.HipAssigner(E.FP32)
    a.set_centroids(np.zeros((k, d), np.float32))
    xd = torch.from_numpy(x).cuda()
    idd = torch.from_numpy(ids.view(np.int32)).cuda()
    ld = torch.from_numpy(labels.view(np.int32)).cuda()
    sums = torch.full((k, d), 7.0, device="cuda")
    counts = torch.full((k,), 9, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    a.accumulate_device(xd, ld, sums, counts, ids=idd)
    a.synchronize()
    want = np.zeros((k, d), np.float32)
    want_c = np.zeros(k, np.int64)
    xf = x.astype(np.float32)
    for r in range(m):
        want[labels[r]] += xf[ids[r]]
        want_c[labels[r]] += 1
    assert np.array_equal(counts.cpu().numpy(), want_c)
    assert np.array_equal(sums.cpu().numpy(), want)


def _job_data(seed, n, d, blobs):
    rng = np.random.default_rng(seed)
    centres = rng.normal(scale=4, size=(blobs, d)).astype(np.float32)
    return (centres[rng.integers(0, blobs, n)] + rng.normal(size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_batched_training_equals_single(hip, prec):
    """Every job of a lockstep batch (one graph replay per step, uneven n, d,
    k and step counts, balance on some) ends with exactly the result of its
    own single-problem training on the same assigner."""
    specs = [(3000, 24, 40, 0.0), (9000, 64, 200, 0.1), (700, 96, 64, 0.0), (20000, 128, 300, 0.1),
             (5000, 33, 17, 0.0)]
    if prec in ("bf16", "fp8"):
        specs.append((4000, 256, 90, 0.1))
    problems = []
    for i, (n, d, k, bal) in enumerate(specs):
        problems.append((_job_data(10 + i, n, d, 2 * k), k, E.Config(nepochs=2 + (i % 2), balance=bal, seed=7 + i)))
    print(f"batch {prec}: {len(problems)} jobs", flush=True)
    got, st = E.train_batch(problems, E.HipAssigner(prec))
    assert st.steps == max(E.resolve_recipe(p[0].shape[0], p[0].shape[1], p[1], p[2]).total_steps for p in problems)
    for (x, k, cfg), g in zip(problems, got):
        ref = E.train(x, k, cfg, E.HipAssigner(prec))
        assert np.array_equal(g.labels, ref.labels)
        assert np.array_equal(g.centroids, ref.centroids)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits


def test_batched_fp8_mixed_magnitudes(hip):
    """FP8 jobs whose clusters sit at scales 1e3 and 1e-3 side by side: the
    batched prep sets each centroid's own e4m3 scale as the single-problem
    prep does (each job ends with exactly its single-problem result), and
    the small clusters stay resolved: the final centroids' FP8 top-8 search
    labels every small-cluster row as FP64 brute force does."""
    problems = []
    for i, (n, d, k) in enumerate([(4000, 32, 40), (3000, 64, 24)]):
        rng = np.random.default_rng(300 + i)
        small = rng.normal(size=(k // 2, d)) * 1e-3
        big = rng.normal(size=(k // 2, d)) * 1e3
        lab = rng.integers(0, k, n)
        centres = np.concatenate([small, big])
        noise = np.where((lab < k // 2)[:, None], 1e-4, 1e2) * rng.normal(size=(n, d))
        problems.append(((centres[lab] + noise).astype(np.float32), k, E.Config(nepochs=2, seed=310 + i)))
    print("batch fp8 mixed magnitudes", flush=True)
    got, _ = E.train_batch(problems, E.HipAssigner("fp8", rescore_m=8))
    for (x, k, cfg), g in zip(problems, got):
        ref = E.train(x, k, cfg, E.HipAssigner("fp8", rescore_m=8))
        assert np.array_equal(g.labels, ref.labels)
        assert np.array_equal(g.centroids, ref.centroids)
        rows = x[np.abs(x).max(axis=1) < 1.0]  # the small clusters' rows
        a = E.HipAssigner("fp8", rescore_m=8)
        a.set_centroids(g.centroids)
        labels, _ = a.assign(rows)
        _, bd = O.brute_f64(rows, g.centroids)
        got_d = O.dist_f64(rows, g.centroids, labels)
        assert len(rows) > 1000 and np.all(got_d <= bd + 1e-6 * (1 + bd))


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_batched_wide_rows_equal_single(hip, prec):
    """Jobs whose operand rows pass 512 bytes (FP32 d > 128, BF16 d > 256,
    FP8 d > 512) run the LDS kernel inside the batched step, rows resident
    in LDS or streamed in d-chunks, beside register-resident jobs; each ends
    with exactly its single-problem result."""
    wide = {"fp32": (200, 1000), "bf16": (300, 1400), "fp8": (600, 2100)}[prec]
    specs = [(3000, wide[0], 60, 0.0), (2500, wide[1], 40, 0.1), (4000, 64, 90, 0.0), (2000, 24, 30, 0.1)]
    problems = [(_job_data(140 + i, n, d, 2 * k), k, E.Config(nepochs=2, balance=bal, seed=150 + i))
                for i, (n, d, k, bal) in enumerate(specs)]
    print(f"batch wide rows {prec}", flush=True)
    got, _ = E.train_batch(problems, E.HipAssigner(prec))
    for (x, k, cfg), g in zip(problems, got):
        ref = E.train(x, k, cfg, E.HipAssigner(prec))
        assert np.array_equal(g.labels, ref.labels)
        assert np.array_equal(g.centroids, ref.centroids)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_batched_slices_and_big_clusters_equal_single(hip, prec):
    """The batched step's own paths: a job whose centroid tiles split into
    two slices (k = 2100, d = 64), jobs whose clusters get hundreds of
    members per step (k = 3 and 5 with batches of 2048: the per-64-dimension
    folds of big clusters, with balance on one), beside an ordinary job; each
    ends with exactly its single-problem result."""
    specs = [(8000, 64, 2100, 0.0, None), (20000, 24, 3, 0.1, 2048),
             (12000, 200 if prec in ("bf16", "fp8") else 100, 5, 0.0, 2048),
             (3000, 40, 50, 0.0, None)]
    problems = []
    for i, (n, d, k, bal, bs) in enumerate(specs):
        cfg = E.Config(nepochs=2, balance=bal, seed=40 + i, batch_size=bs)
        problems.append((_job_data(50 + i, n, d, 2 * k), k, cfg))
    print(f"batch slices/big {prec}", flush=True)
    got, _ = E.train_batch(problems, E.HipAssigner(prec))
    for (x, k, cfg), g in zip(problems, got):
        ref = E.train(x, k, cfg, E.HipAssigner(prec))
        assert np.array_equal(g.labels, ref.labels)
        assert np.array_equal(g.centroids, ref.centroids)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
def test_batched_edge_jobs_equal_single(hip, prec):
    """Edge shapes in one batch: a single cluster, as many rows as clusters,
    three- and five-wide rows (row lengths not a multiple of four), a row
    width whose operand rows round up to 384 bytes (BF16; FP32 stays at
    d <= 128), tiny batches beside a large one."""
    wide = {"bf16": 130, "fp8": 300}.get(prec, 120)
    specs = [(50, 3, 1, 0.0), (64, 7, 64, 0.0), (1000, wide, 700, 0.1), (300, 5, 2, 0.1), (30000, 32, 40, 0.0)]
    problems = [(_job_data(70 + i, n, d, max(2, 2 * k)), k, E.Config(nepochs=2, balance=bal, seed=60 + i))
                for i, (n, d, k, bal) in enumerate(specs)]
    print(f"batch edge jobs {prec}", flush=True)
    got, _ = E.train_batch(problems, E.HipAssigner(prec))
    for (x, k, cfg), g in zip(problems, got):
        ref = E.train(x, k, cfg, E.HipAssigner(prec))
        assert np.array_equal(g.labels, ref.labels)
        assert np.array_equal(g.centroids, ref.centroids)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits


@pytest.mark.parametrize("prec", ["fp32", "i8"])
def test_device_rows_training_equals_host_rows(hip, prec):
    """emh_train_device (rows already in HBM, initial centroids gathered on
    the device) gives exactly the result of emh_train on the same rows."""
    import torch

    print(f"device rows {prec}", flush=True)
    if prec == "i8":
        x = _i8_blobs(6000, 48, 60, 2)
        cfg = E.Config(nepochs=2, seed=5)
        a, b = E.HipAssigner(E.I8), E.HipAssigner(E.I8)
    else:
        x = _job_data(21, 9000, 40, 120)
        cfg = E.Config(nepochs=2, seed=5, balance=0.1)
        a, b = E.HipAssigner(E.FP32), E.HipAssigner(E.FP32)
    ref = E.train(x, 60, cfg, a)
    got = E.train(torch.from_numpy(x).cuda(), 60, cfg, b)
    assert np.array_equal(got.labels, ref.labels)
    assert np.array_equal(got.centroids, ref.centroids)
    assert got.objective == ref.objective and got.n_splits == ref.n_splits


def test_batched_device_rows_equal_host_rows(hip):
    """emh_train_batch with some jobs' rows on the device and some on the host
    gives every job its host-rows result."""
    import torch

    specs = [(3000, 24, 40, 0.0), (9000, 64, 200, 0.1), (5000, 33, 17, 0.0)]
    host, mixed = [], []
    for i, (n, d, k, bal) in enumerate(specs):
        x = _job_data(30 + i, n, d, 2 * k)
        cfg = E.Config(nepochs=2, balance=bal, seed=3 + i)
        host.append((x, k, cfg))
        mixed.append((torch.from_numpy(x).cuda() if i != 1 else x, k, cfg))
    ref, _ = E.train_batch(host, E.HipAssigner("bf16"))
    got, _ = E.train_batch(mixed, E.HipAssigner("bf16"))
    for g, r in zip(got, ref):
        assert np.array_equal(g.labels, r.labels)
        assert np.array_equal(g.centroids, r.centroids)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8", "i8", "i16"])
def test_step_graph_equals_step_launches(hip, prec, monkeypatch):
    """Training without the balance penalty replays each full batch's step as
    one captured graph (step index and per-step rates read on the device);
    the result equals the same training with every step launched kernel by
    kernel (EMH_TRAIN_GRAPH=0) bit for bit: labels, centroids, objective and
    revivals, over a short batch closing each epoch and a fractional epoch."""
    if prec in ("i8", "i16"):
        x = _i8_blobs(9000, 40, 90, 7)
    else:
        x = _job_data(17, 9000, 72, 180)
    cfg = E.Config(nepochs=3.5, seed=21, batch_size=1024)
    print(f"step graph {prec}", flush=True)
    got = E.train(x, 70, cfg, E.HipAssigner(prec))
    monkeypatch.setenv("EMH_TRAIN_GRAPH", "0")
    ref = E.train(x, 70, cfg, E.HipAssigner(prec))
    assert ref.recipe.total_steps == got.recipe.total_steps and got.n_splits == ref.n_splits
    assert np.array_equal(got.labels, ref.labels)
    assert np.array_equal(got.centroids, ref.centroids)
    assert got.objective == ref.objective


def test_step_graph_recaptured_when_search_scratch_moves(hip, monkeypatch, tmp_path):
    """Regression: a captured step addresses the assigner's search scratch
    (the per-slice results of a split search). When a later batch size (the
    short batch closing an epoch) tunes a wider split, that buffer grows and
    moves; the full-batch graph must then be captured again, not replayed
    with the freed address. The full batch's tile comes from the tile cache
    (tuned by an earlier call in this process), the short batch's is tuned
    mid-run, and the result equals the run without graphs."""
    monkeypatch.setenv("EMH_TUNE_CACHE", str(tmp_path / "tiles.json"))
    x = _job_data(23, 20000, 96, 600)
    bs, k = 4096, 1024
    print("step graph recapture", flush=True)
    E.train(x[: 2 * bs], k, E.Config(nepochs=1, batch_size=bs), E.HipAssigner("fp32"))  # tunes the full batch
    cfg = E.Config(nepochs=2, seed=3, batch_size=bs)
    got = E.train(x, k, cfg, E.HipAssigner("fp32"))
    monkeypatch.setenv("EMH_TRAIN_GRAPH", "0")
    ref = E.train(x, k, cfg, E.HipAssigner("fp32"))
    assert np.array_equal(got.labels, ref.labels)
    assert np.array_equal(got.centroids, ref.centroids)
    assert got.objective == ref.objective and got.n_splits == ref.n_splits


def test_batched_i8_jobs_equal_oracle(hip):
    """int8 jobs in one lockstep batch (the packed integer search of every
    job in one launch, i8 rows summed on the device): each job's centroids,
    last-pass labels, objective and revivals equal the CPU oracle's int8
    training bit for bit. Shapes cover row lengths off the 16-byte quad
    (d = 7, 100), every operand width class up to 512 bytes (d = 300 and 512
    also take the sums past 256 dimensions), a job whose centroid tiles
    split into slices (k = 3000), clusters of hundreds of members per step
    (k = 3, batches of 2048) and the balance penalty."""
    specs = [(6000, 48, 40, 0.1, None), (900, 7, 30, 0.0, None), (4000, 100, 64, 0.1, None),
             (3000, 200, 25, 0.0, None), (2500, 300, 20, 0.1, None), (2000, 512, 16, 0.0, None),
             (9000, 64, 3000, 0.0, None), (12000, 24, 3, 0.1, 2048)]
    problems = []
    for i, (n, d, k, bal, bs) in enumerate(specs):
        x = _i8_blobs(n, d, max(2, 2 * k), 80 + i)
        problems.append((x, k, (bal, bs, 90 + i)))
    batch = [(x, k, E.Config(nepochs=2 + (i % 2), balance=bal, seed=sd, batch_size=bs))
             for i, (x, k, (bal, bs, sd)) in enumerate(problems)]
    print(f"batch i8: {len(batch)} jobs", flush=True)
    got, st = E.train_batch(batch, E.HipAssigner(E.I8))
    assert st.steps == max(E.resolve_recipe(x.shape[0], x.shape[1], k, c).total_steps for x, k, c in batch)
    for (x, k, cfg), g in zip(batch, got):
        ocfg = O.Config(nepochs=cfg.nepochs, balance=cfg.balance, seed=cfg.seed,
                        batch_size=cfg.batch_size if cfg.has_batch_size else None)
        ref = O.train(x, k, ocfg, assigner=O.ASSIGNER_I8)
        assert g.recipe.batch_size == ref.recipe.batch_size
        assert np.array_equal(g.labels, ref.labels)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits
        assert np.array_equal(g.centroids, ref.centroids)


def test_batched_i8_wide_rows_equal_oracle(hip):
    """int8 jobs whose rows pass 512 bytes run the LDS kernel's i64 search
    inside the batched step (rows resident in LDS, and streamed in d-chunks
    for the widest), beside a register-resident job; each equals the CPU
    oracle's int8 training bit for bit, balance on one."""
    specs = [(1500, 600, 20, 0.1), (1200, 1100, 12, 0.0), (3000, 64, 30, 0.0)]
    batch = []
    for i, (n, d, k, bal) in enumerate(specs):
        batch.append((_i8_blobs(n, d, 2 * k, 120 + i), k, E.Config(nepochs=2, balance=bal, seed=130 + i)))
    print(f"batch i8 wide: {len(batch)} jobs", flush=True)
    got, _ = E.train_batch(batch, E.HipAssigner(E.I8))
    for (x, k, cfg), g in zip(batch, got):
        ref = O.train(x, k, O.Config(nepochs=cfg.nepochs, balance=cfg.balance, seed=cfg.seed), assigner=O.ASSIGNER_I8)
        assert np.array_equal(g.labels, ref.labels)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits
        assert np.array_equal(g.centroids, ref.centroids)


def test_batched_i16_jobs_equal_oracle(hip):
    """int8 rows with CentroidBits::I16 centroids in one lockstep batch (the
    LDS kernel's exact i64 scores over two i8 planes): each job equals the
    CPU oracle's I16 training bit for bit, balance on some, rows from 48 to
    600 bytes."""
    specs = [(3000, 48, 30, 0.1), (1500, 200, 20, 0.0), (1000, 600, 12, 0.1), (2500, 7, 9, 0.0)]
    batch = []
    for i, (n, d, k, bal) in enumerate(specs):
        batch.append((_i8_blobs(n, d, 2 * k, 160 + i), k, E.Config(nepochs=2, balance=bal, seed=170 + i)))
    print(f"batch i16: {len(batch)} jobs", flush=True)
    got, _ = E.train_batch(batch, E.HipAssigner(E.I16))
    for (x, k, cfg), g in zip(batch, got):
        ref = O.train(x, k, O.Config(nepochs=cfg.nepochs, balance=cfg.balance, seed=cfg.seed), assigner=O.ASSIGNER_I16)
        assert np.array_equal(g.labels, ref.labels)
        assert g.objective == ref.objective
        assert g.n_splits == ref.n_splits
        assert np.array_equal(g.centroids, ref.centroids)


def test_batched_i8_device_rows_equal_host_rows(hip):
    """A batched int8 job whose rows already sit in HBM (initial centroids
    gathered and widened on the device) ends as its host-rows twin."""
    import torch

    xs = [_i8_blobs(5000, 40, 60, 3), _i8_blobs(3000, 130, 40, 4)]
    cfgs = [E.Config(nepochs=2, seed=11), E.Config(nepochs=2, balance=0.1, seed=12)]
    host = [(x, 30, c) for x, c in zip(xs, cfgs)]
    dev = [(torch.from_numpy(xs[0]).cuda(), 30, cfgs[0]), host[1]]
    ref, _ = E.train_batch(host, E.HipAssigner(E.I8))
    got, _ = E.train_batch(dev, E.HipAssigner(E.I8))
    for g, r in zip(got, ref):
        assert np.array_equal(g.labels, r.labels)
        assert np.array_equal(g.centroids, r.centroids)
        assert g.objective == r.objective


def test_batched_i8_bias_out_of_packed_range_is_an_error(hip):
    """A balance penalty so strong that the rounded bias leaves the packed
    int8 score range stops the batch with a clear error (the single-problem
    driver takes the i64 search instead); the kernels all run in bounds."""
    x = _i8_blobs(3000, 512, 20, 6)
    cfg = E.Config(nepochs=2, balance=1000.0, seed=1)
    with pytest.raises(E.EmeansError, match="packed I8 score range"):
        E.train_batch([(x, 10, cfg)], E.HipAssigner(E.I8))
    # the same problem on the single-problem driver: the oracle's result
    ref = O.train(x, 10, O.Config(nepochs=2, balance=1000.0, seed=1), assigner=O.ASSIGNER_I8)
    got = E.train(x, 10, cfg, E.HipAssigner(E.I8))
    assert np.array_equal(got.labels, ref.labels) and np.array_equal(got.centroids, ref.centroids)
