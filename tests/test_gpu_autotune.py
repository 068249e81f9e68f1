"""GPU checks of the variable-size autotuner (magma_amd.autotune): every bucket
of the sizes gets a valid variant, the cache is written and read back without
re-tuning, and a plan built with the tuned variants solves the batch to the
tester's backward-error bound."""
import json

import numpy as np
import pytest

from conftest import TESTER_TOL

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def at():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import autotune
    return autotune


def test_bucket_ids_match_plan(at):
    from magma_amd import batched as mb
    ns = np.array([1, 16, 17, 100, 128, 129, 192, 193, 300, 512, 513, 700, 1024, 1025, 2048], dtype=np.int32)
    plan = mb.VbatchedPlan(ns, np.ones_like(ns))
    got = {b["id"]: b["count"] for b in plan.buckets()}
    plan.close()
    ids, counts = np.unique(at.bucket_of(ns), return_counts=True)
    assert got == dict(zip(ids.tolist(), counts.tolist()))


def test_tune_cache_and_solve(at, tmp_path):
    from magma_amd import batched as mb
    rng = np.random.default_rng(9)
    ns = rng.integers(8, 260, 120).astype(np.int32)
    rs = rng.integers(1, 5, 120).astype(np.int32)
    cache = tmp_path / "tune.json"
    v = at.load_variants(ns, rs, path=cache)
    want = set(int(b) for b in at.bucket_of(ns))
    assert set(v) == want
    assert all(x in at.VARIANT_NAMES for x in v.values())
    table = json.loads(cache.read_text())
    entry = table[at.device_key()]
    assert set(int(k) for k in entry) == want
    for k, rec in entry.items():
        assert rec["name"] in rec["seconds"]
        assert rec["seconds"][rec["name"]] == min(rec["seconds"].values())
    # a second call reads the cache and tunes nothing
    stamp = cache.stat().st_mtime_ns
    assert at.load_variants(ns, rs, path=cache) == v
    assert cache.stat().st_mtime_ns == stamp

    A = mb.VBatch(ns, ns)
    B = mb.VBatch(ns, rs)
    mats, rhs = [], []
    for i, (n, r) in enumerate(zip(ns, rs)):
        M = rng.uniform(-1, 1, (n, n))
        R = rng.uniform(-1, 1, (n, r))
        A.set(i, M)
        B.set(i, R)
        mats.append(M)
        rhs.append(R)
    ipiv = mb.IpivV(ns)
    info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
    plan = mb.VbatchedPlan(ns, rs, variants=v)
    assert {b["id"]: b["variant"] for b in plan.buckets()} == v
    plan.execute(A, B, ipiv, info)
    torch.cuda.synchronize()
    plan.close()
    assert (info.cpu().numpy() == 0).all()
    for i, n in enumerate(ns):
        X = B.get(i)
        r = np.abs(rhs[i] - mats[i] @ X).sum(1).max()           # infinity norms
        be = r / (n * np.abs(mats[i]).sum(1).max() * np.abs(X).sum(1).max())
        assert be < TESTER_TOL, (i, n, be)
