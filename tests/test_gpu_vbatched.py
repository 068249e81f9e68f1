"""GPU parity of the variable-size and single-precision batched paths against
the oracle (testing_zgetrf_vbatched checks: pivots match, LU error)."""
import numpy as np
import pytest

from conftest import EPS, TESTER_TOL

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

LU_FWD_TOL = 1e3 * EPS


@pytest.fixture(scope="module")
def mb():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import batched
    return batched


def cm(M):
    return np.ascontiguousarray(M.T)[None]


def check_one(oracle, A, LU, ipiv, info, m, n, tag):
    LUo, ipo, info_o = oracle.dgetrf(cm(A), m=m)
    assert int(info) == int(info_o[0]), tag
    if oracle.pivot_margin(cm(A), m)[0] > 1e-10:
        assert np.array_equal(ipiv, ipo[0]), tag
    fe = oracle.lu_forward_error(LUo, cm(LU), m)[0]
    assert fe < LU_FWD_TOL, (tag, fe)


def make_vbatch(mb, rng, ms, ns, lds=None):
    A = mb.VBatch(ms, ns, ld=lds)
    mats = []
    for i, (m, n) in enumerate(zip(ms, ns)):
        M = rng.uniform(-1, 1, (m, n))
        mats.append(M)
        A.set(i, M)
    return A, mats


def test_getrf_vbatched_mixed_sizes(mb, oracle):
    rng = np.random.default_rng(3)
    ns = [1, 5, 16, 17, 31, 64, 100, 128, 129, 150, 200, 256, 300, 512, 40, 3, 77]
    A, mats = make_vbatch(mb, rng, ns, ns)
    ipiv, info = mb.dgetrf_vbatched(A)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    for i, n in enumerate(ns):
        check_one(oracle, mats[i], A.get(i), ipiv.get(i, n), info[i], n, n, f"n={n}")


def test_getrf_vbatched_rectangular_and_padded(mb, oracle):
    rng = np.random.default_rng(4)
    ms = [40, 24, 300, 64, 7, 130]
    ns = [24, 40, 200, 96, 33, 130]
    lds = [48, 24, 300, 70, 9, 131]
    A, mats = make_vbatch(mb, rng, ms, ns, lds)
    ipiv, info = mb.dgetrf_vbatched(A)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    for i, (m, n) in enumerate(zip(ms, ns)):
        check_one(oracle, mats[i], A.get(i), ipiv.get(i, min(m, n)), info[i], m, n, f"{m}x{n}")


def test_getrf_vbatched_singular(mb, oracle):
    rng = np.random.default_rng(5)
    ns = [12, 100, 160]
    A = mb.VBatch(ns, ns)
    mats = []
    for i, n in enumerate(ns):
        M = rng.uniform(-1, 1, (n, n))
        M[:, 4] = 0.0
        mats.append(M)
        A.set(i, M)
    ipiv, info = mb.dgetrf_vbatched(A)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 5).all()


def test_getrf_vbatched_argument_codes(mb):
    A = mb.VBatch([4, 4], [4, 4])
    A.d_m[1] = -1
    with pytest.raises(mb.MagmaError) as e:
        mb.dgetrf_vbatched(A)
    assert e.value.code == -1
    A = mb.VBatch([4, 4], [4, 4])
    A.d_ld[0] = 2
    with pytest.raises(mb.MagmaError) as e:
        mb.dgetrf_vbatched(A)
    assert e.value.code == -4


def test_getrf_vbatched_max_nocheck_async(mb, oracle):
    rng = np.random.default_rng(6)
    ns = [8, 33, 128, 140, 250]
    A, mats = make_vbatch(mb, rng, ns, ns)
    ipiv, info = mb.dgetrf_vbatched_max_nocheck(A)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    for i, n in enumerate(ns):
        check_one(oracle, mats[i], A.get(i), ipiv.get(i, n), info[i], n, n, f"nocheck n={n}")


@pytest.mark.parametrize("trans", [0, 1])
def test_getrs_vbatched(mb, oracle, trans):
    rng = np.random.default_rng(7 + trans)
    ns = [3, 20, 64, 128, 200, 333]
    rs = [1, 4, 16, 2, 7, 3]
    LUs = mb.VBatch(ns, ns)
    Bs = mb.VBatch(ns, rs)
    ip = mb.IpivV(ns)
    refs = []
    for i, (n, r) in enumerate(zip(ns, rs)):
        A = rng.uniform(-1, 1, (n, n))
        B = rng.uniform(-1, 1, (n, r))
        LU, piv, _ = oracle.dgetrf(cm(A))
        LUs.set(i, LU[0].T)
        ip.data[int(ip.offs[i]):int(ip.offs[i]) + n] = torch.from_numpy(piv[0]).cuda()
        Bs.set(i, B)
        refs.append((A, B))
    mb.dgetrs_vbatched(111 if trans == 0 else 112, LUs, ip, Bs)
    torch.cuda.synchronize()
    for i, n in enumerate(ns):
        A, B = refs[i]
        X = Bs.get(i)
        Aop = A if trans == 0 else A.T
        be = oracle.gesv_backward_error(cm(Aop), cm(X), cm(B), n)[0]
        assert be < TESTER_TOL, (n, be)


def _random_system_batch(mb, rng, ns, rs):
    A = mb.VBatch(ns, ns)
    B = mb.VBatch(ns, rs)
    ref = []
    for i, (n, r) in enumerate(zip(ns, rs)):
        Am = rng.uniform(-1, 1, (n, n))
        Bm = rng.uniform(-1, 1, (n, r))
        A.set(i, Am)
        B.set(i, Bm)
        ref.append((Am, Bm))
    return A, B, ref


def test_gesv_vbatched(mb, oracle):
    rng = np.random.default_rng(8)
    ns = list(rng.integers(8, 513, 40))
    rs = list(rng.integers(1, 17, 40))
    A, B, ref = _random_system_batch(mb, rng, ns, rs)
    ipiv, info = mb.dgesv_vbatched(A, B)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    for i, n in enumerate(ns):
        Am, Bm = ref[i]
        be = oracle.gesv_backward_error(cm(Am), cm(B.get(i)), cm(Bm), n)[0]
        assert be < TESTER_TOL, (n, be)
        LUo, ipo, _ = oracle.dgetrf(cm(Am))
        if oracle.pivot_margin(cm(Am))[0] > 1e-10:
            assert np.array_equal(ipiv.get(i, n), ipo[0]), n


@pytest.mark.parametrize("variant", [None, 2, 3, 4, 5])
def test_vbatched_plan_graph_replay(mb, oracle, variant):
    """A plan executed directly, then captured into a hipGraph and replayed
    several times with eager work in between; every replay matches."""
    rng = np.random.default_rng(9)
    ns = list(rng.integers(8, 300, 30))
    rs = list(rng.integers(1, 17, 30))
    A, B, ref = _random_system_batch(mb, rng, ns, rs)
    A0 = A.data.clone()
    B0 = B.data.clone()
    variants = None if variant is None else {b: variant for b in range(22)}
    plan = mb.VbatchedPlan(ns, rs, variants)
    ipiv = mb.IpivV(ns)
    info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
    plan.execute(A, B, ipiv, info)
    torch.cuda.synchronize()
    X_eager = B.data.clone()
    A.data.copy_(A0)
    B.data.copy_(B0)
    plan.capture(A, B, ipiv, info)
    for rep in range(3):
        A.data.copy_(A0)
        B.data.copy_(B0)
        _ = (A0 * 2).sum()                  # eager work between replays
        torch.cuda.synchronize()
        plan.replay()
        torch.cuda.synchronize()
        assert torch.equal(B.data, X_eager), rep
    for i, n in enumerate(ns):
        Am, Bm = ref[i]
        be = oracle.gesv_backward_error(cm(Am), cm(B.get(i)), cm(Bm), n)[0]
        assert be < TESTER_TOL, (n, be)
    plan.close()


@pytest.mark.parametrize("n,nrhs", [(16, 1), (64, 3), (128, 4), (129, 2), (200, 2), (241, 3), (256, 4), (300, 5)])
def test_sgesv_batched(mb, oracle, n, nrhs):
    rng = np.random.default_rng(10 + n)
    batch = 16
    A = rng.uniform(-1, 1, (batch, n, n)).astype(np.float32)
    B = rng.uniform(-1, 1, (batch, nrhs, n)).astype(np.float32)
    dA = torch.from_numpy(A.copy()).cuda()
    dB = torch.from_numpy(B.copy()).cuda()
    ipiv, info = mb.sgetrf_batched(dA)
    mb.sgetrs_batched(111, dA, ipiv, dB)
    torch.cuda.synchronize()
    X = dB.cpu().numpy().astype(np.float64)
    be = oracle.gesv_backward_error(A.astype(np.float64), X, B.astype(np.float64), n)
    assert (be < 30 * np.finfo(np.float32).eps).all(), be.max()
    margin = oracle.pivot_margin(A.astype(np.float64))
    LUo, ipo, _ = oracle.dgetrf(A.astype(np.float64))
    tf = margin > 1e-4          # FP32 rounding: compare pivots where the gap is clear
    assert np.array_equal(ipiv.cpu().numpy()[tf], ipo[tf])


@pytest.mark.parametrize("variant", [2, 5])
def test_vbatched_plan_lu_parity(mb, oracle, variant):
    """The blocked variants (32-wide panels, and two 32-wide panels per
    64-wide block) give the oracle's pivots on tie-free systems and its LU
    within the forward-error bound, including orders that end in a partial
    block (n = 129, 177, 300, 449)."""
    rng = np.random.default_rng(40 + variant)
    ns = [129, 140, 177, 192, 200, 256, 257, 300, 449, 512]
    rs = [1] * len(ns)
    A, B, ref = _random_system_batch(mb, rng, ns, rs)
    plan = mb.VbatchedPlan(ns, rs, {b: variant for b in range(22)})
    ipiv = mb.IpivV(ns)
    info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
    plan.execute(A, B, ipiv, info)
    torch.cuda.synchronize()
    plan.close()
    assert (info.cpu().numpy() == 0).all()
    for i, n in enumerate(ns):
        Am, _ = ref[i]
        check_one(oracle, Am, A.get(i), ipiv.get(i, n), 0, n, n, f"variant {variant} n={n}")
