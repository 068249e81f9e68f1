"""GPU checks of the differentiable batched solve (magma_amd.autograd): the
forward solution against LAPACK, the hand-written adjoint (grad_B = A^-T G,
grad_A = -grad_B X^T) against an FP64 CPU reference gradient, against central
finite differences (h = 1e-6, FP64) and through torch.autograd.gradcheck; FP32
against the FP64 reference within an FP32 tolerance."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ag():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from magma_amd import autograd
    return autograd


def make(rng, batch, n, nrhs):
    A = rng.uniform(-1, 1, (batch, n, n)) + 8.0 * np.eye(n)
    B = rng.uniform(-1, 1, (batch, n, nrhs))
    W = rng.uniform(-1, 1, (batch, n, nrhs))
    return A, B, W


def reference_grads(A, B, W):
    """FP64 CPU gradient of sum(W * X), X = A^{-1} B (LAPACK via numpy)."""
    X = np.linalg.solve(A, B)
    gB = np.linalg.solve(np.transpose(A, (0, 2, 1)), W)
    gA = -gB @ np.transpose(X, (0, 2, 1))
    return X, gA, gB


@pytest.mark.parametrize("batch,n,nrhs", [(32, 64, 8), (5, 17, 3), (3, 128, 1)])
def test_adjoint_fp64_matches_cpu_reference(ag, batch, n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    A, B, W = make(rng, batch, n, nrhs)
    tA = torch.tensor(A, device="cuda", requires_grad=True)
    tB = torch.tensor(B, device="cuda", requires_grad=True)
    X = ag.solve(tA, tB)
    loss = (torch.tensor(W, device="cuda") * X).sum()
    loss.backward()
    Xr, gAr, gBr = reference_grads(A, B, W)
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    assert rel(X.detach().cpu().numpy(), Xr) < 1e-13
    assert rel(tB.grad.cpu().numpy(), gBr) < 1e-13
    assert rel(tA.grad.cpu().numpy(), gAr) < 1e-13


def test_adjoint_central_differences(ag):
    rng = np.random.default_rng(7)
    A, B, W = make(rng, 2, 12, 3)
    tA = torch.tensor(A, device="cuda", requires_grad=True)
    tB = torch.tensor(B, device="cuda", requires_grad=True)
    tW = torch.tensor(W, device="cuda")
    (tW * ag.solve(tA, tB)).sum().backward()
    h = 1e-6

    def loss(Av, Bv):
        with torch.no_grad():
            return float((tW * ag.solve(torch.tensor(Av, device="cuda"), torch.tensor(Bv, device="cuda"))).sum())

    for _ in range(6):                    # random directions in A and in B
        EA = rng.standard_normal(A.shape)
        fd = (loss(A + h * EA, B) - loss(A - h * EA, B)) / (2 * h)
        an = float((tA.grad.cpu().numpy() * EA).sum())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
        EB = rng.standard_normal(B.shape)
        fd = (loss(A, B + h * EB) - loss(A, B - h * EB)) / (2 * h)
        an = float((tB.grad.cpu().numpy() * EB).sum())
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)


def test_gradcheck(ag):
    rng = np.random.default_rng(11)
    A, B, _ = make(rng, 2, 6, 2)
    tA = torch.tensor(A, device="cuda", requires_grad=True)
    tB = torch.tensor(B, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: ag.solve(a, b), (tA, tB), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_adjoint_fp32(ag):
    rng = np.random.default_rng(3)
    A, B, W = make(rng, 64, 64, 8)
    tA = torch.tensor(A, device="cuda", dtype=torch.float32, requires_grad=True)
    tB = torch.tensor(B, device="cuda", dtype=torch.float32, requires_grad=True)
    X = ag.solve(tA, tB)
    (torch.tensor(W, device="cuda", dtype=torch.float32) * X).sum().backward()
    # reference on the FP32-rounded inputs, in FP64
    A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    Xr, gAr, gBr = reference_grads(A32, B32, W.astype(np.float32).astype(np.float64))
    rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
    tol = 64 * np.finfo(np.float32).eps
    assert rel(X.detach().double().cpu().numpy(), Xr) < tol
    assert rel(tB.grad.double().cpu().numpy(), gBr) < tol
    assert rel(tA.grad.double().cpu().numpy(), gAr) < tol


def test_singular_raises(ag):
    A = torch.eye(4, dtype=torch.float64, device="cuda").repeat(2, 1, 1)
    A[1, :, 2] = 0.0
    with pytest.raises(RuntimeError, match="singular"):
        ag.solve(A, torch.ones(2, 4, 1, dtype=torch.float64, device="cuda"))
