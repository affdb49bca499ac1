"""Forward / adjoint kernels through torch.autograd, the fused completion
loss+gradient, central finite differences in FP64, and the Adam driver."""
import sys

import numpy as np
import pytest

from dynasor_plus_amd import GpuContext, Tensor, host
from tests._oracle import load as load_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def frob_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def problem():
    """Bernoulli labels from a planted rank-4 model with N(0, 0.5) factors."""
    T, _, _ = host.generate((60, 50, 40), 12_000, seed=21, coords="uniform", values="bernoulli",
                            planted_rank=4, planted_factors="normal", planted_var=0.5)
    return T


def test_forward_matches_oracle(problem):
    say("forward")
    from dynasor_plus_amd.autograd import cp_forward
    orc = load_oracle()
    rng = np.random.default_rng(1)
    A = [rng.standard_normal((d, 6)).astype(np.float32) * 0.5 for d in problem.dims]
    with GpuContext(0) as eng:
        eng.upload(problem)
        At = [torch.tensor(a, device="cuda") for a in A]
        m = cp_forward(eng, At).cpu().numpy()
    assert frob_rel(m, orc.forward(problem, A)) < 1e-5


def test_autograd_adjoint_matches_residual_mttkrp(problem):
    say("adjoint vs residual mttkrp")
    from dynasor_plus_amd.autograd import bernoulli_logit_loss, cp_forward
    orc = load_oracle()
    rng = np.random.default_rng(2)
    R = 6
    A = [rng.standard_normal((d, R)) * 0.5 for d in problem.dims]
    with GpuContext(0) as eng:
        eng.upload(problem)
        At = [torch.tensor(a, device="cuda", dtype=torch.float32, requires_grad=True) for a in A]
        y = torch.tensor(problem.vals, device="cuda")
        L = bernoulli_logit_loss(cp_forward(eng, At), y)
        L.backward()
        grads = [a.grad.cpu().numpy() for a in At]
    # reference: MTTKRP (definition) of the residual tensor sigma(m) - y, in FP64
    m = orc.forward(problem, A)
    resid = 1.0 / (1.0 + np.exp(-m)) - problem.vals.astype(np.float64)
    Tg = Tensor(problem.dims, resid.astype(np.float32), problem.idx)
    for n in range(3):
        want = orc.mttkrp_f64(Tg, A, n)
        assert frob_rel(grads[n], want) < 1e-4, n


def test_adjoint_matches_central_differences(problem):
    """FP64 central differences of the FP64 loss vs the device adjoint (FP32)."""
    say("finite differences")
    from dynasor_plus_amd.autograd import completion_loss_and_grad
    orc = load_oracle()
    rng = np.random.default_rng(3)
    R = 5
    A = [rng.standard_normal((d, R)) * 0.4 for d in problem.dims]
    with GpuContext(0) as eng:
        eng.upload(problem)
        At = [torch.tensor(a, device="cuda", dtype=torch.float32) for a in A]
        loss, grads = completion_loss_and_grad(eng, At, "bernoulli")
        grads = [g.cpu().numpy() for g in grads]
    L0 = orc.completion_loss(problem, A, 1)
    assert abs(loss - L0) / abs(L0) < 1e-5
    h = 1e-4
    fd, ad = [], []
    for _ in range(40):
        n = int(rng.integers(0, 3))
        i = int(rng.integers(0, problem.dims[n]))
        r = int(rng.integers(0, R))
        Ap = [a.copy() for a in A]
        Am = [a.copy() for a in A]
        Ap[n][i, r] += h
        Am[n][i, r] -= h
        fd.append((orc.completion_loss(problem, Ap, 1) - orc.completion_loss(problem, Am, 1)) / (2 * h))
        ad.append(grads[n][i, r])
    fd, ad = np.array(fd), np.array(ad)
    assert np.linalg.norm(ad - fd) / np.linalg.norm(fd) < 1e-4


def test_fused_matches_autograd(problem):
    say("fused vs autograd")
    from dynasor_plus_amd.autograd import (bernoulli_logit_loss, completion_loss_and_grad,
                                           cp_forward)
    rng = np.random.default_rng(4)
    A = [rng.standard_normal((d, 7)).astype(np.float32) * 0.3 for d in problem.dims]
    with GpuContext(0) as eng:
        eng.upload(problem)
        At = [torch.tensor(a, device="cuda", requires_grad=True) for a in A]
        y = torch.tensor(problem.vals, device="cuda")
        L = bernoulli_logit_loss(cp_forward(eng, At), y)
        L.backward()
        loss, grads = completion_loss_and_grad(eng, At, "bernoulli")
        for n in range(3):
            assert frob_rel(grads[n].cpu().numpy(), At[n].grad.cpu().numpy()) < 1e-5
        assert abs(loss - float(L)) / abs(float(L)) < 1e-5
        lsq, _ = completion_loss_and_grad(eng, At, "ls")
        m = cp_forward(eng, [a.detach() for a in At])
        assert abs(lsq - float(0.5 * ((m - y) ** 2).sum())) / lsq < 1e-5


def test_forward_adjoint_bitwise_deterministic(problem):
    say("determinism")
    from dynasor_plus_amd.autograd import bernoulli_logit_loss, cp_forward
    rng = np.random.default_rng(5)
    A = [rng.standard_normal((d, 8)).astype(np.float32) * 0.3 for d in problem.dims]
    outs = []
    with GpuContext(0) as eng:
        eng.upload(problem)
        y = torch.tensor(problem.vals, device="cuda")
        for _ in range(2):
            At = [torch.tensor(a, device="cuda", requires_grad=True) for a in A]
            m = cp_forward(eng, At)
            gm = torch.sigmoid(m.detach()) - y
            m.backward(gm)
            outs.append([m.detach().cpu().numpy()] + [a.grad.cpu().numpy() for a in At])
    for x, z in zip(*outs):
        np.testing.assert_array_equal(x, z)


def test_adam_completion_driver(problem):
    """100 Adam steps on the fused loss/gradient: the loss falls, and the loss
    reported for a step is the FP64 loss (oracle) of the factors of that step."""
    say("adam")
    from dynasor_plus_amd.autograd import completion_loss_and_grad, fit_completion
    orc = load_oracle()
    steps = 100
    with GpuContext(0) as eng:
        eng.upload(problem)
        A, losses = fit_completion(eng, problem.dims, rank=8, steps=steps, lr=1e-2, seed=0)
        final, _ = completion_loss_and_grad(eng, A, "bernoulli")
    assert np.all(np.isfinite(losses)) and len(losses) == steps
    assert losses[-1] < 0.95 * losses[0]
    assert np.mean(np.diff(losses) < 0) > 0.9            # decreasing almost every step
    L64 = orc.completion_loss(problem, [a.cpu().numpy().astype(np.float64) for a in A], 1)
    assert abs(final - L64) / L64 < 1e-5
    assert final < losses[-1]


def test_empty_rows_zeroed_in_caller_buffers():
    """Rows without nonzeros are written (as zeros) by the kernels themselves:
    the caller's NaN-filled output buffers come back finite, with exact zeros
    on the empty rows and the FP64 oracle elsewhere."""
    say("empty rows in caller buffers")
    import ctypes as C
    from dynasor_plus_amd import _lib as L
    from dynasor_plus_amd.autograd import _ptrs
    orc = load_oracle()
    rng = np.random.default_rng(9)
    dims, nnz, R = (40, 30, 20), 3000, 16
    idx = []
    for d in dims:
        allowed = np.array([r for r in range(d) if r % 4 != 1])     # every fourth row empty
        idx.append(rng.choice(allowed, nnz).astype(np.uint32))
    idx[0][:700] = 2                                                 # one heavy row
    T = Tensor(dims, rng.standard_normal(nnz).astype(np.float32), idx)
    A = [(rng.standard_normal((d, R)) * 0.5).astype(np.float32) for d in dims]
    g = rng.standard_normal(nnz).astype(np.float32)
    with GpuContext(0) as eng:
        eng.upload(T)
        At = [torch.tensor(a, device="cuda") for a in A]
        gd = torch.tensor(g, device="cuda")
        adj = [torch.full((d, R), float("nan"), device="cuda") for d in dims]
        comp = [torch.full((d, R), float("nan"), device="cuda") for d in dims]
        loss = C.c_double(0.0)
        with eng.on_stream(torch.cuda.current_stream().cuda_stream):
            for n in range(3):
                eng._check(eng.lib.dyn_gpu_adjoint_dev(eng.h, L._vp(gd.data_ptr()), n, _ptrs(At), R,
                                                       L._vp(adj[n].data_ptr())), "adjoint_dev")
            eng._check(eng.lib.dyn_gpu_completion_grad_dev(eng.h, L.DYN_LOSS_LS, _ptrs(At), R,
                                                           _ptrs(comp), C.byref(loss)),
                       "completion_grad_dev")
        torch.cuda.synchronize()
        adj = [x.cpu().numpy() for x in adj]
        comp = [x.cpu().numpy() for x in comp]
    resid = orc.forward(T, A) - T.vals.astype(np.float64)
    for n, d in enumerate(dims):
        empty = np.setdiff1d(np.arange(d), idx[n])
        assert empty.size >= d // 4
        for got in (adj[n], comp[n]):
            assert np.all(np.isfinite(got))
            assert not got[empty].any()
        assert frob_rel(adj[n], orc.mttkrp_f64(Tensor(dims, g, idx), A, n)) < 1e-5
        assert frob_rel(comp[n], orc.mttkrp_f64(Tensor(dims, resid.astype(np.float32), idx),
                                                A, n)) < 1e-4


@pytest.mark.parametrize("dims", [(40, 30, 20, 10), (20, 16, 12, 10, 8)])
def test_fused_ls_gradient_more_modes(dims):
    """The fused loss/gradient kernel with three and four gathered operands:
    least-squares gradients against the FP64 MTTKRP of the residual tensor,
    and the loss against the oracle's."""
    say("fused ls gradient", dims)
    from dynasor_plus_amd.autograd import completion_loss_and_grad
    orc = load_oracle()
    rng = np.random.default_rng(len(dims))
    T, _, _ = host.generate(dims, 6000, seed=len(dims) + 30, coords="uniform", values="lognormal")
    R = 8
    A = [(rng.standard_normal((d, R)) * 0.5).astype(np.float32) for d in dims]
    with GpuContext(0) as eng:
        eng.upload(T)
        loss, grads = completion_loss_and_grad(eng, [torch.tensor(a, device="cuda") for a in A], "ls")
        grads = [g.cpu().numpy() for g in grads]
    m = orc.forward(T, A)
    resid = m - T.vals.astype(np.float64)
    want_loss = 0.5 * float(np.sum(resid * resid))
    assert abs(loss - want_loss) / want_loss < 1e-5
    Tr = Tensor(T.dims, resid.astype(np.float32), T.idx)
    for n in range(len(dims)):
        assert frob_rel(grads[n], orc.mttkrp_f64(Tr, A, n)) < 1e-4, n
