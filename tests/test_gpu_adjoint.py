"""The hand-written adjoint of the differentiable LJ trajectory (libmdx
mdx_lj_trajectory) against a dense FP64 CPU autograd reference at N=500
and against central finite differences at N=32,000.

Loss: (mean P - 1.5)^2 + (mean PE/N + 4)^2 over the last
100 of 200 drift-kick-drift steps, checkpoints every 20 steps; gradients
with respect to eps, sigma and the initial positions.
"""
import numpy as np
import pytest

from adjoint_reference import loss_and_grads
from mdacp_amd import workloads as W
from mdacp_amd.mdx import MdxError, System

pytestmark = pytest.mark.gpu

OPTS = dict(nsteps=200, nobs=100, checkpoint_every=20, p_target=1.5, pe_target=-4.0)


def gpu_run(state, eps=1.0, sigma=1.0, grad=True, **kw):
    opts = dict(OPTS, **kw)
    with System.from_state(state) as s:
        return s.lj_trajectory(eps=eps, sigma=sigma, grad=grad, **opts)


def test_adjoint_vs_dense_autograd_n500():
    s = W.fcc_state(5, 0.7, 1.0, 5)          # N=500, L=8.94 (2 cells per axis)
    print(f"adjoint vs dense autograd N={s.n}", flush=True)
    g = gpu_run(s)
    loss, ge, gs, gq, Pm, Em = loss_and_grads(s.q, s.p, s.L, nsteps=200, nobs=100, segment=20)
    assert g["loss"] == pytest.approx(loss, rel=1e-10)
    assert g["pressure_mean"] == pytest.approx(Pm, rel=1e-10)
    assert g["pe_mean"] == pytest.approx(Em, rel=1e-10)
    assert g["dloss_deps"] == pytest.approx(ge, rel=1e-8)
    assert g["dloss_dsigma"] == pytest.approx(gs, rel=1e-8)
    err = np.abs(g["dloss_dq0"] - gq).max() / np.abs(gq).max()
    assert err < 1e-8, err


def test_adjoint_vs_finite_differences_n32000():
    s = W.fcc_state(20, 0.7, 1.0, 5)         # N=32000, rho=0.7, T=1
    print(f"adjoint vs central differences N={s.n}", flush=True)
    g = gpu_run(s)
    h = 1e-6
    fd_e = (gpu_run(s, eps=1 + h, grad=False)["loss"] - gpu_run(s, eps=1 - h, grad=False)["loss"]) / (2 * h)
    fd_s = (gpu_run(s, sigma=1 + h, grad=False)["loss"] - gpu_run(s, sigma=1 - h, grad=False)["loss"]) / (2 * h)
    assert g["dloss_deps"] == pytest.approx(fd_e, rel=1e-6)
    assert g["dloss_dsigma"] == pytest.approx(fd_s, rel=1e-6)
    rng = np.random.default_rng(0)
    for _ in range(3):
        i, d = int(rng.integers(s.n)), int(rng.integers(3))
        sp, sm = s.copy(), s.copy()
        sp.q[i, d] += h
        sm.q[i, d] -= h
        fd = (gpu_run(sp, grad=False)["loss"] - gpu_run(sm, grad=False)["loss"]) / (2 * h)
        scale = np.abs(g["dloss_dq0"]).max()
        assert abs(g["dloss_dq0"][i, d] - fd) <= 1e-6 * scale + 1e-6 * abs(fd), (i, d, fd)
    # cost of the reverse sweep relative to the forward pass (recompute included)
    print(f"forward {g['forward_ms']:.1f} ms, backward {g['backward_ms']:.1f} ms", flush=True)


def test_adjoint_at_larger_sigma_with_wider_search():
    """sigma = 1.3 puts the cut at 3.9, past the default search length 3.3:
    the system is created with buffer_length 1.5 (search length 4.5) and
    the run matches the dense autograd reference at eps = 1.2, sigma = 1.3
    (rho sigma^3 = 0.70).  With the default search length the call is
    refused."""
    s = W.fcc_state(5, 0.32, 1.0, 6)         # N=500, L=11.6
    print(f"adjoint at sigma 1.3 N={s.n}", flush=True)
    with System.from_state(s, buffer_length=1.5) as g:
        r = g.lj_trajectory(eps=1.2, sigma=1.3, grad=True, **OPTS)
    loss, ge, gs, gq, Pm, Em = loss_and_grads(s.q, s.p, s.L, eps=1.2, sigma=1.3, nsteps=200,
                                              nobs=100, segment=20)
    assert r["loss"] == pytest.approx(loss, rel=1e-10)
    assert r["pressure_mean"] == pytest.approx(Pm, rel=1e-10)
    assert r["pe_mean"] == pytest.approx(Em, rel=1e-10)
    assert r["dloss_deps"] == pytest.approx(ge, rel=1e-8)
    assert r["dloss_dsigma"] == pytest.approx(gs, rel=1e-8)
    assert np.abs(r["dloss_dq0"] - gq).max() / np.abs(gq).max() < 1e-8
    with System.from_state(s) as g:
        with pytest.raises(MdxError) as ei:
            g.lj_trajectory(eps=1.0, sigma=1.2, grad=False, **OPTS)
        assert ei.value.code == -1


def test_plain_potential_matches_kick():
    """At eps = sigma = 1 the differentiable form is the reference potential:
    its trajectory observables match the plain step's."""
    s = W.perturbed(W.fcc_state(8, 0.6, 1.0, 41), 0.2, 42)
    g = gpu_run(s, grad=False, nsteps=40, nobs=1, checkpoint_every=20)
    with System.from_state(s) as p:
        # plain path, rebuilding at the same steps (every 20)
        for _ in range(2):
            p.make_pairlist()
            for _ in range(20):
                p.update_position_half()
                p.refresh_images()
                p.calculate_force()
                p.update_position_half()
        p.refresh_images()
        ke, pe, vir = p.observe()
    V = float(np.prod(s.L))
    P = (2.0 * ke / 3.0 - vir) / V
    assert g["pressure_mean"] == pytest.approx(P, rel=1e-12)
    assert g["pe_mean"] == pytest.approx(pe / s.n, rel=1e-12)


def test_autograd_function_harness():
    import torch
    from mdacp_amd.autograd import LJTrajectoryLoss
    s = W.fcc_state(5, 0.7, 1.0, 5)
    q0 = torch.tensor(s.q, requires_grad=True)
    eps = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    loss = LJTrajectoryLoss.apply(q0, eps, sig, torch.tensor(s.p), s.L,
                                  dict(nsteps=40, nobs=20, checkpoint_every=10))
    (2.0 * loss).backward()
    g = gpu_run(s, nsteps=40, nobs=20, checkpoint_every=10)
    assert float(loss) == pytest.approx(g["loss"], rel=1e-14)
    assert float(eps.grad) == pytest.approx(2.0 * g["dloss_deps"], rel=1e-12)
    assert float(sig.grad) == pytest.approx(2.0 * g["dloss_dsigma"], rel=1e-12)
    np.testing.assert_allclose(q0.grad.numpy(), 2.0 * g["dloss_dq0"], rtol=1e-12, atol=0)


def test_adjoint_deterministic():
    s = W.perturbed(W.fcc_state(6, 0.7, 1.0, 8), 0.1, 9)
    a = gpu_run(s, nsteps=30, nobs=10, checkpoint_every=10)
    b = gpu_run(s, nsteps=30, nobs=10, checkpoint_every=10)
    assert a["dloss_deps"] == b["dloss_deps"] and a["dloss_dsigma"] == b["dloss_dsigma"]
    np.testing.assert_array_equal(a["dloss_dq0"], b["dloss_dq0"])
