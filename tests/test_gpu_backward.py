"""SVD adjoint kernel (jsvd_plan_backward_*) and the torch.autograd binding,
checked against FP64 central finite differences and against
torch.linalg.svd's gradient (CPU, FP64).

Workload shape: 96 x 64 matrices with sigma_k = 10^(-2k/63) (consecutive
relative gaps 7%), Haar factors, seeded upstream gradients dU, dS, dV ~ N(0,1).
The loss L = <dU, U> + <dS, S> + <dV, V> depends on the gauge (column signs
and order), so every perturbed or foreign solve is aligned to the reference
solve first: columns matched by singular value, signs by <u_j, u_j_ref>.

Tolerances: FD (h = 1e-6, directional, FP64) relative error <= 1e-6;
torch.linalg.svd gradient relative Frobenius error <= 1e-9 (FP64, every one
of the 512 matrices of the workload).  FP32 (input, factors and gradient in
float32; the adjoint kernel evaluates in FP64 and rounds once): against the
FP64 kernel on the same FP32-rounded factors <= 1e-6 (only the final
rounding differs: 6e-8), and against torch's FP64 gradient of the same
FP32-rounded input <= 1e-5.  The second bound: the factors carry FP32
rounding (6e-8 relative), which the adjoint amplifies by about the inverse
relative gap of neighbouring singular values (1 / 0.07 here); ~1e-6 is the
expected error (measured: 7e-7 at most over the 512 matrices).
"""
import sys

import numpy as np
import pytest
import torch

import jsvd
from jsvd import autograd as jag
from jsvd import gen

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _align(u, s, v, u_ref, s_ref):
    perm = np.array([int(np.argmin(np.abs(s - x))) for x in s_ref])
    u2, s2, v2 = u[:, perm], s[perm], v[:, perm]
    sign = np.sign(np.sum(u2 * u_ref, axis=0))
    sign[sign == 0] = 1.0
    return u2 * sign, s2, v2 * sign


def _loss(u, s, v, du, ds, dv):
    return float(np.sum(du * u) + np.sum(ds * s) + np.sum(dv * v))


def _batch(count, seed=20260515):
    mats, sig = gen.gapped_batch(count, seed=seed)
    rng = np.random.default_rng(seed + 1)
    du = [rng.standard_normal((96, 64)) for _ in range(count)]
    ds = [rng.standard_normal(64) for _ in range(count)]
    dv = [rng.standard_normal((64, 64)) for _ in range(count)]
    return mats, du, ds, dv


def _torch_grads(mats, du, ds, dv, u_ref, s_ref):
    """torch.linalg.svd's FP64 CPU gradient of the loss, batched, with
    torch's factors aligned to (u_ref, s_ref) column by column (the gauge the
    loss depends on)."""
    a = torch.tensor(np.stack(mats), dtype=torch.float64, requires_grad=True)
    ut, st, vht = torch.linalg.svd(a, full_matrices=False)
    vt = vht.transpose(1, 2)
    sn, un = st.detach().numpy(), ut.detach().numpy()
    perm = np.stack([np.array([int(np.argmin(np.abs(sn[k] - x))) for x in s_ref[k]]) for k in range(len(mats))])
    idx = torch.tensor(perm)
    up = torch.gather(ut, 2, idx[:, None, :].expand(-1, ut.shape[1], -1))
    vp = torch.gather(vt, 2, idx[:, None, :].expand(-1, vt.shape[1], -1))
    sp = torch.gather(st, 1, idx)
    sign = np.sign(np.sum(np.take_along_axis(un, perm[:, None, :], axis=2) * u_ref, axis=1))
    sign[sign == 0] = 1.0
    sg = torch.tensor(sign)[:, None, :]
    loss = (torch.sum(torch.tensor(np.stack(du)) * up * sg) + torch.sum(torch.tensor(np.stack(ds)) * sp)
            + torch.sum(torch.tensor(np.stack(dv)) * vp * sg))
    loss.backward()
    return a.grad.numpy()


def _grad_gpu(mats, du, ds, dv, dtype=torch.float64):
    dev = torch.device("cuda:0")
    a = torch.tensor(np.stack(mats), dtype=dtype, device=dev, requires_grad=True)
    u, s, v = jag.svd(a)
    loss = (torch.sum(torch.tensor(np.stack(du), dtype=dtype, device=dev) * u)
            + torch.sum(torch.tensor(np.stack(ds), dtype=dtype, device=dev) * s)
            + torch.sum(torch.tensor(np.stack(dv), dtype=dtype, device=dev) * v))
    loss.backward()
    return (a.grad.detach().cpu().double().numpy(), u.detach().cpu().double().numpy(),
            s.detach().cpu().double().numpy(), v.detach().cpu().double().numpy())


def test_backward_against_central_differences():
    _note("backward vs central FD, 4 matrices x 2 directions")
    mats, du, ds, dv = _batch(4)
    ga, u, s, v = _grad_gpu(mats, du, ds, dv)
    rng = np.random.default_rng(7)
    h = 1e-6
    for k in range(len(mats)):
        for _ in range(2):
            e = rng.standard_normal(mats[k].shape)
            e /= np.linalg.norm(e)
            outs = jsvd.svd_batched([mats[k] + h * e, mats[k] - h * e], epsilon=1e-15, max_sweeps=30)
            lp = _loss(*_align(outs[0][0], outs[0][1], outs[0][2], u[k], s[k]), du[k], ds[k], dv[k])
            lm = _loss(*_align(outs[1][0], outs[1][1], outs[1][2], u[k], s[k]), du[k], ds[k], dv[k])
            fd = (lp - lm) / (2 * h)
            an = float(np.sum(ga[k] * e))
            rel = abs(fd - an) / max(abs(an), 1e-3 * np.linalg.norm(ga[k]))
            _note(f"  matrix {k}: fd {fd:.10e} analytic {an:.10e} rel {rel:.2e}")
            assert rel <= 1e-6


def test_backward_against_torch_linalg_svd():
    _note("backward vs torch.linalg.svd gradient, 8 matrices")
    mats, du, ds, dv = _batch(8, seed=11)
    ga, u, s, v = _grad_gpu(mats, du, ds, dv)
    for k in range(len(mats)):
        a = torch.tensor(mats[k], dtype=torch.float64, requires_grad=True)
        ut, st, vht = torch.linalg.svd(a, full_matrices=False)
        vt = vht.T
        un, sn = ut.detach().numpy(), st.detach().numpy()
        perm = np.array([int(np.argmin(np.abs(sn - x))) for x in s[k]])
        sign = np.sign(np.sum(un[:, perm] * u[k], axis=0))
        sign_t = torch.tensor(sign)
        loss = (torch.sum(torch.tensor(du[k]) * ut[:, perm] * sign_t) + torch.sum(torch.tensor(ds[k]) * st[perm])
                + torch.sum(torch.tensor(dv[k]) * vt[:, perm] * sign_t))
        loss.backward()
        g = a.grad.numpy()
        rel = np.linalg.norm(ga[k] - g) / np.linalg.norm(g)
        _note(f"  matrix {k}: relative difference {rel:.2e}")
        assert rel <= 1e-9


def test_backward_full_batch_fp64_and_fp32():
    # The 512-matrix workload in FP64 and FP32.  FP64: every matrix's
    # gradient against torch.linalg.svd's.  FP32: the FP32 adjoint kernel
    # against the FP64 kernel applied to the same FP32-rounded factors, and
    # against torch's FP64 gradient.
    _note("backward 512 matrices, fp64 and fp32")
    mats, du, ds, dv = _batch(512, seed=5)
    ga64, u, s, v = _grad_gpu(mats, du, ds, dv, torch.float64)
    g_torch = _torch_grads(mats, du, ds, dv, u, s)
    rel64 = (np.linalg.norm((ga64 - g_torch).reshape(512, -1), axis=1)
             / np.linalg.norm(g_torch.reshape(512, -1), axis=1))
    _note(f"  fp64 vs torch.linalg.svd, 512 matrices: max relative {rel64.max():.2e}, median {np.median(rel64):.2e}")
    assert rel64.max() <= 1e-9
    ga32, u32, s32, v32 = _grad_gpu(mats, du, ds, dv, torch.float32)
    # torch on the same FP32-rounded input, aligned to the FP32 run's own
    # column order and signs (rounding the input may change both, and the
    # loss depends on them).
    mats32 = [x.astype(np.float32).astype(np.float64) for x in mats]
    g_torch32 = _torch_grads(mats32, du, ds, dv, u32, s32)
    rel_t = (np.linalg.norm((ga32 - g_torch32).reshape(512, -1), axis=1)
             / np.linalg.norm(g_torch32.reshape(512, -1), axis=1))
    _note(f"  fp32 vs torch.linalg.svd (fp64): max relative {rel_t.max():.2e}, median {np.median(rel_t):.2e}")
    assert rel_t.max() <= 1e-5
    plan = jsvd.BatchPlan([96] * 512, [64] * 512)
    dev = torch.device("cuda:0")
    t = lambda x: torch.tensor(np.stack(x) if isinstance(x, list) else x, dtype=torch.float64, device=dev).reshape(-1)
    g_ref = torch.empty(512 * 96 * 64, dtype=torch.float64, device=dev)
    plan.backward(t(u32), t(s32), t(v32), t(du), t(ds), t(dv), g_ref)
    g_ref = g_ref.cpu().numpy().reshape(512, 96, 64)
    rel = np.linalg.norm((ga32 - g_ref).reshape(512, -1), axis=1) / np.linalg.norm(g_ref.reshape(512, -1), axis=1)
    _note(f"  fp32 vs fp64 kernel on the same factors: max relative {rel.max():.2e}")
    assert rel.max() <= 1e-6
    plan.close()


def test_backward_partial_gradients():
    # Upstream gradients that are None (only S used): gA = U diag(dS) V^T.
    _note("backward with only dS")
    mats, _, ds, _ = _batch(4, seed=3)
    dev = torch.device("cuda:0")
    a = torch.tensor(np.stack(mats), dtype=torch.float64, device=dev, requires_grad=True)
    u, s, v = jag.svd(a)
    (torch.tensor(np.stack(ds), device=dev) * s).sum().backward()
    g = a.grad.cpu().numpy()
    un, vn = u.detach().cpu().numpy(), v.detach().cpu().numpy()
    for k in range(4):
        expect = (un[k] * np.array(ds[k])) @ vn[k].T
        assert np.linalg.norm(g[k] - expect) <= 1e-12 * np.linalg.norm(expect)
