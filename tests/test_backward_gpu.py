"""Gradient of the plain multipole integrals on the GPU (N = 15000, 32 pairs,
L = 0..20; FP64 by the O(N) semiseparable recurrences and by the dense Keff
kernel, FP32 dense): against the long-double oracle gradient, against
central finite differences (step 1e-6) of the GPU forward at 64 seeded grid
points, and through torch.autograd.

Tolerances: FP64 |d| <= 1e-12 * m, FP32 |d| <= 2e-6 * m, with m the
cancellation-free magnitude w_i sum_L |c_L| (K_L |w rho|)_i (a CPU emulation of
the FP32 kernel's roundings gives 1.7e-7 * m)."""
import numpy as np
import pytest

from oracle import gradient as OG
from dwba_amd.synthetic import backward_workload

pytestmark = pytest.mark.gpu

N, P, LMAX = 15000, 32, 20


@pytest.fixture(scope="module")
def case():
    import torch
    grid, ra, rb, c = backward_workload(N, P, LMAX, seed=0)
    ga, gb = OG.density_gradients(grid.r, grid.w_simpson, ra, rb, c, LMAX)
    ma, mb = OG.density_gradients(grid.r, grid.w_simpson, ra, rb, c, LMAX, absolute=True)
    dev = torch.device("cuda", 0)
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    return dict(grid=grid, ra=ra, rb=rb, c=c, ga=ga, gb=gb, ma=ma, mb=mb,
                r_d=t(grid.r), w_d=t(grid.w_simpson), ra_d=t(ra), rb_d=t(rb), c_d=t(c))


MODES = [("fp64", "semiseparable", 1e-12), ("fp64", "dense", 1e-12), ("fp32", "dense", 2e-6)]


@pytest.mark.parametrize("precision,contraction,tol", MODES)
def test_gradient_against_oracle(case, precision, contraction, tol):
    import torch
    from dwba_amd.autograd import density_gradients
    print(f"\n[backward] N={N} P={P} L_max={LMAX} precision={precision} contraction={contraction}", flush=True)
    ga, gb = density_gradients(case["ra_d"], case["rb_d"], case["r_d"], case["w_d"], case["c_d"], LMAX, precision,
                               contraction)
    torch.cuda.synchronize()
    ga, gb = ga.cpu().numpy(), gb.cpu().numpy()
    ea = np.abs(ga - case["ga"]) / np.maximum(case["ma"], 1e-300)
    eb = np.abs(gb - case["gb"]) / np.maximum(case["mb"], 1e-300)
    print(f"  max |d|/m: grad_a {ea.max():.2e}  grad_b {eb.max():.2e}", flush=True)
    assert np.all(np.abs(ga - case["ga"]) <= tol * case["ma"] + 1e-300)
    assert np.all(np.abs(gb - case["gb"]) <= tol * case["mb"] + 1e-300)


@pytest.mark.parametrize("precision,contraction,tol", MODES)
def test_autograd_function(case, precision, contraction, tol):
    import torch
    from dwba_amd.autograd import multipole_integrals, density_gradients
    print(f"\n[autograd] N={N} P={P} L_max={LMAX} precision={precision} contraction={contraction}", flush=True)
    ra = case["ra_d"].clone().requires_grad_(True)
    rb = case["rb_d"].clone().requires_grad_(True)
    I = multipole_integrals(ra, rb, case["r_d"], case["w_d"], LMAX, precision, contraction)
    assert I.shape == (P, LMAX + 1)
    loss = (case["c_d"] * I).sum()
    loss.backward()
    ga, gb = density_gradients(case["ra_d"], case["rb_d"], case["r_d"], case["w_d"], case["c_d"], LMAX, precision,
                               contraction)
    torch.cuda.synchronize()
    assert torch.equal(ra.grad, ga) and torch.equal(rb.grad, gb)
    # forward values against the oracle
    ref = OG.plain_forward(case["grid"].r, case["grid"].w_simpson, case["ra"], case["rb"], LMAX)
    mag = OG.plain_forward(case["grid"].r, case["grid"].w_simpson, np.abs(case["ra"]), np.abs(case["rb"]), LMAX)
    got = I.detach().cpu().numpy()
    err = np.abs(got - ref) / mag
    print(f"  forward max |d|/sum|aKb| {err.max():.2e}", flush=True)
    if precision == "fp64":
        assert np.all(np.abs(got - ref) <= 1e-10 * np.abs(ref) + 1e-13 * mag)
    else:
        assert np.all(np.abs(got - ref) <= 1e-6 * mag)


@pytest.mark.parametrize("contraction", ["semiseparable", "dense"])
def test_gradient_against_finite_differences(case, contraction):
    import torch
    from dwba_amd.autograd import multipole_integrals, density_gradients
    print(f"\n[finite differences] N={N} 64 points, step 1e-6, FP64 forward, {contraction}", flush=True)
    ga, gb = density_gradients(case["ra_d"], case["rb_d"], case["r_d"], case["w_d"], case["c_d"], LMAX, "fp64",
                               contraction)
    ga, gb = ga.cpu().numpy(), gb.cpu().numpy()
    r = case["grid"].r
    lo, hi = np.searchsorted(r, 0.1), np.searchsorted(r, 200.0)
    rng = np.random.default_rng(7)
    h = 1e-6
    c = case["c"]
    worst = 0.0
    for k in range(64):
        p = int(rng.integers(0, P))
        i = int(rng.integers(lo, hi))
        which = k % 2
        base_a = case["ra_d"][p:p + 1].clone()
        base_b = case["rb_d"][p:p + 1].clone()
        vals = []
        for s in (+1.0, -1.0):
            a, b = base_a.clone(), base_b.clone()
            (a if which == 0 else b)[0, i] += s * h
            I = multipole_integrals(a, b, case["r_d"], case["w_d"], LMAX, "fp64", contraction)
            vals.append(float((torch.as_tensor(c[p], device=I.device) * I[0]).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        grad = (ga if which == 0 else gb)[p, i]
        # forward rounding ~ 1e-14 of the pair's cancellation-free size S, divided by the step
        S = float(np.sum(np.abs(c[p]) * OG.plain_forward(r, case["grid"].w_simpson, np.abs(case["ra"][p]),
                                                         np.abs(case["rb"][p]), LMAX)[0]))
        tol = 1e-6 * abs(grad) + 1e-14 * S / h
        worst = max(worst, abs(fd - grad) / tol)
        assert abs(fd - grad) <= tol, (k, p, i, which, fd, grad, tol)
    print(f"  worst |fd - grad| / tol = {worst:.2e}", flush=True)


def test_backward_rejects_bad_arguments(case):
    import torch
    from dwba_amd.autograd import density_gradients
    with pytest.raises(ValueError):
        density_gradients(case["ra_d"], case["rb_d"][:, :100], case["r_d"], case["w_d"], case["c_d"], LMAX)
    with pytest.raises(ValueError):
        density_gradients(case["ra_d"], case["rb_d"], case["r_d"], case["w_d"], case["c_d"], 64)
    with pytest.raises(ValueError):
        density_gradients(case["ra_d"], case["rb_d"], case["r_d"], case["w_d"], case["c_d"][:, :3], LMAX)
    with pytest.raises(ValueError):
        density_gradients(case["ra_d"], case["rb_d"], case["r_d"], case["w_d"], case["c_d"], LMAX, "fp64", "fast")
