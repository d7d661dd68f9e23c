"""GPU parity of dwba_outer_tail_f64 (Filon/Levin outer tail + analytic tail)
against the oracle's restatement of oscillatory_integrals.py."""
import ctypes

import numpy as np
import pytest

from oracle import tail as T
from dwba_amd import planner

pytestmark = pytest.mark.gpu


def _branch(ka, la, kb, lb, ea, eb, rm, rmax, sign):
    pf = T.PhaseFunctions(ka, la, 0.3, ea, 0.0, kb, lb, 0.7, eb, 0.0)
    mid = (rm + rmax) / 2
    om = abs(pf.dphi(sign, mid))
    h = (np.pi / 4) / om
    return "filon" if abs(pf.d2phi(sign, mid)) * h * h < 0.1 else "levin"


def _gpu_tail(params, moments, kind=1):
    import torch
    from dwba_amd import _hip
    _hip.ensure_init()
    lib = _hip.load_library()
    dev = torch.device("cuda", 0)
    L_max = len(moments) - 1
    mom = torch.as_tensor(np.asarray(moments, dtype=np.float64), device=dev)
    out = torch.full((L_max + 1,), np.nan, dtype=torch.float64, device=dev)
    p = params
    tj = (_hip.TailJob * 1)()
    tj[0] = _hip.TailJob(p["ka"], p["da"], p["ea"], p["sa"], p["kb"], p["db"], p["eb"], p["sb"], p["la"], p["lb"],
                         p["rm"], p["rmax"], _hip.ptr(mom), _hip.ptr(out), L_max, kind)
    tj_d = _hip.struct_array_to_device(tj, dev)
    D, x = planner.levin_tables(8)
    seg, dmax = planner.filon_limits()
    Dp, Da = _hip.f64_array(D.ravel())
    Xp, Xa = _hip.f64_array(x)
    wsb = lib.dwba_outer_tail_workspace_bytes(1, L_max)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    _hip.check(lib.dwba_outer_tail_f64(1, _hip.ptr(tj_d), L_max, seg, dmax, Dp, Xp, _hip.ptr(ws), wsb,
                                       _hip.stream_handle()), "tail")
    return out.cpu().numpy()


def _oracle_noise(p, moments):
    """Conditioning noise of the reference's Levin segments (oracle estimate)."""
    out = np.zeros(len(moments))
    for L in range(1, len(moments)):
        if abs(moments[L]) > 1e-12:
            out[L] = T.outer_integral_noise(T.multipole_envelope(moments[L], L), p["ka"], p["la"], p["da"], p["ea"],
                                            p["sa"], p["kb"], p["lb"], p["db"], p["eb"], p["sb"], p["rm"], p["rmax"])
    return out


def _oracle_tail(p, moments, kind=1):
    out = np.zeros(len(moments))
    for L in range(1, len(moments)):
        m = moments[L]
        if abs(m) <= 1e-12:
            continue
        if kind == 2:
            out[L] = T.analytical_multipole_tail(p["rm"], p["ka"], p["kb"], p["da"], p["db"], p["la"], p["lb"], L,
                                                 m, p["ea"], p["eb"], p["sa"], p["sb"])
            continue
        env = T.multipole_envelope(m, L)
        v = T.dwba_outer_integral(env, p["ka"], p["la"], p["da"], p["ea"], p["sa"], p["kb"], p["lb"], p["db"],
                                  p["eb"], p["sb"], p["rm"], p["rmax"], np.pi / 4)
        v += T.analytical_multipole_tail(p["rmax"], p["ka"], p["kb"], p["da"], p["db"], p["la"], p["lb"], L,
                                         bound_overlap=m, eta_i=p["ea"], eta_f=p["eb"], sigma_i=p["sa"],
                                         sigma_f=p["sb"])
        out[L] = v
    return out


CASES = {
    # neutral, well separated k: Filon for both phases
    "filon": dict(ka=2.1, la=3, kb=1.3, lb=4, rm=15.0, rmax=200.0, ea=0.0, eb=0.0),
    # equal k, different l: Phi_-' is the centrifugal term only, Levin on Phi_-
    "levin": dict(ka=0.5, la=10, kb=0.5, lb=5, rm=5.0, rmax=55.0, ea=0.0, eb=0.0),
    # Coulomb phases
    "coulomb": dict(ka=1.7, la=2, kb=1.1, lb=5, rm=20.0, rmax=300.0, ea=-0.6, eb=-0.9),
    # fast oscillation over a long range: segment cap + node up-scaling
    "capped": dict(ka=9.0, la=1, kb=8.2, lb=2, rm=10.0, rmax=1000.0, ea=0.0, eb=0.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_outer_tail_matches_oracle(name):
    p = dict(CASES[name], da=0.3, db=0.7, sa=0.1 if CASES[name]["ea"] else 0.0, sb=-0.2 if CASES[name]["eb"] else 0.0)
    if name == "levin":
        assert _branch(p["ka"], p["la"], p["kb"], p["lb"], p["ea"], p["eb"], p["rm"], p["rmax"], -1) == "levin"
    moments = np.array([0.0, 0.8, -0.5, 0.3, 1e-13, 0.05] + [0.02 * (-1) ** k for k in range(20)])
    got = _gpu_tail(p, moments)
    ref = _oracle_tail(p, moments)
    scale = np.max(np.abs(ref)) + 1e-300
    tol = 1e-10 * np.abs(ref) + 1e-13 * scale + 10.0 * _oracle_noise(p, moments)
    assert np.all(np.abs(got - ref) <= tol), (got, ref, np.abs(got - ref) / tol)
    assert got[0] == 0.0 and got[4] == 0.0
    assert len(got) == 26


def test_legacy_analytic_tail():
    p = dict(CASES["coulomb"], da=0.3, db=0.7, sa=0.1, sb=-0.2)
    moments = np.array([0.0, 0.8, -0.5, 0.3])
    np.testing.assert_allclose(_gpu_tail(p, moments, kind=2), _oracle_tail(p, moments, kind=2), rtol=1e-12)


QUAD_EPSABS = 1.49e-8    # scipy.integrate.quad's default absolute (and relative) tolerance


def test_non_oscillatory_branch():
    """Same k, same l, no Coulomb phase: Phi_-' vanishes identically, so the
    reference integrates Phi_- with scipy quad at its default tolerances
    (oscillatory_integrals.py:1418-1419) and the library with 512 panels of
    8-point Gauss-Legendre; Phi_+ goes through Filon.  Phi_- is the constant
    c = delta_a - delta_b there, so that half is exactly
    M e^{ic} (r_m^-L - r_max^-L) / L.  The library must equal the oracle with
    that exact value in place of quad's to 1e-10, and the oracle itself
    (quad's value) to within quad's stated tolerance."""
    p = dict(ka=1.5, la=3, kb=1.5, lb=3, rm=15.0, rmax=200.0, ea=0.0, eb=0.0, da=0.3, db=0.7, sa=0.0, sb=0.0)
    pf = T.PhaseFunctions(p["ka"], p["la"], p["da"], 0.0, 0.0, p["kb"], p["lb"], p["db"], 0.0, 0.0)
    assert abs(pf.dphi(-1, 0.5 * (p["rm"] + p["rmax"]))) < 1e-10
    moments = np.array([0.0, 0.8, -0.5, 0.3, 1e-13, 0.05] + [0.02 * (-1) ** k for k in range(20)])
    got = _gpu_tail(p, moments)
    ref_quad = _oracle_tail(p, moments)
    ref_exact = ref_quad.copy()
    for L in range(1, len(moments)):
        m = moments[L]
        if abs(m) <= 1e-12:
            continue
        env = T.multipole_envelope(m, L)
        q = T.outer_integral_oscillatory(env, lambda r: pf.phi(-1, r), lambda r: pf.dphi(-1, r),
                                         lambda r: pf.d2phi(-1, r), p["rm"], p["rmax"])
        exact = m * np.exp(1j * (p["da"] - p["db"])) * (p["rm"] ** -L - p["rmax"] ** -L) / L
        ref_exact[L] += 0.5 * (exact.real - q.real)
    scale = np.max(np.abs(ref_exact)) + 1e-300
    assert np.all(np.abs(got - ref_exact) <= 1e-10 * np.abs(ref_exact) + 1e-13 * scale), \
        (got, ref_exact, np.abs(got - ref_exact))
    assert np.all(np.abs(got - ref_quad) <= 0.5 * QUAD_EPSABS * np.maximum(1.0, np.abs(ref_quad))), \
        np.abs(got - ref_quad)
    assert got[0] == 0.0 and got[4] == 0.0
