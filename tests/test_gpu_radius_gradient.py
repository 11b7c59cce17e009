"""Radius gradients of the polydisperse operator on the GPU: mpm_op_radius_gradient against
central differences of mpm_op_apply (operators rebuilt at a_i +- h on the same grid and
support), and torch.autograd through PolyOperator (extinction plus EELS with respect to
per-particle radii and Drude parameters) against central differences of the loss (relative
step 1e-6).  No counterpart in the reference (its solvers return dipoles only); the forward
operator is the reference's polydisperse one (polydisperse_ewald_electric_field.cu)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cumpm_amd import _capi as C  # noqa: E402
from gpu_helpers import Op  # noqa: E402

pytestmark = pytest.mark.gpu


def _cluster(N, L, rmin, rmax, seed, gap=0.2):
    """Random sequential addition of N spheres with radii U(rmin, rmax) in a periodic cube."""
    rng = np.random.default_rng(seed)
    pos, rad = [], []
    while len(pos) < N:
        p = rng.uniform(0, L, 3)
        a = rng.uniform(rmin, rmax)
        ok = True
        for q, b in zip(pos, rad):
            d = p - q
            d -= L * np.round(d / L)
            if np.linalg.norm(d) < a + b + gap:
                ok = False
                break
        if ok:
            pos.append(p)
            rad.append(a)
    return np.array(pos), np.array(rad)


def _grad(op, x, g):
    B = x.shape[0]
    xb, gb = C.DeviceBuffer(x.nbytes), C.DeviceBuffer(g.nbytes)
    xb.upload(np.ascontiguousarray(x))
    gb.upload(np.ascontiguousarray(g))
    out = C.DeviceBuffer(8 * op.N)
    C.check(C.load().mpm_op_radius_gradient(op.h, xb.ptr, gb.ptr, B, out.ptr, 0, None))
    o = np.zeros(op.N)
    out.download(o)
    return o


@pytest.mark.parametrize("N,bar", [(80, 1e-6), (12, 1e-5)])
def test_operator_radius_derivative_matches_differences(N, bar):
    """N = 80 distinct radii: pair functions at each pair's own radii; N = 12: the radius-class
    tables (their interpolation is the difference allowed by the looser bar)."""
    print(f"radius derivative N={N}", flush=True)
    L = 20.0 if N == 80 else 16.0
    pos, rad = _cluster(N, L, 0.8, 1.3, seed=N)
    box = [L, L, L]
    op = Op(box, 1e-6, 0.5, radii=rad)
    op.set_positions(pos)
    sc = np.full(N, -0.2 + 0.05j)
    op.set_self_coef(sc)
    info = op.info()
    grid, P = tuple(info.grid), int(info.support)
    rng = np.random.default_rng(3)
    B = 3
    x = rng.standard_normal((B, N, 3)) + 1j * rng.standard_normal((B, N, 3))
    g = rng.standard_normal((B, N, 3)) + 1j * rng.standard_normal((B, N, 3))
    got = _grad(op, x, g)
    assert np.all(np.isfinite(got))
    for i in (0, 5, N - 1):
        h = 1e-5 * rad[i]
        ys = []
        for sgn in (1, -1):
            r = rad.copy()
            r[i] += sgn * h
            o = Op(box, 1e-6, 0.5, grid=grid, support=P, radii=r)
            o.set_positions(pos)
            o.set_self_coef(sc)
            ys.append(o.apply(x))
        fd = np.real(np.sum(np.conj(g) * (ys[0] - ys[1]))) / (2 * h)
        assert abs(got[i] - fd) < bar * np.abs(got).max(), (i, got[i], fd)
    # accumulate adds; the same call twice is deterministic
    xb, gb = C.DeviceBuffer(x.nbytes), C.DeviceBuffer(g.nbytes)
    xb.upload(x)
    gb.upload(g)
    out = C.DeviceBuffer(8 * N)
    C.check(C.load().mpm_op_radius_gradient(op.h, xb.ptr, gb.ptr, B, out.ptr, 0, None))
    C.check(C.load().mpm_op_radius_gradient(op.h, xb.ptr, gb.ptr, B, out.ptr, 1, None))
    o2 = np.zeros(N)
    out.download(o2)
    np.testing.assert_array_equal(o2, 2 * got)


def test_radius_gradient_rejects_other_operators():
    op = Op([20.0] * 3, 1e-6, 0.5)
    op.set_positions(np.random.default_rng(1).uniform(0, 20, (10, 3)))
    x = np.zeros((1, 10, 3), complex)
    xb = C.DeviceBuffer(x.nbytes)
    out = C.DeviceBuffer(80)
    assert C.load().mpm_op_radius_gradient(op.h, xb.ptr, xb.ptr, 1, out.ptr, 0, None) == C.MPM_ERR_UNSUPPORTED


def test_autograd_radii_and_drude_match_central_differences():
    """Loss = extinction over 3 frequencies x 3 polarisations + the EELS sum of one beam, of a
    polydisperse cluster (N = 70, radii U(0.8, 1.3) nm, lengths in nm); gradient with respect
    to (radii, eps_inf, wp, gamma) per particle against central differences with relative
    step 1e-6, per coordinate for a few particles and along a random direction."""
    from cumpm_amd import autograd as A
    print("autograd radii", flush=True)
    pos, rad = _cluster(70, 20.0, 0.8, 1.3, seed=9)
    omega = np.array([0.45, 0.6, 0.75])
    E0 = np.eye(3)
    op = A.PolyOperator([20.0] * 3, pos, rad, tol=1e-13, ewald_tol=1e-6)
    einc = A.ebeam_field(op, omega * 0.1, 0.7, pos[7, :2] + [rad[7] + 2.0, 0.0], sigma=0.5)
    W, K, N = len(omega), 3, op.N
    b = torch.cat([A.uniform_rhs(op, E0, W), einc], 0)

    def loss(radii, eps_inf, wp, gamma):
        eps = A.drude(omega, eps_inf, wp, gamma)
        sc = A.self_coefficient(eps, radii[None, :])
        rows = torch.cat([sc.repeat_interleave(K, 0), sc], 0)
        x = op.solve(b, rows, radii)
        p = op.dipoles(x[:W * K]).reshape(W, K, N, 3)
        return A.extinction(p, E0, omega) + A.eels_loss(x[W * K:], einc, N, 0)

    rng = np.random.default_rng(4)
    dev = torch.device("cuda")
    base = [torch.tensor(rad, dtype=torch.float64, device=dev),
            torch.tensor(4.0 * (1 + rng.uniform(-0.05, 0.05, N)), dtype=torch.float64, device=dev),
            torch.tensor(1.7 * (1 + rng.uniform(-0.05, 0.05, N)), dtype=torch.float64, device=dev),
            torch.tensor(0.08 * (1 + rng.uniform(-0.05, 0.05, N)), dtype=torch.float64, device=dev)]
    params = [t.clone().requires_grad_(True) for t in base]
    Lv = loss(*params)
    Lv.backward()
    grads = [t.grad.detach().cpu().numpy() for t in params]
    assert all(r["converged"] for rs in op.reports for r in rs)
    for i in (0, 7, 33, 69):
        h = 1e-6 * float(base[0][i])
        dp, dm = [t.clone() for t in base], [t.clone() for t in base]
        dp[0][i] += h
        dm[0][i] -= h
        with torch.no_grad():
            fd = (loss(*dp) - loss(*dm)).item() / (2 * h)
        assert abs(grads[0][i] - fd) < 2e-5 * np.abs(grads[0]).max(), (i, grads[0][i], fd)
    dirs = [torch.tensor(rng.choice([-1.0, 1.0], N), device=dev) * t for t in base]
    gdot = sum(float((t.grad * d).sum()) for t, d in zip(params, dirs))
    h = 1e-6
    with torch.no_grad():
        lp = loss(*[t + h * d for t, d in zip(base, dirs)]).item()
        lm = loss(*[t - h * d for t, d in zip(base, dirs)]).item()
    fd = (lp - lm) / (2 * h)
    assert abs(gdot - fd) < 1e-5 * abs(fd), (gdot, fd)
