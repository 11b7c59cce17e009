"""torch.autograd through the GPU solve: gradients of extinction plus EELS with respect to
per-particle Drude parameters (plasma frequency, damping, eps_inf) against central
differences (relative step 1e-6), on a small cluster per coordinate and on the N = 4000 hex
monolayer with quadrupoles along a random direction."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from gpu_helpers import hex_monolayer  # noqa: E402

pytestmark = pytest.mark.gpu

EV_TO_INV_NM = 8065.544 / 1e7  # photon energy in eV -> wavenumber in 1/nm


def _model(op, omega, E0, einc, eps_m=1.0):
    """loss(eps_inf, wp, gamma): summed extinction of the uniform fields E0 plus the EELS sum of
    the beam field einc, from one batched solve of W K + W systems."""
    from cumpm_amd import autograd as A
    W, K = len(omega), len(E0)
    b = torch.cat([A.uniform_rhs(op, E0, W), einc], 0)

    def loss(eps_inf, wp, gamma):
        eps = A.drude(omega, eps_inf, wp, gamma, eps_m)            # (W, N)
        sc = A.self_coefficient(eps)                                # radius 1 after scaling
        rows = torch.cat([sc.repeat_interleave(K, 0), sc], 0)      # (W K + W, N)
        x = op.solve(b, rows)
        p = op.dipoles(x[:W * K]).reshape(W, K, op.N, 3)
        return A.extinction(p, E0, omega, eps_m) + A.eels_loss(x[W * K:], einc, op.N, op.Q)
    return loss


def _params(N, rng):
    dev = torch.device("cuda")
    wp = torch.tensor(1.7 * (1.0 + rng.uniform(-0.05, 0.05, N)), dtype=torch.float64, device=dev)
    gamma = torch.tensor(0.08 * (1.0 + rng.uniform(-0.05, 0.05, N)), dtype=torch.float64, device=dev)
    eps_inf = torch.tensor(4.0 * (1.0 + rng.uniform(-0.05, 0.05, N)), dtype=torch.float64, device=dev)
    return eps_inf, wp, gamma


def test_autograd_small_cluster_matches_central_differences_and_adjoint_module():
    from cumpm_amd import autograd as A
    from cumpm_amd.gradients import extinction_and_wp_gradient
    print("autograd small cluster", flush=True)
    rng = np.random.default_rng(5)
    box = np.array([22.0, 23.0, 24.0])
    pos = rng.uniform(0, 20, (10, 3))
    pos[1] = pos[0] + [1.5, 0.4, 0.3]
    omega = np.array([0.35, 0.55, 0.75])
    E0 = np.eye(3)
    op = A.Operator(box, pos, quadrupoles=[0, 1, 6], tol=1e-13, ewald_tol=1e-8)
    einc = A.ebeam_field(op, omega * 0.1, 0.7, pos[3, :2] + [3.0, 0.0], sigma=0.5)
    loss = _model(op, omega, E0, einc)
    eps_inf, wp, gamma = _params(10, rng)
    params = [t.clone().requires_grad_(True) for t in (eps_inf, wp, gamma)]
    L = loss(*params)
    L.backward()
    grads = [t.grad.detach().cpu().numpy() for t in params]
    assert all(r["converged"] for rs in op.reports for r in rs)
    base = [t.detach() for t in params]
    for which in range(3):
        for i in (0, 1, 6, 9):
            h = 1e-6 * float(base[which][i])
            dp = [t.clone() for t in base]
            dm = [t.clone() for t in base]
            dp[which][i] += h
            dm[which][i] -= h
            with torch.no_grad():
                fd = (loss(*dp) - loss(*dm)).item() / (2 * h)
            g = grads[which][i]
            assert abs(g - fd) < 1e-5 * max(np.abs(grads[which]).max(), 1e-30), (which, i, g, fd)
    # the extinction part alone against the non-torch adjoint module (plasma frequencies)
    E_only = A.uniform_rhs(op, E0, 3)
    wp2 = wp.clone().requires_grad_(True)
    eps = A.drude(omega, eps_inf, wp2, gamma)
    x = op.solve(E_only, A.self_coefficient(eps).repeat_interleave(3, 0))
    Lx = A.extinction(op.dipoles(x).reshape(3, 3, 10, 3), E0, omega)
    Lx.backward()
    # the module takes scalar eps_inf and gamma: compare at uniform values
    eps_inf_u = torch.full_like(eps_inf, 4.0)
    gamma_u = torch.full_like(gamma, 0.08)
    wp3 = wp.clone().requires_grad_(True)
    x3 = op.solve(E_only, A.self_coefficient(A.drude(omega, eps_inf_u, wp3, gamma_u)).repeat_interleave(3, 0))
    L3 = A.extinction(op.dipoles(x3).reshape(3, 3, 10, 3), E0, omega)
    L3.backward()
    L_ref, g_ref, _ = extinction_and_wp_gradient(box, pos, omega=omega, eps_inf=4.0, wp=wp.cpu().numpy(), gamma=0.08,
                                                 quadrupoles=[0, 1, 6], tol=1e-13, ewald_tol=1e-8)
    assert abs(L3.item() - L_ref) < 1e-9 * abs(L_ref)
    np.testing.assert_allclose(wp3.grad.cpu().numpy(), g_ref, rtol=1e-7, atol=1e-9 * np.abs(g_ref).max())


def test_autograd_hex_monolayer_quadrupoles_extinction_plus_eels():
    """The differentiation workload: hex monolayer N = 4000 with quadrupoles, 10 frequencies,
    loss = sum of C_ext over 3 polarisations plus the EELS sum of one beam at 0.7 c passing 2 nm
    outside a particle's surface; per-particle wp perturbed by +-5% (seed 4); the gradient
    with respect to (eps_inf, wp, gamma) along a random direction against the central
    difference with relative step 1e-6 (1e-4)."""
    from cumpm_amd import autograd as A
    print("autograd hex 4000 quads", flush=True)
    pos, box = hex_monolayer(80, 50, 11.0, jitter=0.2, seed=2)
    box = np.array([box[0], box[1], 250.0])
    N = pos.shape[0]
    omega = np.linspace(0.2, 0.8, 10)  # eV
    a = 5.0
    op = A.Operator(box, pos, radius=a, quadrupoles=True, tol=1e-11, ewald_tol=1e-6)
    j = 1234
    epos = pos[j, :2] + np.array([a + 2.0, 0.0])
    einc = A.ebeam_field(op, omega * EV_TO_INV_NM, 0.7, epos)
    loss = _model(op, omega, np.eye(3), einc)
    rng = np.random.default_rng(4)
    eps_inf, wp, gamma = _params(N, rng)
    params = [t.clone().requires_grad_(True) for t in (eps_inf, wp, gamma)]
    L = loss(*params)
    L.backward()
    assert all(r["converged"] for rs in op.reports for r in rs)
    dirs = [torch.tensor(rng.choice([-1.0, 1.0], N), device="cuda") * t.detach() for t in params]
    gdot = sum(float((t.grad * d).sum()) for t, d in zip(params, dirs))
    h = 1e-6
    with torch.no_grad():
        lp = loss(*[t.detach() + h * d for t, d in zip(params, dirs)]).item()
        lm = loss(*[t.detach() - h * d for t, d in zip(params, dirs)]).item()
    fd = (lp - lm) / (2 * h)
    assert abs(gdot - fd) < 1e-4 * abs(fd), (gdot, fd)
