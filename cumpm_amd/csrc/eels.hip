// Electron energy-loss (EELS) driver: the relativistic field of a fast electron as the
// incident field, self-consistent dipoles (and quadrupoles) from the batched solver, and
// the loss probability from the reciprocity sum.
//
// Restates src/cpp/eels_solver.cu:241-880 (set_dims :241-279, nondimensionalize
// :281-303, guesses :316-436, compute :438-880) and the Python layer's argument handling
// (src/cuMPM/eels.py:21-609, split sub-dipoles excepted: they stay in Python).  The
// incident field of every (frequency, particle) is one thread of ebeam_kernel on the
// device (the reference evaluates it on the host); solves go through solve_batched, in
// the reference's order one system at a time (batched = 0) or every frequency of a
// Bloch point in one batched launch (batched = 1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>

#include "mpm_internal.h"

namespace mpm {
namespace {

constexpr double kPiE = 3.14159265358979323846;

// Modified Bessel functions K0(x) and K1(x) for x > 0 by the trapezoidal rule on
// K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt: the integrand extends to an even
// function analytic in the strip |Im t| < pi/2, so equal steps converge geometrically
// (L. N. Trefethen and J. A. C. Weideman, "The exponentially convergent trapezoidal rule",
// SIAM Review 56 (2014) 385-458).  Step 0.2, or 0.5/sqrt(x) where the integrand's width
// sqrt(1/x) is narrower; summation stops once the terms have passed their peak and fall
// below 1e-17 of the sum.  Relative error below 4e-14 against a library evaluation for
// 1e-6 <= x <= 300 (beyond, exp(-x) is at the bottom of the double range).
__host__ __device__ inline void bessel_k01(double x, double& k0, double& k1) {
    const double h = fmin(0.2, 0.5 / sqrt(x));
    double e = exp(-x);
    double s0 = 0.5 * e, s1 = 0.5 * e;
    for (int k = 1; k < 400; ++k) {
        const double ch = cosh(k * h);
        e = exp(-x * ch);
        s0 += e;
        s1 += e * ch;
        if (x * ch > 1.0 && e * ch < 1e-17 * s1) break;
    }
    k0 = h * s0;
    k1 = h * s1;
}

struct BeamGeom {
    double ex, ey;          // beam position in the xy-plane (nondimensional)
    double v, gamma;        // speed / c and Lorentz factor 1/sqrt(1 - eps_m v^2)
    double sigma;           // probe half-width 0.5 nm, nondimensional (eels_solver.cu:616)
    double kx, ky;          // Bloch wave vector of the aperiodic-sampling point (0 without)
    double Lx, Ly;          // box (image translations)
    int N, Q, W;
    int64_t n;              // unknowns per system, 3N + 5Q
};

// Field of the beam at (px, py, pz) for one image translation (Rx, Ry): the space-frequency
// field of a charge moving along z (eels_solver.cu:610-640), E_perp = -A K1(xi) r/r_s,
// E_z = A (i / gamma) K0(xi), A = w / (v^2 gamma) e^{2 pi i w z / v}, xi = 2 pi w r_s /
// (v gamma), r_s = sqrt(r_perp^2 + sigma^2) (the softened distance).
__device__ inline void beam_field(double px, double py, double pz, double w, const BeamGeom& g, double Rx,
                                  double Ry, double E[3][2]) {
    const double dx = px - g.ex - Rx, dy = py - g.ey - Ry;
    const double rs = sqrt(dx * dx + dy * dy + g.sigma * g.sigma);
    const double xi = 2.0 * kPiE * w * rs / (g.v * g.gamma);
    double k0, k1;
    bessel_k01(xi, k0, k1);
    const double amp = w / (g.v * g.v * g.gamma);
    double sph, cph;
    sincos(2.0 * kPiE * w * pz / g.v, &sph, &cph);
    const double ar = amp * cph, ai = amp * sph;
    E[0][0] = -ar * k1 * dx / rs;
    E[0][1] = -ai * k1 * dx / rs;
    E[1][0] = -ar * k1 * dy / rs;
    E[1][1] = -ai * k1 * dy / rs;
    E[2][0] = -ai * k0 / g.gamma;  // A i / gamma
    E[2][1] = ar * k0 / g.gamma;
}

// One thread per (frequency w, row item): items < N are particles (three field rows),
// items >= N quadrupoles (five gradient rows: the volume-averaged gradient <d_i E_j> of the
// beam field by a 24-point surface rule on the unit sphere, 3/(24 R) sum_k E(x + R n_k) n_k,
// R = 1, in the components (Gxx - Gzz, Gxy, Gxz, Gyy - Gzz, Gyz) of the symmetric part;
// eels_solver.cu:646-706).  The surface points n_k are the Fibonacci sphere z_k = 1 -
// (2k + 1)/24, azimuth -(k + 1/2) pi (3 - sqrt 5) (the reference's tabulated normals,
// eels_solver.cu:1-27).  With the aperiodic sampling (asm) the field sums the images
// |R_x| <= rmax[w][0] L_x, |R_y| <= rmax[w][1] L_y with the phase e^{i k.(x - R)}.
__global__ __launch_bounds__(256) void ebeam_kernel(const double* __restrict__ pos, const double* __restrict__ omega,
                                                    const int* __restrict__ rmax, const int* __restrict__ qown,
                                                    BeamGeom g, double2* __restrict__ out) {
    const int items = g.N + g.Q;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)items * g.W) return;
    const int w = (int)(t / items), it = (int)(t % items);
    const bool quad = it >= g.N;
    const int p = quad ? qown[it - g.N] : it;
    const double px = pos[p], py = pos[(size_t)g.N + p], pz = pos[2 * (size_t)g.N + p];
    const double om = omega[w];
    double acc[5][2];
    for (int c = 0; c < 5; ++c) acc[c][0] = acc[c][1] = 0.0;
    const int rx = rmax[2 * w], ry = rmax[2 * w + 1];
    for (int ix = -rx; ix <= rx; ++ix)
        for (int iy = -ry; iy <= ry; ++iy) {
            const double Rx = ix * g.Lx, Ry = iy * g.Ly;
            double fr = 1.0, fi = 0.0;
            if (g.kx != 0.0 || g.ky != 0.0) sincos(g.kx * (px - Rx) + g.ky * (py - Ry), &fi, &fr);
            if (!quad) {
                double E[3][2];
                beam_field(px, py, pz, om, g, Rx, Ry, E);
                for (int c = 0; c < 3; ++c) {
                    acc[c][0] += E[c][0] * fr - E[c][1] * fi;
                    acc[c][1] += E[c][0] * fi + E[c][1] * fr;
                }
            } else {
                double T[3][3][2];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) T[i][j][0] = T[i][j][1] = 0.0;
                const double ga = kPiE * (3.0 - sqrt(5.0));
                for (int k = 0; k < 24; ++k) {
                    const double nz = 1.0 - (2.0 * k + 1.0) / 24.0;
                    const double rr = sqrt(1.0 - nz * nz);
                    double sa, ca;
                    sincos(-(k + 0.5) * ga, &sa, &ca);
                    const double n[3] = {rr * ca, rr * sa, nz};
                    double E[3][2];
                    beam_field(px + n[0], py + n[1], pz + n[2], om, g, Rx, Ry, E);
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            T[i][j][0] += E[j][0] * n[i];
                            T[i][j][1] += E[j][1] * n[i];
                        }
                }
                const double s = 3.0 / 24.0;
                double Gc[5][2];
                for (int h = 0; h < 2; ++h) {
                    Gc[0][h] = s * (T[0][0][h] - T[2][2][h]);
                    Gc[1][h] = s * 0.5 * (T[0][1][h] + T[1][0][h]);
                    Gc[2][h] = s * 0.5 * (T[0][2][h] + T[2][0][h]);
                    Gc[3][h] = s * (T[1][1][h] - T[2][2][h]);
                    Gc[4][h] = s * 0.5 * (T[1][2][h] + T[2][1][h]);
                }
                for (int c = 0; c < 5; ++c) {
                    acc[c][0] += Gc[c][0] * fr - Gc[c][1] * fi;
                    acc[c][1] += Gc[c][0] * fi + Gc[c][1] * fr;
                }
            }
        }
    double2* o = out + (size_t)w * g.n;
    if (!quad) {
        for (int c = 0; c < 3; ++c) o[3 * (size_t)p + c] = make_double2(acc[c][0], acc[c][1]);
    } else {
        const size_t base = 3 * (size_t)g.N + 5 * (size_t)(it - g.N);
        for (int c = 0; c < 5; ++c) o[base + c] = make_double2(acc[c][0], acc[c][1]);
    }
}

}  // namespace
}  // namespace mpm

using mpm::cplx;

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------
struct mpm_eels_solver {
    // configuration (physical until nondimensionalised)
    std::vector<double> box, omega, radius;
    std::vector<std::vector<cplx>> eps_p;  // [W][cols]
    double v = 0.5, eps_m = 1.0, xi = 0.5, tol = 1e-3, integration_step = 0.01;
    bool quiet = false;
    int guess = MPM_GUESS_DERIVATIVE, method = MPM_GMRES, field_type = MPM_FIELD_AUTO;
    bool solve_quads = false;
    std::vector<int> quad_idxs;
    bool asm_flag = false;
    int asm_nx = 1, asm_ny = 1;
    int precision = MPM_PREC_AUTO;
    bool refined = false, batched = false, use_jacobi = true;
    int restart = 50, maxiter = 1000;
    // state
    size_t N = 0, W = 0, Q = 0, frames = 0;
    double length_scale = 1.0, eps_scale = 1.0, vol_frac = 0.0;
    std::unique_ptr<mpm::EwaldOp> op;
    std::vector<double> eels;         // [F][W]
    std::vector<cplx> dips;           // [F][W][N][3]
    std::vector<cplx> quads;          // [F][W][Q][5]
    std::vector<mpm_solve_report> reports;  // [F][kpoints][W]
    std::vector<cplx> last_einc;      // [W][n] of the last frame's last Bloch point
    double solve_ms = 0, setup_ms = 0, total_ms = 0;
    mpm::DevBuf<double> d_pos, d_omega;
    mpm::DevBuf<int> d_rmax, d_qown;
    mpm::DevBuf<double2> d_einc, d_b, d_x;
    int indent = 0;

    void print(const std::string& m) const {
        if (!quiet) std::cout << std::string(indent * 2, ' ') << m << std::endl;
    }
    size_t n_unknowns() const { return 3 * N + 5 * Q; }
    size_t quad_particle(size_t q) const { return quad_idxs.empty() ? q : (size_t)quad_idxs[q]; }

    // eels_solver.cu:241-279
    void set_dims(size_t n) {
        N = n;
        Q = solve_quads ? (quad_idxs.empty() ? N : quad_idxs.size()) : 0;
        if (radius.empty())
            radius.assign(N, 1.0);
        else if (radius.size() == 1)
            radius.assign(N, radius[0]);
        else if (radius.size() != N)
            throw mpm::Error(MPM_ERR_RUNTIME,
                             "The number of particles is inconsistent with the number of radii provided.");
        bool mono = true;
        for (size_t i = 1; i < N; ++i) mono = mono && radius[i] == radius[0];
        if (field_type == MPM_FIELD_MONODISPERSE && !mono)
            throw mpm::Error(MPM_ERR_RUNTIME, "field_type was set to 'monodisperse' but radii are not identical.");
        if (eps_p.empty()) throw mpm::Error(MPM_ERR_RUNTIME, "eps_p is empty.");
        W = eps_p.size();
        MPM_REQUIRE(omega.size() >= W, MPM_ERR_INVALID_ARGUMENT,
                    "omega has fewer entries than eps_p has wavelengths");
        for (auto& row : eps_p) {
            if (row.size() == 1)
                row.assign(N, row[0]);
            else if (row.size() != N)
                throw mpm::Error(MPM_ERR_RUNTIME, "The number of particles is inconsistent with eps_p provided.");
        }
        for (int q : quad_idxs)
            if (q < 0 || (size_t)q >= N) throw mpm::Error(MPM_ERR_INVALID_ARGUMENT, "quadrupole index out of range");
    }

    // eels_solver.cu:281-303 (frequencies scale with the length unit: w a)
    void nondimensionalize() {
        length_scale = *std::min_element(radius.begin(), radius.end());
        eps_scale = eps_m;
        for (auto& b : box) b /= length_scale;
        for (auto& r : radius) r /= length_scale;
        eps_m /= eps_scale;
        for (auto& row : eps_p)
            for (auto& e : row) e /= eps_scale;
        for (auto& w : omega) w *= length_scale;
        integration_step /= length_scale;
        double s3 = 0.0;
        for (double r : radius) s3 += r * r * r;
        const double bv = box[0] * box[1] * box[2];
        vol_frac = bv == 0.0 ? 0.0 : (4.0 / 3.0) * M_PI * s3 / bv;
    }

    void build_operator() {
        bool mono = true;
        for (size_t i = 1; i < N; ++i) mono = mono && radius[i] == radius[0];
        const bool use_direct = field_type == MPM_FIELD_DIRECT;
        const bool use_poly = field_type == MPM_FIELD_POLYDISPERSE || (field_type == MPM_FIELD_AUTO && !mono);
        if (Q > 0 && use_poly && !use_direct)
            throw mpm::Error(MPM_ERR_UNSUPPORTED,
                             "quadrupoles with the polydisperse field are not provided by this library");
        mpm_op_desc od;
        std::memset(&od, 0, sizeof(od));
        for (int d = 0; d < 3; ++d) od.box[d] = box[d];
        od.errortol = tol;  // eels_solver.cu:436-438: the Ewald accuracy is the Krylov tol
        od.xi = xi;
        std::unique_ptr<mpm::EwaldOp> o;
        if (use_direct)
            o = std::make_unique<mpm::EwaldOp>(mpm::EwaldOp::DirectTag{}, radius.data(), (int64_t)N);
        else if (use_poly)
            o = std::make_unique<mpm::EwaldOp>(od, radius.data(), (int64_t)N);
        else
            o = std::make_unique<mpm::EwaldOp>(od);
        if (Q > 0)
            o->set_quads(quad_idxs.empty() ? nullptr : quad_idxs.data(),
                         quad_idxs.empty() ? -1 : (int64_t)quad_idxs.size());
        op = std::move(o);
    }

    mpm_solve_desc solve_desc() const {
        mpm_solve_desc d;
        d.method = method;
        d.restart = restart;
        d.maxiter = maxiter;
        d.tol = tol;
        d.jacobi = use_jacobi ? 1 : 0;
        d.precision = precision == MPM_PREC_DOUBLE ? MPM_SOLVE_DOUBLE : (refined ? MPM_SOLVE_REFINED : MPM_SOLVE_MIXED);
        return d;
    }

    // self coefficient -3 / (4 pi (1 - eps)), without the a^-3 of the dipole driver
    // (eels_solver.cu:730-733)
    std::vector<cplx> self_coef(size_t w) const {
        std::vector<cplx> sc(N);
        for (size_t p = 0; p < N; ++p) sc[p] = -3.0 / (4.0 * M_PI * (1.0 - eps_p[w][p]));
        return sc;
    }

    // largest per-particle norm of the dipole rows and of the quadrupole rows (:736-754)
    double field_norm(const cplx* e) const {
        double m = 0.0;
        for (size_t p = 0; p < N; ++p)
            m = std::max(m, std::sqrt(std::norm(e[3 * p]) + std::norm(e[3 * p + 1]) + std::norm(e[3 * p + 2])));
        for (size_t q = 0; q < Q; ++q) {
            double s = 0.0;
            for (int c = 0; c < 5; ++c) s += std::norm(e[3 * N + 5 * q + c]);
            m = std::max(m, std::sqrt(s));
        }
        return m < 1e-15 ? 1.0 : m;
    }

    // the first frequency's guess without an earlier Bloch point (:802-829): 0.1 alpha_MF
    // |E|-normalised field with the sign of its real part, z component times i
    std::vector<cplx> first_guess(const cplx* e) const {
        std::vector<cplx> g(n_unknowns(), 0.0);
        for (size_t p = 0; p < N; ++p) {
            double nrm = std::sqrt(std::norm(e[3 * p]) + std::norm(e[3 * p + 1]) + std::norm(e[3 * p + 2]));
            if (nrm < 1e-15) nrm = 1.0;
            const cplx beta = (eps_p[0][p] - 1.0) / (eps_p[0][p] + 2.0);
            const cplx val = 4.0 * M_PI * std::pow(radius[p], 3.0) * beta / (1.0 - beta * vol_frac);
            for (int c = 0; c < 3; ++c) {
                const cplx ew = e[3 * p + c] / nrm;
                g[3 * p + c] = 0.1 * val * ew * (ew.real() >= 0.0 ? 1.0 : -1.0);
            }
            g[3 * p + 2] *= cplx(0.0, 1.0);
        }
        return g;
    }

    // mean-field guess of the reference (:327-344): 4 pi beta / (1 - beta phi) times the field
    // it is given (the reference passes the normalised field, then divides by the norm again)
    std::vector<cplx> mean_field(size_t w, const cplx* e, bool with_radius) const {
        std::vector<cplx> g(n_unknowns(), 0.0);
        for (size_t p = 0; p < N; ++p) {
            const cplx beta = (eps_p[w][p] - 1.0) / (eps_p[w][p] + 2.0);
            cplx val = 4.0 * M_PI * beta / (1.0 - beta * vol_frac);
            if (with_radius) val *= std::pow(radius[p], 3.0);
            for (int c = 0; c < 3; ++c) g[3 * p + c] = val * e[3 * p + c];
        }
        return g;
    }

    // previous / derivative guesses (:346-436) from this Bloch point's earlier frequencies
    // fr: [W][n] solutions of this Bloch point (physical normalisation)
    std::vector<cplx> freq_guess(size_t w, const std::vector<cplx>& fr, const cplx* e_norm) const {
        const size_t n = n_unknowns();
        if (guess == MPM_GUESS_MEAN_FIELD) return mean_field(w, e_norm, false);
        if (guess == MPM_GUESS_PREVIOUS || w < 2)
            return std::vector<cplx>(fr.begin() + (w - 1) * n, fr.begin() + w * n);
        auto X = [](cplx e) { return (e - 1.0 == 0.0) ? cplx(0.0) : (e + 2.0) / (e - 1.0); };
        std::vector<cplx> g(n);
        auto extrapolate = [&](size_t p, size_t off, int len) {
            cplx run = X(eps_p[w - 1][p]) - X(eps_p[w - 2][p]);
            cplx new_run = X(eps_p[w][p]) - X(eps_p[w - 1][p]);
            if (run == 0.0) {
                run = 1.0;
                new_run = 0.0;
            }
            for (int c = 0; c < len; ++c) {
                const cplx v1 = fr[(w - 1) * n + off + c], v2 = fr[(w - 2) * n + off + c];
                g[off + c] = v1 + new_run * (v1 - v2) / run;
            }
        };
        for (size_t p = 0; p < N; ++p) extrapolate(p, 3 * p, 3);
        for (size_t q = 0; q < Q; ++q) extrapolate(quad_particle(q), 3 * N + 5 * q, 5);
        return g;
    }

    // previous Bloch point's solutions carried to this one (:768-800): frequency 0 by the
    // phase e^{i dk.x}; later frequencies bilinearly from (w - 1, this point), (w, previous
    // point) and (w - 1, previous point)
    std::vector<cplx> k_guess(size_t w, const std::vector<cplx>& fr, const std::vector<cplx>& prev, double dkx,
                              double dky, const std::vector<double>& xyz) const {
        const size_t n = n_unknowns();
        std::vector<cplx> g(n);
        auto phase = [&](size_t p) { return std::exp(cplx(0.0, dkx * xyz[3 * p] + dky * xyz[3 * p + 1])); };
        auto one = [&](size_t p, size_t off, int len) {
            const cplx ph = phase(p);
            for (int c = 0; c < len; ++c) {
                if (w == 0)
                    g[off + c] = prev[off + c] * ph;
                else
                    g[off + c] = fr[(w - 1) * n + off + c] + (prev[w * n + off + c] - prev[(w - 1) * n + off + c]) * ph;
            }
        };
        for (size_t p = 0; p < N; ++p) one(p, 3 * p, 3);
        for (size_t q = 0; q < Q; ++q) one(quad_particle(q), 3 * N + 5 * q, 5);
        return g;
    }

    // reciprocity sum (:844-880): sum_p Im(p . conj E) + (1/6) sum_q Im(q : conj G) over
    // 2 pi^2, with the off-diagonal quadrupole components counted twice
    double loss(const cplx* x, const cplx* e) const {
        double s = 0.0;
        for (size_t i = 0; i < 3 * N; ++i) s += (x[i] * std::conj(e[i])).imag();
        const double wq[5] = {1.0, 2.0, 2.0, 1.0, 2.0};
        for (size_t q = 0; q < Q; ++q) {
            cplx t = 0.0;
            for (int c = 0; c < 5; ++c) t += wq[c] * x[3 * N + 5 * q + c] * std::conj(e[3 * N + 5 * q + c]);
            s += t.imag() / 6.0;
        }
        return s / (2.0 * M_PI * M_PI);
    }

    void compute(const double* epos, const double* x, const double* y, const double* z, size_t n) {
        if (n == 0) return;
        MPM_REQUIRE(epos, MPM_ERR_INVALID_ARGUMENT, "null beam position");
        auto t0 = std::chrono::steady_clock::now();
        if (N == 0) {
            const std::vector<double> box0 = box, radius0 = radius, omega0 = omega;
            const std::vector<std::vector<cplx>> eps0 = eps_p;
            const double eps_m0 = eps_m, step0 = integration_step;
            try {
                set_dims(n);
                nondimensionalize();
                build_operator();
            } catch (...) {
                box = box0;
                radius = radius0;
                omega = omega0;
                eps_p = eps0;
                eps_m = eps_m0;
                integration_step = step0;
                N = W = Q = 0;
                length_scale = eps_scale = 1.0;
                vol_frac = 0.0;
                op.reset();
                throw;
            }
        } else if (N != n) {
            throw mpm::Error(MPM_ERR_RUNTIME, "The number of particles has changed!");
        }
        const size_t nu = n_unknowns();
        std::vector<double> xyz(3 * N), soa(3 * N);
        for (size_t i = 0; i < N; ++i) {
            xyz[3 * i] = x[i] / length_scale;
            xyz[3 * i + 1] = y[i] / length_scale;
            xyz[3 * i + 2] = z[i] / length_scale;
            for (int d = 0; d < 3; ++d) soa[d * N + i] = xyz[3 * i + d];
        }
        op->set_positions(xyz.data(), (int64_t)N);
        d_pos.upload(soa.data(), soa.size());
        d_omega.upload(omega.data(), W);
        std::vector<int> qown(std::max<size_t>(Q, 1), 0);
        for (size_t q = 0; q < Q; ++q) qown[q] = (int)quad_particle(q);
        d_qown.upload(qown.data(), qown.size());
        d_einc.reserve(W * nu);
        d_b.reserve(W * nu);
        d_x.reserve(W * nu);

        const double gamma = 1.0 / std::sqrt(1.0 - eps_m * v * v);  // eps_m is 1 after scaling
        const int Nx = asm_flag ? asm_nx : 1, Ny = asm_flag ? asm_ny : 1;
        const double dkx = asm_flag ? 2.0 * M_PI / (box[0] * Nx) : 0.0;
        const double dky = asm_flag ? 2.0 * M_PI / (box[1] * Ny) : 0.0;
        std::vector<int> rmax(2 * W, 0);
        if (asm_flag)
            for (size_t w = 0; w < W; ++w) {
                const double r_decay = 12.0 * std::abs(v * gamma / (2.0 * M_PI * omega[w]));
                rmax[2 * w] = (int)std::ceil(r_decay / box[0]);
                rmax[2 * w + 1] = (int)std::ceil(r_decay / box[1]);
            }
        d_rmax.upload(rmax.data(), rmax.size());

        std::vector<double> frame_eels(W, 0.0), eels_sum(W, 0.0);
        std::vector<cplx> frame(W * nu, 0.0), sum(W * nu, 0.0), prev(W * nu, 0.0);
        std::vector<mpm_solve_report> reps;
        bool has_prev = false;
        double pkx = 0.0, pky = 0.0;
        double device_ms = 0.0;
        auto t1 = std::chrono::steady_clock::now();
        print(asm_flag ? "Brillouin Zone point:" : "Wavenumber:");
        indent++;
        std::vector<cplx> einc(W * nu);
        for (int ix = 0; ix < Nx; ++ix)
            for (int iy = 0; iy < Ny; ++iy) {
                const double kx = asm_flag ? -M_PI / box[0] + (ix + 0.5) * dkx : 0.0;
                const double ky = asm_flag ? -M_PI / box[1] + (iy + 0.5) * dky : 0.0;
                if (!op->is_direct()) op->set_bloch(kx, ky);
                // incident fields of every frequency on the device
                mpm::BeamGeom g;
                g.ex = epos[0] / length_scale;
                g.ey = epos[1] / length_scale;
                g.v = v;
                g.gamma = gamma;
                g.sigma = 0.5 / length_scale;
                g.kx = kx;
                g.ky = ky;
                g.Lx = box[0];
                g.Ly = box[1];
                g.N = (int)N;
                g.Q = (int)Q;
                g.W = (int)W;
                g.n = (int64_t)nu;
                {
                    mpm::ProfScope ps("ebeam", nullptr, (int64_t)W);
                    const int64_t threads = (int64_t)(N + Q) * W;
                    hipLaunchKernelGGL(mpm::ebeam_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, nullptr,
                                       d_pos.p, d_omega.p, d_rmax.p, d_qown.p, g, d_einc.p);
                    MPM_HIP(hipGetLastError());
                }
                MPM_HIP(hipMemcpy(einc.data(), d_einc.p, W * nu * sizeof(double2), hipMemcpyDeviceToHost));
                std::vector<double> enorm(W);
                for (size_t w = 0; w < W; ++w) enorm[w] = field_norm(&einc[w * nu]);
                std::vector<cplx> rows(W * N);
                for (size_t w = 0; w < W; ++w) {
                    const auto sc = self_coef(w);
                    std::copy(sc.begin(), sc.end(), rows.begin() + w * N);
                }
                auto guess_for = [&](size_t w) {
                    const cplx* e = &einc[w * nu];
                    // batched: every frequency carries its own previous-point solution over
                    if (has_prev && batched) return k_guess(0, frame, std::vector<cplx>(prev.begin() + w * nu,
                                                            prev.begin() + (w + 1) * nu), kx - pkx, ky - pky, xyz);
                    if (has_prev) return k_guess(w, frame, prev, kx - pkx, ky - pky, xyz);
                    if (w == 0) return first_guess(e);
                    // every frequency at once: the dipole driver's mean field of this field
                    if (batched) return mean_field(w, e, true);
                    std::vector<cplx> en(nu);
                    for (size_t i = 0; i < nu; ++i) en[i] = e[i] / enorm[w];
                    return freq_guess(w, frame, en.data());
                };
                auto timed = [&](const double2* b, double2* xx, int B, mpm_solve_report* r) {
                    MPM_HIP(hipDeviceSynchronize());
                    auto a = std::chrono::steady_clock::now();
                    mpm::solve_batched(*op, solve_desc(), b, xx, B, r, nullptr);
                    MPM_HIP(hipDeviceSynchronize());
                    device_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
                };
                std::vector<mpm_solve_report> kr(W);
                if (batched) {
                    std::vector<cplx> b(W * nu), x0(W * nu);
                    for (size_t w = 0; w < W; ++w) {
                        const auto gg = guess_for(w);
                        for (size_t i = 0; i < nu; ++i) {
                            b[w * nu + i] = einc[w * nu + i] / enorm[w];
                            x0[w * nu + i] = gg[i] / enorm[w];
                        }
                    }
                    op->set_self_coef(reinterpret_cast<const double*>(rows.data()), (int)W);
                    MPM_HIP(hipMemcpy(d_b.p, b.data(), W * nu * sizeof(double2), hipMemcpyHostToDevice));
                    MPM_HIP(hipMemcpy(d_x.p, x0.data(), W * nu * sizeof(double2), hipMemcpyHostToDevice));
                    timed(d_b.p, d_x.p, (int)W, kr.data());
                    MPM_HIP(hipMemcpy(x0.data(), d_x.p, W * nu * sizeof(double2), hipMemcpyDeviceToHost));
                    for (size_t i = 0; i < W * nu; ++i) frame[i] = x0[i] * enorm[i / nu];
                } else {
                    std::vector<cplx> b(nu), x0(nu);
                    for (size_t w = 0; w < W; ++w) {
                        if (!asm_flag) print(std::to_string(w) + " of " + std::to_string(W));
                        else if (w == 0) print(std::to_string(ix * Ny + iy) + " of " + std::to_string(Nx * Ny));
                        op->set_self_coef(reinterpret_cast<const double*>(&rows[w * N]), 1);
                        const auto gg = guess_for(w);
                        for (size_t i = 0; i < nu; ++i) {
                            b[i] = einc[w * nu + i] / enorm[w];
                            x0[i] = gg[i] / enorm[w];
                        }
                        MPM_HIP(hipMemcpy(d_b.p, b.data(), nu * sizeof(double2), hipMemcpyHostToDevice));
                        MPM_HIP(hipMemcpy(d_x.p, x0.data(), nu * sizeof(double2), hipMemcpyHostToDevice));
                        timed(d_b.p, d_x.p, 1, &kr[w]);
                        MPM_HIP(hipMemcpy(x0.data(), d_x.p, nu * sizeof(double2), hipMemcpyDeviceToHost));
                        for (size_t i = 0; i < nu; ++i) frame[w * nu + i] = x0[i] * enorm[w];
                    }
                }
                for (size_t w = 0; w < W; ++w) {
                    const double e = loss(&frame[w * nu], &einc[w * nu]);
                    frame_eels[w] = e;
                    eels_sum[w] += e;
                    for (size_t i = 0; i < nu; ++i) sum[w * nu + i] += frame[w * nu + i];
                }
                reps.insert(reps.end(), kr.begin(), kr.end());
                prev = frame;
                pkx = kx;
                pky = ky;
                has_prev = true;
            }
        indent--;
        if (asm_flag) {
            // average over the Bloch points: dk^2 / (Brillouin-zone area) per point (:858-872)
            const double f = (dkx * dky) / ((2.0 * M_PI / box[0]) * (2.0 * M_PI / box[1]));
            for (size_t w = 0; w < W; ++w) frame_eels[w] = eels_sum[w] * f;
            for (auto& s : sum) s *= f;
            frame = sum;
        }
        MPM_HIP(hipDeviceSynchronize());
        auto t2 = std::chrono::steady_clock::now();
        setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
        solve_ms = device_ms;
        // commit the frame
        eels.insert(eels.end(), frame_eels.begin(), frame_eels.end());
        for (size_t w = 0; w < W; ++w) {
            dips.insert(dips.end(), frame.begin() + w * nu, frame.begin() + w * nu + 3 * N);
            quads.insert(quads.end(), frame.begin() + w * nu + 3 * N, frame.begin() + (w + 1) * nu);
        }
        reports.insert(reports.end(), reps.begin(), reps.end());
        last_einc = einc;
        frames++;
    }
};

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
using mpm::guarded;

extern "C" {

mpm_status mpm_eels_solver_create(const mpm_eels_desc* d, mpm_eels_solver** out) {
    return guarded([&] {
        MPM_REQUIRE(d && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(d->eps_p && d->num_wavevectors > 0 && d->eps_cols > 0, MPM_ERR_RUNTIME, "eps_p is empty.");
        MPM_REQUIRE(d->omega && d->num_omega > 0, MPM_ERR_INVALID_ARGUMENT, "omega is empty");
        MPM_REQUIRE(d->v > 0.0 && d->v < 1.0, MPM_ERR_INVALID_ARGUMENT,
                    "v must lie in (0, 1): the beam speed as a fraction of the speed of light");
        MPM_REQUIRE(d->num_radii >= 0 && (d->num_radii == 0 || d->radius), MPM_ERR_INVALID_ARGUMENT, "null radius");
        MPM_REQUIRE(d->field_type >= 0 && d->field_type <= 3, MPM_ERR_RUNTIME, "Unknown field_type");
        MPM_REQUIRE(d->guess >= 0 && d->guess <= 2, MPM_ERR_RUNTIME, "Guess type not supported.");
        MPM_REQUIRE(d->method == MPM_GMRES || d->method == MPM_BICGSTAB, MPM_ERR_INVALID_ARGUMENT, "unknown method");
        MPM_REQUIRE(d->precision >= 0 && d->precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
        MPM_REQUIRE(!d->asm_flag || (d->asm_nx >= 1 && d->asm_ny >= 1), MPM_ERR_INVALID_ARGUMENT,
                    "asm_Nx and asm_Ny must be >= 1");
        MPM_REQUIRE(!d->asm_flag || d->field_type != MPM_FIELD_DIRECT, MPM_ERR_INVALID_ARGUMENT,
                    "aperiodic sampling needs a periodic (Ewald) field");
        auto s = std::make_unique<mpm_eels_solver>();
        s->box.assign(d->box, d->box + 3);
        s->eps_p.assign(d->num_wavevectors, std::vector<cplx>(d->eps_cols));
        for (int64_t w = 0; w < d->num_wavevectors; ++w)
            for (int64_t c = 0; c < d->eps_cols; ++c)
                s->eps_p[w][c] = cplx(d->eps_p[2 * (w * d->eps_cols + c)], d->eps_p[2 * (w * d->eps_cols + c) + 1]);
        s->omega.assign(d->omega, d->omega + d->num_omega);
        s->radius.assign(d->radius, d->radius + d->num_radii);
        for (double r : s->radius) MPM_REQUIRE(r > 0.0, MPM_ERR_INVALID_ARGUMENT, "radii must be positive");
        s->v = d->v;
        s->eps_m = d->eps_m;
        s->xi = d->xi;
        s->tol = d->tol;
        s->quiet = d->quiet != 0;
        s->guess = d->guess;
        s->method = d->method;
        s->field_type = d->field_type;
        s->integration_step = d->integration_step;
        s->solve_quads = d->solve_quadrupoles != 0;
        if (s->solve_quads && d->num_quad_idxs > 0) {
            MPM_REQUIRE(d->quad_idxs, MPM_ERR_INVALID_ARGUMENT, "null quad_idxs");
            s->quad_idxs.assign(d->quad_idxs, d->quad_idxs + d->num_quad_idxs);
        }
        s->asm_flag = d->asm_flag != 0;
        s->asm_nx = d->asm_nx;
        s->asm_ny = d->asm_ny;
        s->precision = d->precision;
        s->refined = d->refined != 0;
        s->batched = d->batched != 0;
        s->restart = d->restart > 0 ? d->restart : 50;
        s->maxiter = d->maxiter > 0 ? d->maxiter : 1000;
        *out = s.release();
    });
}

void mpm_eels_solver_destroy(mpm_eels_solver* s) { delete s; }

mpm_status mpm_ebeam_field(const double* xyz, int64_t n_particles, const int* quad_particles, int64_t n_quads,
                           const double* omega, int64_t n_omega, double v, const double epos[2], double sigma,
                           double* out) {
    return guarded([&] {
        MPM_REQUIRE(xyz && omega && epos && out && n_particles > 0 && n_omega > 0, MPM_ERR_INVALID_ARGUMENT,
                    "null or empty argument");
        MPM_REQUIRE(n_quads >= 0 && (n_quads == 0 || quad_particles), MPM_ERR_INVALID_ARGUMENT, "null quadrupoles");
        MPM_REQUIRE(v > 0.0 && v < 1.0, MPM_ERR_INVALID_ARGUMENT, "v must lie in (0, 1)");
        const size_t N = (size_t)n_particles, Q = (size_t)n_quads, W = (size_t)n_omega, nu = 3 * N + 5 * Q;
        std::vector<double> soa(3 * N);
        for (size_t i = 0; i < N; ++i)
            for (int d = 0; d < 3; ++d) soa[d * N + i] = xyz[3 * i + d];
        std::vector<int> qown(std::max<size_t>(Q, 1), 0);
        for (size_t q = 0; q < Q; ++q) {
            MPM_REQUIRE(quad_particles[q] >= 0 && (size_t)quad_particles[q] < N, MPM_ERR_INVALID_ARGUMENT,
                        "quadrupole particle out of range");
            qown[q] = quad_particles[q];
        }
        std::vector<int> rmax(2 * W, 0);
        mpm::DevBuf<double> d_pos, d_om;
        mpm::DevBuf<int> d_rmax, d_q;
        mpm::DevBuf<double2> d_out;
        d_pos.upload(soa.data(), soa.size());
        d_om.upload(omega, W);
        d_rmax.upload(rmax.data(), rmax.size());
        d_q.upload(qown.data(), qown.size());
        d_out.resize(W * nu);
        mpm::BeamGeom g;
        g.ex = epos[0];
        g.ey = epos[1];
        g.v = v;
        g.gamma = 1.0 / std::sqrt(1.0 - v * v);
        g.sigma = sigma;
        g.kx = g.ky = 0.0;
        g.Lx = g.Ly = 0.0;
        g.N = (int)N;
        g.Q = (int)Q;
        g.W = (int)W;
        g.n = (int64_t)nu;
        const int64_t threads = (int64_t)(N + Q) * W;
        hipLaunchKernelGGL(mpm::ebeam_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, nullptr, d_pos.p,
                           d_om.p, d_rmax.p, d_q.p, g, d_out.p);
        MPM_HIP(hipGetLastError());
        MPM_HIP(hipMemcpy(out, d_out.p, W * nu * sizeof(double2), hipMemcpyDeviceToHost));
    });
}

mpm_status mpm_eels_solver_compute(mpm_eels_solver* s, const double epos[2], const double* x, const double* y,
                                   const double* z, int64_t n) {
    return guarded([&] {
        MPM_REQUIRE(s, MPM_ERR_INVALID_ARGUMENT, "null solver");
        MPM_REQUIRE(n >= 0 && (n == 0 || (x && y && z)), MPM_ERR_INVALID_ARGUMENT, "null coordinates");
        for (int64_t i = 0; i < n; ++i)
            MPM_REQUIRE(std::isfinite(x[i]) && std::isfinite(y[i]) && std::isfinite(z[i]), MPM_ERR_INVALID_ARGUMENT,
                        "non-finite position");
        MPM_REQUIRE(!epos || (std::isfinite(epos[0]) && std::isfinite(epos[1])), MPM_ERR_INVALID_ARGUMENT,
                    "non-finite beam position");
        s->compute(epos, x, y, z, (size_t)n);
    });
}

mpm_status mpm_eels_solver_get_dims(const mpm_eels_solver* s, int64_t dims[4]) {
    return guarded([&] {
        MPM_REQUIRE(s && dims, MPM_ERR_INVALID_ARGUMENT, "null argument");
        dims[0] = (int64_t)s->frames;
        dims[1] = (int64_t)s->W;
        dims[2] = (int64_t)s->N;
        dims[3] = (int64_t)s->Q;
    });
}

mpm_status mpm_eels_solver_get_eels(const mpm_eels_solver* s, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        std::copy(s->eels.begin(), s->eels.end(), out);
    });
}

mpm_status mpm_eels_solver_get_dipoles(const mpm_eels_solver* s, int physical, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        const double f = physical ? s->eps_scale * std::pow(s->length_scale, 3.0) : 1.0;
        for (size_t i = 0; i < s->dips.size(); ++i) {
            out[2 * i] = s->dips[i].real() * f;
            out[2 * i + 1] = s->dips[i].imag() * f;
        }
    });
}

mpm_status mpm_eels_solver_get_quadrupoles(const mpm_eels_solver* s, int physical, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && (out || s->quads.empty()), MPM_ERR_INVALID_ARGUMENT, "null argument");
        const double f = physical ? s->eps_scale * std::pow(s->length_scale, 4.0) : 1.0;
        for (size_t i = 0; i < s->quads.size(); ++i) {
            out[2 * i] = s->quads[i].real() * f;
            out[2 * i + 1] = s->quads[i].imag() * f;
        }
    });
}

mpm_status mpm_eels_solver_get_incident_field(const mpm_eels_solver* s, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(s->frames > 0, MPM_ERR_RUNTIME, "no frame computed yet");
        std::memcpy(out, s->last_einc.data(), s->last_einc.size() * sizeof(cplx));
    });
}

mpm_status mpm_eels_solver_get_reports(const mpm_eels_solver* s, mpm_solve_report* out, int64_t cap,
                                       int64_t* count) {
    return guarded([&] {
        MPM_REQUIRE(s && count, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *count = (int64_t)s->reports.size();
        if (out && cap >= *count) std::copy(s->reports.begin(), s->reports.end(), out);
    });
}

mpm_status mpm_eels_solver_set_use_jacobi(mpm_eels_solver* s, int enable) {
    return guarded([&] {
        MPM_REQUIRE(s, MPM_ERR_INVALID_ARGUMENT, "null solver");
        s->use_jacobi = enable != 0;
    });
}

mpm_status mpm_eels_solver_get_use_jacobi(const mpm_eels_solver* s, int* enable) {
    return guarded([&] {
        MPM_REQUIRE(s && enable, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *enable = s->use_jacobi ? 1 : 0;
    });
}

mpm_status mpm_eels_solver_get_timing(const mpm_eels_solver* s, double* solve_ms, double* setup_ms,
                                      double* total_ms) {
    return guarded([&] {
        MPM_REQUIRE(s, MPM_ERR_INVALID_ARGUMENT, "null solver");
        if (solve_ms) *solve_ms = s->solve_ms;
        if (setup_ms) *setup_ms = s->setup_ms;
        if (total_ms) *total_ms = s->total_ms;
    });
}

}  // extern "C"
