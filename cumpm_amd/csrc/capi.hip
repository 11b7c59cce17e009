// C ABI of libcumpm_amd (include/cumpm_amd.h) and the Dipole_Solver driver.
//
// The driver restates src/cpp/dipole_solver.cu: set_dims (:131-182),
// nondimensionalize (:184-205), calc_vol_frac (:207-221), the mean-field /
// previous / derivative guesses (:245-346), compute_tensor / compute_spectrum
// (:361-489), compute (:495-607) and the physical rescaling of the getters
// (:609-647).  Solves go through solve_batched(): in the reference order one
// system at a time (batched = 0), or every frequency x polarisation system of
// the frame in one batched launch (batched = 1; the derivative guess needs
// the previous frequencies' solutions, so batched mode starts every system
// from the mean-field guess).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>

#include "mpm_internal.h"

struct mpm_op {
    std::unique_ptr<mpm::EwaldOp> impl;
};

namespace mpm {
void profile_set(bool on);
void profile_reset();
bool profile_get(const std::string& name, double& ms, int64_t& count, int64_t& units);
std::string profile_names();
}  // namespace mpm

namespace {
// Quadrupole block of the metric M = diag(I_3N, W / 2) under which M A is Hermitian:
// W is the Frobenius product of traceless symmetric tensors in the components
// (xx, xy, xz, yy, yz), Q:G = 2 xx Gxx + xx Gyy + yy Gxx + 2 yy Gyy + 2 (xy Gxy + xz Gxz
// + yz Gyz).  inverse 0: v_q <- (W / 2) v_q; 1: v_q <- (W / 2)^-1 v_q (only the (xx, yy)
// pair mixes; the other components are unchanged).
__global__ void __launch_bounds__(256) quad_metric_kernel(double2* __restrict__ v, int B, size_t n, size_t off,
                                                          size_t Q, int inverse) {
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= (size_t)B * Q) return;
    double2* e = v + (t / Q) * n + off + 5 * (t % Q);
    const double2 xx = e[0], yy = e[3];
    const double a = inverse ? 4.0 / 3.0 : 1.0, c = inverse ? -2.0 / 3.0 : 0.5;
    e[0] = make_double2(a * xx.x + c * yy.x, a * xx.y + c * yy.y);
    e[3] = make_double2(c * xx.x + a * yy.x, c * xx.y + a * yy.y);
}


using mpm::cplx;
using mpm::guarded;

}  // namespace

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------
struct mpm_dipole_solver {
    // configuration (physical until nondimensionalised)
    std::vector<double> box;
    std::vector<std::vector<cplx>> eps_p;  // [W][cols]
    std::vector<double> radius;
    double eps_m = 1.0, xi = 0.5, tol = 1e-3;
    bool quiet = false;
    int guess = MPM_GUESS_DERIVATIVE, method = MPM_GMRES, field_type = MPM_FIELD_AUTO;
    std::vector<std::vector<cplx>> E0;
    int precision = MPM_PREC_AUTO;
    bool refined = false, batched = false, use_jacobi = true;
    int restart = 50, maxiter = 1000;
    int grid[3] = {0, 0, 0};
    int support = 0;
    double ewald_tol = 0.0;
    int autotune = 0;         // 0 off, 1 layout and FFT, 2 also grid growth
    mpm_tune_info tune{};     // the first frame's choice
    bool solve_quads = false;
    std::vector<int> quad_idxs;  // empty with solve_quads: every particle
    // state
    size_t N = 0, W = 0, K = 0, Q = 0, frames = 0;
    double length_scale = 1.0, eps_scale = 1.0, vol_frac = 0.0;
    std::unique_ptr<mpm::EwaldOp> op;
    bool use_poly = false, use_direct = false;
    std::vector<cplx> dips;      // [F][W][N][K][3]
    std::vector<cplx> avg_dips;  // [F][W][K][3]
    std::vector<cplx> quads;     // [F][W][Q][K][5]
    std::vector<mpm_solve_report> reports;  // [F][W][K]
    double solve_ms = 0, setup_ms = 0, total_ms = 0;
    mpm::DevBuf<double2> d_b, d_x;
    // batched mode: right-hand sides (on the device), self-coefficient rows (in the operator)
    // and mean-field guesses depend on the configuration only, so they are built once
    bool batch_ready = false;
    std::vector<cplx> mf_all;         // [B][n] mean-field guesses
    cplx* h_stage = nullptr;          // pinned staging for the guesses and solutions
    size_t h_stage_n = 0;
    cplx* stage(size_t n) {
        if (n > h_stage_n) {
            if (h_stage) (void)hipHostFree(h_stage);
            h_stage = nullptr;
            h_stage_n = 0;
            MPM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_stage), n * sizeof(cplx), hipHostMallocDefault));
            h_stage_n = n;
        }
        return h_stage;
    }
    ~mpm_dipole_solver() {
        if (h_stage) (void)hipHostFree(h_stage);
    }
    int indent = 0;

    void print(const std::string& m) const {
        if (!quiet) std::cout << std::string(indent * 2, ' ') << m << std::endl;
    }

    void set_dims(size_t n) {
        N = n;
        // dipole_solver.cu:131-133
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
        // dipole_solver.cu:527-536: auto selects the polydisperse operator iff radii differ
        use_poly = field_type == MPM_FIELD_POLYDISPERSE || (field_type == MPM_FIELD_AUTO && !mono);
        use_direct = field_type == MPM_FIELD_DIRECT;
        if (eps_p.empty()) throw mpm::Error(MPM_ERR_RUNTIME, "eps_p is empty.");
        W = eps_p.size();
        for (auto& row : eps_p) {
            if (row.size() == 1)
                row.assign(N, row[0]);
            else if (row.size() != N)
                throw mpm::Error(MPM_ERR_RUNTIME, "The number of particles is inconsistent with eps_p provided.");
        }
        if (Q > 0 && use_poly)
            throw mpm::Error(MPM_ERR_UNSUPPORTED,
                             "quadrupoles with the polydisperse field are not provided by this library");
        for (int q : quad_idxs)
            if (q < 0 || (size_t)q >= N) throw mpm::Error(MPM_ERR_INVALID_ARGUMENT, "quadrupole index out of range");
        K = E0.size();
        const size_t nu = n_unknowns();
        for (size_t k = 0; k < K; ++k)
            if (E0[k].size() != 3 && E0[k].size() != 3 * N && E0[k].size() != nu)
                throw mpm::Error(MPM_ERR_RUNTIME, "Incident field E0[" + std::to_string(k) + "] has size " +
                                                      std::to_string(E0[k].size()) + ", but expected 3, " +
                                                      std::to_string(3 * N) + ", or " + std::to_string(nu) + ".");
    }

    // unknowns of one system: N dipoles (3 each) then Q quadrupoles (5 each)
    size_t n_unknowns() const { return 3 * N + 5 * Q; }
    // particle carrying quadrupole slot q
    size_t quad_particle(size_t q) const { return quad_idxs.empty() ? q : (size_t)quad_idxs[q]; }

    void nondimensionalize() {
        length_scale = *std::min_element(radius.begin(), radius.end());
        eps_scale = eps_m;
        for (auto& b : box) b /= length_scale;
        for (auto& r : radius) r /= length_scale;
        eps_m /= eps_scale;
        for (auto& row : eps_p)
            for (auto& e : row) e /= eps_scale;
        double s3 = 0.0;
        for (double r : radius) s3 += r * r * r;
        const double bv = box[0] * box[1] * box[2];
        vol_frac = bv == 0.0 ? 0.0 : (4.0 / 3.0) * M_PI * s3 / bv;
    }

    // guess layout [p][K][3] then [q][K][5] (dipole_solver.cu:244-261); quadrupoles start at zero
    std::vector<cplx> mean_field(size_t w) const {
        std::vector<cplx> g(N * K * 3 + Q * K * 5, 0.0);
        for (size_t p = 0; p < N; ++p) {
            const cplx beta = (eps_p[w][p] - 1.0) / (eps_p[w][p] + 2.0);
            const cplx val = 4.0 * M_PI * std::pow(radius[p], 3.0) * beta / (1.0 - beta * vol_frac);
            for (size_t k = 0; k < K; ++k)
                for (int c = 0; c < 3; ++c) {
                    const cplx e = E0[k].size() > 3 ? E0[k][p * 3 + c] : E0[k][c];
                    g[p * K * 3 + k * 3 + c] = val * e;
                }
        }
        return g;
    }
    // fd: [W][N][K][3] and fq: [W][Q][K][5] of the current frame (dipole_solver.cu:233-345)
    std::vector<cplx> guess_for(size_t w, const std::vector<cplx>& fd, const std::vector<cplx>& fq) const {
        const size_t blk = N * K * 3, qblk = Q * K * 5;
        if (guess == MPM_GUESS_MEAN_FIELD || w < 1) return mean_field(w);
        std::vector<cplx> g(blk + qblk);
        if (guess == MPM_GUESS_PREVIOUS || w < 2) {
            std::copy(fd.begin() + (w - 1) * blk, fd.begin() + w * blk, g.begin());
            std::copy(fq.begin() + (w - 1) * qblk, fq.begin() + w * qblk, g.begin() + blk);
            return g;
        }
        // extrapolate in the inverse polarisability X = (eps + 2) / (eps - 1) of the owning particle
        auto X = [](cplx e) { return (e - 1.0 == 0.0) ? cplx(0.0) : (e + 2.0) / (e - 1.0); };
        auto extrapolate = [&](size_t p, const cplx* v1, const cplx* v2, cplx* out, size_t len) {
            cplx run = X(eps_p[w - 1][p]) - X(eps_p[w - 2][p]);
            cplx new_run = X(eps_p[w][p]) - X(eps_p[w - 1][p]);
            if (std::abs(run) < 1e-12) {
                run = 1.0;
                new_run = 0.0;
            }
            for (size_t i = 0; i < len; ++i) out[i] = v1[i] + new_run * (v1[i] - v2[i]) / run;
        };
        for (size_t p = 0; p < N; ++p)
            extrapolate(p, &fd[(w - 1) * blk + p * K * 3], &fd[(w - 2) * blk + p * K * 3], &g[p * K * 3], K * 3);
        for (size_t q = 0; q < Q; ++q)
            extrapolate(quad_particle(q), &fq[(w - 1) * qblk + q * K * 5], &fq[(w - 2) * qblk + q * K * 5],
                        &g[blk + q * K * 5], K * 5);
        return g;
    }

    // one system's start vector from a guess laid out as mean_field()
    std::vector<cplx> guess_column(const std::vector<cplx>& g, size_t k) const {
        std::vector<cplx> x0(n_unknowns());
        for (size_t p = 0; p < N; ++p)
            for (int c = 0; c < 3; ++c) x0[3 * p + c] = g[p * K * 3 + k * 3 + c];
        for (size_t q = 0; q < Q; ++q)
            for (int c = 0; c < 5; ++c) x0[3 * N + 5 * q + c] = g[N * K * 3 + q * K * 5 + k * 5 + c];
        return x0;
    }

    std::vector<cplx> self_coef(size_t w) const {
        std::vector<cplx> sc(N);
        for (size_t p = 0; p < N; ++p) {
            cplx diff = 1.0 - eps_p[w][p];
            if (std::abs(diff) < 1e-12) diff = 1e-12;
            sc[p] = -3.0 / (4.0 * M_PI * diff) / std::pow(radius[p], 3.0);
        }
        return sc;
    }

    // dipole_solver.cu:368-390: E0 of length 3 (uniform), 3N (per particle) or 3N+5Q (full
    // vector); the quadrupole rows are zero unless given
    std::vector<cplx> rhs(size_t k) const {
        std::vector<cplx> e(n_unknowns(), 0.0);
        if (E0[k].size() == e.size() && e.size() != 3)
            std::copy(E0[k].begin(), E0[k].end(), e.begin());
        else
            for (size_t p = 0; p < N; ++p)
                for (int c = 0; c < 3; ++c) e[3 * p + c] = E0[k].size() == 3 ? E0[k][c] : E0[k][3 * p + c];
        return e;
    }

    mpm_solve_desc solve_desc() const {
        mpm_solve_desc d;
        d.method = method;
        d.restart = restart;
        d.maxiter = maxiter;
        d.tol = tol;
        d.jacobi = use_jacobi ? 1 : 0;
        if (precision == MPM_PREC_DOUBLE)
            d.precision = MPM_SOLVE_DOUBLE;
        else
            d.precision = refined ? MPM_SOLVE_REFINED : MPM_SOLVE_MIXED;
        return d;
    }

    void compute(const double* x, const double* y, const double* z, size_t n) {
        if (n == 0) return;
        mpm::HostScope whole("host_compute");
        auto t0 = std::chrono::steady_clock::now();
        if (N == 0) {
            // set-up is all or nothing: on failure (a box too small for the cutoff, bad
            // radii, device errors) the configuration is restored and N stays 0, so a
            // retried compute() starts again from the physical inputs
            const std::vector<double> box0 = box, radius0 = radius;
            const std::vector<std::vector<cplx>> eps0 = eps_p;
            const double eps_m0 = eps_m;
            try {
                set_dims(n);
                nondimensionalize();
                mpm_op_desc od;
                std::memset(&od, 0, sizeof(od));
                for (int d = 0; d < 3; ++d) {
                    od.box[d] = box[d];
                    od.grid[d] = grid[d];
                }
                od.errortol = ewald_tol > 0 ? ewald_tol : tol;
                od.xi = xi;
                od.support = support;
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
                batch_ready = false;
            } catch (...) {
                box = box0;
                radius = radius0;
                eps_p = eps0;
                eps_m = eps_m0;
                N = W = K = Q = 0;
                length_scale = eps_scale = 1.0;
                vol_frac = 0.0;
                use_poly = use_direct = false;
                op.reset();
                throw;
            }
        } else if (N != n) {
            throw mpm::Error(MPM_ERR_RUNTIME, "The number of particles has changed!");
        }
        // frames counts committed frames only: it grows after this frame's results are stored
        std::vector<double> xyz(3 * N);
        for (size_t i = 0; i < N; ++i) {
            xyz[3 * i] = x[i] / length_scale;
            xyz[3 * i + 1] = y[i] / length_scale;
            xyz[3 * i + 2] = z[i] / length_scale;
        }
        {
            mpm::HostScope hs("host_set_positions");
            op->set_positions(xyz.data(), (int64_t)N);
        }
        if (autotune && !use_direct && frames == 0 && tune.candidates == 0) {
            const int B = batched ? (int)std::min<size_t>(W * K, (size_t)mpm::kMaxSlots) : 1;
            mpm::autotune(*op, B, precision != MPM_PREC_DOUBLE, autotune == 2 ? MPM_TUNE_GRID : 0, nullptr, tune);
            print("Autotune: layout " + std::to_string(tune.layout) + ", fft_backend " +
                  std::to_string(tune.fft_backend) + ", grid " + std::to_string(tune.grid[0]) + "x" +
                  std::to_string(tune.grid[1]) + "x" + std::to_string(tune.grid[2]) + " (" +
                  std::to_string(tune.best_ms) + " ms per apply, default " + std::to_string(tune.default_ms) +
                  " ms)");
        }
        const size_t n3 = n_unknowns();
        // initial guess: mean field, or the previous frame's first frequency (dipole_solver.cu:578-592)
        std::vector<cplx> init;
        if (frames == 0) {
            init = mean_field(0);
        } else {
            const size_t off = (frames - 1) * W * N * K * 3, qoff = (frames - 1) * W * Q * K * 5;
            init.assign(dips.begin() + off, dips.begin() + off + N * K * 3);
            init.insert(init.end(), quads.begin() + qoff, quads.begin() + qoff + Q * K * 5);
        }
        std::vector<cplx> fd(W * N * K * 3, 0.0), fq(W * Q * K * 5, 0.0);
        std::vector<mpm_solve_report> reps(W * K);
        MPM_HIP(hipDeviceSynchronize());
        auto t1 = std::chrono::steady_clock::now();
        double device_ms = 0.0;  // wall time of the device solves only (inputs resident)
        auto timed_solve = [&](const double2* b, double2* x, int B, mpm_solve_report* r) {
            MPM_HIP(hipDeviceSynchronize());
            auto a = std::chrono::steady_clock::now();
            mpm::solve_batched(*op, solve_desc(), b, x, B, r, nullptr);
            MPM_HIP(hipDeviceSynchronize());
            device_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
        };
        auto store = [&](size_t w, size_t k, const cplx* sol) {
            for (size_t p = 0; p < N; ++p)
                for (int c = 0; c < 3; ++c) fd[(w * N + p) * K * 3 + k * 3 + c] = sol[3 * p + c];
            for (size_t q = 0; q < Q; ++q)
                for (int c = 0; c < 5; ++c) fq[(w * Q + q) * K * 5 + k * 5 + c] = sol[3 * N + 5 * q + c];
        };
        if (!batched) {
            print("Wavenumber:");
            indent++;
            d_b.reserve(n3);
            d_x.reserve(n3);
            for (size_t w = 0; w < W; ++w) {
                print(std::to_string(w) + " of " + std::to_string(W));
                auto sc = self_coef(w);
                op->set_self_coef(reinterpret_cast<const double*>(sc.data()), 1);
                const std::vector<cplx> g = (w == 0) ? init : guess_for(w, fd, fq);
                for (size_t k = 0; k < K; ++k) {
                    auto b = rhs(k);
                    const std::vector<cplx> x0 = guess_column(g, k);
                    MPM_HIP(hipMemcpy(d_b.p, b.data(), n3 * sizeof(double2), hipMemcpyHostToDevice));
                    MPM_HIP(hipMemcpy(d_x.p, x0.data(), n3 * sizeof(double2), hipMemcpyHostToDevice));
                    timed_solve(d_b.p, d_x.p, 1, &reps[w * K + k]);
                    std::vector<cplx> sol(n3);
                    MPM_HIP(hipMemcpy(sol.data(), d_x.p, n3 * sizeof(double2), hipMemcpyDeviceToHost));
                    store(w, k, sol.data());
                }
            }
            indent--;
        } else {
            const size_t B = W * K;
            mpm::HostLaps laps;
            if (!batch_ready) {
                std::vector<cplx> b_all(B * n3), rows(B * N);
                mf_all.assign(B * n3, 0.0);
                for (size_t w = 0; w < W; ++w) {
                    const std::vector<cplx> sc = self_coef(w), g = mean_field(w);
                    for (size_t k = 0; k < K; ++k) {
                        const size_t sys = w * K + k;
                        const std::vector<cplx> b = rhs(k), x0 = guess_column(g, k);
                        std::copy(b.begin(), b.end(), b_all.begin() + sys * n3);
                        std::copy(x0.begin(), x0.end(), mf_all.begin() + sys * n3);
                        // one self-coefficient row per system (rows repeat across polarisations)
                        std::copy(sc.begin(), sc.end(), rows.begin() + sys * N);
                    }
                }
                d_b.reserve(B * n3);
                MPM_HIP(hipMemcpy(d_b.p, b_all.data(), B * n3 * sizeof(double2), hipMemcpyHostToDevice));
                op->set_self_coef(reinterpret_cast<const double*>(rows.data()), (int)B);
                batch_ready = true;
            }
            // guesses: the mean field, except frequency 0 after the first frame, which starts
            // from the previous frame's solution (dipole_solver.cu:578-592)
            cplx* xs = stage(B * n3);
            std::copy(mf_all.begin(), mf_all.end(), xs);
            if (frames > 0)
                for (size_t k = 0; k < K; ++k) {
                    const std::vector<cplx> x0 = guess_column(init, k);
                    std::copy(x0.begin(), x0.end(), xs + k * n3);
                }
            laps.lap("host_rhs_guesses");
            d_x.reserve(B * n3);
            MPM_HIP(hipMemcpy(d_x.p, xs, B * n3 * sizeof(double2), hipMemcpyHostToDevice));
            laps.lap("host_upload");
            timed_solve(d_b.p, d_x.p, (int)B, reps.data());
            laps.lap("host_solve_wall");
            MPM_HIP(hipMemcpy(xs, d_x.p, B * n3 * sizeof(double2), hipMemcpyDeviceToHost));
            laps.lap("host_download");
            for (size_t w = 0; w < W; ++w)
                for (size_t k = 0; k < K; ++k) store(w, k, xs + (w * K + k) * n3);
            laps.lap("host_store");
        }
        MPM_HIP(hipDeviceSynchronize());
        auto t2 = std::chrono::steady_clock::now();
        setup_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
        solve_ms = device_ms;
        // alpha_eff: mean dipole per (w, k)
        std::vector<cplx> avg(W * K * 3);
        for (size_t w = 0; w < W; ++w)
            for (size_t k = 0; k < K; ++k)
                for (int c = 0; c < 3; ++c) {
                    cplx s = 0.0;
                    for (size_t p = 0; p < N; ++p) s += fd[(w * N + p) * K * 3 + k * 3 + c];
                    avg[(w * K + k) * 3 + c] = s / (double)N;
                }
        // commit the frame
        avg_dips.insert(avg_dips.end(), avg.begin(), avg.end());
        dips.insert(dips.end(), fd.begin(), fd.end());
        quads.insert(quads.end(), fq.begin(), fq.end());
        reports.insert(reports.end(), reps.begin(), reps.end());
        frames++;
    }
};

// ---------------------------------------------------------------------------
// Near field (src/cpp/near_field.cu:14-213): the field that dipoles at given
// positions induce at arbitrary field points, as the intensity |E0 + E_ind|^2.
// The field points join the operator as sites without dipoles, so that row f of
// A [p; 0] (self coefficients zero) is the sum over particles of the pair terms
// at (f, particle) plus the Gaussian contraction of the reciprocal grid at f:
// minus the induced field there, exactly the INTERACTION_FIELD evaluation
// (monodisperse_ewald_electric_field.cu:1631-1869, 2385-2411).
// ---------------------------------------------------------------------------
struct mpm_near_field {
    double box[3];
    cplx E0[3];
    std::vector<double> radius;
    double xi = 0.5, errortol = 1e-3;
    int field_type = MPM_FIELD_AUTO;
    int precision = MPM_PREC_AUTO;  // the reference's fields default to an FP32 reciprocal grid
    std::vector<cplx> dip;                      // [N][3]
    std::vector<double> dpos[3], fpos[3];       // dipole and field-point coordinates
    double length_scale = 1.0;
    bool direct = false, poly = false;
    size_t op_sites = 0;  // direct / polydisperse field: particles + field points the radii cover
    std::unique_ptr<mpm::EwaldOp> op;
    mpm::DevBuf<double2> d_x, d_y;

    void calculate(double* intensity) {
        const size_t N = dip.size() / 3, M = fpos[0].size();
        if (fpos[0].empty() || dpos[0].empty() || dip.empty())
            throw mpm::Error(MPM_ERR_RUNTIME,
                             "Near_Field: must set dipoles, dipole positions, and field points before calculation.");
        if (dpos[0].size() != N)
            throw mpm::Error(MPM_ERR_RUNTIME,
                             "Near_Field: dipole position size is inconsistent with dipole vector size.");
        if (!op) {
            if (radius.empty())
                radius.assign(N, 1.0);
            else if (radius.size() == 1)
                radius.assign(N, radius[0]);
            else if (radius.size() != N)
                throw mpm::Error(MPM_ERR_RUNTIME,
                                 "The number of particles is inconsistent with the number of radii provided.");
            bool mono = true;
            for (size_t i = 1; i < N; ++i) mono = mono && radius[i] == radius[0];
            direct = field_type == MPM_FIELD_DIRECT;
            poly = !direct && (field_type == MPM_FIELD_POLYDISPERSE || (field_type == MPM_FIELD_AUTO && !mono));
            length_scale = radius[0];  // near_field.cu:120 (the first radius)
            if (poly) {
                // The field points are sites without dipoles.  Their radii follow the reference's
                // INTERACTION_FIELD evaluation by index: field point f takes particle f's radius
                // for f < N; beyond, its spread weight takes the first radius
                // (polydisperse_ewald_electric_field.cu:297) and its pair functions the smallest
                // radius class (:402, classes sorted ascending).
                const double rmin = *std::min_element(radius.begin(), radius.end());
                std::vector<double> rr(N + M), wr(N + M);
                for (size_t i = 0; i < N + M; ++i) {
                    const size_t f = i - N;
                    rr[i] = (i < N ? radius[i] : (f < N ? radius[f] : rmin)) / length_scale;
                    wr[i] = (i < N ? radius[i] : (f < N ? radius[f] : radius[0])) / length_scale;
                }
                mpm_op_desc od;
                std::memset(&od, 0, sizeof(od));
                for (int d = 0; d < 3; ++d) od.box[d] = box[d] / length_scale;
                od.errortol = errortol;
                od.xi = xi;
                op = std::make_unique<mpm::EwaldOp>(od, rr.data(), (int64_t)rr.size());
                op->set_weight_radii(wr.data(), (int64_t)wr.size());
                op_sites = N + M;
            } else if (direct) {
                // the particles' radii then unit radii for the field points (sites without dipoles:
                // their radius enters only their own self term, which multiplies a zero dipole)
                std::vector<double> rr(N + M, 1.0);
                for (size_t i = 0; i < N; ++i) rr[i] = radius[i] / length_scale;
                op = std::make_unique<mpm::EwaldOp>(mpm::EwaldOp::DirectTag{}, rr.data(), (int64_t)rr.size());
                op_sites = N + M;
            } else {
                mpm_op_desc od;
                std::memset(&od, 0, sizeof(od));
                for (int d = 0; d < 3; ++d) od.box[d] = box[d] / length_scale;
                od.errortol = errortol;
                od.xi = xi;
                op = std::make_unique<mpm::EwaldOp>(od);
            }
        }
        MPM_REQUIRE(!(direct || poly) || op_sites == N + M, MPM_ERR_RUNTIME,
                    "Near_Field: the direct and polydisperse fields keep the particle and field-point counts of "
                    "their first calculation");
        const size_t T = N + M;
        std::vector<double> xyz(3 * T);
        for (size_t i = 0; i < T; ++i)
            for (int d = 0; d < 3; ++d)
                xyz[3 * i + d] = (i < N ? dpos[d][i] : fpos[d][i - N]) / length_scale;
        op->set_positions(xyz.data(), (int64_t)T);
        std::vector<cplx> zero(T, 0.0);
        op->set_self_coef(reinterpret_cast<const double*>(zero.data()), 1);
        std::vector<cplx> x(3 * T, 0.0);
        std::copy(dip.begin(), dip.end(), x.begin());
        d_x.reserve(3 * T);
        d_y.reserve(3 * T);
        MPM_HIP(hipMemcpy(d_x.p, x.data(), 3 * T * sizeof(double2), hipMemcpyHostToDevice));
        op->apply_contiguous(d_x.p, d_y.p, 1, precision != MPM_PREC_DOUBLE && !direct, MPM_PART_ALL, nullptr);
        MPM_HIP(hipGetLastError());
        std::vector<cplx> y(3 * T);
        MPM_HIP(hipMemcpy(y.data(), d_y.p, 3 * T * sizeof(double2), hipMemcpyDeviceToHost));
        for (size_t f = 0; f < M; ++f) {
            double acc = 0.0;
            for (int c = 0; c < 3; ++c) acc += std::norm(E0[c] - y[3 * (N + f) + c]);
            intensity[f] = acc;
        }
    }
};

// ---------------------------------------------------------------------------
// Scratch of mpm_op_apply_noalloc: the reciprocal grids of one apply live in this
// file-scope device array instead of the operator's own (allocated on demand)
// buffers.  Every GPU the library runs on holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in, not device memory: the
// array's address on each device comes from hipGetSymbolAddress (scratch_on_device)
// and only that address reaches kernels.  Calls on one device share the array, so
// calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on the current device, looked up once per device.
char* scratch_on_device() {
    int dev = 0;
    MPM_HIP(hipGetDevice(&dev));
    MPM_REQUIRE(dev >= 0 && dev < kMaxDevices, MPM_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        MPM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        MPM_REQUIRE(ptr != nullptr, MPM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<char*>(ptr);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

mpm_status mpm_near_field_create(const double box[3], const double* E0, const double* radius, int64_t num_radii,
                                 double xi, double errortol, int field_type, mpm_near_field** out) {
    return guarded([&] {
        MPM_REQUIRE(box && E0 && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(num_radii >= 0 && (num_radii == 0 || radius), MPM_ERR_INVALID_ARGUMENT, "null radius");
        MPM_REQUIRE(field_type >= 0 && field_type <= 3, MPM_ERR_RUNTIME, "Unknown field_type");
        auto nf = std::make_unique<mpm_near_field>();
        for (int d = 0; d < 3; ++d) {
            nf->box[d] = box[d];
            nf->E0[d] = cplx(E0[2 * d], E0[2 * d + 1]);
        }
        nf->radius.assign(radius, radius + num_radii);
        for (double r : nf->radius) MPM_REQUIRE(r > 0.0, MPM_ERR_INVALID_ARGUMENT, "radii must be positive");
        nf->xi = xi;
        nf->errortol = errortol;
        nf->field_type = field_type;
        *out = nf.release();
    });
}

void mpm_near_field_destroy(mpm_near_field* nf) { delete nf; }

mpm_status mpm_near_field_set_dipoles(mpm_near_field* nf, const double* dip, int64_t n) {
    return guarded([&] {
        MPM_REQUIRE(nf && (dip || n == 0) && n >= 0, MPM_ERR_INVALID_ARGUMENT, "null argument");
        nf->dip.resize(3 * n);
        for (int64_t k = 0; k < 3 * n; ++k) nf->dip[k] = cplx(dip[2 * k], dip[2 * k + 1]);
    });
}

mpm_status mpm_near_field_set_dipole_positions(mpm_near_field* nf, const double* x, const double* y, const double* z,
                                               int64_t n) {
    return guarded([&] {
        MPM_REQUIRE(nf && ((x && y && z) || n == 0) && n >= 0, MPM_ERR_INVALID_ARGUMENT, "null argument");
        const double* c[3] = {x, y, z};
        for (int d = 0; d < 3; ++d) nf->dpos[d].assign(c[d], c[d] + n);
    });
}

mpm_status mpm_near_field_set_field_points(mpm_near_field* nf, const double* x, const double* y, const double* z,
                                           int64_t n) {
    return guarded([&] {
        MPM_REQUIRE(nf && ((x && y && z) || n == 0) && n >= 0, MPM_ERR_INVALID_ARGUMENT, "null argument");
        const double* c[3] = {x, y, z};
        for (int d = 0; d < 3; ++d) nf->fpos[d].assign(c[d], c[d] + n);
    });
}

mpm_status mpm_near_field_set_precision(mpm_near_field* nf, int precision) {
    return guarded([&] {
        MPM_REQUIRE(nf, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(precision >= 0 && precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
        nf->precision = precision;
    });
}

mpm_status mpm_near_field_calculate(mpm_near_field* nf, double* intensity) {
    return guarded([&] {
        MPM_REQUIRE(nf && intensity, MPM_ERR_INVALID_ARGUMENT, "null argument");
        nf->calculate(intensity);
    });
}

const char* mpm_status_string(mpm_status s) {
    switch (s) {
        case MPM_OK: return "ok";
        case MPM_ERR_INVALID_ARGUMENT: return "invalid argument";
        case MPM_ERR_RUNTIME: return "runtime error";
        case MPM_ERR_HIP: return "HIP error";
        case MPM_ERR_FFT: return "rocFFT error";
        case MPM_ERR_UNSUPPORTED: return "unsupported";
        case MPM_ERR_OUT_OF_MEMORY: return "out of memory";
    }
    return "unknown status";
}

const char* mpm_last_error(void) { return mpm::g_last_error.c_str(); }

mpm_status mpm_op_create(const mpm_op_desc* desc, mpm_op** out) {
    return guarded([&] {
        MPM_REQUIRE(desc && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        auto h = std::make_unique<mpm_op>();
        h->impl = std::make_unique<mpm::EwaldOp>(*desc);
        *out = h.release();
    });
}

mpm_status mpm_op_create_poly(const mpm_op_desc* desc, const double* radii, int64_t n_radii, mpm_op** out) {
    return guarded([&] {
        MPM_REQUIRE(desc && radii && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(n_radii > 0, MPM_ERR_INVALID_ARGUMENT, "no radii");
        auto h = std::make_unique<mpm_op>();
        h->impl = std::make_unique<mpm::EwaldOp>(*desc, radii, n_radii);
        *out = h.release();
    });
}

mpm_status mpm_op_create_direct(const double* radii, int64_t n_radii, mpm_op** out) {
    return guarded([&] {
        MPM_REQUIRE(out && radii && n_radii > 0, MPM_ERR_INVALID_ARGUMENT, "null argument");
        auto o = std::make_unique<mpm_op>();
        o->impl = std::make_unique<mpm::EwaldOp>(mpm::EwaldOp::DirectTag{}, radii, n_radii);
        *out = o.release();
    });
}

void mpm_op_destroy(mpm_op* op) { delete op; }

mpm_status mpm_op_set_quads(mpm_op* op, const int* idx, int64_t nq) {
    return guarded([&] {
        MPM_REQUIRE(op, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->set_quads(idx, nq);
    });
}

mpm_status mpm_op_get_info(const mpm_op* op, mpm_op_info* info) {
    return guarded([&] {
        MPM_REQUIRE(op && info, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->info(*info);
    });
}

mpm_status mpm_op_get_pair_table(const mpm_op* op, int which, double* r, double* values, int64_t cap,
                                 int64_t* size) {
    return guarded([&] {
        MPM_REQUIRE(op && size, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *size = op->impl->pair_table(which, r, values, cap);
    });
}

mpm_status mpm_op_get_neighbors(const mpm_op* op, int64_t* start, int* j, double* perp, double* para,
                                int64_t cap, int64_t* num_pairs) {
    return guarded([&] {
        MPM_REQUIRE(op && num_pairs, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *num_pairs = op->impl->neighbors(start, j, perp, para, cap);
    });
}

mpm_status mpm_op_set_positions(mpm_op* op, const double* xyz, int64_t n) {
    return guarded([&] {
        MPM_REQUIRE(op && xyz, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->set_positions(xyz, n);
    });
}

mpm_status mpm_op_set_bloch(mpm_op* op, double kx, double ky) {
    return guarded([&] {
        MPM_REQUIRE(op, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->set_bloch(kx, ky);
    });
}

mpm_status mpm_op_set_self_coef(mpm_op* op, const double* coef, int rows) {
    return guarded([&] {
        MPM_REQUIRE(op && coef, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->set_self_coef(coef, rows);
    });
}

mpm_status mpm_op_set_self_coef_device(mpm_op* op, const double* coef, int rows, void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op && coef, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->set_self_coef_device(coef, rows, static_cast<hipStream_t>(stream));
    });
}

mpm_status mpm_op_apply(mpm_op* op, const double* x, double* y, int B, int precision, int parts,
                        void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op && x && y, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(B >= 1, MPM_ERR_INVALID_ARGUMENT, "B must be >= 1");
        MPM_REQUIRE(parts >= 1 && parts <= 3, MPM_ERR_INVALID_ARGUMENT, "parts must be 1, 2 or 3");
        MPM_REQUIRE(precision >= 0 && precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
        auto& E = *op->impl;
        MPM_REQUIRE(E.self_rows == 1 || E.self_rows >= B, MPM_ERR_INVALID_ARGUMENT,
                    "need one self-coefficient row per right-hand side (or a single shared row)");
        E.apply_contiguous(reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), B,
                           precision != MPM_PREC_DOUBLE, parts, static_cast<hipStream_t>(stream));
        MPM_HIP(hipGetLastError());
    });
}

mpm_status mpm_op_grid_bytes(const mpm_op* op, int precision, size_t* bytes) {
    return guarded([&] {
        MPM_REQUIRE(op && bytes, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(op->impl->N > 0, MPM_ERR_RUNTIME, "electricField: particle coordinates have not been set");
        *bytes = op->impl->stage_grid_bytes(precision != MPM_PREC_DOUBLE);
    });
}

mpm_status mpm_op_stage(mpm_op* op, int stage, const double* x, void* grid, double* y, int precision,
                        void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op, MPM_ERR_INVALID_ARGUMENT, "null operator");
        MPM_REQUIRE(precision >= 0 && precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
        op->impl->stage(stage, reinterpret_cast<const double2*>(x), grid, reinterpret_cast<double2*>(y),
                        precision != MPM_PREC_DOUBLE, static_cast<hipStream_t>(stream));
        MPM_HIP(hipGetLastError());
    });
}

mpm_status mpm_op_device_pointer(const mpm_op* op, const char* name, void** ptr, size_t* bytes) {
    return guarded([&] {
        MPM_REQUIRE(op && name && ptr, MPM_ERR_INVALID_ARGUMENT, "null argument");
        size_t b = 0;
        *ptr = op->impl->device_pointer(name, b);
        if (bytes) *bytes = b;
    });
}

mpm_status mpm_op_radius_gradient(mpm_op* op, const double* x, const double* g, int B, double* out, int accumulate,
                                  void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op && x && g && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        op->impl->radius_gradient(reinterpret_cast<const double2*>(x), reinterpret_cast<const double2*>(g), B, out,
                                  accumulate != 0, static_cast<hipStream_t>(stream));
    });
}

mpm_status mpm_op_autotune(mpm_op* op, int B, int precision, int flags, mpm_tune_info* info) {
    return guarded([&] {
        MPM_REQUIRE(op && info, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(precision >= 0 && precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
        MPM_REQUIRE((flags & ~(MPM_TUNE_GRID | MPM_TUNE_FORCE)) == 0, MPM_ERR_INVALID_ARGUMENT, "unknown flags");
        mpm::autotune(*op->impl, B, precision != MPM_PREC_DOUBLE, flags, nullptr, *info);
    });
}

mpm_status mpm_op_apply_noalloc(mpm_op* op, const double* x, double* y, int B, int precision, int parts,
                                void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op && x && y, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(B >= 1, MPM_ERR_INVALID_ARGUMENT, "B must be >= 1");
        MPM_REQUIRE(parts >= 1 && parts <= 3, MPM_ERR_INVALID_ARGUMENT, "parts must be 1, 2 or 3");
        MPM_REQUIRE(precision >= 0 && precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
        auto& E = *op->impl;
        MPM_REQUIRE(E.N > 0, MPM_ERR_RUNTIME, "apply before set_positions");
        MPM_REQUIRE(E.self_rows == 1 || E.self_rows >= B, MPM_ERR_INVALID_ARGUMENT,
                    "need one self-coefficient row per right-hand side (or a single shared row)");
        const bool fp32 = precision != MPM_PREC_DOUBLE;
        const size_t need = E.scratch_bytes(std::min(B, mpm::kMaxSlots), fp32);
        MPM_REQUIRE(need <= kNoAllocScratchBytes, MPM_ERR_INVALID_ARGUMENT,
                    "mpm_op_apply_noalloc: " + std::to_string(need) + " bytes of grid scratch needed, " +
                        std::to_string(kNoAllocScratchBytes) + " available (use fewer right-hand sides or mpm_op_apply)");
        mpm::EwaldOp::Workspace ws;
        ws.base = scratch_on_device();
        ws.bytes = kNoAllocScratchBytes;
        E.apply_contiguous(reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), B, fp32, parts,
                           static_cast<hipStream_t>(stream), nullptr, &ws);
        MPM_HIP(hipGetLastError());
    });
}

mpm_status mpm_solve(mpm_op* op, const mpm_solve_desc* desc, const double* b, double* x, int B,
                     mpm_solve_report* reports, void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op && desc && b && x, MPM_ERR_INVALID_ARGUMENT, "null argument");
        mpm::solve_batched(*op->impl, *desc, reinterpret_cast<const double2*>(b), reinterpret_cast<double2*>(x),
                           B, reports, static_cast<hipStream_t>(stream));
    });
}

mpm_status mpm_solve_adjoint(mpm_op* op, const mpm_solve_desc* desc, const double* g, double* y, int B,
                             mpm_solve_report* reports, void* stream) {
    return guarded([&] {
        MPM_REQUIRE(op && desc && g && y, MPM_ERR_INVALID_ARGUMENT, "null argument");
        hipStream_t s = static_cast<hipStream_t>(stream);
        // A^H = (field operator, Hermitian) + conj(self coefficients)
        struct Conj {
            mpm::EwaldOp& o;
            hipStream_t s;
            Conj(mpm::EwaldOp& op_, hipStream_t st) : o(op_), s(st) { o.conj_self_coef(s); }
            ~Conj() {
                try {
                    o.conj_self_coef(s);
                    (void)hipStreamSynchronize(s);
                } catch (...) {
                }
            }
        } conj(*op->impl, s);
        mpm::EwaldOp& E = *op->impl;
        const size_t Q = E.num_quad_slots();
        if (Q == 0) {
            mpm::solve_batched(E, *desc, reinterpret_cast<const double2*>(g), reinterpret_cast<double2*>(y), B,
                               reports, s);
            return;
        }
        // With quadrupoles A = M^-1 H, H = M A Hermitian up to the self coefficients, so
        // A^H = M conj-self(A) M^-1: solve conj-self(A) z = M^-1 g, then y = M z.
        const size_t n = E.n_unknowns(), off = 3 * (size_t)E.N;
        mpm::DevBuf<double2> gt;
        gt.reserve((size_t)B * n);
        MPM_HIP(hipMemcpyAsync(gt.p, g, (size_t)B * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
        const dim3 grid((unsigned)(((size_t)B * Q + 255) / 256));
        double2* yv = reinterpret_cast<double2*>(y);
        hipLaunchKernelGGL(quad_metric_kernel, grid, dim3(256), 0, s, gt.p, B, n, off, Q, 1);
        hipLaunchKernelGGL(quad_metric_kernel, grid, dim3(256), 0, s, yv, B, n, off, Q, 1);
        MPM_HIP(hipGetLastError());
        mpm::solve_batched(E, *desc, gt.p, yv, B, reports, s);
        hipLaunchKernelGGL(quad_metric_kernel, grid, dim3(256), 0, s, yv, B, n, off, Q, 0);
        MPM_HIP(hipGetLastError());
        MPM_HIP(hipStreamSynchronize(s));
    });
}

mpm_status mpm_dipole_solver_create(const mpm_dipole_desc* d, mpm_dipole_solver** out) {
    return guarded([&] {
        MPM_REQUIRE(d && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(d->num_wavevectors >= 1 && d->eps_cols >= 1 && d->eps_p, MPM_ERR_INVALID_ARGUMENT,
                    "eps_p is empty.");
        MPM_REQUIRE(d->num_quad_idxs >= 0 && (d->num_quad_idxs == 0 || d->quad_idxs), MPM_ERR_INVALID_ARGUMENT,
                    "null quadrupole indices");
        MPM_REQUIRE(d->tol > 0.0, MPM_ERR_INVALID_ARGUMENT, "tol must be positive");
        auto s = std::make_unique<mpm_dipole_solver>();
        s->box.assign(d->box, d->box + 3);
        s->eps_p.resize(d->num_wavevectors);
        for (int64_t w = 0; w < d->num_wavevectors; ++w) {
            s->eps_p[w].resize(d->eps_cols);
            for (int64_t c = 0; c < d->eps_cols; ++c)
                s->eps_p[w][c] = cplx(d->eps_p[2 * (w * d->eps_cols + c)], d->eps_p[2 * (w * d->eps_cols + c) + 1]);
        }
        if (d->num_radii > 0) {
            MPM_REQUIRE(d->radius, MPM_ERR_INVALID_ARGUMENT, "null radius");
            s->radius.assign(d->radius, d->radius + d->num_radii);
            for (double r : s->radius) MPM_REQUIRE(r > 0.0, MPM_ERR_INVALID_ARGUMENT, "radii must be positive");
        }
        s->eps_m = d->eps_m;
        MPM_REQUIRE(s->eps_m > 0.0, MPM_ERR_INVALID_ARGUMENT, "eps_m must be positive");
        s->xi = d->xi;
        s->tol = d->tol;
        s->quiet = d->quiet != 0;
        MPM_REQUIRE(d->guess >= 0 && d->guess <= 2, MPM_ERR_RUNTIME, "Guess type not supported.");
        s->guess = d->guess;
        MPM_REQUIRE(d->method == MPM_GMRES || d->method == MPM_BICGSTAB, MPM_ERR_INVALID_ARGUMENT, "unknown solver");
        s->method = d->method;
        MPM_REQUIRE(d->field_type >= 0 && d->field_type <= 3, MPM_ERR_RUNTIME, "Unknown field_type");
        s->field_type = d->field_type;
        if (d->num_E0 > 0) {
            MPM_REQUIRE(d->E0 && d->E0_len > 0, MPM_ERR_INVALID_ARGUMENT, "null E0");
            s->E0.resize(d->num_E0);
            for (int64_t k = 0; k < d->num_E0; ++k)
                for (int64_t c = 0; c < d->E0_len; ++c)
                    s->E0[k].push_back(cplx(d->E0[2 * (k * d->E0_len + c)], d->E0[2 * (k * d->E0_len + c) + 1]));
        } else {
            s->E0 = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
        }
        s->precision = d->precision;
        s->refined = d->refined != 0;
        s->batched = d->batched != 0;
        if (d->restart > 0) s->restart = d->restart;
        if (d->maxiter > 0) s->maxiter = d->maxiter;
        for (int k = 0; k < 3; ++k) s->grid[k] = d->grid[k];
        s->support = d->support;
        s->ewald_tol = d->ewald_tol;
        MPM_REQUIRE(d->autotune >= 0 && d->autotune <= 2, MPM_ERR_INVALID_ARGUMENT, "autotune must be 0, 1 or 2");
        s->autotune = d->autotune;
        s->solve_quads = d->solve_quadrupoles != 0;
        if (s->solve_quads && d->num_quad_idxs > 0) s->quad_idxs.assign(d->quad_idxs, d->quad_idxs + d->num_quad_idxs);
        *out = s.release();
    });
}

void mpm_dipole_solver_destroy(mpm_dipole_solver* s) { delete s; }

mpm_status mpm_dipole_solver_compute(mpm_dipole_solver* s, const double* x, const double* y, const double* z,
                                     int64_t n) {
    return guarded([&] {
        MPM_REQUIRE(s && ((x && y && z) || n == 0), MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(n >= 0, MPM_ERR_INVALID_ARGUMENT, "negative particle count");
        s->compute(x, y, z, (size_t)n);
    });
}

mpm_status mpm_dipole_solver_get_dims(const mpm_dipole_solver* s, int64_t dims[4]) {
    return guarded([&] {
        MPM_REQUIRE(s && dims, MPM_ERR_INVALID_ARGUMENT, "null argument");
        dims[0] = (int64_t)s->frames;
        dims[1] = (int64_t)(s->N ? s->W : s->eps_p.size());
        dims[2] = (int64_t)s->N;
        dims[3] = (int64_t)s->E0.size();
    });
}

mpm_status mpm_dipole_solver_get_eff_polarizability(const mpm_dipole_solver* s, int physical, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        const size_t W = s->N ? s->W : s->eps_p.size(), K = s->E0.size();
        const double f = physical ? s->eps_scale * std::pow(s->length_scale, 3.0) : 1.0;
        for (size_t q = 0; q < W * K * 3; ++q) {
            cplx acc = 0.0;
            for (size_t fr = 0; fr < s->frames; ++fr) acc += s->avg_dips[fr * W * K * 3 + q];
            const cplx v = s->frames ? acc / (double)s->frames * f : cplx(0.0);
            out[2 * q] = v.real();
            out[2 * q + 1] = v.imag();
        }
    });
}

mpm_status mpm_dipole_solver_get_dipoles(const mpm_dipole_solver* s, int physical, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        const double f = physical ? s->eps_scale * std::pow(s->length_scale, 3.0) : 1.0;
        for (size_t q = 0; q < s->dips.size(); ++q) {
            out[2 * q] = s->dips[q].real() * f;
            out[2 * q + 1] = s->dips[q].imag() * f;
        }
    });
}

mpm_status mpm_dipole_solver_get_num_quads(const mpm_dipole_solver* s, int64_t* num_quads) {
    return guarded([&] {
        MPM_REQUIRE(s && num_quads, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *num_quads = (int64_t)s->Q;
    });
}

mpm_status mpm_dipole_solver_get_quadrupoles(const mpm_dipole_solver* s, int physical, double* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        // dipole_solver.cu:640-647: quadrupoles scale with eps_m a^4
        const double f = physical ? s->eps_scale * std::pow(s->length_scale, 4.0) : 1.0;
        for (size_t q = 0; q < s->quads.size(); ++q) {
            out[2 * q] = s->quads[q].real() * f;
            out[2 * q + 1] = s->quads[q].imag() * f;
        }
    });
}

mpm_status mpm_dipole_solver_get_reports(const mpm_dipole_solver* s, mpm_solve_report* out) {
    return guarded([&] {
        MPM_REQUIRE(s && out, MPM_ERR_INVALID_ARGUMENT, "null argument");
        std::copy(s->reports.begin(), s->reports.end(), out);
    });
}

mpm_status mpm_dipole_solver_get_op_info(const mpm_dipole_solver* s, mpm_op_info* info) {
    return guarded([&] {
        MPM_REQUIRE(s && info, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(s->op, MPM_ERR_RUNTIME, "the operator exists after the first compute()");
        s->op->info(*info);
    });
}

mpm_status mpm_dipole_solver_set_use_jacobi(mpm_dipole_solver* s, int enable) {
    return guarded([&] {
        MPM_REQUIRE(s, MPM_ERR_INVALID_ARGUMENT, "null argument");
        s->use_jacobi = enable != 0;
    });
}

mpm_status mpm_dipole_solver_get_use_jacobi(const mpm_dipole_solver* s, int* enable) {
    return guarded([&] {
        MPM_REQUIRE(s && enable, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *enable = s->use_jacobi ? 1 : 0;
    });
}

mpm_status mpm_dipole_solver_get_tune_info(const mpm_dipole_solver* s, mpm_tune_info* info) {
    return guarded([&] {
        MPM_REQUIRE(s && info, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *info = s->tune;
    });
}

mpm_status mpm_dipole_solver_get_timing(const mpm_dipole_solver* s, double* solve_ms, double* setup_ms,
                                        double* total_ms) {
    return guarded([&] {
        MPM_REQUIRE(s && solve_ms && setup_ms && total_ms, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *solve_ms = s->solve_ms;
        *setup_ms = s->setup_ms;
        *total_ms = s->total_ms;
    });
}

mpm_status mpm_dipole_solver_reset(mpm_dipole_solver* s) {
    return guarded([&] {
        MPM_REQUIRE(s, MPM_ERR_INVALID_ARGUMENT, "null argument");
        s->frames = 0;
        s->dips.clear();
        s->quads.clear();
        s->avg_dips.clear();
        s->reports.clear();
    });
}

mpm_status mpm_memory_stats(int64_t* current_bytes, int64_t* peak_bytes, int reset_peak) {
    return guarded([&] {
        MPM_REQUIRE(current_bytes && peak_bytes, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *current_bytes = mpm::g_dev_bytes.load();
        *peak_bytes = mpm::g_dev_peak.load();
        if (reset_peak) mpm::g_dev_peak.store(*current_bytes);
    });
}

mpm_status mpm_buffer_alloc(size_t bytes, void** dev) {
    return guarded([&] {
        MPM_REQUIRE(dev, MPM_ERR_INVALID_ARGUMENT, "null argument");
        *dev = nullptr;
        if (bytes) MPM_HIP(hipMalloc(dev, bytes));
    });
}

mpm_status mpm_buffer_free(void* dev) {
    return guarded([&] {
        if (dev) MPM_HIP(hipFree(dev));
    });
}

mpm_status mpm_buffer_upload(void* dev, const void* host, size_t bytes) {
    return guarded([&] {
        if (bytes) MPM_HIP(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
    });
}

mpm_status mpm_buffer_download(void* host, const void* dev, size_t bytes) {
    return guarded([&] {
        if (bytes) MPM_HIP(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
    });
}

mpm_status mpm_device_synchronize(void) {
    return guarded([&] { MPM_HIP(hipDeviceSynchronize()); });
}

}  // extern "C"

namespace {
// Plain vector copy, 16 B per thread and iteration: the kernel of known traffic against which
// the HBM counters (FETCH_SIZE, WRITE_SIZE) of the operator's kernels are scaled.
__global__ void __launch_bounds__(256) profile_copy_kernel(const float4* __restrict__ src,
                                                           float4* __restrict__ dst, size_t n) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" {

mpm_status mpm_profile_enable(int enable) {
    return guarded([&] { mpm::profile_set(enable != 0); });
}

mpm_status mpm_profile_reset(void) {
    return guarded([&] { mpm::profile_reset(); });
}

mpm_status mpm_profile_get(const char* name, double* total_ms, int64_t* count, int64_t* units) {
    return guarded([&] {
        MPM_REQUIRE(name && total_ms && count, MPM_ERR_INVALID_ARGUMENT, "null argument");
        int64_t u = 0;
        mpm::profile_get(name, *total_ms, *count, u);
        if (units) *units = u;
    });
}

mpm_status mpm_profile_names(char* buf, size_t cap, size_t* needed) {
    return guarded([&] {
        MPM_REQUIRE(needed && (buf || cap == 0), MPM_ERR_INVALID_ARGUMENT, "null argument");
        const std::string names = mpm::profile_names();
        *needed = names.size() + 1;
        if (cap >= names.size() + 1) std::memcpy(buf, names.c_str(), names.size() + 1);
    });
}

mpm_status mpm_profile_copy(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        MPM_REQUIRE(dst && src, MPM_ERR_INVALID_ARGUMENT, "null argument");
        MPM_REQUIRE(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(dst) % 16 == 0 &&
                        reinterpret_cast<uintptr_t>(src) % 16 == 0,
                    MPM_ERR_INVALID_ARGUMENT, "copy needs 16-byte aligned pointers and size");
        const size_t n = bytes / 16;
        if (!n) return;
        hipLaunchKernelGGL(profile_copy_kernel, dim3(4096), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const float4*>(src), static_cast<float4*>(dst), n);
        MPM_HIP(hipGetLastError());
    });
}

}  // extern "C"
