// CPU oracle for the cuMPM hot path — TEST INFRASTRUCTURE ONLY.
//
// This file is a plain C++17/OpenMP restatement of the reference solver's
// monodisperse spectral-Ewald dipole operator and its Krylov solvers, in FP64.
// It exists to check the HIP product path (tests/, __graft_entry__.smoke()
// and the cpu_baseline leg of bench.py are its only callers).  Nothing in the
// product library links or calls it.
//
// What it restates (reference files, read-only):
//   sizing              src/cpp/electric_field/monodisperse_ewald_electric_field.cu:546-601
//   real-space tables   monodisperse_ewald_electric_field.cu:92-176, 455-500 (dipole tables,
//                       self term `selfo`)
//   table interpolation monodisperse_ewald_electric_field.cu:668-685
//   spread/contract     monodisperse_ewald_electric_field.cu:687-816, 960-1069, 1160-1298, 1488-1525
//   k-space scale       monodisperse_ewald_electric_field.cu:1300-1432
//   real space + self   monodisperse_ewald_electric_field.cu:877-958, 1589-1869
//   neighbour list      src/cpp/electric_field/neighbor_list.cu:20-223, 242-420
//   GMRES(m)            src/cpp/numerical_solver/gmres_solver.cu:56-208
//   BiCGSTAB            src/cpp/numerical_solver/bicgstab_solver.cu:58-192
//   Jacobi / BLAS-1     src/cpp/numerical_solver/numerical_solver.cu:22-276
//   polydisperse        src/cpp/electric_field/polydisperse_ewald_electric_field.cu:27-127
//                       (unique radii), 190-350 (finite-sphere spread/contract weights),
//                       565-640 (scalar k-space coefficient), 654-828 (contract, self,
//                       real space), 1434-1698 (pair functions and tables), 2039-2105 (sizing)
//   quadrupoles         monodisperse_ewald_electric_field.cu:199-470 (seven pair tables and
//                       the gradient self term), 1301-1487 (joint k-space coupling),
//                       1631-1869 (real-space couplings); the cross-term signs follow the
//                       reference's Python formulation tests/quadrupole_field.py:214-215
// The FFT is this file's own mixed-radix + Bluestein transform (the reference
// uses cuFFT Z2Z, unnormalised, forward e^{-i}, inverse e^{+i}).
//
// Pinning (see tests/test_oracle*.py): golden fixtures
// tests/golden/ref_dipole_monodisperse.npz and ref_quadrupoles_monodisperse.npz,
// numpy.fft, a textbook point-dipole Ewald lattice sum (non-overlapping spheres),
// quadrature of the overlapping-sphere field (1 <= r < 2), xi-independence of
// real+recip (every dipole and quadrupole coupling), the quadrupole tables
// (overlap polynomials and the gradient self term included) against an
// electrostatic model of the polarised spheres evaluated by real- and k-space
// quadrature, and a dense np.linalg.solve for the Krylov solvers.  Parity
// unpinned: results at the cuFFT boundary (no reference test holds them; the
// FFT is pinned by numpy.fft).

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using cplx = std::complex<double>;
static const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// 1-D FFT: recursive mixed radix for lengths whose prime factors are <= 31,
// Bluestein (chirp-z over a power-of-two convolution) otherwise.
// ---------------------------------------------------------------------------
namespace fft {

struct Plan1D {
    int n = 0;
    std::vector<int> factors;     // radices, product == n (mixed radix path)
    std::vector<cplx> tw;         // exp(-2 pi i j / n), j < n
    bool bluestein = false;
    int m = 0;                    // Bluestein convolution length (power of two)
    std::vector<cplx> chirp;      // exp(-i pi j^2 / n), j < n
    std::vector<cplx> bhat;       // FFT_m of the conjugate chirp, wrapped
    std::unique_ptr<Plan1D> sub;  // power-of-two plan of length m
};

static void factorize(int n, std::vector<int>& f) {
    f.clear();
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    while (n % 2 == 0) { f.push_back(2); n /= 2; }
    for (int p = 3; p * p <= n; p += 2)
        while (n % p == 0) { f.push_back(p); n /= p; }
    if (n > 1) f.push_back(n);
}

static std::unique_ptr<Plan1D> make_plan(int n);

static void twiddles(int n, std::vector<cplx>& tw) {
    tw.resize(n);
    for (int j = 0; j < n; ++j) {
        double a = -2.0 * kPi * (double)j / (double)n;
        tw[j] = cplx(std::cos(a), std::sin(a));
    }
}

static std::unique_ptr<Plan1D> make_plan(int n) {
    auto p = std::make_unique<Plan1D>();
    p->n = n;
    std::vector<int> f;
    factorize(n, f);
    int maxf = 1;
    for (int x : f) maxf = std::max(maxf, x);
    if (maxf > 31) {
        p->bluestein = true;
        int m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p->m = m;
        p->chirp.resize(n);
        const long long two_n = 2LL * n;
        for (int j = 0; j < n; ++j) {
            long long jj = ((long long)j * j) % two_n;  // keep the phase argument small
            double a = -kPi * (double)jj / (double)n;
            p->chirp[j] = cplx(std::cos(a), std::sin(a));
        }
        p->sub = make_plan(m);
        std::vector<cplx> b(m, cplx(0, 0));
        b[0] = std::conj(p->chirp[0]);
        for (int j = 1; j < n; ++j) b[j] = b[m - j] = std::conj(p->chirp[j]);
        p->bhat = b;
        // forward transform of b (in place through a scratch copy)
        std::vector<cplx> out(m), scratch(m);
        extern void exec_plan(const Plan1D&, const cplx*, int, cplx*, cplx*);
        exec_plan(*p->sub, b.data(), 1, out.data(), scratch.data());
        p->bhat = out;
    } else {
        p->factors = f;
        twiddles(n, p->tw);
    }
    return p;
}

// out[k] = sum_j in[j*stride] * exp(-2 pi i j k / n) for the sub-length n
// reached after `level` radix splits (tw is the full-length table, twstep maps
// the sub-length twiddle index to the full table).
static void rec(const Plan1D& P, int level, int n, const cplx* in, long stride,
                cplx* out, int twstep) {
    if (n == 1) { out[0] = in[0]; return; }
    const int p = P.factors[level];
    const int m = n / p;
    for (int r = 0; r < p; ++r)
        rec(P, level + 1, m, in + r * stride, stride * p, out + r * m, twstep * p);
    // combine: X[k + q m] = sum_r (W_n^{r k} Y_r[k]) W_p^{r q}
    cplx tmp[32];
    cplx y[32];
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) {
            long idx = (long)r * k * twstep;
            y[r] = out[r * m + k] * P.tw[idx % P.n];
        }
        for (int q = 0; q < p; ++q) {
            cplx s(0, 0);
            for (int r = 0; r < p; ++r) {
                long idx = ((long)r * q % p) * (P.n / p);
                s += y[r] * P.tw[idx];
            }
            tmp[q] = s;
        }
        for (int q = 0; q < p; ++q) out[k + q * m] = tmp[q];
    }
}

void exec_plan(const Plan1D& P, const cplx* in, int stride, cplx* out, cplx* scratch) {
    if (!P.bluestein) {
        rec(P, 0, P.n, in, stride, out, 1);
        return;
    }
    const int n = P.n, m = P.m;
    std::vector<cplx> a(m, cplx(0, 0)), A(m), s2(m);
    for (int j = 0; j < n; ++j) a[j] = in[(long)j * stride] * P.chirp[j];
    exec_plan(*P.sub, a.data(), 1, A.data(), s2.data());
    for (int j = 0; j < m; ++j) A[j] *= P.bhat[j];
    // inverse via conjugation
    for (int j = 0; j < m; ++j) A[j] = std::conj(A[j]);
    exec_plan(*P.sub, A.data(), 1, a.data(), s2.data());
    const double inv_m = 1.0 / (double)m;
    for (int k = 0; k < n; ++k) out[k] = std::conj(a[k]) * inv_m * P.chirp[k];
    (void)scratch;
}

// In-place 3-D transform of a row-major [nx][ny][nz] complex array.
// sign = -1: forward e^{-i}; sign = +1: inverse e^{+i}; unnormalised.
static void fft3d(cplx* data, int nx, int ny, int nz, int sign) {
    const int dims[3] = {nx, ny, nz};
    const long strides[3] = {(long)ny * nz, (long)nz, 1};
    for (int ax = 0; ax < 3; ++ax) {
        const int n = dims[ax];
        if (n == 1) continue;
        auto plan = make_plan(n);
        const long stride = strides[ax];
        // enumerate all lines along `ax`
        const int o1 = (ax == 0) ? 1 : 0;          // the two other axes
        const int o2 = (ax == 2) ? 1 : 2;
        const long nlines = (long)dims[o1] * dims[o2];
#pragma omp parallel
        {
            std::vector<cplx> line(n), out(n), scratch(n);
#pragma omp for schedule(static)
            for (long L = 0; L < nlines; ++L) {
                const long a = L / dims[o2], b = L % dims[o2];
                const long base = a * strides[o1] + b * strides[o2];
                for (int j = 0; j < n; ++j) {
                    cplx v = data[base + j * stride];
                    line[j] = (sign > 0) ? std::conj(v) : v;
                }
                exec_plan(*plan, line.data(), 1, out.data(), scratch.data());
                for (int j = 0; j < n; ++j)
                    data[base + j * stride] = (sign > 0) ? std::conj(out[j]) : out[j];
            }
        }
    }
}

}  // namespace fft

// ---------------------------------------------------------------------------
// Ewald sizing and real-space tables (monodisperse, radius 1 after nondim).
// ---------------------------------------------------------------------------
struct Sizing {
    double rc = 0, kcut = 0;
    int n[3] = {0, 0, 0};
    double h[3] = {0, 0, 0};
    double eta[3] = {0, 0, 0};
    int P = 0;       // num_grid_gaussian
    int half = 0;    // P / 2
    int width = 0;   // 2*half + 1 stencil points per axis
};

// monodisperse_ewald_electric_field.cu:546-601.  grid_override / P_override
// replace the computed grid or support when positive (the reference ties them
// to errortol; the overrides decouple them from it, as the GPU operator's do).
static int compute_sizing(const double L[3], double errortol, double xi,
                          const int* grid_override, int P_override, Sizing& S,
                          std::string& err) {
    S.rc = std::sqrt(-std::log(errortol)) / xi;
    if (S.rc < 2.0) S.rc = 2.0;
    for (int d = 0; d < 3; ++d) {
        if (S.rc > L[d] / 2.0) {
            err = "Real space cutoff (" + std::to_string(S.rc) +
                  ") larger than half the box length.";
            return -1;
        }
    }
    S.kcut = 2.0 * xi * xi * S.rc;
    for (int d = 0; d < 3; ++d) {
        S.n[d] = (int)std::ceil(1.0 + L[d] * S.kcut / kPi);
        if (grid_override && grid_override[d] > 0) S.n[d] = grid_override[d];
        S.h[d] = L[d] / S.n[d];
    }
    S.P = (int)std::ceil(-2.0 * std::log(errortol) / kPi);
    if (P_override > 0) S.P = P_override;
    for (int d = 0; d < 3; ++d) S.eta[d] = S.P * std::pow(S.h[d] * xi, 2) / kPi;
    S.half = S.P / 2;
    S.width = 2 * S.half + 1;
    return 0;
}

// Volume-averaged field of a uniformly polarised unit sphere on another unit
// sphere at centre distance r, split into the real-space Ewald part.  Each
// function is a sum over the three "images" of the pair kernel at r-2, r, r+2:
//   f(r) = sum_s [A_s(r) exp(-(r+s)^2 xi^2) + B_s(r) erfc((r+s) xi)] + reg(r) [r<2]
// with s in {+2, -2, 0}.  monodisperse_ewald_electric_field.cu:111-176.
struct DipoleTables {
    std::vector<double> r, perp, para;  // 9001 entries, entry 0 = self (r=0)
    double self_perp = 0;
};

static void dipole_pair_functions(double r, double xi, double& f_perp, double& f_para) {
    const double x2 = xi * xi, x4 = x2 * x2, x5 = x4 * xi, x6 = x4 * x2;
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r4 * r2;
    const double sp = std::pow(kPi, 1.5);
    const double Ep = std::exp(-(r + 2.0) * (r + 2.0) * x2);
    const double Em = std::exp(-(r - 2.0) * (r - 2.0) * x2);
    const double E0 = std::exp(-r2 * x2);
    const double Cp = std::erfc((r + 2.0) * xi);
    const double Cm = std::erfc((r - 2.0) * xi);
    const double C0 = std::erfc(r * xi);
    // common polynomial pieces
    const double c1 = 3.0 - 12.0 * x2 + 32.0 * x4;
    const double c0 = 2.0 * (3.0 + 4.0 * x2 - 32.0 * x4);
    const double q4 = 36.0 * x4 * (1.0 - 4.0 * x2);

    // ---- perpendicular (field_dip_1) ----
    {
        const double pre_e = 1.0 / (1024.0 * sp * x5 * r3);
        const double odd = 4.0 * x4 * r5 + 8.0 * x2 * (2.0 - 7.0 * x2) * r3 + c1 * r;
        const double even = -8.0 * x4 * r4 - 8.0 * x2 * (3.0 + 2.0 * x2) * r2 + c0;
        const double Ap = pre_e * (odd + even);
        const double Am = pre_e * (odd - even);
        const double A0 = (-4.0 * x4 * r4 - 8.0 * x2 * (2.0 - 9.0 * x2) * r2 - 3.0 + 36.0 * x2) /
                          (512.0 * sp * x5 * r2);
        const double pre_b = 1.0 / (2048.0 * kPi * x6 * r3);
        const double bev = -8.0 * x6 * r6 - q4 * r4 - 18.0 * x2 * (1.0 - 8.0 * x2) * r2 + 3.0 -
                           36.0 * x2 + 256.0 * x6;
        const double Bp = pre_b * (bev + 256.0 * x6 * r3);
        const double Bm = pre_b * (bev - 256.0 * x6 * r3);
        const double B0 = (8.0 * x6 * r6 + q4 * r4 + 18.0 * x2 * (1.0 - 8.0 * x2) * r2 - 3.0 + 36.0 * x2) /
                          (1024.0 * kPi * x6 * r3);
        double f = Ap * Ep + Bp * Cp + Am * Em + Bm * Cm + A0 * E0 + B0 * C0;
        if (r < 2.0) f += -1.0 / (4.0 * kPi * r3) + (1.0 - 9.0 * r / 16.0 + r3 / 32.0) / (4.0 * kPi);
        f_perp = f;
    }
    // ---- parallel (field_dip_2) ----
    {
        const double pre_e = 1.0 / (512.0 * sp * x5 * r3);
        const double odd = 8.0 * x4 * r5 + 2.0 * x2 * (7.0 - 20.0 * x2) * r3 - c1 * r;
        const double even = -16.0 * x4 * r4 - 4.0 * x2 * (3.0 - 4.0 * x2) * r2 - c0;
        const double Ap = pre_e * (odd + even);
        const double Am = pre_e * (odd - even);
        const double A0 = (-8.0 * x4 * r4 - 2.0 * x2 * (7.0 - 36.0 * x2) * r2 + 3.0 - 36.0 * x2) /
                          (256.0 * sp * x5 * r2);
        const double pre_b = 1.0 / (1024.0 * kPi * x6 * r3);
        const double bev = -16.0 * x6 * r6 - q4 * r4 - 3.0 + 36.0 * x2 - 256.0 * x6;
        const double Bp = pre_b * (bev + 128.0 * x6 * r3);
        const double Bm = pre_b * (bev - 128.0 * x6 * r3);
        const double B0 = (16.0 * x6 * r6 + q4 * r4 + 3.0 - 36.0 * x2) / (512.0 * kPi * x6 * r3);
        double f = Ap * Ep + Bp * Cp + Am * Em + Bm * Cm + A0 * E0 + B0 * C0;
        if (r < 2.0) f += 1.0 / (2.0 * kPi * r3) + (1.0 - 9.0 * r / 8.0 + r3 / 8.0) / (4.0 * kPi);
        f_para = f;
    }
}

// monodisperse_ewald_electric_field.cu:455-500: 9000 samples r = 1 + 0.001 j,
// preceded by the r = 0 self entry.
static void build_dipole_tables(double xi, DipoleTables& T) {
    const int nsteps = 9000;
    T.r.assign(nsteps + 1, 0.0);
    T.perp.assign(nsteps + 1, 0.0);
    T.para.assign(nsteps + 1, 0.0);
    const double x2 = xi * xi, x3 = x2 * xi;
    const double selfo = (-1.0 + 6.0 * x2 + (1.0 - 2.0 * x2) * std::exp(-4.0 * x2)) /
                             (16.0 * std::pow(kPi, 1.5) * x3) +
                         std::erfc(2.0 * xi) / (4.0 * kPi);
    T.self_perp = selfo;
    T.r[0] = 0.0;
    T.perp[0] = selfo;
    T.para[0] = selfo;
    for (int j = 0; j < nsteps; ++j) {
        const double r = 1.0 + j * 0.001;
        double fp, fa;
        dipole_pair_functions(r, xi, fp, fa);
        T.r[j + 1] = r;
        T.perp[j + 1] = fp;
        T.para[j + 1] = fa;
    }
}

// ---------------------------------------------------------------------------
// Quadrupole pair functions (monodisperse_ewald_electric_field.cu:199-470).
// Each is  A+ e^{-(r+2)^2 xi^2} + B+ erfc((r+2) xi) + A- e^{-(r-2)^2 xi^2}
//        + B- erfc((r-2) xi) + A0 e^{-r^2 xi^2} + B0 erfc(r xi) + [r < 2] reg,
// with A+- = a (O +- E) for the odd and even powers O, E of r and B+- = b (E' +- O').
// Order: field_quad_1..3 (E from Q: perp_Q, para_Q, Q3), grad_quad_1..4 (G1..G4).
// ---------------------------------------------------------------------------
struct Img {
    double Ap, Am, A0, Bp, Bm, B0, reg;
};
static double img_sum(double r, double xi, const Img& c) {
    const double x2 = xi * xi;
    double f = c.Ap * std::exp(-(r + 2.0) * (r + 2.0) * x2) + c.Bp * std::erfc((r + 2.0) * xi) +
               c.Am * std::exp(-(r - 2.0) * (r - 2.0) * x2) + c.Bm * std::erfc((r - 2.0) * xi) +
               c.A0 * std::exp(-r * r * x2) + c.B0 * std::erfc(r * xi);
    if (r < 2.0) f += c.reg;
    return f;
}

static void quad_pair_functions(double r, double xi, double out[7]) {
    const double x = xi, x2 = x * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2, x7 = x6 * x, x8 = x4 * x4,
                 x9 = x8 * x, x10 = x8 * x2;
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r4 * r2, r7 = r6 * r, r8 = r4 * r4,
                 r9 = r8 * r, r10 = r8 * r2;
    const double s15 = std::pow(kPi, 1.5), pi = kPi;
    Img c;
    {   // field_quad_1
        const double a = 15.0 / (16384.0 * s15 * x7 * r4);
        const double O = -24.0 * x6 * r7 - 4.0 * x4 * (9.0 - 8.0 * x2) * r5 + 2.0 * x2 * (21.0 - 80.0 * x2 + 64.0 * x4) * r3 -
                         (45.0 - 120.0 * x2 + 192.0 * x4 - 512.0 * x6) * r;
        const double E = 48.0 * x6 * r6 + 8.0 * x4 * (3.0 - 8.0 * x2) * r4 - 4.0 * x2 * std::pow(3.0 - 8.0 * x2, 2) * r2 -
                         2.0 * (45.0 + 24.0 * x2 - 64.0 * x4 + 512.0 * x6);
        const double b = 15.0 / (32768.0 * pi * x8 * r4);
        const double Be = 48.0 * x8 * r8 + 32.0 * x6 * (3.0 - 8.0 * x2) * r6 - 24.0 * x4 * (3.0 - 16.0 * x2) * r4 +
                          72.0 * x2 * (1.0 - 8.0 * x2) * r2 - 45.0 + 480.0 * x2 + 4096.0 * x8;
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = 15.0 / (8192.0 * s15 * x7 * r3) *
               (24.0 * x6 * r6 + 4.0 * x4 * (9.0 - 32.0 * x2) * r4 - 2.0 * x2 * (21.0 - 128.0 * x2) * r2 + 45.0 - 480.0 * x2);
        c.Bp = c.Bm = b * Be;
        c.B0 = 15.0 / (16384.0 * pi * x8 * r4) * (-(Be - 4096.0 * x8));
        c.reg = -15.0 / (4.0 * pi * r4) + 15.0 * r2 / (64.0 * pi) * (1.0 - 3.0 * r2 / 16.0);
        out[0] = img_sum(r, xi, c);
    }
    {   // field_quad_2
        const double a = 3.0 / (16384.0 * s15 * x7 * r4);
        const double O = -40.0 * x6 * r7 - 20.0 * x4 * (11.0 - 24.0 * x2) * r5 - 2.0 * x2 * (45.0 - 80.0 * x2 + 64.0 * x4) * r3 +
                         (45.0 - 120.0 * x2 + 192.0 * x4 - 512.0 * x6) * r;
        const double E = 80.0 * x6 * r6 + 8.0 * x4 * (45.0 + 8.0 * x2) * r4 - 4.0 * x2 * (15.0 + 48.0 * x2 - 64.0 * x4) * r2 +
                         2.0 * (45.0 + 24.0 * x2 - 64.0 * x4 + 512.0 * x6);
        const double b = 3.0 / (32768.0 * pi * x8 * r4);
        const double Be = 80.0 * x8 * r8 + 160.0 * x6 * (3.0 - 8.0 * x2) * r6 + 120.0 * x4 * (3.0 - 16.0 * x2) * r4 -
                          120.0 * x2 * (1.0 - 8.0 * x2) * r2 + 45.0 - 480.0 * x2 - 4096.0 * x8;
        const double Bo = -2048.0 * x8 * r5;
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = 15.0 / (8192.0 * s15 * x7 * r3) *
               (8.0 * x6 * r6 + 4.0 * x4 * (11.0 - 32.0 * x2) * r4 + 2.0 * x2 * (9.0 - 64.0 * x2) * r2 - 9.0 + 96.0 * x2);
        c.Bp = b * (Be + Bo);
        c.Bm = b * (Be - Bo);
        c.B0 = 15.0 / (16384.0 * pi * x8 * r4) *
               (-16.0 * x8 * r8 - 32.0 * x6 * (3.0 - 8.0 * x2) * r6 - 24.0 * x4 * (3.0 - 16.0 * x2) * r4 +
                24.0 * x2 * (1.0 - 8.0 * x2) * r2 - 9.0 + 96.0 * x2);
        c.reg = 3.0 / (4.0 * pi * r4) - 3.0 * r / (8.0 * pi) * (1.0 - 5.0 * r / 8.0 + 5.0 * r3 / 128.0);
        out[1] = img_sum(r, xi, c);
    }
    {   // field_quad_3
        const double a = 5.0 / (1024.0 * s15 * x5 * r2);
        const double O = 4.0 * x4 * r5 + 16.0 * x2 * (1.0 - 2.0 * x2) * r3 + 3.0 * r;
        const double E = -8.0 * x4 * r4 - 24.0 * x2 * r2 + 6.0;
        const double b = 5.0 / (2048.0 * pi * x6 * r2);
        const double Be = -8.0 * x6 * r6 - 12.0 * x4 * (3.0 - 8.0 * x2) * r4 - 6.0 * x2 * (3.0 - 16.0 * x2) * r2 + 3.0 - 24.0 * x2;
        const double Bo = 128.0 * x6 * r3;
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = 5.0 / (512.0 * s15 * x5 * r) * (-4.0 * x4 * r4 - 16.0 * x2 * (1.0 - 3.0 * x2) * r2 - 3.0 + 24.0 * x2);
        c.Bp = b * (Be + Bo);
        c.Bm = b * (Be - Bo);
        c.B0 = 5.0 / (1024.0 * pi * x6 * r2) * (-Be);
        c.reg = 5.0 * r / (8.0 * pi) * (1.0 - 3.0 * r / 4.0 + r3 / 16.0);
        out[2] = img_sum(r, xi, c);
    }
    {   // grad_quad_1
        const double a = 75.0 / (16384.0 * s15 * x7 * r3);
        const double O = 8.0 * x6 * r7 + (44.0 * x4 - 32.0 * x6) * r5 + (18.0 * x2 + 32.0 * x4) * r3 - (9.0 + 24.0 * x2) * r;
        const double E = -16.0 * x6 * r6 - (72.0 * x4 - 64.0 * x6) * r4 + (12.0 * x2 - 64.0 * x4) * r2 - (18.0 - 48.0 * x2);
        const double Be = 16.0 * x8 * r8 + (96.0 * x6 - 128.0 * x8) * r6 + (72.0 * x4 - 192.0 * x6 + 256.0 * x8) * r4 -
                          (24.0 * x2 - 96.0 * x4 + 256.0 * x6) * r2 + 9.0 - 48.0 * x2 + 192.0 * x4;
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = -75.0 / (8192.0 * s15 * x7 * r2) *
               (8.0 * x6 * r6 + (44.0 * x4 - 64.0 * x6) * r4 + (18.0 * x2 - 64.0 * x4 + 128.0 * x6) * r2 - 9.0 + 48.0 * x2 -
                192.0 * x4);
        c.Bp = c.Bm = -75.0 / (32768.0 * pi * x8 * r3) * Be;
        c.B0 = 75.0 / (16384.0 * pi * x8 * r3) * Be;
        c.reg = 75.0 * r * (4.0 - r2) * (4.0 - r2) / (1024.0 * pi);
        out[3] = img_sum(r, xi, c);
    }
    {   // grad_quad_2
        const double a = 3.0 / (32768.0 * s15 * x9 * r5);
        const double O = -48.0 * x8 * r9 - (576.0 * x6 - 608.0 * x8) * r7 - (1536.0 * x4 - 2576.0 * x6 + 3968.0 * x8) * r5 -
                         (360.0 * x2 + 360.0 * x4 + 640.0 * x6 - 512.0 * x8) * r3 +
                         (135.0 + 180.0 * x2 + 480.0 * x4 - 768.0 * x6 + 2048.0 * x8) * r;
        const double E = 96.0 * x8 * r8 + (1056.0 * x6 - 1216.0 * x8) * r6 + (2160.0 * x4 - 4320.0 * x6 - 256.0 * x8) * r4 -
                         (360.0 * x2 - 1680.0 * x4 - 768.0 * x6 + 1024.0 * x8) * r2 +
                         (270.0 - 1080.0 * x2 - 192.0 * x4 + 512.0 * x6 - 4096.0 * x8);
        const double P = 288.0 * x10 * r10 + (3600.0 * x8 - 4800.0 * x10) * r8 +
                         (10800.0 * x6 - 28800.0 * x8 + 38400.0 * x10) * r6 + (5400.0 * x4 - 21600.0 * x6 + 57600.0 * x8) * r4 -
                         (1350.0 * x2 - 7200.0 * x4 + 28800.0 * x6) * r2 + 405.0 - 2700.0 * x2 + 14400.0 * x4;
        const double b = 1.0 / (65536.0 * pi * x10 * r5);
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = 1.0 / (16384.0 * s15 * x9 * r4) *
               (144.0 * x8 * r8 + (1728.0 * x6 - 2400.0 * x8) * r6 + (4608.0 * x4 - 13200.0 * x6 + 19200.0 * x8) * r4 +
                (1080.0 * x2 - 5400.0 * x4 + 19200.0 * x6) * r2 - 405.0 + 2700.0 * x2 - 14400.0 * x4);
        c.Bp = b * (P + 49152.0 * x10 + 49152.0 * x10 * r5);
        c.Bm = b * (P + 49152.0 * x10 - 49152.0 * x10 * r5);
        c.B0 = -1.0 / (32768.0 * pi * x10 * r5) * P;
        c.reg = -3.0 / (2.0 * pi) * (1.0 / r5 - 1.0 + 25.0 * r / 32.0 - 25.0 * r3 / 256.0 + 3.0 * r5 / 512.0);
        out[4] = img_sum(r, xi, c);
    }
    {   // grad_quad_3
        const double a = -15.0 / (65536.0 * s15 * x9 * r5);
        const double O = 48.0 * x8 * r9 + (336.0 * x6 - 288.0 * x8) * r7 + (216.0 * x4 + 144.0 * x6 + 128.0 * x8) * r5 -
                         (180.0 * x2 - 120.0 * x4 + 640.0 * x6 - 512.0 * x8) * r3 +
                         (135.0 + 180.0 * x2 + 480.0 * x4 - 768.0 * x6 + 2048.0 * x8) * r;
        const double E = -96.0 * x8 * r8 - 576.0 * (x6 - x8) * r6 - (480.0 * x6 + 256.0 * x8) * r4 +
                         (720.0 * x4 + 768.0 * x6 - 1024.0 * x8) * r2 -
                         (-270.0 + 1080.0 * x2 + 192.0 * x4 - 512.0 * x6 + 4096.0 * x8);
        const double P = -96.0 * x10 * r10 + (-720.0 * x8 + 960.0 * x10) * r8 -
                         (720.0 * x6 - 1920.0 * x8 + 2560.0 * x10) * r6 + (360.0 * x4 - 1440.0 * x6 + 3840.0 * x8) * r4 -
                         (270.0 * x2 - 1440.0 * x4 + 5760.0 * x6) * r2 + 135.0 - 900.0 * x2 + 4800.0 * x4;
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = 15.0 / (32768.0 * s15 * x9 * r4) *
               (48.0 * x8 * r8 + (336.0 * x6 - 480.0 * x8) * r6 + (216.0 * x4 - 720.0 * x6 + 1280.0 * x8) * r4 -
                (180.0 * x2 - 840.0 * x4 + 2560.0 * x6) * r2 + 135.0 - 900.0 * x2 + 4800.0 * x4);
        c.Bp = c.Bm = -15.0 / (131072.0 * pi * x10 * r5) * (P + 16384.0 * x10);
        c.B0 = -15.0 / (65536.0 * pi * x10 * r5) * (-P);
        c.reg = -15.0 * std::pow(r2 - 4.0, 3) * (3.0 * r4 + 6.0 * r2 + 8.0) / (2048.0 * pi * r5);
        out[5] = img_sum(r, xi, c);
    }
    {   // grad_quad_4
        const double a = -15.0 / (65536.0 * s15 * x9 * r5);
        const double O = 144.0 * x8 * r9 + (288.0 * x6 + 96.0 * x8) * r7 - (432.0 * x4 - 912.0 * x6 + 896.0 * x8) * r5 +
                         (720.0 * x2 - 2280.0 * x4 + 4480.0 * x6 - 3584.0 * x8) * r3 -
                         (945.0 + 1260.0 * x2 + 3360.0 * x4 - 5376.0 * x6 + 14336.0 * x8) * r;
        const double E = -288.0 * x8 * r8 - (288.0 * x6 + 192.0 * x8) * r6 + (720.0 * x4 - 480.0 * x6 + 1792.0 * x8) * r4 -
                         (1080.0 * x2 + 2160.0 * x4 + 5376.0 * x6 - 7168.0 * x8) * r2 -
                         (1890.0 - 7560.0 * x2 - 1344.0 * x4 + 3584.0 * x6 - 28672.0 * x8);
        const double P = 288.0 * x10 * r10 + (720.0 * x8 - 960.0 * x10) * r8 - (720.0 * x6 - 1920.0 * x8 + 2560.0 * x10) * r6 +
                         (1080.0 * x4 - 4320.0 * x6 + 11520.0 * x8) * r4 - (1350.0 * x2 - 7200.0 * x4 + 28800.0 * x6) * r2 +
                         945.0 - 6300.0 * x2 + 33600.0 * x4;
        c.Ap = a * (O + E);
        c.Am = a * (O - E);
        c.A0 = 15.0 / (32768.0 * s15 * x9 * r4) *
               (144.0 * x8 * r8 + (288.0 * x6 - 480.0 * x8) * r6 - (432.0 * x4 - 1200.0 * x6 + 1280.0 * x8) * r4 +
                (720.0 * x2 - 3000.0 * x4 + 6400.0 * x6) * r2 - 945.0 + 6300.0 * x2 - 33600.0 * x4);
        c.Bp = c.Bm = 15.0 / (131072.0 * pi * x10 * r5) * (P + 114688.0 * x10);
        c.B0 = -15.0 / (65536.0 * pi * x10 * r5) * P;
        c.reg = -15.0 * (3584.0 - 80.0 * r6 - 30.0 * r8 + 9.0 * r10) / (2048.0 * pi * r5);
        out[6] = img_sum(r, xi, c);
    }
}

// Gradient self term, the quadrupole tables' entry 0 (:465-466).
static double quad_self_grad(double xi) {
    const double x2 = xi * xi, x4 = x2 * x2, x5 = x4 * xi, sp = std::sqrt(kPi);
    return 0.5 / kPi *
           (3.0 * (3.0 - 10.0 * x2 + 20.0 * x4) / (8.0 * sp * x5) -
            3.0 * (3.0 + 2.0 * x2 + 4.0 * x4) * std::exp(-4.0 * x2) / (8.0 * sp * x5) + 3.0 * std::erfc(2.0 * xi));
}

// monodisperse_ewald_electric_field.cu:668-685 (clamps at the last sample).
static double interp(double r, const std::vector<double>& rt, const std::vector<double>& y) {
    const size_t ts = y.size();
    if (r <= 0.0) return y[0];
    if (r >= rt[ts - 1]) return y[ts - 1];
    if (r < 1.0) return y[0] * (1.0 - r) + y[1] * r;
    size_t idx = 1 + (size_t)((r - 1.0) / 0.001);
    if (idx >= ts - 1) return y[ts - 1];
    const double r0 = rt[idx], r1 = rt[idx + 1];
    const double t = (r - r0) / (r1 - r0);
    return y[idx] * (1.0 - t) + y[idx + 1] * t;
}

// ---------------------------------------------------------------------------
// Polydisperse pieces (polydisperse_ewald_electric_field.cu)
// ---------------------------------------------------------------------------
// Real-space pair functions of uniformly polarised spheres of radii a, b at
// centre distance r.  Both are odd extensions in the radii:
//   f(r; a, b) = sum_{s, t = +-1} [ A(r; s a, t b) exp(-(r + s a + t b)^2 xi^2)
//                                  + C(r; s a, t b) erfc((r + s a + t b) xi) ] + reg(r; a, b)
// where A and C include the factor 1/(a^3 b^3) of their own (signed) arguments,
// which supplies the signs of the four terms.  polydisperse...cu:1434-1621.
struct PolyCoefs {
    double A, C;
};
static PolyCoefs poly_perp_terms(double r, double a, double b, double xi) {
    const double x2 = xi * xi, x4 = x2 * x2, x5 = x4 * xi, x6 = x4 * x2;
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r4 * r2;
    const double a2 = a * a, b2 = b * b, a3 = a2 * a, b3 = b2 * b, s = a + b;
    const double q = a2 - 4.0 * a * b + b2;
    const double c3 = 4.0 * a3 - 3.0 * a2 * b - 3.0 * a * b2 + 4.0 * b3;
    const double ab3 = a3 * b3;
    const double A = (4.0 * x4 * r5 - 4.0 * s * x4 * r4 + (16.0 * x2 + 8.0 * (-4.0 * a2 + a * b - 4.0 * b2) * x4) * r3 +
                      (-12.0 * s * x2 - 8.0 * c3 * x4) * r2 +
                      (3.0 - 12.0 * (a2 - a * b + b2) * x2 - 4.0 * s * s * q * x4) * r + 3.0 * s + 4.0 * c3 * x2 +
                      4.0 * s * s * s * q * x4) /
                     (1024.0 * std::pow(kPi, 1.5) * ab3 * r3 * x5);
    const double d2 = (a2 - b2) * (a2 - b2);
    const double C = (-8.0 * x6 * r6 + 36.0 * x4 * (-1.0 + 2.0 * (a2 + b2) * x2) * r4 + 128.0 * (a3 + b3) * x6 * r3 +
                      (-18.0 * x2 + 72.0 * (a2 + b2) * x4 + 72.0 * d2 * x6) * r2 + 3.0 - 18.0 * (a2 + b2) * x2 -
                      36.0 * d2 * x4 - 8.0 * s * s * s * s * q * x6) /
                     (2048.0 * kPi * ab3 * r3 * x6);
    return {A, C};
}
static PolyCoefs poly_para_terms(double r, double a, double b, double xi) {
    const double x2 = xi * xi, x4 = x2 * x2, x5 = x4 * xi, x6 = x4 * x2;
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r4 * r2;
    const double a2 = a * a, b2 = b * b, a3 = a2 * a, b3 = b2 * b, s = a + b;
    const double q = a2 - 4.0 * a * b + b2;
    const double c3 = 4.0 * a3 - 3.0 * a2 * b - 3.0 * a * b2 + 4.0 * b3;
    const double e3 = a3 - 3.0 * a2 * b - 3.0 * a * b2 + b3;
    const double ab3 = a3 * b3;
    const double A = (8.0 * x4 * r5 - 8.0 * s * x4 * r4 + (14.0 * x2 - 4.0 * (7.0 * a2 - 4.0 * a * b + 7.0 * b2) * x4) * r3 +
                      (-6.0 * s * x2 - 4.0 * e3 * x4) * r2 +
                      (-3.0 + 12.0 * (a2 - a * b + b2) * x2 + 4.0 * s * s * q * x4) * r - 3.0 * s - 4.0 * c3 * x2 -
                      4.0 * s * s * s * q * x4) /
                     (512.0 * std::pow(kPi, 1.5) * ab3 * r3 * x5);
    const double d2 = (a2 - b2) * (a2 - b2);
    const double C = (-16.0 * x6 * r6 + 36.0 * x4 * (-1.0 + 2.0 * (a2 + b2) * x2) * r4 + 64.0 * (a3 + b3) * x6 * r3 - 3.0 +
                      18.0 * (a2 + b2) * x2 + 36.0 * d2 * x4 + 8.0 * s * s * s * s * q * x6) /
                     (1024.0 * kPi * ab3 * r3 * x6);
    return {A, C};
}

static void poly_pair_functions(double r, double a, double b, double xi, double& fperp, double& fpara) {
    double fp = 0.0, fa = 0.0;
    for (int sa = -1; sa <= 1; sa += 2)
        for (int sb = -1; sb <= 1; sb += 2) {
            const double aa = sa * a, bb = sb * b;
            const double u = r + aa + bb;
            const double e = std::exp(-u * u * xi * xi), c = std::erfc(u * xi);
            const PolyCoefs tp = poly_perp_terms(r, aa, bb, xi), ta = poly_para_terms(r, aa, bb, xi);
            fp += tp.A * e + tp.C * c;
            fa += ta.A * e + ta.C * c;
        }
    // direct (unscreened) overlap of the two spheres
    const double r3 = r * r * r, a3 = a * a * a, b3 = b * b * b;
    if (r < a + b && r >= std::abs(a - b)) {
        const double t = a + b - r;
        fp += t * t * t / (128.0 * kPi * a3 * b3 * r3) *
              (-r3 - 3.0 * (a + b) * r * r + 3.0 * (a * a - 4.0 * a * b + b * b) * r + a3 - 3.0 * a * a * b -
               3.0 * a * b * b + b3);
        fa += -(std::pow(a + b, 4) * (a * a - 4.0 * a * b + b * b) - 8.0 * (a3 + b3) * r3 +
                9.0 * (a * a + b * b) * r * r * r * r - 2.0 * r3 * r3) /
              (64.0 * kPi * a3 * b3 * r3);
    } else if (r < b - a) {  // sphere a inside sphere b
        fp += (r3 - b3) / (4.0 * kPi * b3 * r3);
        fa += (r3 + 2.0 * b3) / (4.0 * kPi * b3 * r3);
    } else if (r < a - b) {  // sphere b inside sphere a
        fp += (r3 - a3) / (4.0 * kPi * a3 * r3);
        fa += (r3 + 2.0 * a3) / (4.0 * kPi * a3 * r3);
    }
    fperp = fp;
    fpara = fa;
}

// r -> 0 limit of the pair functions (concentric spheres): the self term of a
// radius pair; a = b gives the particle's real-space self term.  :1611-1620.
static double poly_self(double a, double b, double xi) {
    const double x2 = xi * xi, x3 = x2 * xi;
    const double a3 = a * a * a, b3 = b * b * b;
    return ((1.0 - 2.0 * a * a * x2 + 2.0 * a * b * x2 - 2.0 * b * b * x2) * std::exp(-(a + b) * (a + b) * x2) +
            (-1.0 + 2.0 * a * a * x2 + 2.0 * a * b * x2 + 2.0 * b * b * x2) * std::exp(-(a - b) * (a - b) * x2)) /
               (16.0 * std::pow(kPi, 1.5) * a3 * b3 * x3) +
           ((a3 - b3) * std::erf((a - b) * xi) - (a3 + b3) * std::erf((a + b) * xi)) / (8.0 * kPi * a3 * b3) +
           std::min(a3, b3) / (4.0 * kPi * a3 * b3);
}

// Finite-sphere spread/contract weight of a particle of radius a at the grid
// offset vector gd (grid point minus particle centre): the vector C(d) gd/d.
// :243-262 (the d < 1e-6 branch included as written there).
static void poly_weight(const double gd[3], double a, double xi, double eta, double w[3]) {
    const double d = std::sqrt(gd[0] * gd[0] + gd[1] * gd[1] + gd[2] * gd[2]);
    const double x2 = xi * xi;
    if (d >= 1e-6) {
        const double k = 3.0 / (8.0 * std::sqrt(2.0) * std::pow(kPi, 1.5) * a * a * a * std::sqrt(eta) * xi * d * d);
        const double tp = (eta + 4.0 * a * x2 * d) * std::exp(-2.0 * (d + a) * (d + a) * x2 / eta);
        const double tm = (eta - 4.0 * a * x2 * d) * std::exp(-2.0 * (d - a) * (d - a) * x2 / eta);
        const double c = k * (tp - tm) / d;
        for (int q = 0; q < 3; ++q) w[q] = c * gd[q];
    } else {
        const double c = 8.0 * std::sqrt(2.0) * std::pow(xi, 5.0) * std::exp(-2.0 * a * x2 / eta) /
                         (std::pow(kPi, 1.5) * std::pow(eta, 2.5));
        for (int q = 0; q < 3; ++q) w[q] = c * gd[q];
    }
}

// Polydisperse sizing: rc = max(sqrt(-ln tol)/xi, 2 a_max), the grid rule of the
// monodisperse operator, P = ceil(2/h_mean - 2 ln tol / pi) points from
// -floor((P-1)/2), eta = P (h_mean xi)^2 / pi.  :2039-2105.
static int compute_sizing_poly(const double L[3], double errortol, double xi, double amax,
                               const int* grid_override, int P_override, Sizing& S, std::string& err) {
    S.rc = std::sqrt(-std::log(errortol)) / xi;
    if (S.rc < 2.0 * amax) S.rc = 2.0 * amax;
    for (int d = 0; d < 3; ++d) {
        if (S.rc > L[d] / 2.0) {
            err = "Real space cutoff (" + std::to_string(S.rc) + ") larger than half the box length.";
            return -1;
        }
    }
    S.kcut = 2.0 * xi * xi * S.rc;
    for (int d = 0; d < 3; ++d) {
        S.n[d] = (int)std::ceil(1.0 + L[d] * S.kcut / kPi);
        if (grid_override && grid_override[d] > 0) S.n[d] = grid_override[d];
        S.h[d] = L[d] / S.n[d];
    }
    const double hm = (S.h[0] + S.h[1] + S.h[2]) / 3.0;
    S.P = (int)std::ceil(2.0 / hm - 2.0 * std::log(errortol) / kPi);
    if (P_override > 0) S.P = P_override;
    const double eta = S.P * std::pow(hm * xi, 2) / kPi;
    for (int d = 0; d < 3; ++d) S.eta[d] = eta;
    S.half = (S.P - 1) / 2;  // offsets -half .. P - 1 - half
    S.width = S.P;
    return 0;
}

// ---------------------------------------------------------------------------
// Operator
// ---------------------------------------------------------------------------
struct Op {
    double L[3];
    double errortol, xi;
    Sizing S;
    DipoleTables T;
    // polydisperse: per-particle radii, unique radii (merged within 1e-5), the
    // pair tables [entry][pair column] and the per-radius self terms
    bool poly = false;
    std::vector<double> radius, uniq;
    std::vector<int> ridx;
    std::vector<double> pr, pperp, ppara, uself;  // r samples, tables [entry][col], self per unique radius
    int ncol = 0;
    // more than 64 radius classes: no tables, each pair's closed forms at its own radii
    bool exact = false;
    std::vector<double> pself;  // [N] self term of each particle
    std::vector<double> pw;  // [i][o][3] spread/contract weight vectors
    // quadrupoles (monodisperse): particles qidx[q] carry 5 traceless components
    // (xx, xy, xz, yy, yz) after the 3N dipole unknowns; qmap[i] = q or -1
    bool quads = false;
    std::vector<int> qidx, qmap;
    std::vector<double> qtab[7];        // field_quad_1..3, grad_quad_1..4 on the dipole table's r
    double self_grad = 0;
    std::vector<double> nb_q;           // [M][7] pair values
    std::vector<double> scQ, scGP, scGQ, Qf, Qd;  // k-space: V, V, V, 5V, 5V
    size_t n_unknowns() const { return 3 * (size_t)N + 5 * qidx.size(); }
    double kbx = 0, kby = 0;  // Bloch wavevector (x, y)
    // geometry
    int N = 0;
    std::vector<double> x, y, z;
    std::vector<int> nb_start, nb_j;   // CSR over source particle i
    std::vector<double> nb_perp, nb_para;
    std::vector<int> sp_idx;           // [i][o] wrapped linear voxel index
    std::vector<double> sp_e;          // [i][o] exp(-2 xi^2 sum d^2/eta)
    double C_spread = 0, C_contract = 0;
    std::vector<double> scale, khat;   // V, 3V
    size_t V = 0;
    std::string err;
};

static void compute_scale(Op& op) {
    const Sizing& S = op.S;
    const int nx = S.n[0], ny = S.n[1], nz = S.n[2];
    op.V = (size_t)nx * ny * nz;
    op.scale.assign(op.V, 0.0);
    op.khat.assign(3 * op.V, 0.0);
    if (op.quads) {
        op.scQ.assign(op.V, 0.0);
        op.scGP.assign(op.V, 0.0);
        op.scGQ.assign(op.V, 0.0);
        op.Qf.assign(5 * op.V, 0.0);
        op.Qd.assign(5 * op.V, 0.0);
    }
    const double xi = op.xi;
#pragma omp parallel for schedule(static)
    for (long v = 0; v < (long)op.V; ++v) {
        const int ix = (int)(v / ((long)ny * nz));
        const int iy = (int)((v / nz) % ny);
        const int iz = (int)(v % nz);
        const double fx = (ix <= (nx - 1) / 2) ? ix : (ix - nx);
        const double fy = (iy <= (ny - 1) / 2) ? iy : (iy - ny);
        const double fz = (iz <= (nz - 1) / 2) ? iz : (iz - nz);
        const double Gx = fx * 2.0 * kPi / op.L[0], Gy = fy * 2.0 * kPi / op.L[1],
                     Gz = fz * 2.0 * kPi / op.L[2];
        const double kx = Gx - op.kbx, ky = Gy - op.kby, kz = Gz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        if (op.poly) {  // exp(-(1 - eta) k^2 / 4 xi^2) / k^2 / V (:565-640)
            if (k2 >= 1e-12)
                op.scale[v] = std::exp(-(1.0 - S.eta[0]) * k2 / (4.0 * xi * xi)) / k2 / (double)op.V;
            continue;
        }
        const double k = std::sqrt(k2);
        if (k < 1e-12) continue;
        op.khat[3 * v + 0] = kx / k;
        op.khat[3 * v + 1] = ky / k;
        op.khat[3 * v + 2] = kz / k;
        const double etak2 = k2 - (Gx * Gx * S.eta[0] + Gy * Gy * S.eta[1] + Gz * Gz * S.eta[2]);
        const double t = std::sin(k) / k - std::cos(k);
        const double ex = std::exp(-etak2 / (4.0 * xi * xi));
        op.scale[v] = 9.0 * t * t * ex / (k2 * k2) / (double)op.V;
        if (op.quads) {  // :1367-1398: sphere form factors j1, j2 and the quadrupole projectors
            const double j1 = t / k;
            const double j2 = (3.0 / k2 - 1.0) * std::sin(k) / k - 3.0 * std::cos(k) / k2;
            const double e2 = ex / k2 / (double)op.V;
            op.scQ[v] = -22.5 * j1 * j2 * e2;
            op.scGP[v] = 45.0 * j1 * j2 * e2;
            op.scGQ[v] = 112.5 * j2 * j2 * e2;
            const double h0 = kx / k, h1 = ky / k, h2 = kz / k;
            const double qf[5] = {h0 * h0 - 1.0 / 3.0, h0 * h1, h0 * h2, h1 * h1 - 1.0 / 3.0, h1 * h2};
            const double qd[5] = {h0 * h0 - h2 * h2, 2.0 * h0 * h1, 2.0 * h0 * h2, h1 * h1 - h2 * h2, 2.0 * h1 * h2};
            for (int c = 0; c < 5; ++c) {
                op.Qf[5 * v + c] = qf[c];
                op.Qd[5 * v + c] = qd[c];
            }
        }
    }
}

// Pair column of two radius classes (upper triangle, row-major) and the table
// lookup with the monodisperse interpolation rule.  :56-68, 168-188.
static int poly_col(const Op& op, int u, int v) {
    if (u > v) std::swap(u, v);
    const int U = (int)op.uniq.size();
    return u * U - u * (u - 1) / 2 + (v - u);
}
static double poly_interp(const Op& op, double r, int u, int v, const std::vector<double>& tab) {
    const int col = poly_col(op, u, v);
    const size_t ts = op.pr.size();
    auto y = [&](size_t e) { return tab[e * op.ncol + col]; };
    if (r <= 0.0) return y(0);
    if (r >= op.pr[ts - 1]) return y(ts - 1);
    if (r < 1.0) return y(0) * (1.0 - r) + y(1) * r;
    size_t idx = 1 + (size_t)((r - 1.0) / 0.001);
    if (idx >= ts - 1) return y(ts - 1);
    const double t = (r - op.pr[idx]) / (op.pr[idx + 1] - op.pr[idx]);
    return y(idx) * (1.0 - t) + y(idx + 1) * t;
}

// Tables of every radius pair: entry 0 = the pair's self term, then samples at
// r = 1 + 0.001 k up to the cutoff.  :1629-1698.
static void build_poly_tables(Op& op) {
    const double rcv = op.S.rc < 1.001 ? 1.001 : op.S.rc;
    const size_t steps = (size_t)std::ceil((rcv - 1.0) / 0.001);
    const int U = (int)op.uniq.size();
    op.ncol = U * (U + 1) / 2;
    op.pr.assign(steps + 1, 0.0);
    for (size_t k = 0; k < steps; ++k) op.pr[k + 1] = 1.0 + k * 0.001;
    op.pperp.assign((steps + 1) * op.ncol, 0.0);
    op.ppara.assign((steps + 1) * op.ncol, 0.0);
    op.uself.assign(U, 0.0);
    for (int u = 0; u < U; ++u)
        for (int v = u; v < U; ++v) {
            const int col = poly_col(op, u, v);
            const double a = op.uniq[u], b = op.uniq[v];
            const double s0 = poly_self(a, b, op.xi);
            op.pperp[col] = s0;
            op.ppara[col] = s0;
            if (u == v) op.uself[u] = s0;
#pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)steps; ++k) {
                double fp, fa;
                poly_pair_functions(op.pr[k + 1], a, b, op.xi, fp, fa);
                op.pperp[(k + 1) * op.ncol + col] = fp;
                op.ppara[(k + 1) * op.ncol + col] = fa;
            }
        }
}

// neighbor_list.cu: all j != i with min-image |r|^2 <= rc^2 (cell list when
// floor(L/rc) >= 3 on every axis, brute force otherwise — same pair set).
static void build_neighbors(Op& op) {
    const int N = op.N;
    const double rc2 = op.S.rc * op.S.rc;
    const double* L = op.L;
    std::vector<std::vector<int>> lists(N);
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = (int)std::floor(L[d] / op.S.rc);
    const bool cells = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
    auto minimg = [&](double dx, double Ld) { return dx - Ld * std::round(dx / Ld); };
    if (cells) {
        const double cs[3] = {L[0] / nc[0], L[1] / nc[1], L[2] / nc[2]};
        const long ncell = (long)nc[0] * nc[1] * nc[2];
        std::vector<int> cell_of(N);
        std::vector<std::vector<int>> members(ncell);
        auto cell_coord = [&](double p, int d) {
            double w = p - L[d] * std::floor(p / L[d]);
            int c = (int)std::floor(w / cs[d]);
            if (c < 0) c = 0;
            if (c >= nc[d]) c = nc[d] - 1;
            return c;
        };
        for (int j = 0; j < N; ++j) {
            int cx = cell_coord(op.x[j], 0), cy = cell_coord(op.y[j], 1), cz = cell_coord(op.z[j], 2);
            members[cx + (long)cy * nc[0] + (long)cz * nc[0] * nc[1]].push_back(j);
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            int cx = cell_coord(op.x[i], 0), cy = cell_coord(op.y[i], 1), cz = cell_coord(op.z[i], 2);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int a = (cx + dx + nc[0]) % nc[0], b = (cy + dy + nc[1]) % nc[1],
                            c = (cz + dz + nc[2]) % nc[2];
                        for (int j : members[a + (long)b * nc[0] + (long)c * nc[0] * nc[1]]) {
                            if (j == i) continue;
                            double rx = minimg(op.x[i] - op.x[j], L[0]);
                            double ry = minimg(op.y[i] - op.y[j], L[1]);
                            double rz = minimg(op.z[i] - op.z[j], L[2]);
                            if (rx * rx + ry * ry + rz * rz <= rc2) lists[i].push_back(j);
                        }
                    }
            std::sort(lists[i].begin(), lists[i].end());
        }
    } else {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double rx = minimg(op.x[i] - op.x[j], L[0]);
                double ry = minimg(op.y[i] - op.y[j], L[1]);
                double rz = minimg(op.z[i] - op.z[j], L[2]);
                if (rx * rx + ry * ry + rz * rz <= rc2) lists[i].push_back(j);
            }
        }
    }
    op.nb_start.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) op.nb_start[i + 1] = op.nb_start[i] + (int)lists[i].size();
    const int M = op.nb_start[N];
    op.nb_j.resize(M);
    op.nb_perp.assign(M, 0.0);
    op.nb_para.assign(M, 0.0);
    op.nb_q.assign(op.quads ? 7 * (size_t)M : 0, 0.0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        int s = op.nb_start[i];
        for (size_t k = 0; k < lists[i].size(); ++k) {
            const int j = lists[i][k];
            op.nb_j[s + k] = j;
            double rx = minimg(op.x[i] - op.x[j], L[0]);
            double ry = minimg(op.y[i] - op.y[j], L[1]);
            double rz = minimg(op.z[i] - op.z[j], L[2]);
            double d = std::sqrt(rx * rx + ry * ry + rz * rz);
            if (d < op.S.rc && op.poly) {
                const double de = d < 1e-6 ? 1e-6 : d;
                if (!op.exact) {
                    op.nb_perp[s + k] = poly_interp(op, de, op.ridx[i], op.ridx[j], op.pperp);
                    op.nb_para[s + k] = poly_interp(op, de, op.ridx[i], op.ridx[j], op.ppara);
                } else {
                    // closed forms at the pair's radii; below r = 1 the tables' linear rule
                    const double a = op.radius[i], b = op.radius[j];
                    double fp, fa;
                    poly_pair_functions(de < 1.0 ? 1.0 : de, a, b, op.xi, fp, fa);
                    if (de < 1.0) {
                        const double s0 = poly_self(a, b, op.xi);
                        fp = s0 * (1.0 - de) + fp * de;
                        fa = s0 * (1.0 - de) + fa * de;
                    }
                    op.nb_perp[s + k] = fp;
                    op.nb_para[s + k] = fa;
                }
            } else if (d < op.S.rc) {
                double de = d < 1.0 ? 1.0 : d;
                op.nb_perp[s + k] = interp(de, op.T.r, op.T.perp);
                op.nb_para[s + k] = interp(de, op.T.r, op.T.para);
                if (op.quads)
                    for (int t = 0; t < 7; ++t) op.nb_q[7 * (size_t)(s + k) + t] = interp(de, op.T.r, op.qtab[t]);
            }
        }
    }
}

// Polydisperse stencils: P^3 points per particle (x-major offsets from
// -floor((P-1)/2)), wrapped voxel index and weight vector.  :190-350.
static void build_stencils_poly(Op& op) {
    const Sizing& S = op.S;
    const int P = S.P, o0 = -S.half;
    const long SS = (long)P * P * P;
    op.sp_idx.assign((size_t)op.N * SS, 0);
    op.pw.assign((size_t)op.N * SS * 3, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < op.N; ++i) {
        const double p[3] = {op.x[i], op.y[i], op.z[i]};
        int g[3];
        double dd[3];
        for (int d = 0; d < 3; ++d) {
            g[d] = (int)std::round(p[d] / S.h[d]);
            dd[d] = g[d] * S.h[d] - p[d];
        }
        long o = 0;
        for (int ox = o0; ox < o0 + P; ++ox)
            for (int oy = o0; oy < o0 + P; ++oy)
                for (int oz = o0; oz < o0 + P; ++oz, ++o) {
                    const int off[3] = {ox, oy, oz};
                    int w[3];
                    double gd[3];
                    for (int d = 0; d < 3; ++d) {
                        w[d] = ((g[d] + off[d]) % S.n[d] + S.n[d]) % S.n[d];
                        gd[d] = dd[d] + off[d] * S.h[d];
                    }
                    op.sp_idx[(size_t)i * SS + o] = (w[0] * S.n[1] + w[1]) * S.n[2] + w[2];
                    poly_weight(gd, op.radius[i], op.xi, S.eta[0], &op.pw[((size_t)i * SS + o) * 3]);
                }
    }
}

// monodisperse_ewald_electric_field.cu:687-816 (weights C*exp(-2 xi^2 sum gd^2/eta)).
static void build_stencils(Op& op) {
    if (op.poly) {
        build_stencils_poly(op);
        return;
    }
    const Sizing& S = op.S;
    const int W = S.width, half = S.half;
    const long SS = (long)W * W * W;
    op.sp_idx.assign((size_t)op.N * SS, 0);
    op.sp_e.assign((size_t)op.N * SS, 0.0);
    const double prod_eta = S.eta[0] * S.eta[1] * S.eta[2];
    op.C_spread = std::pow(2.0 * op.xi * op.xi / kPi, 1.5) * std::sqrt(1.0 / prod_eta);
    op.C_contract = op.C_spread * (S.h[0] * S.h[1] * S.h[2]);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < op.N; ++i) {
        const double p[3] = {op.x[i], op.y[i], op.z[i]};
        int g[3];
        double dd[3];
        for (int d = 0; d < 3; ++d) {
            g[d] = (int)std::round(p[d] / S.h[d]);
            dd[d] = g[d] * S.h[d] - p[d];
        }
        long o = 0;
        for (int ox = -half; ox <= half; ++ox)
            for (int oy = -half; oy <= half; ++oy)
                for (int oz = -half; oz <= half; ++oz, ++o) {
                    const int off[3] = {ox, oy, oz};
                    int w[3];
                    double s = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        w[d] = ((g[d] + off[d]) % S.n[d] + S.n[d]) % S.n[d];
                        const double gd = dd[d] + off[d] * S.h[d];
                        s += gd * gd / S.eta[d];
                    }
                    op.sp_idx[(size_t)i * SS + o] = (w[0] * S.n[1] + w[1]) * S.n[2] + w[2];
                    op.sp_e[(size_t)i * SS + o] = std::exp(-2.0 * op.xi * op.xi * s);
                }
    }
}

// Apply E = A p for one right-hand side (SOLVER_AX mode).  p, E: [N][3] complex.
// self_coef: [N] complex.  parts: bit 0 real space + self, bit 1 reciprocal.
static void apply(const Op& op, const cplx* self_coef, const cplx* p, cplx* E, int parts) {
    const int N = op.N;
    const int Q = (int)op.qidx.size();
    for (size_t q = 0; q < op.n_unknowns(); ++q) E[q] = 0.0;
    const cplx* qv = p + 3 * (size_t)N;  // quadrupoles [Q][5]
    cplx* G = E + 3 * (size_t)N;
    if ((parts & 2) && op.poly) {
        // scalar grid S = sum C.p, scaled in k-space, E_i = h^3 sum C_i S (:444-694)
        const Sizing& S = op.S;
        const size_t V = op.V;
        const long SS = (long)S.P * S.P * S.P;
        std::vector<cplx> grid(V, cplx(0, 0));
        for (int i = 0; i < N; ++i)
            for (long o = 0; o < SS; ++o) {
                const double* w = &op.pw[((size_t)i * SS + o) * 3];
                grid[op.sp_idx[(size_t)i * SS + o]] += w[0] * p[3 * i] + w[1] * p[3 * i + 1] + w[2] * p[3 * i + 2];
            }
        fft::fft3d(grid.data(), S.n[0], S.n[1], S.n[2], -1);
        for (size_t v = 0; v < V; ++v) grid[v] *= op.scale[v];
        fft::fft3d(grid.data(), S.n[0], S.n[1], S.n[2], +1);
        const double h3 = S.h[0] * S.h[1] * S.h[2];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            cplx acc[3] = {0.0, 0.0, 0.0};
            for (long o = 0; o < SS; ++o) {
                const double* w = &op.pw[((size_t)i * SS + o) * 3];
                const cplx g = grid[op.sp_idx[(size_t)i * SS + o]];
                for (int c = 0; c < 3; ++c) acc[c] += g * w[c];
            }
            for (int c = 0; c < 3; ++c) E[3 * i + c] += h3 * acc[c];
        }
    } else if (parts & 2) {
        const Sizing& S = op.S;
        const size_t V = op.V;
        const long SS = (long)S.width * S.width * S.width;
        const int nc = op.quads ? 8 : 3;  // E components, then the 5 quadrupole-gradient components
        std::vector<cplx> grid(nc * V, cplx(0, 0));
        // spread (serial over particles to keep the accumulation race-free); quadrupoles
        // use the dipoles' Gaussian weights (:1871-1999)
        for (int i = 0; i < N; ++i) {
            const int q = op.quads ? op.qmap[i] : -1;
            for (long o = 0; o < SS; ++o) {
                const double w = op.C_spread * op.sp_e[(size_t)i * SS + o];
                const int v = op.sp_idx[(size_t)i * SS + o];
                for (int c = 0; c < 3; ++c) grid[c * V + v] += w * p[3 * i + c];
                if (q >= 0)
                    for (int c = 0; c < 5; ++c) grid[(3 + c) * V + v] += w * qv[5 * q + c];
            }
        }
        for (int c = 0; c < nc; ++c) fft::fft3d(grid.data() + c * V, S.n[0], S.n[1], S.n[2], -1);
#pragma omp parallel for schedule(static)
        for (long v = 0; v < (long)V; ++v) {
            const double kx = op.khat[3 * v], ky = op.khat[3 * v + 1], kz = op.khat[3 * v + 2];
            const cplx dot = grid[v] * kx + grid[V + v] * ky + grid[2 * V + v] * kz;
            cplx s = dot * op.scale[v];
            if (op.quads) {  // scale_kernel_joint (:1435-1487)
                cplx gq = 0.0;
                for (int c = 0; c < 5; ++c) gq += grid[(3 + c) * V + v] * op.Qd[5 * v + c];
                const cplx I(0.0, 1.0);
                s += gq * (I * op.scQ[v]);
                const cplx g = dot * (I * op.scGP[v]) + gq * op.scGQ[v];
                for (int c = 0; c < 5; ++c) grid[(3 + c) * V + v] = g * op.Qf[5 * v + c];
            }
            grid[v] = s * kx;
            grid[V + v] = s * ky;
            grid[2 * V + v] = s * kz;
        }
        for (int c = 0; c < nc; ++c) fft::fft3d(grid.data() + c * V, S.n[0], S.n[1], S.n[2], +1);
        if (op.quads) {
#pragma omp parallel for schedule(static)
            for (int q = 0; q < Q; ++q) {
                const int i = op.qidx[q];
                cplx acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                for (long o = 0; o < SS; ++o) {
                    const double w = op.C_contract * op.sp_e[(size_t)i * SS + o];
                    const int v = op.sp_idx[(size_t)i * SS + o];
                    for (int c = 0; c < 5; ++c) acc[c] += grid[(3 + c) * V + v] * w;
                }
                for (int c = 0; c < 5; ++c) G[5 * q + c] += acc[c];
            }
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            cplx acc[3] = {0.0, 0.0, 0.0};
            for (long o = 0; o < SS; ++o) {
                const double w = op.C_contract * op.sp_e[(size_t)i * SS + o];
                const int v = op.sp_idx[(size_t)i * SS + o];
                for (int c = 0; c < 3; ++c) acc[c] += grid[c * V + v] * w;
            }
            for (int c = 0; c < 3; ++c) E[3 * i + c] += acc[c];
        }
    }
    if (parts & 1) {
        for (int i = 0; i < N; ++i) {
            const cplx f = self_coef[i] + (op.poly ? op.pself[i] : op.T.self_perp);
            for (int c = 0; c < 3; ++c) E[3 * i + c] += f * p[3 * i + c];
        }
        for (int q = 0; q < Q; ++q) {  // :1615-1627
            const cplx f = 2.5 * self_coef[op.qidx[q]] + 0.5 * op.self_grad;
            for (int c = 0; c < 5; ++c) G[5 * q + c] += f * qv[5 * q + c];
        }
        // source-centric scatter, as the reference kernel: E_j += T(r_ij) p_i
        const bool bloch = (op.kbx != 0.0 || op.kby != 0.0);
        for (int i = 0; i < N; ++i) {
            for (int k = op.nb_start[i]; k < op.nb_start[i + 1]; ++k) {
                const int j = op.nb_j[k];
                double r[3] = {op.x[i] - op.x[j], op.y[i] - op.y[j], op.z[i] - op.z[j]};
                for (int d = 0; d < 3; ++d) r[d] -= op.L[d] * std::round(r[d] / op.L[d]);
                const double d = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                if (!(d < op.S.rc)) continue;
                const double de = op.poly ? (d < 1e-6 ? 1e-6 : d) : (d < 1.0 ? 1.0 : d);
                const double u[3] = {r[0] / de, r[1] / de, r[2] / de};
                const cplx rp = p[3 * i] * u[0] + p[3 * i + 1] * u[1] + p[3 * i + 2] * u[2];
                cplx e[3];
                for (int c = 0; c < 3; ++c)
                    e[c] = (p[3 * i + c] - rp * u[c]) * op.nb_perp[k] + (rp * u[c]) * op.nb_para[k];
                cplx f(1.0, 0.0);
                if (bloch) {
                    const double ph = -(op.kbx * r[0] + op.kby * r[1]);
                    f = cplx(std::cos(ph), std::sin(ph));
                }
                // Dipole-quadrupole cross terms carry the signs of the reference's Python
                // formulation (tests/quadrupole_field.py:214-215: E += Q terms, G -= P terms),
                // which keep real + reciprocal independent of xi; the CUDA kernel
                // (monodisperse...cu:1749-1800) has both signs flipped.
                if (op.quads) {  // :1631-1869, source i, target j, u = r_ij / max(d, 1)
                    const double* t = &op.nb_q[7 * (size_t)k];
                    const int qs = op.qmap[i], qt = op.qmap[j];
                    const double I5[5] = {1.0, 0.0, 0.0, 1.0, 0.0};
                    const double rr[5] = {u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2]};
                    cplx Qr[3] = {0.0, 0.0, 0.0}, Qrr = 0.0;
                    if (qs >= 0) {
                        const cplx* qq = qv + 5 * qs;
                        Qr[0] = qq[0] * u[0] + qq[1] * u[1] + qq[2] * u[2];
                        Qr[1] = qq[1] * u[0] + qq[3] * u[1] + qq[4] * u[2];
                        Qr[2] = qq[2] * u[0] + qq[4] * u[1] - (qq[0] + qq[3]) * u[2];
                        Qrr = Qr[0] * u[0] + Qr[1] * u[1] + Qr[2] * u[2];
                        for (int c = 0; c < 3; ++c) e[c] += 0.5 * (Qrr * (t[0] * u[c]) + Qr[c] * (2.0 * t[1]));
                    }
                    if (qt >= 0) {
                        const cplx* pi = p + 3 * i;
                        const cplx PrrP[5] = {pi[0] * (2.0 * u[0]), pi[0] * u[1] + pi[1] * u[0], pi[0] * u[2] + pi[2] * u[0],
                                              pi[1] * (2.0 * u[1]), pi[1] * u[2] + pi[2] * u[1]};
                        cplx g[5];
                        for (int c = 0; c < 5; ++c) g[c] = -(rp * (t[0] * rr[c] + (t[1] + t[2]) * I5[c]) + PrrP[c] * t[1]);
                        if (qs >= 0) {
                            const cplx* qq = qv + 5 * qs;
                            const cplx QrrQ[5] = {Qr[0] * (2.0 * u[0]), Qr[0] * u[1] + Qr[1] * u[0], Qr[0] * u[2] + Qr[2] * u[0],
                                                  Qr[1] * (2.0 * u[1]), Qr[1] * u[2] + Qr[2] * u[1]};
                            for (int c = 0; c < 5; ++c)
                                g[c] += 0.5 * (Qrr * ((t[3] + t[5]) * I5[c] + t[6] * rr[c]) + qq[c] * t[4] +
                                               QrrQ[c] * (2.0 * t[5]));
                        }
                        for (int c = 0; c < 5; ++c) G[5 * qt + c] += g[c] * f;
                    }
                }
                for (int c = 0; c < 3; ++c) E[3 * j + c] += e[c] * f;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Krylov solvers on A (left Jacobi preconditioner M = diag(self_coef)).
// ---------------------------------------------------------------------------
static void jacobi(const Op& op, const cplx* self_coef, const cplx* src, cplx* dst) {
    for (int i = 0; i < op.N; ++i) {
        const cplx sc = self_coef[i];
        const double den = std::norm(sc);
        const cplx inv = den > 1e-30 ? cplx(sc.real() / den, -sc.imag() / den) : cplx(1.0, 0.0);
        for (int c = 0; c < 3; ++c) dst[3 * i + c] = src[3 * i + c] * inv;
    }
    for (size_t q = 0; q < op.qidx.size(); ++q) {  // numerical_solver.cu:223-233
        const cplx sc = self_coef[op.qidx[q]];
        const double den = std::norm(sc);
        const cplx inv = den > 1e-30 ? cplx(sc.real() / den, -sc.imag() / den) : cplx(1.0, 0.0);
        for (int c = 0; c < 5; ++c) dst[3 * (size_t)op.N + 5 * q + c] = src[3 * (size_t)op.N + 5 * q + c] * inv;
    }
}
static double vnorm(const std::vector<cplx>& a) {
    double s = 0;
    for (auto& v : a) s += std::norm(v);
    return std::sqrt(s);
}
static cplx vdot(const cplx* a, const cplx* b, size_t n) {  // conj(a) . b
    cplx s = 0;
    for (size_t i = 0; i < n; ++i) s += std::conj(a[i]) * b[i];
    return s;
}
static cplx vdotu(const cplx* a, const cplx* b, size_t n) {  // a . b
    cplx s = 0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

struct SolveReport {
    int iters;        // operator applies inside the Krylov loop (reference counting)
    int matvecs;      // all operator applies
    int converged;
    double relres_pre;   // ||M^-1 (b - A x)|| / ||b|| (reference criterion)
    double relres_true;  // ||b - A x|| / ||b||
};

// gmres_solver.cu:56-208
static void gmres(const Op& op, const cplx* sc, const cplx* b, cplx* x, double tol, int restart,
                  int maxiter, bool use_jacobi, SolveReport& rep) {
    const size_t n = op.n_unknowns();
    std::vector<cplx> bb(b, b + n), r(n), w(n), t(n);
    rep = SolveReport{0, 0, 0, 0, 0};
    const double bnorm = vnorm(bb);
    if (bnorm == 0.0) {
        std::fill(x, x + n, cplx(0, 0));
        rep.converged = 1;
        return;
    }
    const int m = (int)std::min<size_t>(n, restart);
    std::vector<cplx> V((size_t)(m + 1) * n);
    std::vector<std::vector<cplx>> H(m + 1, std::vector<cplx>(m, 0.0));
    std::vector<cplx> g(m + 1), s(m);
    std::vector<double> c(m);
    int iter = 0;
    while (iter < maxiter) {
        apply(op, sc, x, r.data(), 3);
        rep.matvecs++;
        for (size_t q = 0; q < n; ++q) r[q] = bb[q] - r[q];
        if (use_jacobi) { jacobi(op, sc, r.data(), t.data()); r = t; }
        const double rn = vnorm(r);
        if (rn < tol * bnorm) { rep.converged = 1; break; }
        g.assign(m + 1, 0.0);
        g[0] = rn;
        for (size_t q = 0; q < n; ++q) V[q] = r[q] / rn;
        int k = 0;
        for (k = 0; k < m && iter < maxiter; ++k, ++iter) {
            cplx* Vk = &V[(size_t)k * n];
            cplx* Vk1 = &V[(size_t)(k + 1) * n];
            apply(op, sc, Vk, w.data(), 3);
            rep.matvecs++;
            if (use_jacobi) { jacobi(op, sc, w.data(), t.data()); w = t; }
            for (int i = 0; i <= k; ++i) {
                const cplx* Vi = &V[(size_t)i * n];
                H[i][k] = vdot(Vi, w.data(), n);
                for (size_t q = 0; q < n; ++q) w[q] -= H[i][k] * Vi[q];
            }
            const double wn = vnorm(w);
            H[k + 1][k] = wn;
            if (wn > 0.0)
                for (size_t q = 0; q < n; ++q) Vk1[q] = w[q] / wn;
            for (int i = 0; i < k; ++i) {
                const cplx h1 = H[i][k], h2 = H[i + 1][k];
                H[i][k] = c[i] * h1 + std::conj(s[i]) * h2;
                H[i + 1][k] = -s[i] * h1 + c[i] * h2;
            }
            const cplx h1 = H[k][k], h2 = H[k + 1][k];
            double cr = 1.0;
            cplx sr = 0.0;
            if (h2 != 0.0) {
                const double a1 = std::abs(h1);
                if (a1 == 0.0) {
                    cr = 0.0;
                    sr = 1.0;
                } else {
                    const double scl = a1 + std::abs(h2);
                    const double n1 = std::abs(h1 / scl), n2 = std::abs(h2 / scl);
                    const double hyp = scl * std::sqrt(n1 * n1 + n2 * n2);
                    cr = a1 / hyp;
                    sr = (std::conj(h1) / a1) * (h2 / hyp);
                }
            }
            c[k] = cr;
            s[k] = sr;
            H[k][k] = cr * h1 + std::conj(sr) * h2;
            H[k + 1][k] = 0.0;
            const cplx g1 = g[k], g2 = g[k + 1];
            g[k] = cr * g1 + std::conj(sr) * g2;
            g[k + 1] = -sr * g1 + cr * g2;
            if (std::abs(g[k + 1]) < tol * bnorm) {
                k++;
                iter++;
                break;
            }
        }
        std::vector<cplx> yv(k);
        for (int i = k - 1; i >= 0; --i) {
            yv[i] = g[i];
            for (int j = i + 1; j < k; ++j) yv[i] -= H[i][j] * yv[j];
            yv[i] /= H[i][i];
        }
        for (int i = 0; i < k; ++i)
            for (size_t q = 0; q < n; ++q) x[q] += yv[i] * V[(size_t)i * n + q];
    }
    rep.iters = iter;
    // final residuals (diagnostic only)
    apply(op, sc, x, r.data(), 3);
    for (size_t q = 0; q < n; ++q) r[q] = bb[q] - r[q];
    rep.relres_true = vnorm(r) / bnorm;
    jacobi(op, sc, r.data(), t.data());
    rep.relres_pre = vnorm(t) / bnorm;
    if (rep.relres_pre < tol) rep.converged = 1;
}

// bicgstab_solver.cu:58-192 (unconjugated rho and r0~.v; conjugated omega).
static void bicgstab(const Op& op, const cplx* sc, const cplx* b, cplx* x, double tol, int maxiter,
                     bool use_jacobi, SolveReport& rep) {
    const size_t n = op.n_unknowns();
    std::vector<cplx> bb(b, b + n), r(n), r0(n), p(n), v(n, 0.0), s(n), t(n), tmp(n);
    rep = SolveReport{0, 0, 0, 0, 0};
    const double bnorm = vnorm(bb);
    if (bnorm == 0.0) {
        std::fill(x, x + n, cplx(0, 0));
        rep.converged = 1;
        return;
    }
    auto Aprec = [&](const cplx* in, std::vector<cplx>& out) {
        apply(op, sc, in, out.data(), 3);
        rep.matvecs++;
        if (use_jacobi) { jacobi(op, sc, out.data(), tmp.data()); out = tmp; }
    };
    apply(op, sc, x, tmp.data(), 3);
    rep.matvecs++;
    for (size_t q = 0; q < n; ++q) r[q] = bb[q] - tmp[q];
    if (use_jacobi) { jacobi(op, sc, r.data(), tmp.data()); r = tmp; }
    double rn = vnorm(r);
    int iter = 0;
    if (!(rn < tol * bnorm)) {
        r0 = r;
        p = r;
        cplx rho = 1.0, alpha = 1.0, omega = 1.0;
        for (iter = 0; iter < maxiter; ++iter) {
            const cplx rho_new = vdotu(r0.data(), r.data(), n);
            if (std::abs(rho_new) < 1e-16) break;
            if (iter > 0) {
                const cplx beta = (rho_new / rho) * (alpha / omega);
                for (size_t q = 0; q < n; ++q) p[q] = (p[q] - omega * v[q]) * beta + r[q];
            }
            rho = rho_new;
            Aprec(p.data(), v);
            const cplx rv = vdotu(r0.data(), v.data(), n);
            if (std::abs(rv) < 1e-16) break;
            alpha = rho / rv;
            for (size_t q = 0; q < n; ++q) s[q] = r[q] - alpha * v[q];
            if (vnorm(s) < tol * bnorm) {
                for (size_t q = 0; q < n; ++q) x[q] += alpha * p[q];
                iter++;
                break;
            }
            Aprec(s.data(), t);
            const cplx ts = vdot(t.data(), s.data(), n);
            const cplx tt = vdot(t.data(), t.data(), n);
            if (std::abs(tt) < 1e-16) { iter++; break; }
            omega = ts / tt.real();
            for (size_t q = 0; q < n; ++q) x[q] += alpha * p[q] + omega * s[q];
            for (size_t q = 0; q < n; ++q) r[q] = s[q] - omega * t[q];
            if (vnorm(r) < tol * bnorm) { iter++; break; }
            if (std::abs(omega) < 1e-16) { iter++; break; }
        }
    }
    rep.iters = iter;
    apply(op, sc, x, r.data(), 3);
    for (size_t q = 0; q < n; ++q) r[q] = bb[q] - r[q];
    rep.relres_true = vnorm(r) / bnorm;
    jacobi(op, sc, r.data(), tmp.data());
    rep.relres_pre = vnorm(tmp) / bnorm;
    rep.converged = rep.relres_pre < tol ? 1 : 0;
}

// ---------------------------------------------------------------------------
// C ABI for the Python test harness (oracle/oracle.py).
// ---------------------------------------------------------------------------
extern "C" {

int orc_fft3d(double* data, int nx, int ny, int nz, int sign) {
    fft::fft3d(reinterpret_cast<cplx*>(data), nx, ny, nz, sign);
    return 0;
}

int orc_sizing(const double* L, double errortol, double xi, const int* grid_override,
               int P_override, int* n_out, double* h_out, double* eta_out, int* P_out,
               int* width_out, double* rc_out) {
    Sizing S;
    std::string err;
    if (compute_sizing(L, errortol, xi, grid_override, P_override, S, err)) return -1;
    for (int d = 0; d < 3; ++d) {
        n_out[d] = S.n[d];
        h_out[d] = S.h[d];
        eta_out[d] = S.eta[d];
    }
    *P_out = S.P;
    *width_out = S.width;
    *rc_out = S.rc;
    return 0;
}

int orc_tables(double xi, double* r, double* perp, double* para, double* self_perp) {
    DipoleTables T;
    build_dipole_tables(xi, T);
    std::memcpy(r, T.r.data(), T.r.size() * sizeof(double));
    std::memcpy(perp, T.perp.data(), T.perp.size() * sizeof(double));
    std::memcpy(para, T.para.data(), T.para.size() * sizeof(double));
    *self_perp = T.self_perp;
    return (int)T.r.size();
}

void orc_pair_functions(const double* r, int n, double xi, double* perp, double* para) {
    for (int i = 0; i < n; ++i) dipole_pair_functions(r[i], xi, perp[i], para[i]);
}

void* orc_op_create(const double* L, double errortol, double xi, const int* grid_override,
                    int P_override, char* errbuf, int errlen) {
    auto op = new Op();
    for (int d = 0; d < 3; ++d) op->L[d] = L[d];
    op->errortol = errortol;
    op->xi = xi;
    std::string err;
    if (compute_sizing(L, errortol, xi, grid_override, P_override, op->S, err)) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", err.c_str());
        delete op;
        return nullptr;
    }
    build_dipole_tables(xi, op->T);
    compute_scale(*op);
    return op;
}

// Polydisperse operator: radii[N] (nondimensional), merged within 1e-5 into
// sorted unique radii (:40-83).  Positions must then come with the same N.
void* orc_poly_create(const double* L, double errortol, double xi, const double* radii, int N,
                      const int* grid_override, int P_override, char* errbuf, int errlen) {
    auto op = new Op();
    for (int d = 0; d < 3; ++d) op->L[d] = L[d];
    op->errortol = errortol;
    op->xi = xi;
    op->poly = true;
    op->radius.assign(radii, radii + N);
    for (double r : op->radius) {
        if (op->exact) break;
        bool seen = false;
        for (double u : op->uniq) seen = seen || std::abs(u - r) < 1e-5;
        if (!seen) op->uniq.push_back(r);
        op->exact = op->uniq.size() > 64;
    }
    if (op->exact) {
        const double amax = N > 0 ? *std::max_element(op->radius.begin(), op->radius.end()) : 1.0;
        std::string err;
        if (compute_sizing_poly(L, errortol, xi, amax, grid_override, P_override, op->S, err)) {
            if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", err.c_str());
            delete op;
            return nullptr;
        }
        op->uniq.clear();
        op->pself.resize(N);
        for (int i = 0; i < N; ++i) op->pself[i] = poly_self(op->radius[i], op->radius[i], xi);
        compute_scale(*op);
        return op;
    }
    std::sort(op->uniq.begin(), op->uniq.end());
    op->ridx.assign(N, 0);
    for (int i = 0; i < N; ++i)
        for (size_t u = 0; u < op->uniq.size(); ++u)
            if (std::abs(op->uniq[u] - op->radius[i]) < 1e-5) {
                op->ridx[i] = (int)u;
                break;
            }
    std::string err;
    if (N < 1 || compute_sizing_poly(L, errortol, xi, op->uniq.back(), grid_override, P_override, op->S, err)) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, errlen, "%s", N < 1 ? "no radii" : err.c_str());
        delete op;
        return nullptr;
    }
    build_poly_tables(*op);
    op->pself.resize(N);
    for (int i = 0; i < N; ++i) op->pself[i] = op->uself[op->ridx[i]];
    compute_scale(*op);
    return op;
}

void orc_poly_pair_functions(const double* r, int n, double a, double b, double xi, double* perp, double* para) {
    for (int i = 0; i < n; ++i) poly_pair_functions(r[i], a, b, xi, perp[i], para[i]);
}

double orc_poly_self(double a, double b, double xi) { return poly_self(a, b, xi); }

void orc_poly_weight(const double* gd, double a, double xi, double eta, double* w) { poly_weight(gd, a, xi, eta, w); }

// Quadrupoles on the monodisperse operator: quad_idxs[nq] particle indices (all
// particles when nq < 0).  Call after orc_op_set_positions; unknowns become 3N + 5Q.
int orc_op_set_quads(void* h, const int* quad_idxs, int nq) {
    Op& op = *static_cast<Op*>(h);
    if (op.poly || op.N == 0) return -1;
    op.quads = true;
    op.qidx.clear();
    if (nq < 0)
        for (int i = 0; i < op.N; ++i) op.qidx.push_back(i);
    else
        op.qidx.assign(quad_idxs, quad_idxs + nq);
    op.qmap.assign(op.N, -1);
    for (size_t q = 0; q < op.qidx.size(); ++q) {
        if (op.qidx[q] < 0 || op.qidx[q] >= op.N) return -1;
        op.qmap[op.qidx[q]] = (int)q;
    }
    const size_t ts = op.T.r.size();
    for (int t = 0; t < 7; ++t) op.qtab[t].assign(ts, 0.0);
    op.self_grad = quad_self_grad(op.xi);
    for (int t = 3; t < 7; ++t) op.qtab[t][0] = op.self_grad;
    for (size_t e = 1; e < ts; ++e) {
        double v[7];
        quad_pair_functions(op.T.r[e], op.xi, v);
        for (int t = 0; t < 7; ++t) op.qtab[t][e] = v[t];
    }
    compute_scale(op);
    build_neighbors(op);
    return (int)op.qidx.size();
}

void orc_quad_pair_functions(const double* r, int n, double xi, double* out) {
    for (int i = 0; i < n; ++i) quad_pair_functions(r[i], xi, out + 7 * i);
}

double orc_quad_self_grad(double xi) { return quad_self_grad(xi); }

void orc_op_destroy(void* h) { delete static_cast<Op*>(h); }

void orc_op_set_bloch(void* h, double kx, double ky) {
    Op& op = *static_cast<Op*>(h);
    op.kbx = kx;
    op.kby = ky;
    compute_scale(op);
}

int orc_op_set_positions(void* h, const double* xyz, int N) {
    Op& op = *static_cast<Op*>(h);
    if (op.poly && N != (int)op.radius.size()) return -1;
    op.N = N;
    op.x.resize(N);
    op.y.resize(N);
    op.z.resize(N);
    for (int i = 0; i < N; ++i) {
        op.x[i] = xyz[3 * i];
        op.y[i] = xyz[3 * i + 1];
        op.z[i] = xyz[3 * i + 2];
    }
    build_neighbors(op);
    build_stencils(op);
    return op.nb_start[N];
}

void orc_op_info(void* h, int* n, double* h_out, double* eta, int* P, int* width, double* rc,
                 double* self_perp, long* num_pairs) {
    Op& op = *static_cast<Op*>(h);
    for (int d = 0; d < 3; ++d) {
        n[d] = op.S.n[d];
        h_out[d] = op.S.h[d];
        eta[d] = op.S.eta[d];
    }
    *P = op.S.P;
    *width = op.S.width;
    *rc = op.S.rc;
    if (op.poly && op.exact) {
        const double amin = *std::min_element(op.radius.begin(), op.radius.end());
        *self_perp = poly_self(amin, amin, op.xi);
    } else {
        *self_perp = op.poly ? op.uself[0] : op.T.self_perp;
    }
    *num_pairs = op.N ? op.nb_start[op.N] : 0;
}

void orc_op_apply(void* h, const double* self_coef, const double* p, double* E, int parts) {
    const Op& op = *static_cast<Op*>(h);
    apply(op, reinterpret_cast<const cplx*>(self_coef), reinterpret_cast<const cplx*>(p),
          reinterpret_cast<cplx*>(E), parts);
}

// method: 0 = GMRES(restart), 1 = BiCGSTAB.  x holds x0 on entry.
// report: [iters, matvecs, converged, relres_pre, relres_true] as doubles.
int orc_solve(void* h, int method, const double* self_coef, const double* b, double* x,
              double tol, int restart, int maxiter, int use_jacobi, double* report) {
    const Op& op = *static_cast<Op*>(h);
    SolveReport rep;
    const cplx* sc = reinterpret_cast<const cplx*>(self_coef);
    if (method == 0)
        gmres(op, sc, reinterpret_cast<const cplx*>(b), reinterpret_cast<cplx*>(x), tol, restart,
              maxiter, use_jacobi != 0, rep);
    else
        bicgstab(op, sc, reinterpret_cast<const cplx*>(b), reinterpret_cast<cplx*>(x), tol, maxiter,
                 use_jacobi != 0, rep);
    report[0] = rep.iters;
    report[1] = rep.matvecs;
    report[2] = rep.converged;
    report[3] = rep.relres_pre;
    report[4] = rep.relres_true;
    return 0;
}

int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
