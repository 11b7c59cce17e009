// Spectral-Ewald field operator y = A x for a batch of right-hand sides on gfx950:
// monodisperse dipoles, monodisperse quadrupoles, the polydisperse (finite-sphere)
// operator and the direct open-boundary field.
//
// A = diag(self_coef_b + self term) + G_real + G_recip, the operator of
// src/cpp/electric_field/monodisperse_ewald_electric_field.cu (SOLVER_AX and
// INTERACTION_FIELD modes), polydisperse_ewald_electric_field.cu and
// direct_electric_field.cu.  The arithmetic follows the reference; the data movement
// does not:
//
//  * layout  — thin films run on a slab of the occupied z-planes (the z transform is
//              replaced by a Toeplitz coupling of those planes); other geometries,
//              quadrupoles and the polydisperse operator on the full grid.
//  * spread  — region products on the matrix cores (spread_mfma_kernel, separable
//              weights V^T W^T per region, every voxel written once): x by 16-row
//              y regions over all slab planes, or 4 x 8 x 16 (x, y, z) regions on the
//              full grid (monodisperse dipoles); owner-computes tiles that pull in the
//              binned particles where the regions do not apply (small full grids,
//              quadrupoles, polydisperse); no atomics, in place of the reference's
//              atomic scatter (:1160-1298).
//  * FFT     — the built-in batched LDS line FFT (register radices 2-23, a generic
//              stage beyond; lines held in LDS with padded rows), 2-D on the slab with a
//              transposed k-space buffer and 3-D on the full grid; rocFFT as an
//              alternative back-end.
//  * scale   — full grid: the Green's tensor projection per voxel (:1406-1432), in FP32
//              fused with the forward and inverse z transforms (zline_scale_kernel);
//              slab: xtoeplitz_kernel fuses the x transform, the plane coupling and the
//              inverse x transform (FP32), ztoeplitz / zfused kernels otherwise.
//  * gather  — the transposes of the spread's region products on the matrix cores
//              (contract_mfma_kernel and its narrow-batch form on the slab,
//              contract_full_mfma_kernel on the full grid; 16-byte complex-pair loads)
//              into per-slot partials, then finish_kernel (slot sums, real-space
//              neighbour sum over a symmetric CSR list, self term, Jacobi); elsewhere a
//              wavefront per particle fusing the contraction (:1488-1525), the
//              real-space sum (:1631-1869) and the self term (:1589-1629); one store
//              per particle and right-hand side.
//  * precalc — per frame: the neighbour list and pair coefficients (monodisperse) and
//              the region weight rows on the device; stencil weights, bins and region
//              candidate lists on host threads.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <thread>

#include "mpm_internal.h"

namespace mpm {

static const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Host: sizing, real-space pair functions
// ---------------------------------------------------------------------------

// monodisperse_ewald_electric_field.cu:546-601
void compute_sizing(const double L[3], double errortol, double xi, const int grid_override[3],
                    int P_override, Sizing& S) {
    MPM_REQUIRE(errortol > 0.0 && errortol < 1.0, MPM_ERR_INVALID_ARGUMENT,
                "errortol must lie in (0, 1)");
    MPM_REQUIRE(xi > 0.0, MPM_ERR_INVALID_ARGUMENT, "xi must be positive");
    for (int d = 0; d < 3; ++d)
        MPM_REQUIRE(L[d] > 0.0, MPM_ERR_INVALID_ARGUMENT, "box lengths must be positive");
    S.rc = std::sqrt(-std::log(errortol)) / xi;
    if (S.rc < 2.0) S.rc = 2.0;
    for (int d = 0; d < 3; ++d)
        MPM_REQUIRE(!(S.rc > L[d] / 2.0), MPM_ERR_RUNTIME,
                    "Real space cutoff (" + std::to_string(S.rc) +
                        ") larger than half the box length.");
    S.kcut = 2.0 * xi * xi * S.rc;
    for (int d = 0; d < 3; ++d) {
        S.n[d] = (int)std::ceil(1.0 + L[d] * S.kcut / kPi);
        if (grid_override && grid_override[d] > 0) S.n[d] = grid_override[d];
        MPM_REQUIRE(S.n[d] >= 1 && S.n[d] <= 4096, MPM_ERR_INVALID_ARGUMENT, "grid size out of range");
        S.h[d] = L[d] / S.n[d];
    }
    S.P = (int)std::ceil(-2.0 * std::log(errortol) / kPi);
    if (P_override > 0) S.P = P_override;
    MPM_REQUIRE(S.P >= 1 && S.P <= 40, MPM_ERR_INVALID_ARGUMENT, "support out of range");
    for (int d = 0; d < 3; ++d) S.eta[d] = S.P * std::pow(S.h[d] * xi, 2) / kPi;
    S.half = S.P / 2;
    S.width = 2 * S.half + 1;
}

// Volume-averaged field of a uniformly polarised unit sphere on a unit sphere
// at centre distance r, real-space Ewald part, as a sum over the kernel's three
// images r-2, r, r+2: sum_s [A_s exp(-(r+s)^2 xi^2) + B_s erfc((r+s) xi)] plus
// a short-range polynomial for overlapping spheres (r < 2).
// monodisperse_ewald_electric_field.cu:111-176.
void dipole_pair_functions(double r, double xi, double& f_perp, double& f_para) {
    const double x2 = xi * xi, x4 = x2 * x2, x5 = x4 * xi, x6 = x4 * x2;
    const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r4 * r2;
    const double sp = std::pow(kPi, 1.5);
    const double Ep = std::exp(-(r + 2.0) * (r + 2.0) * x2);
    const double Em = std::exp(-(r - 2.0) * (r - 2.0) * x2);
    const double E0 = std::exp(-r2 * x2);
    const double Cp = std::erfc((r + 2.0) * xi);
    const double Cm = std::erfc((r - 2.0) * xi);
    const double C0 = std::erfc(r * xi);
    const double c1 = 3.0 - 12.0 * x2 + 32.0 * x4;
    const double c0 = 2.0 * (3.0 + 4.0 * x2 - 32.0 * x4);
    const double q4 = 36.0 * x4 * (1.0 - 4.0 * x2);
    {   // perpendicular component (reference table field_dip_1)
        const double pre = 1.0 / (1024.0 * sp * x5 * r3);
        const double odd = 4.0 * x4 * r5 + 8.0 * x2 * (2.0 - 7.0 * x2) * r3 + c1 * r;
        const double even = -8.0 * x4 * r4 - 8.0 * x2 * (3.0 + 2.0 * x2) * r2 + c0;
        const double A0 = (-4.0 * x4 * r4 - 8.0 * x2 * (2.0 - 9.0 * x2) * r2 - 3.0 + 36.0 * x2) /
                          (512.0 * sp * x5 * r2);
        const double preb = 1.0 / (2048.0 * kPi * x6 * r3);
        const double bev = -8.0 * x6 * r6 - q4 * r4 - 18.0 * x2 * (1.0 - 8.0 * x2) * r2 + 3.0 -
                           36.0 * x2 + 256.0 * x6;
        const double B0 = (8.0 * x6 * r6 + q4 * r4 + 18.0 * x2 * (1.0 - 8.0 * x2) * r2 - 3.0 +
                           36.0 * x2) / (1024.0 * kPi * x6 * r3);
        double f = pre * (odd + even) * Ep + preb * (bev + 256.0 * x6 * r3) * Cp +
                   pre * (odd - even) * Em + preb * (bev - 256.0 * x6 * r3) * Cm + A0 * E0 + B0 * C0;
        if (r < 2.0) f += -1.0 / (4.0 * kPi * r3) + (1.0 - 9.0 * r / 16.0 + r3 / 32.0) / (4.0 * kPi);
        f_perp = f;
    }
    {   // parallel component (reference table field_dip_2)
        const double pre = 1.0 / (512.0 * sp * x5 * r3);
        const double odd = 8.0 * x4 * r5 + 2.0 * x2 * (7.0 - 20.0 * x2) * r3 - c1 * r;
        const double even = -16.0 * x4 * r4 - 4.0 * x2 * (3.0 - 4.0 * x2) * r2 - c0;
        const double A0 = (-8.0 * x4 * r4 - 2.0 * x2 * (7.0 - 36.0 * x2) * r2 + 3.0 - 36.0 * x2) /
                          (256.0 * sp * x5 * r2);
        const double preb = 1.0 / (1024.0 * kPi * x6 * r3);
        const double bev = -16.0 * x6 * r6 - q4 * r4 - 3.0 + 36.0 * x2 - 256.0 * x6;
        const double B0 = (16.0 * x6 * r6 + q4 * r4 + 3.0 - 36.0 * x2) / (512.0 * kPi * x6 * r3);
        double f = pre * (odd + even) * Ep + preb * (bev + 128.0 * x6 * r3) * Cp +
                   pre * (odd - even) * Em + preb * (bev - 128.0 * x6 * r3) * Cm + A0 * E0 + B0 * C0;
        if (r < 2.0) f += 1.0 / (2.0 * kPi * r3) + (1.0 - 9.0 * r / 8.0 + r3 / 8.0) / (4.0 * kPi);
        f_para = f;
    }
}

// Table entry 0 (monodisperse_ewald_electric_field.cu:455-457).
double dipole_self_term(double xi) {
    const double x2 = xi * xi, x3 = x2 * xi;
    return (-1.0 + 6.0 * x2 + (1.0 - 2.0 * x2) * std::exp(-4.0 * x2)) /
               (16.0 * std::pow(kPi, 1.5) * x3) +
           std::erfc(2.0 * xi) / (4.0 * kPi);
}

// ---- polydisperse (polydisperse_ewald_electric_field.cu) ----

// rc = max(sqrt(-ln tol)/xi, 2 a_max); the grid rule of the monodisperse operator;
// P = ceil(2/h_mean - 2 ln tol/pi) points from -floor((P-1)/2); eta = P (h_mean xi)^2/pi
// (one value for every axis).  :2039-2105.
static void compute_sizing_poly(const double L[3], double errortol, double xi, double amax,
                                const int grid_override[3], int P_override, Sizing& S) {
    compute_sizing(L, errortol, xi, grid_override, 0, S);  // validation and the grid rule at rc >= 2
    if (S.rc < 2.0 * amax) {
        S.rc = 2.0 * amax;
        for (int d = 0; d < 3; ++d)
            MPM_REQUIRE(!(S.rc > L[d] / 2.0), MPM_ERR_RUNTIME,
                        "Real space cutoff (" + std::to_string(S.rc) + ") larger than half the box length.");
        S.kcut = 2.0 * xi * xi * S.rc;
        for (int d = 0; d < 3; ++d) {
            S.n[d] = (int)std::ceil(1.0 + L[d] * S.kcut / kPi);
            if (grid_override && grid_override[d] > 0) S.n[d] = grid_override[d];
            MPM_REQUIRE(S.n[d] >= 1 && S.n[d] <= 4096, MPM_ERR_INVALID_ARGUMENT, "grid size out of range");
            S.h[d] = L[d] / S.n[d];
        }
    }
    const double hm = (S.h[0] + S.h[1] + S.h[2]) / 3.0;
    S.P = (int)std::ceil(2.0 / hm - 2.0 * std::log(errortol) / kPi);
    if (P_override > 0) S.P = P_override;
    MPM_REQUIRE(S.P >= 1 && S.P <= 40, MPM_ERR_INVALID_ARGUMENT, "support out of range");
    const double eta = S.P * std::pow(hm * xi, 2) / kPi;
    for (int d = 0; d < 3; ++d) S.eta[d] = eta;
    S.half = (S.P - 1) / 2;
    S.width = S.P;
}

// Forward-mode dual number v + d eps (eps^2 = 0): the closed forms below, written once
// for a generic scalar, give their derivative with respect to one radius when that radius
// is seeded with d = 1 (radius gradients of the polydisperse operator).
namespace dual {  // found by argument-dependent lookup only: no overload hides ::exp, ::min, ...
struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double x) : v(x) {}  // NOLINT: constants promote
    Dual(double x, double dx) : v(x), d(dx) {}
};
static inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
static inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
static inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
static inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
static inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
static inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
static inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
static inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
static inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
static inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
static inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
static inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
static inline Dual operator/(double a, Dual b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
static inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
static inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
static inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
static inline Dual exp(Dual a) {
    const double e = std::exp(a.v);
    return {e, e * a.d};
}
static inline Dual erfc(Dual a) { return {std::erfc(a.v), -2.0 / std::sqrt(kPi) * std::exp(-a.v * a.v) * a.d}; }
static inline Dual erf(Dual a) { return {std::erf(a.v), 2.0 / std::sqrt(kPi) * std::exp(-a.v * a.v) * a.d}; }
static inline Dual sqrt(Dual a) {
    const double q = std::sqrt(a.v);
    return {q, 0.5 * a.d / q};
}
static inline Dual abs(Dual a) { return a.v < 0.0 ? -a : a; }
static inline Dual min(Dual a, Dual b) { return a.v <= b.v ? a : b; }
static inline Dual ipow(Dual a, int n) { return {std::pow(a.v, n), n * std::pow(a.v, n - 1) * a.d}; }
}  // namespace dual
using dual::Dual;
static inline double ipow(double a, int n) { return std::pow(a, n); }

// Real-space pair functions of uniformly polarised spheres of radii a, b at
// centre distance r: odd extensions in the radii,
//   f = sum_{s,t = +-1} [A(r; s a, t b) exp(-(r + s a + t b)^2 xi^2) + C(r; s a, t b) erfc((r + s a + t b) xi)]
//       + the direct overlap term,
// A and C carrying the factor 1/(a^3 b^3) of their own signed arguments.  :1434-1621.
template <typename T>
struct PolyTerm {
    T Ap, Cp, Aa, Ca;  // perpendicular and parallel coefficients
};
template <typename T>
static PolyTerm<T> poly_terms(T r, T a, T b, double xi) {
    const double x2 = xi * xi, x4 = x2 * x2, x5 = x4 * xi, x6 = x4 * x2;
    const T r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r4 * r2;
    const T a2 = a * a, b2 = b * b, a3 = a2 * a, b3 = b2 * b, s = a + b, s3 = s * s * s;
    const T q = a2 - 4.0 * a * b + b2, d2 = (a2 - b2) * (a2 - b2);
    const T c3 = 4.0 * a3 - 3.0 * a2 * b - 3.0 * a * b2 + 4.0 * b3;
    const T e3 = a3 - 3.0 * a2 * b - 3.0 * a * b2 + b3;
    const T m2 = a2 - a * b + b2;
    const T de = std::pow(kPi, 1.5) * a3 * b3 * r3 * x5, dc = kPi * a3 * b3 * r3 * x6;
    PolyTerm<T> t;
    t.Ap = (4.0 * x4 * r5 - 4.0 * s * x4 * r4 + (16.0 * x2 + 8.0 * (-4.0 * a2 + a * b - 4.0 * b2) * x4) * r3 +
            (-12.0 * s * x2 - 8.0 * c3 * x4) * r2 + (3.0 - 12.0 * m2 * x2 - 4.0 * s * s * q * x4) * r + 3.0 * s +
            4.0 * c3 * x2 + 4.0 * s3 * q * x4) / (1024.0 * de);
    t.Cp = (-8.0 * x6 * r6 + 36.0 * x4 * (-1.0 + 2.0 * (a2 + b2) * x2) * r4 + 128.0 * (a3 + b3) * x6 * r3 +
            (-18.0 * x2 + 72.0 * (a2 + b2) * x4 + 72.0 * d2 * x6) * r2 + 3.0 - 18.0 * (a2 + b2) * x2 - 36.0 * d2 * x4 -
            8.0 * s3 * s * q * x6) / (2048.0 * dc);
    t.Aa = (8.0 * x4 * r5 - 8.0 * s * x4 * r4 + (14.0 * x2 - 4.0 * (7.0 * a2 - 4.0 * a * b + 7.0 * b2) * x4) * r3 +
            (-6.0 * s * x2 - 4.0 * e3 * x4) * r2 + (-3.0 + 12.0 * m2 * x2 + 4.0 * s * s * q * x4) * r - 3.0 * s -
            4.0 * c3 * x2 - 4.0 * s3 * q * x4) / (512.0 * de);
    t.Ca = (-16.0 * x6 * r6 + 36.0 * x4 * (-1.0 + 2.0 * (a2 + b2) * x2) * r4 + 64.0 * (a3 + b3) * x6 * r3 - 3.0 +
            18.0 * (a2 + b2) * x2 + 36.0 * d2 * x4 + 8.0 * s3 * s * q * x6) / (1024.0 * dc);
    return t;
}

template <typename T>
static void poly_pair_functions(T r, T a, T b, double xi, T& fperp, T& fpara) {
    using std::abs;
    using std::erfc;
    using std::exp;
    T fp = 0.0, fa = 0.0;
    for (int sa = -1; sa <= 1; sa += 2)
        for (int sb = -1; sb <= 1; sb += 2) {
            const T u = r + (double)sa * a + (double)sb * b;
            const T e = exp(-u * u * xi * xi), c = erfc(u * xi);
            const PolyTerm<T> t = poly_terms<T>(r, (double)sa * a, (double)sb * b, xi);
            fp += t.Ap * e + t.Cp * c;
            fa += t.Aa * e + t.Ca * c;
        }
    const T r3 = r * r * r, a3 = a * a * a, b3 = b * b * b;
    if (r < a + b && r >= abs(a - b)) {  // lens-shaped overlap
        const T w = a + b - r;
        fp += w * w * w / (128.0 * kPi * a3 * b3 * r3) *
              (-r3 - 3.0 * (a + b) * r * r + 3.0 * (a * a - 4.0 * a * b + b * b) * r + a3 - 3.0 * a * a * b -
               3.0 * a * b * b + b3);
        fa += -(ipow(a + b, 4) * (a * a - 4.0 * a * b + b * b) - 8.0 * (a3 + b3) * r3 +
                9.0 * (a * a + b * b) * r3 * r - 2.0 * r3 * r3) / (64.0 * kPi * a3 * b3 * r3);
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

// r -> 0 limit (concentric spheres): a radius pair's table entry 0, and for a = b
// a particle's real-space self term.  :1611-1620.
template <typename T>
static T poly_self_term(T a, T b, double xi) {
    using std::erf;
    using std::exp;
    using std::min;
    const double x2 = xi * xi, x3 = x2 * xi;
    const T a3 = a * a * a, b3 = b * b * b;
    return ((1.0 - 2.0 * a * a * x2 + 2.0 * a * b * x2 - 2.0 * b * b * x2) * exp(-(a + b) * (a + b) * x2) +
            (-1.0 + 2.0 * a * a * x2 + 2.0 * a * b * x2 + 2.0 * b * b * x2) * exp(-(a - b) * (a - b) * x2)) /
               (16.0 * std::pow(kPi, 1.5) * a3 * b3 * x3) +
           ((a3 - b3) * erf((a - b) * xi) - (a3 + b3) * erf((a + b) * xi)) / (8.0 * kPi * a3 * b3) +
           min(a3, b3) / (4.0 * kPi * a3 * b3);
}

// Finite-sphere spread/contract weight vector of a particle of radius a at grid
// offset gd (grid point minus centre): C(|gd|) gd / |gd|, with the reference's
// separate form below |gd| = 1e-6.  :243-262.
template <typename T>
static void poly_weight(const double gd[3], T a, double xi, double eta, T w[3]) {
    using std::exp;
    const double d = std::sqrt(gd[0] * gd[0] + gd[1] * gd[1] + gd[2] * gd[2]);
    const double x2 = xi * xi;
    T c;
    if (d >= 1e-6) {
        const T k = 3.0 / (8.0 * std::sqrt(2.0) * std::pow(kPi, 1.5) * a * a * a * std::sqrt(eta) * xi * d * d);
        c = k * ((eta + 4.0 * a * x2 * d) * exp(-2.0 * (d + a) * (d + a) * x2 / eta) -
                 (eta - 4.0 * a * x2 * d) * exp(-2.0 * (d - a) * (d - a) * x2 / eta)) / d;
    } else {
        c = 8.0 * std::sqrt(2.0) * std::pow(xi, 5.0) * exp(-2.0 * a * x2 / eta) /
            (std::pow(kPi, 1.5) * std::pow(eta, 2.5));
    }
    for (int q = 0; q < 3; ++q) w[q] = c * gd[q];
}

// ---- quadrupole pair functions (monodisperse_ewald_electric_field.cu:199-470) ----
// Each is  A+ e^{-(r+2)^2 xi^2} + B+ erfc((r+2) xi) + A- e^{-(r-2)^2 xi^2}
//        + B- erfc((r-2) xi) + A0 e^{-r^2 xi^2} + B0 erfc(r xi) + [r < 2] reg,
// with A+- = a (O +- E) for the odd and even powers O, E of r and B+- = b (E' +- O').
// Order: field_quad_1..3 (field of a quadrupole), grad_quad_1..4 (gradient couplings).
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

// The reference samples the pair functions on r = 1 + 0.001 j (j < 9000) and
// interpolates linearly, clamping at r >= 9.999 (:92-97, :668-685).  The pair
// values are therefore computed here the same way, from the same samples.
struct PairTable {
    std::vector<double> r, perp, para;
    explicit PairTable(double xi) {
        const int n = 9000;
        r.resize(n + 1);
        perp.resize(n + 1);
        para.resize(n + 1);
        const double s = dipole_self_term(xi);
        r[0] = 0.0;
        perp[0] = para[0] = s;
        for (int j = 0; j < n; ++j) {
            r[j + 1] = 1.0 + j * 0.001;
            dipole_pair_functions(r[j + 1], xi, perp[j + 1], para[j + 1]);
        }
    }
    // quadrupole tables on the same samples: field_quad_1..3 (entry 0 = 0) and
    // grad_quad_1..4 (entry 0 = the gradient self term); built on first use
    std::vector<double> q[7];
    double self_grad = 0.0;
    void build_quads(double xi) {
        if (!q[0].empty()) return;
        const size_t ts = r.size();
        for (auto& t : q) t.assign(ts, 0.0);
        self_grad = quad_self_grad(xi);
        for (int t = 3; t < 7; ++t) q[t][0] = self_grad;
        for (size_t e = 1; e < ts; ++e) {
            double v[7];
            quad_pair_functions(r[e], xi, v);
            for (int t = 0; t < 7; ++t) q[t][e] = v[t];
        }
    }
    double interp(const std::vector<double>& y, double x) const {
        const size_t ts = y.size();
        if (x <= 0.0) return y[0];
        if (x >= r[ts - 1]) return y[ts - 1];
        if (x < 1.0) return y[0] * (1.0 - x) + y[1] * x;
        size_t idx = 1 + (size_t)((x - 1.0) / 0.001);
        if (idx >= ts - 1) return y[ts - 1];
        const double t = (x - r[idx]) / (r[idx + 1] - r[idx]);
        return y[idx] * (1.0 - t) + y[idx + 1] * t;
    }
};

// ---------------------------------------------------------------------------
// Profiling
// ---------------------------------------------------------------------------
namespace {
struct ProfEntry {
    double ms = 0;
    int64_t count = 0;
    int64_t units = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::map<std::string, ProfEntry> g_prof;
struct Pending {
    std::string name, kname;
    int64_t units;
    hipEvent_t e0, e1;
};
std::vector<Pending> g_pending;
}  // namespace

bool profiling_enabled() { return g_prof_on; }

ProfScope::ProfScope(const char* n, hipStream_t s, int64_t u) : name(n), units(u), stream(s) {
    if (!g_prof_on) return;
    MPM_HIP(hipEventCreate(&e0));
    MPM_HIP(hipEventCreate(&e1));
    MPM_HIP(hipEventRecord(e0, stream));
}
void ProfScope::kernel(const char* base, std::initializer_list<std::string> targs) {
    if (!e0) return;
    kname = base;
    kname += '<';
    bool first = true;
    for (const std::string& a : targs) {
        if (!first) kname += ", ";
        kname += a;
        first = false;
    }
    kname += '>';
}
ProfScope::~ProfScope() {
    if (!e0) return;
    (void)hipEventRecord(e1, stream);
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_pending.push_back({name, kname, units, e0, e1});
}

HostScope::HostScope(const char* n) : name(n), on(g_prof_on) {
    if (on) t0 = std::chrono::steady_clock::now();
}
HostScope::~HostScope() {
    if (!on) return;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::lock_guard<std::mutex> g(g_prof_mu);
    auto& e = g_prof[name];
    e.ms += ms;
    e.count += 1;
}

HostLaps::HostLaps() : on(g_prof_on) {
    if (on) t = std::chrono::steady_clock::now();
}
void HostLaps::lap(const char* name) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
    std::lock_guard<std::mutex> g(g_prof_mu);
    auto& e = g_prof[name];
    e.ms += ms;
    e.count += 1;
}

void profile_flush() {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& p : g_pending) {
        (void)hipEventSynchronize(p.e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.e0, p.e1);
        for (const std::string* k : {&p.name, &p.kname}) {
            if (k->empty()) continue;
            auto& e = g_prof[*k];
            e.ms += ms;
            e.count += 1;
            e.units += p.units;
        }
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
    g_pending.clear();
}
void profile_set(bool on) {
    profile_flush();
    g_prof_on = on;
}
void profile_reset() {
    profile_flush();
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof.clear();
}
bool profile_get(const std::string& name, double& ms, int64_t& count, int64_t& units) {
    profile_flush();
    std::lock_guard<std::mutex> g(g_prof_mu);
    auto it = g_prof.find(name);
    if (it == g_prof.end()) {
        ms = 0;
        count = 0;
        units = 0;
        return false;
    }
    ms = it->second.ms;
    count = it->second.count;
    units = it->second.units;
    return true;
}
std::string profile_names() {
    profile_flush();
    std::lock_guard<std::mutex> g(g_prof_mu);
    std::string out;
    for (const auto& kv : g_prof) {
        out += kv.first;
        out += '\n';
    }
    return out;
}


// ---------------------------------------------------------------------------
// rocFFT plans
// ---------------------------------------------------------------------------
#define MPM_FFT(call)                                                                          \
    do {                                                                                       \
        rocfft_status s_ = (call);                                                             \
        if (s_ != rocfft_status_success)                                                       \
            throw ::mpm::Error(MPM_ERR_FFT, std::string(#call) + " failed with rocfft status " + \
                                                std::to_string((int)s_));                      \
    } while (0)

static std::once_flag g_rocfft_once;

// Batched in-place complex transforms of `per_rhs` contiguous arrays per right-hand
// side (3 components for the full 3-D grid; 3 x Zs planes for the slab's 2-D
// transforms).  Plans are created per (precision, batch bucket) on first use.
struct FftPlans {
    struct Pair {
        rocfft_plan fwd = nullptr, inv = nullptr;
        size_t work = 0;
    };
    std::vector<size_t> lengths;  // fastest first
    int per_rhs = 3;
    std::map<std::pair<int, int>, Pair> plans;
    DevBuf<char> work;
    rocfft_execution_info info = nullptr;

    FftPlans(std::vector<size_t> len, int per) : lengths(std::move(len)), per_rhs(per) {
        std::call_once(g_rocfft_once, [] { rocfft_setup(); });
        MPM_FFT(rocfft_execution_info_create(&info));
    }
    ~FftPlans() {
        for (auto& kv : plans) {
            if (kv.second.fwd) rocfft_plan_destroy(kv.second.fwd);
            if (kv.second.inv) rocfft_plan_destroy(kv.second.inv);
        }
        if (info) rocfft_execution_info_destroy(info);
    }
    Pair& get(bool fp32, int batch) {
        auto key = std::make_pair((int)fp32, batch);
        auto it = plans.find(key);
        if (it != plans.end()) return it->second;
        Pair p;
        const rocfft_precision prec = fp32 ? rocfft_precision_single : rocfft_precision_double;
        MPM_FFT(rocfft_plan_create(&p.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward, prec,
                                   lengths.size(), lengths.data(), (size_t)batch * per_rhs, nullptr));
        MPM_FFT(rocfft_plan_create(&p.inv, rocfft_placement_inplace, rocfft_transform_type_complex_inverse, prec,
                                   lengths.size(), lengths.data(), (size_t)batch * per_rhs, nullptr));
        size_t w1 = 0, w2 = 0;
        MPM_FFT(rocfft_plan_get_work_buffer_size(p.fwd, &w1));
        MPM_FFT(rocfft_plan_get_work_buffer_size(p.inv, &w2));
        p.work = std::max(w1, w2);
        return plans.emplace(key, p).first->second;
    }
    // Transform B right-hand sides starting at `grid`, decomposing the batch into
    // power-of-two buckets so that a handful of plans cover any B.
    void exec(void* grid, size_t bytes_per_rhs, int B, bool fp32, bool forward, hipStream_t s) {
        char* base = static_cast<char*>(grid);
        int done = 0;
        while (done < B) {
            int bucket = 1;
            while (bucket * 2 <= B - done && bucket * 2 <= kMaxSlots) bucket *= 2;
            Pair& p = get(fp32, bucket);
            work.reserve(p.work);
            MPM_FFT(rocfft_execution_info_set_stream(info, s));
            if (p.work) MPM_FFT(rocfft_execution_info_set_work_buffer(info, work.p, p.work));
            void* buf[1] = {base + (size_t)done * bytes_per_rhs};
            MPM_FFT(rocfft_execute(forward ? p.fwd : p.inv, buf, nullptr, info));
            done += bucket;
        }
    }
};

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------
template <typename R>
struct C2;
template <>
struct C2<double> {
    using T = double2;
};
template <>
struct C2<float> {
    using T = float2;
};

// ---------------------------------------------------------------------------
// Built-in batched line FFT: unnormalised complex DFTs of length n along one
// axis of a strided array, computed in LDS with the Stockham auto-sort
// formulation: a stage of radix R maps butterfly j of a line (j < n/R) to
//   out[(j/Ns) Ns R + j%Ns + q Ns] = sum_r in[j + r n/R] W_{Ns R}^{(j%Ns) r} W_R^{r q}.
// The planner packs the prime factors of n into as few radices as possible
// (composites up to 16, primes up to 23 on their own: 210 -> 14 x 15,
// 114 -> 19 x 6), and every radix runs as a register DFT of compile-time size:
// odd primes with the symmetric-pair direct form, composites as an in-register
// four-step of smaller DFTs with constant twiddles.  Prime factors above 23
// run as a generic stage in which a thread computes whole outputs from the
// twiddle table.  Stages alternate between two LDS buffers of TL lines (32 KB
// together) plus the twiddles; one global read and one global write per element.
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 24;
constexpr int kMaxPrimeRadix = 23;  // larger prime factors use the generic stage
constexpr int kMaxCompositeRadix = 16;

template <typename R>
constexpr int fft_elems() {  // lines x length per block (two LDS buffers of this size)
    return sizeof(R) == 4 ? 2048 : 1024;
}

// Stage constants are computed on the host so that the kernel does no integer
// division: stage st has radix radix[st], m = n / radix butterflies per line,
// Ns = product of the earlier radices, step = n / (Ns radix).
struct LinePlan {
    int n;
    int nstages;
    int radix[kMaxStages];
    int m[kMaxStages];
    int ns[kMaxStages];
    int step[kMaxStages];
    float rm[kMaxStages];
    float rns[kMaxStages];
};

struct LineGeom {
    long long nlines;
    long long inner;         // lines per outer block
    long long inner_stride;  // distance between consecutive lines of a block
    long long outer;         // distance between blocks
    long long elem_stride;   // distance between consecutive elements of a line
};

__device__ inline long long line_base(const LineGeom& g, long long l) {
    return (l / g.inner) * g.outer + (l % g.inner) * g.inner_stride;
}

// x / d and x % d for 0 <= x < 2^22 via a float reciprocal and one correction
// (24-bit multiply: full rate).
__device__ inline void fdivmod(int x, int d, float rcp, int& q, int& r) {
    q = (int)((float)x * rcp);
    r = x - __mul24(q, d);
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
}

template <typename CT>
__device__ inline CT cmul_t(CT a, CT b) {
    CT r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <int P>
struct SmallestFactor {
    static constexpr int v = (P % 2 == 0)    ? 2
                             : (P % 3 == 0)  ? 3
                             : (P % 5 == 0)  ? 5
                             : (P % 7 == 0)  ? 7
                             : (P % 11 == 0) ? 11
                             : (P % 13 == 0) ? 13
                                             : P;
};

// cos and sin of 2 pi m / P; exact on the multiples of pi / 4.  Called with
// compile-time arguments only (inside fully unrolled loops), so it folds to
// constants.
__device__ __forceinline__ void unit_root(int m, int P, double& c, double& s) {
    m = ((m % P) + P) % P;
    if ((8 * m) % P == 0) {
        const int o = 8 * m / P;
        const double h = 0.70710678118654752440;
        const double cs[8] = {1.0, h, 0.0, -h, -1.0, -h, 0.0, h};
        const double sn[8] = {0.0, h, 1.0, h, 0.0, -h, -1.0, -h};
        c = cs[o];
        s = sn[o];
    } else {
        const double a = 2.0 * 3.14159265358979323846 * (double)m / (double)P;
        c = __builtin_cos(a);
        s = __builtin_sin(a);
    }
}

// x * exp(-+2 pi i m / P) (minus for the forward transform), with the trivial
// rotations done as swaps and negations.
template <typename CT, bool INV>
__device__ __forceinline__ CT rot(CT x, int m, int P) {
    m = ((m % P) + P) % P;
    if (m == 0) return x;
    CT r;
    if (2 * m == P) {
        r.x = -x.x;
        r.y = -x.y;
        return r;
    }
    if (4 * m == P || 4 * m == 3 * P) {  // multiply by +-i
        const bool plus_i = (4 * m == P) == INV;
        if (plus_i) {
            r.x = -x.y;
            r.y = x.x;
        } else {
            r.x = x.y;
            r.y = -x.x;
        }
        return r;
    }
    double c, s;
    unit_root(m, P, c, s);
    if (!INV) s = -s;
    using R = decltype(x.x);
    const R cr = (R)c, sr = (R)s;
    r.x = x.x * cr - x.y * sr;
    r.y = x.x * sr + x.y * cr;
    return r;
}

// In-place DFT of P values held in registers, natural order in and out.
template <typename R, int P, bool INV>
__device__ __forceinline__ void dft_reg(typename C2<R>::T* v);

// Odd prime P: X_0 = sum x_j; for k = 1..(P-1)/2 with s_j = x_j + x_{P-j},
// d_j = x_j - x_{P-j}: A_k = x_0 + sum_j s_j cos(2 pi jk/P),
// B_k = sum_j d_j sin(2 pi jk/P), X_k = A_k -+ i B_k, X_{P-k} = A_k +- i B_k.
template <typename R, int P, bool INV>
__device__ __forceinline__ void dft_odd_prime(typename C2<R>::T* v) {
    using CT = typename C2<R>::T;
    constexpr int H = (P - 1) / 2;
    CT sm[H + 1], df[H + 1];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        sm[j].x = v[j].x + v[P - j].x;
        sm[j].y = v[j].y + v[P - j].y;
        df[j].x = v[j].x - v[P - j].x;
        df[j].y = v[j].y - v[P - j].y;
    }
    const CT x0 = v[0];
    CT X0 = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        X0.x += sm[j].x;
        X0.y += sm[j].y;
    }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        R ar = x0.x, ai = x0.y, br = R(0), bi = R(0);
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            double c, s;
            unit_root(j * k, P, c, s);
            ar += sm[j].x * (R)c;
            ai += sm[j].y * (R)c;
            br += df[j].x * (R)s;
            bi += df[j].y * (R)s;
        }
        if (!INV) {
            v[k].x = ar + bi;
            v[k].y = ai - br;
            v[P - k].x = ar - bi;
            v[P - k].y = ai + br;
        } else {
            v[k].x = ar - bi;
            v[k].y = ai + br;
            v[P - k].x = ar + bi;
            v[P - k].y = ai - br;
        }
    }
    v[0] = X0;
}

template <typename R, int P, bool INV>
__device__ __forceinline__ void dft_reg(typename C2<R>::T* v) {
    using CT = typename C2<R>::T;
    if constexpr (P == 1) {
        return;
    } else if constexpr (P == 2) {
        const CT a = v[0], b = v[1];
        v[0].x = a.x + b.x;
        v[0].y = a.y + b.y;
        v[1].x = a.x - b.x;
        v[1].y = a.y - b.y;
    } else if constexpr (P == 4) {
        const CT a = v[0], b = v[1], c = v[2], d = v[3];
        CT t0, t1, t2, t3, u;
        t0.x = a.x + c.x; t0.y = a.y + c.y;
        t1.x = a.x - c.x; t1.y = a.y - c.y;
        t2.x = b.x + d.x; t2.y = b.y + d.y;
        t3.x = b.x - d.x; t3.y = b.y - d.y;
        if (INV) {  // u = +i t3
            u.x = -t3.y;
            u.y = t3.x;
        } else {    // u = -i t3
            u.x = t3.y;
            u.y = -t3.x;
        }
        v[0].x = t0.x + t2.x; v[0].y = t0.y + t2.y;
        v[2].x = t0.x - t2.x; v[2].y = t0.y - t2.y;
        v[1].x = t1.x + u.x; v[1].y = t1.y + u.y;
        v[3].x = t1.x - u.x; v[3].y = t1.y - u.y;
    } else if constexpr (SmallestFactor<P>::v == P) {
        dft_odd_prime<R, P, INV>(v);
    } else {
        // P = A B: input j1 + A j2, output k2 + B k1
        constexpr int A = (P % 4 == 0) ? 4 : SmallestFactor<P>::v;
        constexpr int B = P / A;
        CT y[A][B];
#pragma unroll
        for (int j1 = 0; j1 < A; ++j1) {
#pragma unroll
            for (int j2 = 0; j2 < B; ++j2) y[j1][j2] = v[j1 + A * j2];
            dft_reg<R, B, INV>(y[j1]);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) y[j1][k2] = rot<CT, INV>(y[j1][k2], j1 * k2, P);
        }
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            CT t[A];
#pragma unroll
            for (int j1 = 0; j1 < A; ++j1) t[j1] = y[j1][k2];
            dft_reg<R, A, INV>(t);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = t[k1];
        }
    }
}

// One Stockham stage of radix P (register DFT), src -> dst (LDS ping-pong).
template <typename R, int P, bool INV>
__device__ __forceinline__ void fft_stage(const typename C2<R>::T* src, typename C2<R>::T* dst, int n, int nl,
                                          const LinePlan& pl, int st, const typename C2<R>::T* tw) {
    using CT = typename C2<R>::T;
    const int m = pl.m[st], Ns = pl.ns[st], step = pl.step[st];
    const float rm = pl.rm[st], rns = pl.rns[st];
    const int total = __mul24(nl, m);
    for (int bf = threadIdx.x; bf < total; bf += blockDim.x) {
        int li, j, jd, a;
        fdivmod(bf, m, rm, li, j);
        fdivmod(j, Ns, rns, jd, a);
        const int lb = __mul24(li, n);
        const CT* sp = src + lb + j;
        CT v[P];
#pragma unroll
        for (int r = 0; r < P; ++r) v[r] = sp[r * m];
        if (a > 0) {
            const int ta = __mul24(a, step);
#pragma unroll
            for (int r = 1; r < P; ++r) v[r] = cmul_t(v[r], tw[ta * r]);
        }
        dft_reg<R, P, INV>(v);
        CT* dp = dst + lb + __mul24(jd, Ns * P) + a;
#pragma unroll
        for (int q = 0; q < P; ++q) dp[q * Ns] = v[q];
    }
}

// Generic Stockham stage of any radix R, src -> dst.  Pass 1 multiplies every
// input x[j + r m] of src by its stage twiddle W_{Ns R}^{(j%Ns) r} in place;
// pass 2 computes each output as a length-R DFT of the twiddled inputs with the
// R-th roots held in a compact LDS table (lanes mostly read the same root).
template <typename R>
__device__ inline void stage_generic(typename C2<R>::T* src, typename C2<R>::T* dst, typename C2<R>::T* roots, int n,
                                     int nl, int Ns, int Rx, const typename C2<R>::T* tw) {
    using CT = typename C2<R>::T;
    const int m = n / Rx;
    const int step = n / (Ns * Rx);
    const int tot = nl * n;
    const float rn = 1.0f / (float)n, rns = 1.0f / (float)Ns, rr = 1.0f / (float)Rx, rm = 1.0f / (float)m;
    for (int k = threadIdx.x; k < Rx; k += blockDim.x) roots[k] = tw[k * m];
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        int li, p, r, j, jd, a;
        fdivmod(e, n, rn, li, p);
        fdivmod(p, m, rm, r, j);
        fdivmod(j, Ns, rns, jd, a);
        if (r > 0 && a > 0) src[e] = cmul_t(src[e], tw[a * r * step]);
    }
    __syncthreads();
    if (Rx & 1) {
        // odd radix: inputs r and Rx - r meet the roots w and conj(w), and outputs q and
        // Rx - q the roots w and conj(w): one thread forms both outputs from the sums
        // (x_r + x_{Rx-r}) Re w and (x_r - x_{Rx-r}) Im w
        const int H = (Rx + 1) / 2, per_line = m * H;
        const int items = nl * per_line;
        for (int e = threadIdx.x; e < items; e += blockDim.x) {
            const int li = e / per_line;
            const int t = e - li * per_line;
            const int g = t / H, q = t - g * H;
            const int jd = g / Ns, a = g - jd * Ns;
            const CT* sp = src + li * n + jd * Ns + a;
            const CT x0 = sp[0];
            R sr = R(0), si = R(0), dr = R(0), di = R(0);
            int kq = q;
            for (int r = 1; 2 * r < Rx; ++r) {
                const CT xa = sp[r * m], xb = sp[(Rx - r) * m];
                const CT w = roots[kq];
                sr += (xa.x + xb.x) * w.x;
                si += (xa.y + xb.y) * w.x;
                dr += (xa.y - xb.y) * w.y;
                di += (xa.x - xb.x) * w.y;
                kq += q;
                if (kq >= Rx) kq -= Rx;
            }
            CT o;
            o.x = x0.x + sr - dr;
            o.y = x0.y + si + di;
            dst[li * n + (jd * Rx + q) * Ns + a] = o;  // natural (li, o) order
            if (q > 0) {
                o.x = x0.x + sr + dr;
                o.y = x0.y + si - di;
                dst[li * n + (jd * Rx + Rx - q) * Ns + a] = o;
            }
        }
        return;
    }
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        int li, oi, rest, a, jd, q;
        fdivmod(e, n, rn, li, oi);
        fdivmod(oi, Ns, rns, rest, a);
        fdivmod(rest, Rx, rr, jd, q);
        const CT* sp = src + li * n + jd * Ns + a;
        R ar = R(0), ai = R(0);
        int kq = 0;
        for (int r = 0; r < Rx; ++r) {
            const CT x = sp[r * m];
            const CT w = roots[kq];
            ar += x.x * w.x - x.y * w.y;
            ai += x.x * w.y + x.y * w.x;
            kq += q;
            if (kq >= Rx) kq -= Rx;
        }
        CT o;
        o.x = ar;
        o.y = ai;
        dst[e] = o;  // outputs are in natural (li, o) order
    }
}

// Radix classes: the kernel is compiled for the radices of its class only, so
// a plan of small radices is not charged the registers of the large ones.
//   0: radices up to 8;  1: up to 16;  2: up to 16 and the primes 17, 19, 23.
template <typename R, bool INV, int CLS>
__device__ __forceinline__ void run_stage(typename C2<R>::T* src, typename C2<R>::T* dst, typename C2<R>::T* roots,
                                          int n, int nl, const LinePlan& pl, int st, const typename C2<R>::T* tw) {
    const int Rx = pl.radix[st];
#define MPM_ST(P)                                    \
    case P:                                          \
        fft_stage<R, P, INV>(src, dst, n, nl, pl, st, tw);    \
        return;
#define MPM_ST_C(P, C)                                    \
    case P:                                               \
        if constexpr (CLS >= C) {                         \
            fft_stage<R, P, INV>(src, dst, n, nl, pl, st, tw);     \
            return;                                       \
        }                                                 \
        break;
    switch (Rx) {
        MPM_ST(2)
        MPM_ST(3)
        MPM_ST(4)
        MPM_ST(5)
        MPM_ST(6)
        MPM_ST(7)
        MPM_ST(8)
        MPM_ST_C(9, 1)
        MPM_ST_C(10, 1)
        MPM_ST_C(11, 1)
        MPM_ST_C(12, 1)
        MPM_ST_C(13, 1)
        MPM_ST_C(14, 1)
        MPM_ST_C(15, 1)
        MPM_ST_C(16, 1)
        MPM_ST_C(17, 2)
        MPM_ST_C(19, 2)
        MPM_ST_C(23, 2)
        default:
            break;
    }
#undef MPM_ST
#undef MPM_ST_C
    stage_generic<R>(src, dst, roots, n, nl, pl.ns[st], Rx, tw);
}

// Line base offsets fit in 32 bits: the library limits one transform call to
// fewer than 2^31 elements.
__device__ inline int line_base32(const LineGeom& g, int l) {
    return (int)((l / g.inner) * g.outer + (l % g.inner) * g.inner_stride);
}

// One axis of the transform over TL lines per block: lines are read with the
// input geometry and written with the output geometry (in place when both
// buffers and geometries agree; the slab's y pass writes transposed).
template <typename R, bool INV, int CLS>
__global__ __launch_bounds__(256) void line_fft_kernel(const typename C2<R>::T* in, typename C2<R>::T* out,
                                                       LineGeom gin,
                                                       LineGeom gout, LinePlan pl,
                                                       const typename C2<R>::T* __restrict__ twiddle, int TL) {
    using CT = typename C2<R>::T;
    extern __shared__ unsigned char s_raw[];
    const int n = pl.n;
    // two stage buffers of TL lines, then the twiddles and the generic-stage roots;
    // the buffers are selected by offset so that every access stays an LDS access
    CT* const lds0 = reinterpret_cast<CT*>(s_raw);
    const int boff = TL * n;
    CT* tw = lds0 + 2 * boff;
    CT* roots = tw + n;
    const int tid = threadIdx.x;
    const int l0 = blockIdx.x * TL;
    const int nl = min(TL, (int)gin.nlines - l0);
    for (int k = tid; k < n; k += blockDim.x) {
        CT w = twiddle[k];
        if (INV) w.y = -w.y;
        tw[k] = w;
    }
    // a run: each line contiguous and consecutive lines adjacent
    const bool run_in = gin.elem_stride == 1 && gin.inner == 1 && gin.outer == n;
    const bool run_out = gout.elem_stride == 1 && gout.inner == 1 && gout.outer == n;
    __shared__ int s_lbi[64], s_lbo[64];
    if (tid < nl) {
        s_lbi[tid] = line_base32(gin, l0 + tid);
        s_lbo[tid] = line_base32(gout, l0 + tid);
    }
    __syncthreads();
    const int tot = __mul24(nl, n);
    const float rn = 1.0f / (float)n, rnl = 1.0f / (float)nl;
    CT* buf = lds0;
    if (run_in) {
        const CT* src = in + __mul24(l0, n);
        for (int e = tid; e < tot; e += blockDim.x) buf[e] = src[e];
    } else if (gin.elem_stride == 1) {
        for (int e = tid; e < tot; e += blockDim.x) {
            int li, k;
            fdivmod(e, n, rn, li, k);
            buf[e] = in[s_lbi[li] + k];
        }
    } else {
        const int es = (int)gin.elem_stride;
        for (int e = tid; e < tot; e += blockDim.x) {
            int k, li;
            fdivmod(e, nl, rnl, k, li);
            buf[__mul24(li, n) + k] = in[s_lbi[li] + __mul24(k, es)];
        }
    }
    __syncthreads();
    int cur = 0;
    for (int st = 0; st < pl.nstages; ++st) {
        run_stage<R, INV, CLS>(lds0 + (cur ? boff : 0), lds0 + (cur ? 0 : boff), roots, n, nl, pl, st, tw);
        __syncthreads();
        cur ^= 1;
    }
    buf = lds0 + (cur ? boff : 0);
    if (run_out) {
        CT* dstp = out + __mul24(l0, n);
        for (int e = tid; e < tot; e += blockDim.x) dstp[e] = buf[e];
    } else if (gout.elem_stride == 1) {
        for (int e = tid; e < tot; e += blockDim.x) {
            int li, k;
            fdivmod(e, n, rn, li, k);
            out[s_lbo[li] + k] = buf[e];
        }
    } else {
        const int es = (int)gout.elem_stride;
        for (int e = tid; e < tot; e += blockDim.x) {
            int k, li;
            fdivmod(e, nl, rnl, k, li);
            out[s_lbo[li] + __mul24(k, es)] = buf[__mul24(li, n) + k];
        }
    }
}

// Two-stage transform n = P1 P2 (the common case: 210 = 14 x 15, 114 = 19 x 6)
// with one LDS round trip: stage 1 (radix P1, no twiddles) reads its inputs
// x[j + r P2] straight from global memory and leaves y[j P1 + q] in LDS; stage
// 2 (radix P2) reads y[j' + r P1], applies W_n^{j' r} and writes X[j' + q P1]
// straight to global memory.  Work items are mapped line-fastest when the
// global side of a stage is strided (coalesced across lines), element-fastest
// when it is contiguous.
template <typename R, int P, bool INV>
__device__ __forceinline__ void fft2_s1(const typename C2<R>::T* in, int ib, int istride,
                                        typename C2<R>::T* y, int yb) {
    using CT = typename C2<R>::T;
    CT v[P];
#pragma unroll
    for (int r = 0; r < P; ++r) v[r] = in[ib + r * istride];
    dft_reg<R, P, INV>(v);
#pragma unroll
    for (int q = 0; q < P; ++q) y[yb + q] = v[q];
}

template <typename R, int P, bool INV>
__device__ __forceinline__ void fft2_s2(const typename C2<R>::T* y, int yb, int ystride,
                                        const typename C2<R>::T* tw, int a, typename C2<R>::T* out,
                                        int ob, int ostride) {
    using CT = typename C2<R>::T;
    CT v[P];
#pragma unroll
    for (int r = 0; r < P; ++r) v[r] = y[yb + r * ystride];
    if (a > 0) {
#pragma unroll
        for (int r = 1; r < P; ++r) v[r] = cmul_t(v[r], tw[a * r]);
    }
    dft_reg<R, P, INV>(v);
#pragma unroll
    for (int q = 0; q < P; ++q) out[ob + q * ostride] = v[q];
}

#define MPM_FFT2_CASES(CALL)   \
    CALL(2, 0)                 \
    CALL(3, 0)                 \
    CALL(4, 0)                 \
    CALL(5, 0)                 \
    CALL(6, 0)                 \
    CALL(7, 0)                 \
    CALL(8, 0)                 \
    CALL(9, 1)                 \
    CALL(10, 1)                \
    CALL(11, 1)                \
    CALL(12, 1)                \
    CALL(13, 1)                \
    CALL(14, 1)                \
    CALL(15, 1)                \
    CALL(16, 1)                \
    CALL(17, 2)                \
    CALL(19, 2)                \
    CALL(23, 2)

template <typename R, bool INV, int CLS>
__global__ __launch_bounds__(256) void line_fft2_kernel(const typename C2<R>::T* in, typename C2<R>::T* out,
                                                        LineGeom gin,
                                                        LineGeom gout, int n, int P1, int P2,
                                                        const typename C2<R>::T* __restrict__ twiddle, int TL) {
    using CT = typename C2<R>::T;
    extern __shared__ unsigned char s_raw[];
    CT* const y = reinterpret_cast<CT*>(s_raw);
    // LDS rows of P1 + 1 and an odd line pitch: stage 1 writes rows j (or lines, when
    // lanes run over lines) of consecutive lanes without bank conflicts (a pitch of 16
    // complex would put them all on one bank set)
    const int ln = (n + P2) | 1;  // one line: P2 rows of P1 + 1, padded to odd
    CT* tw = y + TL * ln;
    const int tid = threadIdx.x;
    const int l0 = blockIdx.x * TL;
    const int nl = min(TL, (int)gin.nlines - l0);
    for (int k = tid; k < n; k += blockDim.x) {
        CT w = twiddle[k];
        if (INV) w.y = -w.y;
        tw[k] = w;
    }
    __shared__ int s_lbi[64], s_lbo[64];
    if (tid < nl) {
        s_lbi[tid] = line_base32(gin, l0 + tid);
        s_lbo[tid] = line_base32(gout, l0 + tid);
    }
    __syncthreads();
    const int esi = (int)gin.elem_stride, eso = (int)gout.elem_stride;
    const float rnl = 1.0f / (float)nl;
    // stage 1: butterflies j < P2 of radix P1
    {
        const float rm = 1.0f / (float)P2;
        const int tot = nl * P2;
        for (int w = tid; w < tot; w += blockDim.x) {
            int li, j;
            if (esi == 1) fdivmod(w, P2, rm, li, j);
            else fdivmod(w, nl, rnl, j, li);
            const int ib = s_lbi[li] + j * esi, yb = li * ln + j * (P1 + 1);
            switch (P1) {
#define MPM_S1(P, C)                                                                 \
    case P:                                                                          \
        if constexpr (CLS >= C) fft2_s1<R, P, INV>(in, ib, P2 * esi, y, yb);         \
        break;
                MPM_FFT2_CASES(MPM_S1)
#undef MPM_S1
                default:
                    break;
            }
        }
    }
    __syncthreads();
    // stage 2: butterflies j' < P1 of radix P2
    {
        const float rm = 1.0f / (float)P1;
        const int tot = nl * P1;
        for (int w = tid; w < tot; w += blockDim.x) {
            int li, j;
            if (eso == 1) fdivmod(w, P1, rm, li, j);
            else fdivmod(w, nl, rnl, j, li);
            const int yb = li * ln + j, ob = s_lbo[li] + j * eso;
            switch (P2) {
#define MPM_S2(P, C)                                                                       \
    case P:                                                                                \
        if constexpr (CLS >= C) fft2_s2<R, P, INV>(y, yb, P1 + 1, tw, j, out, ob, P1 * eso); \
        break;
                MPM_FFT2_CASES(MPM_S2)
#undef MPM_S2
                default:
                    break;
            }
        }
    }
}

// Full grid ([B][3][nx][ny][nz], z lines contiguous): the forward z transform, the
// Green's tensor projection f <- sc (f . k^) k^ and the inverse z transform in one
// pass, after the forward x and y transforms and before the inverse ones (the
// axes' transforms commute), so that the grid is read and written once for the
// three steps instead of three times.  A block takes TL z lines of each of the
// three components of one right-hand side (the projection couples them per
// k-point), both register stages of each transform as line_fft2_kernel, all in one
// LDS buffer y: forward stage 1 from the grid into y, forward stage 2 in place
// (each butterfly writes the elements it read), the projection in place, inverse
// stage 1 with every thread's inputs read before any output is written (one
// butterfly per thread: 3 TL P2 <= 256), inverse stage 2 from y to the grid.  kc
// holds (k^x, k^y, k^z, sc) per voxel.
template <typename R, int CLS>
__global__ __launch_bounds__(256) void zline_scale_kernel(typename C2<R>::T* grid, int B, int nxy, int n,
                                                          int P1, int P2,
                                                          const typename C2<R>::T* __restrict__ twiddle, int TL,
                                                          const R* __restrict__ kc) {
    using CT = typename C2<R>::T;
    constexpr int PMAX = CLS == 0 ? 8 : (CLS == 1 ? 16 : kMaxPrimeRadix);
    extern __shared__ unsigned char s_raw[];
    // a line in LDS is P2 rows of P1 + 1 (padded: no bank conflicts in stage 1); after
    // the forward transform element k of a line sits at row k / P1, column k % P1
    const int ln = (n + P2) | 1;  // odd pitch: lines of consecutive lanes on distinct banks
    CT* const y = reinterpret_cast<CT*>(s_raw);  // [3 TL][ln]
    CT* const tw = y + 3 * TL * ln;              // forward twiddles, then their conjugates
    CT* const twi = tw + n;
    const int tid = threadIdx.x;
    const int l0 = blockIdx.x * TL;
    const int nl = min(TL, B * nxy - l0);
    const int nt = 3 * nl;  // block lines t = c nl + li
    for (int k = tid; k < n; k += blockDim.x) {
        const CT w = twiddle[k];
        tw[k] = w;
        CT wc = w;
        wc.y = -wc.y;
        twi[k] = wc;
    }
    // line L = xy B + b: the right-hand sides of one (x, y) follow each other, so
    // that its k-space coefficients are read from memory once, not once per b
    __shared__ int s_base[3 * 64];  // grid offset of each block line
    __shared__ int s_vox[64];       // voxel (x, y, 0) of each line
    if (tid < nt) {
        const int c = tid / nl, li = tid - c * nl;
        const int L = l0 + li, xy = L / B, b = L - xy * B;
        s_base[tid] = ((b * 3 + c) * nxy + xy) * n;
        if (c == 0) s_vox[li] = xy * n;
    }
    __syncthreads();
    const float r2 = 1.0f / (float)P2, r1 = 1.0f / (float)P1;
    // forward stage 1: butterflies j < P2 of radix P1, grid -> y
    for (int w = tid; w < nt * P2; w += blockDim.x) {
        int t, j;
        fdivmod(w, P2, r2, t, j);
        switch (P1) {
#define MPM_S1(P, C)                                                                     \
    case P:                                                                              \
        if constexpr (CLS >= C) fft2_s1<R, P, false>(grid, s_base[t] + j, P2, y, t * ln + j * (P1 + 1)); \
        break;
            MPM_FFT2_CASES(MPM_S1)
#undef MPM_S1
            default:
                break;
        }
    }
    __syncthreads();
    // forward stage 2: butterflies a < P1 of radix P2, in place in y (element k at row
    // k / P1, column k % P1)
    for (int w = tid; w < nt * P1; w += blockDim.x) {
        int t, a;
        fdivmod(w, P1, r1, t, a);
        switch (P2) {
#define MPM_S2(P, C)                                                                           \
    case P:                                                                                    \
        if constexpr (CLS >= C) fft2_s2<R, P, false>(y, t * ln + a, P1 + 1, tw, a, y, t * ln + a, P1 + 1); \
        break;
            MPM_FFT2_CASES(MPM_S2)
#undef MPM_S2
            default:
                break;
        }
    }
    __syncthreads();
    // projection per k-point of each line
    for (int w = tid; w < nl * n; w += blockDim.x) {
        const int li = w / n, k = w - li * n;
        const int v = s_vox[li] + k;
        int kq, ka;
        fdivmod(k, P1, r1, kq, ka);
        const R kx = kc[4 * (size_t)v], ky = kc[4 * (size_t)v + 1], kz = kc[4 * (size_t)v + 2],
                sc = kc[4 * (size_t)v + 3];
        CT* f0 = y + li * ln + kq * (P1 + 1) + ka;
        CT* f1 = f0 + nl * ln;
        CT* f2 = f1 + nl * ln;
        const CT a0 = *f0, a1 = *f1, a2 = *f2;
        const R dr = (a0.x * kx + a1.x * ky + a2.x * kz) * sc;
        const R di = (a0.y * kx + a1.y * ky + a2.y * kz) * sc;
        CT o;
        o.x = dr * kx;
        o.y = di * kx;
        *f0 = o;
        o.x = dr * ky;
        o.y = di * ky;
        *f1 = o;
        o.x = dr * kz;
        o.y = di * kz;
        *f2 = o;
    }
    __syncthreads();
    // inverse stage 1: butterfly (t, j) of radix P1 reads y[t n + j + r P2] and writes
    // y[t n + j P1 + q]; all reads precede all writes
    {
        const bool act = tid < nt * P2;
        int t = 0, j = 0;
        if (act) fdivmod(tid, P2, r2, t, j);
        CT v[PMAX];
#pragma unroll
        for (int r = 0; r < PMAX; ++r)
            if (act && r < P1) {
                int kq, ka;
                fdivmod(j + r * P2, P1, r1, kq, ka);
                v[r] = y[t * ln + kq * (P1 + 1) + ka];
            }
        __syncthreads();
        if (act) {
            switch (P1) {
#define MPM_S1(P, C)                                                        \
    case P:                                                                 \
        if constexpr (CLS >= C && P <= PMAX) {                              \
            dft_reg<R, P, true>(v);                                         \
            _Pragma("unroll") for (int q = 0; q < P; ++q) y[t * ln + j * (P1 + 1) + q] = v[q]; \
        }                                                                   \
        break;
                MPM_FFT2_CASES(MPM_S1)
#undef MPM_S1
                default:
                    break;
            }
        }
    }
    __syncthreads();
    // inverse stage 2: y -> grid
    for (int w = tid; w < nt * P1; w += blockDim.x) {
        int t, a;
        fdivmod(w, P1, r1, t, a);
        switch (P2) {
#define MPM_S2(P, C)                                                                                   \
    case P:                                                                                            \
        if constexpr (CLS >= C) fft2_s2<R, P, true>(y, t * ln + a, P1 + 1, twi, a, grid, s_base[t] + a, P1); \
        break;
            MPM_FFT2_CASES(MPM_S2)
#undef MPM_S2
            default:
                break;
        }
    }
}

// Stage radices for length n: primes above kMaxPrimeRadix stand alone (generic
// stage), primes from 17 up stand alone, and the small primes are packed,
// largest first, into the fullest radix that stays <= kMaxCompositeRadix.
static std::vector<int> plan_radices(int n) {
    std::vector<int> primes;
    int m = n;
    for (int p = 2; p * p <= m; ++p)
        while (m % p == 0) {
            primes.push_back(p);
            m /= p;
        }
    if (m > 1) primes.push_back(m);
    std::sort(primes.begin(), primes.end(), std::greater<int>());
    std::vector<int> out, groups;
    for (int p : primes) {
        if (p > kMaxCompositeRadix) {
            out.push_back(p);
            continue;
        }
        int best = -1;
        for (size_t g = 0; g < groups.size(); ++g)
            if (groups[g] * p <= kMaxCompositeRadix && (best < 0 || groups[g] > groups[best])) best = (int)g;
        if (best >= 0)
            groups[best] *= p;
        else
            groups.push_back(p);
    }
    out.insert(out.end(), groups.begin(), groups.end());
    return out;
}

// One axis of the built-in FFT: plan + twiddle tables in both precisions.
struct LineAxis {
    LinePlan plan{};
    int cls = 0;       // radix class of the compiled kernel
    bool two = false;  // two register stages: line_fft2_kernel
    DevBuf<double2> tw64;
    DevBuf<float2> tw32;
    explicit LineAxis(int n) {
        plan.n = n;
        plan.nstages = 0;
        int ns = 1;
        for (int r : plan_radices(n)) {
            MPM_REQUIRE(plan.nstages < kMaxStages, MPM_ERR_UNSUPPORTED, "FFT length has too many factors");
            const int st = plan.nstages++;
            plan.radix[st] = r;
            plan.m[st] = n / r;
            plan.ns[st] = ns;
            plan.step[st] = n / (ns * r);
            plan.rm[st] = 1.0f / (float)plan.m[st];
            plan.rns[st] = 1.0f / (float)ns;
            ns *= r;
            if (r > kMaxPrimeRadix) continue;  // generic stage
            cls = std::max(cls, r <= 8 ? 0 : (r <= 16 ? 1 : 2));
        }
        two = plan.nstages == 2 && plan.radix[0] <= kMaxPrimeRadix && plan.radix[1] <= kMaxPrimeRadix;
        std::vector<double2> t64(n);
        std::vector<float2> t32(n);
        for (int k = 0; k < n; ++k) {
            // exp(-2 pi i k / n)
            const double a = -2.0 * kPi * (double)k / (double)n;
            const double sv = std::sin(a), cv = std::cos(a);
            t64[k] = make_double2(cv, sv);
            t32[k] = make_float2((float)cv, (float)sv);
        }
        tw64.upload(t64.data(), n);
        tw32.upload(t32.data(), n);
    }
};

// Built-in batched FFT over the axes of one reciprocal layout.  A pass
// transforms one axis, reading buffer `src` with geometry `gin` and writing
// buffer `dst` with `gout` (0: the spread/contraction grid, 1: the k-space
// buffer); the line descriptors count the lines of one right-hand side and
// extend unchanged over B stacked right-hand sides.
struct LinePass {
    int axis;
    LineGeom gin, gout;
    int src, dst;
};

struct LineFft {
    std::vector<std::unique_ptr<LineAxis>> axes;
    std::vector<LinePass> fwd, inv;
    long long rhs_elems = 0;  // elements per RHS (grid size)
    int kt_x = -1;            // transposed slab layout: the axis of the contiguous x lines
    int full_z = -1;          // full grid: the axis of the contiguous z lines

    // lines per block: as many as fit the per-block element budget (at most 64)
    template <typename R>
    static int lines_per_block(int n) {
        return std::max(1, std::min(64, fft_elems<R>() / n));
    }
    static bool fits(int n) { return n <= fft_elems<double>(); }

    template <typename R, bool INV, int CLS>
    static void launch(dim3 grid, size_t lds, hipStream_t s, const typename C2<R>::T* in, typename C2<R>::T* out,
                       const LineGeom& gi, const LineGeom& go, const LinePlan& pl, const typename C2<R>::T* tw,
                       int TL) {
        hipLaunchKernelGGL((line_fft_kernel<R, INV, CLS>), grid, dim3(256), lds, s, in, out, gi, go, pl, tw, TL);
    }

    // lines per block of the two-stage kernel: the most lines per unit of work,
    // counting the rounds of 256 threads each stage takes (radix-weighted), within
    // one LDS buffer of the per-block element budget (two stage buffers' worth)
    template <typename R>
    static int lines_per_block2(int n, int P1, int P2) {
        const int tlmax = std::max(1, std::min(64, 2 * fft_elems<R>() / n));
        int best = 1;
        double best_eff = 0.0;
        for (int tl = 1; tl <= tlmax; ++tl) {
            const double cost = (double)((tl * P2 + 255) / 256) * P1 + (double)((tl * P1 + 255) / 256) * P2;
            const double eff = tl / cost;
            if (eff > best_eff * 1.0001) {
                best_eff = eff;
                best = tl;
            }
        }
        return best;
    }

    template <typename R, bool INV, int CLS>
    static void launch2(dim3 grid, size_t lds, hipStream_t s, const typename C2<R>::T* in, typename C2<R>::T* out,
                        const LineGeom& gi, const LineGeom& go, const LinePlan& pl, const typename C2<R>::T* tw,
                        int TL) {
        hipLaunchKernelGGL((line_fft2_kernel<R, INV, CLS>), grid, dim3(256), lds, s, in, out, gi, go, pl.n,
                           pl.radix[0], pl.radix[1], tw, TL);
    }

    // skip_axis: passes of this axis are left out (run fused with the Green's tensor)
    template <typename R>
    void exec(typename C2<R>::T* grid, typename C2<R>::T* kgrid, int B, bool forward, hipStream_t s,
              int skip_axis = -1) {
        using CT = typename C2<R>::T;
        MPM_REQUIRE((long long)B * rhs_elems < (1LL << 31), MPM_ERR_UNSUPPORTED,
                    "reciprocal grid too large for one transform call");
        static const char* kPassNames[] = {"fft_axis0", "fft_axis1", "fft_axis2"};
        for (const LinePass& ps : (forward ? fwd : inv)) {
            if (ps.axis == skip_axis) continue;
            ProfScope pscope(kPassNames[std::min(ps.axis, 2)], s, B);
            LineAxis& ax = *axes[ps.axis];
            LineGeom gi = ps.gin, go = ps.gout;
            gi.nlines *= B;
            go.nlines *= B;
            const int n = ax.plan.n;
            const CT* tw = std::is_same<R, float>::value ? reinterpret_cast<const CT*>(ax.tw32.p)
                                                         : reinterpret_cast<const CT*>(ax.tw64.p);
            const CT* in = ps.src ? kgrid : grid;
            CT* out = ps.dst ? kgrid : grid;
            if (ax.two) {
                const int TL = lines_per_block2<R>(n, ax.plan.radix[0], ax.plan.radix[1]);
                const size_t lds = ((size_t)TL * ((n + ax.plan.radix[1]) | 1) + n) * sizeof(CT);  // padded rows
                const int nsets = (int)((gi.nlines + TL - 1) / TL);
                const dim3 gd((unsigned)nsets);
                pscope.kernel("line_fft2_kernel", {type_name<R>(), tbool(!forward), std::to_string(ax.cls)});
#define MPM_L2(INV)                                                                       \
    if (ax.cls == 0) launch2<R, INV, 0>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);    \
    else if (ax.cls == 1) launch2<R, INV, 1>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL); \
    else launch2<R, INV, 2>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
                if (forward) {
                    MPM_L2(false)
                } else {
                    MPM_L2(true)
                }
#undef MPM_L2
                MPM_HIP(hipGetLastError());
                continue;
            }
            const int TL = lines_per_block<R>(n);
            const size_t lds = ((size_t)2 * TL * n + 2 * n) * sizeof(CT);
            const long long blocks = (gi.nlines + TL - 1) / TL;
            const dim3 gd((unsigned)blocks);
            pscope.kernel("line_fft_kernel", {type_name<R>(), tbool(!forward), std::to_string(ax.cls)});
            if (forward) {
                if (ax.cls == 0) launch<R, false, 0>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
                else if (ax.cls == 1) launch<R, false, 1>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
                else launch<R, false, 2>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
            } else {
                if (ax.cls == 0) launch<R, true, 0>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
                else if (ax.cls == 1) launch<R, true, 1>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
                else launch<R, true, 2>(gd, lds, s, in, out, gi, go, ax.plan, tw, TL);
            }
            MPM_HIP(hipGetLastError());
        }
    }
};

// k-space Green's tensor at wave vector (Gx - kbx, Gy - kby, Gz): (k^x, k^y, k^z, sc)
// with sc = 9 (sin k / k - cos k)^2 exp(-(k^2 - sum_d G_d^2 eta_d) / (4 xi^2)) / k^4 / V
// and zero at k = 0.  monodisperse_ewald_electric_field.cu:1300-1404.
struct KParams {
    int nx, ny, nz;
    int tcols;  // slab columns ordered ky * nxp + kx (else kx * ny + ky)
    int nxp;    // row stride of the transposed columns (nx rounded up to whole cache lines)
    double Lx, Ly, Lz, kbx, kby, xi, etax, etay, etaz, invV;
};

// (kx, ky) of a slab column; ix >= nx marks a padding column
__device__ inline void col_xy(const KParams& q, int col, int& ix, int& iy) {
    if (q.tcols) {
        iy = col / q.nxp;
        ix = col - iy * q.nxp;
    } else {
        ix = col / q.ny;
        iy = col - ix * q.ny;
    }
}

__device__ inline int kcols(const KParams& q) { return q.tcols ? q.ny * q.nxp : q.nx * q.ny; }

__device__ inline void kcoef(const KParams& q, int ix, int iy, int iz, double o[4]) {
    o[0] = o[1] = o[2] = o[3] = 0.0;
    if (ix >= q.nx) return;  // padding column
    const double fx = (ix <= (q.nx - 1) / 2) ? ix : (ix - q.nx);
    const double fy = (iy <= (q.ny - 1) / 2) ? iy : (iy - q.ny);
    const double fz = (iz <= (q.nz - 1) / 2) ? iz : (iz - q.nz);
    const double twopi = 2.0 * 3.14159265358979323846;
    const double Gx = fx * twopi / q.Lx, Gy = fy * twopi / q.Ly, Gz = fz * twopi / q.Lz;
    const double kx = Gx - q.kbx, ky = Gy - q.kby, kz = Gz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double k = sqrt(k2);
    o[0] = o[1] = o[2] = o[3] = 0.0;
    if (k < 1e-12) return;
    o[0] = kx / k;
    o[1] = ky / k;
    o[2] = kz / k;
    const double etak2 = k2 - (Gx * Gx * q.etax + Gy * Gy * q.etay + Gz * Gz * q.etaz);
    const double t = sin(k) / k - cos(k);
    o[3] = 9.0 * t * t * exp(-etak2 / (4.0 * q.xi * q.xi)) / (k2 * k2) * q.invV;
}

__global__ void kspace_kernel(double* __restrict__ k64, float* __restrict__ k32, KParams q) {
    const size_t V = (size_t)q.nx * q.ny * q.nz;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
        const int ix = (int)(v / ((size_t)q.ny * q.nz));
        const int iy = (int)((v / q.nz) % q.ny);
        const int iz = (int)(v % q.nz);
        double o[4];
        kcoef(q, ix, iy, iz, o);
        for (int c = 0; c < 4; ++c) {
            k64[4 * v + c] = o[c];
            k32[4 * v + c] = (float)o[c];
        }
    }
}

// ---- spread: owner-computes gather into a voxel tile ----
//
// Full grid: tiles of 4 x 8 x 8 voxels (z fastest), layout [rhs][comp][x][y][z].
// Slab: tiles of 8 x 16 x 2 voxels (y fastest), layout [rhs][comp][plane][x][y]
// over the Zs occupied z-planes only.  Particles are binned by base voxel; a
// tile stages (through LDS) every particle of every bin whose stencil can reach
// it, found per axis in voxel space, and each thread sums the contributions to
// its voxel.  Grid values are written exactly once, without atomics.
template <bool SLAB>
struct Tile;
template <>
struct Tile<false> {
    static constexpr int X = 4, Y = 8, Z = 8;
};
template <>
struct Tile<true> {
    static constexpr int X = 8, Y = 16, Z = 2;
};

struct SpreadGeom {
    int nx, ny, nz;
    int half, width;
    int nofold;         // 1 when nx, ny, nz >= width (no stencil point folds onto another)
    int bx, by, bz;     // bin size (voxels)
    int nbx, nby, nbz;  // bins per axis
    int ntx, nty;       // tiles per axis (x, y)
    int Zs;             // slab planes
    int nyp;            // slab row stride (ny rounded up to whole cache lines when padded)
};

constexpr int kChunk = 64;       // particles staged in LDS per round
constexpr int kMaxWidth = 13;    // stencil points per axis
constexpr int kListMax = 12;     // candidate bins per axis
constexpr int kMaxCombos = kListMax * kListMax * kListMax;

template <typename R, int G, bool SLAB>
__global__ __launch_bounds__(256) void spread_kernel(SlotTable tab, int B, SpreadGeom g, const int* __restrict__ base,
                                                     const R* __restrict__ wts, const int* __restrict__ bin_start,
                                                     const int* __restrict__ bin_items, const int* __restrict__ ztiles,
                                                     const int* __restrict__ zmap, int N,
                                                     typename C2<R>::T* __restrict__ out) {
    using CT = typename C2<R>::T;
    constexpr int TX = Tile<SLAB>::X, TY = Tile<SLAB>::Y, TZ = Tile<SLAB>::Z;
    __shared__ int s_list[3][kListMax];
    __shared__ int s_n[3];
    __shared__ int s_pre[kMaxCombos + 1];
    __shared__ int s_part[256];
    __shared__ int s_base[kChunk][3];
    __shared__ R s_w[kChunk][3 * kMaxWidth];
    __shared__ CT s_p[kChunk][G][3];

    const int tid = threadIdx.x;
    const int t = blockIdx.x;
    const int btz = ztiles[t / (g.ntx * g.nty)];
    const int rem = t % (g.ntx * g.nty);
    const int btx = rem / g.nty, bty = rem % g.nty;
    const int g0 = blockIdx.y * G;

    int lx, ly, lz;
    if (SLAB) {
        ly = tid % TY;
        lx = (tid / TY) % TX;
        lz = tid / (TX * TY);
    } else {
        lz = tid % TZ;
        ly = (tid / TZ) % TY;
        lx = tid / (TZ * TY);
    }
    const int vx = btx * TX + lx, vy = bty * TY + ly, vz = btz * TZ + lz;
    bool valid = vx < g.nx && vy < g.ny && vz < g.nz;
    int js = 0;
    if (SLAB && valid) {
        js = zmap[vz];
        valid = js >= 0;
    }

    // candidate bins per axis
    if (tid < 3) {
        const int d = tid;
        const int T = d == 0 ? TX : (d == 1 ? TY : TZ);
        const int n = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
        const int bs = d == 0 ? g.bx : (d == 1 ? g.by : g.bz);
        const int nb = d == 0 ? g.nbx : (d == 1 ? g.nby : g.nbz);
        const int bt = d == 0 ? btx : (d == 1 ? bty : btz);
        const int lo = bt * T - g.half, hi = bt * T + T - 1 + g.half;
        int cnt = 0;
        if (hi - lo + 1 >= n) {
            for (int b = 0; b < nb && cnt < kListMax; ++b) s_list[d][cnt++] = b;
        } else {
            int last = -1;
            for (int v = lo; v <= hi; ++v) {
                const int b = (((v % n) + n) % n) / bs;
                if (b == last) continue;
                bool dup = false;
                for (int k = 0; k < cnt; ++k) dup = dup || (s_list[d][k] == b);
                if (!dup && cnt < kListMax) s_list[d][cnt++] = b;
                last = b;
            }
        }
        s_n[d] = cnt;
    }
    __syncthreads();
    const int ncx = s_n[0], ncy = s_n[1], ncz = s_n[2];
    const int nc = ncx * ncy * ncz;
    // deterministic exclusive scan of the bin sizes over the combos
    {
        const int per = (nc + 255) / 256;
        const int c0 = tid * per;
        int sum = 0;
        for (int k = c0; k < min(nc, c0 + per); ++k) {
            const int a = k / (ncy * ncz), b = (k / ncz) % ncy, c = k % ncz;
            const int bin = (s_list[0][a] * g.nby + s_list[1][b]) * g.nbz + s_list[2][c];
            const int cnt = bin_start[bin + 1] - bin_start[bin];
            s_pre[k] = cnt;  // temporarily the count
            sum += cnt;
        }
        s_part[tid] = sum;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int v = tid >= off ? s_part[tid - off] : 0;
            __syncthreads();
            s_part[tid] += v;
            __syncthreads();
        }
        int run = s_part[tid] - sum;  // exclusive prefix of this thread's first combo
        for (int k = c0; k < min(nc, c0 + per); ++k) {
            const int cnt = s_pre[k];
            s_pre[k] = run;
            run += cnt;
        }
        if (tid == 255) s_pre[nc] = s_part[255];
    }
    __syncthreads();
    const int total = s_pre[nc];

    R acc[G][3][2];
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[j][c][0] = acc[j][c][1] = R(0);

    const int width = g.width, half = g.half;
    for (int c0 = 0; c0 < total; c0 += kChunk) {
        const int cnt = min(kChunk, total - c0);
        __syncthreads();
        for (int q = tid; q < cnt; q += blockDim.x) {
            const int flat = c0 + q;
            int lo = 0, hi = nc;  // largest k with s_pre[k] <= flat
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[mid] <= flat) lo = mid;
                else hi = mid;
            }
            const int k = lo;
            const int a = k / (ncy * ncz), b = (k / ncz) % ncy, c = k % ncz;
            const int bin = (s_list[0][a] * g.nby + s_list[1][b]) * g.nbz + s_list[2][c];
            const int pid = bin_items[bin_start[bin] + (flat - s_pre[k])];
            s_base[q][0] = base[pid];
            s_base[q][1] = base[N + pid];
            s_base[q][2] = base[2 * N + pid];
            for (int kk = 0; kk < 3 * width; ++kk) s_w[q][kk] = wts[(size_t)pid * 3 * width + kk];
            for (int j = 0; j < G; ++j) {
                const int b2 = g0 + j;
                for (int cc = 0; cc < 3; ++cc) {
                    CT v;
                    if (b2 < B) {
                        const double2 dv = tab.x[b2][3 * (size_t)pid + cc];
                        v.x = (R)dv.x;
                        v.y = (R)dv.y;
                    } else {
                        v.x = v.y = R(0);
                    }
                    s_p[q][j][cc] = v;
                }
            }
        }
        __syncthreads();
        if (valid && g.nofold) {
            // every axis is at least as wide as the stencil: one wrap, one weight per axis
            for (int q = 0; q < cnt; ++q) {
                int dx = vx - s_base[q][0];
                dx += (dx < -half) ? g.nx : 0;
                dx -= (dx > half) ? g.nx : 0;
                if (dx < -half || dx > half) continue;
                int dy = vy - s_base[q][1];
                dy += (dy < -half) ? g.ny : 0;
                dy -= (dy > half) ? g.ny : 0;
                if (dy < -half || dy > half) continue;
                int dz = vz - s_base[q][2];
                dz += (dz < -half) ? g.nz : 0;
                dz -= (dz > half) ? g.nz : 0;
                if (dz < -half || dz > half) continue;
                const R w = s_w[q][dx + half] * s_w[q][width + dy + half] * s_w[q][2 * width + dz + half];
#pragma unroll
                for (int j = 0; j < G; ++j)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        acc[j][c][0] += w * s_p[q][j][c].x;
                        acc[j][c][1] += w * s_p[q][j][c].y;
                    }
            }
        } else if (valid) {
            const int vv[3] = {vx, vy, vz};
            const int nn[3] = {g.nx, g.ny, g.nz};
            for (int q = 0; q < cnt; ++q) {
                // folded per-axis weight: sum over stencil offsets o == (v - base) mod n
                R w3[3];
                bool hit = true;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    int dd = vv[d] - s_base[q][d];
                    if (dd < 0) dd += nn[d];
                    int o = dd - nn[d] * ((dd + half) / nn[d]);
                    R ws = R(0);
                    bool any = false;
                    for (; o <= half; o += nn[d]) {
                        ws += s_w[q][d * width + o + half];
                        any = true;
                    }
                    w3[d] = ws;
                    hit = hit && any;
                }
                if (!hit) continue;
                const R w = w3[0] * w3[1] * w3[2];
#pragma unroll
                for (int j = 0; j < G; ++j)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        acc[j][c][0] += w * s_p[q][j][c].x;
                        acc[j][c][1] += w * s_p[q][j][c].y;
                    }
            }
        }
    }
    if (!valid) return;
    const size_t nxy = (size_t)g.nx * g.ny;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const int b = g0 + j;
        if (b >= B) break;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            CT o;
            o.x = acc[j][c][0];
            o.y = acc[j][c][1];
            size_t idx;
            if (SLAB)
                idx = (((size_t)b * 3 + c) * g.Zs + js) * g.nx * g.nyp + (size_t)vx * g.nyp + vy;
            else
                idx = ((size_t)b * 3 + c) * nxy * g.nz + ((size_t)vx * g.ny + vy) * g.nz + vz;
            out[idx] = o;
        }
    }
}

// Full grid: k-space projection f <- sc (f.k^) k^ for every RHS (:1406-1432).
template <typename R>
__global__ __launch_bounds__(256) void scale_kernel(typename C2<R>::T* __restrict__ grid, const R* __restrict__ kc,
                                                    size_t V, int B) {
    using CT = typename C2<R>::T;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
        const R kx = kc[4 * v], ky = kc[4 * v + 1], kz = kc[4 * v + 2], sc = kc[4 * v + 3];
        for (int b = 0; b < B; ++b) {
            CT* gp = grid + (size_t)b * 3 * V + v;
            const CT f0 = gp[0], f1 = gp[V], f2 = gp[2 * V];
            const R dr = (f0.x * kx + f1.x * ky + f2.x * kz) * sc;
            const R di = (f0.y * kx + f1.y * ky + f2.y * kz) * sc;
            CT o;
            o.x = dr * kx;
            o.y = di * kx;
            gp[0] = o;
            o.x = dr * ky;
            o.y = di * ky;
            gp[V] = o;
            o.x = dr * kz;
            o.y = di * kz;
            gp[2 * V] = o;
        }
    }
}

// ---- spread on the matrix cores ----
//
// For a region of RX x (planes) x 16 rows (slab: RX x 16 y over all slab planes;
// full grid: 4 x 8 y-planes x 16 z, see RegionGeom) and 16 real columns
// q = 6 b + 2 c + (0: real, 1: imaginary) of the right-hand sides,
//   grid^T[q][voxel] = sum_k V[k][q] W[k][voxel],
//   W[k][(x, plane, row)] = wx_k(x) wp_k(plane) wr_k(row)
// over the candidate particles k of the region (every particle whose stencil
// footprint meets it, in ascending id order), one 16x16x4 MFMA (FP32 or FP64)
// per four candidates and (x, plane) pair: A = V^T (16 columns x 4 candidates),
// B = W^T (4 candidates x 16 consecutive rows).  The region-local separable weights
// (stencil folding on small grids included) are staged per candidate chunk in
// LDS.  Every voxel is written exactly once, without atomics; the sum order is
// fixed (deterministic).
constexpr int kMfmaRY = 16;    // y voxels per region (the MFMA N dimension)
constexpr size_t kMaxFusedLds = 64 * 1024;  // LDS of one xtoeplitz_kernel block
// Radius classes of the polydisperse operator up to which pair tables are built.
constexpr int kPolyTableClasses = 64;
constexpr int kMfmaKC = 64;    // candidates staged per chunk
// Planes (y) per region on the full grid and candidate blocks of 16 per chunk of
// its contraction: 4 x 8 x 16-voxel regions, about 28 candidates each in a
// packing of volume fraction 0.3 with 9-point stencils on a 256^3 grid.
constexpr int kFullRegionPlanes = 8;
constexpr int kFullContractNB = 2;

typedef float mpm_f32x4 __attribute__((ext_vector_type(4)));
typedef double mpm_f64x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; its
// accumulator entry r is D[drow(l / 16, r)][l % 16], with drow = 4 g + r for
// FP32 and g + 4 r for FP64 (measured on gfx950).  rowcol(i) is the real
// column a spread wave feeds into A row i so that accumulator entries (0, 1)
// and (2, 3) of every lane are the real and imaginary parts of one complex
// column: complex column 2 g + h of the wave sits in entries (2 h, 2 h + 1).
template <typename R>
struct Mfma;
template <>
struct Mfma<float> {
    using V4 = mpm_f32x4;
    static __device__ __forceinline__ V4 mac(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int g, int r) { return 4 * g + r; }
    static __device__ __forceinline__ int rowcol(int i) { return i; }
};
template <>
struct Mfma<double> {
    using V4 = mpm_f64x4;
    static __device__ __forceinline__ V4 mac(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int g, int r) { return g + 4 * r; }
    // row i = g + 4 r -> complex column 2 g + r / 2, part r % 2
    static __device__ __forceinline__ int rowcol(int i) { return 2 * (2 * (i & 3) + (i >> 3)) + ((i >> 2) & 1); }
};

// Block shape of spread_mfma_kernel: 4 waves = (CPB / 16 column groups) x
// (plane groups of ZH planes): 32 columns, the planes split over two waves, which
// halves the accumulator registers and lets more blocks overlap their staging.
template <typename R, int RX, int ZSB>
struct SpreadShape {
    static constexpr int CPB = 32;
    static constexpr int NPG = 4 / (CPB / 16);
    static constexpr int ZH = (ZSB + NPG - 1) / NPG;
};

// Region-local separable weights of a candidate, one LDS row per candidate:
// x weights at [0, RX), slab-plane weights at [4, 4 + ZSB), y weights at
// [4 + ZSB4, 4 + ZSB4 + 16) (ZSB4 = ZSB rounded up to 4), so that 4-wide
// groups start on 16-byte boundaries.  A voxel's weight sums the stencil points
// that land on it (stencils fold onto themselves on grids narrower than the
// stencil).
template <int RX, int ZSB>
struct RegionW {
    static constexpr int ZSB4 = (ZSB + 3) & ~3;
    static constexpr int X = 0, Z = 4, Y = 4 + ZSB4;
    static constexpr int NW = Y + kMfmaRY;
};

// Copies the precomputed region-local weight rows of a chunk of cnt <= KC
// candidates into LDS (zero rows past cnt) and ends with a barrier.
// v where c holds, else +0, by a bit mask rather than a select: the compiler turns a
// select of a loaded value into a branch around the load, which then waits for
// the load at the branch's end.
__device__ __forceinline__ float keep_if(float v, bool c) {
    return __uint_as_float(__float_as_uint(v) & (c ? 0xffffffffu : 0u));
}
__device__ __forceinline__ double keep_if(double v, bool c) {
    return __longlong_as_double(__double_as_longlong(v) & (c ? -1LL : 0LL));
}

// Blocks of 256 threads; every load is issued before the first store (one memory
// latency per chunk, not one per element).
template <typename R, int RX, int ZSB, int KC = kMfmaKC>
__device__ inline void stage_weights(R* s_w, int cnt, const R* __restrict__ regw) {
    using L = RegionW<RX, ZSB>;
    constexpr int T = (KC * L::NW + 255) / 256;
    const int n = cnt * L::NW;
    R v[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int e = (int)threadIdx.x + 256 * t;
        v[t] = e < n ? regw[e] : R(0);
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int e = (int)threadIdx.x + 256 * t;
        if (e < KC * L::NW) s_w[e] = v[t];
    }
    __syncthreads();
}

// Plane loop of the matrix-core contractions, software-pipelined over two operand
// buffers: the loads of the next plane are in flight while the current plane's
// products run.  load(p, buf) issues plane p's loads into buf, mma(buf, p)
// consumes them; planes first, first + step, ... below end.  The loop body has no
// branch between a load and its use (a branch would let the compiler sink the
// loads next to their use): the last pair loads its second plane again, and an
// odd plane count ends with one more product.  Scheduling barriers keep each group
// of loads whole and ahead of the other buffer's products, so that the products
// wait only for their own (older) loads.
template <typename Buf, typename LoadF, typename MmaF>
__device__ __forceinline__ void plane_pipeline(int first, int step, int end, Buf& ga, Buf& gb, LoadF&& load,
                                               MmaF&& mma) {
    if (first >= end) return;
    const int n = (end - first + step - 1) / step;
    const int last = first + (n - 1) * step;
    load(first, ga);
    int i = 0, p = first;
    for (; i + 1 < n; i += 2, p += 2 * step) {
        __builtin_amdgcn_sched_barrier(0);
        load(p + step, gb);
        __builtin_amdgcn_sched_barrier(0);
        mma(ga, p);
        __builtin_amdgcn_sched_barrier(0);
        load(min(p + 2 * step, last), ga);
        __builtin_amdgcn_sched_barrier(0);
        mma(gb, p + step);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (i < n) mma(ga, p);
}

// Vector form of the contractions' grid operands: the four voxels a lane feeds into
// the four products of a plane and x are rows 4 r + u (r = lane / 16, u = 0..3) of
// its real column (even lane) or imaginary column (odd lane); the two lanes of a
// complex column load rows (4 r, 4 r + 1) and (4 r + 2, 4 r + 3) as whole complex
// pairs v = (re, im, re, im) and swap halves with a neighbour-lane move, so that one
// 16-byte load per lane reads the 16 rows of eight complex components (one whole
// 128-byte line each for FP32).  g receives rows 4 r + u of this lane's part.
template <typename R, typename V4>
__device__ __forceinline__ void pair_rows(const V4& v, int part, R (&g)[4]) {
    const R s0 = part ? v[0] : v[1], s1 = part ? v[2] : v[3];
    const R t0 = __shfl_xor(s0, 1, 64), t1 = __shfl_xor(s1, 1, 64);
    g[0] = part ? t0 : v[0];
    g[1] = part ? t1 : v[2];
    g[2] = part ? v[1] : t0;
    g[3] = part ? v[3] : t1;
}

// Where the regions sit in a grid buffer.  A region is RX consecutive x, up to
// pblk consecutive planes and 16 consecutive rows, rows contiguous in memory:
// in the slab layout the planes are the occupied z-planes (all in one region)
// and the rows y; on the full grid the planes are y (blocks of pblk) and the
// rows z.  Regions are numbered (x block, plane block, row block), row block
// fastest.  Strides are in complex elements.
struct RegionGeom {
    int nx, nrow, nplane;  // extents along x, rows, planes
    int nrr, npb, pblk;    // row blocks, plane blocks, planes per region
    unsigned xs, ps, cs;   // strides of x, plane, component (b, c)
};

struct RegionAt {
    int x0, plane0, row0, np;  // origin and number of planes present
};

template <int RX>
__device__ inline RegionAt region_at(const RegionGeom& rg, int reg) {
    const int per_x = rg.npb * rg.nrr;
    const int rx = reg / per_x, rem = reg - rx * per_x;
    const int pb = rem / rg.nrr, rr = rem - pb * rg.nrr;
    RegionAt a;
    a.x0 = rx * RX;
    a.plane0 = pb * rg.pblk;
    a.row0 = rr * kMfmaRY;
    a.np = min(rg.pblk, rg.nplane - a.plane0);
    return a;
}

// FP32: held to 128 registers (four waves per SIMD; at the compiler's 140 three, and the
// N = 4000 monolayer's 1272 blocks took 1.7 rounds instead of 1.2: 39.4 against 37.0 us)
template <typename R, int RX, int ZSB>
__global__ __launch_bounds__(256, sizeof(R) == 4 ? 4 : 2) void spread_mfma_kernel(SlotTable tab, int B, RegionGeom rg,
                                                          const R* __restrict__ regw,
                                                          const int* __restrict__ reg_start,
                                                          const int* __restrict__ reg_items, int N,
                                                          typename C2<R>::T* __restrict__ out) {
    using CT = typename C2<R>::T;
    using M = Mfma<R>;
    using Sh = SpreadShape<R, RX, ZSB>;
    constexpr int KC = kMfmaKC, CPB = Sh::CPB, ZH = Sh::ZH;
    using L = RegionW<RX, ZSB>;
    __shared__ R s_v[KC][CPB + 1];
    __shared__ __align__(16) R s_w[KC][L::NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cg = wave % (CPB / 16), pg = wave / (CPB / 16);
    const int reg = blockIdx.x;
    const RegionAt ra = region_at<RX>(rg, reg);
    const int x0 = ra.x0, y0 = ra.row0;
    const int col_base = blockIdx.y * CPB;
    const int wcol0 = col_base + cg * 16;
    const int ncols = 6 * B;
    const int Zs = ra.np;    // planes of this region
    const int p0 = pg * ZH;  // this wave's planes [p0, p0 + ZH) of the region
    const int k0 = reg_start[reg], k1 = reg_start[reg + 1];

    typename M::V4 acc[RX][ZH];
#pragma unroll
    for (int xi = 0; xi < RX; ++xi)
#pragma unroll
        for (int p = 0; p < ZH; ++p) acc[xi][p] = typename M::V4{R(0), R(0), R(0), R(0)};

    // staged columns of this thread: s_v[q][scol] for q = tid / CPB + (256 / CPB) t
    constexpr int QS = 256 / CPB;
    const int scol = tid % CPB, sqc = col_base + scol;
    // columns past the batch read column 0 and stage zeros
    const bool scok = sqc < ncols;
    const int sq = scok ? sqc : 0;
    const double* xcol = reinterpret_cast<const double*>(tab.x[sq / 6]) + 2 * ((sq >> 1) % 3) + (sq & 1);
    for (int c0 = k0; c0 < k1; c0 += KC) {
        const int cnt = min(KC, k1 - c0);
        __syncthreads();
        // values: s_v[k][col] = component (b, c), real or imaginary part, of candidate k;
        // all candidate ids first, then all values (two memory latencies per chunk)
        int id[KC / QS];
#pragma unroll
        for (int t = 0; t < KC / QS; ++t) id[t] = reg_items[c0 + min(tid / CPB + QS * t, cnt - 1)];
        R vv[KC / QS];
#pragma unroll
        for (int t = 0; t < KC / QS; ++t)
            vv[t] = keep_if((R)xcol[6 * (size_t)id[t]], scok && tid / CPB + QS * t < cnt);
        stage_weights<R, RX, ZSB>(&s_w[0][0], cnt, regw + (size_t)c0 * L::NW);
#pragma unroll
        for (int t = 0; t < KC / QS; ++t) s_v[tid / CPB + QS * t][scol] = vv[t];
        __syncthreads();
        if (wcol0 < ncols && p0 < Zs) {
            const int i = lane & 15;
            for (int ks = 0; ks < cnt; ks += 4) {
                const int kq = ks + (lane >> 4);  // rows >= cnt hold zeros
                const R a = s_v[kq][cg * 16 + M::rowcol(i)];
                const R wy = s_w[kq][L::Y + i];
                R wx[RX], wz[ZH];
#pragma unroll
                for (int xi = 0; xi < RX; ++xi) wx[xi] = s_w[kq][L::X + xi] * wy;
#pragma unroll
                for (int p = 0; p < ZH; ++p) wz[p] = (p0 + p < ZSB) ? s_w[kq][L::Z + p0 + p] : R(0);
                // planes past Zs carry zero weights: no branch around the accumulators
#pragma unroll
                for (int p = 0; p < ZH; ++p) {
#pragma unroll
                    for (int xi = 0; xi < RX; ++xi) acc[xi][p] = M::mac(a, wx[xi] * wz[p], acc[xi][p]);
                }
            }
        }
    }
    // lane holds complex columns 2 (lane >> 4) + h of the wave (entries 2 h, 2 h + 1)
    // at row y0 + (lane & 15)
    const int y = y0 + (lane & 15);
    if (y >= rg.nrow || wcol0 >= ncols) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int qc = wcol0 + 4 * (lane >> 4) + 2 * h;
        if (qc >= ncols) continue;
        const int b = qc / 6, c = (qc >> 1) % 3;
        CT* o = out + ((size_t)b * 3 + c) * rg.cs + (size_t)ra.plane0 * rg.ps + y;
#pragma unroll
        for (int xi = 0; xi < RX; ++xi) {
            const int x = x0 + xi;
            if (x >= rg.nx) continue;
#pragma unroll
            for (int p = 0; p < ZH; ++p)
                if (p0 + p < Zs) {
                    CT v;
                    v.x = acc[xi][p][2 * h];
                    v.y = acc[xi][p][2 * h + 1];
                    o[(size_t)(p0 + p) * rg.ps + (size_t)x * rg.xs] = v;
                }
        }
    }
}

// ---- contraction on the matrix cores ----
//
// The transpose of spread_mfma_kernel over the same regions: for the candidate
// particles k of a region and 16 real columns q,
//   D[k][q] = sum_{voxels v of the region} W[k][v] G[v][q],
// one 16x16x4 MFMA per four voxels and 16 candidates (A = W, B = G), every
// loaded operand feeding all candidate blocks.  Each wave reads its B operands
// straight from the grid: lane (j, u) loads column j at row 4 u + (lane >> 4),
// so the four loads of a 16-voxel row consume whole cache lines, or with VEC
// (pair_rows) one 16-byte complex pair per lane and x; waves run without
// barriers after the one weight-staging step.  D goes to the particle's
// slot of the region (the CSR position of the pair); the per-particle sum over
// its regions happens in the finish step, in a fixed order.
template <typename R, int RX, int ZSB, int NB, bool VEC>
__global__ __launch_bounds__(256, sizeof(R) == 4 ? 3 : 2) void contract_mfma_kernel(int B, RegionGeom rg,
                                                                                    const R* __restrict__ regw,
                                                            const int* __restrict__ reg_start,
                                                            const int* __restrict__ reg_items, int N,
                                                            const typename C2<R>::T* __restrict__ grid,
                                                            R* __restrict__ partial, int ncg) {
    using M = Mfma<R>;
    using V4 = typename M::V4;
    constexpr int KC = 16 * NB;  // candidates per chunk, in NB blocks of 16
    using L = RegionW<RX, ZSB>;
    __shared__ __align__(16) R s_w[KC][L::NW];
    __shared__ V4 s_red[4][NB][64];  // per-wave partial accumulators
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // 4 waves = ncg column groups x (4 / ncg) plane groups
    const int npg = 4 / ncg, cg = wave % ncg, pg = wave / ncg;
    const int reg = blockIdx.x;
    const RegionAt ra = region_at<RX>(rg, reg);
    const int x0 = ra.x0, y0 = ra.row0;
    const int wcol0 = (blockIdx.y * ncg + cg) * 16;
    const int ncols = 6 * B;
    const bool active = wcol0 < ncols;
    const int Zs = ra.np;  // planes of this region
    const unsigned plane = rg.ps;
    const int k0 = reg_start[reg], k1 = reg_start[reg + 1];

    // this lane's B-operand column and its rows of the region, as 32-bit offsets
    // (in R) from the uniform grid base: the launch keeps 6 B x (complex per
    // component) below 2^32
    const R* gf = reinterpret_cast<const R*>(grid);
    const int qc = wcol0 + (lane & 15);
    const bool cvalid = qc < ncols;
    // column offset at the region origin (cbase, an in-grid voxel of every plane of the
    // region) and at this lane's row; bit 4 xi + u of vmask: (x0 + xi, y0 + 4 u + lane / 16)
    // inside the grid.  VEC (pair_rows): the lane's complex pair at rows y0 + 4 r + 2 part
    // and the next, bit xi of vmask: (x0 + xi, y0 + 4 r + 2 part) inside; even rows keep
    // both rows of a pair inside together.
    const int r = lane >> 4, part = lane & 1;
    unsigned cbase = 0, coff = 0, vmask = 0;
    if (cvalid) {
        const int b = qc / 6, c = (qc >> 1) % 3;
        cbase = 2u * (unsigned)(b * 3 + c) * rg.cs + (VEC ? 0u : (unsigned)(qc & 1)) +
                2u * ((unsigned)ra.plane0 * rg.ps + (unsigned)x0 * rg.xs + (unsigned)y0);
        if (VEC) {
            coff = cbase + 2u * (unsigned)(4 * r + 2 * part);
#pragma unroll
            for (int xi = 0; xi < RX; ++xi)
                if (x0 + xi < rg.nx && y0 + 4 * r + 2 * part < rg.nrow) vmask |= 1u << xi;
        } else {
            coff = cbase + 2u * (unsigned)(lane >> 4);
#pragma unroll
            for (int xi = 0; xi < RX; ++xi)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (x0 + xi < rg.nx && y0 + 4 * u + (lane >> 4) < rg.nrow) vmask |= 1u << (4 * xi + u);
        }
    }
    const unsigned xstep = 2u * rg.xs;

    for (int c0 = k0; c0 < k1; c0 += KC) {
        const int cnt = min(KC, k1 - c0);
        __syncthreads();
        stage_weights<R, RX, ZSB, KC>(&s_w[0][0], cnt, regw + (size_t)c0 * L::NW);
        V4 acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = V4{R(0), R(0), R(0), R(0)};
        if (active) {
            // per-lane weights: candidate k = 16 nb + (lane & 15) at x = x0 + xi and rows
            // y0 + 4 u + r (VEC: y0 + 4 r + u)
            R wyl[NB][4], wxl[NB][RX];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int k = nb * 16 + (lane & 15);
#pragma unroll
                for (int u = 0; u < 4; ++u) wyl[nb][u] = s_w[k][L::Y + (VEC ? 4 * r + u : 4 * u + r)];
#pragma unroll
                for (int xi = 0; xi < RX; ++xi) wxl[nb][xi] = s_w[k][L::X + xi];
            }
            if constexpr (VEC) {
                struct PlaneBuf {
                    V4 g[RX];
                };
                auto load_plane = [&](int p, PlaneBuf& dst) {
                    const unsigned poff = coff + 2u * (unsigned)p * plane, pin = cbase + 2u * (unsigned)p * plane;
#pragma unroll
                    for (int xi = 0; xi < RX; ++xi) {
                        const bool in = (vmask >> xi) & 1u;
                        dst.g[xi] = *reinterpret_cast<const V4*>(gf + (in ? poff + xi * xstep : pin));
                    }
                };
                auto mma = [&](PlaneBuf& pb, int p) {
                    R wz[NB];
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) wz[nb] = s_w[nb * 16 + (lane & 15)][L::Z + p];
#pragma unroll
                    for (int xi = 0; xi < RX; ++xi) {
                        R gv[4];
                        pair_rows<R>(pb.g[xi], part, gv);
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb) {
                            const R wxz = wxl[nb][xi] * wz[nb];
#pragma unroll
                            for (int u = 0; u < 4; ++u) acc[nb] = M::mac(wyl[nb][u] * wxz, gv[u], acc[nb]);
                        }
                    }
                };
                PlaneBuf ga, gb;
                plane_pipeline(pg, npg, Zs, ga, gb, load_plane, mma);
            } else {
            // the next plane's operands are loaded while this plane's products run
            auto load_plane = [&](int p, R (&dst)[RX][4]) {
                const unsigned poff = coff + 2u * (unsigned)p * plane, pin = cbase + 2u * (unsigned)p * plane;
#pragma unroll
                for (int xi = 0; xi < RX; ++xi)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        // lanes outside the grid read the region's origin instead, which
                        // meets a zero weight (the weight rows are zero past the grid's
                        // edges): no branch around the load and no use of it before the
                        // products
                        const bool in = (vmask >> (4 * xi + u)) & 1u;
                        dst[xi][u] = gf[in ? poff + xi * xstep + 8u * u : pin];
                    }
            };
            auto mma = [&](R (&gv)[RX][4], int p) {
                R wz[NB];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) wz[nb] = s_w[nb * 16 + (lane & 15)][L::Z + p];
                // every candidate block, filled or not (rows past cnt hold zero weights):
                // a branch here would make the compiler copy the accumulators around it
#pragma unroll
                for (int xi = 0; xi < RX; ++xi)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) {
                        const R wxz = wxl[nb][xi] * wz[nb];
#pragma unroll
                        for (int u = 0; u < 4; ++u) acc[nb] = M::mac(wyl[nb][u] * wxz, gv[xi][u], acc[nb]);
                    }
            };
            R ga[RX][4], gb[RX][4];
            plane_pipeline(pg, npg, Zs, ga, gb, load_plane, mma);
            }
        }
        // plane groups of one column group are summed in plane-group order
        if (npg > 1) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) s_red[wave][nb][lane] = acc[nb];
            __syncthreads();
            if (pg == 0) {
                for (int q = 1; q < npg; ++q)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb) acc[nb] += s_red[q * ncg + cg][nb][lane];
            }
        }
        // lane holds D[16 nb + drow(lane >> 4, r)][lane & 15]
        if (pg == 0 && cvalid) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = nb * 16 + M::drow(lane >> 4, q);
                    if (k < cnt) partial[(size_t)(c0 + k) * ncols + qc] = acc[nb][q];
                }
        }
    }
}

// Narrow-batch variant of the contraction: a block is one column group of 16
// and four waves, one per block of 16 candidates, so a wave holds few registers
// and, when only one or two column groups exist, the region's candidates still
// spread over four waves; the next plane's operands are loaded while the
// current plane's products run.
template <typename R, int RX, int ZSB, bool VEC>
__global__ __launch_bounds__(256) void contract_nb_mfma_kernel(int B, RegionGeom rg, const R* __restrict__ regw,
                                                               const int* __restrict__ reg_start,
                                                               const int* __restrict__ reg_items, int N,
                                                               const typename C2<R>::T* __restrict__ grid,
                                                               R* __restrict__ partial) {
    using M = Mfma<R>;
    using V4 = typename M::V4;
    constexpr int KC = kMfmaKC;
    using L = RegionW<RX, ZSB>;
    __shared__ __align__(16) R s_w[KC][L::NW];
    const int tid = threadIdx.x, lane = tid & 63, nb = tid >> 6;  // candidate block of this wave
    const int reg = blockIdx.x;
    const RegionAt ra = region_at<RX>(rg, reg);
    const int x0 = ra.x0, y0 = ra.row0;
    const int wcol0 = blockIdx.y * 16;
    const int ncols = 6 * B;
    const int Zs = ra.np;  // planes of this region
    const unsigned plane = rg.ps;
    const int k0 = reg_start[reg], k1 = reg_start[reg + 1];
    const R* gf = reinterpret_cast<const R*>(grid);
    const int qc = wcol0 + (lane & 15);
    const bool cvalid = qc < ncols;
    // column offset at the region origin (cbase, an in-grid voxel of every plane of the
    // region) and at this lane's row; bit 4 xi + u of vmask: (x0 + xi, y0 + 4 u + lane / 16)
    // inside the grid.  VEC (pair_rows): the lane's complex pair at rows y0 + 4 r + 2 part
    // and the next, bit xi of vmask: (x0 + xi, y0 + 4 r + 2 part) inside.
    const int r = lane >> 4, part = lane & 1;
    unsigned cbase = 0, coff = 0, vmask = 0;
    if (cvalid) {
        const int b = qc / 6, c = (qc >> 1) % 3;
        cbase = 2u * (unsigned)(b * 3 + c) * rg.cs + (VEC ? 0u : (unsigned)(qc & 1)) +
                2u * ((unsigned)ra.plane0 * rg.ps + (unsigned)x0 * rg.xs + (unsigned)y0);
        if (VEC) {
            coff = cbase + 2u * (unsigned)(4 * r + 2 * part);
#pragma unroll
            for (int xi = 0; xi < RX; ++xi)
                if (x0 + xi < rg.nx && y0 + 4 * r + 2 * part < rg.nrow) vmask |= 1u << xi;
        } else {
            coff = cbase + 2u * (unsigned)(lane >> 4);
#pragma unroll
            for (int xi = 0; xi < RX; ++xi)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (x0 + xi < rg.nx && y0 + 4 * u + (lane >> 4) < rg.nrow) vmask |= 1u << (4 * xi + u);
        }
    }
    const unsigned xstep = 2u * rg.xs;
    for (int c0 = k0; c0 < k1; c0 += KC) {
        const int cnt = min(KC, k1 - c0);
        __syncthreads();
        stage_weights<R, RX, ZSB>(&s_w[0][0], cnt, regw + (size_t)c0 * L::NW);
        if (nb * 16 >= cnt) continue;  // no candidates for this wave (no barrier follows in this chunk)
        const int k = nb * 16 + (lane & 15);
        R wyl[4], wxl[RX];
#pragma unroll
        for (int u = 0; u < 4; ++u) wyl[u] = s_w[k][L::Y + (VEC ? 4 * r + u : 4 * u + r)];
#pragma unroll
        for (int xi = 0; xi < RX; ++xi) wxl[xi] = s_w[k][L::X + xi];
        V4 acc = V4{R(0), R(0), R(0), R(0)};
        if constexpr (VEC) {
            struct PlaneBuf {
                V4 g[RX];
            };
            auto load_plane = [&](int p, PlaneBuf& dst) {
                const unsigned poff = coff + 2u * (unsigned)p * plane, pin = cbase + 2u * (unsigned)p * plane;
#pragma unroll
                for (int xi = 0; xi < RX; ++xi) {
                    const bool in = (vmask >> xi) & 1u;
                    dst.g[xi] = *reinterpret_cast<const V4*>(gf + (in ? poff + xi * xstep : pin));
                }
            };
            auto mma = [&](PlaneBuf& pb, int p) {
                const R wz = s_w[k][L::Z + p];
#pragma unroll
                for (int xi = 0; xi < RX; ++xi) {
                    R gv[4];
                    pair_rows<R>(pb.g[xi], part, gv);
                    const R wxz = wxl[xi] * wz;
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc = M::mac(wyl[u] * wxz, gv[u], acc);
                }
            };
            PlaneBuf ga, gb;
            plane_pipeline(0, 1, Zs, ga, gb, load_plane, mma);
        } else {
            auto load_plane = [&](int p, R (&dst)[RX][4]) {
                const unsigned poff = coff + 2u * (unsigned)p * plane, pin = cbase + 2u * (unsigned)p * plane;
#pragma unroll
                for (int xi = 0; xi < RX; ++xi)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        // lanes outside the grid read the region's origin, which meets a zero weight
                        const bool in = (vmask >> (4 * xi + u)) & 1u;
                        dst[xi][u] = gf[in ? poff + xi * xstep + 8u * u : pin];
                    }
            };
            auto mma = [&](R (&gv)[RX][4], int p) {
                const R wz = s_w[k][L::Z + p];
#pragma unroll
                for (int xi = 0; xi < RX; ++xi) {
                    const R wxz = wxl[xi] * wz;
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc = M::mac(wyl[u] * wxz, gv[xi][u], acc);
                }
            };
            R ga[RX][4], gb[RX][4];
            plane_pipeline(0, 1, Zs, ga, gb, load_plane, mma);
        }
        // lane holds D[16 nb + drow(lane >> 4, q)][lane & 15]
        if (cvalid) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kk = nb * 16 + M::drow(lane >> 4, q);
                if (kk < cnt) partial[(size_t)(c0 + kk) * ncols + qc] = acc[q];
            }
        }
    }
}

// Full-grid contraction on the matrix cores: the products of contract_mfma_kernel
// over RX x PB x 16-voxel regions (x, y planes, z rows), one wave per (region,
// column group of 16), work item = region x column groups + group, so that the
// groups of a region run in neighbouring waves and share its weights in the caches.
// Waves are independent (no LDS, no barriers).
//
// Memory access is in whole 16-byte vectors: the four voxels a lane feeds into the
// four products of a plane and x are rows 4 r + u (r = lane / 16, u = 0..3), and the
// two lanes of a complex column pair (real part j even, imaginary part j + 1) load
// rows (4 r, 4 r + 1) and (4 r + 2, 4 r + 3) as complex pairs, then swap halves with
// a neighbour-lane move; so one load instruction per x and plane reads the 16 rows
// of eight complex components, one whole 128-byte line each.  The candidates'
// weights come the same way, a vector each for x, the planes and the lane's rows.
// Needs an even number of rows (z) and 16-byte aligned grid and weight rows.
// The next plane's operands are loaded while the current plane's products run.
template <typename R, int RX, int PB, int NB>
__global__ __launch_bounds__(256) void contract_full_mfma_kernel(int B, RegionGeom rg, const R* __restrict__ regw,
                                                                 const int* __restrict__ reg_start, int nitems,
                                                                 int ngrp, const typename C2<R>::T* __restrict__ grid,
                                                                 R* __restrict__ partial) {
    using M = Mfma<R>;
    using V4 = typename M::V4;  // four R: a complex pair of the grid, or four weights
    using L = RegionW<RX, PB>;
    static_assert(RX == 4 && PB == 8, "one weight vector for x, two for the planes");
    static_assert(L::Z == 4 && L::Y == 12 && L::NW == 28, "weight row: x at 0, planes at 4, rows at 12");
    constexpr int KC = 16 * NB;
    const int lane = threadIdx.x & 63;
    const int item = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (item >= nitems) return;  // the whole wave leaves: nothing below synchronises
    const int reg = item / ngrp;
    const RegionAt ra = region_at<RX>(rg, reg);
    const int x0 = ra.x0, y0 = ra.row0, np = ra.np;
    const int ncols = 6 * B;
    const int r = lane >> 4, part = lane & 1;
    const int qc = (item - reg * ngrp) * 16 + (lane & 15);
    const bool cvalid = qc < ncols;
    const R* gf = reinterpret_cast<const R*>(grid);
    // 32-bit offsets in R from the uniform grid base (the launch keeps them below 2^32):
    // cbase = the component's voxel (x0, plane0, y0), an in-grid pair of rows of every
    // plane; this lane loads rows y0 + 4 r + 2 part and the next (vmask bit xi: inside)
    unsigned cbase = 2u * ((unsigned)ra.plane0 * rg.ps + (unsigned)x0 * rg.xs + (unsigned)y0), coff = cbase,
             vmask = 0;
    if (cvalid) {
        const int b = qc / 6, c = (qc >> 1) % 3;
        cbase += 2u * (unsigned)(b * 3 + c) * rg.cs;
        coff = cbase + 2u * (unsigned)(4 * r + 2 * part);
#pragma unroll
        for (int xi = 0; xi < RX; ++xi)
            if (x0 + xi < rg.nx && y0 + 4 * r + 2 * part < rg.nrow) vmask |= 1u << xi;
    }
    const unsigned xstep = 2u * rg.xs, pstep = 2u * rg.ps;
    const int k0 = reg_start[reg], k1 = reg_start[reg + 1];
    struct PlaneBuf {
        V4 g[RX];
    };
    for (int c0 = k0; c0 < k1; c0 += KC) {
        const int cnt = min(KC, k1 - c0);
        // this lane's candidates k = 16 nb + (lane & 15): x weights, plane weights (zero
        // past the region's planes and for candidates past cnt) and rows 4 r .. 4 r + 3
        R wyl[NB][4], wxl[NB][RX], wzl[NB][PB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int k = nb * 16 + (lane & 15);
            const bool kin = k < cnt;
            const V4* wr = reinterpret_cast<const V4*>(regw + (size_t)(c0 + (kin ? k : 0)) * L::NW);
            const V4 wx = wr[0], wz0 = wr[1], wz1 = wr[2], wy = wr[3 + r];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                wxl[nb][q] = wx[q];
                wzl[nb][q] = keep_if(wz0[q], kin && q < np);
                wzl[nb][4 + q] = keep_if(wz1[q], kin && 4 + q < np);
                wyl[nb][q] = wy[q];
            }
        }
        // an opaque zero, redefined per chunk: keeps the compiler from hoisting all eight
        // planes' loads out of the chunk loop (128 registers, no overlap with the products)
        unsigned salt = 0;
        asm volatile("" : "+v"(salt));
        auto load_plane = [&](int p, PlaneBuf& dst) {
            const int pc = min(p, np - 1);  // planes past the region's reload its last one
            const unsigned poff = coff + salt + (unsigned)pc * pstep, pin = cbase + salt + (unsigned)pc * pstep;
#pragma unroll
            for (int xi = 0; xi < RX; ++xi) {
                const bool in = (vmask >> xi) & 1u;
                dst.g[xi] = *reinterpret_cast<const V4*>(gf + (in ? poff + xi * xstep : pin));
            }
        };
        V4 acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = V4{R(0), R(0), R(0), R(0)};
        auto mma = [&](PlaneBuf& pb, int p) {
#pragma unroll
            for (int xi = 0; xi < RX; ++xi) {
                R gv[4];
                pair_rows<R>(pb.g[xi], part, gv);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    const R wxz = wxl[nb][xi] * wzl[nb][p];
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc[nb] = M::mac(wyl[nb][u] * wxz, gv[u], acc[nb]);
                }
            }
        };
        PlaneBuf ga, gb;
        load_plane(0, ga);
#pragma unroll
        for (int p = 0; p < PB; p += 2) {
            __builtin_amdgcn_sched_barrier(0);
            load_plane(p + 1, gb);
            __builtin_amdgcn_sched_barrier(0);
            mma(ga, p);
            __builtin_amdgcn_sched_barrier(0);
            if (p + 2 < PB) load_plane(p + 2, ga);
            __builtin_amdgcn_sched_barrier(0);
            mma(gb, p + 1);
        }
        // lane holds D[16 nb + drow(lane >> 4, r)][lane & 15]
        if (cvalid) {
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int k = nb * 16 + M::drow(lane >> 4, q);
                    if (k < cnt) partial[(size_t)(c0 + k) * ncols + qc] = acc[nb][q];
                }
        }
    }
}

// Slab: for every in-plane wave vector (kx, ky) — one column of the slab after
// the 2-D transforms — evaluate the z-transform from the Zs occupied planes to
// all nz wave numbers, apply the Green's tensor projection, and transform back to
// the Zs planes the contraction reads.  The 3-D k-space grid never exists in
// memory.  A block covers (256 / RB) columns x RB right-hand sides; the twiddles
// exp(-2 pi i kz z_j / nz) sit in LDS, the Green's tensor (k^, sc) is read from a
// [kz][column] table shared by the RB right-hand sides of a column.
template <typename R, int ZS, int RB>
__global__ __launch_bounds__(256) void zfused_kernel(typename C2<R>::T* __restrict__ slab, int B, int ncols, int nz,
                                                     int Zs, const int* __restrict__ zlist,
                                                     const R* __restrict__ kc) {
    using CT = typename C2<R>::T;
    extern __shared__ unsigned char s_raw[];
    CT* s_tw = reinterpret_cast<CT*>(s_raw);  // [ZS][nz]
    constexpr int CB = 256 / RB;
    const int tid = threadIdx.x;
    const int tc = tid % CB, tb = tid / CB;
    for (int e = tid; e < ZS * nz; e += 256) {
        const int j = e / nz, kz = e % nz;
        CT w;
        if (j < Zs) {
            const long long m = ((long long)kz * zlist[j]) % nz;
            double sv, cv;
            sincospi(2.0 * (double)m / (double)nz, &sv, &cv);
            w.x = (R)cv;
            w.y = (R)(-sv);
        } else {
            w.x = w.y = R(0);
        }
        s_tw[e] = w;
    }
    __syncthreads();
    const int col = blockIdx.x * CB + tc;
    const int b = blockIdx.y * RB + tb;
    if (col >= ncols || b >= B) return;
    CT f[3][ZS], gacc[3][ZS];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < ZS; ++j) {
            if (j < Zs) {
                f[c][j] = slab[(((size_t)b * 3 + c) * Zs + j) * ncols + col];
            } else {
                f[c][j].x = f[c][j].y = R(0);
            }
            gacc[c][j].x = gacc[c][j].y = R(0);
        }
    for (int kz = 0; kz < nz; ++kz) {
        const R* k4 = kc + ((size_t)kz * ncols + col) * 4;
        const R kx = k4[0], ky = k4[1], kzh = k4[2], sc = k4[3];
        R Fr[3] = {R(0), R(0), R(0)}, Fi[3] = {R(0), R(0), R(0)};
#pragma unroll
        for (int j = 0; j < ZS; ++j) {
            const CT w = s_tw[j * nz + kz];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Fr[c] += f[c][j].x * w.x - f[c][j].y * w.y;
                Fi[c] += f[c][j].x * w.y + f[c][j].y * w.x;
            }
        }
        const R pr = sc * (kx * Fr[0] + ky * Fr[1] + kzh * Fr[2]);
        const R pi = sc * (kx * Fi[0] + ky * Fi[1] + kzh * Fi[2]);
        const R kk[3] = {kx, ky, kzh};
#pragma unroll
        for (int j = 0; j < ZS; ++j) {
            const CT w = s_tw[j * nz + kz];  // inverse uses conj(w)
            const R ur = pr * w.x + pi * w.y;
            const R ui = pi * w.x - pr * w.y;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gacc[c][j].x += kk[c] * ur;
                gacc[c][j].y += kk[c] * ui;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < ZS; ++j)
            if (j < Zs) slab[(((size_t)b * 3 + c) * Zs + j) * ncols + col] = gacc[c][j];
}

// Green's tensor table for the slab path, [kz][column][4].
__global__ void kspace_slab_kernel(double* __restrict__ k64, float* __restrict__ k32, KParams q) {
    const size_t ncols = (size_t)kcols(q);
    const size_t V = ncols * q.nz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < V; e += (size_t)gridDim.x * blockDim.x) {
        const int kz = (int)(e / ncols);
        const int col = (int)(e % ncols);
        int ix, iy;
        col_xy(q, col, ix, iy);
        double o[4];
        kcoef(q, ix, iy, kz, o);
        for (int c = 0; c < 4; ++c) {
            k64[4 * e + c] = o[c];
            k32[4 * e + c] = (float)o[c];
        }
    }
}

// Slab with cyclically contiguous planes z_a = z_0 + a (mod nz): the z-transform,
// Green's tensor and inverse z-transform of one (kx, ky) column collapse into a
// Toeplitz coupling of the planes,
//   g[a] = sum_b T_{a-b} f[b],  T_d = sum_kz exp(+2 pi i kz d / nz) sc k^ k^T,
// a complex symmetric 3x3 per (d, column) with T_{-d} = conj(T_d) (the projector
// sc k^ k^T is real).  Table layout [d][6][column], d = 0..Zs-1, components
// (xx, yy, zz, xy, xz, yz), built in FP64 once per grid / Bloch vector / slab size.
// t32ri (odd nz only, else null): the same table in FP32 as six floats per (d, column),
// [Zs][ncols][6] = (Re xx, Re yy, Re zz, Re xy, Im xz, Im yz), the only nonzero parts then.
__global__ void ztoeplitz_table_kernel(double2* __restrict__ t64, float2* __restrict__ t32, float* __restrict__ t32ri,
                                       KParams q, int Zs) {
    const int ncols = kcols(q);
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int d = blockIdx.y;
    if (col >= ncols || d >= Zs) return;
    double tr[6] = {0, 0, 0, 0, 0, 0}, ti[6] = {0, 0, 0, 0, 0, 0};
    int ix, iy;
    col_xy(q, col, ix, iy);
    for (int kz = 0; kz < q.nz; ++kz) {
        double o[4];
        kcoef(q, ix, iy, kz, o);
        if (o[3] == 0.0) continue;
        double sv, cv;
        sincospi(2.0 * (double)(((long long)kz * d) % q.nz) / (double)q.nz, &sv, &cv);
        const double m[6] = {o[0] * o[0], o[1] * o[1], o[2] * o[2], o[0] * o[1], o[0] * o[2], o[1] * o[2]};
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            tr[c] += o[3] * m[c] * cv;
            ti[c] += o[3] * m[c] * sv;
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const size_t e = ((size_t)d * 6 + c) * ncols + col;
        t64[e] = make_double2(tr[c], ti[c]);
        t32[e] = make_float2((float)tr[c], (float)ti[c]);
    }
    if (t32ri) {
        float* o = t32ri + ((size_t)d * ncols + col) * 6;
        o[0] = (float)tr[0];
        o[1] = (float)tr[1];
        o[2] = (float)tr[2];
        o[3] = (float)tr[3];
        o[4] = (float)ti[4];
        o[5] = (float)ti[5];
    }
}

// Output planes per accumulation pass of ztoeplitz_kernel: all of them in FP32;
// in FP64 as many as keep f (6 ZS doubles), the accumulators and one table line
// within the 256-register budget.
template <typename R, int ZS>
constexpr int ztoe_as() {
    return sizeof(R) == 4 ? (ZS <= 12 ? ZS : 8) : (ZS <= 10 ? (ZS + 1) / 2 : (ZS <= 12 ? 2 : 1));
}

// Toeplitz plane coupling of one column: reads f[c][j] = base[(c Zs + j) ld],
// writes g[c][a] = sum_b T_{a-b} f[c'][b] back over it.  Tc points at the
// column's entry of the table [Zs][6][ncols].  ZS >= Zs is the compile-time plane
// bound (planes Zs..ZS-1 are skipped), AS the number of output planes
// accumulated per pass (register budget).
// RI: with an odd nz every kz != 0 pairs with -kz, and the projector's xx, yy, zz,
// xy entries are even in kz while xz, yz are odd, so T_d's first four entries are
// real and the last two imaginary: each entry then costs two real FMAs per
// complex input instead of four.
template <typename R, int ZS, int AS, bool RI>
__device__ __forceinline__ void toeplitz_column(typename C2<R>::T* base, int ld, int Zs,
                                                const typename C2<R>::T* __restrict__ Tc, int ncols) {
    using CT = typename C2<R>::T;
    CT f[3][ZS];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < ZS; ++j) {
            if (j < Zs) {
                f[c][j] = base[(size_t)(c * Zs + j) * ld];
            } else {
                f[c][j].x = f[c][j].y = R(0);
            }
        }
#pragma unroll
    for (int a0 = 0; a0 < ZS; a0 += AS) {
        if (a0 >= Zs) continue;
        R ar[3][AS], ai[3][AS];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < AS; ++k) ar[c][k] = ai[c][k] = R(0);
#pragma unroll
        for (int dd = -(ZS - 1); dd <= ZS - 1; ++dd) {
            const int ad = dd < 0 ? -dd : dd;
            // pairs (a, a - dd) with a in [a0, a0 + AS) and both planes inside [0, ZS)
            const int alo = a0 > dd ? a0 : dd;                        // a - dd >= 0
            const int ahi = (a0 + AS < ZS + dd ? a0 + AS : ZS + dd);  // a - dd < ZS, a < a0 + AS
            if (alo >= ahi || ad >= Zs) continue;
            const CT* tp = Tc + (size_t)ad * 6 * ncols;
            R tr[6], ti[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const CT v = tp[(size_t)c * ncols];
                tr[c] = v.x;
                ti[c] = dd < 0 ? -v.y : v.y;
            }
            if (RI) {
                // real xx, yy, zz, xy (tr[0..3]); imaginary xz, yz (i ti[4], i ti[5])
#pragma unroll
                for (int k = 0; k < AS; ++k) {
                    const int a = a0 + k;
                    const int bb = a - dd;
                    if (a < alo || a >= ahi) continue;
                    if (a >= Zs || bb >= Zs) continue;
                    const CT fx = f[0][bb], fy = f[1][bb], fz = f[2][bb];
                    ar[0][k] += tr[0] * fx.x + tr[3] * fy.x - ti[4] * fz.y;
                    ai[0][k] += tr[0] * fx.y + tr[3] * fy.y + ti[4] * fz.x;
                    ar[1][k] += tr[3] * fx.x + tr[1] * fy.x - ti[5] * fz.y;
                    ai[1][k] += tr[3] * fx.y + tr[1] * fy.y + ti[5] * fz.x;
                    ar[2][k] += tr[2] * fz.x - ti[4] * fx.y - ti[5] * fy.y;
                    ai[2][k] += tr[2] * fz.y + ti[4] * fx.x + ti[5] * fy.x;
                }
                continue;
            }
#pragma unroll
            for (int k = 0; k < AS; ++k) {
                const int a = a0 + k;
                const int bb = a - dd;
                if (a < alo || a >= ahi) continue;
                if (a >= Zs || bb >= Zs) continue;
                const CT fx = f[0][bb], fy = f[1][bb], fz = f[2][bb];
                // row x: (xx, xy, xz); row y: (xy, yy, yz); row z: (xz, yz, zz)
                ar[0][k] += tr[0] * fx.x - ti[0] * fx.y + tr[3] * fy.x - ti[3] * fy.y + tr[4] * fz.x - ti[4] * fz.y;
                ai[0][k] += tr[0] * fx.y + ti[0] * fx.x + tr[3] * fy.y + ti[3] * fy.x + tr[4] * fz.y + ti[4] * fz.x;
                ar[1][k] += tr[3] * fx.x - ti[3] * fx.y + tr[1] * fy.x - ti[1] * fy.y + tr[5] * fz.x - ti[5] * fz.y;
                ai[1][k] += tr[3] * fx.y + ti[3] * fx.x + tr[1] * fy.y + ti[1] * fy.x + tr[5] * fz.y + ti[5] * fz.x;
                ar[2][k] += tr[4] * fx.x - ti[4] * fx.y + tr[5] * fy.x - ti[5] * fy.y + tr[2] * fz.x - ti[2] * fz.y;
                ai[2][k] += tr[4] * fx.y + ti[4] * fx.x + tr[5] * fy.y + ti[5] * fy.x + tr[2] * fz.y + ti[2] * fz.x;
            }
        }
#pragma unroll
        for (int k = 0; k < AS; ++k) {
            const int a = a0 + k;
            if (a >= Zs) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                CT o;
                o.x = ar[c][k];
                o.y = ai[c][k];
                base[(size_t)(c * Zs + a) * ld] = o;
            }
        }
    }
}

// One plane pair of the odd-nz FP32 coupling: output plane o gets T f[b] with T's real
// entries (xx, yy, zz, xy) and imaginary entries S (xz, yz), S = +1 for T_d, -1 for
// T_{-d} = conj(T_d); every accumulator is a chain of FMAs.
template <int ZS, int S>
__device__ __forceinline__ void couple_ri32(float (&ar)[3][ZS], float (&ai)[3][ZS], const float2 (&f)[3][ZS], int o,
                                            int b, float rxx, float ryy, float rzz, float rxy, float ixz, float iyz) {
    const float2 fx = f[0][b], fy = f[1][b], fz = f[2][b];
    const float sxz = S > 0 ? ixz : -ixz, syz = S > 0 ? iyz : -iyz;
    ar[0][o] = fmaf(rxx, fx.x, ar[0][o]);
    ar[0][o] = fmaf(rxy, fy.x, ar[0][o]);
    ar[0][o] = fmaf(-sxz, fz.y, ar[0][o]);
    ai[0][o] = fmaf(rxx, fx.y, ai[0][o]);
    ai[0][o] = fmaf(rxy, fy.y, ai[0][o]);
    ai[0][o] = fmaf(sxz, fz.x, ai[0][o]);
    ar[1][o] = fmaf(rxy, fx.x, ar[1][o]);
    ar[1][o] = fmaf(ryy, fy.x, ar[1][o]);
    ar[1][o] = fmaf(-syz, fz.y, ar[1][o]);
    ai[1][o] = fmaf(rxy, fx.y, ai[1][o]);
    ai[1][o] = fmaf(ryy, fy.y, ai[1][o]);
    ai[1][o] = fmaf(syz, fz.x, ai[1][o]);
    ar[2][o] = fmaf(rzz, fz.x, ar[2][o]);
    ar[2][o] = fmaf(-sxz, fx.y, ar[2][o]);
    ar[2][o] = fmaf(-syz, fy.y, ar[2][o]);
    ai[2][o] = fmaf(rzz, fz.y, ai[2][o]);
    ai[2][o] = fmaf(sxz, fx.x, ai[2][o]);
    ai[2][o] = fmaf(syz, fy.x, ai[2][o]);
}

// toeplitz_column for FP32 and an odd nz on the compact table: Tc points at the column's
// six floats of [Zs][ncols][6], dstride = 6 ncols.  Each T_d is loaded once and applied to
// the pairs at separations +d and -d; all ZS output planes accumulate in one pass.
template <int ZS>
__device__ __forceinline__ void toeplitz_column_ri32(float2* base, int ld, int Zs, const float* __restrict__ Tc,
                                                     size_t dstride) {
    float2 f[3][ZS];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int j = 0; j < ZS; ++j) f[c][j] = j < Zs ? base[(size_t)(c * Zs + j) * ld] : make_float2(0.f, 0.f);
    float ar[3][ZS], ai[3][ZS];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < ZS; ++k) ar[c][k] = ai[c][k] = 0.f;
#pragma unroll
    for (int d = 0; d < ZS; ++d) {
        if (d < Zs) {
            const float* tp = Tc + (size_t)d * dstride;
            const float2 t01 = *reinterpret_cast<const float2*>(tp);
            const float2 t23 = *reinterpret_cast<const float2*>(tp + 2);
            const float2 t45 = *reinterpret_cast<const float2*>(tp + 4);
#pragma unroll
            for (int a = d; a < ZS; ++a) {
                if (a < Zs) {
                    couple_ri32<ZS, 1>(ar, ai, f, a, a - d, t01.x, t01.y, t23.x, t23.y, t45.x, t45.y);
                    if (d > 0) couple_ri32<ZS, -1>(ar, ai, f, a - d, a, t01.x, t01.y, t23.x, t23.y, t45.x, t45.y);
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < ZS; ++a) {
        if (a < Zs) {
#pragma unroll
            for (int c = 0; c < 3; ++c) base[(size_t)(c * Zs + a) * ld] = make_float2(ar[c][a], ai[c][a]);
        }
    }
}

// Applies the Toeplitz plane coupling in place on the slab [B][3][Zs][column].
// One thread per (column, right-hand side); a block covers (256 / rb) columns x rb
// right-hand sides so the rb threads of a column share the table lines in cache.
template <typename R, int ZS, int AS, bool RI>
__global__ __launch_bounds__(256) void ztoeplitz_kernel(typename C2<R>::T* __restrict__ slab, int B, int ncols,
                                                        int Zs, int rb, const typename C2<R>::T* __restrict__ T) {
    const int cb = 256 / rb;
    const int tc = threadIdx.x % cb, tb = threadIdx.x / cb;
    const int col = blockIdx.x * cb + tc;
    const int b = blockIdx.y * rb + tb;
    if (col >= ncols || b >= B) return;
    toeplitz_column<R, ZS, AS, RI>(slab + (size_t)b * 3 * Zs * ncols + col, ncols, Zs, T + col, ncols);
}

// ---- x transform, Toeplitz coupling and inverse x transform in one pass ----
//
// One block per (ky row, right-hand side) of the transposed k-space buffer
// [B][3][Zs][ky][nxp] (after the y transform): the 3 Zs x-lines of the row are
// transformed along x into LDS with two register stages n = P1 P2 (decimation
// in time: global -> LDS, then in place), coupled across planes column by
// column (toeplitz_column on LDS), transformed back (decimation in frequency:
// in place, then LDS -> global) and written over their input.  With input index
// j = j2 + P2 j1 and output index k = k1 + P1 k2,
//   X[k] = sum_j2 W_P2^{j2 k2} W_n^{j2 k1} sum_j1 W_P1^{j1 k1} x[j2 + P2 j1].
// It replaces the forward x pass, the Toeplitz pass and the inverse x pass
// (three reads and three writes of the slab) by one read and one write.
// Blocks are numbered so that the rows of one XCD are ky = xcd (mod 8)
// (workgroups go to the 8 XCDs round-robin): each XCD's L2 then holds one
// eighth of the plane-coupling table, re-read by every right-hand side.
template <typename R, int P, bool INV>
__device__ __forceinline__ void dft_strided(const typename C2<R>::T* in, int ib, int is, typename C2<R>::T* out,
                                            int ob, int os) {
    using CT = typename C2<R>::T;
    CT v[P];
#pragma unroll
    for (int r = 0; r < P; ++r) v[r] = in[ib + r * is];
    dft_reg<R, P, INV>(v);
#pragma unroll
    for (int q = 0; q < P; ++q) out[ob + q * os] = v[q];
}

// In-place decimation-in-frequency stage: x[base + r stride] (r < P) -> DFT ->
// output k1 multiplied by tw[j2 k1].
template <typename R, int P, bool INV>
__device__ __forceinline__ void dif_stage(typename C2<R>::T* x, int base, int stride,
                                          const typename C2<R>::T* tw, int j2) {
    using CT = typename C2<R>::T;
    CT v[P];
#pragma unroll
    for (int r = 0; r < P; ++r) v[r] = x[base + r * stride];
    dft_reg<R, P, INV>(v);
    if (j2 > 0) {
#pragma unroll
        for (int k = 1; k < P; ++k) v[k] = cmul_t(v[k], tw[j2 * k]);
    }
#pragma unroll
    for (int k = 0; k < P; ++k) x[base + k * stride] = v[k];
}

// Up to 10 planes, three blocks per CU fit the LDS (the 210-point lines of a 9-plane slab
// take 45 KB per block): the register budget is held to the matching three waves per SIMD
// (at most 168 VGPRs).  Thicker slabs hold more planes in registers and run two.
template <typename R, int ZS, int CLS, bool RI>
__global__ __launch_bounds__(256, ZS <= 10 ? 3 : 2) void xtoeplitz_kernel(typename C2<R>::T* __restrict__ kbuf, int Zs, int ny,
                                                        int nxp, int n, int P1, int P2,
                                                        const typename C2<R>::T* __restrict__ twiddle,
                                                        const typename C2<R>::T* __restrict__ T,
                                                        const float* __restrict__ Tri) {
    using CT = typename C2<R>::T;
    constexpr int AS = ztoe_as<R, ZS>();
    extern __shared__ unsigned char s_raw[];
    const int nl = 3 * Zs;
    CT* const X = reinterpret_cast<CT*>(s_raw);  // [nl][n], then the forward and inverse twiddles
    CT* const twf = X + nl * n;
    CT* const twi = twf + n;
    const int tid = threadIdx.x;
    const int nky8 = (ny + 7) >> 3;
    const int slot = blockIdx.x >> 3;
    const int ky = ((slot % nky8) << 3) + (blockIdx.x & 7), b = slot / nky8;
    if (ky >= ny) return;  // whole block
    const size_t plane = (size_t)ny * nxp;
    CT* const row0 = kbuf + (size_t)b * nl * plane + (size_t)ky * nxp;  // line li at row0 + li plane
    for (int k = tid; k < n; k += blockDim.x) {
        CT w = twiddle[k];
        twf[k] = w;
        w.y = -w.y;
        twi[k] = w;
    }
    // forward, radix P1: x[j2 + P2 j1] -> X[li n + P1 j2 + k1]
    for (int w = tid; w < nl * P2; w += blockDim.x) {
        const int li = w / P2, j2 = w - li * P2;
        const CT* in = row0 + li * plane;
        switch (P1) {
#define MPM_XS(P, C)                                                                  \
    case P:                                                                           \
        if constexpr (CLS >= C) dft_strided<R, P, false>(in, j2, P2, X, li * n + P1 * j2, 1); \
        break;
            MPM_FFT2_CASES(MPM_XS)
#undef MPM_XS
            default:
                break;
        }
    }
    __syncthreads();
    // forward, radix P2 with twiddles W_n^{k1 j2}, in place: X[li n + k1 + P1 r]
    for (int w = tid; w < nl * P1; w += blockDim.x) {
        const int li = w / P1, k1 = w - li * P1;
        switch (P2) {
#define MPM_XS(P, C)                                                                                 \
    case P:                                                                                          \
        if constexpr (CLS >= C) fft2_s2<R, P, false>(X, li * n + k1, P1, twf, k1, X, li * n + k1, P1); \
        break;
            MPM_FFT2_CASES(MPM_XS)
#undef MPM_XS
            default:
                break;
        }
    }
    __syncthreads();
    const int ncols = ny * nxp;
    for (int kx = tid; kx < n; kx += blockDim.x) {
        if constexpr (RI && std::is_same<R, float>::value)
            toeplitz_column_ri32<ZS>(X + kx, n, Zs, Tri + ((size_t)ky * nxp + kx) * 6, (size_t)ncols * 6);
        else
            toeplitz_column<R, ZS, AS, RI>(X + kx, n, Zs, T + (size_t)ky * nxp + kx, ncols);
    }
    __syncthreads();
    // inverse, radix P1, in place: X[li n + j2 + P2 r] -> DFT -> * conj W_n^{j2 k1}
    for (int w = tid; w < nl * P2; w += blockDim.x) {
        const int li = w / P2, j2 = w - li * P2;
        switch (P1) {
#define MPM_XS(P, C)                                                              \
    case P:                                                                       \
        if constexpr (CLS >= C) dif_stage<R, P, true>(X, li * n + j2, P2, twi, j2); \
        break;
            MPM_FFT2_CASES(MPM_XS)
#undef MPM_XS
            default:
                break;
        }
    }
    __syncthreads();
    // inverse, radix P2: X[li n + P2 k1 + j2] -> out[k1 + P1 k2]
    for (int w = tid; w < nl * P1; w += blockDim.x) {
        const int li = w / P1, k1 = w - li * P1;
        CT* out = row0 + li * plane;
        switch (P2) {
#define MPM_XS(P, C)                                                                       \
    case P:                                                                                \
        if constexpr (CLS >= C) dft_strided<R, P, true>(X, li * n + P2 * k1, 1, out, k1, P1); \
        break;
            MPM_FFT2_CASES(MPM_XS)
#undef MPM_XS
            default:
                break;
        }
    }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One wavefront per (particle, group of G right-hand sides; blockIdx.y strides the
// groups): Gaussian contraction of the (inverse-transformed)
// grid, real-space gather over the symmetric neighbour list, self term.
template <typename R, int G, bool SLAB>
__global__ __launch_bounds__(256) void gather_kernel(
    SlotTable tab, int B, int parts, const typename C2<R>::T* __restrict__ grid, SpreadGeom g,
    const int* __restrict__ zmap, const int* __restrict__ base, const R* __restrict__ wts, int N, double h3,
    const double* __restrict__ pos, const int* __restrict__ nb_start, const int* __restrict__ nb_j,
    const double2* __restrict__ nb_f, double Lx, double Ly, double Lz, double kbx, double kby,
    const double2* __restrict__ selfc, double self_perp, int jacobi, const R* __restrict__ partial,
    const int* __restrict__ ps_start, const int* __restrict__ ps_items, const int* __restrict__ order) {
    __shared__ int s_off[kMaxWidth * kMaxWidth * kMaxWidth];
    const int lane = threadIdx.x & 63;
    // waves visit the particles in spatial order (neighbouring waves share stencil voxels)
    const int wv = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int i = wv < N ? order[wv] : N;
    const int width = g.width, half = g.half;
    {   // stencil offsets (ox, oy, oz) packed, in the order matching the grid layout
        const int S0 = width * width * width;
        for (int t = threadIdx.x; t < S0; t += blockDim.x) {
            int ox, oy, oz;
            if (SLAB) {
                oy = t % width;
                ox = (t / width) % width;
                oz = t / (width * width);
            } else {
                oz = t % width;
                oy = (t / width) % width;
                ox = t / (width * width);
            }
            s_off[t] = ox | (oy << 8) | (oz << 16);
        }
    }
    __syncthreads();
    if (i >= N) return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int S = width * width * width;
    const size_t nxy = (size_t)nx * ny;
    const size_t plane = SLAB ? (size_t)nx * g.nyp : nxy;
    const size_t rhs_stride = SLAB ? 3 * (size_t)g.Zs * plane : 3 * nxy * nz;
    const size_t comp_stride = SLAB ? (size_t)g.Zs * plane : nxy * nz;
    int sx = 0, sy = 0, sz = 0;
    const R* w = wts + (size_t)i * 3 * width;
    if (parts & MPM_PART_RECIP) {
        sx = ((base[i] - half) % nx + nx) % nx;
        sy = ((base[N + i] - half) % ny + ny) % ny;
        sz = ((base[2 * N + i] - half) % nz + nz) % nz;
    }
    const double xi_ = pos[i], yi_ = pos[N + i], zi_ = pos[2 * N + i];
    const bool bloch = (kbx != 0.0 || kby != 0.0);

    for (int g0 = blockIdx.y * G; g0 < B; g0 += G * gridDim.y) {
        double acc[G][3][2];
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[j][c][0] = acc[j][c][1] = 0.0;

        if (parts & MPM_PART_RECIP) {
            R racc[G][3][2];
#pragma unroll
            for (int j = 0; j < G; ++j)
#pragma unroll
                for (int c = 0; c < 3; ++c) racc[j][c][0] = racc[j][c][1] = R(0);
            // contraction already done per region on the matrix cores: sum the
            // particle's region slots (ascending region order)
            const int s1 = partial ? ps_start[i + 1] : 0;
            for (int t = (partial ? ps_start[i] : 0) + lane; t < s1; t += 64) {
                const R* pp = partial + (size_t)ps_items[t] * (6 * B);
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    if (g0 + j >= B) break;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        racc[j][c][0] += (R)pp[6 * (g0 + j) + 2 * c];
                        racc[j][c][1] += (R)pp[6 * (g0 + j) + 2 * c + 1];
                    }
                }
            }
            for (int t = lane; t < (partial ? 0 : S); t += 64) {
                const int pk = s_off[t];
                const int ox = pk & 255, oy = (pk >> 8) & 255, oz = pk >> 16;
                int vx = sx + ox, vy = sy + oy, vz = sz + oz;
                if (g.nofold) {
                    vx -= (vx >= nx) ? nx : 0;
                    vy -= (vy >= ny) ? ny : 0;
                    vz -= (vz >= nz) ? nz : 0;
                } else {
                    while (vx >= nx) vx -= nx;
                    while (vy >= ny) vy -= ny;
                    while (vz >= nz) vz -= nz;
                }
                const R ww = w[ox] * w[width + oy] * w[2 * width + oz];
                size_t v;
                if (SLAB)
                    v = (size_t)zmap[vz] * plane + (size_t)vx * g.nyp + vy;
                else
                    v = ((size_t)vx * ny + vy) * nz + vz;
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    if (g0 + j >= B) break;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const auto f = grid[(size_t)(g0 + j) * rhs_stride + c * comp_stride + v];
                        racc[j][c][0] += ww * f.x;
                        racc[j][c][1] += ww * f.y;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < G; ++j)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    acc[j][c][0] = (double)racc[j][c][0] * h3;
                    acc[j][c][1] = (double)racc[j][c][1] * h3;
                }
        }
        if (parts & MPM_PART_REAL) {
            const int k0 = nb_start[i], k1 = nb_start[i + 1];
            for (int k = k0 + lane; k < k1; k += 64) {
                const int j = nb_j[k];
                const double2 f = nb_f[k];
                double rx = pos[j] - xi_, ry = pos[N + j] - yi_, rz = pos[2 * N + j] - zi_;
                rx -= Lx * round(rx / Lx);
                ry -= Ly * round(ry / Ly);
                rz -= Lz * round(rz / Lz);
                const double d = sqrt(rx * rx + ry * ry + rz * rz);
                const double de = d < 1.0 ? 1.0 : d;
                const double ux = rx / de, uy = ry / de, uz = rz / de;
                double pc = 1.0, ps = 0.0;
                if (bloch) {
                    const double ph = -(kbx * rx + kby * ry);
                    pc = cos(ph);
                    ps = sin(ph);
                }
#pragma unroll
                for (int jj = 0; jj < G; ++jj) {
                    if (g0 + jj >= B) break;
                    const double2* p = tab.x[g0 + jj] + 3 * (size_t)j;
                    const double2 p0 = p[0], p1 = p[1], p2 = p[2];
                    const double rpr = p0.x * ux + p1.x * uy + p2.x * uz;
                    const double rpi = p0.y * ux + p1.y * uy + p2.y * uz;
                    const double u[3] = {ux, uy, uz};
                    const double2 pp[3] = {p0, p1, p2};
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double er = (pp[c].x - rpr * u[c]) * f.x + rpr * u[c] * f.y;
                        const double ei = (pp[c].y - rpi * u[c]) * f.x + rpi * u[c] * f.y;
                        acc[jj][c][0] += er * pc - ei * ps;
                        acc[jj][c][1] += er * ps + ei * pc;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                acc[j][c][0] = wave_sum(acc[j][c][0]);
                acc[j][c][1] = wave_sum(acc[j][c][1]);
            }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < G; ++j) {
                const int b = g0 + j;
                if (b >= B) break;
                const double2* p = tab.x[b] + 3 * (size_t)i;
                double2 fs = make_double2(0.0, 0.0);
                if (parts & MPM_PART_REAL) {
                    fs = selfc[(size_t)tab.row[b] * N + i];
                    fs.x += self_perp;
                }
                double2* y = tab.y[b] + 3 * (size_t)i;
                // optional left Jacobi: y <- y / self_coef (numerical_solver.cu:207-242)
                double2 inv = make_double2(1.0, 0.0);
                if (jacobi) {
                    const double2 sc = selfc[(size_t)tab.row[b] * N + i];
                    const double den = sc.x * sc.x + sc.y * sc.y;
                    if (den > 1e-30) inv = make_double2(sc.x / den, -sc.y / den);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 pc = p[c];
                    const double yr = acc[j][c][0] + fs.x * pc.x - fs.y * pc.y;
                    const double yi = acc[j][c][1] + fs.x * pc.y + fs.y * pc.x;
                    y[c] = make_double2(yr * inv.x - yi * inv.y, yr * inv.y + yi * inv.x);
                }
            }
        }
    }
}

// kFinishLanes lanes per (particle i, right-hand side b), lanes then b fastest,
// the particles in the spatial visiting order of the wave-per-particle kernels:
// the reciprocal part as the sum of the particle's region contractions, the
// real-space part over the neighbour list with the precomputed pair
// coefficients (ux, uy, uz, perp, para - perp, cos, sin) -- the field of p_j at
// i is perp p_j + (para - perp) u (u . p_j), times the Bloch phase -- each lane
// taking every kFinishLanes-th slot and neighbour, then a butterfly sum over the
// lanes (fixed order), the self term and the optional left Jacobi scaling.
// Used with the matrix-core contraction when the batch is wide enough to fill
// the GPU; splitting the neighbour loop over lanes keeps enough independent
// loads in flight at a few thousand particles.
constexpr int kFinishLanes = 4;

template <typename R>
__global__ __launch_bounds__(256) void finish_kernel(SlotTable tab, int B, int parts, int N,
                                                     const R* __restrict__ partial, const int* __restrict__ ps_start,
                                                     const int* __restrict__ ps_items, double h3,
                                                     const int* __restrict__ nb_start, const int* __restrict__ nb_j,
                                                     const double* __restrict__ pairc,
                                                     const double2* __restrict__ selfc, double self_perp,
                                                     int jacobi, const int* __restrict__ order) {
    constexpr int NL = kFinishLanes;
    __shared__ const double2* s_x[kMaxSlots];
    __shared__ double2* s_y[kMaxSlots];
    __shared__ int s_row[kMaxSlots];
#pragma unroll
    for (int k = 0; k < kMaxSlots; ++k)  // uniform indices: no per-lane walk of the argument
        if (threadIdx.x == k) {
            s_x[k] = tab.x[k];
            s_y[k] = tab.y[k];
            s_row[k] = tab.row[k];
        }
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int item = t / NL, sub = t % NL;  // NL divides the wave: the lanes of an item share it
    const int ii = item / B, b = item - ii * B;
    if (ii >= N) return;  // whole items only: all lanes of an item leave together
    // particles in spatial order: neighbouring items share neighbours' dipoles and pair data
    const int i = order[ii];
    const double2* xb = s_x[b];
    double ar[3] = {0.0, 0.0, 0.0}, ai[3] = {0.0, 0.0, 0.0};
    double br[3] = {0.0, 0.0, 0.0}, bi[3] = {0.0, 0.0, 0.0};
    // the slot sum (the caller always asks for the reciprocal part here) and the
    // neighbour sum; the lane's first two slots and first two neighbours are fetched
    // together (indices, then data, read at clamped in-range positions and dropped when
    // past the end), the rest in the loops that follow
    const int s0 = ps_start[i], s1 = ps_start[i + 1];
    const bool real = (parts & MPM_PART_REAL) != 0;
    const int k0 = real ? nb_start[i] : 0, k1 = real ? nb_start[i + 1] : 0;
    // entries past the end are masked to zero (bit masks, not branches, so that the
    // compiler issues every load before the first use)
    auto slot = [&](int item_slot, bool take) {
        const R* pp = partial + (size_t)item_slot * (6 * B) + 6 * b;
        R v[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) v[e] = keep_if(pp[e], take);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ar[c] += (double)v[2 * c];
            ai[c] += (double)v[2 * c + 1];
        }
    };
    auto neighbour = [&](int k, int j, bool take) {
        const double* pc = pairc + 8 * (size_t)k;
        const double ux = pc[0], uy = pc[1], uz = pc[2], cp = pc[3], cd = pc[4];
        const double ec = keep_if(pc[5], take), es = keep_if(pc[6], take);  // zero phase: no contribution
        const double2 p0 = xb[3 * (size_t)j], p1 = xb[3 * (size_t)j + 1], p2 = xb[3 * (size_t)j + 2];
        const double rr = cd * (p0.x * ux + p1.x * uy + p2.x * uz);
        const double ri = cd * (p0.y * ux + p1.y * uy + p2.y * uz);
        const double u[3] = {ux, uy, uz};
        const double2 pp[3] = {p0, p1, p2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double er = cp * pp[c].x + rr * u[c];
            const double ei = cp * pp[c].y + ri * u[c];
            br[c] += er * ec - ei * es;
            bi[c] += er * es + ei * ec;
        }
    };
    {
        const int sa = s0 + sub, sb = sa + NL, ka = k0 + sub, kb = ka + NL;
        const int sac = max(min(sa, s1 - 1), 0), sbc = max(min(sb, s1 - 1), 0);
        const int kac = max(min(ka, k1 - 1), 0), kbc = max(min(kb, k1 - 1), 0);
        const int ia = ps_items[sac], ib = ps_items[sbc], ja = nb_j[kac], jb = nb_j[kbc];
        slot(ia, sa < s1);
        slot(ib, sb < s1);
        neighbour(kac, ja, ka < k1);
        neighbour(kbc, jb, kb < k1);
    }
    for (int sl = s0 + sub + 2 * NL; sl < s1; sl += NL) slot(ps_items[sl], true);
    for (int k = k0 + sub + 2 * NL; k < k1; k += NL) neighbour(k, nb_j[k], true);
    // butterfly over the item's lanes: every lane ends with the same sums
#pragma unroll
    for (int m = 1; m < NL; m <<= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ar[c] += __shfl_xor(ar[c], m, NL);
            ai[c] += __shfl_xor(ai[c], m, NL);
            br[c] += __shfl_xor(br[c], m, NL);
            bi[c] += __shfl_xor(bi[c], m, NL);
        }
    if (sub >= 3) return;
    const int c = sub;  // lane c of the item writes component c
    const double2* p = xb + 3 * (size_t)i;
    double2 fs = make_double2(0.0, 0.0);
    if (parts & MPM_PART_REAL) {
        fs = selfc[(size_t)s_row[b] * N + i];
        fs.x += self_perp;
    }
    // optional left Jacobi: y <- y / self_coef (numerical_solver.cu:207-242)
    double2 inv = make_double2(1.0, 0.0);
    if (jacobi) {
        const double2 sc = selfc[(size_t)s_row[b] * N + i];
        const double den = sc.x * sc.x + sc.y * sc.y;
        if (den > 1e-30) inv = make_double2(sc.x / den, -sc.y / den);
    }
    const double2 pc = p[c];
    const double fr = c == 0 ? ar[0] : (c == 1 ? ar[1] : ar[2]);
    const double fi = c == 0 ? ai[0] : (c == 1 ? ai[1] : ai[2]);
    const double gr = c == 0 ? br[0] : (c == 1 ? br[1] : br[2]);
    const double gi = c == 0 ? bi[0] : (c == 1 ? bi[1] : bi[2]);
    const double yr = fr * h3 + gr + fs.x * pc.x - fs.y * pc.y;
    const double yi = fi * h3 + gi + fs.x * pc.y + fs.y * pc.x;
    s_y[b][3 * (size_t)i + c] = make_double2(yr * inv.x - yi * inv.y, yr * inv.y + yi * inv.x);
}

// ---------------------------------------------------------------------------
// Polydisperse operator kernels (full grid [B][nx][ny][nz], one complex per voxel)
// ---------------------------------------------------------------------------

// Scalar k-space coefficient exp(-(1 - eta) k^2 / 4 xi^2) / k^2 / V at k = G - k_bloch,
// zero at k = 0 (polydisperse...cu:565-640).
__global__ void poly_kspace_kernel(double* __restrict__ k64, float* __restrict__ k32, KParams q, double eta) {
    const size_t V = (size_t)q.nx * q.ny * q.nz;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
        const int ix = (int)(v / ((size_t)q.ny * q.nz)), iy = (int)((v / q.nz) % q.ny), iz = (int)(v % q.nz);
        const double fx = ix <= (q.nx - 1) / 2 ? ix : ix - q.nx;
        const double fy = iy <= (q.ny - 1) / 2 ? iy : iy - q.ny;
        const double fz = iz <= (q.nz - 1) / 2 ? iz : iz - q.nz;
        const double kx = fx * 2.0 * kPi / q.Lx - q.kbx, ky = fy * 2.0 * kPi / q.Ly - q.kby, kz = fz * 2.0 * kPi / q.Lz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double sc = k2 < 1e-12 ? 0.0 : exp(-(1.0 - eta) * k2 / (4.0 * q.xi * q.xi)) / k2 * q.invV;
        k64[v] = sc;
        k32[v] = (float)sc;
    }
}

// Owner-computes spread of the scalar grid S = sum_i C_i . p_i: a thread per
// voxel and G right-hand sides walks the voxel's (particle, stencil point)
// entries in ascending order (no atomics, a fixed summation order).
template <typename R, int G>
__global__ __launch_bounds__(256) void poly_spread_kernel(SlotTable tab, int B, int V, int SS,
                                                          const int* __restrict__ vox_start,
                                                          const int* __restrict__ vox_items,
                                                          const R* __restrict__ pw,
                                                          typename C2<R>::T* __restrict__ out) {
    using CT = typename C2<R>::T;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int g0 = blockIdx.y * G;
    R ar[G], ai[G];
#pragma unroll
    for (int j = 0; j < G; ++j) ar[j] = ai[j] = R(0);
    const int e0 = vox_start[v], e1 = vox_start[v + 1];
    for (int e = e0; e < e1; ++e) {
        const int item = vox_items[e];
        const int pid = item / SS;
        const R w0 = pw[3 * (size_t)item], w1 = pw[3 * (size_t)item + 1], w2 = pw[3 * (size_t)item + 2];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if (g0 + j < B) {
                const double2* x = tab.x[g0 + j] + 3 * (size_t)pid;
                const double2 p0 = x[0], p1 = x[1], p2 = x[2];
                ar[j] += w0 * (R)p0.x + w1 * (R)p1.x + w2 * (R)p2.x;
                ai[j] += w0 * (R)p0.y + w1 * (R)p1.y + w2 * (R)p2.y;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (g0 + j < B) {
            CT o;
            o.x = ar[j];
            o.y = ai[j];
            out[(size_t)(g0 + j) * V + v] = o;
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void poly_scale_kernel(typename C2<R>::T* __restrict__ grid,
                                                         const R* __restrict__ sc, size_t V, size_t total) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const R f = sc[e % V];
        grid[e].x *= f;
        grid[e].y *= f;
    }
}

// Radius gradient of the polydisperse operator, a wave per particle i (all B systems):
//   out_i = sum_b Re(g_b^H (dA/da_i) x_b)
// with three parts.  Reciprocal: h^3 sum_o dC_i(o) . (conj(g_i) phx[v] + x_i conj(phg[v])),
// phx / phg the convolved spread grids of x / g (the kernel is real and even, so the
// derivative of i's spread weight, seen through g^H, contracts against conj(phg)).  Real
// space: for each neighbour j, D = dperp I + d(para - perp) u u^T (the row's derivative),
// Re(g_i^H D x_j) + Re(g_j^H D x_i).  Self: dself_i Re(g_i^H x_i).  FP64 throughout.
__global__ __launch_bounds__(256) void poly_rgrad_kernel(int B, int N, int V, int SS, size_t n3,
                                                         const double2* __restrict__ x, const double2* __restrict__ g,
                                                         const double2* __restrict__ phx,
                                                         const double2* __restrict__ phg,
                                                         const double* __restrict__ dpw, const int* __restrict__ pidx,
                                                         double h3, const int* __restrict__ nb_start,
                                                         const int* __restrict__ nb_j, const double* __restrict__ pairc,
                                                         const double2* __restrict__ dpair,
                                                         const double* __restrict__ dps, double* __restrict__ out,
                                                         int accumulate) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;  // whole wave
    double acc = 0.0, rec = 0.0;
    for (int o = lane; o < SS; o += 64) {
        const size_t item = (size_t)i * SS + o;
        const int v = pidx[item];
        const double w[3] = {dpw[3 * item], dpw[3 * item + 1], dpw[3 * item + 2]};
        for (int b = 0; b < B; ++b) {
            const double2 fx = phx[(size_t)b * V + v], fg = phg[(size_t)b * V + v];
            const double2* xb = x + (size_t)b * n3 + 3 * (size_t)i;
            const double2* gb = g + (size_t)b * n3 + 3 * (size_t)i;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                rec += w[c] * (gb[c].x * fx.x + gb[c].y * fx.y + xb[c].x * fg.x + xb[c].y * fg.y);
        }
    }
    acc = h3 * rec;
    for (int k = nb_start[i] + lane; k < nb_start[i + 1]; k += 64) {
        const int j = nb_j[k];
        const double* pc = pairc + 8 * (size_t)k;
        const double u[3] = {pc[0], pc[1], pc[2]};
        const double2 dp = dpair[k];
        for (int b = 0; b < B; ++b) {
            const double2* xi_ = x + (size_t)b * n3 + 3 * (size_t)i;
            const double2* gi_ = g + (size_t)b * n3 + 3 * (size_t)i;
            const double2* xj_ = x + (size_t)b * n3 + 3 * (size_t)j;
            const double2* gj_ = g + (size_t)b * n3 + 3 * (size_t)j;
            double uxj_r = 0, uxj_i = 0, uxi_r = 0, uxi_i = 0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                uxj_r += u[c] * xj_[c].x;
                uxj_i += u[c] * xj_[c].y;
                uxi_r += u[c] * xi_[c].x;
                uxi_i += u[c] * xi_[c].y;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double djr = dp.x * xj_[c].x + dp.y * uxj_r * u[c], dji = dp.x * xj_[c].y + dp.y * uxj_i * u[c];
                const double dir = dp.x * xi_[c].x + dp.y * uxi_r * u[c], dii = dp.x * xi_[c].y + dp.y * uxi_i * u[c];
                acc += gi_[c].x * djr + gi_[c].y * dji + gj_[c].x * dir + gj_[c].y * dii;
            }
        }
    }
    if (lane == 0)
        for (int b = 0; b < B; ++b) {
            const double2* xb = x + (size_t)b * n3 + 3 * (size_t)i;
            const double2* gb = g + (size_t)b * n3 + 3 * (size_t)i;
#pragma unroll
            for (int c = 0; c < 3; ++c) acc += dps[i] * (gb[c].x * xb[c].x + gb[c].y * xb[c].y);
        }
    acc = wave_sum(acc);
    if (lane == 0) out[i] = accumulate ? out[i] + acc : acc;
}

// A wave per (particle i, group of G right-hand sides): the reciprocal field
// h^3 sum_o C_i(o) S[v(i, o)] over the particle's P^3 stencil points, the
// real-space sum over its neighbours with the pair coefficients of the
// monodisperse finish step, the self term (coefficient + the particle's own
// real-space self value) and the optional left Jacobi scaling; FP64 sums, one
// fixed-order wave reduction, one store.  polydisperse...cu:654-828.
template <typename R, int G>
__global__ __launch_bounds__(256) void poly_gather_kernel(SlotTable tab, int B, int parts, int N, int V, int SS,
                                                          const typename C2<R>::T* __restrict__ grid,
                                                          const R* __restrict__ pw, const int* __restrict__ pidx,
                                                          double h3, const int* __restrict__ nb_start,
                                                          const int* __restrict__ nb_j,
                                                          const double* __restrict__ pairc,
                                                          const double2* __restrict__ selfc,
                                                          const double* __restrict__ selfp, int jacobi,
                                                          const int* __restrict__ order) {
    using CT = typename C2<R>::T;
    const int lane = threadIdx.x & 63;
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= N) return;  // whole wave
    const int i = order[wv];  // spatial order: neighbouring waves share stencil voxels
    const int g0 = blockIdx.y * G;
    double acc[G][3][2];
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[j][c][0] = acc[j][c][1] = 0.0;
    if (parts & MPM_PART_RECIP) {
        for (int o = lane; o < SS; o += 64) {
            const size_t item = (size_t)i * SS + o;
            const int v = pidx[item];
            const double w[3] = {(double)pw[3 * item], (double)pw[3 * item + 1], (double)pw[3 * item + 2]};
#pragma unroll
            for (int j = 0; j < G; ++j) {
                if (g0 + j < B) {
                    const CT sv = grid[(size_t)(g0 + j) * V + v];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        acc[j][c][0] += w[c] * (double)sv.x;
                        acc[j][c][1] += w[c] * (double)sv.y;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                acc[j][c][0] *= h3;
                acc[j][c][1] *= h3;
            }
    }
    if (parts & MPM_PART_REAL) {
        const int k0 = nb_start[i], k1 = nb_start[i + 1];
        for (int k = k0 + lane; k < k1; k += 64) {
            const int jn = nb_j[k];
            const double* pc = pairc + 8 * (size_t)k;
            const double u[3] = {pc[0], pc[1], pc[2]};
            const double cp = pc[3], cd = pc[4], ec = pc[5], es = pc[6];
#pragma unroll
            for (int j = 0; j < G; ++j) {
                if (g0 + j < B) {
                    const double2* x = tab.x[g0 + j] + 3 * (size_t)jn;
                    const double2 pp[3] = {x[0], x[1], x[2]};
                    const double rr = cd * (pp[0].x * u[0] + pp[1].x * u[1] + pp[2].x * u[2]);
                    const double ri = cd * (pp[0].y * u[0] + pp[1].y * u[1] + pp[2].y * u[2]);
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double er = cp * pp[c].x + rr * u[c];
                        const double ei = cp * pp[c].y + ri * u[c];
                        acc[j][c][0] += er * ec - ei * es;
                        acc[j][c][1] += er * es + ei * ec;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            acc[j][c][0] = wave_sum(acc[j][c][0]);
            acc[j][c][1] = wave_sum(acc[j][c][1]);
        }
    if (lane >= G * 3) return;
    const int j = lane / 3, c = lane % 3;  // lane (j, c) stores component c of right-hand side g0 + j
    const int b = g0 + j;
    if (b >= B) return;
    double yr = 0.0, yi = 0.0;
    const double2* xb = nullptr;
    double2* yb = nullptr;
    int row = 0;
#pragma unroll
    for (int jj = 0; jj < G; ++jj) {  // uniform indices into the argument, selected per lane
        if (jj == j) {
            xb = tab.x[g0 + jj];
            yb = tab.y[g0 + jj];
            row = tab.row[g0 + jj];
        }
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            if (jj == j && cc == c) {
                yr = acc[jj][cc][0];
                yi = acc[jj][cc][1];
            }
    }
    const double2 pcv = xb[3 * (size_t)i + c];
    const double2 sc = selfc[(size_t)row * N + i];
    if (parts & MPM_PART_REAL) {
        const double fr = sc.x + selfp[i], fi = sc.y;
        yr += fr * pcv.x - fi * pcv.y;
        yi += fr * pcv.y + fi * pcv.x;
    }
    double2 inv = make_double2(1.0, 0.0);
    if (jacobi) {
        const double den = sc.x * sc.x + sc.y * sc.y;
        if (den > 1e-30) inv = make_double2(sc.x / den, -sc.y / den);
    }
    yb[3 * (size_t)i + c] = make_double2(yr * inv.x - yi * inv.y, yr * inv.y + yi * inv.x);
}

// ---------------------------------------------------------------------------
// Quadrupole kernels (monodisperse, full grid [B][8][nx*ny*nz]: the field's 3
// components, then the 5 quadrupole-gradient components)
// ---------------------------------------------------------------------------

// Per voxel: khat (3), sc, scQ, scGP, scGQ, Qfactor (5), Qdot (5) (:1301-1404).
__global__ void quad_kspace_kernel(double* __restrict__ k64, float* __restrict__ k32, KParams q) {
    const size_t V = (size_t)q.nx * q.ny * q.nz;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
        const int ix = (int)(v / ((size_t)q.ny * q.nz)), iy = (int)((v / q.nz) % q.ny), iz = (int)(v % q.nz);
        const double fx = ix <= (q.nx - 1) / 2 ? ix : ix - q.nx;
        const double fy = iy <= (q.ny - 1) / 2 ? iy : iy - q.ny;
        const double fz = iz <= (q.nz - 1) / 2 ? iz : iz - q.nz;
        const double Gx = fx * 2.0 * kPi / q.Lx, Gy = fy * 2.0 * kPi / q.Ly, Gz = fz * 2.0 * kPi / q.Lz;
        const double kx = Gx - q.kbx, ky = Gy - q.kby, kz = Gz;
        const double k2 = kx * kx + ky * ky + kz * kz, k = sqrt(k2);
        double o[16];
        for (int c = 0; c < 16; ++c) o[c] = 0.0;
        if (k >= 1e-12) {
            const double h0 = kx / k, h1 = ky / k, h2 = kz / k;
            const double etak2 = k2 - (Gx * Gx * q.etax + Gy * Gy * q.etay + Gz * Gz * q.etaz);
            double sk, ck;
            sincos(k, &sk, &ck);
            const double t = sk / k - ck;
            const double ex = exp(-etak2 / (4.0 * q.xi * q.xi));
            const double j1 = t / k, j2 = (3.0 / k2 - 1.0) * sk / k - 3.0 * ck / k2;
            const double e2 = ex / k2 * q.invV;
            o[0] = h0;
            o[1] = h1;
            o[2] = h2;
            o[3] = 9.0 * t * t * ex / (k2 * k2) * q.invV;
            o[4] = -22.5 * j1 * j2 * e2;
            o[5] = 45.0 * j1 * j2 * e2;
            o[6] = 112.5 * j2 * j2 * e2;
            o[7] = h0 * h0 - 1.0 / 3.0;
            o[8] = h0 * h1;
            o[9] = h0 * h2;
            o[10] = h1 * h1 - 1.0 / 3.0;
            o[11] = h1 * h2;
            o[12] = h0 * h0 - h2 * h2;
            o[13] = 2.0 * h0 * h1;
            o[14] = 2.0 * h0 * h2;
            o[15] = h1 * h1 - h2 * h2;
        }
        // Qdot[4] = 2 h1 h2 does not fit the 16 slots: the scale kernel forms it from khat
        for (int c = 0; c < 16; ++c) {
            k64[16 * v + c] = o[c];
            k32[16 * v + c] = (float)o[c];
        }
    }
}

// Owner-computes spread of dipoles (field grid) and quadrupoles (gradient grid)
// with the Gaussian weights: a thread per voxel and G right-hand sides.
template <typename R, int G>
__global__ __launch_bounds__(256) void quad_spread_kernel(SlotTable tab, int B, int N, int V, int SS,
                                                          const int* __restrict__ vox_start,
                                                          const int* __restrict__ vox_items,
                                                          const R* __restrict__ sw, const int* __restrict__ qmap,
                                                          typename C2<R>::T* __restrict__ out) {
    using CT = typename C2<R>::T;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int g0 = blockIdx.y * G;
    R ar[G][8], ai[G][8];
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 8; ++c) ar[j][c] = ai[j][c] = R(0);
    const int e0 = vox_start[v], e1 = vox_start[v + 1];
    for (int e = e0; e < e1; ++e) {
        const int item = vox_items[e];
        const int pid = item / SS;
        const R w = sw[item];
        const int q = qmap[pid];
#pragma unroll
        for (int j = 0; j < G; ++j) {
            if (g0 + j < B) {
                const double2* x = tab.x[g0 + j];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 pv = x[3 * (size_t)pid + c];
                    ar[j][c] += w * (R)pv.x;
                    ai[j][c] += w * (R)pv.y;
                }
                if (q >= 0) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        const double2 qv = x[3 * (size_t)N + 5 * (size_t)q + c];
                        ar[j][3 + c] += w * (R)qv.x;
                        ai[j][3 + c] += w * (R)qv.y;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
        if (g0 + j < B) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                CT o;
                o.x = ar[j][c];
                o.y = ai[j][c];
                out[((size_t)(g0 + j) * 8 + c) * V + v] = o;
            }
        }
    }
}

// Joint k-space coupling of field and gradient grids (scale_kernel_joint,
// :1435-1487): Edk = E.khat, GQ = G.Qdot; E <- (sc Edk + i scQ GQ) khat,
// G <- (i scGP Edk + scGQ GQ) Qfactor.
template <typename R>
__global__ __launch_bounds__(256) void quad_scale_kernel(typename C2<R>::T* __restrict__ grid,
                                                         const R* __restrict__ kq, size_t V, size_t total) {
    using CT = typename C2<R>::T;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t b = e / V, v = e - b * V;
        CT* g = grid + b * 8 * V + v;
        const R* k = kq + 16 * v;
        const R h0 = k[0], h1 = k[1], h2 = k[2];
        const R qd[5] = {k[12], k[13], k[14], k[15], R(2) * h1 * h2};
        CT f[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) f[c] = g[c * V];
        const R er = f[0].x * h0 + f[1].x * h1 + f[2].x * h2, ei = f[0].y * h0 + f[1].y * h1 + f[2].y * h2;
        R gr = R(0), gi = R(0);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            gr += f[3 + c].x * qd[c];
            gi += f[3 + c].y * qd[c];
        }
        // i a (x + i y) = -a y + i a x
        const R sr = k[3] * er - k[4] * gi, si = k[3] * ei + k[4] * gr;
        const R tr = -k[5] * ei + k[6] * gr, ti = k[5] * er + k[6] * gi;
        const R hh[3] = {h0, h1, h2};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            CT o;
            o.x = sr * hh[c];
            o.y = si * hh[c];
            g[c * V] = o;
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            CT o;
            o.x = tr * k[7 + c];
            o.y = ti * k[7 + c];
            g[(3 + c) * V] = o;
        }
    }
}

// A wave per (particle i, group of G right-hand sides): reciprocal field and,
// for a quadrupole particle, gradient (h^3 sum of the Gaussian weights times the
// grids), the real-space couplings of every neighbour's dipole and quadrupole
// (dipole -> field, quadrupole -> field, dipole -> gradient, quadrupole ->
// gradient, :1631-1869 with the cross-term signs of tests/quadrupole_field.py:214-215),
// the self terms (sc + self_perp for the field, 2.5 sc + self_grad / 2 for the
// gradient) and the optional left Jacobi scaling; FP64 sums, one reduction.
// Pair data [16]: u (3), perp, para - perp, cos, sin, field_quad_1..3, grad_quad_1..4.
template <typename R, int G>
__global__ __launch_bounds__(256) void quad_gather_kernel(SlotTable tab, int B, int parts, int N, int V, int SS,
                                                          const typename C2<R>::T* __restrict__ grid,
                                                          const R* __restrict__ sw, const int* __restrict__ pidx,
                                                          const int* __restrict__ qmap, double h3,
                                                          const int* __restrict__ nb_start,
                                                          const int* __restrict__ nb_j,
                                                          const double* __restrict__ pairq,
                                                          const double2* __restrict__ selfc, double self_perp,
                                                          double self_grad, int jacobi,
                                                          const int* __restrict__ order) {
    using CT = typename C2<R>::T;
    const int lane = threadIdx.x & 63;
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= N) return;  // whole wave
    const int i = order[wv];  // spatial order: neighbouring waves share stencil voxels
    const int g0 = blockIdx.y * G;
    const int qi = qmap[i];
    const int nc = qi >= 0 ? 8 : 3;
    double acc[G][8][2];
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[j][c][0] = acc[j][c][1] = 0.0;
    if (parts & MPM_PART_RECIP) {
        for (int o = lane; o < SS; o += 64) {
            const size_t item = (size_t)i * SS + o;
            const int v = pidx[item];
            const double w = (double)sw[item] * h3;
#pragma unroll
            for (int j = 0; j < G; ++j) {
                if (g0 + j < B) {
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        if (c < nc) {
                            const CT gv = grid[((size_t)(g0 + j) * 8 + c) * V + v];
                            acc[j][c][0] += w * (double)gv.x;
                            acc[j][c][1] += w * (double)gv.y;
                        }
                    }
                }
            }
        }
    }
    if (parts & MPM_PART_REAL) {
        const int k0 = nb_start[i], k1 = nb_start[i + 1];
        const double I5[5] = {1.0, 0.0, 0.0, 1.0, 0.0};
        for (int k = k0 + lane; k < k1; k += 64) {
            const int jn = nb_j[k];
            const int qs = qmap[jn];
            const double* pc = pairq + 16 * (size_t)k;
            const double u[3] = {pc[0], pc[1], pc[2]};
            const double cp = pc[3], cd = pc[4], ec = pc[5], es = pc[6];
            const double t0 = pc[7], t1 = pc[8], t2 = pc[9], g1 = pc[10], g2 = pc[11], g3 = pc[12], g4 = pc[13];
            const double rr[5] = {u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2]};
#pragma unroll
            for (int j = 0; j < G; ++j) {
                if (g0 + j < B) {
                    const double2* x = tab.x[g0 + j];
                    const double2 pp[3] = {x[3 * (size_t)jn], x[3 * (size_t)jn + 1], x[3 * (size_t)jn + 2]};
                    double f[8][2];
                    const double rpr = pp[0].x * u[0] + pp[1].x * u[1] + pp[2].x * u[2];
                    const double rpi = pp[0].y * u[0] + pp[1].y * u[1] + pp[2].y * u[2];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        f[c][0] = cp * pp[c].x + cd * rpr * u[c];
                        f[c][1] = cp * pp[c].y + cd * rpi * u[c];
                    }
#pragma unroll
                    for (int c = 3; c < 8; ++c) f[c][0] = f[c][1] = 0.0;
                    double Qr[3][2] = {{0, 0}, {0, 0}, {0, 0}}, Qrr[2] = {0, 0};
                    double2 qq[5];
                    if (qs >= 0) {
#pragma unroll
                        for (int c = 0; c < 5; ++c) qq[c] = x[3 * (size_t)N + 5 * (size_t)qs + c];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const double q0 = h ? qq[0].y : qq[0].x, q1 = h ? qq[1].y : qq[1].x,
                                         q2 = h ? qq[2].y : qq[2].x, q3 = h ? qq[3].y : qq[3].x,
                                         q4 = h ? qq[4].y : qq[4].x;
                            Qr[0][h] = q0 * u[0] + q1 * u[1] + q2 * u[2];
                            Qr[1][h] = q1 * u[0] + q3 * u[1] + q4 * u[2];
                            Qr[2][h] = q2 * u[0] + q4 * u[1] - (q0 + q3) * u[2];
                            Qrr[h] = Qr[0][h] * u[0] + Qr[1][h] * u[1] + Qr[2][h] * u[2];
#pragma unroll
                            for (int c = 0; c < 3; ++c) f[c][h] += 0.5 * (Qrr[h] * t0 * u[c] + 2.0 * t1 * Qr[c][h]);
                        }
                    }
                    if (qi >= 0) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const double rp = h ? rpi : rpr;
                            const double p0 = h ? pp[0].y : pp[0].x, p1 = h ? pp[1].y : pp[1].x,
                                         p2 = h ? pp[2].y : pp[2].x;
                            const double PrrP[5] = {2.0 * p0 * u[0], p0 * u[1] + p1 * u[0], p0 * u[2] + p2 * u[0],
                                                    2.0 * p1 * u[1], p1 * u[2] + p2 * u[1]};
#pragma unroll
                            for (int c = 0; c < 5; ++c)
                                f[3 + c][h] = -(rp * (t0 * rr[c] + (t1 + t2) * I5[c]) + PrrP[c] * t1);
                            if (qs >= 0) {
                                const double qv[5] = {h ? qq[0].y : qq[0].x, h ? qq[1].y : qq[1].x, h ? qq[2].y : qq[2].x,
                                                      h ? qq[3].y : qq[3].x, h ? qq[4].y : qq[4].x};
                                const double QrrQ[5] = {2.0 * Qr[0][h] * u[0], Qr[0][h] * u[1] + Qr[1][h] * u[0],
                                                        Qr[0][h] * u[2] + Qr[2][h] * u[0], 2.0 * Qr[1][h] * u[1],
                                                        Qr[1][h] * u[2] + Qr[2][h] * u[1]};
#pragma unroll
                                for (int c = 0; c < 5; ++c)
                                    f[3 + c][h] += 0.5 * (Qrr[h] * ((g1 + g3) * I5[c] + g4 * rr[c]) + qv[c] * g2 +
                                                          2.0 * g3 * QrrQ[c]);
                            }
                        }
                    }
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        acc[j][c][0] += f[c][0] * ec - f[c][1] * es;
                        acc[j][c][1] += f[c][0] * es + f[c][1] * ec;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            acc[j][c][0] = wave_sum(acc[j][c][0]);
            acc[j][c][1] = wave_sum(acc[j][c][1]);
        }
    if (lane >= G * 8) return;
    const int j = lane / 8, c = lane % 8;  // lane (j, c) stores component c of right-hand side g0 + j
    if (g0 + j >= B || c >= nc) return;
    double yr = 0.0, yi = 0.0;
    const double2* xb = nullptr;
    double2* yb = nullptr;
    int row = 0;
#pragma unroll
    for (int jj = 0; jj < G; ++jj) {  // uniform indices into the argument, selected per lane
        if (jj == j) {
            xb = tab.x[g0 + jj];
            yb = tab.y[g0 + jj];
            row = tab.row[g0 + jj];
        }
#pragma unroll
        for (int cc = 0; cc < 8; ++cc)
            if (jj == j && cc == c) {
                yr = acc[jj][cc][0];
                yi = acc[jj][cc][1];
            }
    }
    const size_t e = c < 3 ? 3 * (size_t)i + c : 3 * (size_t)N + 5 * (size_t)qi + (c - 3);
    const double2 xv = xb[e];
    const double2 sc = selfc[(size_t)row * N + i];
    if (parts & MPM_PART_REAL) {
        const double fr = c < 3 ? sc.x + self_perp : 2.5 * sc.x + 0.5 * self_grad;
        const double fi = c < 3 ? sc.y : 2.5 * sc.y;
        yr += fr * xv.x - fi * xv.y;
        yi += fr * xv.y + fi * xv.x;
    }
    double2 inv = make_double2(1.0, 0.0);
    if (jacobi) {
        const double den = sc.x * sc.x + sc.y * sc.y;
        if (den > 1e-30) inv = make_double2(sc.x / den, -sc.y / den);
    }
    yb[e] = make_double2(yr * inv.x - yi * inv.y, yr * inv.y + yi * inv.x);
}

// ---------------------------------------------------------------------------
// Direct (open-boundary) field in the solver's mode, after
// src/cpp/electric_field/direct_electric_field.cu:37-300 (direct_field_kernel,
// SOLVER_AX).  A thread per target particle and G right-hand sides; the sources
// pass through LDS in tiles of kDirectTile (positions, radii, dipoles and
// quadrupoles of the G systems), so each source's geometry (r, u, 1/r^n) is
// computed once for all G systems.  For a source i != j: r = x_i - x_j, clamped to
// the source radius when shorter (a vanishing r keeps no direction), the point
// dipole field c (3 (p.u) u - p) / r^3 with c = -1/(4 pi); with quadrupoles the
// point quadrupole field (3c / r^4)(2.5 s u - v), v = Qm u, s = u.v, Qm =
// -[[Q0, Q1, Q2], [Q1, Q3, Q4], [Q2, Q4, -Q0 - Q3]] in the Ewald operator's
// components, and at quadrupole targets the field gradients (d_x E_x, d_x E_y,
// d_x E_z, d_y E_y, d_y E_z): of a dipole (-3c / r^4)((p.u)(I - 5 u u) + p u^T +
// u p^T), of a quadrupole (-3c / r^5)(5 (v u^T + u v^T) + 2.5 s I - 17.5 s u u -
// Qm) (the derivative of its field; the reference kernel doubles the off-diagonal
// source components and differs in sign here).  Self terms (sc + 1/(4 pi a^3)) p
// and (2.5 sc / a^2 + 3/(4 pi a^5)) Q in FP64; left Jacobi divides by sc.
// ---------------------------------------------------------------------------
constexpr int kDirectTile = 128;

template <typename R, int G, bool QUADS>
__global__ __launch_bounds__(256) void direct_kernel(SlotTable tab, int B, int N, const double* __restrict__ pos,
                                                     const double* __restrict__ rad, const int* __restrict__ qmap,
                                                     const double2* __restrict__ selfc, int jacobi) {
    constexpr int T = kDirectTile;
    constexpr int QT = QUADS ? T : 1;
    __shared__ R sx[T], sy[T], sz[T], sa[T];
    __shared__ R sp[G][3][2][T];
    __shared__ R sq[QUADS ? G : 1][5][2][QT];
    __shared__ int sqm[QT];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * blockDim.x + tid;
    const int g0 = blockIdx.y * G;
    const bool act = j < N;
    R tx = 0, ty = 0, tz = 0;
    int qt = -1;
    if (act) {
        tx = (R)pos[j];
        ty = (R)pos[(size_t)N + j];
        tz = (R)pos[2 * (size_t)N + j];
        if (QUADS) qt = qmap[j];
    }
    R E[G][3][2];
    R Gq[G][5][2];
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int c = 0; c < 3; ++c) E[g][c][0] = E[g][c][1] = R(0);
#pragma unroll
        for (int c = 0; c < 5; ++c) Gq[g][c][0] = Gq[g][c][1] = R(0);
    }
    const R cf = R(-1.0 / (4.0 * kPi));
    const R eps_orig = sizeof(R) == 4 ? R(1e-6) : R(1e-9);
    const R eps_sing = sizeof(R) == 4 ? R(1e-12) : R(1e-18);
    for (int t0 = 0; t0 < N; t0 += T) {
        __syncthreads();
        if (tid < T) {
            const int src = t0 + tid;
            const bool ok = src < N;
            sx[tid] = ok ? (R)pos[src] : R(0);
            sy[tid] = ok ? (R)pos[(size_t)N + src] : R(0);
            sz[tid] = ok ? (R)pos[2 * (size_t)N + src] : R(0);
            sa[tid] = ok ? (R)rad[src] : R(0);
            const int qs = (QUADS && ok) ? qmap[src] : -1;
            if (QUADS) sqm[tid] = qs;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const bool has = ok && g0 + g < B;
                const double2* x = has ? tab.x[g0 + g] : nullptr;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double2 v = has ? x[3 * (size_t)src + c] : make_double2(0.0, 0.0);
                    sp[g][c][0][tid] = (R)v.x;
                    sp[g][c][1][tid] = (R)v.y;
                }
                if (QUADS) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        const double2 v = (has && qs >= 0) ? x[3 * (size_t)N + 5 * (size_t)qs + c] : make_double2(0.0, 0.0);
                        sq[g][c][0][tid] = (R)v.x;
                        sq[g][c][1][tid] = (R)v.y;
                    }
                }
            }
        }
        __syncthreads();
        if (!act) continue;
        const int cnt = min(T, N - t0);
        for (int k = 0; k < cnt; ++k) {
            if (t0 + k == j) continue;
            R rx = sx[k] - tx, ry = sy[k] - ty, rz = sz[k] - tz;
            R r2 = rx * rx + ry * ry + rz * rz;
            const R ai = sa[k];
            if (r2 < ai * ai) {
                const R r0 = sqrt(r2);
                if (r0 > eps_orig) {
                    const R f = ai / r0;
                    rx *= f;
                    ry *= f;
                    rz *= f;
                }
                r2 = ai * ai;
            }
            if (r2 < eps_sing) continue;
            const R d = sqrt(r2);
            const R id = R(1) / d;
            const R u[3] = {rx * id, ry * id, rz * id};
            const R id3 = id * id * id;
            const R fd = cf * id3;
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const R p[3] = {sp[g][0][h][k], sp[g][1][h][k], sp[g][2][h][k]};
                    const R pu = p[0] * u[0] + p[1] * u[1] + p[2] * u[2];
#pragma unroll
                    for (int c = 0; c < 3; ++c) E[g][c][h] += fd * (R(3) * pu * u[c] - p[c]);
                }
            if (QUADS) {
                const int qs = sqm[k];
                const R id4 = id3 * id, id5 = id4 * id;
                const R rr[5] = {u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2]};
                const R I5[5] = {R(1), R(0), R(0), R(1), R(0)};
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (qs >= 0) {
                            const R q0 = sq[g][0][h][k], q1 = sq[g][1][h][k], q2 = sq[g][2][h][k],
                                    q3 = sq[g][3][h][k], q4 = sq[g][4][h][k];
                            // v = Qm u with Qm = -[[q0, q1, q2], [q1, q3, q4], [q2, q4, -q0 - q3]]
                            const R v[3] = {-(q0 * u[0] + q1 * u[1] + q2 * u[2]),
                                            -(q1 * u[0] + q3 * u[1] + q4 * u[2]),
                                            -(q2 * u[0] + q4 * u[1] - (q0 + q3) * u[2])};
                            const R sv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
                            const R fq = R(3) * cf * id4;
#pragma unroll
                            for (int c = 0; c < 3; ++c) E[g][c][h] += fq * (R(2.5) * sv * u[c] - v[c]);
                            if (qt >= 0) {
                                const R vu[5] = {R(2) * v[0] * u[0], v[0] * u[1] + v[1] * u[0], v[0] * u[2] + v[2] * u[0],
                                                 R(2) * v[1] * u[1], v[1] * u[2] + v[2] * u[1]};
                                const R Qc[5] = {-q0, -q1, -q2, -q3, -q4};
                                const R fg = R(-3) * cf * id5;
#pragma unroll
                                for (int c = 0; c < 5; ++c)
                                    Gq[g][c][h] += fg * (R(5) * vu[c] + R(2.5) * sv * I5[c] - R(17.5) * sv * rr[c] - Qc[c]);
                            }
                        }
                        if (qt >= 0) {
                            const R p[3] = {sp[g][0][h][k], sp[g][1][h][k], sp[g][2][h][k]};
                            const R pu = p[0] * u[0] + p[1] * u[1] + p[2] * u[2];
                            const R pv[5] = {R(2) * p[0] * u[0], p[0] * u[1] + p[1] * u[0], p[0] * u[2] + p[2] * u[0],
                                             R(2) * p[1] * u[1], p[1] * u[2] + p[2] * u[1]};
                            const R fg = R(-3) * cf * id4;
#pragma unroll
                            for (int c = 0; c < 5; ++c) Gq[g][c][h] += fg * (pu * (I5[c] - R(5) * rr[c]) + pv[c]);
                        }
                    }
            }
        }
    }
    if (!act) return;
    const double aj = rad[j];
    const double a3 = aj * aj * aj;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if (g0 + g >= B) break;
        const double2* x = tab.x[g0 + g];
        double2* y = tab.y[g0 + g];
        const double2 sc = selfc[(size_t)tab.row[g0 + g] * N + j];
        double2 inv = make_double2(1.0, 0.0);
        if (jacobi) {
            const double den = sc.x * sc.x + sc.y * sc.y;
            if (den > 1e-30) inv = make_double2(sc.x / den, -sc.y / den);
        }
        const double sr = sc.x + 1.0 / (4.0 * kPi * a3), si = sc.y;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 xv = x[3 * (size_t)j + c];
            const double yr = (double)E[g][c][0] + sr * xv.x - si * xv.y;
            const double yi = (double)E[g][c][1] + sr * xv.y + si * xv.x;
            y[3 * (size_t)j + c] = make_double2(yr * inv.x - yi * inv.y, yr * inv.y + yi * inv.x);
        }
        if (QUADS && qt >= 0) {
            const double qr = 2.5 * sc.x / (aj * aj) + 3.0 / (4.0 * kPi * a3 * aj * aj), qi = 2.5 * sc.y / (aj * aj);
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                const size_t e = 3 * (size_t)N + 5 * (size_t)qt + c;
                const double2 xv = x[e];
                const double yr = (double)Gq[g][c][0] + qr * xv.x - qi * xv.y;
                const double yi = (double)Gq[g][c][1] + qr * xv.y + qi * xv.x;
                y[e] = make_double2(yr * inv.x - yi * inv.y, yr * inv.y + yi * inv.x);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host threads for the per-frame set-up
// ---------------------------------------------------------------------------
// Threads for n items with at least `grain` items each: up to 16 (one GPU's share of the
// host) and the machine's hardware threads.
static int host_threads(int n, int grain) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    return std::max(1, std::min({hw, 16, n / std::max(1, grain)}));
}
// First item of chunk c of n items split into T contiguous chunks.
static int chunk_begin(int n, int T, int c) { return (int)((int64_t)n * c / T); }
// f(begin, end, chunk) on T threads over contiguous chunks of [0, n); exceptions thrown by
// any chunk are rethrown after every thread has finished.
template <typename F>
static void parallel_for(int n, int T, F&& f) {
    if (T <= 1) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(T);
    th.reserve(T);
    for (int c = 0; c < T; ++c)
        th.emplace_back([&, c] {
            try {
                f(chunk_begin(n, T, c), chunk_begin(n, T, c + 1), c);
            } catch (...) {
                err[c] = std::current_exception();
            }
        });
    for (auto& t : th) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// ---------------------------------------------------------------------------
// EwaldOp
// ---------------------------------------------------------------------------
EwaldOp::EwaldOp(const mpm_op_desc& d) {
    for (int k = 0; k < 3; ++k) L[k] = d.box[k];
    errortol = d.errortol;
    xi = d.xi;
    compute_sizing(L, errortol, xi, d.grid, d.support, S);
    MPM_REQUIRE(d.layout == 0 || d.layout == 1, MPM_ERR_INVALID_ARGUMENT, "layout must be 0 or 1");
    force_full = d.layout;
    MPM_REQUIRE(d.fft_backend == 0 || d.fft_backend == 1, MPM_ERR_INVALID_ARGUMENT, "fft_backend must be 0 or 1");
    fft_backend = d.fft_backend;
    MPM_REQUIRE(S.width <= kMaxWidth, MPM_ERR_UNSUPPORTED,
                "stencil wider than " + std::to_string(kMaxWidth) + " points per axis");
    self_perp = dipole_self_term(xi);
    V = (size_t)S.n[0] * S.n[1] * S.n[2];
    table = std::make_unique<PairTable>(xi);
    build_kspace();
}

EwaldOp::EwaldOp(DirectTag, const double* radii, int64_t nr) {
    MPM_REQUIRE(radii && nr > 0, MPM_ERR_INVALID_ARGUMENT, "no radii");
    for (int64_t i = 0; i < nr; ++i)
        MPM_REQUIRE(std::isfinite(radii[i]) && radii[i] > 0.0, MPM_ERR_INVALID_ARGUMENT, "radii must be positive");
    direct = true;
    rad_h.assign(radii, radii + nr);
    L[0] = L[1] = L[2] = 0.0;
    errortol = 0.0;
    xi = 0.0;
    V = 0;
}

EwaldOp::EwaldOp(const mpm_op_desc& d, const double* radii, int64_t nr) : EwaldOp(d) {
    MPM_REQUIRE(radii && nr > 0, MPM_ERR_INVALID_ARGUMENT, "no radii");
    MPM_REQUIRE(d.layout == 0 || d.layout == 1, MPM_ERR_INVALID_ARGUMENT, "layout must be 0 or 1");
    MPM_REQUIRE(d.fft_backend == 0, MPM_ERR_UNSUPPORTED,
                "the polydisperse operator runs on the built-in FFT only (fft_backend 0)");
    poly = true;
    force_full = 1;
    rad_h.assign(radii, radii + nr);
    // radius classes: merged within 1e-5 in input order (polydisperse...cu:40-83)
    for (double r : rad_h) {
        MPM_REQUIRE(std::isfinite(r) && r > 0.0, MPM_ERR_INVALID_ARGUMENT, "radii must be positive");
        if (poly_exact) continue;
        bool seen = false;
        for (double u : uniq_h) seen = seen || std::abs(u - r) < 1e-5;
        if (!seen) uniq_h.push_back(r);
        poly_exact = uniq_h.size() > (size_t)kPolyTableClasses;
    }
    const double amax = *std::max_element(rad_h.begin(), rad_h.end());
    compute_sizing_poly(L, errortol, xi, amax, d.grid, d.support, S);
    peta = S.eta[0];
    poff = -S.half;
    V = (size_t)S.n[0] * S.n[1] * S.n[2];
    if (poly_exact) {
        // U (U + 1) / 2 tables of ~(rc - 1) / 0.001 samples each would not fit; evaluate
        // every pair at its own radii instead, and count the classes for the record
        std::vector<double> sr(rad_h);
        std::sort(sr.begin(), sr.end());
        poly_classes = 1;
        for (size_t i = 1, k = 0; i < sr.size(); ++i)
            if (sr[i] - sr[k] >= 1e-5) {
                ++poly_classes;
                k = i;
            }
        uniq_h.clear();
        pself_h.resize(nr);
        for (int64_t i = 0; i < nr; ++i) pself_h[i] = poly_self_term(rad_h[i], rad_h[i], xi);
        self_perp = poly_self_term(sr[0], sr[0], xi);
        build_kspace();
        return;
    }
    std::sort(uniq_h.begin(), uniq_h.end());
    poly_classes = (int)uniq_h.size();
    ridx_h.assign(nr, 0);
    for (int64_t i = 0; i < nr; ++i)
        for (size_t u = 0; u < uniq_h.size(); ++u)
            if (std::abs(uniq_h[u] - rad_h[i]) < 1e-5) {
                ridx_h[i] = (int)u;
                break;
            }
    // tables of every radius pair: entry 0 the pair's self value, then r = 1 + 0.001 k up to rc
    const double rcv = S.rc < 1.001 ? 1.001 : S.rc;
    const size_t steps = (size_t)std::ceil((rcv - 1.0) / 0.001);
    const int U = (int)uniq_h.size();
    pt_ncol = U * (U + 1) / 2;
    pt_r.assign(steps + 1, 0.0);
    for (size_t k = 0; k < steps; ++k) pt_r[k + 1] = 1.0 + k * 0.001;
    pt_perp.assign((steps + 1) * pt_ncol, 0.0);
    pt_para.assign((steps + 1) * pt_ncol, 0.0);
    uself_h.assign(U, 0.0);
    int col = 0;
    for (int u = 0; u < U; ++u)
        for (int v = u; v < U; ++v, ++col) {
            const double a = uniq_h[u], b = uniq_h[v];
            const double s0 = poly_self_term(a, b, xi);
            pt_perp[col] = pt_para[col] = s0;
            if (u == v) uself_h[u] = s0;
            for (size_t k = 0; k < steps; ++k)
                poly_pair_functions(pt_r[k + 1], a, b, xi, pt_perp[(k + 1) * pt_ncol + col],
                                    pt_para[(k + 1) * pt_ncol + col]);
        }
    self_perp = uself_h[0];
    pself_h.resize(nr);
    for (int64_t i = 0; i < nr; ++i) pself_h[i] = uself_h[ridx_h[i]];
    build_kspace();
}

// Pair values of particles i, j at distance r (>= 1e-6): the radius classes' tables, or
// with more classes than tables allow the closed forms at the pair's own radii, with
// the tables' rule below r = 1 (linear between the r -> 0 value and the value at 1).
void EwaldOp::poly_pair_value(double r, int i, int j, double& perp, double& para) const {
    if (!poly_exact) {
        perp = poly_lookup(pt_perp, r, ridx_h[i], ridx_h[j]);
        para = poly_lookup(pt_para, r, ridx_h[i], ridx_h[j]);
        return;
    }
    const double a = rad_h[i], b = rad_h[j];
    if (r >= 1.0) {
        poly_pair_functions(r, a, b, xi, perp, para);
        return;
    }
    const double s0 = poly_self_term(a, b, xi);
    double p1, q1;
    poly_pair_functions(1.0, a, b, xi, p1, q1);
    perp = s0 * (1.0 - r) + p1 * r;
    para = s0 * (1.0 - r) + q1 * r;
}

// Pair table `which` (0 perp, 1 para, 2..8 the quadrupole tables) on its samples r; the
// table's length is returned and the arrays written when cap allows.
int64_t EwaldOp::pair_table(int which, double* r, double* v, int64_t cap) const {
    MPM_REQUIRE(!direct, MPM_ERR_UNSUPPORTED, "the direct field has no real-space tables");
    MPM_REQUIRE(!poly, MPM_ERR_UNSUPPORTED, "the polydisperse operator keeps one table per radius pair");
    MPM_REQUIRE(which >= 0 && which <= 8, MPM_ERR_INVALID_ARGUMENT, "table index out of range");
    const std::vector<double>& t = which == 0 ? table->perp : which == 1 ? table->para : table->q[which - 2];
    MPM_REQUIRE(!t.empty(), MPM_ERR_RUNTIME, "quadrupole tables exist after mpm_op_set_quads");
    const int64_t n = (int64_t)table->r.size();
    if (cap >= n) {
        if (r) std::copy(table->r.begin(), table->r.end(), r);
        if (v) std::copy(t.begin(), t.end(), v);
    }
    return n;
}

// Neighbour list of the current positions (CSR over particles, j != i, minimum-image
// distance below rc) and each pair's (perp, para) values; the pair count is returned
// and the arrays written when cap allows.
int64_t EwaldOp::neighbors(int64_t* start, int* j, double* perp, double* para, int64_t cap) const {
    MPM_REQUIRE(!direct, MPM_ERR_UNSUPPORTED, "the direct field has no neighbour list");
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "particle coordinates have not been set");
    fetch_pairs();
    if (cap >= num_pairs) {
        if (start)
            for (size_t i = 0; i < h_nb_start.size(); ++i) start[i] = h_nb_start[i];
        for (int64_t k = 0; k < num_pairs; ++k) {
            if (j) j[k] = h_nb_j[k];
            if (perp) perp[k] = h_pair_f[k].x;
            if (para) para[k] = h_pair_f[k].y;
        }
    }
    return num_pairs;
}

// Table value of radius classes (u, v) at distance r: pair column of the upper
// triangle (row-major), the monodisperse interpolation rule (clamped at the last
// sample).  polydisperse...cu:56-68, 168-188.
double EwaldOp::poly_lookup(const std::vector<double>& tab, double r, int u, int v) const {
    if (u > v) std::swap(u, v);
    const int U = (int)uniq_h.size();
    const int col = u * U - u * (u - 1) / 2 + (v - u);
    const size_t ts = pt_r.size();
    auto y = [&](size_t e) { return tab[e * pt_ncol + col]; };
    if (r <= 0.0) return y(0);
    if (r >= pt_r[ts - 1]) return y(ts - 1);
    if (r < 1.0) return y(0) * (1.0 - r) + y(1) * r;
    const size_t idx = 1 + (size_t)((r - 1.0) / 0.001);
    if (idx >= ts - 1) return y(ts - 1);
    const double t = (r - pt_r[idx]) / (pt_r[idx + 1] - pt_r[idx]);
    return y(idx) * (1.0 - t) + y(idx + 1) * t;
}

// Per-frame stencils of the polydisperse operator: the P^3 points of every
// particle (x-major offsets from -floor((P-1)/2)), their wrapped voxels and
// weight vectors, and the voxel -> entry lists of the owner-computes spread.
void EwaldOp::setup_poly_stencils(const std::vector<double>& pos, int Nn) {
    const int P = S.P;
    const int SS = P * P * P;
    MPM_REQUIRE((int64_t)Nn * SS < (1LL << 31), MPM_ERR_UNSUPPORTED, "too many stencil points");
    std::vector<double> pw((size_t)Nn * SS * 3);
    std::vector<int> pidx((size_t)Nn * SS);
    std::vector<int> cnt(V + 1, 0);
    pst_g.resize(3 * (size_t)Nn);
    pst_dd.resize(3 * (size_t)Nn);
    for (int i = 0; i < Nn; ++i) {
        int g[3];
        double dd[3];
        for (int d = 0; d < 3; ++d) {
            const double p = pos[(size_t)d * Nn + i];
            g[d] = (int)std::round(p / S.h[d]);
            dd[d] = g[d] * S.h[d] - p;
            pst_g[3 * (size_t)i + d] = g[d];
            pst_dd[3 * (size_t)i + d] = dd[d];
        }
        int o = 0;
        for (int ox = poff; ox < poff + P; ++ox)
            for (int oy = poff; oy < poff + P; ++oy)
                for (int oz = poff; oz < poff + P; ++oz, ++o) {
                    const int off[3] = {ox, oy, oz};
                    int w[3];
                    double gd[3];
                    for (int d = 0; d < 3; ++d) {
                        w[d] = ((g[d] + off[d]) % S.n[d] + S.n[d]) % S.n[d];
                        gd[d] = dd[d] + off[d] * S.h[d];
                    }
                    const size_t item = (size_t)i * SS + o;
                    const int v = (w[0] * S.n[1] + w[1]) * S.n[2] + w[2];
                    pidx[item] = v;
                    poly_weight(gd, wrad_h.empty() ? rad_h[i] : wrad_h[i], xi, peta, &pw[3 * item]);
                    cnt[v + 1]++;
                }
    }
    for (size_t v = 0; v < V; ++v) cnt[v + 1] += cnt[v];
    std::vector<int> items((size_t)Nn * SS);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (size_t item = 0; item < (size_t)Nn * SS; ++item) items[fill[pidx[item]]++] = (int)item;
    }
    std::vector<float> pw32(pw.begin(), pw.end());
    std::vector<double> sp(Nn);
    for (int i = 0; i < Nn; ++i) sp[i] = pself_h[i];
    d_pw64.upload(pw.data(), pw.size());
    d_pw32.upload(pw32.data(), pw32.size());
    d_pidx.upload(pidx.data(), pidx.size());
    d_vox_start.upload(cnt.data(), cnt.size());
    d_vox_items.upload(items.data(), items.size());
    d_selfp.upload(sp.data(), sp.size());
    ++pst_serial;
}

void EwaldOp::set_weight_radii(const double* r, int64_t n) {
    MPM_REQUIRE(poly, MPM_ERR_UNSUPPORTED, "weight radii belong to the polydisperse operator");
    MPM_REQUIRE(r && n == (int64_t)rad_h.size(), MPM_ERR_INVALID_ARGUMENT, "one weight radius per site");
    for (int64_t i = 0; i < n; ++i)
        MPM_REQUIRE(std::isfinite(r[i]) && r[i] > 0.0, MPM_ERR_INVALID_ARGUMENT, "radii must be positive");
    wrad_h.assign(r, r + n);
    last_xyz.clear();  // the next set_positions rebuilds the stencils
}

// d/da_i of the pair values of particles i, j at distance r (>= 1e-6), from the closed
// forms at the pair's own radii (the limit of the radius-class tables), with the rule
// below r = 1 (linear between the r -> 0 value and the value at 1).
void EwaldOp::poly_pair_dvalue(double r, int i, int j, double& dperp, double& dpara) const {
    const Dual a(rad_h[i], 1.0), b(rad_h[j], 0.0);
    Dual perp, para;
    if (r >= 1.0) {
        poly_pair_functions<Dual>(Dual(r), a, b, xi, perp, para);
    } else {
        const Dual s0 = poly_self_term<Dual>(a, b, xi);
        Dual p1, q1;
        poly_pair_functions<Dual>(Dual(1.0), a, b, xi, p1, q1);
        perp = s0 * (1.0 - r) + p1 * r;
        para = s0 * (1.0 - r) + q1 * r;
    }
    dperp = perp.d;
    dpara = para.d;
}

// Radius derivatives of the current frame: every stencil point's weight vector, every
// neighbour pair's coefficients (for the row's own radius) and every real-space self term.
void EwaldOp::ensure_rgrad_data() {
    if (rg_serial == pst_serial) return;
    const int Nn = (int)N;
    const int P = S.P, SS = P * P * P;
    std::vector<double> dpw((size_t)Nn * SS * 3);
    for (int i = 0; i < Nn; ++i) {
        const Dual a(rad_h[i], 1.0);
        int o = 0;
        for (int ox = poff; ox < poff + P; ++ox)
            for (int oy = poff; oy < poff + P; ++oy)
                for (int oz = poff; oz < poff + P; ++oz, ++o) {
                    const int off[3] = {ox, oy, oz};
                    double gd[3];
                    for (int d = 0; d < 3; ++d) gd[d] = pst_dd[3 * (size_t)i + d] + off[d] * S.h[d];
                    Dual w[3];
                    poly_weight<Dual>(gd, a, xi, peta, w);
                    for (int q = 0; q < 3; ++q) dpw[((size_t)i * SS + o) * 3 + q] = w[q].d;
                }
    }
    const size_t M = h_nb_j.size();
    std::vector<double2> dpair(std::max<size_t>(M, 1), make_double2(0.0, 0.0));
    for (int i = 0; i < Nn; ++i)
        for (int k = h_nb_start[i]; k < h_nb_start[i + 1]; ++k) {
            const double* r = &h_pair_r[3 * (size_t)k];
            const double d = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            double dp, dq;
            poly_pair_dvalue(d < 1e-6 ? 1e-6 : d, i, h_nb_j[k], dp, dq);
            dpair[k] = make_double2(dp, dq - dp);
        }
    std::vector<double> dps(Nn);
    for (int i = 0; i < Nn; ++i) {
        const Dual a(rad_h[i], 1.0);
        dps[i] = poly_self_term<Dual>(a, a, xi).d;
    }
    d_dpw.upload(dpw.data(), dpw.size());
    d_dpair.upload(dpair.data(), dpair.size());
    d_dps.upload(dps.data(), dps.size());
    rg_serial = pst_serial;
}

void EwaldOp::set_quads(const int* idx, int64_t nq) {
    MPM_REQUIRE(!poly, MPM_ERR_UNSUPPORTED, "quadrupoles on the polydisperse operator are not provided by this library");
    MPM_REQUIRE(N == 0, MPM_ERR_RUNTIME, "set quadrupoles before the particle positions");
    MPM_REQUIRE(nq < 0 || idx || nq == 0, MPM_ERR_INVALID_ARGUMENT, "null quadrupole indices");
    quads = true;
    quads_all = nq < 0;
    qreq_h.assign(idx && nq > 0 ? idx : nullptr, idx && nq > 0 ? idx + nq : nullptr);
    if (direct) return;
    force_full = 1;  // the slab kernels carry dipoles only
    table->build_quads(xi);
    build_kspace();
}

// Per-frame stencils with quadrupoles: the W^3 points of every particle, their
// wrapped voxels and Gaussian weights C exp(-2 xi^2 sum gd^2/eta) (the dipole
// weights, monodisperse...cu:687-816), and the voxel -> entry lists.
void EwaldOp::setup_quad_stencils(const std::vector<double>& pos, int Nn) {
    const int W = S.width, half = S.half;
    const int SS = W * W * W;
    MPM_REQUIRE((int64_t)Nn * SS < (1LL << 31), MPM_ERR_UNSUPPORTED, "too many stencil points");
    const double prod_eta = S.eta[0] * S.eta[1] * S.eta[2];
    const double Cs = std::pow(2.0 * xi * xi / kPi, 1.5) * std::sqrt(1.0 / prod_eta);
    std::vector<double> sw((size_t)Nn * SS);
    std::vector<int> pidx((size_t)Nn * SS), cnt(V + 1, 0);
    for (int i = 0; i < Nn; ++i) {
        int g[3];
        double dd[3];
        for (int d = 0; d < 3; ++d) {
            const double p = pos[(size_t)d * Nn + i];
            g[d] = (int)std::round(p / S.h[d]);
            dd[d] = g[d] * S.h[d] - p;
        }
        int o = 0;
        for (int ox = -half; ox <= half; ++ox)
            for (int oy = -half; oy <= half; ++oy)
                for (int oz = -half; oz <= half; ++oz, ++o) {
                    const int off[3] = {ox, oy, oz};
                    double a = 0.0;
                    int w[3];
                    for (int d = 0; d < 3; ++d) {
                        w[d] = ((g[d] + off[d]) % S.n[d] + S.n[d]) % S.n[d];
                        const double gd = dd[d] + off[d] * S.h[d];
                        a += gd * gd / S.eta[d];
                    }
                    const size_t item = (size_t)i * SS + o;
                    const int v = (w[0] * S.n[1] + w[1]) * S.n[2] + w[2];
                    pidx[item] = v;
                    sw[item] = Cs * std::exp(-2.0 * xi * xi * a);
                    cnt[v + 1]++;
                }
    }
    for (size_t v = 0; v < V; ++v) cnt[v + 1] += cnt[v];
    std::vector<int> items((size_t)Nn * SS);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (size_t item = 0; item < (size_t)Nn * SS; ++item) items[fill[pidx[item]]++] = (int)item;
    }
    std::vector<float> sw32(sw.begin(), sw.end());
    d_sw64.upload(sw.data(), sw.size());
    d_sw32.upload(sw32.data(), sw32.size());
    d_pidx.upload(pidx.data(), pidx.size());
    d_vox_start.upload(cnt.data(), cnt.size());
    d_vox_items.upload(items.data(), items.size());
    build_quad_slots(Nn);
}

void EwaldOp::build_quad_slots(int Nn) {
    qidx_h.clear();
    if (quads_all)
        for (int i = 0; i < Nn; ++i) qidx_h.push_back(i);
    else
        qidx_h = qreq_h;
    std::vector<int> qmap(Nn, -1);
    for (size_t q = 0; q < qidx_h.size(); ++q) {
        MPM_REQUIRE(qidx_h[q] >= 0 && qidx_h[q] < Nn, MPM_ERR_INVALID_ARGUMENT, "quadrupole index out of range");
        MPM_REQUIRE(qmap[qidx_h[q]] < 0, MPM_ERR_INVALID_ARGUMENT, "repeated quadrupole index");
        qmap[qidx_h[q]] = (int)q;
    }
    std::vector<int> qi = qidx_h;
    if (qi.empty()) qi.push_back(0);
    d_qidx.upload(qi.data(), qi.size());
    d_qmap.upload(qmap.data(), qmap.size());
}

EwaldOp::~EwaldOp() = default;

KParams EwaldOp::kparams() const {
    KParams q;
    q.nx = S.n[0];
    q.ny = S.n[1];
    q.nz = S.n[2];
    q.tcols = (slab && kt) ? 1 : 0;
    q.nxp = nxp;
    q.Lx = L[0];
    q.Ly = L[1];
    q.Lz = L[2];
    q.kbx = kb[0];
    q.kby = kb[1];
    q.xi = xi;
    q.etax = S.eta[0];
    q.etay = S.eta[1];
    q.etaz = S.eta[2];
    q.invV = 1.0 / (double)V;
    return q;
}

void EwaldOp::build_kspace() {
    // the full-grid coefficient table is built lazily (only the full-grid path reads it)
    kspace_ready = false;
}

void EwaldOp::ensure_kspace_table() {
    const int key = quads ? 6000 : (poly ? 5000 : (!slab ? 0 : (zcontig ? 2 + Zs : 1) + (kt ? 1000 : 0)));
    if (kspace_ready && kspace_key == key) return;
    const int threads = 256;
    if (quads) {
        d_kq64.resize(16 * V);
        d_kq32.resize(16 * V);
        const int blocks = (int)std::min<size_t>((V + threads - 1) / threads, 65535);
        hipLaunchKernelGGL(quad_kspace_kernel, dim3(blocks), dim3(threads), 0, nullptr, d_kq64.p, d_kq32.p, kparams());
    } else if (poly) {
        d_k64.resize(V);
        d_k32.resize(V);
        const int blocks = (int)std::min<size_t>((V + threads - 1) / threads, 65535);
        hipLaunchKernelGGL(poly_kspace_kernel, dim3(blocks), dim3(threads), 0, nullptr, d_k64.p, d_k32.p, kparams(),
                           peta);
    } else if (slab && zcontig) {
        const size_t ncols = (size_t)kspace_cols();
        d_toe64.resize((size_t)Zs * 6 * ncols);
        d_toe32.resize((size_t)Zs * 6 * ncols);
        const bool ri = (S.n[2] & 1) != 0;
        if (ri) d_toe32ri.resize((size_t)Zs * 6 * ncols);
        hipLaunchKernelGGL(ztoeplitz_table_kernel, dim3((unsigned)((ncols + threads - 1) / threads), Zs),
                           dim3(threads), 0, nullptr, d_toe64.p, d_toe32.p, ri ? d_toe32ri.p : nullptr, kparams(),
                           Zs);
    } else {
        // [kz][column][4] over the slab's k-space columns, or [voxel][4] on the full grid
        const size_t nk = slab ? (size_t)kspace_cols() * S.n[2] : V;
        d_k64.resize(4 * nk);
        d_k32.resize(4 * nk);
        const int blocks = (int)std::min<size_t>((nk + threads - 1) / threads, 65535);
        if (slab)
            hipLaunchKernelGGL(kspace_slab_kernel, dim3(blocks), dim3(threads), 0, nullptr, d_k64.p, d_k32.p,
                               kparams());
        else
            hipLaunchKernelGGL(kspace_kernel, dim3(blocks), dim3(threads), 0, nullptr, d_k64.p, d_k32.p, kparams());
    }
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipDeviceSynchronize());
    kspace_ready = true;
    kspace_key = key;
}

__global__ void conj_kernel(double2* __restrict__ v, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        v[e].y = -v[e].y;
}

void EwaldOp::conj_self_coef(hipStream_t s) {
    const size_t n = (size_t)self_rows * N;
    if (n == 0) return;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(conj_kernel, dim3(blocks), dim3(256), 0, s, d_selfc.p, n);
    MPM_HIP(hipGetLastError());
}

void EwaldOp::set_bloch(double kx, double ky) {
    if (kx == kb[0] && ky == kb[1]) return;
    MPM_REQUIRE(!direct, MPM_ERR_UNSUPPORTED, "the direct field has no Bloch wave vector");
    kb[0] = kx;
    kb[1] = ky;
    build_kspace();
    if (N > 0) {
        build_pair_coefs();
        ensure_kspace_table();
    }
}

// ---- neighbour list and pair data on the device (monodisperse operator) ----
//
// The host keeps the O(N) cell lists (members in ascending index) and each particle's
// cell; the device visits, for every particle i, the 27 cells around its own in the
// host's order (dz, dy, dx = -1..1) and their members in order, so the CSR it builds is
// the host's.  Pass 0 counts each particle's pairs, pass 1 fills them in: the neighbour,
// x_j - x_i (minimum image), the table values (perp, para) at d_eff = max(d, 1) and, with
// quadrupoles, the seven quadrupole table values (monodisperse_ewald_electric_field.cu
// :877-958, neighbor_list.cu:242-420).
struct NbGeom {
    double L[3];
    double rc, rc2;
    int nc[3];
    int cells;
    int ts;  // table samples
    int nq;  // quadrupole tables (0 or 7)
};

__device__ inline double tab_interp(const double* __restrict__ y, int ts, double x) {
    // PairTable::interp on the samples 0, 1 + 0.001 k (the table's r array)
    if (x <= 0.0) return y[0];
    const double rlast = 1.0 + (ts - 2) * 0.001;
    if (x >= rlast) return y[ts - 1];
    if (x < 1.0) return y[0] * (1.0 - x) + y[1] * x;
    const size_t idx = 1 + (size_t)((x - 1.0) / 0.001);
    if (idx >= (size_t)ts - 1) return y[ts - 1];
    const double r0 = 1.0 + (idx - 1) * 0.001, r1 = 1.0 + idx * 0.001;
    const double t = (x - r0) / (r1 - r0);
    return y[idx] * (1.0 - t) + y[idx + 1] * t;
}

__global__ __launch_bounds__(256) void nb_search_kernel(int pass, int N, const double* __restrict__ pos, NbGeom g,
                                                        const int* __restrict__ cellc,
                                                        const int* __restrict__ cell_start,
                                                        const int* __restrict__ cell_items,
                                                        const int* __restrict__ nb_start, int* __restrict__ count,
                                                        int* __restrict__ nb_j, double* __restrict__ pair_r,
                                                        double2* __restrict__ nb_f, double* __restrict__ pair_q,
                                                        const double* __restrict__ tab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double xi = pos[i], yi = pos[N + i], zi = pos[2 * N + i];
    int k = pass ? nb_start[i] : 0;
    int cnt = 0;
    auto consider = [&](int j) {
        double r[3] = {xi - pos[j], yi - pos[N + j], zi - pos[2 * N + j]};
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] -= g.L[d] * round(r[d] / g.L[d]);
        const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        if (!(d2 <= g.rc2)) return;
        const double dist = sqrt(d2);
        if (!(dist < g.rc)) return;
        ++cnt;
        if (!pass) return;
        nb_j[k] = j;
        const double de = dist < 1.0 ? 1.0 : dist;
        nb_f[k] = make_double2(tab_interp(tab, g.ts, de), tab_interp(tab + g.ts, g.ts, de));
        for (int t = 0; t < g.nq; ++t) pair_q[7 * (size_t)k + t] = tab_interp(tab + (size_t)(2 + t) * g.ts, g.ts, de);
        double rj[3] = {pos[j] - xi, pos[N + j] - yi, pos[2 * N + j] - zi};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            rj[d] -= g.L[d] * round(rj[d] / g.L[d]);
            pair_r[3 * (size_t)k + d] = rj[d];
        }
        ++k;
    };
    if (g.cells) {
        const int cx = cellc[3 * i], cy = cellc[3 * i + 1], cz = cellc[3 * i + 2];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int a = (cx + dx + g.nc[0]) % g.nc[0], b = (cy + dy + g.nc[1]) % g.nc[1],
                              c = (cz + dz + g.nc[2]) % g.nc[2];
                    const int cell = (c * g.nc[1] + b) * g.nc[0] + a;
                    for (int m = cell_start[cell]; m < cell_start[cell + 1]; ++m) {
                        const int j = cell_items[m];
                        if (j != i) consider(j);
                    }
                }
    } else {
        for (int j = 0; j < N; ++j)
            if (j != i) consider(j);
    }
    if (!pass) count[i] = cnt;
}

// Per-pair coefficients (u = (x_j - x_i) / d_eff, perp, para - perp, the Bloch phase
// e^{-i k_b.r}, and with quadrupoles the seven table values) from the pair data, the
// device form of build_pair_coefs.
__global__ __launch_bounds__(256) void pair_coefs_kernel(int64_t M, const double* __restrict__ pair_r,
                                                         const double2* __restrict__ nb_f,
                                                         const double* __restrict__ pair_q, int stride, double dmin,
                                                         double kbx, double kby, double* __restrict__ out) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < M; k += (int64_t)gridDim.x * blockDim.x) {
        const double* r = pair_r + 3 * k;
        const double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double de = d < dmin ? dmin : d;
        double* o = out + stride * k;
        if (stride == 16)
            for (int t = 0; t < 7; ++t) o[7 + t] = pair_q[7 * k + t];
        o[0] = r[0] / de;
        o[1] = r[1] / de;
        o[2] = r[2] / de;
        const double2 f = nb_f[k];
        o[3] = f.x;
        o[4] = f.y - f.x;
        o[5] = 1.0;
        o[6] = 0.0;
        if (kbx != 0.0 || kby != 0.0) {
            double sn, cs;
            sincos(-(kbx * r[0] + kby * r[1]), &sn, &cs);
            o[5] = cs;
            o[6] = sn;
        }
    }
}

void EwaldOp::search_pairs_device(const std::vector<double>& pos, int Nn, bool cells, const int nc[3],
                                  const std::vector<int>& cell_start, const std::vector<int>& cell_items,
                                  const std::vector<int>& cellc) {
    const PairTable& tab = *table;
    const int ts = (int)tab.r.size();
    const int nq = quads ? 7 : 0;
    if (ptab_nt != 2 + nq || d_ptab.n != (size_t)(2 + nq) * ts) {
        std::vector<double> t((size_t)(2 + nq) * ts);
        std::copy(tab.perp.begin(), tab.perp.end(), t.begin());
        std::copy(tab.para.begin(), tab.para.end(), t.begin() + ts);
        for (int q = 0; q < nq; ++q) std::copy(tab.q[q].begin(), tab.q[q].end(), t.begin() + (size_t)(2 + q) * ts);
        d_ptab.upload(t.data(), t.size());
        ptab_nt = 2 + nq;
    }
    d_pos.upload(pos.data(), pos.size());
    if (cells) {
        d_cell_start.upload(cell_start.data(), cell_start.size());
        d_cell_items.upload(cell_items.data(), std::max<size_t>(cell_items.size(), 1));
        d_cellc.upload(cellc.data(), cellc.size());
    }
    NbGeom g;
    for (int d = 0; d < 3; ++d) {
        g.L[d] = L[d];
        g.nc[d] = nc[d];
    }
    g.rc = S.rc;
    g.rc2 = S.rc * S.rc;
    g.cells = cells ? 1 : 0;
    g.ts = ts;
    g.nq = nq;
    d_nb_count.resize((size_t)Nn + 1);
    d_nb_start.resize((size_t)Nn + 1);
    const dim3 grid((unsigned)((Nn + 255) / 256));
    hipLaunchKernelGGL(nb_search_kernel, grid, dim3(256), 0, nullptr, 0, Nn, d_pos.p, g, d_cellc.p, d_cell_start.p,
                       d_cell_items.p, (const int*)nullptr, d_nb_count.p, (int*)nullptr, (double*)nullptr,
                       (double2*)nullptr, (double*)nullptr, d_ptab.p);
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipMemsetAsync(d_nb_count.p + Nn, 0, sizeof(int), nullptr));
    size_t tmp = 0;
    MPM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_nb_count.p, d_nb_start.p, Nn + 1, nullptr));
    d_scan_tmp.reserve(std::max<size_t>(tmp, 1));
    MPM_HIP(hipcub::DeviceScan::ExclusiveSum(d_scan_tmp.p, tmp, d_nb_count.p, d_nb_start.p, Nn + 1, nullptr));
    int total = 0;
    MPM_HIP(hipMemcpy(&total, d_nb_start.p + Nn, sizeof(int), hipMemcpyDeviceToHost));
    num_pairs = total;
    const size_t M = std::max<size_t>((size_t)total, 1);
    d_nb_j.resize(M);
    d_pair_r.resize(3 * M);
    d_nb_f.resize(M);
    if (nq) d_pairq.resize(7 * M);
    hipLaunchKernelGGL(nb_search_kernel, grid, dim3(256), 0, nullptr, 1, Nn, d_pos.p, g, d_cellc.p, d_cell_start.p,
                       d_cell_items.p, (const int*)d_nb_start.p, (int*)nullptr, d_nb_j.p, d_pair_r.p, d_nb_f.p,
                       nq ? d_pairq.p : (double*)nullptr, d_ptab.p);
    MPM_HIP(hipGetLastError());
    if (total == 0) {  // the kernels read one entry: a harmless zero pair
        MPM_HIP(hipMemsetAsync(d_nb_j.p, 0, sizeof(int), nullptr));
        MPM_HIP(hipMemsetAsync(d_nb_f.p, 0, sizeof(double2), nullptr));
        MPM_HIP(hipMemsetAsync(d_pair_r.p, 0, 3 * sizeof(double), nullptr));
    }
    pairs_dev = true;
    pairs_host_valid = false;
    h_nb_start.clear();
    h_nb_j.clear();
    h_pair_r.clear();
    h_pair_f.clear();
    h_pair_q.clear();
}

// Host copies of the device-built neighbour list and pair data (getters, diagnostics).
void EwaldOp::fetch_pairs() const {
    if (!pairs_dev || pairs_host_valid) return;
    const int Nn = (int)N;
    const size_t M = (size_t)num_pairs;
    h_nb_start.resize((size_t)Nn + 1);
    h_nb_j.resize(M);
    h_pair_r.resize(3 * M);
    h_pair_f.resize(M);
    MPM_HIP(hipMemcpy(h_nb_start.data(), d_nb_start.p, ((size_t)Nn + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (M) {
        MPM_HIP(hipMemcpy(h_nb_j.data(), d_nb_j.p, M * sizeof(int), hipMemcpyDeviceToHost));
        MPM_HIP(hipMemcpy(h_pair_r.data(), d_pair_r.p, 3 * M * sizeof(double), hipMemcpyDeviceToHost));
        MPM_HIP(hipMemcpy(h_pair_f.data(), d_nb_f.p, M * sizeof(double2), hipMemcpyDeviceToHost));
    }
    if (quads) {
        h_pair_q.resize(7 * M);
        if (M) MPM_HIP(hipMemcpy(h_pair_q.data(), d_pairq.p, 7 * M * sizeof(double), hipMemcpyDeviceToHost));
    }
    pairs_host_valid = true;
}

void EwaldOp::build_pair_coefs() {
    if (pairs_dev) {
        const int64_t M = num_pairs;
        const int stride = quads ? 16 : 8;
        d_pairc.resize((size_t)std::max<int64_t>(M, 1) * stride);
        if (M > 0) {
            const int blocks = (int)std::min<int64_t>((M + 255) / 256, 16384);
            hipLaunchKernelGGL(pair_coefs_kernel, dim3(blocks), dim3(256), 0, nullptr, M, d_pair_r.p, d_nb_f.p,
                               quads ? d_pairq.p : (const double*)nullptr, stride, 1.0, kb[0], kb[1], d_pairc.p);
            MPM_HIP(hipGetLastError());
        } else {
            MPM_HIP(hipMemset(d_pairc.p, 0, (size_t)stride * sizeof(double)));
        }
        MPM_HIP(hipDeviceSynchronize());
        return;
    }
    const size_t M = h_pair_f.size();
    const int stride = (quads && h_pair_q.size() == 7 * M) ? 16 : 8;  // quadrupoles: + 7 table values
    std::vector<double> c(std::max<size_t>(M, 1) * stride, 0.0);
    const bool bloch = kb[0] != 0.0 || kb[1] != 0.0;
    for (size_t k = 0; k < M; ++k) {
        const double* r = &h_pair_r[3 * k];
        const double d = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double dmin = poly ? 1e-6 : 1.0;  // d_eff of the unit vector
        const double de = d < dmin ? dmin : d;
        double* o = &c[stride * k];
        if (stride == 16)
            for (int t = 0; t < 7; ++t) o[7 + t] = h_pair_q[7 * k + t];
        o[0] = r[0] / de;
        o[1] = r[1] / de;
        o[2] = r[2] / de;
        o[3] = h_pair_f[k].x;
        o[4] = h_pair_f[k].y - h_pair_f[k].x;
        o[5] = 1.0;
        o[6] = 0.0;
        if (bloch) {
            const double ph = -(kb[0] * r[0] + kb[1] * r[1]);
            o[5] = std::cos(ph);
            o[6] = std::sin(ph);
        }
    }
    d_pairc.upload(c.data(), c.size());
}

// LDS bytes of zfused_kernel for slab size zs_pad and precision size rs.
static size_t zfused_lds(int zs_pad, int nz, size_t rs) { return 2 * rs * zs_pad * nz; }
// Region-local separable weight rows of the matrix-core kernels, one block per
// region (grid-stride), threads over the (candidate, entry) pairs of its slots:
// entry e of a row is the weight of axis d at voxel v summed over the stencil
// points of the particle that land on v (sum over o = off, off + n_d, ... < W of
// its per-axis weights, off = (v - (base - half)) mod n_d), or zero outside the
// grid and in the row's padding.
struct RegionWeightGeom {
    int nreg, RX, npb, nry, pblk, arow, zs, NW, zsb4, W, half, N;
    int n[3];
};

__global__ __launch_bounds__(256) void region_weights_kernel(RegionWeightGeom g, const int* __restrict__ reg_start,
                                                             const int* __restrict__ reg_items,
                                                             const int* __restrict__ base,
                                                             const double* __restrict__ w64,
                                                             const int* __restrict__ zlist, double* __restrict__ rw64,
                                                             float* __restrict__ rw32) {
    for (int r = blockIdx.x; r < g.nreg; r += gridDim.x) {
        const int per_x = g.npb * g.nry;
        const int x0 = (r / per_x) * g.RX, pl0 = ((r % per_x) / g.nry) * g.pblk, y0 = (r % g.nry) * kMfmaRY;
        const int k0 = reg_start[r], k1 = reg_start[r + 1];
        for (int e = threadIdx.x; e < (k1 - k0) * g.NW; e += blockDim.x) {
            const int q = k0 + e / g.NW, col = e % g.NW;
            const int pid = reg_items[q];
            int d = -1, v = 0;
            if (col < 4) {
                if (col < g.RX && x0 + col < g.n[0]) {
                    d = 0;
                    v = x0 + col;
                }
            } else if (col < 4 + g.zsb4) {
                const int p = col - 4;
                if (g.zs > 0) {
                    if (p < g.zs) {
                        d = 2;
                        v = zlist[p];
                    }
                } else if (p < g.pblk && pl0 + p < g.n[1]) {
                    d = 1;
                    v = pl0 + p;
                }
            } else {
                const int j = col - 4 - g.zsb4;
                if (y0 + j < g.n[g.arow]) {
                    d = g.arow;
                    v = y0 + j;
                }
            }
            double acc = 0.0;
            if (d >= 0) {
                const int nd = g.n[d];
                const int b0 = base[(size_t)d * g.N + pid] - g.half;
                int off = (v - b0) % nd;
                if (off < 0) off += nd;
                for (int o = off; o < g.W; o += nd) acc += w64[((size_t)pid * 3 + d) * g.W + o];
            }
            rw64[(size_t)q * g.NW + col] = acc;
            rw32[(size_t)q * g.NW + col] = (float)acc;
        }
    }
}

static int zs_bucket(int Zs) {
    static const int buckets[] = {4, 6, 8, 10, 12, 14, 16};
    for (int b : buckets)
        if (Zs <= b) return b;
    return 0;
}

void EwaldOp::set_positions(const double* xyz, int64_t n) {
    MPM_REQUIRE(n > 0, MPM_ERR_INVALID_ARGUMENT, "no particles");
    MPM_REQUIRE(n < (1LL << 31) / 1024, MPM_ERR_INVALID_ARGUMENT, "too many particles");
    MPM_REQUIRE(!poly || n == (int64_t)rad_h.size(), MPM_ERR_INVALID_ARGUMENT,
                "the polydisperse operator needs one radius per particle");
    const int Nn = (int)n;
    for (int64_t q = 0; q < 3 * n; ++q)
        MPM_REQUIRE(std::isfinite(xyz[q]), MPM_ERR_INVALID_ARGUMENT, "non-finite position");
    if (direct) {
        MPM_REQUIRE(rad_h.size() == 1 || (int64_t)rad_h.size() == n, MPM_ERR_INVALID_ARGUMENT,
                    "the direct field needs one radius per particle (or one for all)");
        std::vector<double> pos(3 * (size_t)Nn), rr(Nn);
        for (int i = 0; i < Nn; ++i) {
            for (int d = 0; d < 3; ++d) pos[(size_t)d * Nn + i] = xyz[3 * (size_t)i + d];
            rr[i] = rad_h.size() == 1 ? rad_h[0] : rad_h[i];
        }
        N = 0;
        d_pos.upload(pos.data(), pos.size());
        d_drad.upload(rr.data(), rr.size());
        if (quads) build_quad_slots(Nn);
        N = n;
        return;
    }
    // unchanged positions: the per-frame precalcs are still valid
    if (n == N && last_xyz.size() == (size_t)(3 * n) &&
        std::memcmp(last_xyz.data(), xyz, 3 * n * sizeof(double)) == 0)
        return;
    // the cache of precalcs is valid only once this call has finished
    last_xyz.clear();
    N = n;
    HostLaps laps;
    const int W = S.width, half = S.half;
    std::vector<double> pos(3 * (size_t)Nn);
    std::vector<int> base(3 * (size_t)Nn);
    std::vector<double> w64((size_t)Nn * 3 * W);
    std::vector<float> w32((size_t)Nn * 3 * W);
    const double prod_eta = S.eta[0] * S.eta[1] * S.eta[2];
    const double Cs = std::pow(2.0 * xi * xi / kPi, 1.5) * std::sqrt(1.0 / prod_eta);
    parallel_for(Nn, host_threads(Nn, 2048), [&](int i0, int i1, int) {
        for (int i = i0; i < i1; ++i) {
            for (int d = 0; d < 3; ++d) {
                const double p = xyz[3 * (size_t)i + d];
                pos[(size_t)d * Nn + i] = p;
                const double gd = std::round(p / S.h[d]);
                MPM_REQUIRE(std::fabs(gd) < 1e9, MPM_ERR_INVALID_ARGUMENT, "position too far from the box");
                const long long gi = (long long)gd;
                const double dd = gi * S.h[d] - p;
                base[(size_t)d * Nn + i] = (int)(((gi % S.n[d]) + S.n[d]) % S.n[d]);
                for (int o = -half; o <= half; ++o) {
                    const double x = dd + o * S.h[d];
                    double wv = std::exp(-2.0 * xi * xi * x * x / S.eta[d]);
                    if (d == 0) wv *= Cs;
                    w64[((size_t)i * 3 + d) * W + o + half] = wv;
                    w32[((size_t)i * 3 + d) * W + o + half] = (float)wv;
                }
            }
        }
    });
    laps.lap("host_setup_weights");
    // ---- reciprocal layout: slab of occupied z-planes, or the full grid ----
    const int nz = S.n[2];
    std::vector<char> occ(nz, 0);
    for (int i = 0; i < Nn; ++i)
        for (int o = -half; o <= half; ++o) occ[(((base[2 * (size_t)Nn + i] + o) % nz) + nz) % nz] = 1;
    std::vector<int> zl;
    for (int z = 0; z < nz; ++z)
        if (occ[z]) zl.push_back(z);
    // one cyclic run of occupied planes: order the slab from the run's first plane
    // so that slab plane a sits at z0 + a (mod nz)
    int runs = 0, z0 = 0;
    for (int z = 0; z < nz; ++z)
        if (occ[z] && !occ[(z + nz - 1) % nz]) {
            ++runs;
            z0 = z;
        }
    if (runs == 1) {
        const int nzl = (int)zl.size();
        for (int a = 0; a < nzl; ++a) zl[a] = (z0 + a) % nz;
    }
    const int zsb = zs_bucket((int)zl.size());
    const bool want_slab = force_full == 0 && zsb > 0 && (int)zl.size() * 2 <= nz &&
                           zfused_lds(zsb, nz, sizeof(double)) <= 64 * 1024;
    slab = want_slab;
    Zs = slab ? (int)zl.size() : 0;
    zcontig = slab && runs == 1;
    zlist = zl;
    std::vector<int> zmap(nz, -1);
    for (size_t j = 0; j < zl.size(); ++j) zmap[zl[j]] = (int)j;
    // tiles and bins
    if (slab) {
        tile[0] = 8; tile[1] = 16; tile[2] = 2;
        binsz[0] = 4; binsz[1] = 4; binsz[2] = 2;
    } else {
        tile[0] = 4; tile[1] = 8; tile[2] = 8;
        binsz[0] = 4; binsz[1] = 8; binsz[2] = 8;
    }
    for (int d = 0; d < 3; ++d) {
        nbins[d] = (S.n[d] + binsz[d] - 1) / binsz[d];
        ntiles[d] = (S.n[d] + tile[d] - 1) / tile[d];
    }
    // regions of the matrix-core spread and contraction: every particle joins each
    // region its stencil footprint meets.  Slab: x blocks of reg_rx and y blocks of
    // 16, all slab planes in each region; full grid (dipoles of one radius): x blocks
    // of 4, y blocks of kFullRegionPlanes planes and z blocks of 16 rows.
    reg_rx = 0;
    reg_full = false;
    reg_npb = 1;
    reg_pblk = 0;
    if (slab || (!poly && !quads)) {
        const int RX = slab ? (zs_bucket(Zs) <= 10 ? 4 : 2) : 4, RY = kMfmaRY;
        const int arow = slab ? 1 : 2;  // axis of the 16 contiguous rows
        const int pblk = slab ? Zs : kFullRegionPlanes;
        reg_rx = RX;
        reg_full = !slab;
        reg_pblk = pblk;
        reg_nrx = (S.n[0] + RX - 1) / RX;
        reg_nry = (S.n[arow] + RY - 1) / RY;
        reg_npb = slab ? 1 : (S.n[1] + pblk - 1) / pblk;
        const int nreg = reg_nrx * reg_npb * reg_nry;
        // distinct blocks of size R that the stencil of particle i meets along axis d
        // (walking its W voxels with the periodic wrap; one division per axis)
        auto blocks_of = [&](int i, int d, int R, int* outv) {
            int n = 0;
            const int nd = S.n[d];
            int v = (base[(size_t)d * Nn + i] - half) % nd;
            if (v < 0) v += nd;
            int r = v / R, next = (r + 1) * R, last = -1;
            for (int o = 0; o < W; ++o) {
                if (r != last) {
                    bool seen = false;
                    for (int q = 0; q < n; ++q) seen = seen || outv[q] == r;
                    if (!seen) outv[n++] = r;
                    last = r;
                }
                if (++v == nd) {
                    v = 0;
                    r = 0;
                    next = R;
                } else if (v == next) {
                    ++r;
                    next += R;
                }
            }
            return n;
        };
        auto for_regions = [&](int i, auto&& f) {
            int rxs[kMaxWidth], rys[kMaxWidth], pbs[kMaxWidth];
            const int nrx = blocks_of(i, 0, RX, rxs);
            const int nry = blocks_of(i, arow, RY, rys);
            int npb = 1;
            pbs[0] = 0;
            if (!slab) npb = blocks_of(i, 1, pblk, pbs);
            for (int a = 0; a < nrx; ++a)
                for (int c = 0; c < npb; ++c)
                    for (int e = 0; e < nry; ++e) f(((size_t)rxs[a] * reg_npb + pbs[c]) * reg_nry + rys[e]);
        };
        // candidates of every region in ascending particle order: each host thread counts
        // its contiguous chunk of particles, then fills after the chunks before it
        const int nth = host_threads(Nn, 16);
        std::vector<std::vector<int>> cnt(nth, std::vector<int>(nreg, 0));
        parallel_for(Nn, nth, [&](int i0, int i1, int t) {
            for (int i = i0; i < i1; ++i) for_regions(i, [&](size_t r) { cnt[t][r]++; });
        });
        std::vector<int> rs(nreg + 1, 0);
        for (int r = 0; r < nreg; ++r) {
            int acc = rs[r];
            for (int t = 0; t < nth; ++t) {
                const int c = cnt[t][r];
                cnt[t][r] = acc;  // this chunk's first position in region r
                acc += c;
            }
            rs[r + 1] = acc;
        }
        std::vector<int> ri(std::max(rs[nreg], 1));
        parallel_for(Nn, nth, [&](int i0, int i1, int t) {
            for (int i = i0; i < i1; ++i) for_regions(i, [&](size_t r) { ri[cnt[t][r]++] = i; });
        });
        // slots of each particle, in ascending region order
        std::vector<int> pss(Nn + 1, 0), psi(std::max(rs[nreg], 1));
        for (int q = 0; q < rs[nreg]; ++q) pss[ri[q] + 1]++;
        for (int i = 0; i < Nn; ++i) pss[i + 1] += pss[i];
        {
            std::vector<int> fill(pss.begin(), pss.end() - 1);
            for (int q = 0; q < rs[nreg]; ++q) psi[fill[ri[q]]++] = q;
        }
        laps.lap("host_setup_region_lists");
        // grow-only: the slot count changes a little from frame to frame
        d_reg_start.assign(rs.data(), rs.size());
        d_reg_items.assign(ri.data(), ri.size());
        d_ps_start.assign(pss.data(), pss.size());
        d_ps_items.assign(psi.data(), psi.size());
        // one buffer for both precisions (an apply runs in one of them at a time), with an
        // eighth of headroom so that the next frames' slot counts fit without reallocation
        {
            const size_t need = (size_t)std::max(rs[nreg], 1) * 6 * kMaxSlots;
            if (need > d_partial64.n) d_partial64.reserve(need + need / 8);
        }
        reg_slots = rs[nreg];
        // the region-local weight rows are computed on the device once the per-particle
        // weights are there (region_weights_kernel, below)
    }
    laps.lap("host_setup_regions");
    std::vector<int> ztl;  // z-tiles that hold at least one computed voxel
    for (int t = 0; t < ntiles[2]; ++t) {
        bool any = !slab;
        for (int z = t * tile[2]; z < std::min(nz, (t + 1) * tile[2]) && !any; ++z) any = occ[z] != 0;
        if (any) ztl.push_back(t);
    }
    const int nb = nbins[0] * nbins[1] * nbins[2];
    std::vector<int> bin_of(Nn), start(nb + 1, 0), items(Nn);
    for (int i = 0; i < Nn; ++i) {
        const int bx = base[i] / binsz[0], by = base[Nn + i] / binsz[1], bz = base[2 * Nn + i] / binsz[2];
        bin_of[i] = (bx * nbins[1] + by) * nbins[2] + bz;
        start[bin_of[i] + 1]++;
    }
    for (int b = 0; b < nb; ++b) start[b + 1] += start[b];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < Nn; ++i) items[fill[bin_of[i]]++] = i;
    }
    laps.lap("host_setup_bins");
    // real-space neighbours: all j != i with min-image distance < rc (neighbor_list.cu
    // keeps d^2 <= rc^2 and the kernel then skips d >= rc; pairs at exactly rc add zero)
    const PairTable& tab = *table;
    const double rc = S.rc, rc2 = rc * rc;
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = std::max(1, (int)std::floor(L[d] / rc));
    const bool cells = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
    std::vector<std::vector<int>> members;
    auto cell_coord = [&](double p, int d) {
        const double wv = p - L[d] * std::floor(p / L[d]);
        int c = (int)std::floor(wv / (L[d] / nc[d]));
        return std::min(std::max(c, 0), nc[d] - 1);
    };
    std::vector<int> cellc;  // each particle's cell (x, y, z)
    if (cells) {
        members.resize((size_t)nc[0] * nc[1] * nc[2]);
        cellc.resize(3 * (size_t)Nn);
        for (int j = 0; j < Nn; ++j) {
            const int cx = cell_coord(pos[j], 0), cy = cell_coord(pos[Nn + j], 1), cz = cell_coord(pos[2 * Nn + j], 2);
            members[(size_t)(cz * nc[1] + cy) * nc[0] + cx].push_back(j);
            cellc[3 * (size_t)j] = cx;
            cellc[3 * (size_t)j + 1] = cy;
            cellc[3 * (size_t)j + 2] = cz;
        }
    }
    // The monodisperse operator searches its pairs on the device from the host's cell lists
    // (the table values need no host work); the polydisperse operator's pair values come from
    // closed forms at each pair's radii, evaluated on host threads.
    const bool dev_pairs = !poly;
    std::vector<int> nstart(Nn + 1, 0);
    std::vector<int> nj;
    std::vector<double2> nf;
    std::vector<double> pr, pq;
    if (dev_pairs) {
        std::vector<int> cs, ci;
        if (cells) {
            cs.assign(members.size() + 1, 0);
            for (size_t c = 0; c < members.size(); ++c) cs[c + 1] = cs[c] + (int)members[c].size();
            ci.reserve(Nn);
            for (const auto& m : members) ci.insert(ci.end(), m.begin(), m.end());
        }
        search_pairs_device(pos, Nn, cells, nc, cs, ci, cellc);
    } else {
        pairs_dev = false;
        pairs_host_valid = true;
        // particles split into contiguous chunks, one per host thread; each chunk's pairs are
        // appended in particle order, so the list is the same for any thread count
        struct Chunk {
            std::vector<int> nj;
            std::vector<double2> nf;
            std::vector<double> pr, pq;
        };
        auto consider = [&](Chunk& ch, int i, int j) {
            std::vector<int>& nj = ch.nj;
            std::vector<double2>& nf = ch.nf;
            std::vector<double>& pr = ch.pr;
            std::vector<double>& pq = ch.pq;
            double r[3] = {pos[i] - pos[j], pos[Nn + i] - pos[Nn + j], pos[2 * Nn + i] - pos[2 * Nn + j]};
            for (int d = 0; d < 3; ++d) r[d] -= L[d] * std::round(r[d] / L[d]);
            const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (!(d2 <= rc2)) return;
            const double dist = std::sqrt(d2);
            if (!(dist < rc)) return;
            nj.push_back(j);
            if (poly) {  // d_eff = max(d, 1e-6) and the radius pair's table
                const double de = dist < 1e-6 ? 1e-6 : dist;
                double fp, fa;
                poly_pair_value(de, i, j, fp, fa);
                nf.push_back(make_double2(fp, fa));
            } else {
                const double de = dist < 1.0 ? 1.0 : dist;
                nf.push_back(make_double2(tab.interp(tab.perp, de), tab.interp(tab.para, de)));
                if (quads)
                    for (int t = 0; t < 7; ++t) pq.push_back(tab.interp(tab.q[t], de));
            }
            // x_j - x_i, the orientation the gather uses
            double rj[3] = {pos[j] - pos[i], pos[Nn + j] - pos[Nn + i], pos[2 * Nn + j] - pos[2 * Nn + i]};
            for (int d = 0; d < 3; ++d) rj[d] -= L[d] * std::round(rj[d] / L[d]);
            pr.insert(pr.end(), rj, rj + 3);
        };
        const int nthreads = host_threads(Nn, 256);
        std::vector<Chunk> chunks(nthreads);
        parallel_for(Nn, nthreads, [&](int i0, int i1, int t) {
            Chunk& ch = chunks[t];
            const size_t guess = (size_t)(i1 - i0) * 64;
            ch.nj.reserve(guess);
            ch.nf.reserve(guess);
            ch.pr.reserve(3 * guess);
            for (int i = i0; i < i1; ++i) {
                if (cells) {
                    const int cx = cell_coord(pos[i], 0), cy = cell_coord(pos[Nn + i], 1),
                              cz = cell_coord(pos[2 * Nn + i], 2);
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                const int a = (cx + dx + nc[0]) % nc[0], b = (cy + dy + nc[1]) % nc[1],
                                          c = (cz + dz + nc[2]) % nc[2];
                                for (int j : members[(size_t)(c * nc[1] + b) * nc[0] + a])
                                    if (j != i) consider(ch, i, j);
                            }
                } else {
                    for (int j = 0; j < Nn; ++j)
                        if (j != i) consider(ch, i, j);
                }
                nstart[i + 1] = (int)ch.nj.size();  // chunk-local; made global below
            }
        });
        {
            size_t tot = 0;
            for (const Chunk& ch : chunks) tot += ch.nj.size();
            nj.reserve(tot);
            nf.reserve(tot);
            pr.reserve(3 * tot);
            for (int c = 0; c < nthreads; ++c) {
                const int off = (int)nj.size();
                for (int i = chunk_begin(Nn, nthreads, c); i < chunk_begin(Nn, nthreads, c + 1); ++i) nstart[i + 1] += off;
                const Chunk& ch = chunks[c];
                nj.insert(nj.end(), ch.nj.begin(), ch.nj.end());
                nf.insert(nf.end(), ch.nf.begin(), ch.nf.end());
                pr.insert(pr.end(), ch.pr.begin(), ch.pr.end());
                pq.insert(pq.end(), ch.pq.begin(), ch.pq.end());
            }
        }
        num_pairs = (int64_t)nj.size();
        h_nb_start = nstart;
        h_nb_j = nj;
    }
    laps.lap("host_setup_neighbors");
    // visiting order of the wave-per-particle kernels: columns of 8 x 8 voxels in x and
    // y, ascending z within a column, so that consecutive waves share stencil voxels
    {
        std::vector<int64_t> key(Nn);
        const int64_t nby = (S.n[1] + 7) / 8;
        for (int i = 0; i < Nn; ++i) {
            int gv[3];
            for (int d = 0; d < 3; ++d) {
                double x = pos[(size_t)d * Nn + i];
                x -= L[d] * std::floor(x / L[d]);
                gv[d] = std::min(S.n[d] - 1, std::max(0, (int)(x / S.h[d])));
            }
            key[i] = ((int64_t)(gv[0] >> 3) * nby + (gv[1] >> 3)) * S.n[2] + gv[2];
        }
        std::vector<int> ord(Nn);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
        d_order.upload(ord.data(), ord.size());
    }
    if (!dev_pairs) d_pos.upload(pos.data(), pos.size());
    d_base.upload(base.data(), base.size());
    d_w64.upload(w64.data(), w64.size());
    d_w32.upload(w32.data(), w32.size());
    d_bin_start.upload(start.data(), start.size());
    d_bin_items.upload(items.data(), items.size());
    d_ztiles.upload(ztl.data(), ztl.size());
    n_ztiles = (int)ztl.size();
    d_zmap.upload(zmap.data(), zmap.size());
    if (zl.empty()) zl.push_back(0);
    d_zlist.upload(zl.data(), zl.size());
    if (reg_rx > 0) {
        // region-local separable weights of every (region, candidate) slot, in the row
        // layout of the matrix-core kernels: x at [0, RX), planes at [4, 4 + ZSB), rows at
        // [4 + ZSB4, 4 + ZSB4 + 16); sums over stencil points that fold
        const int zsb = slab ? zs_bucket(Zs) : kFullRegionPlanes, zsb4 = (zsb + 3) & ~3;
        const int NW = 4 + zsb4 + kMfmaRY;
        const size_t nw = (size_t)std::max(reg_slots, 1) * NW;
        if (nw > d_regw64.n) d_regw64.reserve(nw + nw / 8);
        if (nw > d_regw32.n) d_regw32.reserve(nw + nw / 8);
        RegionWeightGeom wg;
        wg.nreg = reg_nrx * reg_npb * reg_nry;
        wg.RX = reg_rx;
        wg.npb = reg_npb;
        wg.nry = reg_nry;
        wg.pblk = reg_pblk;
        wg.arow = slab ? 1 : 2;
        wg.zs = slab ? Zs : 0;
        wg.NW = NW;
        wg.zsb4 = zsb4;
        wg.W = W;
        wg.half = half;
        wg.N = Nn;
        for (int d = 0; d < 3; ++d) wg.n[d] = S.n[d];
        const int blocks = std::min(wg.nreg, 65535);
        hipLaunchKernelGGL(region_weights_kernel, dim3((unsigned)std::max(blocks, 1)), dim3(256), 0, nullptr, wg,
                           d_reg_start.p, d_reg_items.p, d_base.p, d_w64.p, d_zlist.p, d_regw64.p, d_regw32.p);
        MPM_HIP(hipGetLastError());
        MPM_HIP(hipStreamSynchronize(nullptr));
    }
    if (!dev_pairs) {
        d_nb_start.upload(nstart.data(), nstart.size());
        if (nj.empty()) {
            nj.push_back(0);
            nf.push_back(make_double2(0, 0));
        }
        h_pair_r = std::move(pr);
        h_pair_q = std::move(pq);
        h_pair_f = nf;
        h_pair_f.resize(num_pairs);
        d_nb_j.upload(nj.data(), nj.size());
        d_nb_f.upload(nf.data(), nf.size());
    }
    build_pair_coefs();
    // FFT plans for the chosen layout.  With the built-in transform the slab rows
    // are padded to whole 128-byte cache lines (16 FP32 / 8 FP64 complex; 16 for
    // both), so that the region tiles and transposed passes touch whole lines.
    const bool kt_ok = slab && fft_backend == 0 && S.n[0] > 1 && S.n[1] > 1 && LineFft::fits(S.n[0]) &&
                       LineFft::fits(S.n[1]);
    nyp = kt_ok ? (S.n[1] + 15) / 16 * 16 : S.n[1];
    nxp = kt_ok ? (S.n[0] + 15) / 16 * 16 : S.n[0];
    const int key = slab ? Zs : -1;
    if (!lfft || fft_key != key) {
        const int nx = S.n[0], ny = S.n[1], nzz = S.n[2];
        const long long nxy = (long long)nx * ny;
        auto lf = std::make_unique<LineFft>();
        bool ok = true;
        auto axis = [&](int n) {
            ok = ok && LineFft::fits(n);
            lf->axes.push_back(std::make_unique<LineAxis>(ok ? n : 2));
            return (int)lf->axes.size() - 1;
        };
        auto pass = [&](std::vector<LinePass>& v, int a, LineGeom gi, LineGeom go, int src, int dst) {
            v.push_back(LinePass{a, gi, go, src, dst});
        };
        if (kt_ok) {
            // [3][Zs][nx][nyp] -> y transform written transposed to the k-space
            // buffer [3][Zs][ky][nxp] -> x transform in place on contiguous lines
            lf->rhs_elems = 3LL * Zs * std::max((long long)nx * nyp, (long long)ny * nxp);
            const int ay = axis(ny), ax = axis(nx);
            lf->kt_x = ax;
            const LineGeom yrun{3LL * Zs * nx, 1, 0, nyp, 1};
            const LineGeom ytr{3LL * Zs * nx, nx, 1, (long long)ny * nxp, nxp};
            const LineGeom xrun{3LL * Zs * ny, 1, 0, nxp, 1};
            pass(lf->fwd, ay, yrun, ytr, 0, 1);
            pass(lf->fwd, ax, xrun, xrun, 1, 1);
            pass(lf->inv, ax, xrun, xrun, 1, 1);
            pass(lf->inv, ay, ytr, yrun, 1, 0);
        } else if (slab) {  // a one-voxel axis: in place on [3][Zs][nx][ny]
            lf->rhs_elems = 3LL * Zs * nxy;
            if (ny > 1) {
                const LineGeom g{3LL * Zs * nx, 1, 0, ny, 1};
                const int a = axis(ny);
                pass(lf->fwd, a, g, g, 0, 0);
                pass(lf->inv, a, g, g, 0, 0);
            }
            if (nx > 1) {
                const LineGeom g{3LL * Zs * ny, ny, 1, nxy, ny};
                const int a = axis(nx);
                pass(lf->fwd, a, g, g, 0, 0);
                pass(lf->inv, a, g, g, 0, 0);
            }
        } else {  // [comps][nx][ny][nz]: z, y, x lines in place (one component for the polydisperse grid)
            const long long nc = poly ? 1 : (quads ? 8 : 3);
            lf->rhs_elems = nc * nxy * nzz;
            const LineGeom gz{nc * nxy, 1, 0, nzz, 1};
            const LineGeom gy{nc * nx * nzz, nzz, 1, (long long)ny * nzz, nzz};
            const LineGeom gx{nc * ny * nzz, (long long)ny * nzz, 1, nxy * nzz, (long long)ny * nzz};
            const int n3[3] = {nzz, ny, nx};
            const LineGeom g3[3] = {gz, gy, gx};
            for (int d = 0; d < 3; ++d) {
                if (n3[d] <= 1) continue;
                const int a = axis(n3[d]);
                if (d == 0) lf->full_z = a;
                pass(lf->fwd, a, g3[d], g3[d], 0, 0);
                pass(lf->inv, a, g3[d], g3[d], 0, 0);
            }
        }
        lfft = ok ? std::move(lf) : nullptr;
    }
    // the slab's k-space data sits transposed ([ky][kx] columns) in its own
    // buffer when the built-in transform runs the slab with two real axes
    kt = slab && lfft && fft_backend == 0 && S.n[0] > 1 && S.n[1] > 1;
    if (!fft || fft_key != key) {
        if (slab)
            fft = std::make_unique<FftPlans>(std::vector<size_t>{(size_t)S.n[1], (size_t)S.n[0]}, 3 * Zs);
        else
            fft = std::make_unique<FftPlans>(std::vector<size_t>{(size_t)S.n[2], (size_t)S.n[1], (size_t)S.n[0]},
                                             poly ? 1 : (quads ? 8 : 3));
        fft_key = key;
    }
    if (poly) {
        MPM_REQUIRE(lfft != nullptr, MPM_ERR_UNSUPPORTED, "grid length not supported by the built-in FFT");
        setup_poly_stencils(pos, Nn);
    }
    if (quads) {
        MPM_REQUIRE(lfft != nullptr && fft_backend == 0, MPM_ERR_UNSUPPORTED,
                    "quadrupoles run on the built-in FFT only");
        setup_quad_stencils(pos, Nn);
        build_pair_coefs();  // with the quadrupole table values
    }
    // default self coefficients: zero (reference base_electric_field.cu allocates them zeroed)
    if (self_rows == 0 || d_selfc.n != (size_t)self_rows * Nn) {
        self_rows = 1;
        d_selfc.resize(Nn);
        MPM_HIP(hipMemset(d_selfc.p, 0, Nn * sizeof(double2)));
    }
    laps.lap("host_setup_uploads_fft");
    // k-space tables of the layout just chosen: built here, so that applies allocate
    // nothing beyond their grids
    ensure_kspace_table();
    laps.lap("host_setup_kspace");
    last_xyz.assign(xyz, xyz + 3 * n);
}

void EwaldOp::set_self_coef(const double* coef, int rows) {
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "set positions before the self coefficients");
    MPM_REQUIRE(rows >= 1, MPM_ERR_INVALID_ARGUMENT, "need at least one row of self coefficients");
    d_selfc.upload(reinterpret_cast<const double2*>(coef), (size_t)rows * N);
    self_rows = rows;
}

void EwaldOp::set_self_coef_device(const double* coef, int rows, hipStream_t s) {
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "set positions before the self coefficients");
    MPM_REQUIRE(rows >= 1, MPM_ERR_INVALID_ARGUMENT, "need at least one row of self coefficients");
    d_selfc.resize((size_t)rows * N);
    MPM_HIP(hipMemcpyAsync(d_selfc.p, coef, (size_t)rows * N * sizeof(double2), hipMemcpyDeviceToDevice, s));
    self_rows = rows;
}

size_t EwaldOp::grid_elems_per_rhs() const {
    if (poly) return V;
    if (quads) return 8 * V;
    return slab ? 3 * (size_t)Zs * S.n[0] * nyp : 3 * V;
}
size_t EwaldOp::kgrid_elems_per_rhs() const { return 3 * (size_t)Zs * kspace_cols(); }
// columns of one slab plane in the k-space buffer (padding columns included)
int EwaldOp::kspace_cols() const { return kt ? S.n[1] * nxp : S.n[0] * S.n[1]; }

template <typename R, bool SLAB>
void EwaldOp::launch(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi) {
    using CT = typename C2<R>::T;
    const bool fp32 = std::is_same<R, float>::value;
    CT* grid = reinterpret_cast<CT*>(cur_grid);
    const R* w = fp32 ? reinterpret_cast<const R*>(d_w32.p) : reinterpret_cast<const R*>(d_w64.p);
    SpreadGeom g;
    g.nx = S.n[0];
    g.ny = S.n[1];
    g.nz = S.n[2];
    g.half = S.half;
    g.width = S.width;
    g.nofold = (S.n[0] >= S.width && S.n[1] >= S.width && S.n[2] >= S.width) ? 1 : 0;
    g.bx = binsz[0];
    g.by = binsz[1];
    g.bz = binsz[2];
    g.nbx = nbins[0];
    g.nby = nbins[1];
    g.nbz = nbins[2];
    g.ntx = ntiles[0];
    g.nty = ntiles[1];
    g.Zs = Zs;
    g.nyp = slab ? nyp : S.n[1];
    const R* regw = fp32 ? reinterpret_cast<const R*>(d_regw32.p) : reinterpret_cast<const R*>(d_regw64.p);
    const int Ni = (int)N;
    RegionGeom rg;
    rg.nx = S.n[0];
    rg.nrr = reg_nry;
    rg.npb = reg_npb;
    rg.pblk = reg_pblk;
    if (SLAB) {
        rg.nrow = S.n[1];
        rg.nplane = Zs;
        rg.xs = (unsigned)nyp;
        rg.ps = (unsigned)(S.n[0] * nyp);
        rg.cs = (unsigned)((size_t)Zs * S.n[0] * nyp);
    } else {
        rg.nrow = S.n[2];
        rg.nplane = S.n[1];
        rg.xs = (unsigned)(S.n[1] * S.n[2]);
        rg.ps = (unsigned)S.n[2];
        rg.cs = (unsigned)V;
    }
    // the matrix-core kernels address the grid with 32-bit element offsets
    const int nreg = reg_nrx * reg_npb * reg_nry;
    // full grid: the region kernels need enough (region, column group) waves to fill the
    // GPU; with fewer than a quarter of the per-particle gather's waves (small grids and
    // batches, e.g. 512 particles on 32^3 with one right-hand side: 64 against 512) the
    // per-particle path is faster
    const long long region_waves = (long long)nreg * ((6 * B + 15) / 16);
    const bool full_ok = reg_full && 4 * region_waves >= (long long)N * ((B + 3) / 4);
    const bool regions = reg_rx > 0 && (SLAB || full_ok) &&
                         2ull * (unsigned long long)B * grid_elems_per_rhs() < (1ull << 32);
    // the matrix-core contractions load whole vectors of two complex when an even row
    // count and stride and an aligned grid keep every pair aligned and both of its rows
    // inside or outside together (16-byte stores in the spread measured slower: kept at
    // one complex per lane)
    const bool aligned = (reinterpret_cast<uintptr_t>(grid) % 32) == 0;
    const bool vec = aligned && (SLAB ? (nyp % 2 == 0 && S.n[1] % 2 == 0) : S.n[2] % 2 == 0);
    if (parts & MPM_PART_RECIP) {
        if (phases & 1) {
            if (regions && !SLAB) {
                ProfScope ps("spread", s, B);
                constexpr int CPB = 32;  // SpreadShape<R, ...>::CPB
                const dim3 gs((unsigned)nreg, (unsigned)((6 * B + CPB - 1) / CPB));
                ps.kernel("spread_mfma_kernel", {type_name<R>(), "4", std::to_string(kFullRegionPlanes)});
                hipLaunchKernelGGL((spread_mfma_kernel<R, 4, kFullRegionPlanes>), gs, dim3(256), 0, s, t, B, rg, regw,
                                   d_reg_start.p, d_reg_items.p, Ni, grid);
                MPM_HIP(hipGetLastError());
            } else if (regions) {
                ProfScope ps("spread", s, B);
                constexpr int CPB = 32;  // SpreadShape<R, ...>::CPB
                const dim3 gs((unsigned)nreg, (unsigned)((6 * B + CPB - 1) / CPB));
    #define MPM_SM(RXV, ZSV)                                                                                          \
        case ZSV:                                                                                                     \
            ps.kernel("spread_mfma_kernel", {type_name<R>(), #RXV, #ZSV});                                            \
            hipLaunchKernelGGL((spread_mfma_kernel<R, RXV, ZSV>), gs, dim3(256), 0, s, t, B, rg, regw, d_reg_start.p, \
                               d_reg_items.p, Ni, grid);                                                             \
            break;
                switch (zs_bucket(Zs)) {
                    MPM_SM(4, 4)
                    MPM_SM(4, 6)
                    MPM_SM(4, 8)
                    MPM_SM(4, 10)
                    MPM_SM(2, 12)
                    MPM_SM(2, 14)
                    MPM_SM(2, 16)
                    default:
                        throw Error(MPM_ERR_RUNTIME, "slab size out of range");
                }
    #undef MPM_SM
                MPM_HIP(hipGetLastError());
            } else {
                ProfScope ps("spread", s, B);
                ps.kernel("spread_kernel", {type_name<R>(), B == 1 ? "1" : "4", tbool(SLAB)});
                const int nt = ntiles[0] * ntiles[1] * n_ztiles;
                if (B == 1)
                    hipLaunchKernelGGL((spread_kernel<R, 1, SLAB>), dim3(nt, 1), dim3(256), 0, s, t, B, g, d_base.p, w,
                                       d_bin_start.p, d_bin_items.p, d_ztiles.p, d_zmap.p, Ni, grid);
                else
                    hipLaunchKernelGGL((spread_kernel<R, 4, SLAB>), dim3(nt, (B + 3) / 4), dim3(256), 0, s, t, B, g,
                                       d_base.p, w, d_bin_start.p, d_bin_items.p, d_ztiles.p, d_zmap.p, Ni, grid);
                MPM_HIP(hipGetLastError());
            }
        }
        if (phases & 2) {
            const size_t bytes_per_rhs = grid_elems_per_rhs() * sizeof(CT);
            CT* kgrid = grid;
            if (kt) kgrid = reinterpret_cast<CT*>(cur_kgrid);
            // contiguous slab, transposed k-space rows, a two-stage x transform whose rows
            // fit in LDS: x transform, plane coupling and inverse x transform in one kernel
            const int xaxis = (SLAB && zcontig && kt && lfft && fft_backend == 0) ? lfft->kt_x : -1;
            const size_t xlds = (size_t)(3 * Zs + 2) * S.n[0] * sizeof(CT);
            const bool fusedx = fp32 && xaxis >= 0 && lfft->axes[xaxis]->two && xlds <= kMaxFusedLds;
            // full grid (FP32, built-in transform, two-stage z lines): the z transforms and
            // the Green's tensor in one pass over z lines of three components
            const int zaxis = (!SLAB && fp32 && lfft && fft_backend == 0) ? lfft->full_z : -1;
            int ztl = 0;
            if (zaxis >= 0 && lfft->axes[zaxis]->two) {
                const LinePlan& pl = lfft->axes[zaxis]->plan;
                const size_t per_line = 3 * (size_t)((pl.n + pl.radix[1]) | 1) * sizeof(CT);  // padded, 3 components
                // twiddles and their conjugates, and 1 KB of the kernel's static line tables
                const size_t fixed = 2 * (size_t)pl.n * sizeof(CT) + 1024;
                const int by_lds = kMaxFusedLds > fixed ? (int)((kMaxFusedLds - fixed) / per_line) : 0;
                const int by_threads = 256 / (3 * pl.radix[1]);  // one inverse stage-1 butterfly per thread
                // about 10 KB of lines per component and block (measured at 256^3 with
                // padded rows: 5 lines 2.39 ms, 3 lines 2.55 ms)
                ztl = std::min({by_lds, by_threads, std::max(1, 1280 / pl.n)});
            }
            const bool fusedz = ztl >= 1;
            const int skip = fusedx ? xaxis : (fusedz ? zaxis : -1);
            {
                ProfScope ps("fft", s, B);
                if (lfft && fft_backend == 0)
                    lfft->exec<R>(grid, kgrid, B, true, s, skip);
                else
                    fft->exec(grid, bytes_per_rhs, B, fp32, true, s);
            }
            if (fusedz) {
                ProfScope ps("scale", s, B);
                const LineAxis& ax = *lfft->axes[zaxis];
                const int nxy = S.n[0] * S.n[1], nlines = nxy * B;
                const size_t lds =
                    (3 * (size_t)ztl * ((ax.plan.n + ax.plan.radix[1]) | 1) + 2 * (size_t)ax.plan.n) * sizeof(CT);
                const CT* tw = reinterpret_cast<const CT*>(ax.tw32.p);
                const R* kc = reinterpret_cast<const R*>(d_k32.p);
                const dim3 gz((unsigned)((nlines + ztl - 1) / ztl));
                ps.kernel("zline_scale_kernel", {type_name<R>(), std::to_string(ax.cls)});
                if (ax.cls == 0)
                    hipLaunchKernelGGL((zline_scale_kernel<R, 0>), gz, dim3(256), lds, s, grid, B, nxy,
                                       ax.plan.n, ax.plan.radix[0], ax.plan.radix[1], tw, ztl, kc);
                else if (ax.cls == 1)
                    hipLaunchKernelGGL((zline_scale_kernel<R, 1>), gz, dim3(256), lds, s, grid, B, nxy,
                                       ax.plan.n, ax.plan.radix[0], ax.plan.radix[1], tw, ztl, kc);
                else
                    hipLaunchKernelGGL((zline_scale_kernel<R, 2>), gz, dim3(256), lds, s, grid, B, nxy,
                                       ax.plan.n, ax.plan.radix[0], ax.plan.radix[1], tw, ztl, kc);
                MPM_HIP(hipGetLastError());
            } else if (fusedx) {
                ProfScope ps("zfused", s, B);
                const CT* T = reinterpret_cast<const CT*>(d_toe32.p);
                const LineAxis& ax = *lfft->axes[xaxis];
                const CT* tw = reinterpret_cast<const CT*>(ax.tw32.p);
                const bool ri = (S.n[2] & 1) != 0;
                const dim3 gx((unsigned)(8 * ((S.n[1] + 7) / 8) * B));  // XCD-major row order
    #define MPM_XT(ZSV)                                                                                              \
        case ZSV:                                                                                                    \
            ps.kernel("xtoeplitz_kernel", {type_name<R>(), #ZSV, "2", tbool(ri)});                                   \
            if (ri)                                                                                                  \
                hipLaunchKernelGGL((xtoeplitz_kernel<R, ZSV, 2, true>), gx, dim3(256), xlds, s, kgrid, Zs, S.n[1],   \
                                   nxp, S.n[0], ax.plan.radix[0], ax.plan.radix[1], tw, T, d_toe32ri.p);            \
            else                                                                                                     \
                hipLaunchKernelGGL((xtoeplitz_kernel<R, ZSV, 2, false>), gx, dim3(256), xlds, s, kgrid, Zs, S.n[1],  \
                                   nxp, S.n[0], ax.plan.radix[0], ax.plan.radix[1], tw, T, nullptr);                \
            break;
                if constexpr (std::is_same<R, float>::value) {
                    switch (zs_bucket(Zs)) {
                        MPM_XT(4)
                        MPM_XT(6)
                        MPM_XT(8)
                        MPM_XT(10)
                        MPM_XT(12)
                        MPM_XT(14)
                        MPM_XT(16)
                        default:
                            throw Error(MPM_ERR_RUNTIME, "slab size out of range");
                    }
                }
    #undef MPM_XT
                MPM_HIP(hipGetLastError());
            } else if (SLAB && zcontig) {
                ProfScope ps("zfused", s, B);
                const int ncols = kspace_cols();
                const CT* T = reinterpret_cast<const CT*>(fp32 ? static_cast<const void*>(d_toe32.p)
                                                               : static_cast<const void*>(d_toe64.p));
                const int rb = B >= 16 ? 16 : (B >= 4 ? 4 : 1);
                const int cb = 256 / rb;
                const dim3 gridz((ncols + cb - 1) / cb, (B + rb - 1) / rb);
                const bool ri = (S.n[2] & 1) != 0;  // odd nz: real / imaginary table entries
    #define MPM_ZT(ZSV)                                                                                           \
        case ZSV:                                                                                                 \
            ps.kernel("ztoeplitz_kernel", {type_name<R>(), #ZSV, std::to_string(ztoe_as<R, ZSV>()), tbool(ri)}); \
            if (ri)                                                                                               \
                hipLaunchKernelGGL((ztoeplitz_kernel<R, ZSV, ztoe_as<R, ZSV>(), true>), gridz, dim3(256), 0, s,   \
                                   kgrid, B, ncols, Zs, rb, T);                                                   \
            else                                                                                                  \
                hipLaunchKernelGGL((ztoeplitz_kernel<R, ZSV, ztoe_as<R, ZSV>(), false>), gridz, dim3(256), 0, s,  \
                                   kgrid, B, ncols, Zs, rb, T);                                                   \
            break;
                switch (zs_bucket(Zs)) {
                    MPM_ZT(4)
                    MPM_ZT(6)
                    MPM_ZT(8)
                    MPM_ZT(10)
                    MPM_ZT(12)
                    MPM_ZT(14)
                    MPM_ZT(16)
                    default:
                        throw Error(MPM_ERR_RUNTIME, "slab size out of range");
                }
    #undef MPM_ZT
                MPM_HIP(hipGetLastError());
            } else if (SLAB) {
                ProfScope ps("zfused", s, B);
                const int ncols = kspace_cols();
                const int zsb = zs_bucket(Zs);
                const size_t lds = zfused_lds(zsb, S.n[2], sizeof(R));
                const R* kc = fp32 ? reinterpret_cast<const R*>(d_k32.p) : reinterpret_cast<const R*>(d_k64.p);
                const int rb = B >= 16 ? 16 : (B >= 4 ? 4 : 1);
                const dim3 gridz((ncols + 256 / rb - 1) / (256 / rb), (B + rb - 1) / rb);
    #define MPM_ZF2(ZSV, RBV)                                                                                     \
        do {                                                                                                      \
            ps.kernel("zfused_kernel", {type_name<R>(), #ZSV, #RBV});                                             \
            hipLaunchKernelGGL((zfused_kernel<R, ZSV, RBV>), gridz, dim3(256), lds, s, kgrid, B, ncols, S.n[2],  \
                               Zs, d_zlist.p, kc);                                                                \
        } while (0)
    #define MPM_ZF(ZSV)                     \
        case ZSV:                           \
            if (rb == 16)                   \
                MPM_ZF2(ZSV, 16);           \
            else if (rb == 4)               \
                MPM_ZF2(ZSV, 4);            \
            else                            \
                MPM_ZF2(ZSV, 1);            \
            break;
                switch (zsb) {
                    MPM_ZF(4)
                    MPM_ZF(6)
                    MPM_ZF(8)
                    MPM_ZF(10)
                    MPM_ZF(12)
                    MPM_ZF(14)
                    MPM_ZF(16)
                    default:
                        throw Error(MPM_ERR_RUNTIME, "slab size out of range");
                }
    #undef MPM_ZF
    #undef MPM_ZF2
                MPM_HIP(hipGetLastError());
            } else {
                ProfScope ps("scale", s, B);
                ps.kernel("scale_kernel", {type_name<R>()});
                const int blocks = (int)std::min<size_t>((V + 255) / 256, 8192);
                const R* kc = fp32 ? reinterpret_cast<const R*>(d_k32.p) : reinterpret_cast<const R*>(d_k64.p);
                hipLaunchKernelGGL(scale_kernel<R>, dim3(blocks), dim3(256), 0, s, grid, kc, V, B);
                MPM_HIP(hipGetLastError());
            }
            {
                ProfScope ps("fft", s, B);
                if (lfft && fft_backend == 0)
                    lfft->exec<R>(grid, kgrid, B, false, s, skip);
                else
                    fft->exec(grid, bytes_per_rhs, B, fp32, false, s);
            }
        }
    }
    if (!(phases & 4)) return;
    const R* part = nullptr;
    if (regions && (parts & MPM_PART_RECIP)) {
        ProfScope ps("contract", s, B);
        // column groups of 16 per block: 4 when there are enough right-hand sides,
        // else fewer, with the region's planes split over the remaining waves; slab
        // with one or two column groups: one wave per candidate block instead
        const int ngrp = (6 * B + 15) / 16;
        const int ncg = ngrp >= 4 ? 4 : (ngrp >= 2 ? 2 : 1);
        const dim3 gs((unsigned)nreg, (unsigned)((ngrp + ncg - 1) / ncg));
        const dim3 gn((unsigned)nreg, (unsigned)ngrp);
        R* pbuf = reinterpret_cast<R*>(d_partial64.p);  // sized for FP64, shared by FP32
        if (!SLAB && vec) {
            ps.kernel("contract_full_mfma_kernel",
                      {type_name<R>(), "4", std::to_string(kFullRegionPlanes), std::to_string(kFullContractNB)});
            const int nitems = nreg * ngrp;
            hipLaunchKernelGGL((contract_full_mfma_kernel<R, 4, kFullRegionPlanes, kFullContractNB>),
                               dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, s, B, rg, regw, d_reg_start.p, nitems,
                               ngrp, grid, pbuf);
        } else if (!SLAB) {
            ps.kernel("contract_mfma_kernel", {type_name<R>(), "4", std::to_string(kFullRegionPlanes),
                                               std::to_string(kFullContractNB), "false"});
            hipLaunchKernelGGL((contract_mfma_kernel<R, 4, kFullRegionPlanes, kFullContractNB, false>), gs, dim3(256),
                               0, s, B, rg, regw, d_reg_start.p, d_reg_items.p, Ni, grid, pbuf, ncg);
        } else {
#define MPM_CM(RXV, ZSV)                                                                                          \
    case ZSV:                                                                                                     \
        if (ngrp <= 2) {                                                                                          \
            ps.kernel("contract_nb_mfma_kernel", {type_name<R>(), #RXV, #ZSV, tbool(vec)});                       \
            if (vec)                                                                                              \
                hipLaunchKernelGGL((contract_nb_mfma_kernel<R, RXV, ZSV, true>), gn, dim3(256), 0, s, B, rg,      \
                                   regw, d_reg_start.p, d_reg_items.p, Ni, grid, pbuf);                          \
            else                                                                                                  \
                hipLaunchKernelGGL((contract_nb_mfma_kernel<R, RXV, ZSV, false>), gn, dim3(256), 0, s, B, rg,     \
                                   regw, d_reg_start.p, d_reg_items.p, Ni, grid, pbuf);                          \
        } else if (vec) {                                                                                         \
            ps.kernel("contract_mfma_kernel", {type_name<R>(), #RXV, #ZSV, "4", "true"});                         \
            hipLaunchKernelGGL((contract_mfma_kernel<R, RXV, ZSV, 4, true>), gs, dim3(256), 0, s, B, rg, regw,    \
                               d_reg_start.p, d_reg_items.p, Ni, grid, pbuf, ncg);                               \
        } else {                                                                                                  \
            ps.kernel("contract_mfma_kernel", {type_name<R>(), #RXV, #ZSV, "4", "false"});                        \
            hipLaunchKernelGGL((contract_mfma_kernel<R, RXV, ZSV, 4, false>), gs, dim3(256), 0, s, B, rg, regw,   \
                               d_reg_start.p, d_reg_items.p, Ni, grid, pbuf, ncg);                               \
        }                                                                                                         \
        break;
        switch (zs_bucket(Zs)) {
            MPM_CM(4, 4)
            MPM_CM(4, 6)
            MPM_CM(4, 8)
            MPM_CM(4, 10)
            MPM_CM(2, 12)
            MPM_CM(2, 14)
            MPM_CM(2, 16)
            default:
                throw Error(MPM_ERR_RUNTIME, "slab size out of range");
        }
#undef MPM_CM
        }
        MPM_HIP(hipGetLastError());
        part = pbuf;
    }
    if (part && B >= 4) {
        ProfScope ps("gather", s, B);
        ps.kernel("finish_kernel", {type_name<R>()});
        const double h3 = S.h[0] * S.h[1] * S.h[2];
        const int blocks = (int)(((size_t)Ni * B * kFinishLanes + 255) / 256);
        hipLaunchKernelGGL((finish_kernel<R>), dim3(blocks), dim3(256), 0, s, t, B, parts, Ni, part, d_ps_start.p,
                           d_ps_items.p, h3, d_nb_start.p, d_nb_j.p, d_pairc.p, d_selfc.p, self_perp, jacobi,
                           d_order.p);
        MPM_HIP(hipGetLastError());
    } else {
        ProfScope ps("gather", s, B);
        ps.kernel("gather_kernel", {type_name<R>(), B == 1 ? "1" : "4", tbool(SLAB)});
        const double h3 = S.h[0] * S.h[1] * S.h[2];
        const int blocks = (Ni + 3) / 4;
        if (B == 1)
            hipLaunchKernelGGL((gather_kernel<R, 1, SLAB>), dim3(blocks, 1), dim3(256), 0, s, t, B, parts, grid, g,
                               d_zmap.p, d_base.p, w, Ni, h3, d_pos.p, d_nb_start.p, d_nb_j.p, d_nb_f.p, L[0], L[1],
                               L[2], kb[0], kb[1], d_selfc.p, self_perp, jacobi, part, d_ps_start.p, d_ps_items.p,
                               d_order.p);
        else
            hipLaunchKernelGGL((gather_kernel<R, 4, SLAB>), dim3(blocks, (B + 3) / 4), dim3(256), 0, s, t, B, parts,
                               grid, g,
                               d_zmap.p, d_base.p, w, Ni, h3, d_pos.p, d_nb_start.p, d_nb_j.p, d_nb_f.p, L[0], L[1],
                               L[2], kb[0], kb[1], d_selfc.p, self_perp, jacobi, part, d_ps_start.p, d_ps_items.p,
                               d_order.p);
        MPM_HIP(hipGetLastError());
    }
}

static_assert(kMaxSlots % 4 == 0, "the polydisperse kernels take right-hand sides in groups of 4");

template <typename R>
void EwaldOp::launch_quads(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi) {
    using CT = typename C2<R>::T;
    const bool fp32 = std::is_same<R, float>::value;
    CT* grid = reinterpret_cast<CT*>(cur_grid);
    const R* sw = fp32 ? reinterpret_cast<const R*>(d_sw32.p) : reinterpret_cast<const R*>(d_sw64.p);
    const int SS = S.width * S.width * S.width;
    const int Vi = (int)V, Ni = (int)N;
    constexpr int G = 2;
    if ((parts & MPM_PART_RECIP) && (phases & 1)) {
        ProfScope ps("spread", s, B);
        ps.kernel("quad_spread_kernel", {type_name<R>(), "2"});
        hipLaunchKernelGGL((quad_spread_kernel<R, G>), dim3((unsigned)((V + 255) / 256), (unsigned)((B + G - 1) / G)),
                           dim3(256), 0, s, t, B, Ni, Vi, SS, d_vox_start.p, d_vox_items.p, sw, d_qmap.p, grid);
        MPM_HIP(hipGetLastError());
    }
    if ((parts & MPM_PART_RECIP) && (phases & 2)) {
        {
            ProfScope ps("fft", s, B);
            lfft->exec<R>(grid, grid, B, true, s);
        }
        {
            ProfScope ps("scale", s, B);
            ps.kernel("quad_scale_kernel", {type_name<R>()});
            const R* kq = fp32 ? reinterpret_cast<const R*>(d_kq32.p) : reinterpret_cast<const R*>(d_kq64.p);
            const size_t total = (size_t)B * V;
            const int blocks = (int)std::min<size_t>((total + 255) / 256, 16384);
            hipLaunchKernelGGL((quad_scale_kernel<R>), dim3(blocks), dim3(256), 0, s, grid, kq, V, total);
            MPM_HIP(hipGetLastError());
        }
        {
            ProfScope ps("fft", s, B);
            lfft->exec<R>(grid, grid, B, false, s);
        }
    }
    if (!(phases & 4)) return;
    ProfScope ps("gather", s, B);
    ps.kernel("quad_gather_kernel", {type_name<R>(), "2"});
    const double h3 = S.h[0] * S.h[1] * S.h[2];
    hipLaunchKernelGGL((quad_gather_kernel<R, G>), dim3((unsigned)((Ni + 3) / 4), (unsigned)((B + G - 1) / G)),
                       dim3(256), 0, s, t, B, parts, Ni, Vi, SS, grid, sw, d_pidx.p, d_qmap.p, h3, d_nb_start.p,
                       d_nb_j.p, d_pairc.p, d_selfc.p, self_perp, table->self_grad, jacobi, d_order.p);
    MPM_HIP(hipGetLastError());
}

template <typename R>
void EwaldOp::launch_poly(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi) {
    using CT = typename C2<R>::T;
    const bool fp32 = std::is_same<R, float>::value;
    CT* grid = reinterpret_cast<CT*>(cur_grid);
    const R* pw = fp32 ? reinterpret_cast<const R*>(d_pw32.p) : reinterpret_cast<const R*>(d_pw64.p);
    const int SS = S.P * S.P * S.P;
    const int Vi = (int)V;
    const int Ni = (int)N;
    constexpr int G = 4;
    if ((parts & MPM_PART_RECIP) && (phases & 1)) {
        ProfScope ps("spread", s, B);
        ps.kernel("poly_spread_kernel", {type_name<R>(), "4"});
        hipLaunchKernelGGL((poly_spread_kernel<R, G>), dim3((unsigned)((V + 255) / 256), (unsigned)((B + G - 1) / G)),
                           dim3(256), 0, s, t, B, Vi, SS, d_vox_start.p, d_vox_items.p, pw, grid);
        MPM_HIP(hipGetLastError());
    }
    if ((parts & MPM_PART_RECIP) && (phases & 2)) {
        {
            ProfScope ps("fft", s, B);
            lfft->exec<R>(grid, grid, B, true, s);
        }
        {
            ProfScope ps("scale", s, B);
            ps.kernel("poly_scale_kernel", {type_name<R>()});
            const R* kc = fp32 ? reinterpret_cast<const R*>(d_k32.p) : reinterpret_cast<const R*>(d_k64.p);
            const size_t total = (size_t)B * V;
            const int blocks = (int)std::min<size_t>((total + 255) / 256, 16384);
            hipLaunchKernelGGL((poly_scale_kernel<R>), dim3(blocks), dim3(256), 0, s, grid, kc, V, total);
            MPM_HIP(hipGetLastError());
        }
        {
            ProfScope ps("fft", s, B);
            lfft->exec<R>(grid, grid, B, false, s);
        }
    }
    if (!(phases & 4)) return;
    ProfScope ps("gather", s, B);
    ps.kernel("poly_gather_kernel", {type_name<R>(), "4"});
    const double h3 = S.h[0] * S.h[1] * S.h[2];
    hipLaunchKernelGGL((poly_gather_kernel<R, G>), dim3((unsigned)((Ni + 3) / 4), (unsigned)((B + G - 1) / G)),
                       dim3(256), 0, s, t, B, parts, Ni, Vi, SS, grid, pw, d_pidx.p, h3, d_nb_start.p, d_nb_j.p,
                       d_pairc.p, d_selfc.p, d_selfp.p, jacobi, d_order.p);
    MPM_HIP(hipGetLastError());
}

void EwaldOp::radius_gradient(const double2* x, const double2* g, int B, double* out, bool accumulate,
                              hipStream_t s) {
    MPM_REQUIRE(poly, MPM_ERR_UNSUPPORTED, "radius gradients need the polydisperse operator");
    MPM_REQUIRE(wrad_h.empty(), MPM_ERR_UNSUPPORTED, "radius gradients need one radius per particle");
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "radius gradient before set_positions");
    MPM_REQUIRE(kb[0] == 0.0 && kb[1] == 0.0, MPM_ERR_UNSUPPORTED,
                "radius gradients are provided without a Bloch wave vector");
    MPM_REQUIRE(B >= 1, MPM_ERR_INVALID_ARGUMENT, "B must be >= 1");
    ensure_rgrad_data();
    ensure_kspace_table();
    const int chunk = 8;
    const int Bc = std::min(B, chunk);
    d_rgx.reserve((size_t)Bc * V);
    d_rgg.reserve((size_t)Bc * V);
    const size_t n3 = n_unknowns();
    const int SS = S.P * S.P * S.P;
    const double h3 = S.h[0] * S.h[1] * S.h[2];
    struct PhaseReset {
        int& p;
        ~PhaseReset() { p = 7; }
    } reset{phases};
    for (int b0 = 0; b0 < B; b0 += chunk) {
        const int cnt = std::min(chunk, B - b0);
        SlotTable tx, tg;
        for (int j = 0; j < cnt; ++j) {
            tx.x[j] = x + (size_t)(b0 + j) * n3;
            tg.x[j] = g + (size_t)(b0 + j) * n3;
            tx.y[j] = tg.y[j] = nullptr;
            tx.row[j] = tg.row[j] = 0;
        }
        phases = 3;  // spread, transform, scale, inverse transform: the convolved grid
        cur_grid = d_rgx.p;
        launch_poly<double>(tx, cnt, MPM_PART_RECIP, s, 0);
        cur_grid = d_rgg.p;
        launch_poly<double>(tg, cnt, MPM_PART_RECIP, s, 0);
        phases = 7;
        hipLaunchKernelGGL(poly_rgrad_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, s, cnt, (int)N, (int)V, SS,
                           n3, x + (size_t)b0 * n3, g + (size_t)b0 * n3, d_rgx.p, d_rgg.p, d_dpw.p, d_pidx.p, h3,
                           d_nb_start.p, d_nb_j.p, d_pairc.p, d_dpair.p, d_dps.p, out, (accumulate || b0 > 0) ? 1 : 0);
        MPM_HIP(hipGetLastError());
    }
}

template <typename R>
void EwaldOp::launch_direct(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi) {
    MPM_REQUIRE(parts & MPM_PART_REAL, MPM_ERR_INVALID_ARGUMENT,
                "the direct field has no reciprocal part (parts must include MPM_PART_REAL)");
    ProfScope ps("direct", s, B);
    const int Ni = (int)N;
    const dim3 blk(256);
    if (quads) {
        constexpr int G = 2;
        ps.kernel("direct_kernel", {type_name<R>(), "2", "true"});
        hipLaunchKernelGGL((direct_kernel<R, G, true>), dim3((unsigned)((Ni + 255) / 256), (unsigned)((B + G - 1) / G)),
                           blk, 0, s, t, B, Ni, d_pos.p, d_drad.p, d_qmap.p, d_selfc.p, jacobi);
    } else {
        constexpr int G = 4;
        ps.kernel("direct_kernel", {type_name<R>(), "4", "false"});
        hipLaunchKernelGGL((direct_kernel<R, G, false>), dim3((unsigned)((Ni + 255) / 256), (unsigned)((B + G - 1) / G)),
                           blk, 0, s, t, B, Ni, d_pos.p, d_drad.p, nullptr, d_selfc.p, jacobi);
    }
    MPM_HIP(hipGetLastError());
}

__global__ void zero_words_kernel(uint4* __restrict__ p, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        p[e] = make_uint4(0u, 0u, 0u, 0u);
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t EwaldOp::scratch_bytes(int B, bool fp32) const {
    const size_t cs = fp32 ? sizeof(float2) : sizeof(double2);
    return align256((size_t)B * grid_elems_per_rhs() * cs) + (kt ? align256((size_t)B * kgrid_elems_per_rhs() * cs) : 0);
}

bool EwaldOp::builtin_fft() const { return lfft && fft_backend == 0; }

void EwaldOp::apply(const SlotTable& t, int B, bool fp32, int parts, hipStream_t s, bool jacobi,
                    const Workspace* ws) {
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "apply before set_positions");
    MPM_REQUIRE(B >= 1 && B <= kMaxSlots, MPM_ERR_INVALID_ARGUMENT, "bad slot count");
    for (int b = 0; b < B; ++b)
        MPM_REQUIRE(t.row[b] >= 0 && t.row[b] < self_rows, MPM_ERR_INVALID_ARGUMENT,
                    "self-coefficient row out of range");
    const size_t elems = (size_t)B * grid_elems_per_rhs();
    const size_t kelems = (size_t)B * kgrid_elems_per_rhs();
    if (direct) {
        // no grids
    } else if (ws) {
        // caller's workspace: nothing may allocate, so the k-space tables must already be
        // current (set_positions and set_bloch build them) and the FFT must be the built-in one
        MPM_REQUIRE(!(parts & MPM_PART_RECIP) || (kspace_ready && builtin_fft()), MPM_ERR_UNSUPPORTED,
                    "an apply without allocation needs the built-in FFT and current k-space tables");
        const size_t need = scratch_bytes(B, fp32);
        MPM_REQUIRE(need <= ws->bytes, MPM_ERR_INVALID_ARGUMENT,
                    "the apply needs " + std::to_string(need) + " bytes of scratch, more than the " +
                        std::to_string(ws->bytes) + " available");
        const size_t cs = fp32 ? sizeof(float2) : sizeof(double2);
        cur_grid = ws->base;
        cur_kgrid = kt ? ws->base + align256(elems * cs) : nullptr;
        // the grids hold the previous call's data: start from zeros as freshly cleared memory
        if (parts & MPM_PART_RECIP) {
            const size_t words = need / sizeof(uint4);
            const int blocks = (int)std::min<size_t>((words + 255) / 256, 8192);
            hipLaunchKernelGGL(zero_words_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<uint4*>(ws->base),
                               words);
            MPM_HIP(hipGetLastError());
        }
    } else {
        if (parts & MPM_PART_RECIP) ensure_kspace_table();
        if (fp32) {
            d_grid32.reserve(elems);
            if (kt) d_kgrid32.reserve(kelems);
            cur_grid = d_grid32.p;
            cur_kgrid = d_kgrid32.p;
        } else {
            d_grid64.reserve(elems);
            if (kt) d_kgrid64.reserve(kelems);
            cur_grid = d_grid64.p;
            cur_kgrid = d_kgrid64.p;
        }
    }
    const int jac = jacobi ? 1 : 0;
    if (direct) {
        if (fp32) launch_direct<float>(t, B, parts, s, jac);
        else launch_direct<double>(t, B, parts, s, jac);
        return;
    }
    if (quads) {
        if (fp32) launch_quads<float>(t, B, parts, s, jac);
        else launch_quads<double>(t, B, parts, s, jac);
        return;
    }
    if (poly) {
        if (fp32) launch_poly<float>(t, B, parts, s, jac);
        else launch_poly<double>(t, B, parts, s, jac);
        return;
    }
    if (fp32) {
        if (slab) launch<float, true>(t, B, parts, s, jac);
        else launch<float, false>(t, B, parts, s, jac);
    } else {
        if (slab) launch<double, true>(t, B, parts, s, jac);
        else launch<double, false>(t, B, parts, s, jac);
    }
}

void EwaldOp::apply_contiguous(const double2* x, double2* y, int B, bool fp32, int parts, hipStream_t s,
                               const int* rows, const Workspace* ws) {
    const size_t n = n_unknowns();
    for (int b0 = 0; b0 < B; b0 += kMaxSlots) {
        const int cnt = std::min(kMaxSlots, B - b0);
        SlotTable t;
        for (int j = 0; j < cnt; ++j) {
            t.x[j] = x + (size_t)(b0 + j) * n;
            t.y[j] = y + (size_t)(b0 + j) * n;
            t.row[j] = rows ? rows[b0 + j] : (self_rows > 1 ? b0 + j : 0);
        }
        apply(t, cnt, fp32, parts, s, false, ws);
    }
}

void EwaldOp::configure(int layout, int backend, const int* grid) {
    MPM_REQUIRE(!direct, MPM_ERR_UNSUPPORTED, "the direct field has no reciprocal-space configuration");
    MPM_REQUIRE(layout == 0 || layout == 1, MPM_ERR_INVALID_ARGUMENT, "layout must be 0 or 1");
    MPM_REQUIRE(backend == 0 || backend == 1, MPM_ERR_INVALID_ARGUMENT, "fft_backend must be 0 or 1");
    const bool regrid = grid && (grid[0] != S.n[0] || grid[1] != S.n[1] || grid[2] != S.n[2]);
    if (poly || quads)
        MPM_REQUIRE(layout == 1 && backend == 0, MPM_ERR_UNSUPPORTED,
                    "the polydisperse and quadrupole operators run on the full grid with the built-in FFT");
    MPM_REQUIRE(!(poly && regrid), MPM_ERR_UNSUPPORTED, "the polydisperse operator keeps its grid");
    if (regrid) {
        Sizing T;
        compute_sizing(L, errortol, xi, grid, S.P, T);  // same support, spacing from the new grid
        MPM_REQUIRE(T.width <= kMaxWidth, MPM_ERR_UNSUPPORTED, "stencil too wide");
        S = T;
        V = (size_t)S.n[0] * S.n[1] * S.n[2];
    }
    force_full = layout;
    fft_backend = backend;
    lfft.reset();
    fft.reset();
    fft_key = -2;
    kspace_ready = false;
    kspace_key = -1;
    if (N > 0) {
        const std::vector<double> xyz = last_xyz;
        const int64_t n = N;
        last_xyz.clear();
        set_positions(xyz.data(), n);
    }
}

__global__ void add_kernel(double2* __restrict__ y, const double2* __restrict__ t, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        y[e].x += t[e].x;
        y[e].y += t[e].y;
    }
}

size_t EwaldOp::stage_grid_bytes(bool fp32) const {
    return grid_elems_per_rhs() * (fp32 ? sizeof(float2) : sizeof(double2));
}

void EwaldOp::stage(int which, const double2* x, void* grid, double2* y, bool fp32, hipStream_t s) {
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "electricField: particle coordinates have not been set");
    MPM_REQUIRE(which >= 1 && which <= 4, MPM_ERR_INVALID_ARGUMENT, "unknown stage");
    MPM_REQUIRE(self_rows >= 1, MPM_ERR_RUNTIME, "set the self coefficients first");
    const bool recip = which != 4;
    MPM_REQUIRE(!(recip && direct), MPM_ERR_UNSUPPORTED, "the direct field has no reciprocal-space stages");
    MPM_REQUIRE(!recip || grid, MPM_ERR_INVALID_ARGUMENT, "stage needs a grid buffer");
    MPM_REQUIRE(which == 2 || (x != nullptr && (which == 1 || y != nullptr)), MPM_ERR_INVALID_ARGUMENT,
                "stage needs the dipoles (and, to add into, the field)");
    if (recip) ensure_kspace_table();
    // spread writes the grid, scale transforms it in place; contract and real space add into y
    SlotTable t;
    t.x[0] = x;
    t.row[0] = 0;
    const size_t n = n_unknowns();
    DevBuf<double2> tmp;
    if (which >= 3) {
        tmp.resize(n);
        t.y[0] = tmp.p;
    } else {
        t.y[0] = nullptr;
    }
    cur_grid = grid;
    cur_kgrid = nullptr;
    if (kt && which == 2) {
        const size_t ke = kgrid_elems_per_rhs();
        if (fp32) {
            d_kgrid32.reserve(ke);
            cur_kgrid = d_kgrid32.p;
        } else {
            d_kgrid64.reserve(ke);
            cur_kgrid = d_kgrid64.p;
        }
    }
    struct Phases {
        int& p;
        Phases(int& q, int v) : p(q) { p = v; }
        ~Phases() { p = 7; }
    } ph(phases, which == 1 ? 1 : which == 2 ? 2 : 4);
    const int parts = which == 4 ? MPM_PART_REAL : MPM_PART_RECIP;
    if (direct) {
        if (fp32) launch_direct<float>(t, 1, parts, s, 0);
        else launch_direct<double>(t, 1, parts, s, 0);
    } else if (quads) {
        if (fp32) launch_quads<float>(t, 1, parts, s, 0);
        else launch_quads<double>(t, 1, parts, s, 0);
    } else if (poly) {
        if (fp32) launch_poly<float>(t, 1, parts, s, 0);
        else launch_poly<double>(t, 1, parts, s, 0);
    } else if (fp32) {
        if (slab) launch<float, true>(t, 1, parts, s, 0);
        else launch<float, false>(t, 1, parts, s, 0);
    } else {
        if (slab) launch<double, true>(t, 1, parts, s, 0);
        else launch<double, false>(t, 1, parts, s, 0);
    }
    if (which >= 3) {
        hipLaunchKernelGGL(add_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, y,
                           tmp.p, n);
        MPM_HIP(hipGetLastError());
        MPM_HIP(hipStreamSynchronize(s));  // tmp is released on return
    }
}

void* EwaldOp::device_pointer(const std::string& name, size_t& bytes) const {
    auto give = [&](const void* p, size_t b) {
        bytes = p ? b : 0;
        return const_cast<void*>(p);
    };
    const size_t Nn = (size_t)N;
    if (name == "x_part") return give(d_pos.p, Nn * sizeof(double));
    if (name == "y_part") return give(d_pos.p ? d_pos.p + Nn : nullptr, Nn * sizeof(double));
    if (name == "z_part") return give(d_pos.p ? d_pos.p + 2 * Nn : nullptr, Nn * sizeof(double));
    if (name == "self_coef") return give(d_selfc.p, d_selfc.n * sizeof(double2));
    if (name == "radius") return give(d_drad.p, d_drad.n * sizeof(double));
    if (name == "spread_coef" || name == "contract_coef") {
        if (poly) return give(d_pw64.p, d_pw64.n * sizeof(double));
        if (quads) return give(d_sw64.p, d_sw64.n * sizeof(double));
        return give(d_w64.p, d_w64.n * sizeof(double));
    }
    if (name == "spread_coef_fp32" || name == "contract_coef_fp32") {
        if (poly) return give(d_pw32.p, d_pw32.n * sizeof(float));
        if (quads) return give(d_sw32.p, d_sw32.n * sizeof(float));
        return give(d_w32.p, d_w32.n * sizeof(float));
    }
    if (name == "spread_idxs" || name == "contract_idxs") {
        if (poly || quads) return give(d_pidx.p, d_pidx.n * sizeof(int));
        return give(d_base.p, d_base.n * sizeof(int));
    }
    if (name == "particle_index") return give(d_order.p, d_order.n * sizeof(int));
    if (name == "neighbor_start") return give(d_nb_start.p, d_nb_start.n * sizeof(int));
    if (name == "neighbor_list") return give(d_nb_j.p, d_nb_j.n * sizeof(int));
    if (name == "perp_para") return give(d_nb_f.p, d_nb_f.n * sizeof(double2));
    if (name == "pair_coef") return give(d_pairc.p, d_pairc.n * sizeof(double));
    if (name == "self_perp") return give(d_selfp.p, d_selfp.n * sizeof(double));
    if (name == "scale_coef") {
        if (quads) return give(d_kq64.p, d_kq64.n * sizeof(double));
        if (slab && zcontig) return give(d_toe64.p, d_toe64.n * sizeof(double2));
        return give(d_k64.p, d_k64.n * sizeof(double));
    }
    if (name == "fe_grid") return give(cur_grid, 0);
    if (name == "fes_grid") return give(d_kgrid64.p ? (void*)d_kgrid64.p : (void*)d_kgrid32.p,
                                        d_kgrid64.p ? d_kgrid64.n * sizeof(double2) : d_kgrid32.n * sizeof(float2));
    throw Error(MPM_ERR_INVALID_ARGUMENT, "unknown device buffer name '" + name + "'");
}

void EwaldOp::info(mpm_op_info& o) const {
    std::memset(&o, 0, sizeof(o));
    o.num_particles = N;
    o.num_pairs = num_pairs;
    for (int d = 0; d < 3; ++d) {
        o.grid[d] = S.n[d];
        o.grid_spacing[d] = S.h[d];
        o.spectral_split[d] = S.eta[d];
    }
    o.support = S.P;
    o.stencil_width = S.width;
    o.rc = S.rc;
    o.self_perp = self_perp;
    o.self_coef_ewald = -4.0 * std::pow(xi, 3.0) / (3.0 * std::sqrt(kPi));
    o.bloch[0] = kb[0];
    o.bloch[1] = kb[1];
    o.slab_planes = slab ? Zs : 0;
    o.fft_backend = (lfft && fft_backend == 0) ? 0 : 1;
    o.polydisperse = poly ? 1 : 0;
    o.num_quads = (int64_t)qidx_h.size();
    o.num_unique_radii = poly ? poly_classes : 1;
}

}  // namespace mpm
