// Monodisperse spectral-Ewald dipole operator y = A x for a batch of
// right-hand sides on gfx950.
//
// A = diag(self_coef_b + self_perp) + G_real + G_recip, the operator of
// src/cpp/electric_field/monodisperse_ewald_electric_field.cu (SOLVER_AX mode,
// dipoles).  The arithmetic follows the reference; the data movement does not:
//
//  * spread  — owner-computes gather: a 256-thread workgroup owns a 4x8x8
//              voxel tile and pulls in the particles binned around it (LDS
//              staged), so every grid value is written exactly once, with no
//              atomics and no separate zero-fill; the reference scatters with
//              LDS/global atomics after a memset (:1160-1298, :2000-2048).
//              Gaussian weights are separable, exp(-2 xi^2 sum_d gd_d^2/eta_d)
//              = prod_d exp(-2 xi^2 gd_d^2/eta_d), so each particle carries
//              3 x width weights instead of width^3.
//  * FFT     — rocFFT, 3-D, in place, one plan per (precision, batch bucket),
//              component-major grids [rhs][component][voxel].
//  * scale   — k-space Green's tensor projection f <- sc (f.k^) k^ (:1406-1432)
//              with (k^, sc) packed per voxel and reused across the batch.
//  * gather  — one wavefront per particle fuses the Gaussian contraction
//              (:1488-1525), the real-space neighbour sum (:1631-1869, written
//              as a gather over a symmetric CSR list so it needs no atomics)
//              and the self term (:1589-1629); a single wave reduction and a
//              single store per particle and RHS.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>

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
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::map<std::string, ProfEntry> g_prof;
struct Pending {
    std::string name;
    hipEvent_t e0, e1;
};
std::vector<Pending> g_pending;
}  // namespace

bool profiling_enabled() { return g_prof_on; }

ProfScope::ProfScope(const char* n, hipStream_t s) : name(n), stream(s) {
    if (!g_prof_on) return;
    MPM_HIP(hipEventCreate(&e0));
    MPM_HIP(hipEventCreate(&e1));
    MPM_HIP(hipEventRecord(e0, stream));
}
ProfScope::~ProfScope() {
    if (!e0) return;
    (void)hipEventRecord(e1, stream);
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_pending.push_back({name, e0, e1});
}

void profile_flush() {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (auto& p : g_pending) {
        (void)hipEventSynchronize(p.e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.e0, p.e1);
        auto& e = g_prof[p.name];
        e.ms += ms;
        e.count += 1;
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
bool profile_get(const std::string& name, double& ms, int64_t& count) {
    profile_flush();
    std::lock_guard<std::mutex> g(g_prof_mu);
    auto it = g_prof.find(name);
    if (it == g_prof.end()) {
        ms = 0;
        count = 0;
        return false;
    }
    ms = it->second.ms;
    count = it->second.count;
    return true;
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
// Radices 2, 3, 4, 5 and 7 run as register codelets; any other prime factor runs
// as a generic stage in which a thread computes whole outputs from the
// twiddle table (a single-stage direct DFT when n is prime).  Each stage stages
// its results in registers and writes back in place, so a block needs one LDS
// buffer of TL lines (<= 32 KB) plus the twiddles; one global read and one
// global write per element.
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 24;
constexpr int kLdsElems32 = 2048;  // lines x length per block, FP32
constexpr int kLdsElems64 = 2048;  // FP64

struct LinePlan {
    int n;
    int nstages;
    int radix[kMaxStages];
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

// x / d and x % d for 0 <= x < 2^22 via a float reciprocal and one correction.
__device__ inline void fdivmod(int x, int d, float rcp, int& q, int& r) {
    q = (int)((float)x * rcp);
    r = x - q * d;
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

// In-place DFT of RAD values, forward (sgn = -1) or inverse (sgn = +1).
template <typename R, int RAD>
__device__ inline void dft_small(typename C2<R>::T* v, R sgn) {
    using CT = typename C2<R>::T;
    if (RAD == 2) {
        const CT a = v[0], b = v[1];
        v[0].x = a.x + b.x; v[0].y = a.y + b.y;
        v[1].x = a.x - b.x; v[1].y = a.y - b.y;
    } else if (RAD == 4) {
        const CT a = v[0], b = v[1], c = v[2], d = v[3];
        CT t0, t1, t2, t3;
        t0.x = a.x + c.x; t0.y = a.y + c.y;
        t1.x = a.x - c.x; t1.y = a.y - c.y;
        t2.x = b.x + d.x; t2.y = b.y + d.y;
        t3.x = b.x - d.x; t3.y = b.y - d.y;
        // multiply t3 by sgn*i
        CT u;
        u.x = -sgn * t3.y;
        u.y = sgn * t3.x;
        v[0].x = t0.x + t2.x; v[0].y = t0.y + t2.y;
        v[2].x = t0.x - t2.x; v[2].y = t0.y - t2.y;
        v[1].x = t1.x + u.x; v[1].y = t1.y + u.y;
        v[3].x = t1.x - u.x; v[3].y = t1.y - u.y;
    } else {
        // direct DFT with compile-time roots cos/sin(2 pi k / RAD)
        R cs[RAD], sn[RAD];
#pragma unroll
        for (int k = 0; k < RAD; ++k) {
            const double ang = 2.0 * 3.14159265358979323846 * k / RAD;
            cs[k] = (R)__builtin_cos(ang);
            sn[k] = (R)__builtin_sin(ang);
        }
        CT out[RAD];
#pragma unroll
        for (int q = 0; q < RAD; ++q) {
            R ar = v[0].x, ai = v[0].y;
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                const int k = (r * q) % RAD;
                const R wr = cs[k], wi = sgn * sn[k];
                ar += v[r].x * wr - v[r].y * wi;
                ai += v[r].x * wi + v[r].y * wr;
            }
            out[q].x = ar;
            out[q].y = ai;
        }
#pragma unroll
        for (int q = 0; q < RAD; ++q) v[q] = out[q];
    }
}

// One Stockham stage with a register codelet of radix RAD, in place on buf.
template <typename R, int RAD>
__device__ inline void stage_codelet(typename C2<R>::T* buf, int n, int nl, int Ns, const typename C2<R>::T* tw,
                                     R sgn) {
    using CT = typename C2<R>::T;
    constexpr int KMAX = (kLdsElems32 / RAD + 255) / 256;
    const int nb = n / RAD;
    const int total = nl * nb;
    const int step = n / (Ns * RAD);
    const float rnb = 1.0f / (float)nb, rns = 1.0f / (float)Ns;
    CT v[KMAX][RAD];
    int dst[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const int bf = threadIdx.x + 256 * k;
        dst[k] = -1;
        if (bf < total) {
            int li, j, jd, a;
            fdivmod(bf, nb, rnb, li, j);
            fdivmod(j, Ns, rns, jd, a);
            const CT* src = buf + li * n + j;
#pragma unroll
            for (int r = 0; r < RAD; ++r) {
                CT x = src[r * nb];
                if (r > 0 && a > 0) x = cmul_t(x, tw[a * r * step]);
                v[k][r] = x;
            }
            dft_small<R, RAD>(v[k], sgn);
            dst[k] = li * n + jd * Ns * RAD + a;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
        if (dst[k] >= 0)
#pragma unroll
            for (int q = 0; q < RAD; ++q) buf[dst[k] + q * Ns] = v[k][q];
    __syncthreads();
}

// Generic Stockham stage of any radix R.  Pass 1 multiplies every input
// x[j + r m] by its stage twiddle W_{Ns R}^{(j%Ns) r} in place; pass 2 computes
// each output as a length-R DFT of the twiddled inputs with the R-th roots held
// in a compact LDS table (lanes mostly read the same root: broadcasts).
template <typename R>
__device__ inline void stage_generic(typename C2<R>::T* buf, typename C2<R>::T* roots, int n, int nl, int Ns,
                                     int Rx, const typename C2<R>::T* tw) {
    using CT = typename C2<R>::T;
    constexpr int KO = kLdsElems32 / 256;
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
        if (r > 0 && a > 0) buf[e] = cmul_t(buf[e], tw[a * r * step]);
    }
    __syncthreads();
    CT o[KO];
#pragma unroll
    for (int k = 0; k < KO; ++k) {
        const int e = threadIdx.x + 256 * k;
        if (e < tot) {
            int li, oi, rest, a, jd, q;
            fdivmod(e, n, rn, li, oi);
            fdivmod(oi, Ns, rns, rest, a);
            fdivmod(rest, Rx, rr, jd, q);
            const CT* src = buf + li * n + jd * Ns + a;
            R ar = R(0), ai = R(0);
            int kq = 0;
            for (int r = 0; r < Rx; ++r) {
                const CT x = src[r * m];
                const CT w = roots[kq];
                ar += x.x * w.x - x.y * w.y;
                ai += x.x * w.y + x.y * w.x;
                kq += q;
                if (kq >= Rx) kq -= Rx;
            }
            o[k].x = ar;
            o[k].y = ai;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KO; ++k) {
        const int e = threadIdx.x + 256 * k;
        if (e < tot) buf[e] = o[k];  // outputs are in natural (li, o) order
    }
    __syncthreads();
}

template <typename R>
__global__ __launch_bounds__(256) void line_fft_kernel(typename C2<R>::T* __restrict__ data, LineGeom lg, LinePlan pl,
                                                       const typename C2<R>::T* __restrict__ twiddle, int TL,
                                                       int inverse) {
    using CT = typename C2<R>::T;
    extern __shared__ unsigned char s_raw[];
    const int n = pl.n;
    CT* buf = reinterpret_cast<CT*>(s_raw);
    CT* tw = buf + (size_t)TL * n;
    CT* roots = tw + n;
    const int tid = threadIdx.x;
    const long long l0 = (long long)blockIdx.x * TL;
    const int nl = (int)min((long long)TL, lg.nlines - l0);
    const R sgn = inverse ? R(1) : R(-1);
    for (int k = tid; k < n; k += blockDim.x) {
        CT w = twiddle[k];
        if (inverse) w.y = -w.y;
        tw[k] = w;
    }
    __shared__ long long s_lb[64];
    if (tid < nl) s_lb[tid] = line_base(lg, l0 + tid);
    __syncthreads();
    const int tot = nl * n;
    const float rn = 1.0f / (float)n, rnl = 1.0f / (float)nl;
    if (lg.elem_stride == 1) {
        for (int e = tid; e < tot; e += blockDim.x) {
            int li, k;
            fdivmod(e, n, rn, li, k);
            buf[e] = data[s_lb[li] + k];
        }
    } else {
        for (int e = tid; e < tot; e += blockDim.x) {
            int k, li;
            fdivmod(e, nl, rnl, k, li);
            buf[li * n + k] = data[s_lb[li] + k * lg.elem_stride];
        }
    }
    __syncthreads();
    int Ns = 1;
    for (int st = 0; st < pl.nstages; ++st) {
        const int Rx = pl.radix[st];
        switch (Rx) {
            case 2: stage_codelet<R, 2>(buf, n, nl, Ns, tw, sgn); break;
            case 3: stage_codelet<R, 3>(buf, n, nl, Ns, tw, sgn); break;
            case 4: stage_codelet<R, 4>(buf, n, nl, Ns, tw, sgn); break;
            case 5: stage_codelet<R, 5>(buf, n, nl, Ns, tw, sgn); break;
            case 7: stage_codelet<R, 7>(buf, n, nl, Ns, tw, sgn); break;
            default: stage_generic<R>(buf, roots, n, nl, Ns, Rx, tw); break;
        }
        Ns *= Rx;
    }
    if (lg.elem_stride == 1) {
        for (int e = tid; e < tot; e += blockDim.x) {
            int li, k;
            fdivmod(e, n, rn, li, k);
            data[s_lb[li] + k] = buf[e];
        }
    } else {
        for (int e = tid; e < tot; e += blockDim.x) {
            int k, li;
            fdivmod(e, nl, rnl, k, li);
            data[s_lb[li] + k * lg.elem_stride] = buf[li * n + k];
        }
    }
}

// One axis of the built-in FFT: plan + twiddle tables in both precisions.
struct LineAxis {
    LinePlan plan{};
    DevBuf<double2> tw64;
    DevBuf<float2> tw32;
    explicit LineAxis(int n) {
        plan.n = n;
        plan.nstages = 0;
        int m = n;
        auto push = [&](int r) {
            MPM_REQUIRE(plan.nstages < kMaxStages, MPM_ERR_UNSUPPORTED, "FFT length has too many factors");
            plan.radix[plan.nstages++] = r;
        };
        while (m % 4 == 0) { push(4); m /= 4; }
        while (m % 2 == 0) { push(2); m /= 2; }
        for (int p = 3; p * p <= m; p += 2)
            while (m % p == 0) { push(p); m /= p; }
        if (m > 1) push(m);
        std::vector<double2> t64(n);
        std::vector<float2> t32(n);
        for (int k = 0; k < n; ++k) {
            double sv, cv;
            // exp(-2 pi i k / n) with the argument reduced exactly
            const double a = -2.0 * kPi * (double)k / (double)n;
            sv = std::sin(a);
            cv = std::cos(a);
            t64[k] = make_double2(cv, sv);
            t32[k] = make_float2((float)cv, (float)sv);
        }
        tw64.upload(t64.data(), n);
        tw32.upload(t32.data(), n);
    }
};

// Built-in batched FFT over the axes of one reciprocal layout.
struct LineFft {
    std::vector<std::unique_ptr<LineAxis>> axes;  // one per transformed axis
    std::vector<LineGeom> geom_per_rhs;            // nlines counted for one RHS
    long long rhs_elems = 0;                       // elements per RHS (grid size)

    // largest power of two TL with TL * n within the per-block element budget
    static int lines_per_block(int n, size_t cs) {
        const int cap = cs == sizeof(float2) ? kLdsElems32 : kLdsElems64;
        int tl = 1;
        while (tl * 2 * n <= cap && tl < 64) tl *= 2;
        return tl;
    }
    static bool fits(int n, size_t cs) { return n <= (cs == sizeof(float2) ? kLdsElems32 : kLdsElems64); }

    template <typename R>
    void exec(typename C2<R>::T* grid, int B, bool forward, hipStream_t s) {
        using CT = typename C2<R>::T;
        for (size_t a = 0; a < axes.size(); ++a) {
            LineAxis& ax = *axes[a];
            // the line descriptors of one RHS extend unchanged over the B stacked RHS
            LineGeom g = geom_per_rhs[a];
            g.nlines *= B;
            const int n = ax.plan.n;
            const int TL = lines_per_block(n, sizeof(CT));
            const size_t lds = ((size_t)TL * n + 2 * n) * sizeof(CT);
            const long long blocks = (g.nlines + TL - 1) / TL;
            const CT* tw = std::is_same<R, float>::value ? reinterpret_cast<const CT*>(ax.tw32.p)
                                                         : reinterpret_cast<const CT*>(ax.tw64.p);
            hipLaunchKernelGGL(line_fft_kernel<R>, dim3((unsigned)blocks), dim3(256), lds, s, grid, g, ax.plan, tw, TL,
                               forward ? 0 : 1);
            MPM_HIP(hipGetLastError());
        }
    }
};

// k-space Green's tensor at wave vector (Gx - kbx, Gy - kby, Gz): (k^x, k^y, k^z, sc)
// with sc = 9 (sin k / k - cos k)^2 exp(-(k^2 - sum_d G_d^2 eta_d) / (4 xi^2)) / k^4 / V
// and zero at k = 0.  monodisperse_ewald_electric_field.cu:1300-1404.
struct KParams {
    int nx, ny, nz;
    double Lx, Ly, Lz, kbx, kby, xi, etax, etay, etaz, invV;
};

__device__ inline void kcoef(const KParams& q, int ix, int iy, int iz, double o[4]) {
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
                idx = (((size_t)b * 3 + c) * g.Zs + js) * nxy + (size_t)vx * g.ny + vy;
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
    const size_t ncols = (size_t)q.nx * q.ny;
    const size_t V = ncols * q.nz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < V; e += (size_t)gridDim.x * blockDim.x) {
        const int kz = (int)(e / ncols);
        const size_t col = e % ncols;
        double o[4];
        kcoef(q, (int)(col / q.ny), (int)(col % q.ny), kz, o);
        for (int c = 0; c < 4; ++c) {
            k64[4 * e + c] = o[c];
            k32[4 * e + c] = (float)o[c];
        }
    }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One wavefront per particle: Gaussian contraction of the (inverse-transformed)
// grid, real-space gather over the symmetric neighbour list, self term.
template <typename R, int G, bool SLAB>
__global__ __launch_bounds__(256) void gather_kernel(
    SlotTable tab, int B, int parts, const typename C2<R>::T* __restrict__ grid, SpreadGeom g,
    const int* __restrict__ zmap, const int* __restrict__ base, const R* __restrict__ wts, int N, double h3,
    const double* __restrict__ pos, const int* __restrict__ nb_start, const int* __restrict__ nb_j,
    const double2* __restrict__ nb_f, double Lx, double Ly, double Lz, double kbx, double kby,
    const double2* __restrict__ selfc, double self_perp, int jacobi) {
    __shared__ int s_off[kMaxWidth * kMaxWidth * kMaxWidth];
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
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
    const size_t rhs_stride = SLAB ? 3 * (size_t)g.Zs * nxy : 3 * nxy * nz;
    const size_t comp_stride = SLAB ? (size_t)g.Zs * nxy : nxy * nz;
    int sx = 0, sy = 0, sz = 0;
    const R* w = wts + (size_t)i * 3 * width;
    if (parts & MPM_PART_RECIP) {
        sx = ((base[i] - half) % nx + nx) % nx;
        sy = ((base[N + i] - half) % ny + ny) % ny;
        sz = ((base[2 * N + i] - half) % nz + nz) % nz;
    }
    const double xi_ = pos[i], yi_ = pos[N + i], zi_ = pos[2 * N + i];
    const bool bloch = (kbx != 0.0 || kby != 0.0);

    for (int g0 = 0; g0 < B; g0 += G) {
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
            for (int t = lane; t < S; t += 64) {
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
                    v = (size_t)zmap[vz] * nxy + (size_t)vx * ny + vy;
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

EwaldOp::~EwaldOp() = default;

KParams EwaldOp::kparams() const {
    KParams q;
    q.nx = S.n[0];
    q.ny = S.n[1];
    q.nz = S.n[2];
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
    const int layout = slab ? 1 : 0;
    if (kspace_ready && kspace_layout == layout) return;
    d_k64.resize(4 * V);
    d_k32.resize(4 * V);
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((V + threads - 1) / threads, 65535);
    if (slab)
        hipLaunchKernelGGL(kspace_slab_kernel, dim3(blocks), dim3(threads), 0, nullptr, d_k64.p, d_k32.p, kparams());
    else
        hipLaunchKernelGGL(kspace_kernel, dim3(blocks), dim3(threads), 0, nullptr, d_k64.p, d_k32.p, kparams());
    MPM_HIP(hipGetLastError());
    MPM_HIP(hipDeviceSynchronize());
    kspace_ready = true;
    kspace_layout = layout;
}

void EwaldOp::set_bloch(double kx, double ky) {
    if (kx == kb[0] && ky == kb[1]) return;
    kb[0] = kx;
    kb[1] = ky;
    build_kspace();
}

// LDS bytes of zfused_kernel for slab size zs_pad and precision size rs.
static size_t zfused_lds(int zs_pad, int nz, size_t rs) { return 2 * rs * zs_pad * nz; }
static int zs_bucket(int Zs) {
    static const int buckets[] = {4, 6, 8, 10, 12, 14, 16};
    for (int b : buckets)
        if (Zs <= b) return b;
    return 0;
}

void EwaldOp::set_positions(const double* xyz, int64_t n) {
    MPM_REQUIRE(n > 0, MPM_ERR_INVALID_ARGUMENT, "no particles");
    MPM_REQUIRE(n < (1LL << 31) / 1024, MPM_ERR_INVALID_ARGUMENT, "too many particles");
    const int Nn = (int)n;
    for (int64_t q = 0; q < 3 * n; ++q)
        MPM_REQUIRE(std::isfinite(xyz[q]), MPM_ERR_INVALID_ARGUMENT, "non-finite position");
    // unchanged positions: the per-frame precalcs are still valid
    if (n == N && last_xyz.size() == (size_t)(3 * n) &&
        std::memcmp(last_xyz.data(), xyz, 3 * n * sizeof(double)) == 0)
        return;
    last_xyz.assign(xyz, xyz + 3 * n);
    N = n;
    const int W = S.width, half = S.half;
    std::vector<double> pos(3 * (size_t)Nn);
    std::vector<int> base(3 * (size_t)Nn);
    std::vector<double> w64((size_t)Nn * 3 * W);
    std::vector<float> w32((size_t)Nn * 3 * W);
    const double prod_eta = S.eta[0] * S.eta[1] * S.eta[2];
    const double Cs = std::pow(2.0 * xi * xi / kPi, 1.5) * std::sqrt(1.0 / prod_eta);
    for (int i = 0; i < Nn; ++i) {
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
    // ---- reciprocal layout: slab of occupied z-planes, or the full grid ----
    const int nz = S.n[2];
    std::vector<char> occ(nz, 0);
    for (int i = 0; i < Nn; ++i)
        for (int o = -half; o <= half; ++o) occ[(((base[2 * (size_t)Nn + i] + o) % nz) + nz) % nz] = 1;
    std::vector<int> zl;
    for (int z = 0; z < nz; ++z)
        if (occ[z]) zl.push_back(z);
    const int zsb = zs_bucket((int)zl.size());
    const bool want_slab = force_full == 0 && zsb > 0 && (int)zl.size() * 2 <= nz &&
                           zfused_lds(zsb, nz, sizeof(double)) <= 64 * 1024;
    slab = want_slab;
    Zs = slab ? (int)zl.size() : 0;
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
    if (cells) {
        members.resize((size_t)nc[0] * nc[1] * nc[2]);
        for (int j = 0; j < Nn; ++j) {
            const int cx = cell_coord(pos[j], 0), cy = cell_coord(pos[Nn + j], 1), cz = cell_coord(pos[2 * Nn + j], 2);
            members[(size_t)(cz * nc[1] + cy) * nc[0] + cx].push_back(j);
        }
    }
    std::vector<int> nstart(Nn + 1, 0), nj;
    std::vector<double2> nf;
    nj.reserve((size_t)Nn * 64);
    nf.reserve((size_t)Nn * 64);
    auto consider = [&](int i, int j) {
        double r[3] = {pos[i] - pos[j], pos[Nn + i] - pos[Nn + j], pos[2 * Nn + i] - pos[2 * Nn + j]};
        for (int d = 0; d < 3; ++d) r[d] -= L[d] * std::round(r[d] / L[d]);
        const double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        if (!(d2 <= rc2)) return;
        const double dist = std::sqrt(d2);
        if (!(dist < rc)) return;
        const double de = dist < 1.0 ? 1.0 : dist;
        nj.push_back(j);
        nf.push_back(make_double2(tab.interp(tab.perp, de), tab.interp(tab.para, de)));
    };
    for (int i = 0; i < Nn; ++i) {
        if (cells) {
            const int cx = cell_coord(pos[i], 0), cy = cell_coord(pos[Nn + i], 1), cz = cell_coord(pos[2 * Nn + i], 2);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int a = (cx + dx + nc[0]) % nc[0], b = (cy + dy + nc[1]) % nc[1],
                                  c = (cz + dz + nc[2]) % nc[2];
                        for (int j : members[(size_t)(c * nc[1] + b) * nc[0] + a])
                            if (j != i) consider(i, j);
                    }
        } else {
            for (int j = 0; j < Nn; ++j)
                if (j != i) consider(i, j);
        }
        nstart[i + 1] = (int)nj.size();
    }
    num_pairs = (int64_t)nj.size();
    d_pos.upload(pos.data(), pos.size());
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
    d_nb_start.upload(nstart.data(), nstart.size());
    if (nj.empty()) {
        nj.push_back(0);
        nf.push_back(make_double2(0, 0));
    }
    d_nb_j.upload(nj.data(), nj.size());
    d_nb_f.upload(nf.data(), nf.size());
    // FFT plans for the chosen layout
    const int key = slab ? Zs : -1;
    if (!lfft || fft_key != key) {
        const int nx = S.n[0], ny = S.n[1], nzz = S.n[2];
        const long long nxy = (long long)nx * ny;
        auto lf = std::make_unique<LineFft>();
        bool ok = true;
        auto add = [&](int n, LineGeom g) {
            ok = ok && LineFft::fits(n, sizeof(double2));
            if (!ok) return;
            lf->axes.push_back(std::make_unique<LineAxis>(n));
            lf->geom_per_rhs.push_back(g);
        };
        if (slab) {  // [3][Zs][nx][ny]: y lines (contiguous), then x lines (stride ny)
            lf->rhs_elems = 3LL * Zs * nxy;
            if (ny > 1) add(ny, LineGeom{3LL * Zs * nx, 1, 0, ny, 1});
            if (nx > 1) add(nx, LineGeom{3LL * Zs * ny, ny, 1, nxy, ny});
        } else {     // [3][nx][ny][nz]: z, y, x lines
            lf->rhs_elems = 3LL * nxy * nzz;
            if (nzz > 1) add(nzz, LineGeom{3LL * nxy, 1, 0, nzz, 1});
            if (ny > 1) add(ny, LineGeom{3LL * nx * nzz, nzz, 1, (long long)ny * nzz, nzz});
            if (nx > 1) add(nx, LineGeom{3LL * ny * nzz, (long long)ny * nzz, 1, nxy * nzz, (long long)ny * nzz});
        }
        lfft = ok ? std::move(lf) : nullptr;
    }
    if (!fft || fft_key != key) {
        if (slab)
            fft = std::make_unique<FftPlans>(std::vector<size_t>{(size_t)S.n[1], (size_t)S.n[0]}, 3 * Zs);
        else
            fft = std::make_unique<FftPlans>(std::vector<size_t>{(size_t)S.n[2], (size_t)S.n[1], (size_t)S.n[0]}, 3);
        fft_key = key;
    }
    // default self coefficients: zero (reference base_electric_field.cu allocates them zeroed)
    if (self_rows == 0 || d_selfc.n != (size_t)self_rows * Nn) {
        self_rows = 1;
        d_selfc.resize(Nn);
        MPM_HIP(hipMemset(d_selfc.p, 0, Nn * sizeof(double2)));
    }
}

void EwaldOp::set_self_coef(const double* coef, int rows) {
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "set positions before the self coefficients");
    MPM_REQUIRE(rows >= 1, MPM_ERR_INVALID_ARGUMENT, "need at least one row of self coefficients");
    d_selfc.upload(reinterpret_cast<const double2*>(coef), (size_t)rows * N);
    self_rows = rows;
}

size_t EwaldOp::grid_elems_per_rhs() const {
    return slab ? 3 * (size_t)Zs * S.n[0] * S.n[1] : 3 * V;
}

template <typename R, bool SLAB>
void EwaldOp::launch(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi) {
    using CT = typename C2<R>::T;
    const bool fp32 = std::is_same<R, float>::value;
    CT* grid = fp32 ? reinterpret_cast<CT*>(d_grid32.p) : reinterpret_cast<CT*>(d_grid64.p);
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
    const int Ni = (int)N;
    if (parts & MPM_PART_RECIP) {
        {
            ProfScope ps("spread", s);
            const int nt = ntiles[0] * ntiles[1] * n_ztiles;
            if (B == 1)
                hipLaunchKernelGGL((spread_kernel<R, 1, SLAB>), dim3(nt, 1), dim3(256), 0, s, t, B, g, d_base.p, w,
                                   d_bin_start.p, d_bin_items.p, d_ztiles.p, d_zmap.p, Ni, grid);
            else
                hipLaunchKernelGGL((spread_kernel<R, 4, SLAB>), dim3(nt, (B + 3) / 4), dim3(256), 0, s, t, B, g,
                                   d_base.p, w, d_bin_start.p, d_bin_items.p, d_ztiles.p, d_zmap.p, Ni, grid);
            MPM_HIP(hipGetLastError());
        }
        const size_t bytes_per_rhs = grid_elems_per_rhs() * sizeof(CT);
        {
            ProfScope ps("fft", s);
            if (lfft && fft_backend == 0)
                lfft->exec<R>(grid, B, true, s);
            else
                fft->exec(grid, bytes_per_rhs, B, fp32, true, s);
        }
        if (SLAB) {
            ProfScope ps("zfused", s);
            const int ncols = S.n[0] * S.n[1];
            const int zsb = zs_bucket(Zs);
            const size_t lds = zfused_lds(zsb, S.n[2], sizeof(R));
            const R* kc = fp32 ? reinterpret_cast<const R*>(d_k32.p) : reinterpret_cast<const R*>(d_k64.p);
            const int rb = B >= 16 ? 16 : (B >= 4 ? 4 : 1);
            const dim3 gridz((ncols + 256 / rb - 1) / (256 / rb), (B + rb - 1) / rb);
#define MPM_ZF2(ZSV, RBV)                                                                                     \
    hipLaunchKernelGGL((zfused_kernel<R, ZSV, RBV>), gridz, dim3(256), lds, s, grid, B, ncols, S.n[2], Zs, \
                       d_zlist.p, kc)
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
            ProfScope ps("scale", s);
            const int blocks = (int)std::min<size_t>((V + 255) / 256, 8192);
            const R* kc = fp32 ? reinterpret_cast<const R*>(d_k32.p) : reinterpret_cast<const R*>(d_k64.p);
            hipLaunchKernelGGL(scale_kernel<R>, dim3(blocks), dim3(256), 0, s, grid, kc, V, B);
            MPM_HIP(hipGetLastError());
        }
        {
            ProfScope ps("fft", s);
            if (lfft && fft_backend == 0)
                lfft->exec<R>(grid, B, false, s);
            else
                fft->exec(grid, bytes_per_rhs, B, fp32, false, s);
        }
    }
    {
        ProfScope ps("gather", s);
        const double h3 = S.h[0] * S.h[1] * S.h[2];
        const int blocks = (Ni + 3) / 4;
        if (B == 1)
            hipLaunchKernelGGL((gather_kernel<R, 1, SLAB>), dim3(blocks), dim3(256), 0, s, t, B, parts, grid, g,
                               d_zmap.p, d_base.p, w, Ni, h3, d_pos.p, d_nb_start.p, d_nb_j.p, d_nb_f.p, L[0], L[1],
                               L[2], kb[0], kb[1], d_selfc.p, self_perp, jacobi);
        else
            hipLaunchKernelGGL((gather_kernel<R, 4, SLAB>), dim3(blocks), dim3(256), 0, s, t, B, parts, grid, g,
                               d_zmap.p, d_base.p, w, Ni, h3, d_pos.p, d_nb_start.p, d_nb_j.p, d_nb_f.p, L[0], L[1],
                               L[2], kb[0], kb[1], d_selfc.p, self_perp, jacobi);
        MPM_HIP(hipGetLastError());
    }
}

void EwaldOp::apply(const SlotTable& t, int B, bool fp32, int parts, hipStream_t s, bool jacobi) {
    MPM_REQUIRE(N > 0, MPM_ERR_RUNTIME, "apply before set_positions");
    MPM_REQUIRE(B >= 1 && B <= kMaxSlots, MPM_ERR_INVALID_ARGUMENT, "bad slot count");
    for (int b = 0; b < B; ++b)
        MPM_REQUIRE(t.row[b] >= 0 && t.row[b] < self_rows, MPM_ERR_INVALID_ARGUMENT,
                    "self-coefficient row out of range");
    if (parts & MPM_PART_RECIP) ensure_kspace_table();
    const size_t elems = (size_t)B * grid_elems_per_rhs();
    if (fp32)
        d_grid32.reserve(elems);
    else
        d_grid64.reserve(elems);
    const int jac = jacobi ? 1 : 0;
    if (fp32) {
        if (slab) launch<float, true>(t, B, parts, s, jac);
        else launch<float, false>(t, B, parts, s, jac);
    } else {
        if (slab) launch<double, true>(t, B, parts, s, jac);
        else launch<double, false>(t, B, parts, s, jac);
    }
}

void EwaldOp::apply_contiguous(const double2* x, double2* y, int B, bool fp32, int parts, hipStream_t s,
                               const int* rows) {
    const size_t n = 3 * (size_t)N;
    for (int b0 = 0; b0 < B; b0 += kMaxSlots) {
        const int cnt = std::min(kMaxSlots, B - b0);
        SlotTable t;
        for (int j = 0; j < cnt; ++j) {
            t.x[j] = x + (size_t)(b0 + j) * n;
            t.y[j] = y + (size_t)(b0 + j) * n;
            t.row[j] = rows ? rows[b0 + j] : (self_rows > 1 ? b0 + j : 0);
        }
        apply(t, cnt, fp32, parts, s);
    }
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
}

}  // namespace mpm
