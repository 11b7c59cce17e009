// Internal declarations shared by the HIP translation units of libcumpm_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <complex>
#include <cstdint>
#include <initializer_list>
#include <map>
#include <atomic>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "cumpm_amd.h"

namespace mpm {

using cplx = std::complex<double>;

// Errors inside the library are C++ exceptions carrying an mpm_status; the
// C ABI layer converts them (no exception crosses the boundary).
struct Error : std::runtime_error {
    mpm_status status;
    Error(mpm_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define MPM_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::mpm::Error(e_ == hipErrorOutOfMemory ? MPM_ERR_OUT_OF_MEMORY           \
                                                         : MPM_ERR_HIP,                   \
                               std::string(#call) + ": " + hipGetErrorString(e_) + " (" + \
                                   __FILE__ + ":" + std::to_string(__LINE__) + ")");      \
    } while (0)

#define MPM_REQUIRE(cond, status, msg)          \
    do {                                        \
        if (!(cond)) throw ::mpm::Error(status, msg); \
    } while (0)

// Message of the last failing C-ABI call on this thread, and the wrapper every entry
// point runs its body in: exceptions become an mpm_status and a message.
inline thread_local std::string g_last_error;

template <typename F>
mpm_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return MPM_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return MPM_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return MPM_ERR_RUNTIME;
    }
}

// Device bytes held by the library's own allocations (operators, solver and driver
// buffers), and their high-water mark since the last reset.
inline std::atomic<int64_t> g_dev_bytes{0};
inline std::atomic<int64_t> g_dev_peak{0};
inline void note_alloc(int64_t bytes) {
    const int64_t now = g_dev_bytes.fetch_add(bytes) + bytes;
    int64_t peak = g_dev_peak.load();
    while (now > peak && !g_dev_peak.compare_exchange_weak(peak, now)) {
    }
}

// Owning device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            g_dev_bytes.fetch_sub((int64_t)(n * sizeof(T)));
        }
        p = nullptr;
        n = 0;
    }
    void resize(size_t count) {
        if (count == n && p) return;
        release();
        if (count) {
            MPM_HIP(hipMalloc(&p, count * sizeof(T)));
            note_alloc((int64_t)(count * sizeof(T)));
        }
        n = count;
    }
    // grow-only variant for scratch buffers
    void reserve(size_t count) {
        if (count > n) resize(count);
    }
    void upload(const T* host, size_t count, hipStream_t s = nullptr) {
        resize(count);
        if (count) MPM_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
        if (count) MPM_HIP(hipStreamSynchronize(s));
    }
    // upload into a grow-only allocation (n is then the capacity): per-frame data whose
    // size changes a little from frame to frame keeps its memory
    void assign(const T* host, size_t count, hipStream_t s = nullptr) {
        reserve(count);
        if (count) MPM_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
        if (count) MPM_HIP(hipStreamSynchronize(s));
    }
};

// ---------------------------------------------------------------------------
// Profiling (HIP events around kernel classes, on the launch stream).
// ---------------------------------------------------------------------------
struct ProfScope {
    // units: right-hand sides the bracketed launches process (summed per class)
    ProfScope(const char* name, hipStream_t s, int64_t units = 0);
    ~ProfScope();
    // Also account the bracketed launch under the kernel's own name, written the way
    // profilers print template instances: base<arg0, arg1, ...>.
    void kernel(const char* base, std::initializer_list<std::string> targs);
    const char* name;
    std::string kname;
    int64_t units;
    hipStream_t stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};
// Host wall time of a bracketed region (per-frame set-up, transfers, driver work), recorded
// under `name` while profiling is on; no device synchronisation of its own.
struct HostScope {
    explicit HostScope(const char* name);
    ~HostScope();
    const char* name;
    std::chrono::steady_clock::time_point t0;
    bool on;
};
// Consecutive host regions: lap(name) records the time since the previous lap (or since
// construction) under `name` while profiling is on.
struct HostLaps {
    HostLaps();
    void lap(const char* name);
    std::chrono::steady_clock::time_point t;
    bool on;
};
template <typename R>
inline std::string type_name() { return std::is_same<R, float>::value ? "float" : "double"; }
inline std::string tbool(bool b) { return b ? "true" : "false"; }
bool profiling_enabled();

// ---------------------------------------------------------------------------
// Ewald operator (monodisperse, dipoles).
// ---------------------------------------------------------------------------
struct Sizing {
    double rc = 0, kcut = 0;
    int n[3] = {0, 0, 0};
    double h[3] = {0, 0, 0};
    double eta[3] = {0, 0, 0};
    int P = 0, half = 0, width = 0;
};

void compute_sizing(const double L[3], double errortol, double xi, const int grid_override[3],
                    int P_override, Sizing& S);
void dipole_pair_functions(double r, double xi, double& f_perp, double& f_para);
double dipole_self_term(double xi);

// Pointer table passed by value to the batched kernels (one entry per RHS slot).
constexpr int kMaxSlots = 32;
struct SlotTable {
    const double2* x[kMaxSlots];
    double2* y[kMaxSlots];
    int row[kMaxSlots];  // self-coefficient row of the slot
};

struct FftPlans;   // rocFFT plans keyed by (precision, batch)
struct PairTable;  // sampled real-space pair functions
struct KParams;    // k-space geometry passed to the kernels
struct LineFft;    // built-in batched line FFT

class EwaldOp {
   public:
    EwaldOp(const mpm_op_desc& d);
    // polydisperse operator: one radius per particle (nondimensional), merged
    // within 1e-5 into unique radii; positions must then come with the same N
    EwaldOp(const mpm_op_desc& d, const double* radii, int64_t n_radii);
    // direct (open-boundary) field: no grid, no cutoff; radii one per particle or one for
    // all (nondimensional)
    struct DirectTag {};
    EwaldOp(DirectTag, const double* radii, int64_t n_radii);
    ~EwaldOp();
    bool is_direct() const { return direct; }
    void set_positions(const double* xyz, int64_t N);
    void set_bloch(double kx, double ky);
    void set_self_coef(const double* coef, int rows);
    // the same from device memory, enqueued on s
    void set_self_coef_device(const double* coef, int rows, hipStream_t s);
    // Polydisperse operator, no Bloch vector: out[i] (+)= sum_b Re(g_b^H (dA/da_i) x_b) for
    // device x, g [B][3N], the derivative of the operator (pair functions, real-space self
    // term, finite-sphere spread/contract weights; not the self coefficients) with respect
    // to particle i's radius.  accumulate: add to out instead of overwriting it.
    void radius_gradient(const double2* x, const double2* g, int B, double* out, bool accumulate, hipStream_t s);
    // polydisperse operator: spread/contract weight radii that differ from the pair-function
    // radii (one per site; set before set_positions)
    void set_weight_radii(const double* r, int64_t n);
    // conjugates the device self coefficients in place (the adjoint operator's)
    void conj_self_coef(hipStream_t s);
    // Device memory a caller lends to an apply for its reciprocal grids (base 256-byte
    // aligned); with one, apply allocates nothing.
    struct Workspace {
        char* base = nullptr;
        size_t bytes = 0;
    };
    // y[b] = A x[b] (or M^-1 A x[b] with the left Jacobi preconditioner) for the slots
    // in `t` (count B <= kMaxSlots); the grids live in `ws` when given (then the k-space
    // tables must be current and the FFT the built-in one), else in the operator's buffers
    void apply(const SlotTable& t, int B, bool fp32, int parts, hipStream_t s, bool jacobi = false,
               const Workspace* ws = nullptr);
    // convenience: contiguous [B][3N] buffers, chunked by kMaxSlots
    void apply_contiguous(const double2* x, double2* y, int B, bool fp32, int parts,
                          hipStream_t s, const int* rows = nullptr, const Workspace* ws = nullptr);
    // One stage of an apply of one right-hand side on caller buffers (the reference's
    // pipeline spread / scale / contract / realSpace): 1 spread x into grid, 2 transform,
    // Green's tensor and inverse transform of grid in place, 3 y += contraction of grid,
    // 4 y += real space and self terms of x (self-coefficient row 0, no preconditioner)
    void stage(int which, const double2* x, void* grid, double2* y, bool fp32, hipStream_t s);
    size_t stage_grid_bytes(bool fp32) const;
    // device address and size of a named internal buffer (null when this layout has none)
    void* device_pointer(const std::string& name, size_t& bytes) const;
    // Reciprocal-space configuration: layout 0 automatic / 1 full grid, FFT backend 0
    // built-in / 1 rocFFT, and (grid != null) the grid dimensions at the same support;
    // rebuilds the per-frame data of the current positions.  Every configuration is the
    // same operator up to rounding, except a new grid (a finer discretisation).
    void configure(int layout, int backend, const int* grid = nullptr);
    int layout_setting() const { return force_full; }
    int backend_setting() const { return fft_backend; }
    const int* grid_dims() const { return S.n; }
    bool is_poly() const { return poly; }
    bool has_quads() const { return quads; }
    bool is_slab() const { return slab; }
    // workspace bytes an apply of B (<= kMaxSlots) right-hand sides needs
    size_t scratch_bytes(int B, bool fp32) const;
    bool builtin_fft() const;
    void info(mpm_op_info& out) const;
    // host copies for inspection: the monodisperse pair tables (r, perp, para; table 2..8 =
    // the quadrupole tables when built) and the current neighbour list with its pair values
    int64_t pair_table(int which, double* r, double* v, int64_t cap) const;
    int64_t neighbors(int64_t* start, int* j, double* perp, double* para, int64_t cap) const;
    // quadrupoles on particles idx[nq] (every particle when nq < 0); before set_positions.
    // Unknowns per right-hand side then are [3N dipole | 5Q quadrupole] components.
    void set_quads(const int* idx, int64_t nq);
    size_t n_unknowns() const { return 3 * (size_t)N + 5 * qidx_h.size(); }
    size_t num_quad_slots() const { return qidx_h.size(); }
    const int* quad_idx_dev() const { return quads ? d_qidx.p : nullptr; }

    int64_t N = 0;
    int self_rows = 0;
    const double2* self_coef() const { return d_selfc.p; }

   private:
    void build_kspace();
    void ensure_kspace_table();
    KParams kparams() const;
    size_t grid_elems_per_rhs() const;
    template <typename R, bool SLAB>
    void launch(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi);
    template <typename R>
    void launch_poly(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi);

    double L[3], errortol, xi;
    Sizing S;
    double self_perp = 0;
    double kb[2] = {0, 0};
    size_t V = 0;
    int64_t num_pairs = 0;
    int force_full = 0;          // 1: never use the slab layout
    // reciprocal layout: the full grid, or the slab of occupied z-planes
    bool slab = false;
    int Zs = 0;
    std::vector<int> zlist;
    int tile[3] = {4, 8, 8};
    int binsz[3] = {4, 8, 8};
    int nbins[3] = {0, 0, 0};
    int ntiles[3] = {0, 0, 0};
    int n_ztiles = 0;
    // geometry on device
    DevBuf<double> d_pos;        // [3][N] SoA
    DevBuf<int> d_base;          // [3][N] wrapped base grid index
    DevBuf<double> d_w64;        // [N][3][width] per-axis weights (spread constant folded in x)
    DevBuf<float> d_w32;
    DevBuf<int> d_bin_start;     // [nbins + 1]
    DevBuf<int> d_bin_items;     // [N] particle ids sorted by bin
    DevBuf<int> d_ztiles;        // z-tile indices the spread visits
    DevBuf<int> d_zmap;          // [nz] slab plane of each z (or -1)
    DevBuf<int> d_zlist;         // [Zs] z of each slab plane
    DevBuf<int> d_nb_start;      // [N + 1] target-centric CSR
    DevBuf<int> d_nb_j;          // [M]
    DevBuf<double2> d_nb_f;      // [M] (perp, para)
    // per pair, for the thread-per-(particle, RHS) finish: (ux, uy, uz, perp,
    // para - perp, cos, sin, 0) with u = r / max(|r|, 1), r = x_j - x_i (min
    // image) and the Bloch phase -(kb . r); rebuilt when positions or kb change
    mutable std::vector<int> h_nb_start, h_nb_j;  // neighbour list CSR of the current positions
    DevBuf<int> d_order;                  // [N] particles in spatial order (gather waves)
    mutable std::vector<double> h_pair_r;  // [M][3]
    mutable std::vector<double2> h_pair_f;
    DevBuf<double> d_pairc;        // [M][8]
    // monodisperse operator: the neighbour list and pair data are built on the device
    // (search_pairs_device); the host copies h_nb_start / h_nb_j / h_pair_* are then
    // fetched only when asked for (fetch_pairs)
    bool pairs_dev = false;
    mutable bool pairs_host_valid = true;
    DevBuf<double> d_pair_r;       // [M][3] x_j - x_i
    DevBuf<double> d_pairq;        // [M][7] quadrupole table values
    DevBuf<double> d_ptab;         // [2 + 7][ts] pair tables (perp, para, quadrupole) on the device
    int ptab_nt = 0;
    DevBuf<int> d_cell_start, d_cell_items, d_cellc, d_nb_count;
    DevBuf<char> d_scan_tmp;
    void search_pairs_device(const std::vector<double>& pos, int Nn, bool cells, const int nc[3],
                             const std::vector<int>& cell_start, const std::vector<int>& cell_items,
                             const std::vector<int>& cellc);
    void fetch_pairs() const;
    void build_pair_coefs();
    DevBuf<double2> d_selfc;     // [rows][N]
    bool zcontig = false;        // slab planes form one cyclic run (zlist[a] = z0 + a mod nz)
    bool kt = false;             // slab k-space data in its own buffer, columns ordered ky * nxp + kx
    int nyp = 0, nxp = 0;        // slab row strides of the grid ([x][nyp]) and k-space ([ky][nxp]) buffers
    size_t kgrid_elems_per_rhs() const;
    int kspace_cols() const;
    // matrix-core spread and contraction: regions of reg_rx x (reg_pblk planes) x 16 rows
    // (slab: every occupied z-plane, rows y; full grid: planes y, rows z) and the candidate
    // particles of each region (CSR, ascending ids); reg_nry row blocks and reg_npb plane
    // blocks per x block of regions
    int reg_rx = 0, reg_nrx = 0, reg_nry = 0, reg_npb = 1, reg_pblk = 0;
    bool reg_full = false;       // the regions tile the full grid (monodisperse dipoles)
    int reg_slots = 0;           // (region, candidate) slots of the current frame
    DevBuf<int> d_reg_start;
    DevBuf<int> d_reg_items;
    DevBuf<int> d_ps_start;      // [N + 1] region slots of each particle
    DevBuf<int> d_ps_items;      // slot ids (CSR positions of d_reg_items)
    DevBuf<double> d_regw64;     // [slot][row] region-local separable weights (FP64, FP32)
    DevBuf<float> d_regw32;
    DevBuf<double> d_partial64;  // [slot][6 B] per-region contractions (FP64, or FP32 in its first half)
    // k-space tables: (khat x, y, z, scale) per voxel for the full grid and the
    // z-DFT slab kernel, or the Toeplitz plane coupling [Zs][6][nx*ny] for a
    // contiguous slab; kspace_key records which one is current
    bool kspace_ready = false;
    int kspace_key = -1;
    DevBuf<double> d_k64;
    DevBuf<float> d_k32;
    DevBuf<double2> d_toe64;
    DevBuf<float2> d_toe32;
    DevBuf<float> d_toe32ri;  // odd nz: [Zs][ncols][6] real xx, yy, zz, xy and imaginary xz, yz
    // grids (cur_grid / cur_kgrid: those of the apply in flight, the operator's own buffers
    // or the caller's workspace)
    void* cur_grid = nullptr;
    void* cur_kgrid = nullptr;
    int phases = 7;  // parts of a launch: 1 spread, 2 transforms and Green's tensor, 4 contraction
    DevBuf<double2> d_grid64;
    DevBuf<float2> d_grid32;
    DevBuf<double2> d_kgrid64;   // slab k-space buffer (transposed columns)
    DevBuf<float2> d_kgrid32;
    std::unique_ptr<FftPlans> fft;
    std::unique_ptr<LineFft> lfft;
    int fft_key = -2;
    int fft_backend = 0;         // 0: built-in LDS FFT, 1: rocFFT
    std::unique_ptr<PairTable> table;
    std::vector<double> last_xyz;  // positions of the current precalcs

    // ---- quadrupoles (monodisperse) ----
    bool quads = false, quads_all = false;
    std::vector<int> qreq_h;          // requested indices (empty with quads_all)
    std::vector<int> qidx_h;          // [Q] particle of each quadrupole
    DevBuf<int> d_qidx, d_qmap;       // [Q], [N] (quadrupole slot or -1)
    mutable std::vector<double> h_pair_q;     // [M][7] quadrupole table values per pair
    DevBuf<double> d_sw64;            // [N][W^3] Gaussian weights (spread constant included)
    DevBuf<float> d_sw32;
    DevBuf<double> d_kq64;            // [V][16] k-space: khat, sc, scQ, scGP, scGQ, Qfactor[5], Qdot[5]
    DevBuf<float> d_kq32;
    void setup_quad_stencils(const std::vector<double>& pos, int Nn);
    template <typename R>
    void launch_quads(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi);

    // ---- direct field ----
    bool direct = false;
    DevBuf<double> d_drad;           // [N] radius of each particle
    template <typename R>
    void launch_direct(const SlotTable& t, int B, int parts, hipStream_t s, int jacobi);
    void build_quad_slots(int Nn);   // qidx_h, d_qidx, d_qmap from the requested indices

    // ---- polydisperse operator ----
    bool poly = false;
    std::vector<double> rad_h;       // [N] radii
    std::vector<double> uniq_h;      // sorted unique radii
    std::vector<int> ridx_h;         // [N] index into uniq_h
    std::vector<double> pt_r;        // table samples: 0, then 1 + 0.001 k
    std::vector<double> pt_perp, pt_para;  // [entry][pair column]
    std::vector<double> uself_h;     // real-space self term per unique radius
    int pt_ncol = 0;
    // more radius classes than kPolyTableClasses: no tables, the pair functions of
    // each pair's own radii in closed form (the tables' limit without interpolation)
    bool poly_exact = false;
    int poly_classes = 0;            // distinct radii (merged within 1e-5)
    std::vector<double> pself_h;     // [N] real-space self term of each particle
    double peta = 0;                 // scalar spectral split of the finite-sphere kernel
    int poff = 0;                    // first stencil offset, -floor((P - 1) / 2)
    DevBuf<double> d_selfp;          // [N] real-space self term of each particle
    DevBuf<double> d_pw64;           // [N][P^3][3] spread/contract weight vectors
    DevBuf<float> d_pw32;
    DevBuf<int> d_pidx;              // [N][P^3] wrapped voxel of each stencil point
    DevBuf<int> d_vox_start;         // [V + 1] per voxel, its (particle, point) entries
    DevBuf<int> d_vox_items;         //   entry = particle * P^3 + point, ascending
    double poly_lookup(const std::vector<double>& tab, double r, int u, int v) const;
    void poly_pair_value(double r, int i, int j, double& perp, double& para) const;
    void setup_poly_stencils(const std::vector<double>& pos, int Nn);
    // radius derivatives of the current frame, built on the first radius_gradient call
    // spread/contract weight radii when they differ from the pair-function radii (the near
    // field's field points, whose two radii follow different rules); empty: rad_h
    std::vector<double> wrad_h;
    std::vector<int> pst_g;          // [N][3] nearest grid point of each particle
    std::vector<double> pst_dd;      // [N][3] its offset from the particle
    int64_t pst_serial = 0, rg_serial = -1;
    DevBuf<double> d_dpw;            // [N][P^3][3] d(weight vector)/da
    DevBuf<double2> d_dpair;         // per CSR entry (d perp, d (para - perp)) / d a_row
    DevBuf<double> d_dps;            // [N] d(real-space self term)/da
    DevBuf<double2> d_rgx, d_rgg;    // reciprocal fields of x and g
    void poly_pair_dvalue(double r, int i, int j, double& dperp, double& dpara) const;
    void ensure_rgrad_data();

    friend struct OpAccess;
};

// Batched Krylov solve (krylov.hip).
void solve_batched(EwaldOp& op, const mpm_solve_desc& d, const double2* b, double2* x, int B,
                   mpm_solve_report* reports, hipStream_t s);

// Times the reciprocal-space configurations of `op` on B right-hand sides and leaves the
// fastest configured (tune.hip); MPM_TUNE_* flags.
void autotune(EwaldOp& op, int B, bool fp32, int flags, hipStream_t s, mpm_tune_info& out);

}  // namespace mpm
