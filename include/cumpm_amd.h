/*
 * cumpm_amd — C ABI of the MI355X mutual-polarization (MPM) solver.
 *
 * Plain C types only (pointers, sizes, enums); no exceptions cross this
 * boundary and every call returns an mpm_status.  Complex numbers are two
 * consecutive doubles (re, im), i.e. the layout of std::complex<double>,
 * numpy complex128 and HIP double2.  Vectors of unknowns use the reference's
 * layout [rhs][particle][component] (dipoles only: n = 3N per right-hand side).
 * Lengths passed to the operator are nondimensional (radius 1); the driver
 * (mpm_dipole_solver_*) takes physical units and nondimensionalises exactly
 * as the reference does.
 *
 * Threading: one handle per host thread; handles hold device memory on the
 * device that was current at creation.  Calls that take a `stream` argument
 * enqueue on that HIP stream (NULL = the default stream) and return once the
 * work is enqueued, except where noted.
 */
#ifndef CUMPM_AMD_H
#define CUMPM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MPM_OK = 0,
    MPM_ERR_INVALID_ARGUMENT = 1, /* maps to ValueError in Python (reference: std::invalid_argument) */
    MPM_ERR_RUNTIME = 2,          /* maps to RuntimeError (reference: std::runtime_error) */
    MPM_ERR_HIP = 3,              /* HIP runtime failure (the reference exit()s the process instead) */
    MPM_ERR_FFT = 4,              /* rocFFT failure */
    MPM_ERR_UNSUPPORTED = 5,      /* feature of the reference not yet provided by this library */
    MPM_ERR_OUT_OF_MEMORY = 6
} mpm_status;

/* Reciprocal-space arithmetic of one operator apply.
 * Reference: PrecisionMode {AUTO, MIXED, DOUBLE} (ewald_electric_field_base.cu:28-34);
 * AUTO and MIXED both select FP32 spread/FFT/scale/contract. */
typedef enum {
    MPM_PREC_DOUBLE = 0, /* FP64 throughout */
    MPM_PREC_MIXED = 1,  /* FP32 reciprocal pipeline, FP64 real space / self / accumulation */
    MPM_PREC_AUTO = 2    /* same as MIXED (reference default) */
} mpm_precision;

/* Precision of a whole solve. */
typedef enum {
    MPM_SOLVE_DOUBLE = 0,  /* every apply FP64 (reference precision="double") */
    MPM_SOLVE_MIXED = 1,   /* every apply FP32-reciprocal (reference AUTO/MIXED) */
    MPM_SOLVE_REFINED = 2  /* GMRES: FP32-reciprocal Arnoldi applies and initial residual; FP64 residual
                              b - Ax at the end of every cycle decides convergence.  BiCGSTAB: FP32-
                              reciprocal iteration; a system whose recurrence residual meets tol restarts
                              from its iterate until the FP64 residual does (at most 20 passes) */
} mpm_solve_precision;

typedef enum { MPM_GMRES = 0, MPM_BICGSTAB = 1 } mpm_method;

/* Operator-apply parts (reference: calculate() = both; compute_real_field() = real only). */
enum { MPM_PART_REAL = 1, MPM_PART_RECIP = 2, MPM_PART_ALL = 3 };

typedef struct mpm_op mpm_op;

typedef struct {
    double box[3];   /* nondimensional box lengths */
    double errortol; /* Ewald accuracy (the reference passes the Krylov tol here) */
    double xi;       /* Ewald splitting parameter */
    int grid[3];     /* 0 = reference sizing rule; >0 overrides that axis */
    int support;     /* num_grid_gaussian P; 0 = reference rule ceil(-2 ln tol / pi) */
    int layout;      /* reciprocal-space layout: 0 = automatic (the z-plane slab when the
                        particle stencils touch at most half of the z-planes), 1 = full grid */
    int fft_backend; /* 0 = built-in batched LDS FFT (rocFFT when a length does not fit),
                        1 = rocFFT */
} mpm_op_desc;

typedef struct {
    int64_t num_particles;
    int64_t num_pairs;   /* real-space neighbour pairs (each ordered pair counted) */
    int grid[3];
    int support;         /* P */
    int stencil_width;   /* 2*floor(P/2)+1 points per axis */
    double grid_spacing[3];
    double spectral_split[3];
    double rc;
    double self_perp;    /* real-space self term (table entry 0) */
    double self_coef_ewald; /* -4 xi^3 / (3 sqrt(pi)) */
    double bloch[2];
    int slab_planes;     /* occupied z-planes of the slab layout; 0 = full grid */
    int fft_backend;     /* FFT in use: 0 = built-in, 1 = rocFFT */
    int polydisperse;    /* 1 for an operator made by mpm_op_create_poly */
    int num_unique_radii;
    int64_t num_quads;   /* Q: particles with quadrupoles (unknowns per right-hand side: 3N + 5Q) */
} mpm_op_info;

/* Monodisperse_Ewald_Electric_Field(box_x, box_y, box_z, errortol, xi, SOLVER_AX, radius=1)
 * — src/cpp/electric_field/monodisperse_ewald_electric_field.cu:28-74, sizing :546-616,
 * tables :92-542.  Fails with MPM_ERR_RUNTIME if rc > L/2 (:553-556). */
mpm_status mpm_op_create(const mpm_op_desc* desc, mpm_op** out);

/* Polydisperse_Ewald_Electric_Field(box_x, box_y, box_z, errortol, xi, SOLVER_AX, particle_radii)
 * — src/cpp/electric_field/polydisperse_ewald_electric_field.cu:27-127, sizing :2039-2105, pair
 * tables :1434-1698.  radii: host [n_radii], nondimensional; radii within 1e-5 share a class.
 * rc = max(sqrt(-ln tol)/xi, 2 a_max); P = ceil(2/h_mean - 2 ln tol/pi) unless desc->support > 0.
 * Full-grid layout and the built-in FFT only (desc->layout and desc->fft_backend must be 0 or
 * request them).  Up to 64 radius classes get pair tables as the reference's; with more, each
 * pair's closed forms are evaluated at its own radii (the reference's tables would need
 * U(U+1)/2 columns). */
mpm_status mpm_op_create_poly(const mpm_op_desc* desc, const double* radii, int64_t n_radii, mpm_op** out);
/* Direct_Electric_Field(radius, SOLVER_AX, solve_quadrupoles, quad_idxs) —
 * src/cpp/electric_field/direct_electric_field.cu:37-300, 418-427: the open-boundary point
 * multipole field, O(N^2) per apply, no grid and no cutoff.  radii: host, one per particle
 * or one for all (nondimensional); pairs closer than the source's radius are evaluated at
 * that radius.  Quadrupoles (mpm_op_set_quads) couple in the Ewald operator's conventions:
 * the reference kernel doubles the off-diagonal components of a quadrupole source and its
 * quadrupole-to-gradient term is not the derivative of the quadrupole's field; here both are
 * the large-box limit of the Ewald operator.  MPM_PREC_MIXED evaluates the pair sums in FP32
 * (as the reference's default), self terms in FP64.  mpm_op_set_bloch: unsupported. */
mpm_status mpm_op_create_direct(const double* radii, int64_t n_radii, mpm_op** out);
void mpm_op_destroy(mpm_op* op);
mpm_status mpm_op_get_info(const mpm_op* op, mpm_op_info* info);
/* getRealSpaceTablesHost (ewald_electric_field_base.cu:101-110), monodisperse operator: the
 * samples r (entry 0 = 0, then 1 + 0.001 k) and table `which` (0 = field_dip_1 / perp,
 * 1 = field_dip_2 / para, 2..4 = field_quad_1..3, 5..8 = grad_quad_1..4 after
 * mpm_op_set_quads).  *size = table length; r and values (either may be NULL) are written
 * when cap >= *size. */
mpm_status mpm_op_get_pair_table(const mpm_op* op, int which, double* r, double* values, int64_t cap,
                                 int64_t* size);
/* getNeighborListHost / getRealSpacePrecalcsHost (ewald_electric_field_base.cu:92-99,
 * monodisperse_ewald_electric_field.cu:1138-1160): the neighbour list of the current positions
 * as CSR, start[N + 1] and j[M] (every j != i at minimum-image distance below rc; the reference
 * keeps the same pairs in padded rows), and each pair's perp/para values.  *num_pairs = M; the
 * arrays (any may be NULL) are written when cap >= M. */
mpm_status mpm_op_get_neighbors(const mpm_op* op, int64_t* start, int* j, double* perp, double* para,
                                int64_t cap, int64_t* num_pairs);

/* updateParticleCoordinates — src/cpp/electric_field/base_electric_field.cu:40-118 plus the
 * lazy per-frame precalcs of monodisperse_ewald_electric_field.cu:2485-2503 (neighbour list,
 * real-space pair functions, spread/contract stencils).  xyz: host, [N][3], nondimensional. */
mpm_status mpm_op_set_positions(mpm_op* op, const double* xyz, int64_t n_particles);

/* solve_quadrupoles / quad_idxs of the Monodisperse_Ewald_Electric_Field constructor
 * (monodisperse_ewald_electric_field.cu:28-74; kernels :1435-1487, :1528-1629, :1631-1869,
 * :1871-1999): quadrupoles on particles idx[nq] (every particle when nq < 0).  Call before
 * mpm_op_set_positions.  Vectors then hold 3N dipole components followed by 5 components
 * (xx, xy, xz, yy, yz of the traceless moment) per quadrupole, in idx order; the operator
 * returns field and field gradient in the same layout.  Full-grid layout and built-in FFT.
 * The dipole-quadrupole real-space couplings take the signs of the reference's Python
 * formulation (tests/quadrupole_field.py:214-215); the CUDA kernel has both flipped. */
mpm_status mpm_op_set_quads(mpm_op* op, const int* idx, int64_t nq);

/* setBlochWavevector — src/cpp/electric_field/ewald_electric_field_base.h:136-142. */
mpm_status mpm_op_set_bloch(mpm_op* op, double kx, double ky);

/* setSelfCoef — src/cpp/electric_field/base_electric_field.cu (setSelfCoef), one row of N
 * complex coefficients per right-hand side: coef is host [n_rows][N] complex. */
mpm_status mpm_op_set_self_coef(mpm_op* op, const double* coef, int n_rows);
/* The same with coef in device memory ([n_rows][N] complex), copied on `stream` (for callers
 * that keep their parameters resident, e.g. an autograd function; no counterpart). */
mpm_status mpm_op_set_self_coef_device(mpm_op* op, const double* coef, int n_rows, void* stream);

/* calculate() — monodisperse_ewald_electric_field.cu:2211-2411 for B right-hand sides at
 * once: y[b] = A x[b] with self-coefficient row b (row 0 if only one row was set).
 * x, y: device, [B][N][3] complex.  parts: MPM_PART_*.  precision: mpm_precision. */
mpm_status mpm_op_apply(mpm_op* op, const double* x, double* y, int B, int precision, int parts,
                        void* stream);

/* The reference's dev-API pipeline (Ewald_Electric_Field_Base::spread / scale / contract /
 * realSpace, electric_field_bindings.cpp:239-258) for one right-hand side on caller buffers:
 * MPM_STAGE_SPREAD writes x (device, [3N + 5Q] complex) onto grid; MPM_STAGE_SCALE transforms
 * grid, applies the Green's tensor and transforms back, in place; MPM_STAGE_CONTRACT adds the
 * contraction of grid at the particles to y; MPM_STAGE_REAL_SPACE adds the real-space pair
 * terms and self terms of x (self-coefficient row 0) to y.  SPREAD + SCALE + CONTRACT +
 * REAL_SPACE on a zeroed y equals mpm_op_apply.  grid: device, mpm_op_grid_bytes bytes, in
 * this library's layout (full grid [3][nx][ny][nz], or the slab [3][Zs][nx][nyp] of
 * mpm_op_info.slab_planes occupied z-planes, rows padded to nyp = 16 ceil(ny / 16); with
 * quadrupoles [8][nx][ny][nz]; polydisperse one scalar grid [nx][ny][nz]); complex float for
 * MIXED/AUTO, complex double for DOUBLE.  The direct field has REAL_SPACE only. */
enum { MPM_STAGE_SPREAD = 1, MPM_STAGE_SCALE = 2, MPM_STAGE_CONTRACT = 3, MPM_STAGE_REAL_SPACE = 4 };
mpm_status mpm_op_grid_bytes(const mpm_op* op, int precision, size_t* bytes);
mpm_status mpm_op_stage(mpm_op* op, int stage, const double* x, void* grid, double* y, int precision,
                        void* stream);
/* getDev* of the reference (electric_field_bindings.cpp:142-164): the device address of an
 * internal buffer by name, valid until the next set_positions / set_bloch / destroy, and its
 * size in bytes (0 and NULL when this operator's layout holds no such buffer):
 * "x_part", "y_part", "z_part" (double [N] each), "self_coef" (complex double [rows][N]),
 * "radius" (direct field), "spread_coef" / "contract_coef" (FP64 stencil weights: per-axis
 * [N][3][width] monodisperse, [N][width^3] with quadrupoles, [N][P^3][3] polydisperse; the
 * same with "_fp32"), "spread_idxs" / "contract_idxs" (int base voxel [3][N], or wrapped voxel
 * per stencil point), "particle_index" (int [N] spatial order of the gathers),
 * "neighbor_start" (int [N + 1] CSR), "neighbor_list" (int [M]), "perp_para" (complex-double
 * pairs (perp, para) [M]), "pair_coef" (double per-pair coefficients), "self_perp"
 * (polydisperse, double [N]), "scale_coef" (k-space table), "fe_grid" (the grid of the last
 * apply), "fes_grid" (the slab's transposed k-space buffer).  Unknown names fail with
 * MPM_ERR_INVALID_ARGUMENT. */
mpm_status mpm_op_device_pointer(const mpm_op* op, const char* name, void** ptr, size_t* bytes);

/* mpm_op_apply that allocates no device memory.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The reciprocal grids of the apply live in a
 * library-owned static device array of 64 MiB (each GPU the library runs on holds its own
 * copy until the process ends); sizes that need more (grid elements x right-hand sides,
 * up to 32 right-hand sides per pass) fail with MPM_ERR_INVALID_ARGUMENT.  The operator's
 * k-space tables are built by mpm_op_set_positions / mpm_op_set_bloch, so after those the
 * apply calls no allocating routine; it needs the built-in FFT (MPM_ERR_UNSUPPORTED with
 * rocFFT).  Results equal mpm_op_apply's.  Calls on one GPU share the array: calls that
 * overlap on one GPU (two streams, two threads) race. */
mpm_status mpm_op_apply_noalloc(mpm_op* op, const double* x, double* y, int B, int precision, int parts,
                                void* stream);

/* Radius derivative of the polydisperse operator, for the radius gradients of an autograd
 * caller (no counterpart in the reference code; the derivatives of its pair functions
 * polydisperse_ewald_electric_field.cu:1434-1621, real-space self term :1611-1620 and
 * finite-sphere weights :243-262 with respect to the radius):
 *   out[i] = sum_b Re(g_b^H (dA/da_i) x_b)      (accumulate = 1: out[i] += the same)
 * for x, g device [B][3N] complex and out device double [N]; A is the operator without its
 * self coefficients (their radius dependence is the caller's).  Pair derivatives come from
 * the closed forms at each pair's own radii (the limit of the radius-class tables).  The
 * polydisperse operator only (else MPM_ERR_UNSUPPORTED), without a Bloch wave vector. */
mpm_status mpm_op_radius_gradient(mpm_op* op, const double* x, const double* g, int B, double* out, int accumulate,
                                  void* stream);

/* Autotuner.  No counterpart in the reference, whose pipeline is fixed: the operator can run
 * on the z-plane slab or the full grid (layout), with the built-in FFT or rocFFT
 * (fft_backend), and with MPM_TUNE_GRID on a grid grown per axis to the next length whose
 * prime factors are at most 7 (a finer discretisation: results change within the Ewald
 * tolerance).  Each valid configuration is timed with HIP events over warm applies of
 * min(B, 32) right-hand sides in `precision`; the fastest stays configured on op.  Choices
 * are cached for the process per (grid, N, power-of-two batch bucket, precision, operator
 * kind, MPM_TUNE_GRID); MPM_TUNE_FORCE times again.  The quadrupole operator runs on the
 * full grid with the built-in FFT, so its only candidate besides the current one is the grown
 * grid (MPM_TUNE_GRID); the polydisperse operator has one configuration.  Needs positions
 * set; the direct field fails with MPM_ERR_UNSUPPORTED. */
enum { MPM_TUNE_GRID = 1, MPM_TUNE_FORCE = 2 };
typedef struct {
    int layout;         /* chosen mpm_op_desc.layout (0 automatic / slab, 1 full grid) */
    int fft_backend;    /* chosen mpm_op_desc.fft_backend */
    int grid[3];        /* grid in use after tuning */
    double best_ms;     /* time of one apply of the batch in the chosen configuration */
    double default_ms;  /* the same in the configuration the operator had before */
    int candidates;     /* configurations considered */
    int cached;         /* 1 when the choice came from the process cache (no timing) */
} mpm_tune_info;
mpm_status mpm_op_autotune(mpm_op* op, int B, int precision, int flags, mpm_tune_info* info);

typedef struct {
    int method;     /* mpm_method */
    int restart;    /* GMRES restart length (reference default 50) */
    int maxiter;    /* reference default 1000 (counts Arnoldi steps / BiCGSTAB iterations) */
    double tol;     /* stop when ||M^-1 (b - A x)|| < tol ||b|| (gmres_solver.cu:99-101) */
    int jacobi;     /* left Jacobi preconditioner M = diag(self_coef) (numerical_solver.cu:207-270) */
    int precision;  /* mpm_solve_precision */
} mpm_solve_desc;

typedef struct {
    int iters;          /* Arnoldi steps (GMRES) or iterations (BiCGSTAB) */
    int matvecs;        /* operator applies of either precision */
    int matvecs_fp64;   /* FP64 applies */
    int converged;
    double relres_pre;  /* ||M^-1(b - A x)|| / ||b||, FP64 operator (reference criterion) */
    double relres_true; /* ||b - A x|| / ||b||, FP64 operator */
} mpm_solve_report;

/* GMRES_Solver::solve / BiCGSTAB_Solver::solve — src/cpp/numerical_solver/gmres_solver.cu:56-208,
 * bicgstab_solver.cu:58-192 — for B systems at once.  System b uses self-coefficient row b.
 * b, x: device [B][N][3] complex; x holds x0 on entry and the solution on return.
 * reports: host, B entries (may be NULL).  Synchronises `stream` before returning. */
mpm_status mpm_solve(mpm_op* op, const mpm_solve_desc* desc, const double* b, double* x, int B,
                     mpm_solve_report* reports, void* stream);

/* Adjoint solve A^H y = g for B systems (the gradient solve of an autograd caller;
 * mpm_solve_adjoint of the proposed C ABI, no counterpart in the reference code).
 * The dipole field operator is Hermitian (real symmetric without a Bloch wave vector),
 * so A^H differs from A only in its conjugated self coefficients: the same batched
 * solver runs with them.  With quadrupoles M A is Hermitian for M = diag(I, W/2), W the
 * Frobenius product of traceless tensors in (xx, xy, xz, yy, yz), and the solve runs on
 * M^-1 g and returns y = M z.  Arguments as mpm_solve, with g for b and y for x (y holds
 * the initial guess on entry); reports refer to the solved system. */
mpm_status mpm_solve_adjoint(mpm_op* op, const mpm_solve_desc* desc, const double* g, double* y, int B,
                             mpm_solve_report* reports, void* stream);

/* ---- Dipole_Solver driver (src/cpp/dipole_solver.cu, dipole_solver_bindings.cpp:80-179) ---- */
typedef struct mpm_dipole_solver mpm_dipole_solver;

typedef enum { MPM_GUESS_DERIVATIVE = 0, MPM_GUESS_MEAN_FIELD = 1, MPM_GUESS_PREVIOUS = 2 } mpm_guess;
typedef enum { MPM_FIELD_AUTO = 0, MPM_FIELD_MONODISPERSE = 1, MPM_FIELD_POLYDISPERSE = 2,
               MPM_FIELD_DIRECT = 3 } mpm_field_type;

typedef struct {
    double box[3];            /* physical */
    const double* eps_p;      /* host complex [W][eps_cols], eps_cols = 1 or N (copied) */
    int64_t num_wavevectors;  /* W */
    int64_t eps_cols;
    const double* radius;     /* host [num_radii] (copied); num_radii 0 => 1.0 */
    int64_t num_radii;
    double eps_m;
    double xi;
    double tol;
    int quiet;
    int guess;                /* mpm_guess */
    int method;               /* mpm_method */
    int field_type;           /* mpm_field_type */
    const double* E0;         /* host complex [K][E0_len] (copied); K = 0 => identity (3 RHS) */
    int64_t num_E0;           /* K */
    int64_t E0_len;           /* 3, 3N or 3N + 5Q (quadrupole rows zero unless given) */
    int solve_quadrupoles;    /* 1 => unknowns [3N dipoles | 5Q quadrupoles] (monodisperse field only) */
    const int* quad_idxs;     /* host [num_quad_idxs] particles carrying quadrupoles (copied); */
    int64_t num_quad_idxs;    /*   0 with solve_quadrupoles => every particle (Q = N)            */
    int precision;            /* mpm_precision: DOUBLE => MPM_SOLVE_DOUBLE, AUTO/MIXED => MPM_SOLVE_MIXED */
    int refined;              /* 1 => MPM_SOLVE_REFINED instead of MPM_SOLVE_MIXED (ignored for DOUBLE) */
    int batched;              /* 0 = reference order (one RHS at a time); 1 = all W x K RHS in one solve */
    int restart;              /* 0 => 50 */
    int maxiter;              /* 0 => 1000 */
    int grid[3];              /* Ewald overrides (0 = reference rule) */
    int support;
    double ewald_tol;         /* 0 => tol (the reference ties the Ewald accuracy to tol) */
    int autotune;             /* 0 off; 1 => mpm_op_autotune on the first frame, batch = the
                                 systems of one solve, precision = the solve's; 2 => the same
                                 with MPM_TUNE_GRID (no counterpart in the reference) */
} mpm_dipole_desc;

mpm_status mpm_dipole_solver_create(const mpm_dipole_desc* desc, mpm_dipole_solver** out);
void mpm_dipole_solver_destroy(mpm_dipole_solver* s);
/* Dipole_Solver::compute (dipole_solver.cu:495-607): one frame of positions (physical units). */
mpm_status mpm_dipole_solver_compute(mpm_dipole_solver* s, const double* x, const double* y,
                                     const double* z, int64_t n_particles);
/* dims = {frames, wavevectors, particles, polarizations} */
mpm_status mpm_dipole_solver_get_dims(const mpm_dipole_solver* s, int64_t dims[4]);
/* get_eff_polarizability: host complex [W][K][3] (dipole_solver.cu:609-625). */
mpm_status mpm_dipole_solver_get_eff_polarizability(const mpm_dipole_solver* s, int physical,
                                                    double* out);
/* get_dipoles: host complex [F][W][N][K][3] (dipole_solver.cu:627-635). */
mpm_status mpm_dipole_solver_get_dipoles(const mpm_dipole_solver* s, int physical, double* out);
/* Number Q of particles carrying quadrupoles (0 unless solve_quadrupoles; get_num_quads). */
mpm_status mpm_dipole_solver_get_num_quads(const mpm_dipole_solver* s, int64_t* num_quads);
/* get_quadrupoles: host complex [F][W][Q][K][5], physical units scale by eps_m a_min^4
 * (dipole_solver.cu:640-647). */
mpm_status mpm_dipole_solver_get_quadrupoles(const mpm_dipole_solver* s, int physical, double* out);
/* Per-solve reports, [F][W][K] (an addition: the reference reports none). */
mpm_status mpm_dipole_solver_get_reports(const mpm_dipole_solver* s, mpm_solve_report* out);
/* Ewald operator of the solver (valid after the first compute), for inspection. */
mpm_status mpm_dipole_solver_get_op_info(const mpm_dipole_solver* s, mpm_op_info* info);
mpm_status mpm_dipole_solver_set_use_jacobi(mpm_dipole_solver* s, int enable);
mpm_status mpm_dipole_solver_get_use_jacobi(const mpm_dipole_solver* s, int* enable);
/* Wall time (ms) of the last compute: device solves only (right-hand sides and guesses
 * already in device memory), set-up (first-call operator construction and per-frame
 * precalcs), and the whole call. */
mpm_status mpm_dipole_solver_get_timing(const mpm_dipole_solver* s, double* solve_ms, double* setup_ms,
                                        double* total_ms);
/* The autotuner's result on the first frame (all zero while desc->autotune is 0 or before
 * the first compute). */
mpm_status mpm_dipole_solver_get_tune_info(const mpm_dipole_solver* s, mpm_tune_info* info);
/* Forget the stored frames (results and warm-start state) but keep the operator, its FFT
 * plans and per-frame precalcs; the next compute() starts from the mean-field guess as a
 * first frame does.  No counterpart in the reference (its results only accumulate). */
mpm_status mpm_dipole_solver_reset(mpm_dipole_solver* s);

/* ---- EELS driver (src/cpp/eels_solver.cu, eels_bindings.cpp:61-194, src/cuMPM/eels.py) ---- */
typedef struct mpm_eels_solver mpm_eels_solver;

typedef struct {
    double box[3];            /* physical */
    const double* eps_p;      /* host complex [W][eps_cols], eps_cols = 1 or N (copied) */
    int64_t num_wavevectors;  /* W */
    int64_t eps_cols;
    const double* omega;      /* host [num_omega >= W] wavenumbers in inverse length units (copied) */
    int64_t num_omega;
    double v;                 /* beam speed as a fraction of c, 0 < v < 1 */
    const double* radius;     /* host [num_radii] (copied); num_radii 0 => 1.0 */
    int64_t num_radii;
    double eps_m;
    double xi;
    double tol;               /* Krylov tolerance and Ewald accuracy */
    int quiet;
    int guess;                /* mpm_guess */
    int method;               /* mpm_method */
    int field_type;           /* mpm_field_type; DIRECT = the open-boundary field */
    double integration_step;  /* kept for the reference's signature (scaled, unused) */
    int solve_quadrupoles;
    const int* quad_idxs;     /* host [num_quad_idxs] (copied); 0 with solve_quadrupoles => all */
    int64_t num_quad_idxs;
    int asm_flag;             /* aperiodic sampling: average over asm_nx x asm_ny Bloch points */
    int asm_nx, asm_ny;
    int precision;            /* mpm_precision; the reference defaults to AUTO = FP32 reciprocal */
    int refined;              /* 1 => MPM_SOLVE_REFINED instead of MPM_SOLVE_MIXED (ignored for DOUBLE) */
    int batched;              /* 0 = reference order (one frequency at a time with its guesses);
                                 1 = every frequency of a Bloch point in one batched solve */
    int restart;              /* 0 => 50 */
    int maxiter;              /* 0 => 1000 */
} mpm_eels_desc;

/* EELS_Solver(box, eps_p, omega, v, radius, eps_m, xi, tol, quiet, guess_type, solver_type,
 * field_type, integration_step, solve_quadrupoles, quad_idxs, asm_flag, asm_Nx, asm_Ny,
 * precision) — eels_solver.cu:118-233.  Fails with MPM_ERR_INVALID_ARGUMENT for v outside
 * (0, 1) (the reference then evaluates Bessel functions of imaginary argument). */
mpm_status mpm_eels_solver_create(const mpm_eels_desc* desc, mpm_eels_solver** out);
void mpm_eels_solver_destroy(mpm_eels_solver* s);
/* EELS_Solver::compute (eels_solver.cu:438-880): one beam position epos (x, y), physical units,
 * and the particle coordinates; appends one frame (the loss spectrum, dipoles and quadrupoles).
 * The incident field is the relativistic field of the beam (Bessel K0/K1 on the device), each
 * quadrupole's incident gradient its 24-point surface average; the loss is the reciprocity sum
 * sum Im(p . conj E) (+ quadrupole terms) / (2 pi^2). */
mpm_status mpm_eels_solver_compute(mpm_eels_solver* s, const double epos[2], const double* x, const double* y,
                                   const double* z, int64_t n_particles);
/* dims = {frames, wavevectors, particles, quadrupoles} */
mpm_status mpm_eels_solver_get_dims(const mpm_eels_solver* s, int64_t dims[4]);
/* get_eels: host [F][W] (eels_bindings.cpp:9-19). */
mpm_status mpm_eels_solver_get_eels(const mpm_eels_solver* s, double* out);
/* get_dipoles: host complex [F][W][N][3]; physical scales by eps_m a_min^3 (eels_solver.h:155-163). */
mpm_status mpm_eels_solver_get_dipoles(const mpm_eels_solver* s, int physical, double* out);
/* get_quadrupoles: host complex [F][W][Q][5]; physical scales by eps_m a_min^4. */
mpm_status mpm_eels_solver_get_quadrupoles(const mpm_eels_solver* s, int physical, double* out);
/* Incident field of the last frame's last Bloch point, host complex [W][3N + 5Q] (nondimensional;
 * no counterpart in the reference). */
mpm_status mpm_eels_solver_get_incident_field(const mpm_eels_solver* s, double* out);
/* Per-solve reports, [F][Bloch points][W]: *count entries, written when cap >= *count. */
mpm_status mpm_eels_solver_get_reports(const mpm_eels_solver* s, mpm_solve_report* out, int64_t cap,
                                       int64_t* count);
/* The EELS driver's incident field on its own (no counterpart in the reference as an entry
 * point; eels_solver.cu:588-706): for nondimensional particle positions xyz [N][3] (host), the
 * particles quad_particles [Q] that carry quadrupoles, W wavenumbers omega (times the length
 * unit), beam speed v / c, beam position epos and probe softening sigma (nondimensional),
 * out: host complex [W][3N + 5Q], field rows then gradient rows as the solver's unknowns. */
mpm_status mpm_ebeam_field(const double* xyz, int64_t n_particles, const int* quad_particles, int64_t n_quads,
                           const double* omega, int64_t n_omega, double v, const double epos[2], double sigma,
                           double* out);
mpm_status mpm_eels_solver_set_use_jacobi(mpm_eels_solver* s, int enable);
mpm_status mpm_eels_solver_get_use_jacobi(const mpm_eels_solver* s, int* enable);
mpm_status mpm_eels_solver_get_timing(const mpm_eels_solver* s, double* solve_ms, double* setup_ms,
                                      double* total_ms);

/* ---- Near field (src/cpp/near_field.h:12-57, near_field_bindings.cpp:8-25) ---- */
typedef struct mpm_near_field mpm_near_field;
/* Near_Field(box, E0, radius, xi, errortol, field_type): box physical [3], E0 host complex [3],
 * radius host [num_radii] (0 => 1.0; lengths are scaled by the first radius), field_type a value
 * of mpm_field_type: AUTO selects the polydisperse field when the radii differ and DIRECT the
 * open-boundary field (near_field.cu:105-147). */
mpm_status mpm_near_field_create(const double box[3], const double* E0, const double* radius, int64_t num_radii,
                                 double xi, double errortol, int field_type, mpm_near_field** out);
void mpm_near_field_destroy(mpm_near_field* nf);
/* set_dipoles: host complex [n][3] (copied). */
mpm_status mpm_near_field_set_dipoles(mpm_near_field* nf, const double* dip, int64_t n);
mpm_status mpm_near_field_set_dipole_positions(mpm_near_field* nf, const double* x, const double* y, const double* z,
                                               int64_t n);
mpm_status mpm_near_field_set_field_points(mpm_near_field* nf, const double* x, const double* y, const double* z,
                                           int64_t n);
/* Reciprocal-space precision of calculate (mpm_precision): AUTO (the default) and MIXED use an
 * FP32 grid as the reference's fields do by default; DOUBLE runs FP64 throughout.  No
 * counterpart in the reference's Near_Field, whose fields take their default. */
mpm_status mpm_near_field_set_precision(mpm_near_field* nf, int precision);
/* calculate (near_field.cu:105-213): intensity[f] = |E0 + E_ind(f)|^2 at every field point, E_ind
 * the periodic Ewald field of the dipoles: the field points join the operator as sites without
 * dipoles (volume-averaged over a sphere at f, the pair functions of the solver).  With the
 * polydisperse field, field point f takes particle f's radius for f < N; beyond, the first
 * radius for its spread weight and the smallest for its pair functions, as the reference's
 * INTERACTION_FIELD kernels index them (polydisperse_ewald_electric_field.cu:297, 402). */
mpm_status mpm_near_field_calculate(mpm_near_field* nf, double* intensity);

/* ---- Plumbing ---- */
const char* mpm_status_string(mpm_status s);
/* Message of the last failing call on this thread. */
const char* mpm_last_error(void);
/* Device buffers for callers without an allocator of their own (the Python layer). */
mpm_status mpm_buffer_alloc(size_t bytes, void** dev);
/* Device memory held by the library's own allocations (operators with their grids, tables
 * and per-frame data, solver workspaces, driver buffers; not mpm_buffer_alloc buffers or
 * rocFFT's internal work areas) and its high-water mark; reset_peak != 0 restarts the mark
 * at the current value.  No counterpart in the reference. */
mpm_status mpm_memory_stats(int64_t* current_bytes, int64_t* peak_bytes, int reset_peak);
mpm_status mpm_buffer_free(void* dev);
mpm_status mpm_buffer_upload(void* dev, const void* host, size_t bytes);
mpm_status mpm_buffer_download(void* host, const void* dev, size_t bytes);
mpm_status mpm_device_synchronize(void);

/* Kernel timing with HIP events on the stream each kernel is launched on.  When enabled,
 * every launch of the named kernel class is bracketed by events and accumulated. */
mpm_status mpm_profile_enable(int enable);
mpm_status mpm_profile_reset(void);
/* name: a kernel class, "spread", "fft" (one 2-D/3-D transform), "fft_axis0".."fft_axis2" (its
 * passes), "scale", "zfused" (slab Green's tensor), "contract" (matrix-core contraction), "gather"
 * (contraction and/or real space, self term), "krylov"; or a kernel's own name with its template
 * arguments as profilers print it (e.g. "contract_mfma_kernel<float, 4, 10>").  total_ms over count
 * launches; units (may be NULL): right-hand sides summed over those launches. */
mpm_status mpm_profile_get(const char* name, double* total_ms, int64_t* count, int64_t* units);
/* Newline-separated names with recorded launches.  Writes them to buf when cap >= *needed
 * (the length including the terminating zero). */
mpm_status mpm_profile_names(char* buf, size_t cap, size_t* needed);
/* Plain device copy of `bytes` (a multiple of 16, 16-byte aligned) from src to dst on `stream`:
 * no counterpart in the reference; a kernel of known traffic for calibrating hardware counters. */
mpm_status mpm_profile_copy(void* dst, const void* src, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUMPM_AMD_H */
