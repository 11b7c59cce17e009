// Batched, device-resident Krylov solvers for A x = b over many systems that
// share one Ewald operator (each with its own self-coefficient row, i.e. its own
// frequency): GMRES streams any number of systems through kMaxSlots slots,
// BiCGSTAB runs them in groups of kMaxSlots.
//
// Semantics follow the reference solvers system by system:
//   GMRES(m)  src/cpp/numerical_solver/gmres_solver.cu:56-208 — left Jacobi,
//             stop when ||M^-1 (b - A x)|| < tol ||b|| (outer check) or the
//             Givens residual estimate |g_{k+1}| < tol ||b|| (inner check; the
//             cycle then ends with k+1 columns); maxiter counts Arnoldi steps.
//   BiCGSTAB  src/cpp/numerical_solver/bicgstab_solver.cu:58-192 —
//             unconjugated rho and r0~.v, conjugated omega = <t,s>/<t,t>,
//             absolute breakdown tests |.| < 1e-16.
// What differs is the execution: all systems advance together; every vector
// operation is one launch over the whole batch; the Hessenberg column,
// Givens rotations and triangular solve stay on the GPU; the host reads back
// one residual per system per Arnoldi step (the reference blocks on a D2H
// copy for every dot product).  Orthogonalisation is classical Gram-Schmidt
// with one re-orthogonalisation pass (CGS2: two batched GEMV sweeps instead
// of k+1 dependent dot/axpy pairs), numerically equivalent to the
// reference's modified Gram-Schmidt.  Converged systems retire: they leave
// the slot list, cost nothing in later applies, and (GMRES) hand their slot to
// the next pending system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "mpm_internal.h"

namespace mpm {

namespace {

constexpr int kRedBlocks = 64;   // max partial blocks per system in reductions
constexpr int kRedChunk = 1024;  // elements per reduction block

struct Ptrs {
    double2* a[kMaxSlots];
    double2* b[kMaxSlots];
    int sys[kMaxSlots];  // system index (row into per-system device arrays)
    // per-slot data of the streaming GMRES (kernels templated on VAR read these):
    // basis vectors to project against, right-hand side, self-coefficient row
    int k[kMaxSlots];
    const double2* c[kMaxSlots];
    int row[kMaxSlots];
};

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmulc(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// partial[slot][blk][i] = sum_{e in chunk} op(A_i[e], b[e]), A_i = a + i*stride.
// mode 0: conj(A_i) . b ; mode 1: A_i . b (unconjugated); mode 2: |b|^2 (real, i = 0)
// Grid (chunks, slots * groups): one block reduces one element chunk of one
// system against kDotGroup consecutive A_i, so a Gram-Schmidt pass against k
// basis vectors launches chunks * systems * ceil(k / kDotGroup) blocks and every
// thread keeps kDotGroup independent loads in flight.
constexpr int kDotGroup = 8;
template <bool VAR>
__global__ __launch_bounds__(256) void dots_kernel_t(Ptrs p, int kmax, size_t stride, size_t n, size_t chunk,
                                                     int mode, double2* __restrict__ partial) {
    __shared__ double s_red[4][2 * kDotGroup];
    const int ng = (kmax + kDotGroup - 1) / kDotGroup;
    const int slot = blockIdx.y / ng, grp = blockIdx.y % ng;
    const int kk = VAR ? p.k[slot] : kmax;  // VAR: each slot its own count (block-uniform exit)
    if (grp * kDotGroup >= kk) return;
    const int blk = blockIdx.x;
    const size_t e0 = (size_t)blk * chunk;
    const size_t e1 = min(n, e0 + chunk);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = grp * kDotGroup;
    const int ni = min(kDotGroup, kk - i0);
    const double2* A = p.a[slot] + (size_t)i0 * stride;
    const double2* Bv = p.b[slot];
    double sr[kDotGroup], si[kDotGroup];
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i) sr[i] = si[i] = 0.0;
    // two elements per thread and step, so that twice the loads are in flight
    for (size_t e = e0 + threadIdx.x; e < e1; e += 512) {
        const size_t e2 = e + 256;
        const bool two = e2 < e1;
        const double2 bv = Bv[e];
        const double2 bv2 = two ? Bv[e2] : make_double2(0.0, 0.0);
        if (mode == 2) {
            sr[0] += (bv.x * bv.x + bv.y * bv.y) + (bv2.x * bv2.x + bv2.y * bv2.y);
        } else {
#pragma unroll
            for (int i = 0; i < kDotGroup; ++i) {
                if (i < ni) {
                    const double2 av = A[(size_t)i * stride + e];
                    const double2 av2 = two ? A[(size_t)i * stride + e2] : make_double2(0.0, 0.0);
                    const double2 t = (mode == 0) ? cmulc(av, bv) : cmul(av, bv);
                    const double2 t2 = (mode == 0) ? cmulc(av2, bv2) : cmul(av2, bv2);
                    sr[i] += t.x + t2.x;
                    si[i] += t.y + t2.y;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i) {
        if (i < ni) {
            sr[i] = wsum(sr[i]);
            si[i] = wsum(si[i]);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < kDotGroup; ++i) {
            s_red[wave][2 * i] = sr[i];
            s_red[wave][2 * i + 1] = si[i];
        }
    }
    __syncthreads();
    if (threadIdx.x < ni) {
        const int i = threadIdx.x;
        double r = 0.0, im = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            r += s_red[w][2 * i];
            im += s_red[w][2 * i + 1];
        }
        partial[((size_t)slot * kRedBlocks + blk) * 64 + i0 + i] = make_double2(r, im);
    }
}

#define dots_kernel dots_kernel_t<false>

// out[sys][i] (+)= sum_blk partial[slot][blk][i]
__global__ void reduce_kernel(Ptrs p, int nslots, int kk, int nblk, const double2* __restrict__ partial,
                              double2* __restrict__ out, int out_stride, int accumulate,
                              double2* __restrict__ last) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslots * kk) return;
    const int slot = t / kk, i = t % kk;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < nblk; ++b) {
        const double2 v = partial[((size_t)slot * kRedBlocks + b) * 64 + i];
        sr += v.x;
        si += v.y;
    }
    double2* o = out + (size_t)p.sys[slot] * out_stride + i;
    if (last) last[(size_t)p.sys[slot] * out_stride + i] = make_double2(sr, si);
    if (accumulate) {
        o->x += sr;
        o->y += si;
    } else {
        *o = make_double2(sr, si);
    }
}

// One classical Gram-Schmidt pass of the Arnoldi step, after dots_kernel:
// c_i = sum_blk partial[slot][blk][i] (every block sums the projections in block
// order), w -= sum_{i<kk} c_i V_i (a = V_0, b = w); block 0 of a system stores c
// into h[sys] (pass 0) or adds it (pass 1).  With pnorm (the second pass), each
// block also writes the sum of |w_new|^2 over its elements to
// pnorm[slot * pstride + block], the partial norms of the next basis vector.
template <bool VAR>
__global__ __launch_bounds__(256) void cgs_axpy_kernel_t(Ptrs p, int kk_arg, size_t stride, size_t n, int nblk,
                                                         const double2* __restrict__ partial, double2* __restrict__ h,
                                                         int hstride, int accumulate, double* __restrict__ pnorm,
                                                         int pstride) {
    __shared__ double2 s_c[64];
    __shared__ double s_n[4];
    const int slot = blockIdx.y;
    const int kk = VAR ? p.k[slot] : kk_arg;
    if (threadIdx.x < kk) {
        double sr = 0.0, si = 0.0;
        for (int b = 0; b < nblk; ++b) {
            const double2 v = partial[((size_t)slot * kRedBlocks + b) * 64 + threadIdx.x];
            sr += v.x;
            si += v.y;
        }
        s_c[threadIdx.x] = make_double2(sr, si);
        if (blockIdx.x == 0) {
            double2* o = h + (size_t)p.sys[slot] * hstride + threadIdx.x;
            if (accumulate) {
                o->x += sr;
                o->y += si;
            } else {
                *o = make_double2(sr, si);
            }
        }
    }
    __syncthreads();
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    double nn = 0.0;
    if (e < n) {
        const double2* V = p.a[slot];
        double2 acc = p.b[slot][e];
#pragma unroll 4
        for (int i = 0; i < kk; ++i) {
            const double2 t = cmul(s_c[i], V[(size_t)i * stride + e]);
            acc.x -= t.x;
            acc.y -= t.y;
        }
        p.b[slot][e] = acc;
        nn = acc.x * acc.x + acc.y * acc.y;
    }
    if (pnorm) {
        nn = wsum(nn);
        if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = nn;
        __syncthreads();
        if (threadIdx.x == 0) pnorm[(size_t)slot * pstride + blockIdx.x] = ((s_n[0] + s_n[1]) + s_n[2]) + s_n[3];
    }
}

#define cgs_axpy_kernel cgs_axpy_kernel_t<false>

// b = a / sqrt(nn) with nn = sum_blk pnorm[slot][blk] (summed by the first wave of
// every block in a fixed order); block 0 stores nn into norm2[sys].
__global__ __launch_bounds__(256) void scale_norm_kernel(Ptrs p, size_t n, const double* __restrict__ pnorm,
                                                         int pstride, int nbx, double* __restrict__ norm2) {
    __shared__ double s_nn;
    const int slot = blockIdx.y;
    if (threadIdx.x < 64) {
        double t = 0.0;
        for (int b = threadIdx.x; b < nbx; b += 64) t += pnorm[(size_t)slot * pstride + b];
        t = wsum(t);
        if (threadIdx.x == 0) {
            s_nn = t;
            if (blockIdx.x == 0) norm2[p.sys[slot]] = t;
        }
    }
    __syncthreads();
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const double nn = sqrt(s_nn);
    if (e < n && nn > 0.0) {
        const double2 v = p.a[slot][e];
        p.b[slot][e] = make_double2(v.x / nn, v.y / nn);
    }
}

// Left Jacobi (numerical_solver.cu:207-242): v[i][c] *= 1/self_coef[row][i]; in place on b.
__device__ inline double2 jac_inv(double2 sc) {
    const double den = sc.x * sc.x + sc.y * sc.y;
    return den > 1e-30 ? make_double2(sc.x / den, -sc.y / den) : make_double2(1.0, 0.0);
}

// r = M^-1 (b - Ax) (or b - Ax without Jacobi); also keeps t = b - Ax for the true residual.
// a = Ax, b = r (out);  rhs/true_out by system.
// VAR: the right-hand side p.c[slot] and self-coefficient row p.row[slot] per slot.
template <bool VAR>
__global__ __launch_bounds__(256) void residual_kernel_t(Ptrs p, size_t n, const double2* __restrict__ rhs_base,
                                                         double2* __restrict__ true_base,
                                                         const double2* __restrict__ selfc,
                                                         const int* __restrict__ rows, int N, int jacobi,
                                                         const int* __restrict__ qidx) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = p.sys[slot];
    const double2 bv = VAR ? p.c[slot][e] : rhs_base[(size_t)s * n + e];
    const double2 ax = p.a[slot][e];
    const double2 t = make_double2(bv.x - ax.x, bv.y - ax.y);
    true_base[(size_t)s * n + e] = t;
    double2 r = t;
    if (jacobi) {  // the particle of component e: dipoles [N][3], then quadrupoles [Q][5]
        const size_t n3 = 3 * (size_t)N;
        const int part = e < n3 ? (int)(e / 3) : qidx[(e - n3) / 5];
        const int row = VAR ? p.row[slot] : rows[s];
        r = cmul(t, jac_inv(selfc[(size_t)row * N + part]));
    }
    p.b[slot][e] = r;
}
#define residual_kernel residual_kernel_t<false>

// b = a * (1/norm[sys]) when norm > 0
__global__ __launch_bounds__(256) void scale_copy_kernel(Ptrs p, size_t n, const double* __restrict__ norm2) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double nn = sqrt(norm2[p.sys[slot]]);
    if (nn > 0.0) {
        const double2 v = p.a[slot][e];
        p.b[slot][e] = make_double2(v.x / nn, v.y / nn);
    }
}

// b = b + alpha[sys] * a   (complex alpha)
__global__ __launch_bounds__(256) void axpy_kernel(Ptrs p, size_t n, const double2* __restrict__ alpha) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double2 t = cmul(alpha[p.sys[slot]], p.a[slot][e]);
    double2 v = p.b[slot][e];
    p.b[slot][e] = make_double2(v.x + t.x, v.y + t.y);
}

// GMRES column update + Givens rotations on the device, one wave per system.
// H: [sys][(m+1) x m] row-major (row i, column k); h: new column coefficients [sys][m+1];
// hn2: squared norm of the orthogonalised vector.  gmres_solver.cu:139-178.
template <bool VAR>
__global__ __launch_bounds__(64) void givens_kernel_t(Ptrs p, int nslots, int k_arg, int m, double2* __restrict__ H,
                                                      const double2* __restrict__ h, const double* __restrict__ hn2,
                                                      double2* __restrict__ g, double* __restrict__ cs,
                                                      double2* __restrict__ sn, double* __restrict__ resid) {
    // one wave per system: the new Hessenberg column and the earlier rotations are
    // staged in LDS, one lane applies the rotations in order
    __shared__ double2 s_col[64 + 2];
    __shared__ double s_c[64];
    __shared__ double2 s_s[64];
    const int slot = blockIdx.x;
    if (slot >= nslots) return;
    const int lane = threadIdx.x;
    const int s = p.sys[slot];
    const int k = VAR ? p.k[slot] - 1 : k_arg;  // VAR: the column this slot adds
    double2* Hs = H + (size_t)s * (m + 1) * m;
    const double2* hs = h + (size_t)s * (m + 1);
    double2* gs = g + (size_t)s * (m + 1);
    double* c = cs + (size_t)s * m;
    double2* sv = sn + (size_t)s * m;
    for (int i = lane; i <= k; i += 64) s_col[i] = hs[i];
    for (int i = lane; i < k; i += 64) {
        s_c[i] = c[i];
        s_s[i] = sv[i];
    }
    if (lane == 0) s_col[k + 1] = make_double2(sqrt(hn2[s]), 0.0);
    __syncthreads();
    if (lane == 0) {
        double2 h1 = s_col[0];  // row i of the column after the rotations before i
        for (int i = 0; i < k; ++i) {
            const double2 h2 = s_col[i + 1];
            const double ci = s_c[i];
            const double2 si = s_s[i];
            const double2 a = cmulc(si, h2);  // conj(s) h2
            const double2 b = cmul(si, h1);   // s h1
            s_col[i] = make_double2(ci * h1.x + a.x, ci * h1.y + a.y);
            h1 = make_double2(-b.x + ci * h2.x, -b.y + ci * h2.y);
        }
        const double2 h2 = s_col[k + 1];
        double cr = 1.0;
        double2 sr = make_double2(0.0, 0.0);
        if (h2.x != 0.0 || h2.y != 0.0) {
            const double a1 = hypot(h1.x, h1.y);
            if (a1 == 0.0) {
                cr = 0.0;
                sr = make_double2(1.0, 0.0);
            } else {
                const double a2 = hypot(h2.x, h2.y);
                const double scl = a1 + a2;
                const double n1 = a1 / scl, n2 = a2 / scl;
                const double hyp = scl * sqrt(n1 * n1 + n2 * n2);
                cr = a1 / hyp;
                // (conj(h1)/|h1|) * (h2 / hyp)
                const double2 u = make_double2(h1.x / a1, -h1.y / a1);
                sr = cmul(u, make_double2(h2.x / hyp, h2.y / hyp));
            }
        }
        c[k] = cr;
        sv[k] = sr;
        const double2 a = cmulc(sr, h2);
        s_col[k] = make_double2(cr * h1.x + a.x, cr * h1.y + a.y);
        s_col[k + 1] = make_double2(0.0, 0.0);
        const double2 g1 = gs[k], g2 = gs[k + 1];
        const double2 ga = cmulc(sr, g2), gb = cmul(sr, g1);
        gs[k] = make_double2(cr * g1.x + ga.x, cr * g1.y + ga.y);
        const double2 gk1 = make_double2(-gb.x + cr * g2.x, -gb.y + cr * g2.y);
        gs[k + 1] = gk1;
        resid[s] = hypot(gk1.x, gk1.y);
    }
    __syncthreads();
    for (int i = lane; i <= k + 1; i += 64) Hs[i * m + k] = s_col[i];
}

#define givens_kernel givens_kernel_t<false>

// y = H^-1 g (upper triangular, kk columns per system); gmres_solver.cu:187-195.
// One wave per system: the upper triangle is staged in LDS (packed by rows), then
// column by column y_i = g_i / H_ii and g_j -= H_ji y_i for j < i across the lanes.
__global__ __launch_bounds__(64) void backsolve_kernel(Ptrs p, int nslots, const int* __restrict__ kk_of, int m,
                                                       const double2* __restrict__ H, const double2* __restrict__ g,
                                                       double2* __restrict__ y) {
    extern __shared__ double2 s_h[];  // row i holds H_ij, j = i..kk-1, from offset tri(i)
    __shared__ double2 s_g[64];
    __shared__ double2 s_y[64];
    const int slot = blockIdx.x;
    if (slot >= nslots) return;
    const int lane = threadIdx.x;
    const int s = p.sys[slot];
    const int kk = kk_of[s];
    const double2* Hs = H + (size_t)s * (m + 1) * m;
    const double2* gs = g + (size_t)s * (m + 1);
    double2* ys = y + (size_t)s * (m + 1);
    auto tri = [&](int i) { return i * kk - i * (i - 1) / 2; };  // start of row i (its diagonal)
    for (int i = 0; i < kk; ++i)
        for (int j = i + lane; j < kk; j += 64) s_h[tri(i) + j - i] = Hs[i * m + j];
    for (int i = lane; i < kk; i += 64) s_g[i] = gs[i];
    __syncthreads();
    for (int i = kk - 1; i >= 0; --i) {
        if (lane == 0) {
            const double2 d = s_h[tri(i)], v = s_g[i];
            const double den = d.x * d.x + d.y * d.y;
            s_y[i] = make_double2((v.x * d.x + v.y * d.y) / den, (v.y * d.x - v.x * d.y) / den);
        }
        __syncthreads();
        const double2 yi = s_y[i];
        for (int j = lane; j < i; j += 64) {
            const double2 t = cmul(s_h[tri(j) + i - j], yi);
            s_g[j] = make_double2(s_g[j].x - t.x, s_g[j].y - t.y);
        }
        __syncthreads();
    }
    for (int i = lane; i < kk; i += 64) ys[i] = s_y[i];
}

// x += sum_{i<kk} y_i V_i  (a = V_0, b = x)
__global__ __launch_bounds__(256) void xupdate_kernel(Ptrs p, size_t stride, size_t n,
                                                      const int* __restrict__ kk_of,
                                                      const double2* __restrict__ y, int ystride) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = p.sys[slot];
    const int kk = kk_of[s];
    const double2* V = p.a[slot];
    const double2* ys = y + (size_t)s * ystride;
    double2 acc = p.b[slot][e];
    for (int i = 0; i < kk; ++i) {
        const double2 t = cmul(ys[i], V[(size_t)i * stride + e]);
        acc.x += t.x;
        acc.y += t.y;
    }
    p.b[slot][e] = acc;
}

// r0 = g0 * V_0 setup: g[sys][0] = rn, rest zero.
__global__ void ginit_kernel(Ptrs p, int nslots, int m, double2* __restrict__ g,
                             const double* __restrict__ rn2) {
    const int slot = blockIdx.x;
    const int s = p.sys[slot];
    for (int i = threadIdx.x; i <= m; i += blockDim.x)
        g[(size_t)s * (m + 1) + i] = make_double2(i == 0 ? sqrt(rn2[s]) : 0.0, 0.0);
}

// BiCGSTAB vector updates (per-system complex scalars in device arrays).
// p = (p - omega v) beta + r
__global__ __launch_bounds__(256) void bicg_p_kernel(Ptrs pp, size_t n, const double2* __restrict__ v_base,
                                                     const double2* __restrict__ r_base,
                                                     const double2* __restrict__ omega,
                                                     const double2* __restrict__ beta) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = pp.sys[slot];
    double2 p = pp.b[slot][e];
    const double2 v = v_base[(size_t)s * n + e];
    const double2 r = r_base[(size_t)s * n + e];
    const double2 ov = cmul(omega[s], v);
    p = cmul(make_double2(p.x - ov.x, p.y - ov.y), beta[s]);
    pp.b[slot][e] = make_double2(p.x + r.x, p.y + r.y);
}

// s = r - alpha v   (a = r, b = s; v by system)
__global__ __launch_bounds__(256) void bicg_s_kernel(Ptrs pp, size_t n, const double2* __restrict__ v_base,
                                                     const double2* __restrict__ alpha) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = pp.sys[slot];
    const double2 r = pp.a[slot][e];
    const double2 av = cmul(alpha[s], v_base[(size_t)s * n + e]);
    pp.b[slot][e] = make_double2(r.x - av.x, r.y - av.y);
}

// x += alpha p + omega s ; r = s - omega t   (a unused)
__global__ __launch_bounds__(256) void bicg_xr_kernel(Ptrs pp, size_t n, double2* __restrict__ x_base,
                                                      double2* __restrict__ r_base,
                                                      const double2* __restrict__ p_base,
                                                      const double2* __restrict__ s_base,
                                                      const double2* __restrict__ t_base,
                                                      const double2* __restrict__ alpha,
                                                      const double2* __restrict__ omega) {
    const int slot = blockIdx.y;
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = pp.sys[slot];
    const size_t o = (size_t)s * n + e;
    const double2 ap = cmul(alpha[s], p_base[o]);
    const double2 os = cmul(omega[s], s_base[o]);
    const double2 ot = cmul(omega[s], t_base[o]);
    const double2 xv = x_base[o];
    x_base[o] = make_double2(xv.x + ap.x + os.x, xv.y + ap.y + os.y);
    const double2 sv = s_base[o];
    r_base[o] = make_double2(sv.x - ot.x, sv.y - ot.y);
}

inline int nblk_for(size_t n) { return (int)std::min<size_t>((n + kRedChunk - 1) / kRedChunk, kRedBlocks); }
inline size_t chunk_for(size_t n) {
    const size_t nb = (size_t)nblk_for(n);
    return (n + nb - 1) / nb;
}
inline dim3 grid_for(size_t n, int nslots) { return dim3((unsigned)((n + 255) / 256), (unsigned)nslots); }
inline dim3 dots_grid(size_t n, int nslots, int kk) {
    return dim3((unsigned)nblk_for(n), (unsigned)(nslots * ((kk + kDotGroup - 1) / kDotGroup)));
}

// Scratch owned by one solve call.
struct Work {
    size_t n = 0;
    int B = 0, m = 0;
    DevBuf<double2> V, w, r, t, H, g, h, y, partial;
    DevBuf<double2> pb, sb, vb, tb, r0b, alpha, omega, beta;
    DevBuf<double> nrm, tn, resid, cs, pnorm;
    DevBuf<int> rows, kk;
    DevBuf<double2> sn;
};

struct Ctx {
    EwaldOp& op;
    const mpm_solve_desc& d;
    hipStream_t s;
    size_t n;
    int N;
    Work W;
    std::vector<int> rows_h;
    double* host_pinned = nullptr;  // 2*kMaxSlots doubles
    Ctx(EwaldOp& o, const mpm_solve_desc& dd, hipStream_t st) : op(o), d(dd), s(st) {
        N = (int)op.N;
        n = op.n_unknowns();
        MPM_HIP(hipHostMalloc((void**)&host_pinned, 4 * kMaxSlots * sizeof(double)));
    }
    ~Ctx() {
        if (host_pinned) (void)hipHostFree(host_pinned);
    }

    // batched apply with per-system pointers; fp32 selects the reciprocal precision,
    // jacobi fuses the left preconditioner into the operator's store
    void apply(const std::vector<int>& sys, const std::vector<const double2*>& x,
               const std::vector<double2*>& y, bool fp32, bool jacobi = false) {
        SlotTable t;
        for (size_t j = 0; j < sys.size(); ++j) {
            t.x[j] = x[j];
            t.y[j] = y[j];
            t.row[j] = rows_h[sys[j]];
        }
        op.apply(t, (int)sys.size(), fp32, MPM_PART_ALL, s, jacobi);
    }

    // squared norms of vectors b[slot] -> out[sys] (W.nrm by default)
    void norms(Ptrs& p, int ns, double* out = nullptr) {
        const int nb = nblk_for(n);
        hipLaunchKernelGGL(dots_kernel, dots_grid(n, ns, 1), dim3(256), 0, s, p, 1, (size_t)0, n, chunk_for(n), 2, W.partial.p);
        MPM_HIP(hipGetLastError());
        // reduce into nrm (stored as double2 real part) via h scratch then copy
        hipLaunchKernelGGL(reduce_kernel, dim3((ns + 255) / 256), dim3(256), 0, s, p, ns, 1, nb,
                           W.partial.p, W.h.p, W.m + 1, 0, (double2*)nullptr);
        MPM_HIP(hipGetLastError());
        // gather real parts into out[sys]
        extract(p, ns, out ? out : W.nrm.p);
    }
    void extract(Ptrs& p, int ns, double* out);

    void read_scalars(const double* dev, const std::vector<int>& sys, std::vector<double>& out) {
        // dev indexed by system; copy the whole [B] array (small)
        MPM_HIP(hipMemcpyAsync(host_pinned, dev, W.B * sizeof(double), hipMemcpyDeviceToHost, s));
        MPM_HIP(hipStreamSynchronize(s));
        out.assign(W.B, 0.0);
        for (int q = 0; q < W.B; ++q) out[q] = host_pinned[q];
        (void)sys;
    }
};

__global__ void extract_kernel(Ptrs p, int ns, const double2* __restrict__ h, int hs, double* __restrict__ out) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= ns) return;
    out[p.sys[slot]] = h[(size_t)p.sys[slot] * hs].x;
}

void Ctx::extract(Ptrs& p, int ns, double* out) {
    hipLaunchKernelGGL(extract_kernel, dim3((ns + 63) / 64), dim3(64), 0, s, p, ns, W.h.p, W.m + 1, out);
    MPM_HIP(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------
// GMRES, streamed over slots
// ---------------------------------------------------------------------------
// Up to kMaxSlots systems occupy the slots at once; every workspace array (Krylov
// basis, Hessenberg matrix, rotations, residual vectors) is indexed by slot, and a
// system that finishes frees its slot for the next pending one, so that batches
// of any size run as one continuous stream.  Each system follows the reference's
// sequence on its own: outer residual M^-1 (b - A x) and its check, then an
// Arnoldi cycle of up to m steps (leaving early on the Givens estimate), the
// update x += V y, and again the outer residual.  A step of the stream advances
// every occupied slot by one stage: one batched apply for the slots that need a
// residual (FP32 and FP64 ones separately) and one for the slots in a cycle (each
// at its own column), then the Gram-Schmidt, Givens and residual work of all of
// them, and one read-back of the per-slot scalars that decide the next stage.
static void gmres_stream(Ctx& C, const double2* b, double2* x, int B, mpm_solve_report* rep) {
    Work& W = C.W;
    const size_t n = C.n;
    const mpm_solve_desc& d = C.d;
    const int m = (int)std::min<size_t>(n, (size_t)std::max(1, d.restart));
    MPM_REQUIRE(m <= 63, MPM_ERR_INVALID_ARGUMENT, "GMRES restart must be <= 63 in this library");
    const int S = std::min(B, kMaxSlots);
    W.n = n;
    W.B = S;
    W.m = m;
    W.V.reserve((size_t)S * (m + 1) * n);
    W.w.reserve((size_t)S * n);
    W.t.reserve((size_t)S * n);
    W.H.reserve((size_t)S * (m + 1) * m);
    W.g.reserve((size_t)S * (m + 1));
    W.h.reserve((size_t)S * (m + 1));
    W.y.reserve((size_t)S * (m + 1));
    W.partial.reserve((size_t)kMaxSlots * kRedBlocks * 64);
    const int nbx = (int)((n + 255) / 256);
    W.pnorm.reserve((size_t)kMaxSlots * nbx);
    // per-slot scalars in one block, read back with one copy per step: ||M^-1 r||^2 (also the
    // squared norm of the new Arnoldi vector), ||r||^2 of the outer residual, Givens estimate
    W.tn.reserve(3 * kMaxSlots);
    double* const s_nrm = W.tn.p;
    double* const s_tn = W.tn.p + kMaxSlots;
    double* const s_res = W.tn.p + 2 * kMaxSlots;
    W.cs.reserve((size_t)S * m);
    W.sn.reserve((size_t)S * m);
    W.kk.reserve(S);
    const bool inner32 = d.precision != MPM_SOLVE_DOUBLE;
    const bool outer32 = d.precision == MPM_SOLVE_MIXED;
    const double tol = d.tol;
    const size_t vstride = (size_t)(m + 1) * n;
    auto Vk = [&](int slot, int k) { return W.V.p + (size_t)slot * vstride + (size_t)k * n; };
    auto row_of = [&](int q) { return C.op.self_rows > 1 ? q : 0; };

    enum { PENDING = 0, RESIDUAL = 1, ARNOLDI = 2, DONE = 3 };
    std::vector<int> st(B, PENDING), kcol(B, 0), first(B, 1), iters(B, 0), matvecs(B, 0), mv64(B, 0), conv(B, 0),
        fresh64(B, 0), res32(B, 0);
    std::vector<double> bnorm(B, 0.0), last_pre(B, 0.0), last_true(B, 0.0);
    std::vector<int> slot_sys(S, -1), slot_of(B, -1);

    // squared norms of the vectors vec(q), q in sys (placed in slots 0.., |sys| <= S)
    auto norms_of = [&](const std::vector<int>& sys, auto vec, std::vector<double>& out) {
        Ptrs p;
        for (size_t j = 0; j < sys.size(); ++j) {
            p.a[j] = p.b[j] = vec(sys[j]);
            p.sys[j] = (int)j;
        }
        C.norms(p, (int)sys.size(), s_nrm);
        MPM_HIP(hipMemcpyAsync(C.host_pinned, s_nrm, sys.size() * sizeof(double), hipMemcpyDeviceToHost, C.s));
        MPM_HIP(hipStreamSynchronize(C.s));
        out.assign(C.host_pinned, C.host_pinned + sys.size());
    };
    // ||b|| in chunks of S; zero right-hand sides give x = 0 at once
    for (int q0 = 0; q0 < B; q0 += S) {
        std::vector<int> sys;
        for (int q = q0; q < std::min(B, q0 + S); ++q) sys.push_back(q);
        std::vector<double> nb2;
        norms_of(sys, [&](int q) { return (double2*)b + (size_t)q * n; }, nb2);
        for (size_t j = 0; j < sys.size(); ++j) {
            const int q = sys[j];
            bnorm[q] = std::sqrt(nb2[j]);
            if (bnorm[q] == 0.0) {
                MPM_HIP(hipMemsetAsync(x + (size_t)q * n, 0, n * sizeof(double2), C.s));
                st[q] = DONE;
                conv[q] = 1;
                fresh64[q] = 1;
            }
        }
    }
    size_t next_pending = 0;
    auto admit = [&]() {
        for (int sl = 0; sl < S; ++sl) {
            if (slot_sys[sl] >= 0) continue;
            while (next_pending < (size_t)B && st[next_pending] != PENDING) ++next_pending;
            if (next_pending >= (size_t)B) return;
            const int q = (int)next_pending++;
            slot_sys[sl] = q;
            slot_of[q] = sl;
            st[q] = RESIDUAL;
        }
    };
    auto release = [&](int q) {
        slot_sys[slot_of[q]] = -1;
        slot_of[q] = -1;
    };
    auto apply_slots = [&](const std::vector<int>& sys, auto x_of, bool fp32, bool jac) {
        if (sys.empty()) return;
        SlotTable t;
        for (size_t j = 0; j < sys.size(); ++j) {
            t.x[j] = x_of(sys[j]);
            t.y[j] = W.w.p + (size_t)slot_of[sys[j]] * n;
            t.row[j] = row_of(sys[j]);
        }
        C.op.apply(t, (int)sys.size(), fp32, MPM_PART_ALL, C.s, jac);
        for (int q : sys) {
            matvecs[q]++;
            if (!fp32) mv64[q]++;
        }
    };
    // slot-indexed pointer tables: a[j] / b[j] of system sys[j], sys field = its slot
    auto ptrs = [&](const std::vector<int>& sys, auto a_of, auto b_of) {
        Ptrs p;
        for (size_t j = 0; j < sys.size(); ++j) {
            const int q = sys[j];
            p.a[j] = a_of(q);
            p.b[j] = b_of(q);
            p.sys[j] = slot_of[q];
            p.k[j] = kcol[q] + 1;
            p.c[j] = b + (size_t)q * n;
            p.row[j] = row_of(q);
        }
        return p;
    };
    const int nb = nblk_for(n);
    // Residual checks are batched: a system whose cycle ended waits while others are still
    // cycling, unless systems are queued for slots and a quarter of the slots wait (a
    // residual apply costs about as much for a few systems as for many).
    const int batch_wait = std::max(1, S / 4);
    while (true) {
        admit();
        std::vector<int> R32, R64, A, waiting;
        for (int sl = 0; sl < S; ++sl) {
            const int q = slot_sys[sl];
            if (q < 0) continue;
            if (st[q] == RESIDUAL)
                waiting.push_back(q);
            else
                A.push_back(q);
        }
        bool queued = false;
        for (size_t q = next_pending; q < (size_t)B && !queued; ++q) queued = st[q] == PENDING;
        if (A.empty() || (queued && (int)waiting.size() >= batch_wait)) {
            for (int q : waiting) {
                res32[q] = outer32 || (inner32 && first[q]);
                (res32[q] ? R32 : R64).push_back(q);
            }
        }
        if (R32.empty() && R64.empty() && A.empty()) break;
        std::vector<int> R(R32);
        R.insert(R.end(), R64.begin(), R64.end());
        // ---- applies: A x for the residual slots, M^-1 A V_k for the cycling ones ----
        apply_slots(R32, [&](int q) { return (const double2*)(x + (size_t)q * n); }, true, false);
        apply_slots(R64, [&](int q) { return (const double2*)(x + (size_t)q * n); }, false, false);
        apply_slots(A, [&](int q) { return (const double2*)Vk(slot_of[q], kcol[q]); }, inner32, d.jacobi != 0);
        const int nr = (int)R.size(), na = (int)A.size();
        // ---- outer residuals: V_0 = M^-1 (b - A x), t = b - A x, their norms ----
        if (nr) {
            Ptrs p = ptrs(R, [&](int q) { return W.w.p + (size_t)slot_of[q] * n; },
                          [&](int q) { return Vk(slot_of[q], 0); });
            hipLaunchKernelGGL(residual_kernel_t<true>, grid_for(n, nr), dim3(256), 0, C.s, p, n, b, W.t.p,
                               C.op.self_coef(), W.rows.p, C.N, d.jacobi, C.op.quad_idx_dev());
            MPM_HIP(hipGetLastError());
            Ptrs pt = ptrs(R, [&](int q) { return W.t.p + (size_t)slot_of[q] * n; },
                           [&](int q) { return W.t.p + (size_t)slot_of[q] * n; });
            C.norms(pt, nr, s_tn);
            Ptrs pn = ptrs(R, [&](int q) { return Vk(slot_of[q], 0); }, [&](int q) { return Vk(slot_of[q], 0); });
            C.norms(pn, nr, s_nrm);
            MPM_HIP(hipGetLastError());
        }
        // ---- Arnoldi step of the cycling slots: CGS2, normalisation, Givens ----
        if (na) {
            int kmax = 0;
            for (int q : A) kmax = std::max(kmax, kcol[q] + 1);
            Ptrs pw = ptrs(A, [&](int q) { return Vk(slot_of[q], 0); }, [&](int q) { return W.w.p + (size_t)slot_of[q] * n; });
            for (int pass = 0; pass < 2; ++pass) {
                ProfScope ps("krylov", C.s);
                hipLaunchKernelGGL(dots_kernel_t<true>, dots_grid(n, na, kmax), dim3(256), 0, C.s, pw, kmax, n, n,
                                   chunk_for(n), 0, W.partial.p);
                hipLaunchKernelGGL(cgs_axpy_kernel_t<true>, grid_for(n, na), dim3(256), 0, C.s, pw, kmax, n, n, nb,
                                   (const double2*)W.partial.p, W.h.p, m + 1, pass,
                                   pass == 1 ? W.pnorm.p : (double*)nullptr, nbx);
                MPM_HIP(hipGetLastError());
            }
            Ptrs pc = ptrs(A, [&](int q) { return W.w.p + (size_t)slot_of[q] * n; },
                           [&](int q) { return Vk(slot_of[q], kcol[q] + 1); });
            hipLaunchKernelGGL(scale_norm_kernel, grid_for(n, na), dim3(256), 0, C.s, pc, n, (const double*)W.pnorm.p,
                               nbx, nbx, s_nrm);
            hipLaunchKernelGGL(givens_kernel_t<true>, dim3(na), dim3(64), 0, C.s, pw, na, 0, m, W.H.p, W.h.p,
                               (const double*)s_nrm, W.g.p, W.cs.p, W.sn.p, s_res);
            MPM_HIP(hipGetLastError());
        }
        // ---- one read-back: ||M^-1 r||^2, ||r||^2 of the residual slots, Givens estimates ----
        double* hp = C.host_pinned;
        MPM_HIP(hipMemcpyAsync(hp, s_nrm, 3 * kMaxSlots * sizeof(double), hipMemcpyDeviceToHost, C.s));
        MPM_HIP(hipStreamSynchronize(C.s));
        // ---- residual slots: converged, out of iterations, or a new cycle ----
        std::vector<int> start;
        for (int q : R) {
            const int sl = slot_of[q];
            const double rn = std::sqrt(hp[sl]);
            last_pre[q] = rn / bnorm[q];
            last_true[q] = std::sqrt(hp[kMaxSlots + sl]) / bnorm[q];
            fresh64[q] = !res32[q];
            // the initial REFINED residual is FP32 and never taken as converged
            if (rn < tol * bnorm[q] && (!(res32[q] && !outer32) || rn == 0.0)) {
                st[q] = DONE;
                conv[q] = 1;
            } else if (iters[q] >= d.maxiter) {
                st[q] = DONE;  // maxiter exhausted (reference returns the partial x silently)
            } else {
                st[q] = ARNOLDI;
                kcol[q] = 0;
                start.push_back(q);
            }
            first[q] = 0;
        }
        // ---- cycling slots: the Givens estimate decides whether the cycle goes on ----
        std::vector<int> finish;
        for (int q : A) {
            const int sl = slot_of[q];
            iters[q]++;
            const bool est = hp[2 * kMaxSlots + sl] < tol * bnorm[q];
            if (est || kcol[q] + 1 >= m || iters[q] >= d.maxiter)
                finish.push_back(q);
            else
                kcol[q]++;
        }
        // new cycles: V_0 = r / ||r||, g = (||r||, 0, ...)
        if (!start.empty()) {
            Ptrs pv = ptrs(start, [&](int q) { return Vk(slot_of[q], 0); }, [&](int q) { return Vk(slot_of[q], 0); });
            const int ns = (int)start.size();
            hipLaunchKernelGGL(scale_copy_kernel, grid_for(n, ns), dim3(256), 0, C.s, pv, n, (const double*)s_nrm);
            hipLaunchKernelGGL(ginit_kernel, dim3(ns), dim3(64), 0, C.s, pv, ns, m, W.g.p, (const double*)s_nrm);
            MPM_HIP(hipGetLastError());
        }
        // ended cycles: y = H^-1 g, x += V y, then the outer residual again
        if (!finish.empty()) {
            const int ns = (int)finish.size();
            // columns per slot, staged in the pinned block after the three read-back rows (the
            // next write to it follows the next step's stream synchronisation)
            int* kk = reinterpret_cast<int*>(hp + 3 * kMaxSlots);
            std::fill(kk, kk + S, 0);
            for (int q : finish) kk[slot_of[q]] = kcol[q] + 1;
            MPM_HIP(hipMemcpyAsync(W.kk.p, kk, S * sizeof(int), hipMemcpyHostToDevice, C.s));
            Ptrs px = ptrs(finish, [&](int q) { return Vk(slot_of[q], 0); }, [&](int q) { return x + (size_t)q * n; });
            hipLaunchKernelGGL(backsolve_kernel, dim3(ns), dim3(64), (size_t)m * (m + 1) / 2 * sizeof(double2), C.s,
                               px, ns, W.kk.p, m, W.H.p, W.g.p, W.y.p);
            hipLaunchKernelGGL(xupdate_kernel, grid_for(n, ns), dim3(256), 0, C.s, px, n, n, W.kk.p, W.y.p, m + 1);
            MPM_HIP(hipGetLastError());
            for (int q : finish) {
                fresh64[q] = 0;
                st[q] = iters[q] < d.maxiter ? RESIDUAL : DONE;
            }
        }
        for (int sl = 0; sl < S; ++sl)
            if (slot_sys[sl] >= 0 && st[slot_sys[sl]] == DONE) release(slot_sys[sl]);
    }

    // Diagnostic FP64 residuals for systems whose last residual was not computed with
    // the FP64 operator at the final x (MIXED solves, maxiter exits).  Not counted.
    std::vector<int> need;
    for (int q = 0; q < B; ++q)
        if (bnorm[q] > 0.0 && !fresh64[q]) need.push_back(q);
    for (size_t q0 = 0; q0 < need.size(); q0 += S) {
        std::vector<int> sys(need.begin() + q0, need.begin() + std::min(need.size(), q0 + S));
        for (size_t j = 0; j < sys.size(); ++j) {
            slot_sys[j] = sys[j];
            slot_of[sys[j]] = (int)j;
        }
        SlotTable t;
        for (size_t j = 0; j < sys.size(); ++j) {
            t.x[j] = x + (size_t)sys[j] * n;
            t.y[j] = W.w.p + j * n;
            t.row[j] = row_of(sys[j]);
        }
        C.op.apply(t, (int)sys.size(), false, MPM_PART_ALL, C.s, false);
        Ptrs p = ptrs(sys, [&](int q) { return W.w.p + (size_t)slot_of[q] * n; },
                      [&](int q) { return W.w.p + (size_t)slot_of[q] * n; });
        hipLaunchKernelGGL(residual_kernel_t<true>, grid_for(n, (int)sys.size()), dim3(256), 0, C.s, p, n, b, W.t.p,
                           C.op.self_coef(), W.rows.p, C.N, d.jacobi, C.op.quad_idx_dev());
        MPM_HIP(hipGetLastError());
        std::vector<double> rn2, tn2;
        norms_of(sys, [&](int q) { return W.w.p + (size_t)slot_of[q] * n; }, rn2);
        norms_of(sys, [&](int q) { return W.t.p + (size_t)slot_of[q] * n; }, tn2);
        for (size_t j = 0; j < sys.size(); ++j) {
            last_pre[sys[j]] = std::sqrt(rn2[j]) / bnorm[sys[j]];
            last_true[sys[j]] = std::sqrt(tn2[j]) / bnorm[sys[j]];
        }
    }
    if (rep) {
        for (int q = 0; q < B; ++q) {
            rep[q].iters = iters[q];
            rep[q].matvecs = matvecs[q];
            rep[q].matvecs_fp64 = mv64[q];
            rep[q].relres_pre = last_pre[q];
            rep[q].relres_true = last_true[q];
            rep[q].converged = conv[q];
        }
    }
}

// ---------------------------------------------------------------------------
// BiCGSTAB (host scalars, batched vector work)
// ---------------------------------------------------------------------------
// Refinement cycles of the refined-precision BiCGSTAB (each restarts from the iterate).
constexpr int kMaxBicgCycles = 20;

static void bicgstab_group(Ctx& C, const double2* b, double2* x, int B, mpm_solve_report* rep) {
    Work& W = C.W;
    const size_t n = C.n;
    const mpm_solve_desc& d = C.d;
    W.n = n;
    W.B = B;
    W.m = 1;
    W.w.reserve((size_t)B * n);
    W.t.reserve((size_t)B * n);
    W.pb.reserve((size_t)B * n);
    W.sb.reserve((size_t)B * n);
    W.vb.reserve((size_t)B * n);
    W.tb.reserve((size_t)B * n);
    W.r0b.reserve((size_t)B * n);
    W.r.reserve((size_t)B * n);
    W.h.reserve((size_t)B * 2);
    W.partial.reserve((size_t)kMaxSlots * kRedBlocks * 64);
    W.nrm.reserve(B);
    W.alpha.reserve(B);
    W.omega.reserve(B);
    W.beta.reserve(B);
    W.rows.upload(C.rows_h.data(), B, C.s);
    const bool fp32 = d.precision != MPM_SOLVE_DOUBLE;
    const bool outer32 = d.precision == MPM_SOLVE_MIXED;
    const double tol = d.tol;

    auto make_ptrs = [&](const std::vector<int>& sys, auto a_of, auto b_of) {
        Ptrs p;
        for (size_t j = 0; j < sys.size(); ++j) {
            p.a[j] = a_of(sys[j]);
            p.b[j] = b_of(sys[j]);
            p.sys[j] = sys[j];
        }
        return p;
    };
    auto vec = [&](DevBuf<double2>& buf, int q) { return buf.p + (size_t)q * n; };
    std::vector<int> all(B);
    std::iota(all.begin(), all.end(), 0);
    std::vector<int> iters(B, 0), matvecs(B, 0), mv64(B, 0);
    std::vector<double> bnorm(B, 0.0), last_pre(B, 0.0), last_true(B, 0.0);

    auto norms_of = [&](const std::vector<int>& sys, DevBuf<double2>& buf, std::vector<double>& out) {
        Ptrs p = make_ptrs(sys, [&](int q) { return vec(buf, q); }, [&](int q) { return vec(buf, q); });
        C.norms(p, (int)sys.size());
        C.read_scalars(W.nrm.p, sys, out);
        for (auto& v : out) v = std::sqrt(v);
    };
    // dot over systems: mode 0 conj(a).b, mode 1 a.b  -> host complex per system
    auto dots_of = [&](const std::vector<int>& sys, DevBuf<double2>& A, DevBuf<double2>& Bv, int mode,
                       std::vector<cplx>& out) {
        Ptrs p = make_ptrs(sys, [&](int q) { return vec(A, q); }, [&](int q) { return vec(Bv, q); });
        const int nb = nblk_for(n);
        const int ns = (int)sys.size();
        hipLaunchKernelGGL(dots_kernel, dots_grid(n, ns, 1), dim3(256), 0, C.s, p, 1, (size_t)0, n, chunk_for(n), mode, W.partial.p);
        hipLaunchKernelGGL(reduce_kernel, dim3((ns + 255) / 256), dim3(256), 0, C.s, p, ns, 1, nb, W.partial.p,
                           W.h.p, 2, 0, (double2*)nullptr);
        MPM_HIP(hipGetLastError());
        MPM_HIP(hipMemcpyAsync(C.host_pinned, W.h.p, (size_t)B * 2 * sizeof(double2), hipMemcpyDeviceToHost, C.s));
        MPM_HIP(hipStreamSynchronize(C.s));
        out.assign(B, cplx(0, 0));
        const double2* hp = reinterpret_cast<const double2*>(C.host_pinned);
        for (int q : sys) out[q] = cplx(hp[q * 2].x, hp[q * 2].y);
    };
    auto upload_scalars = [&](DevBuf<double2>& dst, const std::vector<cplx>& v) {
        for (int q = 0; q < B; ++q) {
            C.host_pinned[2 * q] = v[q].real();
            C.host_pinned[2 * q + 1] = v[q].imag();
        }
        MPM_HIP(hipMemcpyAsync(dst.p, C.host_pinned, B * sizeof(double2), hipMemcpyHostToDevice, C.s));
        MPM_HIP(hipStreamSynchronize(C.s));
    };
    auto applyM = [&](const std::vector<int>& sys, auto in_of, DevBuf<double2>& out, bool use32) {
        std::vector<const double2*> xs;
        std::vector<double2*> ys;
        for (int q : sys) {
            xs.push_back(in_of(q));
            ys.push_back(vec(out, q));
        }
        C.apply(sys, xs, ys, use32, d.jacobi != 0);
        for (int q : sys) {
            matvecs[q]++;
            if (!use32) mv64[q]++;
        }
    };
    // b norms
    {
        std::vector<double> bn;
        Ptrs p = make_ptrs(all, [&](int q) { return (double2*)b + q * n; }, [&](int q) { return (double2*)b + q * n; });
        C.norms(p, B);
        C.read_scalars(W.nrm.p, all, bn);
        for (int q = 0; q < B; ++q) bnorm[q] = std::sqrt(bn[q]);
    }
    std::vector<int> active, pending;
    for (int q = 0; q < B; ++q) {
        if (bnorm[q] == 0.0)
            MPM_HIP(hipMemsetAsync(x + q * n, 0, n * sizeof(double2), C.s));
        else
            pending.push_back(q);
    }
    // Refined precision: FP32 applies inside the iteration, and a system whose recurrence
    // residual met the tolerance restarts from its iterate unless the FP64 residual
    // M^-1 (b - A x) also does (iterative refinement with BiCGSTAB as the inner solver).
    // DOUBLE and MIXED run one pass, as the reference.
    const bool refine = fp32 && !outer32;
    std::vector<char> by_recurrence(B, 0);
    int it_total = 0;
    for (int cycle = 0; !pending.empty(); ++cycle) {
    active.swap(pending);
    pending.clear();
    // r = M^-1 (b - A x0)
    if (!active.empty()) {
        std::vector<const double2*> xs;
        std::vector<double2*> ys;
        for (int q : active) {
            xs.push_back(x + q * n);
            ys.push_back(vec(W.w, q));
        }
        C.apply(active, xs, ys, outer32);
        for (int q : active) {
            matvecs[q]++;
            if (!outer32) mv64[q]++;
        }
        Ptrs p = make_ptrs(active, [&](int q) { return vec(W.w, q); }, [&](int q) { return vec(W.r, q); });
        hipLaunchKernelGGL(residual_kernel, grid_for(n, (int)active.size()), dim3(256), 0, C.s, p, n,
                           b, W.t.p, C.op.self_coef(), W.rows.p, C.N, d.jacobi, C.op.quad_idx_dev());
        MPM_HIP(hipGetLastError());
        std::vector<double> rn;
        norms_of(active, W.r, rn);
        std::vector<int> still;
        for (int q : active)
            if (!(rn[q] < tol * bnorm[q])) still.push_back(q);
        active.swap(still);
    }
    for (int q : active) {
        MPM_HIP(hipMemcpyAsync(vec(W.r0b, q), vec(W.r, q), n * sizeof(double2), hipMemcpyDeviceToDevice, C.s));
        MPM_HIP(hipMemcpyAsync(vec(W.pb, q), vec(W.r, q), n * sizeof(double2), hipMemcpyDeviceToDevice, C.s));
        MPM_HIP(hipMemsetAsync(vec(W.vb, q), 0, n * sizeof(double2), C.s));
    }
    std::vector<cplx> rho(B, 1.0), alpha(B, 1.0), omega(B, 1.0);
    for (int q = 0; q < B; ++q) by_recurrence[q] = 0;
    for (int it = 0; it_total < d.maxiter && !active.empty(); ++it, ++it_total) {
        std::vector<cplx> rho_new;
        dots_of(active, W.r0b, W.r, 1, rho_new);
        std::vector<int> go;
        for (int q : active)
            if (!(std::abs(rho_new[q]) < 1e-16)) go.push_back(q);
        active.swap(go);
        if (active.empty()) break;
        if (it > 0) {
            std::vector<cplx> beta(B, 0.0);
            for (int q : active) beta[q] = (rho_new[q] / rho[q]) * (alpha[q] / omega[q]);
            upload_scalars(W.beta, beta);
            upload_scalars(W.omega, omega);
            Ptrs pp = make_ptrs(active, [&](int q) { return vec(W.pb, q); }, [&](int q) { return vec(W.pb, q); });
            hipLaunchKernelGGL(bicg_p_kernel, grid_for(n, (int)active.size()), dim3(256), 0, C.s, pp, n,
                               W.vb.p, W.r.p, W.omega.p, W.beta.p);
            MPM_HIP(hipGetLastError());
        }
        for (int q : active) rho[q] = rho_new[q];
        applyM(active, [&](int q) { return (const double2*)vec(W.pb, q); }, W.vb, fp32);
        std::vector<cplx> rv;
        dots_of(active, W.r0b, W.vb, 1, rv);
        go.clear();
        for (int q : active) {
            iters[q]++;
            if (!(std::abs(rv[q]) < 1e-16)) {
                alpha[q] = rho[q] / rv[q];
                go.push_back(q);
            }
        }
        active.swap(go);
        if (active.empty()) break;
        upload_scalars(W.alpha, alpha);
        Ptrs ps = make_ptrs(active, [&](int q) { return vec(W.r, q); }, [&](int q) { return vec(W.sb, q); });
        hipLaunchKernelGGL(bicg_s_kernel, grid_for(n, (int)active.size()), dim3(256), 0, C.s, ps, n, W.vb.p, W.alpha.p);
        MPM_HIP(hipGetLastError());
        std::vector<double> sn;
        norms_of(active, W.sb, sn);
        std::vector<int> cont, stop_s;
        for (int q : active) (sn[q] < tol * bnorm[q] ? stop_s : cont).push_back(q);
        for (int q : stop_s) by_recurrence[q] = 1;
        if (!stop_s.empty()) {  // x += alpha p
            Ptrs pa = make_ptrs(stop_s, [&](int q) { return vec(W.pb, q); }, [&](int q) { return x + q * n; });
            hipLaunchKernelGGL(axpy_kernel, grid_for(n, (int)stop_s.size()), dim3(256), 0, C.s, pa, n, W.alpha.p);
            MPM_HIP(hipGetLastError());
        }
        active.swap(cont);
        if (active.empty()) break;
        applyM(active, [&](int q) { return (const double2*)vec(W.sb, q); }, W.tb, fp32);
        std::vector<cplx> ts, tt;
        dots_of(active, W.tb, W.sb, 0, ts);
        dots_of(active, W.tb, W.tb, 0, tt);
        go.clear();
        for (int q : active) {
            if (std::abs(tt[q]) < 1e-16) continue;  // breakdown: stop this system
            omega[q] = ts[q] / tt[q].real();
            go.push_back(q);
        }
        active.swap(go);
        if (active.empty()) break;
        upload_scalars(W.omega, omega);
        Ptrs px = make_ptrs(active, [&](int q) { return vec(W.pb, q); }, [&](int q) { return vec(W.pb, q); });
        hipLaunchKernelGGL(bicg_xr_kernel, grid_for(n, (int)active.size()), dim3(256), 0, C.s, px, n, x, W.r.p,
                           W.pb.p, W.sb.p, W.tb.p, W.alpha.p, W.omega.p);
        MPM_HIP(hipGetLastError());
        std::vector<double> rn;
        norms_of(active, W.r, rn);
        go.clear();
        for (int q : active) {
            if (rn[q] < tol * bnorm[q]) {
                by_recurrence[q] = 1;
                continue;
            }
            if (std::abs(omega[q]) < 1e-16) continue;
            go.push_back(q);
        }
        active.swap(go);
    }
    // systems stopped by their recurrence residual: the next cycle checks them in FP64
    if (refine && it_total < d.maxiter && cycle + 1 < kMaxBicgCycles)
        for (int q = 0; q < B; ++q)
            if (by_recurrence[q]) pending.push_back(q);
    }
    // final FP64 residuals (diagnostic)
    std::vector<int> need;
    for (int q = 0; q < B; ++q)
        if (bnorm[q] > 0.0) need.push_back(q);
    if (!need.empty()) {
        std::vector<const double2*> xs;
        std::vector<double2*> ys;
        for (int q : need) {
            xs.push_back(x + q * n);
            ys.push_back(vec(W.w, q));
        }
        C.apply(need, xs, ys, false);
        Ptrs p = make_ptrs(need, [&](int q) { return vec(W.w, q); }, [&](int q) { return vec(W.r, q); });
        hipLaunchKernelGGL(residual_kernel, grid_for(n, (int)need.size()), dim3(256), 0, C.s, p, n,
                           b, W.t.p, C.op.self_coef(), W.rows.p, C.N, d.jacobi, C.op.quad_idx_dev());
        MPM_HIP(hipGetLastError());
        std::vector<double> rn, tn;
        norms_of(need, W.r, rn);
        norms_of(need, W.t, tn);
        for (int q : need) {
            last_pre[q] = rn[q] / bnorm[q];
            last_true[q] = tn[q] / bnorm[q];
        }
    }
    if (rep) {
        for (int q = 0; q < B; ++q) {
            rep[q].iters = iters[q];
            rep[q].matvecs = matvecs[q];
            rep[q].matvecs_fp64 = mv64[q];
            rep[q].relres_pre = last_pre[q];
            rep[q].relres_true = last_true[q];
            rep[q].converged = (bnorm[q] == 0.0) || (last_pre[q] < tol);
        }
    }
}

void solve_batched(EwaldOp& op, const mpm_solve_desc& d, const double2* b, double2* x, int B,
                   mpm_solve_report* reports, hipStream_t s) {
    MPM_REQUIRE(op.N > 0, MPM_ERR_RUNTIME, "solve before set_positions");
    MPM_REQUIRE(B >= 1, MPM_ERR_INVALID_ARGUMENT, "need at least one system");
    MPM_REQUIRE(d.tol > 0.0, MPM_ERR_INVALID_ARGUMENT, "tol must be positive");
    MPM_REQUIRE(d.maxiter >= 0, MPM_ERR_INVALID_ARGUMENT, "maxiter must be >= 0");
    MPM_REQUIRE(d.method == MPM_GMRES || d.method == MPM_BICGSTAB, MPM_ERR_INVALID_ARGUMENT, "unknown method");
    MPM_REQUIRE(d.precision >= 0 && d.precision <= 2, MPM_ERR_INVALID_ARGUMENT, "unknown precision");
    MPM_REQUIRE(op.self_rows == 1 || op.self_rows >= B, MPM_ERR_INVALID_ARGUMENT,
                "need one self-coefficient row per system (or a single shared row)");
    const size_t n = op.n_unknowns();
    Ctx C(op, d, s);
    if (d.method == MPM_GMRES) {
        gmres_stream(C, b, x, B, reports);
        MPM_HIP(hipStreamSynchronize(s));
        return;
    }
    for (int b0 = 0; b0 < B; b0 += kMaxSlots) {
        const int cnt = std::min(kMaxSlots, B - b0);
        C.rows_h.assign(cnt, 0);
        for (int j = 0; j < cnt; ++j) C.rows_h[j] = op.self_rows > 1 ? b0 + j : 0;
        bicgstab_group(C, b + b0 * n, x + b0 * n, cnt, reports ? reports + b0 : nullptr);
    }
    MPM_HIP(hipStreamSynchronize(s));
}

}  // namespace mpm
