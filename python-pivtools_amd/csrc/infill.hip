// infill.hip -- least-squares infill of invalid vectors on the device.
//
// The reference fills the invalid vectors of each pass with a variational
// least-squares inpainting (pivtools_cli/piv/piv_backend/infilling.py:
// 144-260 _difference_rows / _fill_least_squares, called by infill_harmonic
// :336-370 and infill_biharmonic :373-410, the shipped default for both the
// mid and the final pass): minimise the squared first (harmonic) or second
// (biharmonic) differences along y and x, in physical units (divided by
// pitch^order), over the difference equations that touch a cell to fill and
// no excluded cell, every other valid cell held at its value.  Cells to fill
// are the NaN vectors; excluded cells are the masked windows and any other
// non-finite cell.  The reference solves the normal equations
// (D_F^T D_F) x = -D_F^T D_K u_K with a sparse LU in FP64.
//
// Here one workgroup owns one image and solves the same normal equations in
// FP64 with conjugate gradients, matrix-free: (A p)_i is the sum over the
// difference rows through cell i of its coefficient times the row applied
// to p (p is kept on the grid, zero outside the fill cells, so known and
// excluded cells drop out by themselves).  Both components share the rows
// and run as two CG recurrences in one loop.  The iteration stops at a
// relative residual of 1e-13 (the solve is then as exact as the LU in FP64
// terms) or after 4 nf + 200 iterations; a residual above the reference's
// own acceptance (1e-6) flags *err |= 2.  A group of fill cells that no
// kept row ties to a known cell has no determined value (the reference
// raises): it is detected by propagating "anchored" along kept rows, and
// flags *err |= 4 (the cells are left at 0).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pivamd_dev.h"

namespace pivamd {

enum : uint8_t { kKnown = 0, kFill = 1, kExcl = 2 };

constexpr int kLsThreads = 1024;

struct LsImage {
    double *p;          // (2, G) search directions on the grid
    double *x;          // (2, G) compact solution (first nf used)
    double *r;          // (2, G) compact residuals
    double *ap;         // (2, G) compact A p
    int *idx;           // (G) compact list of fill cells
    uint8_t *cls;       // (G) cell class
    uint8_t *anch;      // (G) anchored flags
};

__host__ __device__ inline size_t ls_image_bytes(int G)
{
    const size_t b = (size_t)G * (8 * 2 * 4 + 4 + 1 + 1);
    return (b + 255) & ~(size_t)255;
}

__device__ inline LsImage ls_image(void *work, int G, int img)
{
    uint8_t *base = (uint8_t *)work + (size_t)img * ls_image_bytes(G);
    LsImage w;
    w.p = (double *)base;
    w.x = w.p + 2 * (size_t)G;
    w.r = w.x + 2 * (size_t)G;
    w.ap = w.r + 2 * (size_t)G;
    w.idx = (int *)(w.ap + 2 * (size_t)G);
    w.cls = (uint8_t *)(w.idx + G);
    w.anch = w.cls + G;
    return w;
}

// block-wide sums of two doubles (all lanes get the result)
__device__ inline void block_sum2(double &a, double &b, double *s_red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        s_red[2 * wv] = a;
        s_red[2 * wv + 1] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
    for (int k = 0; k < nw; ++k) {
        a += s_red[2 * k];
        b += s_red[2 * k + 1];
    }
}

__device__ inline int block_or(int v, int *s_flag)
{
    __syncthreads();
    if (threadIdx.x == 0) *s_flag = 0;
    __syncthreads();
    if (v) atomicOr(s_flag, 1);
    __syncthreads();
    return *s_flag;
}

// The difference rows through cell (r, c): for each, the row's cells, their
// coefficients and the position of (r, c) in it.  ORDER 1: (-1, 1)/h along
// y and x; ORDER 2: (1, -2, 1)/h^2.  At most 4 (ORDER 1) or 6 rows.
template <int ORDER>
struct Rows {
    static constexpr int kLen = ORDER + 1;
    static constexpr int kMax = 2 * (ORDER + 1);
    int n;
    int cell[kMax][kLen];
    int self[kMax];
    double coef[kMax][kLen];
};

template <int ORDER>
__device__ inline void rows_through(int r, int c, int ny, int nx, double py, double px, Rows<ORDER> &R)
{
    R.n = 0;
    for (int axis = 0; axis < 2; ++axis) {
        const int n = axis == 0 ? ny : nx;
        const int pos = axis == 0 ? r : c;
        if (n <= ORDER) continue;
        const double h = axis == 0 ? py : px;
        const double w = ORDER == 1 ? 1.0 / h : 1.0 / (h * h);
        for (int s = pos - ORDER; s <= pos; ++s) {
            if (s < 0 || s + ORDER > n - 1) continue;
            const int k = R.n++;
            for (int j = 0; j <= ORDER; ++j) {
                const int rr = axis == 0 ? s + j : r, cc = axis == 0 ? c : s + j;
                R.cell[k][j] = rr * nx + cc;
                R.coef[k][j] = ORDER == 1 ? (j == 0 ? -w : w) : (j == 1 ? -2.0 * w : w);
            }
            R.self[k] = pos - s;
        }
    }
}

template <int ORDER>
__global__ void __launch_bounds__(kLsThreads) infill_ls_kernel(const float *u, const float *mask, int N, int ny,
                                                               int nx, double py, double px, float *uo,
                                                               void *work, int *err)
{
    __shared__ double s_red[2 * (kLsThreads / 64)];
    __shared__ int s_scan[kLsThreads / 64];
    __shared__ int s_flag;
    const int img = blockIdx.x;
    const int G = ny * nx;
    const size_t total = (size_t)N * G;
    const float *uy = u + (size_t)img * G;
    const float *ux = u + total + (size_t)img * G;
    float *oy = uo + (size_t)img * G;
    float *ox = uo + total + (size_t)img * G;
    LsImage w = ls_image(work, G, img);
    const int tid = threadIdx.x, nt = blockDim.x;

    // classify, copy through, compact the fill cells (ordered block scan)
    int nf = 0;
    for (int base = 0; base < G; base += nt) {
        const int i = base + tid;
        int f = 0;
        if (i < G) {
            const float a = uy[i], b = ux[i];
            const bool fill = (a != a) || (b != b);
            const bool excl = (mask && mask[i] != 0.0f) || (!fill && !(isfinite(a) && isfinite(b)));
            f = fill && !excl;
            w.cls[i] = excl ? kExcl : (fill ? kFill : kKnown);
            w.anch[i] = 0;
            w.p[i] = 0.0;
            w.p[(size_t)G + i] = 0.0;
            oy[i] = a;
            ox[i] = b;
        }
        // exclusive scan of f over the block: wave ballot + per-wave totals
        const unsigned long long bal = __ballot(f);
        const int lane = tid & 63, wv = tid >> 6;
        const int in_wave = __popcll(bal & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) s_scan[wv] = __popcll(bal);
        __syncthreads();
        int off = nf;
        for (int k = 0; k < wv; ++k) off += s_scan[k];
        if (f) w.idx[off + in_wave] = i;
        for (int k = 0; k < (nt >> 6); ++k) nf += s_scan[k];
    }
    __syncthreads();
    if (nf == 0) return;

    // right-hand side b = -D_F^T (D_K u_K) over kept rows; anchored seeds
    double bb0 = 0.0, bb1 = 0.0;
    for (int k = tid; k < nf; k += nt) {
        const int i = w.idx[k];
        Rows<ORDER> R;
        rows_through<ORDER>(i / nx, i % nx, ny, nx, py, px, R);
        double b0 = 0.0, b1 = 0.0;
        int anch = 0;
        for (int q = 0; q < R.n; ++q) {
            bool keep = true;
            for (int j = 0; j <= ORDER; ++j) keep = keep && w.cls[R.cell[q][j]] != kExcl;
            if (!keep) continue;
            double k0 = 0.0, k1 = 0.0;
            for (int j = 0; j <= ORDER; ++j) {
                const int cj = R.cell[q][j];
                if (w.cls[cj] == kKnown) {
                    k0 += R.coef[q][j] * (double)uy[cj];
                    k1 += R.coef[q][j] * (double)ux[cj];
                    anch = 1;
                }
            }
            b0 -= R.coef[q][R.self[q]] * k0;
            b1 -= R.coef[q][R.self[q]] * k1;
        }
        w.r[k] = b0;
        w.r[(size_t)G + k] = b1;
        w.x[k] = 0.0;
        w.x[(size_t)G + k] = 0.0;
        w.p[i] = b0;
        w.p[(size_t)G + i] = b1;
        w.anch[i] = (uint8_t)anch;
        bb0 += b0 * b0;
        bb1 += b1 * b1;
    }
    block_sum2(bb0, bb1, s_red);

    // propagate "anchored" through kept rows between fill cells
    for (;;) {
        int changed = 0;
        for (int k = tid; k < nf; k += nt) {
            const int i = w.idx[k];
            if (w.anch[i]) continue;
            Rows<ORDER> R;
            rows_through<ORDER>(i / nx, i % nx, ny, nx, py, px, R);
            for (int q = 0; q < R.n && !w.anch[i]; ++q) {
                bool keep = true;
                for (int j = 0; j <= ORDER; ++j) keep = keep && w.cls[R.cell[q][j]] != kExcl;
                if (!keep) continue;
                for (int j = 0; j <= ORDER; ++j) {
                    const int cj = R.cell[q][j];
                    if (cj != i && w.cls[cj] == kFill && w.anch[cj]) {
                        w.anch[i] = 1;
                        changed = 1;
                        break;
                    }
                }
            }
        }
        if (!block_or(changed, &s_flag)) break;
    }
    int stuck = 0;
    for (int k = tid; k < nf; k += nt) stuck |= !w.anch[w.idx[k]];
    if (block_or(stuck, &s_flag)) {
        // undetermined values: the reference refuses such a field
        for (int k = tid; k < nf; k += nt) {
            oy[w.idx[k]] = 0.0f;
            ox[w.idx[k]] = 0.0f;
        }
        if (tid == 0) atomicOr(err, 4);
        return;
    }

    // conjugate gradients, both components in one loop
    double rr0 = bb0, rr1 = bb1;
    const double tol2_0 = 1e-26 * bb0, tol2_1 = 1e-26 * bb1;
    bool done0 = rr0 <= tol2_0, done1 = rr1 <= tol2_1;
    const int max_iter = 4 * nf + 200;
    int it = 0;
    for (; it < max_iter && !(done0 && done1); ++it) {
        __syncthreads();          // p fully updated before the matvec reads it
        double pap0 = 0.0, pap1 = 0.0;
        for (int k = tid; k < nf; k += nt) {
            const int i = w.idx[k];
            Rows<ORDER> R;
            rows_through<ORDER>(i / nx, i % nx, ny, nx, py, px, R);
            double a0 = 0.0, a1 = 0.0;
            for (int q = 0; q < R.n; ++q) {
                bool keep = true;
                for (int j = 0; j <= ORDER; ++j) keep = keep && w.cls[R.cell[q][j]] != kExcl;
                if (!keep) continue;
                double d0 = 0.0, d1 = 0.0;
                for (int j = 0; j <= ORDER; ++j) {
                    const int cj = R.cell[q][j];
                    d0 += R.coef[q][j] * w.p[cj];
                    d1 += R.coef[q][j] * w.p[(size_t)G + cj];
                }
                a0 += R.coef[q][R.self[q]] * d0;
                a1 += R.coef[q][R.self[q]] * d1;
            }
            w.ap[k] = a0;
            w.ap[(size_t)G + k] = a1;
            pap0 += w.p[i] * a0;
            pap1 += w.p[(size_t)G + i] * a1;
        }
        block_sum2(pap0, pap1, s_red);
        if ((!done0 && !(pap0 > 0.0)) || (!done1 && !(pap1 > 0.0))) {
            if (tid == 0) atomicOr(err, 2);
            break;
        }
        const double al0 = done0 ? 0.0 : rr0 / pap0, al1 = done1 ? 0.0 : rr1 / pap1;
        double nr0 = 0.0, nr1 = 0.0;
        for (int k = tid; k < nf; k += nt) {
            const int i = w.idx[k];
            const double r0 = w.r[k] - al0 * w.ap[k];
            const double r1 = w.r[(size_t)G + k] - al1 * w.ap[(size_t)G + k];
            w.x[k] += al0 * w.p[i];
            w.x[(size_t)G + k] += al1 * w.p[(size_t)G + i];
            w.r[k] = r0;
            w.r[(size_t)G + k] = r1;
            nr0 += r0 * r0;
            nr1 += r1 * r1;
        }
        block_sum2(nr0, nr1, s_red);
        const double be0 = done0 ? 0.0 : nr0 / rr0, be1 = done1 ? 0.0 : nr1 / rr1;
        for (int k = tid; k < nf; k += nt) {
            const int i = w.idx[k];
            if (!done0) w.p[i] = w.r[k] + be0 * w.p[i];
            if (!done1) w.p[(size_t)G + i] = w.r[(size_t)G + k] + be1 * w.p[(size_t)G + i];
        }
        if (!done0) {
            rr0 = nr0;
            done0 = rr0 <= tol2_0;
        }
        if (!done1) {
            rr1 = nr1;
            done1 = rr1 <= tol2_1;
        }
    }
    // the reference accepts a relative residual up to 1e-6
    if ((rr0 > 1e-12 * bb0 || rr1 > 1e-12 * bb1) && tid == 0) atomicOr(err, 2);
    __syncthreads();
    for (int k = tid; k < nf; k += nt) {
        const int i = w.idx[k];
        oy[i] = (float)w.x[k];
        ox[i] = (float)w.x[(size_t)G + k];
    }
}

}  // namespace pivamd

using namespace pivamd;

extern "C" unsigned long pivamd_infill_ls_workspace(int N, int ny, int nx)
{
    if (N <= 0 || ny <= 0 || nx <= 0) return 0;
    return (unsigned long)((size_t)N * ls_image_bytes(ny * nx));
}

extern "C" int pivamd_infill_ls(const float *u, const float *mask, int N, int ny, int nx, float py, float px,
                                int order, float *uo, void *d_work, unsigned long work_bytes, int *d_err,
                                hipStream_t st)
{
    if (N <= 0 || ny <= 0 || nx <= 0) return 0;
    if ((order != 1 && order != 2) || !(py > 0.0f) || !(px > 0.0f) || !u || !uo || !d_err) return PIVAMD_ERR_ARG;
    if (u == uo) return PIVAMD_ERR_ARG;
    if (!d_work || work_bytes < pivamd_infill_ls_workspace(N, ny, nx)) return PIVAMD_ERR_NOMEM;
    if (order == 1)
        hipLaunchKernelGGL(infill_ls_kernel<1>, dim3(N), dim3(kLsThreads), 0, st, u, mask, N, ny, nx, (double)py,
                           (double)px, uo, d_work, d_err);
    else
        hipLaunchKernelGGL(infill_ls_kernel<2>, dim3(N), dim3(kLsThreads), 0, st, u, mask, N, ny, nx, (double)py,
                           (double)px, uo, d_work, d_err);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

// ---- k-nearest-neighbour regression infill --------------------------------
// infilling.py:609-695 (infill_knn): each NaN vector (not excluded) becomes
// the k-nearest-neighbour regression of the valid vectors in physical
// coordinates (x * px, y * py), weights 1 / distance ('distance') or equal
// ('uniform'), k = min(n_neighbors, valid donors) -- the prediction of
// sklearn's KNeighborsRegressor, in FP64, stored as float32.  Neighbours at
// equal distance are taken in row-major order (a tie at the k-th neighbour
// is where a KD-tree may choose another donor of the same distance).  The
// search grows square rings around the cell and stops once the ring cannot
// hold anything nearer than the current k-th neighbour.  One lane per cell.
namespace pivamd {

constexpr int kKnnMax = 128;

__global__ void knn_fill_kernel(const float *u, const float *excl, int N, int ny, int nx, float py, float px,
                                int k, int uniform, int n_donor_min, const int *n_donors, float *out, int *err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)ny * nx, total = (long long)N * per;
    if (i >= total) return;
    const int n = (int)(i / per), w = (int)(i - (long long)n * per);
    const float *uy = u + (size_t)n * per, *ux = u + (size_t)total + (size_t)n * per;
    float *oy = out + (size_t)n * per, *ox = out + (size_t)total + (size_t)n * per;
    const float y0 = uy[w], x0 = ux[w];
    const bool ex = excl && excl[w] != 0.0f;
    const bool hole = !ex && !(y0 == y0 && x0 == x0);
    oy[w] = y0;
    ox[w] = x0;
    if (!hole) return;
    const int nd = n_donors[n];
    if (nd < n_donor_min) {
        atomicOr(err, 1);
        return;
    }
    const int kk = k < nd ? k : nd;
    double bd[kKnnMax];
    int bi[kKnnMax];
    int cnt = 0;
    const int r0 = w / nx, c0 = w % nx;
    const double dpy = (double)py, dpx = (double)px;
    const double step = dpy < dpx ? dpy : dpx;
    const int rmax = ny > nx ? ny : nx;
    for (int rad = 1; rad <= rmax; ++rad) {
        // nothing in this ring is nearer than rad * min(py, px)
        if (cnt == kk && (double)rad * step > bd[kk - 1]) break;
        for (int dr = -rad; dr <= rad; ++dr) {
            const int rr = r0 + dr;
            if (rr < 0 || rr >= ny) continue;
            const bool edge = dr == -rad || dr == rad;
            for (int dc = -rad; dc <= rad; dc += edge ? 1 : 2 * rad) {
                const int cc = c0 + dc;
                if (cc < 0 || cc >= nx) continue;
                const int q = rr * nx + cc;
                if (excl && excl[q] != 0.0f) continue;
                const float vy = uy[q], vx = ux[q];
                if (!(vy == vy && vx == vx) || isinf(vy) || isinf(vx)) continue;
                const double ey = (double)dr * dpy, ex2 = (double)dc * dpx;
                const double d = sqrt(ex2 * ex2 + ey * ey);
                if (cnt == kk && (d > bd[kk - 1] || (d == bd[kk - 1] && q > bi[kk - 1]))) continue;
                // insert (d, q), ordered by distance then row-major index
                int pos = cnt < kk ? cnt : kk - 1;
                while (pos > 0 && (bd[pos - 1] > d || (bd[pos - 1] == d && bi[pos - 1] > q))) {
                    bd[pos] = bd[pos - 1];
                    bi[pos] = bi[pos - 1];
                    --pos;
                }
                bd[pos] = d;
                bi[pos] = q;
                if (cnt < kk) ++cnt;
            }
        }
    }
    double sy = 0.0, sx = 0.0, sw = 0.0;
    for (int j = 0; j < cnt; ++j) {
        const double wt = uniform ? 1.0 : 1.0 / bd[j];
        sy += (double)uy[bi[j]] * wt;
        sx += (double)ux[bi[j]] * wt;
        sw += wt;
    }
    oy[w] = (float)(sy / sw);
    ox[w] = (float)(sx / sw);
}

// Valid donors per image (finite, not excluded).
__global__ void knn_count_kernel(const float *u, const float *excl, int N, int per, int *n_donors)
{
    const int n = blockIdx.x;
    const long long total = (long long)N * per;
    int c = 0;
    for (int w = threadIdx.x; w < per; w += blockDim.x) {
        const float a = u[(size_t)n * per + w], b = u[(size_t)total + (size_t)n * per + w];
        c += (!(excl && excl[w] != 0.0f) && isfinite(a) && isfinite(b)) ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ int s[16];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += s[k];
        n_donors[n] = t;
    }
}

}  // namespace pivamd

// d_work: N ints.  *d_err |= 1 when an image has holes and no donor.
extern "C" unsigned long pivamd_infill_knn_workspace(int N) { return (unsigned long)(N > 0 ? N : 1) * sizeof(int); }

extern "C" int pivamd_infill_knn(const float *u, const float *d_excl, int N, int ny, int nx, float py, float px,
                                 int n_neighbors, int uniform, float *out, void *d_work, unsigned long work_bytes,
                                 int *d_err, hipStream_t st)
{
    using namespace pivamd;
    if (N < 0 || ny <= 0 || nx <= 0 || !u || !out || !d_err || !d_work) return PIVAMD_ERR_ARG;
    if (n_neighbors < 1 || n_neighbors > kKnnMax) return PIVAMD_ERR_ARG;
    if (work_bytes < pivamd_infill_knn_workspace(N)) return PIVAMD_ERR_ARG;
    if (N == 0) return 0;
    int *nd = (int *)d_work;
    hipLaunchKernelGGL(knn_count_kernel, dim3(N), dim3(1024), 0, st, u, d_excl, N, ny * nx, nd);
    const long long total = (long long)N * ny * nx;
    hipLaunchKernelGGL(knn_fill_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, u, d_excl, N, ny,
                       nx, py, px, n_neighbors, uniform, 1, nd, out, d_err);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// Local RBF infill (infilling.py:505-597 infill_rbf_local: scipy's
// RBFInterpolator with `neighbors`, one interpolator per component).  Each
// hole (a NaN vector, not excluded) is interpolated from its k = min(k, donors)
// nearest donors in physical coordinates (x = col * px, y = row * py), ranked
// by distance and then row-major index (a KD-tree may keep another donor
// where several share the k-th distance):
//   phi(eps |y_i - y_j|) coefficients plus the monomials of degree <= deg of
//   yhat = (y - shift) / scale (shift, scale: mid-point and half-range of the
//   neighbourhood, a zero half-range taken as 1), the donors in ascending
//   index order, smoothing added to the kernel diagonal:
//     [ K + s I   P ] [c]   [d]
//     [ P^T       0 ] [a] = [0]
// solved by LU with partial pivoting in FP64 (LAPACK's dgesv in the
// reference), then evaluated at the hole.  One workgroup per hole: the
// candidates come from a rectangle that must hold the k nearest, the
// neighbourhood from ranking them, and the (k + m) x (k + m + 2) augmented
// system sits in LDS.  Kernels and default degrees are scipy's.
// *err bits: 1 = fewer than 4 donors (the reference raises), or fewer than
// the monomials; 2 = singular system; 4 = the candidate rectangle exceeds
// the LDS buffer (very anisotropic pitch).
// ---------------------------------------------------------------------------
namespace pivamd {

constexpr int kRbfMaxNbr = 128, kRbfMaxMono = 6;
constexpr int kRbfThreads = 256;

enum : int { kRbfLinear = 0, kRbfTps, kRbfCubic, kRbfQuintic, kRbfMq, kRbfImq, kRbfIq, kRbfGauss };

__device__ __forceinline__ double rbf_phi(int kern, double r)
{
    switch (kern) {
    case kRbfLinear: return -r;
    case kRbfTps: return r == 0.0 ? 0.0 : r * r * log(r);
    case kRbfCubic: return r * r * r;
    case kRbfQuintic: return -(r * r * r * r * r);
    case kRbfMq: return -sqrt(r * r + 1.0);
    case kRbfImq: return 1.0 / sqrt(r * r + 1.0);
    case kRbfIq: return 1.0 / (r * r + 1.0);
    default: return exp(-(r * r));
    }
}

__host__ __device__ constexpr int rbf_degree(int kern)
{
    return kern == kRbfTps || kern == kRbfCubic ? 1 : kern == kRbfQuintic ? 2 : 0;
}

// monomial m of degree <= 2 in scipy's order: 1, x, y, x^2, x y, y^2
__device__ __forceinline__ double rbf_mono(int m, double x, double y)
{
    switch (m) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return y;
    case 3: return x * x;
    case 4: return x * y;
    default: return y * y;
    }
}

// Block-wide sum / min / max / argmax helpers (kRbfThreads lanes, 4 waves).
__device__ __forceinline__ double rbf_wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int OP>    // 0 sum, 1 min, 2 max
__device__ double rbf_block_reduce(double v, double *red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = OP == 0 ? v + w : OP == 1 ? fmin(v, w) : fmax(v, w);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int k = 1; k < kRbfThreads / 64; ++k) r = OP == 0 ? r + red[k] : OP == 1 ? fmin(r, red[k]) : fmax(r, red[k]);
    return r;
}

// argmax of |v| over the lanes, first (smallest) row on ties; rows < 0 never win
__device__ int rbf_block_argmax(double v, int row, double *red, int *redi)
{
    double a = row >= 0 ? fabs(v) : -1.0;
    int r = row >= 0 ? row : 0x7fffffff;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double a2 = __shfl_xor(a, o, 64);
        const int r2 = __shfl_xor(r, o, 64);
        if (a2 > a || (a2 == a && r2 < r)) { a = a2; r = r2; }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = a;
        redi[threadIdx.x >> 6] = r;
    }
    __syncthreads();
    a = red[0];
    r = redi[0];
#pragma unroll
    for (int k = 1; k < kRbfThreads / 64; ++k)
        if (red[k] > a || (red[k] == a && redi[k] < r)) { a = red[k]; r = redi[k]; }
    return a > 0.0 ? r : -1;
}

__global__ void rbf_prep_kernel(const float *u, const float *excl, int N, int per, float *out, int *hole_count,
                                int *holes)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)N * per;
    if (i >= total) return;
    const int w = (int)(i % per);
    const float a = u[i], b = u[total + i];
    out[i] = a;
    out[total + i] = b;
    const bool ex = excl && excl[w] != 0.0f;
    if (!ex && !(a == a && b == b)) holes[atomicAdd(hole_count, 1)] = (int)i;
}

__global__ void rbf_zero_kernel(int *p) { *p = 0; }

// NMAX: largest system (k + monomials) the static LDS holds.
template <int NMAX>
__global__ void __launch_bounds__(kRbfThreads) rbf_fill_kernel(const float *u, const float *excl, int N, int ny,
                                                               int nx, double py, double px, int k_req, int kern,
                                                               double eps, double smooth, const int *n_donors,
                                                               const int *hole_count, const int *holes, float *out,
                                                               int *err)
{
    constexpr int PITCH = NMAX + 3;                       // 2 right-hand sides, odd pitch in doubles
    __shared__ double s_a[NMAX * PITCH];                  // candidates, then the augmented system
    __shared__ int s_nbr[kRbfMaxNbr], s_sel[kRbfMaxNbr];
    __shared__ double s_x[kRbfMaxNbr], s_y[kRbfMaxNbr];   // neighbour coordinates times eps
    __shared__ double s_red[kRbfThreads / 64];
    __shared__ int s_redi[kRbfThreads / 64];
    __shared__ int s_cnt;
    __shared__ double s_piv;
    constexpr int CAP = (NMAX * PITCH * 8) / 12;          // candidates: (distance, index) pairs
    double *c_d = s_a;
    int *c_q = (int *)(s_a + CAP);

    const int t = (int)threadIdx.x;
    const long long per = (long long)ny * nx, total = (long long)N * per;
    const int nh = *hole_count;
    const int deg = rbf_degree(kern);
    const int nm = deg == 0 ? 1 : deg == 1 ? 3 : 6;
    for (int h = blockIdx.x; h < nh; h += gridDim.x) {
        const long long i = holes[h];
        const int n = (int)(i / per), w = (int)(i - (long long)n * per);
        const int r0 = w / nx, c0 = w % nx;
        const float *uy = u + (size_t)n * per, *ux = u + (size_t)total + (size_t)n * per;
        const int nd = n_donors[n];
        const int k = k_req < nd ? k_req : nd;
        if (nd < 4 || nm > k) {
            if (t == 0) atomicOr(err, 1);
            continue;                                      // block-uniform
        }
        const int ns = k + nm;
        auto donor = [&](int q) {
            if (excl && excl[q] != 0.0f) return false;
            const float a = uy[q], b = ux[q];
            return isfinite(a) && isfinite(b);
        };
        // ---- 1. Chebyshev radius R whose square holds >= k donors ----------
        int cum = 0, R = 0;
        const int rmax = ny > nx ? ny : nx;
        while (cum < k && R < rmax) {
            ++R;
            int c = 0;
            for (int e = t; e < 8 * R; e += kRbfThreads) {
                // ring cell e: top row, bottom row, then the two side columns
                int dr, dc;
                if (e < 2 * R + 1) { dr = -R; dc = e - R; }
                else if (e < 4 * R + 2) { dr = R; dc = e - (2 * R + 1) - R; }
                else { const int f = e - (4 * R + 2); dr = -R + 1 + (f >> 1); dc = (f & 1) ? R : -R; }
                const int rr = r0 + dr, cc = c0 + dc;
                if (rr >= 0 && rr < ny && cc >= 0 && cc < nx && donor(rr * nx + cc)) ++c;
            }
            cum += (int)rbf_block_reduce<0>((double)c, s_red);
        }
        // ---- 2. candidates: every donor the k nearest can be -----------------
        const double qx = (double)c0 * px, qy = (double)r0 * py;
        // d_k <= R * |(py, px)|; the margin keeps the square's corner donors
        // whose computed distance rounds above it
        const double reach = (double)R * sqrt(py * py + px * px) * (1.0 + 1e-9);
        const int hr = (int)fmin((double)ny, ceil(reach / py)), hc = (int)fmin((double)nx, ceil(reach / px));
        const int rlo = r0 - hr < 0 ? 0 : r0 - hr, rhi = r0 + hr >= ny ? ny - 1 : r0 + hr;
        const int clo = c0 - hc < 0 ? 0 : c0 - hc, chi = c0 + hc >= nx ? nx - 1 : c0 + hc;
        const int rw = chi - clo + 1, cells = (rhi - rlo + 1) * rw;
        if (t == 0) s_cnt = 0;
        __syncthreads();
        for (int e = t; e < cells; e += kRbfThreads) {
            const int rr = rlo + e / rw, cc = clo + e % rw;
            const int q = rr * nx + cc;
            if (!donor(q)) continue;
            const double dx = (double)cc * px - qx, dy = (double)rr * py - qy;
            const double d = sqrt(dx * dx + dy * dy);
            if (d > reach) continue;
            const int slot = atomicAdd(&s_cnt, 1);
            if (slot < CAP) {
                c_d[slot] = d;
                c_q[slot] = q;
            }
        }
        __syncthreads();
        const int M = s_cnt;
        if (M > CAP || M < k) {
            if (t == 0) atomicOr(err, 4);
            __syncthreads();
            continue;
        }
        // ---- 3. the k nearest by (distance, index), then in index order ------
        for (int a = t; a < M; a += kRbfThreads) {
            const double d = c_d[a];
            const int q = c_q[a];
            int rank = 0;
            for (int b2 = 0; b2 < M; ++b2) {
                const double d2 = c_d[b2];
                rank += (d2 < d || (d2 == d && c_q[b2] < q)) ? 1 : 0;
            }
            if (rank < k) s_sel[rank] = q;
        }
        __syncthreads();
        for (int a = t; a < k; a += kRbfThreads) {
            const int q = s_sel[a];
            int rank = 0;
            for (int b2 = 0; b2 < k; ++b2) rank += s_sel[b2] < q ? 1 : 0;
            s_nbr[rank] = q;
        }
        __syncthreads();
        // ---- 4. neighbourhood coordinates, shift and scale ---------------------
        double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY;
        for (int a = t; a < k; a += kRbfThreads) {
            const int q = s_nbr[a];
            const double x = (double)(q % nx) * px, y = (double)(q / nx) * py;
            xmn = fmin(xmn, x); xmx = fmax(xmx, x);
            ymn = fmin(ymn, y); ymx = fmax(ymx, y);
        }
        xmn = rbf_block_reduce<1>(xmn, s_red);
        xmx = rbf_block_reduce<2>(xmx, s_red);
        ymn = rbf_block_reduce<1>(ymn, s_red);
        ymx = rbf_block_reduce<2>(ymx, s_red);
        const double shx = (xmx + xmn) / 2.0, shy = (ymx + ymn) / 2.0;
        double scx = (xmx - xmn) / 2.0, scy = (ymx - ymn) / 2.0;
        if (scx == 0.0) scx = 1.0;
        if (scy == 0.0) scy = 1.0;
        for (int a = t; a < k; a += kRbfThreads) {
            const int q = s_nbr[a];
            s_x[a] = (double)(q % nx) * px * eps;
            s_y[a] = (double)(q / nx) * py * eps;
        }
        __syncthreads();
        // ---- 5. the augmented system [K + sI, P | d; P^T, 0 | 0] ---------------
        const int ncol = ns + 2;
        for (int e = t; e < ns * ncol; e += kRbfThreads) {
            const int r = e / ncol, c = e - r * ncol;
            double v;
            if (r < k) {
                const int q = s_nbr[r];
                if (c < k) {
                    const double dx = s_x[r] - s_x[c], dy = s_y[r] - s_y[c];
                    v = rbf_phi(kern, sqrt(dx * dx + dy * dy));
                    if (r == c) v += smooth;
                } else if (c < ns) {
                    const double x = (double)(q % nx) * px, y = (double)(q / nx) * py;
                    v = rbf_mono(c - k, (x - shx) / scx, (y - shy) / scy);
                } else {
                    v = (double)(c == ns ? ux[q] : uy[q]);
                }
            } else {
                if (c < k) {
                    const int q = s_nbr[c];
                    const double x = (double)(q % nx) * px, y = (double)(q / nx) * py;
                    v = rbf_mono(r - k, (x - shx) / scx, (y - shy) / scy);
                } else {
                    v = 0.0;
                }
            }
            s_a[r * PITCH + c] = v;
        }
        __syncthreads();
        // ---- 6. LU with partial pivoting, eliminating into both right sides ----
        bool singular = false;
        for (int j = 0; j < ns; ++j) {
            // pivot: largest |a_ij|, i >= j, first on ties (idamax)
            double best = -1.0;
            int brow = -1;
            for (int r = j + t; r < ns; r += kRbfThreads) {
                const double a = fabs(s_a[r * PITCH + j]);
                if (a > best) { best = a; brow = r; }
            }
            const int p = rbf_block_argmax(best < 0.0 ? 0.0 : best, brow, s_red, s_redi);
            if (p < 0) { singular = true; break; }          // block-uniform
            if (p != j)
                for (int c = t; c < ncol; c += kRbfThreads) {
                    const double tmp = s_a[j * PITCH + c];
                    s_a[j * PITCH + c] = s_a[p * PITCH + c];
                    s_a[p * PITCH + c] = tmp;
                }
            __syncthreads();
            if (t == 0) s_piv = 1.0 / s_a[j * PITCH + j];
            __syncthreads();
            const double rp = s_piv;
            for (int r = j + 1 + t; r < ns; r += kRbfThreads) s_a[r * PITCH + j] *= rp;
            __syncthreads();
            const int rows = ns - j - 1, cols = ncol - j - 1;
            for (int e = t; e < rows * cols; e += kRbfThreads) {
                const int r = j + 1 + e / cols, c = j + 1 + e % cols;
                s_a[r * PITCH + c] -= s_a[r * PITCH + j] * s_a[j * PITCH + c];
            }
            __syncthreads();
        }
        if (singular) {
            if (t == 0) atomicOr(err, 2);
            __syncthreads();
            continue;
        }
        // ---- 7. back substitution (columns ns, ns + 1 become the coefficients) --
        for (int j = ns - 1; j >= 0; --j) {
            if (t < 2) s_a[j * PITCH + ns + t] /= s_a[j * PITCH + j];
            __syncthreads();
            for (int e = t; e < 2 * j; e += kRbfThreads) {
                const int r = e >> 1, c = ns + (e & 1);
                s_a[r * PITCH + c] -= s_a[r * PITCH + j] * s_a[j * PITCH + c];
            }
            __syncthreads();
        }
        // ---- 8. evaluate at the hole ------------------------------------------
        const double xe = qx * eps, ye = qy * eps;
        const double xh = (qx - shx) / scx, yh = (qy - shy) / scy;
        double vx = 0.0, vy = 0.0;
        for (int a = t; a < ns; a += kRbfThreads) {
            double f;
            if (a < k) {
                const double dx = xe - s_x[a], dy = ye - s_y[a];
                f = rbf_phi(kern, sqrt(dx * dx + dy * dy));
            } else {
                f = rbf_mono(a - k, xh, yh);
            }
            vx += f * s_a[a * PITCH + ns];
            vy += f * s_a[a * PITCH + ns + 1];
        }
        vx = rbf_block_reduce<0>(vx, s_red);
        vy = rbf_block_reduce<0>(vy, s_red);
        if (t == 0) {
            out[i] = (float)vy;
            out[total + i] = (float)vx;
        }
        __syncthreads();
    }
}

}  // namespace pivamd

// d_work: N donor counts, the hole count and up to N ny nx hole indices.
extern "C" unsigned long pivamd_infill_rbf_workspace(int N, int ny, int nx)
{
    const long long cells = (long long)(N > 0 ? N : 1) * (ny > 0 ? ny : 1) * (nx > 0 ? nx : 1);
    return (unsigned long)((N > 0 ? N : 1) + 1 + cells) * sizeof(int);
}

extern "C" int pivamd_infill_rbf(const float *u, const float *d_excl, int N, int ny, int nx, float py, float px,
                                 int neighbors, int kernel, double epsilon, double smoothing, float *out,
                                 void *d_work, unsigned long work_bytes, int *d_err, hipStream_t st)
{
    using namespace pivamd;
    if (N < 0 || ny <= 0 || nx <= 0 || !u || !out || !d_err || !d_work) return PIVAMD_ERR_ARG;
    if (neighbors < 1 || neighbors > kRbfMaxNbr || kernel < kRbfLinear || kernel > kRbfGauss) return PIVAMD_ERR_ARG;
    if (!(py > 0.0f) || !(px > 0.0f) || !(epsilon > 0.0) || !(smoothing >= 0.0)) return PIVAMD_ERR_ARG;
    if (work_bytes < pivamd_infill_rbf_workspace(N, ny, nx)) return PIVAMD_ERR_ARG;
    if (N == 0) return 0;
    int *nd = (int *)d_work, *cnt = nd + N, *holes = cnt + 1;
    const long long total = (long long)N * ny * nx;
    if (total > 0x7fffffffLL) return PIVAMD_ERR_ARG;
    hipLaunchKernelGGL(rbf_zero_kernel, dim3(1), dim3(1), 0, st, cnt);
    hipLaunchKernelGGL(knn_count_kernel, dim3(N), dim3(1024), 0, st, u, d_excl, N, ny * nx, nd);
    hipLaunchKernelGGL(rbf_prep_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, u, d_excl, N,
                       ny * nx, out, cnt, holes);
    const int nm = rbf_degree(kernel) == 0 ? 1 : rbf_degree(kernel) == 1 ? 3 : 6;
    const unsigned grid = (unsigned)(total < 2048 ? total : 2048);
    if (neighbors + nm <= 70)
        hipLaunchKernelGGL(rbf_fill_kernel<70>, dim3(grid), dim3(kRbfThreads), 0, st, u, d_excl, N, ny, nx,
                           (double)py, (double)px, neighbors, kernel, epsilon, smoothing, nd, cnt, holes, out, d_err);
    else
        hipLaunchKernelGGL(rbf_fill_kernel<kRbfMaxNbr + kRbfMaxMono>, dim3(grid), dim3(kRbfThreads), 0, st, u, d_excl,
                           N, ny, nx, (double)py, (double)px, neighbors, kernel, epsilon, smoothing, nd, cnt, holes,
                           out, d_err);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}

// ---------------------------------------------------------------------------
// Linear griddata infill (infilling.py:440-503 infill_griddata_linear: scipy
// griddata 'linear' over the valid vectors in physical coordinates, 'nearest'
// outside their convex hull, only the holes replaced).  Each hole (a NaN
// vector, not excluded) is filled from a Delaunay triangle of donors found
// around it, one wave per hole, for k = 24, 48, 96, 192 in turn:
//   - candidates: the k nearest donors (distance, then row-major index), in
//     coordinates y = row * sy, x = col * sx relative to the hole (sy, sx:
//     the pitches, divided by their common factor when both are integers, so
//     the predicates below are exact), kept in row-major order;
//   - the triangle: the first (a, b, c) in lexicographic order with nonzero
//     area, the hole inside or on an edge, no candidate strictly inside the
//     circumcircle and, unless every donor is a candidate, the circumcircle
//     strictly within the k-th candidate's distance of the hole (so it is
//     empty of every donor: a Delaunay triangle of the whole field);
//   - value: barycentric weights (orientation ratios) in FP64, as float32;
//   - no candidate triangle holds the hole: the nearest donor (outside the
//     convex hull, scipy's nearest fallback);
//   - triangles hold it but none is covered at k = 192: *err |= 2.
// Where four donors are cocircular (common on a regular grid) scipy's Qhull
// takes one of the equally Delaunay diagonals by its own processing order;
// this takes the first in index order.  Both interpolate linear fields
// exactly and agree wherever the Delaunay triangle is unique.
// *err bits: 1 no donor in an image with holes; 2 as above; 4 the candidate
// square outgrew its LDS buffer.
// ---------------------------------------------------------------------------
namespace pivamd {

constexpr int kGdCand = 1024;    // donors a candidate square may hold
constexpr int kGdMax = 192;      // largest k

#pragma clang fp contract(off)

__global__ void __launch_bounds__(64) griddata_fill_kernel(const float *u, const float *excl, int N, int ny, int nx,
                                                           double sy, double sx, const int *n_donors,
                                                           const int *hole_count, const int *holes, float *out,
                                                           int *err)
{
    __shared__ double c_d[kGdCand];
    __shared__ int c_q[kGdCand], c_r[kGdCand];
    __shared__ double s_x[kGdMax], s_y[kGdMax];
    __shared__ float s_f0[kGdMax], s_f1[kGdMax];
    __shared__ int s_cnt;
    __shared__ float s_near[2];
    const int lane = (int)threadIdx.x;
    const long long per = (long long)ny * nx, total = (long long)N * per;
    const int nh = *hole_count;
    const double step = sy < sx ? sy : sx;
    // (distance, row-major index) order of candidates o and e
    auto before = [&](int o, int e) { return c_d[o] < c_d[e] || (c_d[o] == c_d[e] && c_q[o] < c_q[e]); };
    for (int h = blockIdx.x; h < nh; h += gridDim.x) {
        const long long gi = holes[h];
        const int n = (int)(gi / per), w = (int)(gi - (long long)n * per);
        const float *f0 = u + (size_t)n * per, *f1 = u + (size_t)total + (size_t)n * per;
        const int nd = n_donors[n];
        if (nd < 1) {
            if (lane == 0) atomicOr(err, 1);
            continue;
        }
        const int r0 = w / nx, q0 = w % nx;
        bool done = false;
        float v0 = 0.0f, v1 = 0.0f;
        for (int kk0 = 24; kk0 <= kGdMax && !done; kk0 *= 2) {
            const int kk = kk0 < nd ? kk0 : nd;
            // -- donors of a square of half-width R, grown until it surely
            //    holds the kk nearest (nothing outside it is nearer than
            //    (R + 1) * min pitch); their (distance, index) ranks
            int R = 1;
            while ((2 * R + 1) * (2 * R + 1) < 2 * kk) ++R;
            int M = 0;
            double dk = 0.0;
            bool ok = false;
            for (;;) {
                const int ra = max(r0 - R, 0), rb = min(r0 + R, ny - 1);
                const int ca = max(q0 - R, 0), cb = min(q0 + R, nx - 1);
                const int wid = cb - ca + 1, cells = (rb - ra + 1) * wid;
                if (lane == 0) s_cnt = 0;
                __syncthreads();
                for (int e = lane; e < cells; e += 64) {
                    const int rr = ra + e / wid, cc = ca + e % wid;
                    const int q = rr * nx + cc;
                    const float a = f0[q], b = f1[q];
                    if (!(excl && excl[q] != 0.0f) && isfinite(a) && isfinite(b)) {
                        const int slot = atomicAdd(&s_cnt, 1);
                        if (slot < kGdCand) {
                            const double ey = (double)(rr - r0) * sy, ex = (double)(cc - q0) * sx;
                            c_d[slot] = sqrt(ey * ey + ex * ex);
                            c_q[slot] = q;
                        }
                    }
                }
                __syncthreads();
                M = s_cnt;
                if (M > kGdCand) break;
                const bool whole = ra == 0 && ca == 0 && rb == ny - 1 && cb == nx - 1;
                if (M >= kk) {
                    double best = -1.0;
                    for (int e = lane; e < M; e += 64) {
                        int rank = 0;
                        for (int o = 0; o < M; ++o) rank += before(o, e) ? 1 : 0;
                        c_r[e] = rank;
                        if (rank == kk - 1) best = c_d[e];
                    }
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) best = fmax(best, __shfl_xor(best, o, 64));
                    dk = best;
                    if (whole || dk < (double)(R + 1) * step) { ok = true; break; }
                } else if (whole) {
                    break;
                }
                R = R * 2;
                __syncthreads();
            }
            if (!ok) {
                if (lane == 0) atomicOr(err, 4);
                break;
            }
            __syncthreads();
            // -- the kk nearest in row-major order, and the nearest donor
            for (int e = lane; e < M; e += 64) {
                const int rank = c_r[e];
                if (rank >= kk) continue;
                int pos = 0;
                for (int o = 0; o < M; ++o) pos += (c_r[o] < kk && c_q[o] < c_q[e]) ? 1 : 0;
                const int q = c_q[e];
                s_y[pos] = (double)(q / nx - r0) * sy;
                s_x[pos] = (double)(q % nx - q0) * sx;
                s_f0[pos] = f0[q];
                s_f1[pos] = f1[q];
                if (rank == 0) {
                    s_near[0] = f0[q];
                    s_near[1] = f1[q];
                }
            }
            __syncthreads();
            v0 = s_near[0];
            v1 = s_near[1];
            const int m = kk;
            const bool need_cover = nd > kk;
            // -- the first (a, b, c) in lexicographic order: a uniform, b
            //    across the lanes, c in order; the wave keeps the smallest
            //    (b, c) of the first a that has one
            int found = -1, fa = -1;
            bool lane_uncov = false;
            for (int a = 0; a < m - 2 && fa < 0; ++a) {
                found = -1;
                for (int b = a + 1 + lane; b < m - 1 && found < 0; b += 64) {
                    for (int c = b + 1; c < m; ++c) {
                        const double ar = (s_x[b] - s_x[a]) * (s_y[c] - s_y[a]) - (s_y[b] - s_y[a]) * (s_x[c] - s_x[a]);
                        if (ar == 0.0) continue;
                        const double oa = s_x[b] * s_y[c] - s_y[b] * s_x[c];
                        const double ob = s_x[c] * s_y[a] - s_y[c] * s_x[a];
                        const double oc = s_x[a] * s_y[b] - s_y[a] * s_x[b];
                        const bool hold = ar > 0.0 ? (oa >= 0.0 && ob >= 0.0 && oc >= 0.0)
                                                   : (oa <= 0.0 && ob <= 0.0 && oc <= 0.0);
                        if (!hold) continue;
                        bool empty = true;
                        for (int p = 0; p < m && empty; ++p) {
                            const double adx = s_x[a] - s_x[p], ady = s_y[a] - s_y[p];
                            const double bdx = s_x[b] - s_x[p], bdy = s_y[b] - s_y[p];
                            const double cdx = s_x[c] - s_x[p], cdy = s_y[c] - s_y[p];
                            double det = (adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) +
                                         (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy) +
                                         (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady);
                            if (ar < 0.0) det = -det;
                            empty = !(det > 0.0);
                        }
                        if (!empty) continue;
                        if (need_cover) {
                            const double bx = s_x[b] - s_x[a], by = s_y[b] - s_y[a];
                            const double cx = s_x[c] - s_x[a], cy = s_y[c] - s_y[a];
                            const double dd = 2.0 * (bx * cy - by * cx);
                            const double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
                            const double ox = (cy * b2 - by * c2) / dd;
                            const double oy = (bx * c2 - cx * b2) / dd;
                            const double rho = sqrt(ox * ox + oy * oy);
                            const double gx = ox + s_x[a], gy = oy + s_y[a];
                            if (!(sqrt(gx * gx + gy * gy) + rho < dk)) {
                                lane_uncov = true;
                                continue;
                            }
                        }
                        found = b * m + c;
                        break;
                    }
                }
                int best = found < 0 ? 0x7fffffff : found;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) best = min(best, __shfl_xor(best, o, 64));
                if (best != 0x7fffffff) {
                    fa = a;
                    found = best;
                }
            }
            const bool uncovered = __any(lane_uncov);
            if (fa >= 0) {
                const int a = fa, b = found / m, c = found % m;
                const double ar = (s_x[b] - s_x[a]) * (s_y[c] - s_y[a]) - (s_y[b] - s_y[a]) * (s_x[c] - s_x[a]);
                const double oa = s_x[b] * s_y[c] - s_y[b] * s_x[c];
                const double ob = s_x[c] * s_y[a] - s_y[c] * s_x[a];
                const double oc = s_x[a] * s_y[b] - s_y[a] * s_x[b];
                const double la = oa / ar, lb = ob / ar, lc = oc / ar;
                v0 = (float)(la * (double)s_f0[a] + lb * (double)s_f0[b] + lc * (double)s_f0[c]);
                v1 = (float)(la * (double)s_f1[a] + lb * (double)s_f1[b] + lc * (double)s_f1[c]);
                done = true;
            } else if (!uncovered) {
                done = true;            // outside the hull: the nearest donor
            }
            __syncthreads();
        }
        if (lane == 0) {
            if (done) {
                out[(size_t)n * per + w] = v0;
                out[(size_t)total + (size_t)n * per + w] = v1;
            } else {
                atomicOr(err, 2);
            }
        }
        __syncthreads();
    }
}

#pragma clang fp contract(on)

}  // namespace pivamd

extern "C" unsigned long pivamd_infill_griddata_workspace(int N, int ny, int nx)
{
    return pivamd_infill_rbf_workspace(N, ny, nx);
}

extern "C" int pivamd_infill_griddata(const float *u, const float *d_excl, int N, int ny, int nx, float py, float px,
                                      float *out, void *d_work, unsigned long work_bytes, int *d_err, hipStream_t st)
{
    using namespace pivamd;
    if (N < 0 || ny <= 0 || nx <= 0 || !u || !out || !d_err || !d_work) return PIVAMD_ERR_ARG;
    if (!(py > 0.0f) || !(px > 0.0f)) return PIVAMD_ERR_ARG;
    if (work_bytes < pivamd_infill_griddata_workspace(N, ny, nx)) return PIVAMD_ERR_ARG;
    if (N == 0) return 0;
    int *nd = (int *)d_work, *cnt = nd + N, *holes = cnt + 1;
    const long long total = (long long)N * ny * nx;
    if (total > 0x7fffffffLL) return PIVAMD_ERR_ARG;
    // exact predicates for integer pitches: divide both by their common factor
    double sy = py, sx = px;
    if (py == (float)(long long)py && px == (float)(long long)px) {
        long long a = (long long)py, b = (long long)px;
        while (b) { const long long t = a % b; a = b; b = t; }
        sy = (double)py / (double)a;
        sx = (double)px / (double)a;
    }
    hipLaunchKernelGGL(rbf_zero_kernel, dim3(1), dim3(1), 0, st, cnt);
    hipLaunchKernelGGL(knn_count_kernel, dim3(N), dim3(1024), 0, st, u, d_excl, N, ny * nx, nd);
    hipLaunchKernelGGL(rbf_prep_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, u, d_excl, N,
                       ny * nx, out, cnt, holes);
    const unsigned grid = (unsigned)(total < 4096 ? total : 4096);
    hipLaunchKernelGGL(griddata_fill_kernel, dim3(grid), dim3(64), 0, st, u, d_excl, N, ny, nx, sy, sx, nd, cnt,
                       holes, out, d_err);
    return hipGetLastError() == hipSuccess ? 0 : PIVAMD_ERR_LAUNCH;
}
