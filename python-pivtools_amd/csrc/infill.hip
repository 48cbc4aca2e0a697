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
