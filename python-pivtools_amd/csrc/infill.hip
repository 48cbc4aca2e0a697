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
