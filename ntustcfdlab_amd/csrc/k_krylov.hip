// k_krylov.hip -- the Krylov pressure path of solver.f90 on the GPU:
// matmul_ell (the DIA product, solver.f90:146-190) and solver_BICG3
// (unpreconditioned BiCGSTAB, solver.f90:1-144) on the 7-diagonal matrix
// coef(n,7), jcoef(4) that solver_1st_member builds.
//
// The product adds each row's terms in the reference's order (diagonal,
// +1, -1, +nx, -nx, +nxny, -nxny) with no fused multiply-add, so it is
// bit-identical to the CPU restatement.  The BiCGSTAB scalars live on the
// device (BicgState): every kernel of an iteration reads them there and
// returns at once when the stopping rule of solver.f90:94 has fired, so the
// host queues iterations in batches and reads the state once per batch
// (the SOR solve's scheme).  Inner products are two-stage and
// deterministic: per-block partial sums in a fixed order, then one block
// adds the partials in a fixed order -- the same bits on every run, though
// not the serial sum's bits.
//
// Per iteration: p update, v = A p (+ r2.v), alpha, t = A s with s = r -
// alpha v formed on the fly (+ t.s, t.t), omega, x and r update (+ r.r,
// r2.r), norm/stop: 7 launches, about 250 B of HBM traffic per unknown.
#include "ntust_kernels.h"

namespace ntust {

constexpr int KRY_THREADS = 256;

struct DiaMatrix {
    const double *c[7];   // the columns of coef(n,7)
    long col[3];          // jcoef(2..4)
    int m;                // jcoef entries used (1..4)
    long n;
};

// row i of A y in the reference's order
__device__ __forceinline__ double dia_row(const DiaMatrix &A, const double *__restrict__ y,
                                          long i)
{
    double acc = A.c[0][i] * y[i];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (A.m < d + 2) break;
        const long col = A.col[d];
        if (i < A.n - col) acc = acc + A.c[2 * d + 1][i] * y[i + col];
        if (i >= col) acc = acc + A.c[2 * d + 2][i] * y[i - col];
    }
    return acc;
}

// Fixed-order block sum of up to 3 values per thread into part[k*nblk+b].
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double *part, int nblk)
{
    __shared__ double red[K][KRY_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = v[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) red[k][w] = s;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = red[k][0];
            for (int q = 1; q < KRY_THREADS / 64; q++) s += red[k][q];
            part[k * nblk + blockIdx.x] = s;
        }
    }
}

// Fixed-order total of part[k*nblk .. k*nblk+nblk) for k < K, by one block.
template <int K>
__device__ __forceinline__ void totals(const double *part, int nblk, double (&out)[K])
{
    __shared__ double red[K][KRY_THREADS];
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int b = tid; b < nblk; b += KRY_THREADS) s += part[k * nblk + b];
        red[k][tid] = s;
    }
    __syncthreads();
    for (int h = KRY_THREADS / 2; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int k = 0; k < K; k++) red[k][tid] += red[k][tid + h];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = red[k][0];
}

__global__ void __launch_bounds__(KRY_THREADS)
k_dia_matvec(DiaMatrix A, const double *__restrict__ y, double *__restrict__ x)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < A.n;
         i += (long)gridDim.x * blockDim.x)
        x[i] = dia_row(A, y, i);
}

// r = b - A x, r2 = r, p = v = 0; partials b.b, r.r
__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_setup(DiaMatrix A, const double *__restrict__ b, const double *__restrict__ x,
             double *__restrict__ r, double *__restrict__ r2, double *__restrict__ p,
             double *__restrict__ v, double *part)
{
    double s[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < A.n;
         i += (long)gridDim.x * blockDim.x) {
        const double bi = b[i], ri = bi - dia_row(A, x, i);
        r[i] = ri;
        r2[i] = ri;
        p[i] = 0.0;
        v[i] = 0.0;
        s[0] += bi * bi;
        s[1] += ri * ri;
    }
    block_sums<2>(s, part, gridDim.x);
}

__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_setup_done(BicgState *st, const double *part, int nblk)
{
    double t[2];
    totals<2>(part, nblk, t);
    if (threadIdx.x == 0) {
        st->norm0 = sqrt(t[0]);
        st->norm = sqrt(t[1]) / st->norm0;
        st->rho1 = 1.0;
        st->alpha = 1.0;
        st->omega = 1.0;
        st->rho2 = t[1];   // r2.r with r2 = r: the first iteration's rho2
        st->j = 0;
        st->stop = !(st->norm > st->eps && st->j < st->itmax);
    }
}

// p = r + beta (p - omega v), beta = (rho2/rho1) (alpha/omega)
__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_p(const BicgState *st, long n, const double *__restrict__ r, double *__restrict__ p,
         const double *__restrict__ v)
{
    if (load_flag(&st->stop)) return;
    const double beta = (st->rho2 / st->rho1) * (st->alpha / st->omega), om = st->omega;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = r[i] + beta * (p[i] - om * v[i]);
}

// v = A p; partial r2.v
__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_v(const BicgState *st, DiaMatrix A, const double *__restrict__ p,
         const double *__restrict__ r2, double *__restrict__ v, double *part)
{
    if (load_flag(&st->stop)) return;
    double s[1] = {0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < A.n;
         i += (long)gridDim.x * blockDim.x) {
        const double vi = dia_row(A, p, i);
        v[i] = vi;
        s[0] += r2[i] * vi;
    }
    block_sums<1>(s, part, gridDim.x);
}

__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_alpha(BicgState *st, const double *part, int nblk)
{
    if (load_flag(&st->stop)) return;
    double t[1];
    totals<1>(part, nblk, t);
    if (threadIdx.x == 0) st->alpha = st->rho2 / t[0];
}

// s = r - alpha v (own row and, recomputed alike, the neighbours'), t = A s;
// partials t.s, t.t
__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_t(const BicgState *st, DiaMatrix A, const double *__restrict__ r,
         const double *__restrict__ v, double *__restrict__ ss, double *__restrict__ t,
         double *part)
{
    if (load_flag(&st->stop)) return;
    const double al = st->alpha;
    double s[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < A.n;
         i += (long)gridDim.x * blockDim.x) {
        const double si = r[i] - al * v[i];
        double acc = A.c[0][i] * si;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (A.m < d + 2) break;
            const long col = A.col[d];
            if (i < A.n - col) acc = acc + A.c[2 * d + 1][i] * (r[i + col] - al * v[i + col]);
            if (i >= col) acc = acc + A.c[2 * d + 2][i] * (r[i - col] - al * v[i - col]);
        }
        ss[i] = si;
        t[i] = acc;
        s[0] += acc * si;
        s[1] += acc * acc;
    }
    block_sums<2>(s, part, gridDim.x);
}

__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_omega(BicgState *st, const double *part, int nblk)
{
    if (load_flag(&st->stop)) return;
    double t[2];
    totals<2>(part, nblk, t);
    if (threadIdx.x == 0) st->omega = t[0] / t[1];
}

// x = x + alpha p + omega s, r = s - omega t; partials r.r, r2.r
__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_x(const BicgState *st, long n, const double *__restrict__ p,
         const double *__restrict__ ss, const double *__restrict__ t,
         const double *__restrict__ r2, double *__restrict__ x, double *__restrict__ r,
         double *part)
{
    if (load_flag(&st->stop)) return;
    const double al = st->alpha, om = st->omega;
    double s[2] = {0.0, 0.0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double si = ss[i];
        x[i] = x[i] + al * p[i] + om * si;
        const double ri = si - om * t[i];
        r[i] = ri;
        s[0] += ri * ri;
        s[1] += r2[i] * ri;
    }
    block_sums<2>(s, part, gridDim.x);
}

// rho1 = rho2, norm = |r|/|b|, next rho2 = r2.r, j = j+1, the loop test
__global__ void __launch_bounds__(KRY_THREADS)
k_bicg_norm(BicgState *st, const double *part, int nblk)
{
    if (load_flag(&st->stop)) return;
    double t[2];
    totals<2>(part, nblk, t);
    if (threadIdx.x == 0) {
        st->rho1 = st->rho2;
        st->rho2 = t[1];
        st->norm = sqrt(t[0]) / st->norm0;
        st->j = st->j + 1;
        st->stop = !(st->norm > st->eps && st->j < st->itmax);
    }
}

static int kry_blocks(long n)
{
    const long b = (n + KRY_THREADS - 1) / KRY_THREADS;
    return (int)(b < KRY_MAX_BLOCKS ? b : KRY_MAX_BLOCKS);
}

static DiaMatrix dia(const double *coef, const long *col, int m, long n)
{
    DiaMatrix A;
    for (int c = 0; c < 7; c++) A.c[c] = coef + (long)c * n;
    for (int d = 0; d < 3; d++) A.col[d] = col[d];
    A.m = m;
    A.n = n;
    return A;
}

void launch_dia_matvec(const double *coef, const long *col, int m, long n, const double *y,
                       double *x, hipStream_t st)
{
    hipLaunchKernelGGL(k_dia_matvec, dim3(kry_blocks(n)), dim3(KRY_THREADS), 0, st,
                       dia(coef, col, m, n), y, x);
}

void launch_bicg_setup(const double *coef, const long *col, int m, long n, const double *b,
                       const double *x, BicgVectors &w, BicgState *st, hipStream_t s)
{
    const int nb = kry_blocks(n);
    hipLaunchKernelGGL(k_bicg_setup, dim3(nb), dim3(KRY_THREADS), 0, s, dia(coef, col, m, n),
                       b, x, w.r, w.r2, w.p, w.v, w.part);
    hipLaunchKernelGGL(k_bicg_setup_done, dim3(1), dim3(KRY_THREADS), 0, s, st, w.part, nb);
}

void launch_bicg_iteration(const double *coef, const long *col, int m, long n, double *x,
                           BicgVectors &w, BicgState *st, hipStream_t s)
{
    const int nb = kry_blocks(n);
    const DiaMatrix A = dia(coef, col, m, n);
    const dim3 g(nb), blk(KRY_THREADS), one(1);
    hipLaunchKernelGGL(k_bicg_p, g, blk, 0, s, st, n, w.r, w.p, w.v);
    hipLaunchKernelGGL(k_bicg_v, g, blk, 0, s, st, A, w.p, w.r2, w.v, w.part);
    hipLaunchKernelGGL(k_bicg_alpha, one, blk, 0, s, st, w.part, nb);
    hipLaunchKernelGGL(k_bicg_t, g, blk, 0, s, st, A, w.r, w.v, w.ss, w.t, w.part);
    hipLaunchKernelGGL(k_bicg_omega, one, blk, 0, s, st, w.part, nb);
    hipLaunchKernelGGL(k_bicg_x, g, blk, 0, s, st, n, w.p, w.ss, w.t, w.r2, x, w.r, w.part);
    hipLaunchKernelGGL(k_bicg_norm, one, blk, 0, s, st, w.part, nb);
}

}  // namespace ntust
