// k_sor_mixed.hip -- red-black SOR with FP32 sweeps inside FP64 defect
// correction (gfx950).
//
// The red-black SOR step is affine in p: writing p = p0 + e, one sweep on p
// from p0 + e equals p0 plus one sweep on the correction equation
//     aP e = sum_nb a_nb e_nb + r0,   r0 = sum_nb a_nb p0_nb - b - aP p0
// from e.  So the sweeps run in FP32 on e (half the bytes of an FP64 sweep),
// and every m sweeps the FP64 state absorbs the correction, p0 += e, and the
// FP64 residual r0 is recomputed with the FP64 operator of the plain path.
// In exact arithmetic the iterates are those of the FP64 red-black solver;
// the FP32 rounding of e is bounded by one segment of m sweeps, because the
// residual and the state stay FP64.  The stopping rule is GS.f90:11 applied
// to the change max|eNew - e| = max|pNew - p| of each sweep, with the same
// per-sweep gating as the FP64 kernels.
//
// Traffic: an FP32 half-sweep streams one colour of e and r and reads the
// other colour of e: 8 B per cell of the grid (16 per sweep, FP64: 32); a
// correction reads p, e and writes p, e (24 B/cell) and the residual reads
// p, b and writes r (20 B/cell).
#include "ntust_kernels.h"

namespace ntust {

template <int C>
__global__ void __launch_bounds__(256)
k_rb_sweep32(KGrid g, SplitGrid s, FMetric fm, float *__restrict__ e_own,
             const float *__restrict__ e_oth, const float *__restrict__ r_own, float omega,
             double zeta, double itmax, int t, double *slots, SorControl *ctl, int *dirty)
{
    __shared__ double red[4];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = tid & 63;
    double gate = lane < SLOT_BUCKETS ? slots[slot_index(t - 1, lane)] : 0.0;
    const int stopped = ctl->stop;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const long pl = (long)s.mp * s.ry;
    float pcm = 0.0f;
    float nv[RB_KPT];
    long at[RB_KPT];
    bool in[RB_KPT];
#pragma unroll
    for (int r = 0; r < RB_KPT; r++) {
        const int k = blockIdx.z * RB_KPT + r + 1;
        const int sh = (C + 1 + j + k) & 1;
        const int i = 2 * q + 1 + sh;
        in[r] = i <= g.nx && j <= g.ny && k <= g.nz;
        at[r] = 0;
        nv[r] = 0.0f;
        if (in[r]) {
            const int m = q + sh;
            const long row = (long)s.mp * ((long)j + (long)s.ry * (long)k);
            const long o = row + m;
            const float ax = fm.iDx[i], ay = fm.iDy[j], az = fm.iDz[k];
            const float ayz = ay * az, axz = ax * az, axy = ax * ay;
            const float aE = ayz * fm.rx[i], aW = ayz * fm.rx[i - 1];
            const float aN = axz * fm.ry[j], aS = axz * fm.ry[j - 1];
            const float aF = axy * fm.rz[k], aB = axy * fm.rz[k - 1];
            const float sum = fmaf(aE, e_oth[row + m + 1 - sh],
                              fmaf(aW, e_oth[row + m - sh],
                              fmaf(aN, e_oth[o + s.mp],
                              fmaf(aS, e_oth[o - s.mp],
                              fmaf(aF, e_oth[o + pl], aB * e_oth[o - pl])))));
            const float aP = ((((aE + aW) + aN) + aS) + aF) + aB;
            const float eNew = (sum + r_own[o]) / aP;
            const float eold = e_own[o];
            const float d = eNew - eold;
            const float pc = fabsf(d);
            if (pc > pcm) pcm = pc;
            nv[r] = fmaf(omega, d, eold);
            at[r] = o;
        }
    }
    gate = wave_max(gate);
    const bool active = !stopped && ((double)(t - 1) < itmax) && (gate > zeta);
    if (!active) {
        if (blin == 0 && tid == 0) ctl->stop = 1;
        return;
    }
#pragma unroll
    for (int r = 0; r < RB_KPT; r++)
        if (in[r]) e_own[at[r]] = nv[r];
    double pm = wave_max((double)pcm);
    if ((tid & 63) == 0) red[tid >> 6] = pm;
    __syncthreads();
    if (tid == 0) {
        double mm = red[0];
        for (int w = 1; w < (int)((blockDim.x * blockDim.y) >> 6); w++) mm = red[w] > mm ? red[w] : mm;
        double *bk = &slots[slot_index(t, blin % SLOT_BUCKETS)];
        if (mm > 0.0 && mm > *(volatile double *)bk)
            atomicMax((unsigned long long *)bk, (unsigned long long)__double_as_longlong(mm));
        if (blin == 0) {
            *dirty = 1;
            if (C == 1) ctl->ik = t;
        }
    }
}

// p0 += e on the interior, e = 0; skipped when no sweep of the segment ran
__global__ void __launch_bounds__(256)
k_rb_absorb(KGrid g, SplitGrid s, double *__restrict__ p0, double *__restrict__ p1,
            float *__restrict__ e0, float *__restrict__ e1, const int *dirty)
{
    if (!*dirty) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long o = six(s, i, j, k);
    if ((i + j + k) & 1) {
        p1[o] = p1[o] + (double)e1[o];
        e1[o] = 0.0f;
    } else {
        p0[o] = p0[o] + (double)e0[o];
        e0[o] = 0.0f;
    }
}

// r = sum_nb a_nb p_nb - b - aP p with the FP64 coefficients and summation
// of the FP64 red-black kernel, rounded to FP32
__global__ void __launch_bounds__(256)
k_rb_residual(KGrid g, SplitGrid s, const double *__restrict__ rx,
              const double *__restrict__ ry, const double *__restrict__ rz,
              const double *__restrict__ p0, const double *__restrict__ p1,
              const double *__restrict__ b0, const double *__restrict__ b1,
              float *__restrict__ r0, float *__restrict__ r1, const int *dirty)
{
    if (dirty && !*dirty) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const int c = (i + j + k) & 1;
    const double *own = c ? p1 : p0, *oth = c ? p0 : p1;
    const long o = six(s, i, j, k);
    const long pl = (long)s.mp * s.ry;
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const double ayz = ay * az, axz = ax * az, axy = ax * ay;
    const double aE = ayz * rx[i], aW = ayz * rx[i - 1];
    const double aN = axz * ry[j], aS = axz * ry[j - 1];
    const double aF = axy * rz[k], aB = axy * rz[k - 1];
    const double sum = fma(aE, oth[six(s, i + 1, j, k)], fma(aW, oth[six(s, i - 1, j, k)],
                       fma(aN, oth[o + s.mp], fma(aS, oth[o - s.mp],
                       fma(aF, oth[o + pl], aB * oth[o - pl])))));
    const double aP = ((((aE + aW) + aN) + aS) + aF) + aB;
    const double r = (sum - (c ? b1 : b0)[o]) - aP * own[o];
    (c ? r1 : r0)[o] = (float)r;
}

void launch_rb_sweep32(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                       float *e1, const float *r0, const float *r1, float omega, double zeta,
                       double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                       hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    hipLaunchKernelGGL(k_rb_sweep32<0>, grd, blk, 0, st, g, s, fm, e0, e1, r0, omega, zeta,
                       itmax, t, slots, ctl, dirty);
    hipLaunchKernelGGL(k_rb_sweep32<1>, grd, blk, 0, st, g, s, fm, e1, e0, r1, omega, zeta,
                       itmax, t, slots, ctl, dirty);
}

void launch_rb_absorb(const KGrid &g, const SplitGrid &s, double *p0, double *p1, float *e0,
                      float *e1, const int *dirty, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_rb_absorb, grd, blk, 0, st, g, s, p0, p1, e0, e1, dirty);
}

void launch_rb_residual(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                        const double *rz, const double *p0, const double *p1, const double *b0,
                        const double *b1, float *r0, float *r1, const int *dirty,
                        hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_rb_residual, grd, blk, 0, st, g, s, rx, ry, rz, p0, p1, b0, b1, r0, r1,
                       dirty);
}

}  // namespace ntust
