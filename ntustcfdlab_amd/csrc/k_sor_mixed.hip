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

// SPEC: the speculative form (no gate, no control block, no dirty flag),
// as for the FP64 sweep
template <int C, bool SPEC>
__global__ void __launch_bounds__(256)
k_rb_sweep32(KGrid g, SplitGrid s, FMetric fm, float *__restrict__ e_own,
             const float *__restrict__ e_oth, const float *__restrict__ r_own, float omega,
             double zeta, double itmax, int targ, double *slots, SorControl *ctl, int *dirty,
             int first)
{
    __shared__ double red[4];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = tid & 63;
    const int t = SPEC ? targ : sweep_index<C>(targ, ctl);
    double gate = 0.0;
    int stopped = 0;
    if (!SPEC) {
        gate = lane < SLOT_BUCKETS ? load_flag(&slots[slot_index(t - 1, lane)]) : 0.0;
        stopped = load_flag(&ctl->stop);
    }

    // same thread layout and load-first structure as the FP64 sweep
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int q = bx * blockDim.x + threadIdx.x;
    const int jr = by * blockDim.y + threadIdx.y + 1;
    const int j = min(jr, g.ny);
    const int k0 = bz * RB_KPT + 1;
    const int sh0 = (C + 1 + j + k0) & 1;
    const int ia = min(2 * q + 1 + sh0, g.nx + 1), ib = min(2 * q + 2 - sh0, g.nx + 1);
    const float axa = fm.iDx[ia], rEa = fm.rx[ia], rWa = fm.rx[ia - 1];
    const float axb = fm.iDx[ib], rEb = fm.rx[ib], rWb = fm.rx[ib - 1];
    const float ay = fm.iDy[j], rN = fm.ry[j], rS = fm.ry[j - 1];
    const unsigned mp = (unsigned)s.mp, pl = (unsigned)s.mp * (unsigned)s.ry;
    const unsigned q1 = (unsigned)min(q, (g.nx + 1) / 2);
    float vW[RB_KPT], vE[RB_KPT], vN[RB_KPT], vS[RB_KPT], vF[RB_KPT], vB[RB_KPT], vr[RB_KPT],
        vo[RB_KPT];
    unsigned at[RB_KPT];
#pragma unroll
    for (int r = 0; r < RB_KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        const unsigned o = row + q1 + (unsigned)(sh0 ^ (r & 1));
        at[r] = o;
        vW[r] = e_oth[row + q1];
        vE[r] = e_oth[row + q1 + 1];
        vN[r] = e_oth[o + mp];
        vS[r] = e_oth[o - mp];
        vF[r] = e_oth[o + pl];
        vB[r] = e_oth[o - pl];
        vr[r] = r_own[o];
        vo[r] = e_own[o];
    }
    float pcm = 0.0f, nv[RB_KPT];
    bool in[RB_KPT];
#pragma unroll
    for (int r = 0; r < RB_KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const bool odd = r & 1;
        const float ax = odd ? axb : axa, rE = odd ? rEb : rEa, rW = odd ? rWb : rWa;
        const float az = fm.iDz[k], rF = fm.rz[k], rB = fm.rz[k - 1];
        const float ayz = ay * az, axz = ax * az, axy = ax * ay;
        const float t1 = fmaf(rE, vE[r], rW * vW[r]);
        const float t2 = fmaf(rN, vN[r], rS * vS[r]);
        const float t3 = fmaf(rF, vF[r], rB * vB[r]);
        const float sum = fmaf(ayz, t1, fmaf(axz, t2, axy * t3));
        const float aP = fmaf(ayz, rE + rW, fmaf(axz, rN + rS, axy * (rF + rB)));
        const float d = (sum + vr[r]) / aP - vo[r];
        const int i = 2 * q + 1 + (sh0 ^ (int)odd);
        in[r] = jr <= g.ny && k0 + r <= g.nz && i <= g.nx;
        const float pc = in[r] ? fabsf(d) : 0.0f;
        if (pc > pcm) pcm = pc;
        nv[r] = fmaf(omega, d, vo[r]);
    }
    if (!SPEC) {
        gate = wave_max(gate);
        const bool active = !stopped && ((double)(t - 1) < itmax) && (gate > zeta);
        if (!active) {
            if (blin == 0 && tid == 0) {
                ctl->stop = 1;
                if (first) *dirty = 0;   // no sweep of this segment runs
            }
            return;
        }
        if (blin == 0) {
            if (C == 0) clear_next_buckets(slots, t, tid);
            if (tid == 0) pass_sweep_index<C>(t, ctl);
        }
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
        // a posted atomic (no return, no uncached pre-read of the bucket,
        // which held every block open for a memory round trip: 12 % of an
        // FP32 launch at 256^3)
        if (mm > 0.0)
            atomicMax((unsigned long long *)bk, (unsigned long long)__double_as_longlong(mm));
        if (!SPEC && blin == 0) {
            *dirty = 1;
            if (C == 1) ctl->ik = t;
        }
    }
}

// p0 += e on the interior, e = 0; skipped when no sweep of the segment ran
// (dirty NULL: always)
__global__ void __launch_bounds__(256)
k_rb_absorb(KGrid g, SplitGrid s, double *__restrict__ p0, double *__restrict__ p1,
            float *__restrict__ e0, float *__restrict__ e1, const int *dirty)
{
    if (dirty && !load_flag(dirty)) return;
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
    if (dirty && !load_flag(dirty)) return;
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
    const double rE = rx[i], rW = rx[i - 1], rN = ry[j], rS = ry[j - 1], rF = rz[k], rB = rz[k - 1];
    const double t1 = fma(rE, oth[six(s, i + 1, j, k)], rW * oth[six(s, i - 1, j, k)]);
    const double t2 = fma(rN, oth[o + s.mp], rS * oth[o - s.mp]);
    const double t3 = fma(rF, oth[o + pl], rB * oth[o - pl]);
    const double sum = fma(ayz, t1, fma(axz, t2, axy * t3));
    const double aP = fma(ayz, rE + rW, fma(axz, rN + rS, axy * (rF + rB)));
    const double r = (sum - (c ? b1 : b0)[o]) - aP * own[o];
    (c ? r1 : r0)[o] = (float)r;
}

void launch_rb_sweep32(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                       float *e1, const float *r0, const float *r1, float omega, double zeta,
                       double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                       int first, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    hipLaunchKernelGGL((k_rb_sweep32<0, false>), grd, blk, 0, st, g, s, fm, e0, e1, r0, omega, zeta,
                       itmax, t, slots, ctl, dirty, first);
    hipLaunchKernelGGL((k_rb_sweep32<1, false>), grd, blk, 0, st, g, s, fm, e1, e0, r1, omega, zeta,
                       itmax, t, slots, ctl, dirty, 0);
}

void launch_rb_sweep32_spec(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                            float *e1, const float *r0, const float *r1, float omega, int t,
                            double *slots, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    hipLaunchKernelGGL((k_rb_sweep32<0, true>), grd, blk, 0, st, g, s, fm, e0, e1, r0, omega, 0.0,
                       0.0, t, slots, nullptr, nullptr, 0);
    hipLaunchKernelGGL((k_rb_sweep32<1, true>), grd, blk, 0, st, g, s, fm, e1, e0, r1, omega, 0.0,
                       0.0, t, slots, nullptr, nullptr, 0);
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
