// k_sor_mixed.hip -- red-black SOR with FP32 sweeps inside FP64 defect
// correction (gfx950).
//
// The red-black SOR step is affine in p: writing p = p0 + e, one sweep on p
// from p0 + e equals p0 plus one sweep on the correction equation
//     aP e = sum_nb a_nb e_nb + r0,   r0 = sum_nb a_nb p0_nb - b - aP p0
// from e.  So the sweeps run in FP32 on e (half the bytes of an FP64 sweep),
// and every m sweeps (m = 256) the FP64 state absorbs the correction,
// p0 += e, and the FP64 residual r0 is recomputed with the FP64 operator of
// the plain path.
// In exact arithmetic the iterates are those of the FP64 red-black solver;
// the FP32 rounding of e is bounded by one segment of m sweeps, because the
// residual and the state stay FP64.  The stopping rule is GS.f90:11 applied
// to the change max|eNew - e| = max|pNew - p| of each sweep, with the same
// per-sweep gating as the FP64 kernels.
//
// Colour 1 runs shifted: with s1 = r1/aP1 its correction is carried as
// e1' = e1 - s1, which turns its update into the homogeneous one,
//   e1' <- (1-omega) e1' + omega S(e0)/aP1,
// and moves the constant into the colour-0 residual, r0' = r0 + S(s1)
// (k_rb_correct; the segment starts from e1' = -s1, e1 = 0).
// The colour-1 half-sweeps then read no residual.  The change of a sweep,
// e1'New - e1' = e1New - e1, is the same.
//
// Traffic: an FP32 half-sweep streams one colour of e (and of r for colour
// 0) and reads the other colour of e: 8 + 6 B per cell of the grid per
// sweep (FP64: 32); the correction of a segment (absorb and residual,
// k_rb_correct) reads p, e, r1, b and writes p, e, r: 26 + 24 B/cell over
// its two colour launches.
#include "ntust_kernels.h"

namespace ntust {

// MODE 0: the gated sweep.  MODE 1, 2: the speculative forms (no gate, no
// control block, no dirty flag), as for the FP64 sweep; 1 records the
// sweep's maximum change in the buckets, 2 only whether it exceeds zeta
// (flags[t] = 1), which is all the stopping rule asks of a sweep that is
// not the last of its batch.
template <int C, int MODE>
__global__ void __launch_bounds__(256)
k_rb_sweep32(KGrid g, SplitGrid s, FMetric fm, float *__restrict__ e_own,
             const float *__restrict__ e_oth, const float *__restrict__ r_own, float omega,
             double zeta, double itmax, int targ, double *slots, SorControl *ctl, int *dirty,
             int first, int *flags)
{
    constexpr bool SPEC = MODE != 0;
    __shared__ double red[4];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = tid & 63;
    const int t = targ;
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
        vr[r] = C == 0 ? r_own[o] : 0.0f;   // colour 1: shifted, no right-hand side
        vo[r] = e_own[o];
    }
    float pcm = 0.0f, nv[RB_KPT];
    bool in[RB_KPT], above = false;
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
        if (MODE == 2) {
            // max > zeta exactly when some change is > zeta (NaN is neither)
            above = above || (in[r] && (double)fabsf(d) > zeta);
        } else {
            const float pc = in[r] ? fabsf(d) : 0.0f;
            if (pc > pcm) pcm = pc;
        }
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
        if (blin == 0 && C == 0) clear_next_buckets(slots, t, tid);
    }
#pragma unroll
    for (int r = 0; r < RB_KPT; r++)
        if (in[r]) e_own[at[r]] = nv[r];
    if (MODE == 2) {
        // every lane reaches this point; one plain store per wavefront
        if (__any(above) && lane == 0) flags[t] = 1;
        return;
    }
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

// The FP64 correction of a segment in colour-split form, two launches
// (colour 1, then colour 0): each cell absorbs its correction, p = p + e,
// and gets its residual
//   r = (sum_nb a_nb p_nb - b) - aP p
// from the absorbed values of its neighbours, which are all of the other
// colour.  During the colour-1 launch the colour-0 cells are not yet
// absorbed, so their value is formed on the fly as p + e (the same
// operation the absorb performs, so the same bits); during the colour-0
// launch the colour-1 cells are already absorbed and read as they are.
// The result is bit for bit that of an absorb over the grid followed by the
// residual.  ABSORB = false: the residual alone (p untouched).
//
// Colour 1 is shifted: its correction is e1' + s1 with s1 the scaled
// residual the r1 plane holds; its launch leaves s1 = r1/aP1 in that plane
// and e1' = -s1.  Colour 0 adds S(s1) to its
// residual, read as -e1' from the e1 plane, and leaves e0 = 0.
//
// Thread layout of the sweeps (RB_KPT cells along k, other colour's W/E
// rows shared for the +-k neighbours, XCD-slab block order).  Traffic per
// cell of the grid: 26 B (colour 1) + 24 B (colour 0).
template <int C, bool ABSORB>
__global__ void __launch_bounds__(256)
k_rb_correct(KGrid g, SplitGrid s, const double *__restrict__ rx, const double *__restrict__ ry,
             const double *__restrict__ rz, double *__restrict__ p_own,
             const double *__restrict__ p_oth, float *__restrict__ e_own,
             const float *__restrict__ e_oth, const double *__restrict__ b_own,
             float *__restrict__ r_own, const int *dirty)
{
    constexpr int KPT = RB_KPT;
    // FLY: the other colour is formed as p + e; SSUM: the residual gains
    // S(s1) = -S(e1'); LOADE: either reads the other colour's e; SOWN: the
    // own cell is shifted
    constexpr bool FLY = ABSORB && C == 1;
    constexpr bool SSUM = C == 0;
    constexpr bool LOADE = FLY || SSUM;
    constexpr bool SOWN = C == 1;
    if (dirty && !load_flag(dirty)) return;
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int q = bx * blockDim.x + threadIdx.x;
    const int jr = by * blockDim.y + threadIdx.y + 1;
    const int j = min(jr, g.ny);
    const int k0 = bz * KPT + 1;
    const int sh0 = (C + 1 + j + k0) & 1;
    const int ia = min(2 * q + 1 + sh0, g.nx + 1), ib = min(2 * q + 2 - sh0, g.nx + 1);
    const double axa = g.iDx[ia], rEa = rx[ia], rWa = rx[ia - 1];
    const double axb = g.iDx[ib], rEb = rx[ib], rWb = rx[ib - 1];
    const double ay = g.iDy[j], rN = ry[j], rS = ry[j - 1];
    const unsigned mp = (unsigned)s.mp;
    const unsigned q1 = (unsigned)min(q, (g.nx + 1) / 2);
    // Other-colour values: during the first launch they are formed as
    // p + e, but only for interior cells -- the absorb touches interior
    // cells only, and a ghost entry (whose e is zero) is read as it is,
    // since p + 0 would turn a -0 into +0.  Colour c of pair x in row
    // (jj, kk) is cell i = 2x + (((jj+kk)&1)^c).  All loads are issued
    // first and unconditionally; the interior selection follows, so the
    // e loads are not serialised behind per-load branches.
    auto interior = [&](int x, int jj, int kk) -> bool {
        const int i = 2 * x + (((jj + kk) & 1) ^ (C ^ 1));
        return jj >= 1 && jj <= g.ny && kk >= 1 && kk <= g.nz && i >= 1 && i <= g.nx;
    };
    double vW[KPT + 2], vE[KPT + 2], vN[KPT], vS[KPT], vb[KPT], vp[KPT];
    float eW[KPT + 2], eE[KPT + 2], eN[KPT], eS[KPT], ve[KPT], vs[KPT];
    unsigned at[KPT];
#pragma unroll
    for (int r = -1; r <= KPT; r++) {
        const int k = min(k0 + r, g.nz + 1);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        vW[r + 1] = p_oth[row + q1];
        vE[r + 1] = p_oth[row + q1 + 1];
        if (LOADE) {
            eW[r + 1] = e_oth[row + q1];
            eE[r + 1] = e_oth[row + q1 + 1];
        }
    }
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        const unsigned o = row + q1 + (unsigned)(sh0 ^ (r & 1));
        at[r] = o;
        vN[r] = p_oth[o + mp];
        vS[r] = p_oth[o - mp];
        if (LOADE) {
            eN[r] = e_oth[o + mp];
            eS[r] = e_oth[o - mp];
        }
        vb[r] = b_own[o];
        vp[r] = p_own[o];
        ve[r] = ABSORB ? e_own[o] : 0.0f;
        vs[r] = ABSORB && SOWN ? r_own[o] : 0.0f;
    }
    if (FLY) {
#pragma unroll
        for (int r = -1; r <= KPT; r++) {
            const int k = min(k0 + r, g.nz + 1);
            const double w = vW[r + 1] + (double)eW[r + 1], e = vE[r + 1] + (double)eE[r + 1];
            vW[r + 1] = interior((int)q1, j, k) ? w : vW[r + 1];
            vE[r + 1] = interior((int)q1 + 1, j, k) ? e : vE[r + 1];
        }
#pragma unroll
        for (int r = 0; r < KPT; r++) {
            const int k = min(k0 + r, g.nz);
            const int x = (int)q1 + (sh0 ^ (r & 1));
            const double n = vN[r] + (double)eN[r], sv = vS[r] + (double)eS[r];
            vN[r] = interior(x, j + 1, k) ? n : vN[r];
            vS[r] = interior(x, j - 1, k) ? sv : vS[r];
        }
    }
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const bool odd = r & 1;
        const bool sh = sh0 ^ (int)odd;
        const double ax = odd ? axb : axa, rE = odd ? rEb : rEa, rW = odd ? rWb : rWa;
        const double az = g.iDz[k], rF = rz[k], rB = rz[k - 1];
        const double ayz = ay * az, axz = ax * az, axy = ax * ay;
        const double pF = sh ? vE[r + 2] : vW[r + 2], pB = sh ? vE[r] : vW[r];
        const double own = !ABSORB ? vp[r]
                           : SOWN  ? vp[r] + ((double)ve[r] + (double)vs[r])
                                   : vp[r] + (double)ve[r];
        const double t1 = fma(rE, vE[r + 1], rW * vW[r + 1]);
        const double t2 = fma(rN, vN[r], rS * vS[r]);
        const double t3 = fma(rF, pF, rB * pB);
        const double sum = fma(ayz, t1, fma(axz, t2, axy * t3));
        const double aP = fma(ayz, rE + rW, fma(axz, rN + rS, axy * (rF + rB)));
        double res = (sum - vb[r]) - aP * own;
        if (SSUM) {
            // + S(s1), s1 = -e1' (zero on the ghosts)
            const float fF = sh ? eE[r + 2] : eW[r + 2], fB = sh ? eE[r] : eW[r];
            const double u1 = fma(rE, (double)eE[r + 1], rW * (double)eW[r + 1]);
            const double u2 = fma(rN, (double)eN[r], rS * (double)eS[r]);
            const double u3 = fma(rF, (double)fF, rB * (double)fB);
            res -= fma(ayz, u1, fma(axz, u2, axy * u3));
        }
        const int i = 2 * q + 1 + (int)sh;
        if (jr <= g.ny && k0 + r <= g.nz && i <= g.nx) {
            if (ABSORB) p_own[at[r]] = own;
            if (SOWN) {
                const float sv = (float)(res / aP);
                r_own[at[r]] = sv;
                e_own[at[r]] = -sv;
            } else {
                if (ABSORB) e_own[at[r]] = 0.0f;
                r_own[at[r]] = (float)res;
            }
        }
    }
}

void launch_rb_sweep32(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                       float *e1, const float *r0, const float *r1, float omega, double zeta,
                       double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                       int first, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    hipLaunchKernelGGL((k_rb_sweep32<0, 0>), grd, blk, 0, st, g, s, fm, e0, e1, r0, omega, zeta,
                       itmax, t, slots, ctl, dirty, first, nullptr);
    hipLaunchKernelGGL((k_rb_sweep32<1, 0>), grd, blk, 0, st, g, s, fm, e1, e0, r1, omega, zeta,
                       itmax, t, slots, ctl, dirty, 0, nullptr);
}

void launch_rb_sweep32_spec(const KGrid &g, const SplitGrid &s, const FMetric &fm, float *e0,
                            float *e1, const float *r0, const float *r1, float omega, int t,
                            double *slots, double zeta, int *flags, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    if (flags) {
        hipLaunchKernelGGL((k_rb_sweep32<0, 2>), grd, blk, 0, st, g, s, fm, e0, e1, r0, omega,
                           zeta, 0.0, t, slots, nullptr, nullptr, 0, flags);
        hipLaunchKernelGGL((k_rb_sweep32<1, 2>), grd, blk, 0, st, g, s, fm, e1, e0, r1, omega,
                           zeta, 0.0, t, slots, nullptr, nullptr, 0, flags);
    } else {
        hipLaunchKernelGGL((k_rb_sweep32<0, 1>), grd, blk, 0, st, g, s, fm, e0, e1, r0, omega,
                           0.0, 0.0, t, slots, nullptr, nullptr, 0, nullptr);
        hipLaunchKernelGGL((k_rb_sweep32<1, 1>), grd, blk, 0, st, g, s, fm, e1, e0, r1, omega,
                           0.0, 0.0, t, slots, nullptr, nullptr, 0, nullptr);
    }
}

void launch_rb_residual(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                        const double *rz, const double *p0, const double *p1, const double *b0,
                        const double *b1, float *r0, float *r1, float *e1_shift,
                        const int *dirty, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    // p is only read when ABSORB is false
    double *q0 = const_cast<double *>(p0), *q1 = const_cast<double *>(p1);
    hipLaunchKernelGGL((k_rb_correct<1, false>), grd, blk, 0, st, g, s, rx, ry, rz, q1, p0,
                       e1_shift, nullptr, b1, r1, dirty);
    hipLaunchKernelGGL((k_rb_correct<0, false>), grd, blk, 0, st, g, s, rx, ry, rz, q0, p1,
                       nullptr, e1_shift, b0, r0, dirty);
}

// absorb + residual of a segment of the shifted sweeps (k_rb_correct,
// colour 1 then colour 0)
void launch_rb_correct(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                       const double *rz, double *p0, double *p1, float *e0, float *e1,
                       const double *b0, const double *b1, float *r0, float *r1,
                       const int *dirty, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT - 1) / RB_KPT);
    hipLaunchKernelGGL((k_rb_correct<1, true>), grd, blk, 0, st, g, s, rx, ry, rz, p1, p0,
                       e1, e0, b1, r1, dirty);
    hipLaunchKernelGGL((k_rb_correct<0, true>), grd, blk, 0, st, g, s, rx, ry, rz, p0, p1,
                       e0, e1, b0, r0, dirty);
}

}  // namespace ntust
