// k_sor_lex.hip -- GS.f90's lexicographic point SOR, pipelined over sweeps
// on the colour-split planes of the red-black solver (gfx950).
//
// GS.f90:35-68 sweeps k, j, i in place: cell (i,j,k) of sweep t reads the
// sweep-t values of its lower neighbours (i-1, j-1, k-1) and the sweep t-1
// values of its upper ones and of itself.  With s = i+j+k, the lower
// neighbours lie on hyperplane s-1 and the upper ones on s+1.  Update
// hyperplane s of sweep t at step tau = s + 2t:
//   - plane s-1 of sweep t was updated at tau-1 (the lower neighbours),
//   - plane s+1 of sweep t-1 at tau-1 (the upper neighbours),
//   - plane s of sweep t-1 at tau-2 (the cell itself),
// and nothing touches those values in between.  So every read finds the
// value the reference's loop reads, and the results are its bits.  The
// cells of step tau have s = tau mod 2 -- one colour of the red-black
// colouring, no two of them neighbours -- and belong to sweeps
// t = (tau - s)/2: up to (nx+ny+nz)/2 sweeps advance in one launch.  A batch
// of sweeps t0..t1 takes the launches tau = 3+2*t0 .. nx+ny+nz+2*t1 and
// leaves every cell whose sweep lies outside the batch as it is.
//
// The point update is GS.f90:39-51 term by term (the right-hand side
// mChange/dt from the same expression, k_rb_rhs); the stopping rule of
// GS.f90:11 needs, for every sweep but the last, only whether some change
// exceeds zeta, and for the last one its maximum.
#include "ntust_kernels.h"

namespace ntust {

// One step tau of a batch t0..t1 (colour C = tau & 1).  Thread layout of
// k_rb_sweep's FP64 form: RB_KPT64 cells along k of one (q, j) column, the
// other colour's rows k0-1 .. k0+RB_KPT64 shared for the +-k neighbours,
// XCD-slab block order.
//
// The sweep index of a cell is t = tw - l - d, with tw uniform over the
// wavefront (one j row), l its lane and d = 0..4 set by the cell's place
// along k.  A cell of sweep t < t1 whose change exceeds zeta stamps
// flags[t] = epoch (the batch's stamp; a stale copy of a flag is an older
// stamp, so a plain load can only cause a redundant store); the wavefront
// first ORs its cells per t with five ballots.  MAXREC: the cells of sweep
// t1 raise its buckets with their maximum change (a NaN change raises
// nothing, GS.f90:61; the stamp is not set for it either).
template <int C, bool MAXREC>
__global__ void __launch_bounds__(256)
k_lex_sweep(KGrid g, SplitGrid s, double *__restrict__ pc_own,
            const double *__restrict__ pc_oth, const double *__restrict__ b_own, double omega,
            double zeta, int tau, int t0, int t1, int *flags, int epoch, double *slots)
{
    constexpr int KPT = RB_KPT64;
    __shared__ double red[4];
    int bx, by, bz;
    xcd_block(bx, by, bz);
    const int qb = bx * (int)blockDim.x;
    const int jb = by * (int)blockDim.y + 1;
    const int k0 = bz * KPT + 1;
    {
        // the hyperplanes of this step's cells are tau - 2t, t0 <= t <= t1;
        // a block with none of them exits (the test is uniform over it)
        const int smin = 2 * qb + 1 + jb + k0;
        const int smax = 2 * (qb + (int)blockDim.x - 1) + 2 + jb + (int)blockDim.y - 1 + k0 +
                         KPT - 1;
        if (smax < tau - 2 * t1 || smin > tau - 2 * t0) return;
    }
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = tid & 63;
    const int q = qb + threadIdx.x;
    const int jr = jb + threadIdx.y;
    const int j = min(jr, g.ny);
    const int sh0 = (C + 1 + j + k0) & 1;
    const int ia = min(2 * q + 1 + sh0, g.nx + 1), ib = min(2 * q + 2 - sh0, g.nx + 1);
    const double axa = g.iDx[ia], dEa = g.Dxs[ia], dWa = g.Dxs[ia - 1];
    const double axb = g.iDx[ib], dEb = g.Dxs[ib], dWb = g.Dxs[ib - 1];
    const double ay = g.iDy[j], dN = g.Dys[j], dS = g.Dys[j - 1];
    const unsigned mp = (unsigned)s.mp;
    const unsigned q1 = (unsigned)min(q, (g.nx + 1) / 2);
    double vW[KPT + 2], vE[KPT + 2], vN[KPT], vS[KPT], vb[KPT], vo[KPT];
    unsigned at[KPT];
#pragma unroll
    for (int r = -1; r <= KPT; r++) {
        const int k = min(k0 + r, g.nz + 1);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        vW[r + 1] = pc_oth[row + q1];
        vE[r + 1] = pc_oth[row + q1 + 1];
    }
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const unsigned row = mp * ((unsigned)j + (unsigned)s.ry * (unsigned)k);
        const unsigned o = row + q1 + (unsigned)(sh0 ^ (r & 1));
        at[r] = o;
        vN[r] = pc_oth[o + mp];
        vS[r] = pc_oth[o - mp];
        vb[r] = b_own[o];
        vo[r] = pc_own[o];
    }
    // sweep of cell 0 of this thread; cell r is in sweep tc - d(r)
    const int tc = (tau - (2 * q + 1 + sh0 + j + k0)) / 2;
    double pcm = 0.0, nv[KPT];
    bool in[KPT];
    unsigned amask = 0;   // bit d: a cell of sweep tc - d changed by more than zeta
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const bool odd = r & 1;
        const int sh = sh0 ^ (int)odd;
        const int d = (sh + r - sh0) / 2;
        const int t = tc - d;
        const double ax = odd ? axb : axa, dE = odd ? dEb : dEa, dW = odd ? dWb : dWa;
        const double az = g.iDz[k], dF = g.Dzs[k], dB = g.Dzs[k - 1];
        const double pF = sh ? vE[r + 2] : vW[r + 2], pB = sh ? vE[r] : vW[r];
        // GS.f90:43-51, in its order of operations
        const double pNew = (-vE[r + 1] * ay * az / dE - vW[r + 1] * ay * az / dW
                             - vN[r] * ax * az / dN - vS[r] * ax * az / dS
                             - pF * ax * ay / dF - pB * ax * ay / dB + vb[r])
                          / (-ay * az / dE - ay * az / dW - ax * az / dN - ax * az / dS
                             - ax * ay / dF - ax * ay / dB);
        const double pc = fabs(pNew - vo[r]);
        nv[r] = vo[r] + (omega * (pNew - vo[r]));
        const int i = 2 * q + 1 + sh;
        in[r] = jr <= g.ny && k0 + r <= g.nz && i <= g.nx && t >= t0 && t <= t1;
        if (in[r] && t < t1 && pc > zeta) amask |= 1u << d;
        if (MAXREC && in[r] && t == t1 && pc > pcm) pcm = pc;
    }
#pragma unroll
    for (int r = 0; r < KPT; r++)
        if (in[r]) pc_own[at[r]] = nv[r];
    // per-sweep stamps: bit m of (hi:lo) is sweep tw - m, tw the sweep of
    // cell 0 of lane 0
    unsigned long long lo = 0, hi = 0;
#pragma unroll
    for (int d = 0; d < 5; d++) {
        const unsigned long long b = __ballot((amask >> d) & 1u);
        lo |= b << d;
        if (d > 0) hi |= b >> (64 - d);
    }
    const int tw = tc + (int)threadIdx.x;   // tc of lane 0 (blockDim.x is 64)
    if ((lo >> lane) & 1ull) {
        int *f = &flags[tw - lane];
        if (*f != epoch) *f = epoch;
    }
    if (lane < 4 && ((hi >> lane) & 1ull)) {
        int *f = &flags[tw - 64 - lane];
        if (*f != epoch) *f = epoch;
    }
    if (!MAXREC) return;
    pcm = wave_max(pcm);
    if (lane == 0) red[tid >> 6] = pcm;
    __syncthreads();
    if (tid == 0) {
        double mm = red[0];
        for (int w = 1; w < (int)((blockDim.x * blockDim.y) >> 6); w++)
            mm = red[w] > mm ? red[w] : mm;
        if (mm > 0.0)
            atomicMax((unsigned long long *)&slots[slot_index(t1, blin % SLOT_BUCKETS)],
                      (unsigned long long)__double_as_longlong(mm));
    }
}

void launch_lex_sweeps(const KGrid &g, const SplitGrid &s, double *p0, double *p1,
                       const double *b0, const double *b1, double omega, double zeta, int t0,
                       int t1, int *flags, int epoch, double *slots, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT64 - 1) / RB_KPT64);
    const int smax = g.nx + g.ny + g.nz;
    for (int tau = 3 + 2 * t0; tau <= smax + 2 * t1; tau++) {
        // sweep t1 has cells in the steps tau >= 3 + 2*t1
        const bool mx = tau >= 3 + 2 * t1;
        if (tau & 1) {
            if (mx)
                hipLaunchKernelGGL((k_lex_sweep<1, true>), grd, blk, 0, st, g, s, p1, p0, b1,
                                   omega, zeta, tau, t0, t1, flags, epoch, slots);
            else
                hipLaunchKernelGGL((k_lex_sweep<1, false>), grd, blk, 0, st, g, s, p1, p0, b1,
                                   omega, zeta, tau, t0, t1, flags, epoch, slots);
        } else {
            if (mx)
                hipLaunchKernelGGL((k_lex_sweep<0, true>), grd, blk, 0, st, g, s, p0, p1, b0,
                                   omega, zeta, tau, t0, t1, flags, epoch, slots);
            else
                hipLaunchKernelGGL((k_lex_sweep<0, false>), grd, blk, 0, st, g, s, p0, p1, b0,
                                   omega, zeta, tau, t0, t1, flags, epoch, slots);
        }
    }
}

long lex_launches(const KGrid &g, long t0, long t1)
{
    return (long)(g.nx + g.ny + g.nz) + 2 * t1 - 2 * t0 - 2;
}

}  // namespace ntust
