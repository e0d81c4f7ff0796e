// k_sor.hip -- pressure Poisson solve by SOR (gfx950).
//
// Restates GS.f90:1-106: point SOR on
//   sum_faces A_f (p_nb - p_P)/d_f = (1/dt) sum_faces A_f u*_f . n
// with the stopping rule "sweep while max|pNew - p| > zeta and ik < itmax"
// (GS.f90:11, the change measured before relaxation, :53) and the ghost
// values held fixed during the solve.
//
// Two orderings:
//  * lexicographic (exact): the reference's k, j, i order, pipelined over
//    sweeps on the colour-split planes below (k_sor_lex.hip).
//  * red-black: colour (i+j+k)&1, red first, on colour-split copies
//    of p and of the right-hand side b = mChange/dt, so each half-sweep
//    streams one colour plane of p and b and reads the other colour through
//    the caches.  Each half-sweep kernel decides by itself whether its sweep
//    runs, from the previous sweep's max|pNew - p| left in device memory, so
//    the host can queue many sweeps without a round trip and the sweep count
//    still stops exactly where GS.f90:11 stops.
#include "ntust_kernels.h"

namespace ntust {

__global__ void k_sor_init(double *slots, SorControl *ctl)
{
    const int b = threadIdx.x;
    if (b < SLOT_BUCKETS) {
        slots[slot_index(0, b)] = b == 0 ? 1.0 : 0.0;
        slots[slot_index(1, b)] = 0.0;
    }
    if (b == 0) {
        ctl->stop = 0;
        ctl->ik = 0;
    }
}

void launch_sor_init(double *slots, SorControl *ctl, hipStream_t st)
{
    hipLaunchKernelGGL(k_sor_init, dim3(1), dim3(64), 0, st, slots, ctl);
}

// ---------------------------------------------------------------------
// red-black SOR on colour-split storage
// ---------------------------------------------------------------------

// p (halo) -> colour planes, cells 0..n+1 on every axis
__global__ void k_rb_gather(KGrid g, SplitGrid s, const double *__restrict__ p,
                            double *__restrict__ p0, double *__restrict__ p1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > g.nx + 1 || j > g.ny + 1) return;
    const double val = p[hix(g, i, j, k)];
    if ((i + j + k) & 1) p1[six(s, i, j, k)] = val;
    else p0[six(s, i, j, k)] = val;
}

// b = mChange/dt on the interior, GS.f90:39-41 and the mChange/dt of :49
__global__ void k_rb_rhs(KGrid g, SplitGrid s, const double *__restrict__ us,
                         const double *__restrict__ vs, const double *__restrict__ ws,
                         double dt, double *__restrict__ b0, double *__restrict__ b1)
{
    int bx, by, bz;
    xcd_band(bx, by, bz);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    const int k = bz + 1;
    if (i > g.nx || j > g.ny) return;
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const long o = hix(g, i, j, k);
    const double mChange = (us[o] - us[o - 1]) * ay * az
                         + (vs[o] - vs[o - g.sx]) * ax * az
                         + (ws[o] - ws[o - g.sxy]) * ax * ay;
    const double b = mChange / dt;
    if ((i + j + k) & 1) b1[six(s, i, j, k)] = b;
    else b0[six(s, i, j, k)] = b;
}

// colour planes -> p interior and pre (GS.f90:86-94)
__global__ void k_rb_scatter(KGrid g, SplitGrid s, const double *__restrict__ p0,
                             const double *__restrict__ p1, double *__restrict__ p,
                             double *__restrict__ pre)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const double val = ((i + j + k) & 1) ? p1[six(s, i, j, k)] : p0[six(s, i, j, k)];
    p[hix(g, i, j, k)] = val;
    if (pre) pre[iix(g, i, j, k)] = val;
}

// One half-sweep of colour c, sweep t (1-based).  The buckets of sweep t-1
// hold the previous sweep's max|pNew-p| (sweep 0: 1.0, as GS.f90:9 starts);
// the kernel raises the buckets of sweep t.  Point update, the operator
// factored by axis, bit-identical to the CPU red-black restatement:
//   t1 = fma(rE,pE, rW*pW)  t2 = fma(rN,pN, rS*pS)  t3 = fma(rF,pF, rB*pB)
//   sum = fma(ayz, t1, fma(axz, t2, axy*t3))
//   aP  = fma(ayz, rE+rW, fma(axz, rN+rS, axy*(rF+rB)))
//   pNew = (sum-b)/aP;  p = fma(omega, pNew-p, p)
//
// The maximum of sweep t is kept in SLOT_BUCKETS buckets on separate cache
// lines (block b raises bucket b mod SLOT_BUCKETS, and only when its value is
// larger), because thousands of atomics on one address serialise; every
// wavefront of the next sweep reads the buckets back with one load.
//
// Each thread updates RB_KPT64 cells along k of one (q, j) column.  The
// cell pair of a thread alternates between i = 2q+1+sh and 2q+2-sh with the
// parity of k, so the x metrics are loaded once per thread for both.  The
// other colour's rows k0-1 .. k0+RB_KPT64 of the column are loaded once (the
// W and E entries): its (j, k+-1) neighbour of a cell is the W or E entry of
// the adjacent row, so the front and back values come from registers.  All
// loads of all the thread's cells are issued before any of them is used and
// before the gate value is waited on (no per-cell branch, indices clamped
// into the arrays), so a wavefront pays one memory latency; only the stores
// depend on the gate and on the cell being interior.  Offsets are 32-bit.
// Blocks run in the XCD-slab order.  At 256^3 this layout (8 cells, row
// reuse) measured 55.6 us per half-sweep against 62.0 us for 4 cells with
// separate front/back loads in the natural order; at 128^3 7.8 against 8.5.
// MODE 1, 2: the speculative forms -- no gate, no control block: the sweep
// runs and records its maximum (1) or only whether it exceeds zeta (2,
// flags[t] = 1: all the stopping rule asks of a sweep that is not the last
// of its batch); the host decides afterwards whether the reference's loop
// would have run it (and replays from a checkpoint if not).  MODE 0: the
// gated sweep.
template <int C, int MODE>
__global__ void __launch_bounds__(256)
k_rb_sweep(KGrid g, SplitGrid s, const double *__restrict__ rx,
           const double *__restrict__ ry, const double *__restrict__ rz,
           double *__restrict__ pc_own, const double *__restrict__ pc_oth,
           const double *__restrict__ b_own, double omega, double zeta, double itmax, int targ,
           double *slots, SorControl *ctl, int *flags)
{
    constexpr bool SPEC = MODE != 0;
    constexpr int KPT = RB_KPT64;
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
    // vW[r+1], vE[r+1]: other colour, row k0+r (r = -1 .. KPT; rows clamped
    // into the ghost planes 0 .. nz+1)
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
    double pcm = 0.0, nv[KPT];
    bool in[KPT], above = false;
#pragma unroll
    for (int r = 0; r < KPT; r++) {
        const int k = min(k0 + r, g.nz);
        const bool odd = r & 1;
        const double ax = odd ? axb : axa, rE = odd ? rEb : rEa, rW = odd ? rWb : rWa;
        const double az = g.iDz[k], rF = rz[k], rB = rz[k - 1];
        const double ayz = ay * az, axz = ax * az, axy = ax * ay;
        // (j, k+1) and (j, k-1) of the other colour sit at q1 + sh of
        // rows k0+r+1 and k0+r-1
        const bool sh = sh0 ^ (int)odd;
        const double pF = sh ? vE[r + 2] : vW[r + 2], pB = sh ? vE[r] : vW[r];
        const double t1 = fma(rE, vE[r + 1], rW * vW[r + 1]);
        const double t2 = fma(rN, vN[r], rS * vS[r]);
        const double t3 = fma(rF, pF, rB * pB);
        const double sum = fma(ayz, t1, fma(axz, t2, axy * t3));
        const double aP = fma(ayz, rE + rW, fma(axz, rN + rS, axy * (rF + rB)));
        const double d = (sum - vb[r]) / aP - vo[r];
        const int i = 2 * q + 1 + (sh0 ^ (int)odd);
        in[r] = jr <= g.ny && k0 + r <= g.nz && i <= g.nx;
        if (MODE == 2) {
            // max > zeta exactly when some change is > zeta (NaN is neither)
            above = above || (in[r] && fabs(d) > zeta);
        } else {
            const double pc = in[r] ? fabs(d) : 0.0;
            if (pc > pcm) pcm = pc;   // a NaN change never raises the max (GS.f90:61)
        }
        nv[r] = fma(omega, d, vo[r]);
    }
    if (!SPEC) {
        // every wavefront of the block reads the same buckets: the decision
        // is uniform over the block, so returning here cannot split a barrier
        gate = wave_max(gate);
        const bool active = !stopped && ((double)(t - 1) < itmax) && (gate > zeta);
        if (!active) {
            if (blin == 0 && tid == 0) ctl->stop = 1;
            return;
        }
        if (blin == 0 && C == 0) clear_next_buckets(slots, t, tid);
    }
#pragma unroll
    for (int r = 0; r < KPT; r++)
        if (in[r]) pc_own[at[r]] = nv[r];
    if (MODE == 2) {
        // every lane reaches this point; one plain store per wavefront
        if (__any(above) && lane == 0) flags[t] = 1;
        return;
    }
    pcm = wave_max(pcm);
    if ((tid & 63) == 0) red[tid >> 6] = pcm;
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
        if (!SPEC && C == 1 && blin == 0) ctl->ik = t;
    }
}

// ---------------------------------------------------------------------
// members of the Krylov pressure path (solver_1st_member.f90,
// solver_2nd_member.f90, solver_pressure_field.f90)
// ---------------------------------------------------------------------

// The 7-diagonal matrix coef(ndim,7) of solver_1st_member.f90:18-69 (row ik
// of cell (i,j,k) in x-fastest order; 1/a/b evaluated as (1/a)/b), boundary
// rows as the reference edits them: the neighbour coefficient outside the
// domain is zeroed, and on the east face its value is also taken off the
// diagonal.
__global__ void __launch_bounds__(256)
k_poisson_coeffs(KGrid g, double *__restrict__ coef)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long nd = (long)g.nx * g.ny * g.nz;
    const long ik = iix(g, i, j, k);
    double c1 = 1 / g.iDx[i] / g.Dxs[i] + 1 / g.iDx[i] / g.Dxs[i - 1] + 1 / g.iDy[j] / g.Dys[j] +
                1 / g.iDy[j] / g.Dys[j - 1] + 1 / g.iDz[k] / g.Dzs[k] + 1 / g.iDz[k] / g.Dzs[k - 1];
    double c2 = -1 / g.iDx[i] / g.Dxs[i], c3 = -1 / g.iDx[i] / g.Dxs[i - 1];
    double c4 = -1 / g.iDy[j] / g.Dys[j], c5 = -1 / g.iDy[j] / g.Dys[j - 1];
    double c6 = -1 / g.iDz[k] / g.Dzs[k], c7 = -1 / g.iDz[k] / g.Dzs[k - 1];
    double aa = 0.0, bb = 0.0, cc = 0.0;
    if (i == 1) {
        c3 = 0.0;
    } else if (i == g.nx) {
        aa = -1 / g.iDx[i] / g.Dxs[i];
        c2 = 0.0;
    }
    if (j == 1) c5 = 0.0;
    else if (j == g.ny) c4 = 0.0;
    if (k == 1) c7 = 0.0;
    else if (k == g.nz) c6 = 0.0;
    c1 = c1 + aa + bb + cc;
    coef[ik] = c1;
    coef[ik + nd] = c2;
    coef[ik + 2 * nd] = c3;
    coef[ik + 3 * nd] = c4;
    coef[ik + 4 * nd] = c5;
    coef[ik + 5 * nd] = c6;
    coef[ik + 6 * nd] = c7;
}

// b = mChange/dt on the interior, interior layout (GS.f90:39-41,49; the
// right-hand side the red-black solver gathers)
__global__ void __launch_bounds__(256)
k_poisson_rhs(KGrid g, const double *__restrict__ us, const double *__restrict__ vs,
              const double *__restrict__ ws, double dt, double *__restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const double ax = g.iDx[i], ay = g.iDy[j], az = g.iDz[k];
    const long o = hix(g, i, j, k);
    const double mChange = (us[o] - us[o - 1]) * ay * az
                         + (vs[o] - vs[o - g.sx]) * ax * az
                         + (ws[o] - ws[o - g.sxy]) * ax * ay;
    b[iix(g, i, j, k)] = mChange / dt;
}

// solver_pressure_field.f90:8-17: pre = x, p = pre on the interior
__global__ void __launch_bounds__(256)
k_pressure_field(KGrid g, const double *__restrict__ x, double *__restrict__ p,
                 double *__restrict__ pre)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const long ik = iix(g, i, j, k);
    const double v = x[ik];
    pre[ik] = v;
    p[hix(g, i, j, k)] = v;
}

void launch_poisson_coeffs(const KGrid &g, double *coef, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_poisson_coeffs, grd, blk, 0, st, g, coef);
}

void launch_poisson_rhs(const KGrid &g, const double *us, const double *vs, const double *ws,
                        double dt, double *b, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_poisson_rhs, grd, blk, 0, st, g, us, vs, ws, dt, b);
}

void launch_pressure_field(const KGrid &g, const double *x, double *p, double *pre,
                           hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_pressure_field, grd, blk, 0, st, g, x, p, pre);
}

void launch_rb_gather(const KGrid &g, const SplitGrid &s, const double *p, double *p0,
                      double *p1, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 2 + 63) / 64, (g.ny + 2 + 3) / 4, g.nz + 2);
    hipLaunchKernelGGL(k_rb_gather, grd, blk, 0, st, g, s, p, p0, p1);
}

void launch_rb_rhs(const KGrid &g, const SplitGrid &s, const double *us, const double *vs,
                   const double *ws, double dt, double *b0, double *b1, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_rb_rhs, grd, blk, 0, st, g, s, us, vs, ws, dt, b0, b1);
}

void launch_rb_scatter(const KGrid &g, const SplitGrid &s, const double *p0, const double *p1,
                       double *p, double *pre, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_rb_scatter, grd, blk, 0, st, g, s, p0, p1, p, pre);
}

void launch_rb_sweep(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                     const double *rz, double *p0, double *p1, const double *b0,
                     const double *b1, double omega, double zeta, double itmax, int t,
                     double *slots, SorControl *ctl, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT64 - 1) / RB_KPT64);
    hipLaunchKernelGGL((k_rb_sweep<0, 0>), grd, blk, 0, st, g, s, rx, ry, rz, p0, p1, b0,
                       omega, zeta, itmax, t, slots, ctl, nullptr);
    hipLaunchKernelGGL((k_rb_sweep<1, 0>), grd, blk, 0, st, g, s, rx, ry, rz, p1, p0, b1,
                       omega, zeta, itmax, t, slots, ctl, nullptr);
}

void launch_rb_sweep_spec(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                          const double *rz, double *p0, double *p1, const double *b0,
                          const double *b1, double omega, int t, double *slots, double zeta,
                          int *flags, hipStream_t st)
{
    const int nq = (g.nx + 1) / 2;
    dim3 blk(64, 4), grd((nq + 63) / 64, (g.ny + 3) / 4, (g.nz + RB_KPT64 - 1) / RB_KPT64);
    if (flags) {
        hipLaunchKernelGGL((k_rb_sweep<0, 2>), grd, blk, 0, st, g, s, rx, ry, rz, p0, p1, b0,
                           omega, zeta, 0.0, t, slots, nullptr, flags);
        hipLaunchKernelGGL((k_rb_sweep<1, 2>), grd, blk, 0, st, g, s, rx, ry, rz, p1, p0, b1,
                           omega, zeta, 0.0, t, slots, nullptr, flags);
    } else {
        hipLaunchKernelGGL((k_rb_sweep<0, 1>), grd, blk, 0, st, g, s, rx, ry, rz, p0, p1, b0,
                           omega, 0.0, 0.0, t, slots, nullptr, nullptr);
        hipLaunchKernelGGL((k_rb_sweep<1, 1>), grd, blk, 0, st, g, s, rx, ry, rz, p1, p0, b1,
                           omega, 0.0, 0.0, t, slots, nullptr, nullptr);
    }
}

// zero the buckets of sweeps t0 .. t0+n-1
__global__ void __launch_bounds__(256)
k_slots_clear(double *slots, int t0, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n * SLOT_BUCKETS) slots[slot_index(t0 + q / SLOT_BUCKETS, q % SLOT_BUCKETS)] = 0.0;
}

void launch_slots_clear(double *slots, int t0, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_slots_clear, dim3((n * SLOT_BUCKETS + 255) / 256), dim3(256), 0, st,
                       slots, t0, n);
}

}  // namespace ntust
