// k_sor_tile.hip -- one red-black SOR sweep per launch on 3-D tiles (gfx950).
//
// For grids whose half-sweeps are launch- and latency-bound (the working set
// in L2 and the Infinity Cache), two launches per sweep cost more than the
// data.  Here one launch does the whole sweep with no plane loop: a
// workgroup owns a TX x TY x TZ block of cells and
//   1. loads the old black values of the block plus two cells on every side
//      into LDS, and the old red values and right-hand sides of the block
//      plus one cell into registers (every load issued before any is used);
//   2. updates every red cell of the block plus one cell (the ones outside
//      the block belong to neighbouring workgroups, which compute the same
//      values from the same inputs) and keeps them in LDS;
//   3. updates the black cells of the block from those red values.
// Neighbours read old values while owners write new ones, so a sweep reads
// one copy of the colour planes and writes the other (sweep t reads copy
// (t-1)&1 and writes copy t&1), as the z-streaming sweep of k_sor_fused.hip
// does, with the same host-side handling.  The point update, the gate on the
// previous sweep's maximum (GS.f90:11) and the per-sweep maxima in buckets
// are those of the two-pass kernels: the iterates, the stopping sweep and
// the reported change are bit-identical to them.
//
// Traffic per cell and sweep: both colours and the right-hand side read
// once plus the halo ((TX+4)(TY+4)(TZ+4)/(TX TY TZ) of the black plane,
// (TX+2)(TY+2)(TZ+2)/(TX TY TZ) of the red plane and its right-hand side),
// both colours written once.
#include <hip/hip_runtime.h>

#include "ntust_kernels.h"

namespace ntust {

template <typename T>
struct TileMetric {
    const T *iDx, *iDy, *iDz, *rx, *ry, *rz;   // offset by +1
};

__device__ __forceinline__ double tfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float tfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double tabs(double a) { return fabs(a); }
__device__ __forceinline__ float tabs(float a) { return fabsf(a); }

// the point update in the operation order of k_rb_sweep / k_rb_sweep32:
// MIXED = false: d = (sum - b)/aP - p; MIXED = true: d = (sum + r)/aP - e
template <typename T, bool MIXED>
__device__ __forceinline__ T tile_update(T pW, T pE, T pS, T pN, T pB, T pF, T po, T ax, T ay,
                                         T az, T rE, T rW, T rN, T rS, T rF, T rB, T rhs,
                                         T omega, T &change)
{
    const T ayz = ay * az, axz = ax * az, axy = ax * ay;
    const T t1 = tfma(rE, pE, rW * pW);
    const T t2 = tfma(rN, pN, rS * pS);
    const T t3 = tfma(rF, pF, rB * pB);
    const T sum = tfma(ayz, t1, tfma(axz, t2, axy * t3));
    const T aP = tfma(ayz, rE + rW, tfma(axz, rN + rS, axy * (rF + rB)));
    const T d = (MIXED ? (sum + rhs) : (sum - rhs)) / aP - po;
    change = tabs(d);
    return tfma(omega, d, po);
}

// Cell (i,j,k) with i in pair q = i>>1 of row (j,k): the red cell of pair q
// is i = 2q + ((j+k)&1), the black one i = 2q + (((j+k)&1)^1).
//
// LDS regions (pairs x rows x planes):
//   black: pairs q0-1 .. q0+TX/2+1, rows j0-2 .. j0+TY+1, planes k0-2 .. k0+TZ+1
//   red:   pairs q0   .. q0+TX/2,   rows j0-1 .. j0+TY,   planes k0-1 .. k0+TZ
// with i0 = 1 + TX bx (odd, so q0 = i0>>1 and the block's cells are
// i0 .. i0+TX-1 = 2q0+1 .. 2q0+TX).
template <typename T, bool MIXED, int TX, int TY, int TZ, int NT>
__global__ void __launch_bounds__(NT)
k_rb_sweep_tile(KGrid g, SplitGrid s, TileMetric<T> m, const T *__restrict__ in0,
                const T *__restrict__ in1, T *__restrict__ out0, T *__restrict__ out1,
                const T *__restrict__ rhs0, const T *__restrict__ rhs1, T omega, double zeta,
                double itmax, int t, double *slots, SorControl *ctl, int *dirty)
{
    constexpr int BP = TX / 2 + 3, BY = TY + 4, BZ = TZ + 4, BPL = BP * BY;
    constexpr int RP = TX / 2 + 1, RY = TY + 2, RZ = TZ + 2, RPL = RP * RY;
    static_assert(BPL <= NT, "one black column per thread");
    constexpr int MX = TX + 4, MY = TY + 4, MZ = TZ + 4;  // metrics of i0-2 .. i0+TX+1 etc.
    __shared__ T blk[BPL * BZ];
    __shared__ T rn[RPL * RZ];
    __shared__ T mDz[MZ], mRz[MZ];
    __shared__ double redm[NT / 64];

    const int tid = threadIdx.x, lane = tid & 63;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    double gate = lane < SLOT_BUCKETS ? load_flag(&slots[slot_index(t - 1, lane)]) : 0.0;
    const int stopped = load_flag(&ctl->stop);

    const int i0 = 1 + (int)blockIdx.x * TX, j0 = 1 + (int)blockIdx.y * TY;
    const int k0 = 1 + (int)blockIdx.z * TZ, q0 = i0 >> 1;
    const unsigned mp = (unsigned)s.mp, pl = (unsigned)s.mp * (unsigned)s.ry;
    const int qm = s.mp - 1;
    auto kof = [&](int k) { return pl * (unsigned)min(max(k, 0), g.nz + 1); };

    // One thread per pair column of a region; its offsets are computed once
    // and the planes are walked.  Offsets are clamped into the stored cells:
    // an entry outside them is never read by an update, and loads stay
    // unconditional (a load under a branch is waited for at the join).
    //   black region column (bp, by) = (tid mod BP, tid div BP), tid < BP*BY
    //   red region column   (rp, ry) = (tid mod RP, tid div RP), tid < RP*RY
    //   black cells column  (rp, ry) with ry < TY: rows j0 .. j0+TY-1
    const int bp = tid % BP, byy = tid / BP;
    const int rp = tid % RP, ryy = tid / RP;
    const bool bcol = tid < BPL, rcol = tid < RPL, ocol = tid < RP * TY;
    const unsigned bo = (unsigned)min(max(q0 - 1 + bp, 0), qm) +
                        mp * (unsigned)min(max(j0 - 2 + byy, 0), g.ny + 1);
    const unsigned ro = (unsigned)min(max(q0 + rp, 0), qm) +
                        mp * (unsigned)min(max(j0 - 1 + ryy, 0), g.ny + 1);
    const unsigned oo = (unsigned)min(max(q0 + rp, 0), qm) +
                        mp * (unsigned)min(max(j0 + ryy, 0), g.ny + 1);

    // ---- 1. loads ---------------------------------------------------------
    T vb[BZ], vro[RZ], vrr[RZ], vbr[TZ];
#pragma unroll
    for (int kk = 0; kk < BZ; kk++) vb[kk] = in1[bo + kof(k0 - 2 + kk)];
#pragma unroll
    for (int kk = 0; kk < RZ; kk++) {
        const unsigned o = ro + kof(k0 - 1 + kk);
        vro[kk] = in0[o];
        vrr[kk] = rhs0[o];
    }
#pragma unroll
    for (int kk = 0; kk < TZ; kk++) vbr[kk] = MIXED ? (T)0 : rhs1[oo + kof(k0 + kk)];
    // the pair's x metrics (cells a = 2q and a+1), the row's y metrics
    const int q = q0 + rp;
    const int a = 2 * q, ac = min(max(a, 0), g.nx + 1);
    const T dA = m.iDx[ac], dB = m.iDx[ac + 1], rM = m.rx[ac - 1], rA = m.rx[ac],
            rB = m.rx[ac + 1];
    const int jr = j0 - 1 + ryy, jrc = min(max(jr, 0), g.ny + 1);
    const T ayr = m.iDy[jrc], rNr = m.ry[jrc], rSr = m.ry[jrc - 1];
    for (int e = tid; e < MZ; e += NT) {
        const int k = min(max(k0 - 2 + e, -1), g.nz + 2);
        mDz[e] = m.iDz[k];
        mRz[e] = m.rz[k];
    }
    if (bcol) {
#pragma unroll
        for (int kk = 0; kk < BZ; kk++) blk[tid + BPL * kk] = vb[kk];
    }
    __syncthreads();

    // ---- 2. red cells of the block plus one cell -----------------------------
    // red cell of pair q in row j, plane k: i = a + ((j+k)&1); its black
    // neighbours: pair (i-1)>>1 = q-1+par and (i+1)>>1 = q+par of the row,
    // pair q of rows j-1, j+1 and planes k-1, k+1
    T pcm = (T)0;
    if (rcol) {
        const int bcol0 = (rp + 1) + BP * (ryy + 1);   // black LDS entry of pair q, row jr
#pragma unroll
        for (int kk = 0; kk < RZ; kk++) {
            const int k = k0 - 1 + kk;
            const int par = (jr + k) & 1;
            const int i = a + par;
            T nv = vro[kk];   // a ghost red cell keeps its value
            if (i >= 1 && i <= g.nx && jr >= 1 && jr <= g.ny && k >= 1 && k <= g.nz) {
                const int c = bcol0 + BPL * (kk + 1);
                const T pW = blk[c - 1 + par], pE = blk[c + par];
                const T pS = blk[c - BP], pN = blk[c + BP];
                const T pB = blk[c - BPL], pF = blk[c + BPL];
                T ch;
                nv = tile_update<T, MIXED>(pW, pE, pS, pN, pB, pF, vro[kk], par ? dB : dA, ayr,
                                           mDz[kk + 1], par ? rB : rA, par ? rA : rM, rNr, rSr,
                                           mRz[kk + 1], mRz[kk], vrr[kk], omega, ch);
                const bool owned = i >= i0 && i < i0 + TX && jr >= j0 && jr < j0 + TY &&
                                   k >= k0 && k < k0 + TZ;
                // a NaN change never raises the max (GS.f90:61)
                if (owned && ch > pcm) pcm = ch;
            }
            rn[tid + RPL * kk] = nv;
        }
    }
    __syncthreads();

    // every wavefront reads the same buckets: the decision is uniform over
    // the block, so returning here cannot split a barrier
    gate = wave_max(gate);
    const bool active = !stopped && ((double)(t - 1) < itmax) && (gate > zeta);
    if (!active) {
        // the segment's dirty flag stays as earlier sweeps left it
        if (blin == 0 && tid == 0) ctl->stop = 1;
        return;
    }
    if (blin == 0) clear_next_buckets(slots, t, tid);

    // ---- 3. store the owned red cells, update the black cells -----------------
    if (rcol && jr >= j0 && jr < j0 + TY && jr <= g.ny) {
#pragma unroll
        for (int kk = 1; kk <= TZ; kk++) {
            const int k = k0 - 1 + kk;
            const int i = a + ((jr + k) & 1);
            if (i >= i0 && i < i0 + TX && i <= g.nx && k <= g.nz)
                out0[ro + pl * (unsigned)k] = rn[tid + RPL * kk];
        }
    }
    if (ocol) {
        const int j = j0 + ryy, jc = min(j, g.ny + 1);
        const T ay = m.iDy[jc], rN = m.ry[jc], rS = m.ry[jc - 1];
        // red LDS entry of pair q, row j (red rows start at j0-1)
        const int rc0 = rp + RP * (ryy + 1);
        const int bc0 = (rp + 1) + BP * (ryy + 2);   // black LDS entry of pair q, row j
#pragma unroll
        for (int kk = 0; kk < TZ; kk++) {
            const int k = k0 + kk;
            const int par = ((j + k) & 1) ^ 1;
            const int i = a + par;
            if (i < i0 || i >= i0 + TX || i > g.nx || j > g.ny || k > g.nz) continue;
            // red neighbours: pair (i-1)>>1 = q-1+par and (i+1)>>1 = q+par of
            // the row, pair q of rows j-1, j+1 and planes k-1, k+1
            const int c = rc0 + RPL * (kk + 1);
            const T pW = rn[c - 1 + par], pE = rn[c + par];
            const T pS = rn[c - RP], pN = rn[c + RP];
            const T pB = rn[c - RPL], pF = rn[c + RPL];
            const T po = blk[bc0 + BPL * (kk + 2)];
            T ch;
            const T nv = tile_update<T, MIXED>(pW, pE, pS, pN, pB, pF, po, par ? dB : dA, ay,
                                               mDz[kk + 2], par ? rB : rA, par ? rA : rM, rN,
                                               rS, mRz[kk + 2], mRz[kk + 1], vbr[kk], omega, ch);
            if (ch > pcm) pcm = ch;
            out1[oo + pl * (unsigned)k] = nv;
        }
    }

    // ---- the sweep's maximum change --------------------------------------------
    double pm = wave_max((double)pcm);
    if (lane == 0) redm[tid >> 6] = pm;
    __syncthreads();
    if (tid == 0) {
        double mm = redm[0];
        for (int w = 1; w < NT / 64; w++) mm = redm[w] > mm ? redm[w] : mm;
        double *bk = &slots[slot_index(t, blin % SLOT_BUCKETS)];
        if (mm > 0.0)
            atomicMax((unsigned long long *)bk, (unsigned long long)__double_as_longlong(mm));
        if (blin == 0) {
            ctl->ik = t;
            if (MIXED) *dirty = 1;
        }
    }
}

constexpr int T32X = 64, T32Y = 8, T32Z = 8, T32N = 448;   // FP32: 33 KB of LDS
constexpr int T64X = 64, T64Y = 8, T64Z = 4, T64N = 448;   // FP64: 43 KB of LDS

void launch_rb_sweep_tile(const KGrid &g, const SplitGrid &s, const double *rx, const double *ry,
                          const double *rz, const double *in0, const double *in1, double *out0,
                          double *out1, const double *b0, const double *b1, double omega,
                          double zeta, double itmax, int t, double *slots, SorControl *ctl,
                          hipStream_t st)
{
    TileMetric<double> m{g.iDx, g.iDy, g.iDz, rx, ry, rz};
    const dim3 grd((g.nx + T64X - 1) / T64X, (g.ny + T64Y - 1) / T64Y,
                   (g.nz + T64Z - 1) / T64Z);
    hipLaunchKernelGGL((k_rb_sweep_tile<double, false, T64X, T64Y, T64Z, T64N>), grd, dim3(T64N),
                       0, st, g, s, m, in0, in1, out0, out1, b0, b1, omega, zeta, itmax, t,
                       slots, ctl, (int *)nullptr);
}

void launch_rb_sweep32_tile(const KGrid &g, const SplitGrid &s, const FMetric &fm,
                            const float *in0, const float *in1, float *out0, float *out1,
                            const float *r0, const float *r1, float omega, double zeta,
                            double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                            hipStream_t st)
{
    TileMetric<float> m{fm.iDx, fm.iDy, fm.iDz, fm.rx, fm.ry, fm.rz};
    const dim3 grd((g.nx + T32X - 1) / T32X, (g.ny + T32Y - 1) / T32Y,
                   (g.nz + T32Z - 1) / T32Z);
    hipLaunchKernelGGL((k_rb_sweep_tile<float, true, T32X, T32Y, T32Z, T32N>), grd, dim3(T32N), 0,
                       st, g, s, m, in0, in1, out0, out1, r0, r1, omega, zeta, itmax, t, slots,
                       ctl, dirty);
}

}  // namespace ntust
