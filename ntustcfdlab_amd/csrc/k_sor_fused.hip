// k_sor_fused.hip -- one red-black SOR sweep per launch (gfx950).
//
// The two-pass red-black solver (k_sor.hip, k_sor_mixed.hip) launches one
// kernel per colour, and each launch also reads the other colour's plane:
// 32 B per cell and sweep in FP64.  Here one launch does the whole sweep.
// A workgroup owns an x-y tile of TX x TY cells over a chunk of planes and
// streams through the chunk along z with a ring of five planes in LDS
// (both colours, the tile plus two cells in x and y).  At step r it
//   - stores plane r+1 (fetched during the previous step) into the ring
//     and fetches plane r+2 and the next right-hand sides into registers;
//   - updates the red cells of plane r on the tile plus one cell in x and y
//     (from the old black values of planes r-1, r, r+1);
//   - updates the black cells of plane r-1 on the tile (from the new red
//     values of planes r-2, r-1, r).
// The red cells outside the tile, and the red planes just below and above
// the chunk, belong to neighbouring workgroups, which compute the same
// values from the same inputs.  While owners write new values, neighbours
// read the old ones, so a sweep reads one copy of the colour planes and
// writes the other (sweep t reads copy (t-1)&1 and writes copy t&1); after
// ik sweeps the result is copy ik&1.  Traffic per cell and sweep: p of both
// colours and b read once (plus the x-y halo, (TX+4)(TY+4)/(TX TY) of p),
// p written once: about 24 B in FP64, 12 B in FP32.
//
// The point update, the gate on the previous sweep's maximum change
// (GS.f90:11) and the per-sweep maxima in buckets are those of the two-pass
// kernels, so the iterates, the stopping sweep and the reported change are
// bit-identical to them.
#include <hip/hip_runtime.h>

#include "ntust_kernels.h"

namespace ntust {

template <typename T>
struct SweepMetric {
    const T *iDx, *iDy, *iDz, *rx, *ry, *rz;   // offset by +1
};

__device__ __forceinline__ double dfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double dabs(double a) { return fabs(a); }
__device__ __forceinline__ float dabs(float a) { return fabsf(a); }

constexpr int NSLOT = 5;      // planes in the LDS ring
constexpr int STZ_MAX = 64;   // planes per workgroup at most

// the point update (the operation order of k_rb_sweep / k_rb_sweep32)
template <typename T, bool MIXED>
__device__ __forceinline__ T point_update(T pW, T pE, T pS, T pN, T pB, T pF, T po, T ax, T ay,
                                          T az, T rE, T rW, T rN, T rS, T rF, T rB, T rhs,
                                          T omega, T &change)
{
    const T ayz = ay * az, axz = ax * az, axy = ax * ay;
    const T t1 = dfma(rE, pE, rW * pW);
    const T t2 = dfma(rN, pN, rS * pS);
    const T t3 = dfma(rF, pF, rB * pB);
    const T sum = dfma(ayz, t1, dfma(axz, t2, axy * t3));
    const T aP = dfma(ayz, rE + rW, dfma(axz, rN + rS, axy * (rF + rB)));
    const T d = (MIXED ? (sum + rhs) : (sum - rhs)) / aP - po;
    change = dabs(d);
    return dfma(omega, d, po);
}

// MIXED = false: p and b, d = (sum - b)/aP - p   (k_rb_sweep)
// MIXED = true:  e and r, d = (sum + r)/aP - e   (k_rb_sweep32), colour 1
//                shifted as there: r = 0, the r1 plane is not read
//
// Thread t < (TX/2+1)(TY+2) owns the cell pair (x0 + 2p, x0 + 2p + 1) of row
// j0 - 1 + jr (p = t mod (TX/2+1), jr = t div (TX/2+1)) on every plane: the
// red cell of the pair on plane r (tile plus one cell) and, inside the tile,
// the black cell of the pair on plane r-1.  All offsets are fixed per
// thread; only the parity of the plane moves the cells within the pair.
template <typename T, bool MIXED, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT)
k_rb_sweep_stream(KGrid g, SplitGrid s, SweepMetric<T> m, const T *__restrict__ in0,
                  const T *__restrict__ in1, T *__restrict__ out0, T *__restrict__ out1,
                  const T *__restrict__ rhs0, const T *__restrict__ rhs1, T omega, double zeta,
                  double itmax, int t, int tzc, double *slots, SorControl *ctl, int *dirty)
{
    constexpr int W = TX / 2 + 3;            // split entries per row held
    constexpr int HY = TY + 4;               // rows held
    constexpr int PLN = W * HY;              // entries per colour and plane
    constexpr int SL = 2 * PLN;              // entries per ring slot
    constexpr int NLD = (2 * PLN + NT - 1) / NT;
    constexpr int RR = TY + 2;               // rows of the red region
    static_assert(W * RR <= NT, "one pair column per thread");

    __shared__ T ring[NSLOT * SL];
    // z metrics of planes k0-2 .. k0+tzc+1 (a global load inside the plane
    // loop would wait for the prefetch issued before it: vmcnt retires in
    // order)
    __shared__ T mzD[STZ_MAX + 4], mzR[STZ_MAX + 4];
    __shared__ double red[NT / 64];

    const int tid = threadIdx.x;
    const int blin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = tid & 63;
    double gate = lane < SLOT_BUCKETS ? load_flag(&slots[slot_index(t - 1, lane)]) : 0.0;
    const int stopped = load_flag(&ctl->stop);

    const int x0 = blockIdx.x * TX, j0 = blockIdx.y * TY + 1;
    const int k0 = blockIdx.z * tzc + 1;
    const int k1 = min(k0 + tzc - 1, g.nz);
    const int mb = (x0 >> 1) - 1;                  // first split index held
    const unsigned pl = (unsigned)s.mp * (unsigned)s.ry;   // entries per plane
    auto kcl = [&](int k) { return pl * (unsigned)min(max(k, 0), g.nz + 1); };

    // This thread's share of a plane: source (colour plane plus offset within
    // the plane) and LDS offset within a ring slot (-1: none).  Loads are
    // unconditional, from addresses clamped into the arrays: an entry outside
    // the cells (x or y beyond the ghost layer, planes below 0 or above nz+1)
    // is never read by an update, and a load under a branch would make the
    // compiler wait for it at the join.
    const T *ld_p[NLD];
    int ld_l[NLD];
#pragma unroll
    for (int r = 0; r < NLD; r++) {
        const int e = tid + r * NT;
        const int c = e >= PLN;
        const int q = e - c * PLN;
        const int mm = q % W, jj = q / W;
        const int mi = min(max(mb + mm, 0), s.mp - 1), j = min(max(j0 - 2 + jj, 0), g.ny + 1);
        ld_p[r] = (c ? in1 : in0) + (unsigned)(mi + s.mp * j);
        ld_l[r] = e < 2 * PLN ? e : -1;
    }
    auto fetch = [&](int k, T (&v)[NLD]) {
        const unsigned off = kcl(k);
#pragma unroll
        for (int r = 0; r < NLD; r++) v[r] = ld_p[r][off];
    };
    auto put = [&](int sl, const T (&v)[NLD]) {
        T *base = ring + sl;
#pragma unroll
        for (int r = 0; r < NLD; r++)
            if (ld_l[r] >= 0) base[ld_l[r]] = v[r];
    };

    // The pair column of this thread: LDS column pp of row jr + 1, i.e. the
    // cells a = x0 - 2 + 2 pp and a + 1 of row j.  Threads run along the LDS
    // rows (W per row), so the neighbour reads of a wavefront are linear in
    // the thread index (no bank conflicts); columns 0 and W-1 hold halo
    // cells only and have no red cell to update.
    const int pp = tid % W, jr = tid / W;
    const int j = j0 - 1 + jr;
    const bool col = tid < W * RR && pp >= 1 && pp <= TX / 2 + 1 && j >= 1 && j <= g.ny;
    const bool tile_row = jr >= 1 && jr <= TY;                 // black cells (and owned red)
    const int own = pp + W * (jr + 1);                         // LDS entry of the pair
    const int jc = min(max(j, 0), g.ny + 1);
    const int a = x0 - 2 + 2 * pp;
    // split offset of the pair, clamped into the row for the loads of a
    // tile that overhangs the grid (every pair that is stored is inside)
    const unsigned gb = (unsigned)min(max(mb + pp, 0), s.mp - 1) + (unsigned)s.mp * (unsigned)jc;
    const int xpar = (x0 + j) & 1;
    // the metrics of the pair: iDx(a), iDx(a+1), 1/Dxs(a-1..a+1), iDy(j), 1/Dys(j-1..j)
    const int ac = min(max(a, 0), g.nx + 1);
    const T dA = m.iDx[ac], dB = m.iDx[ac + 1], rM = m.rx[ac - 1], rA = m.rx[ac],
            rB = m.rx[ac + 1];
    const T ay = m.iDy[jc], rN = m.ry[jc], rS = m.ry[jc - 1];

    // prologue: z metrics, planes k0-2 and k0-1 into the ring, planes k0 and
    // k0+1 and the first right-hand sides into registers
    T va[NLD], vb[NLD], preA[NLD], preB[NLD];
    fetch(k0 - 2, va);
    fetch(k0 - 1, vb);
    fetch(k0, preA);
    fetch(k0 + 1, preB);
    // Loads run two plane steps ahead.  Steps alternate between two register
    // sets (A on steps r = k0-1, k0+1, ..., B on the others), so a load in
    // flight is never copied: pre holds plane r+1, cR the b of the red cell
    // of plane r, cB the b of the black cell of plane r-1.
    T cRA = rhs0[gb + kcl(k0 - 1)], cRB = rhs0[gb + kcl(k0)];
    T cBA = (T)0, cBB = (T)0;
    // ring slot of plane k is (k + 5) mod 5; sb tracks slot(r)
    int sb = (k0 - 1 + NSLOT) % NSLOT;
    auto sl_of = [&](int d) {
        const int q = sb + d;
        return (q >= NSLOT ? q - NSLOT : (q < 0 ? q + NSLOT : q)) * SL;
    };
    put(sl_of(-1), va);
    put(sl_of(0), vb);
    for (int q = tid; q < 2 * (tzc + 4); q += NT) {
        const int k = k0 - 2 + (q >> 1);
        const T z = k <= g.nz + 2 ? ((q & 1) ? m.rz[k] : m.iDz[k]) : (T)0;
        (q & 1 ? mzR : mzD)[q >> 1] = z;
    }
    // every wavefront reads the same buckets: the decision is uniform over
    // the block, so returning here cannot split a barrier
    gate = wave_max(gate);
    const bool active = !stopped && ((double)(t - 1) < itmax) && (gate > zeta);
    if (!active) {
        if (blin == 0 && tid == 0) ctl->stop = 1;
        return;
    }
    if (blin == 0) clear_next_buckets(slots, t, tid);

    T pcm = (T)0;
    auto step = [&](int r, T (&pre)[NLD], T &cR, T &cB) {
        const int sm = sl_of(-1), sc = sl_of(0), sp = sl_of(1);
        // stage 1: plane r+1 into the ring, fetch plane r+3
        put(sp, pre);
        fetch(r + 3, pre);
        __syncthreads();
        // stage 2: the red cell of the pair on plane r, i = a + par
        const int zr = r - k0 + 2;
        {
            const int par = xpar ^ (r & 1);
            const int i = a + par;
            if (r >= 1 && r <= g.nz && col && i >= 1 && i <= g.nx) {
                const T *B = ring + PLN;               // black of a slot at + PLN
                const T pW = B[sc + own - 1 + par], pE = B[sc + own + par];
                const T pS = B[sc + own - W], pN = B[sc + own + W];
                const T pB = B[sm + own], pF = B[sp + own];
                const T po = ring[sc + own];
                T ch;
                const T nv = point_update<T, MIXED>(
                    pW, pE, pS, pN, pB, pF, po, par ? dB : dA, ay, mzD[zr], par ? rB : rA,
                    par ? rA : rM, rN, rS, mzR[zr], mzR[zr - 1], cR, omega, ch);
                if (ch > pcm) pcm = ch;   // a NaN change never raises the max (GS.f90:61)
                // stage 2 reads black entries and each pair's own red entry
                ring[sc + own] = nv;
                if (r >= k0 && r <= k1 && tile_row && i > x0 && i <= x0 + TX)
                    out0[gb + pl * (unsigned)r] = nv;
            }
        }
        cR = rhs0[gb + kcl(r + 2)];
        __syncthreads();
        // stage 3: the black cell of the pair on plane r-1, i = a + 1 - pb
        const int kb = r - 1;
        {
            const int pb = xpar ^ (kb & 1);
            const int i = a + 1 - pb;
            if (kb >= k0 && kb <= k1 && col && tile_row && i > x0 && i <= x0 + TX &&
                i <= g.nx) {
                const int s2 = sl_of(-2);
                const T *R = ring;                     // red of a slot at + 0
                const T pW = R[sm + own - pb], pE = R[sm + own + 1 - pb];
                const T pS = R[sm + own - W], pN = R[sm + own + W];
                const T pB = R[s2 + own], pF = R[sc + own];
                const T po = ring[sm + PLN + own];
                const int zb = zr - 1;
                T ch;
                const T nv = point_update<T, MIXED>(
                    pW, pE, pS, pN, pB, pF, po, pb ? dA : dB, ay, mzD[zb], pb ? rA : rB,
                    pb ? rM : rA, rN, rS, mzR[zb], mzR[zb - 1], cB, omega, ch);
                if (ch > pcm) pcm = ch;
                out1[gb + pl * (unsigned)kb] = nv;
            }
        }
        cB = MIXED ? (T)0 : rhs1[gb + kcl(r + 1)];
        sb = sb + 1 == NSLOT ? 0 : sb + 1;
    };
    // both steps of a pair always run (a step past k1+1 stores nothing: its
    // red plane is outside the chunk and its black plane is k1+1), so the
    // loads in flight are the same on every path and the compiler's waits
    // can stay two steps behind
    for (int r = k0 - 1; r <= k1 + 1; r += 2) {
        step(r, preA, cRA, cBA);
        step(r + 1, preB, cRB, cBB);
    }

    // the sweep's maximum change
    double pm = wave_max((double)pcm);
    if (lane == 0) red[tid >> 6] = pm;
    __syncthreads();
    if (tid == 0) {
        double mm = red[0];
        for (int w = 1; w < NT / 64; w++) mm = red[w] > mm ? red[w] : mm;
        double *bk = &slots[slot_index(t, blin % SLOT_BUCKETS)];
        // a posted atomic (no return, no uncached pre-read of the bucket,
        // which held every block open for a memory round trip: 12 % of an
        // FP32 launch at 256^3)
        if (mm > 0.0)
            atomicMax((unsigned long long *)bk, (unsigned long long)__double_as_longlong(mm));
        if (blin == 0) {
            ctl->ik = t;
            if (MIXED) *dirty = 1;
        }
    }
}

constexpr int STX = 64, STY = 16, STN = 640;   // (STX/2+3)*(STY+2) = 630 threads used

// planes per workgroup: about two workgroups per CU, at least 4 planes
static int stream_chunk(const KGrid &g)
{
    const int tiles = ((g.nx + STX - 1) / STX) * ((g.ny + STY - 1) / STY);
    const int want = 512;
    int nzc = (want + tiles - 1) / tiles;          // chunks wanted
    if (nzc < 1) nzc = 1;
    int tzc = (g.nz + nzc - 1) / nzc;
    if (tzc < 4) tzc = 4;
    if (tzc > STZ_MAX) tzc = STZ_MAX;
    return tzc;
}

void launch_rb_sweep_fused(const KGrid &g, const SplitGrid &s, const double *rx,
                           const double *ry, const double *rz, const double *in0,
                           const double *in1, double *out0, double *out1, const double *b0,
                           const double *b1, double omega, double zeta, double itmax, int t,
                           double *slots, SorControl *ctl, hipStream_t st)
{
    SweepMetric<double> m{g.iDx, g.iDy, g.iDz, rx, ry, rz};
    const int tzc = stream_chunk(g);
    const dim3 grd((g.nx + STX - 1) / STX, (g.ny + STY - 1) / STY, (g.nz + tzc - 1) / tzc);
    hipLaunchKernelGGL((k_rb_sweep_stream<double, false, STX, STY, STN>), grd, dim3(STN), 0, st, g,
                       s, m, in0, in1, out0, out1, b0, b1, omega, zeta, itmax, t, tzc, slots,
                       ctl, (int *)nullptr);
}

void launch_rb_sweep32_fused(const KGrid &g, const SplitGrid &s, const FMetric &fm,
                             const float *in0, const float *in1, float *out0, float *out1,
                             const float *r0, const float *r1, float omega, double zeta,
                             double itmax, int t, double *slots, SorControl *ctl, int *dirty,
                             hipStream_t st)
{
    SweepMetric<float> m{fm.iDx, fm.iDy, fm.iDz, fm.rx, fm.ry, fm.rz};
    const int tzc = stream_chunk(g);
    const dim3 grd((g.nx + STX - 1) / STX, (g.ny + STY - 1) / STY, (g.nz + tzc - 1) / tzc);
    hipLaunchKernelGGL((k_rb_sweep_stream<float, true, STX, STY, STN>), grd, dim3(STN), 0, st, g, s,
                       m, in0, in1, out0, out1, r0, r1, omega, zeta, itmax, t, tzc, slots, ctl,
                       dirty);
}

// p += e on the interior and e = 0, for the copy of e the last sweep wrote
// (copy ik&1); the shifted colour-1 correction is e1' + s1 (s1 in the r1
// plane); skipped when no sweep of the segment ran
__global__ void __launch_bounds__(256)
k_rb_absorb_pp(KGrid g, SplitGrid s, double *__restrict__ p0, double *__restrict__ p1,
               float *ea0, float *ea1, float *eb0, float *eb1, const float *__restrict__ s1,
               const int *dirty, const SorControl *ctl)
{
    if (!load_flag(dirty)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx || j > g.ny) return;
    const bool second = load_flag(&ctl->ik) & 1;
    float *e0 = second ? eb0 : ea0, *e1 = second ? eb1 : ea1;
    const long o = six(s, i, j, k);
    if ((i + j + k) & 1) {
        p1[o] = p1[o] + ((double)e1[o] + (double)s1[o]);
        e1[o] = 0.0f;
    } else {
        p0[o] = p0[o] + (double)e0[o];
        e0[o] = 0.0f;
    }
}

void launch_rb_absorb_pp(const KGrid &g, const SplitGrid &s, double *p0, double *p1,
                         float *ea0, float *ea1, float *eb0, float *eb1, const float *s1,
                         const int *dirty, const SorControl *ctl, hipStream_t st)
{
    dim3 blk(64, 4), grd((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz);
    hipLaunchKernelGGL(k_rb_absorb_pp, grd, blk, 0, st, g, s, p0, p1, ea0, ea1, eb0, eb1, s1,
                       dirty, ctl);
}

}  // namespace ntust
