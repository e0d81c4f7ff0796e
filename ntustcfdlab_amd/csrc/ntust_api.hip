// ntust_api.hip -- host side of libntust_hip.so: the C ABI declared in
// include/ntust.h, device memory, the stream, and the time-step driver.
//
// Device memory is allocated lazily, by name, the first time a call needs a
// buffer, and kept until ntust_finalize.  Drop-in calls copy whole arrays in
// and out, so untouched elements of an output keep the caller's values, as
// the reference's in-place module arrays do.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/ntust.h"
#include "ntust_kernels.h"

using namespace ntust;

// Scratch of the allocation-free force integral
// (ntust_virtual_force_integrator_noalloc): the line sums FXz, FYz
// ((nx+2)(ny+2) doubles each) and FXy, FYy (nx+2 each).  Each GPU the library
// runs on holds its own copy until the process ends, and every call on that
// GPU shares it.  Host code uses the per-device address that
// hipGetSymbolAddress returns at set-up (g_scratch), never this name.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)1 << 20;   // 8 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};   // device address of scratch, per device

char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(NTUST_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,             \
                        hipGetErrorString(e_));                                            \
    } while (0)

struct Params {
    double dt = 1.0e-3, nu = 0.1, Cs = 0.18, omega = 1.5, zeta = 1.0e-4, itmax = 3000.0,
           lx = 1.0;
    int les = 0, scheme = NTUST_QUICK_CENTRE, sor = NTUST_SOR_LEXICOGRAPHIC, force_every = 1;
};

struct Ctx {
    bool ready = false;
    int dev = 0;                       // the device current at ntust_init
    int nx = 0, ny = 0, nz = 0;
    long nh = 0, ni = 0;
    hipStream_t st = nullptr;
    KGrid g{};
    SplitGrid sg{};
    std::vector<double> h_metric[6];   // iDx Dxs iDy Dys iDz Dzs (host copies)
    double *metric[6] = {};            // device, base (element -1)
    double *recip[3] = {};             // 1/Dxs, 1/Dys, 1/Dzs, base
    float *fmetric[6] = {};            // FP32 iDx, iDy, iDz, 1/Dxs, 1/Dys, 1/Dzs, base
    // FP32 sweeps between FP64 corrections of the forward solve: the first
    // segment of a solve, which carries the solve's initial correction
    // (the whole solution from a cold start), is mix_first_segment long,
    // the later ones, which carry only the change since the last fold,
    // mix_segment.  The gated loop checks the stop on the host every
    // gate_quantum sweeps, independently of the segments.  The adjoint's
    // transposed FP32 sweeps fold after the same segment lengths.
    long mix_first_segment = 64;
    long mix_segment = 256;
    long gate_quantum = 64;
    // arithmetic of the resident step's predictor and eddy viscosity
    // (NTUST_ARITH_*), with their tables (base pointers of device arrays)
    int arith = NTUST_ARITH_EXACT;
    QCoef<double> *qtab64[3][2] = {};
    QCoef<float> *qtab32[3][2] = {};
    double *rmet64[6] = {};            // 1/iDx 1/Dxs 1/iDy 1/Dys 1/iDz 1/Dzs
    float *rmet32[6] = {};
    BcTable bc{};
    Params prm;
    // named device buffers
    struct Buf { std::string name; double *ptr; size_t n; };
    std::vector<Buf> bufs;
    double filt_cs = NAN;              // Cs the filter table was built for
    SorControl *ctl = nullptr;
    double *pmax = nullptr;
    int ik_hint = 0;
    // measurement of the red-black half-sweep; sev: one pair per segment of
    // the FP32-sweep solver's speculative batch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int NSEV = 64;
    hipEvent_t sev[2 * NSEV] = {};
    double prof_ms = 0.0;
    long long prof_launches = 0;
    long long prof_sweeps = 0;         // lexicographic: sweeps in the timed launches
    // lexicographic solver: stamp of the current batch's per-sweep flags,
    // previous solve's sweep count
    int lex_epoch = 0;
    int lex_hint = 0;
    // body force added to the predictor (device halo arrays, "body_*")
    bool body = false;
    double *fu = nullptr, *fv = nullptr, *fw = nullptr;
    // exact FP64 checkpoints of (u, v, w, p), slots of 4 halo arrays
    int nckpt = 0;
    // objective of the adjoint run: weights of the mean kinetic energy,
    // C_D and C_L over the averaging window; ETA gradient wanted
    double adj_w[3] = {1.0, 0.0, 0.0};
    bool adj_eta = false;
    // Lagrangian markers of the direct forcing (device arrays in the named
    // buffers "mk_*"); off when mk.n == 0
    MarkerSet mk{};
};

Ctx C;

// A buffer asked for again with a larger size is reallocated (its contents
// are not kept).  The zero fill is queued on the library stream: the stream
// is non-blocking, so a fill on the null stream would not be ordered before
// the kernels that follow on it (a late fill can wipe their results).
double *buf(const char *name, size_t n, bool zero = false)
{
    for (auto &b : C.bufs) {
        if (b.name != name) continue;
        if (b.n >= n) return b.ptr;
        // the old buffer may still be in use by queued work
        if (hipStreamSynchronize(C.st) != hipSuccess) return nullptr;
        if (hipFree(b.ptr) != hipSuccess) return nullptr;
        b.ptr = nullptr;
        b.n = 0;
        if (hipMalloc(&b.ptr, n * sizeof(double)) != hipSuccess) return nullptr;
        if (zero && hipMemsetAsync(b.ptr, 0, n * sizeof(double), C.st) != hipSuccess)
            return nullptr;
        b.n = n;
        return b.ptr;
    }
    double *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
    if (zero && hipMemsetAsync(p, 0, n * sizeof(double), C.st) != hipSuccess) return nullptr;
    C.bufs.push_back({name, p, n});
    return p;
}

#define NEED(var, name, n, zero)                                                           \
    double *var = buf(name, (n), (zero));                                                  \
    if (!var) return fail(NTUST_E_ALLOC, "device allocation of %s failed", name)

#define NEED_INIT()                                                                        \
    if (!C.ready) return fail(NTUST_E_NOT_INIT, "ntust_init has not been called")

int h2d(double *d, const double *h, size_t n)
{
    HIPCHK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, C.st));
    return 0;
}

int d2h(double *h, const double *d, size_t n)
{
    HIPCHK(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, C.st));
    return 0;
}

#define TRY(x)                                                                             \
    do {                                                                                   \
        int r_ = (x);                                                                      \
        if (r_ < 0) return r_;                                                             \
    } while (0)

int sync() { HIPCHK(hipStreamSynchronize(C.st)); HIPCHK(hipGetLastError()); return 0; }

// (Cs*delta)**2 with delta = (iDx*iDy*iDz)**(1./3.), smagorinsky.f90:20-22:
// libm pow on the host, the exponent the single-precision literal 1./3.
int filter_table(double Cs)
{
    NEED(filt, "filt", (size_t)C.ni, false);
    if (C.filt_cs == Cs) return 0;
    const double third = (double)(1.0f / 3.0f);
    const std::vector<double> &ix = C.h_metric[0], &iy = C.h_metric[2], &iz = C.h_metric[4];
    std::vector<double> t((size_t)C.ni);
    size_t o = 0;
    for (int k = 1; k <= C.nz; k++)
        for (int j = 1; j <= C.ny; j++)
            for (int i = 1; i <= C.nx; i++) {
                const double d = pow(ix[i + 1] * iy[j + 1] * iz[k + 1], third);
                const double cd = Cs * d;
                t[o++] = cd * cd;
            }
    // on the library stream (ordered with the kernels that read the table);
    // t is released only after the copy has finished
    HIPCHK(hipMemcpyAsync(filt, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice,
                          C.st));
    HIPCHK(hipStreamSynchronize(C.st));
    C.filt_cs = Cs;
    return 0;
}

bool kslab_ok(int istart, int iend)
{
    return istart >= 1 && iend <= C.nz && istart <= iend + 1;
}

// ---- pressure solve on device arrays us, vs, ws -> p, pre --------------

// The first batch of a red-black solve runs speculatively: sweeps 1..nb
// without the device gate, each recording its maximum change, for nb = 7/8
// of the previous solve's sweep count (a time step's count changes slowly).
// The host then finds the first sweep after which the reference's loop
// (GS.f90:11) stops; if the batch ran past it, the solve restarts from its
// initial state -- the colour planes are gathered again from p, which the
// solve does not change until the end -- and replays exactly that many
// sweeps.  The rest of the solve uses the device gate.  The arithmetic and
// the sweep count are those of the gated solve: the results are the same
// bits.  The ungated sweeps are 16-18 % faster (FP32 at 128^3 and 256^3).
constexpr long SPEC_MAX = 65536;

long spec_batch(long tmax, long quantum)
{
    const long h = C.ik_hint;
    long nb = h - h / 8;
    nb = nb < tmax ? nb : tmax;
    nb = nb < SPEC_MAX ? nb : SPEC_MAX;
    nb -= nb % quantum;
    return nb >= 2 * quantum ? nb : 0;
}

// A speculative batch of sweeps t0..t1 recorded, for sweeps t0..t1-1, only
// whether each one's maximum change exceeded zeta (flags, zeroed before the
// batch), and for sweep t1 its maximum in the buckets (the next gated sweep
// and the reported maximum read it).  *stop: the first of these sweeps
// whose maximum ends the loop (not > zeta; a NaN maximum ends it too), or
// -1.
int spec_stop(const double *slots, const int *flags, long t0, long t1, double zeta,
              long *stop)
{
    std::vector<int> f((size_t)(t1 - t0 + 1));
    double bk[SLOT_BUCKETS * SLOT_STRIDE];
    HIPCHK(hipMemcpyAsync(f.data(), flags + t0, (t1 - t0 + 1) * sizeof(int),
                          hipMemcpyDeviceToHost, C.st));
    HIPCHK(hipMemcpyAsync(bk, slots + slot_index(t1, 0), sizeof bk, hipMemcpyDeviceToHost, C.st));
    TRY(sync());
    *stop = -1;
    for (long t = t0; t < t1; t++)
        if (f[(size_t)(t - t0)] == 0) {
            *stop = t;
            return 0;
        }
    double m = 0.0;
    for (int b = 0; b < SLOT_BUCKETS; b++) m = bk[b * SLOT_STRIDE] > m ? bk[b * SLOT_STRIDE] : m;
    if (!(m > zeta)) *stop = t1;
    return 0;
}

// end of the second speculative batch: itmax, when the previous solve ran
// to itmax (a solve that the stopping rule ends varies its count from step
// to step, and a tail past its stop would be replayed almost every time);
// else 0, no tail
long spec_tail(long tmax)
{
    if ((long)C.ik_hint < tmax) return 0;
    return tmax < SPEC_MAX ? tmax : SPEC_MAX;
}

// GS.f90's lexicographic SOR (k_sor_lex.hip): sweeps in batches, every
// batch speculative.  Batch t0..t1 ends with the host reading, for sweeps
// t0..t1-1, whether each one changed some cell by more than zeta (its
// stamp), and the maximum change of sweep t1.  The loop of GS.f90:11 stops
// after the first sweep without a stamp (or after t1 when its maximum is
// not above zeta); if that sweep lies inside the batch, the colour planes
// go back to the batch's start -- gathered again from p, which the solve
// leaves alone until the end, or the copy taken before the batch -- and the
// batch runs again, exactly to that sweep.  First batch: 7/8 of the
// previous solve's count, or all of itmax when that solve ran to itmax;
// then a quarter of that count (at least 32 sweeps), doubling each time.
int lex_stop(const int *flags, const double *slots, long t0, long t1, int epoch, double zeta,
             long *stop)
{
    std::vector<int> f((size_t)(t1 - t0 + 1));
    double bk[SLOT_BUCKETS * SLOT_STRIDE];
    if (t1 > t0)
        HIPCHK(hipMemcpyAsync(f.data(), flags + t0, (t1 - t0) * sizeof(int),
                              hipMemcpyDeviceToHost, C.st));
    HIPCHK(hipMemcpyAsync(bk, slots + slot_index(t1, 0), sizeof bk, hipMemcpyDeviceToHost, C.st));
    TRY(sync());
    *stop = -1;
    for (long t = t0; t < t1; t++)
        if (f[(size_t)(t - t0)] != epoch) {
            *stop = t;
            return 0;
        }
    double m = 0.0;
    for (int b = 0; b < SLOT_BUCKETS; b++) m = bk[b * SLOT_STRIDE] > m ? bk[b * SLOT_STRIDE] : m;
    if (!(m > zeta)) *stop = t1;
    return 0;
}

int solve_lexicographic(const double *us, const double *vs, const double *ws, double *p,
                        double *pre, double dt, double omega, double zeta, double itmax,
                        int *ik_out, double *pmax_out)
{
    const KGrid &g = C.g;
    const long tmax = (long)fmin(itmax, 1.0e8) > 0 ? (long)fmin(itmax, 1.0e8) : 0;
    if (!(1.0 > zeta) || tmax < 1) {
        // GS.f90:11 runs no sweep; pre = p (:86-94)
        if (pre) launch_pack_interior(g, p, pre, C.st);
        HIPCHK(hipGetLastError());
        TRY(sync());
        *ik_out = 0;
        *pmax_out = 1.0;
        return 0;
    }
    const size_t np = (size_t)C.sg.plane;
    NEED(P0, "sor_p0", np, true);
    NEED(P1, "sor_p1", np, true);
    NEED(B0, "sor_b0", np, true);
    NEED(B1, "sor_b1", np, true);
    NEED(fl, "lex_flags", (size_t)tmax / 2 + 2, true);
    int *flags = (int *)fl;
    NEED(slots, "sor_slots", (size_t)slot_index(tmax + 2, 0), false);
    launch_rb_gather(g, C.sg, p, P0, P1, C.st);
    launch_rb_rhs(g, C.sg, us, vs, ws, dt, B0, B1, C.st);
    const long hint = C.lex_hint;
    long nb = hint >= tmax ? tmax : (hint > 0 ? hint - hint / 8 : 32);
    long t0 = 1, ik = 0, grow = 0;
    double *K0 = nullptr, *K1 = nullptr;
    for (;;) {
        const long t1 = t0 + (nb > 1 ? nb : 1) - 1 < tmax ? t0 + (nb > 1 ? nb : 1) - 1 : tmax;
        if (t0 > 1) {
            if (!K0) {
                NEED(a, "lex_k0", np, false);
                NEED(b, "lex_k1", np, false);
                K0 = a;
                K1 = b;
            }
            HIPCHK(hipMemcpyAsync(K0, P0, np * sizeof(double), hipMemcpyDeviceToDevice, C.st));
            HIPCHK(hipMemcpyAsync(K1, P1, np * sizeof(double), hipMemcpyDeviceToDevice, C.st));
        }
        int ep = ++C.lex_epoch;
        launch_slots_clear(slots, (int)t1, 1, C.st);
        HIPCHK(hipEventRecord(C.ev0, C.st));
        launch_lex_sweeps(g, C.sg, P0, P1, B0, B1, omega, zeta, (int)t0, (int)t1, flags, ep,
                          slots, C.st);
        HIPCHK(hipEventRecord(C.ev1, C.st));
        HIPCHK(hipGetLastError());
        long q;
        TRY(lex_stop(flags, slots, t0, t1, ep, zeta, &q));
        if (q >= t0 && q < t1) {
            if (t0 == 1) {
                launch_rb_gather(g, C.sg, p, P0, P1, C.st);
            } else {
                HIPCHK(hipMemcpyAsync(P0, K0, np * sizeof(double), hipMemcpyDeviceToDevice,
                                      C.st));
                HIPCHK(hipMemcpyAsync(P1, K1, np * sizeof(double), hipMemcpyDeviceToDevice,
                                      C.st));
            }
            ep = ++C.lex_epoch;
            launch_slots_clear(slots, (int)q, 1, C.st);
            launch_lex_sweeps(g, C.sg, P0, P1, B0, B1, omega, zeta, (int)t0, (int)q, flags, ep,
                              slots, C.st);
            HIPCHK(hipGetLastError());
            ik = q;
            break;
        }
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, C.ev0, C.ev1));
        C.prof_ms += ms;
        C.prof_launches += lex_launches(g, t0, t1);
        C.prof_sweeps += t1 - t0 + 1;
        if (q == t1 || t1 == tmax) {
            ik = t1;
            break;
        }
        // later batches: a quarter of the previous count, doubling, so a
        // solve with no useful count pays O(log itmax) pipeline fills
        t0 = t1 + 1;
        grow = grow ? 2 * grow : (hint / 4 > 32 ? hint / 4 : 32);
        nb = grow;
    }
    C.lex_hint = (int)ik;
    double bk[SLOT_BUCKETS * SLOT_STRIDE];
    HIPCHK(hipMemcpyAsync(bk, slots + slot_index(ik, 0), sizeof bk, hipMemcpyDeviceToHost,
                          C.st));
    launch_rb_scatter(g, C.sg, P0, P1, p, pre, C.st);
    HIPCHK(hipGetLastError());
    TRY(sync());
    double m = 0.0;
    for (int b = 0; b < SLOT_BUCKETS; b++) m = bk[b * SLOT_STRIDE] > m ? bk[b * SLOT_STRIDE] : m;
    *pmax_out = m;
    *ik_out = (int)ik;
    return 0;
}

int solve_pressure(const double *us, const double *vs, const double *ws, double *p,
                   double *pre, double dt, double omega, double zeta, double itmax, int mode,
                   int *ik_out, double *pmax_out)
{
    const KGrid &g = C.g;
    if (mode == NTUST_SOR_LEXICOGRAPHIC)
        return solve_lexicographic(us, vs, ws, p, pre, dt, omega, zeta, itmax, ik_out,
                                   pmax_out);
    if (mode != NTUST_SOR_REDBLACK && mode != NTUST_SOR_MIXED)
        return fail(NTUST_E_ARG, "unsupported SOR mode %d", mode);

    const size_t np = (size_t)C.sg.plane;
    NEED(P0, "sor_p0", np, true);
    NEED(P1, "sor_p1", np, true);
    NEED(B0, "sor_b0", np, true);
    NEED(B1, "sor_b1", np, true);
    const long tmax = (long)fmin(itmax, 1.0e8) > 0 ? (long)fmin(itmax, 1.0e8) : 0;
    NEED(slots, "sor_slots", (size_t)slot_index(tmax + 2, 0), false);
    launch_rb_gather(g, C.sg, p, P0, P1, C.st);
    launch_rb_rhs(g, C.sg, us, vs, ws, dt, B0, B1, C.st);
    // sweep 0: 1.0 (pChangeMax_ before the first sweep), sweep 1: 0, stop
    // and sweep count cleared; each sweep clears the buckets of the next
    launch_sor_init(slots, C.ctl, C.st);
    const double *rx = C.recip[0] + 1, *ry = C.recip[1] + 1, *rz = C.recip[2] + 1;
    SorControl h{};
    long t = 1;
    long batch = C.ik_hint > 16 ? C.ik_hint - 8 : 16;
    if (mode == NTUST_SOR_MIXED) {
        // FP32 correction e and FP32 residual r, colour-split like p
        const size_t nf = np / 2 + 1;   // float planes carved from double buffers
        NEED(e0d, "mix_e0", nf, true);
        NEED(e1d, "mix_e1", nf, true);
        NEED(r0d, "mix_r0", nf, true);
        NEED(r1d, "mix_r1", nf, true);
        float *E0 = (float *)e0d, *E1 = (float *)e1d, *R0 = (float *)r0d, *R1 = (float *)r1d;
        // segment ends: seg1, seg1 + seg, seg1 + 2 seg, ..., and tmax
        const long seg = C.mix_segment, seg1 = C.mix_first_segment;
        auto seg_end = [&](long tt) {
            const long e = tt <= seg1 ? seg1 : seg1 + ((tt - seg1 - 1) / seg + 1) * seg;
            return e < tmax ? e : tmax;
        };
        auto seg_first = [&](long tt) {
            return tt == 1 || (tt > seg1 && (tt - seg1 - 1) % seg == 0);
        };
        auto seg_floor = [&](long x) {   // the last segment end <= x (0: none)
            return x < seg1 ? 0L : seg1 + (x - seg1) / seg * seg;
        };
        const long nseg = (tmax + seg - 1) / seg + 2;
        NEED(dd, "mix_dirty", (size_t)nseg, true);
        int *dirty = (int *)dd;
        HIPCHK(hipMemsetAsync(dirty, 0, nseg * sizeof(int), C.st));
        HIPCHK(hipMemsetAsync(E0, 0, np * sizeof(float), C.st));
        HIPCHK(hipMemsetAsync(E1, 0, np * sizeof(float), C.st));
        // colour 1 runs shifted (e1' = -s1)
        launch_rb_residual(g, C.sg, rx, ry, rz, P0, P1, B0, B1, R0, R1, E1, nullptr, C.st);
        FMetric fm{C.fmetric[0] + 1, C.fmetric[1] + 1, C.fmetric[2] + 1,
                   C.fmetric[3] + 1, C.fmetric[4] + 1, C.fmetric[5] + 1};
        long c = 0;   // the gated loop's segment (its dirty flag)
        const long gq = C.gate_quantum;
        long want = C.ik_hint > gq ? C.ik_hint - gq / 2 : 2 * gq;
        long nb1 = 1.0 > zeta ? seg_floor(spec_batch(tmax, 1)) : 0;
        if (nb1 < 2 * gq) nb1 = 0;
        if (nb1 > 0) {
            // sweeps 1..upto in segments counted from sweep 1, each followed
            // by the FP64 correction, as the gated loop below runs them; all
            // but the last record only whether they exceed zeta
            NEED(fl, "sor_flags", (size_t)SPEC_MAX / 2 + 2, false);
            int *flags = (int *)fl;
            HIPCHK(hipMemsetAsync(flags, 0, (nb1 + 1) * sizeof(int), C.st));
            // the sweeps of every segment but the first (which follows the
            // set-up) are timed, up to NSEV segments
            int nev = 0;
            long tlaunch = 0;
            auto run = [&](long from, long upto, bool timed) {
                for (long tt = from; tt <= upto;) {
                    const long se = seg_end(tt) < upto ? seg_end(tt) : upto;
                    const long n = se - tt + 1;
                    const bool ev = timed && tt > 1 && nev < Ctx::NSEV;
                    if (ev) HIPCHK(hipEventRecord(C.sev[2 * nev], C.st));
                    for (long b = 0; b < n; b++)
                        launch_rb_sweep32_spec(g, C.sg, fm, E0, E1, R0, R1, (float)omega,
                                               (int)(tt + b), slots, zeta,
                                               tt + b < upto ? flags : nullptr, C.st);
                    if (ev) {
                        HIPCHK(hipEventRecord(C.sev[2 * nev + 1], C.st));
                        nev++;
                        tlaunch += 2 * n;
                    }
                    launch_rb_correct(g, C.sg, rx, ry, rz, P0, P1, E0, E1, B0, B1, R0, R1,
                                      nullptr, C.st);
                    tt = se + 1;
                }
                return 0;
            };
            // the state after a segment's correction, from the FP64 planes
            auto restart = [&]() {
                HIPCHK(hipMemsetAsync(E0, 0, np * sizeof(float), C.st));
                HIPCHK(hipMemsetAsync(E1, 0, np * sizeof(float), C.st));
                launch_rb_residual(g, C.sg, rx, ry, rz, P0, P1, B0, B1, R0, R1, E1, nullptr,
                                   C.st);
                return 0;
            };
            launch_slots_clear(slots, 1, (int)nb1 + 1, C.st);
            TRY(run(1, nb1, true));
            long q;
            TRY(spec_stop(slots, flags, 1, nb1, zeta, &q));
            if (q >= 1 && q < nb1) {
                launch_rb_gather(g, C.sg, p, P0, P1, C.st);
                TRY(restart());
                TRY(run(1, q, false));
                HIPCHK(hipGetLastError());
                TRY(sync());
            } else {
                for (int e = 0; e < nev; e++) {
                    float ms = 0.f;
                    HIPCHK(hipEventElapsedTime(&ms, C.sev[2 * e], C.sev[2 * e + 1]));
                    C.prof_ms += ms;
                }
                C.prof_launches += tlaunch;
            }
            h.ik = (int)(q >= 1 ? q : nb1);
            t = q >= 1 ? tmax + 1 : nb1 + 1;
            // after a solve that ran to itmax, the tail to itmax speculative
            // too: from a copy of the FP64 planes, so that a stop inside it
            // replays from there (a tail cut short of itmax ends on a
            // segment end)
            long hi = t <= tmax ? spec_tail(tmax) : 0;
            if (hi < tmax) hi = seg_floor(hi);
            if (hi > nb1) {
                NEED(k0d, "mix_ck0", np, false);
                NEED(k1d, "mix_ck1", np, false);
                HIPCHK(hipMemcpyAsync(k0d, P0, np * sizeof(double), hipMemcpyDeviceToDevice, C.st));
                HIPCHK(hipMemcpyAsync(k1d, P1, np * sizeof(double), hipMemcpyDeviceToDevice, C.st));
                HIPCHK(hipMemsetAsync(flags + nb1 + 1, 0, (hi - nb1) * sizeof(int), C.st));
                launch_slots_clear(slots, (int)nb1 + 1, (int)(hi - nb1) + 1, C.st);
                TRY(run(nb1 + 1, hi, false));
                TRY(spec_stop(slots, flags, nb1 + 1, hi, zeta, &q));
                if (q >= nb1 + 1 && q < hi) {
                    HIPCHK(hipMemcpyAsync(P0, k0d, np * sizeof(double), hipMemcpyDeviceToDevice,
                                          C.st));
                    HIPCHK(hipMemcpyAsync(P1, k1d, np * sizeof(double), hipMemcpyDeviceToDevice,
                                          C.st));
                    TRY(restart());
                    TRY(run(nb1 + 1, q, false));
                    HIPCHK(hipGetLastError());
                    TRY(sync());
                }
                h.ik = (int)(q >= 1 ? q : hi);
                t = q >= 1 ? tmax + 1 : hi + 1;
            }
            want = 2 * gq;
        }
        // Each segment is closed by the FP64 correction (the speculative
        // batch ends on a segment end).  The host checks the stop every
        // `want` sweeps, which may cut a segment in parts: the later parts
        // keep the segment's dirty flag, and a stop inside a segment still
        // closes it with the correction (the sweeps after the stop return
        // at once, so the result is that of the whole segment).
        while (t <= tmax) {
            long queued = 0;
            bool timed = false;
            do {
                const long se = seg_end(t);
                const bool first = seg_first(t);
                const long n = se - t + 1 < want - queued ? se - t + 1 : want - queued;
                if (!timed) HIPCHK(hipEventRecord(C.ev0, C.st));
                for (long b = 0; b < n; b++)
                    launch_rb_sweep32(g, C.sg, fm, E0, E1, R0, R1, (float)omega, zeta, itmax,
                                      (int)(t + b), slots, C.ctl, dirty + c,
                                      b == 0 && first ? 1 : 0, C.st);
                if (!timed) HIPCHK(hipEventRecord(C.ev1, C.st));
                if (t + n - 1 == se) {
                    launch_rb_correct(g, C.sg, rx, ry, rz, P0, P1, E0, E1, B0, B1, R0, R1,
                                      dirty + c, C.st);
                    c++;
                }
                HIPCHK(hipGetLastError());
                if (!timed) { timed = true; batch = n; }
                t += n;
                queued += n;
            } while (queued < want && t <= tmax);
            HIPCHK(hipMemcpyAsync(&h, C.ctl, sizeof h, hipMemcpyDeviceToHost, C.st));
            TRY(sync());
            if (!h.stop && h.ik == t - 1) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, C.ev0, C.ev1));
                C.prof_ms += ms;
                C.prof_launches += 2 * batch;
            }
            if (h.stop) {
                if (t - 1 != seg_end(t - 1)) {
                    launch_rb_correct(g, C.sg, rx, ry, rz, P0, P1, E0, E1, B0, B1, R0, R1,
                                      dirty + c, C.st);
                    HIPCHK(hipGetLastError());
                }
                break;
            }
            want = 2 * gq;
        }
        t = tmax + 1;   // the red-black FP64 loop below does not run
    }
    const long nb64 = mode == NTUST_SOR_REDBLACK && 1.0 > zeta
                          ? spec_batch(tmax, 8)
                          : 0;
    if (nb64 > 0) {
        // all but the last sweep of a batch record only whether they exceed
        // zeta
        NEED(fl, "sor_flags", (size_t)SPEC_MAX / 2 + 2, false);
        int *flags = (int *)fl;
        HIPCHK(hipMemsetAsync(flags, 0, (nb64 + 1) * sizeof(int), C.st));
        auto run = [&](long from, long upto) {
            for (long tt = from; tt <= upto; tt++)
                launch_rb_sweep_spec(g, C.sg, rx, ry, rz, P0, P1, B0, B1, omega, (int)tt, slots,
                                     zeta, tt < upto ? flags : nullptr, C.st);
        };
        launch_slots_clear(slots, 1, (int)nb64 + 1, C.st);
        HIPCHK(hipEventRecord(C.ev0, C.st));
        run(1, nb64);
        HIPCHK(hipEventRecord(C.ev1, C.st));
        long q;
        TRY(spec_stop(slots, flags, 1, nb64, zeta, &q));
        if (q >= 1 && q < nb64) {
            launch_rb_gather(g, C.sg, p, P0, P1, C.st);
            run(1, q);
            HIPCHK(hipGetLastError());
            TRY(sync());
        } else {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, C.ev0, C.ev1));
            C.prof_ms += ms;
            C.prof_launches += 2 * nb64;
        }
        h.ik = (int)(q >= 1 ? q : nb64);
        t = q >= 1 ? tmax + 1 : nb64 + 1;
        // after a solve that ran to itmax, the tail to itmax speculative too:
        // from a copy of the colour planes, so that a stop inside it replays
        // from there
        const long hi = t <= tmax ? spec_tail(tmax) : 0;
        if (hi > nb64) {
            NEED(k0d, "sor_ck0", np, false);
            NEED(k1d, "sor_ck1", np, false);
            HIPCHK(hipMemcpyAsync(k0d, P0, np * sizeof(double), hipMemcpyDeviceToDevice, C.st));
            HIPCHK(hipMemcpyAsync(k1d, P1, np * sizeof(double), hipMemcpyDeviceToDevice, C.st));
            HIPCHK(hipMemsetAsync(flags + nb64 + 1, 0, (hi - nb64) * sizeof(int), C.st));
            launch_slots_clear(slots, (int)nb64 + 1, (int)(hi - nb64) + 1, C.st);
            run(nb64 + 1, hi);
            TRY(spec_stop(slots, flags, nb64 + 1, hi, zeta, &q));
            if (q >= nb64 + 1 && q < hi) {
                HIPCHK(hipMemcpyAsync(P0, k0d, np * sizeof(double), hipMemcpyDeviceToDevice,
                                      C.st));
                HIPCHK(hipMemcpyAsync(P1, k1d, np * sizeof(double), hipMemcpyDeviceToDevice,
                                      C.st));
                run(nb64 + 1, q);
                HIPCHK(hipGetLastError());
                TRY(sync());
            }
            h.ik = (int)(q >= 1 ? q : hi);
            t = q >= 1 ? tmax + 1 : hi + 1;
        }
        batch = 16;
    }
    while (t <= tmax) {
        const long nb = batch < tmax - t + 1 ? batch : tmax - t + 1;
        HIPCHK(hipEventRecord(C.ev0, C.st));
        for (long b = 0; b < nb; b++)
            launch_rb_sweep(g, C.sg, rx, ry, rz, P0, P1, B0, B1, omega, zeta, itmax,
                            (int)(t + b), slots, C.ctl, C.st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(C.ev1, C.st));
        HIPCHK(hipMemcpyAsync(&h, C.ctl, sizeof h, hipMemcpyDeviceToHost, C.st));
        TRY(sync());
        if (!h.stop && h.ik == t + nb - 1) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, C.ev0, C.ev1));
            C.prof_ms += ms;
            C.prof_launches += 2 * nb;
        }
        if (h.stop) break;
        t += nb;
        batch = 16;
    }
    C.ik_hint = h.ik;
    double bk[SLOT_BUCKETS * SLOT_STRIDE];
    HIPCHK(hipMemcpyAsync(bk, slots + slot_index(h.ik, 0), sizeof bk, hipMemcpyDeviceToHost,
                          C.st));
    launch_rb_scatter(g, C.sg, P0, P1, p, pre, C.st);
    HIPCHK(hipGetLastError());
    TRY(sync());
    double m = 0.0;
    for (int b = 0; b < SLOT_BUCKETS; b++) m = bk[b * SLOT_STRIDE] > m ? bk[b * SLOT_STRIDE] : m;
    *pmax_out = m;
    *ik_out = h.ik;
    return 0;
}

// ---- one device-resident step -------------------------------------------

struct StepBufs {
    double *u, *v, *w, *p, *eta, *us, *vs, *ws, *u2, *v2, *w2, *pre, *uc, *vc, *wc, *vis, *filt;
    double *FY = nullptr, *FZ = nullptr, *fxz = nullptr, *fyz = nullptr, *fxy = nullptr,
           *fyy = nullptr, *coef = nullptr;
    // when set, the projection also writes its u1, v1, w1 there (the
    // adjoint's ETA gradient reads them)
    double *u1 = nullptr, *v1 = nullptr, *w1 = nullptr;
};

int step_buffers(StepBufs &b, bool force)
{
    const char *hn[11] = {"u", "v", "w", "p", "eta", "us", "vs", "ws", "u2", "v2", "w2"};
    double **hp[11] = {&b.u, &b.v, &b.w, &b.p, &b.eta, &b.us, &b.vs, &b.ws, &b.u2, &b.v2, &b.w2};
    for (int q = 0; q < 11; q++) {
        *hp[q] = buf(hn[q], C.nh, true);
        if (!*hp[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", hn[q]);
    }
    const char *in[5] = {"pre", "uc", "vc", "wc", "visc"};
    double **ip[5] = {&b.pre, &b.uc, &b.vc, &b.wc, &b.vis};
    for (int q = 0; q < 5; q++) {
        *ip[q] = buf(in[q], C.ni, true);
        if (!*ip[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", in[q]);
    }
    if (force) {
        NEED(a, "FY", C.nh, true);
        NEED(c, "FZ", C.nh, true);
        NEED(d, "force_xz", (size_t)(C.nx + 2) * (C.ny + 2), true);
        NEED(e, "force_yz", (size_t)(C.nx + 2) * (C.ny + 2), true);
        NEED(f, "force_xy", (size_t)(C.nx + 2), true);
        NEED(h, "force_yy", (size_t)(C.nx + 2), true);
        NEED(k, "force_coef", 2, true);
        b.FY = a; b.FZ = c; b.fxz = d; b.fyz = e; b.fxy = f; b.fyy = h; b.coef = k;
    }
    if (C.prm.les) TRY(filter_table(C.prm.Cs));
    NEED(filt, "filt", C.ni, false);
    b.filt = filt;
    return 0;
}

template <typename T>
QTables<T> qtables()
{
    QTables<T> q;
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < 2; k++) {
            if constexpr (sizeof(T) == 8) q.t[a][k] = C.qtab64[a][k];
            else q.t[a][k] = C.qtab32[a][k];
        }
    return q;
}

template <typename T>
RMetric<T> rmetric()
{
    T *const *m;
    if constexpr (sizeof(T) == 8) m = C.rmet64;
    else m = C.rmet32;
    return RMetric<T>{m[0] + 1, m[1] + 1, m[2] + 1, m[3] + 1, m[4] + 1, m[5] + 1};
}

// dJ/dZ_m, dJ/dY_m of the markers (2n doubles), zeroed at allocation
double *marker_gradient()
{
    return C.mk.n > 0 ? buf("mk_G", 2 * (size_t)C.mk.n, true) : nullptr;
}

// One step of main.f90:209-259 on the device state.  force: compute FY, FZ
// and the force integral (synchronises for cd/cl); centre: write uc, vc, wc.
int one_step(const StepBufs &b, bool force, bool centre, int *ik, double *pm, double *cd,
             double *cl)
{
    const Params &q = C.prm;
    const KGrid &g = C.g;
    // main.f90:209 -- the reference evaluates nu_t every step and scales it
    // by LES; with LES = 0 the product is nu + 0 for any finite nu_t, so the
    // evaluation is skipped.
    if (q.les) {
        if (C.arith == NTUST_ARITH_FAST)
            launch_smagorinsky_fast<double>(g, rmetric<double>(), b.u, b.v, b.w, b.filt, b.vis,
                                            1, C.nz, C.st);
        else if (C.arith == NTUST_ARITH_FP32)
            launch_smagorinsky_fast<float>(g, rmetric<float>(), b.u, b.v, b.w, b.filt, b.vis,
                                           1, C.nz, C.st);
        else
            launch_smagorinsky(g, b.u, b.v, b.w, nullptr, nullptr, nullptr, b.filt, b.vis,
                               nullptr, nullptr, nullptr, 1, C.nz, C.st);
    }
    // main.f90:213
    if (q.scheme == NTUST_QUICK_CENTRE && C.arith == NTUST_ARITH_FAST)
        launch_predict_quick_march<double>(g, qtables<double>(), b.u, b.v, b.w,
                                           q.les ? b.vis : nullptr, q.dt, q.nu, q.les, b.us,
                                           b.vs, b.ws, 1, C.nz, C.st);
    else if (q.scheme == NTUST_QUICK_CENTRE && C.arith == NTUST_ARITH_FP32)
        launch_predict_quick_march<float>(g, qtables<float>(), b.u, b.v, b.w,
                                          q.les ? b.vis : nullptr, q.dt, q.nu, q.les, b.us,
                                          b.vs, b.ws, 1, C.nz, C.st);
    else if (q.scheme == NTUST_QUICK_CENTRE)
        launch_predict_quick(g, b.u, b.v, b.w, q.les ? b.vis : nullptr, q.dt, q.nu, q.les, b.us,
                             b.vs, b.ws, 1, C.nz, C.st);
    else
        launch_predict_upwind(g, b.u, b.v, b.w, q.dt, q.nu, b.us, b.vs, b.ws, 1, C.st);
    // body force (not in the reference): u* = u* + dt*f on the interior
    if (C.body) launch_axpy_interior(g, q.dt, C.fu, C.fv, C.fw, b.us, b.vs, b.ws, C.st);
    HIPCHK(hipGetLastError());
    // main.f90:231
    TRY(solve_pressure(b.us, b.vs, b.ws, b.p, b.pre, q.dt, q.omega, q.zeta, q.itmax, q.sor, ik,
                       pm));
    // main.f90:247,255 -- projection, forcing and the update u = u2
    launch_project(g, b.us, b.vs, b.ws, b.p, b.eta, 0.0, 0.0, 0.0, q.dt, b.u1, b.v1, b.w1,
                   nullptr, nullptr, nullptr, nullptr, force ? b.FY : nullptr,
                   force ? b.FZ : nullptr, centre ? b.uc : nullptr, centre ? b.vc : nullptr,
                   centre ? b.wc : nullptr, b.u, b.v, b.w, b.u2, b.v2, b.w2, 1, C.nz, C.st);
    // marker direct forcing on the post-projection velocity (no counterpart
    // in the reference; with ETA = 0 the blend above left u = u1)
    if (C.mk.n > 0) {
        NEED(U, "mk_U", 3 * (size_t)C.mk.n * C.nx, false);
        launch_marker_interp(g, C.mk, b.u, b.v, b.w, U, C.st);
        launch_marker_force(g, C.mk, U, q.dt, b.u, b.v, b.w, force ? b.FY : nullptr,
                            force ? b.FZ : nullptr, C.st);
        if (centre) launch_centre(g, b.u, b.v, b.w, b.uc, b.vc, b.wc, C.st);
    }
    // main.f90:251
    if (force) {
        launch_force_integral(g, b.FY, b.FZ, q.lx, b.fxz, b.fyz, b.fxy, b.fyy, b.coef, C.st);
        double h[2];
        HIPCHK(hipMemcpyAsync(h, b.coef, sizeof h, hipMemcpyDeviceToHost, C.st));
        TRY(sync());
        *cd = h[0];
        *cl = h[1];
    }
    // main.f90:259
    launch_boundary_conditions(g, C.bc, b.u, b.v, b.w, b.p, C.st);
    launch_copy_shell(g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---- discrete adjoint ------------------------------------------------------

struct AdjBufs {
    double *lu, *lv, *lw, *lp, *lus, *lvs, *lws, *lb, *lvisc, *visc, *gu, *gv, *gw;
    double *z0, *z1, *Z0, *Z1;
};

int adj_buffers(AdjBufs &a)
{
    const char *hn[11] = {"adj_lu", "adj_lv", "adj_lw", "adj_lp", "adj_lus", "adj_lvs",
                          "adj_lws", "adj_lb", "adj_gu", "adj_gv", "adj_gw"};
    double **hp[11] = {&a.lu, &a.lv, &a.lw, &a.lp, &a.lus, &a.lvs, &a.lws, &a.lb,
                       &a.gu, &a.gv, &a.gw};
    for (int q = 0; q < 11; q++) {
        *hp[q] = buf(hn[q], C.nh, true);
        if (!*hp[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", hn[q]);
    }
    const char *in[2] = {"adj_lvisc", "adj_visc"};
    double **ip[2] = {&a.lvisc, &a.visc};
    for (int q = 0; q < 2; q++) {
        *ip[q] = buf(in[q], C.ni, true);
        if (!*ip[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", in[q]);
    }
    // colour planes: only interior cells are ever written, the ghost and
    // padding entries stay zero from the allocation
    const char *sn[4] = {"adj_z0", "adj_z1", "adj_Z0", "adj_Z1"};
    double **sp[4] = {&a.z0, &a.z1, &a.Z0, &a.Z1};
    for (int q = 0; q < 4; q++) {
        *sp[q] = buf(sn[q], (size_t)C.sg.plane, true);
        if (!*sp[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", sn[q]);
    }
    return 0;
}

int adjoint_ready()
{
    if (C.prm.sor != NTUST_SOR_REDBLACK && C.prm.sor != NTUST_SOR_MIXED)
        return fail(NTUST_E_ARG, "the adjoint needs the red-black or mixed pressure solver "
                                 "(ntust_configure sor_mode 1 or 2)");
    if (!(C.prm.omega > 0.0)) return fail(NTUST_E_ARG, "the adjoint needs omega > 0");
    return 0;
}

// lambda (adjoint of the state after a step) -> adjoint of the state before
// it.  (ui, vi, wi) is the velocity at the start of the step, sweeps the
// step's SOR sweep count.  Adds dt*lambda_u* to the body-force gradient.
int adjoint_of_step(const AdjBufs &a, const double *eta, const double *filt, const double *ui,
                    const double *vi, const double *wi, int sweeps, const AdjForce &F)
{
    const Params &q = C.prm;
    const KGrid &g = C.g;
    const double *rx = C.recip[0] + 1, *ry = C.recip[1] + 1, *rz = C.recip[2] + 1;
    launch_adj_boundary_conditions(g, C.bc, a.lu, a.lv, a.lw, a.lp, C.st);
    if (C.mk.n > 0) {
        // the marker forcing, transposed: lambda_u1 = lambda_u2 - (S I)^T q;
        // with the step's u1 (geometry gradient wanted) dJ/d(marker position)
        const size_t nm = (size_t)C.mk.n * C.nx;
        NEED(L, "mk_L", 3 * nm, false);
        NEED(GP, "mk_GP", 2 * nm, false);
        double *G = marker_gradient();
        if (!G) return fail(NTUST_E_ALLOC, "device allocation of mk_G failed");
        launch_marker_adjoint(g, C.mk, F, q.dt, a.lu, a.lv, a.lw, F.u1, F.v1, F.w1, L, GP, G,
                              G + C.mk.n, C.st);
    }
    launch_adj_project(g, a.lu, a.lv, a.lw, eta, q.dt, F, a.lus, a.lvs, a.lws, a.lp, a.lu, a.lv,
                       a.lw, C.st);
    // the FP32-sweep solver's adjoint runs its transposed sweeps in FP32
    // too (FP64 folds after segments of the forward's lengths); the FP64
    // solver's in FP64
    AdjMixed mx{};
    const bool mixed = q.sor == NTUST_SOR_MIXED;
    if (mixed) {
        const size_t nf = (size_t)C.sg.plane / 2 + 1;   // float planes in double buffers
        const char *fn[5] = {"adj_e0", "adj_e1", "adj_r0", "adj_r1", "adj_f0"};
        float **fp[5] = {&mx.e0, &mx.e1, &mx.r0, &mx.r1, &mx.f0};
        for (int k = 0; k < 5; k++) {
            double *d = buf(fn[k], nf, true);
            if (!d) return fail(NTUST_E_ALLOC, "device allocation of %s failed", fn[k]);
            *fp[k] = (float *)d;
        }
        NEED(bz, "adj_bzero", (size_t)C.sg.plane, true);
        mx.bzero = bz;
        mx.fm = FMetric{C.fmetric[0] + 1, C.fmetric[1] + 1, C.fmetric[2] + 1,
                        C.fmetric[3] + 1, C.fmetric[4] + 1, C.fmetric[5] + 1};
        mx.seg1 = 2 * C.mix_first_segment;
        mx.seg = 2 * C.mix_segment;
    }
    if (sweeps > 0)
        launch_adj_pressure(g, C.sg, rx, ry, rz, q.omega, sweeps, a.lp, a.lb, a.z0, a.z1, a.Z0,
                            a.Z1, mixed ? &mx : nullptr, C.st);
    else
        HIPCHK(hipMemsetAsync(a.lb, 0, C.nh * sizeof(double), C.st));
    launch_adj_divergence(g, a.lb, q.dt, a.lus, a.lvs, a.lws, a.lu, a.lv, a.lw, C.st);
    launch_axpy_interior(g, q.dt, a.lus, a.lvs, a.lws, a.gu, a.gv, a.gw, C.st);
    const bool quick = q.scheme == NTUST_QUICK_CENTRE;
    const bool les = quick && q.les;
    if (les) {
        launch_smagorinsky(g, ui, vi, wi, nullptr, nullptr, nullptr, filt, a.visc, nullptr,
                           nullptr, nullptr, 1, C.nz, C.st);
        HIPCHK(hipMemsetAsync(a.lvisc, 0, C.ni * sizeof(double), C.st));
    }
    launch_adj_predict(g, quick ? 0 : 1, ui, vi, wi, les ? a.visc : nullptr, q.dt, q.nu, q.les,
                       a.lus, a.lvs, a.lws, a.lu, a.lv, a.lw, les ? a.lvisc : nullptr, C.st);
    if (les) launch_adj_smagorinsky(g, ui, vi, wi, filt, a.lvisc, a.lu, a.lv, a.lw, C.st);
    HIPCHK(hipGetLastError());
    return 0;
}

int kinetic_energy(const double *u, const double *v, const double *w, double *ke)
{
    const size_t nb = (size_t)((C.nx + 63) / 64) * ((C.ny + 3) / 4) * C.nz;
    NEED(part, "ke_part", nb, false);
    const int n = launch_kinetic_energy(C.g, u, v, w, part, C.st);
    HIPCHK(hipGetLastError());
    std::vector<double> h((size_t)n);
    TRY(d2h(h.data(), part, (size_t)n));
    TRY(sync());
    double s = 0.0;
    for (double x : h) s += x;
    *ke = s;
    return 0;
}

int copy3(double *du, double *dv, double *dw, const double *su, const double *sv,
          const double *sw)
{
    const size_t bytes = C.nh * sizeof(double);
    HIPCHK(hipMemcpyAsync(du, su, bytes, hipMemcpyDeviceToDevice, C.st));
    HIPCHK(hipMemcpyAsync(dv, sv, bytes, hipMemcpyDeviceToDevice, C.st));
    HIPCHK(hipMemcpyAsync(dw, sw, bytes, hipMemcpyDeviceToDevice, C.st));
    return 0;
}

// ---- files ------------------------------------------------------------------

// One Fortran unformatted sequential record: 4-byte length, the data, the
// length again (the record layout gfortran and flang write by default).
bool put_record(FILE *f, const void *data, size_t len)
{
    if (len > 0x7fffffffu) return false;
    const int32_t n = (int32_t)len;
    return fwrite(&n, 4, 1, f) == 1 && (len == 0 || fwrite(data, 1, len, f) == len) &&
           fwrite(&n, 4, 1, f) == 1;
}

// Reads one record of exactly `len` bytes (false on any other length).
bool get_record(FILE *f, void *data, size_t len)
{
    int32_t a = 0, b = 0;
    if (fread(&a, 4, 1, f) != 1 || a < 0 || (size_t)a != len) return false;
    if (len && fread(data, 1, len, f) != len) return false;
    return fread(&b, 4, 1, f) == 1 && b == a;
}

// Exact FP64 restart file: a 64-byte header, then the payload arrays in
// the order of restart_arrays(), each as raw little-endian doubles.  The
// hash is FNV-1a over the payload's 64-bit words.
struct RestartHeader {
    char magic[8];            // "NTUSTRS1"
    uint32_t version;         // 1
    int32_t nx, ny, nz;
    int32_t istep;
    int32_t narrays;
    double time;
    uint64_t payload_bytes;
    uint64_t hash;
    uint8_t pad[8];
};
static_assert(sizeof(RestartHeader) == 64, "restart header layout");

constexpr uint64_t FNV_OFFSET = 14695981039346656037ull, FNV_PRIME = 1099511628211ull;

uint64_t fnv_words(uint64_t h, const double *a, size_t n)
{
    for (size_t q = 0; q < n; q++) {
        uint64_t w;
        memcpy(&w, a + q, 8);
        h = (h ^ w) * FNV_PRIME;
    }
    return h;
}

// The resident state a step continues from bit for bit (u, v, w, p, ETA;
// halo), the outputs of the last step (pre, uc, vc, wc) and the steady
// check's previous wc (interior).
struct RestartArray { const char *name; bool halo; };
constexpr RestartArray restart_arrays[] = {
    {"u", true},    {"v", true},    {"w", true},    {"p", true},  {"eta", true},
    {"pre", false}, {"uc", false},  {"vc", false},  {"wc", false}, {"last_velocity", false}};
constexpr int N_RESTART = sizeof(restart_arrays) / sizeof(restart_arrays[0]);

}  // namespace

// =====================================================================
// C ABI
// =====================================================================

extern "C" {

const char *ntust_last_error(void) { return g_err; }

int ntust_finalize(void)
{
    // every release is attempted; the first failure is reported
    hipError_t first = hipSuccess;
    auto keep = [&first](hipError_t e) { if (first == hipSuccess) first = e; };
    if (C.st) keep(hipStreamSynchronize(C.st));
    for (auto &b : C.bufs) keep(hipFree(b.ptr));
    C.bufs.clear();
    for (int m = 0; m < 6; m++)
        if (C.metric[m]) keep(hipFree(C.metric[m]));
    for (int m = 0; m < 3; m++)
        if (C.recip[m]) keep(hipFree(C.recip[m]));
    for (int m = 0; m < 6; m++)
        if (C.fmetric[m]) keep(hipFree(C.fmetric[m]));
    for (int a = 0; a < 3; a++)
        for (int q = 0; q < 2; q++) {
            if (C.qtab64[a][q]) keep(hipFree(C.qtab64[a][q]));
            if (C.qtab32[a][q]) keep(hipFree(C.qtab32[a][q]));
        }
    for (int m = 0; m < 6; m++) {
        if (C.rmet64[m]) keep(hipFree(C.rmet64[m]));
        if (C.rmet32[m]) keep(hipFree(C.rmet32[m]));
    }
    if (C.ctl) keep(hipFree(C.ctl));
    if (C.pmax) keep(hipFree(C.pmax));
    if (C.ev0) keep(hipEventDestroy(C.ev0));
    if (C.ev1) keep(hipEventDestroy(C.ev1));
    for (int e = 0; e < 2 * Ctx::NSEV; e++)
        if (C.sev[e]) keep(hipEventDestroy(C.sev[e]));
    if (C.st) keep(hipStreamDestroy(C.st));
    C = Ctx();
    if (first != hipSuccess)
        return fail(NTUST_E_HIP, "ntust_finalize: %s", hipGetErrorString(first));
    return 0;
}

// the device part of ntust_init; on failure the caller releases what it made
static int init_device(int nx, int ny, int nz, const double *iDx, const double *Dxs,
                       const double *iDy, const double *Dys, const double *iDz,
                       const double *Dzs)
{
    C.nx = nx; C.ny = ny; C.nz = nz;
    HIPCHK(hipGetDevice(&C.dev));
    if (C.dev < 0 || C.dev >= MAX_DEVICES)
        return fail(NTUST_E_ARG, "device %d: at most %d devices", C.dev, MAX_DEVICES);
    // the allocation-free entry point's scratch on this device, looked up
    // once per device with the device current
    if (!g_scratch[C.dev])
        HIPCHK(hipGetSymbolAddress((void **)&g_scratch[C.dev], HIP_SYMBOL(scratch)));
    C.nh = (long)(nx + 4) * (ny + 4) * (nz + 4);
    C.ni = (long)nx * ny * nz;
    HIPCHK(hipStreamCreateWithFlags(&C.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&C.ev0));
    HIPCHK(hipEventCreate(&C.ev1));
    for (int e = 0; e < 2 * Ctx::NSEV; e++) HIPCHK(hipEventCreate(&C.sev[e]));
    const double *src[6] = {iDx, Dxs, iDy, Dys, iDz, Dzs};
    const int len[6] = {nx + 4, nx + 4, ny + 4, ny + 4, nz + 4, nz + 4};
    for (int m = 0; m < 6; m++) {
        C.h_metric[m].assign(src[m], src[m] + len[m]);
        HIPCHK(hipMalloc(&C.metric[m], len[m] * sizeof(double)));
        HIPCHK(hipMemcpy(C.metric[m], src[m], len[m] * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int a = 0; a < 3; a++) {
        const std::vector<double> &D = C.h_metric[2 * a + 1];
        std::vector<double> r(D.size());
        for (size_t m = 0; m < D.size(); m++) r[m] = 1.0 / D[m];
        HIPCHK(hipMalloc(&C.recip[a], r.size() * sizeof(double)));
        HIPCHK(hipMemcpy(C.recip[a], r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice));
        // FP32 copies for the FP32 sweeps: cell width and 1/centre spacing
        const std::vector<double> &W = C.h_metric[2 * a];
        std::vector<float> fw(W.begin(), W.end()), fr(r.begin(), r.end());
        HIPCHK(hipMalloc(&C.fmetric[a], fw.size() * sizeof(float)));
        HIPCHK(hipMemcpy(C.fmetric[a], fw.data(), fw.size() * sizeof(float), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&C.fmetric[3 + a], fr.size() * sizeof(float)));
        HIPCHK(hipMemcpy(C.fmetric[3 + a], fr.data(), fr.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // tables of the fast predictor and eddy viscosity: QUICK face
    // coefficients per axis and kind, reciprocal metrics (FP64 and FP32)
    for (int a = 0; a < 3; a++) {
        const int n = a == 0 ? nx : (a == 1 ? ny : nz);
        const std::vector<double> &iD = C.h_metric[2 * a], &Ds = C.h_metric[2 * a + 1];
        for (int q = 0; q < 2; q++) {
            // element m of a metric is at index m+1 (m = -1 .. n+2)
            auto L = [&](int m) { return q == 0 ? iD[m + 2] : Ds[m + 1]; };
            auto Cw = [&](int m) { return q == 0 ? Ds[m + 1] : iD[m + 1]; };
            std::vector<QCoef<double>> t((size_t)n + 3, QCoef<double>{});
            std::vector<QCoef<float>> tf((size_t)n + 3, QCoef<float>{});
            for (int m = 0; m <= n; m++) {
                QCoef<double> c;
                const double la = 0.125 * L(m) * L(m) / Cw(m), lb = 0.125 * L(m) * L(m) / Cw(m + 1);
                c.am1 = -la / L(m - 1);
                c.a0 = 0.5 + la / L(m) + la / L(m - 1);
                c.ap1 = 0.5 - la / L(m);
                c.b0 = 0.5 - lb / L(m);
                c.bp1 = 0.5 + lb / L(m + 1) + lb / L(m);
                c.bp2 = -lb / L(m + 1);
                c.rL = 1.0 / L(m);
                c.rC = 1.0 / Cw(m);
                t[(size_t)m + 1] = c;
                tf[(size_t)m + 1] = QCoef<float>{(float)c.am1, (float)c.a0, (float)c.ap1,
                                                 (float)c.b0,  (float)c.bp1, (float)c.bp2,
                                                 (float)c.rL,  (float)c.rC};
            }
            HIPCHK(hipMalloc(&C.qtab64[a][q], t.size() * sizeof(t[0])));
            HIPCHK(hipMemcpy(C.qtab64[a][q], t.data(), t.size() * sizeof(t[0]),
                             hipMemcpyHostToDevice));
            HIPCHK(hipMalloc(&C.qtab32[a][q], tf.size() * sizeof(tf[0])));
            HIPCHK(hipMemcpy(C.qtab32[a][q], tf.data(), tf.size() * sizeof(tf[0]),
                             hipMemcpyHostToDevice));
        }
    }
    for (int m = 0; m < 6; m++) {
        const std::vector<double> &M = C.h_metric[m];
        std::vector<double> r(M.size());
        std::vector<float> rf(M.size());
        for (size_t e = 0; e < M.size(); e++) {
            r[e] = 1.0 / M[e];
            rf[e] = (float)r[e];
        }
        HIPCHK(hipMalloc(&C.rmet64[m], r.size() * sizeof(double)));
        HIPCHK(hipMemcpy(C.rmet64[m], r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&C.rmet32[m], rf.size() * sizeof(float)));
        HIPCHK(hipMemcpy(C.rmet32[m], rf.data(), rf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&C.ctl, sizeof(SorControl)));
    HIPCHK(hipMalloc(&C.pmax, sizeof(double)));
    KGrid &g = C.g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.sx = nx + 4; g.sy = ny + 4; g.sxy = (long)g.sx * g.sy;
    g.iDx = C.metric[0] + 1; g.Dxs = C.metric[1] + 1;
    g.iDy = C.metric[2] + 1; g.Dys = C.metric[3] + 1;
    g.iDz = C.metric[4] + 1; g.Dzs = C.metric[5] + 1;
    SplitGrid &s = C.sg;
    s.mp = ((nx + 3) / 2 + 7) / 8 * 8;
    s.ry = ny + 2;
    s.plane = (long)s.mp * s.ry * (nz + 2);
    // the reference's cavity: no-slip everywhere, lid w = 1 on the north wall
    for (int q = 0; q < 18; q++) C.bc.c[q] = BC_NOSLIP;
    C.bc.c[11] = BC_DIRICHLET;
    C.ready = true;
    return 0;
}

int ntust_init(int nx, int ny, int nz, const double *iDx, const double *Dxs,
               const double *iDy, const double *Dys, const double *iDz, const double *Dzs)
{
    if (nx < 2 || ny < 2 || nz < 2)
        return fail(NTUST_E_ARG, "grid %dx%dx%d: every extent must be >= 2", nx, ny, nz);
    if (!iDx || !Dxs || !iDy || !Dys || !iDz || !Dzs)
        return fail(NTUST_E_ARG, "metric array is NULL");
    if (C.ready) ntust_finalize();
    const int rc = init_device(nx, ny, nz, iDx, Dxs, iDy, Dys, iDz, Dzs);
    if (rc < 0) {
        // release whatever was created, keeping the original message
        char msg[sizeof g_err];
        memcpy(msg, g_err, sizeof msg);
        ntust_finalize();
        memcpy(g_err, msg, sizeof msg);
    }
    return rc;
}

int ntust_set_bc(const int *codes)
{
    NEED_INIT();
    if (!codes) return fail(NTUST_E_ARG, "codes is NULL");
    for (int q = 0; q < 18; q++) C.bc.c[q] = codes[q];
    return 0;
}

int ntust_gradient_phi_gauss(const double *Phi, const double *prebase, double *dPhidX,
                             const double *Dxm, const double *Dym, const double *Dzm,
                             int istart, int iend)
{
    NEED_INIT();
    if (!Phi || !dPhidX || !Dxm || !Dym || !Dzm) return fail(NTUST_E_ARG, "NULL array");
    if (!kslab_ok(istart, iend)) return fail(NTUST_E_ARG, "istart/iend out of 1..nz");
    NEED(phi, "grad_phi", C.nh, false);
    NEED(d, "grad_out", 3 * C.ni, false);
    NEED(pre, "grad_pre", (size_t)C.nx * C.ny, false);
    // the three metric arguments are (-1:nx+2) in the reference declaration
    // (smagorinsky.f90:51) and are read at i, j, k only
    const int mlen = 4 + (C.nx > C.ny ? (C.nx > C.nz ? C.nx : C.nz) : (C.ny > C.nz ? C.ny : C.nz));
    NEED(mx, "grad_mx", mlen, false);
    NEED(my, "grad_my", mlen, false);
    NEED(mz, "grad_mz", mlen, false);
    TRY(h2d(phi, Phi, C.nh));
    TRY(h2d(d, dPhidX, 3 * C.ni));
    TRY(h2d(mx, Dxm, C.nx + 4));
    TRY(h2d(my, Dym, C.ny + 4));
    TRY(h2d(mz, Dzm, C.nz + 4));
    if (prebase) TRY(h2d(pre, prebase, (size_t)C.nx * C.ny));
    launch_gradient_flat(C.g, phi, prebase ? pre : nullptr, d, mx + 1, my + 1, mz + 1, istart,
                         iend, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(dPhidX, d, 3 * C.ni));
    return sync();
}

int ntust_smagorinsky(const double *u, const double *v, const double *w,
                      const double *pre_u, const double *pre_v, const double *pre_w,
                      double Cs, int istart, int iend, double *dudx, double *dvdx,
                      double *dwdx, double *Viseff)
{
    NEED_INIT();
    if (!u || !v || !w || !Viseff) return fail(NTUST_E_ARG, "NULL array");
    if ((dudx || dvdx || dwdx) && !(dudx && dvdx && dwdx))
        return fail(NTUST_E_ARG, "dudx, dvdx, dwdx: give all three or none");
    if (!kslab_ok(istart, iend)) return fail(NTUST_E_ARG, "istart/iend out of 1..nz");
    NEED(du, "u", C.nh, true);
    NEED(dv, "v", C.nh, true);
    NEED(dw, "w", C.nh, true);
    NEED(vis, "visc", C.ni, true);
    TRY(filter_table(Cs));
    NEED(filt, "filt", C.ni, false);
    TRY(h2d(du, u, C.nh));
    TRY(h2d(dv, v, C.nh));
    TRY(h2d(dw, w, C.nh));
    TRY(h2d(vis, Viseff, C.ni));
    const size_t nxy = (size_t)C.nx * C.ny;
    double *pu = nullptr, *pv = nullptr, *pw = nullptr;
    if (pre_u || pre_v || pre_w) {
        NEED(a, "pre_u", nxy, true);
        NEED(b, "pre_v", nxy, true);
        NEED(c, "pre_w", nxy, true);
        if (pre_u) { TRY(h2d(a, pre_u, nxy)); pu = a; }
        if (pre_v) { TRY(h2d(b, pre_v, nxy)); pv = b; }
        if (pre_w) { TRY(h2d(c, pre_w, nxy)); pw = c; }
    }
    double *gu = nullptr, *gv = nullptr, *gw = nullptr;
    if (dudx) {
        NEED(a, "dudx", 3 * C.ni, false);
        NEED(b, "dvdx", 3 * C.ni, false);
        NEED(c, "dwdx", 3 * C.ni, false);
        TRY(h2d(a, dudx, 3 * C.ni));
        TRY(h2d(b, dvdx, 3 * C.ni));
        TRY(h2d(c, dwdx, 3 * C.ni));
        gu = a; gv = b; gw = c;
    }
    launch_smagorinsky(C.g, du, dv, dw, pu, pv, pw, filt, vis, gu, gv, gw, istart, iend, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(Viseff, vis, C.ni));
    if (dudx) {
        TRY(d2h(dudx, gu, 3 * C.ni));
        TRY(d2h(dvdx, gv, 3 * C.ni));
        TRY(d2h(dwdx, gw, 3 * C.ni));
    }
    return sync();
}

int ntust_discretisation_quick_centre(const double *u, const double *v, const double *w,
                                      const double *Viseff, double dt, double nu, int LES,
                                      int istart, int iend, double *u_star, double *v_star,
                                      double *w_star)
{
    NEED_INIT();
    if (!u || !v || !w || !Viseff || !u_star || !v_star || !w_star)
        return fail(NTUST_E_ARG, "NULL array");
    if (!kslab_ok(istart, iend)) return fail(NTUST_E_ARG, "istart/iend out of 1..nz");
    NEED(du, "u", C.nh, true);
    NEED(dv, "v", C.nh, true);
    NEED(dw, "w", C.nh, true);
    NEED(vis, "visc", C.ni, true);
    NEED(us, "us", C.nh, true);
    NEED(vs, "vs", C.nh, true);
    NEED(ws, "ws", C.nh, true);
    TRY(h2d(du, u, C.nh));
    TRY(h2d(dv, v, C.nh));
    TRY(h2d(dw, w, C.nh));
    TRY(h2d(vis, Viseff, C.ni));
    TRY(h2d(us, u_star, C.nh));
    TRY(h2d(vs, v_star, C.nh));
    TRY(h2d(ws, w_star, C.nh));
    launch_predict_quick(C.g, du, dv, dw, vis, dt, nu, LES, us, vs, ws, istart, iend, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(u_star, us, C.nh));
    TRY(d2h(v_star, vs, C.nh));
    TRY(d2h(w_star, ws, C.nh));
    return sync();
}

int ntust_discretisation_upwind(const double *u, const double *v, const double *w,
                                double dt, double nu, double *u_star, double *v_star,
                                double *w_star)
{
    NEED_INIT();
    if (!u || !v || !w || !u_star || !v_star || !w_star) return fail(NTUST_E_ARG, "NULL array");
    NEED(du, "u", C.nh, true);
    NEED(dv, "v", C.nh, true);
    NEED(dw, "w", C.nh, true);
    NEED(us, "us", C.nh, true);
    NEED(vs, "vs", C.nh, true);
    NEED(ws, "ws", C.nh, true);
    TRY(h2d(du, u, C.nh));
    TRY(h2d(dv, v, C.nh));
    TRY(h2d(dw, w, C.nh));
    TRY(h2d(us, u_star, C.nh));
    TRY(h2d(vs, v_star, C.nh));
    TRY(h2d(ws, w_star, C.nh));
    launch_predict_upwind(C.g, du, dv, dw, dt, nu, us, vs, ws, 0, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(u_star, us, C.nh));
    TRY(d2h(v_star, vs, C.nh));
    TRY(d2h(w_star, ws, C.nh));
    return sync();
}

int ntust_gauss_seidel(const double *u_star, const double *v_star, const double *w_star,
                       double *p, double *pre, double dt, double omega, double zeta,
                       double itmax, int *ik, double *pChangeMax)
{
    NEED_INIT();
    if (!u_star || !v_star || !w_star || !p) return fail(NTUST_E_ARG, "NULL array");
    NEED(us, "us", C.nh, true);
    NEED(vs, "vs", C.nh, true);
    NEED(ws, "ws", C.nh, true);
    NEED(dp, "p", C.nh, true);
    NEED(dpre, "pre", C.ni, true);
    TRY(h2d(us, u_star, C.nh));
    TRY(h2d(vs, v_star, C.nh));
    TRY(h2d(ws, w_star, C.nh));
    TRY(h2d(dp, p, C.nh));
    int n = 0;
    double pm = 0.0;
    TRY(solve_pressure(us, vs, ws, dp, dpre, dt, omega, zeta, itmax, C.prm.sor, &n, &pm));
    TRY(d2h(p, dp, C.nh));
    TRY(sync());
    if (pre) {
        // pre = p on the interior (GS.f90:86-94)
        const int nx = C.nx, ny = C.ny, nz = C.nz;
        const long sx = nx + 4, sxy = (long)sx * (ny + 4);
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                memcpy(pre + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1)),
                       p + 2 + sx * (j + 1) + sxy * (k + 1), nx * sizeof(double));
    }
    if (ik) *ik = n;
    if (pChangeMax) *pChangeMax = pm;
    return ((double)n >= itmax && !(pm <= zeta)) ? NTUST_WARN_ITMAX : 0;
}

int ntust_calcul_new_velocity(const double *u_star, const double *v_star,
                              const double *w_star, const double *p, const double *ETA,
                              double u_solid, double v_solid, double w_solid, double dt,
                              int istart, int iend, double *u1, double *v1, double *w1,
                              double *u2, double *v2, double *w2, double *FX, double *FY,
                              double *FZ, double *uc, double *vc, double *wc)
{
    NEED_INIT();
    const double *in[5] = {u_star, v_star, w_star, p, ETA};
    double *out[12] = {u1, v1, w1, u2, v2, w2, FX, FY, FZ, uc, vc, wc};
    for (auto a : in) if (!a) return fail(NTUST_E_ARG, "NULL array");
    for (auto a : out) if (!a) return fail(NTUST_E_ARG, "NULL array");
    if (!kslab_ok(istart, iend)) return fail(NTUST_E_ARG, "istart/iend out of 1..nz");
    static const char *hn[14] = {"us", "vs", "ws", "p", "eta", "u1", "v1", "w1",
                                 "u2", "v2", "w2", "FX", "FY", "FZ"};
    double *d[14];
    for (int q = 0; q < 14; q++) {
        d[q] = buf(hn[q], C.nh, true);
        if (!d[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", hn[q]);
    }
    NEED(duc, "uc", C.ni, true);
    NEED(dvc, "vc", C.ni, true);
    NEED(dwc, "wc", C.ni, true);
    for (int q = 0; q < 5; q++) TRY(h2d(d[q], in[q], C.nh));
    for (int q = 0; q < 9; q++) TRY(h2d(d[5 + q], out[q], C.nh));
    launch_project(C.g, d[0], d[1], d[2], d[3], d[4], u_solid, v_solid, w_solid, dt, d[5], d[6],
                   d[7], d[8], d[9], d[10], d[11], d[12], d[13], duc, dvc, dwc, nullptr,
                   nullptr, nullptr, d[8], d[9], d[10], istart, iend, C.st);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 9; q++) TRY(d2h(out[q], d[5 + q], C.nh));
    // uc, vc, wc are written on istart..iend only; keep the rest
    const size_t nxy = (size_t)C.nx * C.ny;
    const size_t off = nxy * (istart - 1), cnt = nxy * (iend - istart + 1);
    TRY(d2h(uc + off, duc + off, cnt));
    TRY(d2h(vc + off, dvc + off, cnt));
    TRY(d2h(wc + off, dwc + off, cnt));
    return sync();
}

int ntust_updating_velocity(const double *u2, const double *v2, const double *w2,
                            int istart, int iend, double *u, double *v, double *w)
{
    NEED_INIT();
    if (!u2 || !v2 || !w2 || !u || !v || !w) return fail(NTUST_E_ARG, "NULL array");
    if (!kslab_ok(istart, iend)) return fail(NTUST_E_ARG, "istart/iend out of 1..nz");
    NEED(a, "u2", C.nh, true);
    NEED(b, "v2", C.nh, true);
    NEED(c, "w2", C.nh, true);
    NEED(x, "u", C.nh, true);
    NEED(y, "v", C.nh, true);
    NEED(z, "w", C.nh, true);
    TRY(h2d(a, u2, C.nh));
    TRY(h2d(b, v2, C.nh));
    TRY(h2d(c, w2, C.nh));
    TRY(h2d(x, u, C.nh));
    TRY(h2d(y, v, C.nh));
    TRY(h2d(z, w, C.nh));
    launch_copy_interior(C.g, a, b, c, x, y, z, istart, iend, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(u, x, C.nh));
    TRY(d2h(v, y, C.nh));
    TRY(d2h(w, z, C.nh));
    return sync();
}

int ntust_virtual_force_integrator(const double *FY, const double *FZ, double lx,
                                   double *cDrag, double *cLift)
{
    NEED_INIT();
    if (!FY || !FZ || !cDrag || !cLift) return fail(NTUST_E_ARG, "NULL argument");
    NEED(fy, "FY", C.nh, true);
    NEED(fz, "FZ", C.nh, true);
    NEED(fxz, "force_xz", (size_t)(C.nx + 2) * (C.ny + 2), true);
    NEED(fyz, "force_yz", (size_t)(C.nx + 2) * (C.ny + 2), true);
    NEED(fxy, "force_xy", (size_t)(C.nx + 2), true);
    NEED(fyy, "force_yy", (size_t)(C.nx + 2), true);
    NEED(coef, "force_coef", 2, true);
    TRY(h2d(fy, FY, C.nh));
    TRY(h2d(fz, FZ, C.nh));
    launch_force_integral(C.g, fy, fz, lx, fxz, fyz, fxy, fyy, coef, C.st);
    HIPCHK(hipGetLastError());
    double h[2];
    TRY(d2h(h, coef, 2));
    TRY(sync());
    *cDrag = h[0];
    *cLift = h[1];
    return 0;
}

int ntust_virtual_force_integrator_noalloc(const double *FY, const double *FZ, double lx,
                                           double *coef)
{
    NEED_INIT();
    if (!FY || !FZ || !coef) return fail(NTUST_E_ARG, "NULL argument");
    const size_t a = (size_t)(C.nx + 2) * (C.ny + 2);
    const size_t need = 2 * a + 2 * (size_t)(C.nx + 2);
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return fail(NTUST_E_ARG,
                    "grid %dx%d: the allocation-free force integral needs %zu scratch doubles "
                    "(2(nx+2)(ny+3)), at most %zu",
                    C.nx, C.ny, need, NOALLOC_SCRATCH_DOUBLES);
    double *s = g_scratch[C.dev];
    if (!s) return fail(NTUST_E_STATE, "scratch of device %d not set up", C.dev);
    double *fxz = s, *fyz = s + a, *fxy = s + 2 * a, *fyy = fxy + (C.nx + 2);
    launch_force_clear(C.g, fxz, fyz, fxy, fyy, C.st);
    launch_force_integral(C.g, FY, FZ, lx, fxz, fyz, fxy, fyy, coef, C.st);
    HIPCHK(hipGetLastError());
    return sync();
}

int ntust_boundary_conditions(double *u, double *v, double *w, double *p, double *u1,
                              double *v1, double *w1, double *u2, double *v2, double *w2,
                              double *u_star, double *v_star, double *w_star)
{
    NEED_INIT();
    if (!u || !v || !w) return fail(NTUST_E_ARG, "NULL velocity array");
    NEED(du, "u", C.nh, true);
    NEED(dv, "v", C.nh, true);
    NEED(dw, "w", C.nh, true);
    TRY(h2d(du, u, C.nh));
    TRY(h2d(dv, v, C.nh));
    TRY(h2d(dw, w, C.nh));
    double *dp = nullptr;
    if (p) {
        NEED(a, "p", C.nh, true);
        TRY(h2d(a, p, C.nh));
        dp = a;
    }
    launch_boundary_conditions(C.g, C.bc, du, dv, dw, dp, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(u, du, C.nh));
    TRY(d2h(v, dv, C.nh));
    TRY(d2h(w, dw, C.nh));
    if (p) TRY(d2h(p, dp, C.nh));
    TRY(sync());
    // boundary_conditions.f90:305-324: whole-array copies (host side, the
    // arrays are the caller's)
    const size_t bytes = C.nh * sizeof(double);
    double *tu[3] = {u1, u2, u_star}, *tv[3] = {v1, v2, v_star}, *tw[3] = {w1, w2, w_star};
    for (int q = 0; q < 3; q++) {
        if (tu[q]) memcpy(tu[q], u, bytes);
        if (tv[q]) memcpy(tv[q], v, bytes);
        if (tw[q]) memcpy(tw[q], w, bytes);
    }
    return 0;
}

// ---------------------------------------------------------------------
// members of the Krylov pressure path
// ---------------------------------------------------------------------

int ntust_poisson_coeffs(double *coef, int *jcoef)
{
    NEED_INIT();
    if (!coef || !jcoef) return fail(NTUST_E_ARG, "NULL argument");
    NEED(c, "poisson_coef", 7 * C.ni, false);
    launch_poisson_coeffs(C.g, c, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(coef, c, 7 * C.ni));
    TRY(sync());
    // solver_1st_member.f90:71-74
    jcoef[0] = 0;
    jcoef[1] = 1;
    jcoef[2] = C.nx;
    jcoef[3] = C.nx * C.ny;
    return 0;
}

int ntust_poisson_rhs(const double *u_star, const double *v_star, const double *w_star,
                      double dt, double *b)
{
    NEED_INIT();
    if (!u_star || !v_star || !w_star || !b) return fail(NTUST_E_ARG, "NULL array");
    NEED(us, "us", C.nh, true);
    NEED(vs, "vs", C.nh, true);
    NEED(ws, "ws", C.nh, true);
    NEED(db, "poisson_b", C.ni, false);
    TRY(h2d(us, u_star, C.nh));
    TRY(h2d(vs, v_star, C.nh));
    TRY(h2d(ws, w_star, C.nh));
    launch_poisson_rhs(C.g, us, vs, ws, dt, db, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(b, db, C.ni));
    return sync();
}

int ntust_pressure_field(const double *x, double *p, double *pre)
{
    NEED_INIT();
    if (!x || !p || !pre) return fail(NTUST_E_ARG, "NULL array");
    NEED(dx, "pressure_x", C.ni, false);
    NEED(dp, "p", C.nh, true);
    NEED(dpre, "pre", C.ni, true);
    TRY(h2d(dx, x, C.ni));
    TRY(h2d(dp, p, C.nh));
    launch_pressure_field(C.g, dx, dp, dpre, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(p, dp, C.nh));
    TRY(d2h(pre, dpre, C.ni));
    return sync();
}

// ---------------------------------------------------------------------
// device-resident stepping
// ---------------------------------------------------------------------

int ntust_configure(double dt, double nu, double Cs, int LES, int scheme, double omega,
                    double zeta, double itmax, int sor_mode, double lx, int force_every)
{
    NEED_INIT();
    if (scheme != NTUST_QUICK_CENTRE && scheme != NTUST_UPWIND)
        return fail(NTUST_E_ARG, "unknown scheme %d", scheme);
    if (sor_mode != NTUST_SOR_LEXICOGRAPHIC && sor_mode != NTUST_SOR_REDBLACK &&
        sor_mode != NTUST_SOR_MIXED)
        return fail(NTUST_E_ARG, "unknown SOR mode %d", sor_mode);
    if (!(itmax >= 0.0)) return fail(NTUST_E_ARG, "itmax must be >= 0");
    Params &q = C.prm;
    q.dt = dt; q.nu = nu; q.Cs = Cs; q.les = LES; q.scheme = scheme; q.omega = omega;
    q.zeta = zeta; q.itmax = itmax; q.sor = sor_mode; q.lx = lx; q.force_every = force_every;
    return 0;
}

int ntust_state_upload(const double *u, const double *v, const double *w, const double *p,
                       const double *ETA)
{
    NEED_INIT();
    if (!u || !v || !w || !p || !ETA) return fail(NTUST_E_ARG, "NULL array");
    NEED(du, "u", C.nh, true);
    NEED(dv, "v", C.nh, true);
    NEED(dw, "w", C.nh, true);
    NEED(dp, "p", C.nh, true);
    NEED(de, "eta", C.nh, true);
    NEED(us, "us", C.nh, true);
    NEED(vs, "vs", C.nh, true);
    NEED(ws, "ws", C.nh, true);
    NEED(u2, "u2", C.nh, true);
    NEED(v2, "v2", C.nh, true);
    NEED(w2, "w2", C.nh, true);
    TRY(h2d(du, u, C.nh));
    TRY(h2d(dv, v, C.nh));
    TRY(h2d(dw, w, C.nh));
    TRY(h2d(dp, p, C.nh));
    TRY(h2d(de, ETA, C.nh));
    launch_copy_shell(C.g, du, dv, dw, us, vs, ws, u2, v2, w2, C.st);
    HIPCHK(hipGetLastError());
    return sync();
}

int ntust_state_download(double *u, double *v, double *w, double *p, double *pre, double *uc,
                         double *vc, double *wc)
{
    NEED_INIT();
    NEED(du, "u", C.nh, true);
    NEED(dv, "v", C.nh, true);
    NEED(dw, "w", C.nh, true);
    NEED(dp, "p", C.nh, true);
    if (u) TRY(d2h(u, du, C.nh));
    if (v) TRY(d2h(v, dv, C.nh));
    if (w) TRY(d2h(w, dw, C.nh));
    if (p) TRY(d2h(p, dp, C.nh));
    if (pre) { NEED(a, "pre", C.ni, true); TRY(d2h(pre, a, C.ni)); }
    if (uc) { NEED(a, "uc", C.ni, true); TRY(d2h(uc, a, C.ni)); }
    if (vc) { NEED(a, "vc", C.ni, true); TRY(d2h(vc, a, C.ni)); }
    if (wc) { NEED(a, "wc", C.ni, true); TRY(d2h(wc, a, C.ni)); }
    return sync();
}

int ntust_step(int nsteps, int *ik, double *pchange, double *cdrag, double *clift)
{
    NEED_INIT();
    if (nsteps < 0) return fail(NTUST_E_ARG, "nsteps < 0");
    const Params &q = C.prm;
    StepBufs b;
    TRY(step_buffers(b, q.force_every > 0));
    int warn = 0;
    for (int s = 0; s < nsteps; s++) {
        const bool force = q.force_every > 0 && ((s + 1) % q.force_every == 0);
        int n = 0;
        double pm = 0.0, cd = NAN, cl = NAN;
        TRY(one_step(b, force, s == nsteps - 1, &n, &pm, &cd, &cl));
        if ((double)n >= q.itmax && !(pm <= q.zeta)) warn = NTUST_WARN_ITMAX;
        if (ik) ik[s] = n;
        if (pchange) pchange[s] = pm;
        if (cdrag) cdrag[s] = cd;
        if (clift) clift[s] = cl;
    }
    TRY(sync());
    return warn;
}

// ---------------------------------------------------------------------
// discrete adjoint of the resident step
// ---------------------------------------------------------------------

int ntust_set_body_force(const double *fu, const double *fv, const double *fw)
{
    NEED_INIT();
    if (!fu && !fv && !fw) {
        C.body = false;
        return 0;
    }
    if (!fu || !fv || !fw) return fail(NTUST_E_ARG, "fu, fv, fw: give all three or none");
    NEED(a, "body_u", C.nh, true);
    NEED(b, "body_v", C.nh, true);
    NEED(c, "body_w", C.nh, true);
    TRY(h2d(a, fu, C.nh));
    TRY(h2d(b, fv, C.nh));
    TRY(h2d(c, fw, C.nh));
    C.fu = a; C.fv = b; C.fw = c;
    C.body = true;
    return sync();
}

int ntust_kinetic_energy(double *ke)
{
    NEED_INIT();
    if (!ke) return fail(NTUST_E_ARG, "ke is NULL");
    NEED(u, "u", C.nh, true);
    NEED(v, "v", C.nh, true);
    NEED(w, "w", C.nh, true);
    return kinetic_energy(u, v, w, ke);
}

int ntust_checkpoint_reserve(int nslots)
{
    NEED_INIT();
    if (nslots < 0) return fail(NTUST_E_ARG, "nslots < 0");
    if (nslots == 0) return 0;
    NEED(ck, "ckpt", (size_t)nslots * 4 * C.nh, false);
    (void)ck;
    if (nslots > C.nckpt) C.nckpt = nslots;
    return 0;
}

int ntust_checkpoint_save(int slot)
{
    NEED_INIT();
    if (slot < 0 || slot >= C.nckpt)
        return fail(NTUST_E_ARG, "checkpoint slot %d outside 0..%d", slot, C.nckpt - 1);
    StepBufs b;
    TRY(step_buffers(b, false));
    NEED(ck, "ckpt", (size_t)C.nckpt * 4 * C.nh, false);
    double *d = ck + (size_t)slot * 4 * C.nh;
    TRY(copy3(d, d + C.nh, d + 2 * C.nh, b.u, b.v, b.w));
    HIPCHK(hipMemcpyAsync(d + 3 * C.nh, b.p, C.nh * sizeof(double), hipMemcpyDeviceToDevice,
                          C.st));
    return sync();
}

int ntust_checkpoint_load(int slot)
{
    NEED_INIT();
    if (slot < 0 || slot >= C.nckpt)
        return fail(NTUST_E_ARG, "checkpoint slot %d outside 0..%d", slot, C.nckpt - 1);
    StepBufs b;
    TRY(step_buffers(b, false));
    NEED(ck, "ckpt", (size_t)C.nckpt * 4 * C.nh, false);
    const double *d = ck + (size_t)slot * 4 * C.nh;
    TRY(copy3(b.u, b.v, b.w, d, d + C.nh, d + 2 * C.nh));
    HIPCHK(hipMemcpyAsync(b.p, d + 3 * C.nh, C.nh * sizeof(double), hipMemcpyDeviceToDevice,
                          C.st));
    launch_copy_shell(C.g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
    HIPCHK(hipGetLastError());
    return sync();
}

int ntust_adjoint_set(const double *lu, const double *lv, const double *lw, const double *lp)
{
    NEED_INIT();
    AdjBufs a;
    TRY(adj_buffers(a));
    const double *src[4] = {lu, lv, lw, lp};
    double *dst[4] = {a.lu, a.lv, a.lw, a.lp};
    for (int q = 0; q < 4; q++) {
        if (src[q]) TRY(h2d(dst[q], src[q], C.nh));
        else HIPCHK(hipMemsetAsync(dst[q], 0, C.nh * sizeof(double), C.st));
    }
    double *gr[3] = {a.gu, a.gv, a.gw};
    for (int q = 0; q < 3; q++) HIPCHK(hipMemsetAsync(gr[q], 0, C.nh * sizeof(double), C.st));
    if (C.adj_eta) {
        NEED(ge, "adj_geta", C.nh, true);
        HIPCHK(hipMemsetAsync(ge, 0, C.nh * sizeof(double), C.st));
    }
    if (C.mk.n > 0) {
        double *G = marker_gradient();
        if (!G) return fail(NTUST_E_ALLOC, "device allocation of mk_G failed");
        HIPCHK(hipMemsetAsync(G, 0, 2 * (size_t)C.mk.n * sizeof(double), C.st));
    }
    return sync();
}

int ntust_adjoint_get(double *lu, double *lv, double *lw, double *lp)
{
    NEED_INIT();
    AdjBufs a;
    TRY(adj_buffers(a));
    if (lu) TRY(d2h(lu, a.lu, C.nh));
    if (lv) TRY(d2h(lv, a.lv, C.nh));
    if (lw) TRY(d2h(lw, a.lw, C.nh));
    if (lp) TRY(d2h(lp, a.lp, C.nh));
    return sync();
}

int ntust_adjoint_gradient(double *gfu, double *gfv, double *gfw)
{
    NEED_INIT();
    AdjBufs a;
    TRY(adj_buffers(a));
    if (gfu) TRY(d2h(gfu, a.gu, C.nh));
    if (gfv) TRY(d2h(gfv, a.gv, C.nh));
    if (gfw) TRY(d2h(gfw, a.gw, C.nh));
    return sync();
}
int ntust_adjoint_objective(double w_ke, double w_cd, double w_cl, int eta_gradient)
{
    NEED_INIT();
    if (!std::isfinite(w_ke) || !std::isfinite(w_cd) || !std::isfinite(w_cl))
        return fail(NTUST_E_ARG, "objective weights must be finite");
    C.adj_w[0] = w_ke;
    C.adj_w[1] = w_cd;
    C.adj_w[2] = w_cl;
    C.adj_eta = eta_gradient != 0;
    return 0;
}
int ntust_adjoint_eta_gradient(double *geta)
{
    NEED_INIT();
    if (!C.adj_eta)
        return fail(NTUST_E_STATE, "no ETA gradient: ntust_adjoint_objective(..., eta_gradient=1) "
                                   "was not set before the adjoint");
    NEED(ge, "adj_geta", C.nh, true);
    if (geta) TRY(d2h(geta, ge, C.nh));
    return sync();
}

int ntust_set_markers(int n, const double *zm, const double *ym, const double *wm, double hx,
                      double hy, double hz, const double *Xn, const double *Yn, const double *Zn)
{
    NEED_INIT();
    if (n < 0) return fail(NTUST_E_ARG, "marker count %d", n);
    C.mk = MarkerSet{};
    if (n == 0) return 0;
    if (!zm || !ym || !wm || !Xn || !Yn || !Zn) return fail(NTUST_E_ARG, "NULL argument");
    if (!(hx > 0.0 && hy > 0.0 && hz > 0.0))
        return fail(NTUST_E_ARG, "kernel widths must be > 0");
    const int nx = C.nx, ny = C.ny, nz = C.nz;
    auto pt = [](const double *N, int i, bool node) {
        return node ? N[i + 2] : 0.5 * (N[i + 1] + N[i + 2]);
    };
    // window starts: the first point inside the kernel's support (|d| < 2 h),
    // and no point past the window's last one inside it
    const double rad = MK_RADIUS;
    std::vector<int> win(6 * (size_t)n);   // jw[a][m] at a*n+m, kw at (3+a)*n+m
    for (int a = 0; a < 3; a++)
        for (int m = 0; m < n; m++) {
            int j = 1;
            while (j <= ny && !(pt(Yn, j, a == 1) > ym[m] - rad * hy)) j++;
            int k = 1;
            while (k <= nz && !(pt(Zn, k, a == 2) > zm[m] - rad * hz)) k++;
            if ((j + MK_WIN <= ny && pt(Yn, j + MK_WIN, a == 1) < ym[m] + rad * hy) ||
                (k + MK_WIN <= nz && pt(Zn, k + MK_WIN, a == 2) < zm[m] + rad * hz))
                return fail(NTUST_E_ARG,
                            "marker %d: the grid is finer than the kernel width there (more "
                            "than %d points within %g h); narrow hy/hz", m, MK_WIN, rad);
            win[(size_t)a * n + m] = j;
            win[(size_t)(3 + a) * n + m] = k;
        }
    // spread lists: per component, the interior points (j, k) of the windows
    // in increasing (k, j), each with its markers in increasing order
    std::vector<int> ints(win);
    int ne[3], off_sj[3], off_sk[3], off_so[3], off_sm[3];
    for (int a = 0; a < 3; a++) {
        std::vector<std::vector<int>> lists((size_t)(ny + 2) * (nz + 2));
        for (int m = 0; m < n; m++)
            for (int dk = 0; dk < MK_WIN; dk++)
                for (int dj = 0; dj < MK_WIN; dj++) {
                    const int j = win[(size_t)a * n + m] + dj, k = win[(size_t)(3 + a) * n + m] + dk;
                    if (j >= 1 && j <= ny && k >= 1 && k <= nz)
                        lists[(size_t)k * (ny + 2) + j].push_back(m);
                }
        std::vector<int> sj, sk, so{0}, sm;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                const auto &l = lists[(size_t)k * (ny + 2) + j];
                if (l.empty()) continue;
                sj.push_back(j);
                sk.push_back(k);
                sm.insert(sm.end(), l.begin(), l.end());
                so.push_back((int)sm.size());
            }
        ne[a] = (int)sj.size();
        off_sj[a] = (int)ints.size(); ints.insert(ints.end(), sj.begin(), sj.end());
        off_sk[a] = (int)ints.size(); ints.insert(ints.end(), sk.begin(), sk.end());
        off_so[a] = (int)ints.size(); ints.insert(ints.end(), so.begin(), so.end());
        off_sm[a] = (int)ints.size(); ints.insert(ints.end(), sm.begin(), sm.end());
    }
    // the kernels take markers and list entries along the launch's y extent
    if (n > 65535 || ne[0] > 65535 || ne[1] > 65535 || ne[2] > 65535)
        return fail(NTUST_E_ARG, "%d markers, %d/%d/%d spread points: at most 65535 each", n,
                    ne[0], ne[1], ne[2]);
    const size_t nd = 3 * (size_t)n + (nx + 5) + (ny + 5) + (nz + 5);
    std::vector<double> dbl;
    dbl.reserve(nd);
    dbl.insert(dbl.end(), zm, zm + n);
    dbl.insert(dbl.end(), ym, ym + n);
    dbl.insert(dbl.end(), wm, wm + n);
    dbl.insert(dbl.end(), Xn, Xn + nx + 5);
    dbl.insert(dbl.end(), Yn, Yn + ny + 5);
    dbl.insert(dbl.end(), Zn, Zn + nz + 5);
    NEED(dd, "mk_dbl", nd, false);
    NEED(di, "mk_int", ints.size() / 2 + 1, false);
    HIPCHK(hipMemcpyAsync(dd, dbl.data(), nd * sizeof(double), hipMemcpyHostToDevice, C.st));
    HIPCHK(hipMemcpyAsync(di, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice,
                          C.st));
    TRY(sync());
    const int *ip = (const int *)di;
    MarkerSet &M = C.mk;
    M.hx = hx; M.hy = hy; M.hz = hz;
    M.zm = dd; M.ym = dd + n; M.wm = dd + 2 * n;
    M.Xn = dd + 3 * n; M.Yn = M.Xn + nx + 5; M.Zn = M.Yn + ny + 5;
    for (int a = 0; a < 3; a++) {
        M.jw[a] = ip + (size_t)a * n;
        M.kw[a] = ip + (size_t)(3 + a) * n;
        M.ne[a] = ne[a];
        M.sj[a] = ip + off_sj[a];
        M.sk[a] = ip + off_sk[a];
        M.so[a] = ip + off_so[a];
        M.sm[a] = ip + off_sm[a];
    }
    M.n = n;
    double *G = marker_gradient();
    if (!G) return fail(NTUST_E_ALLOC, "device allocation of mk_G failed");
    HIPCHK(hipMemsetAsync(G, 0, 2 * (size_t)n * sizeof(double), C.st));
    return sync();
}

int ntust_marker_velocity(double *U)
{
    NEED_INIT();
    if (C.mk.n <= 0) return fail(NTUST_E_STATE, "no markers set");
    if (!U) return fail(NTUST_E_ARG, "NULL argument");
    NEED(dU, "mk_U", 3 * (size_t)C.mk.n * C.nx, false);
    TRY(d2h(U, dU, 3 * (size_t)C.mk.n * C.nx));
    return sync();
}

int ntust_adjoint_marker_gradient(double *gz, double *gy)
{
    NEED_INIT();
    if (C.mk.n <= 0) return fail(NTUST_E_STATE, "no markers set");
    double *G = marker_gradient();
    if (!G) return fail(NTUST_E_ALLOC, "device allocation of mk_G failed");
    if (gz) TRY(d2h(gz, G, C.mk.n));
    if (gy) TRY(d2h(gy, G + C.mk.n, C.mk.n));
    return sync();
}

int ntust_adjoint_add_ke(double scale)
{
    NEED_INIT();
    AdjBufs a;
    TRY(adj_buffers(a));
    StepBufs b;
    TRY(step_buffers(b, false));
    launch_ke_gradient(C.g, b.u, b.v, b.w, scale, a.lu, a.lv, a.lw, C.st);
    HIPCHK(hipGetLastError());
    return sync();
}

int ntust_adjoint_step(int *ik)
{
    NEED_INIT();
    TRY(adjoint_ready());
    StepBufs b;
    TRY(step_buffers(b, false));
    AdjBufs a;
    TRY(adj_buffers(a));
    NEED(iu, "adj_in_u", C.nh, false);
    NEED(iv, "adj_in_v", C.nh, false);
    NEED(iw, "adj_in_w", C.nh, false);
    NEED(ip, "adj_in_p", C.nh, false);
    TRY(copy3(iu, iv, iw, b.u, b.v, b.w));
    HIPCHK(hipMemcpyAsync(ip, b.p, C.nh * sizeof(double), hipMemcpyDeviceToDevice, C.st));
    AdjForce F;
    if (C.adj_eta) {
        NEED(g1, "adj_u1", 3 * C.nh, false);
        NEED(ge, "adj_geta", C.nh, true);
        b.u1 = g1;
        b.v1 = g1 + C.nh;
        b.w1 = g1 + 2 * C.nh;
        F.u1 = b.u1;
        F.v1 = b.v1;
        F.w1 = b.w1;
        F.geta = ge;
    }
    int n = 0;
    double pm = 0.0, cd = NAN, cl = NAN;
    TRY(one_step(b, false, false, &n, &pm, &cd, &cl));
    TRY(adjoint_of_step(a, b.eta, b.filt, iu, iv, iw, n, F));
    // the state is left as it was
    TRY(copy3(b.u, b.v, b.w, iu, iv, iw));
    HIPCHK(hipMemcpyAsync(b.p, ip, C.nh * sizeof(double), hipMemcpyDeviceToDevice, C.st));
    launch_copy_shell(C.g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
    HIPCHK(hipGetLastError());
    if (ik) *ik = n;
    return sync();
}

// doubles of the per-run stores of ntust_adjoint_run for nsteps N and
// checkpoint spacing K: the checkpoints, the velocities of one segment and
// after it, the final state, and per step of a segment the line sums (drag
// or lift objective) and u1, v1, w1 with dJ/dETA (ETA gradient)
static double adjoint_store_doubles(int N, int K, bool force, bool want_eta)
{
    const double nh = (double)C.nh;
    const double pz = (double)(C.nx + 2) * (C.ny + 2), fsz = 2 * pz + 2 * (double)(C.nx + 2);
    const double nseg = (double)((N + K - 1) / K);
    double d = nseg * 4 * nh + (K + 1.0) * 3 * nh + 4 * nh;
    if (force) d += K * fsz;
    if (want_eta) d += K * 3 * nh + nh;
    return d;
}

int ntust_adjoint_checkpoint_plan(int nsteps, double reserve_bytes, int *every, double *bytes)
{
    NEED_INIT();
    if (nsteps < 1) return fail(NTUST_E_ARG, "nsteps must be >= 1");
    if (!every) return fail(NTUST_E_ARG, "NULL every");
    if (!(reserve_bytes >= 0.0) || !std::isfinite(reserve_bytes))
        return fail(NTUST_E_ARG, "reserve_bytes must be finite and >= 0");
    const bool force = C.adj_w[1] != 0.0 || C.adj_w[2] != 0.0, want_eta = C.adj_eta;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    // the run grows these stores in place (free, then allocate): what they
    // hold now is available to it
    double held = 0.0;
    for (const auto &b : C.bufs)
        if (b.name == "ckpt" || b.name == "adj_seg" || b.name == "adj_final" ||
            b.name == "adj_fseg" || b.name == "adj_useg" || b.name == "adj_geta")
            held += 8.0 * (double)b.n;
    const double avail = (double)fr + held - reserve_bytes;
    int K = nsteps;
    if (8.0 * adjoint_store_doubles(nsteps, nsteps, force, want_eta) > avail)
        K = (int)ceil(sqrt((double)nsteps));
    *every = K;
    if (bytes) *bytes = 8.0 * adjoint_store_doubles(nsteps, K, force, want_eta);
    return 0;
}

int ntust_adjoint_run(int nsteps, int checkpoint_every, int navg, double *J, int *ik,
                      double *ms)
{
    NEED_INIT();
    TRY(adjoint_ready());
    if (nsteps < 1) return fail(NTUST_E_ARG, "nsteps must be >= 1");
    if (checkpoint_every < 1) return fail(NTUST_E_ARG, "checkpoint_every must be >= 1");
    if (navg < 0 || navg > nsteps) return fail(NTUST_E_ARG, "navg outside 0..nsteps");
    const int N = nsteps, K = checkpoint_every < N ? checkpoint_every : N;
    const int nseg = (N + K - 1) / K;
    const double wke = C.adj_w[0], wcd = C.adj_w[1], wcl = C.adj_w[2];
    const double wJ = navg > 0 ? 1.0 / navg : 0.0;
    const bool force = wcd != 0.0 || wcl != 0.0, want_eta = C.adj_eta;
    StepBufs b;
    TRY(step_buffers(b, force));
    AdjBufs a;
    TRY(adj_buffers(a));
    // per step of a segment: the force integral's line sums FXz, FYz, FXy,
    // FYy (their zero borders are never written) and the projection's u1,
    // v1, w1, as the reverse step of a drag / lift objective and of the ETA
    // gradient read them (sizes as adjoint_store_doubles counts them)
    const size_t pz = (size_t)(C.nx + 2) * (C.ny + 2), fsz = 2 * pz + 2 * (size_t)(C.nx + 2);
    double *fseg = nullptr, *useg = nullptr, *geta = nullptr;
    if (force) {
        NEED(x, "adj_fseg", (size_t)K * fsz, true);
        fseg = x;
    }
    if (want_eta) {
        NEED(x, "adj_useg", (size_t)K * 3 * C.nh, false);
        NEED(y, "adj_geta", C.nh, true);
        useg = x;
        geta = y;
    }
    NEED(ck, "ckpt", (size_t)nseg * 4 * C.nh, false);
    if (nseg > C.nckpt) C.nckpt = nseg;
    // velocity at the start of every step of one segment, and after its last
    NEED(seg, "adj_seg", (size_t)(K + 1) * 3 * C.nh, false);
    NEED(fin, "adj_final", 4 * C.nh, false);
    const size_t nh = C.nh;
    auto slot = [&](int n) { return ck + (size_t)n * 4 * nh; };
    auto segv = [&](int j) { return seg + (size_t)j * 3 * nh; };
    // direct the step's line sums and u1 to slot j of the segment (j < 0:
    // line sums to slot 0, u1 not kept)
    auto bind = [&](int j) {
        if (force) {
            double *f = fseg + (size_t)(j < 0 ? 0 : j) * fsz;
            b.fxz = f;
            b.fyz = f + pz;
            b.fxy = f + 2 * pz;
            b.fyy = f + 2 * pz + (C.nx + 2);
        }
        if (want_eta) {
            double *u = j < 0 ? nullptr : useg + (size_t)j * 3 * nh;
            b.u1 = u;
            b.v1 = u ? u + nh : nullptr;
            b.w1 = u ? u + 2 * nh : nullptr;
        }
    };
    auto adj_force = [&](int n, int j) {
        AdjForce F;
        if (force && n >= N - navg) {
            const double *f = fseg + (size_t)j * fsz;
            F.sd = wJ * wcd;
            F.sl = wJ * wcl;
            F.fxz = f;
            F.fyz = f + pz;
            F.fxy = f + 2 * pz;
            F.fyy = f + 2 * pz + (C.nx + 2);
            F.lx = C.prm.lx;
        }
        if (want_eta) {
            const double *u = useg + (size_t)j * 3 * nh;
            F.u1 = u;
            F.v1 = u + nh;
            F.w1 = u + 2 * nh;
            F.geta = geta;
        }
        return F;
    };
    std::vector<int> ikf((size_t)N), ikr((size_t)N);
    const int last0 = (nseg - 1) * K;
    double Jsum = 0.0;
    double t_fwd = 0.0, t_rec = 0.0, t_adj = 0.0;
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto msec = [](std::chrono::steady_clock::time_point x, std::chrono::steady_clock::time_point y) {
        return std::chrono::duration<double, std::milli>(y - x).count();
    };

    // forward pass: checkpoints at segment starts; the last segment's
    // velocities are kept, so that segment needs no recomputation
    TRY(sync());
    auto t0 = now();
    // progress lines on stderr, one per segment, when NTUST_PROGRESS is set
    // in the environment (long horizons run for minutes inside this call)
    const bool report = getenv("NTUST_PROGRESS") != nullptr;
    for (int n = 0; n < N; n++) {
        if (report && n % K == 0 && n > 0)
            fprintf(stderr, "ntust: adjoint forward, step %d of %d (%.1f s)\n", n, N,
                    msec(t0, now()) / 1e3);
        if (n % K == 0) {
            double *d = slot(n / K);
            TRY(copy3(d, d + nh, d + 2 * nh, b.u, b.v, b.w));
            HIPCHK(hipMemcpyAsync(d + 3 * nh, b.p, nh * sizeof(double), hipMemcpyDeviceToDevice,
                                  C.st));
        }
        if (n >= last0) {
            double *d = segv(n - last0);
            TRY(copy3(d, d + nh, d + 2 * nh, b.u, b.v, b.w));
        }
        bind(n >= last0 ? n - last0 : -1);
        double pm = 0.0, cd = NAN, cl = NAN;
        const bool win = n >= N - navg;
        TRY(one_step(b, force && win, n == N - 1, &ikf[n], &pm, &cd, &cl));
        if (win) {
            if (wke != 0.0) {
                double ke = 0.0;
                TRY(kinetic_energy(b.u, b.v, b.w, &ke));
                Jsum += wke * ke;
            }
            if (force) Jsum += wcd * cd + wcl * cl;
        }
    }
    {
        double *d = segv(N - last0);
        TRY(copy3(d, d + nh, d + 2 * nh, b.u, b.v, b.w));
        TRY(copy3(fin, fin + nh, fin + 2 * nh, b.u, b.v, b.w));
        HIPCHK(hipMemcpyAsync(fin + 3 * nh, b.p, nh * sizeof(double), hipMemcpyDeviceToDevice,
                              C.st));
    }
    TRY(sync());
    auto t1 = now();
    t_fwd = msec(t0, t1);

    // reverse pass
    double *clear[7] = {a.lu, a.lv, a.lw, a.lp, a.gu, a.gv, a.gw};
    for (double *x : clear) HIPCHK(hipMemsetAsync(x, 0, nh * sizeof(double), C.st));
    if (geta) HIPCHK(hipMemsetAsync(geta, 0, nh * sizeof(double), C.st));
    if (C.mk.n > 0) {
        double *G = marker_gradient();
        if (!G) return fail(NTUST_E_ALLOC, "device allocation of mk_G failed");
        HIPCHK(hipMemsetAsync(G, 0, 2 * (size_t)C.mk.n * sizeof(double), C.st));
    }
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int s0 = sg * K, s1 = s0 + K < N ? s0 + K : N;
        if (sg != nseg - 1) {
            auto r0 = now();
            const double *d = slot(sg);
            TRY(copy3(b.u, b.v, b.w, d, d + nh, d + 2 * nh));
            HIPCHK(hipMemcpyAsync(b.p, d + 3 * nh, nh * sizeof(double), hipMemcpyDeviceToDevice,
                                  C.st));
            launch_copy_shell(C.g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
            for (int n = s0; n < s1; n++) {
                double *e = segv(n - s0);
                TRY(copy3(e, e + nh, e + 2 * nh, b.u, b.v, b.w));
                bind(n - s0);
                double pm = 0.0, cd = NAN, cl = NAN;
                TRY(one_step(b, force && n >= N - navg, false, &ikr[n], &pm, &cd, &cl));
                if (ikr[n] != ikf[n])
                    return fail(NTUST_E_STATE,
                                "step %d recomputed from its checkpoint ran %d SOR sweeps, "
                                "the forward pass %d", n, ikr[n], ikf[n]);
            }
            double *e = segv(s1 - s0);
            TRY(copy3(e, e + nh, e + 2 * nh, b.u, b.v, b.w));
            TRY(sync());
            t_rec += msec(r0, now());
        }
        if (report)
            fprintf(stderr, "ntust: adjoint reverse, segment %d of %d (steps %d-%d, %.1f s)\n",
                    nseg - sg, nseg, s0, s1 - 1, msec(t0, now()) / 1e3);
        auto a0 = now();
        for (int n = s1 - 1; n >= s0; n--) {
            if (n >= N - navg && wke != 0.0) {
                const double *e = segv(n - s0 + 1);
                launch_ke_gradient(C.g, e, e + nh, e + 2 * nh, wJ * wke, a.lu, a.lv, a.lw, C.st);
            }
            const double *e = segv(n - s0);
            TRY(adjoint_of_step(a, b.eta, b.filt, e, e + nh, e + 2 * nh, ikf[n],
                                adj_force(n, n - s0)));
        }
        TRY(sync());
        t_adj += msec(a0, now());
    }
    // leave the final forward state on the device
    TRY(copy3(b.u, b.v, b.w, fin, fin + nh, fin + 2 * nh));
    HIPCHK(hipMemcpyAsync(b.p, fin + 3 * nh, nh * sizeof(double), hipMemcpyDeviceToDevice, C.st));
    launch_copy_shell(C.g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
    HIPCHK(hipGetLastError());
    TRY(sync());
    if (J) *J = Jsum * wJ;
    if (ik) for (int n = 0; n < N; n++) ik[n] = ikf[n];
    if (ms) {
        ms[0] = t_fwd;
        ms[1] = t_rec + t_adj;
        ms[2] = t_rec;
        ms[3] = t_adj;
    }
    return 0;
}

void *ntust_stream(void) { return (void *)C.st; }

int ntust_prof_reset(void)
{
    C.prof_ms = 0.0;
    C.prof_launches = 0;
    C.prof_sweeps = 0;
    return 0;
}

int ntust_set_arithmetic(int mode)
{
    if (mode != NTUST_ARITH_EXACT && mode != NTUST_ARITH_FAST && mode != NTUST_ARITH_FP32)
        return fail(NTUST_E_ARG, "unknown arithmetic mode %d", mode);
    C.arith = mode;
    return 0;
}

int ntust_prof_kernel(double *bytes_per_cell)
{
    // algorithmic bytes per grid cell of one timed launch: per sweep, read
    // and write p and read b (24 B FP64); FP32: read and write e, read r on
    // the colour-0 cells only (colour 1 runs shifted, with no right-hand
    // side): 8 + 2 = 10 B.  A two-pass launch is half a sweep.
    if (C.prm.sor == NTUST_SOR_LEXICOGRAPHIC) {
        // a lexicographic step launch advances about half the cells of the
        // grid by one sweep: 24 B per cell update, averaged over the timed
        // launches (the first and last steps of a batch touch fewer cells)
        if (bytes_per_cell)
            *bytes_per_cell = C.prof_launches ? 24.0 * (double)C.prof_sweeps /
                                                    (double)C.prof_launches
                                              : 12.0;
        return 6;
    }
    const bool mixed = C.prm.sor == NTUST_SOR_MIXED;
    if (bytes_per_cell) *bytes_per_cell = (mixed ? 10.0 : 24.0) / 2.0;
    return mixed ? 2 : 0;
}

int ntust_prof_read(double *ms, long long *launches)
{
    if (ms) *ms = C.prof_ms;
    if (launches) *launches = C.prof_launches;
    return 0;
}

int ntust_prof_copy(double nbytes, double *ms)
{
    NEED_INIT();
    if (!(nbytes >= 16.0) || nbytes > 64.0e9)
        return fail(NTUST_E_ARG, "copy size %g bytes out of range", nbytes);
    const long words = (long)(nbytes / 16.0);
    double *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, words * 16));
    if (hipMalloc(&b, words * 16) != hipSuccess) {
        hipFree(a);
        return fail(NTUST_E_ALLOC, "device allocation of the copy buffers failed");
    }
    int rc = 0;
    float t = 0.f;
    // the source filled by a kernel of its own first, so the timed copy
    // reads memory that is not in flight
    if (hipMemsetAsync(a, 0, words * 16, C.st) != hipSuccess ||
        hipEventRecord(C.ev0, C.st) != hipSuccess) rc = -1;
    if (rc == 0) launch_stream_copy(a, b, words, C.st);
    if (rc == 0 && (hipEventRecord(C.ev1, C.st) != hipSuccess ||
                    hipStreamSynchronize(C.st) != hipSuccess ||
                    hipEventElapsedTime(&t, C.ev0, C.ev1) != hipSuccess ||
                    hipGetLastError() != hipSuccess))
        rc = -1;
    hipFree(a);
    hipFree(b);
    if (rc) return fail(NTUST_E_HIP, "calibration copy failed");
    if (ms) *ms = t;
    return 0;
}

// ---- output, steady check, restart --------------------------------------

int ntust_output_q(const char *path, float *qout)
{
    NEED_INIT();
    StepBufs b;
    TRY(step_buffers(b, false));
    const size_t nq = 5 * (size_t)C.ni;
    NEED(qd, "qout", (nq + 1) / 2, false);
    float *q = (float *)qd;
    launch_pack_q(C.g, b.pre, b.uc, b.vc, b.wc, b.eta, q, C.st);
    HIPCHK(hipGetLastError());
    std::vector<float> own;
    float *h = qout;
    if (!h) {
        own.resize(nq);
        h = own.data();
    }
    HIPCHK(hipMemcpyAsync(h, q, nq * sizeof(float), hipMemcpyDeviceToHost, C.st));
    TRY(sync());
    if (!path) return 0;
    FILE *f = fopen(path, "ab");   // position='append' (filer.f90:177)
    if (!f) return fail(NTUST_E_IO, "cannot open %s for appending", path);
    const int32_t nb = 1, dims[3] = {C.nx, C.ny, C.nz};
    const float temp[4] = {1.0f, 1.0f, 1.0f, 1.0f};   // variables_module.f90:363
    bool ok = put_record(f, &nb, 4) && put_record(f, dims, 12) && put_record(f, temp, 16) &&
              put_record(f, h, nq * sizeof(float));
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(NTUST_E_IO, "writing %s failed (a record over 2 GiB?)", path);
    return 0;
}

int ntust_read_q(const char *path, float *qout)
{
    NEED_INIT();
    if (!path || !qout) return fail(NTUST_E_ARG, "NULL argument");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(NTUST_E_IO, "cannot open %s", path);
    int32_t nb = 0, dims[3] = {0, 0, 0};
    float temp[4];
    bool ok = get_record(f, &nb, 4) && get_record(f, dims, 12);
    if (ok && (dims[0] != C.nx || dims[1] != C.ny || dims[2] != C.nz)) {
        fclose(f);
        return fail(NTUST_E_ARG, "%s holds a %dx%dx%d grid, the library %dx%dx%d", path,
                    dims[0], dims[1], dims[2], C.nx, C.ny, C.nz);
    }
    ok = ok && get_record(f, temp, 16) &&
         get_record(f, qout, 5 * (size_t)C.ni * sizeof(float));
    fclose(f);
    if (!ok) return fail(NTUST_E_IO, "%s is not a one-block PLOT3D solution file", path);
    return 0;
}

int ntust_reading_variables(const float *qout)
{
    NEED_INIT();
    if (!qout) return fail(NTUST_E_ARG, "NULL Qout");
    StepBufs b;
    TRY(step_buffers(b, false));
    const size_t nq = 5 * (size_t)C.ni;
    NEED(qd, "qout", (nq + 1) / 2, false);
    float *q = (float *)qd;
    HIPCHK(hipMemcpyAsync(q, qout, nq * sizeof(float), hipMemcpyHostToDevice, C.st));
    launch_unpack_q(C.g, q, b.u, b.v, b.w, b.p, b.pre, b.uc, b.vc, b.wc, C.st);
    // main.f90:100, the call that follows reading_variables
    launch_boundary_conditions(C.g, C.bc, b.u, b.v, b.w, b.p, C.st);
    launch_copy_shell(C.g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
    HIPCHK(hipGetLastError());
    return sync();
}

int ntust_write_mesh_x(const char *path, const double *Xs, const double *Ys, const double *Zs)
{
    NEED_INIT();
    if (!path || !Xs || !Ys || !Zs) return fail(NTUST_E_ARG, "NULL argument");
    const size_t ni = (size_t)C.ni;
    std::vector<float> xyz(3 * ni);
    size_t c = 0;
    for (int k = 0; k < C.nz; k++)
        for (int j = 0; j < C.ny; j++)
            for (int i = 0; i < C.nx; i++, c++) {
                xyz[c] = (float)Xs[i];
                xyz[c + ni] = (float)Ys[j];
                xyz[c + 2 * ni] = (float)Zs[k];
            }
    FILE *f = fopen(path, "wb");
    if (!f) return fail(NTUST_E_IO, "cannot open %s", path);
    const int32_t nb = 1, dims[3] = {C.nx, C.ny, C.nz};
    bool ok = put_record(f, &nb, 4) && put_record(f, dims, 12) &&
              put_record(f, xyz.data(), xyz.size() * sizeof(float));
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(NTUST_E_IO, "writing %s failed (a record over 2 GiB?)", path);
    return 0;
}

int ntust_check_steady(int istep, int isto, double *velocity_difference)
{
    NEED_INIT();
    if (!velocity_difference) return fail(NTUST_E_ARG, "NULL velocity_difference");
    StepBufs b;
    TRY(step_buffers(b, false));
    NEED(last, "last_velocity", (size_t)C.ni, true);
    NEED(vm, "steady_max", 1, false);
    const int compare = istep != isto;   // check_steady.f90:6
    unsigned long long *vmax = (unsigned long long *)vm;
    if (compare) HIPCHK(hipMemsetAsync(vmax, 0, sizeof *vmax, C.st));
    launch_steady(C.g, b.wc, last, compare, vmax, C.st);
    HIPCHK(hipGetLastError());
    double m = 0.0;
    if (compare) HIPCHK(hipMemcpyAsync(&m, vmax, sizeof m, hipMemcpyDeviceToHost, C.st));
    TRY(sync());
    if (compare) *velocity_difference = m;
    return 0;
}

int ntust_restart_write(const char *path, int istep, double time)
{
    NEED_INIT();
    if (!path) return fail(NTUST_E_ARG, "NULL path");
    StepBufs b;
    TRY(step_buffers(b, false));
    NEED(last, "last_velocity", (size_t)C.ni, true);
    (void)last;
    RestartHeader hd{};
    memcpy(hd.magic, "NTUSTRS1", 8);
    hd.version = 1;
    hd.nx = C.nx;
    hd.ny = C.ny;
    hd.nz = C.nz;
    hd.istep = istep;
    hd.narrays = N_RESTART;
    hd.time = time;
    hd.hash = FNV_OFFSET;
    FILE *f = fopen(path, "wb");
    if (!f) return fail(NTUST_E_IO, "cannot open %s", path);
    bool ok = fwrite(&hd, sizeof hd, 1, f) == 1;   // rewritten with the hash below
    std::vector<double> h((size_t)C.nh);
    for (int a = 0; a < N_RESTART && ok; a++) {
        const size_t n = restart_arrays[a].halo ? (size_t)C.nh : (size_t)C.ni;
        const double *d = buf(restart_arrays[a].name, n, false);
        if (!d) {
            fclose(f);
            return fail(NTUST_E_ALLOC, "device buffer %s", restart_arrays[a].name);
        }
        if (d2h(h.data(), d, n) < 0 || sync() < 0) {
            fclose(f);
            return NTUST_E_HIP;
        }
        hd.hash = fnv_words(hd.hash, h.data(), n);
        hd.payload_bytes += n * sizeof(double);
        ok = fwrite(h.data(), sizeof(double), n, f) == n;
    }
    ok = ok && fseek(f, 0, SEEK_SET) == 0 && fwrite(&hd, sizeof hd, 1, f) == 1;
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(NTUST_E_IO, "writing %s failed", path);
    return 0;
}

int ntust_restart_read(const char *path, int *istep, double *time)
{
    NEED_INIT();
    if (!path) return fail(NTUST_E_ARG, "NULL path");
    FILE *f = fopen(path, "rb");
    if (!f) return fail(NTUST_E_IO, "cannot open %s", path);
    RestartHeader hd{};
    const size_t want = (size_t)5 * C.nh + (size_t)5 * C.ni;
    if (fread(&hd, sizeof hd, 1, f) != 1 || memcmp(hd.magic, "NTUSTRS1", 8) != 0 ||
        hd.version != 1 || hd.narrays != N_RESTART) {
        fclose(f);
        return fail(NTUST_E_IO, "%s is not a version-1 restart file", path);
    }
    if (hd.nx != C.nx || hd.ny != C.ny || hd.nz != C.nz) {
        fclose(f);
        return fail(NTUST_E_ARG, "%s holds a %dx%dx%d grid, the library %dx%dx%d", path, hd.nx,
                    hd.ny, hd.nz, C.nx, C.ny, C.nz);
    }
    if (hd.payload_bytes != want * sizeof(double)) {
        fclose(f);
        return fail(NTUST_E_IO, "%s: payload of %llu bytes, expected %zu", path,
                    (unsigned long long)hd.payload_bytes, want * sizeof(double));
    }
    // first pass: the whole payload must hash right before the state changes
    std::vector<double> h((size_t)C.nh);
    uint64_t hash = FNV_OFFSET;
    bool ok = true;
    for (int a = 0; a < N_RESTART && ok; a++) {
        const size_t n = restart_arrays[a].halo ? (size_t)C.nh : (size_t)C.ni;
        ok = fread(h.data(), sizeof(double), n, f) == n;
        if (ok) hash = fnv_words(hash, h.data(), n);
    }
    if (!ok || hash != hd.hash) {
        fclose(f);
        return fail(NTUST_E_IO, "%s: %s", path, ok ? "payload hash mismatch" : "file truncated");
    }
    StepBufs b;
    TRY(step_buffers(b, false));
    NEED(last, "last_velocity", (size_t)C.ni, true);
    (void)last;
    ok = fseek(f, (long)sizeof hd, SEEK_SET) == 0;
    for (int a = 0; a < N_RESTART && ok; a++) {
        const size_t n = restart_arrays[a].halo ? (size_t)C.nh : (size_t)C.ni;
        double *d = buf(restart_arrays[a].name, n, false);
        ok = d && fread(h.data(), sizeof(double), n, f) == n;
        if (ok && (h2d(d, h.data(), n) < 0 || sync() < 0)) {
            fclose(f);
            return NTUST_E_HIP;
        }
    }
    fclose(f);
    if (!ok) return fail(NTUST_E_IO, "%s: reading the payload failed", path);
    // the copies of u, v, w the step keeps (as ntust_state_upload)
    launch_copy_shell(C.g, b.u, b.v, b.w, b.us, b.vs, b.ws, b.u2, b.v2, b.w2, C.st);
    HIPCHK(hipGetLastError());
    TRY(sync());
    if (istep) *istep = hd.istep;
    if (time) *time = hd.time;
    return 0;
}

// ---- Krylov path ----------------------------------------------------------

// jcoef(2..m) as offsets, checked against n
static int dia_offsets(const int *jcoef, int n, int m, long col[3])
{
    if (!jcoef) return fail(NTUST_E_ARG, "NULL jcoef");
    if (m < 1 || m > 4) return fail(NTUST_E_ARG, "m = %d: the DIA layout has 1..4 offsets", m);
    for (int d = 0; d < 3; d++) {
        col[d] = d + 2 <= m ? jcoef[d + 1] : 0;
        if (d + 2 <= m && (col[d] < 1 || col[d] >= n))
            return fail(NTUST_E_ARG, "jcoef(%d) = %ld outside 1..n-1", d + 2, col[d]);
    }
    return 0;
}

int ntust_matmul_ell(double *x, const double *coef, const int *jcoef, const double *y, int n,
                     int m)
{
    NEED_INIT();
    if (!x || !coef || !y || n < 1) return fail(NTUST_E_ARG, "NULL array or n < 1");
    long col[3];
    TRY(dia_offsets(jcoef, n, m, col));
    NEED(dc, "kry_coef", 7 * (size_t)n, false);
    NEED(dy, "kry_y", (size_t)n, false);
    NEED(dx, "kry_x", (size_t)n, false);
    TRY(h2d(dc, coef, 7 * (size_t)n));
    TRY(h2d(dy, y, (size_t)n));
    launch_dia_matvec(dc, col, m, n, dy, dx, C.st);
    HIPCHK(hipGetLastError());
    TRY(d2h(x, dx, (size_t)n));
    return sync();
}

int ntust_solver_bicg3(const double *coef, const int *jcoef, const double *b, double *x, int n,
                       int m, double eps, int itmax, int *iterations, double *norm)
{
    NEED_INIT();
    if (!coef || !b || !x || n < 1) return fail(NTUST_E_ARG, "NULL array or n < 1");
    long col[3];
    TRY(dia_offsets(jcoef, n, m, col));
    NEED(dc, "kry_coef", 7 * (size_t)n, false);
    NEED(db, "kry_b", (size_t)n, false);
    NEED(dx, "kry_x", (size_t)n, false);
    BicgVectors w;
    double **vec[6] = {&w.p, &w.r, &w.r2, &w.v, &w.ss, &w.t};
    const char *nm[6] = {"kry_p", "kry_r", "kry_r2", "kry_v", "kry_s", "kry_t"};
    for (int q = 0; q < 6; q++) {
        *vec[q] = buf(nm[q], (size_t)n, false);
        if (!*vec[q]) return fail(NTUST_E_ALLOC, "device allocation of %s failed", nm[q]);
    }
    NEED(part, "kry_part", 2 * (size_t)KRY_MAX_BLOCKS, false);
    NEED(sd, "kry_state", (sizeof(BicgState) + 7) / 8, false);
    w.part = part;
    BicgState *st = (BicgState *)sd;
    BicgState h{};
    h.eps = eps;
    h.itmax = itmax;
    TRY(h2d(dc, coef, 7 * (size_t)n));
    TRY(h2d(db, b, (size_t)n));
    TRY(h2d(dx, x, (size_t)n));
    HIPCHK(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, C.st));
    launch_bicg_setup(dc, col, m, n, db, dx, w, st, C.st);
    HIPCHK(hipGetLastError());
    // iterations in batches; the kernels of iterations past the stopping
    // rule return at once
    HIPCHK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, C.st));
    TRY(sync());
    int batch = 8;
    while (!h.stop) {
        for (int q = 0; q < batch; q++) launch_bicg_iteration(dc, col, m, n, dx, w, st, C.st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&h, st, sizeof h, hipMemcpyDeviceToHost, C.st));
        TRY(sync());
        batch = batch < 64 ? 2 * batch : 64;
    }
    TRY(d2h(x, dx, (size_t)n));
    TRY(sync());
    if (iterations) *iterations = h.j;
    if (norm) *norm = h.norm;
    return 0;
}

// ---- airfoil geometry -------------------------------------------------------

int ntust_ray_casting(const double *az0, const double *ay0, int poly, double AOA, double zc,
                      double yc, const double *Yn, const double *Zn, int nsub, double *ETA)
{
    NEED_INIT();
    if (!az0 || !ay0 || !Yn || !Zn) return fail(NTUST_E_ARG, "NULL argument");
    if (poly < 2) return fail(NTUST_E_ARG, "the polygon needs at least 2 vertices (%d)", poly);
    if (nsub < 1 || nsub > RAY_NSUB_MAX)
        return fail(NTUST_E_ARG, "nsub %d outside 1..%d", nsub, RAY_NSUB_MAX);
    // airfoil.f90:597-611 on the host (poly vertices): quarter-chord shift,
    // rotation by -AOA degrees in which ay uses the rotated az, placement
    std::vector<double> az((size_t)poly), ay((size_t)poly);
    const double PI = 4.0 * atan(1.0);
    for (int i = 0; i < poly; i++) {
        double z = az0[i] - 0.25, y = ay0[i];
        z = cos((-AOA) * PI / 180) * z - sin((-AOA) * PI / 180) * y;
        y = sin((-AOA) * PI / 180) * z + cos((-AOA) * PI / 180) * y;
        az[i] = (z + zc) + 0.25;
        ay[i] = y + yc;
    }
    // the bounding-box search of airfoil.f90:614-654 over nodes 1..n
    double ymin = ay[0], ymax = ay[0], zmin = az[0], zmax = az[0];
    for (int i = 1; i < poly; i++) {
        ymin = ay[i] < ymin ? ay[i] : ymin;
        ymax = ay[i] > ymax ? ay[i] : ymax;
        zmin = az[i] < zmin ? az[i] : zmin;
        zmax = az[i] > zmax ? az[i] : zmax;
    }
    // a bound not found keeps the module variable's first value, 0; a
    // polygon wholly outside the grid then gives ETA = 0 (the reference's
    // unit cavity, where the airfoil sits at (4, 4))
    int A = -1, B = -1, Cb = -1, E = -1;
    for (int j = 1; j <= C.ny && A < 0; j++) if (Yn[j + 1] > ymin) A = j - 1;
    for (int j = 1; j <= C.ny && B < 0; j++) if (Yn[j + 1] > ymax) B = j;
    for (int k = 1; k <= C.nz && Cb < 0; k++) if (Zn[k + 1] > zmin) Cb = k - 1;
    for (int k = 1; k <= C.nz && E < 0; k++) if (Zn[k + 1] > zmax) E = k;
    A = A < 0 ? 0 : A;
    B = B < 0 ? 0 : B;
    Cb = Cb < 0 ? 0 : Cb;
    E = E < 0 ? 0 : E;
    const bool boxed = E - 1 >= Cb && B - 1 >= A;
    if (boxed && (A < 1 || Cb < 1))
        return fail(NTUST_E_ARG,
                    "the airfoil (z %g..%g, y %g..%g) straddles the lower edge of the grid: "
                    "the reference's box loops would index outside its arrays",
                    zmin, zmax, ymin, ymax);
    StepBufs b;
    TRY(step_buffers(b, false));
    const size_t n2 = (size_t)C.nz * C.ny;
    NEED(daz, "ray_az", (size_t)poly, false);
    NEED(day, "ray_ay", (size_t)poly, false);
    NEED(dy, "ray_Y", (size_t)C.ny + 5, false);
    NEED(dz, "ray_Z", (size_t)C.nz + 5, false);
    NEED(e1, "ray_eta1", n2, false);
    NEED(pt, "ray_points", n2 / 2 + 1, false);
    NEED(cn, "ray_counts", n2 / 2 + 1, false);
    TRY(h2d(daz, az.data(), (size_t)poly));
    TRY(h2d(day, ay.data(), (size_t)poly));
    TRY(h2d(dy, Yn, (size_t)C.ny + 5));
    TRY(h2d(dz, Zn, (size_t)C.nz + 5));
    HIPCHK(hipMemsetAsync(pt, 0, n2 * sizeof(int), C.st));
    HIPCHK(hipMemsetAsync(cn, 0, n2 * sizeof(int), C.st));
    if (ETA) TRY(h2d(b.eta, ETA, C.nh));
    // an empty box: no vertex cell, no sub-grid (k_ray_subcells launches none)
    RayGeom r{C.nx, C.ny, C.nz, poly, A, boxed ? B : A, Cb, boxed ? E : Cb, nsub, dy, dz, daz,
              day};
    launch_ray_casting(C.g, r, (int *)pt, (int *)cn, e1, b.eta, C.st);
    HIPCHK(hipGetLastError());
    if (ETA) TRY(d2h(ETA, b.eta, C.nh));
    // the host vectors stay alive until the copies have run
    return sync();
}

int ntust_dynamic_aoa(double AOA1, double time, double dt, double StartDynamic, double *AOA)
{
    if (!AOA) return fail(NTUST_E_ARG, "NULL AOA");
    // airfoil.f90:311-321: PI is the default-real literal 3.14159265359
    // widened to real*8; reduce_frequency = 2.0 (variables_module.f90:148)
    const double PI = (double)3.14159265359f, reduce_frequency = 2.0;
    const double frequency = reduce_frequency / PI;
    *AOA = AOA1 + 10 * fabs(sin(2 * PI * frequency * (time - dt * (StartDynamic - 1))));
    return 0;
}

}  // extern "C"
