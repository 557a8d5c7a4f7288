// qed2d_api.cpp -- context management, solver / MD / HMC orchestration and
// the C ABI of libqed2d_hip (declared in include/qed2d.h).
//
// The host side only sequences kernels on the context's stream.  Solvers
// keep their scalars on the device (SolverState) and the host looks at them
// once per chunk of iterations, so a CG iteration costs two kernel launches
// and no synchronisation.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/qed2d.h"
#include "qed2d_internal.h"

using namespace qed2d;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
enum SpinorSlot {
    S_R, S_P0, S_P1, S_T, S_TMP, S_TMP2, S_UCP, S_X, S_F, S_G5DX, S_AUX, S_PB, S_PBA, S_PBL,
    S_RW_X, S_RW_Z, S_RW_IN, S_RW_T1, S_RW_T2, S_UPP, S_NSLOTS
};
enum FloatSlot { FS_R, FS_P0, FS_P1, FS_T, FS_X, FS_U, FS_NSLOTS };
enum RealSlot { R_MOM, R_FACC, R_MPP, R_NSLOTS };
constexpr int kRing = 4;  // state copies of the chunks of a pipelined solve

struct qed2d_ctx {
    Latt L;
    double beta = 0.0;
    int device = 0;
    hipStream_t stream = nullptr;
    int sync = 1;
    // Rows marched per thread (the tile height) of each kernel family; the
    // defaults come from default_tiles(), qed2d_ctx_autotune measures them.
    int rows = 16;    // FP64 solver kernels (CG, multi-shift, residuals)
    int rows32 = 8;   // FP32 solver kernels (mixed-precision CG)
    int rows_k2 = 0;    // the second CG kernel (K2) of the FP64 CG; 0: rows
    int rows32_k2 = 0;  // K2 of the mixed-precision CG; 0: rows32
    int rows_q = 2;   // FP64 Q / Dw applies
    int rows_q32 = 2; // FP32 Q apply of the public entry point
    double ru_delta = kRuDelta;  // mixed CG: reliable update after |r| drops by this factor
    int ru_mode = 1;  // mixed CG: 1 reliable updates decided on the device, 0 by the host
    double last_dh = 0.0;    // dH of the last trajectory, from double-double energies
    std::string err;
    double2 *spin[S_NSLOTS] = {};
    float2 *fspin[FS_NSLOTS] = {};
    double *real[R_NSLOTS] = {};
    std::vector<double2 *> flav;  // RHMC pseudofermions, one per flavour
    std::vector<double2 *> shiftbuf;  // multi-shift CG: internal x_i
    std::vector<double2 *> psbuf;     // multi-shift CG: search directions p_i
    // multi-shift CG of several right-hand sides in one launch per kernel (the
    // pseudofermions of an RHMC force): a chain set of nf copies of this
    // lattice sharing the links
    struct {
        int cap = 0, n = 0, rows = 0;
        BChain *dch = nullptr;
        int *blk0 = nullptr;
        int4 *shape = nullptr;
        MultiShiftState *M = nullptr, *hM = nullptr;
        double *part = nullptr;
        int *cnt = nullptr;
        BParts P{};
        double2 *arena = nullptr;  // r, p0, p1, t of every right-hand side
        std::vector<std::vector<double2 *>> x, ps;  // per right-hand side and shift
        // the mixed-precision force solves: FP32 multi-shift x_i and p_i per
        // right-hand side and shift; the FP64 solutions and right-hand sides of
        // every pole, pole after pole (pcap poles)
        std::vector<std::vector<float2 *>> x32, ps32;
        double2 *px = nullptr, *psrc = nullptr;
        int pcap = 0;
        // the force of every pole solution: Y = Q x (yarena, ycap fields) and a
        // table of x, Y, 2 + m0 and rho per field (device and pinned copies)
        double2 *yarena = nullptr;
        int ycap = 0;
        void *tab = nullptr, *htab = nullptr;
        int tabcap = 0;
    } mr;
    // right-hand-side sets (copies of this lattice sharing the links), indexed by
    // their count: qed2d_cg_mrhs and the RHMC heatbath's shifted solves
    std::vector<qed2d_batch *> rhsets;
    int multishift = 1;               // MultiCG / RHMC force through multi-shift CG
    int rhmc_prec = QED2D_PREC_F64;   // the RHMC force's solves (qed2d_ctx_set_rhmc_precision)
    MultiShiftState *ms = nullptr;    // device
    MultiShiftState *hms = nullptr;   // pinned
    double *part = nullptr;  // partial sums of the generic reductions
    long part_len = 0;
    double *cgpart = nullptr;  // partial sums of the solver kernels
    CGPartials P{};
    SolverState *state = nullptr;
    RedOut *red = nullptr;
    SolverState *hstate = nullptr;  // pinned
    // pipelined solves: copies of the state at the end of queued chunks in
    // host-coherent pinned memory, each with the sequence number written after it
    SolverState *hring = nullptr;
    int *hseq = nullptr;
    int ring_next = 0;
    int seq_no = 0;
    RedOut *hred = nullptr;         // pinned
    // A Q apply (so = Q si) that the CG solvers queue behind each chunk
    // predicted to end the solve, run on the device iff the solve has then
    // converged (the Nf=2 force's Q X, which then needs no host round trip);
    // tail_ran: whether it ran for the last solve.  Unset: U == nullptr.
    struct SolveTail {
        const double *U = nullptr;
        const double *si = nullptr;
        double *so = nullptr;
        double am0 = 0.0;
    } tail;
    bool tail_ran = false;
    // kernel-class timing (qed2d_profile_*)
    int prof = 0;  // 0 off, N: time every N-th launch of each class
    long prof_seen[QED2D_PROF_NCLASS] = {};
    std::vector<hipEvent_t> evpool;
    struct Pending { int cls; hipEvent_t a, b; };
    std::vector<Pending> pending;
    double prof_ms[QED2D_PROF_NCLASS] = {};
    long prof_n[QED2D_PROF_NCLASS] = {};
};

namespace qed2d {
thread_local ProfArm g_prof_arm{nullptr, nullptr};
}
// why the last qed2d_ctx_create of this thread failed (qed2d_last_error(NULL))
static thread_local std::string g_create_err = "null context";

namespace {

int fail(qed2d_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((c), QED2D_EHIP, "%s: %s", #call, hipGetErrorString(e_));         \
    } while (0)

int launched(qed2d_ctx *c)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, QED2D_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return QED2D_OK;
}

void prof_harvest(qed2d_ctx *c);

int finish(qed2d_ctx *c)
{
    int rc = launched(c);
    if (rc) return rc;
    if (c->sync) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        prof_harvest(c);
    }
    return QED2D_OK;
}

long spinor_elems(const qed2d_ctx *c) { return 2 * c->L.V; }  // complex numbers

int get_spin(qed2d_ctx *c, int slot, double2 **out)
{
    if (!c->spin[slot]) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(double2) * spinor_elems(c)) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of a %ld-byte spinor failed",
                        (long)(sizeof(double2) * spinor_elems(c)));
        c->spin[slot] = (double2 *)p;
    }
    *out = c->spin[slot];
    return QED2D_OK;
}

int get_fspin(qed2d_ctx *c, int slot, float2 **out)
{
    if (!c->fspin[slot]) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(float2) * spinor_elems(c)) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of an FP32 spinor failed");
        c->fspin[slot] = (float2 *)p;
    }
    *out = c->fspin[slot];
    return QED2D_OK;
}

int get_real(qed2d_ctx *c, int slot, double **out)
{
    if (!c->real[slot]) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(double) * 2 * c->L.V) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of a real field failed");
        c->real[slot] = (double *)p;
    }
    *out = c->real[slot];
    return QED2D_OK;
}

#define TRY(x)                      \
    do {                            \
        int rc_ = (x);              \
        if (rc_ != QED2D_OK) return rc_; \
    } while (0)

// ---- kernel-class timing ---------------------------------------------------
hipEvent_t prof_event(qed2d_ctx *c)
{
    if (c->evpool.empty()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
}

// Opens a timed region; returns the start event (or null when this launch is
// not sampled).
hipEvent_t prof_begin(qed2d_ctx *c, int cls = QED2D_PROF_OTHER)
{
    if (!c->prof) return nullptr;
    if ((c->prof_seen[cls]++ % c->prof) != 0) return nullptr;
    hipEvent_t a = prof_event(c);
    if (a && hipEventRecord(a, c->stream) != hipSuccess) return nullptr;
    return a;
}

void prof_end(qed2d_ctx *c, int cls, hipEvent_t a)
{
    if (!c->prof || !a) return;
    hipEvent_t b = prof_event(c);
    if (!b || hipEventRecord(b, c->stream) != hipSuccess) return;
    c->pending.push_back({cls, a, b});
}

// Called after the stream has been synchronised: fold finished timings in.
void prof_harvest(qed2d_ctx *c)
{
    for (auto &p : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->prof_ms[p.cls] += ms;
            c->prof_n[p.cls] += 1;
        }
        c->evpool.push_back(p.a);
        c->evpool.push_back(p.b);
    }
    c->pending.clear();
}

// Arms the kernel-exact timing of the next launch of class cls (when this
// launch is sampled); returns false otherwise.
bool prof_arm(qed2d_ctx *c, int cls, hipEvent_t *a, hipEvent_t *b)
{
    if (!c->prof) return false;
    if ((c->prof_seen[cls]++ % c->prof) != 0) return false;
    *a = prof_event(c);
    *b = prof_event(c);
    if (!*a || !*b) {
        if (*a) c->evpool.push_back(*a);
        if (*b) c->evpool.push_back(*b);
        return false;
    }
    g_prof_arm = ProfArm{*a, *b};
    return true;
}

// After the armed launch: keeps the events for harvesting if a kernel took
// them, returns them to the pool otherwise.
void prof_fire(qed2d_ctx *c, int cls, hipEvent_t a, hipEvent_t b)
{
    if (g_prof_arm.a == a) {
        g_prof_arm = ProfArm{nullptr, nullptr};
        c->evpool.push_back(a);
        c->evpool.push_back(b);
        return;
    }
    c->pending.push_back({cls, a, b});
}

// Times one kernel of class cls by its execution (start and end events of
// the first kernel the launch makes).
#define PROF(c, cls, launch)                              \
    do {                                                  \
        hipEvent_t pa_ = nullptr, pb_ = nullptr;          \
        if (prof_arm((c), (cls), &pa_, &pb_)) {           \
            launch;                                       \
            prof_fire((c), (cls), pa_, pb_);              \
        } else {                                          \
            launch;                                       \
        }                                                 \
    } while (0)

int read_state(qed2d_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->hstate, c->state, sizeof(SolverState), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_harvest(c);
    return QED2D_OK;
}

int read_red(qed2d_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->hred, c->red, sizeof(RedOut), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_harvest(c);
    return QED2D_OK;
}

// Host double-double helpers (same TwoSum as the device).
struct hdd {
    double hi = 0.0, lo = 0.0;
    void add(double b)
    {
        dd a{hi, lo};
        a = dd_add(a, b);
        hi = a.hi;
        lo = a.lo;
    }
    void add(double bh, double bl)
    {
        dd a{hi, lo}, b{bh, bl};
        a = dd_add(a, b);
        hi = a.hi;
        lo = a.lo;
    }
    double value() const { return hi + lo; }
};

// a - b of two double-doubles, rounded once: dH resolved below the ulp of H.
double dd_diff(const hdd &a, const hdd &b)
{
    dd x{a.hi, a.lo}, y{-b.hi, -b.lo};
    x = dd_add(x, y);
    return x.hi + x.lo;
}

// ---------------------------------------------------------------------------
// elementary operators
// ---------------------------------------------------------------------------
void q64(qed2d_ctx *c, const double *U, const double *si, double *so, double am0)
{
    PROF(c, QED2D_PROF_Q_F64,
         launch_q_apply_f64(c->L, (const double2 *)U, (const double2 *)si, (double2 *)so, am0,
                            c->rows_q, c->stream));
}

// so = Q^2 si (+ mu^2 si), tmp = Q si
void qsq64(qed2d_ctx *c, const double *U, const double *si, double *so, double *tmp, double am0,
           double mu2)
{
    q64(c, U, si, tmp, am0);
    launch_q_shift_f64(c->L, (const double2 *)U, (const double2 *)tmp, (const double2 *)si,
                       (double2 *)so, am0, mu2, c->rows_q, c->stream);
}

int dot_host(qed2d_ctx *c, const double *a, const double *b, double *re, double *im)
{
    launch_dot((const double2 *)a, (const double2 *)b, spinor_elems(c), c->red, c->part,
               c->stream);
    TRY(launched(c));
    TRY(read_red(c));
    *re = c->hred->v[0];
    if (im) *im = c->hred->v[1];
    return QED2D_OK;
}

int norm2_host(qed2d_ctx *c, const double *a, double *n2)
{
    launch_norm2((const double2 *)a, spinor_elems(c), c->red, c->part, c->stream);
    TRY(launched(c));
    TRY(read_red(c));
    *n2 = c->hred->v[0];
    return QED2D_OK;
}

// Re a^dag b (|a|^2 when b is null) as a double-double, for the fermion terms
// of H.
int dot_dd_host(qed2d_ctx *c, const double *a, const double *b, hdd *re)
{
    launch_dot_dd((const double2 *)a, (const double2 *)b, spinor_elems(c), c->red, c->part,
                  c->stream);
    TRY(launched(c));
    TRY(read_red(c));
    *re = hdd();
    re->add(c->hred->v[0], c->hred->v[1]);
    return QED2D_OK;
}

// Sum Re P, sum arg P, sum p^2 as double-doubles.
int observables(qed2d_ctx *c, const double *U, const double *mom, hdd *sp, hdd *sq, hdd *sm)
{
    launch_observables(c->L, (const double2 *)U, mom, nullptr, nullptr, c->red, c->part,
                       c->stream);
    TRY(launched(c));
    TRY(read_red(c));
    if (sp) { *sp = hdd(); sp->add(c->hred->v[0], c->hred->v[1]); }
    if (sq) { *sq = hdd(); sq->add(c->hred->v[2], c->hred->v[3]); }
    if (sm) { *sm = hdd(); sm->add(c->hred->v[4], c->hred->v[5]); }
    return QED2D_OK;
}

// H = sum p^2 / 2 + beta (V - sum Re P)   (src/QED2dHMC.jl:198-215), returned
// as a double-double so that differences of two H stay exact to ~1e-16 |H|.
hdd hamiltonian_from(const qed2d_ctx *c, const hdd &sp, const hdd &sm)
{
    // beta (V - sum ReP): form V - sumReP in double-double first
    hdd g;
    g.add((double)c->L.V);
    g.add(-sp.hi, -sp.lo);
    double gh = g.hi, gl = g.lo;
    // multiply by beta with an FMA-exact product of the high part
    double ph = c->beta * gh;
    double pe = std::fma(c->beta, gh, -ph);
    hdd h;
    h.add(0.5 * sm.hi, 0.5 * sm.lo);
    h.add(ph, pe + c->beta * gl);
    return h;
}
int hamiltonian_dd(qed2d_ctx *c, const double *mom, const double *U, hdd *H)
{
    hdd sp, sm;
    TRY(observables(c, U, mom, &sp, nullptr, &sm));
    *H = hamiltonian_from(c, sp, sm);
    return QED2D_OK;
}

// ---------------------------------------------------------------------------
// CG drivers
// ---------------------------------------------------------------------------
struct SolveArgs {
    double am0 = 0.0, mu2 = 0.0, eps = 0.0;
    int maxiter = 0, guess = 0;
    int chunk0 = 8;  // iterations queued before the first look at the state
};

int upload_state(qed2d_ctx *c, const SolveArgs &a)
{
    SolverState &s = *c->hstate;
    std::memset(&s, 0, sizeof s);
    s.mu2 = a.mu2;
    s.eps = a.eps;
    s.maxiter = a.maxiter;
    s.guess = a.guess;
    s.delta2 = c->ru_delta * c->ru_delta;
    HIPCHK(c, hipMemcpyAsync(c->state, c->hstate, sizeof(SolverState), hipMemcpyHostToDevice,
                             c->stream));
    return QED2D_OK;
}

// Iterations to queue before the next look at the state: an estimate of what
// is left (to convergence, or for the mixed solver to the next reliable
// update) from the convergence rate seen so far.  A solve stops at most a
// few queued iterations early; the rest of the chunk returns at once.
struct ChunkPredictor {
    double err0 = 0.0, rate = 0.0;
    int it0 = -1;
    int next(double err, int iter, double tol, double ru_target)
    {
        int chunk = 8;
        if (it0 >= 0 && iter > it0 && err > 0.0 && err0 > err) {
            rate = std::log(err / err0) / (double)(iter - it0);  // < 0
            double target = tol;
            if (ru_target > target) target = ru_target;
            const double left = std::log(target / err) / rate;
            if (left > 0.0 && left < 1e6) chunk = (int)std::ceil(left) + 1;
            else chunk = 2;
        }
        it0 = iter;
        err0 = err;
        if (chunk < 2) chunk = 2;
        if (chunk > 64) chunk = 64;
        return chunk;
    }
    // The chunk after a reliable update at |r|^2 = err: 2 (K1 decides at once
    // whether the true residual converged) when the update was triggered near
    // the tolerance or no rate is known yet, else the iterations to the next
    // update (|r|^2 down by delta2) or to the tolerance, whichever comes first.
    int after_ru(double err, double tol, double delta2) const
    {
        if (!(rate < 0.0) || !(err > 100.0 * tol)) return 2;
        double drop = tol / err;
        if (delta2 > drop) drop = delta2;
        const double left = std::log(drop) / rate;
        int chunk = (left > 0.0 && left < 1e6) ? (int)std::ceil(left) + 1 : 2;
        if (chunk < 2) chunk = 2;
        if (chunk > 64) chunk = 64;
        return chunk;
    }
};

// ---------------------------------------------------------------------------
// Pipelined solves.  The last K2 of a queued chunk of CG iterations (its
// block 0, after its last write to the state) copies the solver state into a
// host-coherent pinned slot of the chunk's own and then writes the chunk's
// sequence number beside it (a chunk without a K2 ends with a one-thread
// kernel that does the same); the host queues
// the next chunk before it waits (polling that number) for the previous one,
// so the GPU keeps iterating while the host looks at the state.  Chunks are sized from
// the observed convergence rate to end where the solve is predicted to end; a
// launch past the end of a solve returns at once.
// ---------------------------------------------------------------------------
struct ChunkQueue {
    int slot[kRing];
    bool last[kRing];
    int k0[kRing];   // first iteration of the chunk
    bool ru[kRing];  // the chunk starts with an on-device reliable update
    int want[kRing];  // the sequence number of its state copy
    bool tail[kRing];  // the chunk is followed by the solve's tail (qed2d_ctx::tail)
    bool popped_tail = false;  // the chunk popped last was
    int head = 0, n = 0;
    int back() const { return (head + n - 1) % kRing; }
    bool back_last() const { return n > 0 && last[back()]; }
};

// The state-copy slot and sequence number of the next chunk; the chunk's last
// K2 takes the copy (StateSnap).
struct ChunkSnap {
    int slot;
    StateSnap sn;
};
ChunkSnap cq_reserve(qed2d_ctx *c)
{
    const int s = c->ring_next;
    c->ring_next = (s + 1) % kRing;
    c->seq_no = (c->seq_no == 0x7fffffff) ? 1 : c->seq_no + 1;
    return ChunkSnap{s, StateSnap{c->hring + s, c->hseq + s, c->seq_no}};
}

// Enters a queued chunk; taken: its last kernel took the state copy, else a
// one-thread kernel takes it now.
int cq_push(qed2d_ctx *c, ChunkQueue &q, const ChunkSnap &cs, bool taken, bool last, int k0 = 0,
            bool ru = false, bool tail = false)
{
    if (!taken) launch_state_snap(c->state, cs.sn.out, cs.sn.seq, cs.sn.v, c->stream);
    // the tail sees the state the copy holds: it ran iff the copy says converged
    if (tail)
        launch_q_apply_ru(c->L, (const double2 *)c->tail.U, (const double2 *)c->tail.si,
                          (double2 *)c->tail.so, c->tail.am0, c->state, c->rows_q, c->stream, -1);
    TRY(launched(c));
    const int i = (q.head + q.n) % kRing;
    q.tail[i] = tail;
    q.want[i] = cs.sn.v;
    q.slot[i] = cs.slot;
    q.last[i] = last;
    q.k0[i] = k0;
    q.ru[i] = ru;
    q.n += 1;
    return QED2D_OK;
}

// Waits for the oldest queued chunk; its state lands in *c->hstate.
int cq_pop(qed2d_ctx *c, ChunkQueue &q)
{
    if (q.n < 1) return fail(c, QED2D_EHIP, "internal: no queued CG chunk to wait for");
    const int s = q.slot[q.head], want = q.want[q.head];
    q.popped_tail = q.tail[q.head];
    q.head = (q.head + 1) % kRing;
    q.n -= 1;
    // poll the sequence number; every 2 ms ask the stream whether a launch
    // failed, or whether it drained without the copy arriving (a stream query
    // is not free on the GPU side: asked every few microseconds, it put a gap
    // of about 6 us before the chunk queued next)
    auto t_query = std::chrono::steady_clock::now();
    for (long spin = 1;; ++spin) {
        if (__atomic_load_n(c->hseq + s, __ATOMIC_ACQUIRE) == want) break;
        if ((spin & 255) == 0 &&
            std::chrono::steady_clock::now() - t_query >= std::chrono::milliseconds(2)) {
            t_query = std::chrono::steady_clock::now();
            const hipError_t e = hipStreamQuery(c->stream);
            if (e == hipSuccess && __atomic_load_n(c->hseq + s, __ATOMIC_ACQUIRE) != want)
                return fail(c, QED2D_EHIP, "internal: solver state copy %d did not arrive", want);
            if (e != hipSuccess && e != hipErrorNotReady)
                return fail(c, QED2D_EHIP, "CG chunk: %s", hipGetErrorString(e));
        }
        __builtin_ia32_pause();
    }
    std::memcpy((void *)c->hstate, (const void *)(c->hring + s), sizeof(SolverState));
    return QED2D_OK;
}

// At the end of a solve (the chunk popped last reported it converged, or
// not): the tail ran iff that chunk or one queued after it carried it (the
// state stays converged once it is).  Forgets the queued chunks (their
// launches return at once).
void cq_finish(qed2d_ctx *c, ChunkQueue &q)
{
    bool any = q.popped_tail;
    for (int j = 0; j < q.n; ++j) any = any || q.tail[(q.head + j) % kRing];
    c->tail_ran = any && c->hstate->status == 1;
    q.n = 0;
}

// What the observed states say about the rest of a solve.
struct SolvePlan {
    double rate = 0.0;  // d ln|r|^2 per iteration (< 0); 0: not known yet
    double err = 0.0;
    int iter = 0;
    int nobs = 0;  // observations with a rate between two of them
    void observe(const SolverState &h)
    {
        const double e = h.err;
        const int it = h.iter;
        if (!(e > 0.0) || it < 1 || it <= iter) return;
        if (iter >= 1 && err > e) {
            rate = std::log(e / err) / (double)(it - iter);
            nobs += 1;
        } else if (iter < 1 && h.r0norm > e) {
            rate = std::log(e / h.r0norm) / (double)it;
        }
        iter = it;
        err = e;
    }
    // The K1 predicted to see |r|^2 < target (0: not known).
    int kconv(double target) const
    {
        if (!(rate < 0.0) || iter < 1) return 0;
        if (err < target) return iter + 1;
        const double left = std::log(target / err) / rate;
        if (!(left < 1e6)) return 0;
        return iter + 1 + (int)std::ceil(left);
    }
    // Iterations over which |r|^2 falls by the factor d (0: not known).
    int span(double d) const
    {
        if (!(rate < 0.0)) return 0;
        const double n = std::log(d) / rate;
        return (n > 0.0 && n < 1e6) ? (int)std::ceil(n) : 0;
    }
};

constexpr int kChunkMax = 64;    // iterations per queued chunk at most
constexpr int kChunkBlind = 16;  // chunk length while the convergence rate is unknown

int cg_f64(qed2d_ctx *c, double *x, const double *U, const double *b, const SolveArgs &a,
           int *iters)
{
    double2 *r, *P[2], *t;
    TRY(get_spin(c, S_R, &r));
    TRY(get_spin(c, S_P0, &P[0]));
    TRY(get_spin(c, S_P1, &P[1]));
    TRY(get_spin(c, S_T, &t));
    TRY(upload_state(c, a));
    const Latt &L = c->L;
    hipStream_t st = c->stream;
    if (a.guess) {
        q64(c, U, (const double *)x, (double *)t, a.am0);
        launch_cg_resid(L, (const double2 *)U, (const double2 *)b, t,
                        a.mu2 != 0.0 ? (const double2 *)x : nullptr, r, nullptr, P[0], nullptr,
                        a.am0, c->state, c->P, 1, 0, c->rows, st);
    } else {
        launch_cg_init_f64(L, (const double2 *)b, (double2 *)x, r, P[0], c->P, c->rows, st);
    }
    const int kmax = a.maxiter + 1;  // K1(maxiter + 1) ends a solve that has not converged
    int k = 1;
    ChunkQueue q;
    SolvePlan plan;
    c->tail_ran = false;
    const bool has_tail = c->tail.U != nullptr;
    auto queue = [&](int k1, bool last) -> int {
        const ChunkSnap cs = cq_reserve(c);
        const int kend = std::min(k1, kmax), k0 = k;
        for (; k <= kend; ++k) {
            PROF(c, QED2D_PROF_CG_K1_F64,
                 launch_cg_k1_f64(L, (const double2 *)U, r, P[0], P[1], t, a.am0, c->state, c->P,
                                  k, c->rows, st));
            PROF(c, QED2D_PROF_CG_K2_F64,
                 launch_cg_k2_f64(L, (const double2 *)U, t, P[0], P[1], (double2 *)x, r, a.am0,
                                  c->state, c->P, k, c->rows_k2 ? c->rows_k2 : c->rows, st,
                                  k == kend ? cs.sn : StateSnap{nullptr, nullptr, 0}));
        }
        TRY(launched(c));
        return cq_push(c, q, cs, k > k0, last || k > kmax, 0, false, has_tail && last);
    };
    TRY(queue(a.chunk0, false));
    // (a caller's guess may already satisfy the tolerance: no second chunk
    // before the first look)
    if (!a.guess) TRY(queue(k + kChunkBlind - 1, false));
    for (;;) {
        TRY(cq_pop(c, q));
        const SolverState &h = *c->hstate;
        if (h.status != 0) break;
        plan.observe(h);
        // keep two chunks queued unless the newer one is predicted to end the solve
        while (k <= kmax && (q.n == 0 || (q.n < 2 && !q.back_last()))) {
            const int kc = plan.kconv(h.tol);
            int k1 = k + kChunkBlind - 1;
            bool last = false;
            if (kc > 0) {
                k1 = std::max(kc + 1, k + 1);  // the K1 that sees it converged, and one more
                last = k1 <= k + kChunkMax - 1;
                if (!last) k1 = k + kChunkMax - 1;
            }
            TRY(queue(k1, last));
        }
    }
    cq_finish(c, q);
    *iters = c->hstate->iter;
    if (c->hstate->status == 1) return QED2D_OK;
    return fail(c, QED2D_ENOCONV, "CG not converged after %d iterations (|r|^2 = %g, tol = %g)",
                a.maxiter, c->hstate->err, c->hstate->tol);
}

// Mixed precision: FP32 iterations, FP64 reliable updates (the true residual
// b - (Q^2 + mu^2) x in FP64 replaces the iterated one).  Where the FP32 K2's
// partial-sum grid can take the residual kernel's (soft mode), a K1 whose
// iterated |r| has dropped by delta only requests the update and the solve
// goes on: the update's kernels are queued at the start of the chunks the
// host plans there and run on the device iff a request is pending, so the
// host does not wait for them.  A residual that looks converged, one far past
// the trigger and the iteration limit pause the solve for an update from the
// host, after which K1 tests the true residual (at the iteration the FP64
// solver would stop at).
int cg_mixed(qed2d_ctx *c, double *x, const double *U, const double *b, const SolveArgs &a,
             int *iters)
{
    // (the residual reads x only for its mu^2 x term)
    const double2 *xmu = a.mu2 != 0.0 ? (const double2 *)x : nullptr;
    float2 *rs, *P[2], *ts, *xs, *Us;
    double2 *t;
    TRY(get_fspin(c, FS_R, &rs));
    TRY(get_fspin(c, FS_P0, &P[0]));
    TRY(get_fspin(c, FS_P1, &P[1]));
    TRY(get_fspin(c, FS_T, &ts));
    TRY(get_fspin(c, FS_X, &xs));
    TRY(get_fspin(c, FS_U, &Us));
    TRY(get_spin(c, S_T, &t));
    const Latt &L = c->L;
    hipStream_t st = c->stream;
    const float am0f = (float)a.am0;
    const int rows32_k2 = c->rows32_k2 ? c->rows32_k2 : c->rows32;
    // soft mode: the on-device update's residual partials take the count K2
    // leaves in the slot: folded to one (K2 grids past kPartsUnfolded), or
    // padded, which needs a residual grid no larger than K2's (the largest
    // such grid: at 1024^2 with K2 at 4 rows, 8 rows per thread ran the
    // residual in less time than 16, which leave one block per CU)
    const int nk2 = cg2_f32_blocks(L, rows32_k2);
    int rows_r = nk2 > kPartsUnfolded ? c->rows : 0;
    for (int R : {1, 2, 4, 8, 16})
        if (!rows_r && resid_blocks(L, R) <= nk2) rows_r = R;
    const bool soft = c->ru_mode && rows_r > 0;
    // In soft mode the chunk predicted to end the solve carries the reliable
    // update of its last K1 (which pauses where the iterated residual looks
    // converged) and that K1 and K2 again, each run only if the pause came
    // there: the host's look then finds the solve converged instead of paused.
    // (Site-pair kernels, whose K1 and K2 take the guard, and grids whose
    // partials are never folded: a folded slot must not be refolded by a
    // guarded launch that did not run.)
    const bool spec_ok = soft && pair_ok(L) && cg2_f32_blocks(L, c->rows32) <= kPartsUnfolded &&
                         nk2 <= kPartsUnfolded;
    // The site-pair kernels read the links packed (one float each) when they
    // are U(1); the first look at the state tells, and a solve whose links are
    // not restarts with FP32 complex links (x is untouched until then).
    bool packed = pair_ok(L);
    const float2 *pend0 = pair_ok(L) ? P[0] : nullptr;  // K2 leaves x += alpha p to K1
    const int kmax = a.maxiter + 1;
    int k = 1;
    int kb_last = 1;  // the last update known to have run, or planned (1: the start)
    ChunkQueue q;
    SolvePlan plan;
    auto start = [&]() -> int {
        TRY(upload_state(c, a));
        if (packed)
            launch_link_pack((const double2 *)U, (float *)Us, 2 * L.V, c->state, st);
        else
            launch_to_f32((const double2 *)U, Us, 2 * L.V, st);
        if (a.guess) {
            q64(c, U, x, (double *)t, a.am0);
            launch_cg_resid(L, (const double2 *)U, (const double2 *)b, t, xmu,
                            nullptr, rs, nullptr, P[0], a.am0, c->state, c->P, 1, 1, c->rows, st);
            HIPCHK(c, hipMemsetAsync(xs, 0, sizeof(float2) * 2 * L.V, st));
        } else {
            launch_cg_init_mixed(L, (const double2 *)b, (double2 *)x, rs, xs, P[0], c->P, c->rows,
                                 st);
        }
        k = 1;
        kb_last = 1;
        q.n = 0;
        plan = SolvePlan();
        return QED2D_OK;
    };
    // iterations k .. k1, after the on-device update when ru (soft mode);
    // spec: followed by the update of a pause at K1(k1) and K1(k1), K2(k1)
    // again, each a no-op unless K1(k1) paused
    c->tail_ran = false;
    const bool has_tail = c->tail.U != nullptr;
    // (tl: the solve's tail follows the chunk)
    auto queue = [&](int k1, bool ru, bool last, bool spec = false, bool tl = false) -> int {
        const int k0 = k, kend = std::min(k1, kmax);
        const ChunkSnap cs = cq_reserve(c);
        const StateSnap none{nullptr, nullptr, 0};
        // (the spec tail runs iff at least one iteration is queued)
        const bool tail = spec && spec_ok && kend >= k0;
        if (ru) {
            hipEvent_t ev = prof_begin(c, QED2D_PROF_RELIABLE);
            launch_ru_accum(L, (double2 *)x, xs, pend0, P[1], c->state, st, k);
            launch_q_apply_ru(L, (const double2 *)U, (const double2 *)x, t, a.am0, c->state,
                              c->rows_q, st);
            launch_cg_resid(L, (const double2 *)U, (const double2 *)b, t, xmu,
                            nullptr, rs, nullptr, nullptr, a.am0, c->state, c->P, k, 3, rows_r,
                            st);
            prof_end(c, QED2D_PROF_RELIABLE, ev);
        }
        for (; k <= kend; ++k) {
            PROF(c, QED2D_PROF_CG_K1_F32,
                 launch_cg_k1_f32(L, Us, packed, rs, P[0], P[1], ts, xs, am0f, c->state, c->P, k,
                                  c->rows32, st, soft));
            PROF(c, QED2D_PROF_CG_K2_F32,
                 launch_cg_k2_f32(L, Us, packed, ts, P[0], P[1], xs, rs, am0f, a.mu2 != 0.0,
                                  c->state, c->P, k, rows32_k2, st, false,
                                  (k == kend && !tail) ? cs.sn : none));
        }
        if (tail) {
            const int kc = k - 1;
            hipEvent_t ev = prof_begin(c, QED2D_PROF_RELIABLE);
            launch_ru_accum(L, (double2 *)x, xs, pend0, P[1], c->state, st, 0, kc);
            launch_q_apply_ru(L, (const double2 *)U, (const double2 *)x, t, a.am0, c->state,
                              c->rows_q, st, kc);
            launch_cg_resid(L, (const double2 *)U, (const double2 *)b, t, xmu,
                            nullptr, rs, nullptr, nullptr, a.am0, c->state, c->P, kc, 4, rows_r,
                            st);
            prof_end(c, QED2D_PROF_RELIABLE, ev);
            launch_cg_k1_f32(L, Us, packed, rs, P[0], P[1], ts, xs, am0f, c->state, c->P, kc,
                             c->rows32, st, soft, true);
            launch_cg_k2_f32(L, Us, packed, ts, P[0], P[1], xs, rs, am0f, a.mu2 != 0.0,
                             c->state, c->P, kc, rows32_k2, st, true, cs.sn);
        }
        TRY(launched(c));
        return cq_push(c, q, cs, k > k0, last || k > kmax, k0, ru, has_tail && tl);
    };
    // The next chunk from k: its end, whether it starts with an update.  In soft
    // mode an update is planned at the chunk start after the K1 predicted to
    // request it, from the newest state the host has seen: right away when a
    // request is pending; else where the residual is predicted to have fallen
    // by delta^2 below its largest value since the last update, or, when the
    // chunk in flight starts with an update that will run, span(delta^2)
    // iterations after that.  Chunks last at most half a span, so an update
    // that comes later than planned waits at most that long.  The residual
    // that looks converged pauses the solve (the host's update then lets K1
    // test the true residual at the iteration the FP64 solver would), so the
    // chunk predicted to reach it ends there and is the last queued.
    auto plan_next = [&](const SolverState &h, int *k1, bool *ru, bool *last) {
        *ru = false;
        *last = false;
        // (the rate from the start alone overstates the asymptotic one)
        const int nd = (soft && plan.nobs > 0) ? plan.span(h.delta2) : 0;
        if (nd == 0) {
            *k1 = k + kChunkBlind - 1;
            // (no soft mode: up to the predicted pause, for the update or the end)
            const int kp = soft ? 0 : plan.kconv(std::max(h.tol, h.delta2 *
                                                              h.rmax2[(h.iter + 1) & 1]));
            if (kp > 0) {
                *k1 = std::max(kp + 1, k + 1);
                if (*k1 > k + kChunkMax - 1) *k1 = k + kChunkMax - 1;
            }
            return;
        }
        // an update planned at kb_last that the state shows did not run
        if (h.iter + 1 >= kb_last && h.true_at < kb_last) kb_last = h.true_at;
        if (h.true_at > kb_last) kb_last = h.true_at;
        // the K1 that requests the next update, seen from the state
        int jt = h.ru_req ? h.iter + 1 : plan.kconv(h.delta2 * h.rmax2[(h.iter + 1) & 1]);
        if (q.n > 0 && q.ru[q.back()] && q.k0[q.back()] > kb_last && jt > 0 &&
            jt < q.k0[q.back()])
            kb_last = q.k0[q.back()];  // the update in flight runs
        if (kb_last > h.iter + 1 || jt <= 0) jt = kb_last + nd;
        int due = jt + 2;
        if (k >= due && k > 1) {
            *ru = true;
            kb_last = k;
            due = k + nd + 2;
        }
        // the chunk that ends the solve: K1(kc) pauses it for the host
        const int kc = plan.kconv(h.tol);
        if (kc > 0 && kc < due) {
            *k1 = std::max(kc, k);
            *last = *k1 <= k + kChunkMax - 1;
            if (!*last) *k1 = k + kChunkMax - 1;
            return;
        }
        *k1 = std::min(due - 1, k + std::max(3, nd / 2));
        // (a chunk short of four iterations would leave the host less time than
        // it takes to look and queue the next one; an update then waits a few
        // iterations)
        if (*k1 < k + 3) *k1 = k + 3;
        if (*k1 > k + kChunkMax - 1) *k1 = k + kChunkMax - 1;
    };
    TRY(start());
    TRY(queue(a.chunk0, false, false));
    bool first_look = true;
    // (a caller's guess may already satisfy the tolerance, and its x is updated
    // by the first reliable update: no second chunk before the first look)
    if (!a.guess) TRY(queue(k + kChunkBlind - 1, false, false));
    for (;;) {
        TRY(cq_pop(c, q));
        if (first_look && packed && c->hstate->bad_links) {
            packed = false;
            TRY(start());
            TRY(queue(a.chunk0, false, false));
            continue;
        }
        first_look = false;
        const SolverState &h = *c->hstate;
        if (h.status != 0) break;
        if (h.need_ru) {
            // K1(k) paused: fold the FP32 partial solution into x, recompute the
            // true residual in FP64 and run K1(k) again on it.  (A chunk queued
            // after the pause returns at once.)
            q.n = 0;
            k = h.stop_k;
            hipEvent_t ev = prof_begin(c, QED2D_PROF_RELIABLE);
            // (the site-pair kernels pause with K2(k-1)'s x update pending)
            launch_ru_accum(L, (double2 *)x, xs, pend0, P[1], c->state, st);
            launch_q_apply_f64(L, (const double2 *)U, (const double2 *)x, t, a.am0, c->rows_q, st);
            launch_cg_resid(L, (const double2 *)U, (const double2 *)b, t, xmu,
                            nullptr, rs, nullptr, nullptr, a.am0, c->state, c->P, k, 2,
                            rows_r ? rows_r : c->rows, st);
            prof_end(c, QED2D_PROF_RELIABLE, ev);
            kb_last = k;
            plan.observe(h);
            // K1(k) decides at once whether the true residual converged (the
            // tail follows when the pause came from a residual that looks so)
            TRY(queue(k, false, false, false, h.err < h.tol));
            continue;
        }
        plan.observe(h);
        while (k <= kmax && (q.n == 0 || (q.n < 2 && !q.back_last()))) {
            int k1;
            bool ru, last;
            plan_next(h, &k1, &ru, &last);
            TRY(queue(k1, ru, last, last, last));
        }
    }
    cq_finish(c, q);
    *iters = c->hstate->iter;
    if (c->hstate->status == 1) return QED2D_OK;
    return fail(c, QED2D_ENOCONV,
                "mixed-precision CG not converged after %d iterations (|r|^2 = %g, tol = %g)",
                a.maxiter, c->hstate->err, c->hstate->tol);
}

int cg_any(qed2d_ctx *c, double *x, const double *U, const double *b, double am0, double mu,
           int op, int maxiter, double eps, int guess, int precision, int *iters)
{
    SolveArgs a;
    a.am0 = am0;
    a.mu2 = (op == QED2D_OP_SQR_MUSQ) ? mu * mu : 0.0;
    a.eps = eps;
    a.maxiter = maxiter;
    a.guess = guess;
    int dummy = 0;
    if (!iters) iters = &dummy;
    if (precision == QED2D_PREC_MIXED) return cg_mixed(c, x, U, b, a, iters);
    return cg_f64(c, x, U, b, a, iters);
}

// Multi-shift CG: (Q^2 + mu_i^2) x_i = b for all i at once; x[i] are device
// spinors (may be internal buffers).  Each shift stops on its own
// |r_i|^2 < eps |b|^2, the reference's per-CG rule (src/QED2dSolver.jl:16,88).
int grow(qed2d_ctx *c, std::vector<double2 *> &v, int n)
{
    while ((int)v.size() < n) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(double2) * spinor_elems(c)) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of a multi-shift buffer failed");
        v.push_back((double2 *)p);
    }
    return QED2D_OK;
}

int grow32(qed2d_ctx *c, std::vector<float2 *> &v, int n)
{
    while ((int)v.size() < n) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(float2) * spinor_elems(c)) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of an FP32 multi-shift buffer failed");
        v.push_back((float2 *)p);
    }
    return QED2D_OK;
}

// nx internal solution buffers, np search-direction buffers, the state.
int get_shiftbufs(qed2d_ctx *c, int nx, int np)
{
    TRY(grow(c, c->shiftbuf, nx));
    TRY(grow(c, c->psbuf, np));
    if (!c->ms) {
        if (hipMalloc((void **)&c->ms, sizeof(MultiShiftState)) != hipSuccess ||
            hipHostMalloc((void **)&c->hms, sizeof(MultiShiftState), hipHostMallocDefault) !=
                hipSuccess)
            return fail(c, QED2D_ENOMEM, "allocation of the multi-shift state failed");
        HIPCHK(c, hipMemsetAsync(c->ms, 0, sizeof(MultiShiftState), c->stream));
    }
    return QED2D_OK;
}

int read_ms(qed2d_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->hms, c->ms, sizeof(MultiShiftState), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_harvest(c);
    return QED2D_OK;
}

int multishift_cg(qed2d_ctx *c, double2 *const *x, const double *U, const double *b, double am0,
                  int n, const double *mu, int maxiter, double eps, int *iters)
{
    if (n < 1 || n > kMaxShift)
        return fail(c, QED2D_EINVAL, "multi-shift CG takes 1..%d shifts (got %d)", kMaxShift, n);
    TRY(get_shiftbufs(c, 0, n));
    double2 *r, *P[2], *t;
    TRY(get_spin(c, S_R, &r));
    TRY(get_spin(c, S_P0, &P[0]));
    TRY(get_spin(c, S_P1, &P[1]));
    TRY(get_spin(c, S_T, &t));
    MultiShiftState &h = *c->hms;
    std::memset(&h, 0, sizeof h);
    double sb = mu[0] * mu[0];
    for (int k = 1; k < n; ++k) sb = std::min(sb, mu[k] * mu[k]);
    h.sigma_b = sb;
    h.eps = eps;
    h.maxiter = maxiter;
    h.nshift = n;
    for (int k = 0; k < n; ++k) {
        h.s[k] = mu[k] * mu[k] - sb;
        h.x[k] = x[k];
        h.ps[k] = c->psbuf[k];
    }
    HIPCHK(c, hipMemcpyAsync(c->ms, c->hms, sizeof(MultiShiftState), hipMemcpyHostToDevice,
                             c->stream));
    const Latt &L = c->L;
    hipStream_t st = c->stream;
    launch_ms_init(L, (const double2 *)b, r, P[0], c->ms, c->P, n, c->rows, st);
    int k = 1;
    int chunk = 8;
    ChunkPredictor pred;
    for (;;) {
        for (int j = 0; j < chunk && k <= maxiter + 1; ++j, ++k) {
            PROF(c, QED2D_PROF_OTHER,
                 launch_ms_k1(L, (const double2 *)U, r, P[0], P[1], t, am0, c->ms, c->P, k, n,
                              c->rows, st));
            PROF(c, QED2D_PROF_OTHER, launch_ms_shifts(c->ms, r, spinor_elems(c), k, n, st));
            PROF(c, QED2D_PROF_OTHER,
                 launch_ms_k2(L, (const double2 *)U, t, P[0], P[1], r, am0, c->ms, c->P, k, n,
                              c->rows, st));
        }
        TRY(launched(c));
        TRY(read_ms(c));
        if (c->hms->status != 0 || k > maxiter + 1) break;
        chunk = pred.next(c->hms->err, c->hms->iter, c->hms->tol, 0.0);
    }
    if (iters) *iters = c->hms->iter;
    if (c->hms->status == 1) return QED2D_OK;
    return fail(c, QED2D_ENOCONV, "multi-shift CG not converged after %d iterations", maxiter);
}

// Multi-shift CG of nf right-hand sides b[j] with shifts mu[j][0..npole[j]-1]
// and masses am0[j] at once: every K1/K2 is one launch over all right-hand
// sides (a chain set of nf copies of the lattice sharing U), each with its own
// Krylov scalars and stopping.  The solutions land in c->mr.x[j][i].  Per
// right-hand side the tiling and arithmetic are those of multishift_cg.
int mr_setup(qed2d_ctx *c, int nf)
{
    auto &m = c->mr;
    if (nf > m.cap) {
        for (void *p : {(void *)m.dch, (void *)m.blk0, (void *)m.shape, (void *)m.M,
                        (void *)m.part, (void *)m.cnt, (void *)m.arena})
            if (p) (void)hipFree(p);
        if (m.hM) (void)hipHostFree(m.hM);
        // (only what is reallocated here: the solution, force and pole buffers
        // of the set stay, whoever allocated them)
        m.dch = nullptr;
        m.blk0 = nullptr;
        m.shape = nullptr;
        m.M = m.hM = nullptr;
        m.part = nullptr;
        m.cnt = nullptr;
        m.arena = nullptr;
        m.P = BParts{};
        m.cap = m.n = m.rows = 0;
        const long nbc = march_grid(c->L, 1).nblocks();
        const long n2 = spinor_elems(c);
        if (hipMalloc((void **)&m.dch, sizeof(BChain) * nf) != hipSuccess ||
            hipMalloc((void **)&m.blk0, sizeof(int) * (nf + 1)) != hipSuccess ||
            hipMalloc((void **)&m.shape, sizeof(int4) * nf) != hipSuccess ||
            hipMalloc((void **)&m.M, sizeof(MultiShiftState) * nf) != hipSuccess ||
            hipHostMalloc((void **)&m.hM, sizeof(MultiShiftState) * nf) != hipSuccess ||
            hipMalloc((void **)&m.part, sizeof(double) * 7 * nbc * nf) != hipSuccess ||
            hipMalloc((void **)&m.cnt, sizeof(int) * 5 * nf) != hipSuccess ||
            hipMalloc((void **)&m.arena, sizeof(double2) * 4 * n2 * nf) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "multi-shift right-hand-side set allocation failed");
        const long pb = nbc * nf;
        m.P.tt[0] = m.part;
        m.P.tt[1] = m.part + 2 * pb;
        m.P.rr[0] = m.part + 4 * pb;
        m.P.rr[1] = m.part + 5 * pb;
        m.P.b = m.part + 6 * pb;
        m.P.ntt[0] = m.cnt;
        m.P.ntt[1] = m.cnt + nf;
        m.P.nrr[0] = m.cnt + 2 * nf;
        m.P.nrr[1] = m.cnt + 3 * nf;
        m.P.nb = m.cnt + 4 * nf;
        m.cap = nf;
    }
    if (m.n != nf || m.rows != c->rows) {
        const MarchGrid g = march_grid(c->L, c->rows);
        std::vector<int> blk0(nf + 1);
        std::vector<int4> shape(nf);
        for (int j = 0; j <= nf; ++j) blk0[j] = j * g.nblocks();
        for (int j = 0; j < nf; ++j) shape[j] = make_int4(g.bx, g.by, g.nbx, g.nblocks());
        HIPCHK(c, hipMemcpyAsync(m.blk0, blk0.data(), sizeof(int) * (nf + 1),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(m.shape, shape.data(), sizeof(int4) * nf, hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        m.n = nf;
        m.rows = c->rows;
    }
    if ((int)m.x.size() < nf) {
        m.x.resize(nf);
        m.ps.resize(nf);
    }
    if ((int)m.x32.size() < nf) {
        m.x32.resize(nf);
        m.ps32.resize(nf);
    }
    return QED2D_OK;
}

// f32: the whole solve in FP32 (FP32 links, spinors, x_i and p_i; Krylov
// scalars and sums in FP64), the solutions in c->mr.x32[j][i]; a right-hand
// side that reaches the iteration limit is not an error (its x_i are guesses).
int multishift_multi(qed2d_ctx *c, const double *U, int nf, const double *const *b,
                     const double *am0, const int *npole, const double *const *mu, int maxiter,
                     double eps, int *iters, bool f32 = false)
{
    TRY(mr_setup(c, nf));
    auto &m = c->mr;
    const long n2 = spinor_elems(c);
    const long nbc = march_grid(c->L, 1).nblocks();
    std::vector<BChain> ch(nf);
    int maxshift = 0;
    for (int j = 0; j < nf; ++j) {
        if (npole[j] < 1 || npole[j] > kMaxShift)
            return fail(c, QED2D_EINVAL, "multi-shift CG takes 1..%d shifts", kMaxShift);
        maxshift = std::max(maxshift, npole[j]);
        if (f32) {
            TRY(grow32(c, m.x32[j], npole[j]));
            TRY(grow32(c, m.ps32[j], npole[j]));
        } else {
            TRY(grow(c, m.x[j], npole[j]));
            TRY(grow(c, m.ps[j], npole[j]));
        }
        ch[j].L = c->L;
        ch[j].off = (long)j * c->L.V;
        ch[j].uoff = 0;
        ch[j].pb = (int)(j * nbc);
        ch[j].pad = 0;
        ch[j].am0 = am0[j];
        ch[j].beta = c->beta;
        MultiShiftState &h = m.hM[j];
        std::memset(&h, 0, sizeof h);
        double sb = mu[j][0] * mu[j][0];
        for (int k = 1; k < npole[j]; ++k) sb = std::min(sb, mu[j][k] * mu[j][k]);
        h.sigma_b = sb;
        h.eps = eps;
        h.maxiter = maxiter;
        h.nshift = npole[j];
        for (int k = 0; k < npole[j]; ++k) {
            h.s[k] = mu[j][k] * mu[j][k] - sb;
            h.x[k] = f32 ? (void *)m.x32[j][k] : (void *)m.x[j][k];
            h.ps[k] = f32 ? (void *)m.ps32[j][k] : (void *)m.ps[j][k];
        }
        h.b = (const double2 *)b[j];
        h.kk = 1;
        h.k2 = 1;
    }
    hipStream_t st = c->stream;
    HIPCHK(c, hipMemcpyAsync(m.dch, ch.data(), sizeof(BChain) * nf, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(m.M, m.hM, sizeof(MultiShiftState) * nf, hipMemcpyHostToDevice,
                             st));
    BGeo g;
    g.ch = m.dch;
    g.blk0 = m.blk0;
    g.shape = m.shape;
    g.n = nf;
    g.nblocks = nf * march_grid(c->L, c->rows).nblocks();
    g.inter = 1;  // one lattice shape, shared links
    double2 *r = m.arena, *p0 = r + n2 * nf, *p1 = p0 + n2 * nf, *t = p1 + n2 * nf;
    // (FP32: the same arena holds the FP32 r, p0, p1, t; the FP32 links go to
    // the mixed CG's link buffer)
    float2 *r32 = (float2 *)m.arena, *p032 = r32 + n2 * nf, *p132 = p032 + n2 * nf,
           *t32 = p132 + n2 * nf;
    float2 *U32 = nullptr;
    if (f32) {
        TRY(get_fspin(c, FS_U, &U32));
        launch_to_f32((const double2 *)U, U32, 2 * c->L.V, st);
        launch_b_ms_init_f32(g, c->rows, r32, p032, m.M, m.P, st);
    } else {
        launch_b_ms_init(g, c->rows, r, p0, m.M, m.P, st);
    }
    std::vector<ChunkPredictor> pred(nf);
    int chunk = 8;
    long launched_k = 0;
    for (;;) {
        for (int jj = 0; jj < chunk; ++jj) {
            // (the shift updates run as extra blocks of the same launch)
            if (f32) {
                PROF(c, QED2D_PROF_OTHER,
                     launch_b_ms_k1_f32(g, c->rows, U32, r32, p032, p132, t32, m.M, m.P, n2,
                                        maxshift, st));
                PROF(c, QED2D_PROF_OTHER,
                     launch_b_ms_k2_f32(g, c->rows, U32, t32, p032, p132, r32, m.M, m.P, st));
                continue;
            }
            PROF(c, QED2D_PROF_OTHER,
                 launch_b_ms_k1(g, c->rows, (const double2 *)U, r, p0, p1, t, m.M, m.P, n2,
                                maxshift, st));
            PROF(c, QED2D_PROF_OTHER,
                 launch_b_ms_k2(g, c->rows, (const double2 *)U, t, p0, p1, r, m.M, m.P, st));
        }
        launched_k += chunk;
        TRY(launched(c));
        HIPCHK(c, hipMemcpyAsync(m.hM, m.M, sizeof(MultiShiftState) * nf, hipMemcpyDeviceToHost,
                                 st));
        HIPCHK(c, hipStreamSynchronize(st));
        prof_harvest(c);
        bool running = false;
        int next = 64;
        for (int j = 0; j < nf; ++j) {
            const MultiShiftState &h = m.hM[j];
            if (h.status != 0) continue;
            running = true;
            const int pch = pred[j].next(h.err, h.iter, h.tol, 0.0);
            if (pch < next) next = pch;
        }
        if (!running) break;
        if (launched_k > 4L * (maxiter + 2) + 64)
            return fail(c, QED2D_EHIP, "multi-shift right-hand-side set stopped advancing");
        chunk = next;
    }
    int it = 0;
    for (int j = 0; j < nf; ++j) {
        if (m.hM[j].status != 1 && !f32)
            return fail(c, QED2D_ENOCONV, "multi-shift CG of right-hand side %d not converged "
                        "after %d iterations", j, maxiter);
        it = std::max(it, m.hM[j].iter);
    }
    if (iters) *iters = it;
    return QED2D_OK;
}

// ---------------------------------------------------------------------------
// MD
// ---------------------------------------------------------------------------

void updt(qed2d_ctx *c, double *U, const double *mom, double eps)
{
    PROF(c, QED2D_PROF_UPDT_FLD, launch_updt_fld(c->L, (double2 *)U, mom, eps, c->stream));
}

// update_momenta! for Nf=2 (src/QED2dHMC.jl:452-481)
int update_momenta_nf2(qed2d_ctx *c, double *mom, const double *U, double *X, const double *F,
                       double *g5DX, double am0, double eps, int maxiter, double tol,
                       int precision, int *iters)
{
    int it = 0;
    // g5DX = Q X, queued by the solver behind its last chunk where it can be
    c->tail = qed2d_ctx::SolveTail{U, X, g5DX, am0};
    const int rc = cg_any(c, X, U, F, am0, 0.0, QED2D_OP_SQR_MSQ, maxiter, tol, 0, precision, &it);
    c->tail = qed2d_ctx::SolveTail{};
    TRY(rc);
    *iters += it;
    if (!c->tail_ran) q64(c, U, X, g5DX, am0);
    PROF(c, QED2D_PROF_FORCE_MOM,
         launch_force_mom(c->L, (const double2 *)U, (const double2 *)X, (const double2 *)g5DX,
                          nullptr, mom, c->beta, eps, 1.0, c->stream));
    return launched(c);
}

int md_nf2(qed2d_ctx *c, int integrator, double *mom, double *U, double *X, const double *F,
           double *g5DX, double am0, double eps, int ns, int maxiter, double tol, int precision,
           int *iters)
{
    *iters = 0;
    if (integrator == QED2D_OMF4) {
        // src/QED2dHMC.jl:307-422
        const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
        const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
        const double pc[6] = {r1, r3, r5, r5, r3, r1};
        const double uc[5] = {r2, r4, r6, r4, r2};
        for (int i = 0; i < ns; ++i) {
            for (int s = 0; s < 6; ++s) {
                TRY(update_momenta_nf2(c, mom, U, X, F, g5DX, am0, pc[s] * eps, maxiter, tol,
                                       precision, iters));
                if (s < 5) updt(c, U, mom, eps * uc[s]);
            }
        }
        return launched(c);
    }
    // leapfrog, src/QED2dHMC.jl:426-449
    TRY(update_momenta_nf2(c, mom, U, X, F, g5DX, am0, 0.5 * eps, maxiter, tol, precision,
                           iters));
    for (int i = 0; i < ns - 1; ++i) {
        updt(c, U, mom, eps);
        TRY(update_momenta_nf2(c, mom, U, X, F, g5DX, am0, eps, maxiter, tol, precision, iters));
    }
    updt(c, U, mom, eps);
    TRY(update_momenta_nf2(c, mom, U, X, F, g5DX, am0, 0.5 * eps, maxiter, tol, precision,
                           iters));
    return launched(c);
}

void gauge_mom(qed2d_ctx *c, double *mom, const double *U, double coef)
{
    launch_force_mom(c->L, (const double2 *)U, nullptr, nullptr, nullptr, mom, c->beta, coef, 0.0,
                     c->stream);
}

// The pure-gauge integrators run each link update U(eu) together with the
// momentum update P(ep) that follows it (launch_updt_gauge_mom), out of place:
// (U, mom) and a scratch pair take turns, and the result is copied back when
// it ends in the scratch pair.
struct PurePair {
    double2 *U[2];
    double *m[2];
    int cur = 0;
};

int pure_pair(qed2d_ctx *c, double *mom, double *U, PurePair *pp)
{
    pp->U[0] = (double2 *)U;
    pp->m[0] = mom;
    TRY(get_spin(c, S_UPP, &pp->U[1]));
    TRY(get_real(c, R_MPP, &pp->m[1]));
    pp->cur = 0;
    return QED2D_OK;
}

void updt_gauge(qed2d_ctx *c, PurePair &pp, double eu, double ep)
{
    const int a = pp.cur, b = a ^ 1;
    PROF(c, QED2D_PROF_UPDT_FLD,
         launch_updt_gauge_mom(c->L, pp.U[a], pp.m[a], pp.U[b], pp.m[b], eu, c->beta, ep,
                               c->stream));
    pp.cur = b;
}

void pure_finish(qed2d_ctx *c, PurePair &pp)
{
    if (pp.cur == 0) return;
    launch_copy(pp.U[0], pp.U[1], 2 * c->L.V, c->stream);
    launch_copy((double2 *)pp.m[0], (const double2 *)pp.m[1], c->L.V, c->stream);
    pp.cur = 0;
}

// OMF4! pure gauge (src/QED2dHMC.jl:239-305): per step P(r1) [U(r2) P(r3)]
// [U(r4) P(r5)] [U(r6) P(r5)] [U(r4) P(r3)] [U(r2) P(r1)], each bracket one pass.
int omf4_pure(qed2d_ctx *c, double *mom, double *U, double eps, int ns)
{
    const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
    const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
    const double pc[6] = {r1, r3, r5, r5, r3, r1};
    const double uc[5] = {r2, r4, r6, r4, r2};
    PurePair pp;
    TRY(pure_pair(c, mom, U, &pp));
    for (int i = 0; i < ns; ++i) {
        gauge_mom(c, pp.m[pp.cur], (const double *)pp.U[pp.cur], pc[0] * eps);
        for (int s = 0; s < 5; ++s) updt_gauge(c, pp, eps * uc[s], pc[s + 1] * eps);
    }
    pure_finish(c, pp);
    return launched(c);
}

// Leapfrog for the pure gauge theory (stage order of src/QED2dHMC.jl:426-449):
// P(eps/2) [U(eps) P(eps)]^(ns-1) [U(eps) P(eps/2)].
int leapfrog_pure(qed2d_ctx *c, double *mom, double *U, double eps, int ns)
{
    PurePair pp;
    TRY(pure_pair(c, mom, U, &pp));
    gauge_mom(c, pp.m[pp.cur], (const double *)pp.U[pp.cur], 0.5 * eps);
    for (int i = 0; i < ns - 1; ++i) updt_gauge(c, pp, eps, eps);
    updt_gauge(c, pp, eps, 0.5 * eps);
    pure_finish(c, pp);
    return launched(c);
}

int metropolis(qed2d_ctx *c, double hini, double hfin, int qzero, const double *U, uint64_t seed,
               uint32_t chain, uint32_t traj, int *accepted)
{
    double pacc;
    bool veto = false;
    if (qzero) {
        hdd sq;
        TRY(observables(c, U, nullptr, nullptr, &sq, nullptr));
        veto = std::fabs(sq.value() / (2.0 * M_PI)) > 0.1;
    }
    pacc = veto ? 0.0 : std::exp(-(hfin - hini));
    if (pacc < 1.0) {
        double r = qed2d_metropolis_uniform(seed, chain, traj);
        *accepted = (pacc > r) ? 1 : 0;
    } else {
        *accepted = 1;
    }
    return QED2D_OK;
}

void key_of(uint64_t seed, uint32_t chain, uint32_t *k0, uint32_t *k1)
{
    *k0 = (uint32_t)seed;
    *k1 = (uint32_t)(seed >> 32) ^ chain;
}

// RHMC force (src/QED2dHMC.jl:516-557)
// Room for the force of n pole solutions: n fields Y and the table.
int force_tab(qed2d_ctx *c, int n)
{
    auto &m = c->mr;
    const long n2 = spinor_elems(c);
    if (n > m.ycap) {
        if (m.yarena) (void)hipFree(m.yarena);
        m.yarena = nullptr;
        m.ycap = 0;
        if (hipMalloc((void **)&m.yarena, sizeof(double2) * n2 * n) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of %d force fields failed", n);
        m.ycap = n;
    }
    if (n > m.tabcap) {
        if (m.tab) (void)hipFree(m.tab);
        if (m.htab) (void)hipHostFree(m.htab);
        m.tab = m.htab = nullptr;
        m.tabcap = 0;
        if (hipMalloc(&m.tab, 4 * sizeof(double) * n) != hipSuccess ||
            hipHostMalloc(&m.htab, 4 * sizeof(double) * n) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "force table allocation failed");
        m.tabcap = n;
    }
    return QED2D_OK;
}

int rhs_set(qed2d_ctx *c, int n, qed2d_batch **out);

// |r|^2 / |b|^2 of the FP32 multi-shift phase of the mixed-precision RHMC
// force (iterated residuals).  Below about 1e-12 the FP32 iterated residuals no
// longer follow the true ones, but the later FP32 steps still shorten the FP64-
// corrected solves that follow: RHMC 128^2 (16 poles, 4 pseudofermions, tol
// 1e-16) ran 2.57, 2.67, 2.79 and 2.76 trajectories/s at 1e-10, 1e-12, 1e-13
// and 1e-14.
constexpr double kRhmcF32Eps = 1e-13;

// The mixed-precision RHMC force solves (qed2d_ctx_set_rhmc_precision): every
// pole of every pseudofermion in one FP32 multi-shift solve to kRhmcF32Eps (or
// eps, if looser), its solutions widened to FP64 as the guesses of one batched
// mixed-precision CG of all the poles' systems (Q^2 + mu^2) x = phi (FP32
// iterations, FP64 reliable updates, each system's FP64 stopping rule
// |r|^2 < eps |phi|^2).  Pole k of the loop over flavours and poles lands in
// c->mr.px + k * n2.
int rhmc_solve_mixed(qed2d_ctx *c, const double *U, int nf, const double *const *b,
                     const double *am0, const int *npole, const double *const *mu, int maxiter,
                     double eps, int *iters)
{
    auto &m = c->mr;
    const long n2 = spinor_elems(c);
    hipStream_t st = c->stream;
    int n = 0;
    for (int j = 0; j < nf; ++j) n += npole[j];
    int it32 = 0;
    TRY(multishift_multi(c, U, nf, b, am0, npole, mu, maxiter, std::max(eps, kRhmcF32Eps), &it32,
                         true));
    if (n > m.pcap) {
        if (m.px) (void)hipFree(m.px);
        if (m.psrc) (void)hipFree(m.psrc);
        m.px = m.psrc = nullptr;
        m.pcap = 0;
        if (hipMalloc((void **)&m.px, sizeof(double2) * n2 * n) != hipSuccess ||
            hipMalloc((void **)&m.psrc, sizeof(double2) * n2 * n) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of %d pole solutions failed", n);
        m.pcap = n;
    }
    std::vector<double> mass(n), shift(n);
    int k = 0;
    for (int j = 0; j < nf; ++j)
        for (int i = 0; i < npole[j]; ++i, ++k) {
            launch_to_f64(m.x32[j][i], m.px + (long)k * n2, n2, st);
            launch_copy(m.psrc + (long)k * n2, (const double2 *)b[j], n2, st);
            mass[k] = am0[j];
            shift[k] = mu[j][i] * mu[j][i];
        }
    TRY(launched(c));
    qed2d_batch *set = nullptr;
    TRY(rhs_set(c, n, &set));
    std::vector<int> its(n);
    std::string err;
    const int rc = batch_rhs_cg_shifted(set, U, (double *)m.px, (const double *)m.psrc,
                                        mass.data(), shift.data(), maxiter, eps,
                                        QED2D_PREC_MIXED, its.data(), c->rows_q, c->rows,
                                        c->rows32, c->rows_k2, c->rows32_k2, c->ru_delta, &err,
                                        true);
    if (rc) return fail(c, rc, "%s", err.c_str());
    int itmax = 0;
    for (int v : its) itmax = std::max(itmax, v);
    if (iters) *iters = it32 + itmax;
    return QED2D_OK;
}

int update_momenta_rhmc(qed2d_ctx *c, double *mom, const double *U, double *X,
                        const std::vector<double2 *> &F, double *g5DX, const double *am0, int nf,
                        const int *npole, const double *const *mu, const double *const *rho,
                        double eps, int maxiter, double tol)
{
    double *facc = nullptr;
    TRY(get_real(c, R_FACC, &facc));
    bool first = true;
    bool fused = (c->multishift == 1 && nf > 0);
    for (int j = 0; j < nf; ++j) fused = fused && npole[j] >= 1 && npole[j] <= kMaxShift;
    if (fused) {
        // all poles of all flavours: one multi-shift solve over the right-hand sides
        std::vector<const double *> bj(nf);
        for (int j = 0; j < nf; ++j) bj[j] = (const double *)F[j];
        int it = 0;
        const bool mixed = (c->rhmc_prec == QED2D_PREC_MIXED);
        if (mixed)
            TRY(rhmc_solve_mixed(c, U, nf, bj.data(), am0, npole, mu, maxiter, tol, &it));
        else
            TRY(multishift_multi(c, U, nf, bj.data(), am0, npole, mu, maxiter, tol, &it));
        // every pole solution's Q x in one launch, and the sum of their forces
        // (in the order of the loop over flavours and poles) in another
        int n = 0;
        for (int j = 0; j < nf; ++j) n += npole[j];
        TRY(force_tab(c, n));
        auto &m = c->mr;
        const long n2 = spinor_elems(c);
        const double2 **hx = (const double2 **)m.htab;
        double2 **hy = (double2 **)(hx + n);
        double *hm = (double *)(hy + n), *hr = hm + n;
        int k = 0;
        for (int j = 0; j < nf; ++j)
            for (int i = 0; i < npole[j]; ++i, ++k) {
                hx[k] = mixed ? m.px + (long)k * n2 : m.x[j][i];
                hy[k] = m.yarena + (long)k * n2;
                hm[k] = 2.0 + am0[j];
                hr[k] = rho[j][i];
            }
        HIPCHK(c, hipMemcpyAsync(m.tab, m.htab, 4 * sizeof(double) * n, hipMemcpyHostToDevice,
                                 c->stream));
        const double2 *const *dx = (const double2 *const *)m.tab;
        double2 *const *dy = (double2 *const *)(dx + n);
        const double *dm = (const double *)(dy + n), *dr = dm + n;
        PROF(c, QED2D_PROF_Q_F64,
             launch_q_apply_tab(c->L, (const double2 *)U, dx, dy, dm, n, c->rows_q, c->stream));
        launch_force_acc_tab(c->L, (const double2 *)U, dx, (const double2 *const *)dy, dr, n,
                             facc, c->stream);
        launch_force_mom(c->L, (const double2 *)U, nullptr, nullptr, n ? facc : nullptr, mom,
                         c->beta, eps, 0.0, c->stream);
        return launched(c);
    }
    if (c->multishift) {
        // all poles of a flavour in one multi-shift solve
        for (int j = 0; j < nf; ++j) {
            if (npole[j] < 1 || npole[j] > kMaxShift) continue;
            TRY(get_shiftbufs(c, npole[j], npole[j]));
            int it = 0;
            TRY(multishift_cg(c, c->shiftbuf.data(), U, (const double *)F[j], am0[j], npole[j],
                              mu[j], maxiter, tol, &it));
            for (int i = 0; i < npole[j]; ++i) {
                const double *Xi = (const double *)c->shiftbuf[i];
                q64(c, U, Xi, g5DX, am0[j]);
                launch_force_acc(c->L, (const double2 *)U, (const double2 *)Xi,
                                 (const double2 *)g5DX, facc, rho[j][i], first ? 1 : 0,
                                 c->stream);
                first = false;
            }
        }
        (void)X;
        launch_force_mom(c->L, (const double2 *)U, nullptr, nullptr, first ? nullptr : facc, mom,
                         c->beta, eps, 0.0, c->stream);
        return launched(c);
    }
    for (int j = 0; j < nf; ++j) {
        int guess = 0;
        for (int i = 0; i < npole[j]; ++i) {
            int it = 0;
            TRY(cg_any(c, X, U, (const double *)F[j], am0[j], mu[j][i], QED2D_OP_SQR_MUSQ, maxiter,
                       tol, guess, QED2D_PREC_F64, &it));
            guess = 1;
            q64(c, U, X, g5DX, am0[j]);
            launch_force_acc(c->L, (const double2 *)U, (const double2 *)X, (const double2 *)g5DX,
                             facc, rho[j][i], first ? 1 : 0, c->stream);
            first = false;
        }
    }
    launch_force_mom(c->L, (const double2 *)U, nullptr, nullptr, first ? nullptr : facc, mom,
                     c->beta, eps, 0.0, c->stream);
    return launched(c);
}

int leapfrog_rhmc(qed2d_ctx *c, double *mom, double *U, double *X, const std::vector<double2 *> &F,
                  double *g5DX, const double *am0, int nf, const int *npole,
                  const double *const *mu, const double *const *rho, double eps, int ns,
                  int maxiter, double tol)
{
    TRY(update_momenta_rhmc(c, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, 0.5 * eps, maxiter,
                            tol));
    for (int i = 0; i < ns - 1; ++i) {
        updt(c, U, mom, eps);
        TRY(update_momenta_rhmc(c, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, maxiter,
                                tol));
    }
    updt(c, U, mom, eps);
    return update_momenta_rhmc(c, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, 0.5 * eps,
                               maxiter, tol);
}

// OMF4 with the RHMC force: a build extension (the reference's MD! has no
// OMF4 method for RHMC, src/QED2dHMC.jl:12-16); stage order of :307-422.
int omf4_rhmc(qed2d_ctx *c, double *mom, double *U, double *X, const std::vector<double2 *> &F,
              double *g5DX, const double *am0, int nf, const int *npole, const double *const *mu,
              const double *const *rho, double eps, int ns, int maxiter, double tol)
{
    const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
    const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
    const double pc[6] = {r1, r3, r5, r5, r3, r1};
    const double uc[5] = {r2, r4, r6, r4, r2};
    for (int i = 0; i < ns; ++i) {
        for (int s = 0; s < 6; ++s) {
            TRY(update_momenta_rhmc(c, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, pc[s] * eps,
                                    maxiter, tol));
            if (s < 5) updt(c, U, mom, eps * uc[s]);
        }
    }
    return launched(c);
}

int md_rhmc(qed2d_ctx *c, int integrator, double *mom, double *U, double *X,
            const std::vector<double2 *> &F, double *g5DX, const double *am0, int nf,
            const int *npole, const double *const *mu, const double *const *rho, double eps,
            int ns, int maxiter, double tol)
{
    if (integrator == QED2D_OMF4)
        return omf4_rhmc(c, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, ns, maxiter, tol);
    return leapfrog_rhmc(c, mom, U, X, F, g5DX, am0, nf, npole, mu, rho, eps, ns, maxiter, tol);
}

int multicg_impl(qed2d_ctx *c, double *so, const double *U, const double *si, double am0, int n,
                 const double *mu, const double *rho, int maxiter, double eps)
{
    double2 *aux;
    TRY(get_spin(c, S_AUX, &aux));
    const long n2 = spinor_elems(c);
    if (c->multishift && n >= 1 && n <= kMaxShift) {
        TRY(get_shiftbufs(c, n, n));
        int it = 0;
        TRY(multishift_cg(c, c->shiftbuf.data(), U, si, am0, n, mu, maxiter, eps, &it));
        launch_copy((double2 *)so, (const double2 *)si, n2, c->stream);
        for (int j = 0; j < n; ++j)
            launch_scale_add((double2 *)so, 1.0, c->shiftbuf[j], rho[j], n2, c->stream);
        return launched(c);
    }
    launch_copy((double2 *)so, (const double2 *)si, n2, c->stream);
    for (int j = 0; j < n; ++j) {
        int it;
        TRY(cg_any(c, (double *)aux, U, si, am0, mu[j], QED2D_OP_SQR_MUSQ, maxiter, eps, 0,
                   QED2D_PREC_F64, &it));
        launch_scale_add((double2 *)so, 1.0, aux, rho[j], n2, c->stream);
    }
    return launched(c);
}

int gen_pf_impl(qed2d_ctx *c, double *F, const double *U, const double *X, double am0, int n,
                const double *mu, const double *nu, int maxiter, double tol)
{
    double2 *aux;
    TRY(get_spin(c, S_TMP2, &aux));
    const long n2 = spinor_elems(c);
    hipStream_t st = c->stream;
    launch_copy((double2 *)F, (const double2 *)X, n2, st);
    for (int i = 0; i < n; ++i) {
        launch_copy(aux, (const double2 *)F, n2, st);
        q64(c, U, (const double *)aux, F, am0);
        launch_axpy_c((double2 *)F, aux, 0.0, mu[i], n2, st);
        launch_copy(aux, (const double2 *)F, n2, st);
        int it;
        TRY(cg_any(c, F, U, (const double *)aux, am0, nu[i], QED2D_OP_SQR_MUSQ, maxiter, tol, 0,
                   QED2D_PREC_F64, &it));
        launch_copy(aux, (const double2 *)F, n2, st);
        q64(c, U, (const double *)aux, F, am0);
        launch_axpy_c((double2 *)F, aux, 0.0, -nu[i], n2, st);
    }
    return launched(c);
}

// The set of n right-hand sides on this lattice (created on first use, kept
// until the context is destroyed).
int rhs_set(qed2d_ctx *c, int n, qed2d_batch **out)
{
    if ((int)c->rhsets.size() <= n) c->rhsets.resize(n + 1, nullptr);
    if (!c->rhsets[n]) {
        const int rc = qed2d_batch_create_rhs(&c->rhsets[n], c->L.L1, c->L.L2, n, c->beta,
                                              c->L.bc, (void *)c->stream);
        if (rc) {
            c->rhsets[n] = nullptr;
            return fail(c, rc, "cannot create a set of %d right-hand sides", n);
        }
    }
    // (the set's mixed-precision solves follow the context's reliable-update
    // mode, so qed2d_cg_mrhs reproduces qed2d_cg in either)
    (void)qed2d_batch_set_reliable_mode(c->rhsets[n], c->ru_mode);
    *out = c->rhsets[n];
    return QED2D_OK;
}

// gen_pf_impl for nf flavours at once, F[j] holding X_j on entry: the same
// steps per flavour and pole, with pole i's shifted solves of every flavour
// that has an i-th pole run as one FP64 solve of a right-hand-side set (a mass
// and a shift per right-hand side), which gives each flavour the bits of its
// own solve.
int gen_pf_multi(qed2d_ctx *c, const std::vector<double2 *> &F, const double *U,
                 const double *am0, int nf, const int *npole, const double *const *mu,
                 const double *const *nu, int maxiter, double tol)
{
    const long n2 = spinor_elems(c);
    hipStream_t st = c->stream;
    TRY(force_tab(c, 2 * nf));
    double2 *A = c->mr.yarena, *Y = A + (long)nf * n2;  // right-hand sides, solutions
    int pmax = 0;
    for (int j = 0; j < nf; ++j) pmax = std::max(pmax, npole[j]);
    std::vector<double> m(nf), s(nf);
    for (int i = 0; i < pmax; ++i) {
        int k = 0;
        for (int j = 0; j < nf; ++j) {
            if (npole[j] <= i) continue;
            double2 *a = A + (long)k * n2;
            launch_copy(a, F[j], n2, st);
            q64(c, U, (const double *)a, (double *)F[j], am0[j]);
            launch_axpy_c(F[j], a, 0.0, mu[j][i], n2, st);
            launch_copy(a, F[j], n2, st);
            m[k] = am0[j];
            s[k] = nu[j][i] * nu[j][i];
            ++k;
        }
        qed2d_batch *set = nullptr;
        TRY(rhs_set(c, k, &set));
        std::string err;
        const int rc = batch_rhs_cg_shifted(set, U, (double *)Y, (const double *)A, m.data(),
                                            s.data(), maxiter, tol, QED2D_PREC_F64, nullptr,
                                            c->rows_q, c->rows, c->rows32, c->rows_k2,
                                            c->rows32_k2, c->ru_delta, &err);
        if (rc) return fail(c, rc, "%s", err.c_str());
        k = 0;
        for (int j = 0; j < nf; ++j) {
            if (npole[j] <= i) continue;
            const double2 *y = Y + (long)k * n2;
            q64(c, U, (const double *)y, (double *)F[j], am0[j]);
            launch_axpy_c(F[j], y, 0.0, -nu[j][i], n2, st);
            ++k;
        }
    }
    return launched(c);
}

// so = (A / r_b) (si + sum_j rho_j (Q^2 + mu_j^2)^{-1} si)  (src/QED2dRHMC.jl:261-271)
int r_impl(qed2d_ctx *c, double *so, const double *U, const double *si, double am0, int n,
           const double *mu, const double *rho, double A, double r_b, int maxiter, double eps)
{
    TRY(multicg_impl(c, so, U, si, am0, n, mu, rho, maxiter, eps));
    // so .= r_b^(-1) * A * so  (src/QED2dRHMC.jl:268)
    const double f = std::pow(r_b, -1.0) * A;
    launch_scale_add((double2 *)so, f, (const double2 *)so, 0.0, spinor_elems(c), c->stream);
    return launched(c);
}

// Z = (Q^2 R^2 - 1) X  (src/QED2dRHMC.jl:174-186).  Z must not alias X; the
// last Q application and the subtraction of X share one stencil pass.
int luscher_z_impl(qed2d_ctx *c, double *Z, const double *U, const double *X, double am0,
                   int n, const double *mu, const double *rho, double A, double r_b,
                   int maxiter, double tol)
{
    double2 *t1, *t2;
    TRY(get_spin(c, S_RW_T1, &t1));
    TRY(get_spin(c, S_RW_T2, &t2));
    TRY(r_impl(c, (double *)t1, U, X, am0, n, mu, rho, A, r_b, maxiter, tol));
    TRY(r_impl(c, (double *)t2, U, (const double *)t1, am0, n, mu, rho, A, r_b, maxiter, tol));
    q64(c, U, (const double *)t2, (double *)t1, am0);
    launch_q_shift_f64(c->L, (const double2 *)U, t1, (const double2 *)X, (double2 *)Z, am0, -1.0,
                       c->rows_q, c->stream);
    return launched(c);
}

bool tile_ok(int r) { return r == 1 || r == 2 || r == 4 || r == 8 || r == 16; }

// The tallest of {big, big/2, ..., 1} that leaves a launch of at least 256
// blocks of `sites` sites per row of a tile (one block per CU).
int fill_tile(long V, int big, int sites)
{
    int R = big;
    while (R > 1 && V / ((long)sites * R) < 256) R >>= 1;
    return R;
}

// Tile heights measured best on MI355X (pipelined marches): applies are
// bandwidth bound at any height >= 2; the FP64 CG kernels need tall tiles on
// lattices that fit the 256 MB Infinity Cache and medium ones above.  Below
// 1024^2 a tall tile leaves most CUs idle (128^2 at 16 rows is 4 blocks), so
// the solver tiles shrink until a launch has a block per CU (128^2 RHMC
// trajectory: 0.83 s at 16 rows, 0.43 s at 1).
void default_tiles(qed2d_ctx *c)
{
    const long V = c->L.V;
    c->rows_q = 2;
    c->rows_q32 = 2;
    c->rows = (V <= (1L << 21)) ? 16 : 8;
    c->rows32 = 8;
    if (V < (1L << 20)) {
        c->rows = fill_tile(V, 16, kBlock);
        c->rows32 = fill_tile(V, 8, 2 * kBlock);
    }
}

// Mean duration in microseconds of `reps` runs of fn (after one warm-up run)
// on the context's stream; fn's nonzero statuses are ignored (the timed CG
// runs are cut off by their iteration limit on purpose).
int time_runs(qed2d_ctx *c, int reps, const std::function<void()> &fn, double *us)
{
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    fn();
    HIPCHK(c, hipEventRecord(e0, c->stream));
    for (int i = 0; i < reps; ++i) fn();
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *us = 1e3 * ms / reps;
    return QED2D_OK;
}

// The autotuner: times every tile height of each kernel family on this
// context's lattice (random links and source, scratch buffers freed on
// return) and keeps the fastest.  CG families are timed over `iters`
// iterations of a solve that cannot converge (tolerance 0).
int autotune_impl(qed2d_ctx *c, int iters)
{
    const Latt &L = c->L;
    const long n2 = spinor_elems(c);
    hipStream_t st = c->stream;
    void *mem = nullptr;
    if (hipMalloc(&mem, 3 * sizeof(double2) * n2 + 3 * sizeof(float2) * n2) != hipSuccess)
        return fail(c, QED2D_ENOMEM, "autotune scratch allocation failed");
    double2 *U = (double2 *)mem, *b = U + n2, *x = b + n2;
    float2 *Uf = (float2 *)(x + n2), *bf = Uf + n2, *xf = bf + n2;
    launch_hot_start(L, U, 0x7475u, 0x6e65u, st);
    launch_fill_cgauss(L, b, 0x7475u, 0x6e65u, 0, kPurposeVec, 0, st);
    launch_to_f32(U, Uf, n2, st);
    launch_to_f32(b, bf, n2, st);
    const int cand[5] = {1, 2, 4, 8, 16};
    const int save_sync = c->sync;
    c->sync = 0;
    int rc = QED2D_OK;
    double best, us;
    int pick;
    // FP64 apply
    best = 1e30, pick = c->rows_q;
    for (int R : cand) {
        rc = time_runs(c, 5, [&] { launch_q_apply_f64(L, U, b, x, 0.1, R, st); }, &us);
        if (rc) break;
        if (us < best) best = us, pick = R;
    }
    if (!rc) c->rows_q = pick;
    // FP32 apply
    best = 1e30, pick = c->rows_q32;
    for (int R : cand) {
        if (rc) break;
        rc = time_runs(c, 5, [&] { launch_q_apply_f32(L, Uf, bf, xf, 0.1f, R, st); }, &us);
        if (us < best) best = us, pick = R;
    }
    if (!rc) c->rows_q32 = pick;
    // CG kernels: solves of `iters` iterations queued in one go (one look at
    // the state, at the end; no reliable update before the last iteration),
    // after one untimed solve the mean of three
    SolveArgs a;
    a.am0 = 0.1;
    a.eps = 0.0;
    a.maxiter = iters;
    a.chunk0 = iters + 1;
    const double keep_delta = c->ru_delta;
    c->ru_delta = 1e-30;
    int it;
    // (the two kernels of an iteration are timed together, every pair of
    // heights: K1 and K2 may prefer different ones, and one's height changes
    // the other's time through the caches.  Two passes over all pairs, the
    // times of a pair added: the pairs lie within a few per cent of each
    // other, and a drift of the clocks during one pass would otherwise pick.)
    for (int prec = 0; prec < 2 && !rc; ++prec) {
        int &tile = prec ? c->rows32 : c->rows;
        int &tile2 = prec ? c->rows32_k2 : c->rows_k2;
        const int keep = tile, keep2 = tile2;
        double tsum[5][5] = {};
        for (int pass = 0; pass < 2 && !rc; ++pass) {
            for (int i = 0; i < 5 && !rc; ++i) {
                if (prec == 0 && cand[i] == 1) continue;
                for (int j = 0; j < 5; ++j) {
                    if (prec == 0 && cand[j] == 1) continue;
                    tile = cand[i];
                    tile2 = cand[j];
                    rc = time_runs(c, 3, [&] {
                        if (prec) (void)cg_mixed(c, (double *)x, (const double *)U, (const double *)b, a, &it);
                        else (void)cg_f64(c, (double *)x, (const double *)U, (const double *)b, a, &it);
                    }, &us);
                    if (rc) break;
                    tsum[i][j] += us;
                }
            }
        }
        best = 1e30;
        int pick2 = keep2;
        pick = keep;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j)
                if (tsum[i][j] > 0.0 && tsum[i][j] < best)
                    best = tsum[i][j], pick = cand[i], pick2 = cand[j];
        tile = rc ? keep : pick;
        tile2 = rc ? keep2 : (pick2 == pick ? 0 : pick2);
    }
    c->ru_delta = keep_delta;
    c->sync = save_sync;
    (void)hipStreamSynchronize(st);
    (void)hipFree(mem);
    if (rc) return rc;
    c->err.clear();
    return launched(c);
}

bool overlaps(const void *a, const void *b) { return a && b && a == b; }

}  // namespace

namespace qed2d {
int ctx_hamiltonian_dd(qed2d_ctx *c, const double *mom, const double *U, double *hi, double *lo)
{
    hdd h;
    TRY(hamiltonian_dd(c, mom, U, &h));
    *hi = h.hi;
    *lo = h.lo;
    return QED2D_OK;
}

int ctx_dot_dd(qed2d_ctx *c, const double *a, const double *b, double *hi, double *lo)
{
    hdd h;
    TRY(dot_dd_host(c, a, b, &h));
    *hi = h.hi;
    *lo = h.lo;
    return QED2D_OK;
}

// Queued forms of the two: the reduction's result lands in *dst (pinned host
// memory) when the stream gets there; after a synchronisation the *_from forms
// give the values the synchronous calls return.
int ctx_queue_hamiltonian(qed2d_ctx *c, const double *mom, const double *U, RedOut *dst)
{
    launch_observables(c->L, (const double2 *)U, mom, nullptr, nullptr, c->red, c->part,
                       c->stream);
    TRY(launched(c));
    HIPCHK(c, hipMemcpyAsync(dst, c->red, sizeof(RedOut), hipMemcpyDeviceToHost, c->stream));
    return QED2D_OK;
}
int ctx_queue_dot_dd(qed2d_ctx *c, const double *a, const double *b, RedOut *dst)
{
    launch_dot_dd((const double2 *)a, (const double2 *)b, spinor_elems(c), c->red, c->part,
                  c->stream);
    TRY(launched(c));
    HIPCHK(c, hipMemcpyAsync(dst, c->red, sizeof(RedOut), hipMemcpyDeviceToHost, c->stream));
    return QED2D_OK;
}
void ctx_hamiltonian_from(const qed2d_ctx *c, const RedOut &r, double *hi, double *lo)
{
    hdd sp, sm;
    sp.add(r.v[0], r.v[1]);
    sm.add(r.v[4], r.v[5]);
    const hdd h = hamiltonian_from(c, sp, sm);
    *hi = h.hi;
    *lo = h.lo;
}
void ctx_dot_dd_from(const RedOut &r, double *hi, double *lo)
{
    hdd h;
    h.add(r.v[0], r.v[1]);
    *hi = h.hi;
    *lo = h.lo;
}
}  // namespace qed2d

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int qed2d_ctx_create(qed2d_ctx **out, int L1, int L2, double beta, int bc_flags, void *stream)
{
    if (!out) return QED2D_EINVAL;
    *out = nullptr;
    if (L1 < 2 || L2 < 2 || (long)L1 * L2 > (1L << 31)) return QED2D_EINVAL;
    if (bc_flags & ~QED2D_BC_APERIODIC_T) return QED2D_EINVAL;
    qed2d_ctx *c = new qed2d_ctx();
    c->L.L1 = L1;
    c->L.L2 = L2;
    c->L.V = (long)L1 * L2;
    default_tiles(c);
    c->L.bc = bc_flags;
    c->beta = beta;
    c->stream = (hipStream_t)stream;
    hipError_t he = hipGetDevice(&c->device);
    if (he != hipSuccess) {
        g_create_err = std::string("hipGetDevice: ") + hipGetErrorString(he);
        delete c;
        return QED2D_EHIP;
    }
    // partial sums: the largest marching grid (rows = 1) needs 2 doubles per
    // block, the observables 6 per block of at most 1024.
    // generic reductions: at most 1024 blocks x 6 doubles; solver kernels: per
    // block of the largest marching grid (rows = 1), 2 x 2 (tt) + 2 (rr) + 1 (b).
    c->part_len = 8 * 1024;
    const long nbmax = march_grid(c->L, 1).nblocks();
    bool ok = hipMalloc((void **)&c->part, sizeof(double) * c->part_len) == hipSuccess &&
              hipMalloc((void **)&c->cgpart, sizeof(double) * 7 * nbmax) == hipSuccess &&
              hipMalloc((void **)&c->state, sizeof(SolverState)) == hipSuccess &&
              hipMalloc((void **)&c->red, sizeof(RedOut)) == hipSuccess &&
              hipHostMalloc((void **)&c->hstate, sizeof(SolverState), hipHostMallocDefault) ==
                  hipSuccess &&
              hipHostMalloc((void **)&c->hring, kRing * sizeof(SolverState),
                            hipHostMallocCoherent) == hipSuccess &&
              hipHostMalloc((void **)&c->hseq, kRing * sizeof(int), hipHostMallocCoherent) ==
                  hipSuccess &&
              hipHostMalloc((void **)&c->hred, sizeof(RedOut), hipHostMallocDefault) == hipSuccess;
    if (ok) std::memset(c->hseq, 0, kRing * sizeof(int));
    if (!ok) {
        he = hipGetLastError();
        g_create_err = std::string("allocation of the context's buffers failed: ") +
                       hipGetErrorString(he);
        qed2d_ctx_destroy(c);
        return QED2D_ENOMEM;
    }
    c->P.tt[0] = c->cgpart;
    c->P.tt[1] = c->cgpart + 2 * nbmax;
    c->P.rr[0] = c->cgpart + 4 * nbmax;
    c->P.rr[1] = c->cgpart + 5 * nbmax;
    c->P.b = c->cgpart + 6 * nbmax;
    he = hipMemsetAsync(c->state, 0, sizeof(SolverState), c->stream);
    if (he == hipSuccess) he = hipMemsetAsync(c->red, 0, sizeof(RedOut), c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    if (he != hipSuccess) {
        g_create_err = std::string("clearing the solver state failed: ") + hipGetErrorString(he);
        qed2d_ctx_destroy(c);
        return QED2D_EHIP;
    }
    *out = c;
    return QED2D_OK;
}

void qed2d_ctx_destroy(qed2d_ctx *c)
{
    if (!c) return;
    (void)(c->stream ? hipStreamSynchronize(c->stream) : hipDeviceSynchronize());
    for (auto &p : c->spin) if (p) (void)hipFree(p);
    for (auto &p : c->fspin) if (p) (void)hipFree(p);
    for (auto &p : c->real) if (p) (void)hipFree(p);
    for (auto p : c->flav) if (p) (void)hipFree(p);
    for (auto p : c->shiftbuf) if (p) (void)hipFree(p);
    for (auto p : c->psbuf) if (p) (void)hipFree(p);
    for (auto p : c->rhsets) if (p) qed2d_batch_destroy(p);
    for (auto &v : c->mr.x) for (auto p : v) if (p) (void)hipFree(p);
    for (auto &v : c->mr.ps) for (auto p : v) if (p) (void)hipFree(p);
    for (auto &v : c->mr.x32) for (auto p : v) if (p) (void)hipFree(p);
    for (auto &v : c->mr.ps32) for (auto p : v) if (p) (void)hipFree(p);
    if (c->mr.px) (void)hipFree(c->mr.px);
    if (c->mr.psrc) (void)hipFree(c->mr.psrc);
    for (void *p : {(void *)c->mr.dch, (void *)c->mr.blk0, (void *)c->mr.shape, (void *)c->mr.M,
                    (void *)c->mr.part, (void *)c->mr.cnt, (void *)c->mr.arena,
                    (void *)c->mr.yarena, c->mr.tab})
        if (p) (void)hipFree(p);
    if (c->mr.hM) (void)hipHostFree(c->mr.hM);
    if (c->mr.htab) (void)hipHostFree(c->mr.htab);
    if (c->ms) (void)hipFree(c->ms);
    if (c->hms) (void)hipHostFree(c->hms);
    if (c->part) (void)hipFree(c->part);
    if (c->cgpart) (void)hipFree(c->cgpart);
    if (c->state) (void)hipFree(c->state);
    if (c->red) (void)hipFree(c->red);
    if (c->hstate) (void)hipHostFree(c->hstate);
    if (c->hring) (void)hipHostFree(c->hring);
    if (c->hseq) (void)hipHostFree(c->hseq);
    if (c->hred) (void)hipHostFree(c->hred);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : c->evpool) (void)hipEventDestroy(e);
    delete c;
}

const char *qed2d_last_error(const qed2d_ctx *c)
{
    return c ? c->err.c_str() : g_create_err.c_str();
}

int qed2d_ctx_set_beta(qed2d_ctx *c, double beta)
{
    if (!c) return QED2D_EINVAL;
    c->beta = beta;
    return QED2D_OK;
}

int qed2d_ctx_set_sync(qed2d_ctx *c, int sync)
{
    if (!c) return QED2D_EINVAL;
    c->sync = sync ? 1 : 0;
    return QED2D_OK;
}

int qed2d_ctx_set_tile(qed2d_ctx *c, int rows)
{
    if (!c) return QED2D_EINVAL;
    if (!tile_ok(rows))
        return fail(c, QED2D_EINVAL, "tile rows must be 1, 2, 4, 8 or 16 (got %d)", rows);
    c->rows = rows;
    c->rows_q = rows;
    return QED2D_OK;
}

int qed2d_ctx_set_tiles(qed2d_ctx *c, int q64, int q32, int cg64, int cg32)
{
    if (!c) return QED2D_EINVAL;
    for (int v : {q64, q32, cg64, cg32})
        if (v != 0 && !tile_ok(v))
            return fail(c, QED2D_EINVAL, "tile rows must be 0, 1, 2, 4, 8 or 16 (got %d)", v);
    if (q64) c->rows_q = q64;
    if (q32) c->rows_q32 = q32;
    if (cg64) {
        c->rows = cg64;
        c->rows_k2 = 0;
    }
    if (cg32) {
        c->rows32 = cg32;
        c->rows32_k2 = 0;
    }
    return QED2D_OK;
}

int qed2d_ctx_set_tiles_k2(qed2d_ctx *c, int cg64, int cg32)
{
    if (!c) return QED2D_EINVAL;
    for (int v : {cg64, cg32})
        if (v != 0 && !tile_ok(v))
            return fail(c, QED2D_EINVAL, "tile rows must be 0, 1, 2, 4, 8 or 16 (got %d)", v);
    c->rows_k2 = cg64;
    c->rows32_k2 = cg32;
    return QED2D_OK;
}

int qed2d_ctx_get_tiles_k2(qed2d_ctx *c, int *tiles)
{
    if (!c || !tiles) return QED2D_EINVAL;
    tiles[0] = c->rows_k2 ? c->rows_k2 : c->rows;
    tiles[1] = c->rows32_k2 ? c->rows32_k2 : c->rows32;
    return QED2D_OK;
}

int qed2d_ctx_get_tiles(qed2d_ctx *c, int *tiles)
{
    if (!c || !tiles) return QED2D_EINVAL;
    tiles[0] = c->rows_q;
    tiles[1] = c->rows_q32;
    tiles[2] = c->rows;
    tiles[3] = c->rows32;
    return QED2D_OK;
}

int qed2d_ctx_set_reliable_delta(qed2d_ctx *c, double delta)
{
    if (!c) return QED2D_EINVAL;
    if (!(delta > 0.0 && delta < 1.0))
        return fail(c, QED2D_EINVAL, "reliable-update delta must be in (0, 1)");
    c->ru_delta = delta;
    return QED2D_OK;
}

int qed2d_ctx_set_reliable_mode(qed2d_ctx *c, int on_device)
{
    if (!c) return QED2D_EINVAL;
    if (on_device != 0 && on_device != 1)
        return fail(c, QED2D_EINVAL, "reliable-update mode must be 0 or 1 (got %d)", on_device);
    c->ru_mode = on_device;
    return QED2D_OK;
}

int qed2d_ctx_set_rhmc_precision(qed2d_ctx *c, int precision)
{
    if (!c) return QED2D_EINVAL;
    if (precision != QED2D_PREC_F64 && precision != QED2D_PREC_MIXED)
        return fail(c, QED2D_EINVAL, "RHMC force precision must be QED2D_PREC_F64 or "
                    "QED2D_PREC_MIXED (got %d)", precision);
    if (precision == QED2D_PREC_MIXED && !pair_ok(c->L))
        return fail(c, QED2D_EINVAL, "the mixed-precision RHMC force needs an even L1 >= 4 "
                    "(L1 = %d)", c->L.L1);
    c->rhmc_prec = precision;
    return QED2D_OK;
}

int qed2d_ctx_set_tile_f32(qed2d_ctx *c, int rows)
{
    if (!c) return QED2D_EINVAL;
    if (!tile_ok(rows))
        return fail(c, QED2D_EINVAL, "tile rows must be 1, 2, 4, 8 or 16 (got %d)", rows);
    c->rows32 = rows;
    c->rows_q32 = rows;
    return QED2D_OK;
}

int qed2d_ctx_autotune(qed2d_ctx *c, int iters, int *tiles)
{
    if (!c) return QED2D_EINVAL;
    if (iters < 4) iters = 4;
    TRY(autotune_impl(c, iters));
    if (tiles) qed2d_ctx_get_tiles(c, tiles);
    return finish(c);
}

int qed2d_hbm_stream(qed2d_ctx *c, long nbytes, double *gbs)
{
    if (!c || !gbs) return QED2D_EINVAL;
    const long n = nbytes / (long)sizeof(double2);
    if (n < kBlock) return fail(c, QED2D_EINVAL, "stream arrays too small");
    void *mem = nullptr;
    if (hipMalloc(&mem, 3 * n * sizeof(double2) + kStreamReadBlocks * sizeof(double)) !=
        hipSuccess)
        return fail(c, QED2D_ENOMEM, "hipMalloc of 3 x %ld bytes failed", nbytes);
    double2 *a = (double2 *)mem, *b = a + n, *d = b + n;
    double *part = (double *)(d + n);
    HIPCHK(c, hipMemsetAsync(a, 0x3f, n * sizeof(double2), c->stream));
    HIPCHK(c, hipMemsetAsync(b, 0x40, n * sizeof(double2), c->stream));
    const double moved[4] = {2.0, 3.0, 1.0, 2.0};  // arrays read + written per element
    // (the 16-array copy moves 16 whole pieces of 256-element blocks)
    const long n16 = 16 * (n / (16 * kBlock) * kBlock);
    int rc = QED2D_OK;
    for (int w = 0; w < 4 && !rc; ++w) {
        double us;
        const long nw = (w == 3) ? n16 : n;
        rc = time_runs(c, 10, [&] { launch_stream(w, a, b, d, part, nw, c->stream); }, &us);
        gbs[w] = moved[w] * nw * sizeof(double2) / (us * 1e-6) / 1e9;
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(mem);
    return rc ? rc : finish(c);
}

int qed2d_ctx_last_dh(const qed2d_ctx *c, double *dH)
{
    if (!c || !dH) return QED2D_EINVAL;
    *dH = c->last_dh;
    return QED2D_OK;
}

int qed2d_synchronize(qed2d_ctx *c)
{
    if (!c) return QED2D_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return QED2D_OK;
}

#define CHECK_CTX(c) \
    if (!(c)) return QED2D_EINVAL
#define NEED(c, p)                                                         \
    if (!(p)) return fail((c), QED2D_EINVAL, "%s: null pointer %s", __func__, #p)
#define NOALIAS(c, a, b)                                                            \
    if (overlaps((a), (b)))                                                         \
        return fail((c), QED2D_EINVAL, "%s: %s aliases %s", __func__, #a, #b)

int qed2d_gamm5Dw(qed2d_ctx *c, double *so, const double *U, const double *si, double am0)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si); NOALIAS(c, so, U);
    q64(c, U, si, so, am0);
    return finish(c);
}

int qed2d_gamm5Dw_f32(qed2d_ctx *c, float *so, const float *U, const float *si, float am0)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    launch_q_apply_f32(c->L, (const float2 *)U, (const float2 *)si, (float2 *)so, am0, c->rows_q32,
                       c->stream);
    return finish(c);
}

int qed2d_gamm5Dw_mrhs(qed2d_ctx *c, double *so, const double *U, const double *si, int nrhs,
                       double am0)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    if (nrhs < 1) return fail(c, QED2D_EINVAL, "nrhs must be >= 1");
    if ((long)march_grid(c->L, c->rows_q).nblocks() * nrhs > 0x7fffffffL)
        return fail(c, QED2D_EINVAL, "nrhs %d too large for one launch", nrhs);
    launch_q_apply_mrhs_f64(c->L, (const double2 *)U, (const double2 *)si, (double2 *)so, nrhs,
                            am0, c->rows_q, c->stream);
    return finish(c);
}

int qed2d_gamm5Dw_mrhs_f32(qed2d_ctx *c, float *so, const float *U, const float *si, int nrhs,
                           float am0)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    if (nrhs < 1) return fail(c, QED2D_EINVAL, "nrhs must be >= 1");
    if ((long)march_grid(c->L, c->rows_q32).nblocks() * nrhs > 0x7fffffffL)
        return fail(c, QED2D_EINVAL, "nrhs %d too large for one launch", nrhs);
    launch_q_apply_mrhs_f32(c->L, (const float2 *)U, (const float2 *)si, (float2 *)so, nrhs, am0,
                            c->rows_q32, c->stream);
    return finish(c);
}

int qed2d_gamm5(qed2d_ctx *c, double *so, const double *si)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, si);
    launch_gamm5(c->L, (const double2 *)si, (double2 *)so, c->stream);
    return finish(c);
}

int qed2d_Dw(qed2d_ctx *c, double *so, const double *U, const double *si, double am0)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    double2 *tmp;
    TRY(get_spin(c, S_TMP, &tmp));
    q64(c, U, si, (double *)tmp, am0);
    launch_gamm5(c->L, tmp, (double2 *)so, c->stream);
    return finish(c);
}

int qed2d_gamm5Dw_sqr(qed2d_ctx *c, double *so, const double *U, const double *si, double am0)
{
    return qed2d_gamm5Dw_sqr_musq(c, so, nullptr, U, si, am0, 0.0);
}

int qed2d_gamm5Dw_sqr_msq(qed2d_ctx *c, double *so, double *tmp, const double *U,
                          const double *si, double am0)
{
    return qed2d_gamm5Dw_sqr_musq(c, so, tmp, U, si, am0, 0.0);
}

int qed2d_gamm5Dw_sqr_musq(qed2d_ctx *c, double *so, double *tmp, const double *U,
                           const double *si, double am0, double mu)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si); NOALIAS(c, tmp, si); NOALIAS(c, tmp, so);
    if (!tmp) {
        double2 *t;
        TRY(get_spin(c, S_TMP, &t));
        tmp = (double *)t;
    }
    qsq64(c, U, si, so, tmp, am0, mu * mu);
    return finish(c);
}

int qed2d_gamm5Dw_sqr_sqr(qed2d_ctx *c, double *so, const double *U, const double *si,
                          double am0)
{
    return qed2d_gamm5Dw_sqr_sqr_musq(c, so, nullptr, U, si, am0, 0.0);
}

int qed2d_gamm5Dw_sqr_sqr_musq(qed2d_ctx *c, double *so, double *tmp, const double *U,
                               const double *si, double am0, double mu)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si); NOALIAS(c, tmp, si); NOALIAS(c, tmp, so);
    double2 *t1, *t2;
    TRY(get_spin(c, S_TMP, &t1));
    TRY(get_spin(c, S_TMP2, &t2));
    if (tmp) t1 = (double2 *)tmp;
    q64(c, U, si, (double *)t1, am0);
    q64(c, U, (const double *)t1, (double *)t2, am0);
    q64(c, U, (const double *)t2, (double *)t1, am0);
    launch_q_shift_f64(c->L, (const double2 *)U, t1, (const double2 *)si, (double2 *)so, am0,
                       mu * mu, c->rows_q, c->stream);
    return finish(c);
}

int qed2d_tr_dQwdU(qed2d_ctx *c, double *frc, const double *U, const double *X,
                   const double *g5DwX)
{
    CHECK_CTX(c);
    NEED(c, frc); NEED(c, U); NEED(c, X); NEED(c, g5DwX);
    launch_tr_dQwdU(c->L, (const double2 *)U, (const double2 *)X, (const double2 *)g5DwX, frc,
                    c->stream);
    return finish(c);
}

int qed2d_krnl_force(qed2d_ctx *c, double *frc1, double *frc2, const double *U)
{
    CHECK_CTX(c);
    NEED(c, frc1); NEED(c, frc2); NEED(c, U);
    NOALIAS(c, frc1, frc2);
    launch_krnl_force(c->L, (const double2 *)U, c->beta, frc1, frc2, c->stream);
    return finish(c);
}

int qed2d_updt_fld(qed2d_ctx *c, double *U, const double *mom, double eps)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, mom);
    updt(c, U, mom, eps);
    return finish(c);
}

int qed2d_force_mom_update(qed2d_ctx *c, double *mom, const double *U, const double *X,
                           const double *g5DwX, double eps)
{
    CHECK_CTX(c);
    NEED(c, mom); NEED(c, U);
    if (X && !g5DwX) return fail(c, QED2D_EINVAL, "g5DwX required with X");
    launch_force_mom(c->L, (const double2 *)U, (const double2 *)X, (const double2 *)g5DwX,
                     nullptr, mom, c->beta, eps, 1.0, c->stream);
    return finish(c);
}

int qed2d_plaquette_field(qed2d_ctx *c, double *plx, const double *U)
{
    CHECK_CTX(c);
    NEED(c, plx); NEED(c, U);
    launch_observables(c->L, (const double2 *)U, nullptr, plx, nullptr, c->red, c->part,
                       c->stream);
    return finish(c);
}

int qed2d_qtop_field(qed2d_ctx *c, double *plx, const double *U)
{
    CHECK_CTX(c);
    NEED(c, plx); NEED(c, U);
    launch_observables(c->L, (const double2 *)U, nullptr, nullptr, plx, c->red, c->part,
                       c->stream);
    return finish(c);
}

int qed2d_action(qed2d_ctx *c, const double *U, double *S)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, S);
    hdd sp;
    TRY(observables(c, U, nullptr, &sp, nullptr, nullptr));
    hdd g;
    g.add((double)c->L.V);
    g.add(-sp.hi, -sp.lo);
    *S = c->beta * g.value();
    return QED2D_OK;
}

int qed2d_plaquette(qed2d_ctx *c, const double *U, double *P)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, P);
    hdd sp;
    TRY(observables(c, U, nullptr, &sp, nullptr, nullptr));
    *P = sp.value() / (double)c->L.V;
    return QED2D_OK;
}

int qed2d_qtop(qed2d_ctx *c, const double *U, double *Q)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, Q);
    hdd sq;
    TRY(observables(c, U, nullptr, nullptr, &sq, nullptr));
    *Q = sq.value() / (2.0 * M_PI);
    return QED2D_OK;
}

int qed2d_hamiltonian(qed2d_ctx *c, const double *mom, const double *U, double *H)
{
    CHECK_CTX(c);
    NEED(c, mom); NEED(c, U); NEED(c, H);
    hdd h;
    TRY(hamiltonian_dd(c, mom, U, &h));
    *H = h.value();
    return QED2D_OK;
}

int qed2d_dot(qed2d_ctx *c, const double *a, const double *b, double *re, double *im)
{
    CHECK_CTX(c);
    NEED(c, a); NEED(c, b); NEED(c, re);
    return dot_host(c, a, b, re, im);
}

int qed2d_norm2(qed2d_ctx *c, const double *a, double *n2)
{
    CHECK_CTX(c);
    NEED(c, a); NEED(c, n2);
    return norm2_host(c, a, n2);
}

int qed2d_cg(qed2d_ctx *c, double *so, const double *U, const double *si, double am0, double mu,
             int op, int maxiter, double eps, int so_as_guess, int precision, int *iters)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    if (op != QED2D_OP_SQR_MSQ && op != QED2D_OP_SQR_MUSQ)
        return fail(c, QED2D_EINVAL, "unknown operator %d", op);
    if (precision != QED2D_PREC_F64 && precision != QED2D_PREC_MIXED)
        return fail(c, QED2D_EINVAL, "unknown precision %d", precision);
    if (maxiter < 1) return fail(c, QED2D_EINVAL, "maxiter must be >= 1");
    return cg_any(c, so, U, si, am0, mu, op, maxiter, eps, so_as_guess, precision, iters);
}

int qed2d_cg_mrhs(qed2d_ctx *c, double *x, const double *U, const double *b, int nrhs,
                  double am0, int maxiter, double eps, int precision, int *iters)
{
    CHECK_CTX(c);
    NEED(c, x); NEED(c, U); NEED(c, b);
    NOALIAS(c, x, b);
    if (nrhs < 1) return fail(c, QED2D_EINVAL, "nrhs must be >= 1");
    if (precision != QED2D_PREC_F64 && precision != QED2D_PREC_MIXED)
        return fail(c, QED2D_EINVAL, "unknown precision %d", precision);
    if (maxiter < 1) return fail(c, QED2D_EINVAL, "maxiter must be >= 1");
    const long V2 = spinor_elems(c) * 2;  // doubles per field
    if (precision == QED2D_PREC_MIXED && !pair_ok(c->L)) {
        // the set's FP32 kernels march site pairs: one solve at a time
        for (int j = 0; j < nrhs; ++j)
            TRY(cg_any(c, x + j * V2, U, b + j * V2, am0, 0.0, QED2D_OP_SQR_MSQ, maxiter, eps, 0,
                       precision, iters ? iters + j : nullptr));
        return finish(c);
    }
    qed2d_batch *set = nullptr;
    TRY(rhs_set(c, nrhs, &set));
    std::string err;
    const int rc = batch_rhs_cg(set, U, x, b, am0, maxiter, eps, precision, iters, c->rows_q,
                                c->rows, c->rows32, c->rows_k2, c->rows32_k2, c->ru_delta, &err);
    if (rc) return fail(c, rc, "%s", err.c_str());
    return finish(c);
}

int qed2d_multicg(qed2d_ctx *c, double *so, const double *U, const double *si, double am0,
                  int n, const double *mu, const double *rho, int maxiter, double eps)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    if (n < 0 || (n > 0 && (!mu || !rho))) return fail(c, QED2D_EINVAL, "bad pole arrays");
    TRY(multicg_impl(c, so, U, si, am0, n, mu, rho, maxiter, eps));
    return finish(c);
}

int qed2d_R(qed2d_ctx *c, double *so, const double *U, const double *si, double am0, int n,
            const double *mu, const double *rho, double A, double r_b, int maxiter, double eps)
{
    CHECK_CTX(c);
    NEED(c, so); NEED(c, U); NEED(c, si);
    NOALIAS(c, so, si);
    if (n < 0 || (n > 0 && (!mu || !rho))) return fail(c, QED2D_EINVAL, "bad pole arrays");
    TRY(r_impl(c, so, U, si, am0, n, mu, rho, A, r_b, maxiter, eps));
    return finish(c);
}

int qed2d_luscher_z(qed2d_ctx *c, double *Z, int p, const double *U, const double *X, double am0,
                    int n, const double *mu, const double *rho, double A, double r_b,
                    int maxiter, double tol)
{
    CHECK_CTX(c);
    NEED(c, Z); NEED(c, U); NEED(c, X);
    if (p < 0) return fail(c, QED2D_EINVAL, "power p must be >= 0");
    if (n < 0 || (n > 0 && (!mu || !rho))) return fail(c, QED2D_EINVAL, "bad pole arrays");
    double2 *in;
    TRY(get_spin(c, S_RW_IN, &in));
    // src/QED2dRHMC.jl:189-201: tmp = X; repeat p times Z = Z(tmp), tmp = Z
    launch_copy(in, (const double2 *)X, spinor_elems(c), c->stream);
    for (int i = 0; i < p; ++i) {
        TRY(luscher_z_impl(c, Z, U, (const double *)in, am0, n, mu, rho, A, r_b, maxiter, tol));
        launch_copy(in, (const double2 *)Z, spinor_elems(c), c->stream);
    }
    return finish(c);
}

int qed2d_reweighting_factor(qed2d_ctx *c, const double *U, double am0, int n, const double *mu,
                             const double *rho, double A, double r_b, int N, int taylor,
                             int maxiter, double tol, uint64_t seed, uint32_t chain,
                             uint32_t traj, double *W)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, W);
    if (N < 1 || taylor < 0) return fail(c, QED2D_EINVAL, "need N >= 1 and taylor >= 0");
    if ((long)N > (1L << 30)) return fail(c, QED2D_EINVAL, "N too large");
    if (n < 0 || (n > 0 && (!mu || !rho))) return fail(c, QED2D_EINVAL, "bad pole arrays");
    double2 *X, *Z, *in;
    TRY(get_spin(c, S_RW_X, &X));
    TRY(get_spin(c, S_RW_Z, &Z));
    TRY(get_spin(c, S_RW_IN, &in));
    const long n2 = spinor_elems(c);
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    double w = 0.0;
    // src/QED2dRHMC.jl:215-228
    for (int j = 0; j < N; ++j) {
        double argexp = 0.0, tf = 0.5;
        launch_fill_cgauss(c->L, X, k0, k1, traj, kPurposeRew, 2u * (uint32_t)j, c->stream);
        launch_copy(in, X, n2, c->stream);
        for (int i = 1; i <= taylor; ++i) {
            TRY(luscher_z_impl(c, (double *)Z, U, (const double *)in, am0, n, mu, rho, A, r_b,
                               maxiter, tol));
            double re, im;
            TRY(dot_host(c, (const double *)X, (const double *)Z, &re, &im));
            argexp += tf * re;
            // Tfactor *= (-1)*(2i+1)/2/factorial(i+1)*factorial(i), left to right
            double fi = 1.0;
            for (int q = 2; q <= i; ++q) fi *= q;
            tf *= (double)(-(2 * i + 1)) / 2.0 / (fi * (i + 1)) * fi;
            launch_copy(in, Z, n2, c->stream);
        }
        w += std::exp(argexp) / N;
    }
    *W = w;
    return finish(c);
}

int qed2d_generate_pseudofermion(qed2d_ctx *c, double *F, const double *U, const double *X,
                                 double am0, int n, const double *mu, const double *nu,
                                 int maxiter, double tol)
{
    CHECK_CTX(c);
    NEED(c, F); NEED(c, U); NEED(c, X);
    NOALIAS(c, F, X);
    if (n < 0 || (n > 0 && (!mu || !nu))) return fail(c, QED2D_EINVAL, "bad pole arrays");
    TRY(gen_pf_impl(c, F, U, X, am0, n, mu, nu, maxiter, tol));
    return finish(c);
}

int qed2d_power_method(qed2d_ctx *c, const double *U, const double *b0, double am0, int iter,
                       double *lambda_min, double *lambda_max)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, b0); NEED(c, lambda_min); NEED(c, lambda_max);
    double2 *b, *ba, *bl, *tmp;
    TRY(get_spin(c, S_PB, &b));
    TRY(get_spin(c, S_PBA, &ba));
    TRY(get_spin(c, S_PBL, &bl));
    TRY(get_spin(c, S_TMP, &tmp));
    const long n2 = spinor_elems(c);
    hipStream_t st = c->stream;
    const double shift = 2.0;
    launch_copy(b, (const double2 *)b0, n2, st);
    double im;
    // src/QED2dRHMC.jl:144-149; the normalisation reads dot(ba, ba) on the
    // device (the iterations queue without a host round trip)
    for (int i = 0; i < iter; ++i) {
        qsq64(c, U, (const double *)b, (double *)ba, (double *)tmp, am0, 0.0);
        launch_scale_add(ba, 1.0, b, shift, n2, st);
        launch_dot(ba, ba, n2, c->red, c->part, st);
        launch_scale_add_inv(b, 0.0, ba, 0.0, c->red, 1, n2, st);  // b = ba / dot(ba, ba)
        if ((i & 255) == 255) TRY(launched(c));
    }
    qsq64(c, U, (const double *)b, (double *)ba, (double *)tmp, am0, 0.0);
    launch_scale_add(ba, 1.0, b, shift, n2, st);
    double num, den;
    TRY(dot_host(c, (const double *)b, (const double *)ba, &num, &im));
    TRY(dot_host(c, (const double *)b, (const double *)b, &den, &im));
    const double lmax = num / den - shift;
    // src/QED2dRHMC.jl:157-166 (Q^2 applied to a copy; the reference aliases)
    for (int i = 0; i < iter; ++i) {
        launch_copy(bl, b, n2, st);
        qsq64(c, U, (const double *)bl, (double *)b, (double *)tmp, am0, 0.0);
        launch_scale_add(b, 1.0, bl, -lmax, n2, st);
        launch_dot(b, b, n2, c->red, c->part, st);
        launch_scale_add_inv(b, 0.0, b, 0.0, c->red, 0, n2, st);  // b = b / dot(b, b)
        if ((i & 255) == 255) TRY(launched(c));
    }
    qsq64(c, U, (const double *)b, (double *)ba, (double *)tmp, am0, 0.0);
    launch_scale_add(ba, 1.0, b, -lmax, n2, st);
    TRY(dot_host(c, (const double *)b, (const double *)ba, &num, &im));
    TRY(dot_host(c, (const double *)b, (const double *)b, &den, &im));
    *lambda_min = num / den + lmax;
    *lambda_max = lmax;
    return QED2D_OK;
}

int qed2d_fill_momenta(qed2d_ctx *c, double *mom, uint64_t seed, uint32_t chain, uint32_t traj)
{
    CHECK_CTX(c);
    NEED(c, mom);
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    launch_fill_momenta(c->L, mom, k0, k1, traj, c->stream);
    return finish(c);
}

int qed2d_fill_pseudofermion(qed2d_ctx *c, double *X, uint64_t seed, uint32_t chain,
                             uint32_t traj, uint32_t flavour)
{
    CHECK_CTX(c);
    NEED(c, X);
    if (flavour >= 16) return fail(c, QED2D_EINVAL, "flavour index must be < 16");
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    launch_fill_pseudofermion(c->L, (double2 *)X, k0, k1, traj, flavour, c->stream);
    return finish(c);
}

int qed2d_hot_start(qed2d_ctx *c, double *U, uint64_t seed, uint32_t chain)
{
    CHECK_CTX(c);
    NEED(c, U);
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    launch_hot_start(c->L, (double2 *)U, k0, k1, c->stream);
    return finish(c);
}

double qed2d_metropolis_uniform(uint64_t seed, uint32_t chain, uint32_t traj)
{
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    uint32_t ctr[4] = {0u, 0u, kPurposeMetro, traj};
    philox4x32_10(ctr, k0, k1);
    return u53(ctr[0], ctr[1]);
}

int qed2d_omf4_pure(qed2d_ctx *c, double *mom, double *U, double eps, int ns)
{
    CHECK_CTX(c);
    NEED(c, mom); NEED(c, U);
    TRY(omf4_pure(c, mom, U, eps, ns));
    return finish(c);
}

int qed2d_md_nf2(qed2d_ctx *c, int integrator, double *mom, double *U, double *X, const double *F,
                 double *g5DX, double am0, double eps, int ns, int maxiter, double tol,
                 int precision, int *cg_iters)
{
    CHECK_CTX(c);
    NEED(c, mom); NEED(c, U); NEED(c, X); NEED(c, F); NEED(c, g5DX);
    NOALIAS(c, X, F);
    int it = 0;
    TRY(md_nf2(c, integrator, mom, U, X, F, g5DX, am0, eps, ns, maxiter, tol, precision, &it));
    if (cg_iters) *cg_iters = it;
    return finish(c);
}

// A trajectory that fails after U was saved to Ucp restores U and reports no
// acceptance; the first error's message is kept.
static int restore_fail(qed2d_ctx *c, double *U, const double2 *Ucp, int rc, int *accepted)
{
    std::string msg = c->err;
    launch_copy((double2 *)U, Ucp, spinor_elems(c), c->stream);
    (void)hipStreamSynchronize(c->stream);
    c->err = msg;
    *accepted = 0;
    return rc;
}

#define RTRY(call)                                                            \
    do {                                                                      \
        int rc_ = (call);                                                     \
        if (rc_ != QED2D_OK) return restore_fail(c, U, Ucp, rc_, accepted);   \
    } while (0)

int qed2d_hmc_pure(qed2d_ctx *c, double *U, double eps, int ns, int integrator, int qzero,
                   uint64_t seed, uint32_t chain, uint32_t traj, double *hini, double *hfin,
                   int *accepted)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, hini); NEED(c, hfin); NEED(c, accepted);
    double2 *Ucp;
    double *mom;
    TRY(get_spin(c, S_UCP, &Ucp));
    TRY(get_real(c, R_MOM, &mom));
    const long n2 = spinor_elems(c);
    launch_copy(Ucp, (const double2 *)U, n2, c->stream);
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    launch_fill_momenta(c->L, mom, k0, k1, traj, c->stream);
    hdd h0, h1;
    RTRY(hamiltonian_dd(c, mom, U, &h0));
    if (integrator == QED2D_OMF4) RTRY(omf4_pure(c, mom, U, eps, ns));
    else RTRY(leapfrog_pure(c, mom, U, eps, ns));
    RTRY(hamiltonian_dd(c, mom, U, &h1));
    *hini = h0.value();
    *hfin = h1.value();
    c->last_dh = dd_diff(h1, h0);
    RTRY(metropolis(c, *hini, *hfin, qzero, U, seed, chain, traj, accepted));
    if (!*accepted) launch_copy((double2 *)U, Ucp, n2, c->stream);
    return finish(c);
}

int qed2d_hmc_nf2(qed2d_ctx *c, double *U, double am0, double eps, int ns, int integrator,
                  int maxiter, double tol, int qzero, int precision, double acc_tol,
                  uint64_t seed, uint32_t chain, uint32_t traj, double *hini, double *hfin,
                  int *accepted, int *cg_iters)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, hini); NEED(c, hfin); NEED(c, accepted);
    double2 *Ucp, *X, *F, *g5DX;
    double *mom;
    TRY(get_spin(c, S_UCP, &Ucp));
    TRY(get_spin(c, S_X, &X));
    TRY(get_spin(c, S_F, &F));
    TRY(get_spin(c, S_G5DX, &g5DX));
    TRY(get_real(c, R_MOM, &mom));
    const long n2 = spinor_elems(c);
    hipStream_t st = c->stream;
    launch_copy(Ucp, (const double2 *)U, n2, st);
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    launch_fill_momenta(c->L, mom, k0, k1, traj, st);
    launch_fill_pseudofermion(c->L, X, k0, k1, traj, 0, st);
    // hini = H + X^dag X  (src/QED2dHMC.jl:74)
    hdd h0, h1;
    RTRY(hamiltonian_dd(c, mom, U, &h0));
    hdd xx;
    RTRY(dot_dd_host(c, (const double *)X, nullptr, &xx));
    h0.add(xx.hi, xx.lo);
    q64(c, U, (const double *)X, (double *)F, am0);  // F = Q X  (:76-78)
    int it = 0;
    int rc = md_nf2(c, integrator, mom, U, (double *)X, (const double *)F, (double *)g5DX, am0,
                    eps, ns, maxiter, tol, precision, &it);
    if (cg_iters) *cg_iters = it;
    RTRY(rc);
    RTRY(hamiltonian_dd(c, mom, U, &h1));
    if (acc_tol <= 0.0) {
        // hfin = H + Re X^dag F with X the last MD solve  (:82)
        hdd re;
        RTRY(dot_dd_host(c, (const double *)X, (const double *)F, &re));
        h1.add(re.hi, re.lo);
    } else {
        int ait = 0;
        RTRY(cg_any(c, (double *)X, U, (const double *)F, am0, 0.0, QED2D_OP_SQR_MSQ, maxiter,
                   acc_tol, 1, QED2D_PREC_F64, &ait));
        q64(c, U, (const double *)X, (double *)g5DX, am0);
        hdd re, yy;
        RTRY(dot_dd_host(c, (const double *)X, (const double *)F, &re));
        RTRY(dot_dd_host(c, (const double *)g5DX, nullptr, &yy));
        h1.add(2.0 * re.hi, 2.0 * re.lo);
        h1.add(-yy.hi, -yy.lo);
    }
    *hini = h0.value();
    *hfin = h1.value();
    c->last_dh = dd_diff(h1, h0);
    RTRY(metropolis(c, *hini, *hfin, qzero, U, seed, chain, traj, accepted));
    if (!*accepted) launch_copy((double2 *)U, Ucp, n2, st);
    return finish(c);
}

int qed2d_hmc_rhmc(qed2d_ctx *c, double *U, const double *am0, int nf, const int *npole,
                   const double *const *mu, const double *const *nu, const double *const *rho,
                   double eps, int ns, int integrator, int maxiter, double tol, int qzero,
                   int reversibility_test, uint64_t seed, uint32_t chain, uint32_t traj,
                   double *hini, double *hfin, int *accepted)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, hini); NEED(c, hfin); NEED(c, accepted);
    if (nf < 0 || nf > 16 || (nf > 0 && (!am0 || !npole || !mu || !nu || !rho)))
        return fail(c, QED2D_EINVAL, "bad flavour arrays");
    double2 *Ucp, *X, *g5DX;
    double *mom;
    TRY(get_spin(c, S_UCP, &Ucp));
    TRY(get_spin(c, S_X, &X));
    TRY(get_spin(c, S_G5DX, &g5DX));
    TRY(get_real(c, R_MOM, &mom));
    const long n2 = spinor_elems(c);
    while ((int)c->flav.size() < nf) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(double2) * n2) != hipSuccess)
            return fail(c, QED2D_ENOMEM, "hipMalloc of a pseudofermion failed");
        c->flav.push_back((double2 *)p);
    }
    hipStream_t st = c->stream;
    launch_copy(Ucp, (const double2 *)U, n2, st);
    uint32_t k0, k1;
    key_of(seed, chain, &k0, &k1);
    launch_fill_momenta(c->L, mom, k0, k1, traj, st);
    hdd h0, h1;
    RTRY(hamiltonian_dd(c, mom, U, &h0));
    std::vector<double2 *> F(c->flav.begin(), c->flav.begin() + nf);
    // multi-shift mode 1: every flavour's heatbath and acceptance solves together
    bool together = (c->multishift == 1 && nf > 1);
    for (int j = 0; j < nf; ++j) together = together && npole[j] >= 1 && npole[j] <= kMaxShift;
    for (int j = 0; j < nf; ++j) {
        // src/QED2dHMC.jl:136-144
        launch_fill_pseudofermion(c->L, X, k0, k1, traj, (uint32_t)j, st);
        hdd xx;
        RTRY(dot_dd_host(c, (const double *)X, nullptr, &xx));
        h0.add(xx.hi, xx.lo);
        if (together)
            launch_copy(F[j], X, n2, st);
        else
            RTRY(gen_pf_impl(c, (double *)F[j], U, (const double *)X, am0[j], npole[j], mu[j],
                            nu[j], maxiter, tol));
    }
    if (together) RTRY(gen_pf_multi(c, F, U, am0, nf, npole, mu, nu, maxiter, tol));
    RTRY(md_rhmc(c, integrator, mom, U, (double *)X, F, (double *)g5DX, am0, nf, npole, mu,
                 rho, eps, ns, maxiter, tol));
    RTRY(hamiltonian_dd(c, mom, U, &h1));
    if (together) {
        std::vector<const double *> bj(nf);
        for (int j = 0; j < nf; ++j) bj[j] = (const double *)F[j];
        int it = 0;
        RTRY(multishift_multi(c, U, nf, bj.data(), am0, npole, mu, maxiter, tol, &it));
    }
    for (int j = 0; j < nf; ++j) {
        // src/QED2dHMC.jl:150-156
        if (together) {
            // multicg_impl's sum over the poles, from the joint solve
            launch_copy(X, F[j], n2, st);
            for (int i = 0; i < npole[j]; ++i)
                launch_scale_add(X, 1.0, c->mr.x[j][i], rho[j][i], n2, st);
        } else {
            RTRY(multicg_impl(c, (double *)X, U, (const double *)F[j], am0[j], npole[j], mu[j],
                             rho[j], maxiter, tol));
        }
        hdd re;
        RTRY(dot_dd_host(c, (const double *)X, (const double *)F[j], &re));
        h1.add(re.hi, re.lo);
    }
    *hini = h0.value();
    *hfin = h1.value();
    c->last_dh = dd_diff(h1, h0);
    if (reversibility_test) {
        // src/QED2dHMC.jl:159-164
        launch_negate(mom, 2 * c->L.V, st);
        RTRY(md_rhmc(c, integrator, mom, U, (double *)X, F, (double *)g5DX, am0, nf, npole, mu,
                    rho, eps, ns, maxiter, tol));
        *accepted = -1;
        return finish(c);
    }
    RTRY(metropolis(c, *hini, *hfin, qzero, U, seed, chain, traj, accepted));
    if (!*accepted) launch_copy((double2 *)U, Ucp, n2, st);
    return finish(c);
}

int qed2d_save_gauge(qed2d_ctx *c, const double *U, const char *path)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, path);
    const long n = 2 * c->L.V;
    std::vector<double> h(2 * n);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(h.data(), U, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
    if (qed2d_write_links_text(h.data(), n, path) != QED2D_OK)
        return fail(c, QED2D_EINVAL, "cannot write %s", path);
    return QED2D_OK;
}

int qed2d_load_gauge(qed2d_ctx *c, double *U, const char *path)
{
    CHECK_CTX(c);
    NEED(c, U); NEED(c, path);
    const long n = 2 * c->L.V;
    std::vector<double> h(2 * n);
    if (qed2d_read_links_text(h.data(), n, path) != QED2D_OK)
        return fail(c, QED2D_EINVAL, "cannot read %ld links from %s", n, path);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(U, h.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
    return QED2D_OK;
}

int qed2d_ctx_set_multishift(qed2d_ctx *c, int mode)
{
    CHECK_CTX(c);
    if (mode < 0 || mode > 2) return fail(c, QED2D_EINVAL, "multi-shift mode must be 0, 1 or 2");
    c->multishift = mode;
    return QED2D_OK;
}

int qed2d_multishift_cg(qed2d_ctx *c, double *const *x, const double *U, const double *b,
                        double am0, int n, const double *mu, int maxiter, double eps, int *iters)
{
    CHECK_CTX(c);
    NEED(c, x); NEED(c, U); NEED(c, b); NEED(c, mu);
    if (n < 1 || n > kMaxShift)
        return fail(c, QED2D_EINVAL, "multi-shift CG takes 1..%d shifts (got %d)", kMaxShift, n);
    std::vector<double2 *> xs(n);
    for (int k = 0; k < n; ++k) {
        NEED(c, x[k]);
        NOALIAS(c, (const double *)x[k], b);
        xs[k] = (double2 *)x[k];
    }
    TRY(multishift_cg(c, xs.data(), U, b, am0, n, mu, maxiter, eps, iters));
    return finish(c);
}

int qed2d_profile_enable(qed2d_ctx *c, int every)
{
    CHECK_CTX(c);
    if (every < 0) return fail(c, QED2D_EINVAL, "sampling period must be >= 0");
    c->prof = every;
    for (auto &n : c->prof_seen) n = 0;
    return QED2D_OK;
}

int qed2d_profile_reset(qed2d_ctx *c)
{
    CHECK_CTX(c);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_harvest(c);
    for (int k = 0; k < QED2D_PROF_NCLASS; ++k) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    return QED2D_OK;
}

int qed2d_profile_read(qed2d_ctx *c, int cls, long *count, double *total_ms)
{
    CHECK_CTX(c);
    if (cls < 0 || cls >= QED2D_PROF_NCLASS || !count || !total_ms) return QED2D_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    prof_harvest(c);
    *count = c->prof_n[cls];
    *total_ms = c->prof_ms[cls];
    return QED2D_OK;
}

int qed2d_unfold(qed2d_ctx *c, double *Uout, const double *Uin, int reflect)
{
    CHECK_CTX(c);
    NEED(c, Uout); NEED(c, Uin);
    NOALIAS(c, Uout, Uin);
    launch_unfold(c->L, (const double2 *)Uin, (double2 *)Uout, reflect ? 1 : 0, c->stream);
    return finish(c);
}

}  // extern "C"
