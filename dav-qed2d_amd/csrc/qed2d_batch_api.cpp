// qed2d_batch_api.cpp -- batched ensembles of independent lattices: the C ABI
// (qed2d_batch_*) and its orchestration.  Each chain's trajectory is the
// reference's HMC! for Nf=2 (src/QED2dHMC.jl:64-113) on that chain, drawn from
// the same (seed, chain, trajectory) streams as the single-lattice entry
// point qed2d_hmc_nf2, so a chain of a batch and a single-lattice run with the
// same tiles produce the same bits.  What is batched: every CG iteration, Q
// apply, force + momentum update and link update is one launch over all
// chains; the per-chain reductions of the Hamiltonian and the Metropolis
// decisions are per chain.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/qed2d.h"
#include "qed2d_march.h"

using namespace qed2d;

namespace {

struct GeoBuf {
    int *blk0 = nullptr;
    int4 *shape = nullptr;
    int nblocks = 0;
};

}  // namespace

struct qed2d_batch {
    hipStream_t stream = nullptr;
    int n = 0;
    long vtot = 0;
    int bc = 0;
    bool shared = false;  // right-hand sides of one lattice sharing the caller's links
    long nU = 0;          // elements of the link field(s) U points to
    std::vector<BChain> hch;
    BChain *dch = nullptr;
    std::vector<qed2d_ctx *> sub;  // one context per chain: its reductions
    std::string err;
    int rows = 16, rows32 = 8, rows_q = 4;
    int rows_k2 = 0, rows32_k2 = 0;  // the second CG kernel's heights; 0: rows / rows32
    std::vector<double> mu2;  // per chain: the CG operator's shift (Q^2 + mu2); empty: 0
    bool guess = false;       // the next solve starts from the x it is given
    double ru_delta = kRuDelta;
    // mixed precision: a K1 that sees the residual drop by ru_delta requests a
    // reliable update and the chain iterates on until the next slot (soft), or
    // pauses until the slot (the single-lattice host-decided schedule, bit for
    // bit; qed2d_batch_set_reliable_mode(b, 0))
    bool ru_soft = true;
    std::vector<double> last_dh;  // per chain, from the double-double energies
    std::map<int, GeoBuf> geo;  // key: 2 * rows + pair
    // CG chunks as HIP graphs (qed2d_batch_set_graphs): one executable graph per
    // chunk shape, captured on a private stream and launched on `stream`
    bool graphs = false;
    hipStream_t cap = nullptr;
    using GraphKey = std::tuple<int, int, int, int, int, const void *, const void *, int, int,
                                int, int, int>;
    std::map<GraphKey, hipGraphExec_t> gexec;
    // fields (arena)
    double2 *U = nullptr, *Ucp = nullptr, *X = nullptr, *F = nullptr, *Y = nullptr;
    double2 *r = nullptr, *p0 = nullptr, *p1 = nullptr, *t = nullptr;
    float2 *Us = nullptr, *rs = nullptr, *p0s = nullptr, *p1s = nullptr, *ts = nullptr,
           *xs = nullptr;
    double2 *xn = nullptr;  // a reliable update's new FP64 solution (mixed precision)
    double *mom = nullptr;
    // solver state and partials
    SolverState *S = nullptr, *hS = nullptr;
    RedOut *hred = nullptr;  // pinned: the energy reductions of a trajectory (5 x n)
    // per-chain reductions in one launch: producer blocks of chain c are
    // rblk[k][c] .. rblk[k][c+1]-1 (k = 0: observables over V, 1: dot products
    // over 2V); partials, the finished sums (5 x n) and the chains' random keys
    int *rblk[2] = {nullptr, nullptr};
    int nrblk[2] = {0, 0};
    double *rpart = nullptr;
    RedOut *dred = nullptr;
    uint2 *dkeys = nullptr, *hkeys = nullptr;
    // pipelined solves: two pinned slots of chain states, each with an event
    SolverState *hSring = nullptr;
    hipEvent_t sev[2] = {};
    double *part = nullptr;
    int *cnt = nullptr;
    BParts P{};
};

namespace {

int bfail(qed2d_batch *b, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

#define BTRY(x)                     \
    do {                            \
        int rc_ = (x);              \
        if (rc_ != QED2D_OK) return rc_; \
    } while (0)
#define BHIP(b, x)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return bfail((b), QED2D_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

int blaunched(qed2d_batch *b)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, QED2D_EHIP, "kernel launch: %s", hipGetErrorString(e));
    return QED2D_OK;
}

// A sub-context's error, carried to the batch with the chain it came from.
int subfail(qed2d_batch *b, int c, int rc)
{
    return bfail(b, rc, "chain %d: %s", c, qed2d_last_error(b->sub[c]));
}

Latt arena(const qed2d_batch *b)
{
    Latt L;
    L.L1 = 1;
    L.L2 = 1;
    L.V = b->vtot;
    L.bc = 0;
    return L;
}

template <typename T> int balloc(qed2d_batch *b, T **p, long n)
{
    if (*p) return QED2D_OK;
    void *q = nullptr;
    if (hipMalloc(&q, sizeof(T) * n) != hipSuccess)
        return bfail(b, QED2D_ENOMEM, "hipMalloc of %ld bytes failed", (long)(sizeof(T) * n));
    *p = (T *)q;
    return QED2D_OK;
}

// Launch geometry of one kernel family (rows per thread, pair or one-site
// march) over the chains, built once and cached.
int get_geo(qed2d_batch *b, int rows, bool pair, BGeo *out)
{
    const int key = 2 * rows + (pair ? 1 : 0);
    auto it = b->geo.find(key);
    if (it == b->geo.end()) {
        std::vector<int> blk0(b->n + 1);
        std::vector<int4> shape(b->n);
        long tot = 0;
        for (int c = 0; c < b->n; ++c) {
            const MarchGrid g = pair ? march_grid2(b->hch[c].L, rows) : march_grid(b->hch[c].L, rows);
            blk0[c] = (int)tot;
            shape[c] = make_int4(g.bx, g.by, g.nbx, g.nblocks());
            tot += g.nblocks();
        }
        if (tot > 0x7fffffffL) return bfail(b, QED2D_EINVAL, "batch too large for one launch");
        blk0[b->n] = (int)tot;
        GeoBuf gb;
        BHIP(b, hipMalloc((void **)&gb.blk0, sizeof(int) * (b->n + 1)));
        BHIP(b, hipMalloc((void **)&gb.shape, sizeof(int4) * b->n));
        BHIP(b, hipMemcpy(gb.blk0, blk0.data(), sizeof(int) * (b->n + 1), hipMemcpyHostToDevice));
        BHIP(b, hipMemcpy(gb.shape, shape.data(), sizeof(int4) * b->n, hipMemcpyHostToDevice));
        gb.nblocks = (int)tot;
        it = b->geo.emplace(key, gb).first;
    }
    out->ch = b->dch;
    out->blk0 = it->second.blk0;
    out->shape = it->second.shape;
    out->n = b->n;
    out->nblocks = it->second.nblocks;
    out->inter = b->shared ? 1 : 0;
    return QED2D_OK;
}

int set_masses(qed2d_batch *b, const double *am0)
{
    for (int c = 0; c < b->n; ++c) b->hch[c].am0 = am0[c];
    BHIP(b, hipMemcpyAsync(b->dch, b->hch.data(), sizeof(BChain) * b->n, hipMemcpyHostToDevice,
                           b->stream));
    return QED2D_OK;
}

// Philox key of a chain's random streams: (seed, chain).
void key_of(uint64_t seed, uint32_t chain, uint32_t *k0, uint32_t *k1)
{
    *k0 = (uint32_t)seed;
    *k1 = (uint32_t)(seed >> 32) ^ chain;
}

// Blocks of one launch over all chains at `rows` rows per thread (pair: the
// FP32 site-pair kernels, chains with even L1 >= 4).
long batch_blocks(const qed2d_batch *b, int rows, bool pair)
{
    long n = 0;
    for (const BChain &ch : b->hch)
        n += (pair && pair_ok(ch.L)) ? march_grid2(ch.L, rows).nblocks()
                                     : march_grid(ch.L, rows).nblocks();
    return n;
}

// Default tile heights: the tallest (at most 16 rows FP64, 8 FP32) that leaves
// a solver launch of at least 1024 blocks over all chains (an ensemble of
// small lattices has few blocks per chain: the 64-chain 16^2 ... 256^2
// ensemble runs 597 chain-trajectories/s at 16 / 8 rows, 693 at 2 / 2).
void default_batch_tiles(qed2d_batch *b)
{
    int r = 16;
    while (r > 1 && batch_blocks(b, r, false) < 1024) r >>= 1;
    b->rows = r;
    int r32 = 8;
    while (r32 > 1 && batch_blocks(b, r32, true) < 1024) r32 >>= 1;
    b->rows32 = r32;
    b->rows_q = (batch_blocks(b, 4, false) < 1024) ? 2 : 4;
    b->rows_k2 = b->rows32_k2 = 0;
}

int read_states(qed2d_batch *b)
{
    BHIP(b, hipMemcpyAsync(b->hS, b->S, sizeof(SolverState) * b->n, hipMemcpyDeviceToHost,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return QED2D_OK;
}

struct Predict {  // the single-lattice solvers' chunk predictor, per chain
    double err0 = 0.0, rate = 0.0;
    int it0 = -1;
    int next(double err, int iter, double tol, double ru_target)
    {
        int chunk = 8;
        if (it0 >= 0 && iter > it0 && err > 0.0 && err0 > err) {
            rate = std::log(err / err0) / (double)(iter - it0);
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
    // iterations left to |r|^2 < tol at the observed rate (a large value while
    // the rate is unknown)
    double remaining(double err, double tol) const
    {
        if (!(rate < 0.0) || !(err > 0.0)) return 1e9;
        const double left = std::log(tol / err) / rate;
        return left > 0.0 ? left : 0.0;
    }
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

int ring_setup(qed2d_batch *b)
{
    if (b->hSring) return QED2D_OK;
    if (hipHostMalloc((void **)&b->hSring, 2 * sizeof(SolverState) * b->n) != hipSuccess ||
        hipEventCreateWithFlags(&b->sev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->sev[1], hipEventDisableTiming) != hipSuccess)
        return bfail(b, QED2D_ENOMEM, "pinned state slots of the batched solver");
    return QED2D_OK;
}

int red_setup(qed2d_batch *b)
{
    if (b->dred) return QED2D_OK;
    std::vector<int> t[2];
    for (int k = 0; k < 2; ++k) {
        t[k].assign(b->n + 1, 0);
        for (int c = 0; c < b->n; ++c)
            t[k][c + 1] = t[k][c] + chain_red_blocks((k + 1) * b->hch[c].L.V);
        b->nrblk[k] = t[k][b->n];
    }
    const long np = std::max(6L * b->nrblk[0], 4L * b->nrblk[1]);
    if (hipMalloc((void **)&b->rblk[0], sizeof(int) * (b->n + 1)) != hipSuccess ||
        hipMalloc((void **)&b->rblk[1], sizeof(int) * (b->n + 1)) != hipSuccess ||
        hipMalloc((void **)&b->rpart, sizeof(double) * np) != hipSuccess ||
        hipMalloc((void **)&b->dkeys, sizeof(uint2) * b->n) != hipSuccess ||
        hipHostMalloc((void **)&b->hkeys, sizeof(uint2) * b->n) != hipSuccess ||
        hipHostMalloc((void **)&b->hred, 5 * sizeof(RedOut) * b->n) != hipSuccess ||
        hipMalloc((void **)&b->dred, sizeof(RedOut) * 5 * b->n) != hipSuccess) {
        // (release what was allocated: a later call starts over)
        for (void *p : {(void *)b->rblk[0], (void *)b->rblk[1], (void *)b->rpart,
                        (void *)b->dkeys, (void *)b->dred})
            if (p) (void)hipFree(p);
        if (b->hkeys) (void)hipHostFree(b->hkeys);
        if (b->hred) (void)hipHostFree(b->hred);
        b->rblk[0] = b->rblk[1] = nullptr;
        b->rpart = nullptr;
        b->dkeys = b->hkeys = nullptr;
        b->dred = b->hred = nullptr;
        return bfail(b, QED2D_ENOMEM, "reduction buffers of the batch");
    }
    BHIP(b, hipMemcpy(b->rblk[0], t[0].data(), sizeof(int) * (b->n + 1), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->rblk[1], t[1].data(), sizeof(int) * (b->n + 1), hipMemcpyHostToDevice));
    return QED2D_OK;
}

// (Q^2) x_c = src_c on every chain from x = 0, to |r|^2 < tol |b|^2 (the
// single-lattice cg_f64 / cg_mixed, one launch per kernel over all chains).
int batch_cg(qed2d_batch *b, double2 *x, const double2 *src, int maxiter, double tol,
             int precision, std::vector<int> &iters, int chunk0 = 8)
{
    const long nel = 2 * b->vtot;
    hipStream_t st = b->stream;
    BTRY(balloc(b, &b->r, nel));
    BTRY(balloc(b, &b->p0, nel));
    BTRY(balloc(b, &b->p1, nel));
    BTRY(balloc(b, &b->t, nel));
    BGeo g64, g64b, gq;
    const int rows64b = b->rows_k2 ? b->rows_k2 : b->rows;
    const int rows32b = b->rows32_k2 ? b->rows32_k2 : b->rows32;
    BTRY(get_geo(b, b->rows, false, &g64));
    BTRY(get_geo(b, rows64b, false, &g64b));
    BTRY(get_geo(b, b->rows_q, false, &gq));
    const bool mixed = (precision == QED2D_PREC_MIXED);
    bool shifted = false;  // (the FP32 K2 reads p_k only for chains with a shift)
    for (double m2 : b->mu2) shifted = shifted || m2 != 0.0;
    BGeo g32, g32b;
    if (mixed) {
        for (int c = 0; c < b->n; ++c)
            if (!pair_ok(b->hch[c].L))
                return bfail(b, QED2D_EINVAL, "mixed-precision batches need even L1 >= 4 "
                             "(chain %d has L1 = %d)", c, b->hch[c].L.L1);
        BTRY(balloc(b, &b->Us, b->nU));
        BTRY(balloc(b, &b->rs, nel));
        BTRY(balloc(b, &b->p0s, nel));
        BTRY(balloc(b, &b->p1s, nel));
        BTRY(balloc(b, &b->ts, nel));
        BTRY(balloc(b, &b->xs, nel));
        BTRY(balloc(b, &b->xn, nel));
        BTRY(get_geo(b, b->rows32, true, &g32));
        BTRY(get_geo(b, rows32b, true, &g32b));
    }
    // FP32 links packed (one float each) when every link is U(1), as the
    // single-lattice solver; the first look at the states tells, and a solve
    // whose links are not restarts with FP32 complex links.
    bool packed = mixed;
    std::vector<Predict> pred;
    std::vector<int> ev(b->n);        // per chain: iterations to its next event
    std::vector<double> rem(b->n);    // per chain: iterations its solve has left
    int chunk = chunk0;
    long launched_k = 0;
    auto start = [&]() -> int {
        for (int c = 0; c < b->n; ++c) {
            SolverState &s = b->hS[c];
            std::memset(&s, 0, sizeof s);
            s.eps = tol;
            s.maxiter = maxiter;
            s.delta2 = b->ru_delta * b->ru_delta;
            s.mu2 = b->mu2.empty() ? 0.0 : b->mu2[c];
            s.kk = 1;
            s.k2 = 1;
        }
        BHIP(b, hipMemcpyAsync(b->S, b->hS, sizeof(SolverState) * b->n, hipMemcpyHostToDevice,
                               st));
        if (mixed) {
            if (packed)
                launch_link_pack(b->U, (float *)b->Us, b->nU, b->S, st);
            else
                launch_to_f32(b->U, b->Us, b->nU, st);
        }
        if (b->guess) {
            // from the guess in x: the true residual in FP64 (mixed: into rs,
            // the FP32 partial solution xs starting at 0)
            launch_b_q_apply(gq, b->rows_q, b->U, x, b->t, nullptr, st);
            launch_b_cg_resid(g64, b->rows, b->U, src, b->t, x, mixed ? nullptr : b->r,
                              mixed ? b->rs : nullptr, mixed ? nullptr : b->p0,
                              mixed ? b->p0s : nullptr, b->S, b->P, 0, st);
            if (mixed) BHIP(b, hipMemsetAsync(b->xs, 0, sizeof(float2) * nel, st));
        } else if (mixed) {
            launch_b_cg_init(g64, b->rows, src, x, b->rs, b->p0s, b->xs, 1, b->P, st);
        } else {
            launch_b_cg_init(g64, b->rows, src, x, b->r, b->p0, nullptr, 0, b->P, st);
        }
        BTRY(blaunched(b));
        pred.assign(b->n, Predict());
        chunk = chunk0;
        launched_k = 0;
        return QED2D_OK;
    };
    // Pipelined: behind the chunk whose states the host waits for, the next one
    // is already queued (not behind the first, whose look may restart the solve
    // with other links).  Each chunk's states land in a pinned slot of two, with
    // an event.  In mixed precision a chunk may start with a reliable-update
    // slot: its kernels act on the chains whose K1 paused for an update
    // (need_ru) and return at once on the others, so a chain that paused in the
    // chunk before resumes without waiting for the host to see it, and takes
    // the same steps as a single-lattice solve whose host performs the update
    // at the pause.  A chunk gets a slot when the states the host last read
    // show a chain paused, or predict a chain to pause in the chunk in flight
    // (a slot costs about one iteration over all chains even when idle); a
    // chain that pauses unannounced is seen by the next look and served by
    // the chunk queued then.
    BTRY(ring_setup(b));
    int slot = 0, nq = 0, qhead = 0;
    int qslot[2], qlen[2];
    // One chunk's kernels on stream s.  (The partial-sum counts that the
    // single-lattice launch functions track on the host live on the device
    // here, so a chunk's launches do not depend on the iteration: the same
    // sequence serves every chunk of that length and can be replayed as a graph.)
    auto issue = [&](hipStream_t s, int len, bool ru_slot) {
        if (ru_slot) {
            // the paused chains: fold the FP32 partial solution into x,
            // recompute the true residual in FP64, resume at the same K1
            launch_b_ru_apply(gq, b->rows_q, b->U, x, b->xs, b->p0s, b->p1s, b->t, b->xn, b->S,
                              s);
            launch_b_cg_resid(g64, b->rows, b->U, src, b->t, b->xn, nullptr, b->rs, nullptr,
                              nullptr, b->S, b->P, 2, s, x, b->xs);
        }
        for (int j = 0; j < len; ++j) {
            if (mixed) {
                launch_b_cg_k1_f32(g32, b->rows32, b->Us, packed, b->rs, b->p0s, b->p1s, b->ts,
                                   b->xs, b->S, b->P, s, b->ru_soft);
                launch_b_cg_k2_f32(g32b, rows32b, b->Us, packed, b->ts, b->p0s, b->p1s, b->xs,
                                   b->rs, b->S, b->P, shifted, s);
            } else {
                launch_b_cg_k1_f64(g64, b->rows, b->U, b->r, b->p0, b->p1, b->t, b->S, b->P, s);
                launch_b_cg_k2_f64(g64b, rows64b, b->U, b->t, b->p0, b->p1, x, b->r, b->S, b->P,
                                   s);
            }
        }
    };
    auto queue_chunk = [&](int len, bool ru_slot) -> int {
        ru_slot = ru_slot && mixed;
        if (b->graphs) {
            const qed2d_batch::GraphKey key{len, ru_slot + 2 * b->ru_soft, mixed, packed, shifted,
                                            (const void *)x,
                                            (const void *)src, b->rows, rows64b, b->rows32,
                                            rows32b, b->rows_q};
            auto it = b->gexec.find(key);
            if (it == b->gexec.end()) {
                if (!b->cap)
                    BHIP(b, hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
                BHIP(b, hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
                issue(b->cap, len, ru_slot);
                hipGraph_t g = nullptr;
                const hipError_t e1 = hipStreamEndCapture(b->cap, &g);
                hipGraphExec_t ex = nullptr;
                const hipError_t e2 =
                    e1 == hipSuccess ? hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) : e1;
                if (g) (void)hipGraphDestroy(g);
                if (e2 != hipSuccess)
                    return bfail(b, QED2D_EHIP, "capturing a CG chunk: %s", hipGetErrorString(e2));
                it = b->gexec.emplace(key, ex).first;
            }
            BHIP(b, hipGraphLaunch(it->second, st));
        } else {
            issue(st, len, ru_slot);
        }
        launched_k += len;
        BTRY(blaunched(b));
        BHIP(b, hipMemcpyAsync(b->hSring + (long)slot * b->n, b->S, sizeof(SolverState) * b->n,
                               hipMemcpyDeviceToHost, st));
        BHIP(b, hipEventRecord(b->sev[slot], st));
        qslot[(qhead + nq) % 2] = slot;
        qlen[(qhead + nq) % 2] = len;
        nq += 1;
        slot ^= 1;
        return QED2D_OK;
    };
    BTRY(start());
    BTRY(queue_chunk(chunk, false));
    bool first_look = true;
    for (;;) {
        if (!first_look && nq == 1) BTRY(queue_chunk(chunk, true));
        const int s0 = qslot[qhead];
        qhead = (qhead + 1) % 2;
        nq -= 1;
        BHIP(b, hipEventSynchronize(b->sev[s0]));
        std::memcpy((void *)b->hS, (const void *)(b->hSring + (long)s0 * b->n),
                    sizeof(SolverState) * b->n);
        if (first_look && packed && b->hS[0].bad_links) {
            packed = false;
            BTRY(start());
            BTRY(queue_chunk(chunk, false));
            continue;
        }
        first_look = false;
        // Each running chain predicts the iteration of its next event (its next
        // reliable update or its convergence) and how many iterations its solve
        // has left.  A chunk ends at the first event of a chain whose remaining
        // count is within kSlack of the longest one's; any other chain's event
        // ends it no earlier than kSlack iterations on (a chain that finishes or
        // pauses mid-chunk idles, which delays the batch only if it is near the
        // longest).
        constexpr int kSlack = 16;
        bool running = false, ru_slot = false;
        double rlong = 0.0;
        const int inflight = nq > 0 ? qlen[qhead] : 0;  // iterations of the chunk in flight
        for (int c = 0; c < b->n; ++c) {
            const SolverState &h = b->hS[c];
            ev[c] = 0;
            if (h.status != 0) continue;
            running = true;
            if (h.need_ru || h.ru_req) {
                ev[c] = pred[c].after_ru(h.err, h.tol, h.delta2);
                ru_slot = true;
            } else {
                const double rut = mixed ? h.delta2 * h.rmax2[(h.iter + 1) & 1] : 0.0;
                ev[c] = pred[c].next(h.err, h.iter, h.tol, rut);
                // (its pause, at the K1 ev[c] - 1 iterations on, falls in the chunk in flight)
                if (ev[c] <= inflight) ru_slot = true;
            }
            rem[c] = pred[c].remaining(h.err, h.tol);
            rlong = std::max(rlong, rem[c]);
        }
        if (!running) break;  // (a chunk still queued finds every chain stopped)
        int next = 64;
        for (int c = 0; c < b->n; ++c) {
            if (ev[c] == 0) continue;
            const int p = (rem[c] + kSlack < rlong) ? std::max(ev[c], kSlack) : ev[c];
            if (p < next) next = p;
        }
        if (launched_k > 4L * (maxiter + 2) + 64)
            return bfail(b, QED2D_EHIP, "batched CG: chains stopped advancing");
        chunk = next;
        BTRY(queue_chunk(chunk, ru_slot));
    }
    iters.assign(b->n, 0);
    int bad = -1;
    for (int c = 0; c < b->n; ++c) {
        iters[c] = b->hS[c].iter;
        if (b->hS[c].status != 1 && bad < 0) bad = c;
    }
    if (bad >= 0)
        return bfail(b, QED2D_ENOCONV, "CG on chain %d not converged after %d iterations "
                     "(|r|^2 = %g, tol = %g)", bad, maxiter, b->hS[bad].err, b->hS[bad].tol);
    return QED2D_OK;
}

// The batch's autotuner: times each tile height of the Q apply and every
// (K1, K2) pair of heights of the FP64 and the FP32 CG on the batch's own
// lattices and links (a Gaussian source per chain, solves of `iters`
// iterations that cannot converge: tolerance 0) and keeps the fastest.  The
// batch's pseudofermion fields (X, F) serve as scratch.
int batch_autotune(qed2d_batch *b, int iters)
{
    const long nel = 2 * b->vtot;
    hipStream_t st = b->stream;
    BTRY(balloc(b, &b->X, nel));
    BTRY(balloc(b, &b->F, nel));
    for (int c = 0; c < b->n; ++c) {
        const BChain &ch = b->hch[c];
        uint32_t k0, k1;
        key_of(0x74756e65ull + (uint64_t)c, (uint32_t)c, &k0, &k1);
        launch_fill_pseudofermion(ch.L, b->F + 2 * ch.off, k0, k1, 0, 0, st);
    }
    std::vector<double> am0(b->n, 0.1);
    BTRY(set_masses(b, am0.data()));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BHIP(b, hipEventCreate(&e0));
    BHIP(b, hipEventCreate(&e1));
    int rc = QED2D_OK;
    // one warm-up run, then the mean of three (statuses of the cut-off solves
    // ignored)
    auto timed = [&](const std::function<int()> &fn, double *us) -> int {
        int r = fn();
        if (r != QED2D_OK && r != QED2D_ENOCONV) return r;
        if (hipEventRecord(e0, st) != hipSuccess) return QED2D_EHIP;
        for (int rep = 0; rep < 3; ++rep) {
            r = fn();
            if (r != QED2D_OK && r != QED2D_ENOCONV) return r;
        }
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return QED2D_EHIP;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return QED2D_EHIP;
        *us = 1e3 * ms / 3;
        return QED2D_OK;
    };
    const int cand[5] = {1, 2, 4, 8, 16};
    double best = 1e30, us = 0.0;
    int pick = b->rows_q;
    for (int R : cand) {
        if (rc) break;
        BGeo gq;
        rc = get_geo(b, R, false, &gq);
        if (rc) break;
        rc = timed([&] {
            launch_b_q_apply(gq, R, b->U, b->F, b->X, nullptr, st);
            return blaunched(b);
        }, &us);
        if (!rc && us < best) best = us, pick = R;
    }
    if (!rc) b->rows_q = pick;
    bool pairs = true;
    for (const BChain &ch : b->hch) pairs = pairs && pair_ok(ch.L);
    std::vector<int> it;
    // (the solves are queued in one chunk, one look at the states at the end,
    // and make no reliable update before their last iteration)
    const double keep_delta = b->ru_delta;
    b->ru_delta = 1e-30;
    // (direct launches: a graph per tile pair and chunk would only fill the cache)
    const bool keep_graphs = b->graphs;
    b->graphs = false;
    for (int prec = 0; prec < 2 && !rc; ++prec) {
        if (prec == 1 && !pairs) break;  // the FP32 batch needs even L1 >= 4 everywhere
        int &t1 = prec ? b->rows32 : b->rows;
        int &t2 = prec ? b->rows32_k2 : b->rows_k2;
        const int keep1 = t1, keep2 = t2;
        int p1 = keep1, p2 = keep2;
        // (two passes over all pairs, the times of a pair added: the pairs lie
        // within a few per cent of each other)
        double tsum[5][5] = {};
        for (int pass = 0; pass < 2 && !rc; ++pass) {
            for (int i = 0; i < 5 && !rc; ++i) {
                for (int j = 0; j < 5; ++j) {
                    t1 = cand[i];
                    t2 = cand[j];
                    rc = timed([&] {
                        return batch_cg(b, b->X, b->F, iters, 0.0,
                                        prec ? QED2D_PREC_MIXED : QED2D_PREC_F64, it, iters + 1);
                    }, &us);
                    if (rc) break;
                    tsum[i][j] += us;
                }
            }
        }
        best = 1e30;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j)
                if (tsum[i][j] > 0.0 && tsum[i][j] < best)
                    best = tsum[i][j], p1 = cand[i], p2 = cand[j];
        t1 = rc ? keep1 : p1;
        t2 = rc ? keep2 : (p2 == p1 ? 0 : p2);
    }
    b->ru_delta = keep_delta;
    b->graphs = keep_graphs;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return bfail(b, rc, "batch autotune failed: %s", b->err.c_str());
    b->err.clear();
    BHIP(b, hipStreamSynchronize(st));
    return QED2D_OK;
}

int batch_update_momenta(qed2d_batch *b, double eps, int maxiter, double tol, int precision,
                         std::vector<long> &iters)
{
    std::vector<int> it;
    BTRY(batch_cg(b, b->X, b->F, maxiter, tol, precision, it));
    for (int c = 0; c < b->n; ++c) iters[c] += it[c];
    BGeo gq;
    BTRY(get_geo(b, b->rows_q, false, &gq));
    launch_b_q_apply(gq, b->rows_q, b->U, b->X, b->Y, nullptr, b->stream);
    launch_b_force_mom(b->dch, b->n, b->vtot, b->U, b->X, b->Y, b->mom, eps, b->stream);
    return blaunched(b);
}

void batch_updt(qed2d_batch *b, double eps)
{
    launch_updt_fld(arena(b), b->U, b->mom, eps, b->stream);
}

int batch_md(qed2d_batch *b, int integrator, double eps, int ns, int maxiter, double tol,
             int precision, std::vector<long> &iters)
{
    if (integrator == QED2D_OMF4) {
        const double r1 = OMF_R1, r2 = OMF_R2, r3 = OMF_R3, r4 = OMF_R4;
        const double r5 = 0.5 - r1 - r3, r6 = 1.0 - 2.0 * (r2 + r4);
        const double pc[6] = {r1, r3, r5, r5, r3, r1};
        const double uc[5] = {r2, r4, r6, r4, r2};
        for (int i = 0; i < ns; ++i)
            for (int s = 0; s < 6; ++s) {
                BTRY(batch_update_momenta(b, pc[s] * eps, maxiter, tol, precision, iters));
                if (s < 5) batch_updt(b, eps * uc[s]);
            }
        return blaunched(b);
    }
    BTRY(batch_update_momenta(b, 0.5 * eps, maxiter, tol, precision, iters));
    for (int i = 0; i < ns - 1; ++i) {
        batch_updt(b, eps);
        BTRY(batch_update_momenta(b, eps, maxiter, tol, precision, iters));
    }
    batch_updt(b, eps);
    BTRY(batch_update_momenta(b, 0.5 * eps, maxiter, tol, precision, iters));
    return blaunched(b);
}


}  // namespace

extern "C" {

int qed2d_batch_create(qed2d_batch **out, int nchain, const int *L1, const int *L2,
                       const double *beta, int bc_flags, void *stream)
{
    if (!out) return QED2D_EINVAL;
    *out = nullptr;
    if (nchain < 1 || !L1 || !L2 || !beta) return QED2D_EINVAL;
    if (bc_flags & ~QED2D_BC_APERIODIC_T) return QED2D_EINVAL;
    qed2d_batch *b = new qed2d_batch();
    b->stream = (hipStream_t)stream;
    b->n = nchain;
    b->bc = bc_flags;
    b->hch.resize(nchain);
    long off = 0, pb = 0;
    for (int c = 0; c < nchain; ++c) {
        if (L1[c] < 2 || L2[c] < 2 || (long)L1[c] * L2[c] > (1L << 31)) {
            delete b;
            return QED2D_EINVAL;
        }
        BChain &ch = b->hch[c];
        ch.L.L1 = L1[c];
        ch.L.L2 = L2[c];
        ch.L.V = (long)L1[c] * L2[c];
        ch.L.bc = bc_flags;
        ch.off = off;
        ch.uoff = off;
        ch.pb = (int)pb;
        ch.pad = 0;
        ch.am0 = 0.0;
        ch.beta = beta[c];
        off += ch.L.V;
        pb += march_grid(ch.L, 1).nblocks();
    }
    b->vtot = off;
    b->nU = 2 * off;
    default_batch_tiles(b);
    b->last_dh.assign(nchain, 0.0);
    const long nel = 2 * off;
    bool ok = hipMalloc((void **)&b->dch, sizeof(BChain) * nchain) == hipSuccess &&
              hipMalloc((void **)&b->U, sizeof(double2) * nel) == hipSuccess &&
              hipMalloc((void **)&b->S, sizeof(SolverState) * nchain) == hipSuccess &&
              hipHostMalloc((void **)&b->hS, sizeof(SolverState) * nchain) == hipSuccess &&
              hipMalloc((void **)&b->part, sizeof(double) * 7 * pb) == hipSuccess &&
              hipMalloc((void **)&b->cnt, sizeof(int) * 5 * nchain) == hipSuccess;
    if (ok)
        ok = hipMemcpy(b->dch, b->hch.data(), sizeof(BChain) * nchain, hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMemset(b->cnt, 0, sizeof(int) * 5 * nchain) == hipSuccess;
    if (!ok) {
        qed2d_batch_destroy(b);
        return QED2D_ENOMEM;
    }
    b->P.tt[0] = b->part;
    b->P.tt[1] = b->part + 2 * pb;
    b->P.rr[0] = b->part + 4 * pb;
    b->P.rr[1] = b->part + 5 * pb;
    b->P.b = b->part + 6 * pb;
    b->P.ntt[0] = b->cnt;
    b->P.ntt[1] = b->cnt + nchain;
    b->P.nrr[0] = b->cnt + 2 * nchain;
    b->P.nrr[1] = b->cnt + 3 * nchain;
    b->P.nb = b->cnt + 4 * nchain;
    b->sub.assign(nchain, nullptr);
    for (int c = 0; c < nchain; ++c) {
        int rc = qed2d_ctx_create(&b->sub[c], L1[c], L2[c], beta[c], bc_flags, stream);
        if (rc != QED2D_OK) {
            qed2d_batch_destroy(b);
            return rc;
        }
        qed2d_ctx_set_sync(b->sub[c], 0);
    }
    *out = b;
    return QED2D_OK;
}

void qed2d_batch_destroy(qed2d_batch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    for (qed2d_ctx *c : b->sub)
        if (c) qed2d_ctx_destroy(c);
    for (auto &kv : b->geo) {
        (void)hipFree(kv.second.blk0);
        (void)hipFree(kv.second.shape);
    }
    if (b->shared) b->U = nullptr;  // the caller's links
    void *dev[] = {b->dch, b->U, b->Ucp, b->X, b->F, b->Y, b->r, b->p0, b->p1, b->t, b->Us,
                   b->rs, b->p0s, b->p1s, b->ts, b->xs, b->mom, b->S, b->part, b->cnt, b->rblk[0], b->rblk[1], b->rpart,
                   b->dred, b->dkeys, b->xn};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (b->hS) (void)hipHostFree(b->hS);
    if (b->hSring) (void)hipHostFree(b->hSring);
    if (b->hred) (void)hipHostFree(b->hred);
    if (b->hkeys) (void)hipHostFree(b->hkeys);
    for (hipEvent_t e : b->sev) if (e) (void)hipEventDestroy(e);
    for (auto &kv : b->gexec) (void)hipGraphExecDestroy(kv.second);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
}

const char *qed2d_batch_last_error(const qed2d_batch *b)
{
    return b ? b->err.c_str() : "null batch";
}

int qed2d_batch_size(const qed2d_batch *b, int *nchain, long *sites)
{
    if (!b) return QED2D_EINVAL;
    if (nchain) *nchain = b->n;
    if (sites) *sites = b->vtot;
    return QED2D_OK;
}

int qed2d_batch_offset(const qed2d_batch *b, int chain, long *site_offset)
{
    if (!b || !site_offset || chain < 0 || chain >= b->n) return QED2D_EINVAL;
    *site_offset = b->hch[chain].off;
    return QED2D_OK;
}

int qed2d_batch_links(qed2d_batch *b, int chain, double **U)
{
    if (!b || !U) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set has no links of its own");
    if (chain < 0 || chain >= b->n) return bfail(b, QED2D_EINVAL, "no chain %d", chain);
    *U = (double *)(b->U + 2 * b->hch[chain].off);
    return QED2D_OK;
}

int qed2d_batch_get_links(qed2d_batch *b, int chain, double *dst)
{
    if (!b || !dst) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set has no links of its own");
    if (chain < 0 || chain >= b->n) return bfail(b, QED2D_EINVAL, "no chain %d", chain);
    const BChain &ch = b->hch[chain];
    BHIP(b, hipMemcpyAsync(dst, b->U + 2 * ch.off, sizeof(double2) * 2 * ch.L.V, hipMemcpyDefault,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return QED2D_OK;
}

int qed2d_batch_set_links(qed2d_batch *b, int chain, const double *src)
{
    if (!b || !src) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set has no links of its own");
    if (chain < 0 || chain >= b->n) return bfail(b, QED2D_EINVAL, "no chain %d", chain);
    const BChain &ch = b->hch[chain];
    BHIP(b, hipMemcpyAsync(b->U + 2 * ch.off, src, sizeof(double2) * 2 * ch.L.V, hipMemcpyDefault,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return QED2D_OK;
}

int qed2d_batch_set_tiles(qed2d_batch *b, int q64, int cg64, int cg32)
{
    if (!b) return QED2D_EINVAL;
    auto ok = [](int r) { return r == 0 || r == 1 || r == 2 || r == 4 || r == 8 || r == 16; };
    if (!ok(q64) || !ok(cg64) || !ok(cg32))
        return bfail(b, QED2D_EINVAL, "tile rows must be 0, 1, 2, 4, 8 or 16");
    if (q64) b->rows_q = q64;
    if (cg64) {
        b->rows = cg64;
        b->rows_k2 = 0;
    }
    if (cg32) {
        b->rows32 = cg32;
        b->rows32_k2 = 0;
    }
    return QED2D_OK;
}

int qed2d_batch_set_tiles_k2(qed2d_batch *b, int cg64, int cg32)
{
    if (!b) return QED2D_EINVAL;
    auto ok = [](int r) { return r == 0 || r == 1 || r == 2 || r == 4 || r == 8 || r == 16; };
    if (!ok(cg64) || !ok(cg32))
        return bfail(b, QED2D_EINVAL, "tile rows must be 0, 1, 2, 4, 8 or 16");
    b->rows_k2 = cg64;
    b->rows32_k2 = cg32;
    return QED2D_OK;
}

int qed2d_batch_set_reliable_mode(qed2d_batch *b, int on_device)
{
    if (!b) return QED2D_EINVAL;
    b->ru_soft = on_device != 0;
    return QED2D_OK;
}

int qed2d_batch_set_graphs(qed2d_batch *b, int on)
{
    if (!b) return QED2D_EINVAL;
    b->graphs = on != 0;
    return QED2D_OK;
}

int qed2d_batch_get_tiles(qed2d_batch *b, int *tiles)
{
    if (!b || !tiles) return QED2D_EINVAL;
    tiles[0] = b->rows_q;
    tiles[1] = b->rows;
    tiles[2] = b->rows32;
    tiles[3] = b->rows_k2 ? b->rows_k2 : b->rows;
    tiles[4] = b->rows32_k2 ? b->rows32_k2 : b->rows32;
    return QED2D_OK;
}

int qed2d_batch_autotune(qed2d_batch *b, int iters, int *tiles)
{
    if (!b) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set tunes through its context");
    if (iters < 4) iters = 4;
    BTRY(batch_autotune(b, iters));
    if (tiles) qed2d_batch_get_tiles(b, tiles);
    return QED2D_OK;
}

int qed2d_batch_hot_start(qed2d_batch *b, const uint64_t *seed)
{
    if (!b || !seed) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set has no links of its own");
    for (int c = 0; c < b->n; ++c) {
        uint32_t k0, k1;
        key_of(seed[c], (uint32_t)c, &k0, &k1);
        launch_hot_start(b->hch[c].L, b->U + 2 * b->hch[c].off, k0, k1, b->stream);
    }
    BTRY(blaunched(b));
    BHIP(b, hipStreamSynchronize(b->stream));
    return QED2D_OK;
}

int qed2d_batch_create_rhs(qed2d_batch **out, int L1, int L2, int nrhs, double beta,
                           int bc_flags, void *stream)
{
    if (!out) return QED2D_EINVAL;
    *out = nullptr;
    if (nrhs < 1 || L1 < 2 || L2 < 2 || (long)L1 * L2 * nrhs > (1L << 33)) return QED2D_EINVAL;
    if (bc_flags & ~QED2D_BC_APERIODIC_T) return QED2D_EINVAL;
    qed2d_batch *b = new qed2d_batch();
    b->stream = (hipStream_t)stream;
    b->n = nrhs;
    b->bc = bc_flags;
    b->shared = true;
    b->hch.resize(nrhs);
    Latt L;
    L.L1 = L1;
    L.L2 = L2;
    L.V = (long)L1 * L2;
    L.bc = bc_flags;
    const long nbc = march_grid(L, 1).nblocks();
    for (int c = 0; c < nrhs; ++c) {
        BChain &ch = b->hch[c];
        ch.L = L;
        ch.off = c * L.V;
        ch.uoff = 0;
        ch.pb = (int)(c * nbc);
        ch.pad = 0;
        ch.am0 = 0.0;
        ch.beta = beta;
    }
    b->vtot = nrhs * L.V;
    b->nU = 2 * L.V;
    b->last_dh.assign(nrhs, 0.0);
    const long pb = nbc * nrhs;
    bool ok = hipMalloc((void **)&b->dch, sizeof(BChain) * nrhs) == hipSuccess &&
              hipMalloc((void **)&b->S, sizeof(SolverState) * nrhs) == hipSuccess &&
              hipHostMalloc((void **)&b->hS, sizeof(SolverState) * nrhs) == hipSuccess &&
              hipMalloc((void **)&b->part, sizeof(double) * 7 * pb) == hipSuccess &&
              hipMalloc((void **)&b->cnt, sizeof(int) * 5 * nrhs) == hipSuccess;
    if (ok)
        ok = hipMemcpy(b->dch, b->hch.data(), sizeof(BChain) * nrhs, hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMemset(b->cnt, 0, sizeof(int) * 5 * nrhs) == hipSuccess;
    if (!ok) {
        qed2d_batch_destroy(b);
        return QED2D_ENOMEM;
    }
    b->P.tt[0] = b->part;
    b->P.tt[1] = b->part + 2 * pb;
    b->P.rr[0] = b->part + 4 * pb;
    b->P.rr[1] = b->part + 5 * pb;
    b->P.b = b->part + 6 * pb;
    b->P.ntt[0] = b->cnt;
    b->P.ntt[1] = b->cnt + nrhs;
    b->P.nrr[0] = b->cnt + 2 * nrhs;
    b->P.nrr[1] = b->cnt + 3 * nrhs;
    b->P.nb = b->cnt + 4 * nrhs;
    *out = b;
    return QED2D_OK;
}

}  // extern "C"

namespace {

int cg_entry(qed2d_batch *b, double *x, const double *src, const double *am0, int maxiter,
             double tol, int precision, int *iters)
{
    if (!x || !src || !am0) return bfail(b, QED2D_EINVAL, "null pointer");
    if (x == src) return bfail(b, QED2D_EINVAL, "x aliases src");
    if (precision != QED2D_PREC_F64 && precision != QED2D_PREC_MIXED)
        return bfail(b, QED2D_EINVAL, "unknown precision %d", precision);
    if (maxiter < 1) return bfail(b, QED2D_EINVAL, "maxiter must be >= 1");
    BTRY(set_masses(b, am0));
    std::vector<int> it;
    const int rc = batch_cg(b, (double2 *)x, (const double2 *)src, maxiter, tol, precision, it);
    if (iters)
        for (int c = 0; c < b->n && c < (int)it.size(); ++c) iters[c] = it[c];
    BHIP(b, hipStreamSynchronize(b->stream));
    return rc;
}

}  // namespace

extern "C" {

int qed2d_batch_cg_shared(qed2d_batch *b, const double *U, double *x, const double *src,
                          const double *am0, int maxiter, double tol, int precision, int *iters)
{
    if (!b) return QED2D_EINVAL;
    if (!b->shared) return bfail(b, QED2D_EINVAL, "not a right-hand-side set");
    if (!U) return bfail(b, QED2D_EINVAL, "null links");
    b->U = (double2 *)U;
    const int rc = cg_entry(b, x, src, am0, maxiter, tol, precision, iters);
    b->U = nullptr;
    return rc;
}

int qed2d_batch_cg(qed2d_batch *b, double *x, const double *src, const double *am0,
                   int maxiter, double tol, int precision, int *iters)
{
    if (!b) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set: use qed2d_batch_cg_shared");
    return cg_entry(b, x, src, am0, maxiter, tol, precision, iters);
}

}  // extern "C"

namespace qed2d {

int batch_rhs_cg(qed2d_batch *b, const double *U, double *x, const double *src, double am0,
                 int maxiter, double eps, int precision, int *iters, int rows_q, int rows,
                 int rows32, int rows_k2, int rows32_k2, double ru_delta, std::string *err)
{
    std::vector<double> m(b->n, am0);
    return batch_rhs_cg_shifted(b, U, x, src, m.data(), nullptr, maxiter, eps, precision, iters,
                                rows_q, rows, rows32, rows_k2, rows32_k2, ru_delta, err);
}

int batch_rhs_cg_shifted(qed2d_batch *b, const double *U, double *x, const double *src,
                         const double *am0, const double *mu2, int maxiter, double eps,
                         int precision, int *iters, int rows_q, int rows, int rows32,
                         int rows_k2, int rows32_k2, double ru_delta, std::string *err,
                         bool guess)
{
    b->rows_q = rows_q;
    b->rows = rows;
    b->rows32 = rows32;
    b->rows_k2 = rows_k2;
    b->rows32_k2 = rows32_k2;
    b->ru_delta = ru_delta;
    if (mu2) b->mu2.assign(mu2, mu2 + b->n);
    b->guess = guess;
    const int rc = qed2d_batch_cg_shared(b, U, x, src, am0, maxiter, eps, precision, iters);
    b->guess = false;
    b->mu2.clear();
    if (rc && err) *err = b->err;
    return rc;
}

}  // namespace qed2d

extern "C" {

int qed2d_batch_hmc_nf2(qed2d_batch *b, const double *am0, double eps, int ns, int integrator,
                        int maxiter, double tol, int precision, double acc_tol,
                        const uint64_t *seed, uint32_t traj, double *hini, double *hfin,
                        int *accepted, int *cg_iters)
{
    if (!b) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set has no links of its own");
    if (!am0 || !seed || !hini || !hfin || !accepted)
        return bfail(b, QED2D_EINVAL, "null pointer");
    if (precision != QED2D_PREC_F64 && precision != QED2D_PREC_MIXED)
        return bfail(b, QED2D_EINVAL, "unknown precision %d", precision);
    const long nel = 2 * b->vtot;
    hipStream_t st = b->stream;
    BTRY(balloc(b, &b->Ucp, nel));
    BTRY(balloc(b, &b->X, nel));
    BTRY(balloc(b, &b->F, nel));
    BTRY(balloc(b, &b->Y, nel));
    BTRY(balloc(b, &b->mom, nel));
    BTRY(set_masses(b, am0));
    BTRY(red_setup(b));
    launch_copy(b->Ucp, b->U, nel, st);
    std::vector<double> h0h(b->n), h0l(b->n);
    // momenta and pseudofermions of every chain (its own key) in one launch; the
    // pinned keys are free again: the previous call ended with a synchronisation
    for (int c = 0; c < b->n; ++c) {
        uint32_t k0, k1;
        key_of(seed[c], (uint32_t)c, &k0, &k1);
        b->hkeys[c] = make_uint2(k0, k1);
    }
    BHIP(b, hipMemcpyAsync(b->dkeys, b->hkeys, sizeof(uint2) * b->n, hipMemcpyHostToDevice, st));
    launch_b_fill_mom_pf(b->dch, b->n, b->vtot, b->dkeys, traj, b->mom, b->X, st);
    BTRY(blaunched(b));
    // hini = H + X^dag X  (src/QED2dHMC.jl:74), per chain, each reduction one
    // launch over all chains into dred[k * n + c] (k = 0, 1: H and X^dag X before
    // the MD; 2, 3, 4: H, Re X^dag F and |Y|^2 after it), read after the
    // trajectory's last launch
    RedOut *hr = b->hred;
    const int n = b->n;
    launch_b_observables(b->dch, b->rblk[0], n, b->nrblk[0], b->U, b->mom, b->rpart, b->dred,
                         st);
    launch_b_dot_dd(b->dch, b->rblk[1], n, b->nrblk[1], b->X, nullptr, b->rpart, b->dred + n,
                    st);
    BTRY(blaunched(b));
    BGeo gq;
    BTRY(get_geo(b, b->rows_q, false, &gq));
    launch_b_q_apply(gq, b->rows_q, b->U, b->X, b->F, nullptr, st);  // F = Q X  (:76-78)
    std::vector<long> its(b->n, 0);
    int rc = batch_md(b, integrator, eps, ns, maxiter, tol, precision, its);
    if (cg_iters)
        for (int c = 0; c < b->n; ++c) cg_iters[c] = (int)its[c];
    if (rc != QED2D_OK) {
        std::string msg = b->err;
        launch_copy(b->U, b->Ucp, nel, st);
        (void)hipStreamSynchronize(st);
        b->err = msg;
        for (int c = 0; c < b->n; ++c) accepted[c] = 0;
        return rc;
    }
    if (acc_tol > 0.0) {
        // FP64 check of the last MD solve: re-solve from it where its true
        // residual is above acc_tol (single-lattice acc_tol path)
        for (int c = 0; c < b->n; ++c) {
            const long o = 2 * b->hch[c].off;
            int it2 = 0;
            rc = qed2d_cg(b->sub[c], (double *)(b->X + o), (const double *)(b->U + o),
                          (const double *)(b->F + o), am0[c], 0.0, QED2D_OP_SQR_MSQ, maxiter,
                          acc_tol, 1, QED2D_PREC_F64, &it2);
            if (rc) return subfail(b, c, rc);
        }
        launch_b_q_apply(gq, b->rows_q, b->U, b->X, b->Y, nullptr, st);
    }
    launch_b_observables(b->dch, b->rblk[0], n, b->nrblk[0], b->U, b->mom, b->rpart,
                         b->dred + 2 * n, st);
    launch_b_dot_dd(b->dch, b->rblk[1], n, b->nrblk[1], b->X, b->F, b->rpart, b->dred + 3 * n,
                    st);
    if (acc_tol > 0.0)
        launch_b_dot_dd(b->dch, b->rblk[1], n, b->nrblk[1], b->Y, nullptr, b->rpart,
                        b->dred + 4 * n, st);
    BTRY(blaunched(b));
    BHIP(b, hipMemcpyAsync(hr, b->dred, sizeof(RedOut) * (acc_tol > 0.0 ? 5 : 4) * n,
                           hipMemcpyDeviceToHost, st));
    BHIP(b, hipStreamSynchronize(st));
    for (int c = 0; c < b->n; ++c) {
        const BChain &ch = b->hch[c];
        const long o = 2 * ch.off;
        dd h0, xx;
        ctx_hamiltonian_from(b->sub[c], hr[c], &h0.hi, &h0.lo);
        ctx_dot_dd_from(hr[n + c], &xx.hi, &xx.lo);
        h0 = dd_add(h0, xx);
        h0h[c] = h0.hi;
        h0l[c] = h0.lo;
        dd h1, re;
        ctx_hamiltonian_from(b->sub[c], hr[2 * n + c], &h1.hi, &h1.lo);
        ctx_dot_dd_from(hr[3 * n + c], &re.hi, &re.lo);
        if (acc_tol <= 0.0) {
            h1 = dd_add(h1, re);  // hfin = H + Re X^dag F  (:82)
        } else {
            dd yy;
            ctx_dot_dd_from(hr[4 * n + c], &yy.hi, &yy.lo);
            h1 = dd_add(h1, dd{2.0 * re.hi, 2.0 * re.lo});
            h1 = dd_add(h1, dd{-yy.hi, -yy.lo});
        }
        hini[c] = h0h[c] + h0l[c];
        hfin[c] = h1.hi + h1.lo;
        const dd dh = dd_add(h1, dd{-h0h[c], -h0l[c]});
        b->last_dh[c] = dh.hi + dh.lo;
        // Metropolis (src/QED2dHMC.jl:95-102)
        const double pacc = std::exp(-(hfin[c] - hini[c]));
        if (pacc < 1.0)
            accepted[c] = (pacc > qed2d_metropolis_uniform(seed[c], (uint32_t)c, traj)) ? 1 : 0;
        else
            accepted[c] = 1;
        if (!accepted[c])
            launch_copy(b->U + o, b->Ucp + o, 2 * ch.L.V, st);
    }
    BTRY(blaunched(b));
    BHIP(b, hipStreamSynchronize(st));
    return QED2D_OK;
}

int qed2d_batch_last_dh(const qed2d_batch *b, double *dH)
{
    if (!b || !dH) return QED2D_EINVAL;
    for (int c = 0; c < b->n; ++c) dH[c] = b->last_dh[c];
    return QED2D_OK;
}

int qed2d_batch_observables(qed2d_batch *b, double *plaquette, double *qtop, double *action)
{
    if (!b) return QED2D_EINVAL;
    if (b->shared) return bfail(b, QED2D_EINVAL, "a right-hand-side set has no links of its own");
    for (int c = 0; c < b->n; ++c) {
        const double *Uc = (const double *)(b->U + 2 * b->hch[c].off);
        int rc = QED2D_OK;
        if (plaquette && (rc = qed2d_plaquette(b->sub[c], Uc, &plaquette[c]))) return subfail(b, c, rc);
        if (qtop && (rc = qed2d_qtop(b->sub[c], Uc, &qtop[c]))) return subfail(b, c, rc);
        if (action && (rc = qed2d_action(b->sub[c], Uc, &action[c]))) return subfail(b, c, rc);
    }
    return QED2D_OK;
}

}  // extern "C"
