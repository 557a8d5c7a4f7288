// qed2d_noalloc.hip -- qed2d_hamiltonian_noalloc: H, S_g, <P> and Q of one
// field with no context and no device allocation, for callers that cannot
// allocate (code under HIP graph capture, serving loops).
//
// The reduction's scratch (one double-double partial triple per block and the
// finished sums) lives in `scratch`, a file-scope device array.  In host code
// that name is only a host-side stand-in: the array's device address is taken
// once per device with hipGetSymbolAddress and kept in g_scratch[device].
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/qed2d.h"
#include "qed2d_internal.h"

using namespace qed2d;

// k_observables writes 6 doubles per block of at most 1024 blocks (the
// reduction grid is capped there, so any lattice fits); k_finish_dd writes
// one RedOut after them: 49,216 bytes per device, held until the process ends.
constexpr int kPartDoubles = 6 * 1024;
constexpr int kScratchDoubles = kPartDoubles + (int)(sizeof(RedOut) / sizeof(double));

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device (resolved on the
// first call there).
int scratch_of_device(double **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return QED2D_EHIP;
    if (dev < 0 || dev >= kMaxDevices) return QED2D_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return QED2D_EHIP;
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return QED2D_OK;
}

// The host arithmetic of qed2d_hamiltonian / _action / _plaquette / _qtop
// (qed2d_api.cpp) on the finished double-double sums, operation by operation
// (the __d*_rn forms keep the compiler from fusing a multiply and an add the
// host does separately), so the results carry the same bits.
__global__ void k_energies_finish(const RedOut *red, double V, double beta, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    // each finished sum is first added to a zero double-double (observables())
    const dd z = {0.0, 0.0};
    const dd sp = dd_add(z, dd{red->v[0], red->v[1]});
    const dd sq = dd_add(z, dd{red->v[2], red->v[3]});
    const dd sm = dd_add(z, dd{red->v[4], red->v[5]});
    // g = V - sum Re P as a double-double
    dd g = dd_add(z, V);
    g = dd_add(g, dd{-sp.hi, -sp.lo});
    // H = sum p^2 / 2 + beta g   (hamiltonian_dd; sum p^2 = 0 without momenta)
    const double ph = __dmul_rn(beta, g.hi);
    const double pe = __fma_rn(beta, g.hi, -ph);
    dd h = dd_add(z, dd{__dmul_rn(0.5, sm.hi), __dmul_rn(0.5, sm.lo)});
    h = dd_add(h, dd{ph, __dadd_rn(pe, __dmul_rn(beta, g.lo))});
    out[0] = __dadd_rn(h.hi, h.lo);                           // Hamiltonian
    out[1] = __dmul_rn(beta, __dadd_rn(g.hi, g.lo));          // Action
    out[2] = __ddiv_rn(__dadd_rn(sp.hi, sp.lo), V);           // Plaquette
    out[3] = __ddiv_rn(__dadd_rn(sq.hi, sq.lo), 2.0 * M_PI);  // Qtop
}

}  // namespace

extern "C" int qed2d_hamiltonian_noalloc(int L1, int L2, double beta, const double *mom,
                                         const double *U, double *out, void *stream)
{
    if (!U || !out) return QED2D_EINVAL;
    if (L1 < 2 || L2 < 2 || (long)L1 * L2 > (1L << 31)) return QED2D_EINVAL;
    Latt L;
    L.L1 = L1;
    L.L2 = L2;
    L.V = (long)L1 * L2;
    L.bc = 0;
    // the reduction grid (red_blocks: at most 1024 blocks) must fit the scratch
    const long nb = (L.V + kBlock - 1) / kBlock;
    if (6 * (nb < 1024 ? nb : 1024) > kPartDoubles) return QED2D_EINVAL;
    double *part = nullptr;
    int rc = scratch_of_device(&part);
    if (rc != QED2D_OK) return rc;
    RedOut *red = (RedOut *)(part + kPartDoubles);
    hipStream_t st = (hipStream_t)stream;
    launch_observables(L, (const double2 *)U, mom, nullptr, nullptr, red, part, st);
    launch_k(k_energies_finish, dim3(1), dim3(64), 0, st, (const RedOut *)red, (double)L.V,
             beta, out);
    return hipGetLastError() == hipSuccess ? QED2D_OK : QED2D_EHIP;
}
