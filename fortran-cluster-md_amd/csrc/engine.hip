// engine.hip -- host side of the engine and the C ABI (include/fcmd.h).
//
// Owns the device arena (fcmd_dev.hpp), turns the reference's parameters into
// device constants with the reference's operand order, maps the drop-in entry
// points onto kernel sequences, and replays a captured step graph for the
// ensemble run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/fcmd.h"
#include "fcmd_kernels.hpp"

using namespace fcmd;

namespace {

long long roundup(long long v, long long m) { return (v + m - 1) / m * m; }

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "fcmd: %s failed: %s\n", #call, hipGetErrorString(e_)); \
            return FCMD_ERR_HIP;                                                     \
        }                                                                            \
    } while (0)

// ---------------------------------------------------------------------------
// small kernels of the host interface
__global__ void k_scatter_d(double *dst, const double *src, const int *inv, int abase, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[inv[abase + k]] = src[k];
}
__global__ void k_gather_d(double *dst, const double *src, const int *inv, int abase, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[inv[abase + k]];
}
__global__ void k_scatter_f(float *dst, const double *src, const int *inv, int abase, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[inv[abase + k]] = (float)src[k];
}
__global__ void k_gather_f(double *dst, const float *src, const int *inv, int abase, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = (double)src[inv[abase + k]];
}
__global__ void k_scatter_typ(int *typ, const int *api, const int *inv, int abase, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) typ[inv[abase + k]] = api[k] - 1;
}
__global__ void k_init_slots(int *oid, int *inv, int *typ, int n_total, const int *blk_traj,
                             const TrajParams *tp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_total) return;
    const TrajParams &p = tp[blk_traj[i / BLK]];
    const int rel = i - p.abase;
    oid[i] = rel < p.natom ? rel + 1 : 0;
    inv[i] = i;
    if (rel >= p.natom) typ[i] = -1;
}
// per-trajectory: record the largest cluster of this analysis (-1 for a
// trajectory stopped by an error) and the first passage to fpt_size, at the
// trajectory's own step number (the step about to be integrated: the
// analysis sees the state before it, as main.f:133-164 does)
__global__ void k_record(const ClusterScalars *cls, const TrajScalars *ts, int ntraj, int *series,
                         int *fpt, int fpt_size)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntraj) return;
    const bool ok = !ts[t].err;
    const int big = cls[t].cbig;
    series[t] = ok ? big : -1;
    if (ok && fpt[t] < 0 && big >= fpt_size) fpt[t] = ts[t].step;
}

// which trajectories the next cluster analysis takes: traj, or all (-1)
__global__ void k_set_want(ClusterScalars *cls, int ntraj, int traj)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntraj) cls[t].want = (traj < 0 || t == traj);
}

// the error code of the lowest-numbered trajectory that has one (0: none),
// so that a check of every trajectory costs one 4-byte copy
__global__ __launch_bounds__(256) void k_first_err(const TrajScalars *ts, int ntraj, int *out)
{
    __shared__ int first;
    if (threadIdx.x == 0) first = INT_MAX;
    __syncthreads();
    for (int t = threadIdx.x; t < ntraj; t += blockDim.x)
        if (ts[t].err) {
            atomicMin(&first, t);
            break;
        }
    __syncthreads();
    if (threadIdx.x == 0) *out = first == INT_MAX ? 0 : ts[first].err;
}

// k_fill: n ints of p set to v
__global__ void k_fill(int *p, int n, int v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// Scratch of the allocation-free run (fcmd_run_noalloc): its first-passage
// times and largest-cluster series are staged in this fixed device array
// instead of buffers the library allocates.  Every GPU holds its own copy at
// its own address, kept until the process ends; host code never passes the
// array's name to a kernel or a copy (in host code it is a stand-in, not a
// device address) but the address hipGetSymbolAddress gave for the device
// (scratch_ptr, set up once per device by fcmd_create).
static __device__ __align__(256) int scratch[FCMD_NOALLOC_INTS];

constexpr int MAX_DEVICES = 64;
int *scratch_ptr[MAX_DEVICES] = {};
std::mutex scratch_mu;

// the scratch address on the current device (looked up once per device)
int scratch_on_current_device(int **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return FCMD_ERR_HIP;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "fcmd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return FCMD_ERR_HIP;
        }
        scratch_ptr[dev] = (int *)p;
    }
    *out = scratch_ptr[dev];
    return FCMD_OK;
}

}  // namespace

namespace fcmd {
void launch_neighbor_entry(const DevPtrs &d, int traj, int ntraj, int step, int *step_out,
                           hipStream_t s);
}

// ---------------------------------------------------------------------------
struct fcmd_system {
    int ntraj = 0;
    bool mixed = false;
    int list_width = 256;
    std::vector<fcmd_params> hp;
    std::vector<TrajParams> tp;
    LaunchDims g{};
    long long natoms_total = 0, ncells_total = 0, nlist_total = 0, nhist_total = 0;
    bool single_species = true;
    bool any_fre = false;
    DevPtrs d{};
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                  // the rebuild beside the forces (enqueue_step)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // rebuild beside the forces: for ensembles (one trajectory either
    // rebuilds or not, so it has nothing to overlap); FCMD_OVERLAP=0 or 1
    // in the environment decides instead
    int overlap = 1;
    // every resort-th list rebuild of a trajectory re-sorts its atoms by
    // cell; the others keep them in their slots (FCMD_RESORT, read by
    // fcmd_create; 1 re-sorts at every rebuild)
    int resort = 8;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double *stage = nullptr;    // natom-sized staging (device)
    int *stage_i = nullptr;
    int *dstep = nullptr;       // device int for step round trips
    int *derr = nullptr;        // device int: k_first_err's result
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    int count_nmax = 0, count_warmup = 0;  // fcmd_size_counts (0: off)
    double last_ms = 0.0;
    // fcmd_run's largest-cluster series and first-passage times on the
    // device, kept between runs and grown when a run needs more
    int *dseries = nullptr, *dfpt = nullptr;
    long long series_cap = 0;
    int *scratch = nullptr;     // this device's fixed scratch (fcmd_run_noalloc)
    // HIP events around the first NAN_TIMED analyses of the last run
    static constexpr int NAN_TIMED = 16;
    hipEvent_t an_ev[2 * NAN_TIMED] = {};
    double an_ms = 0.0;
    int an_count = 0;

    template <typename T>
    int alloc(T **p, long long n)
    {
        void *q = nullptr;
        const size_t bytes = (size_t)std::max<long long>(n, 1) * sizeof(T);
        if (hipMalloc(&q, bytes) != hipSuccess) return FCMD_ERR_HIP;
        if (hipMemset(q, 0, bytes) != hipSuccess) return FCMD_ERR_HIP;
        allocs.push_back(q);
        *p = (T *)q;
        return FCMD_OK;
    }
    ~fcmd_system()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (void *p : allocs) (void)hipFree(p);
        if (dseries) (void)hipFree(dseries);
        if (dfpt) (void)hipFree(dfpt);
        for (hipEvent_t e : an_ev)
            if (e) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// pair constants with the reference's operand order
void pair_consts(const fcmd_params &p, TrajParams &t)
{
    auto six = [](double s) { return s * s * s * s * s * s; };
    const int ns = std::max(1, p.nsorts);
    for (int a = 0; a < MAXS; ++a)
        for (int b = 0; b < MAXS; ++b) {
            const int k = a * MAXS + b;
            double c1 = 0, eps = 0;
            if (a < ns && b < ns) {
                if (p.accatom == FCMD_ACC_TABLE) {
                    c1 = six(p.sig_ab[k]);
                    eps = p.eps_ab[k];
                } else if (a == b) {
                    c1 = six(p.sig[a]);  // for-a-lj-n-pbc-m.f:62, :329
                    eps = p.eps[a];
                } else {
                    c1 = six(0.5 * (p.sig[a] + p.sig[b]));  // :334-335 Lorentz-Berthelot
                    eps = std::sqrt(p.eps[a] * p.eps[b]);   // :338
                }
            }
            t.c1[k] = c1;
            t.c2[k] = c1 * c1;
            t.c3[k] = 0.5 * c1;
            t.e48[k] = eps * 48.0;
            t.e4[k] = eps * 4.0;
            t.c1f[k] = (float)t.c1[k];
            t.c2f[k] = (float)t.c2[k];
            t.c3f[k] = (float)t.c3[k];
        }
}

int fill_params(fcmd_system *s, int k, const fcmd_params &p, TrajParams &t)
{
    if (p.natom <= 0 || p.nsorts < 1 || p.nsorts > MAXS) return FCMD_ERR_ARG;
    for (int d = 0; d < 3; ++d) {
        if (!(p.box[d] > 0)) return FCMD_ERR_ARG;
        t.box[d] = p.box[d];
        t.ibox[d] = 1.0 / p.box[d];
    }
    t.atrcut = p.atrcut;
    t.atskin = p.atskin;
    t.atr = p.sig[0];
    t.sig2f = (float)(p.sig[0] * p.sig[0]);
    for (int d = 0; d < 3; ++d) {
        t.boxf[d] = (float)t.box[d];
        t.iboxf[d] = (float)t.ibox[d];
    }
    {
        // the mixed force kernels' cutoff decision band, formed as they form it
        const double rc2 = t.atrcut * t.atrcut;
        t.cut_in2f = (float)(rc2 * (1.0 - 1e-5));
        t.cut_out2f = (float)(rc2 * (1.0 + 1e-5));
    }
    t.eps1 = p.eps[0];
    t.dt = p.dt;
    t.Q = p.Q;
    t.att = p.att;
    for (int a = 0; a < MAXS; ++a) {
        t.mas[a] = a < p.nsorts ? p.mas[a] : 1.0;
        t.imas[a] = 1.0 / t.mas[a];
    }
    for (int a = 0; a < MAXS; ++a) t.mrel[a] = t.mas[a] / t.mas[0];
    pair_consts(p, t);
    t.natom = p.natom;
    t.nsorts = p.nsorts;
    t.nbtim = p.nbtim;
    t.btim = p.btim;
    t.atnenum = p.atnenum;
    t.ethst_nh = p.ethst_nh;
    t.integrator = p.integrator;
    t.neighbor = p.neighbor;
    t.accatom = p.accatom;
    t.stoddard = p.stoddard;
    if (p.nh_mask == 0) {
        t.nh_mask = (1 << MAXS) - 1;
        t.nh_refmass = 1;
        t.nkt = 3.0 * p.natom * CIKB * p.att;  // int-verlet-nh-mav.f:65
    } else {
        t.nh_mask = p.nh_mask;
        t.nh_refmass = 0;
        int nc = 0;
        for (int a = 0; a < p.nsorts; ++a)
            if ((p.nh_mask >> a) & 1) nc += p.num[a];
        t.nkt = 3.0 * nc * CIKB * p.att;
    }
    t.resort = s->resort;
    (void)k;
    return FCMD_OK;
}

int upload_params(fcmd_system *s)
{
    HIPCHK(hipMemcpy(s->d.tp, s->tp.data(), sizeof(TrajParams) * s->ntraj, hipMemcpyHostToDevice));
    return FCMD_OK;
}

int check_traj(fcmd_system *s, int traj) { return (s && traj >= 0 && traj < s->ntraj) ? FCMD_OK : FCMD_ERR_ARG; }

int enqueue_step(fcmd_system *s, bool energy = true);

int sync_err(fcmd_system *s)
{
    hipLaunchKernelGGL(k_first_err, dim3(1), dim3(256), 0, s->stream, s->d.ts, s->ntraj, s->derr);
    HIPCHK(hipGetLastError());
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, s->derr, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return err;
}

void set_want(fcmd_system *s, int traj)
{
    hipLaunchKernelGGL(k_set_want, dim3((s->ntraj + 255) / 256), dim3(256), 0, s->stream, s->d.cls, s->ntraj,
                       traj);
}

// Capture `gsteps` steps into the run graph (replacing any).  The launch
// shapes and tunables are baked into the graph, so it is recaptured whenever
// they change.
int capture_graph(fcmd_system *s, int gsteps)
{
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
        s->graph_steps = 0;
    }
    hipGraph_t gr;
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = FCMD_OK;
    // the energy and virial of the graph's last step only: its scalars are the
    // ones a run can end with
    for (int k = 0; k < gsteps && !rc; ++k) rc = enqueue_step(s, k == gsteps - 1);
    HIPCHK(hipStreamEndCapture(s->stream, &gr));
    if (rc) {
        (void)hipGraphDestroy(gr);
        return rc;
    }
    HIPCHK(hipGraphInstantiate(&s->graph, gr, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(gr));
    s->graph_steps = gsteps;
    return FCMD_OK;
}

// one integrator step of every trajectory.  The Nose-Hoover final velocity
// update is left pending and done by the next drift; finalize() applies it
// wherever v is read or replaced outside the step sequence.
//
// The rebuild is a chain of small dependent kernels over the ~1/nbtim of the
// atoms whose trajectories rebuild now; it leaves most of the GPU idle.  So
// it runs on the side stream, followed there by the rebuilt trajectories'
// forces, while the forces of the trajectories that keep their lists (which
// the rebuild does not touch) run on the engine stream beside it.  In a step
// where few trajectories rebuild, the chain takes about as long as those
// forces, and the cross-stream wait falls on the chain rather than on the
// forces.  Every atom block is computed by the same kernel code either way,
// so the results are the same bits as the forces of all blocks in one launch
// after the rebuild (s->overlap = 0).
//
// energy: the step also forms epot, the virial and the pressure.  A run
// leaves them out of every step whose scalars the next step overwrites
// before anyone can read them; positions, velocities, forces and every other
// scalar are the same bits either way.
int enqueue_step(fcmd_system *s, bool energy)
{
    launch_drift(s->d, s->g, s->mixed, s->stream);
    if (!s->overlap) {
        launch_rebuild(s->d, s->g, s->mixed, s->stream, {}, true);
        launch_force(s->d, s->g, s->mixed, s->single_species, FORCE_STEP, s->stream, energy);
        return FCMD_OK;
    }
    HIPCHK(hipEventRecord(s->ev_fork, s->stream));
    HIPCHK(hipStreamWaitEvent(s->side, s->ev_fork, 0));
    launch_rebuild(s->d, s->g, s->mixed, s->side, {}, true);
    launch_force_atoms(s->d, s->g, s->mixed, s->single_species, FORCE_STEP, s->side, {}, FP_REBUILT, energy);
    HIPCHK(hipEventRecord(s->ev_join, s->side));
    launch_force_atoms(s->d, s->g, s->mixed, s->single_species, FORCE_STEP, s->stream, {}, FP_KEPT, energy);
    HIPCHK(hipStreamWaitEvent(s->stream, s->ev_join, 0));
    launch_traj_post(s->d, s->g, FORCE_STEP, s->stream, {}, energy);
    return FCMD_OK;
}

void finalize(fcmd_system *s) { launch_kick_nh(s->d, s->g, s->mixed, s->stream); }

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int fcmd_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *fcmd_error_text(int code)
{
    switch (code) {
    case FCMD_OK: return "ok";
    case FCMD_ERR_SKIN: return "*** FATAL ERROR! ***  Skin radius exceeded!!";
    case FCMD_ERR_LISTFULL: return "*** FATAL ERROR *** Neighborlist is nearly full!";
    case FCMD_ERR_NH: return "Numerical integration did not converge!";
    case FCMD_ERR_EDGES: return "*** FATAL ERROR *** too many cluster neighbours per atom";
    case FCMD_ERR_CELLS: return "*** FATAL ERROR *** cell grid exceeds its allocation";
    case FCMD_ERR_SCRATCH:
        return "run too large for the allocation-free entry point: (analyses + 1) x trajectories "
               "exceeds FCMD_NOALLOC_INTS";
    case FCMD_ERR_ARG: return "invalid argument";
    case FCMD_ERR_HIP: return "HIP runtime error";
    case FCMD_ERR_NODEVICE: return "no HIP device";
    default: return "unknown error";
    }
}

int fcmd_create(fcmd_system **out, int ntraj, const fcmd_params *params, int precision, int list_width)
{
    if (!out || ntraj <= 0 || !params) return FCMD_ERR_ARG;
    if (precision != FCMD_FP64 && precision != FCMD_MIXED && precision != FCMD_MIXED_KAHAN) return FCMD_ERR_ARG;
    *out = nullptr;
    if (fcmd_device_count() <= 0) return FCMD_ERR_NODEVICE;
    fcmd_system *s = new (std::nothrow) fcmd_system();
    if (!s) return FCMD_ERR_ARG;
    // set-up of the allocation-free run's scratch on this device
    if (int rc0 = scratch_on_current_device(&s->scratch)) {
        delete s;
        return rc0;
    }
    s->ntraj = ntraj;
    s->mixed = precision == FCMD_MIXED || precision == FCMD_MIXED_KAHAN;
    s->g.kahan = precision == FCMD_MIXED_KAHAN;
    if (const char *rs = std::getenv("FCMD_RESORT")) s->resort = std::max(1, std::atoi(rs));
    s->list_width = list_width > 0 ? list_width : 256;
    s->hp.assign(params, params + ntraj);
    s->tp.resize(ntraj);
    int rc;
    long long abase = 0, cbase = 0, lbase = 0, hbase = 0;
    for (int k = 0; k < ntraj; ++k) {
        const fcmd_params &p = params[k];
        TrajParams &t = s->tp[k];
        std::memset(&t, 0, sizeof(t));
        if ((rc = fill_params(s, k, p, t)) != FCMD_OK) {
            delete s;
            return rc;
        }
        if (p.nsorts > 1 || p.accatom != FCMD_ACC_1) s->single_species = false;
        if (p.stoddard == FCMD_ST_FRE) s->any_fre = true;
        const long long npad = roundup(p.natom, BLK);
        // smallest cell side the list radius can reach: min(atrskin, atrcut)
        const double smin = std::min(p.atrskin, p.atrcut) * (1.0 - 1e-6);
        long long nc = 1;
        for (int d = 0; d < 3; ++d) nc *= std::max<long long>(1, (long long)std::floor(p.box[d] / smin));
        const long long ncpad = roundup(nc + 1, SCAN_CHUNK);  // + the end cell past the grid
        t.abase = (int)abase;
        t.npad = (int)npad;
        t.bfirst = (int)(abase / BLK);
        t.nblk = (int)(npad / BLK);
        t.cbase = (int)cbase;
        t.ncpad = (int)ncpad;
        t.cfirst = (int)(cbase / SCAN_CHUNK);
        t.ncblk = (int)(ncpad / SCAN_CHUNK);
        t.lbase = lbase;
        t.lwidth = s->list_width;
        t.lstride = (int)roundup(s->list_width, LIST_ROUND);
        t.hbase = (int)hbase;
        abase += npad;
        cbase += ncpad;
        lbase += npad * t.lstride;
        hbase += p.natom + 1;
    }
    if (abase >= (1LL << 31) || cbase >= (1LL << 31) || lbase >= (1LL << 40)) {
        delete s;
        return FCMD_ERR_ARG;
    }
    s->natoms_total = abase;
    s->ncells_total = cbase;
    s->nlist_total = lbase;
    s->nhist_total = hbase;
    s->g.nblk = (int)(abase / BLK);
    s->g.ncblk = (int)(cbase / SCAN_CHUNK);
    s->g.ntraj = ntraj;
    // 16-bit list entries when every trajectory's slots fit, unless
    // FCMD_LIST_BITS=32 asks for 32-bit ones (same lists, same results)
    s->g.list16 = 1;
    for (auto &t : s->tp)
        if (t.npad > LIST16_MAX) s->g.list16 = 0;
    if (const char *lb = std::getenv("FCMD_LIST_BITS"))
        if (std::strcmp(lb, "32") == 0) s->g.list16 = 0;
    s->g.force_u = 4;
    s->g.nbr_cu = 2;
    s->g.wl_grid = 2048;

    DevPtrs &d = s->d;
    const long long N = abase;
    d.nblk = s->g.nblk;
#define A(ptr, n)                                  \
    if ((rc = s->alloc(&(ptr), (n))) != FCMD_OK) { \
        delete s;                                  \
        return rc;                                 \
    }
    A(d.x, N) A(d.y, N) A(d.z, N) A(d.vx, N) A(d.vy, N) A(d.vz, N);
    A(d.x2, N) A(d.y2, N) A(d.z2, N) A(d.vx2, N) A(d.vy2, N) A(d.vz2, N);
    if (s->mixed) {
        A(d.axf, N) A(d.ayf, N) A(d.azf, N) A(d.axf2, N) A(d.ayf2, N) A(d.azf2, N);
        A(d.cellid, N);
    } else {
        A(d.ax, N) A(d.ay, N) A(d.az, N) A(d.ax2, N) A(d.ay2, N) A(d.az2, N);
    }
    A(d.typ, N) A(d.oid, N) A(d.typ2, N) A(d.oid2, N) A(d.inv, N);
    A(d.blk_traj, s->g.nblk) A(d.cblk_traj, s->g.ncblk);
    A(d.tp, ntraj) A(d.ts, ntraj) A(d.cls, ntraj);
    A(d.part, (long long)s->g.nblk * NPART);
    A(d.wl_blk, s->g.nblk) A(d.wl_cchk, s->g.ncblk) A(d.wl_traj, ntraj) A(d.wl_cnt, WL_LISTS);
    A(d.wl_size, (long long)WL_LISTS * ntraj);
    A(d.cell_cnt, cbase) A(d.cell_start, cbase) A(d.cell_first, cbase) A(d.atom_cell, N) A(d.atom_slot, N) A(d.sorted, N);
    A(d.scan_tmp_ll, s->g.ncblk);
    A(d.nn, N) A(d.posf, N) A(d.cpos, N) A(d.cslot, N);
    if (s->g.list16) {
        A(d.list16, lbase);
    } else {
        A(d.list32, lbase);
    }
    A(d.edeg, N) A(d.edeg1, N) A(d.edges, N * EMAX) A(d.parent, N) A(d.parentf, N) A(d.liq, N);
    A(d.clpart, (long long)s->g.nblk * 4);
    A(d.cs_tmp, N) A(d.cgs_tmp, N) A(d.csf_tmp, N) A(d.cek_tmp, N) A(d.cekf_tmp, N);
    A(d.liqo, N) A(d.rank, N) A(d.o_cl, N) A(d.o_cs, N) A(d.o_cgs, N) A(d.o_clf, N) A(d.o_ccf, N);
    A(d.o_cek, N);
    A(d.o_csi, hbase) A(d.o_csf, hbase) A(d.o_csif, hbase) A(d.o_cgsf, hbase) A(d.o_cgsfa, hbase);
    A(d.o_ceks, hbase) A(d.o_cekf, hbase) A(d.o_ceksf, hbase);
    A(d.o_hacc, hbase) A(d.hacc_stop, ntraj);
    long long maxn = 1;
    for (auto &p : s->hp) maxn = std::max<long long>(maxn, p.natom + 1);
    A(s->stage, maxn) A(s->stage_i, maxn) A(s->dstep, 4) A(s->derr, 4);
#undef A

    std::vector<int> bt(s->g.nblk), ct(s->g.ncblk), wsz((size_t)WL_LISTS * ntraj);
    for (int k = 0; k < ntraj; ++k) {
        const TrajParams &t = s->tp[k];
        for (int b = 0; b < t.nblk; ++b) bt[t.bfirst + b] = k;
        for (int b = 0; b < t.ncblk; ++b) ct[t.cfirst + b] = k;
        wsz[WL_LISTS * k + WL_BLK] = t.nblk;
        wsz[WL_LISTS * k + WL_CCHK] = t.ncblk;
        wsz[WL_LISTS * k + WL_TRAJ] = 1;
    }
    s->overlap = ntraj > 1;
    if (const char *ov = std::getenv("FCMD_OVERLAP")) s->overlap = std::strcmp(ov, "0") != 0;
    if (hipStreamCreate(&s->stream) != hipSuccess || hipEventCreate(&s->ev0) != hipSuccess ||
        hipEventCreate(&s->ev1) != hipSuccess || hipStreamCreate(&s->side) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete s;
        return FCMD_ERR_HIP;
    }
    for (hipEvent_t &e : s->an_ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
            delete s;
            return FCMD_ERR_HIP;
        }
    if (hipMemcpy(d.blk_traj, bt.data(), sizeof(int) * bt.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.cblk_traj, ct.data(), sizeof(int) * ct.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.wl_size, wsz.data(), sizeof(int) * wsz.size(), hipMemcpyHostToDevice) != hipSuccess ||
        upload_params(s) != FCMD_OK) {
        delete s;
        return FCMD_ERR_HIP;
    }
    // scalars: reference start values (read-3d.f:263-268), psi = s = 0
    std::vector<TrajScalars> ts(ntraj);
    for (int k = 0; k < ntraj; ++k) {
        const fcmd_params &p = params[k];
        TrajScalars &t = ts[k];
        std::memset(&t, 0, sizeof(t));
        t.atrskin = p.atrskin;
        const double w = (double)(float)p.nbtim;
        for (int i = 0; i < 256; ++i) t.dmmh[i] = (p.atrskin - p.atrcut) / w;
        t.dmmsu = 256.0 * (p.atrskin - p.atrcut) / w;
        t.step = p.btim;
    }
    if (hipMemcpy(d.ts, ts.data(), sizeof(TrajScalars) * ntraj, hipMemcpyHostToDevice) != hipSuccess) {
        delete s;
        return FCMD_ERR_HIP;
    }
    hipLaunchKernelGGL(k_init_slots, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s->stream, d.oid,
                       d.inv, d.typ, (int)N, d.blk_traj, d.tp);
    if (hipStreamSynchronize(s->stream) != hipSuccess) {
        delete s;
        return FCMD_ERR_HIP;
    }
    *out = s;
    return FCMD_OK;
}

void fcmd_destroy(fcmd_system *s) { delete s; }

int fcmd_set_params(fcmd_system *s, int traj, const fcmd_params *p)
{
    int rc;
    if ((rc = check_traj(s, traj)) || !p) return FCMD_ERR_ARG;
    if (p->natom != s->hp[traj].natom) return FCMD_ERR_ARG;
    TrajParams t = s->tp[traj];
    if ((rc = fill_params(s, traj, *p, t)) != FCMD_OK) return rc;
    s->hp[traj] = *p;
    s->tp[traj] = t;
    if (p->nsorts > 1 || p->accatom != FCMD_ACC_1) s->single_species = false;
    if (p->stoddard == FCMD_ST_FRE) s->any_fre = true;
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
    return upload_params(s);
}

int fcmd_upload(fcmd_system *s, int traj, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *ax,
                const double *ay, const double *az, const int *api)
{
    if (check_traj(s, traj)) return FCMD_ERR_ARG;
    finalize(s);
    const TrajParams &t = s->tp[traj];
    const int n = t.natom;
    const dim3 grid((n + 255) / 256), blk(256);
    const double *src[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    double *dd[9] = {s->d.x, s->d.y, s->d.z, s->d.vx, s->d.vy, s->d.vz, s->d.ax, s->d.ay, s->d.az};
    float *df[3] = {s->d.axf, s->d.ayf, s->d.azf};
    for (int k = 0; k < 9; ++k) {
        if (!src[k]) continue;
        HIPCHK(hipMemcpyAsync(s->stage, src[k], sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
        if (k >= 6 && s->mixed)
            hipLaunchKernelGGL(k_scatter_f, grid, blk, 0, s->stream, df[k - 6], s->stage, s->d.inv, t.abase, n);
        else
            hipLaunchKernelGGL(k_scatter_d, grid, blk, 0, s->stream, dd[k], s->stage, s->d.inv, t.abase, n);
        HIPCHK(hipGetLastError());
    }
    if (api) {
        HIPCHK(hipMemcpyAsync(s->stage_i, api, sizeof(int) * n, hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(k_scatter_typ, grid, blk, 0, s->stream, s->d.typ, s->stage_i, s->d.inv, t.abase, n);
    }
    if (x || y || z) launch_refresh_posf(s->d, traj, n, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return FCMD_OK;
}

int fcmd_download(fcmd_system *s, int traj, double *x, double *y, double *z, double *vx, double *vy,
                  double *vz, double *ax, double *ay, double *az)
{
    if (check_traj(s, traj)) return FCMD_ERR_ARG;
    finalize(s);
    const TrajParams &t = s->tp[traj];
    const int n = t.natom;
    const dim3 grid((n + 255) / 256), blk(256);
    double *dst[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    double *dd[9] = {s->d.x, s->d.y, s->d.z, s->d.vx, s->d.vy, s->d.vz, s->d.ax, s->d.ay, s->d.az};
    float *df[3] = {s->d.axf, s->d.ayf, s->d.azf};
    for (int k = 0; k < 9; ++k) {
        if (!dst[k]) continue;
        if (k >= 6 && s->mixed)
            hipLaunchKernelGGL(k_gather_f, grid, blk, 0, s->stream, s->stage, df[k - 6], s->d.inv, t.abase, n);
        else
            hipLaunchKernelGGL(k_gather_d, grid, blk, 0, s->stream, s->stage, dd[k], s->d.inv, t.abase, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dst[k], s->stage, sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return FCMD_OK;
}

int fcmd_set_scalars(fcmd_system *s, int traj, const fcmd_scalars *sc)
{
    if (check_traj(s, traj) || !sc) return FCMD_ERR_ARG;
    finalize(s);  // the pending update uses the current psi
    HIPCHK(hipStreamSynchronize(s->stream));
    TrajScalars t;
    HIPCHK(hipMemcpy(&t, s->d.ts + traj, sizeof(t), hipMemcpyDeviceToHost));
    t.atrskin = sc->atrskin;
    t.psi = sc->psi;
    t.s = sc->s;
    t.ekin = sc->ekin;
    t.epot = sc->epot;
    t.sumv2 = sc->sumv2;
    std::memcpy(t.dmmh, sc->dmmh, sizeof(t.dmmh));
    t.dmms = sc->dmms;
    t.dmmsu = sc->dmmsu;
    t.dmmax = sc->dmmax;
    t.dmmav = sc->dmmav;
    t.dmmst = sc->dmmst;
    t.dmmi = sc->dmmi;
    t.step = sc->step;
    t.err = sc->err;
    HIPCHK(hipMemcpy(s->d.ts + traj, &t, sizeof(t), hipMemcpyHostToDevice));
    return FCMD_OK;
}

int fcmd_get_scalars(fcmd_system *s, int traj, fcmd_scalars *sc)
{
    if (check_traj(s, traj) || !sc) return FCMD_ERR_ARG;
    TrajScalars t;
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(&t, s->d.ts + traj, sizeof(t), hipMemcpyDeviceToHost));
    sc->atrskin = t.atrskin;
    sc->psi = t.psi;
    sc->s = t.s;
    sc->ekin = t.ekin;
    sc->epot = t.epot;
    sc->sumv2 = t.sumv2;
    std::memcpy(sc->dmmh, t.dmmh, sizeof(t.dmmh));
    sc->dmms = t.dmms;
    sc->dmmsu = t.dmmsu;
    sc->dmmax = t.dmmax;
    sc->dmmav = t.dmmav;
    sc->dmmsr = t.dmmsr;
    sc->dmmst = t.dmmst;
    sc->dmmi = t.dmmi;
    sc->step = t.step;
    sc->npairs = t.npairs;
    sc->nrebuild = t.nrebuild;
    sc->nh_iters = t.nh_iters;
    sc->err = t.err;
    sc->nlist = t.nlist;
    sc->virial = t.virial;
    sc->pressure = t.pressure;
    return FCMD_OK;
}

int fcmd_neighbor(fcmd_system *s, int traj, int *step)
{
    if (check_traj(s, traj) || !step) return FCMD_ERR_ARG;
    launch_neighbor_entry(s->d, traj, s->ntraj, *step, s->dstep, s->stream);
    launch_rebuild(s->d, s->g, s->mixed, s->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(step, s->dstep, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    return sync_err(s);
}

int fcmd_accatom(fcmd_system *s)
{
    if (!s) return FCMD_ERR_ARG;
    finalize(s);  // the pending update reads a, which AccAtom changes
    launch_force(s->d, s->g, s->mixed, s->single_species, FORCE_ACCATOM, s->stream);
    HIPCHK(hipGetLastError());
    return sync_err(s);
}

int fcmd_integrator(fcmd_system *s, int *step)
{
    if (!s || !step) return FCMD_ERR_ARG;
    // the reference passes integrator(i) the step number i
    for (int k = 0; k < s->ntraj; ++k)
        HIPCHK(hipMemcpyAsync((char *)(s->d.ts + k) + offsetof(TrajScalars, step), step, sizeof(int),
                              hipMemcpyHostToDevice, s->stream));
    if (int rc = enqueue_step(s)) return rc;
    finalize(s);  // integrator(step) returns with the step's final velocities
    HIPCHK(hipGetLastError());
    return sync_err(s);
}

int fcmd_stoddard(fcmd_system *s, int traj, int step, int *cl, int *cs, int *cc, int *csi, double *cek,
                  double *ceks, int *clf, int *csf, int *ccf, int *csif, double *cekf, double *ceksf,
                  int *cgs, int *cgsf, int *cgsfa, int *cfpt, int *cfptf, int *counters)
{
    if (check_traj(s, traj)) return FCMD_ERR_ARG;
    const TrajParams &t = s->tp[traj];
    const bool fre = t.stoddard == ST_FRE;
    const int n = t.natom;
    set_want(s, traj);
    finalize(s);  // cluster kinetic energies need the final velocities
    launch_cluster(s->d, s->g, fre, true, s->stream);
    HIPCHK(hipGetLastError());
    int rc = sync_err(s);
    if (rc) return rc;
    ClusterScalars c;
    HIPCHK(hipMemcpy(&c, s->d.cls + traj, sizeof(c), hipMemcpyDeviceToHost));
    auto geti = [&](int *dst, const int *src, long long cnt) -> int {
        if (dst) HIPCHK(hipMemcpy(dst, src, sizeof(int) * cnt, hipMemcpyDeviceToHost));
        return FCMD_OK;
    };
    auto getd = [&](double *dst, const double *src, long long cnt) -> int {
        if (dst) HIPCHK(hipMemcpy(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost));
        return FCMD_OK;
    };
    const int a = t.abase, h = t.hbase;
    // csi(0) and ceks(0) are never reset by the reference; keep what the caller had
    const int csi0 = csi ? csi[0] : 0;
    const double ceks0 = ceks ? ceks[0] : 0.0;
    if ((rc = geti(cl, s->d.o_cl + a, n)) || (rc = geti(cs, s->d.o_cs + a, n)) ||
        (rc = geti(csi, s->d.o_csi + h, n + 1)) || (rc = getd(cek, s->d.o_cek + a, n)) ||
        (rc = getd(ceks, s->d.o_ceks + h, n + 1)))
        return rc;
    if (cc)
        for (int i = 0; i < n; ++i) cc[i] = 1;
    if (csi) csi[0] = csi0 + (n - c.npos);  // stoddard-pbc-n.f:145-146 with cs = 0 rows
    if (ceks) ceks[0] = ceks0;
    int *cnt = counters;
    if (cnt) cnt[0] = n;  // cnum: one number per visited atom
    if (fre) {
        if ((rc = geti(clf, s->d.o_clf + a, n)) || (rc = geti(ccf, s->d.o_ccf + a, n)) ||
            (rc = geti(csf, s->d.o_csf + h, n + 1)) || (rc = geti(csif, s->d.o_csif + h, n + 1)) ||
            (rc = getd(cekf, s->d.o_cekf + h, n + 1)) || (rc = getd(ceksf, s->d.o_ceksf + h, n + 1)) ||
            (rc = geti(cgs, s->d.o_cgs + a, n)) || (rc = geti(cgsf, s->d.o_cgsf + h, n + 1)) ||
            (rc = geti(cgsfa, s->d.o_cgsfa + h, n + 1)))
            return rc;
        if (csf) csf[0] = 0;
        if (cekf) cekf[0] = c.cekf0;
        if (ceksf) ceksf[0] = 0.0;
        if (cgsf) cgsf[0] = 0;
        if (cgsfa) cgsfa[0] = 0;
        if (csif) csif[0] = s->hp[traj].num[0] - c.nliq;  // stoddard-pbc-n-m-fre.f:251
        if (cnt) {
            cnt[1] = c.nliq;  // cnumf
            // first-passage bookkeeping, stoddard-pbc-n-m-fre.f:229-258
            if (c.cbig > cnt[2]) cnt[2] = c.cbig;
            if (cnt[2] >= cnt[4]) {
                if (cfpt)
                    for (int i = cnt[4]; i <= cnt[2]; ++i) cfpt[i - 1] = step;
                cnt[4] += 1;
            }
            if (c.cbigf > cnt[3]) cnt[3] = c.cbigf;
            if (cnt[3] >= cnt[5]) {
                if (cfptf)
                    for (int i = cnt[5]; i <= cnt[3]; ++i) cfptf[i - 1] = step;
                cnt[5] += 1;
            }
        }
    }
    return FCMD_OK;
}

int fcmd_halflist(fcmd_system *s, int traj, int *atnlist, long long cap, int *atnidx, long long *len)
{
    if (check_traj(s, traj) || !atnlist || !atnidx || !len) return FCMD_ERR_ARG;
    const TrajParams &t = s->tp[traj];
    HIPCHK(hipStreamSynchronize(s->stream));
    const int n = t.natom;
    std::vector<int> nn(t.npad), oid(t.npad);
    HIPCHK(hipMemcpy(nn.data(), s->d.nn + t.abase, sizeof(int) * t.npad, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(oid.data(), s->d.oid + t.abase, sizeof(int) * t.npad, hipMemcpyDeviceToHost));
    const long long lo = t.lbase, nslots = (long long)t.npad * t.lstride;
    std::vector<int> lst(nslots);
    if (s->g.list16) {
        std::vector<unsigned short> l16(nslots);
        HIPCHK(hipMemcpy(l16.data(), s->d.list16 + lo, sizeof(unsigned short) * nslots, hipMemcpyDeviceToHost));
        for (long long k = 0; k < nslots; ++k) lst[k] = l16[k];
    } else {
        HIPCHK(hipMemcpy(lst.data(), s->d.list32 + lo, sizeof(int) * nslots, hipMemcpyDeviceToHost));
    }
    // neighbours (reference numbers, j > i) of each reference atom i
    std::vector<std::vector<int>> nb(n);
    for (int r = 0; r < n; ++r) {
        const int oi = oid[r];
        const long long base = list_slot(t, r) - lo;
        for (int q = 0; q < nn[r]; ++q) {
            const int j = lst[base + (long long)q * TILE];  // trajectory-relative
            const int oj = oid[j];
            if (oj > oi) nb[oi - 1].push_back(oj);
        }
    }
    long long ind = 0;
    for (int i = 1; i <= n; ++i) {
        auto &v = nb[i - 1];
        std::sort(v.begin(), v.end());
        if (ind + 1 + (long long)v.size() > cap) return FCMD_ERR_LISTFULL;
        atnlist[ind] = -i;
        atnidx[i - 1] = (int)(ind + 1);
        ++ind;
        for (int j : v) atnlist[ind++] = j;
    }
    *len = ind;
    return FCMD_OK;
}

int fcmd_run(fcmd_system *s, int nsteps, int cluster_every, int fpt_size, int *series, int *fpt, int use_graph)
{
    if (!s || nsteps < 0 || cluster_every < 0) return FCMD_ERR_ARG;
    const int nsnap = cluster_every > 0 ? (nsteps + cluster_every - 1) / cluster_every : 0;
    std::vector<int> init(s->ntraj, -1);  // lives until the copy below has run
    if (nsnap > 0) {
        // series and first-passage buffers: kept by the system, grown here
        const long long need = (long long)nsnap * s->ntraj;
        if (need > s->series_cap) {
            if (s->dseries) (void)hipFree(s->dseries);
            s->dseries = nullptr;
            s->series_cap = 0;
            HIPCHK(hipMalloc(&s->dseries, sizeof(int) * (size_t)need));
            s->series_cap = need;
        }
        if (!s->dfpt) HIPCHK(hipMalloc(&s->dfpt, sizeof(int) * s->ntraj));
        if (fpt) std::copy(fpt, fpt + s->ntraj, init.begin());
        HIPCHK(hipMemcpyAsync(s->dfpt, init.data(), sizeof(int) * s->ntraj, hipMemcpyHostToDevice, s->stream));
        set_want(s, -1);
    }
    // the run graph: g steps (g = 10, or the analysis interval if shorter)
    // captured once and replayed; a graph made for a longer interval than
    // this run's analyses allow is recaptured
    const int gsteps = cluster_every > 0 ? std::min(cluster_every, 10) : 10;
    if (use_graph && (!s->graph || (cluster_every > 0 && s->graph_steps > cluster_every)))
        if (int rc = capture_graph(s, gsteps)) return rc;
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    int done = 0, snap = 0;
    s->an_count = 0;
    while (done < nsteps) {
        if (nsnap > 0 && done % cluster_every == 0) {
            const bool timed = snap < fcmd_system::NAN_TIMED;
            if (timed) HIPCHK(hipEventRecord(s->an_ev[2 * snap], s->stream));
            finalize(s);
            launch_cluster(s->d, s->g, s->any_fre, false, s->stream);
            hipLaunchKernelGGL(k_record, dim3((s->ntraj + 255) / 256), dim3(256), 0, s->stream, s->d.cls, s->d.ts,
                               s->ntraj, s->dseries + (size_t)snap * s->ntraj, s->dfpt, fpt_size);
            if (s->count_nmax > 0) launch_size_counts(s->d, s->g, s->count_nmax, s->count_warmup, s->stream);
            if (timed) HIPCHK(hipEventRecord(s->an_ev[2 * snap + 1], s->stream));
            ++snap;
        }
        const int todo = std::min(nsteps - done, nsnap > 0 ? cluster_every - done % cluster_every : nsteps - done);
        int k = 0;
        if (use_graph && s->graph)
            for (; k + s->graph_steps <= todo; k += s->graph_steps) HIPCHK(hipGraphLaunch(s->graph, s->stream));
        for (; k < todo; ++k)
            if (int rc = enqueue_step(s, done + k + 1 == nsteps)) return rc;
        HIPCHK(hipGetLastError());
        done += todo;
    }
    finalize(s);
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    s->an_ms = 0.0;
    s->an_count = std::min(snap, (int)fcmd_system::NAN_TIMED);
    for (int k = 0; k < s->an_count; ++k) {
        float a = 0.f;
        HIPCHK(hipEventElapsedTime(&a, s->an_ev[2 * k], s->an_ev[2 * k + 1]));
        s->an_ms += a;
    }
    if (nsnap > 0) {
        if (series)
            HIPCHK(hipMemcpy(series, s->dseries, sizeof(int) * (size_t)nsnap * s->ntraj, hipMemcpyDeviceToHost));
        if (fpt) HIPCHK(hipMemcpy(fpt, s->dfpt, sizeof(int) * s->ntraj, hipMemcpyDeviceToHost));
    }
    return sync_err(s);
}

int fcmd_run_noalloc(fcmd_system *s, int nsteps, int cluster_every, int fpt_size, int *series, int *fpt)
{
    if (!s || nsteps < 0 || cluster_every < 0) return FCMD_ERR_ARG;
    if (!s->scratch) return FCMD_ERR_HIP;
    const int nsnap = cluster_every > 0 ? (nsteps + cluster_every - 1) / cluster_every : 0;
    if ((long long)(nsnap + 1) * s->ntraj > FCMD_NOALLOC_INTS) return FCMD_ERR_SCRATCH;
    // scratch layout: fpt[ntraj], then series[nsnap][ntraj]; both written
    // before they are read (fpt from the caller or -1, series by each
    // analysis for every trajectory), whatever an earlier call left there
    int *dfpt = s->scratch, *dser = s->scratch + s->ntraj;
    if (nsnap > 0) {
        if (fpt)
            HIPCHK(hipMemcpyAsync(dfpt, fpt, sizeof(int) * s->ntraj, hipMemcpyHostToDevice, s->stream));
        else
            hipLaunchKernelGGL(k_fill, dim3((s->ntraj + 255) / 256), dim3(256), 0, s->stream, dfpt, s->ntraj, -1);
        set_want(s, -1);
    }
    int done = 0, snap = 0;
    while (done < nsteps) {
        if (nsnap > 0 && done % cluster_every == 0) {
            finalize(s);
            launch_cluster(s->d, s->g, s->any_fre, false, s->stream);
            hipLaunchKernelGGL(k_record, dim3((s->ntraj + 255) / 256), dim3(256), 0, s->stream, s->d.cls, s->d.ts,
                               s->ntraj, dser + (size_t)snap * s->ntraj, dfpt, fpt_size);
            if (s->count_nmax > 0) launch_size_counts(s->d, s->g, s->count_nmax, s->count_warmup, s->stream);
            ++snap;
        }
        const int todo = std::min(nsteps - done, nsnap > 0 ? cluster_every - done % cluster_every : nsteps - done);
        for (int k = 0; k < todo; ++k)
            if (int rc = enqueue_step(s, done + k + 1 == nsteps)) return rc;
        HIPCHK(hipGetLastError());
        done += todo;
    }
    finalize(s);
    if (nsnap > 0) {
        if (series)
            HIPCHK(hipMemcpyAsync(series, dser, sizeof(int) * (size_t)nsnap * s->ntraj, hipMemcpyDeviceToHost,
                                  s->stream));
        if (fpt) HIPCHK(hipMemcpyAsync(fpt, dfpt, sizeof(int) * s->ntraj, hipMemcpyDeviceToHost, s->stream));
    }
    return sync_err(s);  // also waits for the copies above
}

int fcmd_size_counts(fcmd_system *s, int nmax, int warmup)
{
    if (!s) return FCMD_ERR_ARG;
    finalize(s);
    HIPCHK(hipMemsetAsync(s->d.o_hacc, 0, sizeof(int) * s->nhist_total, s->stream));
    HIPCHK(hipMemsetAsync(s->d.hacc_stop, 0, sizeof(int) * s->ntraj, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->count_nmax = nmax > 0 ? nmax : 0;
    s->count_warmup = warmup;
    return FCMD_OK;
}

int fcmd_get_size_counts(fcmd_system *s, int traj, int *counts, int *stopped)
{
    if (check_traj(s, traj) || !counts) return FCMD_ERR_ARG;
    const TrajParams &t = s->tp[traj];
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(counts, s->d.o_hacc + t.hbase, sizeof(int) * (t.natom + 1), hipMemcpyDeviceToHost));
    counts[0] = 0;
    if (stopped) HIPCHK(hipMemcpy(stopped, s->d.hacc_stop + traj, sizeof(int), hipMemcpyDeviceToHost));
    return FCMD_OK;
}

int fcmd_last_run_ms(fcmd_system *s, double *total_ms)
{
    if (!s || !total_ms) return FCMD_ERR_ARG;
    *total_ms = s->last_ms;
    return FCMD_OK;
}

int fcmd_last_run_analysis_ms(fcmd_system *s, double *mean_ms, int *count)
{
    if (!s || !mean_ms) return FCMD_ERR_ARG;
    *mean_ms = s->an_count > 0 ? s->an_ms / s->an_count : 0.0;
    if (count) *count = s->an_count;
    return FCMD_OK;
}

// ---------------------------------------------------------------------------
// Tuning.  The variants differ only in how many loads are in flight and in
// launch widths: every one sums the same terms in the same order.

namespace {
const int FORCE_U[] = {2, 4, 8};
const int NBR_CU[] = {2, 4};
const int WL_GRIDS[] = {1024, 2048, 4096};

bool one_of(int v, const int *set, int n)
{
    for (int k = 0; k < n; ++k)
        if (set[k] == v) return true;
    return false;
}

// after the launch shapes change: a captured run graph is captured again
// with the new ones (same length), so that the next run replays at once
int recapture_graph(fcmd_system *s)
{
    if (!s->graph) return FCMD_OK;
    return capture_graph(s, s->graph_steps);
}
}  // namespace

int fcmd_set_tuning(fcmd_system *s, const int *tuning)
{
    if (!s || !tuning || !one_of(tuning[0], FORCE_U, 3) || !one_of(tuning[1], NBR_CU, 2) ||
        !one_of(tuning[2], WL_GRIDS, 3))
        return FCMD_ERR_ARG;
    s->g.force_u = tuning[0];
    s->g.nbr_cu = tuning[1];
    s->g.wl_grid = tuning[2];
    return recapture_graph(s);
}

int fcmd_get_tuning(fcmd_system *s, int *tuning)
{
    if (!s || !tuning) return FCMD_ERR_ARG;
    tuning[0] = s->g.force_u;
    tuning[1] = s->g.nbr_cu;
    tuning[2] = s->g.wl_grid;
    return FCMD_OK;
}

int fcmd_autotune(fcmd_system *s, int rounds, int *tuning)
{
    if (!s || rounds <= 0) return FCMD_ERR_ARG;
    finalize(s);
    const int nf = 3, nr = 2 * 3;
    // one step per candidate and round, the force and rebuild candidates
    // cycling independently so that each sees a like mix of rebuilds
    const int nsteps = rounds * nr;
    std::vector<hipEvent_t> ev((size_t)nsteps * 4);
    for (auto &e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    const LaunchDims keep = s->g;
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t *e = ev.data() + (size_t)k * 4;
        LaunchDims g = keep;
        g.force_u = FORCE_U[k % nf];
        g.nbr_cu = NBR_CU[(k % nr) / 3];
        g.wl_grid = WL_GRIDS[(k % nr) % 3];
        launch_drift(s->d, g, s->mixed, s->stream);
        HIPCHK(hipEventRecord(e[0], s->stream));
        launch_rebuild(s->d, g, s->mixed, s->stream, {}, true);
        HIPCHK(hipEventRecord(e[1], s->stream));
        // the run's kernels: the energy on the last step only
        const bool en = k == nsteps - 1;
        launch_force_atoms(s->d, g, s->mixed, s->single_species, FORCE_STEP, s->stream, {}, FP_ALL, en);
        HIPCHK(hipEventRecord(e[2], s->stream));
        launch_traj_post(s->d, g, FORCE_STEP, s->stream, {}, en);
        HIPCHK(hipEventRecord(e[3], s->stream));
    }
    HIPCHK(hipGetLastError());
    finalize(s);
    HIPCHK(hipStreamSynchronize(s->stream));
    double tf[3] = {0, 0, 0}, tr[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t *e = ev.data() + (size_t)k * 4;
        float a = 0.f, b = 0.f;
        HIPCHK(hipEventElapsedTime(&a, e[0], e[1]));
        HIPCHK(hipEventElapsedTime(&b, e[1], e[2]));
        tr[k % nr] += a;
        tf[k % nf] += b;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    int bf = 0, br = 0;
    for (int c = 1; c < nf; ++c)
        if (tf[c] < tf[bf]) bf = c;
    for (int c = 1; c < nr; ++c)
        if (tr[c] < tr[br]) br = c;
    s->g.force_u = FORCE_U[bf];
    s->g.nbr_cu = NBR_CU[br / 3];
    s->g.wl_grid = WL_GRIDS[br % 3];
    if (int rc = recapture_graph(s)) return rc;
    if (tuning) fcmd_get_tuning(s, tuning);
    return sync_err(s);
}

int fcmd_time_kernels(fcmd_system *s, int nsteps, double *avg_ms)
{
    if (!s || nsteps <= 0 || !avg_ms) return FCMD_ERR_ARG;
    const int nk = FCMD_NTIMED;  // k_drift, k_traj_pre, rebuild sequence, k_force, k_traj_post, kick
    // per step, the start and end of each of the first nk - 1 groups,
    // recorded by the launches themselves (no event packets between kernels,
    // so the step runs as in a run); the kick once at the end, bracketed
    const int ne = 2 * (nk - 1);
    std::vector<hipEvent_t> ev((size_t)nsteps * ne + 2);
    // timing-only events: no system-scope fence (cache writeback) at each
    // record, which would otherwise be charged to the timed kernel
    for (auto &e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t *e = ev.data() + (size_t)k * ne;
        const KTimer kt[5] = {{e[0], e[1]}, {e[2], e[3]}, {e[4], e[5]}, {e[6], e[7]}, {e[8], e[9]}};
        launch_drift_atoms(s->d, s->g, s->mixed, s->stream, kt[0]);
        launch_traj_pre(s->d, s->g, s->stream, kt[1]);
        launch_rebuild(s->d, s->g, s->mixed, s->stream, kt[2], true);
        // as a run's captured graph does: the energy every 10th step
        const bool en = (k + 1) % 10 == 0 || k == nsteps - 1;
        launch_force_atoms(s->d, s->g, s->mixed, s->single_species, FORCE_STEP, s->stream, kt[3], FP_ALL, en);
        launch_traj_post(s->d, s->g, FORCE_STEP, s->stream, kt[4], en);
    }
    // the run loop defers the NH update to the next drift; only the last step
    // of a run (or an analysis) pays it, so it is timed once
    hipEvent_t *ek = ev.data() + (size_t)nsteps * ne;
    HIPCHK(hipEventRecord(ek[0], s->stream));
    finalize(s);
    HIPCHK(hipEventRecord(ek[1], s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int c = 0; c < nk; ++c) avg_ms[c] = 0.0;
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t *e = ev.data() + (size_t)k * ne;
        for (int c = 0; c < nk - 1; ++c) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e[2 * c], e[2 * c + 1]));
            avg_ms[c] += ms / nsteps;
        }
    }
    float mk = 0.f;
    HIPCHK(hipEventElapsedTime(&mk, ek[0], ek[1]));
    avg_ms[nk - 1] = mk / nsteps;
    for (auto &e : ev) (void)hipEventDestroy(e);
    return sync_err(s);
}

int fcmd_list_entries(fcmd_system *s, long long *entries)
{
    if (!s || !entries) return FCMD_ERR_ARG;
    std::vector<TrajScalars> ts(s->ntraj);
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(ts.data(), s->d.ts, sizeof(TrajScalars) * s->ntraj, hipMemcpyDeviceToHost));
    long long n = 0;
    for (auto &t : ts) n += 2LL * t.npairs;
    *entries = n;
    return FCMD_OK;
}

}  // extern "C"
