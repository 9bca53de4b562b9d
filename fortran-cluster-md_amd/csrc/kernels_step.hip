// kernels_step.hip -- one MD step: drift, LJ forces, velocity update.
//
// Reference routines followed (cluster/):
//   drift      int-verlet.f:34-41, int-verlet-mav.f:44-56, int-verlet-nh-mav.f:49-61
//   mav/skin   int-verlet-mav.f:64-74, neighbor-pbc-mav.f:48-63, neighbor-pbc.f:40-59
//   forces     for-a-lj-n-pbc.f:54-127, for-a-lj-n-pbc-m.f:61-152
//   kick       int-verlet.f:53-59, int-verlet-mav.f:86-92, int-verlet-nh-mav.f:91-180
//
// Element-wise updates use the reference's operand order and are compiled
// without FMA contraction, so each per-atom update rounds as the reference
// does; only sums (forces over neighbours, energies) are re-associated.
#include "fcmd_common.hpp"
#include "fcmd_kernels.hpp"

namespace fcmd {

__device__ __forceinline__ bool nh_coupled(const TrajParams &tp, int typ)
{
    return (tp.nh_mask >> typ) & 1;
}

// Cell grid for a list radius: side >= atrskin*(1+1e-9), so every pair within
// atrskin lies in neighbouring cells even after rounding of x/side; at most
// CELL_MAX cells per side (wider cells only add candidates).
__device__ void set_cell_grid(const TrajParams &tp, TrajScalars &ts)
{
    int n[3];
    for (int k = 0; k < 3; ++k) {
        n[k] = (int)floor(tp.box[k] / (ts.atrskin * (1.0 + 1e-9)));
        n[k] = n[k] < 1 ? 1 : (n[k] > CELL_MAX ? CELL_MAX : n[k]);
        ts.cside[k] = tp.box[k] / (double)n[k];
        ts.icside[k] = 1.0 / ts.cside[k];
        ts.ncf[k] = (float)n[k];
        // 1/n to 1 ulp: dc/n only has to round to the right integer
        ts.incf[k] = __builtin_amdgcn_rcpf(ts.ncf[k]);
        ts.csf[k] = (float)ts.cside[k];
    }
    if ((long long)n[0] * n[1] * n[2] >= (long long)tp.ncpad) {  // one cell past the grid is kept
        ts.err = ERR_CELLS;
        return;
    }
    ts.ncx = n[0];
    ts.ncy = n[1];
    ts.ncz = n[2];
}

// Scalar part of neighbor(step): skin checks and adaptation, then flag the
// list rebuild.  neighbor-pbc-mav.f:48-63 / neighbor-pbc.f:40-59, :96-100.
__device__ void neighbor_scalars(const TrajParams &tp, TrajScalars &ts, int *step)
{
    const double rc = tp.atrcut;
    if (tp.neighbor == NB_PBC) {
        if (*step < 0) *step = -*step;
        ts.dmmav = ts.dmmsu * (double)tp.nbtim / (double)(*step - tp.btim);
        if (ts.dmms > (ts.atrskin - rc)) {
            ts.err = ERR_SKIN;
            return;
        }
        if (ts.dmmst == tp.nbtim && ts.dmms < ts.dmmav && ts.dmmav < 0.95 * (ts.atrskin - rc))
            ts.atrskin = ts.atrskin - 0.01 * (ts.atrskin - rc);
    } else {
        if (ts.dmms > (ts.atrskin - rc)) {
            ts.err = ERR_SKIN;
            return;
        }
        ts.dmmav = (double)(float)ts.dmmst * ts.dmmsu / 256.0;
        if (*step != tp.btim && *step > 0) {
            ts.atrskin = rc + (ts.atrskin - rc) * (double)tp.nbtim / (double)ts.dmmst;
            ts.atrskin = fmin(tp.box[0] / 2.0, ts.atrskin);
            ts.atrskin = fmin(tp.box[1] / 2.0, ts.atrskin);
            ts.atrskin = fmin(tp.box[2] / 2.0, ts.atrskin);
        }
        if (*step < 0) *step = -*step;
    }
    ts.rs2 = ts.atrskin * ts.atrskin;
    set_cell_grid(tp, ts);
    if (ts.err) return;
    ts.dmmax = fmax(ts.dmms, ts.dmmax);
    ts.dmmsr = ts.dmms;
    ts.dmms = 0.0;
    ts.dmmst = 0;
    ts.npairs = 0;
    ts.nlist = 0;
    ts.rebuild = 1;
    ts.nrebuild += 1;
    // the first rebuild and every resort-th after it re-sort the atoms by
    // cell (locality of the gathers); the others keep them in their slots
    ts.move = (ts.nrebuild - 1) % tp.resort == 0;
}

// Fixed-order reduction of one trajectory's per-block partial slots
// slot[0..N-1] by one block (deterministic: thread k takes blocks k, k+BLK,
// ...; then each value's wavefront reduction and (w0 + w1) + (w2 + w3), or
// the max, with one barrier for all of them).  Results in thread 0.
template <int N>
__device__ void reduce_partials(const TrajParams &tp, const double *part, const int (&slot)[N],
                                const bool (&is_max)[N], double (*lds)[4], double (&r)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k) r[k] = 0.0;
    for (int q = threadIdx.x; q < tp.nblk; q += BLK) {
        const double *p = part + (size_t)(tp.bfirst + q) * NPART;
#pragma unroll
        for (int k = 0; k < N; ++k) r[k] = is_max[k] ? fmax(r[k], p[slot[k]]) : r[k] + p[slot[k]];
    }
    block_reduce_end<N>(r, is_max, lds);
}

// After the drift: thermostat half update, displacement bookkeeping and the
// rebuild decision for one trajectory (runs on one thread).
__device__ void traj_pre(const TrajParams &tp, TrajScalars &ts, double dmmr, double sumv2)
{
    ts.rebuild = 0;
    ts.pending = 0;  // the drift has applied it
    if (tp.integrator == INT_NH) {
        ts.sumv2 = sumv2;
        if (tp.ethst_nh) {
            const double dt = tp.dt, dt2 = dt * dt / 2, dth = dt / 2;
            const double m1 = tp.mas[0];
            ts.s = ts.s + ts.psi * dt + (m1 * sumv2 - tp.nkt) * dt2 / tp.Q;
            ts.psi = ts.psi + (m1 * sumv2 - tp.nkt) * dth / tp.Q;
        }
    }
    if (tp.integrator != INT_VV) {
        dmmr = 2.2 * sqrt(dmmr);
        ts.dmms = ts.dmms + dmmr;
        ts.dmmsu = ts.dmmsu + dmmr - ts.dmmh[ts.dmmi];
        ts.dmmh[ts.dmmi] = dmmr;
        ts.dmmi = (ts.dmmi + 1) & 255;
        ts.dmmst = ts.dmmst + 1;
        if (ts.dmms > (ts.atrskin - tp.atrcut - 2.0 * ts.dmms / (double)(float)ts.dmmst)) {
            int step = ts.step;
            neighbor_scalars(tp, ts, &step);
        }
    }
}

template <typename AccT>
__device__ __forceinline__ void load_acc(const DevPtrs &d, int i, double &ax, double &ay, double &az);
template <>
__device__ __forceinline__ void load_acc<double>(const DevPtrs &d, int i, double &ax, double &ay, double &az)
{
    ax = d.ax[i];
    ay = d.ay[i];
    az = d.az[i];
}
template <>
__device__ __forceinline__ void load_acc<float>(const DevPtrs &d, int i, double &ax, double &ay, double &az)
{
    ax = (double)d.axf[i];
    ay = (double)d.ayf[i];
    az = (double)d.azf[i];
}

template <typename AccT>
__device__ __forceinline__ void store_acc(const DevPtrs &d, int i, double ax, double ay, double az);
template <>
__device__ __forceinline__ void store_acc<double>(const DevPtrs &d, int i, double ax, double ay, double az)
{
    d.ax[i] = ax;
    d.ay[i] = ay;
    d.az[i] = az;
}
template <>
__device__ __forceinline__ void store_acc<float>(const DevPtrs &d, int i, double ax, double ay, double az)
{
    d.axf[i] = (float)ax;
    d.ayf[i] = (float)ay;
    d.azf[i] = (float)az;
}

// ---------------------------------------------------------------------------
template <typename AccT>
__global__ __launch_bounds__(BLK) void k_drift(DevPtrs d)
{
    __shared__ double lds[2][4];
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    TrajScalars &ts = d.ts[t];
    if (ts.err) return;
    const int i = b * BLK + threadIdx.x;
    const bool act = (i - tp.abase) < tp.natom;
    double d2 = 0.0, sv = 0.0;
    if (act) {
        const double dt = tp.dt;
        double vx = __builtin_nontemporal_load(d.vx + i), vy = __builtin_nontemporal_load(d.vy + i),
               vz = __builtin_nontemporal_load(d.vz + i);
        double ax, ay, az;
        if (sizeof(AccT) == 4) {
            ax = __builtin_nontemporal_load(d.axf + i);
            ay = __builtin_nontemporal_load(d.ayf + i);
            az = __builtin_nontemporal_load(d.azf + i);
        } else {
            ax = __builtin_nontemporal_load(d.ax + i);
            ay = __builtin_nontemporal_load(d.ay + i);
            az = __builtin_nontemporal_load(d.az + i);
        }
        double dx, dy, dz;
        if (tp.integrator == INT_NH) {
            const double dt2 = dt * dt / 2, dth = dt / 2;
            const int ty = tp.nsorts > 1 ? d.typ[i] : 0;  // one species: no load
            const bool cpl = nh_coupled(tp, ty);
            const double psi = cpl ? ts.psi : 0.0;
            if (ts.pending) {
                // the previous step's final update, deferred to here (same
                // expression as k_kick_nh, so the same bits)
                const double den = cpl ? 1.0 + dth * ts.psi : 1.0;
                vx = (vx + dth * ax) / den;
                vy = (vy + dth * ay) / den;
                vz = (vz + dth * az) / den;
            }
            dx = dt * vx + dt2 * (ax - psi * vx);
            dy = dt * vy + dt2 * (ay - psi * vy);
            dz = dt * vz + dt2 * (az - psi * vz);
            if (cpl) {
                // reference: m1*sum(v^2) with m1 for every atom (NH is single species)
                sv = vx * vx + vy * vy + vz * vz;
                if (!tp.nh_refmass) sv *= tp.mrel[ty];
            }
            vx = vx + dth * (ax - psi * vx);
            vy = vy + dth * (ay - psi * vy);
            vz = vz + dth * (az - psi * vz);
        } else {
            dx = (vx + 0.5 * ax * dt) * dt;
            dy = (vy + 0.5 * ay * dt) * dt;
            dz = (vz + 0.5 * az * dt) * dt;
            vx = vx + 0.5 * ax * dt;
            vy = vy + 0.5 * ay * dt;
            vz = vz + 0.5 * az * dt;
        }
        const double xn = __builtin_nontemporal_load(d.x + i) + dx, yn = __builtin_nontemporal_load(d.y + i) + dy,
                     zn = __builtin_nontemporal_load(d.z + i) + dz;
        __builtin_nontemporal_store(xn, d.x + i);
        __builtin_nontemporal_store(yn, d.y + i);
        __builtin_nontemporal_store(zn, d.z + i);
        if (sizeof(AccT) == 4) {
            int cx, cy, cz;
            unpack_cell(__int_as_float(__builtin_nontemporal_load(d.cellid + i)), cx, cy, cz);
            d.posf[i] = make_posf(xn, yn, zn, cx, cy, cz, ts);
        }
        // cached: the force kernel reads v next (see the force kernel's streams)
        d.vx[i] = vx;
        d.vy[i] = vy;
        d.vz[i] = vz;
        d2 = dx * dx + dy * dy + dz * dz;
    }
    double r[2] = {d2, sv};
    block_reduce_end<2>(r, {true, false}, lds);
    if (threadIdx.x == 0) {
        d.part[(size_t)b * NPART + P_DMAX] = r[0];
        d.part[(size_t)b * NPART + P_SUMV2] = r[1];
    }
}

// one block per trajectory, after the drift
// lists (a single trajectory): the block also writes the rebuild's work
// lists, which k_wl_fill would otherwise make in a launch of its own (every
// block of the trajectory, every cell chunk, the trajectory itself; empty
// when it keeps its list)
__global__ __launch_bounds__(BLK) void k_traj_pre(DevPtrs d, int lists)
{
    __shared__ double lds[2][4];
    const int t = blockIdx.x;
    const TrajParams &tp = d.tp[t];
    TrajScalars &ts = d.ts[t];
    const bool err = ts.err;
    if (!err) {
        double r[2];
        reduce_partials<2>(tp, d.part, {P_SUMV2, P_DMAX}, {false, true}, lds, r);
        if (threadIdx.x == 0) traj_pre(tp, ts, r[1], r[0]);
    }
    if (!lists) return;
    __syncthreads();  // thread 0's rebuild decision
    const bool rb = ts.rebuild && !ts.err;
    const int nb = rb ? tp.nblk : 0, nc = rb ? tp.ncblk : 0;
    for (int q = threadIdx.x; q < nb; q += BLK) d.wl_blk[q] = tp.bfirst + q;
    for (int q = threadIdx.x; q < nc; q += BLK) d.wl_cchk[q] = tp.cfirst + q;
    if (threadIdx.x == 0) {
        d.wl_traj[0] = t;
        d.wl_cnt[WL_BLK] = nb;
        d.wl_cnt[WL_CCHK] = nc;
        d.wl_cnt[WL_TRAJ] = rb ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Nose-Hoover implicit update, solved on one thread per trajectory.
//
// int-verlet-nh-mav.f:106-147 solves  v' = v_h + dth (a - psi' v'),
// psi' = psi_h + dth (m1 sum v'^2 - 3NkT)/Q  for all 3N+1 unknowns by Newton.
// Every solution has v' = (v_h + dth a)/(1 + dth psi'), so with
// S = sum (v_h + dth a)^2 the system is the scalar equation
//   g(psi) = psi_h + dth (m1 S/(1 + dth psi)^2 - 3NkT)/Q - psi = 0,
// solved here by Newton to machine precision; the reference iteration stops
// at a relative velocity change below 1e-10, i.e. (quadratic convergence)
// at the same root.
__device__ void nh_solve(const TrajParams &tp, TrajScalars &ts, double S)
{
    const double dth = tp.dt / 2;
    const double A = tp.mas[0] * S;
    const double psih = ts.psi;
    // g is a difference of terms of this size; its rounding noise (and so
    // the smallest meaningful Newton step) is a few ulps of it
    const double gscale = fabs(psih) + dth * (A + tp.nkt) / tp.Q;
    double psi = psih;
    int it = 0;
    for (it = 1; it <= 100; ++it) {
        const double den = 1.0 + dth * psi;
        if (!(den > 0.0)) {
            ts.err = ERR_NH;
            return;
        }
        const double w = 1.0 / den;
        const double g = psih + dth * (A * w * w - tp.nkt) / tp.Q - psi;
        const double dg = -2.0 * dth * dth * A * w * w * w / tp.Q - 1.0;
        const double step = -g / dg;
        psi += step;
        if (fabs(step) <= 8e-16 * (fabs(psi) + gscale)) break;
    }
    if (it > 100) {
        ts.err = ERR_NH;
        return;
    }
    const double w = 1.0 / (1.0 + dth * psi);
    ts.psi = psi;
    ts.sumv2 = S * w * w;
    ts.ekin = 0.5 * tp.mas[0] * ts.sumv2;
    ts.nh_iters = it;
}

// Finish of AccAtom/the integrator for one trajectory (one thread).  esum:
// per-atom pair energies over the full list; eks: VV kinetic energy, or the
// Nose-Hoover S; ekfree: kinetic energy of atoms not coupled to the
// thermostat (species-selective coupling only); wsum: per-atom pair virial
// terms over the full list.
__device__ void traj_post(const TrajParams &tp, TrajScalars &ts, double esum, double eks, double ekfree,
                          double wsum, int mode, bool energy)
{
    if (energy) {
        // every pair is seen from both atoms of the full list
        esum = 0.5 * esum;
        double w = 0.5 * wsum;
        if (tp.accatom == ACC_1) {
            ts.epot = esum * tp.eps1 * 4.0;
            w = w * 48.0 * tp.eps1;  // for-a-lj-n-pbc.f:114
        } else {
            ts.epot = esum;
        }
        // for-a-lj-n-pbc.f:116-117: 1e13/V (N k T - W/3), T the target temperature
        ts.virial = w;
        ts.pressure =
            (1.0e13 / (tp.box[0] * tp.box[1] * tp.box[2])) * ((double)tp.natom * CIKB * tp.att - w / 3.0);
    }
    if (mode == FORCE_ACCATOM) return;
    if (tp.integrator == INT_NH) {
        nh_solve(tp, ts, eks);
        ts.ekin += ekfree;
        ts.pending = 1;  // v' = (v_h + dth a)/(1 + dth psi') is applied later
    } else {
        ts.ekin = eks;
    }
    ts.step += 1;
}

// one block per trajectory, after the forces
__global__ __launch_bounds__(BLK) void k_traj_post(DevPtrs d, int mode, int energy)
{
    __shared__ double lds[4][4];
    const int t = blockIdx.x;
    const TrajParams &tp = d.tp[t];
    TrajScalars &ts = d.ts[t];
    if (ts.err) return;
    double r[4];
    reduce_partials<4>(tp, d.part, {P_EPOT, P_EKS, P_SUMV2, P_VIR}, {false, false, false, false}, lds, r);
    if (threadIdx.x == 0) traj_post(tp, ts, r[0], r[1], r[2], r[3], mode, energy != 0);
}

// ---------------------------------------------------------------------------
// Pair sum of one atom over its full-list neighbours.  FORM = ACC_1:
// for-a-lj-n-pbc.f:95-105 (sum of r^-2 r^-6 (c2 r^-6 - c3) d, scaled by
// 48 eps / m afterwards); FORM = ACC_M: for-a-lj-n-pbc-m.f:115-138 (per-pair
// 48 eps_ab r^-2 r^-6 (c2_ab r^-6 - c3_ab) d).  Neighbours are taken U at a
// time, all loads of a chunk issued before its arithmetic; pairs are summed
// in list order.
//
// FP64: the reference's arithmetic, minimum image in its form, so every pair
// term and cutoff decision is the reference's, bit for bit.
// list slots per atom staged in LDS by the force kernel
constexpr int STAGE = 16;

// Streaming data of the force kernel (the atom's own count and list slots) is
// read once per step: nontemporal loads keep it from displacing the
// neighbours' records, which the same and nearby wavefronts gather again from
// L2.  The velocities (which the drift has just written) and the
// accelerations (which the next drift reads) go through the caches: measured
// on the ensemble, that made the drift ~7 us faster for ~1 us on this kernel.

// list slots loaded together with the count, before it is known
constexpr int SU = 8;

// An atom's list (entries of type LT, trajectory-relative): its count, its
// first slot and its first SU entries, all loaded at once (every atom owns
// lstride >= LIST_ROUND slots, whatever its count), before anything waits on
// memory.
// Entries travel as the loaded dwords ("words") and are decoded only once
// every load of a batch has been issued: a load under a per-lane condition
// whose result were used at once would make the compiler wait for it inside
// the condition, one load at a time.
template <typename LT>
struct ListHead {
    int nn, base;
    const unsigned *lp;  // the dword holding the atom's entry of list position 0
    int sh;              // the entry's bit offset in its dword (16-bit entries)
    unsigned fill;       // a word that decodes to the atom itself
    unsigned head[SU];   // words
    __device__ __forceinline__ ListHead(const DevPtrs &d, const TrajParams &tp, int i, int rel)
    {
        nn = __builtin_nontemporal_load(d.nn + i);
        base = tp.abase;
        const LT *p = list_base<LT>(d) + list_slot(tp, rel);
        // 16-bit entries are read as the aligned dword that holds them (an
        // atom pair shares it): whole-dword loads, not sub-dword ones, which
        // are slow here.  Slots are even-aligned: lbase and the tile offsets
        // are even, and rel & 1 is the lane's parity.
        sh = sizeof(LT) == 2 ? 16 * (rel & 1) : 0;
        lp = (const unsigned *)(sizeof(LT) == 2 ? p - (rel & 1) : p);
        fill = (unsigned)rel << sh;
#pragma unroll
        for (int s = 0; s < SU; ++s) head[s] = word(s);
    }
    // the word of list position s, from HBM
    __device__ __forceinline__ unsigned word(int s) const
    {
        constexpr int STRIDE = TILE * (int)sizeof(LT) / 4;  // dwords per list position
        return __builtin_nontemporal_load(lp + (size_t)s * STRIDE);
    }
    // the trajectory-relative neighbour slot a word holds (below 65536 for
    // 16-bit entries, so that a gather's byte offset from the trajectory's
    // first slot is a 32-bit one)
    __device__ __forceinline__ unsigned decode_rel(unsigned w) const
    {
        return sizeof(LT) == 2 ? (w >> sh) & 0xffffu : w;
    }
    // the neighbour slot a word holds
    __device__ __forceinline__ int decode(unsigned w) const { return base + (int)decode_rel(w); }
};

// Walks an atom's nn list slots in batches of U (U <= SU): load(w) issues the
// gathers of a batch's words w (the fill word, the atom itself, past nn),
// sum(s0) consumes them.  Slots below STAGE: the head slots come with the
// count, so the first batch's gathers wait on one round trip instead of two;
// the other STAGE - SU are loaded meanwhile and kept in LDS.  Slots from
// STAGE on (long lists): each batch's words are loaded while the batch
// before it is summed, so a batch waits on its gathers only.  No word load
// needs a bound check: an atom owns lstride >= nn slots, a multiple of
// LIST_ROUND.
template <int U, typename H, typename Load, typename Sum>
__device__ __forceinline__ void walk_list(unsigned (*stage)[BLK], const H &h, Load load, Sum sum)
{
    static_assert(U <= SU && SU <= STAGE && STAGE % U == 0 && LIST_ROUND % STAGE == 0,
                  "batches must tile the staged slots and the allocated ones");
    const int nn = h.nn;
#pragma unroll
    for (int s = 0; s < SU; ++s) stage[s][threadIdx.x] = s < nn ? h.head[s] : h.fill;
    // the other staged words, where any lane of the wavefront has that many
    // neighbours (a wave-uniform branch: the tile row a load reads is fetched
    // whole anyway, and a row no lane needs is not fetched at all)
    unsigned tail[STAGE - SU];
#pragma unroll
    for (int s = SU; s < STAGE; ++s) {
        tail[s - SU] = h.fill;
        if (__builtin_amdgcn_ballot_w64(s < nn)) tail[s - SU] = h.word(s);
    }
    unsigned w[U];
    if (nn > 0) {
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = u < nn ? h.head[u] : h.fill;
        load(w);
#pragma unroll
        for (int s = SU; s < STAGE; ++s) stage[s][threadIdx.x] = s < nn ? tail[s - SU] : h.fill;
        sum(0);
    }
    for (int s0 = U; s0 < STAGE && s0 < nn; s0 += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = stage[s0 + u][threadIdx.x];
        load(w);
        sum(s0);
    }
    if (nn > STAGE) {
        unsigned wn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) wn[u] = h.word(STAGE + u);
        for (int s0 = STAGE; s0 < nn; s0 += U) {
#pragma unroll
            for (int u = 0; u < U; ++u) w[u] = s0 + u < nn ? wn[u] : h.fill;
            load(w);
            if (s0 + U < nn) {
#pragma unroll
                for (int u = 0; u < U; ++u) wn[u] = h.word(s0 + U + u);
            }
            sum(s0);
        }
    }
}

template <int FORM, bool EN, int U, typename H>
__device__ __forceinline__ void pair_sum_fp64(const DevPtrs &d, const TrajParams &tp, const TrajScalars &ts,
                                              int i, const H &h, int ti, double &fx, double &fy, double &fz,
                                              double &e, double &w, unsigned (*stage)[BLK])
{
    const double xi = d.x[i], yi = d.y[i], zi = d.z[i];
    const double bx = tp.box[0], by = tp.box[1], bz = tp.box[2];
    const double ibx = tp.ibox[0], iby = tp.ibox[1], ibz = tp.ibox[2];
    const double bf1 = 2.0 + fint(ts.atrskin / tp.atr);
    const double bf2 = bf1 + 0.5;
    const double rc2 = tp.atrcut * tp.atrcut;
    const int nn = h.nn;
    fx = fy = fz = 0.0;
    int jj[U], tj[U];
    double xj[U], yj[U], zj[U];
    auto load = [&](const unsigned (&wd)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = h.decode(wd[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xj[u] = d.x[jj[u]];
            yj[u] = d.y[jj[u]];
            zj[u] = d.z[jj[u]];
            tj[u] = (FORM == ACC_1) ? 0 : d.typ[jj[u]];
        }
    };
    auto sum = [&](int s0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (s0 + u >= nn) break;
            double dx = xi - xj[u];
            double dy = yi - yj[u];
            double dz = zi - zj[u];
            dx = dx + bx * (bf1 - fint(bf2 + dx * ibx));
            dy = dy + by * (bf1 - fint(bf2 + dy * iby));
            dz = dz + bz * (bf1 - fint(bf2 + dz * ibz));
            double rij2 = dx * dx + dy * dy + dz * dz;
            if (rij2 <= rc2) {
                rij2 = 1.0 / rij2;
                const double rij6 = rij2 * rij2 * rij2;
                double fij;
                if (FORM == ACC_1) {
                    // q: the bracket the force and the virial term
                    // (for-a-lj-n-pbc.f:106) share
                    const double q = tp.c2[0] * rij6 - tp.c3[0];
                    fij = rij2 * rij6 * q;
                    if (EN) {
                        e += rij6 * (tp.c2[0] * rij6 - tp.c1[0]);
                        w += rij6 * q;
                    }
                } else {
                    const int pij = ti * MAXS + tj[u];
                    const double q = tp.c2[pij] * rij6 - tp.c3[pij];
                    fij = tp.e48[pij] * rij2 * rij6 * q;
                    if (EN) {
                        e += tp.e4[pij] * rij6 * (tp.c2[pij] * rij6 - tp.c1[pij]);
                        w += tp.e48[pij] * rij6 * q;
                    }
                }
                fx = fx + fij * dx;
                fy = fy + fij * dy;
                fz = fz + fij * dz;
            }
        }
    };
    walk_list<U>(stage, h, load, sum);
}

// Mixed: separations from the FP32 cell-offset records (make_posf: one 16-byte
// gather per neighbour, no FP64 except the cutoff test of the rare pair within
// 1e-5 of the cutoff, which is the reference's): the cell difference wrapped to a neighbour
// (sep_wrap; every listed pair was in neighbouring cells at the list build),
// or for a box of fewer than 3 cells on some side (SMALL, a loop of its own)
// the rounding image.  Pair distance and LJ terms in FP32 with fused
// multiply-adds (1/r^2 by the hardware reciprocal, 1 ulp); each pair's FP32
// force and energy terms are summed in FP64 (an FP64 fused multiply-add costs
// what an FP32 one does here, and fewer instructions than a compensated FP32
// sum).  The single-species form works in sigma units, so that its constants
// are instruction immediates.
// The reference's cutoff test of one pair on the FP64 positions
// (for-a-lj-n-pbc.f:89-95): its minimum image d + L (bf1 - int(bf2 + d/L)),
// r2 <= atrcut^2.  For the rare pair near the cutoff on the mixed path.
__device__ __forceinline__ bool reference_cut_test(const DevPtrs &d, const TrajParams &tp,
                                                   const TrajScalars &ts, int i, int j, double rc2)
{
    const double bf1 = 2.0 + fint(ts.atrskin / tp.atr);
    const double bf2 = bf1 + 0.5;
    double dx = d.x[i] - d.x[j];
    double dy = d.y[i] - d.y[j];
    double dz = d.z[i] - d.z[j];
    dx = dx + tp.box[0] * (bf1 - fint(bf2 + dx * tp.ibox[0]));
    dy = dy + tp.box[1] * (bf1 - fint(bf2 + dy * tp.ibox[1]));
    dz = dz + tp.box[2] * (bf1 - fint(bf2 + dz * tp.ibox[2]));
    return dx * dx + dy * dy + dz * dz <= rc2;
}

// Kahan-compensated FP32 addition: s + x with the rounding error of earlier
// additions carried in c (Kahan 1965)
__device__ __forceinline__ void kahan_add(float &s, float &c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// KAHAN: each pair's FP32 force components are accumulated per atom in FP32
// with Kahan compensation (FCMD_MIXED_KAHAN) instead of in FP64.
template <int FORM, bool EN, bool SMALL, bool KAHAN, int U, typename H>
__device__ __forceinline__ void pair_sum_mixed(const DevPtrs &d, const TrajParams &tp, const TrajScalars &ts,
                                               int i, const float4 &pi, const H &h, int ti, double &fx,
                                               double &fy, double &fz, double &e, double &w,
                                               unsigned (*stage)[BLK])
{
    const float cxi = cell_xf(pi.w), cyi = cell_yf(pi.w), czi = cell_zf(pi.w);
    // FP32 grid and box constants, converted where they are set so that they
    // are scalar operands here (cells a side, 1/cells to 1 ulp, cell sides,
    // box sides and inverses)
    const float ncx = ts.ncf[0], ncy = ts.ncf[1], ncz = ts.ncf[2];
    const float icx = ts.incf[0], icy = ts.incf[1], icz = ts.incf[2];
    const float sx = ts.csf[0], sy = ts.csf[1], sz = ts.csf[2];
    const float bx = tp.boxf[0], by = tp.boxf[1], bz = tp.boxf[2];
    const float ibx = tp.iboxf[0], iby = tp.iboxf[1], ibz = tp.iboxf[2];
    // cutoff decision band: r2 from the cell-offset records is within
    // ~2e-6 relative of the reference's FP64 r2, so outside rc^2 (1 -+ 1e-5)
    // the FP32 verdict is the reference's; inside, the reference's FP64 test
    // decides (the truncated force jumps at rc, so one pair decided the other
    // way would be the largest error of a large system)
    const double rc2 = tp.atrcut * tp.atrcut;
    const float in2 = tp.cut_in2f, out2 = tp.cut_out2f;
    const float sig2 = tp.sig2f;  // sigma^2 of the single-species form
    const int nn = h.nn;
    double sfx = 0.0, sfy = 0.0, sfz = 0.0;
    float kfx = 0.0f, kfy = 0.0f, kfz = 0.0f, kcx = 0.0f, kcy = 0.0f, kcz = 0.0f;
    // this atom's energy and virial terms, summed in FP32 (its own few
    // pairs) and reduced over atoms in FP64 by the caller
    float ef = 0.0f, wf = 0.0f;
    // the FP32 separation of i from the neighbour record q (its cell
    // difference wrapped to a neighbour) and its square
    auto separation = [&](const float4 &q, float &dxf, float &dyf, float &dzf) {
        const float cxj = cell_xf(q.w), cyj = cell_yf(q.w), czj = cell_zf(q.w);
        if (SMALL) {
            dxf = ncx < 3.0f ? sep_small(cxi, cxj, sx, pi.x - q.x, bx, ibx)
                             : sep_wrap(cxi, cxj, ncx, icx, sx, pi.x - q.x);
            dyf = ncy < 3.0f ? sep_small(cyi, cyj, sy, pi.y - q.y, by, iby)
                             : sep_wrap(cyi, cyj, ncy, icy, sy, pi.y - q.y);
            dzf = ncz < 3.0f ? sep_small(czi, czj, sz, pi.z - q.z, bz, ibz)
                             : sep_wrap(czi, czj, ncz, icz, sz, pi.z - q.z);
        } else {
            dxf = sep_wrap(cxi, cxj, ncx, icx, sx, pi.x - q.x);
            dyf = sep_wrap(cyi, cyj, ncy, icy, sy, pi.y - q.y);
            dzf = sep_wrap(czi, czj, ncz, icz, sz, pi.z - q.z);
        }
        return __builtin_fmaf(dxf, dxf, __builtin_fmaf(dyf, dyf, dzf * dzf));
    };
    // one pair's force, energy and virial terms added to the sums (sgn = 1),
    // or taken out again (sgn = -1)
    auto pair_terms = [&](float dxf, float dyf, float dzf, float r2, int tjj, float sgn) {
        const float ri2 = __builtin_amdgcn_rcpf(r2);
        float fij;
        if (FORM == ACC_1) {
            // in sigma units, s6 = (sigma/r)^6: r^-2 r^-6 (sigma^12 r^-6 -
            // sigma^6/2) = r^-2 s6 (s6 - 1/2), the energy term r^-6
            // (sigma^12 r^-6 - sigma^6) = s6 (s6 - 1) and the virial term
            // r^-6 (sigma^12 r^-6 - sigma^6/2) = s6 (s6 - 1/2); the constants
            // are immediates
            const float u2 = sig2 * ri2;
            const float s6 = u2 * u2 * u2;
            const float t = s6 - 0.5f;
            fij = ri2 * s6 * t;
            if (EN) {
                ef = __builtin_fmaf(sgn * s6, s6 - 1.0f, ef);
                wf = __builtin_fmaf(sgn * s6, t, wf);
            }
        } else {
            const float r6 = ri2 * ri2 * ri2;
            const int pij = ti * MAXS + tjj;
            const float q = __builtin_fmaf(tp.c2f[pij], r6, -tp.c3f[pij]);
            fij = (float)tp.e48[pij] * ri2 * r6 * q;
            if (EN) {
                ef = __builtin_fmaf(sgn * (float)tp.e4[pij], r6 * __builtin_fmaf(tp.c2f[pij], r6, -tp.c1f[pij]),
                                    ef);
                wf = __builtin_fmaf(sgn * fij, r2, wf);
            }
        }
        if constexpr (KAHAN) {
            const float fs = sgn * fij;
            kahan_add(kfx, kcx, dxf * fs);
            kahan_add(kfy, kcy, dyf * fs);
            kahan_add(kfz, kcz, dzf * fs);
        } else {
            const double fd = (double)(sgn * fij);
            sfx = __builtin_fma((double)dxf, fd, sfx);
            sfy = __builtin_fma((double)dyf, fd, sfy);
            sfz = __builtin_fma((double)dzf, fd, sfz);
        }
    };
    int tj[U];
    float4 pj[U];
    // the largest r2 taken, as its bit pattern (r2 >= 0, so integer order is
    // float order, and the maximum is one integer instruction): above in2, a
    // pair in the band was
    int r2max = 0;
    // the trajectory's records and species: the gathers' offsets are
    // trajectory-relative
    const float4 *pt = d.posf + h.base;
    const int *tt = d.typ + h.base;
    auto load = [&](const unsigned (&wd)[U]) {
        unsigned jr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) jr[u] = h.decode_rel(wd[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            pj[u] = pt[jr[u]];
            tj[u] = (FORM == ACC_1) ? 0 : tt[jr[u]];
        }
    };
    auto sum = [&](int s0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (s0 + u >= nn) break;
            float dxf, dyf, dzf;
            const float r2 = separation(pj[u], dxf, dyf, dzf);
            if (r2 <= out2) {
                pair_terms(dxf, dyf, dzf, r2, tj[u], 1.0f);
                r2max = max(r2max, __float_as_int(r2));
            }
        }
    };
    walk_list<U>(stage, h, load, sum);
    // the rare atom that took a pair in the band walks its list once more
    // and takes out again each such pair the reference's FP64 test rejects
    if (r2max > __float_as_int(in2)) {
        for (int q = 0; q < nn; ++q) {
            const int j = h.decode(q < STAGE ? stage[q][threadIdx.x] : h.word(q));
            float dxf, dyf, dzf;
            const float r2 = separation(d.posf[j], dxf, dyf, dzf);
            if (r2 <= out2 && r2 > in2 && !reference_cut_test(d, tp, ts, i, j, rc2))
                pair_terms(dxf, dyf, dzf, r2, (FORM == ACC_1) ? 0 : d.typ[j], -1.0f);
        }
    }
    if constexpr (KAHAN) {
        fx = kfx;
        fy = kfy;
        fz = kfz;
    } else {
        fx = sfx;
        fy = sfy;
        fz = sfz;
    }
    e += (double)ef;
    w += (double)wf;
}

template <typename AccT, int FORM, bool EN, bool KAHAN, int U, typename LT>
__device__ __forceinline__ void pair_sum(const DevPtrs &d, const TrajParams &tp, const TrajScalars &ts,
                                         int i, int rel, int ti, double &fx, double &fy, double &fz,
                                         double &e, double &w, unsigned (*stage)[BLK])
{
    // the list's first loads go out before anything waits on memory
    const ListHead<LT> h(d, tp, i, rel);
    if constexpr (sizeof(AccT) == 4) {
        const float4 pi = d.posf[i];
        if (ts.ncx < 3 || ts.ncy < 3 || ts.ncz < 3)
            pair_sum_mixed<FORM, EN, true, KAHAN, U>(d, tp, ts, i, pi, h, ti, fx, fy, fz, e, w, stage);
        else
            pair_sum_mixed<FORM, EN, false, KAHAN, U>(d, tp, ts, i, pi, h, ti, fx, fy, fz, e, w, stage);
    } else
        pair_sum_fp64<FORM, EN, U>(d, tp, ts, i, h, ti, fx, fy, fz, e, w, stage);
}

// The force kernels' epilogue for atom slot i of block b: accelerations from
// the pair sums (for-a-lj-n-pbc.f:121-127 / for-a-lj-n-pbc-m.f:146-152), the
// integrator's update fused behind them, and the block's partial sums (energy,
// kinetic terms, virial).  Every thread of the block calls it (act: slot i
// holds an atom).
template <typename AccT, bool EN>
__device__ __forceinline__ void force_finish(const DevPtrs &d, const TrajParams &tp, int b, int i, bool act,
                                             bool form1, int ti, int mode, double fx, double fy, double fz,
                                             double e, double w, double (*lds)[4])
{
    constexpr bool MIXED = sizeof(AccT) == 4;
    double eks = 0.0, ekfree = 0.0;
    if (act) {
        // the velocities after the pair sum: loaded before it, they would
        // hold six registers through the pair loop, which the single-species
        // kernel (held to 72) would spill
        double vx = 0.0, vy = 0.0, vz = 0.0;
        if (mode == FORCE_STEP) {
            vx = d.vx[i];
            vy = d.vy[i];
            vz = d.vz[i];
        }
        double ax, ay, az;
        double ax0 = 0.0, ay0 = 0.0, az0 = 0.0;
        if (mode == FORCE_ACCATOM) load_acc<AccT>(d, i, ax0, ay0, az0);
        if (form1) {
            const double iatm = tp.imas[0];
            ax = ax0 + tp.e48[0] * fx * iatm;
            ay = ay0 + tp.e48[0] * fy * iatm;
            az = az0 + tp.e48[0] * fz * iatm;
        } else {
            const double iatm = tp.imas[ti];
            ax = ax0 + fx * iatm;
            ay = ay0 + fy * iatm;
            az = az0 + fz * iatm;
        }
        if (MIXED) {
            // integrate what is stored
            const float axs = (float)ax, ays = (float)ay, azs = (float)az;
            d.axf[i] = axs;  // cached: the next drift reads them
            d.ayf[i] = ays;
            d.azf[i] = azs;
            ax = axs;
            ay = ays;
            az = azs;
        } else {
            d.ax[i] = ax;
            d.ay[i] = ay;
            d.az[i] = az;
        }
        if (mode == FORCE_STEP) {
            const double dt = tp.dt;
            if (tp.integrator == INT_NH) {
                const double dth = dt / 2;
                const double wx = vx + dth * ax, wy = vy + dth * ay, wz = vz + dth * az;
                if (nh_coupled(tp, ti)) {
                    eks = wx * wx + wy * wy + wz * wz;
                    if (!tp.nh_refmass) eks *= tp.mrel[ti];
                } else {
                    // not thermostatted: its final velocity is already known
                    ekfree = 0.5 * tp.mas[ti] * (wx * wx + wy * wy + wz * wz);
                }
            } else {
                vx = vx + 0.5 * ax * dt;
                vy = vy + 0.5 * ay * dt;
                vz = vz + 0.5 * az * dt;
                d.vx[i] = vx;
                d.vy[i] = vy;
                d.vz[i] = vz;
                const double m = (tp.integrator == INT_VV) ? tp.mas[0] : tp.mas[ti];
                eks = 0.5 * m * (vx * vx + vy * vy + vz * vz);
            }
        }
    }
    if constexpr (EN) {
        double r[4] = {e, eks, ekfree, w};
        block_reduce_end<4>(r, {false, false, false, false}, lds);
        if (threadIdx.x == 0) {
            d.part[(size_t)b * NPART + P_EPOT] = r[0];
            d.part[(size_t)b * NPART + P_EKS] = r[1];
            d.part[(size_t)b * NPART + P_SUMV2] = r[2];
            d.part[(size_t)b * NPART + P_VIR] = r[3];
        }
    } else {
        double r[2] = {eks, ekfree};
        block_reduce_end<2>(r, {false, false}, lds);
        if (threadIdx.x == 0) {
            d.part[(size_t)b * NPART + P_EKS] = r[0];
            d.part[(size_t)b * NPART + P_SUMV2] = r[1];
        }
    }
}

// The atom block a force launch's workgroup takes, or -1.  FP_REBUILT: one
// work-list item per workgroup, split over the XCDs in runs; else the
// XCD-aware block order.
__device__ __forceinline__ int force_block(const DevPtrs &d, int part)
{
    if (part == FP_REBUILT) {
        const int n = d.wl_cnt[WL_BLK], c = (n + NXCD - 1) / NXCD;
        const int w = (int)(blockIdx.x % NXCD) * c + (int)(blockIdx.x / NXCD);
        if ((int)(blockIdx.x / NXCD) >= c || w >= n) return -1;
        return d.wl_blk[w];
    }
    const int b = xcd_block(blockIdx.x, d.nblk);
    return b < d.nblk ? b : -1;
}

// LJ forces over the full neighbour list plus the integrator's fused
// epilogue.  MULTI = false: every trajectory of the launch is single-species
// (ACC_1); MULTI = true: each trajectory uses its own form.  U: neighbours
// per load batch (a tunable; the pairs are summed in list order whatever U).
// LT: list entry type.  Single-species kernels of up to 4 neighbours a
// batch are held to 64 VGPRs (8 wavefronts a SIMD, without spills: the
// FP32 grid and box constants are scalars).
//
// EN: the energy and virial sums are formed (see launch_force).
//
// part: which atom blocks the launch takes.  FP_ALL: every block.
// FP_KEPT: the blocks of trajectories that keep their list this step (this
// launch can run beside the rebuild of the others).  FP_REBUILT: the blocks
// of the trajectories that rebuild, from the rebuild's work list (after it).
template <typename AccT, bool MULTI, int U, typename LT, bool EN, bool KAHAN>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(!MULTI && U <= 4 ? 8 : 1))) void k_force(
    DevPtrs d, int mode, int part)
{
    __shared__ double lds[4][4];
    __shared__ unsigned stage[STAGE][BLK];  // each thread's first list slots
    const int b = force_block(d, part);
    if (b < 0) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    const TrajScalars &ts = d.ts[t];
    if (ts.err) return;
    if (part == FP_KEPT && ts.rebuild) return;
    const int i = b * BLK + threadIdx.x;
    const int rel = i - tp.abase;
    const bool act = rel < tp.natom;
    const bool form1 = !MULTI || tp.accatom == ACC_1;
    double e = 0.0, w = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    int ti = 0;
    if (act) {
        ti = tp.nsorts > 1 ? d.typ[i] : 0;  // one species: no load
        if (form1)
            pair_sum<AccT, ACC_1, EN, KAHAN, U, LT>(d, tp, ts, i, rel, ti, fx, fy, fz, e, w, stage);
        else
            pair_sum<AccT, ACC_M, EN, KAHAN, U, LT>(d, tp, ts, i, rel, ti, fx, fy, fz, e, w, stage);
    }
    force_finish<AccT, EN>(d, tp, b, i, act, form1, ti, mode, fx, fy, fz, e, w, lds);
}

// Nose-Hoover final velocities v' = (v_h + dth a)/(1 + dth psi'), for
// trajectories whose update is still pending (the run loop leaves it to the
// next drift; everything that reads v calls this first).
template <typename AccT>
__global__ __launch_bounds__(BLK) void k_kick_nh(DevPtrs d)
{
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    const TrajScalars &ts = d.ts[t];
    if (ts.err || tp.integrator != INT_NH || !ts.pending) return;
    const int i = b * BLK + threadIdx.x;
    if (i - tp.abase >= tp.natom) return;
    const double dth = tp.dt / 2;
    double ax, ay, az;
    load_acc<AccT>(d, i, ax, ay, az);
    const double den = nh_coupled(tp, tp.nsorts > 1 ? d.typ[i] : 0) ? 1.0 + dth * ts.psi : 1.0;
    d.vx[i] = (d.vx[i] + dth * ax) / den;
    d.vy[i] = (d.vy[i] + dth * ay) / den;
    d.vz[i] = (d.vz[i] + dth * az) / den;
}

// positions replaced from the host: recompute the mixed-path cell offsets of
// trajectory t's atoms against the cells of its last list build
__global__ __launch_bounds__(BLK) void k_refresh_posf(DevPtrs d, int t)
{
    const TrajParams &tp = d.tp[t];
    const TrajScalars &ts = d.ts[t];
    const int rel = blockIdx.x * BLK + threadIdx.x;
    if (rel >= tp.natom) return;
    const int i = tp.abase + rel;
    int cx, cy, cz;
    unpack_cell(__int_as_float(d.cellid[i]), cx, cy, cz);
    d.posf[i] = make_posf(d.x[i], d.y[i], d.z[i], cx, cy, cz, ts);
}

void launch_refresh_posf(const DevPtrs &d, int t, int natom, hipStream_t s)
{
    // the mixed path only: elsewhere the records are rewritten at each list
    // build before anything reads them
    if (d.cellid) hipLaunchKernelGGL(k_refresh_posf, dim3((natom + BLK - 1) / BLK), dim3(BLK), 0, s, d, t);
}

// neighbor(step) called directly (init-cube1.f:113, read-rst.f:63): only
// trajectory `traj` rebuilds
__global__ void k_neighbor_entry(DevPtrs d, int traj, int ntraj, int step, int *step_out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int u = 0; u < ntraj; ++u) d.ts[u].rebuild = 0;
    TrajScalars &ts = d.ts[traj];
    if (ts.err) return;
    int s = step;
    neighbor_scalars(d.tp[traj], ts, &s);
    *step_out = s;
}

// ---------------------------------------------------------------------------
void launch_neighbor_entry(const DevPtrs &d, int traj, int ntraj, int step, int *step_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_neighbor_entry, dim3(1), dim3(64), 0, s, d, traj, ntraj, step, step_out);
}

void launch_drift_atoms(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s, const KTimer &kt)
{
    if (mixed)
        launch_timed(k_drift<float>, dim3(xcd_grid(g.nblk)), dim3(BLK), s, kt, d);
    else
        launch_timed(k_drift<double>, dim3(xcd_grid(g.nblk)), dim3(BLK), s, kt, d);
}

void launch_traj_pre(const DevPtrs &d, const LaunchDims &g, hipStream_t s, const KTimer &kt)
{
    launch_timed(k_traj_pre, dim3(g.ntraj), dim3(BLK), s, kt, d, (int)(g.ntraj == 1));
}

void launch_drift(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s)
{
    launch_drift_atoms(d, g, mixed, s);
    launch_traj_pre(d, g, s);
}

template <typename AccT, bool MULTI, typename LT, bool EN, bool KAHAN>
static void launch_force_k(const DevPtrs &d, const LaunchDims &g, int mode, int part, hipStream_t s,
                           const KTimer &kt)
{
    const dim3 grid(xcd_grid(g.nblk)), blk(BLK);
    if (g.force_u == 2)
        launch_timed(k_force<AccT, MULTI, 2, LT, EN, KAHAN>, grid, blk, s, kt, d, mode, part);
    else if (g.force_u == 8)
        launch_timed(k_force<AccT, MULTI, 8, LT, EN, KAHAN>, grid, blk, s, kt, d, mode, part);
    else
        launch_timed(k_force<AccT, MULTI, 4, LT, EN, KAHAN>, grid, blk, s, kt, d, mode, part);
}

// the Kahan-compensated FP32 sums exist on the mixed path only
template <typename AccT, bool MULTI, typename LT, bool EN>
static void launch_force_lt(const DevPtrs &d, const LaunchDims &g, int mode, int part, hipStream_t s,
                            const KTimer &kt)
{
    if (sizeof(AccT) == 4 && g.kahan)
        launch_force_k<AccT, MULTI, LT, EN, sizeof(AccT) == 4>(d, g, mode, part, s, kt);
    else
        launch_force_k<AccT, MULTI, LT, EN, false>(d, g, mode, part, s, kt);
}

template <typename AccT, bool MULTI>
static void launch_force_u(const DevPtrs &d, const LaunchDims &g, int mode, int part, hipStream_t s,
                           const KTimer &kt, bool energy)
{
    if (g.list16) {
        if (energy)
            launch_force_lt<AccT, MULTI, unsigned short, true>(d, g, mode, part, s, kt);
        else
            launch_force_lt<AccT, MULTI, unsigned short, false>(d, g, mode, part, s, kt);
    } else {
        if (energy)
            launch_force_lt<AccT, MULTI, int, true>(d, g, mode, part, s, kt);
        else
            launch_force_lt<AccT, MULTI, int, false>(d, g, mode, part, s, kt);
    }
}

void launch_force_atoms(const DevPtrs &d, const LaunchDims &g, bool mixed, bool single_species, int mode,
                        hipStream_t s, const KTimer &kt, int part, bool energy)
{
    if (mode == FORCE_ACCATOM) energy = true;  // AccAtom() always sets epot
    if (mixed) {
        if (single_species)
            launch_force_u<float, false>(d, g, mode, part, s, kt, energy);
        else
            launch_force_u<float, true>(d, g, mode, part, s, kt, energy);
    } else {
        if (single_species)
            launch_force_u<double, false>(d, g, mode, part, s, kt, energy);
        else
            launch_force_u<double, true>(d, g, mode, part, s, kt, energy);
    }
}

void launch_traj_post(const DevPtrs &d, const LaunchDims &g, int mode, hipStream_t s, const KTimer &kt,
                      bool energy)
{
    if (mode == FORCE_ACCATOM) energy = true;
    launch_timed(k_traj_post, dim3(g.ntraj), dim3(BLK), s, kt, d, mode, (int)energy);
}

void launch_force(const DevPtrs &d, const LaunchDims &g, bool mixed, bool single_species, int mode,
                  hipStream_t s, bool energy)
{
    launch_force_atoms(d, g, mixed, single_species, mode, s, {}, FP_ALL, energy);
    launch_traj_post(d, g, mode, s, {}, energy);
}

__global__ void k_clear_pending(DevPtrs d, int ntraj)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntraj) d.ts[t].pending = 0;
}

void launch_kick_nh(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s)
{
    if (mixed)
        hipLaunchKernelGGL(k_kick_nh<float>, dim3(xcd_grid(g.nblk)), dim3(BLK), 0, s, d);
    else
        hipLaunchKernelGGL(k_kick_nh<double>, dim3(xcd_grid(g.nblk)), dim3(BLK), 0, s, d);
    hipLaunchKernelGGL(k_clear_pending, dim3((g.ntraj + 255) / 256), dim3(256), 0, s, d, g.ntraj);
}

}  // namespace fcmd
