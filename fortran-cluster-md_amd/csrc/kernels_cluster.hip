// kernels_cluster.hip -- Stillinger and ten Wolde-Frenkel cluster analysis.
//
// Replaces the sequential label sweep of stoddard-pbc-n.f:69-148,
// stoddard-pbc-n-m.f:70-155 and stoddard-pbc-n-m-fre.f:84-258.
//
// Edges are neighbour-list pairs with r2 <= atskin^2, r2 formed with the
// reference's minimum image  d + L*(bf1 - int(bf2 + d/L)),  bf1 = 2 +
// int(atrskin/sigma), bf2 = bf1 + 0.5, so the edge set is the reference's
// bit for bit.  Components come from a lock-free union-find whose link rule
// always hangs the root with the larger reference atom number under the
// smaller one; the final root of a component is therefore its lowest
// atom number, which is exactly the reference's label cl(i).  The ten
// Wolde-Frenkel label is the 1-based rank, among liquid-like atoms in atom
// order, of the component's lowest member (stoddard-pbc-n-m-fre.f:198-225).
#include <climits>

#include "fcmd_common.hpp"
#include "fcmd_kernels.hpp"

namespace fcmd {

__device__ __forceinline__ int ld_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ int uf_find(int *p, int x)
{
    for (;;) {
        const int px = ld_relaxed(p + x);
        if (px == x) return x;
        const int ppx = ld_relaxed(p + px);
        if (ppx == px) return px;
        atomicCAS(p + x, px, ppx);  // path halving; losing the race is harmless
        x = ppx;
    }
}

__device__ void uf_unite(int *p, const int *key, int a, int b)
{
    for (;;) {
        a = uf_find(p, a);
        b = uf_find(p, b);
        if (a == b) return;
        if (key[a] > key[b]) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(p + b, b, a) == b) return;  // b was still a root
    }
}

__device__ __forceinline__ double kin(const DevPtrs &d, const TrajParams &tp, int i, int ty)
{
    const double m = (tp.stoddard == ST_N) ? tp.mas[0] : tp.mas[ty];
    const double vx = d.vx[i], vy = d.vy[i], vz = d.vz[i];
    return 0.5 * m * (vx * vx + vy * vy + vz * vz);
}

// reset per-slot work arrays and this trajectory's histograms
__global__ __launch_bounds__(BLK) void k_cl_init(DevPtrs d)
{
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    if (!d.cls[t].want || d.ts[t].err) return;
    const int i = b * BLK + threadIdx.x;
    const int rel = i - tp.abase;
    const bool fre = tp.stoddard == ST_FRE;  // the tWF arrays are read for tWF only
    d.parent[i] = i;
    d.cs_tmp[i] = 0;
    d.cgs_tmp[i] = 0;
    d.cek_tmp[i] = 0.0;
    if (fre) {
        d.parentf[i] = i;
        d.csf_tmp[i] = 0;
        d.cekf_tmp[i] = 0.0;
        d.liq[i] = 0;
    }
    if (rel < tp.natom) {
        if (fre) d.liqo[tp.abase + rel] = 0;
        for (int k = (rel == 0 ? 0 : rel + 1); k <= rel + 1; ++k) {
            const int h = tp.hbase + k;
            d.o_csi[h] = 0;
            d.o_ceks[h] = 0.0;
            if (fre) {
                d.o_csf[h] = 0;
                d.o_csif[h] = 0;
                d.o_cgsf[h] = 0;
                d.o_cgsfa[h] = 0;
                d.o_cekf[h] = 0.0;
                d.o_ceksf[h] = 0.0;
            }
        }
    }
    if (b == tp.bfirst && threadIdx.x == 0) {
        ClusterScalars &c = d.cls[t];
        c.cekf0 = 0.0;
        c.npos = 0;
        c.cbig = 0;
        c.nliq = 0;
        c.cbigf = 0;
    }
}

// edges within atskin; degrees; liquid-like flag (tWF).  LT: list entry type.
//
// REC (the mixed path, whose cell-offset records the drift keeps current):
// a neighbour's separation comes from its 16-byte record in FP32, as in the
// force kernel, one gather instead of four.  That r2 is within ~1e-6
// relative of the reference's FP64 r2, so outside a band of +-1e-4 around
// atskin^2 its verdict is the reference's; inside it the reference's FP64
// test on the FP64 positions decides, so the edge set is the reference's
// bit for bit either way.
template <typename LT, bool REC>
__global__ __launch_bounds__(BLK) void k_cl_edges(DevPtrs d)
{
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    const TrajScalars &ts = d.ts[t];
    if (!d.cls[t].want || ts.err) return;
    const int i = b * BLK + threadIdx.x;
    const int rel = i - tp.abase;
    if (rel >= tp.natom) return;
    const int ti = d.typ[i];
    const double bx = tp.box[0], by = tp.box[1], bz = tp.box[2];
    const double ibx = tp.ibox[0], iby = tp.ibox[1], ibz = tp.ibox[2];
    const double bf1 = fint(ts.atrskin / tp.atr) + 2.0;
    const double bf2 = bf1 + 0.5;
    const double rijc2 = tp.atskin * tp.atskin;
    const LT *lp = list_base<LT>(d) + list_slot(tp, rel);
    const int nn = d.nn[i];
    int deg = 0, deg1 = 0;
    int *ed = d.edges + (size_t)i * EMAX;
    const bool only_a = tp.stoddard == ST_NM;
    // the reference's pair test (stoddard-pbc-n.f:84-89)
    auto ref_r2 = [&](double xi, double yi, double zi, int j) {
        double dx = xi - d.x[j];
        double dy = yi - d.y[j];
        double dz = zi - d.z[j];
        dx = dx + bx * (bf1 - fint(bf2 + dx * ibx));
        dy = dy + by * (bf1 - fint(bf2 + dy * iby));
        dz = dz + bz * (bf1 - fint(bf2 + dz * ibz));
        return dx * dx + dy * dy + dz * dz;
    };
    if (!(only_a && ti != 0)) {
        // EB neighbours a batch: the batch's list entries, then their species
        // and records (or positions), are all loaded before the first edge is
        // stored (the compiler keeps no load behind a store it cannot tell
        // apart, so a neighbour at a time was two round trips each).  The
        // entries are loaded whatever the count (an atom owns lstride slots,
        // a multiple of LIST_ROUND, so a batch never reads past them: a load
        // under the lane's own condition would be waited for on the spot),
        // and a slot past the list then names the atom itself.
        constexpr int EB = 4;
        static_assert(LIST_ROUND % EB == 0, "batches must tile the allocated slots");
        auto load_batch = [&](int s0, int (&jj)[EB]) {
            LT w[EB];
#pragma unroll
            for (int u = 0; u < EB; ++u) w[u] = lp[(size_t)(s0 + u) * TILE];
#pragma unroll
            for (int u = 0; u < EB; ++u) jj[u] = s0 + u < nn ? tp.abase + (int)w[u] : i;
        };
        if constexpr (REC) {
            const float4 pi = d.posf[i];
            const float cxi = cell_xf(pi.w), cyi = cell_yf(pi.w), czi = cell_zf(pi.w);
            const float ncx = ts.ncf[0], ncy = ts.ncf[1], ncz = ts.ncf[2];
            const float icx = ts.incf[0], icy = ts.incf[1], icz = ts.incf[2];
            const float sx = ts.csf[0], sy = ts.csf[1], sz = ts.csf[2];
            const float fbx = tp.boxf[0], fby = tp.boxf[1], fbz = tp.boxf[2];
            const float fibx = tp.iboxf[0], fiby = tp.iboxf[1], fibz = tp.iboxf[2];
            const float lo2 = (float)(rijc2 * (1.0 - 1e-4)), hi2 = (float)(rijc2 * (1.0 + 1e-4));
            for (int s0 = 0; s0 < nn; s0 += EB) {
                int jj[EB], tj[EB];
                float4 pj[EB];
                load_batch(s0, jj);
#pragma unroll
                for (int u = 0; u < EB; ++u) {
                    pj[u] = d.posf[jj[u]];
                    tj[u] = d.typ[jj[u]];
                }
#pragma unroll
                for (int u = 0; u < EB; ++u) {
                    if (s0 + u >= nn || (only_a && tj[u] != 0)) continue;
                    const float4 q = pj[u];
                    const float cxj = cell_xf(q.w), cyj = cell_yf(q.w), czj = cell_zf(q.w);
                    const float dxf = ncx < 3.0f ? sep_small(cxi, cxj, sx, pi.x - q.x, fbx, fibx)
                                                 : sep_wrap(cxi, cxj, ncx, icx, sx, pi.x - q.x);
                    const float dyf = ncy < 3.0f ? sep_small(cyi, cyj, sy, pi.y - q.y, fby, fiby)
                                                 : sep_wrap(cyi, cyj, ncy, icy, sy, pi.y - q.y);
                    const float dzf = ncz < 3.0f ? sep_small(czi, czj, sz, pi.z - q.z, fbz, fibz)
                                                 : sep_wrap(czi, czj, ncz, icz, sz, pi.z - q.z);
                    const float r2f = __builtin_fmaf(dxf, dxf, __builtin_fmaf(dyf, dyf, dzf * dzf));
                    bool in = r2f <= lo2;
                    if (!in && r2f <= hi2) in = ref_r2(d.x[i], d.y[i], d.z[i], jj[u]) <= rijc2;
                    if (in) {
                        if (deg < EMAX) ed[deg] = jj[u];
                        ++deg;
                        deg1 += (tj[u] == 0);
                    }
                }
            }
        } else {
            const double xi = d.x[i], yi = d.y[i], zi = d.z[i];
            for (int s0 = 0; s0 < nn; s0 += EB) {
                int jj[EB], tj[EB];
                double xj[EB], yj[EB], zj[EB];
                load_batch(s0, jj);
#pragma unroll
                for (int u = 0; u < EB; ++u) {
                    tj[u] = d.typ[jj[u]];
                    xj[u] = d.x[jj[u]];
                    yj[u] = d.y[jj[u]];
                    zj[u] = d.z[jj[u]];
                }
#pragma unroll
                for (int u = 0; u < EB; ++u) {
                    if (s0 + u >= nn || (only_a && tj[u] != 0)) continue;
                    double dx = xi - xj[u];
                    double dy = yi - yj[u];
                    double dz = zi - zj[u];
                    dx = dx + bx * (bf1 - fint(bf2 + dx * ibx));
                    dy = dy + by * (bf1 - fint(bf2 + dy * iby));
                    dz = dz + bz * (bf1 - fint(bf2 + dz * ibz));
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 <= rijc2) {
                        if (deg < EMAX) ed[deg] = jj[u];
                        ++deg;
                        deg1 += (tj[u] == 0);
                    }
                }
            }
        }
    }
    if (deg > EMAX) {
        d.ts[t].err = ERR_EDGES;
        deg = EMAX;
    }
    d.edeg[i] = deg;
    d.edeg1[i] = deg1;
    if (tp.stoddard == ST_FRE && ti == 0 && deg >= tp.atnenum) {
        if (deg1 >= tp.atnenum) {
            d.liq[i] = 1;
            d.liqo[tp.abase + d.oid[i] - 1] = 1;
            atomicAdd(&d.cls[t].nliq, 1);
        } else {
            atomicAdd(&d.cls[t].cekf0, kin(d, tp, i, ti));
        }
    }
}

__global__ __launch_bounds__(BLK) void k_cl_hook(DevPtrs d)
{
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    if (!d.cls[t].want || d.ts[t].err) return;
    const int i = b * BLK + threadIdx.x;
    if (i - tp.abase >= tp.natom) return;
    const int oi = d.oid[i];
    const int deg = d.edeg[i];
    const int *ed = d.edges + (size_t)i * EMAX;
    const bool fre = tp.stoddard == ST_FRE;
    const bool li = fre && d.liq[i];
    // HB edges a batch: their ends and those ends' numbers loaded before the
    // batch's first union (whose atomics keep later loads behind them); an
    // atom's EMAX edge slots are whole batches
    constexpr int HB = 4;
    static_assert(EMAX % HB == 0, "batches must tile the edge slots");
    for (int k0 = 0; k0 < deg; k0 += HB) {
        int jj[HB], oj[HB], lj[HB];
#pragma unroll
        for (int u = 0; u < HB; ++u) jj[u] = ed[k0 + u];
#pragma unroll
        for (int u = 0; u < HB; ++u) jj[u] = k0 + u < deg ? jj[u] : i;
#pragma unroll
        for (int u = 0; u < HB; ++u) {
            oj[u] = d.oid[jj[u]];
            lj[u] = li ? d.liq[jj[u]] : 0;
        }
#pragma unroll
        for (int u = 0; u < HB; ++u)
            if (oi < oj[u]) {  // false past the edges (the atom itself)
                uf_unite(d.parent, d.oid, i, jj[u]);
                if (lj[u]) uf_unite(d.parentf, d.oid, i, jj[u]);
            }
    }
}

// per_atom: also the reference-order label array (fcmd_stoddard reads it; a
// device-driven run needs only the histograms and the largest clusters)
__global__ __launch_bounds__(BLK) void k_cl_count(DevPtrs d, int per_atom)
{
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    if (!d.cls[t].want || d.ts[t].err) return;
    const int i = b * BLK + threadIdx.x;
    if (i - tp.abase >= tp.natom) return;
    const int ti = d.typ[i];
    // an atom without edges is a component of its own (edges are symmetric:
    // no other atom hangs under it), so it takes plain stores, not atomics
    const bool alone = d.edeg[i] == 0;
    const int r = alone ? i : uf_find(d.parent, i);
    if (per_atom) d.o_cl[tp.abase + d.oid[i] - 1] = d.oid[r];
    const double ke = kin(d, tp, i, ti);
    if (tp.stoddard == ST_FRE && ti != 0) {
        if (alone)
            d.cgs_tmp[i] = 1;
        else
            atomicAdd(d.cgs_tmp + r, 1);
    } else if (alone) {
        d.cs_tmp[i] = 1;
        d.cek_tmp[i] = ke;
    } else {
        atomicAdd(d.cs_tmp + r, 1);
        atomicAdd(d.cek_tmp + r, ke);
    }
    if (tp.stoddard == ST_FRE && d.liq[i]) {
        const int rf = uf_find(d.parentf, i);
        atomicAdd(d.csf_tmp + rf, 1);
        atomicAdd(d.cekf_tmp + rf, ke);
    }
}

// rank of each liquid-like atom among liquid-like atoms in atom order
// (one block per trajectory; runs for tWF only)
__global__ __launch_bounds__(BLK) void k_cl_rank(DevPtrs d)
{
    __shared__ int lds[4];
    const int t = blockIdx.x;
    const TrajParams &tp = d.tp[t];
    if (!d.cls[t].want || d.ts[t].err || tp.stoddard != ST_FRE) return;
    int carry = 0;
    for (int b0 = 0; b0 < tp.natom; b0 += BLK) {
        const int k = b0 + threadIdx.x;
        const int v = k < tp.natom ? d.liqo[tp.abase + k] : 0;
        // inclusive wave scan, then across the 4 wavefronts
        int inc = v;
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(inc, o, 64);
            if (lane >= o) inc += u;
        }
        const int w = threadIdx.x >> 6;
        if (lane == 63) lds[w] = inc;
        __syncthreads();
        int off = carry;
        for (int q = 0; q < w; ++q) off += lds[q];
        const int tot = lds[0] + lds[1] + lds[2] + lds[3];
        if (k < tp.natom) d.rank[tp.abase + k] = off + inc - v;
        carry += tot;
        __syncthreads();
    }
}

// per_atom: also the reference-order per-atom arrays (cs, cek, cgs, clf, ccf)
// Cluster sizes below HLDS are counted per block in LDS and added to the
// histograms once per block and size: every root of a size-2 cluster in a
// trajectory adding to the same two histogram entries had made this kernel
// wait on one long chain of atomics per entry.
constexpr int HLDS = 64;

__global__ __launch_bounds__(BLK) void k_cl_finish(DevPtrs d, int per_atom)
{
    __shared__ double ldsd[4];
    __shared__ int ldsi[4];
    __shared__ int hcnt[HLDS];
    __shared__ double hke[HLDS];
    const int b = xcd_block(blockIdx.x, d.nblk);
    if (b >= d.nblk) return;
    const int t = d.blk_traj[b];
    const TrajParams &tp = d.tp[t];
    if (!d.cls[t].want || d.ts[t].err) return;
    const int i = b * BLK + threadIdx.x;
    const bool act = i - tp.abase < tp.natom;
    ClusterScalars &c = d.cls[t];
    for (int k = threadIdx.x; k < HLDS; k += BLK) {
        hcnt[k] = 0;
        hke[k] = 0.0;
    }
    __syncthreads();
    // monomers dominate a vapour: their histogram row, the count of clusters
    // and the largest one are per-block partials here, reduced per trajectory
    // in a fixed order by k_cl_sums (no atomics on one address from every block)
    int mono = 0, pos = 0, big = 0;
    double mono_ke = 0.0;
    if (act) {
        const int oi = d.oid[i];
        const int o = tp.abase + oi - 1;
        const bool root = d.edeg[i] == 0 || ld_relaxed(d.parent + i) == i;  // no edges: its own root
        if (root) {
            const int sz = d.cs_tmp[i];
            if (per_atom) {
                d.o_cs[o] = sz;
                d.o_cek[o] = d.cek_tmp[i];
                d.o_cgs[o] = d.cgs_tmp[i];
            }
            if (sz == 1) {
                mono = 1;
                mono_ke = d.cek_tmp[i];
            } else if (sz > 1 && sz < HLDS) {
                atomicAdd(hcnt + sz, 1);
                atomicAdd(hke + sz, d.cek_tmp[i]);
            } else if (sz > 1) {
                atomicAdd(d.o_csi + tp.hbase + sz, 1);
                atomicAdd(d.o_ceks + tp.hbase + sz, d.cek_tmp[i]);
            }
            pos = sz > 0;
            big = sz;
        } else if (per_atom) {
            d.o_cs[o] = 0;
            d.o_cek[o] = 0.0;
            d.o_cgs[o] = 0;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < HLDS; k += BLK)
        if (hcnt[k]) {
            atomicAdd(d.o_csi + tp.hbase + k, hcnt[k]);
            atomicAdd(d.o_ceks + tp.hbase + k, hke[k]);
        }
    const int bm = block_sum_i(mono, ldsi);
    const int bp = block_sum_i(pos, ldsi);
    const double be = block_sum(mono_ke, ldsd);
    const double bb = block_max((double)big, ldsd);
    if (threadIdx.x == 0) {
        double *q = d.clpart + (size_t)b * 4;
        q[0] = (double)bm;
        q[1] = be;
        q[2] = (double)bp;
        q[3] = bb;
    }
    if (!act || tp.stoddard != ST_FRE) return;
    const int oi = d.oid[i];
    const int o = tp.abase + oi - 1;
    if (d.liq[i]) {
        const int rf = uf_find(d.parentf, i);
        const int label = d.rank[tp.abase + d.oid[rf] - 1] + 1;
        if (per_atom) {
            d.o_clf[o] = label;
            d.o_ccf[o] = 1;
        }
        if (rf == i) {
            const int sz = d.csf_tmp[i];
            d.o_csf[tp.hbase + label] = sz;
            d.o_cekf[tp.hbase + label] = d.cekf_tmp[i];
            atomicAdd(d.o_csif + tp.hbase + sz, 1);
            atomicAdd(d.o_ceksf + tp.hbase + sz, d.cekf_tmp[i]);
            atomicMax(&c.cbigf, sz);
        }
    } else if (per_atom) {
        // last cluster to touch a non-liquid atom in the reference sweep: the
        // largest label among its liquid-like neighbours (0 if none)
        int lab = 0;
        const int deg = d.edeg[i];
        const int *ed = d.edges + (size_t)i * EMAX;
        for (int k = 0; k < deg; ++k) {
            const int j = ed[k];
            if (d.liq[j]) {
                const int rf = uf_find(d.parentf, j);
                lab = max(lab, d.rank[tp.abase + d.oid[rf] - 1] + 1);
            }
        }
        d.o_clf[o] = lab;
        d.o_ccf[o] = 0;
    }
}

// One block per trajectory: its blocks' k_cl_finish partials in a fixed order
// (thread k takes blocks k, k+BLK, ...; then a fixed tree): the monomer row
// of the histograms, the count of clusters and the largest one.
__global__ __launch_bounds__(BLK) void k_cl_sums(DevPtrs d)
{
    __shared__ double lds[4];
    const int t = blockIdx.x;
    const TrajParams &tp = d.tp[t];
    ClusterScalars &c = d.cls[t];
    if (!c.want || d.ts[t].err) return;
    double mono = 0.0, ke = 0.0, pos = 0.0, big = 0.0;
    for (int k = threadIdx.x; k < tp.nblk; k += BLK) {
        const double *q = d.clpart + (size_t)(tp.bfirst + k) * 4;
        mono += q[0];
        ke += q[1];
        pos += q[2];
        big = fmax(big, q[3]);
    }
    mono = block_sum(mono, lds);
    ke = block_sum(ke, lds);
    pos = block_sum(pos, lds);
    big = block_max(big, lds);
    if (threadIdx.x == 0) {
        d.o_csi[tp.hbase + 1] += (int)mono;
        d.o_ceks[tp.hbase + 1] += ke;
        c.npos += (int)pos;
        c.cbig = max(c.cbig, (int)big);
    }
}

// Cluster-size probability counts, helpers/sizer.f:59-103 (Stillinger
// histogram, SWITCH = 1): one block per trajectory adds this analysis's
// csi(g) to its counts for g = 1 .. the first size >= nmax present
// (inclusive: sizer counts that row, then sets its stop flag), and stops the
// trajectory there.  Snapshots before step `warmup` of the trajectory's own
// step counter add nothing but can stop the counting, as in sizer (its flag
// is set whatever the step).
__global__ __launch_bounds__(BLK) void k_size_counts(DevPtrs d, int nmax, int warmup)
{
    const int t = blockIdx.x;
    if (d.hacc_stop[t] || !d.cls[t].want || d.ts[t].err) return;
    const int step = d.ts[t].step;
    const TrajParams &tp = d.tp[t];
    const int *csi = d.o_csi + tp.hbase;
    __shared__ int first;
    if (threadIdx.x == 0) first = INT_MAX;
    __syncthreads();
    for (int g = max(nmax, 1) + threadIdx.x; g <= tp.natom; g += BLK)
        if (csi[g] > 0) {
            atomicMin(&first, g);
            break;
        }
    __syncthreads();
    const int upto = first == INT_MAX ? tp.natom : first;
    if (step >= warmup)
        for (int g = 1 + threadIdx.x; g <= upto; g += BLK) d.o_hacc[tp.hbase + g] += csi[g];
    if (threadIdx.x == 0 && first != INT_MAX) d.hacc_stop[t] = 1;
}

void launch_size_counts(const DevPtrs &d, const LaunchDims &g, int nmax, int warmup, hipStream_t s)
{
    hipLaunchKernelGGL(k_size_counts, dim3(g.ntraj), dim3(BLK), 0, s, d, nmax, warmup);
}

void launch_cluster(const DevPtrs &d, const LaunchDims &g, bool fre, bool per_atom, hipStream_t s)
{
    hipLaunchKernelGGL(k_cl_init, dim3(xcd_grid(g.nblk)), dim3(BLK), 0, s, d);
    // the mixed path keeps its cell-offset records current (d.cellid: the
    // cells they are taken against)
    const dim3 ge(xcd_grid(g.nblk));
    if (d.cellid) {
        if (g.list16)
            hipLaunchKernelGGL((k_cl_edges<unsigned short, true>), ge, dim3(BLK), 0, s, d);
        else
            hipLaunchKernelGGL((k_cl_edges<int, true>), ge, dim3(BLK), 0, s, d);
    } else {
        if (g.list16)
            hipLaunchKernelGGL((k_cl_edges<unsigned short, false>), ge, dim3(BLK), 0, s, d);
        else
            hipLaunchKernelGGL((k_cl_edges<int, false>), ge, dim3(BLK), 0, s, d);
    }
    hipLaunchKernelGGL(k_cl_hook, dim3(xcd_grid(g.nblk)), dim3(BLK), 0, s, d);
    hipLaunchKernelGGL(k_cl_count, dim3(xcd_grid(g.nblk)), dim3(BLK), 0, s, d, (int)per_atom);
    if (fre) hipLaunchKernelGGL(k_cl_rank, dim3(g.ntraj), dim3(BLK), 0, s, d);
    hipLaunchKernelGGL(k_cl_finish, dim3(xcd_grid(g.nblk)), dim3(BLK), 0, s, d, (int)per_atom);
    hipLaunchKernelGGL(k_cl_sums, dim3(g.ntraj), dim3(BLK), 0, s, d);
}

}  // namespace fcmd
