// kernels_nbr.hip -- cell-binned neighbour-list rebuild.
//
// Replaces the O(N^2) half-list build of neighbor-pbc.f:62-94 /
// neighbor-pbc-mav.f:197-229.  The drift kernel's per-trajectory epilogue
// raises a trajectory's rebuild flag; k_worklist then lists the atom blocks,
// cell chunks and tile chunks of the trajectories that rebuild now, and every
// rebuild kernel loops over that list with a fixed grid.  A step in which no
// trajectory rebuilds costs a few empty launches, and the sequence can be
// captured in a graph because no host decision is involved.
//
//   worklist -> zero_cells -> bin (wrap + cell + slot) -> scan(cell counts)
//   -> scatter -> cellsort (deterministic order inside a cell) -> permute
//   -> copyback -> count (neighbours per atom, tile widths) -> scan(tile
//   sizes) -> fill
//
// The pair test is the reference's: on wrapped positions,
//   d = x(i) - x(j);  d = d + L - L*int(1.5 + d/L);  r2 <= atrskin^2
// with i the lower reference atom number, so every pair gets the same r2, bit
// for bit, as in the reference's i<j loop and the pair set is identical.
#include <algorithm>

#include "fcmd_common.hpp"
#include "fcmd_kernels.hpp"

namespace fcmd {

constexpr int WL_GRID = 2048;  // blocks of the work-list kernels (8 per CU)

enum { WL_BLK = 0, WL_CCHK = 1, WL_TCHK = 2, WL_TRAJ = 3 };

__device__ __forceinline__ bool rebuilding(const TrajScalars &ts) { return ts.rebuild && !ts.err; }

// ---------------------------------------------------------------------------
__device__ __forceinline__ long long wave_incl_scan(long long v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive scan across the block; returns the thread's prefix and the block
// total (in *total).
__device__ __forceinline__ long long block_excl_scan(long long v, long long *lds, long long *total)
{
    const long long inc = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) lds[w] = inc;
    __syncthreads();
    long long off = 0;
    for (int k = 0; k < w; ++k) off += lds[k];
    *total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return off + inc - v;
}

// one block: where each rebuilding trajectory's items go in the work lists
__global__ __launch_bounds__(BLK) void k_wl_scan(DevPtrs d, int ntraj)
{
    __shared__ long long lds[4];
    long long carry[4] = {0, 0, 0, 0};
    for (int base = 0; base < ntraj; base += BLK) {
        const int t = base + threadIdx.x;
        const bool f = t < ntraj && rebuilding(d.ts[t]);
        const TrajParams *tp = f ? d.tp + t : nullptr;
        const long long n[4] = {f ? tp->nblk : 0, f ? tp->ncblk : 0, f ? tp->ntblk : 0, f ? 1 : 0};
        for (int k = 0; k < 4; ++k) {
            long long tot;
            const long long pre = block_excl_scan(n[k], lds, &tot) + carry[k];
            if (t < ntraj) d.wl_off[4 * t + k] = f ? (int)pre : -1;
            carry[k] += tot;
        }
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) d.wl_cnt[k] = (int)carry[k];
}

// one block per trajectory: write its items
__global__ __launch_bounds__(BLK) void k_wl_fill(DevPtrs d)
{
    const int t = blockIdx.x;
    const int *off = d.wl_off + 4 * t;
    if (off[0] < 0) return;
    const TrajParams &tp = d.tp[t];
    for (int q = threadIdx.x; q < tp.nblk; q += BLK) d.wl_blk[off[0] + q] = tp.bfirst + q;
    for (int q = threadIdx.x; q < tp.ncblk; q += BLK) d.wl_cchk[off[1] + q] = tp.cfirst + q;
    for (int q = threadIdx.x; q < tp.ntblk; q += BLK) d.wl_tchk[off[2] + q] = tp.tfirst + q;
    if (threadIdx.x == 0) d.wl_traj[off[3]] = t;
}

// grid-stride loop over one work list; the body sees a uniform item per block
#define FOR_ITEMS(list, k, item)                                                    \
    for (int w_ = blockIdx.x, n_ = d.wl_cnt[k]; w_ < n_; w_ += gridDim.x)          \
        if (const int item = (list)[w_]; true)

// ---------------------------------------------------------------------------
// segmented exclusive scans over SCAN_CHUNK-element chunks.
// kind 0: cells (in = cell_cnt, out = cell_start); kind 1: tiles (in = tile_w
// scaled by TILE, out = tile_off)
template <int KIND>
__global__ __launch_bounds__(BLK) void k_scan_chunk(DevPtrs d)
{
    __shared__ long long lds[4];
    FOR_ITEMS((KIND == 0 ? d.wl_cchk : d.wl_tchk), KIND == 0 ? WL_CCHK : WL_TCHK, c)
    {
        const int base = c * SCAN_CHUNK + threadIdx.x * 4;
        long long v[4];
        long long s = 0;
        for (int k = 0; k < 4; ++k) {
            v[k] = KIND == 0 ? (long long)d.cell_cnt[base + k] : (long long)d.tile_w[base + k] * TILE;
            s += v[k];
        }
        long long total;
        long long pre = block_excl_scan(s, lds, &total);
        for (int k = 0; k < 4; ++k) {
            if (KIND == 0)
                d.cell_start[base + k] = (int)pre;
            else
                d.tile_off[base + k] = pre;
            pre += v[k];
        }
        if (threadIdx.x == 0) d.scan_tmp_ll[c] = total;
    }
}

// one block per rebuilding trajectory: exclusive scan of its chunk totals
template <int KIND>
__global__ __launch_bounds__(BLK) void k_scan_totals(DevPtrs d)
{
    __shared__ long long lds[4];
    FOR_ITEMS(d.wl_traj, WL_TRAJ, t)
    {
        const TrajParams &tp = d.tp[t];
        TrajScalars &ts = d.ts[t];
        const int first = KIND == 0 ? tp.cfirst : tp.tfirst;
        const int n = KIND == 0 ? tp.ncblk : tp.ntblk;
        long long carry = 0;
        for (int b0 = 0; b0 < n; b0 += 4 * BLK) {
            long long v[4];
            long long s = 0;
            for (int k = 0; k < 4; ++k) {
                const int idx = b0 + threadIdx.x * 4 + k;
                v[k] = idx < n ? d.scan_tmp_ll[first + idx] : 0;
                s += v[k];
            }
            long long total;
            long long pre = block_excl_scan(s, lds, &total);
            pre += carry;
            for (int k = 0; k < 4; ++k) {
                const int idx = b0 + threadIdx.x * 4 + k;
                if (idx < n) d.scan_tmp_ll[first + idx] = pre;
                pre += v[k];
            }
            carry += total;
        }
        if (KIND == 1 && threadIdx.x == 0) {
            ts.nlist = (int)carry;
            if (carry > tp.lcap) ts.err = ERR_LISTFULL;
        }
        __syncthreads();
    }
}

template <int KIND>
__global__ __launch_bounds__(BLK) void k_scan_add(DevPtrs d)
{
    FOR_ITEMS((KIND == 0 ? d.wl_cchk : d.wl_tchk), KIND == 0 ? WL_CCHK : WL_TCHK, c)
    {
        const int t = KIND == 0 ? d.cblk_traj[c] : d.tblk_traj[c];
        const long long add = d.scan_tmp_ll[c] + (KIND == 1 ? d.tp[t].lbase : 0);
        const int base = c * SCAN_CHUNK + threadIdx.x * 4;
        for (int k = 0; k < 4; ++k) {
            if (KIND == 0)
                d.cell_start[base + k] += (int)add;
            else
                d.tile_off[base + k] += add;
        }
    }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_zero_cells(DevPtrs d)
{
    FOR_ITEMS(d.wl_cchk, WL_CCHK, c)
    {
        const int base = c * SCAN_CHUNK + threadIdx.x * 4;
        for (int k = 0; k < 4; ++k) d.cell_cnt[base + k] = 0;
    }
}

__device__ __forceinline__ int cell_coord(double p, double side, int n)
{
    int c = (int)(p / side);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// wrap in place (neighbor-pbc.f:62-66) and bin
__global__ __launch_bounds__(BLK) void k_bin(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const TrajParams &tp = d.tp[d.blk_traj[b]];
        const TrajScalars &ts = d.ts[d.blk_traj[b]];
        const int i = b * BLK + threadIdx.x;
        if (i - tp.abase < tp.natom) {
            const double bx = tp.box[0], by = tp.box[1], bz = tp.box[2];
            const double x = fmod(fmod(d.x[i], bx) + bx, bx);
            const double y = fmod(fmod(d.y[i], by) + by, by);
            const double z = fmod(fmod(d.z[i], bz) + bz, bz);
            d.x[i] = x;
            d.y[i] = y;
            d.z[i] = z;
            const int cx = cell_coord(x, ts.cside[0], ts.ncx);
            const int cy = cell_coord(y, ts.cside[1], ts.ncy);
            const int cz = cell_coord(z, ts.cside[2], ts.ncz);
            const int c = (cz * ts.ncy + cy) * ts.ncx + cx;
            d.atom_cell[i] = c;
            d.atom_slot[i] = atomicAdd(d.cell_cnt + tp.cbase + c, 1);
        }
    }
}

__global__ __launch_bounds__(BLK) void k_scatter(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const TrajParams &tp = d.tp[d.blk_traj[b]];
        const int i = b * BLK + threadIdx.x;
        if (i - tp.abase < tp.natom) {
            const int c = d.atom_cell[i];
            d.sorted[tp.abase + d.cell_start[tp.cbase + c] + d.atom_slot[i]] = i;
        }
    }
}

// insertion sort of each cell's atoms by reference atom number, so the
// order (and every later floating-point sum) does not depend on atomics
__global__ __launch_bounds__(BLK) void k_cellsort(DevPtrs d)
{
    FOR_ITEMS(d.wl_cchk, WL_CCHK, c)
    {
        const TrajParams &tp = d.tp[d.cblk_traj[c]];
        for (int k = 0; k < 4; ++k) {
            const int cell = c * SCAN_CHUNK + threadIdx.x * 4 + k;
            const int n = d.cell_cnt[cell];
            if (n < 2) continue;
            int *a = d.sorted + tp.abase + d.cell_start[cell];
            for (int p = 1; p < n; ++p) {
                const int v = a[p];
                const int key = d.oid[v];
                int q = p - 1;
                while (q >= 0 && d.oid[a[q]] > key) {
                    a[q + 1] = a[q];
                    --q;
                }
                a[q + 1] = v;
            }
        }
    }
}

template <typename AccT>
__global__ __launch_bounds__(BLK) void k_permute(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const TrajParams &tp = d.tp[d.blk_traj[b]];
        const int i = b * BLK + threadIdx.x;
        const int o = (i - tp.abase < tp.natom) ? d.sorted[i] : i;
        d.x2[i] = d.x[o];
        d.y2[i] = d.y[o];
        d.z2[i] = d.z[o];
        d.vx2[i] = d.vx[o];
        d.vy2[i] = d.vy[o];
        d.vz2[i] = d.vz[o];
        if (sizeof(AccT) == 8) {
            d.ax2[i] = d.ax[o];
            d.ay2[i] = d.ay[o];
            d.az2[i] = d.az[o];
        } else {
            d.axf2[i] = d.axf[o];
            d.ayf2[i] = d.ayf[o];
            d.azf2[i] = d.azf[o];
        }
        d.typ2[i] = d.typ[o];
        d.oid2[i] = d.oid[o];
    }
}

template <typename AccT>
__global__ __launch_bounds__(BLK) void k_copyback(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const TrajParams &tp = d.tp[d.blk_traj[b]];
        const int i = b * BLK + threadIdx.x;
        d.x[i] = d.x2[i];
        d.y[i] = d.y2[i];
        d.z[i] = d.z2[i];
        d.vx[i] = d.vx2[i];
        d.vy[i] = d.vy2[i];
        d.vz[i] = d.vz2[i];
        if (sizeof(AccT) == 8) {
            d.ax[i] = d.ax2[i];
            d.ay[i] = d.ay2[i];
            d.az[i] = d.az2[i];
        } else {
            d.axf[i] = d.axf2[i];
            d.ayf[i] = d.ayf2[i];
            d.azf[i] = d.azf2[i];
        }
        d.typ[i] = d.typ2[i];
        const int o = d.oid2[i];
        d.oid[i] = o;
        if (i - tp.abase < tp.natom) d.inv[tp.abase + o - 1] = i;
    }
}

// ---------------------------------------------------------------------------
// The candidate search: every atom j != i of the 27 (or fewer, without
// duplicates) cells around i's cell whose reference r2 is within the list
// radius, in a fixed order; sink(j, cnt) is called for each.
template <typename Sink>
__device__ __forceinline__ int candidates(const DevPtrs &d, const TrajParams &tp, const TrajScalars &ts,
                                          int i, int &half, Sink sink)
{
    const double bx = tp.box[0], by = tp.box[1], bz = tp.box[2];
    const double ibx = tp.ibox[0], iby = tp.ibox[1], ibz = tp.ibox[2];
    const double rs2 = ts.rs2;
    const double xi = d.x[i], yi = d.y[i], zi = d.z[i];
    const int oi = d.oid[i];
    const int ncx = ts.ncx, ncy = ts.ncy, ncz = ts.ncz;
    const int cx = cell_coord(xi, ts.cside[0], ncx);
    const int cy = cell_coord(yi, ts.cside[1], ncy);
    const int cz = cell_coord(zi, ts.cside[2], ncz);
    const int lx = ncx >= 3 ? -1 : 0, hx = ncx >= 3 ? 1 : ncx - 1;
    const int ly = ncy >= 3 ? -1 : 0, hy = ncy >= 3 ? 1 : ncy - 1;
    const int lz = ncz >= 3 ? -1 : 0, hz = ncz >= 3 ? 1 : ncz - 1;
    int cnt = 0;
    for (int oz = lz; oz <= hz; ++oz) {
        int zz = cz + oz;
        zz += (zz < 0) ? ncz : (zz >= ncz ? -ncz : 0);
        for (int oy = ly; oy <= hy; ++oy) {
            int yy = cy + oy;
            yy += (yy < 0) ? ncy : (yy >= ncy ? -ncy : 0);
            const int rowbase = (zz * ncy + yy) * ncx;
            for (int ox = lx; ox <= hx; ++ox) {
                int xx = cx + ox;
                xx += (xx < 0) ? ncx : (xx >= ncx ? -ncx : 0);
                const int cell = tp.cbase + rowbase + xx;
                const int j0 = tp.abase + d.cell_start[cell];
                const int j1 = j0 + d.cell_cnt[cell];
                for (int j = j0; j < j1; ++j) {
                    if (j == i) continue;
                    const int oj = d.oid[j];
                    double dx, dy, dz;
                    if (oi < oj) {
                        dx = xi - d.x[j];
                        dy = yi - d.y[j];
                        dz = zi - d.z[j];
                    } else {
                        dx = d.x[j] - xi;
                        dy = d.y[j] - yi;
                        dz = d.z[j] - zi;
                    }
                    dx = dx + bx - bx * fint(1.5 + dx * ibx);
                    dy = dy + by - by * fint(1.5 + dy * iby);
                    dz = dz + bz - bz * fint(1.5 + dz * ibz);
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 <= rs2) {
                        sink(j, cnt);
                        ++cnt;
                        half += (oi < oj);
                    }
                }
            }
        }
    }
    return cnt;
}

// scratch slot of candidate k of atom slot i (tile-major, coalesced per wave)
__device__ __forceinline__ size_t tmp_index(int i, int k)
{
    return ((size_t)(i / TILE) * NB_TMP + k) * TILE + (i & (TILE - 1));
}

// count: neighbours per atom, tile widths, pair count; the first NB_TMP
// neighbours of each atom are kept so that the fill only copies them
__global__ __launch_bounds__(BLK) void k_nbr_count(DevPtrs d)
{
    __shared__ int ldsi[4];
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const int t = d.blk_traj[b];
        const TrajParams &tp = d.tp[t];
        TrajScalars &ts = d.ts[t];
        const int i = b * BLK + threadIdx.x;
        const int rel = i - tp.abase;
        int cnt = 0, half = 0;
        if (rel < tp.natom) {
            int *tmp = d.nbtmp;
            cnt = candidates(d, tp, ts, i, half, [&](int j, int k) {
                if (k < NB_TMP) tmp[tmp_index(i, k)] = j;
            });
        }
        d.nn[i] = cnt;
        const int w = wave_max_i(cnt);
        if ((threadIdx.x & (TILE - 1)) == 0) d.tile_w[tp.tbase + rel / TILE] = w;
        const int h = block_sum_i(half, ldsi);
        if (threadIdx.x == 0 && h) atomicAdd(&ts.npairs, h);
    }
}

// fill the sliced-ELL list: copy the kept neighbours, search again only for
// atoms with more than NB_TMP (same order, so the same list)
__global__ __launch_bounds__(BLK) void k_nbr_fill(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const int t = d.blk_traj[b];
        const TrajParams &tp = d.tp[t];
        const TrajScalars &ts = d.ts[t];
        if (ts.err) continue;  // list overflow found by the tile scan
        const int i = b * BLK + threadIdx.x;
        const int rel = i - tp.abase;
        if (rel >= tp.natom) continue;
        const int nn = d.nn[i];
        int *out = d.list + d.tile_off[tp.tbase + rel / TILE] + (i & (TILE - 1));
        if (nn <= NB_TMP) {
            for (int k = 0; k < nn; ++k) out[(size_t)k * TILE] = d.nbtmp[tmp_index(i, k)];
        } else {
            int half = 0;
            candidates(d, tp, ts, i, half, [&](int j, int k) { out[(size_t)k * TILE] = j; });
        }
    }
}

// ---------------------------------------------------------------------------
void launch_rebuild(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s)
{
    const dim3 blk(BLK);
    const dim3 ga(std::min(g.nblk, WL_GRID)), gc(std::min(g.ncblk, WL_GRID));
    const dim3 gt(std::min(g.ntblk, WL_GRID)), gj(std::min(g.ntraj, WL_GRID));
    hipLaunchKernelGGL(k_wl_scan, dim3(1), blk, 0, s, d, g.ntraj);
    hipLaunchKernelGGL(k_wl_fill, dim3(g.ntraj), blk, 0, s, d);
    hipLaunchKernelGGL(k_zero_cells, gc, blk, 0, s, d);
    hipLaunchKernelGGL(k_bin, ga, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_chunk<0>, gc, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_totals<0>, gj, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_add<0>, gc, blk, 0, s, d);
    hipLaunchKernelGGL(k_scatter, ga, blk, 0, s, d);
    hipLaunchKernelGGL(k_cellsort, gc, blk, 0, s, d);
    if (mixed) {
        hipLaunchKernelGGL(k_permute<float>, ga, blk, 0, s, d);
        hipLaunchKernelGGL(k_copyback<float>, ga, blk, 0, s, d);
    } else {
        hipLaunchKernelGGL(k_permute<double>, ga, blk, 0, s, d);
        hipLaunchKernelGGL(k_copyback<double>, ga, blk, 0, s, d);
    }
    hipLaunchKernelGGL(k_nbr_count, ga, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_chunk<1>, gt, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_totals<1>, gj, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_add<1>, gt, blk, 0, s, d);
    hipLaunchKernelGGL(k_nbr_fill, ga, blk, 0, s, d);
}

}  // namespace fcmd
