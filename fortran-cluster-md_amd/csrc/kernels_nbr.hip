// kernels_nbr.hip -- cell-binned neighbour-list rebuild.
//
// Replaces the O(N^2) half-list build of neighbor-pbc.f:62-94 /
// neighbor-pbc-mav.f:67-99.  The drift kernel's per-trajectory epilogue
// raises a trajectory's rebuild flag; k_wl_fill then lists the atom blocks,
// cell chunks and trajectories of the trajectories that rebuild now, and
// every rebuild kernel loops over those lists with a fixed grid.  A step in
// which no trajectory rebuilds costs a few empty launches, and the sequence
// can be captured in a graph because no host decision is involved.
//
//   wl_fill -> bin (wrap + cell + slot) -> scan (cell counts, which it
//   clears for the next rebuild) -> scatter -> cellsort (deterministic order
//   inside a cell) -> permute (into the alternate buffers) -> nbr (search +
//   list in one pass, then the copy back)
//
// Cell counts are zero outside the bin -> scan window (zeroed at allocation;
// the scan clears every count of the trajectory's cell range after reading
// it).  After the cellsort, cell_first holds trajectory-relative starts for
// every cell of the range; a cell's atoms are [first[c], first[c + 1]), and
// the range keeps at least one cell past the grid, whose start is natom.
//
// The pair test is the reference's: on wrapped positions,
//   d = x(i) - x(j);  d = d + L - L*int(1.5 + d/L);  r2 <= atrskin^2
// with i the lower reference atom number, so every pair gets the same r2, bit
// for bit, as in the reference's i<j loop and the pair set is identical.
#include <algorithm>

#include "fcmd_common.hpp"
#include "fcmd_kernels.hpp"

namespace fcmd {


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

// One block per trajectory.  Each block sums the item counts of the
// rebuilding trajectories before its own (all blocks repeat this small scan
// over the trajectory flags instead of waiting for a separate launch), then
// writes its trajectory's items at that offset; block 0 writes the lengths.
__global__ __launch_bounds__(BLK) void k_wl_fill(DevPtrs d, int ntraj)
{
    __shared__ int lds[4][2 * WL_LISTS];
    const int t = blockIdx.x;
    int below[WL_LISTS] = {}, total[WL_LISTS] = {};
    for (int base = 0; base < ntraj; base += BLK) {
        const int u = base + threadIdx.x;
        const bool f = u < ntraj && rebuilding(d.ts[u]);
        for (int k = 0; k < WL_LISTS; ++k) {
            const int v = f ? d.wl_size[WL_LISTS * u + k] : 0;
            below[k] += u < t ? v : 0;
            total[k] += v;
        }
    }
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < WL_LISTS; ++k) {
        below[k] = wave_sum_i(below[k]);
        total[k] = wave_sum_i(total[k]);
        if ((threadIdx.x & 63) == 0) {
            lds[w][k] = below[k];
            lds[w][WL_LISTS + k] = total[k];
        }
    }
    __syncthreads();
    int off[WL_LISTS];
    for (int k = 0; k < WL_LISTS; ++k) {
        off[k] = lds[0][k] + lds[1][k] + lds[2][k] + lds[3][k];
        if (t == 0 && threadIdx.x == 0)
            d.wl_cnt[k] = lds[0][WL_LISTS + k] + lds[1][WL_LISTS + k] + lds[2][WL_LISTS + k] +
                          lds[3][WL_LISTS + k];
    }
    if (!rebuilding(d.ts[t])) return;
    const TrajParams &tp = d.tp[t];
    for (int q = threadIdx.x; q < tp.nblk; q += BLK) d.wl_blk[off[WL_BLK] + q] = tp.bfirst + q;
    for (int q = threadIdx.x; q < tp.ncblk; q += BLK) d.wl_cchk[off[WL_CCHK] + q] = tp.cfirst + q;
    if (threadIdx.x == 0) d.wl_traj[off[WL_TRAJ]] = t;
}

// Grid-stride loop over one work list, XCD-aware: the items are split into
// NXCD contiguous runs, one per XCD (workgroup IDs go to the XCDs
// round-robin), so the blocks one XCD runs are neighbours in space.
// gridDim.x must be a multiple of NXCD.  The body sees a uniform item per block.
#define FOR_ITEMS(list, k, item)                                                     \
    for (int n_ = d.wl_cnt[k], c_ = (n_ + NXCD - 1) / NXCD,                           \
             e_ = min(n_, (int)(blockIdx.x % NXCD + 1) * c_),                         \
             w_ = (int)(blockIdx.x % NXCD) * c_ + (int)(blockIdx.x / NXCD);          \
         w_ < e_; w_ += gridDim.x / NXCD)                                             \
        if (const int item = (list)[w_]; true)

// One item per workgroup, the same XCD-aware split as FOR_ITEMS, for kernels
// launched with at least as many workgroups as the list can hold (a loop over
// items keeps loop-invariant values live across the body and costs registers
// that a heavy body needs).  Workgroups past the list's length return.
#define ONE_ITEM(list, k, item)                                                      \
    const int n_ = d.wl_cnt[k], c_ = (n_ + NXCD - 1) / NXCD;                         \
    const int w_ = (int)(blockIdx.x % NXCD) * c_ + (int)(blockIdx.x / NXCD);         \
    if ((int)(blockIdx.x / NXCD) >= c_ || w_ >= n_) return;                          \
    const int item = (list)[w_]

// ---------------------------------------------------------------------------
// cell_start = exclusive scan of cell_cnt within each trajectory: a scan of
// each SCAN_CHUNK-cell chunk (k_scan_chunk, which also clears the counts), a
// scan of each trajectory's chunk totals (k_scan_totals); the chunk offset is
// added where a start is first used (k_scatter reads start + offset,
// k_cellsort writes the final start to cell_first)
__global__ __launch_bounds__(BLK) void k_scan_chunk(DevPtrs d)
{
    __shared__ long long lds[4];
    FOR_ITEMS(d.wl_cchk, WL_CCHK, c)
    {
        const int base = c * SCAN_CHUNK + threadIdx.x * 4;
        int v[4];
        long long s = 0;
        for (int k = 0; k < 4; ++k) {
            v[k] = d.cell_cnt[base + k];
            s += v[k];
        }
        for (int k = 0; k < 4; ++k) d.cell_cnt[base + k] = 0;
        long long total;
        long long pre = block_excl_scan(s, lds, &total);
        for (int k = 0; k < 4; ++k) {
            d.cell_start[base + k] = (int)pre;
            pre += v[k];
        }
        if (threadIdx.x == 0) d.scan_tmp_ll[c] = total;
    }
}

// one block per rebuilding trajectory: exclusive scan of its chunk totals
__global__ __launch_bounds__(BLK) void k_scan_totals(DevPtrs d)
{
    __shared__ long long lds[4];
    FOR_ITEMS(d.wl_traj, WL_TRAJ, t)
    {
        const TrajParams &tp = d.tp[t];
        const int first = tp.cfirst;
        const int n = tp.ncblk;
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
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// cell of a wrapped coordinate: p * (1/side), whose rounding is far below
// the margin the cell side keeps over the list radius
__device__ __forceinline__ int cell_coord(double p, double inv_side, int n)
{
    int c = (int)(p * inv_side);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// dmod(dmod(x,L)+L,L), neighbor-pbc.f:62-66.  For |x| < L the inner dmod is x
// itself and the outer one takes a = x+L in [0, 2L] down by L at most twice,
// exactly (Sterbenz), so the usual case needs no fmod; the result is the same
// double either way.
__device__ __forceinline__ double wrap_box(double x, double L)
{
    if (fabs(x) < L) {
        const double a = x + L;
        if (a < L) return a;
        const double r = a - L;
        return r < L ? r : r - L;
    }
    return fmod(fmod(x, L) + L, L);
}

// wrap in place (neighbor-pbc.f:62-66) and bin.  Every read comes before the
// first store, and the slot's atomic before the other stores (the compiler
// keeps no load ahead of a store it cannot tell apart).
//
// A wavefront's atoms lie close to cell order (the slots were sorted by cell
// at the last re-sort), so neighbouring lanes often bin into the same cell:
// each run of equal cells in adjacent lanes takes one atomic for the run, its
// first lane adding the run's length, and the others count on from the
// result.  The slots within a cell are only distinct positions: the cell
// sort orders every cell by reference number afterwards.
__global__ __launch_bounds__(BLK) void k_bin(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const int t = d.blk_traj[b];
        const TrajParams &tp = d.tp[t];
        const TrajScalars &ts = d.ts[t];
        const int i = b * BLK + threadIdx.x;
        const bool act = i - tp.abase < tp.natom;  // the active lanes are a prefix of the wavefront
        const int lane = threadIdx.x & 63;
        int c = -1, cx = 0, cy = 0, cz = 0;
        double x = 0.0, y = 0.0, z = 0.0;
        if (act) {
            const int ncx = ts.ncx, ncy = ts.ncy, ncz = ts.ncz;
            const double bx = tp.box[0], by = tp.box[1], bz = tp.box[2];
            const double icx = ts.icside[0], icy = ts.icside[1], icz = ts.icside[2];
            const double x0 = d.x[i], y0 = d.y[i], z0 = d.z[i];  // issued together, before the wraps' branches
            x = wrap_box(x0, bx), y = wrap_box(y0, by), z = wrap_box(z0, bz);
            cx = cell_coord(x, icx, ncx);
            cy = cell_coord(y, icy, ncy);
            cz = cell_coord(z, icz, ncz);
            c = (cz * ncy + cy) * ncx + cx;
        }
        const int cprev = __shfl_up(c, 1, 64);
        const bool head = act && (lane == 0 || cprev != c);
        const unsigned long long heads = __builtin_amdgcn_ballot_w64(head);
        const int nact = __builtin_popcountll(__builtin_amdgcn_ballot_w64(act));
        const unsigned long long below = heads & (~0ull >> (63 - lane));  // heads at or below this lane
        const unsigned long long above = heads & ~(~0ull >> (63 - lane));
        const int hl = 63 - __builtin_clzll(below | 1ull);                 // this lane's run head
        int base = 0;
        if (head) base = atomicAdd(d.cell_cnt + tp.cbase + c, (above ? __builtin_ctzll(above) : nact) - lane);
        const int slot = __shfl(base, hl, 64) + (lane - hl);
        if (act) {
            const int packed = cx | (cy << CELL_BITS) | (cz << (2 * CELL_BITS));
            d.x[i] = x;
            d.y[i] = y;
            d.z[i] = z;
            d.atom_cell[i] = packed;
            d.atom_slot[i] = slot;
            if (!ts.move) {
                // the atom keeps its slot: its cell-offset record for the new
                // grid here (a re-sorting rebuild writes it in k_permute)
                d.posf[i] = make_posf(x, y, z, cx, cy, cz, ts);
                if (d.cellid) d.cellid[i] = packed;
            }
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
            const TrajScalars &ts = d.ts[d.blk_traj[b]];
            const int pc = d.atom_cell[i], slot = d.atom_slot[i];
            const int c = tp.cbase + ((pc >> (2 * CELL_BITS)) * ts.ncy + ((pc >> CELL_BITS) & CELL_MAX)) * ts.ncx +
                          (pc & CELL_MAX);
            const int start = d.cell_start[c] + (int)d.scan_tmp_ll[c / SCAN_CHUNK];
            d.sorted[tp.abase + start + slot] = i;
        }
    }
}

// compare-exchange of (key, value) pairs by key
__device__ __forceinline__ void cmp_swap(int &ka, int &va, int &kb, int &vb)
{
    if (ka > kb) {
        const int k = ka, v = va;
        ka = kb, va = vb;
        kb = k, vb = v;
    }
}

// sort of each cell's atoms by reference atom number, so the order (and every
// later floating-point sum) does not depend on atomics; also writes each
// cell's final start to cell_first.  One thread per cell: a chunk's
// SCAN_CHUNK cells are CS_PARTS workgroups of BLK threads (no wider than the
// force kernel's, so that they find room on a CU beside the force launch that
// runs during the rebuild), and with the final starts in an array of their
// own no thread waits for another.  A cell of up to CS_REG atoms loads its
// members and their numbers at once and sorts them in registers (odd-even
// transposition, sort_small); a larger one is sorted afterwards by its
// whole wavefront (a thread sorting 15-30 atoms in place, a chain of
// dependent loads per step, held the whole launch once droplets had formed).
constexpr int CS_PARTS = SCAN_CHUNK / BLK;  // workgroups per chunk

// insertion sort of a cell's n atoms by reference number, by one thread
__device__ __forceinline__ void sort_cell_serial(const DevPtrs &d, int *a, int n)
{
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

// A cell of n <= W atoms sorted in registers by an odd-even transposition
// network of W phases (the members past n padded with the largest key).  The
// cell sort picks the narrowest W that holds every cell of its wavefront:
// in a vapour most cells hold 0-3 atoms, and the 8-wide network had cost
// every thread 28 compare-exchanges whatever its cell.
template <int W>
__device__ __forceinline__ void sort_small(const DevPtrs &d, int *a, int n)
{
    if (n < 2) return;
    int v[W], key[W];
#pragma unroll
    for (int p = 0; p < W; ++p) v[p] = p < n ? a[p] : 0;
#pragma unroll
    for (int p = 0; p < W; ++p) key[p] = p < n ? d.oid[v[p]] : INT_MAX;
#pragma unroll
    for (int ph = 0; ph < W; ++ph)
#pragma unroll
        for (int p = ph & 1; p + 1 < W; p += 2) cmp_swap(key[p], v[p], key[p + 1], v[p + 1]);
#pragma unroll
    for (int p = 0; p < W; ++p)
        if (p < n) a[p] = v[p];
}

// A large cell (more than CS_REG atoms) sorted by the whole wavefront: lane
// L holds member L and its reference number, counts the numbers below its
// own over the cell (one shuffle a member) and stores the member at that
// rank.  The numbers are distinct, so this is the order the one-thread sort
// gives.  A cell of more than 64 atoms (never met at liquid density and the
// list radii here, but tested) falls back to lane 0's insertion sort.
// Wave-uniform: a and n are the same in every lane.
__device__ __forceinline__ void sort_cell_wave(const DevPtrs &d, int *a, int n)
{
    const int lane = threadIdx.x & 63;
    if (n > 64) {
        if (lane == 0) sort_cell_serial(d, a, n);
        return;
    }
    const int v = lane < n ? a[lane] : 0;
    const int key = lane < n ? d.oid[v] : INT_MAX;
    int rank = 0;
    for (int m = 0; m < n; ++m) rank += __shfl(key, m, 64) < key;
    if (lane < n) a[rank] = v;
}

__global__ __launch_bounds__(BLK) void k_cellsort(DevPtrs d)
{
    // the work list's chunks, CS_PARTS items each, split over the XCDs as
    // FOR_ITEMS does (a chunk's parts stay on one XCD)
    const int n = d.wl_cnt[WL_CCHK] * CS_PARTS, per = (n + NXCD - 1) / NXCD;
    const int end = min(n, (int)(blockIdx.x % NXCD + 1) * per);
    for (int w = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD); w < end; w += gridDim.x / NXCD) {
        const int c = d.wl_cchk[w / CS_PARTS];
        const TrajParams &tp = d.tp[d.cblk_traj[c]];
        const int add = (int)d.scan_tmp_ll[c], abase = tp.abase;
        const int local = (w % CS_PARTS) * BLK + threadIdx.x;
        const int cell = c * SCAN_CHUNK + local;
        // chunk-relative start and end (the next cell's start, or for the
        // chunk's last cell the next chunk's offset or natom)
        const int s0 = d.cell_start[cell];
        const int e0 = local + 1 < SCAN_CHUNK
                           ? d.cell_start[cell + 1]
                           : (c + 1 < tp.cfirst + tp.ncblk ? (int)d.scan_tmp_ll[c + 1] : tp.natom) - add;
        d.cell_first[cell] = s0 + add;
        const int cnt = e0 - s0, first = abase + s0 + add;
        const int small = cnt <= CS_REG ? cnt : 0;
        if (!__builtin_amdgcn_ballot_w64(small > 2))
            sort_small<2>(d, d.sorted + first, small);
        else if (!__builtin_amdgcn_ballot_w64(small > 4))
            sort_small<4>(d, d.sorted + first, small);
        else
            sort_small<CS_REG>(d, d.sorted + first, small);
        // the wavefront's large cells, one after another (w is uniform over
        // the workgroup, so every lane of the wavefront is here)
        for (unsigned long long big = __builtin_amdgcn_ballot_w64(cnt > CS_REG); big; big &= big - 1) {
            const int l = __builtin_ctzll(big);
            sort_cell_wave(d, d.sorted + __shfl(first, l, 64), __shfl(cnt, l, 64));
        }
    }
}

// A re-sorting rebuild (ts.move) gathers every per-atom array into cell
// order (into the alternate buffers) and writes each atom's cell-offset
// record for its cell (on the mixed path also the cell alone, for the
// drift).  Any other rebuild leaves the atoms in their slots and writes the
// list search's view of them in cell order (cpos, cslot), 20 bytes an atom
// instead of moving every array there and back.
// the acceleration array of the precision in use: FP64, or FP32 on the mixed path
template <typename AccT>
__device__ __forceinline__ AccT *acc_array(double *p64, float *p32)
{
    if constexpr (sizeof(AccT) == 8)
        return p64;
    else
        return p32;
}

// Both kernels that move per-atom arrays (here and the list search's copy
// back) issue every load before any store: the compiler cannot tell the arrays
// apart, so a store ahead of a later load would make each array a round trip
// of its own.
template <typename AccT>
__global__ __launch_bounds__(BLK) void k_permute(DevPtrs d)
{
    FOR_ITEMS(d.wl_blk, WL_BLK, b)
    {
        const int t = d.blk_traj[b];
        const TrajParams &tp = d.tp[t];
        const TrajScalars &ts = d.ts[t];
        const int i = b * BLK + threadIdx.x;
        const bool act = i - tp.abase < tp.natom;
        const int o = act ? d.sorted[i] : i;
        if (!ts.move) {
            // atoms stay in their slots: the list search's candidates in
            // cell order, position i of the order holding slot o
            d.cpos[i] = d.posf[o];
            d.cslot[i] = o;
            continue;
        }
        const AccT *ax = acc_array<AccT>(d.ax, d.axf), *ay = acc_array<AccT>(d.ay, d.ayf),
                   *az = acc_array<AccT>(d.az, d.azf);
        const double x = d.x[o], y = d.y[o], z = d.z[o];
        const double vx = d.vx[o], vy = d.vy[o], vz = d.vz[o];
        const AccT a0 = ax[o], a1 = ay[o], a2 = az[o];
        const int typ = d.typ[o], oid = d.oid[o], cw = d.atom_cell[o];
        int cx, cy, cz;
        unpack_cell(__int_as_float(cw), cx, cy, cz);
        const float4 r = make_posf(x, y, z, cx, cy, cz, ts);
        if (act) {
            d.posf[i] = r;
            if (sizeof(AccT) == 4) d.cellid[i] = __float_as_int(r.w);
        }
        d.x2[i] = x;
        d.y2[i] = y;
        d.z2[i] = z;
        d.vx2[i] = vx;
        d.vy2[i] = vy;
        d.vz2[i] = vz;
        acc_array<AccT>(d.ax2, d.axf2)[i] = a0;
        acc_array<AccT>(d.ay2, d.ayf2)[i] = a1;
        acc_array<AccT>(d.az2, d.azf2)[i] = a2;
        d.typ2[i] = typ;
        d.oid2[i] = oid;
    }
}

// ---------------------------------------------------------------------------
// The reference's pair test on the FP64 positions of the slots (after a
// re-sorting rebuild the permuted arrays, else the arrays in place): on
// wrapped positions, d = x(i) - x(j) with i the lower reference atom number,
// d = d + L - L*int(1.5 + d/L), r2 <= atrskin^2 (neighbor-pbc.f:79-84).
// b, ib: box sides and their inverses, read once by the caller (read here,
// after the list stores, they would be reloaded for every candidate).
struct RefArrays {
    const double *x, *y, *z;
    const int *oid;
};
__device__ __forceinline__ bool reference_pair_test(const RefArrays &r, int i, int j, double rs2,
                                                   const double (&b)[3], const double (&ib)[3])
{
    const double bx = b[0], by = b[1], bz = b[2];
    const double ibx = ib[0], iby = ib[1], ibz = ib[2];
    const bool fwd = r.oid[i] < r.oid[j];
    const int a = fwd ? i : j, c = fwd ? j : i;
    double dx = r.x[a] - r.x[c];
    double dy = r.y[a] - r.y[c];
    double dz = r.z[a] - r.z[c];
    dx = dx + bx - bx * fint(1.5 + dx * ibx);
    dy = dy + by - by * fint(1.5 + dy * iby);
    dz = dz + bz - bz * fint(1.5 + dz * ibz);
    return dx * dx + dy * dy + dz * dz <= rs2;
}

// ---------------------------------------------------------------------------
// The candidate search, on the permuted slots (cell order): every atom j != i
// of the 27 (or fewer, without duplicates) cells around i's cell whose
// reference r2 is within the list radius, in a fixed order (cells z-major,
// then y, then x; atoms of one cell by reference number); sink(j, cnt) is
// called for each.  Candidates are read as their 16-byte cell-offset records
// (one load, FP32 arithmetic); only the rare pair within 1e-4 of the radius
// takes the reference's FP64 test.
//
// Atoms are sorted by cell, so the x-neighbouring cells of one (y, z) row
// are one contiguous slot range, or two where the row wraps round the box.
// The bounds of a z-plane's rows are loaded together before its first
// candidate; the plane's rows are then walked as one sequence (each row's
// second range following its first), CU candidates (a tunable) per iteration
// with their loads issued together.  The visiting order is the same as cell
// by cell.
//
// Separations: FP32 from the cell-offset records.  On a grid of at least 3
// cells a side (!SMALL) the candidate's y and z cell offsets are the row's
// and the plane's (-1, 0 or 1), and x is wrapped from the records' cells
// (sep_wrap); with fewer than 3 cells on some side (SMALL) every coordinate
// takes the records' cells and, on such a side, the rounding image.
template <int CU, bool SMALL, typename Sink>
__device__ __forceinline__ int candidates(const DevPtrs &d, const TrajParams &tp, const TrajScalars &ts,
                                          int i, const float4 *cand, bool view, const int *cslot,
                                          const RefArrays &ref,
                                          int &half, Sink sink)
{
    const double rs2 = ts.rs2;
    const double bd[3] = {tp.box[0], tp.box[1], tp.box[2]}, ibd[3] = {tp.ibox[0], tp.ibox[1], tp.ibox[2]};
    // FP32 decision band: r2 computed from the cell-offset records is within
    // ~2e-6 relative of the reference's r2 (offsets rounded once to FP32,
    // |offset| <= a cell side), so outside +-1e-4 its verdict is the
    // reference's; inside, the reference's FP64 test decides
    const float in2 = (float)(rs2 * (1.0 - 1e-4)), out2 = (float)(rs2 * (1.0 + 1e-4));
    const float4 pi = d.posf[i];
    const int ncx = ts.ncx, ncy = ts.ncy, ncz = ts.ncz;
    int cx, cy, cz;
    unpack_cell(pi.w, cx, cy, cz);
    const float cxf = (float)cx, cyf = (float)cy, czf = (float)cz;
    // FP32 grid and box constants as scalars (converted where they are set)
    const float ncxf = ts.ncf[0], ncyf = ts.ncf[1], nczf = ts.ncf[2];
    const float icx = ts.incf[0], icy = ts.incf[1], icz = ts.incf[2];
    const float sx = ts.csf[0], sy = ts.csf[1], sz = ts.csf[2];
    const float bxf = tp.boxf[0], byf = tp.boxf[1], bzf = tp.boxf[2];
    const float ibxf = tp.iboxf[0], ibyf = tp.iboxf[1], ibzf = tp.iboxf[2];
    const int lx = ncx >= 3 ? -1 : 0, hx = ncx >= 3 ? 1 : ncx - 1;
    const int ly = ncy >= 3 ? -1 : 0, hy = ncy >= 3 ? 1 : ncy - 1;
    const int lz = ncz >= 3 ? -1 : 0, hz = ncz >= 3 ? 1 : ncz - 1;
    // the x run cx+lx .. cx+hx (cyclic) as runs [xa0, xa1] then [xb0, xb1]
    int xa0 = cx + lx, xa1 = cx + hx, xb0 = 0, xb1 = -1;
    if (xa0 < 0) {
        xa0 = xa0 + ncx;
        xb1 = xa1;
        xa1 = ncx - 1;
    } else if (xa1 >= ncx) {
        xb1 = xa1 - ncx;
        xa1 = ncx - 1;
    }
    const int *cstart = d.cell_first + tp.cbase;
    int cnt = 0;
#pragma unroll 1
    for (int oz = lz; oz <= hz; ++oz) {
        int zz = cz + oz;
        zz += (zz < 0) ? ncz : (zz >= ncz ? -ncz : 0);
        // per row q: first range [A[q], A[q] + la[q]), then the rest from
        // B[q] + la[q] (absolute slots); e0, e1, n: the plane's row ends.
        // All twelve bounds are loaded before any is used (a missing row or
        // second range reads a harmless in-range start and gets length 0),
        // so the plane costs one round trip.
        const bool wrap = xb1 >= xb0;
        const int xs0 = wrap ? xb0 : 0, xs1 = wrap ? xb1 : -1;
        int sa[3], ea[3], sb[3], eb[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int oy = ly + q <= hy ? ly + q : 0;
            int yy = cy + oy;
            yy += (yy < 0) ? ncy : (yy >= ncy ? -ncy : 0);
            const int row = (zz * ncy + yy) * ncx;
            sa[q] = cstart[row + xa0];
            ea[q] = cstart[row + xa1 + 1];
            sb[q] = cstart[row + xs0];
            eb[q] = cstart[row + xs1 + 1];
        }
        // The plane's three rows walked as one sequence k = 0 .. n-1 (lanes
        // whose rows differ in length stay busy across row ends): six
        // segments, row q's first range then its second, ending at T[0..4]
        // and n; candidate k is slot k + off[segment], the segment being the
        // number of ends <= k (an empty segment's end equals the next one's
        // start, and the later select wins).
        int T[5], off[6], n = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const bool ok = ly + q <= hy;
            const int la = ok ? ea[q] - sa[q] : 0;
            const int lb = ok && wrap ? eb[q] - sb[q] : 0;
            off[2 * q] = tp.abase + sa[q] - n;
            off[2 * q + 1] = tp.abase + sb[q] - la - n;
            if (q < 2) T[2 * q] = n + la;
            n += la + lb;
            if (q < 2) T[2 * q + 1] = n;
            else T[4] = n - lb;
        }
        const float mozf = -(float)oz;  // wrapped cell difference cz_i - cz_j of the plane
        for (int k0 = 0; k0 < n; k0 += CU) {
            int jj[CU];
            float4 pj[CU];
            float moyf[CU];
#pragma unroll
            for (int u = 0; u < CU; ++u) {
                const int k = min(k0 + u, n - 1);
                const bool g1 = k >= T[1], g3 = k >= T[3];  // past row 0, past row 1
                int o = off[0];
                o = k >= T[0] ? off[1] : o;
                o = g1 ? off[2] : o;
                o = k >= T[2] ? off[3] : o;
                o = g3 ? off[4] : o;
                o = k >= T[4] ? off[5] : o;
                // the candidate's record and slot (its position in the cell
                // order, or the slot the view holds there)
                jj[u] = k + o;
                pj[u] = cand[jj[u]];
                if (view) jj[u] = cslot[jj[u]];
                moyf[u] = g3 ? -1.0f : (g1 ? 0.0f : 1.0f);  // rows oy = -1, 0, 1 (!SMALL)
            }
#pragma unroll
            for (int u = 0; u < CU; ++u) {
                if (k0 + u >= n) break;
                const int j = jj[u];
                if (j == i) continue;
                float fx, fy, fz;
                if (SMALL) {
                    const float cxj = cell_xf(pj[u].w), cyj = cell_yf(pj[u].w), czj = cell_zf(pj[u].w);
                    fx = ncx < 3 ? sep_small(cxf, cxj, sx, pi.x - pj[u].x, bxf, ibxf)
                                 : sep_wrap(cxf, cxj, ncxf, icx, sx, pi.x - pj[u].x);
                    fy = ncy < 3 ? sep_small(cyf, cyj, sy, pi.y - pj[u].y, byf, ibyf)
                                 : sep_wrap(cyf, cyj, ncyf, icy, sy, pi.y - pj[u].y);
                    fz = ncz < 3 ? sep_small(czf, czj, sz, pi.z - pj[u].z, bzf, ibzf)
                                 : sep_wrap(czf, czj, nczf, icz, sz, pi.z - pj[u].z);
                } else {
                    fx = sep_wrap(cxf, cell_xf(pj[u].w), ncxf, icx, sx, pi.x - pj[u].x);
                    fy = __builtin_fmaf(moyf[u], sy, pi.y - pj[u].y);
                    fz = __builtin_fmaf(mozf, sz, pi.z - pj[u].z);
                }
                const float r2f = __builtin_fmaf(fx, fx, __builtin_fmaf(fy, fy, fz * fz));
                bool in = r2f <= in2;
                if (!in && r2f <= out2) in = reference_pair_test(ref, i, j, rs2, bd, ibd);
                if (in) {
                    sink(j, cnt);
                    ++cnt;
                    half += (i < j);
                }
            }
        }
    }
    return cnt;
}

// The list, in one pass: search, write each atom's neighbours to its lwidth
// slots, count the pairs and the widest atom, then copy the block's permuted
// arrays back.  An atom with more than lwidth
// neighbours stops the trajectory with LISTFULL (the reference stops when its
// list array is nearly full, neighbor-pbc.f:89-93); its count is clamped so
// that no later kernel reads past its slots.  The 2-candidate form is held to
// 8 wavefronts a SIMD (its 60 VGPRs allow it; its scalar registers, 106,
// would hold it to 7): it waits on its loads more than it issues.
template <int CU, typename LT>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(CU == 2 ? 8 : 1))) void k_nbr(DevPtrs d)
{
    __shared__ int red[2];
    ONE_ITEM(d.wl_blk, WL_BLK, b);
    {
        if (threadIdx.x == 0) red[0] = red[1] = 0;
        __syncthreads();
        const int t = d.blk_traj[b];
        const TrajParams &tp = d.tp[t];
        TrajScalars &ts = d.ts[t];
        const int i = b * BLK + threadIdx.x;
        const int abase = tp.abase;
        const bool live = i - abase < tp.natom;
        const int width = tp.lwidth;
        int cnt = 0, half = 0;
        // a re-sorting rebuild searches the permuted records and arrays
        // (slot = position in the cell order); any other the view
        const bool move = ts.move;
        const float4 *cand = move ? d.posf : d.cpos;
        const RefArrays ref = move ? RefArrays{d.x2, d.y2, d.z2, d.oid2} : RefArrays{d.x, d.y, d.z, d.oid};
        if (live) {
            LT *out = list_base<LT>(d) + list_slot(tp, i - abase);
            // plain (cached) stores: each lane appends at its own position, so
            // one store instruction covers parts of many list rows, and the L2
            // merges those pieces into whole lines (nontemporal stores sent the
            // pieces on one by one: the rebuild took 10 % longer).  The store
            // address advances a row per entry (sink is called with k = 0, 1, ...)
            auto sink = [&](int j, int k) {
                if (k < width) {
                    *out = (LT)(j - abase);
                    out += TILE;
                }
            };
            if (ts.ncx < 3 || ts.ncy < 3 || ts.ncz < 3)
                cnt = candidates<CU, true>(d, tp, ts, i, cand, !move, d.cslot, ref, half, sink);
            else
                cnt = candidates<CU, false>(d, tp, ts, i, cand, !move, d.cslot, ref, half, sink);
        }
        if (cnt > width) {
            ts.err = ERR_LISTFULL;
            cnt = width;
        }
        d.nn[i] = cnt;
        // block max and sum through LDS atomics (integers: order-free), which
        // keep no shuffle lane addresses live across the search
        atomicMax(&red[0], cnt);
        atomicAdd(&red[1], half);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (red[0] > 0) atomicMax(&ts.nlist, red[0]);
            if (red[1]) atomicAdd(&ts.npairs, red[1]);
        }
        if (!move) return;
        // copy back: every slot of the block takes its permuted arrays (no
        // thread of this kernel reads the primary ones); streamed
        // (nontemporal), and every load issued before any store
#define FCMD_LD(a) __builtin_nontemporal_load(d.a##2 + i)
#define FCMD_ST(v, a) __builtin_nontemporal_store(v, d.a + i)
        const double x = FCMD_LD(x), y = FCMD_LD(y), z = FCMD_LD(z);
        const double vx = FCMD_LD(vx), vy = FCMD_LD(vy), vz = FCMD_LD(vz);
        const int typ = FCMD_LD(typ), o = d.oid2[i];
        const bool f32 = d.axf != nullptr;
        float fa0 = 0.f, fa1 = 0.f, fa2 = 0.f;
        double da0 = 0.0, da1 = 0.0, da2 = 0.0;
        if (f32) {
            fa0 = FCMD_LD(axf), fa1 = FCMD_LD(ayf), fa2 = FCMD_LD(azf);
        } else {
            da0 = FCMD_LD(ax), da1 = FCMD_LD(ay), da2 = FCMD_LD(az);
        }
        FCMD_ST(x, x), FCMD_ST(y, y), FCMD_ST(z, z);
        FCMD_ST(vx, vx), FCMD_ST(vy, vy), FCMD_ST(vz, vz);
        if (f32) {
            FCMD_ST(fa0, axf), FCMD_ST(fa1, ayf), FCMD_ST(fa2, azf);
        } else {
            FCMD_ST(da0, ax), FCMD_ST(da1, ay), FCMD_ST(da2, az);
        }
        FCMD_ST(typ, typ);
#undef FCMD_LD
#undef FCMD_ST
        d.oid[i] = o;
        if (live) d.inv[abase + o - 1] = i;
    }
}

// ---------------------------------------------------------------------------
constexpr int REBUILD_LOOP_GRID = 2048;
// The cell sort: a thread a cell, and a dilute vapour has about two cells an
// atom, so a step that rebuilds most trajectories gives it millions of
// cells, each a short chain of dependent loads; 8192 workgroups keep more of
// those chains in flight than the rebuild's other loops need (4096-8192
// measured best: rebuild -2.5 us a step against 1024).
constexpr int CELLSORT_GRID = 8192;

void launch_rebuild(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s, const KTimer &kt,
                    bool after_pre)
{
    const KTimer first{kt.t0, nullptr}, last{nullptr, kt.t1};
    const dim3 blk(BLK);
    const int wg = g.wl_grid;  // workgroups of the cell-count scans (a tunable)
    const dim3 gc(xcd_grid(std::min(g.ncblk, wg)));
    const dim3 gj(xcd_grid(std::min(g.ntraj, wg)));
    // a single trajectory's k_traj_pre has written the work lists already
    const bool lists = after_pre && g.ntraj == 1;
    if (!lists) launch_timed(k_wl_fill, dim3(g.ntraj), blk, s, first, d, g.ntraj);
    // The list search: one workgroup per possible atom block (it needs the
    // registers a loop over blocks would keep).  The binning, the scatter and
    // the permutation: REBUILD_LOOP_GRID workgroups looping over the rebuilt
    // blocks (one round of resident wavefronts), which in a step where few
    // trajectories rebuild spares the dispatch of a workgroup per atom block
    // of the arena that has nothing to do.
    const dim3 gall(xcd_grid(g.nblk));
    const dim3 gl(xcd_grid(std::min(g.nblk, REBUILD_LOOP_GRID)));
    launch_timed(k_bin, gl, blk, s, lists ? first : KTimer{}, d);
    hipLaunchKernelGGL(k_scan_chunk, gc, blk, 0, s, d);
    hipLaunchKernelGGL(k_scan_totals, gj, blk, 0, s, d);
    hipLaunchKernelGGL(k_scatter, gl, blk, 0, s, d);
    hipLaunchKernelGGL(k_cellsort, dim3(xcd_grid(std::min(CS_PARTS * g.ncblk, CELLSORT_GRID))), blk, 0, s, d);
    if (mixed)
        hipLaunchKernelGGL(k_permute<float>, gl, blk, 0, s, d);
    else
        hipLaunchKernelGGL(k_permute<double>, gl, blk, 0, s, d);
    if (g.list16) {
        if (g.nbr_cu == 4)
            launch_timed(k_nbr<4, unsigned short>, gall, blk, s, last, d);
        else
            launch_timed(k_nbr<2, unsigned short>, gall, blk, s, last, d);
    } else {
        if (g.nbr_cu == 4)
            launch_timed(k_nbr<4, int>, gall, blk, s, last, d);
        else
            launch_timed(k_nbr<2, int>, gall, blk, s, last, d);
    }
}

}  // namespace fcmd
