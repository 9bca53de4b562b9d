// fcmd_dev.hpp -- device-side data layout of the LJ cluster-nucleation engine.
//
// One allocation holds an ensemble of independent trajectories (a single
// system is an ensemble of one).  Trajectory k owns
//   atoms  [abase, abase + npad)           npad = natom rounded up to BLK
//   cells  [cbase, cbase + ncpad)          enough for the smallest cell side
//   list   [lbase, lbase + npad * lstride) sliced-ELL neighbour list: per
//          64-atom tile (one wavefront) lstride slots x 64 lanes, slot-major
// so that every thread block belongs to exactly one trajectory and every
// per-trajectory reduction is a fixed-order sum over its own blocks.
//
// Atoms are kept sorted by cell (re-sorted at each list rebuild); oid holds
// the reference atom number (1-based, as in COMMON /ATOMC/) of each slot.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fcmd {

constexpr int BLK = 256;    // atoms per block (4 wavefronts of 64)
constexpr int TILE = 64;    // atoms per neighbour-list tile (one wavefront)
constexpr int MAXS = 10;    // species, const.inc:12 mxsorts
constexpr int EMAX = 48;    // cluster edges (r <= atskin) kept per atom
constexpr int NPART = 5;    // per-block partial sums
constexpr int NXCD = 8;     // XCDs of one MI355X, each with its own L2
// per coordinate in a packed cell: a byte, which one v_cvt_f32_ubyte turns
// into a float (the pair loops work with cell coordinates as floats)
constexpr int CELL_BITS = 8;
constexpr int CELL_MAX = (1 << CELL_BITS) - 1;  // cells per box side at most

// XCD-aware block order.  Workgroup IDs are handed to the XCDs round-robin;
// logical block (p % NXCD) * chunk + p / NXCD for workgroup p gives every XCD
// one contiguous run of blocks.  Atoms are sorted by cell, so the neighbour
// gathers of a block then hit the L2 that has just loaded its neighbours in
// space, instead of missing into memory from another XCD.  Launch
// xcd_grid(n) workgroups for n logical blocks; ids >= n have no block.
__host__ __device__ inline int xcd_block(int p, int n)
{
    const int chunk = (n + NXCD - 1) / NXCD;
    return (p % NXCD) * chunk + p / NXCD;
}
__host__ __device__ inline int xcd_grid(int n) { return (n + NXCD - 1) / NXCD * NXCD; }

// plugin variants (which reference file each entry point follows)
enum { INT_VV = 0, INT_MAV = 1, INT_NH = 2 };     // int-verlet.f, -mav.f, -nh-mav.f
enum { NB_PBC = 0, NB_MAV = 1 };                  // neighbor-pbc.f, neighbor-pbc-mav.f
enum { ACC_1 = 0, ACC_M = 1, ACC_TABLE = 2 };     // for-a-lj-n-pbc.f, -m.f, explicit table
enum { ST_N = 0, ST_NM = 1, ST_FRE = 2 };         // stoddard-pbc-n.f, -n-m.f, -n-m-fre.f

enum { ERR_OK = 0, ERR_SKIN = 1, ERR_LISTFULL = 2, ERR_NH = 3, ERR_EDGES = 4, ERR_CELLS = 5 };

// partial-sum slots
enum { P_DMAX = 0, P_SUMV2 = 1, P_EPOT = 2, P_EKS = 3, P_VIR = 4 };

constexpr double CIKB = 1.38066e-6;  // const.inc:17, Boltzmann constant (internal units)

struct TrajParams {
    double box[3], ibox[3];
    double atrcut, atskin, atr, eps1; // cutoff, cluster radius, sigma and eps of species 1
    double dt, Q, att, nkt;          // nkt = 3.0d0*natom*cikb*att (int-verlet-nh-mav.f:65)
    double mas[MAXS];
    // 1/mas[a] and mas[a]/mas[0], divided on the host (IEEE division, so the
    // same doubles a kernel would get) to keep FP64 divisions out of the kernels
    double imas[MAXS], mrel[MAXS];
    // pair constants, index a*MAXS+b (0-based species):
    //   c1 = sigma^6, c2 = c1*c1, c3 = 0.5*c1, e48 = 48 eps, e4 = 4 eps
    double c1[MAXS * MAXS], c2[MAXS * MAXS], c3[MAXS * MAXS];
    double e48[MAXS * MAXS], e4[MAXS * MAXS];
    float c1f[MAXS * MAXS], c2f[MAXS * MAXS], c3f[MAXS * MAXS];
    float sig2f;  // sigma^2 of species 1 (the single-species mixed pair sum)
    // FP32 forms for the mixed pair loops, converted on the host so that the
    // kernels read them as scalars: box sides, their inverses, and the cutoff
    // decision band atrcut^2 (1 -+ 1e-5)
    float boxf[3], iboxf[3], cut_in2f, cut_out2f;
    int natom, nsorts, nbtim, btim, atnenum, ethst_nh;
    int integrator, neighbor, accatom, stoddard;
    int abase, npad, bfirst, nblk;   // atoms, atom blocks
    int cbase, ncpad, cfirst, ncblk; // cells, cell blocks
    long long lbase;                 // neighbour-list segment: npad * lstride slots
    int lwidth;                      // list slots per atom (its maximum neighbour count)
    int lstride;                     // list slots allocated per atom: lwidth rounded up to LIST_ROUND
    int hbase;                       // cluster-size histograms, natom+1 entries
    int nh_mask;                     // species bits coupled to Nose-Hoover
    int nh_refmass;                  // 1: m1 for every coupled atom (reference)
    int resort;                      // every resort-th list rebuild re-sorts the atoms by cell
};

// work lists of the rebuild
enum { WL_BLK = 0, WL_CCHK = 1, WL_TRAJ = 2, WL_LISTS = 3 };

// the cell sort: a cell of up to CS_REG atoms is sorted by one thread, a
// larger one (a liquid droplet's cells hold 15-30) by a wavefront
constexpr int CS_REG = 8;

// Every atom owns lstride list slots, lwidth rounded up to a multiple of
// LIST_ROUND: the force kernel reads its list in batches that may run past
// the count, and with whole batches allocated no such load needs a bound
// check (what the slots past the count hold is never used).
constexpr int LIST_ROUND = 16;

// first list slot of trajectory-relative atom slot rel; its k-th neighbour is
// at + k*TILE (the 64 lanes of a tile read 256 contiguous bytes per slot)
__host__ __device__ inline long long list_slot(const TrajParams &tp, int rel)
{
    return tp.lbase + (long long)(rel / TILE) * tp.lstride * TILE + (rel & (TILE - 1));
}

// the list array of entry type LT
template <typename LT>
__host__ __device__ inline LT *list_base(const struct DevPtrs &d);

// per-trajectory results of one cluster analysis
struct ClusterScalars {
    double cekf0;                    // cekf(0): KE of dense but not liquid-like A atoms
    int npos;                        // Stillinger clusters with cs > 0
    int cbig;                        // largest cs of this call
    int nliq;                        // liquid-like atoms (= cnumf)
    int cbigf;                       // largest tWF cluster of this call
    int want;                        // analyse this trajectory
    int pad;
};

struct TrajScalars {
    double atrskin, psi, s, ekin, epot, sumv2;
    double dmmh[256];
    double dmms, dmmsu, dmmax, dmmav;
    double dmmsr;                    // dmms when the last rebuild was decided
    // pair virial W = sum over pairs of r_ij . f_ij and the pressure
    // 1e13/V (N k T - W/3) of for-a-lj-n-pbc.f:106,114-117 (commented out
    // there), from the last force evaluation
    double virial, pressure;
    double rs2;                      // list radius^2 of the current list
    double cside[3];                 // cell side used by the current list
    double icside[3];                // 1 / cside
    // FP32 forms of the current cell grid for the mixed pair loops (scalar
    // operands there): cells a side, 1/cells to 1 ulp, cell sides
    float ncf[3], incf[3], csf[3];
    int dmmst, dmmi, step, rebuild, err, npairs, nrebuild, nlist;
    int ncx, ncy, ncz, pad0;
    int nh_iters;
    int pending;                     // v holds v_h: the NH kick (1+dth psi) is still to apply
    int move;                        // this rebuild re-sorts the atoms (else: a sorted view)
    int pad1;
};

// Device pointers of one engine.  Accelerations are double (FP64 path) or
// float (mixed path); both sets exist, only one is used.
struct DevPtrs {
    double *x, *y, *z, *vx, *vy, *vz;
    double *ax, *ay, *az;
    float *axf, *ayf, *azf;
    int *typ, *oid;
    // alternate buffers for the permutation at rebuild
    double *x2, *y2, *z2, *vx2, *vy2, *vz2, *ax2, *ay2, *az2;
    float *axf2, *ayf2, *azf2;
    int *typ2, *oid2;
    int *inv;                        // inv[abase + oid - 1] = slot
    // per block / per trajectory
    int nblk;                        // atom blocks of the whole arena
    int *blk_traj;                   // atom block -> trajectory
    int *cblk_traj;                  // cell block -> trajectory
    TrajParams *tp;
    TrajScalars *ts;
    double *part;                    // NPART per atom block
    // rebuild work lists: the atom blocks, cell chunks and trajectories of the
    // trajectories that rebuild now; lengths in wl_cnt[WL_LISTS]
    int *wl_blk, *wl_cchk, *wl_traj, *wl_cnt;
    int *wl_size;                    // WL_LISTS per trajectory: its item counts
    // cells
    int *cell_cnt, *cell_start, *sorted;
    int *cell_first;                 // final trajectory-relative cell starts (cell sort -> list search)
    int *atom_cell, *atom_slot;      // per slot at binning: packed cell (CELL_BITS each), slot in cell
    long long *scan_tmp_ll;          // chunk totals for segmented scans
    // neighbour list: nn[slot] entries at list[list_slot(tp, rel) + k*TILE],
    // each the neighbour's trajectory-relative slot: 16-bit (list16) when
    // every trajectory has at most 65536 slots (half the list traffic), else
    // 32-bit (list32); the other pointer is null
    int *nn;
    unsigned short *list16;
    int *list32;
    // per slot: the FP32 offset of the atom from the centre of its cell at the
    // last list build, and that cell packed in w (make_posf).  Written at each
    // list build (the list search reads it); on the mixed path also kept
    // current by the drift (the force kernel's positions).
    float4 *posf;
    int *cellid;                     // mixed path: the packed cell of posf.w, alone (the drift reads it)
    // the list search's candidates in cell order: at a rebuild that keeps
    // the atoms in their slots, cpos[k] = posf[cslot[k]] for the k-th atom of
    // the cell order (a re-sorting rebuild reads posf itself, slot = k)
    float4 *cpos;
    int *cslot;
    // clusters: per slot
    ClusterScalars *cls;
    int *edeg, *edeg1, *edges, *parent, *parentf, *liq;
    int *cs_tmp, *cgs_tmp, *csf_tmp;
    double *clpart;                  // per atom block: monomers, their KE, clusters, largest (cluster finish -> sums)
    double *cek_tmp, *cekf_tmp;
    // clusters: per reference atom (abase + oid - 1)
    int *liqo, *rank;
    int *o_cl, *o_cs, *o_cgs, *o_clf, *o_ccf;
    double *o_cek;
    // clusters: per size (hbase + size)
    int *o_csi, *o_csf, *o_csif, *o_cgsf, *o_cgsfa;
    double *o_ceks, *o_cekf, *o_ceksf;
    int *o_hacc;                     // cluster-size counts over analyses (hbase + size)
    int *hacc_stop;                  // per trajectory: counting has stopped
};

template <>
__host__ __device__ inline unsigned short *list_base<unsigned short>(const DevPtrs &d)
{
    return d.list16;
}
template <>
__host__ __device__ inline int *list_base<int>(const DevPtrs &d)
{
    return d.list32;
}

// the largest trajectory (in slots) whose list entries fit 16 bits
constexpr long long LIST16_MAX = 65536;

}  // namespace fcmd
