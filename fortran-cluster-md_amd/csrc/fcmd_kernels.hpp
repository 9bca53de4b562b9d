// fcmd_kernels.hpp -- host-side launchers of the device kernels.
#pragma once

#include <hip/hip_ext.h>

#include "fcmd_dev.hpp"

namespace fcmd {

constexpr int SCAN_CHUNK = 4 * BLK;  // cells per scan chunk

enum { FORCE_STEP = 0, FORCE_ACCATOM = 1 };

// which atom blocks a force launch takes: all; those of trajectories that
// keep their list this step; those of trajectories that rebuild (work list)
enum { FP_ALL = 0, FP_KEPT = 1, FP_REBUILT = 2 };

struct LaunchDims {
    int nblk;    // atom blocks
    int ncblk;   // cell chunks
    int nhblk;   // histogram chunks
    int ntraj;
    int list16;  // list entries are 16-bit (DevPtrs::list16)
    int kahan;   // mixed path: FP32 per-atom force sums with Kahan compensation (FCMD_MIXED_KAHAN)
    // tunables (none changes a result bit; fcmd_autotune picks them):
    int force_u;  // neighbours per load batch in the force kernel: 2, 4 or 8
    int nbr_cu;   // candidates per load batch in the list search: 2 or 4
    int wl_grid;  // workgroups of the rebuild's cell-count scans: 1024, 2048 or 4096
};

// Events recorded at a kernel's start and end (hipExtLaunchKernelGGL), for
// timing it alone; a launch sequence records t0 at its first kernel's start
// and t1 at its last one's end.  None: ordinary launches.
struct KTimer {
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

// a kernel launch that records kt's events, if any
template <typename F, typename... Args>
inline void launch_timed(F kern, dim3 grid, dim3 blk, hipStream_t s, const KTimer &kt, Args... args)
{
    if (kt.t0 || kt.t1)
        hipExtLaunchKernelGGL(kern, grid, blk, 0, s, kt.t0, kt.t1, 0, args...);
    else
        hipLaunchKernelGGL(kern, grid, blk, 0, s, args...);
}

// launch_drift = launch_drift_atoms (k_drift) + launch_traj_pre (k_traj_pre);
// launch_force = launch_force_atoms (k_force) + launch_traj_post (k_traj_post).
// energy: the force launch also sums the potential energy and the virial, and
// k_traj_post sets epot, virial and pressure from them; a run step whose
// scalars nobody can read before the next step overwrites them (every step of
// a run but its last and the last of each captured graph) leaves them out.
void launch_drift(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s);
void launch_drift_atoms(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s, const KTimer &kt = {});
void launch_traj_pre(const DevPtrs &d, const LaunchDims &g, hipStream_t s, const KTimer &kt = {});
void launch_force(const DevPtrs &d, const LaunchDims &g, bool mixed, bool single_species, int mode,
                  hipStream_t s, bool energy = true);
void launch_force_atoms(const DevPtrs &d, const LaunchDims &g, bool mixed, bool single_species, int mode,
                        hipStream_t s, const KTimer &kt = {}, int part = FP_ALL, bool energy = true);
void launch_traj_post(const DevPtrs &d, const LaunchDims &g, int mode, hipStream_t s, const KTimer &kt = {},
                      bool energy = true);
// applies any pending Nose-Hoover final update and clears the flags
void launch_kick_nh(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s);
// after_pre: the step's k_traj_pre ran just before (for a single trajectory
// it wrote the work lists, so k_wl_fill is left out)
void launch_rebuild(const DevPtrs &d, const LaunchDims &g, bool mixed, hipStream_t s, const KTimer &kt = {},
                    bool after_pre = false);
// after positions of trajectory t are replaced from the host (mixed path)
void launch_refresh_posf(const DevPtrs &d, int t, int natom, hipStream_t s);
// per_atom: also the reference-order per-atom outputs (o_cl, o_cs, o_cek, o_cgs,
// o_clf, o_ccf); the histograms and the largest clusters are always made
void launch_cluster(const DevPtrs &d, const LaunchDims &g, bool fre, bool per_atom, hipStream_t s);
// after launch_cluster: sizer-style cluster-size counts of this analysis (each
// trajectory at its own step number)
void launch_size_counts(const DevPtrs &d, const LaunchDims &g, int nmax, int warmup, hipStream_t s);

}  // namespace fcmd
