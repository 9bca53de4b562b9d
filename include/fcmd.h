/*
 * fcmd.h -- C ABI of the MI355X LJ cluster-nucleation engine.
 *
 * Drop-in boundary for the reference's plugin subroutines (janw76/
 * fortran-cluster-md, cluster/), which are external F77 subroutines linked
 * by name and sharing state through COMMON blocks (Makefile:8-114):
 *
 *   AccAtom()          for-a-lj-n-pbc.f:14, for-a-lj-n-pbc-m.f:18  -> fcmd_accatom
 *   neighbor(step)     neighbor-pbc.f:11, neighbor-pbc-mav.f:13    -> fcmd_neighbor
 *   integrator(step)   int-verlet.f:16, int-verlet-mav.f:18,
 *                      int-verlet-nh-mav.f:13                       -> fcmd_integrator
 *   stoddard(step)     stoddard-pbc-n.f:10, stoddard-pbc-n-m.f:11,
 *                      stoddard-pbc-n-m-fre.f:11                    -> fcmd_stoddard
 *
 * The Fortran shims in fortran-cluster-md_amd/fortran/ keep those names and
 * argument lists and call these functions through ISO_C_BINDING.  All arrays
 * are passed in the reference's order (atom i of COMMON /ATOMC/ is element
 * i-1); units are the reference's internal units (Angstrom, fs, 1e-27 kg,
 * 1e-17 J, K).  State stays resident on the GPU between calls; fcmd_upload /
 * fcmd_download move it.  Every function returns 0 or an FCMD_ERR_* code;
 * fcmd_error_text gives the reference's message for it.
 *
 * One fcmd_system holds one or more independent trajectories (an ensemble);
 * the drop-in uses one.  Calls on one fcmd_system must not overlap.
 */
#ifndef FCMD_H
#define FCMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FCMD_MAXSORTS 10

enum { FCMD_FP64 = 0, FCMD_MIXED = 1, FCMD_MIXED_KAHAN = 2 };
enum { FCMD_INT_VV = 0, FCMD_INT_MAV = 1, FCMD_INT_NH = 2 };
enum { FCMD_NB_PBC = 0, FCMD_NB_MAV = 1 };
enum { FCMD_ACC_1 = 0, FCMD_ACC_M = 1, FCMD_ACC_TABLE = 2 };
enum { FCMD_ST_N = 0, FCMD_ST_NM = 1, FCMD_ST_FRE = 2 };

enum {
    FCMD_OK = 0,
    FCMD_ERR_SKIN = 1,      /* neighbor-pbc*.f: "Skin radius exceeded!!" */
    FCMD_ERR_LISTFULL = 2,  /* neighbor-pbc*.f: "Neighborlist is nearly full!" */
    FCMD_ERR_NH = 3,        /* int-verlet-nh-mav.f:151 "did not converge" */
    FCMD_ERR_EDGES = 4,     /* more cluster edges per atom than the engine keeps */
    FCMD_ERR_CELLS = 5,     /* cell grid larger than allocated */
    FCMD_ERR_SCRATCH = 6,   /* fcmd_run_noalloc: the run needs more than FCMD_NOALLOC_INTS */
    FCMD_ERR_ARG = 10,
    FCMD_ERR_HIP = 11,
    FCMD_ERR_NODEVICE = 12
};

typedef struct fcmd_system fcmd_system;

/* Parameters of one trajectory: the COMMON values the hot path reads. */
typedef struct {
    int natom, nsorts;                     /* ATOMP natom, atsorts */
    double box[3];                         /* BOXPP boxx, boxy, boxz */
    double mas[FCMD_MAXSORTS];             /* apmas */
    double sig[FCMD_MAXSORTS];             /* apsig */
    double eps[FCMD_MAXSORTS];             /* apeps */
    int num[FCMD_MAXSORTS];                /* apnum */
    double atrcut, atrskin, atskin;        /* in length units (read-3d.f:259-261) */
    double dt, Q, att;                     /* TIMEP dt; ATOMP Q (internal), att */
    int nbtim, btim, atnenum, ethst_nh;    /* ethst_nh: ENERG ethst == 'N' */
    int integrator, neighbor, accatom, stoddard;   /* which reference file */
    double eps_ab[FCMD_MAXSORTS * FCMD_MAXSORTS];  /* FCMD_ACC_TABLE only */
    double sig_ab[FCMD_MAXSORTS * FCMD_MAXSORTS];
    int nh_mask;    /* species bits coupled to Nose-Hoover; 0 = every atom with m1 (reference) */
    int pad;
} fcmd_params;

/* Scalars of one trajectory (ATOMP atrskin, psi, s; ENERG ekin, epot,
 * sumv2; DMMPP), plus engine counters: dmmsr is dmms as it stood when the
 * last rebuild was decided (what neighbor-pbc-mav.f:102 prints), npairs the
 * pairs in the current list, nrebuild the rebuilds so far, nlist the largest
 * neighbour count of any atom in the current list (at most list_width).
 * virial and pressure (read only) come from the last force evaluation: the
 * pair virial W = sum over pairs of r_ij . f_ij (1e-17 J) and the pressure
 * 1e13/V (N k T - W/3) in Pa, T the target temperature att, as the
 * reference's commented-out lines for-a-lj-n-pbc.f:106,114-117 define them;
 * the two-species forms sum 48 eps_ab r^-6 (sigma_ab^12 r^-6 - sigma_ab^6/2)
 * per pair. */
typedef struct {
    double atrskin, psi, s, ekin, epot, sumv2;
    double dmmh[256];
    double dmms, dmmsu, dmmax, dmmav, dmmsr;
    int dmmst, dmmi, step, npairs, nrebuild, nh_iters, err, nlist;
    double virial, pressure;
} fcmd_scalars;

int fcmd_device_count(void);
const char *fcmd_error_text(int code);

/* Allocate an ensemble of ntraj trajectories on the current HIP device.
 * precision: FCMD_FP64 (the reference's arithmetic); FCMD_MIXED (pair
 * separations and LJ terms in FP32, each atom's pair forces summed in FP64;
 * state, integrator and reductions over atoms in FP64); FCMD_MIXED_KAHAN
 * (the same, each atom's pair forces summed in FP32 with Kahan
 * compensation); any other value is FCMD_ERR_ARG.  list_width: neighbour-list slots per
 * atom, i.e. the most neighbours (within atrskin, both directions) any atom
 * may have; 0 selects 256 (a liquid droplet at the adaptive skin's usual
 * 2.8-3.3 sigma holds 75-130).  It takes the role of const.inc mxnlist: a list
 * rebuild that finds an atom with more neighbours fails with
 * FCMD_ERR_LISTFULL, as the reference stops when its list array is nearly
 * full (neighbor-pbc.f:89-93).  Memory: 2 bytes per slot when every
 * trajectory has at most 65536 atom slots (natom rounded up to 256), else 4,
 * for list_width rounded up to a multiple of 16 slots per atom (the force
 * kernel reads whole batches of slots); the environment variable
 * FCMD_LIST_BITS=32, read
 * here, asks for the 4-byte entries regardless (the results are the same). */
int fcmd_create(fcmd_system **out, int ntraj, const fcmd_params *params, int precision,
                int list_width);
void fcmd_destroy(fcmd_system *s);

/* Replace the parameters of trajectory traj (same natom, box within the
 * allocated cell grid). */
int fcmd_set_params(fcmd_system *s, int traj, const fcmd_params *p);

/* ATOMC x,v,a and ATOMP api (species 1..) of one trajectory, reference
 * order.  Pointers may be NULL to leave that array as it is. */
int fcmd_upload(fcmd_system *s, int traj, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *ax,
                const double *ay, const double *az, const int *api);
int fcmd_download(fcmd_system *s, int traj, double *x, double *y, double *z, double *vx,
                  double *vy, double *vz, double *ax, double *ay, double *az);

int fcmd_set_scalars(fcmd_system *s, int traj, const fcmd_scalars *sc);
int fcmd_get_scalars(fcmd_system *s, int traj, fcmd_scalars *sc);

/* neighbor(step): neighbor-pbc.f:11 / neighbor-pbc-mav.f:13.  Wraps the
 * positions, applies the skin rules and rebuilds the list of trajectory traj;
 * *step is made non-negative as the reference does. */
int fcmd_neighbor(fcmd_system *s, int traj, int *step);

/* AccAtom(): for-a-lj-n-pbc.f:14 / for-a-lj-n-pbc-m.f:18.  Adds the LJ
 * accelerations to a and sets epot, for every trajectory. */
int fcmd_accatom(fcmd_system *s);

/* integrator(step): one step of every trajectory (int-verlet.f:16,
 * int-verlet-mav.f:18, int-verlet-nh-mav.f:13), including the list rebuild
 * the displacement rule asks for.  *step is this step's number. */
int fcmd_integrator(fcmd_system *s, int *step);

/* stoddard(step): stoddard-pbc-n.f:10 / -n-m.f:11 / -n-m-fre.f:11 for
 * trajectory traj.  Arrays in the CVTPP layout; histogram arrays hold
 * natom+1 entries starting at index 0.  csi[0] and ceks[0] are updated the
 * way the reference leaves them (csi(0) is never reset there).  counters =
 * {cnum, cnumf, cbiggest, cbiggestf, clookfor, clookforf} in/out.  tWF
 * arrays may be NULL unless the trajectory uses FCMD_ST_FRE. */
int fcmd_stoddard(fcmd_system *s, int traj, int step, int *cl, int *cs, int *cc, int *csi,
                  double *cek, double *ceks, int *clf, int *csf, int *ccf, int *csif,
                  double *cekf, double *ceksf, int *cgs, int *cgsf, int *cgsfa, int *cfpt,
                  int *cfptf, int *counters);

/* The current list of trajectory traj in the reference's ATOMP layout:
 * atnlist = -1, j.., -2, j.., ..., -natom (ascending j > i) and atnidx.
 * *len receives the number of entries written (<= cap). */
int fcmd_halflist(fcmd_system *s, int traj, int *atnlist, long long cap, int *atnidx,
                  long long *len);

/* Ensemble run: nsteps integrator steps of every trajectory with no host
 * decision inside (HIP-graph replay when use_graph), a Stillinger analysis
 * every cluster_every steps (0: none) recording each trajectory's largest
 * cluster in series[snapshot*ntraj + traj] (may be NULL) and the first step
 * at which it reached fpt_size in fpt[traj] (-1 if never; may be NULL).
 * The analysis at step i uses the state before step i is integrated, as the
 * reference's main loop (main.f:133-164).  It keeps the histograms and the
 * largest clusters only; fcmd_stoddard makes the per-atom arrays on demand. */
int fcmd_run(fcmd_system *s, int nsteps, int cluster_every, int fpt_size, int *series,
             int *fpt, int use_graph);

/* fcmd_run without any allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Same steps, analyses,
 * outputs and bits as fcmd_run(..., use_graph = 0): every step is launched
 * on the stream (no graph is captured or instantiated), and the run's
 * first-passage times and largest-cluster series are staged in a fixed
 * device array of FCMD_NOALLOC_INTS ints that each GPU holds once for the
 * process (fcmd_create finds its address), instead of buffers fcmd_run
 * allocates.  A run needs (ceil(nsteps / cluster_every) + 1) x ntraj ints
 * of it; a larger one returns FCMD_ERR_SCRATCH before doing anything.  The
 * array keeps the last call's contents; every call writes what it reads.
 * Calls on systems of the same GPU share the array: calls that overlap in
 * time on one GPU race and must not be made. */
#define FCMD_NOALLOC_INTS 65536
int fcmd_run_noalloc(fcmd_system *s, int nsteps, int cluster_every, int fpt_size, int *series,
                     int *fpt);

/* Seeded initial state: natom positions uniform in the box with no pair
 * closer than rmin (minimum image), Maxwell-Boltzmann velocities at
 * temperature att for masses mas[api-1], zero total momentum, rescaled to
 * exactly att (3 N k T / 2 kinetic energy, k = 1.38066e-6).  Deterministic in
 * seed.  Host only. */
int fcmd_init_random(unsigned long long seed, int natom, const double *box, double rmin,
                     double att, const double *mas, const int *api, double *x, double *y,
                     double *z, double *vx, double *vy, double *vz);

/* Cluster-size probability counts (helpers/sizer.f:59-103, Stillinger
 * histogram): from now on every analysis of fcmd_run at a step >= warmup adds
 * csi(g) to counts(g) for g = 1 up to and including the first size >= nmax
 * present in that analysis, after which the trajectory stops counting (sizer
 * stops at its first cluster of nmax or more, also before warmup).  Resets
 * the counts and the stop flags; nmax <= 0 switches counting off. */
int fcmd_size_counts(fcmd_system *s, int nmax, int warmup);

/* counts[natom+1] of trajectory traj (counts[0] = 0) and *stopped (may be
 * NULL): 1 once it has met a cluster of nmax or more. */
int fcmd_get_size_counts(fcmd_system *s, int traj, int *counts, int *stopped);

/* Time of the last fcmd_run in milliseconds (HIP events on the run stream). */
int fcmd_last_run_ms(fcmd_system *s, double *total_ms);

/* Mean time in milliseconds of one cluster analysis of the last fcmd_run
 * (the pending Nose-Hoover update, the Stillinger analysis of every
 * trajectory and the recording of the series), from HIP events around each
 * of its first 16 analyses; *count (may be NULL) receives how many were
 * timed.  No counterpart in the reference (it has no timers). */
int fcmd_last_run_analysis_ms(fcmd_system *s, double *mean_ms, int *count);

/* Run nsteps steps (no graph) and time them with HIP events on the engine
 * stream; avg_ms[FCMD_NTIMED] receives the mean milliseconds per step of
 * {k_drift, k_traj_pre, list-rebuild sequence, k_force, k_traj_post,
 * Nose-Hoover kick}.  The events are recorded by the launches themselves,
 * at the start and end of each kernel (of the rebuild sequence: its first
 * kernel's start, its last one's end), so the steps run as in a run.  The
 * kick is deferred to the next drift inside a run, so it is timed once after
 * the nsteps and its entry is that time / nsteps. */
#define FCMD_NTIMED 6
int fcmd_time_kernels(fcmd_system *s, int nsteps, double *avg_ms);

/* Launch tunables: tuning[3] = {neighbours per load batch in the force
 * kernel (2, 4, 8), candidates per load batch in the list search (2, 4),
 * workgroups of the list rebuild's cell-count scans (1024, 2048, 4096)}.  No
 * tunable changes any
 * result bit.  Defaults {4, 2, 2048}.  No counterpart in the reference (its
 * loops have no launch shape). */
int fcmd_set_tuning(fcmd_system *s, const int *tuning);
int fcmd_get_tuning(fcmd_system *s, int *tuning);

/* Autotuner: runs rounds x 6 integrator steps of every trajectory (real
 * steps: the state advances as in fcmd_run), the candidates taking turns
 * step by step, times the force kernel and the rebuild sequence of each on
 * the engine stream, and keeps the fastest; tuning (may be NULL) receives
 * the choice.  No counterpart in the reference. */
int fcmd_autotune(fcmd_system *s, int rounds, int *tuning);

/* Full-list entries (2 x pairs) summed over all trajectories' current lists. */
int fcmd_list_entries(fcmd_system *s, long long *entries);

#ifdef __cplusplus
}
#endif
#endif
