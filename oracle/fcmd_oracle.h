/*
 * fcmd_oracle.h -- CPU restatement of the reference's LJ cluster-nucleation
 * step (janw76/fortran-cluster-md, cluster/).  TEST INFRASTRUCTURE ONLY: the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it as
 * the checker.  Nothing in the product links or calls it.
 *
 * Every routine follows one reference file, operation by operation, in FP64
 * with no FMA contraction, so that on the same inputs it reproduces the
 * reference bit for bit (pinned against the compiled reference by the golden
 * replays and test_oracle_vs_compiled_reference in tests/test_oracle.py).
 * Atom indices in the neighbour list are 1-based and headers are negative,
 * exactly as in COMMON /ATOMP/ atnlist.
 */
#ifndef FCMD_ORACLE_H
#define FCMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXSORTS 10

/* scalar state: the COMMON-block scalars the hot path reads or writes */
typedef struct {
    int natom, nsorts;
    double box[3];                              /* BOXPP boxx, boxy, boxz */
    double mas[ORC_MAXSORTS], sig[ORC_MAXSORTS], eps[ORC_MAXSORTS];
    int num[ORC_MAXSORTS];                      /* apnum */
    double atrcut, atrskin, atskin, dt, Q, att; /* ATOMP / TIMEP */
    int nbtim, btim, atnenum, ethst_nh;         /* ethst_nh: ethst == 'N' */
    double dmmh[256], dmms, dmmsu, dmmax, dmmav;/* DMMPP */
    int dmmst, dmmi;
    double psi, s, ekin, epot, sumv2;           /* ATOMP psi, s ; ENERG */
    int cnum, cnumf, cbiggest, cbiggestf, clookfor, clookforf, cstep;
    int nh_iters;                               /* Newton iterations used */
    int npairs;                                 /* pairs in the last list */
    /* pair virial and pressure of the last AccAtom (for-a-lj-n-pbc.f:106,
     * 114-117, commented out in the reference) */
    double virial, pressure;
    /* ORC_ACC_TABLE only: eps and sigma of each species pair, a*10+b */
    double eps_ab[ORC_MAXSORTS * ORC_MAXSORTS], sig_ab[ORC_MAXSORTS * ORC_MAXSORTS];
} orc_state;

/* per-atom arrays; sizes are natom unless marked [0..natom] */
typedef struct {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;
    const int *api;
    int *atnlist;  long list_cap;               /* reference list format */
    int *atnidx;
    int *cl, *cs, *cc, *csi;                    /* csi [0..natom] */
    double *cek, *ceks;                         /* ceks [0..natom] */
    int *clf, *csf, *ccf, *csif;                /* csf, csif [0..natom] */
    double *cekf, *ceksf;                       /* [0..natom] */
    int *cgs, *cgsf, *cgsfa;                    /* cgsf, cgsfa [0..natom] */
    int *cfpt, *cfptf;
} orc_arrays;

enum { ORC_OK = 0, ORC_ERR_SKIN = 1, ORC_ERR_LISTFULL = 2, ORC_ERR_NH = 3 };
enum { ORC_NB_PBC = 0, ORC_NB_MAV = 1 };                  /* neighbor */
/* AccAtom: for-a-lj-n-pbc.f, for-a-lj-n-pbc-m.f, and ORC_ACC_TABLE, the
 * -m form with eps and sigma of every species pair from eps_ab/sig_ab instead
 * of Lorentz-Berthelot (an extension: the reference has no such file) */
enum { ORC_ACC_1 = 0, ORC_ACC_M = 1, ORC_ACC_TABLE = 2 };
enum { ORC_INT_VV = 0, ORC_INT_MAV = 1, ORC_INT_NH = 2 }; /* integrator */
enum { ORC_ST_N = 0, ORC_ST_NM = 1, ORC_ST_FRE = 2 };     /* stoddard */

int orc_neighbor(int variant, orc_state *st, orc_arrays *a, int *step);
int orc_accatom(int variant, orc_state *st, orc_arrays *a);
int orc_integrator(int ivar, int nbvar, int accvar, orc_state *st,
                   orc_arrays *a, int *step);
int orc_stoddard(int variant, orc_state *st, orc_arrays *a, int step);

#ifdef __cplusplus
}
#endif
#endif
