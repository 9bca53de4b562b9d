/*
 * fcmd_oracle.c -- CPU restatement of the reference LJ cluster-nucleation step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Each function follows one reference
 * routine statement by statement: same operand order, same Fortran INT()
 * truncation, same DMOD wrap, FP64 throughout and built with
 * -ffp-contract=off, so that it matches the compiled reference bit for bit.
 * Arrays are 0-based here; atom numbers stored in the lists are 1-based, as in
 * the reference.  Helper macros X(i) etc. take the 1-based atom number.
 *
 * Pinned against the compiled reference (oracle/_ref harness) by the golden
 * replays and test_oracle_vs_compiled_reference in tests/test_oracle.py:
 * lists, positions, velocities, energies, mav bookkeeping, labels and
 * histograms bit-exact on the same inputs.
 */
#include "fcmd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CIKB 1.38066e-6 /* const.inc:17 internal Boltzmann constant */

#define X(i)  a->x[(i) - 1]
#define Y(i)  a->y[(i) - 1]
#define Z(i)  a->z[(i) - 1]
#define VX(i) a->vx[(i) - 1]
#define VY(i) a->vy[(i) - 1]
#define VZ(i) a->vz[(i) - 1]
#define AX(i) a->ax[(i) - 1]
#define AY(i) a->ay[(i) - 1]
#define AZ(i) a->az[(i) - 1]
#define API(i) a->api[(i) - 1]
#define NL(k) a->atnlist[(k) - 1]
#define NIDX(i) a->atnidx[(i) - 1]

/* Fortran INT() of a double: truncation toward zero */
static inline double fint(double v) { return (double)(int)v; }

/* Fortran MAX/MIN for the non-NaN values that occur here */
static inline double fmaxd(double p, double q) { return p > q ? p : q; }
static inline double fmind(double p, double q) { return p < q ? p : q; }

/* ------------------------------------------------------------------------- */
/* list build shared by neighbor-pbc.f:62-94 and neighbor-pbc-mav.f:67-99   */
static int build_halflist(orc_state *st, orc_arrays *a)
{
    const int natom = st->natom;
    const double boxx = st->box[0], boxy = st->box[1], boxz = st->box[2];
    const double ibx = 1.0 / boxx, iby = 1.0 / boxy, ibz = 1.0 / boxz;
    const double rs2 = st->atrskin * st->atrskin;
    long ind = 1;

    /* wrap in place: dmod(dmod(x,L)+L,L) */
    for (int i = 1; i <= natom; i++) {
        X(i) = fmod(fmod(X(i), boxx) + boxx, boxx);
        Y(i) = fmod(fmod(Y(i), boxy) + boxy, boxy);
        Z(i) = fmod(fmod(Z(i), boxz) + boxz, boxz);
    }
    for (int i = 1; i <= natom; i++) {
        if (ind > a->list_cap) return ORC_ERR_LISTFULL;
        NL(ind) = -i;
        NIDX(i) = (int)ind;
        ind = ind + 1;
        const double tx = X(i), ty = Y(i), tz = Z(i);
        for (int j = i + 1; j <= natom; j++) {
            double dx = tx - X(j);
            double dy = ty - Y(j);
            double dz = tz - Z(j);
            dx = dx + boxx - boxx * fint(1.5 + dx * ibx);
            dy = dy + boxy - boxy * fint(1.5 + dy * iby);
            dz = dz + boxz - boxz * fint(1.5 + dz * ibz);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= rs2) {
                if (ind > a->list_cap) return ORC_ERR_LISTFULL;
                NL(ind) = j;
                ind = ind + 1;
            }
        }
        /* the reference stops when ind+natom > mxnlist; list_cap plays mxnlist */
        if (ind + natom > a->list_cap) return ORC_ERR_LISTFULL;
    }
    st->npairs = (int)(ind - 1 - natom);
    return ORC_OK;
}

int orc_neighbor(int variant, orc_state *st, orc_arrays *a, int *step)
{
    const double rc = st->atrcut;
    if (variant == ORC_NB_PBC) {
        /* neighbor-pbc.f:40-59 */
        if (*step < 0) *step = -*step;
        st->dmmav = st->dmmsu * (double)st->nbtim / (double)(*step - st->btim);
        if (st->dmms > (st->atrskin - rc)) return ORC_ERR_SKIN;
        if (st->dmmst == st->nbtim && st->dmms < st->dmmav &&
            st->dmmav < 0.95 * (st->atrskin - rc)) {
            st->atrskin = st->atrskin - 0.01 * (st->atrskin - rc);
        }
    } else {
        /* neighbor-pbc-mav.f:48-63 */
        if (st->dmms > (st->atrskin - rc)) return ORC_ERR_SKIN;
        st->dmmav = (double)(float)st->dmmst * st->dmmsu / 256.0;
        if (*step != st->btim && *step > 0) {
            st->atrskin = rc + (st->atrskin - rc) * (double)st->nbtim / (double)st->dmmst;
            st->atrskin = fmind(st->box[0] / 2.0, st->atrskin);
            st->atrskin = fmind(st->box[1] / 2.0, st->atrskin);
            st->atrskin = fmind(st->box[2] / 2.0, st->atrskin);
        }
        if (*step < 0) *step = -*step;
    }
    int rc_ = build_halflist(st, a);
    if (rc_ != ORC_OK) return rc_;
    /* neighbor-pbc.f:96-100 / neighbor-pbc-mav.f:101-105 */
    st->dmmax = fmaxd(st->dmms, st->dmmax);
    st->dmms = 0.0;
    st->dmmst = 0;
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* for-a-lj-n-pbc.f:54-127 (one species) and for-a-lj-n-pbc-m.f:61-152      */
int orc_accatom(int variant, orc_state *st, orc_arrays *a)
{
    const int natom = st->natom;
    const double boxx = st->box[0], boxy = st->box[1], boxz = st->box[2];
    const double atr = st->sig[0];
    const double atr2 = st->atrcut * st->atrcut;
    const double bf1 = 2.0 + fint(st->atrskin / atr);
    const double bf2 = bf1 + 0.5;
    const double ibx = 1.0 / boxx, iby = 1.0 / boxy, ibz = 1.0 / boxz;
    double esum = 0.0;
    double virial = 0.0;
    double *fx = calloc((size_t)natom, sizeof(double));
    double *fy = calloc((size_t)natom, sizeof(double));
    double *fz = calloc((size_t)natom, sizeof(double));

    /* constants: single species (c1,c2,c3) or the three LB pairs */
    const double c1 = atr * atr * atr * atr * atr * atr;
    const double c2 = c1 * c1;
    const double c3 = 0.5 * c1;
    double c111 = 0, c112 = 0, c113 = 0, c221 = 0, c222 = 0, c223 = 0;
    double c121 = 0, c122 = 0, c123 = 0, eps11 = 0, eps22 = 0, eps12 = 0;
    if (variant == ORC_ACC_M) {
        const double s1 = st->sig[0], s2 = st->sig[1];
        c111 = s1 * s1 * s1 * s1 * s1 * s1;
        c112 = c111 * c111;
        c113 = 0.5 * c111;
        eps11 = st->eps[0];
        c221 = s2 * s2 * s2 * s2 * s2 * s2;
        c222 = c221 * c221;
        c223 = 0.5 * c221;
        eps22 = st->eps[1];
        c121 = 0.5 * (s1 + s2);
        c121 = c121 * c121 * c121 * c121 * c121 * c121;
        c122 = c121 * c121;
        c123 = 0.5 * c121;
        eps12 = sqrt(st->eps[0] * st->eps[1]);
    }

    long k = 1;
    for (;;) {
        const int i = -NL(k);
        if (i == natom) break;
        k = k + 1;
        const double tx = X(i), ty = Y(i), tz = Z(i);
        for (;;) {
            const int j = NL(k);
            if (j <= 0) break;
            k = k + 1;
            double dx = tx - X(j);
            double dy = ty - Y(j);
            double dz = tz - Z(j);
            dx = dx + boxx * (bf1 - fint(bf2 + dx * ibx));
            dy = dy + boxy * (bf1 - fint(bf2 + dy * iby));
            dz = dz + boxz * (bf1 - fint(bf2 + dz * ibz));
            double rij2 = dx * dx + dy * dy + dz * dz;
            if (rij2 <= atr2) {
                rij2 = 1.0 / rij2;
                const double rij6 = rij2 * rij2 * rij2;
                double fij;
                if (variant == ORC_ACC_1) {
                    fij = rij2 * rij6 * (c2 * rij6 - c3);
                    esum = esum + rij6 * (c2 * rij6 - c1);
                    virial = virial + rij6 * (c2 * rij6 - c3); /* :106 */
                } else if (variant == ORC_ACC_M) {
                    const int ij = API(i) + API(j);
                    if (ij == 2) {
                        fij = eps11 * 48.0 * rij2 * rij6 * (c112 * rij6 - c113);
                        esum = esum + eps11 * 4.0 * rij6 * (c112 * rij6 - c111);
                        virial = virial + eps11 * 48.0 * rij6 * (c112 * rij6 - c113);
                    } else if (ij == 4) {
                        fij = eps22 * 48.0 * rij2 * rij6 * (c222 * rij6 - c223);
                        esum = esum + eps22 * 4.0 * rij6 * (c222 * rij6 - c221);
                        virial = virial + eps22 * 48.0 * rij6 * (c222 * rij6 - c223);
                    } else {
                        fij = eps12 * 48.0 * rij2 * rij6 * (c122 * rij6 - c123);
                        esum = esum + eps12 * 4.0 * rij6 * (c122 * rij6 - c121);
                        virial = virial + eps12 * 48.0 * rij6 * (c122 * rij6 - c123);
                    }
                } else {
                    /* explicit pair table: the -m expressions with the
                     * pair's own eps and sigma */
                    const int ab = (API(i) - 1) * ORC_MAXSORTS + (API(j) - 1);
                    const double sg = st->sig_ab[ab], ep = st->eps_ab[ab];
                    const double t1 = sg * sg * sg * sg * sg * sg;
                    const double t2 = t1 * t1;
                    const double t3 = 0.5 * t1;
                    fij = ep * 48.0 * rij2 * rij6 * (t2 * rij6 - t3);
                    esum = esum + ep * 4.0 * rij6 * (t2 * rij6 - t1);
                    virial = virial + ep * 48.0 * rij6 * (t2 * rij6 - t3);
                }
                fx[i - 1] = fx[i - 1] + fij * dx;
                fx[j - 1] = fx[j - 1] - fij * dx;
                fy[i - 1] = fy[i - 1] + fij * dy;
                fy[j - 1] = fy[j - 1] - fij * dy;
                fz[i - 1] = fz[i - 1] + fij * dz;
                fz[j - 1] = fz[j - 1] - fij * dz;
            }
        }
    }

    /* for-a-lj-n-pbc.f:114-117 (the -m and table forms fold 48 eps per pair) */
    if (variant == ORC_ACC_1) virial = virial * 48.0 * st->eps[0];
    st->virial = virial;
    st->pressure = (1.0e13 / (boxx * boxy * boxz)) * ((double)natom * CIKB * st->att - virial / 3.0);
    if (variant == ORC_ACC_1) {
        const double ate = st->eps[0];
        const double iatm = 1.0 / st->mas[0];
        st->epot = esum * ate * 4.0;
        for (int i = 1; i <= natom; i++) {
            AX(i) = AX(i) + ate * 48.0 * fx[i - 1] * iatm;
            AY(i) = AY(i) + ate * 48.0 * fy[i - 1] * iatm;
            AZ(i) = AZ(i) + ate * 48.0 * fz[i - 1] * iatm;
        }
    } else {
        st->epot = esum;
        for (int i = 1; i <= natom; i++) {
            const double iatm = 1.0 / st->mas[API(i) - 1];
            AX(i) = AX(i) + fx[i - 1] * iatm;
            AY(i) = AY(i) + fy[i - 1] * iatm;
            AZ(i) = AZ(i) + fz[i - 1] * iatm;
        }
    }
    free(fx);
    free(fy);
    free(fz);
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
static void zero_acc(orc_state *st, orc_arrays *a)
{
    st->epot = 0.0;
    for (int i = 1; i <= st->natom; i++) {
        AX(i) = 0.0;
        AY(i) = 0.0;
        AZ(i) = 0.0;
    }
}

/* moving-average displacement bookkeeping and rebuild trigger:
 * int-verlet-mav.f:64-74 == int-verlet-nh-mav.f:70-78 */
static int mav_trigger(orc_state *st, double dmmr)
{
    dmmr = 2.2 * sqrt(dmmr);
    st->dmms = st->dmms + dmmr;
    st->dmmsu = st->dmmsu + dmmr - st->dmmh[st->dmmi];
    st->dmmh[st->dmmi] = dmmr;
    st->dmmi = (st->dmmi + 1) & 255;
    st->dmmst = st->dmmst + 1;
    return st->dmms > (st->atrskin - st->atrcut - 2.0 * st->dmms / (double)(float)st->dmmst);
}

int orc_integrator(int ivar, int nbvar, int accvar, orc_state *st,
                   orc_arrays *a, int *step)
{
    const int natom = st->natom;
    const double dt = st->dt;
    int rc_;

    if (ivar == ORC_INT_VV) {
        /* int-verlet.f:34-59 */
        for (int i = 1; i <= natom; i++) {
            X(i) = X(i) + (VX(i) + 0.5 * AX(i) * dt) * dt;
            Y(i) = Y(i) + (VY(i) + 0.5 * AY(i) * dt) * dt;
            Z(i) = Z(i) + (VZ(i) + 0.5 * AZ(i) * dt) * dt;
            VX(i) = VX(i) + 0.5 * AX(i) * dt;
            VY(i) = VY(i) + 0.5 * AY(i) * dt;
            VZ(i) = VZ(i) + 0.5 * AZ(i) * dt;
        }
        zero_acc(st, a);
        if ((rc_ = orc_accatom(accvar, st, a)) != ORC_OK) return rc_;
        const double atm = st->mas[0];
        st->ekin = 0.0;
        for (int i = 1; i <= natom; i++) {
            VX(i) = VX(i) + 0.5 * AX(i) * dt;
            VY(i) = VY(i) + 0.5 * AY(i) * dt;
            VZ(i) = VZ(i) + 0.5 * AZ(i) * dt;
            st->ekin = st->ekin + 0.5 * atm * (VX(i) * VX(i) + VY(i) * VY(i) + VZ(i) * VZ(i));
        }
        return ORC_OK;
    }

    if (ivar == ORC_INT_MAV) {
        /* int-verlet-mav.f:44-92 */
        double dmmr = 0.0;
        for (int i = 1; i <= natom; i++) {
            const double dx = (VX(i) + 0.5 * AX(i) * dt) * dt;
            const double dy = (VY(i) + 0.5 * AY(i) * dt) * dt;
            const double dz = (VZ(i) + 0.5 * AZ(i) * dt) * dt;
            X(i) = X(i) + dx;
            Y(i) = Y(i) + dy;
            Z(i) = Z(i) + dz;
            VX(i) = VX(i) + 0.5 * AX(i) * dt;
            VY(i) = VY(i) + 0.5 * AY(i) * dt;
            VZ(i) = VZ(i) + 0.5 * AZ(i) * dt;
            dmmr = fmaxd(dmmr, dx * dx + dy * dy + dz * dz);
        }
        if (mav_trigger(st, dmmr)) {
            if ((rc_ = orc_neighbor(nbvar, st, a, step)) != ORC_OK) return rc_;
        }
        zero_acc(st, a);
        if ((rc_ = orc_accatom(accvar, st, a)) != ORC_OK) return rc_;
        st->ekin = 0.0;
        for (int i = 1; i <= natom; i++) {
            VX(i) = VX(i) + 0.5 * AX(i) * dt;
            VY(i) = VY(i) + 0.5 * AY(i) * dt;
            VZ(i) = VZ(i) + 0.5 * AZ(i) * dt;
            st->ekin = st->ekin + 0.5 * st->mas[API(i) - 1] *
                                      (VX(i) * VX(i) + VY(i) * VY(i) + VZ(i) * VZ(i));
        }
        return ORC_OK;
    }

    /* int-verlet-nh-mav.f:43-180 */
    const double dt2 = dt * dt / 2;
    const double dth = dt / 2;
    const double m1 = st->mas[0];
    const double Q = st->Q;
    const double nkt = 3.0 * natom * CIKB * st->att;     /* 3.0d0*natom*cikb*att */
    const double nkt_i = (double)(3 * natom) * CIKB * st->att; /* 3*natom*cikb*att */
    double psi = st->psi;
    double dmmr = 0.0;
    st->sumv2 = 0.0;
    for (int i = 1; i <= natom; i++) {
        const double dx = dt * VX(i) + dt2 * (AX(i) - psi * VX(i));
        const double dy = dt * VY(i) + dt2 * (AY(i) - psi * VY(i));
        const double dz = dt * VZ(i) + dt2 * (AZ(i) - psi * VZ(i));
        X(i) = X(i) + dx;
        Y(i) = Y(i) + dy;
        Z(i) = Z(i) + dz;
        st->sumv2 = st->sumv2 + VX(i) * VX(i) + VY(i) * VY(i) + VZ(i) * VZ(i);
        VX(i) = VX(i) + dth * (AX(i) - psi * VX(i));
        VY(i) = VY(i) + dth * (AY(i) - psi * VY(i));
        VZ(i) = VZ(i) + dth * (AZ(i) - psi * VZ(i));
        dmmr = fmaxd(dmmr, dx * dx + dy * dy + dz * dz);
    }
    if (st->ethst_nh) {
        st->s = st->s + psi * dt + (m1 * st->sumv2 - nkt) * dt2 / Q;
        st->psi = psi + (m1 * st->sumv2 - nkt) * dth / Q;
    }
    if (mav_trigger(st, dmmr)) {
        if ((rc_ = orc_neighbor(nbvar, st, a, step)) != ORC_OK) return rc_;
    }
    zero_acc(st, a);
    if ((rc_ = orc_accatom(accvar, st, a)) != ORC_OK) return rc_;

    /* implicit velocity / friction solve by Newton iteration :91-166 */
    const double err = 1.0e-10;
    double *vxn = malloc(sizeof(double) * (size_t)natom * 9);
    double *vyn = vxn + natom, *vzn = vyn + natom;
    double *vxo = vzn + natom, *vyo = vxo + natom, *vzo = vyo + natom;
    double *bx = vzo + natom, *by = bx + natom, *bz = by + natom;
    st->ekin = 0.0;
    st->sumv2 = 0.0;
    for (int i = 0; i < natom; i++) {
        vxn[i] = a->vx[i];
        vyn[i] = a->vy[i];
        vzn[i] = a->vz[i];
        st->sumv2 = st->sumv2 + vxn[i] * vxn[i] + vyn[i] * vyn[i] + vzn[i] * vzn[i];
    }
    const double PSI = st->psi; /* COMMON PSI after the half update */
    double psn = PSI, pso = PSI;
    int ready = 0, iter;
    for (iter = 1; iter <= 1000; iter++) {
        pso = psn;
        double delps = 0;
        for (int i = 0; i < natom; i++) {
            vxo[i] = vxn[i];
            vyo[i] = vyn[i];
            vzo[i] = vzn[i];
            bx[i] = -dth * (a->ax[i] - pso * vxo[i]) - (a->vx[i] - vxo[i]);
            delps = delps + bx[i] * m1 * vxo[i] * dt / Q;
            by[i] = -dth * (a->ay[i] - pso * vyo[i]) - (a->vy[i] - vyo[i]);
            delps = delps + by[i] * m1 * vyo[i] * dt / Q;
            bz[i] = -dth * (a->az[i] - pso * vzo[i]) - (a->vz[i] - vzo[i]);
            delps = delps + bz[i] * m1 * vzo[i] * dt / Q;
        }
        const double di = -(pso * dth + 1);
        delps = delps - di * ((-m1 * st->sumv2 + nkt_i) * dth / Q - (PSI - pso));
        delps = delps / (-dt2 * m1 * st->sumv2 / Q + di);
        st->sumv2 = 0;
        for (int i = 0; i < natom; i++) {
            vxn[i] = vxn[i] + (bx[i] + dth * vxo[i] * delps) / di;
            vyn[i] = vyn[i] + (by[i] + dth * vyo[i] * delps) / di;
            vzn[i] = vzn[i] + (bz[i] + dth * vzo[i] * delps) / di;
            st->sumv2 = st->sumv2 + vxn[i] * vxn[i] + vyn[i] * vyn[i] + vzn[i] * vzn[i];
        }
        psn = pso + delps;
        ready = 1;
        for (int i = 0; i < natom; i++) {
            if (fabs((vxn[i] - vxo[i]) / vxn[i]) > err) ready = 0;
            if (fabs((vyn[i] - vyo[i]) / vyn[i]) > err) ready = 0;
            if (fabs((vzn[i] - vzo[i]) / vzn[i]) > err) ready = 0;
        }
        if (ready) break;
    }
    st->nh_iters = ready ? iter : 1000;
    if (!ready) {
        free(vxn);
        return ORC_ERR_NH;
    }
    st->sumv2 = 0;
    for (int i = 0; i < natom; i++) {
        a->vx[i] = vxn[i];
        a->vy[i] = vyn[i];
        a->vz[i] = vzn[i];
        st->sumv2 = st->sumv2 + a->vx[i] * a->vx[i] + a->vy[i] * a->vy[i] + a->vz[i] * a->vz[i];
    }
    st->psi = psn;
    st->ekin = 0.5 * m1 * st->sumv2;
    free(vxn);
    (void)nkt;
    return ORC_OK;
}

/* ------------------------------------------------------------------------- */
/* Stillinger sweep.  The reference (stoddard-pbc-n.f:123-143) visits a
 * component by repeatedly taking the lowest-numbered atom that is labelled
 * with the current cluster number and not yet visited; a binary min-heap
 * reproduces that visit order (and so the order of the cek sums) without the
 * O(N^2) rescan. */
typedef struct { int *h; int n; } heap_t;

static void hpush(heap_t *q, int v)
{
    int c = q->n++;
    q->h[c] = v;
    while (c > 0) {
        int p = (c - 1) / 2;
        if (q->h[p] <= q->h[c]) break;
        int t = q->h[p]; q->h[p] = q->h[c]; q->h[c] = t;
        c = p;
    }
}

static int hpop(heap_t *q)
{
    int top = q->h[0];
    q->h[0] = q->h[--q->n];
    int c = 0;
    for (;;) {
        int l = 2 * c + 1, r = l + 1, m = c;
        if (l < q->n && q->h[l] < q->h[m]) m = l;
        if (r < q->n && q->h[r] < q->h[m]) m = r;
        if (m == c) break;
        int t = q->h[m]; q->h[m] = q->h[c]; q->h[c] = t;
        c = m;
    }
    return top;
}

int orc_stoddard(int variant, orc_state *st, orc_arrays *a, int step)
{
    const int natom = st->natom;
    const double boxx = st->box[0], boxy = st->box[1], boxz = st->box[2];
    const double atr = st->sig[0];
    const double rijc2 = st->atskin * st->atskin;
    const double bf1 = fint(st->atrskin / atr) + 2.0;
    const double bf2 = bf1 + 0.5;
    const double ibx = 1.0 / boxx, iby = 1.0 / boxy, ibz = 1.0 / boxz;
    const int fre = (variant == ORC_ST_FRE);

    st->cstep = step;
    for (int i = 1; i <= natom; i++) {
        a->cl[i - 1] = 0;
        a->cc[i - 1] = 0;
        a->cs[i - 1] = 0;
        a->csi[i] = 0;
        a->cek[i - 1] = 0;
        a->ceks[i] = 0;
    }
    if (fre) {
        for (int i = 1; i <= natom; i++) {
            a->clf[i - 1] = 0;
            a->ccf[i - 1] = 0;
            a->csf[i] = 0;
            a->csif[i] = 0;
            a->cekf[i] = 0;
            a->ceksf[i] = 0;
            a->cgs[i - 1] = 0;
            a->cgsf[i] = 0;
            a->cgsfa[i] = 0;
        }
        a->csif[0] = 0;
        a->cgsf[0] = 0;
        a->cgsfa[0] = 0;
        a->cekf[0] = 0.0;
        a->ceksf[0] = 0.0;
    }

    /* edges: list pairs within atskin (stoddard-pbc-n.f:69-98) as CSR */
    int *deg = calloc((size_t)natom + 1, sizeof(int));
    long nedge = 0, cap = 1024;
    int *ei = malloc(sizeof(int) * cap), *ej = malloc(sizeof(int) * cap);
    for (int i = 1; i <= natom; i++) {
        if (i == natom) break;
        if (variant == ORC_ST_NM && API(i) != 1) continue;
        const double tx = X(i), ty = Y(i), tz = Z(i);
        long k = NIDX(i);
        for (;;) {
            k = k + 1;
            const int j = NL(k);
            if (j < 0) break;
            if (variant == ORC_ST_NM && API(j) != 1) break;
            double dx = tx - X(j);
            double dy = ty - Y(j);
            double dz = tz - Z(j);
            dx = dx + boxx * (bf1 - fint(bf2 + dx * ibx));
            dy = dy + boxy * (bf1 - fint(bf2 + dy * iby));
            dz = dz + boxz * (bf1 - fint(bf2 + dz * ibz));
            const double rij2 = dx * dx + dy * dy + dz * dz;
            if (rij2 <= rijc2) {
                if (nedge == cap) {
                    cap *= 2;
                    ei = realloc(ei, sizeof(int) * cap);
                    ej = realloc(ej, sizeof(int) * cap);
                }
                ei[nedge] = i;
                ej[nedge] = j;
                nedge++;
                deg[i]++;
                deg[j]++;
            }
        }
    }
    int *off = malloc(sizeof(int) * ((size_t)natom + 2));
    off[1] = 0;
    for (int i = 1; i <= natom; i++) off[i + 1] = off[i] + deg[i];
    int *adj = malloc(sizeof(int) * (size_t)(2 * nedge + 1));
    int *fill = malloc(sizeof(int) * ((size_t)natom + 2));
    memcpy(fill, off, sizeof(int) * ((size_t)natom + 2));
    for (long e = 0; e < nedge; e++) {
        adj[fill[ei[e]]++] = ej[e];
        adj[fill[ej[e]]++] = ei[e];
    }
    free(ei);
    free(ej);
    free(fill);

    heap_t q;
    q.h = malloc(sizeof(int) * (size_t)(2 * nedge + natom + 1));

    /* ten Wolde-Frenkel liquid filter (stoddard-pbc-n-m-fre.f:148-168) */
    int *liq = NULL, nliq = 0;
    if (fre) {
        liq = malloc(sizeof(int) * ((size_t)natom + 1));
        for (int k = 1; k <= natom; k++) {
            if (API(k) == 1 && deg[k] >= st->atnenum) {
                int filter = 0;
                for (int p = off[k]; p < off[k + 1]; p++)
                    if (API(adj[p]) == 1) filter = filter + 1;
                if (filter >= st->atnenum) {
                    liq[nliq++] = k;
                } else {
                    a->cekf[0] = a->cekf[0] + 0.5 * st->mas[API(k) - 1] *
                                                  (VX(k) * VX(k) + VY(k) * VY(k) + VZ(k) * VZ(k));
                }
            }
        }
    }

    /* Stillinger clusters: label = lowest atom number of the component */
    int cnum = 0;
    for (int k = 1; k <= natom; k++) {
        cnum = cnum + 1;
        if (a->cc[k - 1] != 0) continue;
        q.n = 0;
        hpush(&q, k);
        while (q.n > 0) {
            const int j = hpop(&q);
            if (a->cc[j - 1] != 0) continue;
            a->cl[j - 1] = cnum;
            a->cc[j - 1] = 1;
            const double ke = (variant == ORC_ST_N ? 0.5 * st->mas[0] : 0.5 * st->mas[API(j) - 1]) *
                              (VX(j) * VX(j) + VY(j) * VY(j) + VZ(j) * VZ(j));
            if (fre && API(j) != 1) {
                a->cgs[cnum - 1] = a->cgs[cnum - 1] + 1;
            } else {
                a->cs[cnum - 1] = a->cs[cnum - 1] + 1;
                a->cek[cnum - 1] = a->cek[cnum - 1] + ke;
            }
            for (int p = off[j]; p < off[j + 1]; p++) {
                const int m = adj[p];
                a->cl[m - 1] = cnum;
                if (a->cc[m - 1] == 0) hpush(&q, m);
            }
        }
    }
    st->cnum = cnum;

    if (fre) {
        /* tWF clusters over the liquid-like atoms, numbered by liqmon rank */
        int *rank = calloc((size_t)natom + 1, sizeof(int));
        for (int r = 0; r < nliq; r++) rank[liq[r]] = r + 1;
        int cnumf = 0;
        for (int k = 1; k <= nliq; k++) {
            cnumf = cnumf + 1;
            const int m0 = liq[k - 1];
            if (a->ccf[m0 - 1] != 0) continue;
            q.n = 0;
            hpush(&q, k);
            while (q.n > 0) {
                const int r = hpop(&q);
                const int m = liq[r - 1];
                if (a->ccf[m - 1] != 0) continue;
                a->clf[m - 1] = cnumf;
                a->ccf[m - 1] = 1;
                if (API(m) == 1) {
                    a->csf[cnumf] = a->csf[cnumf] + 1;
                    a->cekf[cnumf] = a->cekf[cnumf] + 0.5 * st->mas[API(m) - 1] *
                                                          (VX(m) * VX(m) + VY(m) * VY(m) + VZ(m) * VZ(m));
                } else {
                    a->cgsf[cnumf] = a->cgsf[cnumf] + 1;
                }
                for (int p = off[m]; p < off[m + 1]; p++) {
                    const int nb = adj[p];
                    a->clf[nb - 1] = cnumf;
                    if (rank[nb] > k && a->ccf[nb - 1] == 0) hpush(&q, rank[nb]);
                }
            }
        }
        st->cnumf = cnumf;
        free(rank);

        /* statistics and first-passage bookkeeping (:229-258) */
        for (int i = 1; i <= cnum; i++) {
            if (a->cs[i - 1] > st->cbiggest) st->cbiggest = a->cs[i - 1];
            a->csi[a->cs[i - 1]] = a->csi[a->cs[i - 1]] + 1;
            a->ceks[a->cs[i - 1]] = a->ceks[a->cs[i - 1]] + a->cek[i - 1];
        }
        if (st->cbiggest >= st->clookfor) {
            for (int i = st->clookfor; i <= st->cbiggest; i++) a->cfpt[i - 1] = step;
            st->clookfor = st->clookfor + 1;
        }
        int catoms = 0;
        for (int i = 1; i <= cnumf; i++) {
            if (a->csf[i] > st->cbiggestf) st->cbiggestf = a->csf[i];
            a->csif[a->csf[i]] = a->csif[a->csf[i]] + 1;
            a->ceksf[a->csf[i]] = a->ceksf[a->csf[i]] + a->cekf[i];
            a->cgsfa[a->csf[i]] = a->cgsfa[a->csf[i]] + a->cgsf[i];
            catoms = catoms + a->csf[i];
        }
        a->csif[0] = st->num[0] - catoms;
        if (st->cbiggestf >= st->clookforf) {
            for (int i = st->clookforf; i <= st->cbiggestf; i++) a->cfptf[i - 1] = step;
            st->clookforf = st->clookforf + 1;
        }
        free(liq);
    } else {
        /* stoddard-pbc-n.f:145-148 (csi(0) collects the non-roots: never reset) */
        for (int i = 1; i <= cnum; i++) {
            a->csi[a->cs[i - 1]] = a->csi[a->cs[i - 1]] + 1;
            a->ceks[a->cs[i - 1]] = a->ceks[a->cs[i - 1]] + a->cek[i - 1];
        }
    }
    free(q.h);
    free(adj);
    free(off);
    free(deg);
    return ORC_OK;
}
