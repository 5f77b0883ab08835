/*
 * mpm_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see mpm_oracle.h).
 *
 * A restatement, in plain C, of the reference's explicit MPM step.  Every
 * function cites the reference lines it follows and keeps the reference's
 * evaluation order (so that, built without FMA contraction, it reproduces the
 * reference's glm/FP64 rounding as far as glm's own order is known).
 *
 * Pinning (what fixes each function independently of this restatement; the
 * checks live in tests/test_oracle.py):
 *   or_bases / or_find_adjacent : partition of unity, linear reproduction,
 *                                 hand-computed corner cases (sign(0)=0).
 *   or_elastic                  : Hooke's law against an independent
 *                                 Lame-constant stiffness matrix.
 *   or_perfectly_plastic        : sqrt(3 J2) == sigma_y after a plastic step,
 *                                 elastic branch == or_elastic, pressure kept.
 *   or_vonmises_hardening       : surface sqrt(J2) == k + alpha/sqrt(2),
 *                                 alpha increment closed form, reduces to
 *                                 perfectly plastic when C0 == 0.
 *   or_neohookean               : F = I gives 0; uniaxial stretch closed form.
 *   or_j2                       : uniaxial sigma^2/3, pure shear tau^2,
 *                                 hydrostatic 0, 1/2 dev:dev of the tensor
 *                                 (numpy) and rotation invariance.
 *   or_det3                     : numpy.linalg.det on random matrices.
 *   or_timeline_velocity        : interpolation closed form, outside = 0.
 *   or_step_plain               : mass / momentum invariants, free flight,
 *                                 uniform-stress patch, 1-point hand step,
 *                                 committed golden fixtures; phase 5 damping
 *                                 (one-point closed form with and without
 *                                 velocity), the y reaction into Force_Temp
 *                                 (bottom-node share of the weights), phase 6
 *                                 platen override (point sharing a platen
 *                                 point's cell moves with the platen, keeps
 *                                 v and F).
 *   or_step_contact             : with no surface points it reduces to a
 *                                 single-field step with sqrt(3) mass gates
 *                                 (checked against or_step_plain); contact
 *                                 correction conserves node momentum; kernel
 *                                 density special values (w(0) = 1, w(1) = 0,
 *                                 w(1/2) = 1/2, cell centre w(sqrt(3)/2));
 *                                 kernel gradient exact on a linear field,
 *                                 zero on a constant one, and the x - 1 row
 *                                 wrap at ix = 0.
 *   glm rounding order (determinant, outerProduct, mat*mat): parity unpinned
 *   beyond this restatement -- the reference cannot be compiled here (glm is
 *   absent), so no reference build exists to compare ULPs against.
 */
#include "mpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double sgn(double a) { return (double)((0.0 < a) - (a < 0.0)); }

/* GridPoint_Factory.h:23-32 (linear index) and :33-67 (node positions) */
static inline int nidx(int ix, int iy, int iz, const int nn[3])
{
    return ix + iy * nn[0] + iz * (nn[0] * nn[1]);
}

void or_node_position(const double L[3], const int ncell[3], int ix, int iy, int iz,
                      double out[3])
{
    /* createGrid: dx = i_x * d3Length[0]/i3Cells[0]  -> (i*L)/n */
    out[0] = (double)ix * L[0] / (double)ncell[0];
    out[1] = (double)iy * L[1] / (double)ncell[1];
    out[2] = (double)iz * L[2] / (double)ncell[2];
}

static inline void cell_size(const double L[3], const int ncell[3], double h[3])
{
    /* d3_Length_Cell = d3_Length_Grid / dvec3(i3_Cells)  (…PolygonCell.cpp:12) */
    h[0] = L[0] / (double)ncell[0];
    h[1] = L[1] / (double)ncell[1];
    h[2] = L[2] / (double)ncell[2];
}

/* GridPoint_Mediator::findAdjacentGridPoints, GridPoint_Mediator.h:31-62.
 * Returns the number of in-range nodes written to idx_out (x-outer, z-inner).
 * Points outside [0, ncell*h) on any axis are reported by the caller as an
 * out-of-grid error (the reference would reuse stale cache entries there). */
int or_find_adjacent(const double L[3], const int ncell[3], const double x[3],
                     int idx_out[8])
{
    double h[3];
    int nn[3] = {ncell[0] + 1, ncell[1] + 1, ncell[2] + 1};
    cell_size(L, ncell, h);
    int i0[3];
    for (int a = 0; a < 3; ++a) i0[a] = (int)(x[a] / h[a]);   /* trunc toward 0 */
    int n = 0;
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int iz = 0; iz < 2; ++iz) {
                int j[3] = {i0[0] + ix, i0[1] + iy, i0[2] + iz};
                if (j[0] < 0 || j[1] < 0 || j[2] < 0) continue;
                if (j[0] > nn[0] - 1 || j[1] > nn[1] - 1 || j[2] > nn[2] - 1) continue;
                idx_out[n++] = nidx(j[0], j[1], j[2], nn);
            }
    return n;
}

/* GridPoint_Mediator::calculateBases, GridPoint_Mediator.h:64-88 */
void or_bases(const double h[3], const double xp[3], const double xi[3],
              double *N, double G[3])
{
    double d[3] = {xp[0] - xi[0], xp[1] - xi[1], xp[2] - xi[2]};
    double s[3];
    for (int a = 0; a < 3; ++a) {
        s[a] = 1.0 - fabs(d[a]) / h[a];
        if (s[a] < 0.0 || s[a] > 1.0) s[a] = 0.0;
    }
    *N = s[0] * s[1] * s[2];
    G[0] = -s[1] * s[2] * sgn(d[0]) / h[0];
    G[1] = -s[0] * s[2] * sgn(d[1]) / h[1];
    G[2] = -s[0] * s[1] * sgn(d[2]) / h[2];
}

/* ConstitutiveRelation_calculateIncrement_Elastic.cpp:4-15 */
void or_elastic(double E, double nu, const double de[6], double ds[6])
{
    double C = E / (1.0 + nu) / (1.0 - 2.0 * nu);
    ds[0] = C * ((1.0 - nu) * de[0] + nu * de[1] + nu * de[2]);
    ds[1] = C * ((1.0 - nu) * de[1] + nu * de[0] + nu * de[2]);
    ds[2] = C * ((1.0 - nu) * de[2] + nu * de[0] + nu * de[1]);
    ds[3] = C * ((0.5 - nu) * de[3]);
    ds[4] = C * ((0.5 - nu) * de[4]);
    ds[5] = C * ((0.5 - nu) * de[5]);
}

/* ConstitutiveRelation_calculateState_J2.cpp:4-26 (J2 of a Voigt 6-vector) */
double or_j2(const double s[6])
{
    double j2 = 0.0;
    j2 += 1.0 / 6.0 * (s[0] - s[1]) * (s[0] - s[1]);
    j2 += 1.0 / 6.0 * (s[1] - s[2]) * (s[1] - s[2]);
    j2 += 1.0 / 6.0 * (s[2] - s[0]) * (s[2] - s[0]);
    j2 += s[3] * s[3];
    j2 += s[4] * s[4];
    j2 += s[5] * s[5];
    return j2;
}

/* shared body of …PerfectlyPlastic.cpp:4-72 and …VonMisesHardening.cpp:4-86:
 * radius = k (perfectly plastic) or k + R (hardening). Returns 1 if plastic. */
static int radial_return(double E, double nu, double radius, const double s[6],
                         const double de[6], double ds[6], double dep[6])
{
    double dst[6], st[6], dev[6];
    or_elastic(E, nu, de, dst);
    for (int i = 0; i < 6; ++i) st[i] = s[i] + dst[i];
    double p = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
    for (int i = 0; i < 6; ++i) dev[i] = st[i];
    dev[0] -= p; dev[1] -= p; dev[2] -= p;
    double j2 = or_j2(st);
    if (j2 <= radius * radius) {
        for (int i = 0; i < 6; ++i) { ds[i] = dst[i]; dep[i] = 0.0; }
        return 0;
    }
    for (int i = 0; i < 6; ++i) dev[i] *= radius / sqrt(j2);
    for (int i = 0; i < 6; ++i) ds[i] = dev[i] - s[i];
    ds[0] += p; ds[1] += p; ds[2] += p;
    double G = 0.5 * E / (1.0 + nu);
    double dee[6];
    dee[0] = ds[0] / E - ds[1] * nu / E - ds[2] * nu / E;
    dee[1] = ds[1] / E - ds[0] * nu / E - ds[2] * nu / E;
    dee[2] = ds[2] / E - ds[0] * nu / E - ds[1] * nu / E;
    dee[3] = ds[3] / G;
    dee[4] = ds[4] / G;
    dee[5] = ds[5] / G;
    for (int i = 0; i < 6; ++i) dep[i] = de[i] - dee[i];
    return 1;
}

/* ConstitutiveRelation_calculateIncrement_PerfectlyPlastic.cpp:4-72 */
void or_perfectly_plastic(double E, double nu, double sy, const double s[6],
                          const double de[6], double ds[6], double dep[6])
{
    double k0 = sy / sqrt(3.0);
    radial_return(E, nu, k0, s, de, ds, dep);
}

/* ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:4-86 */
void or_vonmises_hardening(double E, double nu, double sy, double alpha, double C0,
                           double C1, const double s[6], const double de[6],
                           double ds[6], double dep[6], double *dalpha)
{
    double k0 = sy / sqrt(3.0);
    double R = alpha / sqrt(2.0);
    *dalpha = 0.0;
    if (!radial_return(E, nu, k0 + R, s, de, ds, dep)) return;
    double eff = 0.0;
    for (int i = 0; i < 6; ++i) eff += dep[i] * dep[i];
    eff *= 2.0 / 3.0;
    eff = sqrt(eff);
    if ((sqrt(2.0 / 3.0) * C1 - alpha) > 0.0)
        *dalpha = C0 * (sqrt(2.0 / 3.0) * C1 - alpha) * eff;
    else
        *dalpha = 0.0;
}

/* glm determinant of a column-major 3x3 (used at …_MP.cpp:301, …NeoHookean.cpp:9) */
double or_det3(const double m[9])
{
#define M(c, r) m[3 * (c) + (r)]
    return +M(0, 0) * (M(1, 1) * M(2, 2) - M(2, 1) * M(1, 2))
           - M(1, 0) * (M(0, 1) * M(2, 2) - M(2, 1) * M(0, 2))
           + M(2, 0) * (M(0, 1) * M(1, 2) - M(1, 1) * M(0, 2));
#undef M
}

/* ConstitutiveRelation_calculateIncrement_NeoHookean.cpp:4-22 */
void or_neohookean(double E, double nu, const double s[6], const double F[9],
                   double ds[6])
{
    double mu = 0.5 * E / (1.0 + nu);
    double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double J = or_det3(F);
    double a = lam * log(J) / J;
    double b = mu / J;
    double P[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            /* (F * transpose(F))[c][r] = sum_k F[k][r] * F[k][c] */
            double fft = F[3 * 0 + r] * F[3 * 0 + c] + F[3 * 1 + r] * F[3 * 1 + c] +
                         F[3 * 2 + r] * F[3 * 2 + c];
            double id = (c == r) ? 1.0 : 0.0;
            P[3 * c + r] = a * id + b * (fft - id);
        }
    ds[0] = P[0] - s[0];
    ds[1] = P[4] - s[1];
    ds[2] = P[8] - s[2];
    ds[3] = P[3 * 0 + 1] - s[3];
    ds[4] = P[3 * 1 + 2] - s[4];
    ds[5] = P[3 * 2 + 0] - s[5];
}

/* TimeLine::getVelocity, TimeLine.h:33-46 */
void or_timeline_velocity(int n, const double *t, const double *v, double time,
                          double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        if (t[i] <= time && time < t[i + 1]) {
            double sc = (time - t[i]) / (t[i + 1] - t[i]);
            for (int a = 0; a < 3; ++a)
                out[a] = (1.0 - sc) * v[3 * i + a] + sc * v[3 * (i + 1) + a];
        }
    }
}

static long long nnodes_of(const or_world *w)
{
    return (long long)(w->ncell[0] + 1) * (w->ncell[1] + 1) * (w->ncell[2] + 1);
}

/* phase 2 of both steps: neighbours and bases for every point
 * (…_MP.cpp:72-94, …_MP_Contact.cpp:94-115).  Returns OR_E_OUT_OF_GRID when a
 * point is outside the grid box. */
static int build_agp(or_world *w)
{
    double h[3];
    int nn[3] = {w->ncell[0] + 1, w->ncell[1] + 1, w->ncell[2] + 1};
    cell_size(w->L, w->ncell, h);
    for (int p = 0; p < w->n_mp; ++p) {
        const double *x = &w->x[3 * p];
        for (int a = 0; a < 3; ++a)
            if (!(x[a] >= 0.0) || !((int)(x[a] / h[a]) < w->ncell[a])) return OR_E_OUT_OF_GRID;
        int idx[8];
        int n = or_find_adjacent(w->L, w->ncell, x, idx);
        if (n != 8) return OR_E_OUT_OF_GRID;
        for (int k = 0; k < 8; ++k) {
            int I = idx[k];
            w->gactive[I] = 1;
            int iz = I / (nn[0] * nn[1]);
            int iy = (I - iz * nn[0] * nn[1]) / nn[0];
            int ix = I - iz * nn[0] * nn[1] - iy * nn[0];
            double xi[3];
            or_node_position(w->L, w->ncell, ix, iy, iz, xi);
            w->agp_idx[8 * p + k] = I;
            or_bases(h, x, xi, &w->agp_N[8 * p + k], &w->agp_G[24 * p + 3 * k]);
        }
    }
    return OR_OK;
}


/* ------------------------------------------------------------------------ */
/* constitutive dispatch + state update, …_MP.cpp:318-353 (…_MP_Contact.cpp:596-631) */
static void constitutive(or_world *w, int p, const double de[6])
{
    double ds[6] = {0}, dep[6] = {0}, da = 0.0;
    double E = w->E[p], nu = w->nu[p], sy = w->sy[p];
    double *s = &w->stress[6 * p];
    for (int i = 0; i < 6; ++i) w->strain[6 * p + i] += de[i];
    int t = w->type[p];
    if (t == OR_ELASTIC)
        or_elastic(E, nu, de, ds);
    else if (t == OR_PLASTIC)
        or_perfectly_plastic(E, nu, sy, s, de, ds, dep);
    else if (t == OR_VONMISESHARDENING)
        or_vonmises_hardening(E, nu, sy, w->alpha[p], w->C0[p], w->C1[p], s, de, ds, dep, &da);
    else if (t == OR_NEOHOOKEAN && w->neohookean_dispatch)
        or_neohookean(E, nu, s, &w->F[9 * p], ds);
    else
        or_perfectly_plastic(E, nu, sy, s, de, ds, dep);
    for (int i = 0; i < 6; ++i) s[i] += ds[i];
    for (int i = 0; i < 6; ++i) w->eps_p[6 * p + i] += dep[i];
    w->alpha[p] += da;
}

/* kinematics after the node loop: F update, volume, strain increment
 * (…_MP.cpp:299-312) */
static void kinematics(or_world *w, int p, const double L[9], double dt, double de[6])
{
    double *F = &w->F[9 * p];
    double T[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            T[3 * c + r] = L[3 * 0 + r] * F[3 * c + 0] + L[3 * 1 + r] * F[3 * c + 1] +
                           L[3 * 2 + r] * F[3 * c + 2];
    for (int i = 0; i < 9; ++i) F[i] += T[i] * dt;
    double det = or_det3(F);
    w->vol[p] = det * w->vol0[p];
    if (w->strain_form == 0) {
        double Fi[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Fi[3 * c + r] = (c == r ? 1.0 : 0.0) + L[3 * c + r] * dt;
        de[0] = Fi[0] - 1.0;
        de[1] = Fi[4] - 1.0;
        de[2] = Fi[8] - 1.0;
        de[3] = Fi[3 * 0 + 1] + Fi[3 * 1 + 0];
        de[4] = Fi[3 * 1 + 2] + Fi[3 * 2 + 1];
        de[5] = Fi[3 * 2 + 0] + Fi[3 * 0 + 2];
    } else {
        de[0] = L[0] * dt;
        de[1] = L[4] * dt;
        de[2] = L[8] * dt;
        de[3] = L[3 * 0 + 1] * dt + L[3 * 1 + 0] * dt;
        de[4] = L[3 * 1 + 2] * dt + L[3 * 2 + 1] * dt;
        de[5] = L[3 * 2 + 0] * dt + L[3 * 0 + 2] * dt;
    }
}

static void grid_zero_active(or_world *w)
{
    long long nn = nnodes_of(w);
    for (long long I = 0; I < nn; ++I) {
        if (!w->gactive[I]) continue;
        w->gactive[I] = 0;
        w->gm[I] = 0.0;
        for (int a = 0; a < 3; ++a) { w->gv[3 * I + a] = 0.0; w->gf[3 * I + a] = 0.0; w->gft[3 * I + a] = 0.0; }
    }
}

/* PhysicsEngine::runSimulation_Classic_SinglePass_MP, one step (…_MP.cpp:31-370) */
/* phases 1-6 of the plain step: the grid of a step from the current points */
static int plain_grid(or_world *w, double dt)
{
    long long nn = nnodes_of(w);
    int T = w->nthreads > 1 ? w->nthreads : 1;
    /* phase 1: reset active nodes (…_MP.cpp:41-63) */
    grid_zero_active(w);
    /* phase 2: neighbours + bases (…_MP.cpp:72-94) */
    int rc = build_agp(w);
    if (rc) return rc;
    /* phase 3: P2G mass, skipping displacement-controlled points (…_MP.cpp:106-145) */
    if (T == 1) {
        for (int p = 0; p < w->n_mp; ++p) {
            if (w->dc[p]) continue;
            for (int k = 0; k < 8; ++k) w->gm[w->agp_idx[8 * p + k]] += w->agp_N[8 * p + k] * w->mass[p];
        }
    } else {
        double *tg = w->thread_grid;
        memset(tg, 0, sizeof(double) * (size_t)T * (size_t)nn * 7);
#ifdef _OPENMP
#pragma omp parallel num_threads(T)
#endif
        {
#ifdef _OPENMP
            int t = omp_get_thread_num();
#pragma omp for schedule(static)
#else
            int t = 0;
#endif
            for (int p = 0; p < w->n_mp; ++p) {
                if (w->dc[p]) continue;
                double *tm = tg + (size_t)t * nn * 7;
                for (int k = 0; k < 8; ++k) tm[w->agp_idx[8 * p + k]] += w->agp_N[8 * p + k] * w->mass[p];
            }
        }
        for (long long I = 0; I < nn; ++I)
            if (w->gactive[I])
                for (int t = 0; t < T; ++t) w->gm[I] += tg[(size_t)t * nn * 7 + I];
    }
    /* phase 4: P2G velocity + force (…_MP.cpp:151-202) */
    {
        double *tg = w->thread_grid;
#ifdef _OPENMP
#pragma omp parallel num_threads(T) if (T > 1)
#endif
        {
#ifdef _OPENMP
            int t = omp_get_thread_num();
#pragma omp for schedule(static)
#else
            int t = 0;
#endif
            for (int p = 0; p < w->n_mp; ++p) {
                if (w->dc[p]) continue;
                double *tv = (T == 1) ? w->gv : tg + (size_t)t * nn * 7 + nn;
                double *tf = (T == 1) ? w->gf : tg + (size_t)t * nn * 7 + 4 * nn;
                const double *s = &w->stress[6 * p];
                double vol = w->vol[p];
                for (int k = 0; k < 8; ++k) {
                    int I = w->agp_idx[8 * p + k];
                    double N = w->agp_N[8 * p + k];
                    const double *G = &w->agp_G[24 * p + 3 * k];
                    if (w->gm[I] > w->m_min)
                        for (int a = 0; a < 3; ++a)
                            tv[3 * I + a] += N * (w->mass[p] * w->v[3 * p + a]) / w->gm[I];
                    tf[3 * I + 0] += -vol * (G[0] * s[0] + G[1] * s[3] + G[2] * s[5]);
                    tf[3 * I + 1] += -vol * (G[1] * s[1] + G[0] * s[3] + G[2] * s[4]);
                    tf[3 * I + 2] += -vol * (G[2] * s[2] + G[0] * s[5] + G[1] * s[4]);
                    for (int a = 0; a < 3; ++a) tf[3 * I + a] += N * w->fext[3 * p + a];
                }
            }
        }
        if (T > 1)
            for (long long I = 0; I < nn; ++I)
                if (w->gactive[I])
                    for (int t = 0; t < T; ++t)
                        for (int a = 0; a < 3; ++a) {
                            w->gv[3 * I + a] += tg[(size_t)t * nn * 7 + nn + 3 * I + a];
                            w->gf[3 * I + a] += tg[(size_t)t * nn * 7 + 4 * nn + 3 * I + a];
                        }
    }
    /* phase 5: grid momentum, damping, BC, reaction (…_MP.cpp:209-239) */
    for (long long I = 0; I < nn; ++I) {
        if (!w->gactive[I]) continue;
        double *v = &w->gv[3 * I], *f = &w->gf[3 * I];
        double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (sqrt(vv) > 1.0e-9) {
            double lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
            double inv = 1.0 / sqrt(vv);
            double sc = w->damping * lf;
            for (int a = 0; a < 3; ++a) f[a] -= sc * (v[a] * inv);
        }
        if (w->gm[I] > w->m_min)
            for (int a = 0; a < 3; ++a) v[a] += f[a] / w->gm[I] * dt;
        unsigned char fx = w->fixed[I];
        if (fx & OR_FIX_X) { v[0] = 0.0; f[0] = 0.0; }
        if (fx & OR_FIX_Y) { v[1] = 0.0; w->gft[3 * I + 1] += f[1]; f[1] = 0.0; }
        if (fx & OR_FIX_Z) { v[2] = 0.0; f[2] = 0.0; }
    }
    /* phase 6: displacement-controlled override (…_MP.cpp:245-265) */
    for (int j = 0; j < w->n_dc; ++j) {
        int p = w->dc_list[j];
        int idx[8];
        int n = or_find_adjacent(w->L, w->ncell, &w->x[3 * p], idx);
        for (int k = 0; k < n; ++k) {
            int I = idx[k];
            for (int a = 0; a < 3; ++a) { w->gv[3 * I + a] = w->v[3 * p + a]; w->gf[3 * I + a] = 0.0; }
        }
    }
    (void)nn;
    return OR_OK;
}

int or_grid_plain(or_world *w, double dt)
{
    return plain_grid(w, dt);
}

int or_step_plain(or_world *w, double dt)
{
    int T = w->nthreads > 1 ? w->nthreads : 1;
    int rc = plain_grid(w, dt);
    if (rc) return rc;
    /* phase 7: G2P + F + volume + strain + constitutive (…_MP.cpp:271-354) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(T) schedule(static) if (T > 1)
#endif
    for (int p = 0; p < w->n_mp; ++p) {
        if (w->dc[p]) continue;
        double L[9] = {0};
        double *v = &w->v[3 * p], *x = &w->x[3 * p];
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            double N = w->agp_N[8 * p + k];
            const double *G = &w->agp_G[24 * p + 3 * k];
            const double *vI = &w->gv[3 * I], *fI = &w->gf[3 * I];
            double m = w->gm[I];
            if (fabs(m) > w->m_min)
                for (int a = 0; a < 3; ++a) v[a] += N * (fI[a] / m) * dt;
            for (int a = 0; a < 3; ++a) x[a] += N * vI[a] * dt;
            if (w->u)
                for (int a = 0; a < 3; ++a) w->u[3 * p + a] += N * vI[a] * dt;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) L[3 * c + r] += vI[r] * G[c];
        }
        double de[6];
        kinematics(w, p, L, dt, de);
        constitutive(w, p, de);
    }
    /* phase 8: platen velocity from the TimeLine and motion (…_MP.cpp:360-368) */
    for (int j = 0; j < w->n_dc; ++j) {
        int p = w->dc_list[j];
        if (w->n_tl != 0) or_timeline_velocity(w->n_tl, w->tl_t, w->tl_v, w->time, &w->v[3 * p]);
        for (int a = 0; a < 3; ++a) w->x[3 * p + a] += w->v[3 * p + a] * dt;
        if (w->u)
            for (int a = 0; a < 3; ++a) w->u[3 * p + a] += w->v[3 * p + a] * dt;
    }
    return OR_OK;
}

static inline double len3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* glm::length of the 3-vector mass (m, m, m) used by the contact path */
static inline double len_m3(double m) { return sqrt(m * m + m * m + m * m); }

/* PhysicsEngine::runSimulation_Classic_SinglePass_MP_Contact, one step
 * (…_MP_Contact.cpp:30-646).  The reference stores the point mass as a
 * 3-vector (m, m, m) here, so every mass gate compares sqrt(3)*m (len_m3).
 * Kernel grid = same lattice as the main grid (…PolygonCell.cpp:33-40). */
int or_step_contact(or_world *w, double dt)
{
    long long nn = nnodes_of(w);
    int nnv[3] = {w->ncell[0] + 1, w->ncell[1] + 1, w->ncell[2] + 1};
    double h[3];
    cell_size(w->L, w->ncell, h);
    /* C1: reset (…_MP_Contact.cpp:41-87) */
    for (long long I = 0; I < nn; ++I) {
        if (w->gactive[I]) {
            w->gactive[I] = 0;
            w->gm[I] = 0.0;
            for (int a = 0; a < 3; ++a) {
                w->gv[3 * I + a] = w->gf[3 * I + a] = w->gft[3 * I + a] = 0.0;
                w->gmg[3 * I + a] = w->gmgn[3 * I + a] = 0.0;
                w->gvn[3 * I + a] = w->gfn[3 * I + a] = 0.0;
                w->gkg[3 * I + a] = 0.0;  /* main-grid kernel-gradient field */
            }
            w->gmn[I] = 0.0;
            w->gcpos[I] = w->gcneg[I] = 0;
            w->gnear[I] = 0;
        }
        if (w->gkactive[I]) {
            w->gkactive[I] = 0;
            w->gk[I] = 0.0;
            for (int a = 0; a < 3; ++a) w->gkgrad[3 * I + a] = 0.0;
        }
    }
    double *kgrad = w->gkgrad;
    /* C2: neighbours + bases (…_MP_Contact.cpp:94-115) */
    int rc = build_agp(w);
    if (rc) return rc;
    /* C3: kernel accumulation on the kernel grid (…_MP_Contact.cpp:121-153) */
    for (int p = 0; p < w->n_mp; ++p) {
        double R = fabs(1.0 * h[0]);
        const double *x = &w->x[3 * p];
        int i0[3];
        for (int a = 0; a < 3; ++a) i0[a] = (int)(x[a] / h[a]);
        for (int ix = -1; ix < 3; ++ix)
            for (int iy = -1; iy < 3; ++iy)
                for (int iz = -1; iz < 3; ++iz) {
                    int j[3] = {i0[0] + ix, i0[1] + iy, i0[2] + iz};
                    if (j[0] < 0 || j[1] < 0 || j[2] < 0) continue;
                    if (j[0] > nnv[0] - 1 || j[1] > nnv[1] - 1 || j[2] > nnv[2] - 1) continue;
                    int I = nidx(j[0], j[1], j[2], nnv);
                    double xi[3];
                    or_node_position(w->L, w->ncell, j[0], j[1], j[2], xi);
                    double d[3] = {x[0] - xi[0], x[1] - xi[1], x[2] - xi[2]};
                    double dist = len3(d);
                    if (dist < R) {
                        w->gkactive[I] = 1;
                        double q = dist / R;
                        w->gk[I] += 1.0 - 3.0 * q * q + 2.0 * q * q * q;
                    }
                }
    }
    /* C4: kernel gradient by central differences (…_MP_Contact.cpp:159-191).
     * Neighbour lookup range-checks only the linear index
     * (GridPoint_Factory.h:27-31), so x-1 at ix=0 wraps to the previous row. */
    for (long long I = 0; I < nn; ++I) {
        if (w->gk[I] < 1.0) continue;
        if (!w->gkactive[I]) continue;
        int iz = (int)(I / (nnv[0] * nnv[1]));
        int iy = (int)((I - (long long)iz * nnv[0] * nnv[1]) / nnv[0]);
        int ix = (int)(I - (long long)iz * nnv[0] * nnv[1] - (long long)iy * nnv[0]);
        long long nb[6] = {nidx(ix - 1, iy, iz, nnv), nidx(ix + 1, iy, iz, nnv),
                           nidx(ix, iy - 1, iz, nnv), nidx(ix, iy + 1, iz, nnv),
                           nidx(ix, iy, iz - 1, nnv), nidx(ix, iy, iz + 1, nnv)};
        double g[3] = {0.0, 0.0, 0.0};
        double k = w->gk[I];
        if (nb[0] >= 0 && nb[0] < nn) g[0] += +0.5 / h[0] * (k - w->gk[nb[0]]);
        if (nb[1] >= 0 && nb[1] < nn) g[0] += -0.5 / h[0] * (k - w->gk[nb[1]]);
        if (nb[2] >= 0 && nb[2] < nn) g[1] += +0.5 / h[1] * (k - w->gk[nb[2]]);
        if (nb[3] >= 0 && nb[3] < nn) g[1] += -0.5 / h[1] * (k - w->gk[nb[3]]);
        if (nb[4] >= 0 && nb[4] < nn) g[2] += +0.5 / h[2] * (k - w->gk[nb[4]]);
        if (nb[5] >= 0 && nb[5] < nn) g[2] += -0.5 / h[2] * (k - w->gk[nb[5]]);
        for (int a = 0; a < 3; ++a) kgrad[3 * I + a] = g[a];
        if (len3(g) < 1.0 / h[0]) kgrad[3 * I] = kgrad[3 * I + 1] = kgrad[3 * I + 2] = 0.0;
    }
    /* C5: point kernel, nearest point per node, surface override (…:199-252) */
    for (int p = 0; p < w->n_mp; ++p) {
        const double *x = &w->x[3 * p];
        double *kp = &w->kernel[p], *kgp = &w->kernel_grad[3 * p];
        *kp = 0.0; kgp[0] = kgp[1] = kgp[2] = 0.0;
        int idx[8];
        int n = or_find_adjacent(w->L, w->ncell, x, idx);
        for (int q = 0; q < n; ++q) {
            int I = idx[q];
            int iz = I / (nnv[0] * nnv[1]);
            int iy = (I - iz * nnv[0] * nnv[1]) / nnv[0];
            int ix = I - iz * nnv[0] * nnv[1] - iy * nnv[0];
            double xi[3], N, G[3];
            or_node_position(w->L, w->ncell, ix, iy, iz, xi);
            or_bases(h, x, xi, &N, G);
            *kp += N * w->gk[I];
            for (int a = 0; a < 3; ++a) kgp[a] += N * kgrad[3 * I + a];
        }
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            double xi[3], d1[3], d2[3];
            int iz = I / (nnv[0] * nnv[1]);
            int iy = (I - iz * nnv[0] * nnv[1]) / nnv[0];
            int ix = I - iz * nnv[0] * nnv[1] - iy * nnv[0];
            or_node_position(w->L, w->ncell, ix, iy, iz, xi);
            const double *xn = &w->x[3 * w->gnear[I]];
            for (int a = 0; a < 3; ++a) { d1[a] = x[a] - xi[a]; d2[a] = xn[a] - xi[a]; }
            if (len3(d1) < len3(d2)) w->gnear[I] = p;
        }
        int surf = w->surface ? w->surface[p] : 0;
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            if (surf && len3(kgp) > len3(&w->gkg[3 * I]))
                for (int a = 0; a < 3; ++a) w->gkg[3 * I + a] = 1.0 * kgp[a];
        }
    }
    /* C6: contact detection (…_MP_Contact.cpp:258-282) */
    for (int p = 0; p < w->n_mp; ++p) {
        if (!(w->surface && w->surface[p])) continue;
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            if (!(w->surface[w->gnear[I]])) continue;
            if (dot3(&w->gkg[3 * I], &w->kernel_grad[3 * p]) >= 0.0) w->gcpos[I] = 1;
            else w->gcneg[I] = 1;
        }
    }
    /* C6: split P2G mass (…_MP_Contact.cpp:288-318).  Every point takes part. */
    for (int p = 0; p < w->n_mp; ++p) {
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            double N = w->agp_N[8 * p + k];
            const double *G = &w->agp_G[24 * p + 3 * k];
            if (dot3(&w->gkg[3 * I], &w->kernel_grad[3 * p]) >= 0.0) {
                w->gm[I] += N * w->mass[p];
                for (int a = 0; a < 3; ++a) w->gmg[3 * I + a] += G[a] * w->mass[p];
            } else {
                w->gmn[I] += N * w->mass[p];
                for (int a = 0; a < 3; ++a) w->gmgn[3 * I + a] += G[a] * w->mass[p];
            }
        }
    }
    /* C6: split P2G momentum + force (…_MP_Contact.cpp:324-385) */
    for (int p = 0; p < w->n_mp; ++p) {
        const double *s = &w->stress[6 * p];
        double vol = w->vol[p];
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            double N = w->agp_N[8 * p + k];
            const double *G = &w->agp_G[24 * p + 3 * k];
            int pos = dot3(&w->gkg[3 * I], &w->kernel_grad[3 * p]) >= 0.0;
            double m = pos ? w->gm[I] : w->gmn[I];
            double *tv = pos ? &w->gv[3 * I] : &w->gvn[3 * I];
            double *tf = pos ? &w->gf[3 * I] : &w->gfn[3 * I];
            if (len_m3(m) > w->m_min)
                for (int a = 0; a < 3; ++a) tv[a] += N * (w->mass[p] * w->v[3 * p + a]) / m;
            tf[0] += -vol * (G[0] * s[0] + G[1] * s[3] + G[2] * s[5]);
            tf[1] += -vol * (G[1] * s[1] + G[0] * s[3] + G[2] * s[4]);
            tf[2] += -vol * (G[2] * s[2] + G[0] * s[5] + G[1] * s[4]);
            for (int a = 0; a < 3; ++a) tf[a] += N * w->fext[3 * p + a];
        }
    }
    /* C7: two-field grid update + contact correction (…_MP_Contact.cpp:392-487) */
    for (long long I = 0; I < nn; ++I) {
        if (!w->gactive[I]) continue;
        double *v = &w->gv[3 * I], *f = &w->gf[3 * I];
        double *vn = &w->gvn[3 * I], *fn = &w->gfn[3 * I];
        double m = w->gm[I], mn = w->gmn[I];
        unsigned char fx = w->fixed[I];
        if (len3(v) > 1.0e-9) {
            double lf = len3(f), inv = 1.0 / sqrt(dot3(v, v)), sc = w->damping * lf;
            for (int a = 0; a < 3; ++a) f[a] -= sc * (v[a] * inv);
        }
        if (len_m3(m) > w->m_min) for (int a = 0; a < 3; ++a) v[a] += f[a] / m * dt;
        if (fx & OR_FIX_X) { v[0] = 0.0; f[0] = 0.0; }
        if (fx & OR_FIX_Y) { v[1] = 0.0; f[1] = 0.0; }
        if (fx & OR_FIX_Z) { v[2] = 0.0; f[2] = 0.0; }
        if (len3(vn) > 1.0e-16) {
            double lf = len3(fn), inv = 1.0 / sqrt(dot3(vn, vn)), sc = w->damping * lf;
            for (int a = 0; a < 3; ++a) fn[a] -= sc * (vn[a] * inv);
        }
        if (len_m3(mn) > w->m_min) for (int a = 0; a < 3; ++a) vn[a] += fn[a] / mn * dt;
        if (fx & OR_FIX_X) { vn[0] = 0.0; fn[0] = 0.0; }
        if (fx & OR_FIX_Y) { vn[1] = 0.0; fn[1] = 0.0; }
        if (fx & OR_FIX_Z) { vn[2] = 0.0; fn[2] = 0.0; }
        if (w->gcneg[I]) {
            double vcm[3], nrm[3], fc[3] = {0, 0, 0}, fcn[3] = {0, 0, 0};
            for (int a = 0; a < 3; ++a) vcm[a] = (m * v[a] + mn * vn[a]) / (m + mn);
            for (int a = 0; a < 3; ++a) nrm[a] = -0.5 * (w->gmg[3 * I + a] - w->gmgn[3 * I + a]);
            double inv = 1.0 / sqrt(dot3(nrm, nrm));
            for (int a = 0; a < 3; ++a) nrm[a] = nrm[a] * inv;
            double nneg[3] = {-nrm[0], -nrm[1], -nrm[2]};
            if (dot3(v, nrm) + dot3(vn, nneg) < 0.0) {
                double d1[3], d2[3];
                for (int a = 0; a < 3; ++a) { d1[a] = vcm[a] - v[a]; d2[a] = vcm[a] - vn[a]; }
                double s1 = m / dt * dot3(d1, nrm), s2 = mn / dt * dot3(d2, nneg);
                for (int a = 0; a < 3; ++a) { fc[a] = s1 * nrm[a]; fcn[a] = s2 * nneg[a]; }
            }
            if (len_m3(m) > w->m_min)
                for (int a = 0; a < 3; ++a) { f[a] += fc[a]; v[a] += fc[a] / m * dt; }
            if (len_m3(mn) > w->m_min)
                for (int a = 0; a < 3; ++a) { fn[a] += fcn[a]; vn[a] += fcn[a] / mn * dt; }
        }
    }
    /* C8: platen override of both fields with reaction (…_MP_Contact.cpp:494-518) */
    for (int j = 0; j < w->n_dc; ++j) {
        int p = w->dc_list[j];
        int idx[8];
        int n = or_find_adjacent(w->L, w->ncell, &w->x[3 * p], idx);
        for (int k = 0; k < n; ++k) {
            int I = idx[k];
            for (int a = 0; a < 3; ++a) {
                w->gv[3 * I + a] = w->v[3 * p + a];
                w->gvn[3 * I + a] = w->v[3 * p + a];
                w->gft[3 * I + a] += w->gf[3 * I + a];
                w->gft[3 * I + a] += w->gfn[3 * I + a];
                w->gf[3 * I + a] = 0.0;
                w->gfn[3 * I + a] = 0.0;
            }
        }
    }
    /* C8: split G2P + constitutive (…_MP_Contact.cpp:524-632) */
    for (int p = 0; p < w->n_mp; ++p) {
        if (w->dc[p]) continue;
        double L[9] = {0};
        double *v = &w->v[3 * p], *x = &w->x[3 * p];
        for (int k = 0; k < 8; ++k) {
            int I = w->agp_idx[8 * p + k];
            double N = w->agp_N[8 * p + k];
            const double *G = &w->agp_G[24 * p + 3 * k];
            int pos = dot3(&w->gkg[3 * I], &w->kernel_grad[3 * p]) >= 0.0;
            double m = pos ? w->gm[I] : w->gmn[I];
            const double *vI = pos ? &w->gv[3 * I] : &w->gvn[3 * I];
            const double *fI = pos ? &w->gf[3 * I] : &w->gfn[3 * I];
            if (len_m3(m) > w->m_min)
                for (int a = 0; a < 3; ++a) v[a] += N * (fI[a] / m) * dt;
            for (int a = 0; a < 3; ++a) x[a] += N * vI[a] * dt;
            if (w->u)
                for (int a = 0; a < 3; ++a) w->u[3 * p + a] += N * vI[a] * dt;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) L[3 * c + r] += vI[r] * G[c];
        }
        double de[6];
        kinematics(w, p, L, dt, de);
        constitutive(w, p, de);
    }
    /* C8: platen moves with its stored velocity (…_MP_Contact.cpp:638-644) */
    for (int j = 0; j < w->n_dc; ++j) {
        int p = w->dc_list[j];
        for (int a = 0; a < 3; ++a) w->x[3 * p + a] += w->v[3 * p + a] * dt;
        if (w->u)
            for (int a = 0; a < 3; ++a) w->u[3 * p + a] += w->v[3 * p + a] * dt;
    }
    return OR_OK;
}

/* the sub-stepping loop of runSimulation_Classic_SinglePass_MP(_Contact)
 * (…_MP.cpp:18-27,381-393) */
int or_run(or_world *w, double dt_total, int path, double *hist, int cap,
           int *n_steps, int *finished)
{
    double acc = 0.0;
    int ns = 0;
    long long nn = nnodes_of(w);
    while (acc < dt_total) {
        double dt = w->dt_max;
        if (w->dt_max > (dt_total - acc)) dt = dt_total - acc;
        acc += dt;
        int rc = path ? or_step_contact(w, dt) : or_step_plain(w, dt);
        if (rc) { if (n_steps) *n_steps = ns; return rc; }
        if (hist && ns < cap) {
            double s = 0.0;
            for (long long I = 0; I < nn; ++I) s += w->gft[3 * I + 1];
            hist[ns] = s;
        }
        ++ns;
        w->time += dt;
        w->cycle++;
    }
    if (n_steps) *n_steps = ns;
    if (finished) *finished = !(w->time < w->time_end);
    return OR_OK;
}

/* PhysicsEngine::reportConsole sums (PhysicsEngine.cpp:33-97) */
void or_report(const or_world *w, double out[8])
{
    long long nn = nnodes_of(w);
    double m = 0.0, ft[3] = {0, 0, 0}, ke = 0.0, mom[3] = {0, 0, 0};
    for (long long I = 0; I < nn; ++I) {
        m += w->gm[I];
        for (int a = 0; a < 3; ++a) ft[a] += w->gft[3 * I + a];
    }
    for (int p = 0; p < w->n_mp; ++p) {
        const double *v = &w->v[3 * p];
        ke += w->mass[p] * pow(sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), 2.0);
        for (int a = 0; a < 3; ++a) mom[a] += w->mass[p] * v[a];
    }
    out[0] = m; out[1] = ft[0]; out[2] = ft[1]; out[3] = ft[2]; out[4] = ke;
    out[5] = mom[0]; out[6] = mom[1]; out[7] = mom[2];
}
