/*
 * mpm_oracle.h -- FP64 CPU oracle for the explicit MPM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU step, used as the checker for the HIP implementation.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library never links, loads or calls it.
 *
 * Followed reference files (line numbers per file):
 *   PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:4-394   plain step
 *   PhysicsEngine_runSimulation_Classic_SinglePass_MP_Contact.cpp:4-676  contact step
 *   GridPoint_Mediator.h:31-88,123-155   neighbour search, linear-hat bases
 *   GridPoint_Factory.h:18-67            node index and node positions
 *   ConstitutiveRelation_calculateIncrement_{Elastic,PerfectlyPlastic,
 *     VonMisesHardening,NeoHookean}.cpp, ConstitutiveRelation_calculateState_J2.cpp
 *   TimeLine.h:33-46                     platen velocity schedule
 *   PhysicsEngine.cpp:33-97              FP64 report sums
 *
 * glm semantics restated (glm is not vendored by the reference):
 *   matrices are column-major, m[col][row] is stored at m[3*col+row];
 *   outerProduct(c,r)[j][i] = c[i]*r[j]; mat*mat sums k=0,1,2 left to right;
 *   determinant is the first-column cofactor expansion; sign(0) = 0;
 *   ivec3(dvec3) truncates toward zero; normalize(v) = v*(1/sqrt(dot(v,v))).
 *
 * Pinning status of each function is listed in mpm_oracle.c's header.
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* material types, same numbering as the reference enum (MaterialPoint.h:8-16) */
enum { OR_ELASTIC = 0, OR_NEOHOOKEAN = 1, OR_VISCOELASTIC = 2, OR_PLASTIC = 3,
       OR_VONMISESHARDENING = 4, OR_GASS = 5 };

/* node fixity bits */
enum { OR_FIX_X = 1, OR_FIX_Y = 2, OR_FIX_Z = 4 };

/* status codes shared with the product C-ABI numbering */
enum { OR_OK = 0, OR_E_ARG = -1, OR_E_OUT_OF_GRID = -5 };

typedef struct or_world {
    /* ---- material points (caller-owned, n_mp entries) ---- */
    int n_mp;
    double *x, *v, *fext;          /* 3 per point                          */
    double *F;                     /* 9 per point, column-major [3*c+r]    */
    double *strain, *stress, *eps_p;  /* 6 per point, Voigt xx yy zz xy yz zx */
    double *alpha;                 /* isotropic back stress                 */
    double *mass, *vol, *vol0;
    double *E, *nu, *sy, *C0, *C1;
    const int *type;
    const unsigned char *dc;       /* displacement-controlled flag          */
    int n_dc; const int *dc_list;  /* displacement-controlled points, in order */
    /* contact-path per-point state (may be NULL for the plain path) */
    double *kernel, *kernel_grad;  /* 1 and 3 per point                      */
    const unsigned char *surface;  /* b_Surface flag, may be NULL (all false) */

    /* ---- background grid ---- */
    double L[3]; int ncell[3];
    const unsigned char *fixed;    /* per node OR_FIX_* bits                 */
    /* grid work arrays (caller-owned, nnodes entries; 3 per node for vectors) */
    double *gm, *gv, *gf, *gft; unsigned char *gactive;

    /* ---- contact path grids (caller-owned, may be NULL for plain path) ---- */
    double *gmn;                   /* negative-field mass, 1 per node         */
    double *gmg, *gmgn;            /* mass gradient, +/- field, 3 per node    */
    double *gvn, *gfn;             /* negative-field velocity, force, 3/node  */
    double *gkg;                   /* main-grid kernel-gradient field, 3/node */
    int *gnear; unsigned char *gcpos, *gcneg;
    double *gk, *gkgrad;           /* kernel grid (same lattice): value, grad */
    unsigned char *gkactive;

    /* ---- AGP cache (8 per point) ---- */
    int *agp_idx; double *agp_N, *agp_G;

    /* ---- parameters ---- */
    double dt_max, damping, m_min;
    double time, time_end;
    long long cycle;
    int n_tl; const double *tl_t, *tl_v;   /* TimeLine knots; tl_v 3 per knot */
    int nthreads;                  /* >1: privatized per-thread grids + reduction */
    double *thread_grid;           /* nthreads * nnodes * 7 doubles (m, v3, f3) */
    int strain_form;               /* 0: reference ((I+L dt)-I); 1: L dt      */
    int neohookean_dispatch;       /* 0: reference (falls to PerfectlyPlastic) */
    /* displacement u = x - x0 carried as its own state (3 per point, may be
     * NULL): the same increments as x, so that small displacements keep the
     * digits a difference of large positions loses */
    double *u;
} or_world;

/* shape function and neighbours ------------------------------------------ */
int  or_find_adjacent(const double L[3], const int ncell[3], const double x[3],
                      int idx_out[8]);
void or_bases(const double h[3], const double xp[3], const double xi[3],
              double *N, double G[3]);
void or_node_position(const double L[3], const int ncell[3], int ix, int iy, int iz,
                      double out[3]);

/* constitutive increments ------------------------------------------------ */
void or_elastic(double E, double nu, const double de[6], double ds[6]);
void or_perfectly_plastic(double E, double nu, double sy, const double s[6],
                          const double de[6], double ds[6], double dep[6]);
void or_vonmises_hardening(double E, double nu, double sy, double alpha, double C0,
                           double C1, const double s[6], const double de[6],
                           double ds[6], double dep[6], double *dalpha);
void or_neohookean(double E, double nu, const double s[6], const double F[9],
                   double ds[6]);
double or_j2(const double s[6]);
double or_det3(const double F[9]);

/* TimeLine::getVelocity -------------------------------------------------- */
void or_timeline_velocity(int n, const double *t, const double *v, double time,
                          double out[3]);

/* full steps -------------------------------------------------------------- */
/* one step of runSimulation_Classic_SinglePass_MP, without the time advance */
int or_step_plain(or_world *w, double dt);
/* phases 1-6 of that step only (…_MP.cpp:41-265): the grid arrays (gm, gv,
 * gf, gft, gactive) then hold the grid of a step taken from the current
 * points; the points are not changed */
int or_grid_plain(or_world *w, double dt);
/* one step of runSimulation_Classic_SinglePass_MP_Contact, without time advance */
int or_step_contact(or_world *w, double dt);
/* the sub-stepping loop of runSimulation_Classic_SinglePass_MP(_Contact);
 * reaction_hist[k] receives sum of node Force_Temp.y after step k (cap entries). */
int or_run(or_world *w, double dt_total, int path, double *reaction_hist, int cap,
           int *n_steps, int *finished);

/* FP64 sums of reportConsole (PhysicsEngine.cpp:33-97):
 * out[0] grid mass, out[1..3] sum Force_Temp, out[4] sum m|v|^2 (no 1/2),
 * out[5..7] momentum sum m v over all points                                 */
void or_report(const or_world *w, double out[8]);

#ifdef __cplusplus
}
#endif
#endif
