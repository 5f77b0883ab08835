/*
 * mpm_auxetics.h -- C ABI of the MI355X explicit MPM step.
 *
 * Plain C, plain pointers and sizes.  The host C++ PhysicsEngine shim
 * (code_mpm_auxetics_amd/host/PhysicsEngine.h), the ctypes test harness and
 * any other FFI bind exactly these symbols.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the reference tree Sina-mon/code_MPM_Auxetics).
 */
#ifndef MPM_AUXETICS_H
#define MPM_AUXETICS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 5

/* ---- status codes (the reference has none; it exits via fatalError,
 *      Errors.cpp:5-12) ---------------------------------------------------- */
enum mpm_status {
    MPM_OK = 0,
    MPM_E_ARG = -1,          /* bad argument / config                         */
    MPM_E_HIP = -2,          /* HIP runtime error (see mpm_last_error)        */
    MPM_E_NAN = -3,          /* non-finite particle state detected            */
    MPM_E_STATE = -4,        /* call order violated (e.g. step before upload) */
    MPM_E_OUT_OF_GRID = -5,  /* a point left the background grid box          */
    MPM_E_ARENA = -6,        /* a point moved further than one cell beyond its
                                brick between two drift checks                */
    MPM_E_NOMEM = -7,
    MPM_E_NO_DEVICE = -8
};

/* precision of the device state */
enum mpm_precision {
    MPM_FP64 = 0,            /* FP64 "plain": every field and operation FP64   */
    MPM_FP32 = 1             /* FP32 state; FP64 det(F)/log J, J2 return map,
                                global checks                                 */
};

/* how a batch of plain steps reaches the GPU: captured once into a hipGraph
 * and replayed while the brick layout is unchanged (default), or launched
 * kernel by kernel.  Results are identical.  Profiling (cfg.profile) and the
 * contact path always launch directly. */
enum mpm_launch { MPM_LAUNCH_GRAPH = 0, MPM_LAUNCH_DIRECT = 1 };

/* step paths: runSimulation_Classic_SinglePass_MP / _MP_Contact
 * (PhysicsEngine.h:45-46) */
enum mpm_path { MPM_PATH_PLAIN = 0, MPM_PATH_CONTACT = 1 };

/* material types: same numbering as the reference enum MaterialType
 * (MaterialPoint.h:8-16) */
enum mpm_material {
    MPM_ELASTIC = 0, MPM_NEOHOOKEAN = 1, MPM_VISCOELASTIC = 2,
    MPM_PLASTIC = 3, MPM_VONMISESHARDENING = 4, MPM_GASS = 5
};

/* node fixity bits: GridPoint::b3_Fixed (GridPoint.h:18) */
enum { MPM_FIX_X = 1, MPM_FIX_Y = 2, MPM_FIX_Z = 4 };

/* readback field mask bits.  MPM_F_ALL covers the MaterialPoint fields;
 * MPM_F_J2 and MPM_F_CELL are derived quantities outside it. */
enum {
    MPM_F_X = 1u << 0, MPM_F_V = 1u << 1, MPM_F_F = 1u << 2, MPM_F_STRESS = 1u << 3,
    MPM_F_STRAIN = 1u << 4, MPM_F_EPSP = 1u << 5, MPM_F_ALPHA = 1u << 6,
    MPM_F_VOL = 1u << 7, MPM_F_KERNEL = 1u << 8, MPM_F_ALL = 0x1FFu,
    MPM_F_J2 = 1u << 9,      /* j2: calculateState_J2 of the stress and of the
                                plastic strain, 2 per point (the quantities
                                GraphicsEngine_drawGame.cpp:164,178,252 colour by) */
    MPM_F_CELL = 1u << 10    /* cell (3 int32 per point): the grid cell the step
                                bins the point into; order (1 uint32 per point):
                                its position in the binned device order
                                (0xFFFFFFFF for platen points) */
};

/* Solver configuration.  Replaces the protected members that
 * PhysicsEngine::initializeWorld_* sets (PhysicsEngine.h:68-93,
 * PhysicsEngine_initializeWorld_AuxeticPolygonCell.cpp:10-40,269-274). */
typedef struct mpm_config {
    double length[3];          /* d3_Length_Grid                                */
    int32_t cells[3];          /* i3_Cells; nodes = cells + 1 per axis          */
    double dt_max;             /* d_TimeIncrement_Maximum                       */
    double time_start;         /* d_Time at creation                            */
    double time_end;           /* d_TimeEnd                                     */
    double damping;            /* d_DampingCoefficient                          */
    double mass_min;           /* d_Mass_Minimum                                */
    double console_interval;   /* d_TimeConsole_Interval (<=0: no reports)      */
    int32_t precision;         /* enum mpm_precision                            */
    int32_t strain_form;       /* 0: (I+L dt)-I as the reference (…_MP.cpp:304);
                                  1: L dt (always used in MPM_FP32)             */
    int32_t neohookean;        /* 0: _NEOHOOKEAN falls to PerfectlyPlastic as in
                                  the reference dispatch (…_MP.cpp:332-339);
                                  1: dispatch calculateIncrement_NeoHookean_6D  */
    int32_t track_strain;      /* keep the total strain d6_Strain (…_MP.cpp:318) */
    int32_t n_timeline;        /* TimeLine knots (TimeLine.h:27-31), may be 0   */
    const double *timeline_t;  /* n_timeline times                              */
    const double *timeline_v;  /* 3*n_timeline velocities                       */
    const uint8_t *node_fixed; /* (cells+1) product entries of MPM_FIX_* bits, or
                                  NULL; copied at mpm_create                    */
    int32_t brick[3];          /* cells per brick; 0 = default (4,4,10), clamped
                                  to the grid                                  */
    int32_t drift_check;       /* steps between drift checks; 0 = default (8)   */
    double friction;           /* contact path: Coulomb coefficient (extension;
                                  0 reproduces the frictionless reference)     */
    int32_t device;            /* HIP device ordinal                            */
    int32_t verbose;           /* 1: print reportConsole lines to stdout        */
    int32_t profile;           /* 1: HIP events around every kernel launch, read
                                  with mpm_kernel_times                         */
    int32_t launch;            /* enum mpm_launch                               */
    const uint16_t *node_group;/* batched worlds: group (world) id per node, or
                                  NULL; copied at mpm_create.  Groups must not
                                  share nodes (keep 2 empty cells between them) */
    int32_t n_groups;          /* number of groups when node_group != NULL      */
    int32_t persistent;        /* 1: plain steps as a persistent launch -- a
                                  resident grid of blocks pulls bricks from
                                  per-XCD work queues (batched worlds: all
                                  cells' bricks in one queue set); 0: one
                                  block per brick.  Results are identical.   */
} mpm_config;

/* Host point arrays, FP64, caller-allocated, original point order.  Mirrors
 * the MaterialPoint fields (MaterialPoint.h:24-56). Vectors are 3 per point,
 * F is 9 per point column-major (glm m[col][row]), Voigt 6-vectors are
 * [xx, yy, zz, xy, yz, zx].  Pointers a call does not need may be NULL. */
typedef struct mpm_points {
    int32_t n;
    double *x, *v, *fext, *F, *strain, *stress, *eps_p, *alpha;
    double *mass, *vol, *vol0;
    double *E, *nu, *sy, *C0, *C1;
    int32_t *type;
    uint8_t *dc;               /* b_DisplacementControl                         */
    uint8_t *surface;          /* b_Surface (contact path)                      */
    double *kernel, *kernel_grad;  /* d_Kernel, d3_Kernel_Gradient (readback)   */
    double *j2;                /* MPM_F_J2 readback: 2 per point               */
    int32_t *cell;             /* MPM_F_CELL readback: 3 per point             */
    uint32_t *order;           /* MPM_F_CELL readback: 1 per point             */
} mpm_points;

/* Grid nodes, caller-allocated, (cells+1) product entries in the reference's
 * linear order ix + iy nx + iz nx ny (GridPoint_Factory.h:23-32).  Mirrors the
 * GridPoint fields (GridPoint.h:12-28).  Pointers a caller does not need may
 * be NULL. */
typedef struct mpm_nodes {
    int64_t n;
    uint8_t *active;           /* b_Active                                      */
    uint8_t *fixed;            /* b3_Fixed as MPM_FIX_* bits                    */
    double *x;                 /* d3_Position, 3 per node                       */
    double *mass;              /* d_Mass                                        */
    double *v;                 /* d3_Velocity, 3 per node                       */
    double *f;                 /* d3_Force, 3 per node                          */
    double *f_temp;            /* d3_Force_Temp, 3 per node                     */
    double *kernel;            /* kernel grid d_Kernel (contact path)           */
    double *kernel_grad;       /* kernel grid d3_Kernel_Gradient, 3 per node    */
} mpm_nodes;

typedef struct mpm_handle mpm_handle;

/* ABI version (for the FFI binding to check) */
int mpm_abi_version(void);

/* Create a solver on one device.  Replaces PhysicsEngine::PhysicsEngine()
 * plus the grid part of initializeWorld_* (GridPoint_Factory::createGrid,
 * GridPoint_Factory.h:33-67; node BCs …PolygonCell.cpp:42-79). */
int mpm_create(const mpm_config *cfg, mpm_handle **out);

/* Upload the material points.  Replaces the allMaterialPoint /
 * v_MarkedMaterialPoints_Displacement_Control vectors that initializeWorld_*
 * fills (PhysicsEngine.h:78-80, …PolygonCell.cpp:137-191).  Points with dc!=0
 * are displacement-controlled (platens), in index order.  At most 2^28 points
 * (MPM_E_ARG beyond); a non-finite real input is MPM_E_ARG as well.  The
 * brick arenas (8 (Sx+3)(Sy+3)(Sz+3) reals per occupied brick, for 1.25x the
 * occupied bricks) must stay below 2^32 elements: beyond, this call or the
 * mpm_step whose rebin needs more returns MPM_E_ARG. */
int mpm_upload_points(mpm_handle *h, const mpm_points *pts);

/* Advance by dt_total in steps of at most dt_max.  Replaces
 * int PhysicsEngine::runSimulation_Classic_SinglePass_MP(double) and
 * …_MP_Contact(double) (PhysicsEngine.h:45-46,
 * PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:4-394).
 * *finished receives the reference's return value: 1 if d_Time >= d_TimeEnd,
 * else 0 (…_MP.cpp:385-393).  Blocks until the steps completed.  The path may
 * change from call to call (…_MP_Contact.cpp:4-676 for MPM_PATH_CONTACT). */
int mpm_step(mpm_handle *h, double dt_total, int32_t path, int32_t *finished);

/* The same advance without blocking (the threading model of the step's
 * caller, GraphicsEngine_gameLoop.cpp:7-36, alternates step and draw on one
 * thread; this lets a caller draw the previous readback while the GPU steps).
 * mpm_step_async checks its arguments, starts the steps on a worker thread of
 * the handle and returns.  Until mpm_step_wait collects the result, every
 * other call on the handle returns MPM_E_STATE (mpm_last_error says why);
 * mpm_step_poll sets *done to 1 once the steps have finished.  mpm_step_wait
 * blocks until then and returns the status mpm_step would have returned, with
 * *finished as in mpm_step.  mpm_destroy waits for a step in flight. */
int mpm_step_async(mpm_handle *h, double dt_total, int32_t path);
int mpm_step_poll(mpm_handle *h, int32_t *done);
int mpm_step_wait(mpm_handle *h, int32_t *finished);

/* Copy point fields (mask of MPM_F_*) back to host arrays in the original
 * point order.  Replaces PhysicsEngine::getMaterialPoints (PhysicsEngine.h:64),
 * whose callers read the live MaterialPoint objects.  MPM_F_KERNEL fills
 * kernel / kernel_grad with d_Kernel / d3_Kernel_Gradient of the last contact
 * step (…_MP_Contact.cpp:199-219), zeros before one. */
int mpm_readback(mpm_handle *h, uint32_t field_mask, mpm_points *out);

/* Grid node state.  Replaces PhysicsEngine::getGridPoints and
 * getGridPoints_Kernel (PhysicsEngine.h:65-66), whose callers read the live
 * GridPoint objects (GraphicsEngine_drawGame.cpp:98,302,373,454).  The fused
 * plain step keeps no grid between steps, so after a plain step (or upload)
 * the nodes are the grid of the next step: the P2G of the current points
 * (…_MP.cpp:41-202) and the grid update at the current time with dt_max
 * (…_MP.cpp:209-265) -- velocity after the update, force after damping and
 * boundary conditions, Force_Temp.y the reaction of y-fixed nodes; kernel
 * fields zero.  After a contact step they are that step's main (first) field
 * (…_MP_Contact.cpp:392-518) with its kernel grid (…_MP_Contact.cpp:121-191),
 * Force_Temp zero (the platen reaction is in mpm_reaction_history).
 * out->n must equal the node count.  Positions are (i L)/n per axis
 * (GridPoint_Factory.h:40-42). */
int mpm_grid_nodes(mpm_handle *h, mpm_nodes *out);

/* FP64 global checks, computed on the device.  Replaces the sums of
 * PhysicsEngine::reportConsole (PhysicsEngine.cpp:33-97):
 * out[0] grid mass, out[1] reaction (sum of node Force_Temp.y), out[2]
 * sum m|v|^2 (no 1/2, as the reference), out[3..5] momentum sum m v,
 * out[6] non-finite point count, out[7] current time, out[8] step count,
 * out[9] number of rebins, out[10] occupied bricks, out[11] grid nodes with
 * mass > 0 after the last grid update. */
int mpm_checks(mpm_handle *h, double out[12]);

/* Per-step reaction force (sum of Force_Temp.y after each step; the
 * "Force_y" of reportConsole, PhysicsEngine.cpp:65-71,195) of the steps of the
 * last mpm_step call; *n receives the count (<= cap). */
int mpm_reaction_history(mpm_handle *h, double *out, int32_t cap, int32_t *n);

/* Per-group reaction (the Force_y sum of each world's y-fixed nodes) of every
 * step of the last mpm_step call on the plain path, out[step * n_groups + g];
 * *n receives the step count (<= cap).  Zeros for contact-path steps.  No
 * reference counterpart: the reference runs one world per PhysicsEngine
 * (PhysicsEngine.cpp:65-71 is the single-world sum). */
int mpm_group_reaction_history(mpm_handle *h, double *out, int32_t cap, int32_t *n);

/* getTime_Current / getTime_End / getTime_Increment (PhysicsEngine.h:57-60) */
int mpm_get_time(mpm_handle *h, double *time, double *time_end, double *dt_max);

/* Device-time of the last mpm_step call in milliseconds (HIP events on the
 * handle's stream).  Replaces getTime_Runtime (PhysicsEngine.h:57). */
int mpm_last_step_ms(mpm_handle *h, double *ms);

/* Per-kernel device time of the last mpm_step call (needs cfg.profile=1):
 * out[0] fused particle kernel total ms, out[1] its launches, out[2] grid
 * kernel total ms, out[3] its launches, out[4] rebin total ms, out[5] rebins.
 * Replaces the per-phase a_Runtime timers (PhysicsEngine.h:73,
 * …_MP.cpp:39-369). */
int mpm_kernel_times(mpm_handle *h, double out[6]);

/* reportConsole(sDescription) (PhysicsEngine.cpp:33-120): the console line of
 * the current state, "<description>\ttime: T\tRuntime: R[\tPosition_y: Y]
 * \tForce_y: F\tKinetic Energy: K\n", numbers in the reference's Script()
 * format (Definitions.h:43-56: %+.6e, Runtime %+.3e).  Position_y is the first
 * platen point's y (the displacement monitor of …PolygonCell.cpp:189), Force_y
 * the last step's reaction, K = sum m|v|^2 over all points, R the device time
 * of all mpm_step calls.  Printed to stdout when cfg.verbose, appended to the
 * log file of mpm_set_log, copied to buf (NUL-terminated, at most cap bytes)
 * when buf != NULL; *len (if not NULL) receives the full length.  mpm_step
 * emits the same line (empty description) every console_interval and at the
 * end, as …_MP.cpp:374-391 does. */
int mpm_report(mpm_handle *h, const char *description, char *buf, int32_t cap, int32_t *len);

/* The log file every report line is appended to: replaces the fixed
 * _STR_LOGFILE "./bmp/log.txt" (Definitions.h:22).  NULL or "" (the default)
 * writes no file. */
int mpm_set_log(mpm_handle *h, const char *path);

/* Human-readable last error of the handle (or of mpm_create when h is NULL). */
const char *mpm_last_error(const mpm_handle *h);

/* Release device memory.  Replaces PhysicsEngine::~PhysicsEngine
 * (PhysicsEngine.cpp:8-31). */
void mpm_destroy(mpm_handle *h);

/* ---- stable radix sort of particle keys (exposed for parity tests of the
 *      deterministic binning).  Sorts n 32-bit keys ascending, ties by input
 *      position, on the device; writes the sorted keys and the permutation
 *      (perm[j] = input index of the j-th output).  No reference counterpart:
 *      the reference keeps points unsorted (PhysicsEngine.h:78). ----------- */
int mpm_sort_keys(const uint32_t *keys_host, int32_t n, uint32_t *sorted_host,
                  uint32_t *perm_host, int32_t device);

/* ---- constitutive increments of a batch (replaces the per-point
 *      ConstitutiveRelation object of …_MP.cpp:322-353; interface of
 *      ConstitutiveRelation.h:16-34).  The step's own device functions, FP64,
 *      the reference's evaluation order, applied to n independent inputs:
 *        law MPM_LAW_ELASTIC           calculateIncrement_Elastic(E, nu, de)
 *                                      (…Elastic.cpp:4-15)
 *        law MPM_LAW_NEOHOOKEAN        calculateIncrement_NeoHookean_6D(E, nu, s, F)
 *                                      (…NeoHookean.cpp:4-22)
 *        law MPM_LAW_PLASTIC           calculateIncrement_PerfectlyPlastic_6D(E, nu, sy, s, de)
 *                                      (…PerfectlyPlastic.cpp:4-72)
 *        law MPM_LAW_VONMISES          calculateIncrement_VonMisesHardening_6D(E, nu, sy,
 *                                      alpha, C0, C1, s, de) (…VonMisesHardening.cpp:4-86)
 *        law MPM_LAW_J2_STATE          calculateState_J2(s) (…calculateState_J2.cpp:4-26)
 *      Inputs the law does not use may be NULL.  Outputs: stress_inc and
 *      plastic_inc (6 per input, Voigt xx yy zz xy yz zx), alpha_inc (1), or j2
 *      (1) for MPM_LAW_J2_STATE.  F is column-major (F[3*col+row]).  The
 *      ViscoElastic and IdealGas laws (…ViscoElastic.cpp, …IdealGas.cpp) are
 *      never dispatched by the reference step and return MPM_E_ARG. ------ */
#define MPM_LAW_ELASTIC 0
#define MPM_LAW_NEOHOOKEAN 1
#define MPM_LAW_VISCOELASTIC 2
#define MPM_LAW_PLASTIC 3
#define MPM_LAW_VONMISES 4
#define MPM_LAW_IDEALGAS 5
#define MPM_LAW_J2_STATE 16

typedef struct mpm_cr_batch {
    int32_t n;
    const double *E, *nu, *yield, *alpha, *C0, *C1;   /* n each */
    const double *stress, *strain_inc;                /* 6n each */
    const double *F;                                  /* 9n */
    double *stress_inc, *plastic_inc;                 /* 6n each */
    double *alpha_inc, *j2;                           /* n each */
} mpm_cr_batch;

int mpm_constitutive(int32_t law, const mpm_cr_batch *b, int32_t device);

/* mpm_constitutive without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The batch is staged
 * in a fixed device array of the library (41 x 8192 doubles, 2.6 MiB, one
 * per GPU, held until the process ends) instead of a buffer the library
 * allocates; the results equal mpm_constitutive's bit for bit.  At most 8192
 * inputs per call (MPM_E_ARG beyond).  Calls on one GPU share the array:
 * this process serialises its own calls, and any other user of the same
 * library copy on that GPU at the same time races with them. */
int mpm_constitutive_noalloc(int32_t law, const mpm_cr_batch *b, int32_t device);

/* ---- differentiable run (no reference counterpart: the reference has no
 *      gradients).  n_steps plain steps of cfg->dt_max, FP64 arithmetic
 *      (…_MP.cpp:106-353, strain increment L dt, damping and TimeLine
 *      platen as the reference; elastic, NeoHookean when cfg->neohookean,
 *      perfectly plastic and J2-hardening points, each point's yield
 *      decision held fixed in the adjoint), on a dense node
 *      grid with deterministic fixed-point node sums; the backward pass is the
 *      adjoint of every phase, replayed from checkpoints every `checkpoint`
 *      steps.  Parameters are the per-point Young's modulus and the initial
 *      point positions; the output is the final displacement u = x - x0,
 *      carried as its own state so that it keeps full FP64 precision.
 *      cfg->precision: MPM_FP64 state, or MPM_FP32 state (u, v, F - I,
 *      sigma, alpha in float, rounded once per step, positions x0 + u;
 *      the adjoint takes the rounding as the identity).  Host
 *      arrays are in the original point order of `pts` (platen points move
 *      as in the reference and get zero gradients). ------------------------ */
typedef struct mpm_diff mpm_diff;

/* Copies the scene: masses, vol0 (or vol), nu, sy, C0, C1, initial alpha
 * (0 if null), initial v, fext, type, dc flags, platen positions; node
 * fixity, damping, TimeLine and the NeoHookean dispatch from cfg.  Types map
 * as the reference's dispatch does (…_MP.cpp:332-339).  MPM_E_ARG for an
 * unknown type, for plastic points without sy (or C0, C1 for hardening), or
 * for a precision other than MPM_FP64 and MPM_FP32. */
int mpm_diff_create(const mpm_config *cfg, const mpm_points *pts, int32_t n_steps,
                    int32_t checkpoint, mpm_diff **out);

/* Runs the forward from F = I, stress 0, the scene's initial velocities and
 * x0 (3 per point) with Young's moduli E (1 per point); writes the final
 * displacements u_final (3 per point) and keeps the checkpoints.
 * MPM_E_OUT_OF_GRID if a point leaves the grid box. */
int mpm_diff_forward(mpm_diff *d, const double *E, const double *x0, double *u_final);

/* Given dLoss/du_final (3 per point), writes dLoss/dE (1 per point) and
 * dLoss/dx0 (3 per point) of the last forward.  MPM_E_STATE before a
 * forward. */
int mpm_diff_backward(mpm_diff *d, const double *g_ufinal, double *g_E, double *g_x0);

/* Last error message of the run. */
const char *mpm_diff_last_error(const mpm_diff *d);

/* Releases the run's device memory. */
void mpm_diff_destroy(mpm_diff *d);

#ifdef __cplusplus
}
#endif
#endif /* MPM_AUXETICS_H */
