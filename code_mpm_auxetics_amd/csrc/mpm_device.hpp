// mpm_device.hpp -- shared device-side types and helpers for the MI355X MPM step.
//
// Layout in HBM (one handle):
//   deformable points: structure-of-arrays, one array per scalar field, kept in
//     brick-binned order (bricks of S = (Sx,Sy,Sz) grid cells); FP64 mode stores
//     absolute positions, FP32 mode stores positions as Q6.26 fixed point in
//     cell units relative to the origin of the point's brick (uniform 1.5e-8
//     cell resolution; re-basing at a rebin is an exact integer shift).
//   displacement-controlled points (platens): FP64 SoA, unsorted.
//   grid nodes: dense lattice (cells+1 per axis); per node an 8-word record
//     {v.xyz, a.x | a.y, a.z, -, -} written by the grid kernel, read by G2P.
//   brick arenas: per occupied brick, the P2G sums (m, p.xyz, f.xyz) of the
//     nodes its points touched, as 8-real node records over the touched box
//     of the (S+3)^3 node window (box-compact, x fastest); written by the
//     fused particle kernel, merged in a fixed order by the grid kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_laws.hpp"   // det3 and the FP64 constitutive laws

namespace mpm {

constexpr int kBlock = 256;          // threads per particle workgroup
constexpr int kQ = 7;                // node fields: m, p.xyz, f.xyz
constexpr uint32_t kNoBrick = 0xFFFFFFFFu;
constexpr int kQFrac = 26;                       // fractional bits of FP32 positions
constexpr double kQOne = 67108864.0;             // 2^26
// FP32 state: largest brick edge in cells.  Positions are Q6.26 cells from
// the brick origin (|u| < 32); a point lies within [-1, S + 1] of the origin
// (the arena window) plus one step's motion before the drift check flags it.
constexpr int kMaxBrickFp32 = 28;

enum : int {
    FLAG_REBIN = 1,        // a point drifted past half a cell outside its brick
    FLAG_ARENA = 2,        // a point left its brick's arena window (fatal)
    FLAG_OUT_OF_GRID = 4,  // a point left the grid box (fatal)
    FLAG_NAN = 8           // non-finite state (fatal)
};

struct GridParams {
    double L[3];
    double h[3];
    float hf[3];
    float ihf[3];  // 1/h in FP32
    int cells[3];
    int nodes[3];
    int S[3];      // brick size in cells
    int nb[3];     // brick lattice size
    int A[3];      // arena window size in nodes = S + 3
    int arenaN;    // A[0]*A[1]*A[2]
    int cbits;     // bits for the cell-in-brick index in sort keys
    double m_min, damping;
    int strain_form, neohookean, track_strain, has_fext, has_pointE;
    double ih[3];  // 1/h in FP64 (FP32-state contact arithmetic multiplies by it)
};

// Per-step record filled on the host (time, dt, platen velocity, epoch, and
// the ping-pong parity the step reads: a step reads buffers [par] and writes
// [par ^ 1], so one captured batch serves either starting parity).
struct StepRec {
    double dt;
    double t;
    double vtl[3];
    int32_t has_tl;
    uint32_t epoch;
    int32_t par;
    int32_t skip_sums;   // 1: nothing reads this step's point check sums or grid mass and node
                         // count (a step of a batch before its last); 0: compute them
};
static_assert(sizeof(StepRec) % 4 == 0, "step records are copied as 32-bit words");

// Material table entry (reference per-point fields E, nu, sigma_y, C0, C1 and
// type; MaterialPoint.h:28-56), deduplicated on upload.
struct Material {
    int32_t type;
    int32_t pad;
    double E, nu, sy, C0, C1;
    // derived on the host with the reference's own expressions
    double C;      // E/(1+nu)/(1-2nu)   (…Elastic.cpp:6)
    double G;      // 0.5*E/(1+nu)       (…PerfectlyPlastic.cpp:54)
    double k0;     // sy/sqrt(3)         (…PerfectlyPlastic.cpp:39)
    double invE, invG, hs;   // 1/E, 1/G, sqrt(2/3)*C1
    double Cn;               // 1/((1+nu)(1-2nu)): C = E*Cn for per-point E
};

template <typename real>
struct PArrays {
    real *x[3];        // FP64 state: absolute position
    int32_t *xq[3];    // FP32 state: Q6.26 cells relative to the brick origin
    real *v[3], *F[9], *s[6], *ep[6], *eps[6];
    real *alpha, *mass, *vol0, *E, *fext[3];
    uint8_t *mat;
    uint8_t *surf;     // b_Surface (contact path)
    uint32_t *id, *pbrick;
};

// Element i of a point array through a 32-bit byte offset: the address is an
// SGPR base plus a zero-extended 32-bit VGPR offset, which global loads and
// stores take directly (no per-array 64-bit address arithmetic).  Point
// counts are limited so that every offset fits 32 bits.
constexpr uint32_t kMaxPoints = 1u << 28;
template <typename T>
__device__ __forceinline__ T &elt(T *base, uint32_t i)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T));
}

// Deformation gradient storage.  FP64 state keeps F; FP32 state keeps
// D = F - I, so that per-step increments (L F dt ~ 1e-8) are not rounded
// away against the 1.0 on the diagonal.  fd() returns component q of F in
// FP64 from the stored value; fstore() converts F to the stored form.
template <typename real>
__device__ __host__ __forceinline__ double fd(real stored, int q)
{
    return (double)stored + ((sizeof(real) == 4 && (q % 4) == 0) ? 1.0 : 0.0);
}
template <typename real>
__device__ __host__ __forceinline__ real fstore(double F, int q)
{
    return (real)(F - ((sizeof(real) == 4 && (q % 4) == 0) ? 1.0 : 0.0));
}

template <typename real>
struct NodeRec {
    real d[8];
};

// ---------------------------------------------------------------------------
// DPP helpers (row = 16 lanes).  bound_ctrl=true: lanes without a source read 0.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = dpp_i<CTRL>(p.x);
    p.y = dpp_i<CTRL>(p.y);
    return *reinterpret_cast<double *>(&p);
}
template <int CTRL>
__device__ __forceinline__ float dpp_r(float v) { return dpp_f<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double dpp_r(double v) { return dpp_d<CTRL>(v); }

// row_shr:d = 0x110+d, row_shl:d = 0x100+d
#define MPM_ROW_SHR(d) (0x110 + (d))
#define MPM_ROW_SHL(d) (0x100 + (d))

__device__ __forceinline__ double sgn(double a) { return (double)((0.0 < a) - (a < 0.0)); }
__device__ __forceinline__ float sgn(float a) { return (float)((0.0f < a) - (a < 0.0f)); }

}  // namespace mpm
