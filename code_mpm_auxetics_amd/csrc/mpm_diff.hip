// mpm_diff.hip -- differentiable run of the plain step and its adjoint, with
// FP64 or FP32 point state (FP64 arithmetic in every step and the adjoint).
//
// The forward is the plain step of PhysicsEngine_runSimulation_Classic_-
// SinglePass_MP.cpp:41-368 with the strain increment L dt, on a dense node
// grid: P2G (…:106-202), grid update with damping, boundary conditions and the
// platen override (…:209-265), G2P with F, volume and the constitutive update
// (…:271-353) for elastic (…Elastic.cpp:4-15), NeoHookean (…NeoHookean.cpp:
// 4-22, when dispatched), perfectly plastic (…PerfectlyPlastic.cpp:4-72) and
// J2-hardening points (…VonMisesHardening.cpp:4-86; the hardening variable
// alpha is part of the state).  The backward pass is the
// hand-written adjoint of each phase in reverse order; the forward keeps
// checkpoints every C steps and the backward recomputes one segment at a time.
//
// Node sums are 64-bit fixed point with a per-node, per-field exponent: a
// first pass over the points takes each node's largest term and term count,
// the second adds the terms scaled so that the count of them fits in 62 bits.
// The sums are exact, so a replay is bit-identical to the forward, and every
// node keeps ~54 bits relative to its own largest term (small far-field nodes
// are not quantised against the loaded region's forces).  Each point also
// carries its displacement u = x - x0 as state, so that displacements of
// 1e-9 m on positions of 1e-2 m keep full FP64 precision; the run's output
// is u.  FP32 state stores u, v, F - I, sigma and alpha in float (positions
// x0 + u, x0 FP64 and fixed for the run) and rounds them once per step; the
// adjoint replays the same rounded states and takes the rounding as the
// identity.
//
// Adjoint of one step (state s = x, u, v, F, sigma, alpha; parameters E; N and
// grad N evaluated at the step's starting positions; piecewise-constant
// choices -- cells, mass gates, boundary masks, platen nodes, each point's
// yield and hardening decision -- have zero derivative):
//   G2P      x' = x + dt sum_I N vt_I (u' likewise), v' = v + dt sum_I N a_I,
//            L = sum_I vt_I (x) grad N,  F' = F + dt L F,
//            sigma' = sigma + C(E) eps(L dt) (elastic), P(F') (NeoHookean),
//            or the radial return of the trial stress (plastic laws)
//   grid     vh = P/m, vt = mask(vh + f/m dt) (or the platen velocity),
//            a = mask(f)/m (0 on platen nodes)
//   P2G      m_I = sum N m, P_I = sum N m v, f_I = -sum V sigma.grad N,
//            V = det(F) V0
// Position adjoints use grad N and the second derivatives of the linear hats
// (d2N/dxa dxb = ga gb s_c for a != b, 0 for a == b).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpm_auxetics.h"
#include "mpm_device.hpp"

using namespace mpm;

namespace {

struct DGrid {
    double L[3], h[3];
    int cells[3], nodes[3];
    long long nn;
    double m_min;
};

struct DArgs {
    DGrid g;
    int n, n_dc;
    double dt, damping;
    // point state (SoA, component-major: x[a * n + p])
    double *x, *u, *v, *F, *s, *al;   // al: hardening variable alpha (n)
    // FP32 state (fp32 != 0): u, v, D = F - I, sigma, alpha in float, the
    // positions x0 + u (x0 FP64, fixed for the run); the FP64 pointers above
    // are then unused.  Arithmetic stays FP64.
    int fp32;
    const double *x0;
    float *u32, *v32, *D32, *s32, *al32;
    const double *m, *V0, *E, *nu, *fext;
    const uint8_t *law;              // per point: LAW_* after the reference's dispatch
    const double *sy, *C0, *C1;      // yield stress, hardening constants
    const uint8_t *fixed;
    double *xdc, *vdc;             // platen positions / velocities (3 * n_dc)
    double vtl[3];                 // TimeLine velocity of this step (platen move)
    int has_tl;
    unsigned long long *marks;     // per node: 1 + highest platen index, 0 = none
    long long *acc;                // 7 x nn: m | P xyz | f xyz
    double *nm, *nP, *nf, *nvt, *na;
    uint8_t *nflag;                // 1 gate (m > m_min), 2 platen node
    unsigned long long *nb;        // 5 x nn node bounds (positive doubles as bits):
                                   // m | P | f | vt-bar | a-bar
    unsigned *cnt;                 // per node: number of point terms
    const int2 *items;             // scatter work items: (first point, count <= 256)
    double *nsc;                   // 5 x nn fixed-point scales of the node fields
    int32_t *err;
    // adjoint
    double *xb, *vb, *Fb, *sb, *Eb, *alb;
    double *scratch;               // 12 per point: L-bar (9), dt x-bar' (3)
    long long *aacc;               // 6 x nn: vt-bar xyz | a-bar xyz
    double *nmb, *nPb, *nfb;
};

// ---------------------------------------------------------------- helpers
// Linear-hat bases of a point in cell i0 (GridPoint_Mediator.h:64-88), per
// axis for the two node planes of the cell: s = 1 - |x - x_i| / h (0 outside
// [0, 1]) and ds/dx = -sign(x - x_i) / h, node position x_i = (i L) / cells.
// Corner k (x-outer, z-inner) takes N = s_x s_y s_z and G_a = ds_a * the other
// two s.
struct Bases {
    double w[3][2], gr[3][2];
};
__device__ __forceinline__ void bases(const DGrid &g, const double x[3], const int i0[3], Bases &B)
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double xi = (double)(i0[a] + c) * g.L[a] / (double)g.cells[a];
            const double d = x[a] - xi;
            double sv = 1.0 - fabs(d) / g.h[a];
            if (sv < 0.0 || sv > 1.0) sv = 0.0;
            B.w[a][c] = sv;
            B.gr[a][c] = -sgn(d) / g.h[a];
        }
}
__device__ __forceinline__ void corner(const Bases &B, int k, double &N, double G[3], double w[3], double gr[3])
{
    const int c[3] = {(k >> 2) & 1, (k >> 1) & 1, k & 1};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w[a] = B.w[a][c[a]];
        gr[a] = B.gr[a][c[a]];
    }
    N = w[0] * w[1] * w[2];
    G[0] = gr[0] * w[1] * w[2];
    G[1] = w[0] * gr[1] * w[2];
    G[2] = w[0] * w[1] * gr[2];
}
// d2N / dx_a dx_b
__device__ __forceinline__ double hess(const double w[3], const double gr[3], int a, int b)
{
    if (a == b) return 0.0;
    const int c = 3 - a - b;
    return gr[a] * gr[b] * w[c];
}
__device__ __forceinline__ bool cell_of(const DGrid &g, const double x[3], int i0[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!(x[a] >= 0.0)) return false;
        i0[a] = (int)(x[a] / g.h[a]);
        if (!(i0[a] < g.cells[a])) return false;
    }
    return true;
}
__device__ __forceinline__ long long nidx(const DGrid &g, const int i0[3], int k)
{
    return (i0[0] + ((k >> 2) & 1)) + (long long)g.nodes[0] * ((i0[1] + ((k >> 1) & 1)) + (long long)g.nodes[1] * (i0[2] + (k & 1)));
}
// Fixed-point scale of node field q: 2^(62 - H - e) with b < 2^e the node's
// largest term and 2^H > 2 (count + 1), so that count terms fit in 62 bits.
__device__ __forceinline__ double node_scale(const DArgs &a, int q, long long I)
{
    const double b = __longlong_as_double((long long)a.nb[q * a.g.nn + I]);
    if (!(b > 0.0)) return 0.0;
    int e;
    frexp(b, &e);
    const unsigned c = a.cnt[I];
    int H = 1;
    while ((1ull << H) <= 2ull * (c + 1ull)) ++H;
    return ldexp(1.0, 62 - H - e);
}
// 1 / a power-of-two scale (exact), 0 for an empty field
__device__ __forceinline__ double inv_scale(double sc) { return sc > 0.0 ? 1.0 / sc : 0.0; }
// column-major 3x3 (m[3c + r]) products
__device__ __forceinline__ void matmul(const double A[9], const double B[9], double C[9])
{
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) C[3 * c + r] = A[r] * B[3 * c] + A[3 + r] * B[3 * c + 1] + A[6 + r] * B[3 * c + 2];
}
// elastic stress increment of eps (engineering shear): Hooke (…Elastic.cpp:4-15)
__device__ __host__ __forceinline__ void hooke(double E, double nu, const double de[6], double ds[6])
{
    const double C = E / (1.0 + nu) / (1.0 - 2.0 * nu);
    ds[0] = C * ((1.0 - nu) * de[0] + nu * de[1] + nu * de[2]);
    ds[1] = C * ((1.0 - nu) * de[1] + nu * de[0] + nu * de[2]);
    ds[2] = C * ((1.0 - nu) * de[2] + nu * de[0] + nu * de[1]);
    ds[3] = C * ((0.5 - nu) * de[3]);
    ds[4] = C * ((0.5 - nu) * de[4]);
    ds[5] = C * ((0.5 - nu) * de[5]);
}
// symmetric Voigt eps of L dt (…_MP.cpp:304-312 with L dt)
__device__ __forceinline__ void eps_of(const double L[9], double dt, double de[6])
{
    de[0] = L[0] * dt;
    de[1] = L[4] * dt;
    de[2] = L[8] * dt;
    de[3] = L[1] * dt + L[3] * dt;
    de[4] = L[5] * dt + L[7] * dt;
    de[5] = L[6] * dt + L[2] * dt;
}
// Material laws of the run, after the reference's dispatch (…_MP.cpp:332-339):
// elastic, NeoHookean (when dispatched), perfectly plastic, J2 with
// hardening.  Forward in the order of the plain step's FP64 functions.
enum { LAW_ELASTIC = 0, LAW_NEOHOOKEAN = 1, LAW_PLASTIC = 3, LAW_VONMISES = 4 };

// J2 of a Voigt stress (…calculateState_J2.cpp:4-26)
__device__ __host__ __forceinline__ double j2v(const double s[6])
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

// Radial return of …PerfectlyPlastic.cpp:4-72 (k = sigma_y / sqrt 3) and
// …VonMisesHardening.cpp:4-86 (radius k + alpha / sqrt 2; on the plastic
// branch d alpha = C0 (sqrt(2/3) C1 - alpha) eps_bar when positive, eps_bar =
// sqrt(2/3 sum dep_i^2)), with every intermediate kept for the adjoint.
struct RR {
    double del[6], st[6], p, dev[6], j2, radius, sc, ds[6], dee[6], dep[6], eff, hs, da;
    bool plastic, harden;
};
__device__ __host__ __forceinline__ void rr_eval(double E, double nu, double sy, double C0, double C1, bool vm,
                                        const double s[6], double al, const double de[6], RR &r)
{
    hooke(E, nu, de, r.del);
    for (int i = 0; i < 6; ++i) r.st[i] = s[i] + r.del[i];
    r.p = 1.0 / 3.0 * (r.st[0] + r.st[1] + r.st[2]);
    for (int i = 0; i < 6; ++i) r.dev[i] = r.st[i];
    r.dev[0] -= r.p;
    r.dev[1] -= r.p;
    r.dev[2] -= r.p;
    r.j2 = j2v(r.st);
    const double k0 = sy / sqrt(3.0);
    r.radius = vm ? k0 + al / sqrt(2.0) : k0;
    r.plastic = !(r.j2 <= r.radius * r.radius);
    r.harden = false;
    r.sc = 1.0;
    r.eff = r.hs = r.da = 0.0;
    for (int i = 0; i < 6; ++i) r.dee[i] = r.dep[i] = 0.0;
    if (!r.plastic) {
        for (int i = 0; i < 6; ++i) r.ds[i] = r.del[i];
        return;
    }
    r.sc = r.radius / sqrt(r.j2);
    for (int i = 0; i < 6; ++i) r.ds[i] = r.dev[i] * r.sc - s[i];
    r.ds[0] += r.p;
    r.ds[1] += r.p;
    r.ds[2] += r.p;
    const double G = 0.5 * E / (1.0 + nu);
    r.dee[0] = r.ds[0] / E - r.ds[1] * nu / E - r.ds[2] * nu / E;
    r.dee[1] = r.ds[1] / E - r.ds[0] * nu / E - r.ds[2] * nu / E;
    r.dee[2] = r.ds[2] / E - r.ds[0] * nu / E - r.ds[1] * nu / E;
    r.dee[3] = r.ds[3] / G;
    r.dee[4] = r.ds[4] / G;
    r.dee[5] = r.ds[5] / G;
    for (int i = 0; i < 6; ++i) r.dep[i] = de[i] - r.dee[i];
    if (vm) {
        double eff = 0.0;
        for (int q = 0; q < 6; ++q) eff += r.dep[q] * r.dep[q];
        eff *= 2.0 / 3.0;
        r.eff = sqrt(eff);
        r.hs = sqrt(2.0 / 3.0) * C1 - al;
        r.harden = r.hs > 0.0;
        if (r.harden) r.da = C0 * r.hs * r.eff;
    }
}

// Adjoint of rr_eval with its branches held (s' = s + ds, alpha' = alpha + da).
// In: sb = d/ds', ab = d/dalpha'.  Out: sbo = d/ds, abo = d/dalpha, eb = d/dde
// (engineering shear), gE = d/dE.
__device__ __host__ __forceinline__ void rr_adjoint(double E, double nu, double C0, bool vm, const double de[6], const RR &r,
                                           const double sb[6], double ab, double sbo[6], double &abo, double eb[6],
                                           double &gE)
{
    for (int i = 0; i < 6; ++i) { sbo[i] = sb[i]; eb[i] = 0.0; }
    abo = ab;
    gE = 0.0;
    double delb[6];
    if (!r.plastic) {
        for (int i = 0; i < 6; ++i) delb[i] = sb[i];
    } else {
        // alpha' = alpha + C0 hs eff, hs = sqrt(2/3) C1 - alpha
        double effb = 0.0;
        if (vm && r.harden) {
            abo -= ab * C0 * r.eff;
            effb = ab * C0 * r.hs;
        }
        // eff = sqrt(2/3 sum dep^2)
        double depb[6];
        for (int i = 0; i < 6; ++i) depb[i] = (r.eff > 0.0) ? effb * (2.0 / 3.0) * r.dep[i] / r.eff : 0.0;
        // dep = de - dee(ds), dee = compliance(E) ds
        double dsb[6];
        for (int i = 0; i < 6; ++i) {
            eb[i] += depb[i];
            gE += depb[i] * r.dee[i] / E;
        }
        const double G = 0.5 * E / (1.0 + nu);
        dsb[0] = sb[0] - (depb[0] - nu * depb[1] - nu * depb[2]) / E;
        dsb[1] = sb[1] - (depb[1] - nu * depb[0] - nu * depb[2]) / E;
        dsb[2] = sb[2] - (depb[2] - nu * depb[0] - nu * depb[1]) / E;
        for (int i = 3; i < 6; ++i) dsb[i] = sb[i] - depb[i] / G;
        // ds = dev sc - s + p (normal components)
        double devb[6], scb = 0.0, pb = dsb[0] + dsb[1] + dsb[2];
        for (int i = 0; i < 6; ++i) {
            sbo[i] -= dsb[i];
            devb[i] = r.sc * dsb[i];
            scb += dsb[i] * r.dev[i];
        }
        // sc = radius / sqrt(j2), radius = k0 (+ alpha / sqrt 2)
        if (vm) abo += scb / sqrt(r.j2) / sqrt(2.0);
        const double j2b = -0.5 * scb * r.sc / r.j2;
        // j2(st): d/dst_i = dev_i (normal), 2 st_i (shear); dev = st - p, p = tr st / 3
        double stb[6];
        for (int i = 0; i < 3; ++i) {
            stb[i] = j2b * r.dev[i] + devb[i];
            pb -= devb[i];
        }
        for (int i = 3; i < 6; ++i) stb[i] = j2b * 2.0 * r.st[i] + devb[i];
        for (int i = 0; i < 3; ++i) stb[i] += pb / 3.0;
        // st = s + del
        for (int i = 0; i < 6; ++i) {
            sbo[i] += stb[i];
            delb[i] = stb[i];
        }
    }
    // del = C(E) de
    double t[6];
    hooke(E, nu, delb, t);
    for (int i = 0; i < 6; ++i) {
        eb[i] += t[i];
        gE += delb[i] * r.del[i] / E;
    }
}

// NeoHookean total stress (…NeoHookean.cpp:4-22) of the updated F; the
// increment P - s added to s gives P (rounded)
__device__ __host__ __forceinline__ void nh_stress(double E, double nu, const double F[9], double P[9])
{
    const double mu = 0.5 * E / (1.0 + nu);
    const double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double J = det3(F);
    const double a = lam * log(J) / J;
    const double b = mu / J;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            const double fft = F[r] * F[c] + F[3 + r] * F[3 + c] + F[6 + r] * F[6 + c];
            const double id = (c == r) ? 1.0 : 0.0;
            P[3 * c + r] = a * id + b * (fft - id);
        }
}
// cofactor matrix, cof[3c + r] = d det / d F[3c + r] (column-major F)
__device__ __host__ __forceinline__ void cofactor(const double F[9], double cof[9])
{
    cof[0] = F[4] * F[8] - F[5] * F[7];
    cof[1] = F[5] * F[6] - F[3] * F[8];
    cof[2] = F[3] * F[7] - F[4] * F[6];
    cof[3] = F[2] * F[7] - F[1] * F[8];
    cof[4] = F[0] * F[8] - F[2] * F[6];
    cof[5] = F[1] * F[6] - F[0] * F[7];
    cof[6] = F[1] * F[5] - F[2] * F[4];
    cof[7] = F[2] * F[3] - F[0] * F[5];
    cof[8] = F[0] * F[4] - F[1] * F[3];
}
// Adjoint of s' = P(F) (Voigt xx yy zz xy yz zx of P): Fb += dP/dF^T sb,
// gE = d/dE (P is linear in E).  sigma = lam ln J / J I + mu / J (F F^T - I).
__device__ __host__ __forceinline__ void nh_adjoint(double E, double nu, const double F[9], const double sb[6], double Fb[9],
                                           double &gE)
{
    const double mu = 0.5 * E / (1.0 + nu);
    const double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double J = det3(F), lJ = log(J);
    const double b = mu / J;
    // symmetric adjoint of the matrix entries: the shear value stands for both
    const double S[9] = {sb[0], 0.5 * sb[3], 0.5 * sb[5], 0.5 * sb[3], sb[1], 0.5 * sb[4], 0.5 * sb[5], 0.5 * sb[4], sb[2]};
    double P[9];
    nh_stress(E, nu, F, P);
    gE = (sb[0] * P[0] + sb[1] * P[4] + sb[2] * P[8] + sb[3] * P[1] + sb[4] * P[5] + sb[5] * P[6]) / E;
    double bb = 0.0;   // d/db = S : (F F^T - I)
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            const double fft = F[r] * F[c] + F[3 + r] * F[3 + c] + F[6 + r] * F[6 + c];
            bb += S[3 * c + r] * (fft - ((c == r) ? 1.0 : 0.0));
        }
    const double ab = S[0] + S[4] + S[8];   // d/da = tr S
    const double Jb = ab * lam * (1.0 - lJ) / (J * J) - bb * mu / (J * J);
    double cof[9];
    cofactor(F, cof);
    // F F^T: d/dF = 2 S F (S symmetric), times b
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            Fb[3 * c + r] += 2.0 * b * (S[r] * F[3 * c] + S[3 + r] * F[3 * c + 1] + S[6 + r] * F[3 * c + 2]) + Jb * cof[3 * c + r];
}

// f contribution of stress s (Voigt) and gradient G: -V (sigma . G)
__device__ __forceinline__ void sigma_dot(const double s[6], const double G[3], double o[3])
{
    o[0] = G[0] * s[0] + G[1] * s[3] + G[2] * s[5];
    o[1] = G[1] * s[1] + G[0] * s[3] + G[2] * s[4];
    o[2] = G[2] * s[2] + G[0] * s[5] + G[1] * s[4];
}

// ------------------------------------------------------ block node windows
// The points are kept sorted by the brick of 8 x 8 x 4 cells that holds their
// starting cell, then in Morton order inside it (fixed at creation; the node
// sums are exact integers, so the order does not change any result).  A
// scatter block takes one work item: up to 256 consecutive points of one
// brick.  They touch a small box of nodes (the brick plus one node layer,
// more if points drift out of it): the block adds its terms there in LDS and
// each node's total reaches global memory with one atomic.  Lane t takes
// point (t % 8) * 32 + t / 8 of the item, so the 8 points of one cell
// (adjacent in the order) fall on lanes 8 apart instead of hitting the same
// LDS words in one instruction.  A block whose box exceeds the window adds
// straight to global memory.
constexpr int kWinNodes = 640;
constexpr int kBrick[3] = {8, 8, 4};

struct Win {
    int lo[3], dim[3], nodes;
    bool lds;
};

__device__ __forceinline__ int block_point(const int2 *items, bool &in)
{
    const int2 it = items[blockIdx.x];
    const int l = (threadIdx.x & 7) * 32 + (threadIdx.x >> 3);
    in = l < it.y;
    return it.x + l;
}

__device__ __forceinline__ Win block_window(bool valid, const int i0[3], int *sbox)
{
    if (threadIdx.x < 3) {
        sbox[threadIdx.x] = INT_MAX;
        sbox[3 + threadIdx.x] = -1;
    }
    __syncthreads();
    if (valid)
        for (int q = 0; q < 3; ++q) {
            atomicMin(&sbox[q], i0[q]);
            atomicMax(&sbox[3 + q], i0[q]);
        }
    __syncthreads();
    Win w{};
    if (sbox[3] < 0) return w;   // no point of the block in the grid
    long long nodes = 1;
    for (int q = 0; q < 3; ++q) {
        w.lo[q] = sbox[q];
        w.dim[q] = sbox[3 + q] - sbox[q] + 2;
        nodes *= w.dim[q];
    }
    w.lds = nodes <= kWinNodes;
    w.nodes = w.lds ? (int)nodes : 0;
    return w;
}
// window slot of corner k of cell i0
__device__ __forceinline__ int win_slot(const Win &w, const int i0[3], int k)
{
    return (i0[0] + ((k >> 2) & 1) - w.lo[0]) +
           w.dim[0] * ((i0[1] + ((k >> 1) & 1) - w.lo[1]) + w.dim[1] * (i0[2] + (k & 1) - w.lo[2]));
}
// global node of window slot j
__device__ __forceinline__ long long win_node(const DGrid &g, const Win &w, int j)
{
    const int jx = j % w.dim[0], r = j / w.dim[0];
    const int jy = r % w.dim[1], jz = r / w.dim[1];
    return (w.lo[0] + jx) + (long long)g.nodes[0] * ((w.lo[1] + jy) + (long long)g.nodes[1] * (w.lo[2] + jz));
}
__device__ __forceinline__ unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }

// ----------------------------------------------------- point state access
// FP64 state reads and writes the arrays as they are.  FP32 state widens
// every read to FP64 (positions x0 + u, F = I + D) and rounds each write
// once; the adjoint reads the same rounded states, so replays match the
// forward bit for bit and the rounding is taken as the identity.
__device__ __forceinline__ void ld_x(const DArgs &a, int p, double x[3])
{
    const int n = a.n;
    if (a.fp32)
        for (int q = 0; q < 3; ++q) x[q] = a.x0[q * n + p] + (double)a.u32[q * n + p];
    else
        for (int q = 0; q < 3; ++q) x[q] = a.x[q * n + p];
}
__device__ __forceinline__ double ld_v(const DArgs &a, int q, int p)
{
    return a.fp32 ? (double)a.v32[q * a.n + p] : a.v[q * a.n + p];
}
__device__ __forceinline__ double ld_F(const DArgs &a, int q, int p)
{
    return a.fp32 ? (double)a.D32[q * a.n + p] + ((q % 4 == 0) ? 1.0 : 0.0) : a.F[q * a.n + p];
}
__device__ __forceinline__ double ld_s(const DArgs &a, int q, int p)
{
    return a.fp32 ? (double)a.s32[q * a.n + p] : a.s[q * a.n + p];
}
__device__ __forceinline__ double ld_al(const DArgs &a, int p) { return a.fp32 ? (double)a.al32[p] : a.al[p]; }
__device__ __forceinline__ long long fixp(double v, double sc) { return __double2ll_rn(v * sc); }

// --------------------------------------------------------------- forward
__global__ __launch_bounds__(256) void kd_clear(DArgs a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I < a.g.nn) {
        for (int q = 0; q < 7; ++q) a.acc[q * a.g.nn + I] = 0;
        for (int q = 0; q < 3; ++q) a.nb[q * a.g.nn + I] = 0ull;
        a.cnt[I] = 0u;
        a.marks[I] = 0ull;
    }
}
__global__ __launch_bounds__(256) void kd_a_clear(DArgs a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I < a.g.nn) {
        for (int q = 0; q < 6; ++q) a.aacc[q * a.g.nn + I] = 0;
        for (int q = 3; q < 5; ++q) a.nb[q * a.g.nn + I] = 0ull;
    }
}

// P2G of mass, momentum and force (…_MP.cpp:106-202).  BOUNDS: count the
// terms of every node and take each field's largest term; else add them.
template <bool BOUNDS>
__global__ __launch_bounds__(256) void kd_p2g(DArgs a)
{
    constexpr int NF = BOUNDS ? 4 : 7;   // count | max m, P, f  or  m | P xyz | f xyz
    __shared__ int sbox[6];
    __shared__ unsigned long long lds[kWinNodes * NF];
    const int n = a.n;
    const long long nn = a.g.nn;
    bool valid;
    const int p = block_point(a.items, valid);
    double x[3] = {0.0, 0.0, 0.0};
    int i0[3] = {0, 0, 0};
    if (valid) {
        ld_x(a, p, x);
        valid = cell_of(a.g, x, i0);
        if (!valid && BOUNDS) atomicOr(a.err, 1);
    }
    const Win w = block_window(valid, i0, sbox);
    for (int j = threadIdx.x; j < w.nodes * NF; j += 256) lds[j] = 0ull;
    __syncthreads();
    if (valid) {
        Bases B;
        bases(a.g, x, i0, B);
        double F[9], s[6], v[3], fe[3];
        for (int q = 0; q < 9; ++q) F[q] = ld_F(a, q, p);
        for (int q = 0; q < 6; ++q) s[q] = ld_s(a, q, p);
        for (int q = 0; q < 3; ++q) {
            v[q] = ld_v(a, q, p);
            fe[q] = a.fext[q * n + p];
        }
        const double m = a.m[p], V = det3(F) * a.V0[p];
        for (int k = 0; k < 8; ++k) {
            double N, G[3], wt[3], gr[3], sg[3], tP[3], tf[3];
            corner(B, k, N, G, wt, gr);
            const long long I = nidx(a.g, i0, k);
            const int j = w.lds ? win_slot(w, i0, k) : 0;
            sigma_dot(s, G, sg);
            const double tm = N * m;
            for (int q = 0; q < 3; ++q) {
                tP[q] = N * m * v[q];
                tf[q] = -V * sg[q] + N * fe[q];
            }
            if (BOUNDS) {
                const unsigned long long b[3] = {dbits(fabs(tm)),
                                                 dbits(fmax(fabs(tP[0]), fmax(fabs(tP[1]), fabs(tP[2])))),
                                                 dbits(fmax(fabs(tf[0]), fmax(fabs(tf[1]), fabs(tf[2]))))};
                if (w.lds) {
                    atomicAdd(&lds[j], 1ull);
                    for (int q = 0; q < 3; ++q) atomicMax(&lds[(1 + q) * w.nodes + j], b[q]);
                } else {
                    atomicAdd(&a.cnt[I], 1u);
                    for (int q = 0; q < 3; ++q) atomicMax(&a.nb[q * nn + I], b[q]);
                }
            } else {
                const double sm = a.nsc[I], sp = a.nsc[nn + I], sf = a.nsc[2 * nn + I];
                long long iv[7];
                iv[0] = fixp(tm, sm);
                for (int q = 0; q < 3; ++q) {
                    iv[1 + q] = fixp(tP[q], sp);
                    iv[4 + q] = fixp(tf[q], sf);
                }
                for (int q = 0; q < 7; ++q) {
                    if (w.lds) atomicAdd(&lds[q * w.nodes + j], (unsigned long long)iv[q]);
                    else atomicAdd((unsigned long long *)&a.acc[q * nn + I], (unsigned long long)iv[q]);
                }
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < w.nodes; j += 256) {
        const long long I = win_node(a.g, w, j);
        if (BOUNDS) {
            const unsigned long long c = lds[j];
            if (c) atomicAdd(&a.cnt[I], (unsigned)c);
            for (int q = 0; q < 3; ++q) {
                const unsigned long long b = lds[(1 + q) * w.nodes + j];
                if (b) atomicMax(&a.nb[q * nn + I], b);
            }
        } else {
            for (int q = 0; q < 7; ++q) {
                const unsigned long long v = lds[q * w.nodes + j];
                if (v) atomicAdd((unsigned long long *)&a.acc[q * nn + I], v);
            }
        }
    }
}

// per node: the scales of fields q0 .. q0 + nq - 1 from their bounds
__global__ __launch_bounds__(256) void kd_scales(DArgs a, int q0, int nq)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I >= a.g.nn) return;
    for (int q = q0; q < q0 + nq; ++q) a.nsc[q * a.g.nn + I] = node_scale(a, q, I);
}

__global__ __launch_bounds__(256) void kd_marks(DArgs a)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.n_dc) return;
    int i0[3];
    for (int q = 0; q < 3; ++q) i0[q] = (int)(a.xdc[3 * j + q] / a.g.h[q]);
    for (int k = 0; k < 8; ++k) {
        const int n0 = i0[0] + ((k >> 2) & 1), n1 = i0[1] + ((k >> 1) & 1), n2 = i0[2] + (k & 1);
        if (n0 < 0 || n1 < 0 || n2 < 0 || n0 >= a.g.nodes[0] || n1 >= a.g.nodes[1] || n2 >= a.g.nodes[2]) continue;
        atomicMax(&a.marks[n0 + (long long)a.g.nodes[0] * (n1 + (long long)a.g.nodes[1] * n2)], (unsigned long long)j + 1ull);
    }
}

// grid momentum, damping, boundary conditions and platen override
// (…_MP.cpp:209-265)
__global__ __launch_bounds__(256) void kd_grid(DArgs a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I >= a.g.nn) return;
    const long long nn = a.g.nn;
    const double um = inv_scale(a.nsc[I]), up = inv_scale(a.nsc[nn + I]), uf = inv_scale(a.nsc[2 * nn + I]);
    const double m = (double)a.acc[I] * um;
    double P[3], f[3], vh[3], fd[3], vt[3], ac[3];
    for (int q = 0; q < 3; ++q) {
        P[q] = (double)a.acc[(1 + q) * nn + I] * up;
        f[q] = (double)a.acc[(4 + q) * nn + I] * uf;
    }
    const bool gate = m > a.g.m_min;
    for (int q = 0; q < 3; ++q) {
        vh[q] = gate ? P[q] / m : 0.0;
        fd[q] = f[q];
    }
    const double vv = vh[0] * vh[0] + vh[1] * vh[1] + vh[2] * vh[2];
    if (sqrt(vv) > 1.0e-9) {
        const double lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        const double inv = 1.0 / sqrt(vv), sc = a.damping * lf;
        for (int q = 0; q < 3; ++q) fd[q] -= sc * (vh[q] * inv);
    }
    for (int q = 0; q < 3; ++q) vt[q] = gate ? vh[q] + fd[q] / m * a.dt : vh[q];
    const uint8_t fx = a.fixed[I];
    for (int q = 0; q < 3; ++q)
        if (fx & (1u << q)) { vt[q] = 0.0; fd[q] = 0.0; }
    uint8_t flag = gate ? 1 : 0;
    const unsigned long long mk = a.marks[I];
    if (mk) {
        const int j = (int)(mk - 1ull);
        for (int q = 0; q < 3; ++q) { vt[q] = a.vdc[3 * j + q]; fd[q] = 0.0; }
        flag |= 2;
    }
    for (int q = 0; q < 3; ++q) ac[q] = (fabs(m) > a.g.m_min) ? fd[q] / m : 0.0;
    a.nm[I] = m;
    for (int q = 0; q < 3; ++q) {
        a.nP[q * nn + I] = P[q];
        a.nf[q * nn + I] = f[q];
        a.nvt[q * nn + I] = vt[q];
        a.na[q * nn + I] = ac[q];
    }
    a.nflag[I] = flag;
}

__global__ __launch_bounds__(256) void kd_g2p(DArgs a)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.n) return;
    const int n = a.n;
    const long long nn = a.g.nn;
    double x[3];
    ld_x(a, p, x);
    int i0[3];
    if (!cell_of(a.g, x, i0)) return;
    Bases B;
    bases(a.g, x, i0, B);
    double dx[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; ++k) {
        double N, G[3], w[3], gr[3];
        corner(B, k, N, G, w, gr);
        const long long I = nidx(a.g, i0, k);
        for (int q = 0; q < 3; ++q) {
            const double vt = a.nvt[q * nn + I];
            dx[q] += N * vt;
            dv[q] += N * a.na[q * nn + I];
            for (int c = 0; c < 3; ++c) L[3 * c + q] += vt * G[c];
        }
    }
    double F[9], T[9], Fn[9], s[6], de[6], ds[6];
    for (int q = 0; q < 9; ++q) F[q] = ld_F(a, q, p);
    matmul(L, F, T);
    for (int q = 0; q < 9; ++q) {
        Fn[q] = F[q] + T[q] * a.dt;
        if (a.fp32) a.D32[q * n + p] = (float)((double)a.D32[q * n + p] + T[q] * a.dt);
        else a.F[q * n + p] = Fn[q];
    }
    for (int q = 0; q < 3; ++q) {
        if (a.fp32) {
            a.u32[q * n + p] = (float)((double)a.u32[q * n + p] + dx[q] * a.dt);
            a.v32[q * n + p] = (float)((double)a.v32[q * n + p] + dv[q] * a.dt);
        } else {
            a.x[q * n + p] = x[q] + dx[q] * a.dt;
            a.u[q * n + p] += dx[q] * a.dt;
            a.v[q * n + p] += dv[q] * a.dt;
        }
    }
    eps_of(L, a.dt, de);
    for (int q = 0; q < 6; ++q) s[q] = ld_s(a, q, p);
    // constitutive update (…_MP.cpp:322-353)
    const int law = a.law[p];
    if (law == LAW_ELASTIC) {
        hooke(a.E[p], a.nu[p], de, ds);
    } else if (law == LAW_NEOHOOKEAN) {
        double P[9];
        nh_stress(a.E[p], a.nu[p], Fn, P);
        ds[0] = P[0] - s[0];
        ds[1] = P[4] - s[1];
        ds[2] = P[8] - s[2];
        ds[3] = P[1] - s[3];
        ds[4] = P[5] - s[4];
        ds[5] = P[6] - s[5];
    } else {
        RR r;
        const double al = ld_al(a, p);
        rr_eval(a.E[p], a.nu[p], a.sy[p], a.C0[p], a.C1[p], law == LAW_VONMISES, s, al, de, r);
        for (int q = 0; q < 6; ++q) ds[q] = r.ds[q];
        if (law == LAW_VONMISES && r.plastic) {
            if (a.fp32) a.al32[p] = (float)(al + r.da);
            else a.al[p] = al + r.da;
        }
    }
    for (int q = 0; q < 6; ++q) {
        if (a.fp32) a.s32[q * n + p] = (float)(s[q] + ds[q]);
        else a.s[q * n + p] = s[q] + ds[q];
    }
}

// platen motion at the end of the step (…_MP.cpp:360-368)
__global__ __launch_bounds__(256) void kd_dc(DArgs a)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.n_dc) return;
    for (int q = 0; q < 3; ++q) {
        if (a.has_tl) a.vdc[3 * j + q] = a.vtl[q];
        a.xdc[3 * j + q] += a.vdc[3 * j + q] * a.dt;
    }
}

// FP32 state at step 0: u = 0, v = v0, D = 0, sigma = 0, alpha = alpha0
__global__ __launch_bounds__(256) void kd_init32(DArgs a, const double *v0, const double *al0)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.n) return;
    const int n = a.n;
    for (int q = 0; q < 3; ++q) {
        a.u32[q * n + p] = 0.0f;
        a.v32[q * n + p] = (float)v0[q * n + p];
    }
    for (int q = 0; q < 9; ++q) a.D32[q * n + p] = 0.0f;
    for (int q = 0; q < 6; ++q) a.s32[q * n + p] = 0.0f;
    a.al32[p] = (float)al0[p];
}

// --------------------------------------------------------------- adjoint
// G2P and the material update, per point: in xb, vb, Fb, sb = adjoints of
// the state after the step; out = adjoints of the state before it (partial
// in x: the P2G terms are added by kd_a_p2g).  Writes the node scatter
// inputs.
__global__ __launch_bounds__(256) void kd_a_g2p(DArgs a)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.n) return;
    const int n = a.n;
    const long long nn = a.g.nn;
    const double dt = a.dt;
    double x[3];
    ld_x(a, p, x);
    int i0[3];
    if (!cell_of(a.g, x, i0)) return;
    Bases B;
    bases(a.g, x, i0, B);
    double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; ++k) {
        double N, G[3], w[3], gr[3];
        corner(B, k, N, G, w, gr);
        const long long I = nidx(a.g, i0, k);
        for (int q = 0; q < 3; ++q)
            for (int c = 0; c < 3; ++c) L[3 * c + q] += a.nvt[q * nn + I] * G[c];
    }
    double F[9], Fb[9], sbv[6], xb[3], vb[3];
    for (int q = 0; q < 9; ++q) { F[q] = ld_F(a, q, p); Fb[q] = a.Fb[q * n + p]; }
    for (int q = 0; q < 6; ++q) sbv[q] = a.sb[q * n + p];
    for (int q = 0; q < 3; ++q) { xb[q] = a.xb[q * n + p]; vb[q] = a.vb[q * n + p]; }
    // the constitutive update: E-bar, eps-bar, and sigma-bar, alpha-bar before
    // the update (elastic: sigma-bar is unchanged)
    const double E = a.E[p], nu = a.nu[p];
    double de[6], eb[6], gE = 0.0;
    eps_of(L, dt, de);
    const int law = a.law[p];
    if (law == LAW_ELASTIC) {
        // sigma' = sigma + C(E) eps(L dt): eps-bar = C^T sigma-bar'
        double ds[6];
        hooke(E, nu, de, ds);
        for (int q = 0; q < 6; ++q) gE += sbv[q] * ds[q] / E;
        hooke(E, nu, sbv, eb);   // the normal block is symmetric; shear: C (1/2 - nu)
    } else if (law == LAW_NEOHOOKEAN) {
        // sigma' = P(F'), F' = F + dt L F: F'-bar gains dP/dF'^T sigma-bar';
        // sigma' does not depend on sigma or on eps
        double T[9], Fp[9];
        matmul(L, F, T);
        for (int q = 0; q < 9; ++q) Fp[q] = F[q] + T[q] * dt;
        nh_adjoint(E, nu, Fp, sbv, Fb, gE);
        for (int q = 0; q < 6; ++q) {
            eb[q] = 0.0;
            a.sb[q * n + p] = 0.0;
        }
    } else {
        double s[6], sbo[6], abo;
        for (int q = 0; q < 6; ++q) s[q] = ld_s(a, q, p);
        RR r;
        rr_eval(E, nu, a.sy[p], a.C0[p], a.C1[p], law == LAW_VONMISES, s, ld_al(a, p), de, r);
        rr_adjoint(E, nu, a.C0[p], law == LAW_VONMISES, de, r, sbv, a.alb[p], sbo, abo, eb, gE);
        for (int q = 0; q < 6; ++q) a.sb[q * n + p] = sbo[q];
        a.alb[p] = abo;
    }
    a.Eb[p] += gE;
    // L-bar from eps (engineering shear) and from F' = F + dt L F
    double Lb[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Lb[3 * c + r] = dt * (Fb[r] * F[c] + Fb[3 + r] * F[3 + c] + Fb[6 + r] * F[6 + c]);
    Lb[0] += dt * eb[0];
    Lb[4] += dt * eb[1];
    Lb[8] += dt * eb[2];
    Lb[1] += dt * eb[3];
    Lb[3] += dt * eb[3];
    Lb[5] += dt * eb[4];
    Lb[7] += dt * eb[4];
    Lb[6] += dt * eb[5];
    Lb[2] += dt * eb[5];
    // F-bar = F-bar' + dt L^T F-bar'
    double Fn[9];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Fn[3 * c + r] = Fb[3 * c + r] + dt * (L[3 * r] * Fb[3 * c] + L[3 * r + 1] * Fb[3 * c + 1] + L[3 * r + 2] * Fb[3 * c + 2]);
    // x-bar through the bases of x', v' and L
    double xn[3] = {xb[0], xb[1], xb[2]};
    for (int k = 0; k < 8; ++k) {
        double N, G[3], w[3], gr[3];
        corner(B, k, N, G, w, gr);
        const long long I = nidx(a.g, i0, k);
        double vt[3], ac[3];
        for (int q = 0; q < 3; ++q) { vt[q] = a.nvt[q * nn + I]; ac[q] = a.na[q * nn + I]; }
        const double cx = dt * (vt[0] * xb[0] + vt[1] * xb[1] + vt[2] * xb[2] + ac[0] * vb[0] + ac[1] * vb[1] + ac[2] * vb[2]);
        for (int q = 0; q < 3; ++q) xn[q] += cx * G[q];
        // L_rc = sum vt_r dN/dx_c: x_a-bar += sum_rc Lb_rc vt_r d2N/dx_a dx_c
        for (int aa = 0; aa < 3; ++aa) {
            double t = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double hc = hess(w, gr, aa, c);
                if (hc != 0.0) t += hc * (Lb[3 * c] * vt[0] + Lb[3 * c + 1] * vt[1] + Lb[3 * c + 2] * vt[2]);
            }
            xn[aa] += t;
        }
    }
    for (int q = 0; q < 9; ++q) a.Fb[q * n + p] = Fn[q];
    for (int q = 0; q < 3; ++q) a.xb[q * n + p] = xn[q];
    // scatter inputs: L-bar and dt x-bar'; v-bar is unchanged (v-bar = v-bar')
    double *sc = a.scratch + 12 * (size_t)p;
    for (int q = 0; q < 9; ++q) sc[q] = Lb[q];
    for (int q = 0; q < 3; ++q) sc[9 + q] = dt * xb[q];
}

// scatter: vt-bar_I += N dt x-bar' + L-bar . grad N,  a-bar_I += N dt v-bar
// (BOUNDS: each node's largest term, as in kd_p2g; the term counts are the
// forward's, same points and cells)
template <bool BOUNDS>
__global__ __launch_bounds__(256) void kd_a_scatter(DArgs a)
{
    constexpr int NF = BOUNDS ? 2 : 6;   // max vt-bar, a-bar  or  vt-bar xyz | a-bar xyz
    __shared__ int sbox[6];
    __shared__ unsigned long long lds[kWinNodes * NF];
    const int n = a.n;
    const long long nn = a.g.nn;
    bool valid;
    const int p = block_point(a.items, valid);
    double x[3] = {0.0, 0.0, 0.0};
    int i0[3] = {0, 0, 0};
    if (valid) {
        ld_x(a, p, x);
        valid = cell_of(a.g, x, i0);
    }
    const Win w = block_window(valid, i0, sbox);
    for (int j = threadIdx.x; j < w.nodes * NF; j += 256) lds[j] = 0ull;
    __syncthreads();
    if (valid) {
        Bases B;
        bases(a.g, x, i0, B);
        const double *sc = a.scratch + 12 * (size_t)p;
        double vb[3];
        for (int q = 0; q < 3; ++q) vb[q] = a.vb[q * n + p] * a.dt;
        for (int k = 0; k < 8; ++k) {
            double N, G[3], wt[3], gr[3], t[3], ta[3];
            corner(B, k, N, G, wt, gr);
            const long long I = nidx(a.g, i0, k);
            const int j = w.lds ? win_slot(w, i0, k) : 0;
            for (int r = 0; r < 3; ++r) {
                t[r] = N * sc[9 + r] + sc[r] * G[0] + sc[3 + r] * G[1] + sc[6 + r] * G[2];
                ta[r] = N * vb[r];
            }
            if (BOUNDS) {
                const unsigned long long b[2] = {dbits(fmax(fabs(t[0]), fmax(fabs(t[1]), fabs(t[2])))),
                                                 dbits(fmax(fabs(ta[0]), fmax(fabs(ta[1]), fabs(ta[2]))))};
                for (int q = 0; q < 2; ++q) {
                    if (w.lds) atomicMax(&lds[q * w.nodes + j], b[q]);
                    else atomicMax(&a.nb[(3 + q) * nn + I], b[q]);
                }
            } else {
                const double s3 = a.nsc[3 * nn + I], s4 = a.nsc[4 * nn + I];
                long long iv[6];
                for (int r = 0; r < 3; ++r) {
                    iv[r] = fixp(t[r], s3);
                    iv[3 + r] = fixp(ta[r], s4);
                }
                for (int q = 0; q < 6; ++q) {
                    if (w.lds) atomicAdd(&lds[q * w.nodes + j], (unsigned long long)iv[q]);
                    else atomicAdd((unsigned long long *)&a.aacc[q * nn + I], (unsigned long long)iv[q]);
                }
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < w.nodes; j += 256) {
        const long long I = win_node(a.g, w, j);
        for (int q = 0; q < NF; ++q) {
            const unsigned long long v = lds[q * w.nodes + j];
            if (!v) continue;
            if (BOUNDS) atomicMax(&a.nb[(3 + q) * nn + I], v);
            else atomicAdd((unsigned long long *)&a.aacc[q * nn + I], v);
        }
    }
}

// grid update reversed, per node: (vt-bar, a-bar) -> (m-bar, P-bar, f-bar)
__global__ __launch_bounds__(256) void kd_a_grid(DArgs a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I >= a.g.nn) return;
    const long long nn = a.g.nn;
    const double u3 = inv_scale(a.nsc[3 * nn + I]), u4 = inv_scale(a.nsc[4 * nn + I]);
    double vtb[3], ab[3], P[3], f[3];
    for (int q = 0; q < 3; ++q) {
        vtb[q] = (double)a.aacc[q * nn + I] * u3;
        ab[q] = (double)a.aacc[(3 + q) * nn + I] * u4;
        P[q] = a.nP[q * nn + I];
        f[q] = a.nf[q * nn + I];
    }
    const double m = a.nm[I];
    double mb = 0.0, Pb[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
    const uint8_t flag = a.nflag[I];
    if (!(flag & 2)) {   // platen nodes take the platen velocity: no dependence
        const bool gate = flag & 1;
        const uint8_t fx = a.fixed[I];
        for (int q = 0; q < 3; ++q)
            if (fx & (1u << q)) { vtb[q] = 0.0; ab[q] = 0.0; }
        // forward recomputed: vh = P/m, fd = f - c |f| vh/|vh|
        double vh[3], fd[3], hat_v[3] = {0, 0, 0};
        for (int q = 0; q < 3; ++q) {
            vh[q] = gate ? P[q] / m : 0.0;
            fd[q] = f[q];
        }
        const double vv = vh[0] * vh[0] + vh[1] * vh[1] + vh[2] * vh[2];
        const bool damped = sqrt(vv) > 1.0e-9;
        const double lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        const double inv = damped ? 1.0 / sqrt(vv) : 0.0;
        if (damped)
            for (int q = 0; q < 3; ++q) {
                hat_v[q] = vh[q] * inv;
                fd[q] -= a.damping * lf * hat_v[q];
            }
        double fdb[3] = {0, 0, 0}, vhb[3] = {0, 0, 0};
        // a = fd / m
        if (fabs(m) > a.g.m_min)
            for (int q = 0; q < 3; ++q) {
                fdb[q] += ab[q] / m;
                mb -= ab[q] * fd[q] / (m * m);
            }
        // vt = vh + fd / m dt
        if (gate)
            for (int q = 0; q < 3; ++q) {
                vhb[q] += vtb[q];
                fdb[q] += vtb[q] / m * a.dt;
                mb -= vtb[q] * fd[q] * a.dt / (m * m);
            }
        for (int q = 0; q < 3; ++q) fb[q] = fdb[q];
        if (damped) {
            const double fdv = fdb[0] * hat_v[0] + fdb[1] * hat_v[1] + fdb[2] * hat_v[2];
            if (lf > 0.0)
                for (int q = 0; q < 3; ++q) fb[q] -= a.damping * fdv * f[q] / lf;
            // hat-bar = -c |f| fd-bar; vh-bar += (I - hat hat^T) hat-bar / |vh|
            for (int q = 0; q < 3; ++q) vhb[q] += -a.damping * lf * (fdb[q] - fdv * hat_v[q]) * inv;
        }
        if (gate)
            for (int q = 0; q < 3; ++q) {
                Pb[q] = vhb[q] / m;
                mb -= vhb[q] * P[q] / (m * m);
            }
    }
    a.nmb[I] = mb;
    for (int q = 0; q < 3; ++q) {
        a.nPb[q * nn + I] = Pb[q];
        a.nfb[q * nn + I] = fb[q];
    }
}

// P2G reversed, per point (gather): v-bar, sigma-bar, F-bar (through V) and
// the P2G part of x-bar
__global__ __launch_bounds__(256) void kd_a_p2g(DArgs a)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= a.n) return;
    const int n = a.n;
    const long long nn = a.g.nn;
    double x[3];
    ld_x(a, p, x);
    int i0[3];
    if (!cell_of(a.g, x, i0)) return;
    Bases B;
    bases(a.g, x, i0, B);
    double F[9], s[6], v[3], fe[3];
    for (int q = 0; q < 9; ++q) F[q] = ld_F(a, q, p);
    for (int q = 0; q < 6; ++q) s[q] = ld_s(a, q, p);
    for (int q = 0; q < 3; ++q) {
        v[q] = ld_v(a, q, p);
        fe[q] = a.fext[q * n + p];
    }
    const double m = a.m[p], det = det3(F), V = det * a.V0[p];
    double xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0}, sb[6] = {0, 0, 0, 0, 0, 0}, Vb = 0.0;
    for (int k = 0; k < 8; ++k) {
        double N, G[3], w[3], gr[3];
        corner(B, k, N, G, w, gr);
        const long long I = nidx(a.g, i0, k);
        const double mb = a.nmb[I];
        double Pb[3], fb[3];
        for (int q = 0; q < 3; ++q) { Pb[q] = a.nPb[q * nn + I]; fb[q] = a.nfb[q * nn + I]; }
        for (int q = 0; q < 3; ++q) vb[q] += N * m * Pb[q];
        const double cs = mb * m + m * (v[0] * Pb[0] + v[1] * Pb[1] + v[2] * Pb[2]) +
                          (fb[0] * fe[0] + fb[1] * fe[1] + fb[2] * fe[2]);
        for (int q = 0; q < 3; ++q) xb[q] += cs * G[q];
        // f_I = -V sigma.G
        double sg[3];
        sigma_dot(s, G, sg);
        Vb -= fb[0] * sg[0] + fb[1] * sg[1] + fb[2] * sg[2];
        sb[0] -= V * fb[0] * G[0];
        sb[1] -= V * fb[1] * G[1];
        sb[2] -= V * fb[2] * G[2];
        sb[3] -= V * (fb[0] * G[1] + fb[1] * G[0]);
        sb[4] -= V * (fb[1] * G[2] + fb[2] * G[1]);
        sb[5] -= V * (fb[0] * G[2] + fb[2] * G[0]);
        // x-bar from d(sigma.G)/dx: f_i = -V sum_j sigma_ij dN/dx_j
        const double S[9] = {s[0], s[3], s[5], s[3], s[1], s[4], s[5], s[4], s[2]};   // S[3i + j]
        for (int aa = 0; aa < 3; ++aa) {
            double t = 0.0;
            for (int j = 0; j < 3; ++j) {
                const double hj = hess(w, gr, aa, j);
                if (hj != 0.0) t += hj * (fb[0] * S[j] + fb[1] * S[3 + j] + fb[2] * S[6 + j]);
            }
            xb[aa] -= V * t;
        }
    }
    for (int q = 0; q < 3; ++q) {
        a.xb[q * n + p] += xb[q];
        a.vb[q * n + p] += vb[q];
    }
    for (int q = 0; q < 6; ++q) a.sb[q * n + p] += sb[q];
    // V = det(F) V0: F-bar += V-bar V0 det(F) F^-T (cofactor matrix)
    const double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                           F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                           F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    // cof[3c + r] = d det / d F[3c + r]
    for (int q = 0; q < 9; ++q) a.Fb[q * n + p] += Vb * a.V0[p] * cof[q];
}

}  // namespace

// ================================================================ handle
struct mpm_diff {
    hipStream_t stream = nullptr;
    std::string err;
    DGrid g{};
    int device = 0, n = 0, n_dc = 0, n_total = 0, steps = 0, ckpt = 1, n_ckpt = 0;
    double dt = 0.0, damping = 0.0;
    std::vector<double> tl_t, tl_v, times;   // times[k]: start time of step k
    std::vector<int> def_ids, dc_ids;
    // device; the state arrays x, u, v, F, s, xdc, vdc are slices of st
    // (FP32 state: u32, v32, D32, s32, al32, xdc, vdc; x0d holds x0)
    bool fp32 = false;
    double *st = nullptr, *x = nullptr, *u = nullptr, *v = nullptr, *F = nullptr, *s = nullptr, *al = nullptr;
    float *u32 = nullptr, *v32 = nullptr, *D32 = nullptr, *s32 = nullptr, *al32 = nullptr;
    double *x0d = nullptr;
    double *m = nullptr, *V0 = nullptr, *E = nullptr, *nu = nullptr, *fext = nullptr;
    uint8_t *law = nullptr;
    double *sy = nullptr, *C0 = nullptr, *C1 = nullptr, *al0 = nullptr, *alb = nullptr;
    uint8_t *fixed = nullptr;
    double *xdc = nullptr, *vdc = nullptr;
    double *x0dc = nullptr, *v0dc = nullptr;   // platen state at step 0
    double *v0 = nullptr, *F0 = nullptr;       // initial point velocities, identity F
    unsigned long long *marks = nullptr, *nb = nullptr;
    double *nsc = nullptr;
    long long *acc = nullptr, *aacc = nullptr;
    double *nmb = nullptr, *nPb = nullptr, *nfb = nullptr;
    // node quantities of a step (m, P, f, vt, a, term counts, flags) in slots
    // of node_bytes: slot 0 in the forward, slot k - k0 for step k of the
    // segment being replayed, which the adjoint of step k then reads
    uint8_t *nodebuf = nullptr;
    size_t node_bytes = 0;
    int2 *items = nullptr;
    int n_items = 0;
    int32_t *errd = nullptr;
    double *xb = nullptr, *vb = nullptr, *Fb = nullptr, *sb = nullptr, *Eb = nullptr, *scratch = nullptr;
    double *ck = nullptr;      // checkpoints: n_ckpt x state
    double *seg = nullptr;     // one segment: ckpt x state
    bool have_forward = false;
    // doubles of one saved state: x, u, v (3 n each), F (9 n), sigma (6 n),
    // alpha (n), platen x and v (3 n_dc each); FP32 state: 22 floats per
    // point (u, v, D, sigma, alpha) = 11 doubles, platen x and v in FP64
    size_t state() const { return (fp32 ? 11 : 25) * (size_t)n + 6 * (size_t)n_dc; }
};

namespace {

int dfail(mpm_diff *d, const char *msg, int code)
{
    d->err = msg;
    return code;
}
#define DCHK(x)                                                        \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) return dfail(d, hipGetErrorString(e_), MPM_E_HIP); \
    } while (0)
template <typename T>
int dalloc_(mpm_diff *d, T **p, size_t n)
{
    if (hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return dfail(d, "hipMalloc", MPM_E_NOMEM);
    return MPM_OK;
}
inline unsigned gsz(size_t n) { return (unsigned)((n + 255) / 256); }

DArgs args_of(mpm_diff *d, int step, int slot)
{
    DArgs a{};
    a.g = d->g;
    a.n = d->n;
    a.n_dc = d->n_dc;
    a.items = d->items;
    a.dt = d->dt;
    a.damping = d->damping;
    a.fext = d->fext;
    a.x = d->x; a.u = d->u; a.v = d->v; a.F = d->F; a.s = d->s; a.al = d->al;
    a.fp32 = d->fp32 ? 1 : 0;
    a.x0 = d->x0d;
    a.u32 = d->u32; a.v32 = d->v32; a.D32 = d->D32; a.s32 = d->s32; a.al32 = d->al32;
    a.m = d->m; a.V0 = d->V0; a.E = d->E; a.nu = d->nu;
    a.law = d->law; a.sy = d->sy; a.C0 = d->C0; a.C1 = d->C1;
    a.fixed = d->fixed;
    a.xdc = d->xdc; a.vdc = d->vdc;
    // TimeLine::getVelocity at the step's start time (TimeLine.h:33-46)
    const double t = d->times[step];
    a.vtl[0] = a.vtl[1] = a.vtl[2] = 0.0;
    const int nt = (int)d->tl_t.size();
    for (int i = 0; i < nt - 1; ++i)
        if (d->tl_t[i] <= t && t < d->tl_t[i + 1]) {
            const double sc = (t - d->tl_t[i]) / (d->tl_t[i + 1] - d->tl_t[i]);
            for (int q = 0; q < 3; ++q) a.vtl[q] = (1.0 - sc) * d->tl_v[3 * i + q] + sc * d->tl_v[3 * (i + 1) + q];
        }
    a.has_tl = nt > 0;
    a.marks = d->marks; a.acc = d->acc;
    const size_t nn = d->g.nn;
    double *nd = (double *)(d->nodebuf + (size_t)slot * d->node_bytes);
    a.nm = nd; a.nP = nd + nn; a.nf = nd + 4 * nn; a.nvt = nd + 7 * nn; a.na = nd + 10 * nn;
    a.cnt = (unsigned *)(nd + 13 * nn);
    a.nflag = (uint8_t *)(a.cnt + nn);
    a.nb = d->nb; a.nsc = d->nsc; a.err = d->errd;
    a.xb = d->xb; a.vb = d->vb; a.Fb = d->Fb; a.sb = d->sb; a.Eb = d->Eb; a.alb = d->alb; a.scratch = d->scratch;
    a.aacc = d->aacc; a.nmb = d->nmb; a.nPb = d->nPb; a.nfb = d->nfb;
    return a;
}

// node quantities of the step from the current point state (P2G + grid)
int forward_nodes(mpm_diff *d, const DArgs &a)
{
    const long long nn = d->g.nn;
    kd_clear<<<gsz(nn), 256, 0, d->stream>>>(a);
    if (d->n > 0) {
        kd_p2g<true><<<d->n_items, 256, 0, d->stream>>>(a);
        kd_scales<<<gsz(nn), 256, 0, d->stream>>>(a, 0, 3);
        kd_p2g<false><<<d->n_items, 256, 0, d->stream>>>(a);
    }
    if (d->n_dc > 0) kd_marks<<<gsz(d->n_dc), 256, 0, d->stream>>>(a);
    kd_grid<<<gsz(nn), 256, 0, d->stream>>>(a);
    DCHK(hipGetLastError());
    return MPM_OK;
}

// one step: node quantities into node slot `slot`, then (advance) G2P and
// the platen motion
int forward_step(mpm_diff *d, int step, int slot, bool advance)
{
    DArgs a = args_of(d, step, slot);
    int rc = forward_nodes(d, a);
    if (rc || !advance) return rc;
    kd_g2p<<<gsz(d->n), 256, 0, d->stream>>>(a);
    if (d->n_dc > 0) kd_dc<<<gsz(d->n_dc), 256, 0, d->stream>>>(a);
    DCHK(hipGetLastError());
    return MPM_OK;
}

// copy the state (x, u, v, F, sigma, alpha, platen x and v: one contiguous block)
// to buf (save) or from it (load), stream-ordered
int move_state(mpm_diff *d, double *buf, bool save)
{
    DCHK(hipMemcpyAsync(save ? buf : d->st, save ? d->st : buf, sizeof(double) * d->state(), hipMemcpyDeviceToDevice,
                        d->stream));
    return MPM_OK;
}

// 3D Morton key of a cell (21 bits per axis)
uint64_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    auto spread = [](uint64_t v) {
        v &= 0x1FFFFFull;
        v = (v | v << 32) & 0x1F00000000FFFFull;
        v = (v | v << 16) & 0x1F0000FF0000FFull;
        v = (v | v << 8) & 0x100F00F00F00F00Full;
        v = (v | v << 4) & 0x10C30C30C30C30C3ull;
        v = (v | v << 2) & 0x1249249249249249ull;
        return v;
    };
    return spread(x) | spread(y) << 1 | spread(z) << 2;
}

void free_all(mpm_diff *d)
{
    void *ptrs[] = {d->st, d->x0d, d->m, d->V0, d->E, d->nu, d->fext, d->fixed, d->x0dc, d->v0dc, d->v0, d->F0, d->marks, d->nb, d->nsc, d->acc, d->aacc, d->nodebuf, d->items,
                    d->nmb, d->nPb, d->nfb, d->errd, d->xb, d->vb, d->Fb, d->sb, d->Eb,
                    d->scratch, d->ck, d->seg, d->law, d->sy, d->C0, d->C1, d->al0, d->alb};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (d->stream) (void)hipStreamDestroy(d->stream);
}

}  // namespace

extern "C" {

int mpm_diff_create(const mpm_config *cfg, const mpm_points *pts, int32_t n_steps, int32_t checkpoint, mpm_diff **out)
{
    if (!cfg || !pts || !out || n_steps < 1 || checkpoint < 1 || pts->n < 1 || !pts->x || !pts->mass || !pts->nu ||
        (!pts->vol0 && !pts->vol) || (cfg->n_timeline > 0 && (!cfg->timeline_t || !cfg->timeline_v)))
        return MPM_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MPM_E_NO_DEVICE;
    if (cfg->device < 0 || cfg->device >= ndev || hipSetDevice(cfg->device) != hipSuccess) return MPM_E_ARG;
    if (cfg->precision != MPM_FP64 && cfg->precision != MPM_FP32) return MPM_E_ARG;
    mpm_diff *d = new mpm_diff();
    d->device = cfg->device;
    d->fp32 = cfg->precision == MPM_FP32;
    auto fail = [&](int code) {
        free_all(d);
        delete d;
        return code;
    };
    for (int a = 0; a < 3; ++a) {
        if (!(cfg->length[a] > 0.0) || cfg->cells[a] < 1) return fail(MPM_E_ARG);
        d->g.L[a] = cfg->length[a];
        d->g.cells[a] = cfg->cells[a];
        d->g.nodes[a] = cfg->cells[a] + 1;
        d->g.h[a] = cfg->length[a] / (double)cfg->cells[a];
    }
    if (!(cfg->dt_max > 0.0)) return fail(MPM_E_ARG);
    d->g.nn = (long long)d->g.nodes[0] * d->g.nodes[1] * d->g.nodes[2];
    d->g.m_min = cfg->mass_min;
    d->dt = cfg->dt_max;
    d->damping = cfg->damping;
    // start time of each step, accumulated as the reference's d_Time (…_MP.cpp:381-393)
    d->times.resize(n_steps);
    double t = cfg->time_start;
    for (int k = 0; k < n_steps; ++k) {
        d->times[k] = t;
        t += cfg->dt_max;
    }
    d->steps = n_steps;
    d->ckpt = std::min(checkpoint, n_steps);
    d->n_ckpt = (n_steps + d->ckpt - 1) / d->ckpt;
    d->tl_t.assign(cfg->timeline_t, cfg->timeline_t + cfg->n_timeline);
    d->tl_v.assign(cfg->timeline_v, cfg->timeline_v + 3 * cfg->n_timeline);
    d->n_total = pts->n;
    for (int i = 0; i < pts->n; ++i) ((pts->dc && pts->dc[i]) ? d->dc_ids : d->def_ids).push_back(i);
    d->n = (int)d->def_ids.size();
    d->n_dc = (int)d->dc_ids.size();
    if (d->n < 1) {
        d->err = "differentiable run: no deformable points";
        return fail(MPM_E_ARG);
    }
    // laws after the reference's dispatch (…_MP.cpp:332-339): NeoHookean only
    // when cfg->neohookean, ViscoElastic and IdealGas fall to PerfectlyPlastic
    std::vector<uint8_t> hlaw(d->n);
    for (int k = 0; k < d->n; ++k) {
        const int t = pts->type ? pts->type[d->def_ids[k]] : MPM_ELASTIC;
        if (t < MPM_ELASTIC || t > MPM_GASS) {
            d->err = "differentiable run: unknown material type";
            return fail(MPM_E_ARG);
        }
        hlaw[k] = (t == MPM_ELASTIC) ? LAW_ELASTIC : (t == MPM_NEOHOOKEAN && cfg->neohookean) ? LAW_NEOHOOKEAN
                  : (t == MPM_VONMISESHARDENING) ? LAW_VONMISES : LAW_PLASTIC;
        if ((hlaw[k] == LAW_PLASTIC && !pts->sy) || (hlaw[k] == LAW_VONMISES && (!pts->sy || !pts->C0 || !pts->C1))) {
            d->err = "differentiable run: plastic points need sy (and C0, C1 for hardening)";
            return fail(MPM_E_ARG);
        }
    }
    // device order: by brick of the starting cell, then Morton order of the
    // cell inside it (ties: input order); scatter work items of at most 256
    // points of one brick
    std::vector<int2> items;
    {
        std::vector<uint64_t> bkey(pts->n, 0), ckey(pts->n, 0);
        for (int i : d->def_ids) {
            uint32_t c[3];
            for (int a = 0; a < 3; ++a) {
                const double f = pts->x[3 * i + a] / d->g.h[a];
                c[a] = f > 0.0 ? (uint32_t)std::min<double>(f, (double)(d->g.cells[a] - 1)) : 0u;
            }
            bkey[i] = morton3(c[0] / kBrick[0], c[1] / kBrick[1], c[2] / kBrick[2]);
            ckey[i] = morton3(c[0] % kBrick[0], c[1] % kBrick[1], c[2] % kBrick[2]);
        }
        std::stable_sort(d->def_ids.begin(), d->def_ids.end(), [&](int p, int q) {
            return bkey[p] != bkey[q] ? bkey[p] < bkey[q] : ckey[p] < ckey[q];
        });
        for (int k = 0; k < d->n; ++k) {
            const bool same = !items.empty() && bkey[d->def_ids[k]] == bkey[d->def_ids[k - 1]] && items.back().y < 256;
            if (same) ++items.back().y;
            else items.push_back(make_int2(k, 1));
        }
        d->n_items = (int)items.size();
    }
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) return fail(MPM_E_HIP);
    const size_t n = d->n, nd = d->n_dc, nn = d->g.nn, state = d->state();
    int rc;
    // node slot: m, P, f, vt, a (13 doubles), term count (u32), flags (u8)
    d->node_bytes = ((13 * 8 + 4 + 1) * nn + 255) / 256 * 256;
    if ((rc = dalloc_(d, &d->st, state)) || (rc = dalloc_(d, &d->m, n)) || (rc = dalloc_(d, &d->V0, n)) || (rc = dalloc_(d, &d->E, n)) || (rc = dalloc_(d, &d->nu, n)) ||
        (rc = dalloc_(d, &d->fext, 3 * n)) || (rc = dalloc_(d, &d->fixed, nn)) || (rc = dalloc_(d, &d->x0dc, 3 * nd)) || (rc = dalloc_(d, &d->v0dc, 3 * nd)) ||
        (rc = dalloc_(d, &d->v0, 3 * n)) || (rc = dalloc_(d, &d->F0, 9 * n)) || (rc = dalloc_(d, &d->marks, nn)) ||
        (rc = dalloc_(d, &d->nb, 5 * nn)) || (rc = dalloc_(d, &d->nsc, 5 * nn)) || (rc = dalloc_(d, &d->acc, 7 * nn)) ||
        (rc = dalloc_(d, &d->aacc, 6 * nn)) || (rc = dalloc_(d, &d->nodebuf, (size_t)d->ckpt * d->node_bytes)) ||
        (rc = dalloc_(d, &d->items, items.size())) || (rc = dalloc_(d, &d->nmb, nn)) || (rc = dalloc_(d, &d->nPb, 3 * nn)) || (rc = dalloc_(d, &d->nfb, 3 * nn)) ||
        (rc = dalloc_(d, &d->errd, 1)) || (rc = dalloc_(d, &d->xb, 3 * n)) ||
        (rc = dalloc_(d, &d->vb, 3 * n)) || (rc = dalloc_(d, &d->Fb, 9 * n)) || (rc = dalloc_(d, &d->sb, 6 * n)) ||
        (rc = dalloc_(d, &d->Eb, n)) || (rc = dalloc_(d, &d->scratch, 12 * n)) || (rc = dalloc_(d, &d->law, n)) ||
        (rc = dalloc_(d, &d->sy, n)) || (rc = dalloc_(d, &d->C0, n)) || (rc = dalloc_(d, &d->C1, n)) ||
        (rc = dalloc_(d, &d->al0, n)) || (rc = dalloc_(d, &d->alb, n)) ||
        (rc = dalloc_(d, &d->ck, (size_t)d->n_ckpt * state)) || (rc = dalloc_(d, &d->seg, (size_t)d->ckpt * state)) ||
        (d->fp32 && (rc = dalloc_(d, &d->x0d, 3 * n))))
        return fail(rc);
    if (d->fp32) {
        d->u32 = reinterpret_cast<float *>(d->st);
        d->v32 = d->u32 + 3 * n;
        d->D32 = d->v32 + 3 * n;
        d->s32 = d->D32 + 9 * n;
        d->al32 = d->s32 + 6 * n;
        d->xdc = d->st + 11 * n;
    } else {
        d->x = d->st;
        d->u = d->x + 3 * n;
        d->v = d->u + 3 * n;
        d->F = d->v + 3 * n;
        d->s = d->F + 9 * n;
        d->al = d->s + 6 * n;
        d->xdc = d->al + n;
    }
    d->vdc = d->xdc + 3 * nd;
    // static point data, component-major (blocking copies from pageable memory)
    std::vector<double> tmp(9 * n);
    auto up = [&](double *dst, size_t count) { return hipMemcpy(dst, tmp.data(), sizeof(double) * count, hipMemcpyHostToDevice); };
    for (size_t k = 0; k < n; ++k) tmp[k] = pts->mass[d->def_ids[k]];
    if (up(d->m, n) != hipSuccess) return fail(MPM_E_HIP);
    for (size_t k = 0; k < n; ++k) tmp[k] = pts->vol0 ? pts->vol0[d->def_ids[k]] : pts->vol[d->def_ids[k]];
    if (up(d->V0, n) != hipSuccess) return fail(MPM_E_HIP);
    for (size_t k = 0; k < n; ++k) tmp[k] = pts->nu[d->def_ids[k]];
    if (up(d->nu, n) != hipSuccess) return fail(MPM_E_HIP);
    const double *pp[4] = {pts->sy, pts->C0, pts->C1, pts->alpha};
    double *dp[4] = {d->sy, d->C0, d->C1, d->al0};
    for (int f = 0; f < 4; ++f) {
        for (size_t k = 0; k < n; ++k) tmp[k] = pp[f] ? pp[f][d->def_ids[k]] : 0.0;
        if (up(dp[f], n) != hipSuccess) return fail(MPM_E_HIP);
    }
    if (hipMemcpy(d->law, hlaw.data(), n, hipMemcpyHostToDevice) != hipSuccess) return fail(MPM_E_HIP);
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) tmp[a * n + k] = pts->v ? pts->v[3 * d->def_ids[k] + a] : 0.0;
    if (up(d->v0, 3 * n) != hipSuccess) return fail(MPM_E_HIP);
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) tmp[a * n + k] = pts->fext ? pts->fext[3 * d->def_ids[k] + a] : 0.0;
    if (up(d->fext, 3 * n) != hipSuccess) return fail(MPM_E_HIP);
    std::fill(tmp.begin(), tmp.end(), 0.0);
    for (size_t k = 0; k < n; ++k) tmp[k] = tmp[4 * n + k] = tmp[8 * n + k] = 1.0;
    if (up(d->F0, 9 * n) != hipSuccess) return fail(MPM_E_HIP);
    std::vector<double> dx(3 * nd), dv(3 * nd);
    for (size_t j = 0; j < nd; ++j)
        for (int a = 0; a < 3; ++a) {
            dx[3 * j + a] = pts->x[3 * d->dc_ids[j] + a];
            dv[3 * j + a] = pts->v ? pts->v[3 * d->dc_ids[j] + a] : 0.0;
        }
    if (nd && (hipMemcpy(d->x0dc, dx.data(), sizeof(double) * 3 * nd, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(d->v0dc, dv.data(), sizeof(double) * 3 * nd, hipMemcpyHostToDevice) != hipSuccess))
        return fail(MPM_E_HIP);
    if (hipMemcpy(d->items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(MPM_E_HIP);
    std::vector<uint8_t> fx(nn, 0);
    if (cfg->node_fixed) std::memcpy(fx.data(), cfg->node_fixed, nn);
    if (hipMemcpy(d->fixed, fx.data(), nn, hipMemcpyHostToDevice) != hipSuccess) return fail(MPM_E_HIP);
    *out = d;
    return MPM_OK;
}

int mpm_diff_forward(mpm_diff *d, const double *E, const double *x0, double *u_final)
{
    if (!d || !E || !x0 || !u_final) return MPM_E_ARG;
    const size_t n = d->n, nd = d->n_dc, state = d->state();
    d->have_forward = false;
    DCHK(hipSetDevice(d->device));
    DCHK(hipStreamSynchronize(d->stream));
    // parameters and initial state: blocking uploads, then stream-ordered
    // device copies and clears
    std::vector<double> hE(n), hx(3 * n);
    for (size_t k = 0; k < n; ++k) {
        hE[k] = E[d->def_ids[k]];
        for (int a = 0; a < 3; ++a) hx[a * n + k] = x0[3 * d->def_ids[k] + a];
    }
    DCHK(hipMemcpy(d->E, hE.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    if (d->fp32) {
        // u = 0, v = v0, D = 0, sigma = 0, alpha = alpha0, rounded once
        DCHK(hipMemcpy(d->x0d, hx.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        kd_init32<<<gsz(n), 256, 0, d->stream>>>(args_of(d, 0, 0), d->v0, d->al0);
        DCHK(hipGetLastError());
    } else {
        DCHK(hipMemcpy(d->x, hx.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        DCHK(hipMemsetAsync(d->u, 0, sizeof(double) * 3 * n, d->stream));
        DCHK(hipMemcpyAsync(d->v, d->v0, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, d->stream));
        DCHK(hipMemcpyAsync(d->F, d->F0, sizeof(double) * 9 * n, hipMemcpyDeviceToDevice, d->stream));
        DCHK(hipMemsetAsync(d->s, 0, sizeof(double) * 6 * n, d->stream));
        DCHK(hipMemcpyAsync(d->al, d->al0, sizeof(double) * n, hipMemcpyDeviceToDevice, d->stream));
    }
    if (nd) {
        DCHK(hipMemcpyAsync(d->xdc, d->x0dc, sizeof(double) * 3 * nd, hipMemcpyDeviceToDevice, d->stream));
        DCHK(hipMemcpyAsync(d->vdc, d->v0dc, sizeof(double) * 3 * nd, hipMemcpyDeviceToDevice, d->stream));
    }
    DCHK(hipMemsetAsync(d->errd, 0, 4, d->stream));
    for (int k = 0; k < d->steps; ++k) {
        int rc;
        if (k % d->ckpt == 0 && (rc = move_state(d, d->ck + (size_t)(k / d->ckpt) * state, true))) return rc;
        if ((rc = forward_step(d, k, 0, true))) return rc;
    }
    int32_t err = 0;
    std::vector<double> pdc(3 * nd), p0(3 * nd);
    DCHK(hipMemcpyAsync(&err, d->errd, 4, hipMemcpyDeviceToHost, d->stream));
    std::vector<float> hu32(d->fp32 ? 3 * n : 0);
    if (d->fp32) DCHK(hipMemcpyAsync(hu32.data(), d->u32, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, d->stream));
    else DCHK(hipMemcpyAsync(hx.data(), d->u, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, d->stream));
    if (nd) {
        DCHK(hipMemcpyAsync(pdc.data(), d->xdc, sizeof(double) * 3 * nd, hipMemcpyDeviceToHost, d->stream));
        DCHK(hipMemcpyAsync(p0.data(), d->x0dc, sizeof(double) * 3 * nd, hipMemcpyDeviceToHost, d->stream));
    }
    DCHK(hipStreamSynchronize(d->stream));
    if (err) return dfail(d, "a material point left the grid box", MPM_E_OUT_OF_GRID);
    if (d->fp32)
        for (size_t k = 0; k < 3 * n; ++k) hx[k] = (double)hu32[k];
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) u_final[3 * d->def_ids[k] + a] = hx[a * n + k];
    for (size_t j = 0; j < nd; ++j)   // platen displacements (no parameters)
        for (int a = 0; a < 3; ++a) u_final[3 * d->dc_ids[j] + a] = pdc[3 * j + a] - p0[3 * j + a];
    d->have_forward = true;
    return MPM_OK;
}

int mpm_diff_backward(mpm_diff *d, const double *g_ufinal, double *g_E, double *g_x0)
{
    if (!d || !g_ufinal || !g_E || !g_x0) return MPM_E_ARG;
    if (!d->have_forward) return dfail(d, "backward before forward", MPM_E_STATE);
    const size_t n = d->n, state = d->state();
    // The running position adjoint starts at dL/du_final: positions and
    // displacements take the same increments, so with y = x-bar + u-bar the
    // step adjoint is that of x alone, and dL/dx0 = y_0 - dL/du_final.
    std::vector<double> g(3 * n);
    for (size_t k = 0; k < n; ++k)
        for (int a = 0; a < 3; ++a) g[a * n + k] = g_ufinal[3 * d->def_ids[k] + a];
    DCHK(hipSetDevice(d->device));
    DCHK(hipStreamSynchronize(d->stream));
    DCHK(hipMemcpy(d->xb, g.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    DCHK(hipMemsetAsync(d->vb, 0, sizeof(double) * 3 * n, d->stream));
    DCHK(hipMemsetAsync(d->Fb, 0, sizeof(double) * 9 * n, d->stream));
    DCHK(hipMemsetAsync(d->sb, 0, sizeof(double) * 6 * n, d->stream));
    DCHK(hipMemsetAsync(d->Eb, 0, sizeof(double) * n, d->stream));
    DCHK(hipMemsetAsync(d->alb, 0, sizeof(double) * n, d->stream));
    for (int c = d->n_ckpt - 1; c >= 0; --c) {
        const int k0 = c * d->ckpt, k1 = std::min(d->steps, k0 + d->ckpt);
        // replay the segment from its checkpoint, keeping every step's state
        int rc = move_state(d, d->ck + (size_t)c * state, false);
        if (rc) return rc;
        for (int k = k0; k < k1; ++k) {
            if ((rc = move_state(d, d->seg + (size_t)(k - k0) * state, true))) return rc;
            if ((rc = forward_step(d, k, k - k0, k + 1 < k1))) return rc;
        }
        // adjoint steps, last first, on the replayed node quantities
        for (int k = k1 - 1; k >= k0; --k) {
            if ((rc = move_state(d, d->seg + (size_t)(k - k0) * state, false))) return rc;
            DArgs a = args_of(d, k, k - k0);
            kd_a_clear<<<gsz(d->g.nn), 256, 0, d->stream>>>(a);
            kd_a_g2p<<<gsz(n), 256, 0, d->stream>>>(a);
            kd_a_scatter<true><<<d->n_items, 256, 0, d->stream>>>(a);
            kd_scales<<<gsz(d->g.nn), 256, 0, d->stream>>>(a, 3, 2);
            kd_a_scatter<false><<<d->n_items, 256, 0, d->stream>>>(a);
            kd_a_grid<<<gsz(d->g.nn), 256, 0, d->stream>>>(a);
            kd_a_p2g<<<gsz(n), 256, 0, d->stream>>>(a);
            DCHK(hipGetLastError());
        }
    }
    std::vector<double> ge(n), gx(3 * n);
    DCHK(hipMemcpyAsync(ge.data(), d->Eb, sizeof(double) * n, hipMemcpyDeviceToHost, d->stream));
    DCHK(hipMemcpyAsync(gx.data(), d->xb, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, d->stream));
    DCHK(hipStreamSynchronize(d->stream));
    for (int i = 0; i < d->n_total; ++i) g_E[i] = 0.0;
    for (int i = 0; i < 3 * d->n_total; ++i) g_x0[i] = 0.0;
    for (size_t k = 0; k < n; ++k) {
        g_E[d->def_ids[k]] = ge[k];
        for (int a = 0; a < 3; ++a) g_x0[3 * d->def_ids[k] + a] = gx[a * n + k] - g[a * n + k];
    }
    return MPM_OK;
}

const char *mpm_diff_last_error(const mpm_diff *d) { return d ? d->err.c_str() : "null handle"; }

void mpm_diff_destroy(mpm_diff *d)
{
    if (!d) return;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    free_all(d);
    delete d;
}

}  // extern "C"
