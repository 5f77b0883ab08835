// mpm_contact_kernels.hpp -- the two-field contact step
// (PhysicsEngine_runSimulation_Classic_SinglePass_MP_Contact.cpp:30-646).
//
// One launch per phase over points or grid nodes.  Every node sum is a 64-bit
// fixed-point integer add (global atomics), with per-field scales fixed for
// the step by bounds over the points' contributions, formed before the step's
// first scatter (by the previous step's G2P, or by the kernel-density pass in
// the first step of a call), so the result is independent of scheduling.  Order-dependent choices of the reference's
// sequential loops are restated as order-free rules:
//   nearest point of a node   = smallest (distance, point index), starting
//                               from point 0 (…_MP_Contact.cpp:228-233)
//   surface kernel gradient   = largest |grad k| of a surface point, ties to
//                               the smallest index (…:246-249)
//   platen override           = the highest platen index wins, the node's
//                               force is moved to the reaction once (…:494-518)
// Node values are FP64 for both state precisions; point state is read and
// written in its own precision.
#pragma once
#include "mpm_step_kernels.hpp"

namespace mpm {

constexpr uint32_t CF_ACTIVE = 1, CF_KACTIVE = 2, CF_CPOS = 4, CF_CNEG = 8;
constexpr uint32_t CF_GATE0 = 16;   // CF_GATE0 << f: field f's node mass passes the m_min gate
constexpr int kCB = 6;   // bounds: m, m/h, m|v|, |force|, 1 (kernel), spare

template <typename real>
struct CArgs {
    GridParams g;
    PArrays<real> p;
    int32_t n_def, n_dc, n_all;
    const uint8_t *surf;          // per deformable point (current order)
    const double *dc_x, *dc_v, *dc_m, *dc_vol, *dc_s, *dc_fext;
    double *dc_x_out;             // platen positions after the step (may alias dc_x)
    const uint8_t *dc_surf;
    const uint32_t *dc_id;
    const uint32_t *def_id0;      // (is platen, index) of original point 0: [0] flag, [1] index
    uint8_t surf0;                // surface flag of original point 0
    long long nn;                 // node count
    const uint8_t *fixed;         // node fixity bits
    // node data
    long long *acc;               // 21 x nn int64: k | m[2] | mg[2][3] | mv[2][3] | f[2][3]
    double *gkgrad, *gkg;         // 3 x nn each
    double *gm, *gv, *gf;         // 2 x nn, 6 x nn, 6 x nn (field-major: field, component)
    uint32_t *flags;              // CF_* per node
    unsigned long long *near_d, *sg_d;
    unsigned long long *near_i, *sg_i;   // (point id << 32) | current index of the winner
    uint8_t *near_surf;
    uint8_t *fmask;               // per point (current order, platens after deformable): field of each corner
    double *ga;                   // 6 x nn: gated node acceleration f / m per field (field-major)
    unsigned long long *cmarks;   // platen tags of this step
    unsigned long long *bounds;   // kCB positive doubles as bits, this step's
    unsigned long long *bounds_next;   // the next step's, maxed by the G2P over the new point state
    int32_t def_bounds;           // 1: the kernel pass forms the deformable points' bounds itself
    double *kp, *kg;              // per point (current order, platens after deformable): k, grad k (3 x n_all)
    const Material *mats;
    const StepRec *steps;
    int32_t *gflags;
    double *cpart;                // per node block: mass, reaction, active nodes
    double *ppart;                // per point block: sum m|v|^2, m v.xyz, nonfinite
    const uint32_t *slot_order, *occ_start, *occ_count;   // brick blocks: slot per block, point range per slot
    int32_t n_bb;                 // brick blocks (occupied bricks); platen blocks of 256 points follow
    int32_t n_xcd;                // XCDs the dispatcher deals workgroups to, round robin by block index
    int32_t H;                    // headroom bits of the fixed-point sums
    double mu;                    // Coulomb coefficient of the contact correction (0: frictionless)
};

// ---- point access ---------------------------------------------------------
template <typename real>
__device__ __forceinline__ void c_pos(const CArgs<real> &a, int j, double x[3])
{
    if (j >= a.n_def) {
        const int d = j - a.n_def;
        x[0] = a.dc_x[3 * d];
        x[1] = a.dc_x[3 * d + 1];
        x[2] = a.dc_x[3 * d + 2];
        return;
    }
    if constexpr (sizeof(real) == 8) {
        x[0] = a.p.x[0][j];
        x[1] = a.p.x[1][j];
        x[2] = a.p.x[2][j];
    } else {
        const uint32_t b = a.p.pbrick[j];
        const int bc[3] = {(int)(b % a.g.nb[0]), (int)((b / a.g.nb[0]) % a.g.nb[1]), (int)(b / (a.g.nb[0] * a.g.nb[1]))};
#pragma unroll
        for (int q = 0; q < 3; ++q)
            x[q] = ((double)(bc[q] * a.g.S[q]) + (double)a.p.xq[q][j] * (1.0 / kQOne)) * a.g.h[q];
    }
}
template <typename real>
__device__ __forceinline__ uint32_t c_id(const CArgs<real> &a, int j)
{
    return j >= a.n_def ? a.dc_id[j - a.n_def] : a.p.id[j];
}
template <typename real>
__device__ __forceinline__ uint8_t c_surf(const CArgs<real> &a, int j)
{
    return j >= a.n_def ? a.dc_surf[j - a.n_def] : a.surf[j];
}

__device__ __forceinline__ double c_len3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
__device__ __forceinline__ double c_dot3(const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }
__device__ __forceinline__ double c_lenm3(double m) { return sqrt(m * m + m * m + m * m); }

// FAST (FP32 state): the contact helpers multiply by 1/h and place nodes at
// i h instead of dividing; FP64 state keeps the reference's expressions.
template <typename real>
constexpr bool kCFast = sizeof(real) == 4;

// GridPoint_Factory.h:40-42 node position (i*L)/n
template <bool FAST>
__device__ __forceinline__ void c_node_pos(const GridParams &g, int ix, int iy, int iz, double xi[3])
{
    const int ii[3] = {ix, iy, iz};
#pragma unroll
    for (int q = 0; q < 3; ++q) xi[q] = FAST ? (double)ii[q] * g.h[q] : (double)ii[q] * g.L[q] / (double)g.cells[q];
}

// linear-hat N and grad N of point x at node xi (GridPoint_Mediator.h:64-88)
__device__ __forceinline__ void c_bases(const GridParams &g, const double x[3], const double xi[3], double &N, double G[3])
{
    double s[3], sg[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double d = x[q] - xi[q];
        double v = 1.0 - fabs(d) / g.h[q];
        if (v < 0.0 || v > 1.0) v = 0.0;
        s[q] = v;
        sg[q] = sgn(d);
    }
    N = s[0] * s[1] * s[2];
    G[0] = -s[1] * s[2] * sg[0] / g.h[0];
    G[1] = -s[0] * s[2] * sg[1] / g.h[1];
    G[2] = -s[0] * s[1] * sg[2] / g.h[2];
}

// the cell of x (its lowest node i0); false if out of grid
template <bool FAST>
__device__ __forceinline__ bool c_cell(const GridParams &g, const double x[3], int i0[3])
{
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (!(x[q] >= 0.0)) return false;
        i0[q] = (int)(FAST ? x[q] * g.ih[q] : x[q] / g.h[q]);
        if (!(i0[q] < g.cells[q])) return false;
    }
    return true;
}
// corner k of a cell (x-outer, z-inner): offsets 0/1 along x, y, z
__device__ __forceinline__ int c_kx(int k) { return (k >> 2) & 1; }
__device__ __forceinline__ int c_ky(int k) { return (k >> 1) & 1; }
__device__ __forceinline__ int c_kz(int k) { return k & 1; }
// linear node index of the cell's lowest node
__device__ __forceinline__ long long c_node0(const GridParams &g, const int i0[3])
{
    return i0[0] + (long long)g.nodes[0] * (i0[1] + (long long)g.nodes[1] * i0[2]);
}
// node index of corner k from the lowest node's
__device__ __forceinline__ long long c_cnode(const GridParams &g, long long I0, int k)
{
    return I0 + c_kx(k) + (long long)g.nodes[0] * (c_ky(k) + (long long)g.nodes[1] * c_kz(k));
}

// Per axis, the two node planes of the cell (node index i0 and i0 + 1): the
// node coordinate (i L) / n, the hat value and sign(x - x_i), each computed
// with the same expressions as c_node_pos / c_bases, once per point instead
// of once per corner.
struct CAx {
    double xi[3][2], s[3][2], sg[3][2];
};
template <bool FAST>
__device__ __forceinline__ void c_axes(const GridParams &g, const double x[3], const int i0[3], CAx &A)
{
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const double xi = FAST ? (double)(i0[q] + c) * g.h[q] : (double)(i0[q] + c) * g.L[q] / (double)g.cells[q];
            const double d = x[q] - xi;
            double v = 1.0 - (FAST ? fabs(d) * g.ih[q] : fabs(d) / g.h[q]);
            if (v < 0.0 || v > 1.0) v = 0.0;
            A.xi[q][c] = xi;
            A.s[q][c] = v;
            A.sg[q][c] = sgn(d);
        }
}
// node position, N and grad N of corner k (x-outer, z-inner), as c_bases
template <bool FAST>
__device__ __forceinline__ void c_corner(const GridParams &g, const CAx &A, int k, double xi[3], double &N, double G[3])
{
    const int c[3] = {(k >> 2) & 1, (k >> 1) & 1, k & 1};
    double s[3], sg[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        xi[q] = A.xi[q][c[q]];
        s[q] = A.s[q][c[q]];
        sg[q] = A.sg[q][c[q]];
    }
    N = s[0] * s[1] * s[2];
    if (FAST) {
        G[0] = -s[1] * s[2] * sg[0] * g.ih[0];
        G[1] = -s[0] * s[2] * sg[1] * g.ih[1];
        G[2] = -s[0] * s[1] * sg[2] * g.ih[2];
    } else {
        G[0] = -s[1] * s[2] * sg[0] / g.h[0];
        G[1] = -s[0] * s[2] * sg[1] / g.h[1];
        G[2] = -s[0] * s[1] * sg[2] / g.h[2];
    }
}

// fixed-point scale of bound slot q: contributions |c| <= bound < 2^e are
// stored as c * 2^(FRAC - e) with FRAC = 62 - H, H headroom bits
template <typename real>
__device__ __forceinline__ double c_scale(const CArgs<real> &a, int q)
{
    const double b = __longlong_as_double((long long)a.bounds[q]);
    if (!(b > 0.0)) return 0.0;
    int e;
    frexp(b * 1.0001, &e);   // b*1.0001 < 2^e
    return ldexp(1.0, 62 - a.H - e);
}
template <typename real>
__device__ __forceinline__ double c_unscale(const CArgs<real> &a, int q)
{
    const double b = __longlong_as_double((long long)a.bounds[q]);
    if (!(b > 0.0)) return 0.0;
    int e;
    frexp(b * 1.0001, &e);
    return ldexp(1.0, e - 62 + a.H);
}
// round-to-nearest-even of a scaled contribution to int64.  With H >= 11
// headroom bits every scaled contribution is below 2^51 in magnitude, and
// adding 1.5 * 2^52 leaves rint(x) in the low mantissa bits (one add and one
// integer subtract instead of the general conversion).  The add must not be
// fused with the product that feeds it.
template <bool SMALL>
__device__ __forceinline__ long long c_toll(double x)
{
#pragma clang fp contract(off)
    if constexpr (SMALL) {
        const double y = x + 6755399441055744.0;   // 1.5 * 2^52
        return __double_as_longlong(y) - 0x4338000000000000ll;
    }
    return __double2ll_rn(x);
}
__device__ __forceinline__ void c_add(long long *p, double v, double sc)
{
    atomicAdd((unsigned long long *)p, (unsigned long long)__double2ll_rn(v * sc));
}
__device__ __forceinline__ void c_umax(unsigned long long *p, double v)
{
    if (v > 0.0) atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ double wave_dmax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Node window of a block: the points of one launch are in the plain step's
// brick order, so a block's 256 points touch a small box of nodes.  The block
// adds its fixed-point terms there in LDS and flushes each node once with a
// global atomic (the sums are exact integers, so this equals adding every
// term globally).  A box beyond the window falls back to global atomics.
constexpr int kCWin = 512;
// the P2G passes' window: 416 nodes keep their LDS under 40 KB, four blocks
// per CU (a brick of 4 x 4 x 10 cells touches 275 nodes, 330 with a drifted
// layer)
constexpr int kCWinP = 416;
struct CWin {
    int lo[3], dim[3], nodes;
    bool lds;
};
// nodes lo - below .. hi + above over the cells [lo, hi] of the block's
// points (per thread: the cells of its points, any = it has one), at most
// cap.  The bounds are reduced over each wave first (DPP), so one lane per
// wave updates the block's box: per-lane atomics on one LDS word would be
// serialized lane by lane.  Call with every thread of the block.
__device__ __forceinline__ CWin c_window_box(bool any, const int lo_in[3], const int hi_in[3], int *sbox, int below,
                                             int above, int cap)
{
    if (threadIdx.x < 3) {
        sbox[threadIdx.x] = 0x7FFFFFFF;
        sbox[3 + threadIdx.x] = -1;
    }
    int lo[3], hi[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        lo[q] = wave_imin_dpp(any ? lo_in[q] : (1 << 20));
        hi[q] = wave_imax_dpp(any ? hi_in[q] : -1);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && hi[0] >= 0)
        for (int q = 0; q < 3; ++q) {
            atomicMin(&sbox[q], lo[q]);
            atomicMax(&sbox[3 + q], hi[q]);
        }
    __syncthreads();
    CWin w{};
    if (sbox[3] < 0) return w;
    long long nodes = 1;
    for (int q = 0; q < 3; ++q) {
        w.lo[q] = sbox[q] - below;
        w.dim[q] = sbox[3 + q] - sbox[q] + 1 + below + above;
        nodes *= w.dim[q];
    }
    w.lds = nodes <= cap;
    w.nodes = w.lds ? (int)nodes : 0;
    return w;
}

// Chunk of 256 consecutive points taken by this block in a launch of nblk
// chunk blocks: block b runs on XCD b mod X, so XCD x takes the x-th run of
// consecutive chunks, in order.  The chunks of one brick then share an L2
// (each stages the brick's node box).  A bijection of [0, nblk).
__device__ __forceinline__ int c_chunk(int nblk, int X)
{
    const int b = (int)blockIdx.x, x = b % X, k = b / X, q = nblk / X, r = nblk % X;
    return x * q + min(x, r) + k;
}

// The points of a block: brick blocks (the first n_bb, in the plain step's
// launch order, longest first) take all points of one occupied brick, which
// are contiguous in the binned order; platen blocks take 256 platen points
// each (indices n_def.. after the deformable points).
struct CRange {
    int j0, n;
};
template <typename real>
__device__ __forceinline__ CRange c_range(const CArgs<real> &a)
{
    const int b = (int)blockIdx.x;
    if (b < a.n_bb) {
        const uint32_t sl = a.slot_order[b];
        return CRange{(int)a.occ_start[sl], (int)a.occ_count[sl]};
    }
    const int j0 = a.n_def + (b - a.n_bb) * 256;
    return CRange{j0, min(256, a.n_all - j0)};
}
// the cell range of this thread's points of the block (first loop of a pass)
template <typename real>
__device__ __forceinline__ bool c_range_box(const CArgs<real> &a, const CRange &r, int lo[3], int hi[3])
{
    bool any = false;
    for (int q = 0; q < 3; ++q) {
        lo[q] = 1 << 20;
        hi[q] = -1;
    }
#pragma clang loop unroll(disable)
    for (int j = r.j0 + (int)threadIdx.x; j < r.j0 + r.n; j += 256) {
        double x[3];
        int i0[3];
        c_pos(a, j, x);
        if (!c_cell<kCFast<real>>(a.g, x, i0)) continue;
        any = true;
        for (int q = 0; q < 3; ++q) {
            lo[q] = min(lo[q], i0[q]);
            hi[q] = max(hi[q], i0[q]);
        }
    }
    return any;
}
__device__ __forceinline__ int c_wslot(const CWin &w, const int n[3])
{
    return (n[0] - w.lo[0]) + w.dim[0] * ((n[1] - w.lo[1]) + w.dim[1] * (n[2] - w.lo[2]));
}
// window slot of corner k from the slot of the cell's lowest node
__device__ __forceinline__ int c_cslot(const CWin &w, int s0, int k)
{
    return s0 + c_kx(k) + w.dim[0] * (c_ky(k) + w.dim[1] * c_kz(k));
}
__device__ __forceinline__ long long c_wnode(const GridParams &g, const CWin &w, int t)
{
    const int jx = t % w.dim[0], r = t / w.dim[0];
    const int jy = r % w.dim[1], jz = r / w.dim[1];
    return (w.lo[0] + jx) + (long long)g.nodes[0] * ((w.lo[1] + jy) + (long long)g.nodes[1] * (w.lo[2] + jz));
}

// accumulator offsets inside CArgs::acc (in units of nn)
__device__ __forceinline__ long long off_k() { return 0; }
__device__ __forceinline__ long long off_m(int f) { return 1 + f; }
__device__ __forceinline__ long long off_mg(int f, int c) { return 3 + 3 * f + c; }
__device__ __forceinline__ long long off_mv(int f, int c) { return 9 + 3 * f + c; }
__device__ __forceinline__ long long off_f(int f, int c) { return 15 + 3 * f + c; }

// ---- C1: reset ---------------------------------------------------------------
template <typename real>
__global__ void __launch_bounds__(256) k_c_clear(CArgs<real> a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I < a.nn) {
        for (int q = 0; q < 21; ++q) a.acc[q * a.nn + I] = 0;
        for (int q = 0; q < 3; ++q) a.gkgrad[q * a.nn + I] = a.gkg[q * a.nn + I] = 0.0;
        a.flags[I] = 0u;
        a.near_d[I] = 0x7FF0000000000000ull;   // +inf
        a.near_i[I] = ~0ull;
        a.sg_d[I] = 0ull;
        a.sg_i[I] = ~0ull;
        a.near_surf[I] = 0;
    }
    // bound slot 4 is the kernel density's (contributions <= 1).  This
    // step's bounds already hold the deformable points' maxima when the
    // previous step's G2P formed them; the next step's start from zero.
    if (blockIdx.x == 0 && threadIdx.x < kCB) {
        const unsigned long long b0 = threadIdx.x == 4 ? (unsigned long long)__double_as_longlong(1.0) : 0ull;
        if (a.def_bounds) a.bounds[threadIdx.x] = b0;
        a.bounds_next[threadIdx.x] = b0;
    }
}

// ---- bounds of the contributions of this step: per point, then block maxima
// (one atomic per block and bound, only when the block beats the current one)
__device__ __forceinline__ double c_ihm(const GridParams &g)
{
    return fmax(1.0 / g.h[0], fmax(1.0 / g.h[1], 1.0 / g.h[2]));
}
// the four bounds of a point's contributions from its mass, velocity,
// volume, stress and external force (the same expressions wherever they are
// formed, unfused, so both places give the same bits)
__device__ __forceinline__ void c_bounds_of(double ihm, double m, const double v[3], double vol, const double s[6],
                                            const double fe[3], double b[4])
{
#pragma clang fp contract(off)
    double smax = 0.0;
    for (int q = 0; q < 6; ++q) smax = fmax(smax, fabs(s[q]));
    b[0] = fabs(m);
    b[1] = fabs(m) * ihm;
    b[2] = fabs(m) * fmax(fabs(v[0]), fmax(fabs(v[1]), fabs(v[2])));
    b[3] = fabs(vol) * 3.0 * smax * ihm + fmax(fabs(fe[0]), fmax(fabs(fe[1]), fabs(fe[2])));
}
template <typename real>
__device__ __forceinline__ void c_point_bounds(const CArgs<real> &a, int j, double b[4])
{
    const GridParams &g = a.g;
    double m, v[3], vol, s[6], fe[3] = {0, 0, 0};
    if (j >= a.n_def) {
        const int d = j - a.n_def;
        m = a.dc_m[d];
        vol = a.dc_vol[d];
        for (int q = 0; q < 3; ++q) { v[q] = a.dc_v[3 * d + q]; fe[q] = a.dc_fext[3 * d + q]; }
        for (int q = 0; q < 6; ++q) s[q] = a.dc_s[6 * d + q];
    } else {
        m = (double)a.p.mass[j];
        double Fd[9];
        for (int q = 0; q < 9; ++q) Fd[q] = fd(a.p.F[q][j], q);
        vol = det3(Fd) * (double)a.p.vol0[j];
        for (int q = 0; q < 3; ++q) {
            v[q] = (double)a.p.v[q][j];
            if (g.has_fext) fe[q] = (double)a.p.fext[q][j];
        }
        for (int q = 0; q < 6; ++q) s[q] = (double)a.p.s[q][j];
    }
    c_bounds_of(c_ihm(g), m, v, vol, s, fe, b);
}
template <typename real>
__device__ __forceinline__ void c_bounds_flush(const CArgs<real> &a, const double b[4], double (*s_b)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double w = wave_dmax(b[q]);
        if ((threadIdx.x & 63) == 0) s_b[threadIdx.x >> 6][q] = w;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        const double w = fmax(fmax(s_b[0][q], s_b[1][q]), fmax(s_b[2][q], s_b[3][q]));
        const unsigned long long wb = (unsigned long long)__double_as_longlong(w);
        if (w > 0.0 && wb > __hip_atomic_load(&a.bounds[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&a.bounds[q], wb);
    }
}

// ---- C2 + C3: active nodes, kernel on the grid (…_MP_Contact.cpp:94-153),
// and the bounds of the platen points' P2G contributions (of every point's
// in the first step of a call).  The block's kernel sums
// and node flags go through an LDS window of the nodes i0 - 1 .. i0 + 2 of
// its points' cells.  Per axis only the node planes within R of the point
// (a slightly inclusive prefilter) are visited, in ascending order as the
// reference's -1..2 loops; the exact distance test decides.
constexpr int kCWinK = 1024;
template <typename real, bool SMALL>
__global__ void __launch_bounds__(256) k_c_kernel(CArgs<real> a)
{
    __shared__ int sbox[6];
    __shared__ unsigned long long s_k[kCWinK];
    __shared__ uint32_t s_f[kCWinK];
    __shared__ double s_b[4][4];
    const GridParams &g = a.g;
    const CRange r = c_range(a);
    // first loop: the bounds of the P2G contributions where formed here, the
    // box of cells
    double b[4] = {0.0, 0.0, 0.0, 0.0};
    int lo[3], hi[3];
    bool any = false;
    for (int q = 0; q < 3; ++q) {
        lo[q] = 1 << 20;
        hi[q] = -1;
    }
    int oog = 0;
#pragma clang loop unroll(disable)
    for (int j = r.j0 + (int)threadIdx.x; j < r.j0 + r.n; j += 256) {
        double bp[4], x[3];
        // the deformable points' bounds come from the previous step's G2P
        // unless this is the first step of a run
        if (j >= a.n_def || a.def_bounds) {
            c_point_bounds(a, j, bp);
            for (int q = 0; q < 4; ++q) b[q] = fmax(b[q], bp[q]);
        }
        int i0[3];
        c_pos(a, j, x);
        if (!c_cell<kCFast<real>>(g, x, i0)) {
            oog = 1;
            continue;
        }
        any = true;
        for (int q = 0; q < 3; ++q) {
            lo[q] = min(lo[q], i0[q]);
            hi[q] = max(hi[q], i0[q]);
        }
    }
    if (oog) atomicOr(a.gflags, FLAG_OUT_OF_GRID);
    c_bounds_flush(a, b, s_b);
    const CWin w = c_window_box(any, lo, hi, sbox, 1, 2, kCWinK);
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        s_k[t] = 0ull;
        s_f[t] = 0u;
    }
    __syncthreads();
    const double R = fabs(1.0 * g.h[0]);
    const double sck = c_scale(a, 4);
#pragma clang loop unroll(disable)
    for (int j = r.j0 + (int)threadIdx.x; j < r.j0 + r.n; j += 256) {
        double x[3];
        int i0[3];
        c_pos(a, j, x);
        if (!c_cell<kCFast<real>>(g, x, i0)) continue;
        {
            const long long I0 = c_node0(g, i0);
            const int s0 = w.lds ? c_wslot(w, i0) : 0;
            for (int k = 0; k < 8; ++k) {
                if (w.lds) atomicOr(&s_f[c_cslot(w, s0, k)], CF_ACTIVE);
                else atomicOr(&a.flags[c_cnode(g, I0, k)], CF_ACTIVE);
            }
        }
        // per axis, the candidate planes i0 - 1 + t (t = 0..3) inside the grid
        // and within R on that axis, as bits
        uint32_t near[3];
        for (int q = 0; q < 3; ++q) {
            near[q] = 0u;
            for (int t = 0; t < 4; ++t) {
                const int n = i0[q] + t - 1;
                const double xc = kCFast<real> ? (double)n * g.h[q] : (double)n * g.L[q] / (double)g.cells[q];
                if (n >= 0 && n <= g.nodes[q] - 1 && fabs(x[q] - xc) < R * (1.0 + 1.0e-12)) near[q] |= 1u << t;
            }
        }
        for (uint32_t mx = near[0]; mx; mx &= mx - 1) {
            const int n0 = i0[0] + __ffs(mx) - 2;
            const double d0 = x[0] - (kCFast<real> ? (double)n0 * g.h[0] : (double)n0 * g.L[0] / (double)g.cells[0]);
            for (uint32_t my = near[1]; my; my &= my - 1) {
                const int n1 = i0[1] + __ffs(my) - 2;
                const double d1 = x[1] - (kCFast<real> ? (double)n1 * g.h[1] : (double)n1 * g.L[1] / (double)g.cells[1]);
                for (uint32_t mz = near[2]; mz; mz &= mz - 1) {
                    const int n2 = i0[2] + __ffs(mz) - 2;
                    const double d2 = x[2] - (kCFast<real> ? (double)n2 * g.h[2] : (double)n2 * g.L[2] / (double)g.cells[2]);
                    const double d[3] = {d0, d1, d2};
                    const double dist = c_len3(d);
                    if (dist < R) {
                        const double q = kCFast<real> ? dist * g.ih[0] : dist / R;
                        const unsigned long long iv =
                            (unsigned long long)c_toll<SMALL>((1.0 - 3.0 * q * q + 2.0 * q * q * q) * sck);
                        const int n[3] = {n0, n1, n2};
                        if (w.lds) {
                            const int ws = c_wslot(w, n);
                            atomicOr(&s_f[ws], CF_KACTIVE);
                            atomicAdd(&s_k[ws], iv);
                        } else {
                            const long long I = n0 + (long long)g.nodes[0] * (n1 + (long long)g.nodes[1] * n2);
                            atomicOr(&a.flags[I], CF_KACTIVE);
                            atomicAdd((unsigned long long *)&a.acc[off_k() * a.nn + I], iv);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        const uint32_t f = s_f[t];
        if (!f) continue;
        const long long I = c_wnode(g, w, t);
        atomicOr(&a.flags[I], f);
        const unsigned long long v = s_k[t];
        if (v) atomicAdd((unsigned long long *)&a.acc[off_k() * a.nn + I], v);
    }
}

// ---- C4: kernel gradient by central differences (…_MP_Contact.cpp:159-191).
// Neighbours are found by linear index, range-checked on the linear index only
// (GridPoint_Factory.h:27-31): x-1 at ix = 0 wraps to the previous row.
template <typename real>
__global__ void __launch_bounds__(256) k_c_kgrad(CArgs<real> a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I >= a.nn) return;
    const double usk = c_unscale(a, 4);
    const double k = (double)a.acc[I] * usk;
    if (k < 1.0 || !(a.flags[I] & CF_KACTIVE)) return;
    const GridParams &g = a.g;
    const long long sx = 1, sy = g.nodes[0], sz = (long long)g.nodes[0] * g.nodes[1];
    const long long nb[6] = {I - sx, I + sx, I - sy, I + sy, I - sz, I + sz};
    double kn[6];
    for (int q = 0; q < 6; ++q) kn[q] = (nb[q] >= 0 && nb[q] < a.nn) ? (double)a.acc[nb[q]] * usk : 0.0;
    double gr[3] = {0.0, 0.0, 0.0};
    if (nb[0] >= 0 && nb[0] < a.nn) gr[0] += +0.5 / g.h[0] * (k - kn[0]);
    if (nb[1] >= 0 && nb[1] < a.nn) gr[0] += -0.5 / g.h[0] * (k - kn[1]);
    if (nb[2] >= 0 && nb[2] < a.nn) gr[1] += +0.5 / g.h[1] * (k - kn[2]);
    if (nb[3] >= 0 && nb[3] < a.nn) gr[1] += -0.5 / g.h[1] * (k - kn[3]);
    if (nb[4] >= 0 && nb[4] < a.nn) gr[2] += +0.5 / g.h[2] * (k - kn[4]);
    if (nb[5] >= 0 && nb[5] < a.nn) gr[2] += -0.5 / g.h[2] * (k - kn[5]);
    if (c_len3(gr) < 1.0 / g.h[0]) gr[0] = gr[1] = gr[2] = 0.0;
    for (int q = 0; q < 3; ++q) a.gkgrad[q * a.nn + I] = gr[q];
}

// ---- C5: point kernel and its gradient; nearest point and surface gradient,
// pass 1 (minimum distance, maximum |grad k|) (…_MP_Contact.cpp:199-252),
// through an LDS window of the block's nodes.  The window's kernel values and
// gradients are staged in LDS first (one load per node instead of one per
// point corner).
template <typename real>
__global__ void __launch_bounds__(256) k_c_point1(CArgs<real> a)
{
    __shared__ int sbox[6];
    __shared__ unsigned long long s_near[kCWin], s_sg[kCWin];
    __shared__ double s_kv[4][kCWin];
    const GridParams &g = a.g;
    const CRange r = c_range(a);
    int lo[3], hi[3];
    const bool any = c_range_box(a, r, lo, hi);
    const CWin w = c_window_box(any, lo, hi, sbox, 0, 1, kCWin);
    const double usk = c_unscale(a, 4);
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        s_near[t] = 0x7FF0000000000000ull;   // +inf
        s_sg[t] = 0ull;
        const long long I = c_wnode(g, w, t);
        s_kv[0][t] = (double)a.acc[I] * usk;
        for (int q = 0; q < 3; ++q) s_kv[1 + q][t] = a.gkgrad[q * a.nn + I];
    }
    __syncthreads();
#pragma clang loop unroll(disable)
    for (int j = r.j0 + (int)threadIdx.x; j < r.j0 + r.n; j += 256) {
        double x[3];
        int i0[3];
        c_pos(a, j, x);
        if (!c_cell<kCFast<real>>(g, x, i0)) continue;
        // corner nodes and window slots are formed per corner from the
        // cell's lowest node, which keeps fewer registers live
        const long long I0 = c_node0(g, i0);
        const int s0 = w.lds ? c_wslot(w, i0) : 0;
        const bool surf = c_surf(a, j) != 0;
        double kp = 0.0, kg[3] = {0.0, 0.0, 0.0};
        CAx A;
        c_axes<kCFast<real>>(g, x, i0, A);
        for (int k = 0; k < 8; ++k) {
            double xi[3], N, G[3];
            c_corner<kCFast<real>>(g, A, k, xi, N, G);
            const double d[3] = {x[0] - xi[0], x[1] - xi[1], x[2] - xi[2]};
            const unsigned long long dk = (unsigned long long)__double_as_longlong(c_len3(d));
            if (w.lds) {
                const int ws = c_cslot(w, s0, k);
                kp += N * s_kv[0][ws];
                for (int q = 0; q < 3; ++q) kg[q] += N * s_kv[1 + q][ws];
                atomicMin(&s_near[ws], dk);
            } else {
                const long long I = c_cnode(g, I0, k);
                kp += N * ((double)a.acc[I] * usk);
                for (int q = 0; q < 3; ++q) kg[q] += N * a.gkgrad[q * a.nn + I];
                atomicMin(&a.near_d[I], dk);
            }
        }
        a.kp[j] = kp;
        for (int q = 0; q < 3; ++q) a.kg[q * a.n_all + j] = kg[q];
        if (surf) {
            const double lk = c_len3(kg);
            if (lk > 0.0) {
                const unsigned long long lb = (unsigned long long)__double_as_longlong(lk);
                for (int k = 0; k < 8; ++k) {
                    if (w.lds) atomicMax(&s_sg[c_cslot(w, s0, k)], lb);
                    else atomicMax(&a.sg_d[c_cnode(g, I0, k)], lb);
                }
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        const unsigned long long dn = s_near[t], sg = s_sg[t];
        if (dn == 0x7FF0000000000000ull && sg == 0ull) continue;
        const long long I = c_wnode(g, w, t);
        if (dn != 0x7FF0000000000000ull) atomicMin(&a.near_d[I], dn);
        if (sg) atomicMax(&a.sg_d[I], sg);
    }
}

// pass 2: the smallest point id among the points at the minimum distance
// (largest |grad k|), kept with the winner's current index as the 64-bit key
// (id << 32) | index; the window's minima are staged in LDS
template <typename real>
__global__ void __launch_bounds__(256) k_c_point2(CArgs<real> a)
{
    __shared__ int sbox[6];
    __shared__ unsigned long long s_nd[kCWin], s_sd[kCWin], s_ni[kCWin], s_si[kCWin];
    const int j = c_chunk((int)gridDim.x, a.n_xcd) * 256 + threadIdx.x;
    const GridParams &g = a.g;
    double x[3] = {0.0, 0.0, 0.0};
    int i0[3] = {0, 0, 0};
    bool valid = j < a.n_all;
    if (valid) {
        c_pos(a, j, x);
        valid = c_cell<kCFast<real>>(g, x, i0);
    }
    // the point's own inputs first: their latency overlaps the window set-up
    unsigned long long key = 0ull, lk = 0ull;
    bool surf = false;
    if (valid) {
        key = ((unsigned long long)c_id(a, j) << 32) | (uint32_t)j;
        const double kg[3] = {a.kg[j], a.kg[a.n_all + j], a.kg[2 * a.n_all + j]};
        surf = c_surf(a, j) != 0;
        lk = (unsigned long long)__double_as_longlong(c_len3(kg));
    }
    const CWin w = c_window_box(valid, i0, i0, sbox, 0, 1, kCWin);
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        const long long I = c_wnode(g, w, t);
        s_nd[t] = a.near_d[I];
        s_sd[t] = a.sg_d[I];
        s_ni[t] = ~0ull;
        s_si[t] = ~0ull;
    }
    __syncthreads();
    if (valid) {
        CAx A;
        c_axes<kCFast<real>>(g, x, i0, A);
        const long long I0 = c_node0(g, i0);
        const int s0 = w.lds ? c_wslot(w, i0) : 0;
        for (int k = 0; k < 8; ++k) {
            double xi[3], N, G[3];
            c_corner<kCFast<real>>(g, A, k, xi, N, G);
            const double d[3] = {x[0] - xi[0], x[1] - xi[1], x[2] - xi[2]};
            const unsigned long long db = (unsigned long long)__double_as_longlong(c_len3(d));
            if (w.lds) {
                const int ws = c_cslot(w, s0, k);
                if (db == s_nd[ws]) atomicMin(&s_ni[ws], key);
                if (surf && lk != 0ull && lk == s_sd[ws]) atomicMin(&s_si[ws], key);
            } else {
                const long long I = c_cnode(g, I0, k);
                if (db == a.near_d[I]) atomicMin(&a.near_i[I], key);
                if (surf && lk != 0ull && lk == a.sg_d[I]) atomicMin(&a.sg_i[I], key);
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        const unsigned long long ni = s_ni[t], si = s_si[t];
        if (ni == ~0ull && si == ~0ull) continue;
        const long long I = c_wnode(g, w, t);
        if (ni != ~0ull) atomicMin(&a.near_i[I], ni);
        if (si != ~0ull) atomicMin(&a.sg_i[I], si);
    }
}

// pass 3, per node: the winners' surface flag and kernel gradient.  The
// reference starts every node's nearest point at point 0, so point 0 wins
// where it is at least as close as the nearest touching point.
template <typename real>
__global__ void __launch_bounds__(256) k_c_publish(CArgs<real> a)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I >= a.nn) return;
    const unsigned long long ni = a.near_i[I], si = a.sg_i[I];
    if (si != ~0ull) {
        const int j = (int)(uint32_t)si;
        for (int q = 0; q < 3; ++q) a.gkg[q * a.nn + I] = a.kg[q * a.n_all + j];
    }
    if (ni != ~0ull) {
        const GridParams &g = a.g;
        const int j = (int)(uint32_t)ni;
        const long long r = I / g.nodes[0];
        double xi[3], x0[3];
        c_node_pos<kCFast<real>>(g, (int)(I % g.nodes[0]), (int)(r % g.nodes[1]), (int)(r / g.nodes[1]), xi);
        c_pos(a, a.def_id0[0] ? a.n_def + (int)a.def_id0[1] : (int)a.def_id0[1], x0);
        const double d0[3] = {x0[0] - xi[0], x0[1] - xi[1], x0[2] - xi[2]};
        const bool zero_wins = (unsigned long long)__double_as_longlong(c_len3(d0)) <= a.near_d[I];
        a.near_surf[I] = zero_wins ? a.surf0 : c_surf(a, j);
    }
}

// ---- C6: contact detection (…_MP_Contact.cpp:258-282) and split P2G
// (…_MP_Contact.cpp:288-385).  MASS=true: mass and grad m; the field of each
// point corner (sign of grad k_I . grad k_p) is stored as a bit mask per
// point, and surface points flag the nodes whose nearest point is a surface
// point (CF_CPOS / CF_CNEG, the same sign).  MASS=false: momentum N m v
// (divided by the field mass, sqrt(3) gate, in the grid update) and force,
// fields from the mask.  Window slot s of a
// node: MASS: field * 4 + {m, mg.xyz}; else field * 6 + {mv.xyz, f.xyz}.
template <bool MASS>
__device__ __forceinline__ long long c_p2g_off(int s)
{
    if (MASS) {
        const int f = s >> 2, q = s & 3;
        return q == 0 ? off_m(f) : off_mg(f, q - 1);
    }
    const int f = s / 6, q = s % 6;
    return q < 3 ? off_mv(f, q) : off_f(f, q - 3);
}
// The inputs of one point for the split P2G.  MASS needs the mass, the point
// kernel gradient and the surface flag; momentum and force need v,
// V = det F V0, stress, f_ext and the field mask.
struct CP2gIn {
    double x[3], m, kg[3], v[3], vol, s[6], fe[3];
    bool surf;
    uint32_t fm;
};
// Raw stored values of a deformable point, loaded one iteration ahead of
// their use (the loads of the next point overlap the current point's work).
template <typename real>
struct CP2gRaw {
    int32_t xq[3];   // FP32 state
    uint32_t pb;     // FP32 state
    real x[3];       // FP64 state
    real m, v0, F[9], v[3], s[6], fe[3];
    double kg[3];
    uint8_t surf, fm;
};
template <typename real, bool MASS>
__device__ __forceinline__ void c_p2g_raw(const CArgs<real> &a, int j, CP2gRaw<real> &r)
{
    const PArrays<real> &P = a.p;
    if constexpr (sizeof(real) == 4) {
        for (int q = 0; q < 3; ++q) r.xq[q] = P.xq[q][j];
        r.pb = P.pbrick[j];
    } else {
        for (int q = 0; q < 3; ++q) r.x[q] = P.x[q][j];
    }
    r.m = P.mass[j];
    if constexpr (MASS) {
        for (int q = 0; q < 3; ++q) r.kg[q] = a.kg[q * a.n_all + j];
        r.surf = a.surf[j];
    } else {
        r.v0 = P.vol0[j];
        for (int q = 0; q < 9; ++q) r.F[q] = P.F[q][j];
        for (int q = 0; q < 3; ++q) {
            r.v[q] = P.v[q][j];
            r.fe[q] = a.g.has_fext ? P.fext[q][j] : real(0);
        }
        for (int q = 0; q < 6; ++q) r.s[q] = P.s[q][j];
        r.fm = a.fmask[j];
    }
}
// the same quantities as c_pos and the P2G's point loads, from the raw values
template <typename real, bool MASS>
__device__ __forceinline__ void c_p2g_in_raw(const CArgs<real> &a, const CP2gRaw<real> &r, CP2gIn &in)
{
    const GridParams &g = a.g;
    if constexpr (sizeof(real) == 4) {
        const uint32_t b = r.pb;
        const int bc[3] = {(int)(b % g.nb[0]), (int)((b / g.nb[0]) % g.nb[1]), (int)(b / (g.nb[0] * g.nb[1]))};
        for (int q = 0; q < 3; ++q) in.x[q] = ((double)(bc[q] * g.S[q]) + (double)r.xq[q] * (1.0 / kQOne)) * g.h[q];
    } else {
        for (int q = 0; q < 3; ++q) in.x[q] = r.x[q];
    }
    in.m = (double)r.m;
    if constexpr (MASS) {
        for (int q = 0; q < 3; ++q) in.kg[q] = r.kg[q];
        in.surf = r.surf != 0;
    } else {
        double Fd[9];
        for (int q = 0; q < 9; ++q) Fd[q] = fd(r.F[q], q);
        in.vol = (double)(real)(det3(Fd) * (double)r.v0);
        for (int q = 0; q < 3; ++q) {
            in.v[q] = (double)r.v[q];
            in.fe[q] = g.has_fext ? (double)r.fe[q] : 0.0;
        }
        for (int q = 0; q < 6; ++q) in.s[q] = (double)r.s[q];
        in.fm = r.fm;
    }
}
// a platen point's inputs (FP64 arrays)
template <typename real, bool MASS>
__device__ __forceinline__ void c_p2g_in_dc(const CArgs<real> &a, int j, CP2gIn &in)
{
    const int d = j - a.n_def;
    for (int q = 0; q < 3; ++q) in.x[q] = a.dc_x[3 * d + q];
    in.m = a.dc_m[d];
    if constexpr (MASS) {
        for (int q = 0; q < 3; ++q) in.kg[q] = a.kg[q * a.n_all + j];
        in.surf = a.dc_surf[d] != 0;
    } else {
        in.vol = a.dc_vol[d];
        for (int q = 0; q < 3; ++q) { in.v[q] = a.dc_v[3 * d + q]; in.fe[q] = a.dc_fext[3 * d + q]; }
        for (int q = 0; q < 6; ++q) in.s[q] = a.dc_s[6 * d + q];
        in.fm = a.fmask[j];
    }
}
// one point's split P2G into the block's window (or global memory when the
// window did not fit)
template <typename real, bool MASS, bool SMALL>
__device__ __forceinline__ void c_p2g_point(const CArgs<real> &a, const CWin &w, unsigned long long *lds,
                                            const double *s_gk, uint32_t *s_ns, int j, const CP2gIn &in,
                                            double sc0, double sc1)
{
    const GridParams &g = a.g;
    int i0[3];
    if (!c_cell<kCFast<real>>(g, in.x, i0)) return;
    uint32_t mask = 0u;
    CAx A;
    c_axes<kCFast<real>>(g, in.x, i0, A);
    const long long I0 = c_node0(g, i0);
    const int s0 = w.lds ? c_wslot(w, i0) : 0;
    for (int k = 0; k < 8; ++k) {
        const long long I = c_cnode(g, I0, k);
        double xi[3], N, G[3];
        c_corner<kCFast<real>>(g, A, k, xi, N, G);
        const int ws = w.lds ? c_cslot(w, s0, k) : 0;
        int f;
        bool ns = false;
        if constexpr (MASS) {
            double gk[3];
            if (w.lds) {
                for (int q = 0; q < 3; ++q) gk[q] = s_gk[q * kCWinP + ws];
                ns = (s_ns[ws] & 1u) != 0;
            } else {
                for (int q = 0; q < 3; ++q) gk[q] = a.gkg[q * a.nn + I];
                ns = a.near_surf[I] != 0;
            }
            f = (c_dot3(gk, in.kg) >= 0.0) ? 0 : 1;
            mask |= (uint32_t)f << k;
            if (in.surf && ns) {
                const uint32_t cf = f == 0 ? CF_CPOS : CF_CNEG;
                if (w.lds) atomicOr(&s_ns[ws], cf);
                else atomicOr(&a.flags[I], cf);
            }
        } else {
            f = (int)((in.fm >> k) & 1u);
        }
        long long iv[6];
        int base, cnt = 0;
        if (MASS) {
            base = 4 * f;
            iv[0] = c_toll<SMALL>((N * in.m) * sc0);
            for (int q = 0; q < 3; ++q) iv[1 + q] = c_toll<SMALL>((G[q] * in.m) * sc1);
            // grad m is read only by the contact correction, at nodes
            // flagged CF_CNEG, which requires a surface nearest point:
            // elsewhere its sums are never used and are not formed
            cnt = ns ? 4 : 1;
        } else {
            base = 6 * f;
            // momentum N m v; the grid update divides the sum by the field's
            // node mass where it passes the gate, which equals the
            // reference's per-contribution N m v / m_I (…_MP_Contact.cpp:
            // 355-360) summed over the node's points
            for (int q = 0; q < 3; ++q) iv[q] = c_toll<SMALL>((N * (in.m * in.v[q])) * sc0);
            const double *sv = in.s;
            const double fr[3] = {-in.vol * (G[0] * sv[0] + G[1] * sv[3] + G[2] * sv[5]) + N * in.fe[0],
                                  -in.vol * (G[1] * sv[1] + G[0] * sv[3] + G[2] * sv[4]) + N * in.fe[1],
                                  -in.vol * (G[2] * sv[2] + G[0] * sv[5] + G[1] * sv[4]) + N * in.fe[2]};
            for (int q = 0; q < 3; ++q) iv[3 + q] = c_toll<SMALL>(fr[q] * sc1);
            cnt = 6;
        }
        for (int q = 0; q < cnt; ++q) {
            if (w.lds) atomicAdd(&lds[(base + q) * w.nodes + ws], (unsigned long long)iv[q]);
            else atomicAdd((unsigned long long *)&a.acc[c_p2g_off<MASS>(base + q) * a.nn + I], (unsigned long long)iv[q]);
        }
    }
    if constexpr (MASS) a.fmask[j] = (uint8_t)mask;
}

template <typename real, bool MASS, bool SMALL>
__global__ void __launch_bounds__(256) k_c_p2g(CArgs<real> a)
{
    constexpr int NF = MASS ? 8 : 12;
    constexpr int NS = MASS ? kCWinP : 1;   // staged surface gradients and flags (MASS only)
    __shared__ int sbox[6];
    __shared__ unsigned long long lds[kCWinP * NF];
    __shared__ double s_gk[3 * NS];   // MASS: staged surface gradients, component q at q * kCWinP
    __shared__ uint32_t s_ns[NS];   // MASS: bit 0 near_surf (staged); bits CF_CPOS | CF_CNEG (detection)
    const GridParams &g = a.g;
    const CRange r = c_range(a);
    const int end = r.j0 + r.n;
    const bool brick = (int)blockIdx.x < a.n_bb;
    // brick blocks: the first point's inputs are loaded before the window
    // set-up, so their latency overlaps it
    CP2gRaw<real> cur, nxt;
    if (brick && r.j0 + (int)threadIdx.x < end) c_p2g_raw<real, MASS>(a, r.j0 + (int)threadIdx.x, cur);
    int lo[3], hi[3];
    const bool any = c_range_box(a, r, lo, hi);
    const CWin w = c_window_box(any, lo, hi, sbox, 0, 1, kCWinP);
    for (int t = threadIdx.x; t < w.nodes * NF; t += 256) lds[t] = 0ull;
    if constexpr (MASS) {
        for (int t = threadIdx.x; t < w.nodes; t += 256) {
            const long long I = c_wnode(g, w, t);
            for (int q = 0; q < 3; ++q) s_gk[q * kCWinP + t] = a.gkg[q * a.nn + I];
            s_ns[t] = a.near_surf[I] ? 1u : 0u;
        }
    }
    __syncthreads();
    const double sc0 = c_scale(a, MASS ? 0 : 2), sc1 = c_scale(a, MASS ? 1 : 3);
    if (brick) {
        // deformable points of one brick, the next point's inputs in flight
        int j = r.j0 + (int)threadIdx.x;
#pragma clang loop unroll(disable)
        for (; j < end; j += 256) {
            if (j + 256 < end) c_p2g_raw<real, MASS>(a, j + 256, nxt);
            CP2gIn in;
            c_p2g_in_raw<real, MASS>(a, cur, in);
            c_p2g_point<real, MASS, SMALL>(a, w, lds, s_gk, s_ns, j, in, sc0, sc1);
            cur = nxt;
        }
    } else {
#pragma clang loop unroll(disable)
        for (int j = r.j0 + (int)threadIdx.x; j < end; j += 256) {
            CP2gIn in;
            c_p2g_in_dc<real, MASS>(a, j, in);
            c_p2g_point<real, MASS, SMALL>(a, w, lds, s_gk, s_ns, j, in, sc0, sc1);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        const long long I = c_wnode(g, w, t);
        for (int q = 0; q < NF; ++q) {
            const unsigned long long v = lds[q * w.nodes + t];
            if (v) atomicAdd((unsigned long long *)&a.acc[c_p2g_off<MASS>(q) * a.nn + I], v);
        }
        if constexpr (MASS) {
            const uint32_t cf = s_ns[t] & (CF_CPOS | CF_CNEG);
            if (cf) atomicOr(&a.flags[I], cf);
        }
    }
}

// platen tags of this step: highest platen index per node (…_MP_Contact.cpp:494-518)
template <typename real>
__global__ void __launch_bounds__(256) k_c_dcmarks(CArgs<real> a, int step)
{
    // per node the largest tag, reduced per block in LDS (mark_corners_max)
    __shared__ unsigned long long s_mk[4096];
    __shared__ int s_pbox[4][6];
    const int d = blockIdx.x * 256 + threadIdx.x;
    bool on = false;
    int i0[3] = {0, 0, 0};
    unsigned long long tag = 0ull;
    if (d < a.n_dc) {
        const double x[3] = {a.dc_x[3 * d], a.dc_x[3 * d + 1], a.dc_x[3 * d + 2]};
        on = c_cell<kCFast<real>>(a.g, x, i0);
        tag = ((unsigned long long)(a.steps[step].epoch + 1u) << 32) | (unsigned)d;
    }
    mark_corners_max<256>(on, i0, tag, a.g, a.cmarks, s_mk, 4096, s_pbox);
}

// ---- C7 + C8 (nodes): two-field update, contact correction, platen override
// with reaction (…_MP_Contact.cpp:392-518)
template <typename real>
__global__ void __launch_bounds__(256) k_c_grid(CArgs<real> a, int step)
{
    __shared__ double s_r[4][3];
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    double rm = 0.0, rr = 0.0, ra = 0.0;
    const uint32_t flI = I < a.nn ? a.flags[I] : 0u;
    if (flI & CF_ACTIVE) {
        const GridParams &g = a.g;
        const StepRec st = a.steps[step];
        const double dt = st.dt;
        const double um = c_unscale(a, 0), umg = c_unscale(a, 1), umv = c_unscale(a, 2), uf = c_unscale(a, 3);
        double mm[2], v[2][3], f[2][3], mg[2][3];
        for (int fl = 0; fl < 2; ++fl) {
            mm[fl] = (double)a.acc[off_m(fl) * a.nn + I] * um;
            // field velocity: momentum over mass where the mass passes the
            // gate (elsewhere the momentum sums are not read); the mass
            // gradient only where the contact correction reads it
            const bool gate = c_lenm3(mm[fl]) > g.m_min;
            const double rm = (kCFast<real> && gate) ? 1.0 / mm[fl] : 0.0;
            for (int q = 0; q < 3; ++q) {
                f[fl][q] = (double)a.acc[off_f(fl, q) * a.nn + I] * uf;
                v[fl][q] = 0.0;
                if (gate) {
                    const double pq = (double)a.acc[off_mv(fl, q) * a.nn + I] * umv;
                    v[fl][q] = kCFast<real> ? pq * rm : pq / mm[fl];
                }
                mg[fl][q] = (flI & CF_CNEG) ? (double)a.acc[off_mg(fl, q) * a.nn + I] * umg : 0.0;
            }
        }
        const uint8_t fx = a.fixed[I];
        for (int s = 0; s < 2; ++s) {
            double *vv = v[s], *ff = f[s];
            if (c_len3(vv) > (s == 0 ? 1.0e-9 : 1.0e-16)) {
                const double lf = c_len3(ff), inv = 1.0 / sqrt(c_dot3(vv, vv)), sc = g.damping * lf;
                for (int q = 0; q < 3; ++q) ff[q] -= sc * (vv[q] * inv);
            }
            if (c_lenm3(mm[s]) > g.m_min)
                for (int q = 0; q < 3; ++q) vv[q] += ff[q] / mm[s] * dt;
        }
        a.gm[I] = mm[0];
        a.gm[a.nn + I] = mm[1];
        // boundary conditions on both fields (no reaction here: the contact
        // step's reaction is the platen's, …_MP_Contact.cpp:510-511)
        for (int s = 0; s < 2; ++s)
            for (int q = 0; q < 3; ++q)
                if (fx & (1u << q)) { v[s][q] = 0.0; f[s][q] = 0.0; }
        if (flI & CF_CNEG) {
            const double m = mm[0], mn = mm[1];
            double vcm[3], nrm[3], fc[3] = {0, 0, 0}, fcn[3] = {0, 0, 0};
            for (int q = 0; q < 3; ++q) vcm[q] = (m * v[0][q] + mn * v[1][q]) / (m + mn);
            for (int q = 0; q < 3; ++q) nrm[q] = -0.5 * (mg[0][q] - mg[1][q]);
            const double inv = 1.0 / sqrt(c_dot3(nrm, nrm));
            for (int q = 0; q < 3; ++q) nrm[q] = nrm[q] * inv;
            const double nneg[3] = {-nrm[0], -nrm[1], -nrm[2]};
            if (c_dot3(v[0], nrm) + c_dot3(v[1], nneg) < 0.0) {
                double d1[3], d2[3];
                for (int q = 0; q < 3; ++q) { d1[q] = vcm[q] - v[0][q]; d2[q] = vcm[q] - v[1][q]; }
                const double s1 = m / dt * c_dot3(d1, nrm), s2 = mn / dt * c_dot3(d2, nneg);
                for (int q = 0; q < 3; ++q) { fc[q] = s1 * nrm[q]; fcn[q] = s2 * nneg[q]; }
                if (a.mu > 0.0) {
                    // Coulomb friction (no reference counterpart: the reference
                    // corrects the normal velocity only).  The tangential part
                    // of field 0's stick correction d1 is capped at mu times its
                    // normal part; field 1 takes the opposite impulse.
                    const double dn = c_dot3(d1, nrm);
                    double dtg[3];
                    for (int q = 0; q < 3; ++q) dtg[q] = d1[q] - dn * nrm[q];
                    const double lt = c_len3(dtg);
                    if (lt > 0.0) {
                        const double scl = fmin(1.0, a.mu * fabs(dn) / lt);
                        for (int q = 0; q < 3; ++q) {
                            const double ft = m / dt * scl * dtg[q];
                            fc[q] += ft;
                            fcn[q] -= ft;
                        }
                    }
                }
            }
            if (c_lenm3(m) > g.m_min)
                for (int q = 0; q < 3; ++q) { f[0][q] += fc[q]; v[0][q] += fc[q] / m * dt; }
            if (c_lenm3(mn) > g.m_min)
                for (int q = 0; q < 3; ++q) { f[1][q] += fcn[q]; v[1][q] += fcn[q] / mn * dt; }
        }
        const unsigned long long mk = a.cmarks[I];
        if ((uint32_t)(mk >> 32) == st.epoch + 1u) {
            const int d = (int)(mk & 0xFFFFFFFFull);
            rr += f[0][1] + f[1][1];
            for (int q = 0; q < 3; ++q) {
                v[0][q] = v[1][q] = a.dc_v[3 * d + q];
                f[0][q] = f[1][q] = 0.0;
            }
        }
        // G2P's node acceleration per field, f / m where the field mass
        // passes the gate (…_MP_Contact.cpp:560-566), with the gate as a flag
        uint32_t gates = 0u;
        for (int s = 0; s < 2; ++s) {
            const bool gt = c_lenm3(mm[s]) > g.m_min;
            if (gt) gates |= CF_GATE0 << s;
            const double rms = 1.0 / mm[s];
            for (int q = 0; q < 3; ++q) {
                a.gv[(3 * s + q) * a.nn + I] = v[s][q];
                a.gf[(3 * s + q) * a.nn + I] = f[s][q];
                a.ga[(3 * s + q) * a.nn + I] = gt ? (kCFast<real> ? f[s][q] * rms : f[s][q] / mm[s]) : 0.0;
            }
        }
        if (gates) a.flags[I] = flI | gates;
        rm += mm[0];
        ra += 1.0;
    }
    rm = wave_sum_dpp(rm);
    rr = wave_sum_dpp(rr);
    ra = wave_sum_dpp(ra);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_r[wv][0] = rm; s_r[wv][1] = rr; s_r[wv][2] = ra; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double r = 0.0;
        for (int w = 0; w < 4; ++w) r += s_r[w][threadIdx.x];
        a.cpart[3 * (size_t)blockIdx.x + threadIdx.x] = r;
    }
}

// ---- C8: split G2P, kinematics, constitutive update (…_MP_Contact.cpp:524-632)
template <typename real>
__global__ void __launch_bounds__(256) k_c_g2p(CArgs<real> a, int step)
{
    __shared__ double s_r[4][5];
    __shared__ double s_bn[4][4];   // per wave: maxima of the next step's bounds
    __shared__ int sbox[6];
    __shared__ double s_n[12][kCWin];   // per window node: v.xyz of field f at 3f, a.xyz at 6 + 3f
    __shared__ uint32_t s_fl[kCWin];
    const int chunk = c_chunk((int)gridDim.x, a.n_xcd);
    const int j = chunk * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const GridParams &g = a.g;
    double red[5] = {0, 0, 0, 0, 0};
    double bn[4] = {0.0, 0.0, 0.0, 0.0};   // bounds of the new state's contributions in the next step
    int flags = 0;
    double x[3] = {0.0, 0.0, 0.0};
    int i0[3] = {0, 0, 0};
    bool valid = false;
    uint32_t fm = 0u;
    double v[3] = {0.0, 0.0, 0.0};
    real F[9], s[6], mass = real(0);
    int umat = -1;
    if (j < a.n_def) {
        c_pos(a, j, x);
        valid = c_cell<kCFast<real>>(g, x, i0);
        if (!valid) flags |= FLAG_OUT_OF_GRID;
        // the point's own state is loaded before the window set-up: the
        // latency overlaps it
        fm = a.fmask[j];
        for (int q = 0; q < 3; ++q) v[q] = (double)a.p.v[q][j];
        for (int q = 0; q < 9; ++q) F[q] = a.p.F[q][j];
        for (int q = 0; q < 6; ++q) s[q] = a.p.s[q][j];
        mass = a.p.mass[j];
        if constexpr (sizeof(real) == 4) umat = a.p.mat[j];
    }
    // the block's node window (velocity, gated acceleration, flags) staged in LDS
    const CWin w = c_window_box(valid, i0, i0, sbox, 0, 1, kCWin);
    for (int t = threadIdx.x; t < w.nodes; t += 256) {
        const long long I = c_wnode(g, w, t);
        s_fl[t] = a.flags[I];
        for (int q = 0; q < 6; ++q) {
            s_n[q][t] = a.gv[q * a.nn + I];
            s_n[6 + q][t] = a.ga[q * a.nn + I];
        }
    }
    __syncthreads();
    if (valid) {
        constexpr bool FP32 = sizeof(real) == 4;
        const PArrays<real> &P = a.p;
        const double dt = a.steps[step].dt;
        {
            double xn[3] = {x[0], x[1], x[2]}, L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            CAx A;
            c_axes<kCFast<real>>(g, x, i0, A);
            const long long I0 = c_node0(g, i0);
            const int s0 = w.lds ? c_wslot(w, i0) : 0;
            for (int k = 0; k < 8; ++k) {
                const long long I = c_cnode(g, I0, k);
                double xi[3], N, G[3];
                c_corner<kCFast<real>>(g, A, k, xi, N, G);
                const int f = (int)((fm >> k) & 1u);
                uint32_t fl;
                double vI[3], aI[3];
                if (w.lds) {
                    const int ws = c_cslot(w, s0, k);
                    fl = s_fl[ws];
                    for (int q = 0; q < 3; ++q) {
                        vI[q] = s_n[3 * f + q][ws];
                        aI[q] = s_n[6 + 3 * f + q][ws];
                    }
                } else {
                    fl = a.flags[I];
                    for (int q = 0; q < 3; ++q) {
                        vI[q] = a.gv[(3 * f + q) * a.nn + I];
                        aI[q] = a.ga[(3 * f + q) * a.nn + I];
                    }
                }
                if (fl & (CF_GATE0 << f))
                    for (int q = 0; q < 3; ++q) v[q] += N * aI[q] * dt;
                for (int q = 0; q < 3; ++q) xn[q] += N * vI[q] * dt;
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r) L[3 * c + r] += vI[r] * G[c];
            }
            // F += L F dt, V = det(F) V0 (…_MP.cpp:299-302), then the material update
            real Lr[9];
            for (int q = 0; q < 9; ++q) Lr[q] = (real)L[q];
            real T[9];   // FP32 state stores D = F - I: D += (L + L D) dt
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    T[3 * c + r] = (FP32 ? Lr[3 * c + r] : real(0)) +
                                   (Lr[r] * F[3 * c] + Lr[3 + r] * F[3 * c + 1] + Lr[6 + r] * F[3 * c + 2]);
            for (int q = 0; q < 9; ++q) F[q] += T[q] * (real)dt;
            for (int q = 0; q < 9; ++q) P.F[q][j] = F[q];
            for (int q = 0; q < 3; ++q) P.v[q][j] = (real)v[q];
            if constexpr (FP32) {
                const uint32_t b = P.pbrick[j];
                const int bc[3] = {(int)(b % g.nb[0]), (int)((b / g.nb[0]) % g.nb[1]), (int)(b / (g.nb[0] * g.nb[1]))};
                for (int q = 0; q < 3; ++q) {
                    const double u = xn[q] * g.ih[q] - (double)(bc[q] * g.S[q]);
                    P.xq[q][j] = (int32_t)rint(u * kQOne);
                    if (u < -0.5 || u >= (double)g.S[q] + 0.5) flags |= FLAG_REBIN;
                }
            } else {
                for (int q = 0; q < 3; ++q) P.x[q][j] = (real)xn[q];
                const uint32_t b = P.pbrick[j];
                const int bc[3] = {(int)(b % g.nb[0]), (int)((b / g.nb[0]) % g.nb[1]), (int)(b / (g.nb[0] * g.nb[1]))};
                for (int q = 0; q < 3; ++q) {
                    const double u = xn[q] / g.h[q] - (double)(bc[q] * g.S[q]);
                    if (u < -0.5 || u >= (double)g.S[q] + 0.5) flags |= FLAG_REBIN;
                }
            }
            // FP32 state: the point's own material index, loaded above, as
            // the block-uniform one (the same table entry)
            material_update<real>(g, P, a.mats, (uint32_t)j, Lr, (real)dt, F, s, umat);
            // checks of the new state (PhysicsEngine.cpp:73-79)
            const double md = (double)mass;
            const double vd[3] = {(double)(real)v[0], (double)(real)v[1], (double)(real)v[2]};
            const double vv = vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2];
            red[0] = md * vv;
            for (int q = 0; q < 3; ++q) red[1 + q] = md * vd[q];
            const real ssum = s[0] + s[1] + s[2] + s[3] + s[4] + s[5];
            if (!(isfinite(vv) && isfinite((double)ssum) && isfinite(xn[0] + xn[1] + xn[2]))) {
                red[4] = 1.0;
                flags |= FLAG_NAN;
            }
            // the next step's P2G bounds of this point, from the stored new
            // state (the values the kernel pass would read)
            {
                double Fd[9], sd[6], fe[3] = {0.0, 0.0, 0.0};
                for (int q = 0; q < 9; ++q) Fd[q] = fd(F[q], q);
                for (int q = 0; q < 6; ++q) sd[q] = (double)s[q];
                if (g.has_fext)
                    for (int q = 0; q < 3; ++q) fe[q] = (double)P.fext[q][j];
                const double vol = det3(Fd) * (double)P.vol0[j];
                c_bounds_of(c_ihm(g), md, vd, vol, sd, fe, bn);
            }
        }
    }
    for (int k = 0; k < 5; ++k) {
        const double wsum = wave_sum_dpp(red[k]);
        if (lane == 0) s_r[wv][k] = wsum;
    }
    for (int q = 0; q < 4; ++q) {
        const double wm = wave_dmax(bn[q]);
        if (lane == 0) s_bn[wv][q] = wm;
    }
    int fl = flags;
    for (int o = 32; o > 0; o >>= 1) fl |= __shfl_xor(fl, o, 64);
    if (lane == 0 && fl) atomicOr(a.gflags, fl);
    __syncthreads();
    if (threadIdx.x < 5) {
        double r = 0.0;
        for (int q = 0; q < 4; ++q) r += s_r[q][threadIdx.x];
        a.ppart[5 * (size_t)chunk + threadIdx.x] = r;
    }
    if (threadIdx.x >= 64 && threadIdx.x < 68) {
        const int q = threadIdx.x - 64;
        const double wm = fmax(fmax(s_bn[0][q], s_bn[1][q]), fmax(s_bn[2][q], s_bn[3][q]));
        const unsigned long long wb = (unsigned long long)__double_as_longlong(wm);
        if (wm > 0.0 && wb > __hip_atomic_load(&a.bounds_next[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&a.bounds_next[q], wb);
    }
}

// the platen moves with its stored velocity, no TimeLine (…_MP_Contact.cpp:638-644)
template <typename real>
__global__ void __launch_bounds__(256) k_c_dcmove(CArgs<real> a, int step)
{
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= a.n_dc) return;
    const double dt = a.steps[step].dt;
    for (int q = 0; q < 3; ++q) a.dc_x_out[3 * d + q] = a.dc_x[3 * d + q] + a.dc_v[3 * d + q] * dt;
}

// point kernel and its gradient by original point index (for readback)
template <typename real>
__global__ void __launch_bounds__(256) k_c_kernel_out(CArgs<real> a, double *k_id, double *g_id, int n_total)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.n_all) return;
    const uint32_t id = c_id(a, j);
    k_id[id] = a.kp[j];
    for (int q = 0; q < 3; ++q) g_id[3 * (size_t)id + q] = a.kg[q * a.n_all + j];
    (void)n_total;
}

// per-step reaction of a batch of contact steps: block k sums its step's node-block partials
__global__ void __launch_bounds__(256) k_c_step_sums(const double *cpart, int nblk, double *hist)
{
    __shared__ double s[256];
    const int t = threadIdx.x;
    const double *cp = cpart + (size_t)blockIdx.x * nblk * 3;
    double r = 0.0;
    for (int i = t; i < nblk; i += 256) r += cp[3 * i + 1];
    s[t] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) hist[blockIdx.x] = s[0];
}

// Node readback after a contact step (mpm_grid_nodes): the step's first
// field (mass, velocity, force of …_MP_Contact.cpp:392-518), its active flag,
// and the kernel grid (density and central-difference gradient,
// …_MP_Contact.cpp:121-191).  One thread per node.
template <typename real>
__global__ void __launch_bounds__(256) k_c_nodes(CArgs<real> a, NodeOut o)
{
    const long long I = blockIdx.x * 256ll + threadIdx.x;
    if (I >= a.nn) return;
    // the step writes node fields of active nodes only; the others read as
    // the reference's reset node, zero (…_MP_Contact.cpp:41-67)
    const bool act = (a.flags[I] & CF_ACTIVE) != 0;
    o.active[I] = act ? 1 : 0;
    o.mass[I] = act ? a.gm[I] : 0.0;
    for (int q = 0; q < 3; ++q) {
        o.v[3 * I + q] = act ? a.gv[q * a.nn + I] : 0.0;
        o.f[3 * I + q] = act ? a.gf[q * a.nn + I] : 0.0;
        o.kg[3 * I + q] = a.gkgrad[q * a.nn + I];
    }
    o.k[I] = (double)a.acc[off_k() * a.nn + I] * c_unscale(a, 4);
}

}  // namespace mpm
