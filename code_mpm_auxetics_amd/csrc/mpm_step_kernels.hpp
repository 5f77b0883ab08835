// mpm_step_kernels.hpp -- the per-step kernels of the explicit MPM step.
//
//   k_particle<real, G2P>: one launch per time step, one workgroup per
//     occupied brick.  Phase A: the grid update of step n for the nodes of
//     the brick's touched box -- the P2G sums of the (at most 8) brick arenas
//     covering each node, in a canonical brick order, then momentum, damping,
//     boundary conditions, reaction and the platen override -- into an LDS
//     node window.  Phase B, for each point of the brick: G2P of step n
//     (gather 8 window nodes), F / volume / strain update, constitutive
//     update, state write-back, then the P2G contributions of step n+1
//     (m, m v, -V sigma.grad N + N f_ext to the 8 corner nodes).
//     Every contribution is converted to fixed point with a per-brick,
//     per-field exponent and added to the brick's LDS arena with a 64-bit
//     integer add.  Integer addition is associative, so the arena is
//     independent of wave scheduling: the step is bit-reproducible with no
//     float atomics anywhere.  With G2P=false the kernel only rebuilds the
//     arena (after upload or a rebin).
//   Blocks of one launch read step n's arenas, brick records, platen marks and
//   platen velocities and write step n+1's: all four are ping-pong buffers.
//
// Reference: PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:41-368.
#pragma once
#include "mpm_device.hpp"

// register budget of the fused particle kernel: waves per SIMD (3 -> 168
// VGPRs, 2 -> 256); FP64 state needs the larger budget to avoid spills
#ifndef MPM_KP_WAVES32
#define MPM_KP_WAVES32 3
#endif
#ifndef MPM_KP_WAVES64
#define MPM_KP_WAVES64 2
#endif
// threads per particle workgroup (a multiple of 64)
#ifndef MPM_KP_BLOCK
#define MPM_KP_BLOCK 256
#endif

namespace mpm {

// ------------------------------------------------------------- constitutive
// The FP64 laws (elastic_inc, j2_of, radial_return, perfectly_plastic_inc,
// vonmises_inc, neohookean_inc) are in mpm_laws.hpp.

// The same radial return for FP32 state: FP64 arithmetic, elastic constants
// precomputed (C = E/(1+nu)/(1-2nu), 1/E, 1/G: reciprocals instead of the
// reference's divisions).  Returns true on the plastic branch; s is updated in
// place (rounded once).
__device__ __forceinline__ bool radial_return_fp32(double nu, double C, double ie, double ig, double radius,
                                                   float s[6], const float de[6], double dep[6])
{
    // FP32 elastic trial first.  Where its J2 lies inside the yield surface by
    // more than the FP32 evaluation error (bounded by 2^-17 of the squared
    // trial norm: the component differences of J2 carry absolute errors of a
    // few ulp of the largest component), the FP64 test below would find the
    // point elastic too; the FP32 trial is then the update.  Otherwise the
    // trial, the yield test and the return map run in FP64.
    {
        const float Cf = (float)C, nf = (float)nu, a0 = 1.0f - nf, a1 = 0.5f - nf;
        float t[6];
        t[0] = s[0] + Cf * (a0 * de[0] + nf * de[1] + nf * de[2]);
        t[1] = s[1] + Cf * (a0 * de[1] + nf * de[0] + nf * de[2]);
        t[2] = s[2] + Cf * (a0 * de[2] + nf * de[0] + nf * de[1]);
        t[3] = s[3] + Cf * (a1 * de[3]);
        t[4] = s[4] + Cf * (a1 * de[4]);
        t[5] = s[5] + Cf * (a1 * de[5]);
        const float d01 = t[0] - t[1], d12 = t[1] - t[2], d20 = t[2] - t[0];
        const float j2 = (1.0f / 6.0f) * (d01 * d01 + d12 * d12 + d20 * d20) + t[3] * t[3] + t[4] * t[4] + t[5] * t[5];
        const float nrm = t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3] + t[4] * t[4] + t[5] * t[5];
        if (j2 + 0x1p-17f * nrm <= (float)(radius * radius) * (1.0f - 0x1p-20f)) {
#pragma unroll
            for (int i = 0; i < 6; ++i) s[i] = t[i];
            return false;
        }
    }
    double st[6];
    st[0] = (double)s[0] + C * ((1.0 - nu) * de[0] + nu * de[1] + nu * de[2]);
    st[1] = (double)s[1] + C * ((1.0 - nu) * de[1] + nu * de[0] + nu * de[2]);
    st[2] = (double)s[2] + C * ((1.0 - nu) * de[2] + nu * de[0] + nu * de[1]);
    st[3] = (double)s[3] + C * ((0.5 - nu) * de[3]);
    st[4] = (double)s[4] + C * ((0.5 - nu) * de[4]);
    st[5] = (double)s[5] + C * ((0.5 - nu) * de[5]);
    const double j2 = j2_of(st);
    if (j2 <= radius * radius) {
#pragma unroll
        for (int i = 0; i < 6; ++i) s[i] = (float)st[i];
        return false;
    }
    const double p = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
    const double sc = radius / sqrt(j2);
    double ds[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double snew = (i < 3) ? (st[i] - p) * sc + p : st[i] * sc;
        ds[i] = snew - (double)s[i];
        s[i] = (float)snew;
    }
    dep[0] = de[0] - (ds[0] - nu * (ds[1] + ds[2])) * ie;
    dep[1] = de[1] - (ds[1] - nu * (ds[0] + ds[2])) * ie;
    dep[2] = de[2] - (ds[2] - nu * (ds[0] + ds[1])) * ie;
    dep[3] = de[3] - ds[3] * ig;
    dep[4] = de[4] - ds[4] * ig;
    dep[5] = de[5] - ds[5] * ig;
    return true;
}

// ------------------------------------------------------------------ locate
// Per-axis linear-hat data of one point: shape values of the lower/upper node
// (s0, s1), their gradient factors (dN/dx of that axis' hat) and the cell.
template <typename real>
struct Axis {
    real s0, s1, sg0, sg1, h;   // shape values, sign(d) of lower/upper node, cell size
    real ih;                     // 1/h (FP32 state multiplies instead of dividing)
    int cell;                    // global cell index
    real u;                      // continuous cell coordinate relative to the brick origin
};

// FP64: absolute position, neighbour search and node positions exactly as
// GridPoint_Mediator.h:35 (trunc of x/h), GridPoint_Factory.h:40-42 (i*L/n) and
// the bases of GridPoint_Mediator.h:64-88.
__device__ __forceinline__ void axis_of(double x, int a, const GridParams &g, int O, Axis<double> &r)
{
    const double h = g.h[a];
    const int i0 = (int)(x / h);
    const double xi0 = ((double)i0 * g.L[a]) / (double)g.cells[a];
    const double xi1 = ((double)(i0 + 1) * g.L[a]) / (double)g.cells[a];
    const double d0 = x - xi0, d1 = x - xi1;
    double s0 = 1.0 - fabs(d0) / h, s1 = 1.0 - fabs(d1) / h;
    if (s0 < 0.0 || s0 > 1.0) s0 = 0.0;
    if (s1 < 0.0 || s1 > 1.0) s1 = 0.0;
    r.s0 = s0;
    r.s1 = s1;
    r.sg0 = sgn(d0);
    r.sg1 = sgn(d1);
    r.h = h;
    r.ih = 1.0 / h;
    r.cell = (x >= 0.0) ? i0 : -1;   // out of grid (or NaN) -> -1
    r.u = x / h - (double)O;
}

// FP32: Q6.26 fixed-point position in cells relative to the brick origin O.
// Lower node at distance frac*h (s = 1 - frac), upper at (frac-1)*h (s = frac).
__device__ __forceinline__ void axis_of(int32_t q, int a, const GridParams &g, int O, Axis<float> &r)
{
    const int ci = q >> kQFrac;   // floor
    const float frac = (float)(q & ((1 << kQFrac) - 1)) * (float)(1.0 / kQOne);
    r.s0 = 1.0f - frac;
    r.s1 = frac;
    r.sg0 = frac > 0.0f ? 1.0f : 0.0f;
    r.sg1 = -1.0f;
    r.h = g.hf[a];
    r.ih = g.ihf[a];
    r.cell = O + ci;
    r.u = (float)ci + frac;
}

// N and grad N of corner k (x-outer, z-inner as GridPoint_Mediator.h:39-45),
// evaluation order of GridPoint_Mediator.h:83-87.
// (Axes are passed by value: an indexed array of them would be put on the
// stack once the corner index is a run-time value.)
template <typename real>
__device__ __forceinline__ void corner(const Axis<real> bx, const Axis<real> by, const Axis<real> bz, int k,
                                       real &N, real G[3])
{
    const bool ax = (k >> 2) & 1, ay = (k >> 1) & 1, az = k & 1;
    const real sx = ax ? bx.s1 : bx.s0, sy = ay ? by.s1 : by.s0, sz = az ? bz.s1 : bz.s0;
    const real gx = ax ? bx.sg1 : bx.sg0, gy = ay ? by.sg1 : by.sg0, gz = az ? bz.sg1 : bz.sg0;
    N = sx * sy * sz;
    if constexpr (sizeof(real) == 4) {
        G[0] = -sy * sz * gx * bx.ih;
        G[1] = -sx * sz * gy * by.ih;
        G[2] = -sx * sy * gz * bz.ih;
    } else {   // FP64 state: the reference's division (GridPoint_Mediator.h:85-87)
        G[0] = -sy * sz * gx / bx.h;
        G[1] = -sx * sz * gy / by.h;
        G[2] = -sx * sy * gz / bz.h;
    }
}

template <typename real>
__device__ __forceinline__ void axes_of(const real x[3], const int32_t xq[3], const GridParams &g, const int O[3],
                                        Axis<real> &bx, Axis<real> &by, Axis<real> &bz)
{
    if constexpr (sizeof(real) == 4) {
        axis_of(xq[0], 0, g, O[0], bx);
        axis_of(xq[1], 1, g, O[1], by);
        axis_of(xq[2], 2, g, O[2], bz);
    } else {
        axis_of(x[0], 0, g, O[0], bx);
        axis_of(x[1], 1, g, O[1], by);
        axis_of(x[2], 2, g, O[2], bz);
    }
}

// FP32 linear-hat weights of one point arranged for the 8-corner loops:
// N_k = wxy[ax,ay] wz[az], dN_k/dx = gx[ax] wyz[ay,az], dN_k/dy = gy[ay]
// wxz[ax,az], dN_k/dz = gz[az] wxy[ax,ay]; g*[0] = -sign(d0)/h, g*[1] = 1/h
// (GridPoint_Mediator.h:83-87 with d1 = (frac - 1) h < 0).
struct Hat32 {
    float wxy[4], wyz[4], wxz[4], wz[2], gx[2], gy[2], gz[2];
};
__device__ __forceinline__ Hat32 hat32(const Axis<float> &bx, const Axis<float> &by, const Axis<float> &bz)
{
    Hat32 w;
    const float sx[2] = {bx.s0, bx.s1}, sy[2] = {by.s0, by.s1}, sz[2] = {bz.s0, bz.s1};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            w.wxy[2 * i + j] = sx[i] * sy[j];
            w.wyz[2 * i + j] = sy[i] * sz[j];
            w.wxz[2 * i + j] = sx[i] * sz[j];
        }
    w.wz[0] = sz[0];
    w.wz[1] = sz[1];
    w.gx[0] = -bx.sg0 * bx.ih;
    w.gx[1] = bx.ih;
    w.gy[0] = -by.sg0 * by.ih;
    w.gy[1] = by.ih;
    w.gz[0] = -bz.sg0 * bz.ih;
    w.gz[1] = bz.ih;
    return w;
}

// 8-real node records, 32 B (FP32) or 64 B (FP64) aligned
__device__ __forceinline__ void store8(float *p, const float r[8])
{
    reinterpret_cast<float4 *>(p)[0] = make_float4(r[0], r[1], r[2], r[3]);
    reinterpret_cast<float4 *>(p)[1] = make_float4(r[4], r[5], r[6], r[7]);
}
__device__ __forceinline__ void store8(double *p, const double r[8])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) reinterpret_cast<double2 *>(p)[k] = make_double2(r[2 * k], r[2 * k + 1]);
}
__device__ __forceinline__ void load8(const float *p, float r[8])
{
    const float4 a = reinterpret_cast<const float4 *>(p)[0], b = reinterpret_cast<const float4 *>(p)[1];
    r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
    r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
}
__device__ __forceinline__ void load8(const double *p, double r[8])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double2 v = reinterpret_cast<const double2 *>(p)[k];
        r[2 * k] = v.x;
        r[2 * k + 1] = v.y;
    }
}

// ------------------------------------------------------- block reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_umax(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_imin(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_imax(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
// Max of an unsigned value over the wave without the LDS crossbar: inclusive
// DPP row scans (lanes without a source read 0, the identity), then the four
// row maxima.  The result is wave-uniform.
__device__ __forceinline__ uint32_t wave_umax_dpp(uint32_t v)
{
    v = max(v, (uint32_t)dpp_i<MPM_ROW_SHR(1)>((int)v));
    v = max(v, (uint32_t)dpp_i<MPM_ROW_SHR(2)>((int)v));
    v = max(v, (uint32_t)dpp_i<MPM_ROW_SHR(4)>((int)v));
    v = max(v, (uint32_t)dpp_i<MPM_ROW_SHR(8)>((int)v));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 47);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
    return max(max(r0, r1), max(r2, r3));
}
// Box bounds over the wave through wave_umax_dpp: lo in [0, 2^20], hi >= -1.
__device__ __forceinline__ int wave_imin_dpp(int v) { return (1 << 20) - (int)wave_umax_dpp((uint32_t)((1 << 20) - v)); }
__device__ __forceinline__ int wave_imax_dpp(int v) { return (int)wave_umax_dpp((uint32_t)(v + 1)) - 1; }

// Sum of a double over the wave: inclusive DPP row scans of both 32-bit
// halves (lanes without a source read +0.0), then the four row totals added
// as (r0 + r1) + (r2 + r3).  A fixed order, so the result is reproducible;
// wave-uniform.  Call with every lane of the wave active.
__device__ __forceinline__ double lane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_d<MPM_ROW_SHR(1)>(v);
    v += dpp_d<MPM_ROW_SHR(2)>(v);
    v += dpp_d<MPM_ROW_SHR(4)>(v);
    v += dpp_d<MPM_ROW_SHR(8)>(v);
    return (lane_d(v, 15) + lane_d(v, 31)) + (lane_d(v, 47) + lane_d(v, 63));
}

// One point's state as the step reads it: position (FP64 absolute or FP32
// Q6.26 fixed point), velocity, F (F - I in FP32 state), stress, mass,
// initial volume and, when the caller asks for it, alpha.
template <typename real>
struct PointState {
    real x[3];
    int32_t xq[3];
    real v[3], F[9], s[6], m, V0, al;
};
template <typename real>
__device__ __forceinline__ PointState<real> load_point(const PArrays<real> &P, uint32_t i, bool with_alpha)
{
    PointState<real> t;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if constexpr (sizeof(real) == 4) {
            t.xq[q] = elt(P.xq[q], i);
            t.x[q] = real(0);
        } else {
            t.x[q] = elt(P.x[q], i);
            t.xq[q] = 0;
        }
        t.v[q] = elt(P.v[q], i);
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) t.F[q] = elt(P.F[q], i);
#pragma unroll
    for (int q = 0; q < 6; ++q) t.s[q] = elt(P.s[q], i);
    t.m = elt(P.mass, i);
    t.V0 = elt(P.vol0, i);
    t.al = with_alpha ? elt(P.alpha, i) : real(0);
    return t;
}

// Node velocity, damped force and f / m from a node's P2G sums {m, p.xyz,
// f.xyz} (…_MP.cpp:209-222), shared by the step's grid update and the grid
// readback.  FP32 state: one hardware reciprocal of the mass (1 ulp) and a
// reciprocal square root instead of IEEE divisions.  A zero mass gives
// non-finite f / m, which the callers gate by the mass.
template <typename real>
__device__ __forceinline__ void node_damp(const real sum[kQ], real m_min, real damp, real vn[3], real fn[3],
                                          real ac[3])
{
    const real mn = sum[0];
#pragma unroll
    for (int q = 0; q < 3; ++q) fn[q] = sum[4 + q];
    if constexpr (sizeof(real) == 4) {
        const float imn = __builtin_amdgcn_rcpf(mn);
#pragma unroll
        for (int q = 0; q < 3; ++q) vn[q] = (mn > m_min) ? sum[1 + q] * imn : 0.0f;
        const float vv = vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2];
        if (vv > 1.0e-18f) {   // |v| > 1e-9
            const float lf = __builtin_amdgcn_sqrtf(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
            const float sc = damp * lf * __builtin_amdgcn_rsqf(vv);
#pragma unroll
            for (int q = 0; q < 3; ++q) fn[q] -= sc * vn[q];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) ac[q] = fn[q] * imn;
    } else {
#pragma unroll
        for (int q = 0; q < 3; ++q) vn[q] = (mn > m_min) ? sum[1 + q] / mn : real(0);
        const real vv = vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2];
        if (sqrt(vv) > real(1.0e-9)) {
            const real lf = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
            const real inv = real(1) / sqrt(vv);
            const real sc = damp * lf;
#pragma unroll
            for (int q = 0; q < 3; ++q) fn[q] -= sc * (vn[q] * inv);
        }
        // f / m after damping: the velocity update and the stored
        // acceleration share it; a force zeroed later gives 0 / m = 0
#pragma unroll
        for (int q = 0; q < 3; ++q) ac[q] = fn[q] / mn;
    }
}

// Override marks of platen points (…_MP.cpp:245-265): each node touched by
// a point (corners i0 .. i0 + 1; those outside the grid are skipped) ends
// with the largest tag of the points touching it.  Many platen points share
// a node, and same-address global atomics serialize at the L2, so the block
// first takes the maximum per node in LDS (s_mk, cap entries) over the
// bounding box of its points' base nodes (platen points are stored in a
// spatially compact order) and then issues one global atomicMax per marked
// node; a box larger than cap falls back to one global atomic per point and
// corner.  Called by every thread of an NT-thread block (on: the thread has
// a point).  s_pbox: LDS scratch of the block's wave bounds (NT / 64 rows).
template <int NT>
__device__ __forceinline__ void mark_corners_max(bool on, const int i0[3], unsigned long long tag, const GridParams &g,
                                                 unsigned long long *marks, unsigned long long *s_mk, int cap,
                                                 int (*s_pbox)[6])
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int lo[3], hi[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int c = min(max(i0[q], -1), g.nodes[q]);
        lo[q] = wave_imin(on ? c : (1 << 20));
        hi[q] = wave_imax(on ? c + 1 : -(1 << 20));
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 3; ++q) { s_pbox[wv][2 * q] = lo[q]; s_pbox[wv][2 * q + 1] = hi[q]; }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q)
        for (int w = 0; w < NT / 64; ++w) {
            lo[q] = min(lo[q], s_pbox[w][2 * q]);
            hi[q] = max(hi[q], s_pbox[w][2 * q + 1]);
        }
    if (hi[0] < lo[0]) return;   // no point in this block (uniform)
    const int e0 = hi[0] - lo[0] + 1, e1 = hi[1] - lo[1] + 1, e2 = hi[2] - lo[2] + 1;
    const long long nbox = (long long)e0 * e1 * e2;
    const bool in_lds = nbox <= (long long)cap;   // uniform
    if (in_lds) {
        for (int r = tid; r < (int)nbox; r += NT) s_mk[r] = 0ull;
        __syncthreads();
    }
    if (on)
        for (int k = 0; k < 8; ++k) {
            const int n0 = i0[0] + ((k >> 2) & 1), n1 = i0[1] + ((k >> 1) & 1), n2 = i0[2] + (k & 1);
            if (n0 < 0 || n1 < 0 || n2 < 0 || n0 >= g.nodes[0] || n1 >= g.nodes[1] || n2 >= g.nodes[2]) continue;
            if (in_lds)
                atomicMax(&s_mk[(n0 - lo[0]) + e0 * ((n1 - lo[1]) + e1 * (n2 - lo[2]))], tag);
            else
                atomicMax(&marks[n0 + (long long)g.nodes[0] * (n1 + (long long)g.nodes[1] * n2)], tag);
        }
    if (in_lds) {
        __syncthreads();
        for (int r = tid; r < (int)nbox; r += NT) {
            const unsigned long long t = s_mk[r];
            if (!t) continue;
            const int n2 = lo[2] + r / (e0 * e1), rem = r % (e0 * e1);
            const int n1 = lo[1] + rem / e0, n0 = lo[0] + rem % e0;
            atomicMax(&marks[n0 + (long long)g.nodes[0] * (n1 + (long long)g.nodes[1] * n2)], t);
        }
    }
}

constexpr int kUnsetExp = -100000;
constexpr int kMinExp = -90;   // contributions below 2^-90 of a unit are dropped

template <typename real>
struct KpArgs {
    GridParams g;
    PArrays<real> p;
    // ping-pong pairs, indexed by parity: a step reads [par] (step n) and
    // writes [par ^ 1] (step n+1), par from its step record; the arena
    // rebuild (G2P = false) writes [rebuild_par]
    real *arena[2];             // P2G sums per slot (box-compact 8-real records)
    int8_t *exps[2];            // FP32 state: per slot, 8 bytes: each field's exponent of its largest
                                // contribution bound in the step that wrote the arena (kExpNone: all zero)
    int4 *binfo[2];             // per brick {slot or -1, packed touched box}
    unsigned long long *marks[2];         // platen override tags
    double *dc_v[2];                      // platen velocities
    int32_t rebuild_par;
    uint32_t zero_off;          // element offset in an arena of a record of 8 zeros (after the last slot)
    const uint8_t *fixed;       // node fixity bits
    const uint32_t *occ_list, *occ_start, *occ_count;
    const uint32_t *slot_order;       // brick block b works on slot slot_order[b]
    const int4 *idesc;                // per block b: its slot's descriptor (k_item_desc), 2 int4
    const uint8_t *slot_mat;          // per slot: the material of all its points, or 0xFF if mixed
    const int32_t *n_occ;
    int32_t n_pblk;             // platen blocks at the front of the grid
    const Material *mats;
    const StepRec *steps;
    int32_t *flags;
    double *kp_part;      // per slot and wave: 5 doubles (sum m|v|^2, m v.xyz, nonfinite)
    double *gpart;        // per slot: 3 doubles of this step's grid update (mass, reaction, active nodes)
    int32_t n_dc;
    const int32_t *fy_slot;             // batched worlds: node -> y-fixed slot or -1 (null: off)
    double *fy_val;                     // this step's reaction per y-fixed slot
    double *dc_x;                       // 3 per platen point, FP64
    // persistent launch: XCD count, per-step work counters (kQueueStride
    // per step, zeroed before the batch)
    int32_t n_xcd;
    uint32_t *queue;
};
constexpr int kQueueStride = 64;   // work counters per step (>= XCDs)

// Strain increment, total strain, constitutive update and stress write of
// point i (…_MP.cpp:304-353), shared by the plain and the contact step.
// F is the stored deformation gradient (F - I in FP32 state, see fd()).
template <typename real>
__device__ __forceinline__ void material_update(const GridParams &g, const PArrays<real> &P, const Material *mats,
                                                uint32_t i, const real L[9], real dt, const real F[9], real s[6],
                                                int umat = -1, bool have_alpha = false, real alpha_pre = real(0))
{
    constexpr bool FP32 = sizeof(real) == 4;
    // strain increment (…_MP.cpp:304-312); FP32 state uses L dt
    real de[6];
    if (!FP32 && g.strain_form == 0) {
        de[0] = (real(1) + L[0] * dt) - real(1);
        de[1] = (real(1) + L[4] * dt) - real(1);
        de[2] = (real(1) + L[8] * dt) - real(1);
    } else {
        de[0] = L[0] * dt;
        de[1] = L[4] * dt;
        de[2] = L[8] * dt;
    }
    de[3] = L[1] * dt + L[3] * dt;
    de[4] = L[5] * dt + L[7] * dt;
    de[5] = L[6] * dt + L[2] * dt;
    if (g.track_strain)
#pragma unroll
        for (int q = 0; q < 6; ++q) elt(P.eps[q], i) += de[q];
    // constitutive dispatch (…_MP.cpp:322-353)
    // FP32 state: umat >= 0 means every point of the block has that material,
    // read with one block-uniform load instead of the point's index and then
    // its entry.  FP64 state reads the point's entry.
    Material mtv{};
    if constexpr (FP32) {
        if (umat >= 0)
            mtv = mats[umat];
        else
            mtv = mats[elt(P.mat, i)];
    }
    const Material &mt = FP32 ? mtv : mats[elt(P.mat, i)];
    int type = mt.type;
    if (type == 1 && !g.neohookean) type = 3;
    if (type == 2 || type == 5) type = 3;
    const double E = g.has_pointE ? (double)elt(P.E, i) : mt.E;
    // FP32 state: elastic constants of this point without divisions
    // on the common path (per-point E only when materials overflow
    // the table)
    double C = mt.C, ie = mt.invE, ig = mt.invG;
    if (FP32 && g.has_pointE) {
        C = E * mt.Cn;
        ie = 1.0 / E;
        ig = (2.0 * (1.0 + mt.nu)) * ie;
    }
    if (type == 0) {
        if constexpr (FP32) {
            const float Cf = (float)C, nu = (float)mt.nu, a0 = 1.0f - nu, a1 = 0.5f - nu;
            s[0] += Cf * (a0 * de[0] + nu * de[1] + nu * de[2]);
            s[1] += Cf * (a0 * de[1] + nu * de[0] + nu * de[2]);
            s[2] += Cf * (a0 * de[2] + nu * de[0] + nu * de[1]);
            s[3] += Cf * (a1 * de[3]);
            s[4] += Cf * (a1 * de[4]);
            s[5] += Cf * (a1 * de[5]);
        } else {
            real ds[6];
            elastic_inc<real>((real)E, (real)mt.nu, de, ds);
#pragma unroll
            for (int q = 0; q < 6; ++q) s[q] += ds[q];
        }
    } else if (type == 1) {
        double Fd[9], sd[6], ds[6];
        for (int q = 0; q < 9; ++q) Fd[q] = fd(F[q], q);
        for (int q = 0; q < 6; ++q) sd[q] = (double)s[q];
        neohookean_inc(E, mt.nu, sd, Fd, ds);
        for (int q = 0; q < 6; ++q) s[q] = (real)(sd[q] + ds[q]);
    } else if constexpr (FP32) {
        // FP64 island for FP32 state
        double dep[6];
        // have_alpha: the caller loaded alpha with the point's other state
        const double al = (type == 4) ? (double)(have_alpha ? alpha_pre : elt(P.alpha, i)) : 0.0;
        const double radius = mt.k0 + al * 0.70710678118654752440;   // al / sqrt(2)
        if (radial_return_fp32(mt.nu, C, ie, ig, radius, s, de, dep)) {
            if (type == 4) {
                double eff = 0.0;
#pragma unroll
                for (int q = 0; q < 6; ++q) eff += dep[q] * dep[q];
                eff = sqrt(eff * (2.0 / 3.0));
                const double da = ((mt.hs - al) > 0.0) ? mt.C0 * (mt.hs - al) * eff : 0.0;
                elt(P.alpha, i) = (real)(al + da);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) elt(P.ep[q], i) += (real)dep[q];
        }
    } else {
        // FP64 state: reference evaluation order
        double sd[6], ded[6], ds[6], dep[6], da = 0.0;
        for (int q = 0; q < 6; ++q) { sd[q] = (double)s[q]; ded[q] = (double)de[q]; }
        const double k0 = mt.sy / sqrt(3.0);
        if (type == 4) {
            const double al = (double)elt(P.alpha, i);
            if (radial_return(E, mt.nu, k0 + al / sqrt(2.0), sd, ded, ds, dep)) {
                double eff = 0.0;
                for (int q = 0; q < 6; ++q) eff += dep[q] * dep[q];
                eff *= 2.0 / 3.0;
                eff = sqrt(eff);
                if ((sqrt(2.0 / 3.0) * mt.C1 - al) > 0.0)
                    da = mt.C0 * (sqrt(2.0 / 3.0) * mt.C1 - al) * eff;
                elt(P.alpha, i) = (real)(al + da);
                for (int q = 0; q < 6; ++q) elt(P.ep[q], i) += (real)dep[q];
            }
        } else if (radial_return(E, mt.nu, k0, sd, ded, ds, dep)) {
            for (int q = 0; q < 6; ++q) elt(P.ep[q], i) += (real)dep[q];
        }
        for (int q = 0; q < 6; ++q) s[q] = (real)(sd[q] + ds[q]);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) elt(P.s[q], i) = s[q];
}

// Per occupied brick: the material index shared by all its points, or 0xFF
// if they differ (material indices are below 255).
__global__ void __launch_bounds__(256) k_slot_mat(const uint8_t *mat, const uint32_t *occ_start,
                                                  const uint32_t *occ_count, uint8_t *slot_mat)
{
    __shared__ int smin, smax;
    const int s = blockIdx.x;
    if (threadIdx.x == 0) {
        smin = 255;
        smax = 0;
    }
    __syncthreads();
    const uint32_t st = occ_start[s], cnt = occ_count[s];
    int mn = 255, mx = 0;
    for (uint32_t t = threadIdx.x; t < cnt; t += 256) {
        const int v = mat[st + t];
        mn = min(mn, v);
        mx = max(mx, v);
    }
    atomicMin(&smin, mn);
    atomicMax(&smax, mx);
    __syncthreads();
    if (threadIdx.x == 0) slot_mat[s] = (smin == smax) ? (uint8_t)smin : (uint8_t)0xFF;
}

// Per launch-order item of the step kernel (block b works on slot
// slot_order[b]): {slot, brick, first point, point count} and {the brick's one
// material (FP32 state) or -1, alpha loaded with the point state (FP32: a
// VonMises point may be present), brick x | brick y << 16, brick z}.
__global__ void __launch_bounds__(256) k_item_desc(GridParams g, const uint32_t *slot_order, const uint32_t *occ_list,
                                                   const uint32_t *occ_start, const uint32_t *occ_count,
                                                   const uint8_t *slot_mat, const Material *mats, int fp32, int n_occ,
                                                   int4 *idesc)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_occ) return;
    const uint32_t slot = slot_order[i], brick = occ_list[slot];
    const int umat = (fp32 && slot_mat[slot] != 0xFF) ? (int)slot_mat[slot] : -1;
    const int pre_al = (fp32 && (umat < 0 || mats[umat].type == 4)) ? 1 : 0;
    const int bx = (int)(brick % (uint32_t)g.nb[0]), by = (int)((brick / (uint32_t)g.nb[0]) % (uint32_t)g.nb[1]);
    const int bz = (int)(brick / ((uint32_t)g.nb[0] * (uint32_t)g.nb[1]));
    idesc[2 * i] = make_int4((int)slot, (int)brick, (int)occ_start[slot], (int)occ_count[slot]);
    idesc[2 * i + 1] = make_int4(umat, pre_al, bx | (by << 16), bz);
}

// ---------------------------------------------------- batched constitutive
// One thread per input of an mpm_constitutive batch (body: cr_apply).
__global__ void __launch_bounds__(256) k_constitutive(int law, CrBatch b)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= b.n) return;
    cr_apply(law, b, i);
}

// Fixed-point fraction bits.  FP32 state: each contribution is converted to a
// 32-bit integer scaled so that the brick's largest contribution bound so far
// is < 2^26, then added to the 64-bit arena.  FP64 state: 64-bit conversion
// with FRAC = 61 - H, count < 2^(H-1), so every node sum stays below 2^60.
constexpr int kFrac32 = 26;
constexpr int kFracApri = 27;   // FP32 state, a priori exponents (4 bits of headroom in 32 bits)
constexpr int kExpNone = -128;  // stored exponent of a field whose contributions were all zero
constexpr int kPBlock = MPM_KP_BLOCK;

__device__ __forceinline__ uint32_t fbits(float x) { return __float_as_uint(x); }


// One brick per workgroup.  Dynamic LDS: long long arena[7 * arenaN].
template <typename real, bool G2P, bool PERSIST = false>
__global__ void __launch_bounds__(kPBlock, sizeof(real) == 4 ? MPM_KP_WAVES32 : MPM_KP_WAVES64) k_particle(KpArgs<real> a, int step)
{
    extern __shared__ long long s_arena[];   // node-major: kQ int64 per arena node, then
                                             // (G2P) the node window: 6 reals {v, a} per node
    __shared__ uint32_t s_max[3][8];         // running max bound bits per field, one buffer per
                                             // chunk index mod 3 (see the chunk loop)
    __shared__ int4 s_nb[27][2];             // neighbour arenas (dz, dy, dx ascending): {element
                                             // base, window low corner in this brick's node
                                             // coordinates}, {window extent, ext x * ext y}
    __shared__ int s_box[kPBlock / 64][6];
    __shared__ uint32_t s_rmax[kPBlock / 64][kQ];   // a priori: wave maxima of the bounds
    __shared__ double s_red[kPBlock / 64][3];

    constexpr bool FP32 = sizeof(real) == 4;
    const GridParams &g = a.g;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, rl = lane & 15;
    const StepRec st = a.steps[step];
    // the point check sums and the grid mass and node count of this step are
    // read (the last step of a batch, and an arena rebuild)
    const bool sums = !G2P || st.skip_sums == 0;
    const double dtd = st.dt;
    const real dt = (real)dtd;
    // this step's ping-pong parity (a rebuild writes the current parity)
    const int par = G2P ? (st.par & 1) : (a.rebuild_par & 1);
    const real *const arena_in = a.arena[par];
    const int4 *const binfo_in = a.binfo[par];
    real *const arena_out = a.arena[G2P ? par ^ 1 : par];
    int4 *const binfo_out = a.binfo[G2P ? par ^ 1 : par];
    const unsigned long long *const marks_in = a.marks[par];
    unsigned long long *const marks_out = a.marks[par ^ 1];
    const double *const dc_v_in = a.dc_v[par];
    double *const dc_v_out = a.dc_v[par ^ 1];

    // ---- platen (displacement-controlled) blocks: phase 8 + override marks.
    // Platen blocks come first in the grid (n_pblk of them, a multiple of the
    // XCD count); brick block bb = blockIdx - n_pblk keeps its XCD.
    const int bb = (int)blockIdx.x - a.n_pblk;
    if (bb < 0) {
        if (!G2P) return;
        const int j = (int)blockIdx.x * kPBlock + tid;
        const bool on = j < a.n_dc;
        int i0[3] = {0, 0, 0};
        unsigned long long tag = 0;
        if (on) {
            double *x = &a.dc_x[3 * j], v[3];
            // …_MP.cpp:360-368: velocity from the TimeLine at the step's start time
            for (int q = 0; q < 3; ++q) {
                v[q] = st.has_tl ? st.vtl[q] : dc_v_in[3 * j + q];
                dc_v_out[3 * j + q] = v[q];
                x[q] += v[q] * dtd;
            }
            // override marks for step n+1 (…_MP.cpp:245-265): the highest
            // platen index touching a node wins, as in the reference's
            // sequential loop.  The tag carries epoch (n+1)+1 so that 0 means
            // "no mark".
            for (int q = 0; q < 3; ++q) i0[q] = (int)(x[q] / g.h[q]);
            tag = ((unsigned long long)(st.epoch + 2u) << 32) | (unsigned)j;
        }
        __shared__ int s_pbox[kPBlock / 64][6];
        mark_corners_max<kPBlock>(on, i0, tag, g, marks_out, reinterpret_cast<unsigned long long *>(s_arena), kQ * g.arenaN,
                         s_pbox);
        return;
    }

    // Brick work.  One launch per step: block b works on slot slot_order[b]
    // (brick blocks are launched longest first, most point chunks, which
    // shortens the tail of the launch).  Persistent launch (PERSIST): a
    // fixed grid of resident blocks; block b pulls the next item of work
    // queue b mod X (X = XCDs: group x holds slot_order[X i + x], the bricks
    // the one-launch order puts on XCD x, longest first) from a per-step
    // counter, so each group's bricks still run on one XCD.  A slot's work is
    // the same whichever block takes it: results do not depend on the order.
    __shared__ int s_item;
    for (int it = 0;; ++it) {
        int item;
        if constexpr (!PERSIST) {
            if (bb >= *a.n_occ) return;
            item = bb;
        } else {
            const int X = a.n_xcd, x = bb % X;
            if (tid == 0) s_item = (int)atomicAdd(&a.queue[(size_t)step * kQueueStride + x], 1u);
            __syncthreads();
            const int i = s_item;   // rewritten only after the barrier that ends this item
            const int no = *a.n_occ, nx = no / X + ((x < no % X) ? 1 : 0);
            if (i >= nx) return;
            item = X * i + x;
        }
        // the item's descriptor (k_item_desc): one uniform 32-byte load instead
        // of a chain of dependent lookups (slot, brick, start, count, material)
        const int4 d0 = a.idesc[2 * item], d1 = a.idesc[2 * item + 1];
        const int slot = d0.x;
        const uint32_t brick = (uint32_t)d0.y;
        const int bc[3] = {d1.z & 0xFFFF, d1.z >> 16, d1.w};
        const int O[3] = {bc[0] * g.S[0], bc[1] * g.S[1], bc[2] * g.S[2]};
        const uint32_t start = (uint32_t)d0.z, count = (uint32_t)d0.w;
        const int umat = d1.x;           // FP32: the brick's one material, or -1
        const bool pre_al = d1.y != 0;   // FP32: load alpha with the point state (VonMises points may be present)
        const int A0 = g.A[0], A01 = g.A[0] * g.A[1], AN = g.arenaN;

        // The first chunk's point state, loaded after the gather pass has issued
        // (and waited for) its arena loads, so that its HBM latency overlaps the
        // update pass, the arena clear and the barrier before the chunks.
        PointState<real> pt0;

        // ---- phase A: grid update of step n over the node window = this brick's
        // touched box of the last P2G (…_MP.cpp:209-265)
        real *s_node = reinterpret_cast<real *>(s_arena + kQ * AN);
        int wlo[3] = {0, 0, 0}, wext[3] = {0, 0, 0};
        int stash = 0;   // reals of the arena the grid update uses as its stash
        if (G2P) {
            if (tid < 27) {
                const int d[3] = {tid % 3 - 1, (tid / 3) % 3 - 1, tid / 9 - 1};
                const int x = bc[0] + d[0], y = bc[1] + d[1], z = bc[2] + d[2];
                int4 bi = make_int4(-1, 0, 0, 0);
                if (x >= 0 && y >= 0 && z >= 0 && x < g.nb[0] && y < g.nb[1] && z < g.nb[2])
                    bi = binfo_in[x + (size_t)g.nb[0] * (y + (size_t)g.nb[1] * z)];
                // a missing neighbour gets an empty window (extent 0)
                const int pk[3] = {bi.y, bi.z, bi.w};
                int lo[3], ext[3];
    #pragma unroll
                for (int q = 0; q < 3; ++q) {
                    lo[q] = (pk[q] & 0xFFFF) + d[q] * g.S[q];
                    ext[q] = (bi.x >= 0) ? (pk[q] >> 16) - (pk[q] & 0xFFFF) + 1 : 0;
                }
                s_nb[tid][0] = make_int4((int)((uint32_t)max(bi.x, 0) * 8u * (uint32_t)AN), lo[0], lo[1], lo[2]);
                s_nb[tid][1] = make_int4(ext[0], ext[1], ext[2], ext[0] * ext[1]);
            }
            __syncthreads();
            const int4 own0 = s_nb[13][0], own1 = s_nb[13][1];
            wlo[0] = own0.y; wlo[1] = own0.z; wlo[2] = own0.w;
            wext[0] = own1.x; wext[1] = own1.y; wext[2] = own1.z;
            const int nwin = wext[0] * wext[1] * wext[2];
            // window index -> node: quotients through a float reciprocal, exact
            // for r < 2^20 (the fraction of (r + 1/2) / e stays 1/(2e) away from
            // an integer, far above the float error)
            const float rx = 1.0f / (float)max(wext[0], 1), rxy = 1.0f / (float)max(own1.w, 1);
            const real dtg = (real)st.dt, m_min = (real)g.m_min, damp = (real)g.damping;
            // Pass 1: for each window node, the sum of the covering arenas'
            // records (at most 8 candidates, ascending brick order, uncovered ones
            // read the zero record), its fixity and override tag, stashed in LDS:
            // sums in the (not yet used) arena, {tag, fixity | counted} in the
            // node's window slot.  Branch-free, so that all loads of a node are in
            // flight at once.  Each thread later reads back only the slots it
            // wrote: no barrier.
            real *s_sum = reinterpret_cast<real *>(s_arena);
            stash = kQ * nwin;
            for (int r = tid; r < nwin; r += kPBlock) {
                const int lz = (int)(((float)r + 0.5f) * rxy), rem = r - lz * own1.w;
                const int ly = (int)(((float)rem + 0.5f) * rx), lx = rem - ly * wext[0];
                const int l[3] = {wlo[0] + lx, wlo[1] + ly, wlo[2] + lz};
                const long long nid = (O[0] - 1 + l[0]) + (long long)g.nodes[0] * ((O[1] - 1 + l[1]) + (long long)g.nodes[1] * (O[2] - 1 + l[2]));
                // per axis the arenas covering this node: own and the neighbour
                // below (l <= 2) or above (l >= S); the candidates are the 2 x 2 x 2
                // neighbours from `lower` up, and a candidate counts when the node
                // lies in its window (brick sizes >= 3 keep the two sides apart)
                int lower[3];
    #pragma unroll
                for (int q = 0; q < 3; ++q) lower[q] = (l[q] <= 2 && bc[q] > 0) ? -1 : 0;
                const int nb0 = (lower[2] + 1) * 9 + (lower[1] + 1) * 3 + (lower[0] + 1);
                uint32_t eo[8];
                int first = -1;
    #pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int4 w0 = s_nb[nb0 + (c >> 2) * 9 + ((c >> 1) & 1) * 3 + (c & 1)][0];
                    const int4 w1 = s_nb[nb0 + (c >> 2) * 9 + ((c >> 1) & 1) * 3 + (c & 1)][1];
                    const uint32_t t0 = (uint32_t)(l[0] - w0.y), t1 = (uint32_t)(l[1] - w0.z), t2 = (uint32_t)(l[2] - w0.w);
                    const bool ok = t0 < (uint32_t)w1.x && t1 < (uint32_t)w1.y && t2 < (uint32_t)w1.z;
                    eo[c] = ok ? (uint32_t)w0.x + 8u * (t0 + (uint32_t)w1.x * t1 + (uint32_t)w1.w * t2) : a.zero_off;
                    first = (ok && first < 0) ? c : first;
                }
                const int cown = (lower[0] < 0 ? 1 : 0) + (lower[1] < 0 ? 2 : 0) + (lower[2] < 0 ? 4 : 0);
                // counted once in the grid sums: by the lowest covering brick
                const uint32_t fw = (uint32_t)a.fixed[nid] | ((first == cown) ? 0x100u : 0u);
                const unsigned long long mk = marks_in[nid];
                real sum[kQ];
    #pragma unroll
                for (int q = 0; q < kQ; ++q) sum[q] = real(0);
    #pragma unroll
                for (int c = 0; c < 8; ++c) {
                    real rec[8];
                    load8(arena_in + eo[c], rec);
    #pragma unroll
                    for (int q = 0; q < kQ; ++q) sum[q] += rec[q];
                }
    #pragma unroll
                for (int q = 0; q < kQ; ++q) s_sum[kQ * r + q] = sum[q];
                unsigned long long *wt = reinterpret_cast<unsigned long long *>(s_node + 6 * r);
                wt[0] = mk;
                wt[1] = fw;
            }
            if ((uint32_t)tid < count) pt0 = load_point(a.p, start + tid, pre_al);
            // Pass 2: the grid update of each window node (…_MP.cpp:209-265)
            double gs[3] = {0.0, 0.0, 0.0};   // owned nodes: mass, reaction, active count
            for (int r = tid; r < nwin; r += kPBlock) {
                const unsigned long long *wt = reinterpret_cast<const unsigned long long *>(s_node + 6 * r);
                const unsigned long long mk = wt[0];
                const uint32_t fw = (uint32_t)wt[1], fx = fw & 0xFFu;
                // each stash slot is read, then zeroed for the P2G arena, by the
                // thread that wrote it
                real sum[kQ];
    #pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    sum[q] = s_sum[kQ * r + q];
                    s_sum[kQ * r + q] = real(0);
                }
                // the rest of the last 8-byte arena word that the stash ends in
                if (FP32 && r == nwin - 1 && (stash & 1)) s_sum[stash] = real(0);
                const real mn = sum[0];
                real vn[3], fn[3], ac[3];
                node_damp<real>(sum, m_min, damp, vn, fn, ac);
                if (mn > m_min)
    #pragma unroll
                    for (int q = 0; q < 3; ++q) vn[q] += ac[q] * dtg;
                double react = 0.0;
                if (fx & 1) { vn[0] = real(0); ac[0] = real(0); }
                if (fx & 2) { vn[1] = real(0); react = (double)fn[1]; ac[1] = real(0); }
                if (fx & 4) { vn[2] = real(0); ac[2] = real(0); }
                if ((uint32_t)(mk >> 32) == st.epoch + 1u) {   // platen override (…_MP.cpp:245-265)
                    const int j = (int)(mk & 0xFFFFFFFFull);
    #pragma unroll
                    for (int q = 0; q < 3; ++q) { vn[q] = (real)dc_v_in[3 * j + q]; ac[q] = real(0); }
                }
                real *wn = s_node + 6 * r;
    #pragma unroll
                for (int q = 0; q < 3; ++q) {
                    wn[q] = vn[q];
                    wn[3 + q] = (fabs(mn) > m_min) ? ac[q] : real(0);
                }
                if (fw & 0x100u) {
                    gs[0] += (double)mn;
                    gs[1] += react;
                    gs[2] += (mn > real(0)) ? 1.0 : 0.0;
                    if (a.fy_slot && (fx & 2)) {
                        const int lz = (int)(((float)r + 0.5f) * rxy), rem = r - lz * own1.w;
                        const int ly = (int)(((float)rem + 0.5f) * rx), lx = rem - ly * wext[0];
                        const int l[3] = {wlo[0] + lx, wlo[1] + ly, wlo[2] + lz};
                        const long long nid = (O[0] - 1 + l[0]) + (long long)g.nodes[0] * ((O[1] - 1 + l[1]) + (long long)g.nodes[1] * (O[2] - 1 + l[2]));
                        a.fy_val[a.fy_slot[nid]] = react;
                    }
                }
            }
            // wave sums to LDS; the block's sum is taken after the barrier that
            // precedes the point chunks (no barrier of its own)
    #pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k != 1 && !sums) continue;   // only the reaction is read every step
                const double w = wave_sum_dpp(gs[k]);
                if (lane == 0) s_red[wv][k] = w;
            }
        }

        // the arena past the stash (the stash zeroes itself): no barrier between
        // the grid update and this clear
        const int clear0 = (stash * (int)sizeof(real) + 7) / 8;
        for (int i = clear0 + tid; i < kQ * AN; i += kPBlock) s_arena[i] = 0;
        if (tid < 24) s_max[tid >> 3][tid & 7] = 0u;
        int H = 6;   // a node takes at most one contribution per point: count < 2^(H-1)
        while ((1u << (H - 1)) <= count) ++H;
        // Fixed-point exponents.  Agreement (FP64 state, the arena rebuild):
        // per chunk, a block-wide running maximum of the contribution bounds
        // sets each field's exponent, and the arena is rescaled when it grows.
        // A priori (FP32 state, a step): each field's exponent is the brick's
        // largest bound of the previous step plus one, with 27 fraction bits,
        // so that contributions up to 2^4 times that bound still convert
        // exactly to 32 bits; no per-chunk agreement.  The block checks its
        // bounds after the last chunk and, when one outgrew the exponent (or
        // fell 8x below it, or the field was zero before), redoes its P2G
        // from the new state with agreement.
        constexpr bool APRI = FP32 && G2P;
        int FRAC = FP32 ? (APRI ? kFracApri : kFrac32) : 61 - H;
        int ecur[kQ];                        // block-uniform field exponents (value < 2^e)
        uint32_t run_max[kQ];                // block-uniform running max of the bound bits
        uint32_t rmax[kQ];                   // a priori: this lane's running max of the bound bits
    #pragma unroll
        for (int q = 0; q < kQ; ++q) {
            run_max[q] = 0u;
            rmax[q] = 0u;
            ecur[q] = kUnsetExp;
        }
        if constexpr (APRI) {
            const unsigned long long ep = *reinterpret_cast<const unsigned long long *>(a.exps[par] + 8 * (size_t)slot);
    #pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const int e = (int)(int8_t)(uint8_t)(ep >> (8 * q));
                ecur[q] = (e == kExpNone) ? kUnsetExp : e + 1;
            }
        }
        double red[5] = {0, 0, 0, 0, 0};     // this lane's running check sums (chunk order)
        int flags = 0;
        int bmin[3] = {1 << 20, 1 << 20, 1 << 20}, bmax[3] = {-1, -1, -1};
        const PArrays<real> &P = a.p;
        // FP32 position increment scale: dt / h * 2^26 per axis
        const float qscale[3] = {(float)(dtd * (double)g.ihf[0] * kQOne), (float)(dtd * (double)g.ihf[1] * kQOne),
                                 (float)(dtd * (double)g.ihf[2] * kQOne)};

        // P2G inputs of step n+1 for one point (…_MP.cpp:106-185): arena key
        // (or -1), bases, m v, V sigma, f_ext and the per-field bounds of
        // |contribution| (N <= 1, |dN/dx_j| <= 1/h_j); updates the flags and
        // the touched box.
        auto p2g_inputs = [&](const real x[3], const int32_t xq[3], real m, const real v[3], real V, const real s[6],
                              uint32_t i, int &key, Axis<real> &bx, Axis<real> &by, Axis<real> &bz, real mv[3],
                              real vs[6], real fe[3], uint32_t bq[kQ]) {
            axes_of(x, xq, g, O, bx, by, bz);
            const int cl[3] = {bx.cell, by.cell, bz.cell};
            const real cu[3] = {bx.u, by.u, bz.u};
            int lc[3];
            bool ok = true;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                lc[q] = cl[q] - (O[q] - 1);
                if (cl[q] < 0 || cl[q] >= g.cells[q]) { flags |= FLAG_OUT_OF_GRID; ok = false; }
                if (lc[q] < 0 || lc[q] > g.S[q] + 1) { flags |= FLAG_ARENA; ok = false; }
                if (cu[q] < real(-0.5) || cu[q] >= (real)g.S[q] + real(0.5)) flags |= FLAG_REBIN;
            }
            if (ok) {
                key = lc[0] + A0 * lc[1] + A01 * lc[2];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    bmin[q] = min(bmin[q], lc[q]);
                    bmax[q] = max(bmax[q], lc[q] + 1);
                }
                if (g.has_fext)
#pragma unroll
                    for (int q = 0; q < 3; ++q) fe[q] = elt(P.fext[q], i);
#pragma unroll
                for (int q = 0; q < 3; ++q) mv[q] = m * v[q];
#pragma unroll
                for (int q = 0; q < 6; ++q) vs[q] = V * s[q];
                // bounds of |contribution| per field (N <= 1, |dN/dx_j| <= 1/h_j)
                const float ih0 = (float)bx.ih, ih1 = (float)by.ih, ih2 = (float)bz.ih;
                bq[0] = fbits(fabsf((float)m) * 1.0001f);
                bq[1] = fbits(fabsf((float)mv[0]) * 1.0001f);
                bq[2] = fbits(fabsf((float)mv[1]) * 1.0001f);
                bq[3] = fbits(fabsf((float)mv[2]) * 1.0001f);
                bq[4] = fbits((fabsf((float)vs[0]) * ih0 + fabsf((float)vs[3]) * ih1 + fabsf((float)vs[5]) * ih2 + fabsf((float)fe[0])) * 1.0001f);
                bq[5] = fbits((fabsf((float)vs[3]) * ih0 + fabsf((float)vs[1]) * ih1 + fabsf((float)vs[4]) * ih2 + fabsf((float)fe[1])) * 1.0001f);
                bq[6] = fbits((fabsf((float)vs[5]) * ih0 + fabsf((float)vs[4]) * ih1 + fabsf((float)vs[2]) * ih2 + fabsf((float)fe[2])) * 1.0001f);
            }
        };
        // Block-wide exponent agreement of one chunk (see above).
        auto agree = [&](uint32_t c0, const uint32_t bq[kQ]) {
            // ---- block-wide fixed-point exponents: running max of the bounds.
            // Chunk c writes and reads buffer c % 3.  After the barrier below a
            // wave may run ahead into chunk c + 1 and write its maxima while a
            // slower wave still reads chunk c's: the buffers differ, and no wave
            // reaches chunk c + 2 (buffer c + 2 = c - 1 mod 3) before every wave
            // has passed chunk c + 1's barrier, long after its chunk-c read.  Each
            // wave adds the running max of the chunks before (block-uniform), so
            // the buffer holds the running max; wave 0 clears buffer c + 2 after
            // this chunk's barrier, when chunk c - 1's reads of it are complete.
            const int mbuf = (int)((c0 / kPBlock) % 3u);
    #pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const uint32_t w = max(wave_umax_dpp(bq[q]), run_max[q]);
                if (lane == 0 && w) atomicMax(&s_max[mbuf][q], w);
            }
            __syncthreads();
            if (tid < 8) s_max[mbuf == 0 ? 2 : mbuf - 1][tid] = 0u;
            {
                bool any = false;
                int sh[kQ];
    #pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    const uint32_t mb = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_max[mbuf][q]);
                    run_max[q] = mb;
                    // positive float bits: biased exponent + 1 bounds the value
                    const int e = (mb == 0u) ? kUnsetExp : max((int)(mb >> 23) - 126, kMinExp);
                    const int en = __builtin_amdgcn_readfirstlane(e);
                    sh[q] = (ecur[q] != kUnsetExp && en > ecur[q]) ? en - ecur[q] : 0;
                    any |= sh[q] != 0;
                    ecur[q] = en;
                }
                if (any) {   // the exponent grew: rescale what the arena holds
                    for (int idx = tid; idx < kQ * AN; idx += kPBlock) {
                        int shq = 0;
    #pragma unroll
                        for (int q = 0; q < kQ; ++q) shq = (idx % kQ == q) ? sh[q] : shq;
                        if (shq) s_arena[idx] = (shq >= 63) ? (s_arena[idx] >> 63) : (s_arena[idx] >> shq);
                    }
                    __syncthreads();
                }
            }
        };
        // P2G adds of one point: its 8 x 7 contributions in fixed point.
        auto p2g_adds = [&](int key, const Axis<real> &bx, const Axis<real> &by, const Axis<real> &bz, real m,
                            const real mv[3], const real vs[6], const real fe[3]) {
                real sc[kQ];
    #pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    if constexpr (FP32)
                        sc[q] = (ecur[q] == kUnsetExp) ? 0.f : __int_as_float((127 + FRAC - ecur[q]) << 23);
                    else
                        sc[q] = (ecur[q] == kUnsetExp) ? 0.0 : ldexp(1.0, FRAC - ecur[q]);
                }
                // scaled per-point factors (power-of-two scaling is exact)
                const real ms = m * sc[0], mvs[3] = {mv[0] * sc[1], mv[1] * sc[2], mv[2] * sc[3]};
                const real f4[4] = {-vs[0] * sc[4], -vs[3] * sc[4], -vs[5] * sc[4], fe[0] * sc[4]};
                const real f5[4] = {-vs[3] * sc[5], -vs[1] * sc[5], -vs[4] * sc[5], fe[1] * sc[5]};
                const real f6[4] = {-vs[5] * sc[6], -vs[4] * sc[6], -vs[2] * sc[6], fe[2] * sc[6]};
                long long *rowp[4];   // arena records of the (ay, az) corner rows
    #pragma unroll
                for (int r = 0; r < 4; ++r) rowp[r] = s_arena + kQ * (key + A0 * (r >> 1) + A01 * (r & 1));
                Hat32 w;
                if constexpr (FP32) w = hat32(bx, by, bz);
    #pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int ax = (k >> 2) & 1, ay = (k >> 1) & 1, az = k & 1;
                    real N, G[3];
                    if constexpr (FP32) {
                        N = w.wxy[2 * ax + ay] * w.wz[az];
                        G[0] = w.gx[ax] * w.wyz[2 * ay + az];
                        G[1] = w.gy[ay] * w.wxz[2 * ax + az];
                        G[2] = w.gz[az] * w.wxy[2 * ax + ay];
                    } else {
                        corner(bx, by, bz, k, N, G);
                    }
                    real c[kQ];
                    c[0] = N * ms;
                    c[1] = N * mvs[0];
                    c[2] = N * mvs[1];
                    c[3] = N * mvs[2];
                    c[4] = (G[0] * f4[0] + G[1] * f4[1] + G[2] * f4[2]) + N * f4[3];
                    c[5] = (G[1] * f5[1] + G[0] * f5[0] + G[2] * f5[2]) + N * f5[3];
                    c[6] = (G[2] * f6[2] + G[0] * f6[0] + G[1] * f6[1]) + N * f6[3];
                    unsigned long long *rec = reinterpret_cast<unsigned long long *>(rowp[2 * ay + az] + kQ * ax);
    #pragma unroll
                    for (int q = 0; q < kQ; ++q) {
                        long long iv;
                        // round to nearest even: the one-instruction floor(x + 1/2)
                        // (v_cvt_rpi_i32_f32) rounds odd integers above 2^23 up,
                        // since its add rounds, which biased the node sums
                        if constexpr (FP32) iv = (long long)__float2int_rn((float)c[q]);
                        else iv = __double2ll_rn((double)c[q]);
                        atomicAdd(rec + q, (unsigned long long)iv);
                    }
                }
        };
        __syncthreads();
        if (G2P && tid < 3) {
            double t = 0.0;
            for (int w = 0; w < kPBlock / 64; ++w) t += s_red[w][tid];
            a.gpart[3 * (size_t)slot + tid] = (tid == 1 || sums) ? t : 0.0;
        }

        for (uint32_t c0 = 0; c0 < count; c0 += kPBlock) {
            const bool active = (c0 + tid) < count;
            const uint32_t i = start + c0 + tid;
            int key = -1;
            Axis<real> bx, by, bz;
            real m = real(0), mv[3] = {real(0), real(0), real(0)}, vs[6], fe[3] = {real(0), real(0), real(0)};
            uint32_t bq[kQ] = {0, 0, 0, 0, 0, 0, 0};
    #pragma unroll
            for (int q = 0; q < 6; ++q) vs[q] = real(0);

            if (active) {
                const PointState<real> pt = (G2P && c0 == 0) ? pt0 : load_point(P, i, pre_al);
                real x[3], v[3], F[9], s[6];
                int32_t xq[3];
    #pragma unroll
                for (int q = 0; q < 3; ++q) {
                    x[q] = pt.x[q];
                    xq[q] = pt.xq[q];
                    v[q] = pt.v[q];
                }
    #pragma unroll
                for (int q = 0; q < 9; ++q) F[q] = pt.F[q];
    #pragma unroll
                for (int q = 0; q < 6; ++q) s[q] = pt.s[q];
                m = pt.m;
                const real V0 = pt.V0;
                const real al_pre = pt.al;
                real V = real(0);

                if (G2P) {
                    // ---- G2P of step n (…_MP.cpp:279-297)
                    axes_of(x, xq, g, O, bx, by, bz);
                    const bool oob = bx.cell < 0 || bx.cell >= g.cells[0] || by.cell < 0 || by.cell >= g.cells[1] ||
                                     bz.cell < 0 || bz.cell >= g.cells[2];
                    // the 8 corner nodes in the window (the P2G of step n put them there)
                    const int wl[3] = {bx.cell - (O[0] - 1) - wlo[0], by.cell - (O[1] - 1) - wlo[1], bz.cell - (O[2] - 1) - wlo[2]};
                    const bool outw = wl[0] < 0 || wl[0] + 1 >= wext[0] || wl[1] < 0 || wl[1] + 1 >= wext[1] ||
                                      wl[2] < 0 || wl[2] + 1 >= wext[2];
                    if (oob || outw) {
                        flags |= oob ? FLAG_OUT_OF_GRID : FLAG_ARENA;
                    } else {
                        const real *w0 = s_node + 6 * (wl[0] + wext[0] * (wl[1] + wext[1] * wl[2]));
                        const int sy_ = wext[0], sz_ = wext[0] * wext[1];
                        real L[9];
    #pragma unroll
                        for (int q = 0; q < 9; ++q) L[q] = real(0);
                        real dx[3] = {real(0), real(0), real(0)};
                        if constexpr (FP32) {
                            // 8 corners fully unrolled; node record {v.xyz a.x | a.y a.z m -}
                            const Hat32 w = hat32(bx, by, bz);
                            float av[3] = {0.f, 0.f, 0.f};
    #pragma unroll
                            for (int k = 0; k < 8; ++k) {
                                const int ax = (k >> 2) & 1, ay = (k >> 1) & 1, az = k & 1;
                                const float N = w.wxy[2 * ax + ay] * w.wz[az];
                                const float G[3] = {w.gx[ax] * w.wyz[2 * ay + az], w.gy[ay] * w.wxz[2 * ax + az],
                                                    w.gz[az] * w.wxy[2 * ax + ay]};
                                const real *wp = w0 + 6 * (ax + sy_ * ay + sz_ * az);
                                const float vn[3] = {wp[0], wp[1], wp[2]}, an[3] = {wp[3], wp[4], wp[5]};
    #pragma unroll
                                for (int q = 0; q < 3; ++q) {
                                    av[q] += N * an[q];
                                    dx[q] += N * vn[q];
                                }
    #pragma unroll
                                for (int cc = 0; cc < 3; ++cc)
    #pragma unroll
                                    for (int r = 0; r < 3; ++r) L[3 * cc + r] += vn[r] * G[cc];
                            }
    #pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                v[q] += av[q] * dt;
                                xq[q] += __float2int_rn((float)dx[q] * qscale[q]);
                            }
                        } else {
    #pragma unroll 2
                            for (int k = 0; k < 8; ++k) {
                                real N, G[3];
                                corner(bx, by, bz, k, N, G);
                                const real *wp = w0 + 6 * (((k >> 2) & 1) + sy_ * ((k >> 1) & 1) + sz_ * (k & 1));
    #pragma unroll
                                for (int q = 0; q < 3; ++q) v[q] += N * wp[3 + q] * dt;
    #pragma unroll
                                for (int q = 0; q < 3; ++q) x[q] += N * wp[q] * dt;
    #pragma unroll
                                for (int cc = 0; cc < 3; ++cc)
    #pragma unroll
                                    for (int r = 0; r < 3; ++r) L[3 * cc + r] += wp[r] * G[cc];
                            }
                        }
                        // F += (L F) dt  (…_MP.cpp:299); FP32 state: D += (L + L D) dt
                        real T[9];
    #pragma unroll
                        for (int cc = 0; cc < 3; ++cc)
    #pragma unroll
                            for (int r = 0; r < 3; ++r)
                                T[3 * cc + r] = (FP32 ? L[3 * cc + r] : real(0)) +
                                                (L[r] * F[3 * cc] + L[3 + r] * F[3 * cc + 1] + L[6 + r] * F[3 * cc + 2]);
    #pragma unroll
                        for (int q = 0; q < 9; ++q) F[q] += T[q] * dt;
                        // volume from the new F (…_MP.cpp:301-302), FP64
                        {
                            double Fd[9];
    #pragma unroll
                            for (int q = 0; q < 9; ++q) Fd[q] = fd(F[q], q);
                            V = (real)(det3(Fd) * (double)V0);
                        }
                        // write back the kinematic state before the material update
    #pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            if (FP32) elt(P.xq[q], i) = xq[q];
                            else elt(P.x[q], i) = x[q];
                            elt(P.v[q], i) = v[q];
                        }
    #pragma unroll
                        for (int q = 0; q < 9; ++q) elt(P.F[q], i) = F[q];
                        material_update<real>(g, P, a.mats, i, L, dt, F, s, umat, pre_al, al_pre);
                    }
                } else {
                    double Fd[9];
    #pragma unroll
                    for (int q = 0; q < 9; ++q) Fd[q] = fd(F[q], q);
                    V = (real)(det3(Fd) * (double)V0);
                }
                // ---- FP64 checks of the new state (PhysicsEngine.cpp:73-79), when
                // the step's sums are read; the non-finite flag in every step
                if (!sums) {
                    const real sv = fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
                    const real ssum = s[0] + s[1] + s[2] + s[3] + s[4] + s[5];
                    if (!(isfinite(sv) && isfinite(ssum) && (FP32 || isfinite(x[0] + x[1] + x[2])))) flags |= FLAG_NAN;
                } else {
                    const double md = (double)m, vd[3] = {(double)v[0], (double)v[1], (double)v[2]};
                    const double vv = vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2];
                    red[0] += md * vv;
                    red[1] += md * vd[0];
                    red[2] += md * vd[1];
                    red[3] += md * vd[2];
                    const real ssum = s[0] + s[1] + s[2] + s[3] + s[4] + s[5];
                    const bool fin = isfinite(vv) && isfinite(ssum) && (FP32 || isfinite(x[0] + x[1] + x[2]));
                    if (!fin) { red[4] += 1.0; flags |= FLAG_NAN; }
                }
                p2g_inputs(x, xq, m, v, V, s, i, key, bx, by, bz, mv, vs, fe, bq);
            }

            if constexpr (APRI) {
    #pragma unroll
                for (int q = 0; q < kQ; ++q) rmax[q] = max(rmax[q], bq[q]);
            } else {
                agree(c0, bq);
            }

            // ---- P2G: every lane adds its 8 x 7 contributions (fixed point)
            if (key >= 0) p2g_adds(key, bx, by, bz, m, mv, vs, fe);
        }

        // ---- touched node box of the brick (and, a priori, the wave maxima
        // of the bounds): wave values first, so that one barrier covers both
        // them and the arena adds
    #pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int lo = wave_imin_dpp(bmin[q]), hi = wave_imax_dpp(bmax[q]);
            if (lane == 0) { s_box[wv][2 * q] = lo; s_box[wv][2 * q + 1] = hi; }
        }
        if constexpr (APRI) {
    #pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const uint32_t w = wave_umax_dpp(rmax[q]);
                if (lane == 0) s_rmax[wv][q] = w;
            }
        }
        __syncthreads();
        // the block's exponents of this step (the arena's record for the next)
        int enext[kQ];
    #pragma unroll
        for (int q = 0; q < kQ; ++q) enext[q] = ecur[q];
        if constexpr (APRI) {
            bool redo = false;
    #pragma unroll
            for (int q = 0; q < kQ; ++q) {
                uint32_t mb = 0u;
                for (int w = 0; w < kPBlock / 64; ++w) mb = max(mb, s_rmax[w][q]);
                const int e = (mb == 0u) ? kUnsetExp : max((int)(mb >> 23) - 126, kMinExp);
                enext[q] = e;
                if (e != kUnsetExp && (ecur[q] == kUnsetExp || e > ecur[q] + 4 || e < ecur[q] - 3)) redo = true;
            }
            if (redo) {
                // P2G again from the state just written, with agreement; the
                // flags, check sums and touched box of the pass above stand
                for (int i = tid; i < kQ * AN; i += kPBlock) s_arena[i] = 0;
                if (tid < 24) s_max[tid >> 3][tid & 7] = 0u;
                FRAC = kFrac32;
    #pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    run_max[q] = 0u;
                    ecur[q] = kUnsetExp;
                }
                const int fl0 = flags;
                const int bsave[6] = {bmin[0], bmin[1], bmin[2], bmax[0], bmax[1], bmax[2]};
                __syncthreads();
                for (uint32_t c0 = 0; c0 < count; c0 += kPBlock) {
                    const uint32_t i = start + c0 + tid;
                    int key = -1;
                    Axis<real> bx, by, bz;
                    real m = real(0), mv[3] = {real(0), real(0), real(0)}, vs[6] = {real(0), real(0), real(0), real(0), real(0), real(0)};
                    real fe[3] = {real(0), real(0), real(0)};
                    uint32_t bq[kQ] = {0, 0, 0, 0, 0, 0, 0};
                    if (c0 + tid < count) {
                        const PointState<real> pt = load_point(P, i, false);
                        double Fd[9];
    #pragma unroll
                        for (int q = 0; q < 9; ++q) Fd[q] = fd(pt.F[q], q);
                        const real V = (real)(det3(Fd) * (double)pt.V0);
                        m = pt.m;
                        p2g_inputs(pt.x, pt.xq, pt.m, pt.v, V, pt.s, i, key, bx, by, bz, mv, vs, fe, bq);
                    }
                    agree(c0, bq);
                    if (key >= 0) p2g_adds(key, bx, by, bz, m, mv, vs, fe);
                }
                flags = fl0;
                bmin[0] = bsave[0]; bmin[1] = bsave[1]; bmin[2] = bsave[2];
                bmax[0] = bsave[3]; bmax[1] = bsave[4]; bmax[2] = bsave[5];
                __syncthreads();
            }
        }
        if (tid == 0 && FP32) {
            unsigned long long ep = 0ull;
    #pragma unroll
            for (int q = 0; q < kQ; ++q)
                ep |= (unsigned long long)(uint8_t)(int8_t)(enext[q] == kUnsetExp ? kExpNone : enext[q]) << (8 * q);
            *reinterpret_cast<unsigned long long *>(a.exps[G2P ? par ^ 1 : par] + 8 * (size_t)slot) = ep;
        }
        int box[6];
    #pragma unroll
        for (int q = 0; q < 3; ++q) {
            int lo = 1 << 20, hi = -1;
            for (int w = 0; w < kPBlock / 64; ++w) { lo = min(lo, s_box[w][2 * q]); hi = max(hi, s_box[w][2 * q + 1]); }
            box[2 * q] = lo;
            box[2 * q + 1] = hi;
        }
        if (tid == 0) {
            const bool empty = box[1] < box[0] || box[3] < box[2] || box[5] < box[4];
            binfo_out[brick] = make_int4(empty ? -1 : slot, box[0] | (box[1] << 16), box[2] | (box[3] << 16),
                                       box[4] | (box[5] << 16));
        }

        // ---- arena out (fixed point -> real) over the touched box only, as
        // box-compact node records of 8 reals {m, p.xyz, f.xyz, 0}
        if (box[1] >= box[0]) {
            const int bx = box[1] - box[0] + 1, by = box[3] - box[2] + 1, bz = box[5] - box[4] + 1;
            const int nbox = bx * by * bz;
            real *out = arena_out + (size_t)slot * 8 * AN;
            double scl[kQ];
    #pragma unroll
            for (int q = 0; q < kQ; ++q) scl[q] = (ecur[q] == kUnsetExp) ? 0.0 : ldexp(1.0, ecur[q] - FRAC);
            // record index -> node as in the grid update (float quotients, exact here)
            const float rbx = 1.0f / (float)bx, rbxy = 1.0f / (float)(bx * by);
            for (int r = tid; r < nbox; r += kPBlock) {
                const int qz = (int)(((float)r + 0.5f) * rbxy), rem = r - qz * (bx * by);
                const int qy = (int)(((float)rem + 0.5f) * rbx);
                const int lx = box[0] + rem - qy * bx, ly = box[2] + qy, lz = box[4] + qz;
                const long long *src = s_arena + kQ * (lx + A0 * ly + A01 * lz);
                real rec[8];
    #pragma unroll
                for (int q = 0; q < kQ; ++q) rec[q] = (real)((double)src[q] * scl[q]);
                rec[7] = real(0);
                store8(out + 8 * (size_t)r, rec);
            }
        }
        // ---- check sums: one fixed-order partial per wave, written by the wave
        // itself (no block barrier at the end); flags: OR over the wave, only when
        // a lane has one
        if (sums)
    #pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double wsum = wave_sum_dpp(red[k]);
                if (lane == 0) a.kp_part[5 * ((size_t)slot * (kPBlock / 64) + wv) + k] = wsum;
            }
        if (__ballot(flags != 0)) {
            int fl = flags;
    #pragma unroll
            for (int o = 32; o > 0; o >>= 1) fl |= __shfl_xor(fl, o, 64);
            if (lane == 0) atomicOr(a.flags, fl);
        }
        if constexpr (!PERSIST) {
            return;
        } else {
            __syncthreads();   // the next item reuses the LDS
        }
    }
}

}  // namespace mpm

namespace mpm {

// ------------------------------------------------------------ grid readback
// Dense node arrays (x fastest) written by k_grid_state / k_c_nodes.
struct NodeOut {
    uint8_t *active;
    double *mass, *v, *f, *ft;     // 1, 3, 3, 3 per node
    double *k, *kg;                // contact kernel grid: 1, 3 per node (may be null)
};

// The grid of the next plain step from the current arenas (the P2G sums of
// the current points, written by the last step or the rebuild): for every
// node of every occupied brick's window that the brick owns (it is the
// lowest covering brick, as in the step's grid sums), the covering arenas'
// sums in ascending brick order and the grid update of k_particle's phase A
// (…_MP.cpp:209-265) with step record `st`.  The candidate order and the
// arithmetic are those of phase A, so the values are the ones the next step
// computes.  Output: mass, velocity after the update, force after damping,
// boundary conditions and the platen override, Force_Temp.y (the reaction
// of y-fixed nodes).  One block per occupied slot.
template <typename real>
__global__ void __launch_bounds__(256) k_grid_state(GridParams g, const real *arena, const int4 *binfo,
                                                    uint32_t zero_off, const uint8_t *fixed,
                                                    const unsigned long long *marks, const double *dc_v,
                                                    const uint32_t *occ_list, const int32_t *n_occ, StepRec st,
                                                    NodeOut o)
{
    const int slot = blockIdx.x;
    if (slot >= *n_occ) return;
    const uint32_t brick = occ_list[slot];
    const int bc[3] = {(int)(brick % g.nb[0]), (int)((brick / g.nb[0]) % g.nb[1]), (int)(brick / (g.nb[0] * g.nb[1]))};
    const int O[3] = {bc[0] * g.S[0], bc[1] * g.S[1], bc[2] * g.S[2]};
    const int AN = g.arenaN;
    // the 27 neighbour windows (dz, dy, dx ascending), as phase A stages them
    __shared__ int4 s_nb[27][2];
    if (threadIdx.x < 27) {
        const int t = threadIdx.x;
        const int d[3] = {t % 3 - 1, (t / 3) % 3 - 1, t / 9 - 1};
        const int x = bc[0] + d[0], y = bc[1] + d[1], z = bc[2] + d[2];
        int4 bi = make_int4(-1, 0, 0, 0);
        if (x >= 0 && y >= 0 && z >= 0 && x < g.nb[0] && y < g.nb[1] && z < g.nb[2])
            bi = binfo[x + (size_t)g.nb[0] * (y + (size_t)g.nb[1] * z)];
        const int pk[3] = {bi.y, bi.z, bi.w};
        int lo[3], ext[3];
        for (int q = 0; q < 3; ++q) {
            lo[q] = (pk[q] & 0xFFFF) + d[q] * g.S[q];
            ext[q] = (bi.x >= 0) ? (pk[q] >> 16) - (pk[q] & 0xFFFF) + 1 : 0;
        }
        s_nb[t][0] = make_int4((int)((uint32_t)max(bi.x, 0) * 8u * (uint32_t)AN), lo[0], lo[1], lo[2]);
        s_nb[t][1] = make_int4(ext[0], ext[1], ext[2], ext[0] * ext[1]);
    }
    __syncthreads();
    const int4 own0 = s_nb[13][0], own1 = s_nb[13][1];
    const int wlo[3] = {own0.y, own0.z, own0.w}, wext[3] = {own1.x, own1.y, own1.z};
    const int nwin = wext[0] * wext[1] * wext[2];
    const real dtg = (real)st.dt, m_min = (real)g.m_min, damp = (real)g.damping;
    for (int r = threadIdx.x; r < nwin; r += 256) {
        const int lz = r / own1.w, rem = r - lz * own1.w;
        const int ly = rem / wext[0], lx = rem - ly * wext[0];
        const int l[3] = {wlo[0] + lx, wlo[1] + ly, wlo[2] + lz};
        int lower[3];
        for (int q = 0; q < 3; ++q) lower[q] = (l[q] <= 2 && bc[q] > 0) ? -1 : 0;
        const int nb0 = (lower[2] + 1) * 9 + (lower[1] + 1) * 3 + (lower[0] + 1);
        uint32_t eo[8];
        int first = -1;
        for (int c = 0; c < 8; ++c) {
            const int4 w0 = s_nb[nb0 + (c >> 2) * 9 + ((c >> 1) & 1) * 3 + (c & 1)][0];
            const int4 w1 = s_nb[nb0 + (c >> 2) * 9 + ((c >> 1) & 1) * 3 + (c & 1)][1];
            const uint32_t t0 = (uint32_t)(l[0] - w0.y), t1 = (uint32_t)(l[1] - w0.z), t2 = (uint32_t)(l[2] - w0.w);
            const bool ok = t0 < (uint32_t)w1.x && t1 < (uint32_t)w1.y && t2 < (uint32_t)w1.z;
            eo[c] = ok ? (uint32_t)w0.x + 8u * (t0 + (uint32_t)w1.x * t1 + (uint32_t)w1.w * t2) : zero_off;
            first = (ok && first < 0) ? c : first;
        }
        const int cown = (lower[0] < 0 ? 1 : 0) + (lower[1] < 0 ? 2 : 0) + (lower[2] < 0 ? 4 : 0);
        if (first != cown) continue;   // another brick owns this node
        const long long nid = (O[0] - 1 + l[0]) + (long long)g.nodes[0] * ((O[1] - 1 + l[1]) + (long long)g.nodes[1] * (O[2] - 1 + l[2]));
        real sum[kQ];
        for (int q = 0; q < kQ; ++q) sum[q] = real(0);
        for (int c = 0; c < 8; ++c) {
            real rec[8];
            load8(arena + eo[c], rec);
            for (int q = 0; q < kQ; ++q) sum[q] += rec[q];
        }
        const uint32_t fx = fixed[nid];
        const unsigned long long mk = marks[nid];
        const real mn = sum[0];
        real vn[3], fn[3], ac[3];
        node_damp<real>(sum, m_min, damp, vn, fn, ac);   // damping (…_MP.cpp:209-222)
        if (mn > m_min)
            for (int q = 0; q < 3; ++q) vn[q] += ac[q] * dtg;
        double react = 0.0;
        if (fx & 1) { vn[0] = real(0); fn[0] = real(0); }
        if (fx & 2) { vn[1] = real(0); react = (double)fn[1]; fn[1] = real(0); }
        if (fx & 4) { vn[2] = real(0); fn[2] = real(0); }
        if ((uint32_t)(mk >> 32) == st.epoch + 1u) {   // platen override (…_MP.cpp:245-265)
            const int j = (int)(mk & 0xFFFFFFFFull);
            for (int q = 0; q < 3; ++q) { vn[q] = (real)dc_v[3 * j + q]; fn[q] = real(0); }
        }
        o.active[nid] = mn > real(0) ? 1 : 0;
        o.mass[nid] = (double)mn;
        for (int q = 0; q < 3; ++q) {
            o.v[3 * nid + q] = (double)vn[q];
            o.f[3 * nid + q] = (double)fn[q];
        }
        o.ft[3 * nid + 1] = react;
    }
}

}  // namespace mpm
