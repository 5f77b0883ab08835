// mpm_step_kernels.hpp -- the per-step kernels of the explicit MPM step.
//
//   k_particle<real, G2P>: one workgroup per occupied brick.  For each point of
//     the brick: G2P of step n (gather 8 nodes), F / volume / strain update,
//     constitutive update, state write-back, then the P2G contributions of
//     step n+1 (m, m v, -V sigma.grad N + N f_ext to the 8 corner nodes).
//     Contributions are summed per run of points sharing a cell inside each
//     16-lane DPP row (segmented shuffle scan, lane order), and the run totals
//     are added to the brick's LDS arena as 64-bit fixed point with a
//     per-brick exponent.  Integer addition is associative, so the arena is
//     independent of wave scheduling: the step is bit-reproducible with no
//     float atomics anywhere.  With G2P=false the kernel only rebuilds the
//     arena (after upload or a rebin).
//   k_grid<real>: one workgroup per brick of the dilated active set: merges the
//     arenas covering each owned node in a fixed order, then momentum,
//     damping, boundary conditions, reaction and the platen override.
//
// Reference: PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:41-368.
#pragma once
#include "mpm_device.hpp"

namespace mpm {

// ------------------------------------------------------------- constitutive
// ConstitutiveRelation_calculateIncrement_Elastic.cpp:4-15
template <typename T>
__device__ __forceinline__ void elastic_inc(T E, T nu, const T de[6], T ds[6])
{
    T C = E / (T(1) + nu) / (T(1) - T(2) * nu);
    ds[0] = C * ((T(1) - nu) * de[0] + nu * de[1] + nu * de[2]);
    ds[1] = C * ((T(1) - nu) * de[1] + nu * de[0] + nu * de[2]);
    ds[2] = C * ((T(1) - nu) * de[2] + nu * de[0] + nu * de[1]);
    ds[3] = C * ((T(0.5) - nu) * de[3]);
    ds[4] = C * ((T(0.5) - nu) * de[4]);
    ds[5] = C * ((T(0.5) - nu) * de[5]);
}

// ConstitutiveRelation_calculateState_J2.cpp:4-26
__device__ __forceinline__ double j2_of(const double s[6])
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

// Radial return shared by …PerfectlyPlastic.cpp:4-72 and
// …VonMisesHardening.cpp:4-86; always evaluated in FP64.
__device__ __forceinline__ bool radial_return(double E, double nu, double radius,
                                              const double s[6], const double de[6],
                                              double ds[6], double dep[6])
{
    double st[6], dev[6];
    elastic_inc<double>(E, nu, de, ds);
    for (int i = 0; i < 6; ++i) st[i] = s[i] + ds[i];
    double p = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
    for (int i = 0; i < 6; ++i) dev[i] = st[i];
    dev[0] -= p;
    dev[1] -= p;
    dev[2] -= p;
    double j2 = j2_of(st);
    if (j2 <= radius * radius) {
        for (int i = 0; i < 6; ++i) dep[i] = 0.0;
        return false;
    }
    double sc = radius / sqrt(j2);
    for (int i = 0; i < 6; ++i) dev[i] *= sc;
    for (int i = 0; i < 6; ++i) ds[i] = dev[i] - s[i];
    ds[0] += p;
    ds[1] += p;
    ds[2] += p;
    double G = 0.5 * E / (1.0 + nu);
    double dee[6];
    dee[0] = ds[0] / E - ds[1] * nu / E - ds[2] * nu / E;
    dee[1] = ds[1] / E - ds[0] * nu / E - ds[2] * nu / E;
    dee[2] = ds[2] / E - ds[0] * nu / E - ds[1] * nu / E;
    dee[3] = ds[3] / G;
    dee[4] = ds[4] / G;
    dee[5] = ds[5] / G;
    for (int i = 0; i < 6; ++i) dep[i] = de[i] - dee[i];
    return true;
}

// ConstitutiveRelation_calculateIncrement_NeoHookean.cpp:4-22, FP64.
__device__ __forceinline__ void neohookean_inc(double E, double nu, const double s[6],
                                               const double F[9], double ds[6])
{
    double mu = 0.5 * E / (1.0 + nu);
    double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double J = det3(F);
    double a = lam * log(J) / J;
    double b = mu / J;
    double P[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double fft = F[r] * F[c] + F[3 + r] * F[3 + c] + F[6 + r] * F[6 + c];
            double id = (c == r) ? 1.0 : 0.0;
            P[3 * c + r] = a * id + b * (fft - id);
        }
    ds[0] = P[0] - s[0];
    ds[1] = P[4] - s[1];
    ds[2] = P[8] - s[2];
    ds[3] = P[1] - s[3];
    ds[4] = P[5] - s[4];
    ds[5] = P[6] - s[5];
}

// ------------------------------------------------------------------ locate
// FP64: absolute position, neighbour search and node positions exactly as
// GridPoint_Mediator.h:35 (trunc of x/h) and GridPoint_Factory.h:40-42 (i*L/n).
__device__ __forceinline__ void locate_axis(double x, int a, const GridParams &g, int /*O*/,
                                            int &cell, double &d0, double &d1, double &u)
{
    int i0 = (int)(x / g.h[a]);
    cell = i0;
    double xi0 = ((double)i0 * g.L[a]) / (double)g.cells[a];
    double xi1 = ((double)(i0 + 1) * g.L[a]) / (double)g.cells[a];
    d0 = x - xi0;
    d1 = x - xi1;
    u = x / g.h[a];
    if (!(x >= 0.0)) cell = -1;   // out of grid (or NaN)
}
// FP32: position relative to the brick origin O (in cells).
__device__ __forceinline__ void locate_axis(float xr, int a, const GridParams &g, int O,
                                            int &cell, float &d0, float &d1, float &u)
{
    float h = g.hf[a];
    u = xr / h;
    int ci = (int)floorf(u);
    cell = O + ci;
    d0 = xr - (float)ci * h;
    d1 = xr - (float)(ci + 1) * h;
    u = u + (float)O;
    if (!(u == u)) cell = -1;
}

template <typename real>
struct Bases {
    real s[3][2];   // per-axis shape values for node offset 0/1
    real sg[3][2];  // sign(d)
    real h[3];      // cell size per axis (divided, as the reference)
    int cell[3];
    real u[3];      // continuous cell coordinate
};

template <typename real>
__device__ __forceinline__ void make_bases(const real x[3], const GridParams &g,
                                           const int O[3], Bases<real> &b)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        real d0, d1, u;
        locate_axis(x[a], a, g, O[a], b.cell[a], d0, d1, u);
        b.u[a] = u;
        real h = (sizeof(real) == 8) ? (real)g.h[a] : (real)g.hf[a];
        real s0 = real(1) - fabs(d0) / h;
        real s1 = real(1) - fabs(d1) / h;
        if (s0 < real(0) || s0 > real(1)) s0 = real(0);
        if (s1 < real(0) || s1 > real(1)) s1 = real(0);
        b.s[a][0] = s0;
        b.s[a][1] = s1;
        b.sg[a][0] = sgn(d0);
        b.sg[a][1] = sgn(d1);
        b.h[a] = h;
    }
}

// N and grad N of corner k (x-outer, z-inner as GridPoint_Mediator.h:39-45);
// GridPoint_Mediator.h:83-87 evaluation order.
template <typename real>
__device__ __forceinline__ void corner(const Bases<real> &b, int k, real &N, real G[3])
{
    int ax = (k >> 2) & 1, ay = (k >> 1) & 1, az = k & 1;
    real sx = b.s[0][ax], sy = b.s[1][ay], sz = b.s[2][az];
    N = sx * sy * sz;
    G[0] = -sy * sz * b.sg[0][ax] / b.h[0];
    G[1] = -sx * sz * b.sg[1][ay] / b.h[1];
    G[2] = -sx * sy * b.sg[2][az] / b.h[2];
}

// ------------------------------------------------------- block reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

constexpr int kUnsetExp = -100000;

template <typename real>
struct KpArgs {
    GridParams g;
    PArrays<real> p;
    const NodeRec<real> *nodes;
    real *arena_out;
    const uint32_t *occ_list, *occ_start, *occ_count;
    const int32_t *n_occ;
    int32_t occ_cap;
    const Material *mats;
    const StepRec *steps;
    int32_t *flags;
    double *kp_part;      // per slot: 5 doubles (sum m|v|^2, m v.xyz, nonfinite)
    int32_t n_dc;
    double *dc_x, *dc_v;  // 3 per platen point, FP64
    unsigned long long *marks;
    const double *k2_part;  // per grid block: 3 doubles (mass, reaction, active nodes)
    const int32_t *n_k2;
    double *hist;           // per step reaction history
    int32_t hist_index;
};

template <int D, typename real>
__device__ __forceinline__ void seg_step(real (&c)[8 * kQ], int rl, int st)
{
    const bool take = (rl - D) >= st;
#pragma unroll
    for (int i = 0; i < 8 * kQ; ++i) {
        real t = dpp_r<MPM_ROW_SHR(D)>(c[i]);
        c[i] += take ? t : real(0);
    }
}

// One brick per workgroup.  Dynamic LDS: long long arena[7 * arenaN].
template <typename real, bool G2P>
__global__ void __launch_bounds__(kBlock) k_particle(KpArgs<real> a, int step)
{
    extern __shared__ long long s_arena[];
    __shared__ double s_wmax[kBlock / 64][kQ];
    __shared__ int s_e[kQ], s_shift[kQ], s_any;
    __shared__ double s_red[kBlock / 64][5];

    const GridParams &g = a.g;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, rl = lane & 15;
    const StepRec st = a.steps[step];
    const double dtd = st.dt;
    const real dt = (real)dtd;

    // ---- platen (displacement-controlled) blocks: phase 8 + override marks
    if ((int)blockIdx.x >= a.occ_cap) {
        if (!G2P) return;
        int j = ((int)blockIdx.x - a.occ_cap) * kBlock + tid;
        if (j >= a.n_dc) return;
        double *x = &a.dc_x[3 * j], *v = &a.dc_v[3 * j];
        // …_MP.cpp:360-368: velocity from the TimeLine at the step's start time
        if (st.has_tl) { v[0] = st.vtl[0]; v[1] = st.vtl[1]; v[2] = st.vtl[2]; }
        for (int q = 0; q < 3; ++q) x[q] += v[q] * dtd;
        // override marks for step n+1 (…_MP.cpp:245-265): the highest platen
        // index touching a node wins, as in the reference's sequential loop.
        int i0[3];
        for (int q = 0; q < 3; ++q) i0[q] = (int)(x[q] / g.h[q]);
        // mark for step n+1 carries epoch tag (n+1)+1; 0 means "no mark"
        unsigned long long tag = ((unsigned long long)(st.epoch + 2u) << 32) | (unsigned)j;
        for (int k = 0; k < 8; ++k) {
            int n0 = i0[0] + ((k >> 2) & 1), n1 = i0[1] + ((k >> 1) & 1), n2 = i0[2] + (k & 1);
            if (n0 < 0 || n1 < 0 || n2 < 0 || n0 >= g.nodes[0] || n1 >= g.nodes[1] || n2 >= g.nodes[2]) continue;
            long long nid = n0 + (long long)g.nodes[0] * (n1 + (long long)g.nodes[1] * n2);
            atomicMax(&a.marks[nid], tag);
        }
        return;
    }

    const int slot = blockIdx.x;
    if (slot >= *a.n_occ) return;
    const uint32_t brick = a.occ_list[slot];
    const int bx = brick % g.nb[0], by = (brick / g.nb[0]) % g.nb[1], bz = brick / (g.nb[0] * g.nb[1]);
    const int O[3] = {bx * g.S[0], by * g.S[1], bz * g.S[2]};
    const uint32_t start = a.occ_start[slot], count = a.occ_count[slot];
    const int A0 = g.A[0], A01 = g.A[0] * g.A[1], AN = g.arenaN;

    // reaction history of the previous grid update (…PhysicsEngine.cpp:65-71)
    if (G2P && slot == 0 && wv == 0) {
        double r = 0.0;
        int nk = *a.n_k2;
        for (int i = lane; i < nk; i += 64) r += a.k2_part[3 * i + 1];
        r = wave_sum(r);
        if (lane == 0) a.hist[a.hist_index] = r;
    }

    for (int i = tid; i < kQ * AN; i += kBlock) s_arena[i] = 0;
    if (tid < kQ) s_e[tid] = kUnsetExp;
    int H = 5;
    while ((1u << H) <= count) ++H;
    const int FRAC = 61 - H;
    double red[5] = {0, 0, 0, 0, 0};
    int flags = 0;
    __syncthreads();

    const PArrays<real> &P = a.p;
    for (uint32_t c0 = 0; c0 < count; c0 += kBlock) {
        const bool active = (c0 + tid) < count;
        const uint32_t i = start + c0 + tid;
        real c[8 * kQ];
#pragma unroll
        for (int k = 0; k < 8 * kQ; ++k) c[k] = real(0);
        int key = -1;

        if (active) {
            real x[3], v[3], F[9], s[6];
#pragma unroll
            for (int q = 0; q < 3; ++q) { x[q] = P.x[q][i]; v[q] = P.v[q][i]; }
#pragma unroll
            for (int q = 0; q < 9; ++q) F[q] = P.F[q][i];
#pragma unroll
            for (int q = 0; q < 6; ++q) s[q] = P.s[q][i];
            const real m = P.mass[i], V0 = P.vol0[i];

            if (G2P) {
                // ---- G2P of step n (…_MP.cpp:279-297)
                Bases<real> b;
                make_bases(x, g, O, b);
                bool oob = false;
#pragma unroll
                for (int q = 0; q < 3; ++q) oob |= (b.cell[q] < 0) || (b.cell[q] >= g.cells[q]);
                if (oob) {
                    flags |= FLAG_OUT_OF_GRID;
                } else {
                    const long long nb0 = b.cell[0] + (long long)g.nodes[0] * (b.cell[1] + (long long)g.nodes[1] * b.cell[2]);
                    real L[9];
#pragma unroll
                    for (int q = 0; q < 9; ++q) L[q] = real(0);
                    real xn[3] = {x[0], x[1], x[2]};
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        real N, G[3];
                        corner(b, k, N, G);
                        const long long nid = nb0 + ((k >> 2) & 1) + (long long)g.nodes[0] * (((k >> 1) & 1) + (long long)g.nodes[1] * (k & 1));
                        const NodeRec<real> nr = a.nodes[nid];
#pragma unroll
                        for (int q = 0; q < 3; ++q) v[q] += N * nr.d[3 + q] * dt;
#pragma unroll
                        for (int q = 0; q < 3; ++q) xn[q] += N * nr.d[q] * dt;
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                            for (int r = 0; r < 3; ++r) L[3 * cc + r] += nr.d[r] * G[cc];
                    }
#pragma unroll
                    for (int q = 0; q < 3; ++q) x[q] = xn[q];
                    // F += (L F) dt  (…_MP.cpp:299)
                    real T[9];
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                        for (int r = 0; r < 3; ++r)
                            T[3 * cc + r] = L[r] * F[3 * cc] + L[3 + r] * F[3 * cc + 1] + L[6 + r] * F[3 * cc + 2];
#pragma unroll
                    for (int q = 0; q < 9; ++q) F[q] += T[q] * dt;
                    // strain increment (…_MP.cpp:304-312)
                    real de[6];
                    if (sizeof(real) == 8 && g.strain_form == 0) {
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
                        for (int q = 0; q < 6; ++q) P.eps[q][i] += de[q];
                    // constitutive dispatch (…_MP.cpp:322-353)
                    const Material mt = a.mats[P.mat[i]];
                    const double E = g.has_pointE ? (double)P.E[i] : mt.E;
                    const double nu = mt.nu;
                    int type = mt.type;
                    if (type == 1 && !g.neohookean) type = 3;
                    if (type == 2 || type == 5) type = 3;
                    if (type == 0) {
                        real ds[6];
                        elastic_inc<real>((real)E, (real)nu, de, ds);
#pragma unroll
                        for (int q = 0; q < 6; ++q) s[q] += ds[q];
                    } else if (type == 1) {
                        double Fd[9], sd[6], ds[6];
                        for (int q = 0; q < 9; ++q) Fd[q] = (double)F[q];
                        for (int q = 0; q < 6; ++q) sd[q] = (double)s[q];
                        neohookean_inc(E, nu, sd, Fd, ds);
                        for (int q = 0; q < 6; ++q) s[q] = (real)(sd[q] + ds[q]);
                    } else {
                        double sd[6], ded[6], ds[6], dep[6], da = 0.0;
                        for (int q = 0; q < 6; ++q) { sd[q] = (double)s[q]; ded[q] = (double)de[q]; }
                        const double k0 = mt.sy / sqrt(3.0);
                        if (type == 4) {
                            const double al = (double)P.alpha[i];
                            const double R = al / sqrt(2.0);
                            if (radial_return(E, nu, k0 + R, sd, ded, ds, dep)) {
                                double eff = 0.0;
                                for (int q = 0; q < 6; ++q) eff += dep[q] * dep[q];
                                eff *= 2.0 / 3.0;
                                eff = sqrt(eff);
                                if ((sqrt(2.0 / 3.0) * mt.C1 - al) > 0.0)
                                    da = mt.C0 * (sqrt(2.0 / 3.0) * mt.C1 - al) * eff;
                            }
                            P.alpha[i] = (real)(al + da);
                        } else {
                            radial_return(E, nu, k0, sd, ded, ds, dep);
                        }
                        for (int q = 0; q < 6; ++q) {
                            s[q] = (real)(sd[q] + ds[q]);
                            P.ep[q][i] += (real)dep[q];
                        }
                    }
                    // write back
#pragma unroll
                    for (int q = 0; q < 3; ++q) { P.x[q][i] = x[q]; P.v[q][i] = v[q]; }
#pragma unroll
                    for (int q = 0; q < 9; ++q) P.F[q][i] = F[q];
#pragma unroll
                    for (int q = 0; q < 6; ++q) P.s[q][i] = s[q];
                }
            }
            // ---- FP64 checks of the new state (PhysicsEngine.cpp:73-79)
            {
                double vv = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
                double lv = sqrt(vv);
                red[0] += (double)m * (lv * lv);
                red[1] += (double)m * v[0];
                red[2] += (double)m * v[1];
                red[3] += (double)m * v[2];
                bool fin = isfinite((double)x[0] + x[1] + x[2]) && isfinite(vv) &&
                           isfinite((double)s[0] + s[1] + s[2] + s[3] + s[4] + s[5]);
                if (!fin) { red[4] += 1.0; flags |= FLAG_NAN; }
            }
            // ---- P2G contributions of step n+1 (…_MP.cpp:106-185)
            double Fd[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) Fd[q] = (double)F[q];
            const real V = (real)(det3(Fd) * (double)V0);
            Bases<real> b;
            make_bases(x, g, O, b);
            int lc[3];
            bool ok = true;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                lc[q] = b.cell[q] - (O[q] - 1);
                if (b.cell[q] < 0 || b.cell[q] >= g.cells[q]) { flags |= FLAG_OUT_OF_GRID; ok = false; }
                if (lc[q] < 0 || lc[q] > g.S[q] + 1) { flags |= FLAG_ARENA; ok = false; }
                const real uu = b.u[q] - (real)O[q];
                if (uu < real(-0.5) || uu >= (real)g.S[q] + real(0.5)) flags |= FLAG_REBIN;
            }
            if (ok) {
                key = lc[0] + A0 * lc[1] + A01 * lc[2];
                real fe[3] = {real(0), real(0), real(0)};
                if (g.has_fext)
#pragma unroll
                    for (int q = 0; q < 3; ++q) fe[q] = P.fext[q][i];
                const real mv[3] = {m * v[0], m * v[1], m * v[2]};
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    real N, G[3];
                    corner(b, k, N, G);
                    real *ck = &c[kQ * k];
                    ck[0] = N * m;
                    ck[1] = N * mv[0];
                    ck[2] = N * mv[1];
                    ck[3] = N * mv[2];
                    ck[4] = -V * (G[0] * s[0] + G[1] * s[3] + G[2] * s[5]) + N * fe[0];
                    ck[5] = -V * (G[1] * s[1] + G[0] * s[3] + G[2] * s[4]) + N * fe[1];
                    ck[6] = -V * (G[2] * s[2] + G[0] * s[5] + G[1] * s[4]) + N * fe[2];
                }
            }
        }

        // ---- per-chunk fixed-point exponent (max |contribution| per field)
        {
            double mx[kQ];
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                real t = real(0);
#pragma unroll
                for (int k = 0; k < 8; ++k) t = fmax(t, fabs(c[kQ * k + q]));
                mx[q] = wave_max((double)t);
            }
            if (lane == 0)
#pragma unroll
                for (int q = 0; q < kQ; ++q) s_wmax[wv][q] = mx[q];
        }
        __syncthreads();
        if (tid < kQ) {
            double m = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) m = fmax(m, s_wmax[w][tid]);
            int e = (m > 0.0) ? ilogb(m) + 1 : kUnsetExp;
            int sh = 0;
            if (e > s_e[tid]) {
                if (s_e[tid] != kUnsetExp) sh = e - s_e[tid];
                s_e[tid] = e;
            }
            s_shift[tid] = sh;
        }
        __syncthreads();
        {
            bool any = false;
#pragma unroll
            for (int q = 0; q < kQ; ++q) any |= s_shift[q] != 0;
            if (any) {
                for (int idx = tid; idx < kQ * AN; idx += kBlock) {
                    int sh = s_shift[idx / AN];
                    if (sh) s_arena[idx] = (sh >= 63) ? (s_arena[idx] >> 63) : (s_arena[idx] >> sh);
                }
                __syncthreads();
            }
        }

        // ---- run-segmented sums within each 16-lane row (lane order)
        const int kprev = dpp_i<MPM_ROW_SHR(1)>(key);
        const int knext = dpp_i<MPM_ROW_SHL(1)>(key);
        const bool head = (rl == 0) || (kprev != key);
        const bool tail = (rl == 15) || (knext != key);
        int sst = head ? rl : 0;
        sst = max(sst, dpp_i<MPM_ROW_SHR(1)>(sst));
        sst = max(sst, dpp_i<MPM_ROW_SHR(2)>(sst));
        sst = max(sst, dpp_i<MPM_ROW_SHR(4)>(sst));
        sst = max(sst, dpp_i<MPM_ROW_SHR(8)>(sst));
        const int len = rl - sst + 1;
        if (__ballot(len > 1)) seg_step<1>(c, rl, sst);
        if (__ballot(len > 2)) seg_step<2>(c, rl, sst);
        if (__ballot(len > 4)) seg_step<4>(c, rl, sst);
        if (__ballot(len > 8)) seg_step<8>(c, rl, sst);

        // ---- run totals -> LDS arena (64-bit fixed point, order independent)
        if (tail && key >= 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int node = key + ((k >> 2) & 1) + A0 * ((k >> 1) & 1) + A01 * (k & 1);
#pragma unroll
                for (int q = 0; q < kQ; ++q) {
                    const real val = c[kQ * k + q];
                    if (val != real(0)) {
                        long long iv = __double2ll_rn(ldexp((double)val, FRAC - s_e[q]));
                        atomicAdd((unsigned long long *)&s_arena[q * AN + node], (unsigned long long)iv);
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---- arena out (fixed point -> real), fixed-order block sums of checks
    real *out = a.arena_out + (size_t)slot * kQ * AN;
    for (int idx = tid; idx < kQ * AN; idx += kBlock) {
        const int e = s_e[idx / AN];
        out[idx] = (e == kUnsetExp) ? real(0) : (real)ldexp((double)s_arena[idx], e - FRAC);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k] = wave_sum(red[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 5; ++k) s_red[wv][k] = red[k];
    int fl = flags;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fl |= __shfl_xor(fl, o, 64);
    if (lane == 0 && fl) atomicOr(a.flags, fl);
    __syncthreads();
    if (tid < 5) {
        double r = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) r += s_red[w][tid];
        a.kp_part[5 * slot + tid] = r;
    }
}

// ------------------------------------------------------------------ grid
template <typename real>
struct K2Args {
    GridParams g;
    const real *arena_in;
    const int32_t *brick_slot;   // dense brick -> occupied slot or -1
    const uint32_t *k2_list;
    const int32_t *n_k2;
    int32_t k2_cap;
    NodeRec<real> *nodes;
    const uint8_t *fixed;
    const unsigned long long *marks;
    const double *dc_v;
    const StepRec *steps;
    double *k2_part;             // per block: grid mass, reaction, active nodes
};

// Grid momentum, damping, BC, reaction, platen override (…_MP.cpp:209-265).
template <typename real>
__global__ void __launch_bounds__(64) k_grid(K2Args<real> a, int step)
{
    const GridParams &g = a.g;
    const int kb = blockIdx.x;
    if (kb >= *a.n_k2) return;
    const uint32_t brick = a.k2_list[kb];
    const int b[3] = {(int)(brick % g.nb[0]), (int)((brick / g.nb[0]) % g.nb[1]), (int)(brick / (g.nb[0] * g.nb[1]))};
    const StepRec st = a.steps[step];
    const real dt = (real)st.dt;
    const real m_min = (real)g.m_min, damp = (real)g.damping;
    const int nS = g.S[0] * g.S[1] * g.S[2];
    const int AN = g.arenaN;
    double rm = 0.0, rr = 0.0, ra = 0.0;
    for (int t = threadIdx.x; t < nS; t += blockDim.x) {
        const int o[3] = {t % g.S[0], (t / g.S[0]) % g.S[1], t / (g.S[0] * g.S[1])};
        int n[3];
        bool in = true;
#pragma unroll
        for (int q = 0; q < 3; ++q) { n[q] = b[q] * g.S[q] + o[q]; in &= n[q] < g.nodes[q]; }
        if (!in) continue;
        real sum[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) sum[q] = real(0);
        for (int dz = -1; dz <= 1; ++dz) {
            const int bz = b[2] + dz;
            if (bz < 0 || bz >= g.nb[2]) continue;
            const int lz = n[2] - (bz * g.S[2] - 1);
            if (lz < 0 || lz >= g.A[2]) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int by = b[1] + dy;
                if (by < 0 || by >= g.nb[1]) continue;
                const int ly = n[1] - (by * g.S[1] - 1);
                if (ly < 0 || ly >= g.A[1]) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int bx = b[0] + dx;
                    if (bx < 0 || bx >= g.nb[0]) continue;
                    const int lx = n[0] - (bx * g.S[0] - 1);
                    if (lx < 0 || lx >= g.A[0]) continue;
                    const int sl = a.brick_slot[bx + g.nb[0] * (by + g.nb[1] * bz)];
                    if (sl < 0) continue;
                    const real *ar = a.arena_in + (size_t)sl * kQ * AN + lx + g.A[0] * (ly + g.A[1] * lz);
#pragma unroll
                    for (int q = 0; q < kQ; ++q) sum[q] += ar[(size_t)q * AN];
                }
            }
        }
        const long long nid = n[0] + (long long)g.nodes[0] * (n[1] + (long long)g.nodes[1] * n[2]);
        const real m = sum[0];
        real v[3], f[3] = {sum[4], sum[5], sum[6]};
#pragma unroll
        for (int q = 0; q < 3; ++q) v[q] = (m > m_min) ? sum[1 + q] / m : real(0);
        const real vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (sqrt(vv) > real(1.0e-9)) {
            const real lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
            const real inv = real(1) / sqrt(vv);
            const real sc = damp * lf;
#pragma unroll
            for (int q = 0; q < 3; ++q) f[q] -= sc * (v[q] * inv);
        }
        if (m > m_min)
#pragma unroll
            for (int q = 0; q < 3; ++q) v[q] += f[q] / m * dt;
        const uint8_t fx = a.fixed[nid];
        if (fx & 1) { v[0] = real(0); f[0] = real(0); }
        if (fx & 2) { v[1] = real(0); rr += (double)f[1]; f[1] = real(0); }
        if (fx & 4) { v[2] = real(0); f[2] = real(0); }
        const unsigned long long mk = a.marks[nid];
        if ((uint32_t)(mk >> 32) == st.epoch + 1u) {
            const int j = (int)(mk & 0xFFFFFFFFull);
#pragma unroll
            for (int q = 0; q < 3; ++q) { v[q] = (real)a.dc_v[3 * j + q]; f[q] = real(0); }
        }
        NodeRec<real> rec;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            rec.d[q] = v[q];
            rec.d[3 + q] = (fabs(m) > m_min) ? f[q] / m : real(0);
        }
        rec.d[6] = m;
        rec.d[7] = real(0);
        a.nodes[nid] = rec;
        rm += (double)m;
        ra += (m > real(0)) ? 1.0 : 0.0;
    }
    rm = wave_sum(rm);
    rr = wave_sum(rr);
    ra = wave_sum(ra);
    if (threadIdx.x == 0) {
        a.k2_part[3 * kb] = rm;
        a.k2_part[3 * kb + 1] = rr;
        a.k2_part[3 * kb + 2] = ra;
    }
}

}  // namespace mpm
