// mpm_bin_kernels.hpp -- deterministic binning of points into bricks.
//
// A rebin runs after upload and whenever a point drifts more than half a cell
// outside its brick.  Keys are (dense brick id << cbits) | cell-in-brick; a
// stable LSD radix sort (8-bit digits) orders them, ties kept in input order,
// so the permutation equals a CPU stable sort of the same keys bit for bit.
// Stability inside a block comes from wave-level match masks (ballot over the
// distinct digits of a wave) and a fixed wave order; no atomics decide order.
#pragma once
#include "mpm_device.hpp"

namespace mpm {

constexpr int kSortItems = 8;                    // keys per thread per tile
constexpr int kSortTile = kBlock * kSortItems;   // 2048 keys per block

// ---------------------------------------------------------------- scans
// exclusive scan of n u32, tile of 1024 per block (256 threads x 4)
__global__ void __launch_bounds__(256) k_scan_tiles(const uint32_t *in, uint32_t *out,
                                                   uint32_t *tile_sums, uint32_t n)
{
    __shared__ uint32_t s[256];
    const uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    uint32_t v[4], t = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) { v[k] = (base + k < n) ? in[base + k] : 0u; t += v[k]; }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t x = (threadIdx.x >= (unsigned)o) ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - t;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) tile_sums[blockIdx.x] = s[255];
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t *out, const uint32_t *tile_off, uint32_t n)
{
    const uint32_t base = blockIdx.x * 1024u + threadIdx.x * 4u;
    const uint32_t add = tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

// ----------------------------------------------------------- radix sort
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint32_t *keys, uint32_t n, int shift,
                                                      uint32_t *hist, uint32_t nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];   // digit-major
}

__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint32_t *keys, const uint32_t *vals,
                                                         uint32_t *keys_out, uint32_t *vals_out,
                                                         uint32_t n, int shift,
                                                         const uint32_t *offsets, uint32_t nblocks)
{
    __shared__ uint32_t run_off[256];
    __shared__ uint32_t wcnt[kBlock / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    run_off[tid] = offsets[tid * nblocks + blockIdx.x];
    for (int w = 0; w < kBlock / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * (uint32_t)kSortTile;
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t i = base + r * kBlock + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const uint32_t val = valid ? vals[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        // match mask of lanes with the same digit
        unsigned long long rem = __ballot(valid), mine = 0ull;
        while (rem) {
            const int leader = __ffsll((long long)rem) - 1;
            const uint32_t ld = __shfl(d, leader, 64);
            const unsigned long long m = __ballot(valid && d == ld);
            if (valid && d == ld) mine = m;
            rem &= ~m;
        }
        const uint32_t rank = __popcll(mine & lt);
        if (valid && rank == 0) wcnt[wv][d] = __popcll(mine);
        __syncthreads();
        if (valid) {
            uint32_t pos = run_off[d] + rank;
            for (int w = 0; w < wv; ++w) pos += wcnt[w][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        {
            uint32_t tot = 0;
            for (int w = 0; w < kBlock / 64; ++w) { tot += wcnt[w][tid]; wcnt[w][tid] = 0; }
            run_off[tid] += tot;
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- rebin
// Key of every deformable point from its current position.  FP64 state:
// cell = trunc(x/h) (GridPoint_Mediator.h:35).  FP32 state: the brick origin
// of the point plus the integer part of its fixed-point position; from_abs
// uses the FP64 staging positions.
template <typename real>
__global__ void __launch_bounds__(kBlock) k_rebin_keys(GridParams g, PArrays<real> p,
                                                      const double *xabs0, const double *xabs1,
                                                      const double *xabs2, uint32_t n,
                                                      uint32_t *keys, uint32_t *vals, int32_t *flags)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int cell[3];
    bool bad = false;
    const double *xa[3] = {xabs0, xabs1, xabs2};
    for (int a = 0; a < 3; ++a) {
        int c;
        if (xabs0) {
            const double x = xa[a][i];
            c = (int)(x / g.h[a]);
            if (!(x >= 0.0)) bad = true;
        } else if (sizeof(real) == 8) {
            const double x = (double)p.x[a][i];
            c = (int)(x / g.h[a]);
            if (!(x >= 0.0)) bad = true;
        } else {
            const uint32_t pb = p.pbrick[i];
            const int bc = (a == 0) ? (int)(pb % g.nb[0]) : (a == 1) ? (int)((pb / g.nb[0]) % g.nb[1])
                                                                   : (int)(pb / (g.nb[0] * g.nb[1]));
            c = bc * g.S[a] + (p.xq[a][i] >> kQFrac);   // exact floor of the fixed-point cell
        }
        if (c < 0 || c >= g.cells[a]) { bad = true; c = c < 0 ? 0 : g.cells[a] - 1; }
        cell[a] = c;
    }
    if (bad) atomicOr(flags, FLAG_OUT_OF_GRID);
    const uint32_t b = (uint32_t)(cell[0] / g.S[0]) +
                       (uint32_t)g.nb[0] * ((uint32_t)(cell[1] / g.S[1]) + (uint32_t)g.nb[1] * (uint32_t)(cell[2] / g.S[2]));
    const uint32_t lc = (uint32_t)(cell[0] % g.S[0]) +
                        (uint32_t)g.S[0] * ((uint32_t)(cell[1] % g.S[1]) + (uint32_t)g.S[1] * (uint32_t)(cell[2] % g.S[2]));
    keys[i] = (b << g.cbits) | lc;
    vals[i] = i;
}

struct GatherDesc {
    const void *src;
    void *dst;
    int esize;
};

// gather every state array through the permutation (one array per blockIdx.y)
__global__ void __launch_bounds__(kBlock) k_gather(const GatherDesc *d, const uint32_t *perm, uint32_t n)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const GatherDesc gd = d[blockIdx.y];
    const uint32_t i = perm[j];
    if (gd.esize == 8) ((uint64_t *)gd.dst)[j] = ((const uint64_t *)gd.src)[i];
    else if (gd.esize == 4) ((uint32_t *)gd.dst)[j] = ((const uint32_t *)gd.src)[i];
    else ((uint8_t *)gd.dst)[j] = ((const uint8_t *)gd.src)[i];
}

// positions into the new order; FP32 re-bases to the new brick origin
template <typename real>
__global__ void __launch_bounds__(kBlock) k_rebin_pos(GridParams g, PArrays<real> src, PArrays<real> dst,
                                                     const double *xabs0, const double *xabs1,
                                                     const double *xabs2, const uint32_t *perm,
                                                     const uint32_t *skeys, uint32_t n)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = perm[j];
    const uint32_t nbk = skeys[j] >> g.cbits;
    dst.pbrick[j] = nbk;
    const int nbc[3] = {(int)(nbk % g.nb[0]), (int)((nbk / g.nb[0]) % g.nb[1]), (int)(nbk / (g.nb[0] * g.nb[1]))};
    const double *xa[3] = {xabs0, xabs1, xabs2};
    for (int a = 0; a < 3; ++a) {
        if (sizeof(real) == 8) {
            dst.x[a][j] = xabs0 ? (real)xa[a][i] : src.x[a][i];
        } else if (xabs0) {
            const double u = xa[a][i] / g.h[a] - (double)(nbc[a] * g.S[a]);
            dst.xq[a][j] = (int32_t)rint(u * kQOne);
        } else {
            const uint32_t pb = src.pbrick[i];
            const int bc = (a == 0) ? (int)(pb % g.nb[0]) : (a == 1) ? (int)((pb / g.nb[0]) % g.nb[1])
                                                                   : (int)(pb / (g.nb[0] * g.nb[1]));
            // exact re-basing: shift by whole cells
            dst.xq[a][j] = src.xq[a][i] + (int32_t)((unsigned)((bc - nbc[a]) * g.S[a]) << kQFrac);
        }
    }
}

__global__ void k_heads(const uint32_t *skeys, uint32_t n, int cbits, uint32_t *head)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    head[j] = (j == 0 || (skeys[j] >> cbits) != (skeys[j - 1] >> cbits)) ? 1u : 0u;
}

__global__ void k_occ(const uint32_t *skeys, const uint32_t *head, const uint32_t *pos, uint32_t n,
                      int cbits, uint32_t *occ_list, uint32_t *occ_start)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || !head[j]) return;
    occ_list[pos[j]] = skeys[j] >> cbits;
    occ_start[pos[j]] = j;
}

__global__ void k_occ_count(const uint32_t *occ_start, uint32_t n_occ, uint32_t n, uint32_t *occ_count)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_occ) return;
    occ_count[s] = ((s + 1 < n_occ) ? occ_start[s + 1] : n) - occ_start[s];
}

// Spread the points of each brick over the lanes of a wave.  Sorted position
// t = q*P + r of a brick holding n points moves to sum_{r'<r} n_r' + q, where
// n_r' = #{t < n : t mod P = r'}: neighbouring lanes then hold points P apart
// in the cell order, so the P2G adds of one instruction rarely share a node.
constexpr uint32_t kSpread = 8;   // >= typical points per cell
__global__ void k_spread(const uint32_t *perm, const uint32_t *head, const uint32_t *pos, const uint32_t *occ_start,
                         const uint32_t *occ_count, uint32_t n, uint32_t *perm_out)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t s = pos[j] + head[j] - 1u;   // occupied slot of position j
    const uint32_t st = occ_start[s], cnt = occ_count[s], t = j - st;
    const uint32_t r = t % kSpread;
    uint32_t p = t / kSpread;
    for (uint32_t rr = 0; rr < r; ++rr) p += (cnt > rr) ? (cnt - rr - 1u) / kSpread + 1u : 0u;
    perm_out[st + p] = perm[j];
}



// initial platen override marks for the first step after upload
__global__ void k_dc_marks(GridParams g, const double *dc_x, int n_dc, uint32_t epoch,
                           unsigned long long *marks)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_dc) return;
    int i0[3];
    for (int q = 0; q < 3; ++q) i0[q] = (int)(dc_x[3 * j + q] / g.h[q]);
    const unsigned long long tag = ((unsigned long long)(epoch + 1u) << 32) | (unsigned)j;
    for (int k = 0; k < 8; ++k) {
        const int n0 = i0[0] + ((k >> 2) & 1), n1 = i0[1] + ((k >> 1) & 1), n2 = i0[2] + (k & 1);
        if (n0 < 0 || n1 < 0 || n2 < 0 || n0 >= g.nodes[0] || n1 >= g.nodes[1] || n2 >= g.nodes[2]) continue;
        atomicMax(&marks[n0 + (long long)g.nodes[0] * (n1 + (long long)g.nodes[1] * n2)], tag);
    }
}

// position of the point with original index id (contact path: point 0)
__global__ void k_find_id(const uint32_t *ids, uint32_t n, uint32_t id, uint32_t *out)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && ids[j] == id) out[1] = j;
}

// ------------------------------------------------------------ readback
struct ReadbackOut {
    double *x, *v, *F, *s, *eps, *ep, *alpha, *vol;
    double *j2;         // 2 per point: J2 of the stress, J2 of the plastic strain
    int32_t *cell;      // 3 per point: the binning cell
    uint32_t *order;    // 1 per point: position in the binned order
};

template <typename real>
__global__ void __launch_bounds__(kBlock) k_readback(GridParams g, PArrays<real> p, uint32_t n,
                                                    ReadbackOut o, uint32_t mask)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t id = p.id[j];
    if (mask & 1u) {
        for (int a = 0; a < 3; ++a) {
            double xa;
            if (sizeof(real) == 4) {
                const uint32_t pb = p.pbrick[j];
                const int bc = (a == 0) ? (int)(pb % g.nb[0]) : (a == 1) ? (int)((pb / g.nb[0]) % g.nb[1])
                                                                       : (int)(pb / (g.nb[0] * g.nb[1]));
                xa = ((double)(bc * g.S[a]) + (double)p.xq[a][j] / kQOne) * g.h[a];
            } else {
                xa = (double)p.x[a][j];
            }
            o.x[3 * (size_t)id + a] = xa;
        }
    }
    if (mask & 2u) for (int a = 0; a < 3; ++a) o.v[3 * (size_t)id + a] = (double)p.v[a][j];
    double Fd[9];
    for (int a = 0; a < 9; ++a) Fd[a] = fd(p.F[a][j], a);
    if (mask & 4u) for (int a = 0; a < 9; ++a) o.F[9 * (size_t)id + a] = Fd[a];
    if (mask & 8u) for (int a = 0; a < 6; ++a) o.s[6 * (size_t)id + a] = (double)p.s[a][j];
    if ((mask & 16u) && g.track_strain) for (int a = 0; a < 6; ++a) o.eps[6 * (size_t)id + a] = (double)p.eps[a][j];
    if (mask & 32u) for (int a = 0; a < 6; ++a) o.ep[6 * (size_t)id + a] = (double)p.ep[a][j];
    if (mask & 64u) o.alpha[id] = (double)p.alpha[j];
    if (mask & 128u) o.vol[id] = det3(Fd) * (double)p.vol0[j];
    if (mask & 512u) {   // calculateState_J2 (…calculateState_J2.cpp:4-26) of stress and plastic strain
        double sd[6], ed[6];
        for (int a = 0; a < 6; ++a) {
            sd[a] = (double)p.s[a][j];
            ed[a] = (double)p.ep[a][j];
        }
        o.j2[2 * (size_t)id] = j2_of(sd);
        o.j2[2 * (size_t)id + 1] = j2_of(ed);
    }
    if (mask & 1024u) {  // the cell of k_rebin_keys, from the same state
        for (int a = 0; a < 3; ++a) {
            int c;
            if (sizeof(real) == 4) {
                const uint32_t pb = p.pbrick[j];
                const int bc = (a == 0) ? (int)(pb % g.nb[0]) : (a == 1) ? (int)((pb / g.nb[0]) % g.nb[1])
                                                                       : (int)(pb / (g.nb[0] * g.nb[1]));
                c = bc * g.S[a] + (p.xq[a][j] >> kQFrac);
            } else {
                c = (int)((double)p.x[a][j] / g.h[a]);
            }
            o.cell[3 * (size_t)id + a] = c;
        }
        o.order[id] = j;
    }
}

// Reaction of each step of a batch (PhysicsEngine.cpp:65-71): block k sums the
// per-slot grid partials of step k -- strided lane sums, then a fixed tree.
__global__ void __launch_bounds__(256) k_step_sums(const double *gpart, int occ_cap, const int32_t *n_occ,
                                                  double *hist)
{
    __shared__ double s[256];
    const int t = threadIdx.x, no = *n_occ;
    const double *gp = gpart + (size_t)blockIdx.x * occ_cap * 3;
    double r = 0.0;
    for (int i = t; i < no; i += 256) r += gp[3 * i + 1];
    s[t] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) hist[blockIdx.x] = s[0];
}

// The step records of a batch, read from the host's pinned record set
// (mapped host memory) into device memory; one block.
__global__ void __launch_bounds__(64) k_load_steps(const uint32_t *src, uint32_t *dst, int nwords)
{
    for (int i = threadIdx.x; i < nwords; i += 64) dst[i] = src[i];
}

// The per-step reactions of a batch and the last step's grid totals in one
// launch: blocks 0..nsteps-1 are k_step_sums, block nsteps is k_sum3 of the
// last step (the same fixed trees, so the same bits).  The reactions and the
// flags word go straight to the host's pinned result set `out` (mapped host
// memory): out[0] flags, out[1 + k] the reaction of step k.
__global__ void __launch_bounds__(256) k_batch_sums(const double *gpart, int occ_cap, const int32_t *n_occ,
                                                   int nsteps, double *out, double *grid3, const int32_t *flags)
{
    double *const hist = out + 1;
    __shared__ double s[3][256];
    const int t = threadIdx.x, no = *n_occ;
    if ((int)blockIdx.x < nsteps) {
        const double *gp = gpart + (size_t)blockIdx.x * occ_cap * 3;
        double r = 0.0;
        for (int i = t; i < no; i += 256) r += gp[3 * i + 1];
        s[0][t] = r;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w) s[0][t] += s[0][t + w];
            __syncthreads();
        }
        if (t == 0) hist[blockIdx.x] = s[0][0];
        return;
    }
    const double *part = gpart + (size_t)(nsteps - 1) * occ_cap * 3;
    double r[3] = {0, 0, 0};
    for (int i = t; i < no; i += 256)
        for (int k = 0; k < 3; ++k) r[k] += part[3 * i + k];
    for (int k = 0; k < 3; ++k) s[k][t] = r[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 3; ++k) s[k][t] += s[k][t + w];
        __syncthreads();
    }
    if (t < 3) grid3[t] = s[t][0];
    if (t == 0) {
        int32_t *o = reinterpret_cast<int32_t *>(out);
        o[0] = *flags;
        o[1] = 0;
    }
}

// Per-group reaction of a batch: block (step k, group g) sums the group's
// y-fixed slots [off[g], off[g+1]) of step k in a fixed tree.
__global__ void __launch_bounds__(256) k_group_sums(const double *fy_val, int n_fy, const int32_t *off, int n_groups,
                                                   double *ghist)
{
    __shared__ double s[256];
    const int k = blockIdx.x / n_groups, gr = blockIdx.x % n_groups, t = threadIdx.x;
    const double *v = fy_val + (size_t)k * n_fy;
    double r = 0.0;
    for (int i = off[gr] + t; i < off[gr + 1]; i += 256) r += v[i];
    s[t] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] += s[t + w];
        __syncthreads();
    }
    if (t == 0) ghist[(size_t)k * n_groups + gr] = s[0];
}

// fill a u32 array (graph-safe replacement of a memset)
__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// fill a double array (graph-safe replacement of a memset)
__global__ void k_fill_f64(double *p, size_t n, double v)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sum of 3-double partials (grid mass, reaction, active nodes of one step):
// one block of 256, strided lane sums then a fixed tree.  Count: *n_d or n.
__global__ void __launch_bounds__(256) k_sum3(const double *part, const int32_t *n_d, int n, double *out3)
{
    __shared__ double s[3][256];
    const int t = threadIdx.x, cnt = n_d ? *n_d : n;
    double r[3] = {0, 0, 0};
    for (int i = t; i < cnt; i += 256)
        for (int k = 0; k < 3; ++k) r[k] += part[3 * i + k];
    for (int k = 0; k < 3; ++k) s[k][t] = r[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 3; ++k) s[k][t] += s[k][t + w];
        __syncthreads();
    }
    if (t < 3) out3[t] = s[t][0];
}

// FP64 check sums (mpm_checks): point partials (count mul * *n_pp_d, or
// n_pp) in a fixed tree, the last grid update's totals (grid3: mass, reaction,
// active nodes) and the platen points.
__global__ void __launch_bounds__(256) k_finalize(const double *kp_part, const int32_t *n_pp_d, int n_pp,
                                                 const double *grid3, const double *dc_v, const double *dc_m,
                                                 int n_dc, double *out, int mul = 1)
{
    __shared__ double s[5][256];
    const int t = threadIdx.x;
    double r[5] = {0, 0, 0, 0, 0};
    const int npp = n_pp_d ? mul * *n_pp_d : n_pp;
    for (int i = t; i < npp; i += 256)
        for (int k = 0; k < 5; ++k) r[k] += kp_part[5 * i + k];
    for (int j = t; j < n_dc; j += 256) {
        const double *v = &dc_v[3 * j];
        const double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        r[0] += dc_m[j] * vv;
        r[1] += dc_m[j] * v[0];
        r[2] += dc_m[j] * v[1];
        r[3] += dc_m[j] * v[2];
    }
    for (int k = 0; k < 5; ++k) s[k][t] = r[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 5; ++k) s[k][t] += s[k][t + w];
        __syncthreads();
    }
    if (t < 5) out[2 + t] = s[t][0];
    if (t == 0) {
        out[0] = grid3[0];
        out[1] = grid3[1];
        out[7] = grid3[2];
    }
}

}  // namespace mpm
