// mpm_laws.hpp -- the FP64 constitutive laws of the step, the batch layout of
// mpm_constitutive, and the per-input body of its kernel.
//
// Everything here compiles for the device (hipcc) and for the host (any
// C++ compiler, with no HIP headers), so that the batch packing and the law
// bodies can run under host bounds checks.
//
// Reference: ConstitutiveRelation_calculateIncrement_{Elastic,
// PerfectlyPlastic,VonMisesHardening,NeoHookean}.cpp and
// ConstitutiveRelation_calculateState_J2.cpp (ConstitutiveRelation.h:16-34).
#pragma once
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/mpm_auxetics.h"

#if defined(__HIPCC__)
#define MPM_HD __host__ __device__ __forceinline__
#else
#define MPM_HD inline
#endif

namespace mpm {

// glm determinant of a column-major 3x3 (…_MP.cpp:301), always in FP64.
MPM_HD double det3(const double m[9])
{
    return +m[0] * (m[4] * m[8] - m[7] * m[5])
           - m[3] * (m[1] * m[8] - m[7] * m[2])
           + m[6] * (m[1] * m[5] - m[4] * m[2]);
}

// ConstitutiveRelation_calculateIncrement_Elastic.cpp:4-15
template <typename T>
MPM_HD void elastic_inc(T E, T nu, const T de[6], T ds[6])
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
MPM_HD double j2_of(const double s[6])
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
MPM_HD bool radial_return(double E, double nu, double radius, const double s[6], const double de[6], double ds[6],
                          double dep[6])
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

// …PerfectlyPlastic.cpp:4-72: k = sigma_y / sqrt(3).  True on the plastic branch.
MPM_HD bool perfectly_plastic_inc(double E, double nu, double sy, const double s[6], const double de[6], double ds[6],
                                  double dep[6])
{
    return radial_return(E, nu, sy / sqrt(3.0), s, de, ds, dep);
}

// …VonMisesHardening.cpp:4-86: radius k + alpha / sqrt(2); on the plastic
// branch the equivalent plastic strain sqrt(2/3 sum dep_i^2) (engineering
// shear, no factor 2) drives d alpha = C0 (sqrt(2/3) C1 - alpha) eps_bar when
// positive.  True on the plastic branch.
MPM_HD bool vonmises_inc(double E, double nu, double sy, double al, double C0, double C1, const double s[6],
                         const double de[6], double ds[6], double dep[6], double &da)
{
    da = 0.0;
    const double k0 = sy / sqrt(3.0);
    if (!radial_return(E, nu, k0 + al / sqrt(2.0), s, de, ds, dep)) return false;
    double eff = 0.0;
    for (int q = 0; q < 6; ++q) eff += dep[q] * dep[q];
    eff *= 2.0 / 3.0;
    eff = sqrt(eff);
    if ((sqrt(2.0 / 3.0) * C1 - al) > 0.0) da = C0 * (sqrt(2.0 / 3.0) * C1 - al) * eff;
    return true;
}

// ConstitutiveRelation_calculateIncrement_NeoHookean.cpp:4-22, FP64.
MPM_HD void neohookean_inc(double E, double nu, const double s[6], const double F[9], double ds[6])
{
    double mu = 0.5 * E / (1.0 + nu);
    double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double J = det3(F);
    double a = lam * log(J) / J;
    double b = mu / J;
    double P[9];
    for (int c = 0; c < 3; ++c)
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

// ---------------------------------------------------- batched constitutive
// The laws applied to n independent inputs (ConstitutiveRelation.h:16-34):
// law 0 Elastic, 1 NeoHookean, 3 PerfectlyPlastic, 4 VonMisesHardening (the
// MaterialPoint.h:8-16 type codes), 16 calculateState_J2 of the stress.
// FP64, the reference's evaluation order.  Pointers a law does not use are
// null.
struct CrBatch {
    int32_t n;
    const double *E, *nu, *sy, *alpha, *C0, *C1;
    const double *s, *de, *F;    // 6n, 6n, 9n (column-major)
    double *ds, *dep, *da, *j2;  // 6n, 6n, n, n
};

// True when law is one mpm_constitutive evaluates.
MPM_HD bool cr_law_known(int law)
{
    return law == MPM_LAW_ELASTIC || law == MPM_LAW_NEOHOOKEAN || law == MPM_LAW_PLASTIC ||
           law == MPM_LAW_VONMISES || law == MPM_LAW_J2_STATE;
}

// Input i of the batch: reads only the arrays its law uses, writes only its
// outputs (ds, dep, da for the increments; j2 for the J2 state).
MPM_HD void cr_apply(int law, const CrBatch &b, int i)
{
    double s[6], de[6], ds[6] = {0, 0, 0, 0, 0, 0}, dep[6] = {0, 0, 0, 0, 0, 0}, da = 0.0;
    for (int q = 0; q < 6; ++q) {
        s[q] = b.s ? b.s[6 * (size_t)i + q] : 0.0;
        de[q] = b.de ? b.de[6 * (size_t)i + q] : 0.0;
    }
    if (law == MPM_LAW_J2_STATE) {
        b.j2[i] = j2_of(s);
        return;
    }
    if (law == MPM_LAW_ELASTIC) {
        elastic_inc<double>(b.E[i], b.nu[i], de, ds);
    } else if (law == MPM_LAW_NEOHOOKEAN) {
        double F[9];
        for (int q = 0; q < 9; ++q) F[q] = b.F[9 * (size_t)i + q];
        neohookean_inc(b.E[i], b.nu[i], s, F, ds);
    } else if (law == MPM_LAW_PLASTIC) {
        perfectly_plastic_inc(b.E[i], b.nu[i], b.sy[i], s, de, ds, dep);
    } else if (law == MPM_LAW_VONMISES) {
        vonmises_inc(b.E[i], b.nu[i], b.sy[i], b.alpha[i], b.C0[i], b.C1[i], s, de, ds, dep, da);
    }
    for (int q = 0; q < 6; ++q) {
        b.ds[6 * (size_t)i + q] = ds[q];
        b.dep[6 * (size_t)i + q] = dep[q];
    }
    b.da[i] = da;
}

// Layout of one mpm_constitutive call in a single staging buffer: the
// inputs the law reads, then the outputs it writes, each at an offset in
// doubles.  A part the law does not use has len 0 and is left null.
struct CrPart {
    const double *src;   // host input, or null
    double *dst;         // host output, or null
    size_t len;          // doubles
    size_t off;          // offset in the staging buffer (doubles)
};
enum { kCrParts = 13 };

// Checks the batch for law (every array the law reads or writes present) and
// fills the parts in the order E, nu, sy, alpha, C0, C1, stress, strain_inc,
// F, stress_inc, plastic_inc, alpha_inc, j2.  Returns the buffer size in
// doubles, or -1 with *why set when an array is missing.
inline long long cr_layout(int law, const mpm_cr_batch *b, CrPart parts[kCrParts], const char **why)
{
    const size_t n = (size_t)b->n;
    const bool j2 = law == MPM_LAW_J2_STATE, el = law == MPM_LAW_ELASTIC, nh = law == MPM_LAW_NEOHOOKEAN;
    const bool pl = law == MPM_LAW_PLASTIC || law == MPM_LAW_VONMISES, vm = law == MPM_LAW_VONMISES;
    // what the law reads (the kernel body reads exactly these)
    const double *in[9] = {j2 ? nullptr : b->E, j2 ? nullptr : b->nu, pl ? b->yield : nullptr,
                           vm ? b->alpha : nullptr, vm ? b->C0 : nullptr, vm ? b->C1 : nullptr,
                           b->stress, (j2 || nh) ? nullptr : b->strain_inc, nh ? b->F : nullptr};
    const size_t inlen[9] = {n, n, n, n, n, n, 6 * n, 6 * n, 9 * n};
    const bool need[9] = {!j2, !j2, pl, vm, vm, vm, !el, !(j2 || nh), nh};
    for (int k = 0; k < 9; ++k)
        if (need[k] && !in[k]) { *why = "constitutive batch: missing input array"; return -1; }
    double *out[4] = {j2 ? nullptr : b->stress_inc, j2 ? nullptr : b->plastic_inc, j2 ? nullptr : b->alpha_inc,
                      j2 ? b->j2 : nullptr};
    const size_t outlen[4] = {6 * n, 6 * n, n, n};
    const bool oneed[4] = {!j2, !j2, !j2, j2};
    for (int k = 0; k < 4; ++k)
        if (oneed[k] && !out[k]) { *why = "constitutive batch: missing output array"; return -1; }
    size_t total = 0;
    for (int k = 0; k < kCrParts; ++k) {
        CrPart &p = parts[k];
        p.src = k < 9 ? in[k] : nullptr;
        p.dst = k < 9 ? nullptr : out[k - 9];
        p.len = (p.src || p.dst) ? (k < 9 ? inlen[k] : outlen[k - 9]) : 0;
        p.off = total;
        total += p.len;
    }
    return (long long)total;
}

// The kernel's view of a staging buffer laid out by cr_layout.
inline CrBatch cr_views(int32_t n, double *buf, const CrPart parts[kCrParts])
{
    double *dp[kCrParts];
    for (int k = 0; k < kCrParts; ++k) dp[k] = parts[k].len ? buf + parts[k].off : nullptr;
    CrBatch cb{};
    cb.n = n;
    cb.E = dp[0]; cb.nu = dp[1]; cb.sy = dp[2]; cb.alpha = dp[3]; cb.C0 = dp[4]; cb.C1 = dp[5];
    cb.s = dp[6]; cb.de = dp[7]; cb.F = dp[8];
    cb.ds = dp[9]; cb.dep = dp[10]; cb.da = dp[11]; cb.j2 = dp[12];
    return cb;
}

}  // namespace mpm
