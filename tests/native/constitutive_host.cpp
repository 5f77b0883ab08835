// Host build of mpm_constitutive's batch packing (cr_layout, cr_views) and of
// its kernel body (cr_apply), for runs under AddressSanitizer: the staging
// buffer is a heap block of exactly the size cr_layout asks for, so a read or
// write outside any part is reported.
//
// stdin:  int32 law, int32 n, uint32 mask of the arrays given (bit k = input
//         k of E, nu, yield, alpha, C0, C1, stress, strain_inc, F), then the
//         given arrays in that order (n, n, n, n, n, n, 6n, 6n, 9n doubles).
// stdout: int32 status (0, or -1 when cr_layout refuses the batch), then the
//         law's outputs: stress_inc 6n, plastic_inc 6n, alpha_inc n; or j2 n.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../code_mpm_auxetics_amd/csrc/mpm_laws.hpp"

using namespace mpm;

static bool rd(void *p, size_t bytes) { return std::fread(p, 1, bytes, stdin) == bytes; }

int main()
{
    int32_t law = 0, n = 0;
    uint32_t mask = 0;
    if (!rd(&law, 4) || !rd(&n, 4) || !rd(&mask, 4) || n < 0) return 2;
    const size_t len[9] = {(size_t)n, (size_t)n, (size_t)n, (size_t)n, (size_t)n, (size_t)n,
                           6 * (size_t)n, 6 * (size_t)n, 9 * (size_t)n};
    std::vector<std::vector<double>> in(9);
    for (int k = 0; k < 9; ++k)
        if (mask & (1u << k)) {
            in[k].resize(len[k]);
            if (!rd(in[k].data(), len[k] * sizeof(double))) return 2;
        }
    const bool j2 = law == MPM_LAW_J2_STATE;
    std::vector<double> ds(6 * (size_t)n), dep(6 * (size_t)n), da(n), jj(n);
    mpm_cr_batch b{};
    b.n = n;
    const double **ip[9] = {&b.E, &b.nu, &b.yield, &b.alpha, &b.C0, &b.C1, &b.stress, &b.strain_inc, &b.F};
    for (int k = 0; k < 9; ++k) *ip[k] = (mask & (1u << k)) ? in[k].data() : nullptr;
    if (j2) b.j2 = jj.data();
    else { b.stress_inc = ds.data(); b.plastic_inc = dep.data(); b.alpha_inc = da.data(); }
    if (n == 0) {   // mpm_constitutive returns before the layout
        const int32_t ok = 0;
        std::fwrite(&ok, 4, 1, stdout);
        return 0;
    }
    CrPart parts[kCrParts];
    const char *why = nullptr;
    const long long total = cr_layout(law, &b, parts, &why);
    int32_t st = total < 0 ? -1 : 0;
    std::fwrite(&st, 4, 1, stdout);
    if (total < 0) return 0;
    // the "device" buffer: exactly total doubles (at least one element so
    // that new[] returns a real block)
    double *buf = new double[total > 0 ? total : 1];
    for (int k = 0; k < kCrParts; ++k)
        if (parts[k].src) std::memcpy(buf + parts[k].off, parts[k].src, parts[k].len * sizeof(double));
    const CrBatch cb = cr_views(n, buf, parts);
    for (int i = 0; i < n; ++i) cr_apply(law, cb, i);
    for (int k = 0; k < kCrParts; ++k)
        if (parts[k].dst) std::memcpy(parts[k].dst, buf + parts[k].off, parts[k].len * sizeof(double));
    delete[] buf;
    if (j2) std::fwrite(jj.data(), sizeof(double), jj.size(), stdout);
    else {
        std::fwrite(ds.data(), sizeof(double), ds.size(), stdout);
        std::fwrite(dep.data(), sizeof(double), dep.size(), stdout);
        std::fwrite(da.data(), sizeof(double), da.size(), stdout);
    }
    return 0;
}
