// Central finite differences of the constitutive laws' adjoints used by the
// differentiable run (mpm_diff.hip), evaluated on the host: radial return
// (perfectly plastic and J2 with hardening, both yield branches) and the
// NeoHookean total stress.  Perturbations that move the yield or hardening
// decision are skipped (the adjoint holds it).  Prints the worst relative
// error of each law; the test reads them.
#include "../../code_mpm_auxetics_amd/csrc/mpm_diff.hip"

#include <cstdio>
#include <random>

int main()
{
    std::mt19937_64 g(3);
    std::uniform_real_distribution<double> U(-1, 1);
    double worst_rr = 0.0, worst_nh = 0.0;
    int plastic = 0, harden = 0;
    for (int t = 0; t < 4000; ++t) {
        const bool vm = t % 2;
        const double E = 1e9 * (1 + 0.5 * U(g)), nu = 0.3 + 0.1 * U(g), sy = 5e6 * (1 + 0.3 * U(g));
        const double C1 = (t % 4 < 2) ? 1e15 : 3e6, C0 = (C1 > 1e14) ? 50e6 / C1 : 20.0;
        const double al = vm ? 1e6 * (0.5 + 0.5 * U(g)) : 0.0;
        double s[6], de[6], ws[6];
        const double wa = U(g);
        for (int i = 0; i < 6; ++i) {
            s[i] = 3e6 * U(g);
            de[i] = 4e-3 * U(g);
            ws[i] = U(g);
        }
        RR r;
        rr_eval(E, nu, sy, C0, C1, vm, s, al, de, r);
        plastic += r.plastic;
        harden += r.harden;
        auto same = [&](const RR &o) { return o.plastic == r.plastic && o.harden == r.harden; };
        // loss = ws . sigma' + wa alpha'
        auto loss = [&](double E_, const double *s_, double al_, const double *de_, RR &o) {
            rr_eval(E_, nu, sy, C0, C1, vm, s_, al_, de_, o);
            double L = 0.0;
            for (int i = 0; i < 6; ++i) L += ws[i] * (s_[i] + o.ds[i]);
            return L + wa * ((vm && o.plastic) ? al_ + o.da : al_);
        };
        double sbo[6], abo, eb[6], gE;
        rr_adjoint(E, nu, C0, vm, de, r, ws, wa, sbo, abo, eb, gE);
        auto check = [&](double ad, double fd, double floor) {
            worst_rr = fmax(worst_rr, fabs(ad - fd) / (fabs(fd) + floor));
        };
        for (int i = 0; i < 6; ++i) {
            double p[6], m[6];
            RR a, b;
            const double hs = 1e-4 * 3e6;
            for (int k = 0; k < 6; ++k) p[k] = m[k] = s[k];
            p[i] += hs;
            m[i] -= hs;
            double fp = loss(E, p, al, de, a), fm = loss(E, m, al, de, b);
            if (same(a) && same(b)) check(sbo[i], (fp - fm) / (2 * hs), 1e-3);
            const double hd = 1e-4 * 4e-3;
            for (int k = 0; k < 6; ++k) p[k] = m[k] = de[k];
            p[i] += hd;
            m[i] -= hd;
            fp = loss(E, s, al, p, a);
            fm = loss(E, s, al, m, b);
            if (same(a) && same(b)) check(eb[i] * 1e-9, (fp - fm) / (2 * hd) * 1e-9, 1e-3);
        }
        {
            RR a, b;
            const double hE = 1e-4 * E;
            const double fp = loss(E + hE, s, al, de, a), fm = loss(E - hE, s, al, de, b);
            if (same(a) && same(b)) check(gE * 1e3, (fp - fm) / (2 * hE) * 1e3, 1e-3);
        }
        if (vm) {
            RR a, b;
            const double ha = 1e-4 * 1e6;
            const double fp = loss(E, s, al + ha, de, a), fm = loss(E, s, al - ha, de, b);
            if (same(a) && same(b)) check(abo, (fp - fm) / (2 * ha), 1e-3);
        }
    }
    for (int t = 0; t < 2000; ++t) {
        const double E = 1e9 * (1 + 0.5 * U(g)), nu = 0.3 + 0.1 * U(g);
        double F[9], ws[6];
        for (int q = 0; q < 9; ++q) F[q] = ((q % 4) == 0 ? 1.0 : 0.0) + 0.15 * U(g);
        for (int i = 0; i < 6; ++i) ws[i] = U(g);
        auto loss = [&](double E_, const double *F_) {
            double P[9];
            nh_stress(E_, nu, F_, P);
            return ws[0] * P[0] + ws[1] * P[4] + ws[2] * P[8] + ws[3] * P[1] + ws[4] * P[5] + ws[5] * P[6];
        };
        double Fb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gE;
        nh_adjoint(E, nu, F, ws, Fb, gE);
        for (int q = 0; q < 9; ++q) {
            double p[9], m[9];
            const double h = 1e-6;
            for (int k = 0; k < 9; ++k) p[k] = m[k] = F[k];
            p[q] += h;
            m[q] -= h;
            const double fd = (loss(E, p) - loss(E, m)) / (2 * h);
            worst_nh = fmax(worst_nh, fabs(Fb[q] - fd) / (fabs(fd) + 1e6));
        }
        const double hE = 1e-4 * E, fd = (loss(E + hE, F) - loss(E - hE, F)) / (2 * hE);
        worst_nh = fmax(worst_nh, fabs(gE - fd) / (fabs(fd) + 1e-3));
    }
    printf("plastic %d\nharden %d\nradial_return %.3e\nneohookean %.3e\n", plastic, harden, worst_rr, worst_nh);
    return 0;
}
