// ConstitutiveRelation.h -- C++ host mirror of the reference's
// ConstitutiveRelation (ConstitutiveRelation.h:16-34) on top of
// mpm_constitutive in include/mpm_auxetics.h.
//
// Method names, arguments and output members are the reference's; each call
// evaluates the step's own device function for one input (FP64, the
// reference's evaluation order) through mpm_constitutive_noalloc, so the
// per-point use of the reference allocates no device memory.  Batches go
// through mpm_constitutive directly.  glm is not used: the deformation gradient is a column-major
// std::array<double, 9> (m[col][row] at [3*col+row], as glm stores it).
// The ViscoElastic and IdealGas laws are never dispatched by the reference
// step (…_MP.cpp:332-339): their methods keep the signatures and throw.
// Errors: a status other than MPM_OK throws std::runtime_error.
#pragma once

#include <array>
#include <stdexcept>
#include <string>

#include "../../include/mpm_auxetics.h"

namespace mpm_host {

class ConstitutiveRelation {
  public:
    explicit ConstitutiveRelation(int device = 0) : device_(device) {}

    double d6StressIncrement[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double d6PlasticStrainIncrement[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double dBackstress_IsotropicIncrement = 0.0;
    double d_J2 = 0.0;

    void calculateIncrement_Elastic(double dE, double dNu, const double d6StrainIncrement[6])
    {
        mpm_cr_batch b = batch(&dE, &dNu);
        b.strain_inc = d6StrainIncrement;
        run(MPM_LAW_ELASTIC, b);
    }
    void calculateIncrement_NeoHookean_6D(double dE, double dNu, const double d6Stress_Current[6],
                                          const std::array<double, 9> &d33DeformationGradient)
    {
        mpm_cr_batch b = batch(&dE, &dNu);
        b.stress = d6Stress_Current;
        b.F = d33DeformationGradient.data();
        run(MPM_LAW_NEOHOOKEAN, b);
    }
    void calculateIncrement_PerfectlyPlastic_6D(double dE, double dNu, double dYield, const double d6StressCurrent[6],
                                                const double d6StrainIncrement[6])
    {
        mpm_cr_batch b = batch(&dE, &dNu);
        b.yield = &dYield;
        b.stress = d6StressCurrent;
        b.strain_inc = d6StrainIncrement;
        run(MPM_LAW_PLASTIC, b);
    }
    void calculateIncrement_VonMisesHardening_6D(double dE, double dNu, double dYield, double dBackstress_Isotropic,
                                                 double dHardening_Isotropic_C0, double dHardening_Isotropic_C1,
                                                 const double d6StressCurrent[6], const double d6StrainIncrement[6])
    {
        mpm_cr_batch b = batch(&dE, &dNu);
        b.yield = &dYield;
        b.alpha = &dBackstress_Isotropic;
        b.C0 = &dHardening_Isotropic_C0;
        b.C1 = &dHardening_Isotropic_C1;
        b.stress = d6StressCurrent;
        b.strain_inc = d6StrainIncrement;
        run(MPM_LAW_VONMISES, b);
    }
    void calculateState_J2(const double d6State[6])
    {
        mpm_cr_batch b{};
        b.n = 1;
        b.stress = d6State;
        b.j2 = &d_J2;
        check(mpm_constitutive_noalloc(MPM_LAW_J2_STATE, &b, device_));
    }
    void calculateIncrement_ViscoElastic_6D(double, double, double, const double[6], const double[6], const double[6])
    {
        throw std::runtime_error("ViscoElastic is never dispatched by the reference step");
    }
    void calculateIncrement_IdealGass(double, double, double, double, const double[6], double, double, const double[6])
    {
        throw std::runtime_error("IdealGas is never dispatched by the reference step");
    }

  private:
    int device_;

    mpm_cr_batch batch(const double *E, const double *nu)
    {
        mpm_cr_batch b{};
        b.n = 1;
        b.E = E;
        b.nu = nu;
        b.stress_inc = d6StressIncrement;
        b.plastic_inc = d6PlasticStrainIncrement;
        b.alpha_inc = &dBackstress_IsotropicIncrement;
        return b;
    }
    void run(int32_t law, const mpm_cr_batch &b) { check(mpm_constitutive_noalloc(law, &b, device_)); }
    static void check(int rc)
    {
        if (rc != MPM_OK) throw std::runtime_error("mpm_constitutive_noalloc: " + std::string(mpm_last_error(nullptr)));
    }
};

}  // namespace mpm_host
