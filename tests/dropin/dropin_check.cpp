// Exercises the reference-named entry points of include/tcym_dropin.hpp the way
// a TopologicalCYMSolver driver calls them, on fields read from binary files.
// usage: dropin_check nc N U.bin E.bin out_prefix log_dir
// writes out_prefix_U.bin / out_prefix_E.bin after 4 Dynamics::Update steps
// (dTau 0.05), then 4 more steps inside a Resident scope into *_U2/_E2, and
// prints "ncsdot T00 gauss_max unitarity_max TXX TZZ LambdaXX LambdaZZ" for
// the first state.  Then cools a blocked copy into *_C (printing its extent,
// spacing, blocking number and flow clock) and calibrates the cooling method
// (printing DeltaNCsCalibration, blocking number and flow clock; the log goes
// to log_dir), saves the hot fields as the reference's text files in log_dir,
// loads them into new containers and prints their NCsDot; finally thermalizes
// a zero-field lattice into *_TU / *_TE (ThermalEnergy rows in log_dir) and
// sets quasi-particle fields into *_QU / *_QE.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tcym_dropin.hpp"

DOUBLE Qs = 1.0;
DOUBLE n0 = 1.0;

static void rd(const char* f, double* p, size_t n) {
    FILE* fp = fopen(f, "rb");
    if (!fp || fread(p, sizeof(double), n, fp) != n) { fprintf(stderr, "cannot read %s\n", f); exit(3); }
    fclose(fp);
}
static void wr(const std::string& f, const double* p, size_t n) {
    FILE* fp = fopen(f.c_str(), "wb");
    if (!fp || fwrite(p, sizeof(double), n, fp) != n) { fprintf(stderr, "cannot write %s\n", f.c_str()); exit(3); }
    fclose(fp);
}

int main(int argc, char** argv) {
    if (argc != 7) { fprintf(stderr, "usage: dropin_check nc N U.bin E.bin out_prefix log_dir\n"); return 2; }
    const int nc = atoi(argv[1]), N = atoi(argv[2]);
    tcym_dropin::Configure(nc, TCYM_FP64, 0);
    GaugeLinks U(N, N, N, 1.0, 1.0, 1.0);
    ElectricFields E(N, N, N, 1.0, 1.0, 1.0);
    const size_t nu = (size_t)3*tcym_dropin::LinkSize()*U.Volume, ne = (size_t)3*tcym_dropin::VectorSize()*E.Volume;
    rd(argv[3], U.U, nu);
    rd(argv[4], E.E, ne);
    try {
        Dynamics::dTau = 0.05;
        for (int k = 0; k < 4; k++) Dynamics::Update(&U, &E);
        const double ncs = ChernSimonsNumber::NCsDot(&U, &E);
        Observables::Bulk::Update(&U, &E);
        Observables::GaussLaw::UpdateViolation(0, N - 1, 0, N - 1, 0, N - 1, &U, &E);
        Observables::Unitarity::UpdateViolation(0, N - 1, 0, N - 1, 0, N - 1, &U);
        Observables::HardScales::Update(&U);
        printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", ncs, Observables::Bulk::T00(),
               Observables::GaussLaw::GlobalMaxViolation, Observables::Unitarity::GlobalMaxViolation,
               Observables::Bulk::TXX(), Observables::Bulk::TZZ(), Observables::HardScales::LambdaXX(),
               Observables::HardScales::LambdaZZ());
        wr(std::string(argv[5]) + "_U.bin", U.U, nu);
        wr(std::string(argv[5]) + "_E.bin", E.E, ne);
        {
            tcym_dropin::Resident keep;
            for (int k = 0; k < 4; k++) Dynamics::Update(&U, &E);
        }
        wr(std::string(argv[5]) + "_U2.bin", U.U, nu);
        wr(std::string(argv[5]) + "_E2.bin", E.E, ne);

        // a cold copy blocked once after flow step 3 of 8
        GradientFlow::SqrtDcTauOverSpacing = 0.2;
        Cooling::BlockFrequency = 3;
        GaugeLinks UC(N, N, N, 1.0, 1.0, 1.0);
        Cooling::CoolNSave("cold", 8, 1, &U, &UC);
        wr(std::string(argv[5]) + "_C.bin", UC.U, (size_t)3*tcym_dropin::LinkSize()*UC.Volume);
        printf("%d %.17g %d %d %.17g\n", UC.N[0], UC.a[0], Cooling::DynamicFields::BlockingNumber, GradientFlow::CoolingSteps,
               GradientFlow::CoolingTime());
        // CoolingMethod::Start with 4 flow steps, then Calibrate: 5 more calibration
        // steps, blocked after step 6, logged to <dir>/VacuumCoolingT0.4ID7.log
        ChernSimonsNumber::CoolingMethod::StandardCoolingMaxSteps = 4;
        ChernSimonsNumber::CoolingMethod::CalibrationCoolingMaxSteps = 5;
        ChernSimonsNumber::CoolingMethod::CalibrationBlockingLevel = 1;
        GLinks::U = &U;
        EFields::E = &E;
        ChernSimonsNumber::CoolingMethod::Init();
        ChernSimonsNumber::CoolingMethod::Start();
        Cooling::BlockFrequency = 6;
        IO::SetOutputDirectory(argv[6]);
        RandomNumberGenerator::MySEED = 7;
        ChernSimonsNumber::CoolingMethod::Calibrate();
        printf("%.17g %d %d %.17g\n", ChernSimonsNumber::DeltaNCsCalibration, Cooling::DynamicFields::BlockingNumber,
               GradientFlow::CoolingSteps, GradientFlow::CoolingTime());
        // the hot fields as text (<dir>/UOutT0.4ID7.txt, EOutT0.4ID7.txt), read
        // back into new containers
        IO::SaveConfiguration("UOutT0.4", "EOutT0.4", &U, &E);
        IO::SetInputDirectory(argv[6]);
        IO::SetInputFile(0.4, 7);
        GaugeLinks U3(N, N, N, 1.0, 1.0, 1.0);
        ElectricFields E3(N, N, N, 1.0, 1.0, 1.0);
        IO::LoadConfiguration(&U3, &E3);
        printf("%.17g\n", ChernSimonsNumber::NCsDot(&U3, &E3));
        // the thermal initial condition: SetZero, then 2 cycles to time 1.6 at
        // beta 4 (the rows go to <dir>/ThermalEnergyID7.txt)
        GaugeLinks UT(N, N, N, 1.0, 1.0, 1.0);
        ElectricFields ET(N, N, N, 1.0, 1.0, 1.0);
        InitialConditions::SetZero(&UT, &ET);
        ThermalDynamics::beta = 4.0;
        ThermalDynamics::Thermalize(2, 1.6, &UT, &ET);
        wr(std::string(argv[5]) + "_TU.bin", UT.U, nu);
        wr(std::string(argv[5]) + "_TE.bin", ET.E, ne);
        // quasi-particles (Qs 1.5, n0 1, key MySEED = 7), Gauss law restored
        Qs = 1.5;
        InitialConditions::SetQuasiParticles(&UT, &ET);
        wr(std::string(argv[5]) + "_QU.bin", UT.U, nu);
        wr(std::string(argv[5]) + "_QE.bin", ET.E, ne);
    } catch (const tcym_dropin::Error& e) {
        fprintf(stderr, "%s (code %d)\n", e.what(), e.code);
        return 1;
    }
    return 0;
}
