// A sphaleron-run driver written against the reference's interface only: it
// makes the calls of SphaleronSimulation.cpp's set-up and time loop
// (:230-237 parameters, :277 CoolingMethod::Init, :377 Bulk::Update,
// :405-419 the N_CS resets, Start and Calibrate, :439-560 the loop with
// Dynamics::Update, the energy output every 20 steps, CoolingMethod::Update(1),
// the E.B maps of the hot and cooled fields, Calibrate and the precision
// checks), then a few Langevin steps (LangevinSimulation.cpp's
// LangevinDynamics::Update and UpdateRandomGaussianNoise + UpdateFields), and
// checks that a partial box is refused; last, two cooled copies in a row
// through one host object (CoolNSave then NCsDot).  It compiles against
// include/tcym_dropin.hpp unchanged.
//
// usage: sphaleron_driver nc N U.bin E.bin out_dir resident
//   writes out_dir/{EnergyID7.txt, CoolNCsID7.txt, CalibNCsID7.txt} (15 digits),
//   the library's HotEdotBMapC0ID7.txt, CoolEdotBMapC0ID7.txt and VacuumCooling
//   logs, and out_dir/final_{U,E}.bin; resident = 1 runs inside a Resident scope.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "tcym_dropin.hpp"

DOUBLE Qs = 1.0;
DOUBLE n0 = 1.0;

namespace Run {

INT CoolingFrequency, CalibFrequency;
std::ofstream EnergyOut, CoolOut, CalibOut;

void Init(INT N) {
    // measurement cadence of a small lattice (the reference: 10 / 2000 steps, 128 flow steps)
    CoolingFrequency = 5; ChernSimonsNumber::CoolingMethod::StandardCoolingMaxSteps = 12;
    CalibFrequency = CoolingFrequency*4; ChernSimonsNumber::CoolingMethod::CalibrationCoolingMaxSteps = 6;
    Cooling::BlockFrequency = 8;
    ChernSimonsNumber::CoolingMethod::StandardBlockingLevel = 0;
    ChernSimonsNumber::CoolingMethod::CalibrationBlockingLevel = 1;
    Lattice::N[0] = Lattice::N[1] = Lattice::N[2] = N;
    Lattice::Init();
    ChernSimonsNumber::CoolingMethod::Init();
}

void Energy() {
    Observables::Bulk::Update();
    Observables::HardScales::Update();
    EnergyOut << Dynamics::Time() << " " << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " "
              << Observables::Bulk::TYY() << " " << Observables::Bulk::TZZ() << " " << Observables::Bulk::ELECTRIC() << " "
              << Observables::Bulk::MAGNETIC() << " " << Observables::HardScales::LambdaXX() << " "
              << Observables::HardScales::LambdaYY() << " " << Observables::HardScales::LambdaZZ() << std::endl;
}

void Calibration() {
    CalibOut << Dynamics::Time() << " " << ChernSimonsNumber::DeltaNCsCalibration << " "
             << ChernSimonsNumber::DeltaNCsPreviousCalibration << " " << ChernSimonsNumber::DeltaNCsRealTimeCalibration << " "
             << ChernSimonsNumber::DeltaNCsRealTimePreviousCalibration << std::endl;
}

void Evolve(DOUBLE MaxTime) {
    const std::string dir = IO::OutputDirectory;
    EnergyOut.open((dir + "EnergyID7.txt").c_str());
    CoolOut.open((dir + "CoolNCsID7.txt").c_str());
    CalibOut.open((dir + "CalibNCsID7.txt").c_str());
    EnergyOut.precision(15); CoolOut.precision(15); CalibOut.precision(15);

    Energy();
    ChernSimonsNumber::DeltaNCsRealTime = DOUBLE(0.0);
    ChernSimonsNumber::DeltaNCsCoolRealTime = DOUBLE(0.0);
    ChernSimonsNumber::DeltaNCsCooling = DOUBLE(0.0);
    ChernSimonsNumber::CoolingMethod::Start();
    CoolOut << Dynamics::Time() << " " << ChernSimonsNumber::DeltaNCsCoolRealTime << " " << ChernSimonsNumber::DeltaNCsCooling << std::endl;
    ChernSimonsNumber::CoolingMethod::Calibrate();
    Calibration();

    INT MovieCount = 0;
    while (Dynamics::Time() < MaxTime) {
        Dynamics::Update();
        if (Dynamics::tSteps%20 == 0) Energy();
        if (Dynamics::tSteps%CoolingFrequency == 0) {
            ChernSimonsNumber::CoolingMethod::Update(1);
            CoolOut << Dynamics::Time() << " " << ChernSimonsNumber::DeltaNCsCoolRealTime << " " << ChernSimonsNumber::DeltaNCsCooling << std::endl;
            if (MovieCount == 0) {
                Observables::EdotBMap::CreateMap("HotEdotBMapC0");
                Observables::EdotBMap::CreateMap("CoolEdotBMapC0", ChernSimonsNumber::CoolingMethod::UMid, ChernSimonsNumber::CoolingMethod::EMid);
                MovieCount++;
            }
        }
        if (Dynamics::tSteps%CalibFrequency == 0) {
            ChernSimonsNumber::CoolingMethod::Calibrate();
            Calibration();
        }
        if (Dynamics::tSteps%25 == 0) {
            Observables::GaussLaw::CheckViolation();
            Observables::Unitarity::CheckViolation();
        }
    }
    EnergyOut.close(); CoolOut.close(); CalibOut.close();
}

}  // namespace Run

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
    if (argc != 7) { fprintf(stderr, "usage: sphaleron_driver nc N U.bin E.bin out_dir resident\n"); return 2; }
    const int nc = atoi(argv[1]), N = atoi(argv[2]), resident = atoi(argv[6]);
    tcym_dropin::Configure(nc, TCYM_FP64, 0);
    try {
        Run::Init(N);
        const size_t nu = (size_t)3*tcym_dropin::LinkSize()*GLinks::U->Volume;
        const size_t ne = (size_t)3*tcym_dropin::VectorSize()*EFields::E->Volume;
        rd(argv[3], GLinks::U->U, nu);
        rd(argv[4], EFields::E->E, ne);
        RandomNumberGenerator::Init(7);
        IO::SetOutputDirectory(argv[5]);
        Dynamics::Reset();
        {
            tcym_dropin::Resident* keep = resident ? new tcym_dropin::Resident() : nullptr;
            Run::Evolve(1.0);   // 40 steps of dTau 0.025
            // Langevin steps of LangevinSimulation.cpp: fused noise, then the noise
            // field and UpdateFields, on the same fields
            LangevinVariables::Init();
            LangevinDynamics::beta = 2.0;
            LangevinDynamics::dTauL = 0.01;
            for (int k = 0; k < 3; k++) LangevinDynamics::Update();
            LangevinDynamics::UpdateRandomGaussianNoise();
            LangevinDynamics::UpdateFields();
            delete keep;
        }
        wr(std::string(argv[5]) + "/final_U.bin", GLinks::U->U, nu);
        wr(std::string(argv[5]) + "/final_E.bin", EFields::E->E, ne);
        // a box short of the lattice is refused, as is an unknown order of GetDeltaNCs
        int refused = 0;
        try { Dynamics::UpdateElectricFields(0, N - 2, 0, N - 1, 0, N - 1, GLinks::U, EFields::E); }
        catch (const tcym_dropin::Error& e) { refused += e.code == TCYM_E_UNSUPPORTED; }
        try { Observables::GaussLaw::UpdateViolation(0, N - 1, 1, N - 1, 0, N - 1, GLinks::U, EFields::E); }
        catch (const tcym_dropin::Error& e) { refused += e.code == TCYM_E_UNSUPPORTED; }
        try { ChernSimonsNumber::CoolingMethod::GetDeltaNCs(2); }
        catch (const tcym_dropin::Error& e) { refused += e.code == TCYM_E_UNSUPPORTED; }
        // a sub-box map writes the sites of the box only
        Observables::EdotBMap::CreateMap("BoxMap", 1, 2, 0, 0, 3, 3, GLinks::U, EFields::E);
        // two cooled copies in a row measured through the same host object
        GaugeLinks UC(N, N, N, 1.0, 1.0, 1.0);
        Cooling::BlockFrequency = 0;
        double v[2];
        {
            tcym_dropin::Resident* keep = resident ? new tcym_dropin::Resident() : nullptr;
            for (int k = 0; k < 2; k++) {
                Cooling::CoolNSave("c", 6, 0, GLinks::U, &UC);
                v[k] = ChernSimonsNumber::NCsDot(&UC, EFields::E);
                Dynamics::Update();
            }
            delete keep;
        }
        printf("refused %d tSteps %d tauLSteps %d\n", refused, Dynamics::tSteps, LangevinDynamics::tauLSteps);
        printf("%.17g %.17g\n", v[0], v[1]);
    } catch (const tcym_dropin::Error& e) {
        fprintf(stderr, "%s (code %d)\n", e.what(), e.code);
        return 1;
    }
    return 0;
}
