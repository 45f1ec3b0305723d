// Exercises the reference-named entry points of include/tcym_dropin.hpp the way
// a TopologicalCYMSolver driver calls them, on fields read from binary files.
// usage: dropin_check nc N U.bin E.bin out_prefix
// writes out_prefix_U.bin / out_prefix_E.bin after 4 Dynamics::Update steps
// (dTau 0.05), then 4 more steps inside a Resident scope into *_U2/_E2, and
// prints "ncsdot T00 gauss_max unitarity_max" for the first state.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tcym_dropin.hpp"

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
    if (argc != 6) { fprintf(stderr, "usage: dropin_check nc N U.bin E.bin out_prefix\n"); return 2; }
    const int nc = atoi(argv[1]), N = atoi(argv[2]);
    tcym_dropin::Configure(nc, TCYM_FP64, 0);
    GaugeLinks U(N, N, N, 1.0, 1.0, 1.0);
    ElectricFields E(N, N, N, 1.0, 1.0, 1.0);
    const size_t nu = (size_t)3*tcym_dropin::LinkSize()*U.Volume, ne = (size_t)3*tcym_dropin::VectorSize()*E.Volume;
    rd(argv[3], U.U, nu);
    rd(argv[4], E.E, ne);
    try {
        Dynamics::dTau() = 0.05;
        for (int k = 0; k < 4; k++) Dynamics::Update(&U, &E);
        const double ncs = ChernSimonsNumber::NCsDot(&U, &E);
        Observables::Bulk::Update(&U, &E);
        Observables::GaussLaw::UpdateViolation(0, N - 1, 0, N - 1, 0, N - 1, &U, &E);
        Observables::Unitarity::UpdateViolation(0, N - 1, 0, N - 1, 0, N - 1, &U);
        Observables::HardScales::Update(&U);
        printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", ncs, Observables::Bulk::T00(),
               Observables::GaussLaw::GlobalMaxViolation(), Observables::Unitarity::GlobalMaxViolation(),
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
    } catch (const tcym_dropin::Error& e) {
        fprintf(stderr, "%s (code %d)\n", e.what(), e.code);
        return 1;
    }
    return 0;
}
