// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// threaded text reader/writer and the binary checkpoints of
// topologicalcymsolver_amd/csrc/tcym_io.cpp, and the oracle's sweeps and
// observables (oracle/tcym_oracle.c), on small lattices.  Built with g++ by
// tests/test_sanitize.py; any sanitizer report aborts the program.
//
// tcym_io.cpp's device-slot entry points (tcym_save_text, ...) call into the
// HIP library, which a CPU build cannot link; the definitions below make
// those calls fail, and the harness checks that the error reaches the caller.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "tcym.h"
#include "tcym_io.h"
#include "tcym_oracle.h"

extern "C" {
const char* tcym_last_error(const tcym_ctx*) { return "no device in the sanitizer build"; }
int tcym_get_params(const tcym_ctx*, tcym_params*) { return TCYM_E_UNSUPPORTED; }
int tcym_level(tcym_ctx*, int, int, int*, int*) { return TCYM_E_UNSUPPORTED; }
int tcym_get_links(tcym_ctx*, int, double*) { return TCYM_E_UNSUPPORTED; }
int tcym_get_efield(tcym_ctx*, int, double*) { return TCYM_E_UNSUPPORTED; }
int tcym_set_links(tcym_ctx*, int, const double*) { return TCYM_E_UNSUPPORTED; }
int tcym_set_efield(tcym_ctx*, int, const double*) { return TCYM_E_UNSUPPORTED; }
}

static int failures = 0;
#define CHECK(c) do { if (!(c)) { fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); failures++; } } while (0)

static std::string dir;

static void io_checks(int nc) {
    const int N[3] = {5, 4, 3};
    const int V = N[0]*N[1]*N[2], su = nc == 2 ? 4 : 18, ne = nc == 2 ? 3 : 8;
    std::mt19937_64 rng(nc);
    std::normal_distribution<double> z(0.0, 1.0);
    std::vector<double> U(3*V*su), E(3*V*ne), U2(U.size()), E2(E.size());
    for (auto& v : U) v = z(rng)*1e3;
    for (auto& v : E) v = z(rng)*1e-7;
    U[0] = -0.0; U[1] = 1e23; E[0] = 1e-5;
    const std::string uf = dir + "/U" + std::to_string(nc) + ".txt", ef = dir + "/E" + std::to_string(nc) + ".txt";
    for (int threads : {1, 3, 8}) {
        CHECK(tcym_write_text_links(nc, N, U.data(), uf.c_str(), 17, threads) == TCYM_OK);
        CHECK(tcym_write_text_efield(nc, N, E.data(), ef.c_str(), 17, threads) == TCYM_OK);
        CHECK(tcym_read_text_links(nc, N, U2.data(), uf.c_str(), threads) == TCYM_OK);
        CHECK(tcym_read_text_efield(nc, N, E2.data(), ef.c_str(), threads) == TCYM_OK);
        CHECK(memcmp(U.data(), U2.data(), U.size()*sizeof(double)) == 0);
        CHECK(memcmp(E.data(), E2.data(), E.size()*sizeof(double)) == 0);
        CHECK(tcym_write_text_links(nc, N, U.data(), uf.c_str(), 15, threads) == TCYM_OK);
        CHECK(tcym_read_text_links(nc, N, U2.data(), uf.c_str(), threads) == TCYM_OK);
    }
    // malformed files: a truncated copy, a duplicated line, an empty file, a missing file
    {
        FILE* in = fopen(uf.c_str(), "rb");
        std::vector<char> buf(1 << 20);
        const size_t n = fread(buf.data(), 1, buf.size(), in);
        fclose(in);
        const std::string bad = dir + "/bad.txt";
        FILE* out = fopen(bad.c_str(), "wb");
        fwrite(buf.data(), 1, n/2, out);
        fclose(out);
        CHECK(tcym_read_text_links(nc, N, U2.data(), bad.c_str(), 4) == TCYM_E_IO);
        const char* nl = (const char*)memchr(buf.data(), '\n', n);
        out = fopen(bad.c_str(), "wb");
        fwrite(buf.data(), 1, (size_t)(nl - buf.data()) + 1, out);
        fwrite(buf.data(), 1, n, out);
        fclose(out);
        CHECK(tcym_read_text_links(nc, N, U2.data(), bad.c_str(), 4) == TCYM_E_IO);
        out = fopen(bad.c_str(), "wb");
        fclose(out);
        CHECK(tcym_read_text_links(nc, N, U2.data(), bad.c_str(), 2) == TCYM_E_IO);
        CHECK(tcym_read_text_links(nc, N, U2.data(), (dir + "/none.txt").c_str(), 2) == TCYM_E_IO);
    }
    // binary checkpoints: round trip, links only, a corrupted byte
    tcym_checkpoint_info info{}, back{};
    info.nc = nc; info.precision = TCYM_FP64;
    for (int k = 0; k < 3; k++) { info.N[k] = N[k]; info.a[k] = 1.0 + 0.1*k; }
    info.aScale = 0.9; info.step = 12345; info.seed = 77; info.time = 1.5;
    for (int k = 0; k < 8; k++) info.acc[k] = k*0.25;
    const std::string cf = dir + "/ck" + std::to_string(nc) + ".bin";
    CHECK(tcym_write_checkpoint(cf.c_str(), &info, U.data(), E.data()) == TCYM_OK);
    CHECK(tcym_read_checkpoint(cf.c_str(), &back, U2.data(), E2.data()) == TCYM_OK);
    CHECK(back.step == 12345 && back.has_efield == 1 && memcmp(U.data(), U2.data(), U.size()*sizeof(double)) == 0);
    CHECK(tcym_write_checkpoint(cf.c_str(), &info, U.data(), nullptr) == TCYM_OK);
    CHECK(tcym_read_checkpoint(cf.c_str(), &back, U2.data(), E2.data()) == TCYM_OK && back.has_efield == 0);
    {
        FILE* f = fopen(cf.c_str(), "r+b");
        fseek(f, 600, SEEK_SET);
        fputc(0x5a, f);
        fclose(f);
        CHECK(tcym_read_checkpoint(cf.c_str(), &back, U2.data(), E2.data()) == TCYM_E_IO);
    }
    // the device-slot forms report the failure of their device calls
    tcym_ctx* none = nullptr;
    CHECK(tcym_save_text(none, 0, 0, uf.c_str(), ef.c_str(), 15, 2) != TCYM_OK);
    CHECK(tcym_load_checkpoint(none, 0, 0, cf.c_str(), nullptr) != TCYM_OK);
}

static void oracle_checks(int nc) {
    or_params p{};
    p.nc = nc; p.N[0] = 4; p.N[1] = 3; p.N[2] = 5;
    p.a[0] = 1.0; p.a[1] = 1.3; p.a[2] = 0.7; p.aScale = 0.9; p.nthreads = 4;
    const int V = 60, su = or_su(nc), ne = or_ne(nc);
    std::vector<double> U(3*V*su), E(3*V*ne), Un(U.size()), Um(U.size()), Em(E.size()), G(V*ne), map(V);
    or_random_links(&p, 5, 0.5, U.data());
    or_random_efield(&p, 5, 0.4, E.data());
    double fm = 0.0, a = 0.0, b = 0.0, out6[6], out3[3];
    CHECK(or_gauss_restore(&p, U.data(), E.data(), 10000, 1e-14, &fm) > 0 && fm <= 1e-14);
    or_leapfrog(&p, U.data(), E.data(), 3, 0.05);
    or_langevin(&p, U.data(), E.data(), 3, 0.05, 0.1, 0.5, 9, 0);
    Un = U;
    or_flow(&p, Un.data(), Em.data(), 3, 0.02);
    const double dn = or_delta_ncs_cooling(&p, U.data(), Un.data(), Um.data(), Em.data());
    CHECK(std::isfinite(dn));
    std::vector<double> W = U, F(E.size());
    CHECK(std::isfinite(or_flow_calibration(&p, W.data(), F.data(), 2, 0.02)));
    CHECK(std::isfinite(or_edotb(&p, U.data(), E.data())) && std::isfinite(or_edotb_unimproved(&p, U.data(), E.data())));
    or_edotb_map(&p, U.data(), E.data(), map.data());
    or_bulk(&p, U.data(), E.data(), out6);
    or_hard_scales(&p, U.data(), out3);
    or_gauss(&p, U.data(), E.data(), &a, &b);
    or_gauss_field(&p, U.data(), E.data(), G.data());
    or_unitarity(&p, U.data(), &a, &b);
    CHECK(std::isfinite(or_vacuum_ncs(&p, U.data())));
    or_reunitarize_field(&p, U.data());
    // blocking needs even extents
    or_params q = p;
    q.N[0] = 4; q.N[1] = 4; q.N[2] = 6;
    std::vector<double> Uq(3*96*su), Uc(3*12*su);
    or_random_links(&q, 3, 0.5, Uq.data());
    or_block_links(&q, Uq.data(), Uc.data());
    std::vector<double> log(7*64);
    CHECK(or_thermalize(&q, Uq.data(), std::vector<double>(3*96*ne).data(), 1, 0.5, 0.0375, 4.0, 7, 10000, 1e-14,
                        log.data(), 64) >= 0);
}

int main(int argc, char** argv) {
    if (argc != 2) { fprintf(stderr, "usage: san_check dir\n"); return 2; }
    dir = argv[1];
    for (int nc : {2, 3}) {
        io_checks(nc);
        oracle_checks(nc);
    }
    if (failures) { fprintf(stderr, "%d checks failed\n", failures); return 1; }
    printf("sanitizer checks passed\n");
    return 0;
}
