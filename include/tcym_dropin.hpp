// tcym_dropin.hpp -- the reference's hot-path entry points and globals, with
// their original namespaces, names and signatures, over GaugeLinks /
// ElectricFields host containers in the reference's own memory layout,
// forwarding to the MI355X library through the C ABI of tcym.h.
//
// A driver written against TopologicalCYMSolver, such as the time loop of
// SphaleronSimulation.cpp:230-560, includes this header instead of the
// reference's sources and links libtcym.so.  It finds:
//   the globals                 Lattice::{N,a,aScale,Volume,aCube,Init}, GLinks::U, EFields::E,
//                               LangevinVariables::Xi (LATTICE/3DGrid.cpp, LATTICE/GLinksEFields.cpp,
//                               DYNAMICS/LangevinDynamics.cpp:4-15), Dynamics::{dTau,tSteps},
//                               GradientFlow::{SqrtDcTauOverSpacing,cTau,CoolingSteps},
//                               Cooling::BlockFrequency, ChernSimonsNumber::DeltaNCs*,
//                               ChernSimonsNumber::CoolingMethod::{Standard,Calibration}{CoolingMaxSteps,
//                               BlockingLevel}, {UOld,UMid,UNew,EMid}, IO::OutputDirectory,
//                               RandomNumberGenerator::MySEED -- plain assignable variables
//   Dynamics::Update()/Update(U,E), UpdateElectricFields/UpdateGaugeLinks   DYNAMICS/Dynamics.cpp:97-223
//   LangevinDynamics::Update(), UpdateFields()/UpdateFields(U,E),
//     UpdateRandomGaussianNoise()                                           DYNAMICS/LangevinDynamics.cpp:260-347
//   GradientFlow::StandardUpdate / CalibrationUpdate                        GRADIENTFLOW/GradientFlow.cpp:150,168
//   Cooling::CoolNSave / ContinueCooling                                    GRADIENTFLOW/Cooling.cpp:156,181
//   ChernSimonsNumber::NCsDot                                               TOPOLOGY/ChernSimonsDerivative.cpp:31
//   ChernSimonsNumber::CoolingMethod::{Init,PerformInterpolation,GetDeltaNCs,Start,Update,Calibrate}
//                                                                           TOPOLOGY/CoolingMethod.cpp:19-196
//   ImprovedOperators / UnimprovedOperators::ComputeEDotB                   CALC/ChernSimonsOperators.cpp:49,384
//   Observables::{Bulk,HardScales}::Update(), GaussLaw / Unitarity
//     ::{UpdateViolation,CheckViolation}, EdotBMap::CreateMap               OBSERVABLES/*.cpp
//   GaussLawRestoration::Restore                                            GAUSSLAWFIXING/GaussRestore.cpp:131,169
//   IO::SaveConfiguration / LoadConfiguration                               IO/SaveConfiguration.cpp:3,58, IO/LoadConfiguration.cpp:3
//   InitialConditions::SetZero / SetQuasiParticles, ThermalDynamics::Thermalize
//                                                                           INITIALCONDITIONS/*.cpp, DYNAMICS/ThermalDynamics.cpp:260
// The driver defines the globals Qs and n0 itself, as the reference's drivers
// do (SphaleronSimulation.cpp:76-77); this header declares them extern.
//
// Group and precision are run-time values (tcym_dropin::Configure) instead of
// the SU_Nc_FLAG macro.  The step sizes the reference fixes as constants
// (Dynamics::dTau, LangevinDynamics::dTauL) are variables with the reference's
// values.  Sub-box arguments (xLow .. zHigh) must span the whole lattice, as
// every reference caller passes them, except for EdotBMap::CreateMap, which
// writes any box; a partial box throws TCYM_E_UNSUPPORTED.
//
// Residency: by default every call uploads the host fields it reads and
// downloads the ones it writes, so host arrays are always current.  Inside a
// tcym_dropin::Resident scope the fields stay on the device between calls and
// are downloaded when the scope ends (or by tcym_dropin::SyncToHost()).  The
// cooling method's UOld, UMid, UNew and EMid mirror device slots: the cooling
// method reads the slots, and the host objects receive their contents.
// Failures throw tcym_dropin::Error instead of calling exit(0).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

#include "tcym.h"
#include "tcym_io.h"

typedef int INT;
typedef double DOUBLE;

// defined by the driver (SphaleronSimulation.cpp:76-77)
extern DOUBLE Qs;
extern DOUBLE n0;

namespace tcym_dropin {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Config {
    int nc = 3;
    int precision = TCYM_FP64;
    int device = 0;
};
inline Config& config() { static Config c; return c; }
inline void Configure(int nc, int precision = TCYM_FP64, int device = 0) {
    config().nc = nc; config().precision = precision; config().device = device;
}
inline int LinkSize() { return config().nc == 2 ? 4 : 18; }
inline int VectorSize() { return config().nc == 2 ? 3 : 8; }

}  // namespace tcym_dropin

// ---------------------------------------------------------------- containers
// Same members and index functions as LATTICE/GaugeLinks.cpp:36-114 (spacings
// kept as DOUBLE rather than truncated to INT) and LATTICE/ElectricFields.cpp:9-88.
class GaugeLinks {
public:
    INT Volume;
    static const INT Dimension = 3;
    DOUBLE a[Dimension];
    DOUBLE aCube;
    INT N[Dimension];
    DOUBLE* U;
    INT Index3D(INT x, INT y, INT z) const {
        return md(x, N[0]) + N[0]*(md(y, N[1]) + N[1]*md(z, N[2]));
    }
    INT Index(INT x, INT y, INT z, INT mu) const { return tcym_dropin::LinkSize()*(mu + Dimension*Index3D(x, y, z)); }
    DOUBLE* Get(INT x, INT y, INT z, INT mu) { return &U[Index(x, y, z, mu)]; }
    GaugeLinks(INT Nx, INT Ny, INT Nz, DOUBLE ax, DOUBLE ay, DOUBLE az) {
        Volume = Nx*Ny*Nz; N[0] = Nx; N[1] = Ny; N[2] = Nz;
        a[0] = ax; a[1] = ay; a[2] = az; aCube = ax*ay*az;
        U = new DOUBLE[(size_t)Dimension*tcym_dropin::LinkSize()*Volume]();
    }
    ~GaugeLinks() { delete[] U; }
    GaugeLinks(const GaugeLinks&) = delete;
    GaugeLinks& operator=(const GaugeLinks&) = delete;
private:
    static INT md(INT x, INT n) { return x >= 0 ? x % n : (x + n) % n; }
};

class ElectricFields {
public:
    INT Volume;
    static const INT Dimension = 3;
    DOUBLE a[Dimension];
    DOUBLE aCube;
    INT N[Dimension];
    DOUBLE* E;
    INT Index3D(INT x, INT y, INT z) const {
        return md(x, N[0]) + N[0]*(md(y, N[1]) + N[1]*md(z, N[2]));
    }
    INT Index(INT x, INT y, INT z, INT mu, INT a_) const { return a_ + tcym_dropin::VectorSize()*(mu + Dimension*Index3D(x, y, z)); }
    DOUBLE* Get(INT x, INT y, INT z, INT mu, INT a_) { return &E[Index(x, y, z, mu, a_)]; }
    void SetZero() { std::memset(E, 0, sizeof(DOUBLE)*(size_t)Dimension*tcym_dropin::VectorSize()*Volume); }
    ElectricFields(INT Nx, INT Ny, INT Nz, DOUBLE ax, DOUBLE ay, DOUBLE az) {
        Volume = Nx*Ny*Nz; N[0] = Nx; N[1] = Ny; N[2] = Nz;
        a[0] = ax; a[1] = ay; a[2] = az; aCube = ax*ay*az;
        E = new DOUBLE[(size_t)Dimension*tcym_dropin::VectorSize()*Volume]();
    }
    ~ElectricFields() { delete[] E; }
    ElectricFields(const ElectricFields&) = delete;
    ElectricFields& operator=(const ElectricFields&) = delete;
private:
    static INT md(INT x, INT n) { return x >= 0 ? x % n : (x + n) % n; }
};

// ---------------------------------------------------------------- lattice globals
namespace Lattice {
static const int Dimension = 3;                 // LATTICE/3DGrid.cpp:11
inline INT N[3] = {64, 64, 64};                 // :14
inline DOUBLE a[3] = {1.0, 1.0, 1.0};           // :15
inline DOUBLE aScale = 1.0;                     // :18
inline INT Volume = 64*64*64;                   // :21
inline DOUBLE aCube = 1.0;                      // LATTICE/GLinksEFields.cpp:34
}  // namespace Lattice

namespace GLinks {
inline GaugeLinks* U = nullptr;                 // LATTICE/GLinksEFields.cpp:6
inline void Init() { U = new GaugeLinks(Lattice::N[0], Lattice::N[1], Lattice::N[2], Lattice::a[0], Lattice::a[1], Lattice::a[2]); }
}
namespace EFields {
inline ElectricFields* E = nullptr;             // LATTICE/GLinksEFields.cpp:20
inline void Init() { E = new ElectricFields(Lattice::N[0], Lattice::N[1], Lattice::N[2], Lattice::a[0], Lattice::a[1], Lattice::a[2]); }
}
namespace Lattice {
// LATTICE/GLinksEFields.cpp:36-45
inline void Init() {
    GLinks::Init();
    EFields::Init();
    Volume = N[0]*N[1]*N[2];
    aCube = GLinks::U->a[0]*GLinks::U->a[1]*GLinks::U->a[2];
}
}  // namespace Lattice
namespace LangevinVariables {
inline ElectricFields* Xi = nullptr;            // DYNAMICS/LangevinDynamics.cpp:6
inline void Init() {                            // :9-13
    Xi = new ElectricFields(EFields::E->N[0], EFields::E->N[1], EFields::E->N[2], EFields::E->a[0], EFields::E->a[1], EFields::E->a[2]);
}
}

namespace RandomNumberGenerator {
inline long int MySEED = 0;                     // MISC/RNG/GSLRandomNumberGenerator.cpp:9
// seeds the Philox streams that replace the GSL generator (:12-23)
inline void Init(long int SEED) { MySEED = SEED; }
}

namespace IO {
inline std::string OutputDirectory;             // IO/OutputManagement.cpp:5
template <typename A>
inline void SetOutputDirectory(A x) {           // :7-13
    std::ostringstream s;
    s << x << "/";
    OutputDirectory = s.str();
    std::cerr << "#OUTPUT DIRECTORY IS " << OutputDirectory << std::endl;
}
}  // namespace IO

namespace tcym_dropin {

// One device context per lattice geometry.  Host objects are bound to device
// slots; a binding remembers whether the device copy is newer than the host.
class Device {
public:
    static std::map<std::string, Device*>& registry() { static std::map<std::string, Device*> m; return m; }
    static std::string key(const INT* N, const DOUBLE* a) {
        char k[160];
        snprintf(k, sizeof k, "%d %d %d %d %d %.17g %.17g %.17g %.17g", config().nc, config().precision,
                 N[0], N[1], N[2], a[0], a[1], a[2], Lattice::aScale);
        return k;
    }
    // the context of a lattice geometry, created on first use
    static Device& get(const INT* N, const DOUBLE* a) {
        std::map<std::string, Device*>& all = registry();
        const std::string k = key(N, a);
        auto it = all.find(k);
        if (it != all.end()) return *it->second;
        Device* d = new Device(N, a);
        all[k] = d;
        return *d;
    }
    // the context of a lattice geometry if one exists
    static Device* find(const INT* N, const DOUBLE* a) {
        auto it = registry().find(key(N, a));
        return it == registry().end() ? nullptr : it->second;
    }
    tcym_ctx* ctx = nullptr;
    bool resident = false;
    static bool& resident_flag() { static bool r = false; return r; }

    void check(int r) const {
        if (r != TCYM_OK) throw Error(r, std::string("tcym: ") + tcym_last_error(ctx));
    }
    // bind host links to a device slot.  Resident: reuse any slot that holds a
    // current copy of the object; otherwise upload into `slot` (first saving a
    // device-newer object that the slot held, and any device-newer copy of
    // the object held by this or another context).
    int links(GaugeLinks* U, int slot) {
        if (resident)
            for (int s = 0; s < TCYM_NUM_U; s++) if (ub[s].host == U && ub[s].valid) return s;
        for (auto& kv : registry()) if (kv.second != this) kv.second->forget(U, true);
        flush_links(slot);
        for (int s = 0; s < TCYM_NUM_U; s++) if (ub[s].host == U) { flush_links(s); ub[s] = Bind(); }
        check(tcym_set_links(ctx, slot, U->U));
        ub[slot] = Bind{U, true, false};
        return slot;
    }
    int field(ElectricFields* E, int slot) {
        if (resident)
            for (int s = 0; s < TCYM_NUM_E; s++) if (eb[s].host == E && eb[s].valid) return s;
        for (auto& kv : registry()) if (kv.second != this) kv.second->forget(E, true);
        flush_field(slot);
        for (int s = 0; s < TCYM_NUM_E; s++) if (eb[s].host == E) { flush_field(s); eb[s] = Bind(); }
        check(tcym_set_efield(ctx, slot, E->E));
        eb[slot] = Bind{E, true, false};
        return slot;
    }
    // a device operation is about to overwrite `slot` with the new contents of
    // host object h: a different object the slot held is saved first; copies
    // of h elsewhere are superseded and forgotten
    void will_write_links(int slot, const void* h) {
        if (ub[slot].host != h) flush_links(slot);
        ub[slot] = Bind();
        for (auto& kv : registry()) kv.second->forget(h, false);
    }
    void will_write_field(int slot, const void* h) {
        if (eb[slot].host != h) flush_field(slot);
        eb[slot] = Bind();
        for (auto& kv : registry()) kv.second->forget(h, false);
    }
    // `slot` now holds the newest contents of h: download now unless resident
    void wrote_links(int slot, GaugeLinks* h) {
        ub[slot] = Bind{h, true, true};
        if (!resident) flush_links(slot);
    }
    void wrote_field(int slot, ElectricFields* h) {
        eb[slot] = Bind{h, true, true};
        if (!resident) flush_field(slot);
    }
    // drop every binding of host object h (saving device-newer contents first if save)
    void forget(const void* h, bool save) {
        for (int s = 0; s < TCYM_NUM_U; s++) if (ub[s].host == h) { if (save) flush_links(s); ub[s] = Bind(); }
        for (int s = 0; s < TCYM_NUM_E; s++) if (eb[s].host == h) { if (save) flush_field(s); eb[s] = Bind(); }
    }
    void flush() {
        for (int s = 0; s < TCYM_NUM_U; s++) flush_links(s);
        for (int s = 0; s < TCYM_NUM_E; s++) flush_field(s);
    }
    void invalidate() {
        for (auto& b : ub) b.valid = false;
        for (auto& b : eb) b.valid = false;
    }
private:
    struct Bind { void* host = nullptr; bool valid = false, dirty = false; };
    Bind ub[TCYM_NUM_U], eb[TCYM_NUM_E];
    void flush_links(int s) {
        if (ub[s].dirty && ub[s].host) check(tcym_get_links(ctx, s, ((GaugeLinks*)ub[s].host)->U));
        ub[s].dirty = false;
    }
    void flush_field(int s) {
        if (eb[s].dirty && eb[s].host) check(tcym_get_efield(ctx, s, ((ElectricFields*)eb[s].host)->E));
        eb[s].dirty = false;
    }
    Device(const INT* N, const DOUBLE* a) {
        tcym_params p;
        p.nc = config().nc; p.precision = config().precision; p.device = config().device;
        for (int k = 0; k < 3; k++) { p.N[k] = N[k]; p.a[k] = a[k]; }
        p.aScale = Lattice::aScale;
        int r = tcym_create(&p, &ctx);
        if (r != TCYM_OK) throw Error(r, "tcym_create failed");
    }
};

inline Device& dev(const GaugeLinks* U) { Device& d = Device::get(U->N, U->a); d.resident = Device::resident_flag(); return d; }

// RAII residency scope: fields stay on the device until the scope ends, then
// every device-newer field is downloaded to its host object (which must still
// exist at that point)
struct Resident {
    Resident() { Device::resident_flag() = true; }
    ~Resident() {
        for (auto& kv : Device::registry()) kv.second->flush();
        Device::resident_flag() = false;
    }
};
// download every device-newer field of every context to its host object
inline void SyncToHost() { for (auto& kv : Device::registry()) kv.second->flush(); }
inline void SyncToHost(GaugeLinks* U) { if (Device* d = Device::find(U->N, U->a)) d->flush(); }
// host arrays were modified by the caller: re-upload on next use
inline void HostModified(GaugeLinks* U) { if (Device* d = Device::find(U->N, U->a)) d->invalidate(); }

// sub-box arguments must cover the whole lattice
inline void full_box(const INT* N, INT xl, INT xh, INT yl, INT yh, INT zl, INT zh, const char* who) {
    if (xl != 0 || yl != 0 || zl != 0 || xh != N[0] - 1 || yh != N[1] - 1 || zh != N[2] - 1)
        throw Error(TCYM_E_UNSUPPORTED, std::string(who) + ": only the full lattice box is supported");
}

}  // namespace tcym_dropin

// ---------------------------------------------------------------- dynamics
namespace Dynamics {
inline DOUBLE tau0 = 0.005;                     // DYNAMICS/Dynamics.cpp:10
inline DOUBLE tau = 0.0;                        // :13
inline INT tSteps = 0;                          // :14
inline DOUBLE dTau = 0.025;                     // :17 (a constant in the reference)
inline DOUBLE Time() { return Qs*(tSteps*dTau); }   // :20-22
inline void Reset() { tau = tau0; tSteps = 0; }     // :80-86 (Minkowski)
inline void UpdateElectricFields(INT xl, INT xh, INT yl, INT yh, INT zl, INT zh, GaugeLinks* U, ElectricFields* E) {   // :145-197
    tcym_dropin::full_box(U->N, xl, xh, yl, yh, zl, zh, "Dynamics::UpdateElectricFields");
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.will_write_field(e, E);
    d.check(tcym_update_efield(d.ctx, u, e, dTau));
    d.wrote_field(e, E);
}
inline void UpdateGaugeLinks(INT xl, INT xh, INT yl, INT yh, INT zl, INT zh, GaugeLinks* U, ElectricFields* E) {   // :97-138
    tcym_dropin::full_box(U->N, xl, xh, yl, yh, zl, zh, "Dynamics::UpdateGaugeLinks");
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.will_write_links(u, U);
    d.check(tcym_update_links(d.ctx, u, e, dTau));
    d.wrote_links(u, U);
}
inline void Update(GaugeLinks* U, ElectricFields* E) {   // :204-218
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.will_write_links(u, U);
    d.will_write_field(e, E);
    d.check(tcym_leapfrog(d.ctx, u, e, 1, dTau));
    d.wrote_links(u, U);
    d.wrote_field(e, E);
    tSteps++;
}
inline void Update() { Update(GLinks::U, EFields::E); }   // :220-223
}  // namespace Dynamics

namespace ThermalDynamics {
inline DOUBLE beta = 2.0;                       // DYNAMICS/ThermalDynamics.cpp:10
inline DOUBLE dTau = 0.0375;                    // :23
}

// ---------------------------------------------------------------- observables (used by the updates below)
namespace Observables {
namespace Bulk {
inline DOUBLE ESqr[3], BSqr[3], cB[3], cE[3];   // OBSERVABLES/BulkObservables.cpp:9-14
inline void SetConstants(GaugeLinks* U) {       // :17-25 (Minkowski)
    for (int mu = 0; mu < 3; mu++) {
        const double c = U->a[mu]*Lattice::aScale*Lattice::aScale/U->aCube;
        cB[mu] = cE[mu] = c*c;
    }
}
inline DOUBLE T00() { return 0.5*(cE[0]*ESqr[0] + cE[1]*ESqr[1] + cE[2]*ESqr[2] + cB[0]*BSqr[0] + cB[1]*BSqr[1] + cB[2]*BSqr[2]); }
inline DOUBLE TXX() { return 0.5*((cE[1]*ESqr[1] + cE[2]*ESqr[2] + cB[1]*BSqr[1] + cB[2]*BSqr[2]) - (cE[0]*ESqr[0] + cB[0]*BSqr[0])); }
inline DOUBLE TYY() { return 0.5*((cE[0]*ESqr[0] + cE[2]*ESqr[2] + cB[0]*BSqr[0] + cB[2]*BSqr[2]) - (cE[1]*ESqr[1] + cB[1]*BSqr[1])); }
inline DOUBLE TZZ() { return 0.5*((cE[0]*ESqr[0] + cE[1]*ESqr[1] + cB[0]*BSqr[0] + cB[1]*BSqr[1]) - (cE[2]*ESqr[2] + cB[2]*BSqr[2])); }
inline DOUBLE ELECTRIC() { return 0.5*(cE[0]*ESqr[0] + cE[1]*ESqr[1] + cE[2]*ESqr[2]); }
inline DOUBLE MAGNETIC() { return 0.5*(cB[0]*BSqr[0] + cB[1]*BSqr[1] + cB[2]*BSqr[2]); }
inline void Update(GaugeLinks* U, ElectricFields* E) {   // :133-143
    auto& d = tcym_dropin::dev(U);
    double o[6];
    d.check(tcym_bulk(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), o));
    for (int mu = 0; mu < 3; mu++) { ESqr[mu] = o[mu]; BSqr[mu] = o[3 + mu]; }
    SetConstants(U);
}
inline void Update() { Update(GLinks::U, EFields::E); }   // :145-147
}  // namespace Bulk

namespace HardScales {
inline DOUBLE H[3], cE[3];                      // OBSERVABLES/HardScales.cpp:9-12
inline void Update(GaugeLinks* U) {             // :104-110
    auto& d = tcym_dropin::dev(U);
    d.check(tcym_hard_scales(d.ctx, d.links(U, TCYM_U_AUX), H));
    for (int mu = 0; mu < 3; mu++) {
        const double c = U->a[mu]*Lattice::aScale*Lattice::aScale/U->aCube;   // :15-22 (Minkowski)
        cE[mu] = c*c;
    }
}
inline void Update() { Update(GLinks::U); }     // :114-116
inline DOUBLE LambdaXX() { return cE[1]*H[1] + cE[2]*H[2] - cE[0]*H[0]; }   // :26-36
inline DOUBLE LambdaYY() { return cE[2]*H[2] + cE[0]*H[0] - cE[1]*H[1]; }
inline DOUBLE LambdaZZ() { return cE[0]*H[0] + cE[1]*H[1] - cE[2]*H[2]; }
}  // namespace HardScales

namespace GaussLaw {
inline DOUBLE GlobalMaxViolation = 0.0, GlobalAvgViolation = 0.0;   // OBSERVABLES/GaussLawViolation.cpp:9-10
inline void UpdateViolation(INT xl, INT xh, INT yl, INT yh, INT zl, INT zh, GaugeLinks* U, ElectricFields* E) {   // :13-61
    tcym_dropin::full_box(U->N, xl, xh, yl, yh, zl, zh, "GaussLaw::UpdateViolation");
    auto& d = tcym_dropin::dev(U);
    d.check(tcym_gauss(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), &GlobalMaxViolation, &GlobalAvgViolation));
}
inline void CheckViolation(GaugeLinks* U, ElectricFields* E) {   // :63-68
    UpdateViolation(0, U->N[0] - 1, 0, U->N[1] - 1, 0, U->N[2] - 1, U, E);
    std::cerr << "#GAUSS LAW VIOLATION AT T=" << Dynamics::Time() << " MAX=" << GlobalMaxViolation << " AVG=" << GlobalAvgViolation << std::endl;
}
inline void CheckViolation() { CheckViolation(GLinks::U, EFields::E); }   // :70-72
}  // namespace GaussLaw

namespace Unitarity {
inline DOUBLE GlobalMaxViolation = 0.0, GlobalAvgViolation = 0.0;   // OBSERVABLES/UnitarityViolation.cpp:9-10
inline void UpdateViolation(INT xl, INT xh, INT yl, INT yh, INT zl, INT zh, GaugeLinks* U) {   // :13-51
    tcym_dropin::full_box(U->N, xl, xh, yl, yh, zl, zh, "Unitarity::UpdateViolation");
    auto& d = tcym_dropin::dev(U);
    d.check(tcym_unitarity(d.ctx, d.links(U, TCYM_U_AUX), &GlobalMaxViolation, &GlobalAvgViolation));
}
inline void CheckViolation(GaugeLinks* U) {     // :53-58
    UpdateViolation(0, U->N[0] - 1, 0, U->N[1] - 1, 0, U->N[2] - 1, U);
    std::cerr << "#UNITARITY VIOLATION AT T=" << Dynamics::Time() << " MAX=" << GlobalMaxViolation << " AVG=" << GlobalAvgViolation << std::endl;
}
inline void CheckViolation(DOUBLE CoolTime, GaugeLinks* U) {   // :60-65
    UpdateViolation(0, U->N[0] - 1, 0, U->N[1] - 1, 0, U->N[2] - 1, U);
    std::cerr << "#UNITARITY VIOLATION AT T=" << Dynamics::Time() << " Tc=" << CoolTime << " MAX=" << GlobalMaxViolation
              << " AVG=" << GlobalAvgViolation << std::endl;
}
inline void CheckViolation() { CheckViolation(GLinks::U); }   // :67-69
// the same check on a device slot (the cooled links of Cooling::Perform)
inline void CheckSlot(tcym_dropin::Device& d, int slot, DOUBLE CoolTime) {
    d.check(tcym_unitarity(d.ctx, slot, &GlobalMaxViolation, &GlobalAvgViolation));
    std::cerr << "#UNITARITY VIOLATION AT T=" << Dynamics::Time() << " Tc=" << CoolTime << " MAX=" << GlobalMaxViolation
              << " AVG=" << GlobalAvgViolation << std::endl;
}
}  // namespace Unitarity
}  // namespace Observables

// ---------------------------------------------------------------- Langevin dynamics
namespace LangevinDynamics {
inline DOUBLE beta = ThermalDynamics::beta;     // DYNAMICS/LangevinDynamics.cpp:23
inline DOUBLE sigmac = 1.0;                     // :28
inline DOUBLE tau0L = 0.005;                    // :34
inline DOUBLE tauL = 0.0;
inline INT tauLSteps = 0;                       // :38
inline DOUBLE dTauL = 0.001;                    // :41 (a constant in the reference)
inline DOUBLE damping = 1.0;                    // fixed at 1 in the reference (:186)
inline DOUBLE Time() { return tauLSteps*dTauL; }   // :44-46
// LangevinDynamics::UpdateRandomGaussianNoise (:260-293): Xi ~ N(0, 2 damping/beta)
// from Philox (stream 2, counter tauLSteps, key MySEED) in place of GSL
inline void UpdateRandomGaussianNoise() {
    auto& d = tcym_dropin::dev(GLinks::U);
    ElectricFields* X = LangevinVariables::Xi;
    d.will_write_field(TCYM_E_XI, X);
    d.check(tcym_noise(d.ctx, TCYM_E_XI, (uint64_t)RandomNumberGenerator::MySEED, (uint32_t)tauLSteps,
                       std::sqrt(2.0*damping/beta)));
    d.wrote_field(TCYM_E_XI, X);
}
// LangevinDynamics::UpdateFields with LangevinVariables::Xi (:300-311)
inline void UpdateFields(GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT), x = d.field(LangevinVariables::Xi, TCYM_E_XI);
    d.will_write_links(u, U);
    d.will_write_field(e, E);
    d.check(tcym_langevin_xi(d.ctx, u, e, x, dTauL, damping));
    d.wrote_links(u, U);
    d.wrote_field(e, E);
    tauLSteps++;
}
inline void UpdateFields() { UpdateFields(GLinks::U, EFields::E); }   // :313-317
// LangevinDynamics::Update (:320-347): the noise of UpdateRandomGaussianNoise
// drawn inside the update kernel (the fields are bit-identical to noise-field
// then UpdateFields; Xi is not written), then the checks every 500 steps
inline void Update() {
    GaugeLinks* U = GLinks::U;
    ElectricFields* E = EFields::E;
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.will_write_links(u, U);
    d.will_write_field(e, E);
    d.check(tcym_langevin(d.ctx, u, e, 1, dTauL, damping, 1.0/beta, (uint64_t)RandomNumberGenerator::MySEED,
                          (uint32_t)tauLSteps));
    d.wrote_links(u, U);
    d.wrote_field(e, E);
    tauLSteps++;
    if (tauLSteps%500 == 0) {
        Observables::GaussLaw::CheckViolation();
        Observables::Unitarity::CheckViolation();
    }
}
}  // namespace LangevinDynamics

namespace ChernSimonsNumber {
inline DOUBLE DeltaNCsRealTime = 0.0;                   // TOPOLOGY/ChernSimonsDerivative.cpp:8
inline DOUBLE DeltaNCsCoolRealTime = 0.0;               // :15
inline DOUBLE DeltaNCsCooling = 0.0;                    // :18
inline DOUBLE DeltaNCsCalibration = 0.0, DeltaNCsPreviousCalibration = 0.0;                   // :25
inline DOUBLE DeltaNCsRealTimeCalibration = 0.0, DeltaNCsRealTimePreviousCalibration = 0.0;   // :28
inline DOUBLE NCsDot(GaugeLinks* U, ElectricFields* E) {   // :31-49
    auto& d = tcym_dropin::dev(U);
    double v;
    d.check(tcym_ncsdot(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), &v));
    return v;
}
}  // namespace ChernSimonsNumber

namespace ImprovedOperators {
inline DOUBLE ComputeEDotB(GaugeLinks* U, ElectricFields* E) {   // CALC/ChernSimonsOperators.cpp:384-388
    auto& d = tcym_dropin::dev(U);
    double v;
    d.check(tcym_edotb(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), 1, &v));
    return v;
}
}  // namespace ImprovedOperators

namespace UnimprovedOperators {
inline DOUBLE ComputeEDotB(GaugeLinks* U, ElectricFields* E) {   // CALC/ChernSimonsOperators.cpp:49-53
    auto& d = tcym_dropin::dev(U);
    double v;
    d.check(tcym_edotb(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), 0, &v));
    return v;
}
}  // namespace UnimprovedOperators

// ---------------------------------------------------------------- gradient flow and cooling
namespace GradientFlow {
inline DOUBLE cTau = 0.0;                               // GRADIENTFLOW/GradientFlow.cpp:11
inline INT CoolingSteps = 0;                            // :13
inline DOUBLE SqrtDcTauOverSpacing = 0.5/std::sqrt(2.0);   // :15
inline void Reset() { cTau = 0.0; CoolingSteps = 0; }   // :19-24
inline DOUBLE CoolingTime() { return cTau; }            // :27-29
inline DOUBLE dtau() { return SqrtDcTauOverSpacing*SqrtDcTauOverSpacing; }
// CoolingSteps++ and cTau += DcTau of one flow step at spacings a (:150-163)
inline void advance(const DOUBLE* a) {
    cTau += dtau()*std::pow(a[0]*a[1]*a[2], 2.0/3.0)/(Lattice::aScale*Lattice::aScale);
    CoolingSteps++;
}
inline void StandardUpdate(GaugeLinks* U, ElectricFields* E) {   // :150-165
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_COOL), e = d.field(E, TCYM_E_COOL);
    d.will_write_links(u, U);
    d.will_write_field(e, E);
    d.check(tcym_flow(d.ctx, u, e, 1, dtau(), 0, nullptr));
    d.wrote_links(u, U);
    d.wrote_field(e, E);
    advance(U->a);
}
inline void CalibrationUpdate(GaugeLinks* U, ElectricFields* E) {   // :168-188
    auto& d = tcym_dropin::dev(U);
    double dn = 0.0;
    const int u = d.links(U, TCYM_U_COOL), e = d.field(E, TCYM_E_COOL);
    d.will_write_links(u, U);
    d.will_write_field(e, E);
    d.check(tcym_flow(d.ctx, u, e, 1, dtau(), 1, &dn));
    ChernSimonsNumber::DeltaNCsCooling += dn;
    d.wrote_links(u, U);
    d.wrote_field(e, E);
    advance(U->a);
}
}  // namespace GradientFlow

namespace Cooling {
inline INT BlockFrequency = 0;                  // GRADIENTFLOW/Cooling.cpp:22 (0 never blocks)
namespace DynamicFields {
inline INT BlockingNumber = 0;                  // :14
}

// the step after which the next blocking is due, or 0 (Cooling.cpp:123)
inline int next_block(int k, int maxb, int nb) {
    if (BlockFrequency <= 0 || nb >= maxb) return 0;
    long every = BlockFrequency;
    for (int l = 0; l < nb; l++) every *= 8;
    return (int)(((long)k/every + 1)*every);
}

// Cooling::Perform with Calibrate = 0 on device slot `slot` from flow step
// step0 to nsteps (GRADIENTFLOW/Cooling.cpp:69-147), a the spacing at nb
// blockings: the cooling continues in pieces so that the cooled links are
// checked for unitarity every 100 steps as the reference does (:113-119);
// GradientFlow::CoolingSteps / cTau advance step by step.
inline int perform(tcym_dropin::Device& d, int slot, const DOUBLE* a0, int step0, int nsteps, int maxb, int nb) {
    DOUBLE a[3] = {a0[0], a0[1], a0[2]};
    for (int k = step0; k < nsteps;) {
        const int k1 = (k/100 + 1)*100 < nsteps ? (k/100 + 1)*100 : nsteps;
        int nbk = nb;
        d.check(tcym_cool_continue(d.ctx, slot, k, k1, GradientFlow::dtau(), BlockFrequency, maxb, &nbk));
        for (; k < k1; k++) {
            GradientFlow::advance(a);
            if (next_block(k, maxb, nb) == k + 1) {
                std::cerr << "#BLOCKING LINKS AT COOLING TIME " << GradientFlow::CoolingTime() << std::endl;
                nb++;
                for (int m = 0; m < 3; m++) a[m] *= 2;   // Block(GaugeLinks**) (LATTICE/GaugeLinks.cpp:125-151)
                std::cerr << "#" << nb << " BLOCKING COMPLETED" << std::endl;
            }
        }
        if (nbk != nb) throw tcym_dropin::Error(TCYM_E_ARG, "Cooling::Perform: blocking count mismatch");
        if (k%100 == 0) Observables::Unitarity::CheckSlot(d, slot, GradientFlow::CoolingTime());
    }
    return nb;
}

// Cooling::CoolNSave (GRADIENTFLOW/Cooling.cpp:156-179): the hot links are
// cooled in the slot of Cooling::DynamicFields::U (the cooling method's UNew
// in its own slot), and UCold receives them.  With blocking UCold becomes the
// coarse lattice, as the reference's Copy(UCold, U) sets its extents,
// spacings and volume (LATTICE/GaugeLinks.cpp:91-107); its array must hold
// the coarse links, which the fine allocation does.
inline void CoolNSaveSlot(DOUBLE MaxCoolingSteps, INT MaxBlockingNumber, GaugeLinks* UHot, GaugeLinks* UCold, int slot) {
    auto& d = tcym_dropin::dev(UHot);
    const int u = d.links(UHot, TCYM_U_HOT);
    d.will_write_links(slot, UCold);
    GradientFlow::Reset();
    d.check(tcym_copy_links(d.ctx, slot, u));
    const int nb = perform(d, slot, UHot->a, 0, (int)MaxCoolingSteps, MaxBlockingNumber, 0);
    DynamicFields::BlockingNumber = nb;
    const int f = 1 << nb;
    for (int m = 0; m < 3; m++) { UCold->N[m] = UHot->N[m]/f; UCold->a[m] = UHot->a[m]*f; }
    UCold->Volume = UCold->N[0]*UCold->N[1]*UCold->N[2];
    UCold->aCube = UCold->a[0]*UCold->a[1]*UCold->a[2];
    d.wrote_links(slot, UCold);
}
inline void CoolNSave(std::string, DOUBLE MaxCoolingSteps, INT MaxBlockingNumber, GaugeLinks* UHot, GaugeLinks* UCold) {
    CoolNSaveSlot(MaxCoolingSteps, MaxBlockingNumber, UHot, UCold, TCYM_U_COOL);
}

// Cooling::ContinueCooling on device slot src of context d, whose spacing is
// a (GRADIENTFLOW/Cooling.cpp:181-195): E of the working copy zero, calibration
// flow from the current GradientFlow::CoolingSteps up to MaxCoolingSteps,
// blocking counted on from DynamicFields::BlockingNumber, the unitarity check
// every 100 steps, and the log "CoolingTime MAGNETIC ELECTRIC DeltaNCsCooling
// VacEst" written to IO::OutputDirectory + fname + "ID" + MySEED + ".log"
// (Cooling.cpp:69-111).
inline void ContinueCoolingSlot(const std::string& fname, DOUBLE MaxCoolingSteps, INT MaxBlockingNumber,
                                tcym_dropin::Device& d, int src, const DOUBLE* a0) {
    // with no steps left Perform still logs the starting row (Cooling.cpp:85-90)
    const int step0 = GradientFlow::CoolingSteps, nsteps = (int)MaxCoolingSteps > step0 ? (int)MaxCoolingSteps : step0;
    // the working slots are overwritten: save whatever host objects they held
    d.will_write_links(TCYM_U_COOL, nullptr);
    d.will_write_field(TCYM_E_COOL, nullptr);
    std::ostringstream name;
    name << IO::OutputDirectory << fname << "ID" << RandomNumberGenerator::MySEED << ".log";
    std::ofstream out(name.str().c_str());
    DOUBLE a[3] = {a0[0], a0[1], a0[2]};
    int nb = DynamicFields::BlockingNumber;
    double* log = new double[(size_t)5*101];
    for (int k = step0, first = 1; k < nsteps || first;) {
        const int k1 = (k/100 + 1)*100 < nsteps ? (k/100 + 1)*100 : nsteps;
        int nbk = nb;
        double dn = 0.0;
        const int r = tcym_continue_cooling(d.ctx, first ? src : TCYM_U_COOL, TCYM_U_COOL, TCYM_E_COOL, k, k1,
                                            GradientFlow::dtau(), GradientFlow::cTau, BlockFrequency, MaxBlockingNumber,
                                            &nbk, &dn, log);
        if (r != TCYM_OK) { delete[] log; d.check(r); }
        const double n0c = ChernSimonsNumber::DeltaNCsCooling;
        for (int j = first ? 0 : 1; j <= k1 - k; j++) {   // each piece logs its start again: skip it
            const double* w = log + 5*j;
            out << w[0] << " " << w[1] << " " << w[2] << " " << n0c + w[3] << " " << w[4] << std::endl;
        }
        ChernSimonsNumber::DeltaNCsCooling += dn;
        for (; k < k1; k++) {
            GradientFlow::advance(a);
            if (next_block(k, MaxBlockingNumber, nb) == k + 1) {
                std::cerr << "#BLOCKING LINKS AT COOLING TIME " << GradientFlow::CoolingTime() << std::endl;
                nb++;
                for (int m = 0; m < 3; m++) a[m] *= 2;
                std::cerr << "#" << nb << " BLOCKING COMPLETED" << std::endl;
            }
        }
        if (nbk != nb) { delete[] log; throw tcym_dropin::Error(TCYM_E_ARG, "Cooling::ContinueCooling: blocking count mismatch"); }
        if (k1 > step0 && k%100 == 0) Observables::Unitarity::CheckSlot(d, TCYM_U_COOL, GradientFlow::CoolingTime());
        first = 0;
    }
    delete[] log;
    DynamicFields::BlockingNumber = nb;
}
inline void ContinueCooling(std::string fname, DOUBLE MaxCoolingSteps, INT MaxBlockingNumber, GaugeLinks* UHot) {
    auto& d = tcym_dropin::dev(UHot);
    ContinueCoolingSlot(fname, MaxCoolingSteps, MaxBlockingNumber, d, d.links(UHot, TCYM_U_AUX), UHot->a);
}
}  // namespace Cooling

namespace ChernSimonsNumber {
namespace CoolingMethod {
inline INT StandardBlockingLevel = 0, CalibrationBlockingLevel = 0;       // TOPOLOGY/CoolingMethod.cpp:5
inline INT StandardCoolingMaxSteps = 0, CalibrationCoolingMaxSteps = 0;   // :7
inline GaugeLinks* UOld = nullptr;                                        // :13-15
inline GaugeLinks* UMid = nullptr;
inline GaugeLinks* UNew = nullptr;
inline ElectricFields* EMid = nullptr;                                    // :17
// CoolingMethod::Init (:19-29): the cooled copies at StandardBlockingLevel
inline void Init() {
    const INT f = 1 << StandardBlockingLevel;
    const GaugeLinks* U = GLinks::U;
    UOld = new GaugeLinks(U->N[0]/f, U->N[1]/f, U->N[2]/f, U->a[0]*f, U->a[1]*f, U->a[2]*f);
    UMid = new GaugeLinks(U->N[0]/f, U->N[1]/f, U->N[2]/f, U->a[0]*f, U->a[1]*f, U->a[2]*f);
    UNew = new GaugeLinks(U->N[0]/f, U->N[1]/f, U->N[2]/f, U->a[0]*f, U->a[1]*f, U->a[2]*f);
    const ElectricFields* E = EFields::E;
    EMid = new ElectricFields(E->N[0]/f, E->N[1]/f, E->N[2]/f, E->a[0]*f, E->a[1]*f, E->a[2]*f);
}
inline tcym_dropin::Device& device() {
    if (!UOld) throw tcym_dropin::Error(TCYM_E_ARG, "ChernSimonsNumber::CoolingMethod::Init was not called");
    return tcym_dropin::dev(GLinks::U);
}
// Copy(UOld, UNew) on the device (:137,163)
inline void keep_new(tcym_dropin::Device& d) {
    d.will_write_links(TCYM_U_OLD, UOld);
    d.check(tcym_copy_links(d.ctx, TCYM_U_OLD, TCYM_U_NEW));
    for (int m = 0; m < 3; m++) { UOld->N[m] = UNew->N[m]; UOld->a[m] = UNew->a[m]; }
    UOld->Volume = UNew->Volume; UOld->aCube = UNew->aCube;
    d.wrote_links(TCYM_U_OLD, UOld);
}
// CoolingMethod::PerformInterpolation(c) (:37-75): UMid and EMid on the geodesic
inline void PerformInterpolation(DOUBLE c) {
    auto& d = device();
    d.will_write_links(TCYM_U_MID, UMid);
    d.will_write_field(TCYM_E_MID, EMid);
    d.check(tcym_geodesic(d.ctx, TCYM_U_OLD, TCYM_U_NEW, TCYM_U_MID, TCYM_E_MID, c));
    d.wrote_links(TCYM_U_MID, UMid);
    d.wrote_field(TCYM_E_MID, EMid);
}
// CoolingMethod::GetDeltaNCs(1) (:80-121), leaving UMid, EMid at c = 1/2 as the reference does
inline DOUBLE GetDeltaNCs(INT NOrder) {
    if (NOrder != 1) throw tcym_dropin::Error(TCYM_E_UNSUPPORTED, "CoolingMethod::GetDeltaNCs: order 1 only (every caller)");
    auto& d = device();
    double dn = 0.0;
    d.will_write_links(TCYM_U_MID, UMid);
    d.will_write_field(TCYM_E_MID, EMid);
    d.check(tcym_delta_ncs_cooling(d.ctx, TCYM_U_OLD, TCYM_U_NEW, &dn));
    d.wrote_links(TCYM_U_MID, UMid);
    d.wrote_field(TCYM_E_MID, EMid);
    return dn;
}
// CoolingMethod::Start (:127-137): GLinks::U cooled into UNew, UOld <- UNew
inline void Start() {
    auto& d = device();
    std::cerr << "#BEGINNING INITIAL COOLING AT T=" << Dynamics::Time() << std::endl;
    Cooling::CoolNSaveSlot(StandardCoolingMaxSteps, StandardBlockingLevel, GLinks::U, UNew, TCYM_U_NEW);
    keep_new(d);
}
// CoolingMethod::Update(Measure) (:143-165)
inline void Update(INT Measure) {
    auto& d = device();
    DeltaNCsCooling = 0.0;
    Cooling::CoolNSaveSlot(StandardCoolingMaxSteps, StandardBlockingLevel, GLinks::U, UNew, TCYM_U_NEW);
    if (Measure == 1) DeltaNCsCoolRealTime += GetDeltaNCs(1);
    keep_new(d);
}
// CoolingMethod::Calibrate (:171-196): cool UNew on to the vacuum with CalibrationUpdate
inline void Calibrate() {
    auto& d = device();
    std::cerr << "#CALIBRATING COOLING AT T=" << Dynamics::Time() << " Tc=" << GradientFlow::CoolingTime() << std::endl;
    DeltaNCsCooling = 0.0;
    DeltaNCsPreviousCalibration = DeltaNCsCalibration;
    DeltaNCsRealTimePreviousCalibration = DeltaNCsRealTimeCalibration;
    std::ostringstream fname;
    fname << "VacuumCoolingT" << Dynamics::Time();
    Cooling::ContinueCoolingSlot(fname.str(), StandardCoolingMaxSteps + CalibrationCoolingMaxSteps,
                                 CalibrationBlockingLevel, d, TCYM_U_NEW, UNew->a);
    DeltaNCsCalibration = DeltaNCsCooling;
    DeltaNCsRealTimeCalibration = DeltaNCsCoolRealTime;
    std::cerr << "#CALIBRATION COMPLETED" << std::endl;
}
}  // namespace CoolingMethod
}  // namespace ChernSimonsNumber

// ---------------------------------------------------------------- E.B map
namespace Observables {
namespace EdotBMap {
// "x y z EdotB" for the sites of the box, to IO::OutputDirectory + fname +
// "ID" + MySEED + ".txt" (OBSERVABLES/EdotBMap.cpp:20-335); the device
// evaluates every site and the box selects the lines
inline void CreateMap(std::string fname, INT xLow, INT xHigh, INT yLow, INT yHigh, INT zLow, INT zHigh, GaugeLinks* U,
                      ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    double* m = new double[(size_t)U->Volume];
    const int r = tcym_edotb_map(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), m);
    if (r != TCYM_OK) { delete[] m; d.check(r); }
    std::ostringstream name;
    name << IO::OutputDirectory << fname << "ID" << RandomNumberGenerator::MySEED << ".txt";
    std::ofstream out(name.str().c_str());
    for (INT z = zLow; z <= zHigh; z++)
        for (INT y = yLow; y <= yHigh; y++)
            for (INT x = xLow; x <= xHigh; x++) out << x << " " << y << " " << z << " " << m[U->Index3D(x, y, z)] << std::endl;
    delete[] m;
    std::cerr << "#E.B MAP CREATED AT T=" << Dynamics::Time() << std::endl;
}
inline void CreateMap(std::string fname, GaugeLinks* U, ElectricFields* E) {   // :338-342
    CreateMap(fname, 0, U->N[0] - 1, 0, U->N[1] - 1, 0, U->N[2] - 1, U, E);
}
inline void CreateMap(std::string fname) { CreateMap(fname, GLinks::U, EFields::E); }   // :344-348
}  // namespace EdotBMap
}  // namespace Observables

namespace GaussLawRestoration {
inline DOUBLE GlobalMaxViolation = 0.0;
inline INT Iterations = 0;
inline void Restore(GaugeLinks* U, ElectricFields* E) {   // GAUSSLAWFIXING/GaussRestore.cpp:131-167
    auto& d = tcym_dropin::dev(U);
    int it = 0;
    double fm = 0.0;
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.will_write_field(e, E);
    int r = tcym_gauss_restore(d.ctx, u, e, 10000, 1e-14, &it, &fm);
    if (r != TCYM_OK && r != TCYM_E_NOCONV) d.check(r);
    Iterations = it;
    GlobalMaxViolation = fm;
    d.wrote_field(e, E);
}
inline void Restore() { Restore(GLinks::U, EFields::E); }   // :169-171
}  // namespace GaussLawRestoration

// ---------------------------------------------------------------- text configurations
namespace IO {
inline std::string InputDirectory;              // IO/InputManagement.cpp:3
inline std::string InputUFile, InputEFile;      // :16-17
template <typename A>
inline void SetInputDirectory(A x) {            // :5-12
    std::ostringstream s;
    s << x << "/";
    InputDirectory = s.str();
    std::cerr << "#INPUT DIRECTORY IS " << InputDirectory << std::endl;
}
template <typename A, typename B>
inline void SetInputFile(A x, B y) {            // :19-28
    std::ostringstream u, e;
    u << "UOutT" << x << "ID" << y << ".txt";
    e << "EOutT" << x << "ID" << y << ".txt";
    InputUFile = u.str();
    InputEFile = e.str();
    std::cerr << "#INPUT GAUGE LINK FILE IS " << InputUFile << std::endl;
    std::cerr << "#INPUT ELECTRIC FIELD FILE IS " << InputEFile << std::endl;
}
inline void io_check(int r) {
    if (r != TCYM_OK) throw tcym_dropin::Error(r, std::string("tcym: ") + tcym_io_last_error());
}
// IO::SaveConfiguration (IO/SaveConfiguration.cpp:3-52): OutputDirectory +
// name + "ID" + MySEED + ".txt", 15 significant digits
inline void SaveConfiguration(std::string Ufname, std::string Efname, GaugeLinks* U, ElectricFields* E) {
    tcym_dropin::SyncToHost();
    std::ostringstream uf, ef;
    uf << OutputDirectory << Ufname << "ID" << RandomNumberGenerator::MySEED << ".txt";
    ef << OutputDirectory << Efname << "ID" << RandomNumberGenerator::MySEED << ".txt";
    const int nc = tcym_dropin::config().nc;
    io_check(tcym_write_text_links(nc, U->N, U->U, uf.str().c_str(), 15, 0));
    io_check(tcym_write_text_efield(nc, E->N, E->E, ef.str().c_str(), 15, 0));
}
inline void SaveConfiguration(std::string Ufname, std::string Efname) {   // :58-60
    SaveConfiguration(Ufname, Efname, GLinks::U, EFields::E);
}
// IO::LoadConfiguration (IO/LoadConfiguration.cpp:3-136): InputDirectory +
// InputUFile / InputEFile into the host fields, then the initial energy
// densities and pressures on stderr
inline void LoadConfiguration(GaugeLinks* U, ElectricFields* E) {
    const int nc = tcym_dropin::config().nc;
    io_check(tcym_read_text_links(nc, U->N, U->U, (InputDirectory + InputUFile).c_str(), 0));
    io_check(tcym_read_text_efield(nc, E->N, E->E, (InputDirectory + InputEFile).c_str(), 0));
    for (auto& kv : tcym_dropin::Device::registry()) { kv.second->forget(U, false); kv.second->forget(E, false); }
    Observables::Bulk::Update(U, E);
    std::cerr << "#INITIAL ENERGY DENSITIES AND PRESSURES -- T00 TXX TYY TZZ" << std::endl;
    std::cerr << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " " << Observables::Bulk::TYY() << " "
              << Observables::Bulk::TZZ() << std::endl;
}
}  // namespace IO

// ---------------------------------------------------------------- initial conditions
namespace InitialConditions {
// InitialConditions::SetZero (INITIALCONDITIONS/SetZero.cpp:3-31): unit links, zero E
inline void SetZero(GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    d.will_write_links(TCYM_U_HOT, U);
    d.will_write_field(TCYM_E_HOT, E);
    d.check(tcym_unit_links(d.ctx, TCYM_U_HOT));
    d.check(tcym_zero_efield(d.ctx, TCYM_E_HOT));
    d.wrote_links(TCYM_U_HOT, U);
    d.wrote_field(TCYM_E_HOT, E);
}
inline void SetZero() {
    std::cerr << "#SETTING ZERO FIELD INITIAL CONDITIONS" << std::endl;
    SetZero(GLinks::U, EFields::E);
}
// InitialConditions::SetQuasiParticles (INITIALCONDITIONS/SetQuasiParticles.cpp:24-174)
// with the driver's Qs and n0: the modes, inverse FFT and exponentials on the
// device, then the reference's observable printout and Gauss-law restoration;
// the complex normals use RandomNumberGenerator::MySEED as their Philox key
inline void SetQuasiParticles(GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    std::cerr << "#SETTING QUASI-PARTICLE INITIAL CONDITIONS" << std::endl;
    d.will_write_links(TCYM_U_HOT, U);
    d.will_write_field(TCYM_E_HOT, E);
    d.check(tcym_quasi_particles(d.ctx, TCYM_U_HOT, TCYM_E_HOT, Qs, n0, (uint64_t)RandomNumberGenerator::MySEED));
    d.wrote_links(TCYM_U_HOT, U);
    d.wrote_field(TCYM_E_HOT, E);
    Observables::Bulk::Update(U, E);
    std::cerr << "#INITIAL ENERGY DENSITIES AND PRESSURES -- T00 TXX TYY TZZ" << std::endl;
    std::cerr << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " " << Observables::Bulk::TYY() << " "
              << Observables::Bulk::TZZ() << std::endl;
    GaussLawRestoration::Restore(U, E);
    Observables::Bulk::Update(U, E);
    std::cerr << "#INITIAL ENERGY DENSITIES AND PRESSURES -- T00 TXX TYY TZZ" << std::endl;
    std::cerr << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " " << Observables::Bulk::TYY() << " "
              << Observables::Bulk::TZZ() << std::endl;
}
inline void SetQuasiParticles() { SetQuasiParticles(GLinks::U, EFields::E); }   // :168-174
}  // namespace InitialConditions

namespace ThermalDynamics {
// ThermalDynamics::Thermalize(MaxSteps, MaxTime) (DYNAMICS/ThermalDynamics.cpp:260-316):
// the ThermalEnergy rows go to OutputDirectory + "ThermalEnergyID" + MySEED +
// ".txt"; the thermal E draws use MySEED as their Philox key
inline void Thermalize(INT MaxSteps, DOUBLE MaxTime, GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.will_write_links(u, U);
    d.will_write_field(e, E);
    const int cap = MaxSteps*((int)(MaxTime/dTau)/20 + 2);
    double* log = new double[7*(size_t)(cap > 0 ? cap : 1)];
    int n = 0;
    const int r = tcym_thermalize(d.ctx, u, e, MaxSteps, MaxTime, dTau, beta,
                                  (uint64_t)RandomNumberGenerator::MySEED, 10000, 1e-14, log, cap, &n);
    if (r != TCYM_OK) { delete[] log; d.check(r); }
    std::ostringstream name;
    name << IO::OutputDirectory << "ThermalEnergyID" << RandomNumberGenerator::MySEED << ".txt";
    std::ofstream out(name.str().c_str());
    for (int k = 0; k < n && k < cap; k++) {
        const double* w = log + 7*k;
        out << w[0] << " " << w[1] << " " << w[2] << " " << w[3] << " " << w[4] << " " << w[5] << " " << w[6] << std::endl;
    }
    delete[] log;
    d.wrote_links(u, U);
    d.wrote_field(e, E);
    std::cerr << "#CHECKING INITIAL STATE OBSERVABLES" << std::endl;
    Observables::GaussLaw::CheckViolation(U, E);
    Observables::Unitarity::CheckViolation(U);
    Observables::Bulk::Update(U, E);
    std::cerr << "#INITIAL ENERGY DENSITIES AND PRESSURES -- T00 TXX TYY TZZ" << std::endl;
    std::cerr << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " " << Observables::Bulk::TYY() << " "
              << Observables::Bulk::TZZ() << std::endl;
}
inline void Thermalize(INT MaxSteps, DOUBLE MaxTime) { Thermalize(MaxSteps, MaxTime, GLinks::U, EFields::E); }
}  // namespace ThermalDynamics
