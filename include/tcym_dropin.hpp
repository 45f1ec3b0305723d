// tcym_dropin.hpp -- the reference's hot-path entry points, with their
// original namespaces, names and signatures, over GaugeLinks/ElectricFields
// host containers in the reference's own memory layout, forwarding to the
// MI355X library through the C ABI of tcym.h.
//
// A program written against TopologicalCYMSolver's
//   Dynamics::Update(GaugeLinks*, ElectricFields*)           DYNAMICS/Dynamics.cpp:204
//   LangevinDynamics::UpdateFields(GaugeLinks*, ElectricFields*)  DYNAMICS/LangevinDynamics.cpp:300
//   GradientFlow::StandardUpdate / CalibrationUpdate           GRADIENTFLOW/GradientFlow.cpp:150,168
//   Cooling::CoolNSave / ContinueCooling                       GRADIENTFLOW/Cooling.cpp:156,181
//   ChernSimonsNumber::NCsDot                                  TOPOLOGY/ChernSimonsDerivative.cpp:31
//   ChernSimonsNumber::CoolingMethod::{Start,Update,Calibrate} TOPOLOGY/CoolingMethod.cpp:127,143,171
//   ImprovedOperators::ComputeEDotB                            CALC/ChernSimonsOperators.cpp:384
//   Observables::{Bulk,GaussLaw,Unitarity}, EdotBMap::CreateMap OBSERVABLES/*.cpp
//   GaussLawRestoration::Restore                               GAUSSLAWFIXING/GaussRestore.cpp:131
//   IO::SaveConfiguration / LoadConfiguration                  IO/SaveConfiguration.cpp:3, IO/LoadConfiguration.cpp:3
//   InitialConditions::SetZero, ThermalDynamics::Thermalize    INITIALCONDITIONS/SetZero.cpp:3, DYNAMICS/ThermalDynamics.cpp:260
//   InitialConditions::SetQuasiParticles                       INITIALCONDITIONS/SetQuasiParticles.cpp:174
// includes this header instead of those sources and links libtcym.so.
//
// Group and precision are run-time values (tcym_dropin::Configure) instead of
// the SU_Nc_FLAG macro.  Step sizes that the reference hard-codes
// (Dynamics::dTau, LangevinDynamics::dTauL, GradientFlow::SqrtDcTauOverSpacing)
// are plain variables here with the reference's default values.
//
// Residency: by default every call uploads the host fields it reads and
// downloads the ones it writes, so host arrays are always current.  Inside a
// tcym_dropin::Resident scope the fields stay on the device between calls and
// are downloaded when the scope ends (or by tcym_dropin::SyncToHost()).
// Failures throw tcym_dropin::Error instead of calling exit(0).
#pragma once

#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <fstream>
#include <iostream>
#include <sstream>

#include "tcym.h"
#include "tcym_io.h"

typedef int INT;
typedef double DOUBLE;

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

namespace Lattice {
inline DOUBLE& aScale() { static DOUBLE s = 1.0; return s; }   // LATTICE/3DGrid.cpp:18
}

namespace tcym_dropin {

// One device context per lattice geometry.  Host objects are bound to device
// slots; a binding remembers whether the device copy is newer than the host.
class Device {
public:
    static std::map<std::string, Device*>& registry() { static std::map<std::string, Device*> m; return m; }
    static std::string key(const INT* N, const DOUBLE* a) {
        char k[160];
        snprintf(k, sizeof k, "%d %d %d %d %d %.17g %.17g %.17g %.17g", config().nc, config().precision,
                 N[0], N[1], N[2], a[0], a[1], a[2], Lattice::aScale());
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
    void check(int r) const {
        if (r != TCYM_OK) throw Error(r, std::string("tcym: ") + tcym_last_error(ctx));
    }
    // bind host links to a device slot.  Resident: reuse any slot that holds a
    // current copy of the object; otherwise upload into `slot` (first saving a
    // device-newer object that the slot held).
    int links(GaugeLinks* U, int slot) {
        if (resident)
            for (int s = 0; s < TCYM_NUM_U; s++) if (ub[s].host == U && ub[s].valid) return s;
        flush_links(slot);
        for (int s = 0; s < TCYM_NUM_U; s++) if (ub[s].host == U) { flush_links(s); ub[s] = Bind(); }
        check(tcym_set_links(ctx, slot, U->U));
        ub[slot].host = U; ub[slot].valid = true; ub[slot].dirty = false;
        return slot;
    }
    int field(ElectricFields* E, int slot) {
        if (resident)
            for (int s = 0; s < TCYM_NUM_E; s++) if (eb[s].host == E && eb[s].valid) return s;
        flush_field(slot);
        for (int s = 0; s < TCYM_NUM_E; s++) if (eb[s].host == E) { flush_field(s); eb[s] = Bind(); }
        check(tcym_set_efield(ctx, slot, E->E));
        eb[slot].host = E; eb[slot].valid = true; eb[slot].dirty = false;
        return slot;
    }
    // the device copy was modified: download now unless resident
    void wrote_links(int slot) {
        ub[slot].dirty = true;
        if (!resident) flush_links(slot);
    }
    void wrote_field(int slot) {
        eb[slot].dirty = true;
        if (!resident) flush_field(slot);
    }
    void flush() {
        for (int s = 0; s < TCYM_NUM_U; s++) flush_links(s);
        for (int s = 0; s < TCYM_NUM_E; s++) flush_field(s);
    }
    // the device slot is about to be overwritten: download a device-newer host
    // object it held and forget the binding
    void release_links(int s) { flush_links(s); ub[s] = Bind(); }
    void release_field(int s) { flush_field(s); eb[s] = Bind(); }
    void invalidate() {
        for (auto& b : ub) b.valid = false;
        for (auto& b : eb) b.valid = false;
    }
    static bool& resident_flag() { static bool r = false; return r; }
    bool resident = false;
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
        p.aScale = Lattice::aScale();
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
// download every device-newer field of the lattice's context to its host object
inline void SyncToHost(GaugeLinks* U) { if (Device* d = Device::find(U->N, U->a)) d->flush(); }
// host arrays were modified by the caller: re-upload on next use
inline void HostModified(GaugeLinks* U) { if (Device* d = Device::find(U->N, U->a)) d->invalidate(); }

}  // namespace tcym_dropin

// ---------------------------------------------------------------- dynamics
namespace Dynamics {
inline DOUBLE& dTau() { static DOUBLE v = 0.025; return v; }   // DYNAMICS/Dynamics.cpp:17
inline INT& tSteps() { static INT v = 0; return v; }
inline DOUBLE Time(DOUBLE Qs = 1.0) { return Qs*(tSteps()*dTau()); }   // :20-22
inline void Reset() { tSteps() = 0; }
inline void UpdateElectricFields(INT, INT, INT, INT, INT, INT, GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.check(tcym_update_efield(d.ctx, u, e, dTau()));
    d.wrote_field(e);
}
inline void UpdateGaugeLinks(INT, INT, INT, INT, INT, INT, GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.check(tcym_update_links(d.ctx, u, e, dTau()));
    d.wrote_links(u);
}
inline void Update(GaugeLinks* U, ElectricFields* E) {   // DYNAMICS/Dynamics.cpp:204-218
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.check(tcym_leapfrog(d.ctx, u, e, 1, dTau()));
    d.wrote_links(u);
    d.wrote_field(e);
    tSteps()++;
}
}  // namespace Dynamics

namespace ThermalDynamics {
inline DOUBLE& beta() { static DOUBLE v = 2.0; return v; }      // DYNAMICS/ThermalDynamics.cpp:10
inline DOUBLE& dTau() { static DOUBLE v = 0.0375; return v; }   // :23
}

namespace LangevinDynamics {
inline DOUBLE& dTauL() { static DOUBLE v = 0.001; return v; }   // DYNAMICS/LangevinDynamics.cpp:41
inline DOUBLE& beta() { return ThermalDynamics::beta(); }       // the Langevin temperature is ThermalDynamics::beta
inline DOUBLE& damping() { static DOUBLE v = 1.0; return v; }   // fixed at 1 in the reference
inline INT& tauLSteps() { static INT v = 0; return v; }
inline uint64_t& seed() { static uint64_t v = 1; return v; }
// LangevinDynamics::UpdateFields with LangevinVariables::Xi (DYNAMICS/LangevinDynamics.cpp:300-311)
inline void UpdateFields(GaugeLinks* U, ElectricFields* E, ElectricFields* Xi) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT), x = d.field(Xi, TCYM_E_XI);
    d.check(tcym_langevin_xi(d.ctx, u, e, x, dTauL(), damping()));
    d.wrote_links(u);
    d.wrote_field(e);
    tauLSteps()++;
}
// LangevinDynamics::Update: Philox noise of variance 2 damping/beta, then UpdateFields (:320-347)
inline void Update(GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.check(tcym_langevin(d.ctx, u, e, 1, dTauL(), damping(), 1.0/beta(), seed(), (uint32_t)tauLSteps()));
    d.wrote_links(u);
    d.wrote_field(e);
    tauLSteps()++;
}
}  // namespace LangevinDynamics

namespace ChernSimonsNumber {
inline DOUBLE& DeltaNCsCooling() { static DOUBLE v = 0.0; return v; }        // TOPOLOGY/ChernSimonsDerivative.cpp:18
inline DOUBLE& DeltaNCsCoolRealTime() { static DOUBLE v = 0.0; return v; }   // :15
inline DOUBLE NCsDot(GaugeLinks* U, ElectricFields* E) {                      // :31-49
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

namespace GradientFlow {
inline DOUBLE& SqrtDcTauOverSpacing() { static DOUBLE v = 0.5/std::sqrt(2.0); return v; }   // GRADIENTFLOW/GradientFlow.cpp:15
inline DOUBLE& cTau() { static DOUBLE v = 0.0; return v; }
inline INT& CoolingSteps() { static INT v = 0; return v; }
inline void Reset() { cTau() = 0.0; CoolingSteps() = 0; }
inline DOUBLE CoolingTime() { return cTau(); }
inline DOUBLE dtau() { return SqrtDcTauOverSpacing()*SqrtDcTauOverSpacing(); }
// CoolingSteps++ and cTau += DcTau of one flow step at spacings a (:150-163)
inline void advance(const DOUBLE* a) {
    cTau() += dtau()*std::pow(a[0]*a[1]*a[2], 2.0/3.0)/(Lattice::aScale()*Lattice::aScale());
    CoolingSteps()++;
}
inline void advance(const GaugeLinks* U) { advance(U->a); }
inline void StandardUpdate(GaugeLinks* U, ElectricFields* E) {   // :150-165
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_COOL), e = d.field(E, TCYM_E_COOL);
    d.check(tcym_flow(d.ctx, u, e, 1, dtau(), 0, nullptr));
    d.wrote_links(u);
    d.wrote_field(e);
    advance(U);
}
inline void CalibrationUpdate(GaugeLinks* U, ElectricFields* E) {   // :168-188
    auto& d = tcym_dropin::dev(U);
    double dn = 0.0;
    const int u = d.links(U, TCYM_U_COOL), e = d.field(E, TCYM_E_COOL);
    d.check(tcym_flow(d.ctx, u, e, 1, dtau(), 1, &dn));
    ChernSimonsNumber::DeltaNCsCooling() += dn;
    d.wrote_links(u);
    d.wrote_field(e);
    advance(U);
}
}  // namespace GradientFlow

namespace IO {
// output directory of the calibration logs (IO/OutputManagement.cpp:3-12)
inline std::string& OutputDirectory() { static std::string d; return d; }
inline void SetOutputDirectory(const std::string& x) { OutputDirectory() = x + "/"; }
}  // namespace IO

namespace RandomNumberGenerator {
// the seed that names the calibration logs (MISC/RNG/GSLRandomNumberGenerator.cpp:9,19)
inline long& MySEED() { static long v = 0; return v; }
}

namespace ChernSimonsNumber {
inline DOUBLE& DeltaNCsCalibration() { static DOUBLE v = 0.0; return v; }                   // TOPOLOGY/ChernSimonsDerivative.cpp:25
inline DOUBLE& DeltaNCsPreviousCalibration() { static DOUBLE v = 0.0; return v; }
inline DOUBLE& DeltaNCsRealTimeCalibration() { static DOUBLE v = 0.0; return v; }           // :28
inline DOUBLE& DeltaNCsRealTimePreviousCalibration() { static DOUBLE v = 0.0; return v; }
}  // namespace ChernSimonsNumber

namespace Cooling {
// Cooling::BlockFrequency (GRADIENTFLOW/Cooling.cpp:22); 0 never blocks (the
// reference would divide by zero)
inline INT& BlockFrequency() { static INT v = 0; return v; }
// Cooling::DynamicFields::BlockingNumber: blockings of the last cooled copy (:14)
inline INT& BlockingNumber() { static INT v = 0; return v; }

// GradientFlow::CoolingSteps / cTau over flow steps step0+1 .. nsteps of
// Cooling::Perform (Cooling.cpp:92-141): step k flows at the spacing of the
// blockings done before it; a is the spacing at nb blockings.
inline void clock(const DOUBLE* a, int step0, int nsteps, int maxb, int nb) {
    DOUBLE s[3] = {a[0], a[1], a[2]};
    for (int k = step0 + 1; k <= nsteps; k++) {
        GradientFlow::advance(s);
        if (BlockFrequency() > 0 && nb < maxb) {
            long every = BlockFrequency();
            for (int l = 0; l < nb; l++) every *= 8;
            if (k % every == 0) {
                nb++;
                for (int m = 0; m < 3; m++) s[m] *= 2;   // Block(GaugeLinks**) (LATTICE/GaugeLinks.cpp:125-151)
            }
        }
    }
}

// Cooling::CoolNSave (GRADIENTFLOW/Cooling.cpp:156-179).  With blocking, UCold
// becomes the coarse lattice, as the reference's Copy(UCold, U) sets its
// extents, spacings and volume (LATTICE/GaugeLinks.cpp:91-107): its array must
// hold the coarse links, which the fine allocation does.
inline void CoolNSave(std::string, DOUBLE MaxCoolingSteps, INT MaxBlockingNumber, GaugeLinks* UHot, GaugeLinks* UCold) {
    auto& d = tcym_dropin::dev(UHot);
    GradientFlow::Reset();
    int nb = 0;
    d.check(tcym_cool_blocked(d.ctx, d.links(UHot, TCYM_U_HOT), TCYM_U_NEW, (int)MaxCoolingSteps, GradientFlow::dtau(),
                              BlockFrequency(), MaxBlockingNumber, &nb));
    clock(UHot->a, 0, (int)MaxCoolingSteps, MaxBlockingNumber, 0);
    BlockingNumber() = nb;
    d.check(tcym_get_links(d.ctx, TCYM_U_NEW, UCold->U));
    const int f = 1 << nb;
    for (int m = 0; m < 3; m++) { UCold->N[m] = UHot->N[m]/f; UCold->a[m] = UHot->a[m]*f; }
    UCold->Volume = UCold->N[0]*UCold->N[1]*UCold->N[2];
    UCold->aCube = UCold->a[0]*UCold->a[1]*UCold->a[2];
}

// Cooling::ContinueCooling on device slot src of context d, whose spacing is
// a (GRADIENTFLOW/Cooling.cpp:181-195): E of the working copy zero, calibration
// flow from the current GradientFlow::CoolingSteps up to MaxCoolingSteps,
// blocking counted on from BlockingNumber, and the log
// "CoolingTime MAGNETIC ELECTRIC DeltaNCsCooling VacEst" written to
// IO::OutputDirectory + fname + "ID" + MySEED + ".log" (Cooling.cpp:69-111).
inline void ContinueCoolingSlot(const std::string& fname, DOUBLE MaxCoolingSteps, INT MaxBlockingNumber,
                                tcym_dropin::Device& d, int src, const DOUBLE* a) {
    const int step0 = GradientFlow::CoolingSteps(), nsteps = (int)MaxCoolingSteps;
    if (nsteps < step0) return;
    // the working slots are overwritten: save and release whatever host objects they held
    d.release_links(TCYM_U_COOL);
    d.release_field(TCYM_E_COOL);
    const int nb0 = BlockingNumber();
    int nb = nb0;
    double dn = 0.0;
    double* log = new double[(size_t)5*(nsteps - step0 + 1)];
    const int r = tcym_continue_cooling(d.ctx, src, TCYM_U_COOL, TCYM_E_COOL, step0, nsteps, GradientFlow::dtau(),
                                        GradientFlow::cTau(), BlockFrequency(), MaxBlockingNumber, &nb, &dn, log);
    if (r != TCYM_OK) { delete[] log; d.check(r); }
    std::ostringstream name;
    name << IO::OutputDirectory() << fname << "ID" << RandomNumberGenerator::MySEED() << ".log";
    std::ofstream out(name.str().c_str());
    const double n0 = ChernSimonsNumber::DeltaNCsCooling();
    for (int j = 0; j <= nsteps - step0; j++) {
        const double* w = log + 5*j;
        out << w[0] << " " << w[1] << " " << w[2] << " " << n0 + w[3] << " " << w[4] << std::endl;
    }
    delete[] log;
    clock(a, step0, nsteps, MaxBlockingNumber, nb0);
    BlockingNumber() = nb;
    ChernSimonsNumber::DeltaNCsCooling() += dn;
}
inline void ContinueCooling(std::string fname, DOUBLE MaxCoolingSteps, INT MaxBlockingNumber, GaugeLinks* UHot) {
    auto& d = tcym_dropin::dev(UHot);
    d.release_links(TCYM_U_COOL);
    ContinueCoolingSlot(fname, MaxCoolingSteps, MaxBlockingNumber, d, d.links(UHot, TCYM_U_AUX), UHot->a);
}
}  // namespace Cooling

namespace ChernSimonsNumber {
namespace CoolingMethod {
inline INT& StandardCoolingMaxSteps() { static INT v = 128; return v; }   // SphaleronSimulation.cpp:230
inline INT& CalibrationCoolingMaxSteps() { static INT v = 0; return v; }  // TOPOLOGY/CoolingMethod.cpp:7
inline INT& StandardBlockingLevel() { static INT v = 0; return v; }       // :5
inline INT& CalibrationBlockingLevel() { static INT v = 0; return v; }
// Cooling::CoolNSave(GLinks::U -> UNew) on the device (:133,152)
inline void cool_new(tcym_dropin::Device& d, GaugeLinks* UHot) {
    GradientFlow::Reset();
    int nb = 0;
    d.check(tcym_cool_blocked(d.ctx, d.links(UHot, TCYM_U_HOT), TCYM_U_NEW, StandardCoolingMaxSteps(), GradientFlow::dtau(),
                              Cooling::BlockFrequency(), StandardBlockingLevel(), &nb));
    Cooling::clock(UHot->a, 0, StandardCoolingMaxSteps(), StandardBlockingLevel(), 0);
    Cooling::BlockingNumber() = nb;
}
// CoolingMethod::Start (TOPOLOGY/CoolingMethod.cpp:127-137): cools GLinks::U into UNew, UOld <- UNew
inline void Start(GaugeLinks* UHot) {
    auto& d = tcym_dropin::dev(UHot);
    cool_new(d, UHot);
    d.check(tcym_copy_links(d.ctx, TCYM_U_OLD, TCYM_U_NEW));
}
// CoolingMethod::Update(Measure) (:143-165)
inline void Update(GaugeLinks* UHot, INT Measure) {
    auto& d = tcym_dropin::dev(UHot);
    DeltaNCsCooling() = 0.0;
    cool_new(d, UHot);
    if (Measure == 1) {
        double dn = 0.0;
        d.check(tcym_delta_ncs_cooling(d.ctx, TCYM_U_OLD, TCYM_U_NEW, &dn));
        DeltaNCsCoolRealTime() += dn;
    }
    d.check(tcym_copy_links(d.ctx, TCYM_U_OLD, TCYM_U_NEW));
}
// CoolingMethod::Calibrate (:171-195): cool UNew (= UOld after Start/Update)
// on to the vacuum with CalibrationUpdate; UHot names the lattice
inline void Calibrate(GaugeLinks* UHot) {
    auto& d = tcym_dropin::dev(UHot);
    DeltaNCsCooling() = 0.0;
    DeltaNCsPreviousCalibration() = DeltaNCsCalibration();
    DeltaNCsRealTimePreviousCalibration() = DeltaNCsRealTimeCalibration();
    int lev = 0, n[3];
    d.check(tcym_level(d.ctx, TCYM_U_OLD, -1, &lev, n));
    DOUBLE a[3];
    for (int m = 0; m < 3; m++) a[m] = UHot->a[m]*(1 << lev);
    std::ostringstream fname;
    fname << "VacuumCoolingT" << Dynamics::Time();
    Cooling::ContinueCoolingSlot(fname.str(), StandardCoolingMaxSteps() + CalibrationCoolingMaxSteps(),
                                 CalibrationBlockingLevel(), d, TCYM_U_OLD, a);
    DeltaNCsCalibration() = DeltaNCsCooling();
    DeltaNCsRealTimeCalibration() = DeltaNCsCoolRealTime();
}
}  // namespace CoolingMethod
}  // namespace ChernSimonsNumber

namespace Observables {
namespace Bulk {
inline DOUBLE* ESqr() { static DOUBLE v[3]; return v; }
inline DOUBLE* BSqr() { static DOUBLE v[3]; return v; }
inline DOUBLE* cE() { static DOUBLE v[3]; return v; }
inline DOUBLE* cB() { static DOUBLE v[3]; return v; }
inline void Update(GaugeLinks* U, ElectricFields* E) {   // OBSERVABLES/BulkObservables.cpp:133-143
    auto& d = tcym_dropin::dev(U);
    double o[6];
    d.check(tcym_bulk(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), o));
    for (int mu = 0; mu < 3; mu++) {
        ESqr()[mu] = o[mu]; BSqr()[mu] = o[3 + mu];
        const double c = U->a[mu]*Lattice::aScale()*Lattice::aScale()/U->aCube;   // :17-25 (Minkowski)
        cE()[mu] = cB()[mu] = c*c;
    }
}
inline DOUBLE T00() { return 0.5*(cE()[0]*ESqr()[0] + cE()[1]*ESqr()[1] + cE()[2]*ESqr()[2] + cB()[0]*BSqr()[0] + cB()[1]*BSqr()[1] + cB()[2]*BSqr()[2]); }
inline DOUBLE ELECTRIC() { return 0.5*(cE()[0]*ESqr()[0] + cE()[1]*ESqr()[1] + cE()[2]*ESqr()[2]); }
inline DOUBLE MAGNETIC() { return 0.5*(cB()[0]*BSqr()[0] + cB()[1]*BSqr()[1] + cB()[2]*BSqr()[2]); }
// diagonal pressure components T_ii, in the reference's order of terms
// (OBSERVABLES/BulkObservables.cpp:33-44)
inline DOUBLE Tii(int i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    const int lo = j < k ? j : k, hi = j < k ? k : j;
    return 0.5*((cE()[lo]*ESqr()[lo] + cE()[hi]*ESqr()[hi] + cB()[lo]*BSqr()[lo] + cB()[hi]*BSqr()[hi])
                - (cE()[i]*ESqr()[i] + cB()[i]*BSqr()[i]));
}
inline DOUBLE TXX() { return Tii(0); }
inline DOUBLE TYY() { return Tii(1); }
inline DOUBLE TZZ() { return Tii(2); }
}  // namespace Bulk
namespace HardScales {
inline DOUBLE* H() { static DOUBLE v[3]; return v; }
inline DOUBLE* cE() { static DOUBLE v[3]; return v; }
inline void Update(GaugeLinks* U) {   // OBSERVABLES/HardScales.cpp:40-110
    auto& d = tcym_dropin::dev(U);
    d.check(tcym_hard_scales(d.ctx, d.links(U, TCYM_U_AUX), H()));
    for (int mu = 0; mu < 3; mu++) {
        const double c = U->a[mu]*Lattice::aScale()*Lattice::aScale()/U->aCube;   // :16-22 (Minkowski)
        cE()[mu] = c*c;
    }
}
// characteristic momentum scales (OBSERVABLES/HardScales.cpp:26-36)
inline DOUBLE LambdaXX() { return cE()[1]*H()[1] + cE()[2]*H()[2] - cE()[0]*H()[0]; }
inline DOUBLE LambdaYY() { return cE()[2]*H()[2] + cE()[0]*H()[0] - cE()[1]*H()[1]; }
inline DOUBLE LambdaZZ() { return cE()[0]*H()[0] + cE()[1]*H()[1] - cE()[2]*H()[2]; }
}  // namespace HardScales
namespace GaussLaw {
inline DOUBLE& GlobalMaxViolation() { static DOUBLE v; return v; }
inline DOUBLE& GlobalAvgViolation() { static DOUBLE v; return v; }
inline void UpdateViolation(INT, INT, INT, INT, INT, INT, GaugeLinks* U, ElectricFields* E) {   // OBSERVABLES/GaussLawViolation.cpp:13-61
    auto& d = tcym_dropin::dev(U);
    d.check(tcym_gauss(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), &GlobalMaxViolation(), &GlobalAvgViolation()));
}
}  // namespace GaussLaw
namespace Unitarity {
inline DOUBLE& GlobalMaxViolation() { static DOUBLE v; return v; }
inline DOUBLE& GlobalAvgViolation() { static DOUBLE v; return v; }
inline void UpdateViolation(INT, INT, INT, INT, INT, INT, GaugeLinks* U) {   // OBSERVABLES/UnitarityViolation.cpp:13-51
    auto& d = tcym_dropin::dev(U);
    d.check(tcym_unitarity(d.ctx, d.links(U, TCYM_U_AUX), &GlobalMaxViolation(), &GlobalAvgViolation()));
}
}  // namespace Unitarity
namespace EdotBMap {
// writes "x y z EdotB" per site (OBSERVABLES/EdotBMap.cpp:20-348, output :323)
inline void CreateMap(std::string fname, GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    double* m = new double[(size_t)U->Volume];
    d.check(tcym_edotb_map(d.ctx, d.links(U, TCYM_U_AUX), d.field(E, TCYM_E_AUX), m));
    std::ofstream out(fname.c_str());
    out.precision(15);
    for (INT z = 0; z < U->N[2]; z++)
        for (INT y = 0; y < U->N[1]; y++)
            for (INT x = 0; x < U->N[0]; x++) out << x << " " << y << " " << z << " " << m[U->Index3D(x, y, z)] << "\n";
    delete[] m;
}
}  // namespace EdotBMap
}  // namespace Observables

namespace GaussLawRestoration {
inline DOUBLE& GlobalMaxViolation() { static DOUBLE v; return v; }
inline INT& Iterations() { static INT v; return v; }
inline void Restore(GaugeLinks* U, ElectricFields* E) {   // GAUSSLAWFIXING/GaussRestore.cpp:131-167
    auto& d = tcym_dropin::dev(U);
    int it = 0;
    double fm = 0.0;
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    int r = tcym_gauss_restore(d.ctx, u, e, 10000, 1e-14, &it, &fm);
    if (r != TCYM_OK && r != TCYM_E_NOCONV) d.check(r);
    Iterations() = it;
    GlobalMaxViolation() = fm;
    d.wrote_field(e);
}
}  // namespace GaussLawRestoration

// ---------------------------------------------------------------- text configurations
namespace IO {
// IO::InputDirectory / InputUFile / InputEFile (IO/InputManagement.cpp:3-28)
inline std::string& InputDirectory() { static std::string d; return d; }
inline void SetInputDirectory(const std::string& x) { InputDirectory() = x + "/"; }
inline std::string& InputUFile() { static std::string f; return f; }
inline std::string& InputEFile() { static std::string f; return f; }
template <class A, class B>
inline void SetInputFile(A x, B y) {
    std::ostringstream u, e;
    u << "UOutT" << x << "ID" << y << ".txt";
    e << "EOutT" << x << "ID" << y << ".txt";
    InputUFile() = u.str();
    InputEFile() = e.str();
}
inline void io_check(int r) {
    if (r != TCYM_OK) throw tcym_dropin::Error(r, std::string("tcym: ") + tcym_io_last_error());
}
// IO::SaveConfiguration (IO/SaveConfiguration.cpp:3-52): OutputDirectory +
// name + "ID" + MySEED + ".txt", 15 significant digits
inline void SaveConfiguration(std::string Ufname, std::string Efname, GaugeLinks* U, ElectricFields* E) {
    tcym_dropin::SyncToHost(U);
    std::ostringstream uf, ef;
    uf << OutputDirectory() << Ufname << "ID" << RandomNumberGenerator::MySEED() << ".txt";
    ef << OutputDirectory() << Efname << "ID" << RandomNumberGenerator::MySEED() << ".txt";
    const int nc = tcym_dropin::config().nc;
    io_check(tcym_write_text_links(nc, U->N, U->U, uf.str().c_str(), 15, 0));
    io_check(tcym_write_text_efield(nc, E->N, E->E, ef.str().c_str(), 15, 0));
}
// IO::LoadConfiguration (IO/LoadConfiguration.cpp:3-136): InputDirectory +
// InputUFile / InputEFile into the host fields, then the initial energy
// densities and pressures on stderr
inline void LoadConfiguration(GaugeLinks* U, ElectricFields* E) {
    const int nc = tcym_dropin::config().nc;
    io_check(tcym_read_text_links(nc, U->N, U->U, (InputDirectory() + InputUFile()).c_str(), 0));
    io_check(tcym_read_text_efield(nc, E->N, E->E, (InputDirectory() + InputEFile()).c_str(), 0));
    tcym_dropin::HostModified(U);
    Observables::Bulk::Update(U, E);
    std::cerr << "#INITIAL ENERGY DENSITIES AND PRESSURES -- T00 TXX TYY TZZ" << std::endl;
    std::cerr << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " " << Observables::Bulk::TYY() << " "
              << Observables::Bulk::TZZ() << std::endl;
}
}  // namespace IO

// ---------------------------------------------------------------- initial conditions
namespace InitialConditions {
// the quasi-particle parameters, globals of the drivers (SphaleronSimulationSU3.cpp:76-77)
inline DOUBLE& Qs() { static DOUBLE v = 1.0; return v; }
inline DOUBLE& n0() { static DOUBLE v = 1.0; return v; }
// InitialConditions::SetQuasiParticles (INITIALCONDITIONS/SetQuasiParticles.cpp:24-174):
// the modes, inverse FFT and exponentials on the device, then the
// reference's observable printout and Gauss-law restoration; the complex
// normals use RandomNumberGenerator::MySEED as their Philox key
inline void SetQuasiParticles(GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    std::cerr << "#SETTING QUASI-PARTICLE INITIAL CONDITIONS" << std::endl;
    d.check(tcym_quasi_particles(d.ctx, u, e, Qs(), n0(), (uint64_t)RandomNumberGenerator::MySEED()));
    d.wrote_links(u);
    d.wrote_field(e);
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
// InitialConditions::SetZero (INITIALCONDITIONS/SetZero.cpp:3-24): unit links, zero E
inline void SetZero(GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    d.check(tcym_unit_links(d.ctx, u));
    d.check(tcym_zero_efield(d.ctx, e));
    d.wrote_links(u);
    d.wrote_field(e);
}
}  // namespace InitialConditions

namespace ThermalDynamics {
// ThermalDynamics::Thermalize(MaxSteps, MaxTime) on the given fields
// (DYNAMICS/ThermalDynamics.cpp:260-316): the ThermalEnergy rows go to
// OutputDirectory + "ThermalEnergyID" + MySEED + ".txt"; the thermal E draws
// use MySEED as their Philox key
inline void Thermalize(INT MaxSteps, DOUBLE MaxTime, GaugeLinks* U, ElectricFields* E) {
    auto& d = tcym_dropin::dev(U);
    const int u = d.links(U, TCYM_U_HOT), e = d.field(E, TCYM_E_HOT);
    const int cap = MaxSteps*((int)(MaxTime/dTau())/20 + 2);
    double* log = new double[7*(size_t)(cap > 0 ? cap : 1)];
    int n = 0;
    const int r = tcym_thermalize(d.ctx, u, e, MaxSteps, MaxTime, dTau(), beta(),
                                  (uint64_t)RandomNumberGenerator::MySEED(), 10000, 1e-14, log, cap, &n);
    if (r != TCYM_OK) { delete[] log; d.check(r); }
    std::ostringstream name;
    name << IO::OutputDirectory() << "ThermalEnergyID" << RandomNumberGenerator::MySEED() << ".txt";
    std::ofstream out(name.str().c_str());
    for (int k = 0; k < n && k < cap; k++) {
        const double* w = log + 7*k;
        out << w[0] << " " << w[1] << " " << w[2] << " " << w[3] << " " << w[4] << " " << w[5] << " " << w[6] << std::endl;
    }
    delete[] log;
    d.wrote_links(u);
    d.wrote_field(e);
    Observables::Bulk::Update(U, E);
    std::cerr << "#INITIAL ENERGY DENSITIES AND PRESSURES -- T00 TXX TYY TZZ" << std::endl;
    std::cerr << Observables::Bulk::T00() << " " << Observables::Bulk::TXX() << " " << Observables::Bulk::TYY() << " "
              << Observables::Bulk::TZZ() << std::endl;
}
}  // namespace ThermalDynamics
