// C ABI (include/tcym.h): device context, field slots, orchestration of the
// update/measurement kernels, deterministic reductions and kernel timing.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include <rocfft/rocfft.h>

#include <mutex>

#include "tcym.h"
#include "tcym_backend.hpp"
#include "tcym_reduce.hpp"
#include "tcym_device.hpp"

using tcym::Backend;
using tcym::Geom;
using tcym::kBlock;

namespace {

// final reduction: block q reduces quantity q of nblk partials in a fixed order
__global__ __launch_bounds__(256) void k_final(const double* __restrict__ part, long nblk, int nq, unsigned is_max,
                                               double* __restrict__ out) {
    const int q = blockIdx.x;
    const double r = tcym::final_reduce_block(part, nblk, nq, q, (is_max >> q) & 1);
    if (threadIdx.x == 0) out[q] = r;
}

// a field of nr reals of `bytes` each per (site, mu) fits the 32-bit buffer offsets
bool field_fits(const int* N, int nr, int bytes) {
    const double V = (double)tcym::padded_sites((long)N[0]*N[1]*N[2]);
    return 3.0*V*nr*bytes < 4294967296.0;
}

// hot-trajectory N_CS: acc += half_dt (NCsDot_a + NCsDot_b), NCsDot = E.B / (8 pi^2)
// (no fused multiply-add: the product is rounded before the sum, as in the host composition)
__global__ void k_hot_ncs(const double* __restrict__ eb, double* __restrict__ acc, double half_dt) {
#pragma clang fp contract(off)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double a = 0.125*eb[0]/(M_PI*M_PI), b = 0.125*eb[1]/(M_PI*M_PI);
        acc[0] += half_dt*(a + b);
    }
}

// ---- scratch of tcym_ncsdot_noalloc ----
// The block partials of the improved E.B live in this fixed device array:
// the entry point allocates nothing.  Each GPU holds its own copy at its own
// address, looked up once per device (scratch_ptr) while tcym_create has that
// device current; host code uses only that address, never the array's name.
// Calls that overlap on one GPU share the array and race.
constexpr long kScratchDoubles = 131072;   // 1 MiB: the partials of lattices up to about 220^3
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on `device` (the current device), looked up on first use
hipError_t scratch_ptr(int device, double** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = (double*)p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

// final reduction of the E.B partials, written as NCsDot = E.B / (8 pi^2)
// (TOPOLOGY/ChernSimonsDerivative.cpp:47); same sum order and rounding as
// k_final followed by the host division of tcym_ncsdot
__global__ __launch_bounds__(256) void k_final_ncsdot(const double* __restrict__ part, long nblk, double* __restrict__ out) {
    const double r = tcym::final_reduce_block(part, nblk, 1, 0, false);
    if (threadIdx.x == 0) out[0] = 0.125*r/(M_PI*M_PI);
}

}  // namespace

namespace tcym {
void final_reduce(const Launcher& L, const double* part, long nblk, int nq, const int* is_max, double* out) {
    unsigned m = 0;
    for (int q = 0; q < nq; q++) if (is_max[q]) m |= 1u << q;
    L("reduce", [&](hipStream_t st) { hipLaunchKernelGGL(k_final, dim3(nq), dim3(256), 0, st, part, nblk, nq, m, out); });
}
}  // namespace tcym

struct tcym_ctx : tcym::DeviceQueue {
    Geom g{};
    int nc = 0, prec = 0;
    Backend* be = nullptr;     // storage precision
    Backend* be64 = nullptr;   // FP64 (Gauss-law restoration of the FP32 form)
    void* U[TCYM_NUM_U] = {};
    void* E[TCYM_NUM_E] = {};
    void* Uscr = nullptr;
    double* part = nullptr;
    double* red = nullptr;       // device results
    double* red_host = nullptr;  // pinned
    double* stage = nullptr;     // AoS staging, 3 V max(ns, ne) doubles
    double* G64 = nullptr;
    double* U64 = nullptr;
    double* E64 = nullptr;
    double* map = nullptr;
    void* Bc = nullptr;          // field strength of one link slot (tcym_evolve_ncs, within a call)
    // Field strength B kept by the cooling measurement (tcym_delta_ncs_cooling)
    // for the links of slot bk_slot: valid while that slot's access count
    // uver (bumped by useU, which every reader and writer of a slot passes)
    // still equals bk_ver.  A copy of those links to another slot moves B's
    // validity with them (tcym_copy_links), so the E.B at U_OLD of the next
    // measurement reads the B formed for U_NEW by the previous one.
    void* Bcool = nullptr;
    uint64_t uver[TCYM_NUM_U] = {};
    uint64_t bk_ver = 0;
    int bk_slot = -1;
    bool kept_for(int u) const { return u >= 0 && u == bk_slot && bk_ver == uver[u]; }

    // Resolution level of each slot: 0 is the full lattice; level l is the
    // lattice blocked l times (extent N >> l, spacing a << l), which a
    // blocked cooled copy holds (Block, LATTICE/GaugeLinks.cpp:125-151).
    // A level-l field occupies the start of its slot's buffer.
    int ulev[TCYM_NUM_U] = {};
    int elev[TCYM_NUM_E] = {};
    Geom geo(int lev) const {
        Geom q = g;
        for (int k = 0; k < 3; k++) { q.n[k] = g.n[k] >> lev; q.a[k] = g.a[k]*(double)(1 << lev); }
        q.V = (long)q.n[0]*q.n[1]*q.n[2];
        return q;
    }
    // the geometry of the link slots u, u2 and field slots e, e2 (-1: unused)
    // an entry point reads together; they must share one resolution
    int at(Geom& q, int u, int e = -1, int u2 = -1, int e2 = -1) {
        int lev = -1;
        bool mixed = false;
        auto take = [&](int l) { if (lev < 0) lev = l; else if (l != lev) mixed = true; };
        if (u >= 0 && u < TCYM_NUM_U) take(ulev[u]);
        if (u2 >= 0 && u2 < TCYM_NUM_U) take(ulev[u2]);
        if (e >= 0 && e < TCYM_NUM_E) take(elev[e]);
        if (e2 >= 0 && e2 < TCYM_NUM_E) take(elev[e2]);
        if (mixed) return fail(TCYM_E_ARG, "slots of different resolutions (blocking levels)");
        q = geo(lev < 0 ? 0 : lev);
        return TCYM_OK;
    }
    // device fields (sites padded to whole tiles) and host AoS arrays
    long link_elems(const Geom& q) const { return 3*tcym::padded_sites(q.V)*be->ns; }
    long field_elems(const Geom& q) const { return 3*tcym::padded_sites(q.V)*be->ne; }
    long link_elems() const { return link_elems(g); }
    long field_elems() const { return field_elems(g); }

    int useU(int slot) {
        if (slot < 0 || slot >= TCYM_NUM_U) return fail(TCYM_E_ARG, "link slot out of range");
        uver[slot]++;   // every reader and writer of a slot passes here
        return alloc(&U[slot], (size_t)link_elems()*be->elsize, "hipMalloc(links)");
    }
    int useE(int slot) {
        if (slot < 0 || slot >= TCYM_NUM_E) return fail(TCYM_E_ARG, "field slot out of range");
        return alloc(&E[slot], (size_t)field_elems()*be->elsize, "hipMalloc(efield)");
    }
    int scratch() { return alloc(&Uscr, (size_t)link_elems()*be->elsize, "hipMalloc(link scratch)"); }

    // reduction: run partial kernel, final-reduce nq quantities, copy to host
    template <class F>
    int reduce(int nq, const int* is_max, F partial, double* out) {
        long nblk = partial();
        tcym::final_reduce(L(), part, nblk, nq, is_max, red);
        if (last != hipSuccess) return TCYM_E_HIP;
        int r = hip(hipMemcpyAsync(red_host, red, sizeof(double)*nq, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(result)");
        if (r) return r;
        if ((r = sync())) return r;
        for (int q = 0; q < nq; q++) {
            out[q] = red_host[q];
            if (!isfinite(out[q])) { err = "non-finite reduction result"; return TCYM_E_NONFINITE; }
        }
        return TCYM_OK;
    }
    int edotb(int u, int e, int improved, double* out, bool map_out = false) {
        int r;
        Geom q;
        if ((r = useU(u)) || (r = useE(e)) || (r = at(q, u, e))) return r;
        if (map_out && (r = alloc((void**)&map, sizeof(double)*(size_t)g.V, "hipMalloc(map)"))) return r;
        const int mx[1] = {0};
        return reduce(1, mx, [&] { return be->edotb(L(), q, improved, U[u], E[e], part, map_out ? map : nullptr); }, out);
    }
    static double flow_dctau(const Geom& q, double dtau) {   // GRADIENTFLOW/GradientFlow.cpp:152
        const double ac = q.a[0]*q.a[1]*q.a[2];
        return dtau*pow(ac, 2.0/3.0)/(q.aScale*q.aScale);
    }
    static void flow_cu(const Geom& q, double dtau, double* cU) {  // GradientFlow.cpp:109-113
        const double ac = q.a[0]*q.a[1]*q.a[2], D = dtau*pow(ac, 2.0/3.0);
        for (int mu = 0; mu < 3; mu++) cU[mu] = -(q.a[mu]*q.a[mu]*D)/(ac*q.aScale);
    }
    int ncsdot(int u, int e, double* out) {
        double v;
        int r = edotb(u, e, 1, &v);
        if (r) return r;
        *out = 0.125*v/(M_PI*M_PI);
        return TCYM_OK;
    }
    // NCsDot with the field strength of u kept in Bcool (bmode 1: formed and
    // stored; 2: read from Bcool, where a bmode-1 call on the same links left
    // it); the same sum to the bit as ncsdot
    int ncsdot_kept(int u, int e, int bmode, double* out) {
        int r;
        Geom q;
        if ((r = useU(u)) || (r = useE(e)) || (r = at(q, u, e))) return r;
        if ((r = alloc(&Bcool, (size_t)field_elems()*be->elsize, "hipMalloc(field strength)"))) return r;
        const int mx[1] = {0};
        double v;
        bk_slot = -1;
        if ((r = reduce(1, mx, [&] { return be->edotb_cached(L(), q, U[u], E[e], part, Bcool, bmode); }, &v))) return r;
        bk_slot = u;            // B describes u's links (formed, or read and unchanged)
        bk_ver = uver[u];
        *out = 0.125*v/(M_PI*M_PI);
        return TCYM_OK;
    }
};

extern "C" {

const char* tcym_version(void) { return "tcym-mi355x 0.1 (gfx950)"; }

const char* tcym_last_error(const tcym_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tcym_create(const tcym_params* p, tcym_ctx** out) {
    if (!p || !out) return TCYM_E_ARG;
    *out = nullptr;
    if ((p->nc != 2 && p->nc != 3) || (p->precision != TCYM_FP64 && p->precision != TCYM_FP32)) return TCYM_E_ARG;
    for (int k = 0; k < 3; k++) if (p->N[k] < 3 || !(p->a[k] > 0)) return TCYM_E_ARG;
    if (!(p->aScale > 0)) return TCYM_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p->device || p->device < 0) return TCYM_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) return TCYM_E_HIP;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return TCYM_E_UNSUPPORTED;
    if (hipSetDevice(p->device) != hipSuccess) return TCYM_E_HIP;
    const int f32 = p->precision == TCYM_FP32;
    // the kernels address each field through one buffer descriptor with
    // 32-bit offsets: every field (3 V s_U reals, sites padded to tiles) must
    // stay below 4 GiB
    if (!field_fits(p->N, p->nc == 2 ? 4 : 18, f32 ? 4 : 8)) return TCYM_E_UNSUPPORTED;
    tcym_ctx* c = new tcym_ctx();
    c->nc = p->nc; c->prec = p->precision; c->device = p->device;
    for (int k = 0; k < 3; k++) { c->g.n[k] = p->N[k]; c->g.a[k] = p->a[k]; }
    c->g.aScale = p->aScale;
    c->g.V = (long)p->N[0]*p->N[1]*p->N[2];
    c->be = p->nc == 2 ? tcym::make_backend_su2(f32) : tcym::make_backend_su3(f32);
    c->be64 = f32 ? (p->nc == 2 ? tcym::make_backend_su2(0) : tcym::make_backend_su3(0)) : c->be;
    int r = c->hip(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    double* scr = nullptr;
    if (!r) r = c->hip(scratch_ptr(p->device, &scr), "hipGetSymbolAddress(scratch)");
    long nblk = 0;
    for (int lev = 0; lev < 8 && c->geo(lev).n[0] >= 1 && c->geo(lev).n[1] >= 1 && c->geo(lev).n[2] >= 1; lev++)
        nblk = std::max(nblk, tcym::walk_shape(c->geo(lev)).blocks);
    // partials: up to 8 quantities per site block, or 3 per link block (3 per site block)
    if (!r) r = c->alloc((void**)&c->part, sizeof(double)*(size_t)nblk*9, "hipMalloc(partials)");
    if (!r) r = c->alloc((void**)&c->red, sizeof(double)*16, "hipMalloc(results)");
    if (!r) r = c->hip(hipHostMalloc((void**)&c->red_host, sizeof(double)*16, hipHostMallocDefault), "hipHostMalloc");
    if (!r) r = c->alloc((void**)&c->stage, sizeof(double)*(size_t)(3*c->g.V*std::max(c->be->ns, c->be->ne)), "hipMalloc(stage)");
    if (!r) r = c->sync();
    if (r) { tcym_destroy(c); return r; }
    *out = c;
    return TCYM_OK;
}

int tcym_get_params(const tcym_ctx* c, tcym_params* p) {
    if (!c || !p) return TCYM_E_ARG;
    p->nc = c->nc; p->precision = c->prec; p->device = c->device;
    for (int k = 0; k < 3; k++) { p->N[k] = c->g.n[k]; p->a[k] = c->g.a[k]; }
    p->aScale = c->g.aScale;
    return TCYM_OK;
}

int tcym_destroy(tcym_ctx* c) {
    if (!c) return TCYM_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->red_host) hipHostFree(c->red_host);
    c->release();
    if (c->be64 != c->be) delete c->be64;
    delete c->be;
    delete c;
    return TCYM_OK;
}

int tcym_sync(tcym_ctx* c) { return c ? c->sync() : TCYM_E_ARG; }

int tcym_device_bytes(tcym_ctx* c, long long* b) {
    if (!c || !b) return TCYM_E_ARG;
    *b = c->bytes;
    return TCYM_OK;
}

// ---------------------------------------------------------------- fields
int tcym_set_links(tcym_ctx* c, int slot, const double* U) {
    if (!c || !U) return TCYM_E_ARG;
    int r = c->useU(slot);
    if (r) return r;
    c->ulev[slot] = 0;
    if ((r = c->hip(hipMemcpyAsync(c->stage, U, sizeof(double)*(size_t)(3*c->g.V*c->be->ns), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(links)"))) return r;
    c->be->pack_links(c->L(), c->g, c->stage, c->U[slot]);
    return c->sync();
}
int tcym_get_links(tcym_ctx* c, int slot, double* U) {
    if (!c || !U) return TCYM_E_ARG;
    int r = c->useU(slot);
    Geom q;
    if (r || (r = c->at(q, slot))) return r;
    c->be->unpack_links(c->L(), q, c->U[slot], c->stage);
    if ((r = c->hip(hipMemcpyAsync(U, c->stage, sizeof(double)*(size_t)(3*q.V*c->be->ns), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(links)"))) return r;
    return c->sync();
}
int tcym_set_efield(tcym_ctx* c, int slot, const double* E) {
    if (!c || !E) return TCYM_E_ARG;
    int r = c->useE(slot);
    if (r) return r;
    c->elev[slot] = 0;
    if ((r = c->hip(hipMemcpyAsync(c->stage, E, sizeof(double)*(size_t)(3*c->g.V*c->be->ne), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(efield)"))) return r;
    c->be->pack_efield(c->L(), c->g, c->stage, c->E[slot]);
    return c->sync();
}
int tcym_get_efield(tcym_ctx* c, int slot, double* E) {
    if (!c || !E) return TCYM_E_ARG;
    int r = c->useE(slot);
    Geom q;
    if (r || (r = c->at(q, -1, slot))) return r;
    c->be->unpack_efield(c->L(), q, c->E[slot], c->stage);
    if ((r = c->hip(hipMemcpyAsync(E, c->stage, sizeof(double)*(size_t)(3*q.V*c->be->ne), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(efield)"))) return r;
    return c->sync();
}
int tcym_copy_links(tcym_ctx* c, int dst, int src) {
    if (!c) return TCYM_E_ARG;
    int r;
    const bool kept = c->kept_for(src);
    if ((r = c->useU(dst)) || (r = c->useU(src))) return r;
    if (dst == src) return TCYM_OK;
    c->ulev[dst] = c->ulev[src];
    c->be->copy(c->L(), c->link_elems(c->geo(c->ulev[src])), c->U[src], c->U[dst]);
    if ((r = c->sync())) return r;
    if (kept) {                 // the kept field strength now describes dst
        c->bk_slot = dst;
        c->bk_ver = c->uver[dst];
    }
    return TCYM_OK;
}
int tcym_copy_efield(tcym_ctx* c, int dst, int src) {
    if (!c) return TCYM_E_ARG;
    int r;
    if ((r = c->useE(dst)) || (r = c->useE(src))) return r;
    if (dst == src) return TCYM_OK;
    c->elev[dst] = c->elev[src];
    c->be->copy(c->L(), c->field_elems(c->geo(c->elev[src])), c->E[src], c->E[dst]);
    return c->sync();
}
int tcym_zero_efield(tcym_ctx* c, int slot) {
    if (!c) return TCYM_E_ARG;
    int r = c->useE(slot);
    if (r) return r;
    if ((r = c->hip(hipMemsetAsync(c->E[slot], 0, (size_t)c->field_elems()*c->be->elsize, c->stream), "zero efield"))) return r;
    return c->sync();
}
int tcym_unit_links(tcym_ctx* c, int slot) {
    if (!c) return TCYM_E_ARG;
    int r = c->useU(slot);
    if (r) return r;
    c->ulev[slot] = 0;
    c->be->random_links(c->L(), c->g, 0, 0.0, c->U[slot]);   // exp(0) = 1
    return c->sync();
}
int tcym_random_links(tcym_ctx* c, int slot, uint64_t seed, double sigma) {
    if (!c) return TCYM_E_ARG;
    int r = c->useU(slot);
    if (r) return r;
    c->ulev[slot] = 0;
    c->be->random_links(c->L(), c->g, seed, sigma, c->U[slot]);
    return c->sync();
}
int tcym_random_efield(tcym_ctx* c, int slot, uint64_t seed, double sigma) {
    if (!c) return TCYM_E_ARG;
    int r = c->useE(slot);
    if (r) return r;
    c->elev[slot] = 0;
    c->be->random_efield(c->L(), c->g, seed, 1u, 0u, sigma, c->E[slot]);
    return c->sync();
}
int tcym_noise(tcym_ctx* c, int xi, uint64_t seed, uint32_t step, double stddev) {
    if (!c) return TCYM_E_ARG;
    int r = c->useE(xi);
    if (r) return r;
    c->elev[xi] = 0;
    c->be->random_efield(c->L(), c->g, seed, 2u, step, stddev, c->E[xi]);
    return c->sync();
}
int tcym_level(tcym_ctx* c, int u, int e, int* level, int N[3]) {
    if (!c || (u >= 0) == (e >= 0) || u >= TCYM_NUM_U || e >= TCYM_NUM_E) return TCYM_E_ARG;
    const int lev = u >= 0 ? c->ulev[u] : c->elev[e];
    if (level) *level = lev;
    if (N) for (int k = 0; k < 3; k++) N[k] = c->geo(lev).n[k];
    return TCYM_OK;
}

// ---------------------------------------------------------------- updates
// nsteps update steps of slot u; with src >= 0 the first step reads slot src
// instead (slot u takes its level and ends with the result; src is unchanged)
static int run_update(tcym_ctx* c, int mode, int u, int e, int xi, int nsteps, double dt, double damping,
                      double noise_sd, uint64_t seed, uint32_t step0, bool write_e, int src = -1) {
    int r;
    if ((r = c->useU(u)) || (r = c->scratch())) return r;
    if (src >= 0) {
        if ((r = c->useU(src))) return r;
        c->ulev[u] = c->ulev[src];
    }
    if (e >= 0 && (r = c->useE(e))) return r;
    if (mode != 3 && e < 0) return c->fail(TCYM_E_ARG, "E slot required");
    if (mode == 2 && (r = c->useE(xi))) return r;
    if (nsteps < 0) return c->fail(TCYM_E_ARG, "nsteps < 0");
    Geom q;
    if ((r = c->at(q, u, e, -1, mode == 2 ? xi : -1))) return r;
    void* E = e >= 0 ? c->E[e] : nullptr;
    const void* X = mode == 2 ? c->E[xi] : nullptr;
    for (int k = 0; k < nsteps; k++) {
        c->be->update(c->L(), q, mode, true, write_e, dt, damping, noise_sd, seed, step0 + (uint32_t)k,
                      k == 0 && src >= 0 ? c->U[src] : c->U[u], c->Uscr, E, X, k == nsteps - 1);
        std::swap(c->U[u], c->Uscr);
        if (c->last != hipSuccess) return TCYM_E_HIP;
    }
    return c->sync();
}

int tcym_leapfrog(tcym_ctx* c, int u, int e, int nsteps, double dt) {
    if (!c) return TCYM_E_ARG;
    return run_update(c, 0, u, e, -1, nsteps, dt, 0.0, 0.0, 0, 0, true);
}
int tcym_update_efield(tcym_ctx* c, int u, int e, double dt) {
    if (!c) return TCYM_E_ARG;
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->at(q, u, e))) return r;
    c->be->update(c->L(), q, 0, false, true, dt, 0.0, 0.0, 0, 0, c->U[u], nullptr, c->E[e], nullptr, true);
    return c->sync();
}
int tcym_update_links(tcym_ctx* c, int u, int e, double dt) {
    if (!c) return TCYM_E_ARG;
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->scratch()) || (r = c->at(q, u, e))) return r;
    const double ac = q.a[0]*q.a[1]*q.a[2];
    double cU[3];
    for (int mu = 0; mu < 3; mu++) cU[mu] = -dt*q.a[mu]*q.a[mu]/ac;   // DYNAMICS/Dynamics.cpp:102-104
    c->be->link_update(c->L(), q, cU, c->U[u], c->Uscr, c->E[e]);
    std::swap(c->U[u], c->Uscr);
    return c->sync();
}
int tcym_langevin(tcym_ctx* c, int u, int e, int nsteps, double dt, double damping, double T,
                  uint64_t seed, uint32_t step0) {
    if (!c) return TCYM_E_ARG;
    if (!(T >= 0) || !(damping >= 0)) return c->fail(TCYM_E_ARG, "negative temperature or damping");
    return run_update(c, 1, u, e, -1, nsteps, dt, damping, sqrt(2.0*damping*T), seed, step0, true);
}
int tcym_evolve_ncs(tcym_ctx* c, int mode, int u, int e, int nsteps, double dt, double damping, double T,
                    uint64_t seed, uint32_t step0, double* dncs) {
    if (!c || !dncs) return TCYM_E_ARG;
    if (mode != 0 && mode != 1) return c->fail(TCYM_E_ARG, "evolve_ncs: mode 0 (leapfrog) or 1 (Langevin)");
    if (mode == 1 && (!(T >= 0) || !(damping >= 0))) return c->fail(TCYM_E_ARG, "negative temperature or damping");
    if (nsteps < 0) return c->fail(TCYM_E_ARG, "nsteps < 0");
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->scratch()) || (r = c->at(q, u, e))) return r;
    const double noise = mode == 1 ? sqrt(2.0*damping*T) : 0.0;
    const int mx = 0;
    double* eb = c->red + 8;     // E.B of U_n and U_n+1
    double* acc = c->red + 10;   // running sum
    const double zero = 0.0;
    if ((r = c->hip(hipMemcpyAsync(acc, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync(acc)"))) return r;
    const double half_dt = 0.5*dt;
    // E.B is linear in E: the field strength B of U_n+1, formed for the
    // second E.B of a step, is kept and serves the first E.B of the next
    // step (on the same links), so each step forms B once
    if (nsteps > 1 && (r = c->alloc(&c->Bc, (size_t)c->field_elems()*c->be->elsize, "hipMalloc(field strength)"))) return r;
    for (int k = 0; k < nsteps; k++) {
        c->be->update(c->L(), q, mode, true, true, dt, damping, noise, seed, step0 + (uint32_t)k, c->U[u], c->Uscr, c->E[e], nullptr,
                      true);
        std::swap(c->U[u], c->Uscr);     // Uscr now holds U_n
        long nb = c->be->edotb_cached(c->L(), q, c->Uscr, c->E[e], c->part, c->Bc, k == 0 ? 0 : 2);
        tcym::final_reduce(c->L(), c->part, nb, 1, &mx, eb);
        nb = c->be->edotb_cached(c->L(), q, c->U[u], c->E[e], c->part, c->Bc, k + 1 < nsteps ? 1 : 0);
        tcym::final_reduce(c->L(), c->part, nb, 1, &mx, eb + 1);
        c->launch("hot_ncs", [&](hipStream_t st) { hipLaunchKernelGGL(k_hot_ncs, dim3(1), dim3(64), 0, st, (const double*)eb, acc, half_dt); });
        if (c->last != hipSuccess) return TCYM_E_HIP;
    }
    double h = 0.0;
    if ((r = c->hip(hipMemcpyAsync(c->red_host, acc, sizeof(double), hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync(acc)"))) return r;
    if ((r = c->sync())) return r;
    h = c->red_host[0];
    if (!isfinite(h)) return c->fail(TCYM_E_NONFINITE, "non-finite N_CS integral");
    *dncs += h;
    return TCYM_OK;
}
int tcym_langevin_xi(tcym_ctx* c, int u, int e, int xi, double dt, double damping) {
    if (!c) return TCYM_E_ARG;
    return run_update(c, 2, u, e, xi, 1, dt, damping, 0.0, 0, 0, true);
}
int tcym_flow(tcym_ctx* c, int u, int e, int nsteps, double dtau, int calibrate, double* dncs) {
    if (!c) return TCYM_E_ARG;
    if (!calibrate) return run_update(c, 3, u, e, -1, nsteps, dtau, 0.0, 0.0, 0, 0, e >= 0);
    int r;
    if (e < 0) return c->fail(TCYM_E_ARG, "calibration flow needs an E slot");
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->scratch()) || (r = c->at(q, u, e))) return r;
    const double D = tcym_ctx::flow_dctau(q, dtau);
    double cU[3], acc = 0.0, v;
    tcym_ctx::flow_cu(q, dtau, cU);
    for (int k = 0; k < nsteps; k++) {
        // GradientFlow::CalibrationUpdate (GRADIENTFLOW/GradientFlow.cpp:168-188)
        c->be->update(c->L(), q, 3, false, true, dtau, 0.0, 0.0, 0, 0, c->U[u], nullptr, c->E[e], nullptr, true);
        if ((r = c->ncsdot(u, e, &v))) return r;
        acc += 0.5*D*v;
        c->be->link_update(c->L(), q, cU, c->U[u], c->Uscr, c->E[e]);
        std::swap(c->U[u], c->Uscr);
        if ((r = c->ncsdot(u, e, &v))) return r;
        acc += 0.5*D*v;
    }
    if (dncs) *dncs += acc;
    return c->sync();
}
int tcym_cool(tcym_ctx* c, int u_hot, int u_cold, int nsteps, double dtau) {
    if (!c) return TCYM_E_ARG;
    int r;
    // the flowed E field of Cooling::DynamicFields is discarded by CoolNSave, so
    // the fused flow keeps the force in registers (GRADIENTFLOW/Cooling.cpp:156-179).
    // The first flow step reads the hot links in place of a copy of them.
    if (nsteps > 0 && u_hot != u_cold)
        return run_update(c, 3, u_cold, -1, -1, nsteps, dtau, 0.0, 0.0, 0, 0, false, u_hot);
    if ((r = tcym_copy_links(c, u_cold, u_hot))) return r;
    return run_update(c, 3, u_cold, -1, -1, nsteps, dtau, 0.0, 0.0, 0, 0, false);
}
int tcym_reunitarize(tcym_ctx* c, int u) {
    if (!c) return TCYM_E_ARG;
    int r = c->useU(u);
    Geom q;
    if (r || (r = c->at(q, u))) return r;
    c->be->reunitarize(c->L(), q, c->U[u]);
    return c->sync();
}

// ---------------------------------------------------------------- observables
int tcym_edotb(tcym_ctx* c, int u, int e, int improved, double* out) {
    if (!c || !out) return TCYM_E_ARG;
    return c->edotb(u, e, improved ? 1 : 0, out);
}
int tcym_ncsdot(tcym_ctx* c, int u, int e, double* out) {
    if (!c || !out) return TCYM_E_ARG;
    return c->ncsdot(u, e, out);
}
int tcym_ncsdot_noalloc(tcym_ctx* c, int u, int e, double* out) {
    if (!c || !out) return TCYM_E_ARG;
    if (u < 0 || u >= TCYM_NUM_U || e < 0 || e >= TCYM_NUM_E) return c->fail(TCYM_E_ARG, "slot out of range");
    if (c->last != hipSuccess) return TCYM_E_HIP;
    int r;
    Geom q;
    if ((r = c->at(q, u, e))) return r;
    const long nblk = 3*tcym::walk_shape(q).blocks;    // partials of the improved E.B (one per block)
    if (nblk > kScratchDoubles)
        return c->fail(TCYM_E_ARG, "tcym_ncsdot_noalloc: the lattice needs " + std::to_string(nblk) +
                                       " partial sums, beyond the fixed scratch of " + std::to_string(kScratchDoubles));
    if (!c->U[u] || !c->E[e])
        return c->fail(TCYM_E_ARG, "tcym_ncsdot_noalloc: the slot holds no field (this entry point allocates no slot)");
    double* scr = nullptr;
    if ((r = c->hip(scratch_ptr(c->device, &scr), "hipGetSymbolAddress(scratch)"))) return r;
    // launches go straight to the stream: no timing events, no host copy, no synchronisation
    const long nb = c->be->edotb_on(c->stream, q, c->U[u], c->E[e], scr);
    hipLaunchKernelGGL(k_final_ncsdot, dim3(1), dim3(256), 0, c->stream, (const double*)scr, nb, out);
    return c->hip(hipGetLastError(), "tcym_ncsdot_noalloc launch");
}
int tcym_edotb_map(tcym_ctx* c, int u, int e, double* map) {
    if (!c || !map) return TCYM_E_ARG;
    double tot;
    int r = c->edotb(u, e, 1, &tot, true);
    Geom q;
    if (r || (r = c->at(q, u, e))) return r;
    if ((r = c->hip(hipMemcpyAsync(map, c->map, sizeof(double)*(size_t)q.V, hipMemcpyDeviceToHost, c->stream), "copy map"))) return r;
    return c->sync();
}
int tcym_bulk(tcym_ctx* c, int u, int e, double out[6]) {
    if (!c || !out) return TCYM_E_ARG;
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->at(q, u, e))) return r;
    const int mx[6] = {0, 0, 0, 0, 0, 0};
    double s[6];
    if ((r = c->reduce(6, mx, [&] { return c->be->bulk(c->L(), q, c->U[u], c->E[e], c->part); }, s))) return r;
    for (int k = 0; k < 6; k++) out[k] = s[k]/(double)q.V;
    return TCYM_OK;
}
int tcym_gauss(tcym_ctx* c, int u, int e, double* smax, double* rms) {
    if (!c) return TCYM_E_ARG;
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->at(q, u, e))) return r;
    const int mx[2] = {1, 0};
    double s[2];
    if ((r = c->reduce(2, mx, [&] { return c->be->gauss(c->L(), q, c->U[u], c->E[e], c->part, nullptr); }, s))) return r;
    if (smax) *smax = s[0];
    if (rms) *rms = sqrt(s[1]/((double)c->be->ne*q.V));
    return TCYM_OK;
}
int tcym_hard_scales(tcym_ctx* c, int u, double out[3]) {
    if (!c || !out) return TCYM_E_ARG;
    int r = c->useU(u);
    Geom q;
    if (r || (r = c->at(q, u))) return r;
    const int mx[3] = {0, 0, 0};
    double s[3];
    if ((r = c->reduce(3, mx, [&] { return c->be->hard_scales(c->L(), q, c->U[u], c->part); }, s))) return r;
    for (int k = 0; k < 3; k++) out[k] = s[k]/(double)q.V;   // HardScales.cpp:98
    return TCYM_OK;
}
int tcym_unitarity(tcym_ctx* c, int u, double* umax, double* rms) {
    if (!c) return TCYM_E_ARG;
    int r = c->useU(u);
    Geom q;
    if (r || (r = c->at(q, u))) return r;
    const int mx[2] = {1, 0};
    double s[2];
    if ((r = c->reduce(2, mx, [&] { return c->be->unitarity(c->L(), q, c->U[u], c->part); }, s))) return r;
    if (umax) *umax = s[0];
    if (rms) *rms = sqrt(s[1]/(3.0*q.V));
    return TCYM_OK;
}
int tcym_gauss_restore(tcym_ctx* c, int u, int e, int maxiter, double tol, int* iters, double* final_max) {
    if (!c) return TCYM_E_ARG;
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e)) || (r = c->at(q, u, e))) return r;
    if ((r = c->alloc((void**)&c->G64, sizeof(double)*(size_t)(c->g.V*c->be->ne), "hipMalloc(G)"))) return r;
    double* Uw;
    double* Ew;
    if (c->prec == TCYM_FP32) {
        if (!field_fits(c->g.n, c->be->ns, 8))
            return c->fail(TCYM_E_UNSUPPORTED, "tcym_gauss_restore: the FP64 copy of the links would exceed 4 GiB");
        if ((r = c->alloc((void**)&c->U64, sizeof(double)*(size_t)c->link_elems(), "hipMalloc(U64)"))) return r;
        if ((r = c->alloc((void**)&c->E64, sizeof(double)*(size_t)c->field_elems(), "hipMalloc(E64)"))) return r;
        c->be->to_f64(c->L(), q, c->be->ns, c->U[u], c->U64);
        c->be->to_f64(c->L(), q, c->be->ne, c->E[e], c->E64);
        Uw = c->U64; Ew = c->E64;
    } else {
        Uw = (double*)c->U[u]; Ew = (double*)c->E[e];
    }
    const int mx[2] = {1, 0};
    double s[2];
    auto dev = [&] { return c->be64->gauss(c->L(), q, Uw, Ew, c->part, c->G64); };
    if ((r = c->reduce(2, mx, dev, s))) return r;
    double smax = s[0];
    int n = 0;
    const double gam[3] = {7.0/24.0, 7.0/48.0, 7.0/72.0};   // GAUSSLAWFIXING/GaussRestore.cpp:152-154
    while (n < maxiter && smax > tol) {
        for (int k = 0; k < 3; k++) {
            if (k < 2) dev();
            else if ((r = c->reduce(2, mx, dev, s))) return r;
            c->be64->gauss_correct(c->L(), q, Uw, Ew, c->G64, gam[k]);
        }
        smax = s[0];
        n++;
        if (c->last != hipSuccess) return TCYM_E_HIP;
    }
    if (c->prec == TCYM_FP32) c->be->from_f64(c->L(), q, c->be->ne, c->E64, c->E[e]);
    if ((r = c->sync())) return r;
    if (iters) *iters = n;
    if (final_max) *final_max = smax;
    if (smax > tol) return c->fail(TCYM_E_NOCONV, "Gauss-law restoration reached its iteration cap");
    return TCYM_OK;
}
int tcym_geodesic(tcym_ctx* c, int uold, int unew, int umid, int emid, double cc) {
    if (!c) return TCYM_E_ARG;
    int r;
    Geom q;
    if ((r = c->useU(uold)) || (r = c->useU(unew)) || (r = c->useU(umid)) || (r = c->useE(emid))) return r;
    if ((r = c->at(q, uold, -1, unew))) return r;
    c->ulev[umid] = c->elev[emid] = c->ulev[uold];
    c->be->geodesic(c->L(), q, cc, c->U[uold], c->U[unew], c->U[umid], c->E[emid]);
    return c->sync();
}
// CoolingMethod::GetDeltaNCs(1): Simpson (1,4,1)/6 over the geodesic
// (TOPOLOGY/CoolingMethod.cpp:80-121).  E.B is linear in E, so the field
// strength of unew formed here is kept: when the caller then copies unew to
// uold (CoolingMethod::Update does, :143-165) and touches uold no more, the
// next call reads it for its E.B at uold instead of forming it (two field
// strengths per measurement instead of three; the same sums to the bit).
int tcym_delta_ncs_cooling(tcym_ctx* c, int uold, int unew, double* out) {
    if (!c || !out) return TCYM_E_ARG;
    // checked before the geodesic touches uold (and writes the interpolation slot)
    const bool kept_old = c->kept_for(uold) && uold != TCYM_U_MID;
    int r = tcym_geodesic(c, uold, unew, TCYM_U_MID, TCYM_E_MID, 0.5);
    if (r) return r;
    double v = 0.0, w = 0.0, x;
    if ((r = c->ncsdot(TCYM_U_MID, TCYM_E_MID, &x))) return r;
    v += 4.0*x; w += 4.0;
    if ((r = kept_old ? c->ncsdot_kept(uold, TCYM_E_MID, 2, &x) : c->ncsdot(uold, TCYM_E_MID, &x))) return r;
    v += x; w += 1.0;
    if ((r = c->ncsdot_kept(unew, TCYM_E_MID, 1, &x))) return r;
    v += x; w += 1.0;
    *out = v/w;
    return TCYM_OK;
}
int tcym_cooling_start(tcym_ctx* c, int u_hot, int nsteps, double dtau) {
    if (!c) return TCYM_E_ARG;
    int r = tcym_cool(c, u_hot, TCYM_U_NEW, nsteps, dtau);
    if (r) return r;
    return tcym_copy_links(c, TCYM_U_OLD, TCYM_U_NEW);
}
int tcym_cooling_update(tcym_ctx* c, int u_hot, int nsteps, double dtau, int measure, double* dncs) {
    if (!c) return TCYM_E_ARG;
    int r = tcym_cool(c, u_hot, TCYM_U_NEW, nsteps, dtau);
    if (r) return r;
    if (measure) {
        double d;
        if ((r = tcym_delta_ncs_cooling(c, TCYM_U_OLD, TCYM_U_NEW, &d))) return r;
        if (dncs) *dncs = d;
    }
    return tcym_copy_links(c, TCYM_U_OLD, TCYM_U_NEW);
}
int tcym_vacuum_ncs(tcym_ctx* c, int u, double* out) {
    if (!c || !out) return TCYM_E_ARG;
    int r = c->useU(u);
    Geom q;
    if (r || (r = c->at(q, u))) return r;
    const int mx[1] = {0};
    return c->reduce(1, mx, [&] { return c->be->vacuum_ncs(c->L(), q, c->U[u], c->part); }, out);
}

// ---------------------------------------------------------------- blocking and calibration
int tcym_block(tcym_ctx* c, int u, int e) {
    if (!c) return TCYM_E_ARG;
    int r;
    if ((r = c->useU(u)) || (r = c->scratch())) return r;
    if (e >= 0 && (r = c->useE(e))) return r;
    const Geom q = c->geo(c->ulev[u]);
    for (int k = 0; k < 3; k++)
        if (q.n[k] % 2 != 0 || q.n[k]/2 < 2) return c->fail(TCYM_E_ARG, "blocking needs even extents of at least 4");
    // Block(GaugeLinks**) (LATTICE/GaugeLinks.cpp:125-151) into the scratch buffer
    c->be->block_links(c->L(), q, c->U[u], c->Uscr);
    std::swap(c->U[u], c->Uscr);
    c->ulev[u]++;
    if (e >= 0) {
        // Block(ElectricFields**) sets the coarse field to zero (LATTICE/ElectricFields.cpp:113-127)
        if ((r = c->hip(hipMemsetAsync(c->E[e], 0, (size_t)c->field_elems()*c->be->elsize, c->stream), "zero efield"))) return r;
        c->elev[e] = c->ulev[u];
    }
    return c->sync();
}
// flow steps step0+1 .. nsteps of Cooling::Perform with blocking (GRADIENTFLOW/Cooling.cpp:69-147);
// calibrate: GradientFlow::CalibrationUpdate with the Delta N_CS sum and the log rows
static int perform_cooling(tcym_ctx* c, int u, int e, int step0, int nsteps, double dtau, int bfreq, int maxb,
                           int* nblock, int calibrate, double ctau, double* dncs, double* log) {
    int r;
    double acc = 0.0;
    auto log_row = [&](double* row) -> int {
        // CoolingTime, Bulk::MAGNETIC, Bulk::ELECTRIC, DeltaNCsCooling, VacuumEstimator::NCS (Cooling.cpp:85-111)
        Geom q;
        int rr;
        if ((rr = c->at(q, u, e))) return rr;
        double b[6], vac;
        if ((rr = tcym_bulk(c, u, e, b)) || (rr = tcym_vacuum_ncs(c, u, &vac))) return rr;
        const double a3 = q.a[0]*q.a[1]*q.a[2];
        double el = 0.0, mg = 0.0;
        for (int mu = 0; mu < 3; mu++) {
            const double cc = q.a[mu]*q.aScale*q.aScale/a3, cb = cc*cc;   // BulkObservables.cpp:17-25 (Minkowski)
            el += cb*b[mu];
            mg += cb*b[3 + mu];
        }
        row[0] = ctau; row[1] = 0.5*mg; row[2] = 0.5*el; row[3] = acc; row[4] = vac;
        return TCYM_OK;
    };
    double* row = log;
    if (calibrate && row) { if ((r = log_row(row))) return r; row += 5; }
    // the step after which the next blocking is due (BlockFrequency 8^BlockingNumber,
    // Cooling.cpp:123), or 0 when no more blocking happens
    auto next_block = [&](int k) -> long {
        if (bfreq <= 0 || *nblock >= maxb) return 0;
        long every = bfreq;
        for (int l = 0; l < *nblock; l++) every *= 8;
        return ((long)k/every + 1)*every;
    };
    for (int k = step0; k < nsteps;) {
        Geom q;
        if ((r = c->at(q, u, calibrate ? e : -1))) return r;
        const double D = tcym_ctx::flow_dctau(q, dtau);
        const long nb = next_block(k);
        // standard cooling runs the steps up to the next blocking in one call
        const int k1 = calibrate ? k + 1 : (int)(nb > 0 && nb < nsteps ? nb : nsteps);
        if (calibrate) {
            double d = 0.0;
            if ((r = tcym_flow(c, u, e, 1, dtau, 1, &d))) return r;
            acc += d;
        } else if ((r = run_update(c, 3, u, -1, -1, k1 - k, dtau, 0.0, 0.0, 0, 0, false))) {
            return r;
        }
        for (int j = k; j < k1; j++) ctau += D;
        k = k1;
        if (calibrate && row) { if ((r = log_row(row))) return r; row += 5; }
        if (nb > 0 && k == nb) {
            if ((r = tcym_block(c, u, calibrate ? e : -1))) return r;
            (*nblock)++;
        }
    }
    if (dncs) *dncs = acc;
    return c->sync();
}
int tcym_cool_blocked(tcym_ctx* c, int u_hot, int u_cold, int nsteps, double dtau, int block_frequency,
                      int max_blocking, int* nblocked) {
    if (!c) return TCYM_E_ARG;
    if (nsteps < 0 || block_frequency < 0 || max_blocking < 0) return c->fail(TCYM_E_ARG, "negative step or blocking count");
    int r;
    if ((r = tcym_copy_links(c, u_cold, u_hot))) return r;
    int nb = 0;
    if ((r = perform_cooling(c, u_cold, -1, 0, nsteps, dtau, block_frequency, max_blocking, &nb, 0, 0.0, nullptr, nullptr))) return r;
    if (nblocked) *nblocked = nb;
    return TCYM_OK;
}
int tcym_cool_continue(tcym_ctx* c, int u, int step0, int nsteps, double dtau, int block_frequency, int max_blocking,
                       int* nblocked) {
    if (!c || !nblocked) return TCYM_E_ARG;
    if (step0 < 0 || nsteps < step0 || block_frequency < 0 || max_blocking < 0 || *nblocked < 0)
        return c->fail(TCYM_E_ARG, "bad step or blocking counts");
    int r = c->useU(u);
    if (r) return r;
    return perform_cooling(c, u, -1, step0, nsteps, dtau, block_frequency, max_blocking, nblocked, 0, 0.0, nullptr, nullptr);
}
int tcym_continue_cooling(tcym_ctx* c, int u_src, int u_work, int e_work, int step0, int nsteps, double dtau,
                          double ctau0, int block_frequency, int max_blocking, int* nblocked, double* dncs, double* log) {
    if (!c || !nblocked) return TCYM_E_ARG;
    if (step0 < 0 || nsteps < step0 || block_frequency < 0 || max_blocking < 0) return c->fail(TCYM_E_ARG, "bad step or blocking counts");
    int r;
    // Cooling::Setup: copy the links, E zero at their resolution (GRADIENTFLOW/Cooling.cpp:38-45)
    if ((r = tcym_copy_links(c, u_work, u_src)) || (r = c->useE(e_work))) return r;
    if ((r = c->hip(hipMemsetAsync(c->E[e_work], 0, (size_t)c->field_elems()*c->be->elsize, c->stream), "zero efield"))) return r;
    c->elev[e_work] = c->ulev[u_work];
    return perform_cooling(c, u_work, e_work, step0, nsteps, dtau, block_frequency, max_blocking, nblocked, 1, ctau0, dncs, log);
}

// ---------------------------------------------------------------- thermal initial condition
int tcym_thermalize(tcym_ctx* c, int u, int e, int cycles, double max_time, double dt, double beta, uint64_t seed,
                    int maxiter, double tol, double* log, int log_rows, int* nrows) {
    if (!c) return TCYM_E_ARG;
    if (cycles < 0 || !(dt > 0) || !(beta > 0) || !(max_time >= 0) || log_rows < 0 || (log_rows > 0 && !log))
        return c->fail(TCYM_E_ARG, "tcym_thermalize: bad cycle count, step, beta, time or log");
    int r;
    Geom q;
    if ((r = c->useU(u)) || (r = c->useE(e))) return r;
    c->elev[e] = 0;
    if ((r = c->at(q, u, e))) return r;
    if (c->ulev[u] != 0) return c->fail(TCYM_E_ARG, "tcym_thermalize: the links must be at full resolution");
    // Bulk::SetConstants, Minkowski (OBSERVABLES/BulkObservables.cpp:17-25)
    double cb[3];
    for (int mu = 0; mu < 3; mu++) {
        const double x = q.a[mu]*q.aScale*q.aScale/(q.a[0]*q.a[1]*q.a[2]);
        cb[mu] = x*x;
    }
    const double sigma = 1.0/sqrt(beta);
    int rows = 0;
    for (int i = 0; i < cycles; i++) {
        // ReassignElectricFields: E^a ~ N(0, 1/beta), Philox stream 3 with the
        // cycle as counter step, then GaussLawRestoration::Restore (DYNAMICS/ThermalDynamics.cpp:205-226)
        c->be->random_efield(c->L(), q, seed, 3u, (uint32_t)i, sigma, c->E[e]);
        if ((r = tcym_gauss_restore(c, u, e, maxiter, tol, nullptr, nullptr))) return r;
        // evolve while Time() = tSteps dTau < MaxTime; Bulk every 20 steps (:274-292), Reset (:85-91)
        long t = 0;
        while ((double)t*dt < max_time) {
            int n = 0;
            do { t++; n++; } while (t % 20 != 0 && (double)t*dt < max_time);
            if ((r = tcym_leapfrog(c, u, e, n, dt))) return r;
            if (t % 20 == 0) {
                double b[6];
                if ((r = tcym_bulk(c, u, e, b))) return r;
                if (rows < log_rows) {
                    double* w = log + 7*(size_t)rows;
                    const double E0 = cb[0]*b[0], E1 = cb[1]*b[1], E2 = cb[2]*b[2];
                    const double B0 = cb[0]*b[3], B1 = cb[1]*b[4], B2 = cb[2]*b[5];
                    // Time() + i MaxTime, T00, TXX, TYY, TZZ, ELECTRIC, MAGNETIC (BulkObservables.cpp:28-52)
                    w[0] = (double)t*dt + i*max_time;
                    w[1] = 0.5*(E0 + E1 + E2 + B0 + B1 + B2);
                    w[2] = 0.5*((E1 + E2 + B1 + B2) - (E0 + B0));
                    w[3] = 0.5*((E0 + E2 + B0 + B2) - (E1 + B1));
                    w[4] = 0.5*((E0 + E1 + B0 + B1) - (E2 + B2));
                    w[5] = 0.5*(E0 + E1 + E2);
                    w[6] = 0.5*(B0 + B1 + B2);
                }
                rows++;
            }
        }
    }
    if (nrows) *nrows = rows;
    return TCYM_OK;
}

// ---------------------------------------------------------------- quasi-particle initial condition
int tcym_quasi_particles(tcym_ctx* c, int u, int e, double Qs, double n0, uint64_t seed) {
    if (!c) return TCYM_E_ARG;
    if (!(Qs > 0) || !(n0 >= 0)) return c->fail(TCYM_E_ARG, "tcym_quasi_particles: Qs must be positive and n0 not negative");
    int r;
    if ((r = c->useU(u)) || (r = c->useE(e))) return r;
    c->ulev[u] = 0;
    c->elev[e] = 0;
    const Geom& g = c->g;
    const size_t nmodes = (size_t)6*c->be->ne*(size_t)g.V;   // A_mu and E_mu of every colour
    double* P = nullptr;
    void* work = nullptr;
    if ((r = c->hip(hipMalloc((void**)&P, sizeof(double)*2*nmodes), "hipMalloc(quasi-particle modes)"))) return r;
    auto done = [&](int rr) {
        hipStreamSynchronize(c->stream);
        if (work) hipFree(work);
        hipFree(P);
        return rr;
    };
    c->be->qp_modes(c->L(), g, Qs, n0, seed, P);
    // FourierSpace::*->ExecutePtoX (FFT/FFT3D.cpp, FFTW_BACKWARD: unnormalised,
    // +i exponent): rocFFT's complex inverse transform, batched over 6 NE modes
    static std::once_flag setup;
    std::call_once(setup, [] { rocfft_setup(); });
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    const size_t len[3] = {(size_t)g.n[0], (size_t)g.n[1], (size_t)g.n[2]};   // fastest (x) first
    size_t wbytes = 0;
    bool ok = rocfft_plan_create(&plan, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                                 rocfft_precision_double, 3, len, (size_t)6*c->be->ne, nullptr) == rocfft_status_success &&
              rocfft_plan_get_work_buffer_size(plan, &wbytes) == rocfft_status_success &&
              rocfft_execution_info_create(&info) == rocfft_status_success &&
              rocfft_execution_info_set_stream(info, c->stream) == rocfft_status_success;
    if (ok && wbytes > 0)
        ok = hipMalloc(&work, wbytes) == hipSuccess &&
             rocfft_execution_info_set_work_buffer(info, work, wbytes) == rocfft_status_success;
    if (ok) {
        void* in[1] = {P};
        ok = rocfft_execute(plan, in, nullptr, info) == rocfft_status_success;
    }
    if (info) rocfft_execution_info_destroy(info);
    if (plan) rocfft_plan_destroy(plan);
    if (!ok) return done(c->fail(TCYM_E_HIP, "tcym_quasi_particles: rocFFT plan or execution failed"));
    c->be->qp_fields(c->L(), g, P, c->U[u], c->E[e]);
    return done(c->sync());
}

// ---------------------------------------------------------------- timing
int tcym_set_update_form(tcym_ctx* c, int form) {
    if (!c) return TCYM_E_ARG;
    if (form != TCYM_FORM_AUTO && form != TCYM_FORM_LINK && form != TCYM_FORM_PLAQUETTE)
        return c->fail(TCYM_E_ARG, "update form " + std::to_string(form) + ": TCYM_FORM_AUTO, _LINK or _PLAQUETTE");
    c->g.form = form;
    return TCYM_OK;
}

int tcym_timing(tcym_ctx* c, int enable) {
    if (!c) return TCYM_E_ARG;
    int r = c->resolve();
    c->timing = enable != 0;
    return r;
}
int tcym_timing_reset(tcym_ctx* c) {
    if (!c) return TCYM_E_ARG;
    int r = c->resolve();
    c->stats.clear();
    return r;
}
int tcym_kernel_time(tcym_ctx* c, const char* name, double* total_ms, long long* launches) {
    if (!c || !name) return TCYM_E_ARG;
    int r = c->resolve();
    if (r) return r;
    auto it = c->stats.find(name);
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.second;
    return TCYM_OK;
}

}  // extern "C"
