// C ABI of batches of lattices (include/tcym_batch.h): one device allocation
// per field slot holding every lattice, per-lattice offset tables, persistent
// kernels over all lattices, HIP-graph capture of update+measure segments and
// an autotuner of the update kernels' walk shape and grid.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "tcym.h"
#include "tcym_batch.h"
#include "tcym_backend.hpp"
#include "tcym_device.hpp"
#include "tcym_reduce.hpp"

using tcym::Backend;
using tcym::BatchEns;
using tcym::Geom;
using tcym::KGeom;
using tcym::WalkShape;
using tcym::kBlock;

namespace {

// per-lattice final reductions: block (q, e) reduces quantity q of lattice e's
// partials part[(tile0[e] + b)*nq + q], b < nblk[e], in the order of k_final
__global__ __launch_bounds__(256) void k_final_batch(const double* __restrict__ part, const uint32_t* __restrict__ tile0,
                                                     const uint32_t* __restrict__ nblk, int nq, unsigned is_max,
                                                     double* __restrict__ out) {
    const int q = blockIdx.x, e = blockIdx.y;
    const double r = tcym::final_reduce_block(part + (long)tile0[e]*nq, nblk[e], nq, q, (is_max >> q) & 1);
    if (threadIdx.x == 0) out[(long)e*nq + q] = r;
}
// NCsDot = E.B / (8 pi^2) of every lattice into column counters[1] of the
// trajectory, then advance the column, and the Philox step counter by `steps`
// (one block: every thread reads the column before thread 0 advances it)
__global__ void k_segment_end(const double* __restrict__ red, double* __restrict__ traj, long cap, int count,
                              uint32_t* __restrict__ counters, uint32_t steps) {
    const uint32_t col = counters[1];
    if (traj && col < (uint32_t)cap)
        for (int e = threadIdx.x; e < count; e += blockDim.x) traj[(long)e*cap + col] = 0.125*red[e]/(M_PI*M_PI);
    __syncthreads();
    if (threadIdx.x == 0) { counters[1] = col + 1u; counters[0] += steps; }
}
__global__ void k_advance(uint32_t* __restrict__ counters, uint32_t steps) {
    if (blockIdx.x == 0 && threadIdx.x == 0) counters[0] += steps;
}

KGeom kgeom(const Geom& g, const WalkShape& w) { return make_kgeom(g, w); }

// Shapes of an update pass: -1 the canonical walk, 0 the linear walk, 4..6
// patch walks of the per-link form; 7, 8, 9 the plaquette form (FP32) with
// units of 2, 4 or 8 planes.
constexpr int kShapePlaq = 7;
constexpr int kPlaqBlock = 256;      // threads of a plaquette block (kBatchPlaqThreads)
int plaq_planes(int shape) { return 2 << (shape - kShapePlaq); }
// rows of a plaquette unit of lattice q, 0 if the form does not take it
// (rows of more than half a block need the whole plane in one block)
int plaq_rows_of(const Geom& q) {
    const int py = tcym::plaq_rows(q.n[0], q.n[1], kPlaqBlock);
    return py && (py == q.n[1] || 2*q.n[0] <= kPlaqBlock) ? py : 0;
}
long plaq_units(const Geom& q, int shape) {
    const int py = plaq_rows_of(q), pz = plaq_planes(shape);
    return (long)(q.n[1]/py)*((q.n[2] + pz - 1)/pz);
}

// walk shape of the update kernels: -1 the canonical one, 0 linear, 4..6 patches
WalkShape update_shape(const Geom& g, int shape) {
    if (shape < 0) return tcym::walk_shape(g);
    if (shape == 0) return tcym::linear_shape(g);
    return tcym::patch_shape(g, shape);
}

}  // namespace

struct tcym_batch : tcym::DeviceQueue {
    int nc = 0, prec = 0, count = 0;
    Geom g{};                       // spacings; extents unused
    std::vector<Geom> geo;          // per lattice
    Backend* be = nullptr;
    std::vector<long> u_off, e_off; // element offsets per lattice
    long u_elems = 0, e_elems = 0;  // elements per slot
    void* U[TCYM_NUM_U] = {};
    void* E[TCYM_NUM_E] = {};
    void* Uscr = nullptr;
    std::vector<BatchEns> ens;
    BatchEns* ens_dev = nullptr;
    uint32_t* utiles = nullptr;
    uint32_t* ftiles = nullptr;
    uint32_t* ltiles = nullptr;
    uint32_t* stiles = nullptr;
    uint32_t n_utiles = 0, n_ftiles = 0, n_ltiles = 0, n_stiles = 0;
    uint32_t* red_l0 = nullptr;     // per lattice: first tile and tile count of the link / site measurement passes
    uint32_t* red_ln = nullptr;
    uint32_t* red_s0 = nullptr;
    uint32_t* red_sn = nullptr;
    double* part = nullptr;
    double* red = nullptr;          // count x 6 device results
    double* red_host = nullptr;     // pinned
    double* stage = nullptr;        // AoS staging of the largest lattice
    uint32_t* counters = nullptr;   // [0] Philox step of the next update, [1] next trajectory column
    uint32_t step_host = 0;
    double* traj = nullptr;
    long traj_cap = 0;
    int cus = 256;
    // per pass (0: leapfrog and Langevin, 1: flow): walk shape of the update
    // kernel (see update_shape), persistent blocks per CU, XCD ranges
    int shape[2] = {-1, -1};
    int per_cu[2] = {8, 8};
    bool xcd_split[2] = {true, true};
    uint64_t seed_set = ~0ull;      // key base of the Langevin noise in the table
    struct Graph { hipGraph_t graph = nullptr; hipGraphExec_t exec = nullptr; };
    std::map<std::string, Graph> graphs;

    long max_link_host() const {
        long m = 0;
        for (const Geom& q : geo) m = std::max(m, 3*q.V*std::max(be->ns, be->ne));
        return m;
    }
    int useU(int slot) {
        if (slot < 0 || slot >= TCYM_NUM_U) return fail(TCYM_E_ARG, "link slot out of range");
        return alloc(&U[slot], (size_t)u_elems*be->elsize, "hipMalloc(batch links)");
    }
    int useE(int slot) {
        if (slot < 0 || slot >= TCYM_NUM_E) return fail(TCYM_E_ARG, "field slot out of range");
        return alloc(&E[slot], (size_t)e_elems*be->elsize, "hipMalloc(batch efield)");
    }
    int scratch() { return alloc(&Uscr, (size_t)u_elems*be->elsize, "hipMalloc(batch link scratch)"); }
    int lattice(int e) { return (e < 0 || e >= count) ? fail(TCYM_E_ARG, "lattice index out of range") : TCYM_OK; }
    char* at_u(void* base, int e) const { return (char*)base + (size_t)u_off[e]*be->elsize; }
    char* at_e(void* base, int e) const { return (char*)base + (size_t)e_off[e]*be->elsize; }

    int grid_for(uint32_t ntiles, int blocks_per_cu, bool split) const {
        long G = (long)cus*blocks_per_cu;
        if (!split) G -= 1;                         // not a multiple of 8: no XCD ranges
        G = std::min<long>(G, (long)ntiles);
        return (int)std::max<long>(G, 1);
    }
    tcym::BatchDev dev() const {
        tcym::BatchDev d{};
        d.g = g;
        d.ens = ens_dev;
        d.utiles = utiles; d.n_utiles = n_utiles;
        d.ftiles = ftiles; d.n_ftiles = n_ftiles;
        d.ltiles = ltiles; d.n_ltiles = n_ltiles;
        d.stiles = stiles; d.n_stiles = n_stiles;
        d.grid_update = grid_for(n_utiles, per_cu[0], xcd_split[0]);
        d.grid_flow = grid_for(n_ftiles, per_cu[1], xcd_split[1]);
        d.grid_measure = grid_for(std::max(n_ltiles, n_stiles), 8, true);
        d.step = counters;
        d.plaq_update = shape[0] >= kShapePlaq;
        d.plaq_flow = shape[1] >= kShapePlaq;
        return d;
    }
    void drop_graphs() {
        for (auto& kv : graphs) {
            if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) hipGraphDestroy(kv.second.graph);
        }
        graphs.clear();
    }
    // tile tables for the walk shapes su (leapfrog, Langevin) and sf (flow)
    // (measurement passes always use the canonical shape of each lattice)
    // every lattice takes the plaquette form (FP32 storage only)
    bool plaq_ok() const {
        if (be->elsize != 4) return false;
        for (const Geom& q : geo)
            if (!plaq_rows_of(q)) return false;
        return true;
    }
    int build_tables(int su, int sf) {
        if ((su >= kShapePlaq || sf >= kShapePlaq) && !plaq_ok())
            return fail(TCYM_E_UNSUPPORTED, "the plaquette form needs FP32 storage and rows that fit its blocks");
        std::vector<uint32_t> ut, ft, lt, st, l0, ln, s0, sn;
        ens.assign(count, BatchEns{});
        for (int e = 0; e < count; e++) {
            const Geom& q = geo[e];
            const WalkShape wu = update_shape(q, su < kShapePlaq ? su : -1), wf = update_shape(q, sf < kShapePlaq ? sf : -1),
                            wr = tcym::walk_shape(q);
            BatchEns& b = ens[e];
            b.gu = kgeom(q, wu);
            b.gf = kgeom(q, wf);
            b.gr = kgeom(q, wr);
            b.u_off = (uint64_t)u_off[e];
            b.e_off = (uint64_t)e_off[e];
            b.seed = seed_set == ~0ull ? 0 : seed_set + (uint64_t)e;
            b.utile0 = (uint32_t)ut.size();
            b.ftile0 = (uint32_t)ft.size();
            b.ltile0 = (uint32_t)lt.size();
            b.stile0 = (uint32_t)st.size();
            l0.push_back(b.ltile0); ln.push_back((uint32_t)(3*wr.blocks));
            s0.push_back(b.stile0); sn.push_back((uint32_t)wr.blocks);
            // a pass's table lists walk tiles or, in the plaquette form, units
            // (band-major within a slab: unit k is band k % nyb of slab k / nyb)
            if (su >= kShapePlaq) {
                b.upy = plaq_rows_of(q); b.upz = plaq_planes(su); b.unyb = q.n[1]/b.upy;
                b.udnyb = tcym::fast_div((uint32_t)b.unyb);
                ut.insert(ut.end(), (size_t)plaq_units(q, su), (uint32_t)e);
            } else {
                ut.insert(ut.end(), (size_t)(3*wu.blocks), (uint32_t)e);
            }
            if (sf >= kShapePlaq) {
                b.fpy = plaq_rows_of(q); b.fpz = plaq_planes(sf); b.fnyb = q.n[1]/b.fpy;
                b.fdnyb = tcym::fast_div((uint32_t)b.fnyb);
                ft.insert(ft.end(), (size_t)plaq_units(q, sf), (uint32_t)e);
            } else {
                ft.insert(ft.end(), (size_t)(3*wf.blocks), (uint32_t)e);
            }
            lt.insert(lt.end(), (size_t)(3*wr.blocks), (uint32_t)e);
            st.insert(st.end(), (size_t)wr.blocks, (uint32_t)e);
        }
        if (ut.size() >= 0xffffffffull || ft.size() >= 0xffffffffull || lt.size() >= 0xffffffffull)
            return fail(TCYM_E_UNSUPPORTED, "batch too large");
        // the update tables are sized for the largest candidate shape
        size_t ucap = 0;
        for (const Geom& q : geo) {
            long m = 0;
            for (int s : {-1, 0, 4, 5, 6}) m = std::max(m, 3*update_shape(q, s).blocks);
            if (plaq_rows_of(q))
                for (int s = kShapePlaq; s < kShapePlaq + 3; s++) m = std::max(m, plaq_units(q, s));
            ucap += (size_t)m;
        }
        int r;
        if ((r = alloc((void**)&utiles, sizeof(uint32_t)*ucap, "hipMalloc(tiles)"))) return r;
        if ((r = alloc((void**)&ftiles, sizeof(uint32_t)*ucap, "hipMalloc(tiles)"))) return r;
        if ((r = alloc((void**)&ltiles, sizeof(uint32_t)*lt.size(), "hipMalloc(tiles)"))) return r;
        if ((r = alloc((void**)&stiles, sizeof(uint32_t)*st.size(), "hipMalloc(tiles)"))) return r;
        if ((r = alloc((void**)&ens_dev, sizeof(BatchEns)*count, "hipMalloc(lattice table)"))) return r;
        for (uint32_t** p : {&red_l0, &red_ln, &red_s0, &red_sn})
            if ((r = alloc((void**)p, sizeof(uint32_t)*count, "hipMalloc(lattice table)"))) return r;
        const size_t np = std::max((size_t)lt.size(), (size_t)st.size()*6);
        if ((r = alloc((void**)&part, sizeof(double)*np, "hipMalloc(partials)"))) return r;
        n_utiles = (uint32_t)ut.size(); n_ftiles = (uint32_t)ft.size();
        n_ltiles = (uint32_t)lt.size(); n_stiles = (uint32_t)st.size();
        auto up = [&](void* d, const void* h, size_t n) { return hip(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, stream), "hipMemcpyAsync(tables)"); };
        if ((r = up(utiles, ut.data(), sizeof(uint32_t)*ut.size()))) return r;
        if ((r = up(ftiles, ft.data(), sizeof(uint32_t)*ft.size()))) return r;
        if ((r = up(ltiles, lt.data(), sizeof(uint32_t)*lt.size()))) return r;
        if ((r = up(stiles, st.data(), sizeof(uint32_t)*st.size()))) return r;
        if ((r = up(ens_dev, ens.data(), sizeof(BatchEns)*count))) return r;
        if ((r = up(red_l0, l0.data(), sizeof(uint32_t)*count))) return r;
        if ((r = up(red_ln, ln.data(), sizeof(uint32_t)*count))) return r;
        if ((r = up(red_s0, s0.data(), sizeof(uint32_t)*count))) return r;
        if ((r = up(red_sn, sn.data(), sizeof(uint32_t)*count))) return r;
        shape[0] = su;
        shape[1] = sf;
        drop_graphs();
        return sync();
    }
    // Langevin keys seed0 + e in the lattice table
    int set_seed(uint64_t seed0) {
        if (seed0 == seed_set) return TCYM_OK;
        seed_set = seed0;
        for (int e = 0; e < count; e++) ens[e].seed = seed0 + (uint64_t)e;
        int r = hip(hipMemcpyAsync(ens_dev, ens.data(), sizeof(BatchEns)*count, hipMemcpyHostToDevice, stream), "hipMemcpyAsync(lattice table)");
        if (r) return r;
        drop_graphs();
        return sync();
    }
    // one update step of every lattice: U[u] (or U[src]) -> Uscr, then swap
    // with U[u]; `complete` false for a step whose links only the next update
    // step reads
    void step(int mode, int u, int e, bool write_e, double dt, double damping, double noise_sd, uint32_t step_off, bool complete,
              int src = -1) {
        be->update_batch(L(), dev(), mode, write_e, dt, damping, noise_sd, step_off, src >= 0 ? U[src] : U[u], Uscr,
                         e >= 0 ? E[e] : nullptr, nullptr, complete);
        std::swap(U[u], Uscr);
    }
    void advance(uint32_t steps) {
        launch("batch_counter", [&](hipStream_t st) { hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, counters, steps); });
        step_host += steps;
    }
    // per-lattice reductions: run the partial kernel, reduce nq quantities per lattice into red
    void final_reduce(bool links, int nq, unsigned is_max) {
        const uint32_t* t0 = links ? red_l0 : red_s0;
        const uint32_t* tn = links ? red_ln : red_sn;
        launch("batch_reduce", [&](hipStream_t st) {
            hipLaunchKernelGGL(k_final_batch, dim3(nq, count), dim3(256), 0, st, (const double*)part, t0, tn, nq, is_max, red);
        });
    }
    int fetch(int nq, double* out) {
        int r = hip(hipMemcpyAsync(red_host, red, sizeof(double)*nq*count, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(results)");
        if (r || (r = sync())) return r;
        for (int i = 0; i < nq*count; i++) {
            out[i] = red_host[i];
            if (!isfinite(out[i])) { err = "non-finite reduction result"; return TCYM_E_NONFINITE; }
        }
        return TCYM_OK;
    }
    int ncsdot(int u, int e, double* out) {
        int r;
        if ((r = useU(u)) || (r = useE(e))) return r;
        be->edotb_batch(L(), dev(), U[u], E[e], part);
        final_reduce(true, 1, 0u);
        if (last != hipSuccess) return TCYM_E_HIP;
        if ((r = fetch(1, out))) return r;
        for (int i = 0; i < count; i++) out[i] = 0.125*out[i]/(M_PI*M_PI);
        return TCYM_OK;
    }
    int copy_links(int dst, int src) {
        int r;
        if ((r = useU(dst)) || (r = useU(src))) return r;
        be->copy(L(), u_elems, U[src], U[dst]);
        return status();
    }
    // nsteps flow steps of slot u; with src >= 0 the first reads slot src
    // (unchanged) in place of a copy of it in u
    int flow(int u, int nsteps, double dtau, int src = -1) {
        int r;
        if ((r = useU(u)) || (r = scratch())) return r;
        if (src >= 0 && (r = useU(src))) return r;
        if (src >= 0 && nsteps <= 0) return copy_links(u, src);
        for (int k = 0; k < nsteps; k++) step(3, u, -1, false, dtau, 0.0, 0.0, 0, k == nsteps - 1, k == 0 ? src : -1);
        return status();
    }
};

extern "C" {

const char* tcym_batch_last_error(const tcym_batch* b) { return b ? b->err.c_str() : "null batch"; }

int tcym_batch_create(const tcym_batch_params* p, tcym_batch** out) {
    if (!p || !out || !p->N) return TCYM_E_ARG;
    *out = nullptr;
    if ((p->nc != 2 && p->nc != 3) || (p->precision != TCYM_FP64 && p->precision != TCYM_FP32)) return TCYM_E_ARG;
    if (p->count < 1) return TCYM_E_ARG;
    for (int k = 0; k < 3; k++) if (!(p->a[k] > 0)) return TCYM_E_ARG;
    if (!(p->aScale > 0)) return TCYM_E_ARG;
    for (int e = 0; e < p->count; e++)
        for (int k = 0; k < 3; k++) if (p->N[3*e + k] < 3) return TCYM_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p->device || p->device < 0) return TCYM_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) return TCYM_E_HIP;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return TCYM_E_UNSUPPORTED;
    if (hipSetDevice(p->device) != hipSuccess) return TCYM_E_HIP;
    const int f32 = p->precision == TCYM_FP32;
    const int ns = p->nc == 2 ? 4 : 18, ne = p->nc == 2 ? 3 : 8, el = f32 ? 4 : 8;
    tcym_batch* b = new tcym_batch();
    b->nc = p->nc; b->prec = p->precision; b->count = p->count; b->device = p->device;
    b->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    for (int k = 0; k < 3; k++) b->g.a[k] = p->a[k];
    b->g.aScale = p->aScale;
    for (int e = 0; e < p->count; e++) {
        Geom q = b->g;
        for (int k = 0; k < 3; k++) q.n[k] = p->N[3*e + k];
        q.V = (long)q.n[0]*q.n[1]*q.n[2];
        // each lattice's field is addressed through its own 32-bit buffer offsets
        if (3.0*(double)tcym::padded_sites(q.V)*ns*el >= 4294967296.0) { delete b; return TCYM_E_UNSUPPORTED; }
        b->geo.push_back(q);
        b->u_off.push_back(b->u_elems);
        b->e_off.push_back(b->e_elems);
        b->u_elems += 3*tcym::padded_sites(q.V)*ns;
        b->e_elems += 3*tcym::padded_sites(q.V)*ne;
    }
    b->be = p->nc == 2 ? tcym::make_backend_su2(f32) : tcym::make_backend_su3(f32);
    int r = b->hip(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!r) r = b->alloc((void**)&b->red, sizeof(double)*6*(size_t)b->count, "hipMalloc(results)");
    if (!r) r = b->hip(hipHostMalloc((void**)&b->red_host, sizeof(double)*6*(size_t)b->count, hipHostMallocDefault), "hipHostMalloc");
    if (!r) r = b->alloc((void**)&b->stage, sizeof(double)*(size_t)b->max_link_host(), "hipMalloc(stage)");
    if (!r) r = b->alloc((void**)&b->counters, sizeof(uint32_t)*4, "hipMalloc(counters)");
    if (!r) r = b->build_tables(-1, -1);
    if (!r) r = b->sync();
    if (r) { tcym_batch_destroy(b); return r; }
    *out = b;
    return TCYM_OK;
}

int tcym_batch_destroy(tcym_batch* b) {
    if (!b) return TCYM_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    b->drop_graphs();
    if (b->red_host) hipHostFree(b->red_host);
    b->release();
    delete b->be;
    delete b;
    return TCYM_OK;
}

int tcym_batch_sync(tcym_batch* b) { return b ? b->sync() : TCYM_E_ARG; }

int tcym_batch_device_bytes(tcym_batch* b, long long* bytes) {
    if (!b || !bytes) return TCYM_E_ARG;
    *bytes = b->bytes;
    return TCYM_OK;
}

// ---------------------------------------------------------------- fields
int tcym_batch_set_links(tcym_batch* b, int slot, int e, const double* U) {
    if (!b || !U) return TCYM_E_ARG;
    int r;
    if ((r = b->lattice(e)) || (r = b->useU(slot))) return r;
    const Geom& q = b->geo[e];
    if ((r = b->hip(hipMemcpyAsync(b->stage, U, sizeof(double)*3*q.V*b->be->ns, hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync(links)"))) return r;
    b->be->pack_links(b->L(), q, b->stage, b->at_u(b->U[slot], e));
    return b->sync();
}
int tcym_batch_get_links(tcym_batch* b, int slot, int e, double* U) {
    if (!b || !U) return TCYM_E_ARG;
    int r;
    if ((r = b->lattice(e)) || (r = b->useU(slot))) return r;
    const Geom& q = b->geo[e];
    b->be->unpack_links(b->L(), q, b->at_u(b->U[slot], e), b->stage);
    if ((r = b->hip(hipMemcpyAsync(U, b->stage, sizeof(double)*3*q.V*b->be->ns, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync(links)"))) return r;
    return b->sync();
}
int tcym_batch_set_efield(tcym_batch* b, int slot, int e, const double* E) {
    if (!b || !E) return TCYM_E_ARG;
    int r;
    if ((r = b->lattice(e)) || (r = b->useE(slot))) return r;
    const Geom& q = b->geo[e];
    if ((r = b->hip(hipMemcpyAsync(b->stage, E, sizeof(double)*3*q.V*b->be->ne, hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync(efield)"))) return r;
    b->be->pack_efield(b->L(), q, b->stage, b->at_e(b->E[slot], e));
    return b->sync();
}
int tcym_batch_get_efield(tcym_batch* b, int slot, int e, double* E) {
    if (!b || !E) return TCYM_E_ARG;
    int r;
    if ((r = b->lattice(e)) || (r = b->useE(slot))) return r;
    const Geom& q = b->geo[e];
    b->be->unpack_efield(b->L(), q, b->at_e(b->E[slot], e), b->stage);
    if ((r = b->hip(hipMemcpyAsync(E, b->stage, sizeof(double)*3*q.V*b->be->ne, hipMemcpyDeviceToHost, b->stream), "hipMemcpyAsync(efield)"))) return r;
    return b->sync();
}

int tcym_batch_initial_condition(tcym_batch* b, uint64_t seed0, double sigma_u, double sigma_e,
                                 int maxiter, double tol, int* iters) {
    if (!b) return TCYM_E_ARG;
    // lattice e is built by an FP64 single-lattice context exactly as a caller
    // of the single-lattice ABI would build it, then copied in (AoS FP64)
    std::vector<double> Uh, Eh;
    int status = TCYM_OK;
    for (int e = 0; e < b->count; e++) {
        const Geom& q = b->geo[e];
        tcym_params p{};
        p.nc = b->nc; p.precision = TCYM_FP64; p.aScale = b->g.aScale; p.device = b->device;
        for (int k = 0; k < 3; k++) { p.N[k] = q.n[k]; p.a[k] = b->g.a[k]; }
        tcym_ctx* c = nullptr;
        int r = tcym_create(&p, &c);
        if (r) return b->fail(r, "initial condition: tcym_create failed");
        int it = 0;
        double fm = 0.0;
        Uh.resize((size_t)(3*q.V*b->be->ns));
        Eh.resize((size_t)(3*q.V*b->be->ne));
        if (!r) r = tcym_random_links(c, TCYM_U_HOT, seed0 + (uint64_t)e, sigma_u);
        if (!r) r = tcym_random_efield(c, TCYM_E_HOT, seed0 + (uint64_t)e, sigma_e);
        if (!r && maxiter > 0) {
            r = tcym_gauss_restore(c, TCYM_U_HOT, TCYM_E_HOT, maxiter, tol, &it, &fm);
            if (r == TCYM_E_NOCONV) { status = TCYM_E_NOCONV; r = TCYM_OK; }
        }
        if (!r) r = tcym_get_links(c, TCYM_U_HOT, Uh.data());
        if (!r) r = tcym_get_efield(c, TCYM_E_HOT, Eh.data());
        std::string cerr = r ? tcym_last_error(c) : "";
        tcym_destroy(c);
        if (r) return b->fail(r, "initial condition: " + cerr);
        if (iters) iters[e] = it;
        if ((r = tcym_batch_set_links(b, TCYM_U_HOT, e, Uh.data()))) return r;
        if ((r = tcym_batch_set_efield(b, TCYM_E_HOT, e, Eh.data()))) return r;
    }
    if (status) b->err = "initial condition: GaussRestore did not converge on every lattice";
    return status;
}

int tcym_batch_set_step(tcym_batch* b, uint32_t step) {
    if (!b) return TCYM_E_ARG;
    uint32_t c[2] = {step, 0};
    int r = b->hip(hipMemcpyAsync(b->counters, c, sizeof(c), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync(step)");
    if (r || (r = b->sync())) return r;
    b->step_host = step;
    return TCYM_OK;
}
int tcym_batch_get_step(tcym_batch* b, uint32_t* step) {
    if (!b || !step) return TCYM_E_ARG;
    *step = b->step_host;
    return TCYM_OK;
}

// ---------------------------------------------------------------- updates
static int run_updates(tcym_batch* b, int mode, int nsteps, double dt, double damping, double noise_sd) {
    int r;
    if ((r = b->useU(TCYM_U_HOT)) || (r = b->useE(TCYM_E_HOT)) || (r = b->scratch())) return r;
    if (nsteps < 0) return b->fail(TCYM_E_ARG, "nsteps < 0");
    for (int k = 0; k < nsteps; k++) {
        b->step(mode, TCYM_U_HOT, TCYM_E_HOT, true, dt, damping, noise_sd, (uint32_t)k, k == nsteps - 1);
        if (b->last != hipSuccess) return TCYM_E_HIP;
    }
    if (mode == 1) b->advance((uint32_t)nsteps);
    return b->sync();
}
int tcym_batch_leapfrog(tcym_batch* b, int nsteps, double dt) {
    if (!b) return TCYM_E_ARG;
    return run_updates(b, 0, nsteps, dt, 0.0, 0.0);
}
int tcym_batch_langevin(tcym_batch* b, int nsteps, double dt, double damping, double T, uint64_t seed0) {
    if (!b) return TCYM_E_ARG;
    if (!(T >= 0) || !(damping >= 0)) return b->fail(TCYM_E_ARG, "negative temperature or damping");
    int r = b->set_seed(seed0);
    if (r) return r;
    return run_updates(b, 1, nsteps, dt, damping, sqrt(2.0*damping*T));
}
int tcym_batch_flow(tcym_batch* b, int u, int nsteps, double dtau) {
    if (!b) return TCYM_E_ARG;
    if (nsteps < 0) return b->fail(TCYM_E_ARG, "nsteps < 0");
    int r = b->flow(u, nsteps, dtau);
    return r ? r : b->sync();
}

int tcym_batch_evolve(tcym_batch* b, int mode, int nseg, int steps, double dt, double damping, double T,
                      uint64_t seed0, double* ncsdot) {
    if (!b) return TCYM_E_ARG;
    if ((mode != 0 && mode != 1) || nseg < 0 || steps < 1) return b->fail(TCYM_E_ARG, "evolve: mode 0 or 1, nseg >= 0, steps >= 1");
    if (mode == 1 && (!(T >= 0) || !(damping >= 0))) return b->fail(TCYM_E_ARG, "negative temperature or damping");
    int r;
    if ((r = b->useU(TCYM_U_HOT)) || (r = b->useE(TCYM_E_HOT)) || (r = b->scratch())) return r;
    if (mode == 1 && (r = b->set_seed(seed0))) return r;
    if (nseg == 0) return TCYM_OK;
    // trajectory columns and the column counter
    if (nseg > b->traj_cap) {
        if (b->traj) {
            auto it = std::find(b->owned.begin(), b->owned.end(), (void*)b->traj);
            if (it != b->owned.end()) b->owned.erase(it);
            hipFree(b->traj);
            b->traj = nullptr;
            b->drop_graphs();
        }
        if ((r = b->alloc((void**)&b->traj, sizeof(double)*(size_t)nseg*b->count, "hipMalloc(trajectory)"))) return r;
        b->traj_cap = nseg;
    }
    uint32_t c[2] = {b->step_host, 0};
    if ((r = b->hip(hipMemcpyAsync(b->counters, c, sizeof(c), hipMemcpyHostToDevice, b->stream), "hipMemcpyAsync(counters)"))) return r;
    if ((r = b->sync())) return r;
    const double noise = mode == 1 ? sqrt(2.0*damping*T) : 0.0;
    for (int j = 0; j < nseg; j++) {
        // the graph bakes the buffer pair and the coefficients into its nodes
        char key[256];
        snprintf(key, sizeof(key), "%d %d %.17g %.17g %.17g %p %p %p", mode, steps, dt, damping, noise,
                 b->U[TCYM_U_HOT], b->Uscr, (void*)b->traj);
        auto it = b->graphs.find(key);
        if (it == b->graphs.end()) {
            tcym_batch::Graph gr;
            void* u0 = b->U[TCYM_U_HOT];
            void* s0 = b->Uscr;
            if ((r = b->hip(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))) return r;
            b->capturing = true;
            for (int k = 0; k < steps; k++) b->step(mode, TCYM_U_HOT, TCYM_E_HOT, true, dt, damping, noise, (uint32_t)k, k == steps - 1);
            b->be->edotb_batch(b->L(), b->dev(), b->U[TCYM_U_HOT], b->E[TCYM_E_HOT], b->part);
            b->final_reduce(true, 1, 0u);
            const long cap = b->traj_cap;
            b->launch("batch_segment_end", [&](hipStream_t st) {
                hipLaunchKernelGGL(k_segment_end, dim3(1), dim3(256), 0, st, (const double*)b->red, b->traj, cap,
                                   b->count, b->counters, mode == 1 ? (uint32_t)steps : 0u);
            });
            b->capturing = false;
            // capture leaves the host's buffer pair as after `steps` swaps; restore it for the replay
            b->U[TCYM_U_HOT] = u0;
            b->Uscr = s0;
            hipError_t ce = hipStreamEndCapture(b->stream, &gr.graph);
            if ((r = b->hip(ce, "hipStreamEndCapture"))) return r;
            if (b->last != hipSuccess) { hipGraphDestroy(gr.graph); return TCYM_E_HIP; }
            if ((r = b->hip(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0), "hipGraphInstantiate"))) {
                hipGraphDestroy(gr.graph);
                return r;
            }
            it = b->graphs.emplace(key, gr).first;
        }
        if ((r = b->hip(hipGraphLaunch(it->second.exec, b->stream), "hipGraphLaunch"))) return r;
        if (steps & 1) std::swap(b->U[TCYM_U_HOT], b->Uscr);
        if (mode == 1) b->step_host += (uint32_t)steps;   // only Langevin steps draw noise
    }
    if ((r = b->sync())) return r;
    if (ncsdot) {
        std::vector<double> h((size_t)nseg*b->count);
        for (int e = 0; e < b->count; e++)
            if ((r = b->hip(hipMemcpy(&h[(size_t)e*nseg], b->traj + (size_t)e*b->traj_cap, sizeof(double)*nseg, hipMemcpyDeviceToHost), "hipMemcpy(trajectory)"))) return r;
        memcpy(ncsdot, h.data(), sizeof(double)*h.size());
    }
    return TCYM_OK;
}

// ---------------------------------------------------------------- cooling
int tcym_batch_cooling_start(tcym_batch* b, int nsteps, double dtau) {
    if (!b) return TCYM_E_ARG;
    int r;
    if ((r = b->flow(TCYM_U_NEW, nsteps, dtau, TCYM_U_HOT))) return r;
    if ((r = b->copy_links(TCYM_U_OLD, TCYM_U_NEW))) return r;
    return b->sync();
}
int tcym_batch_cooling_update(tcym_batch* b, int nsteps, double dtau, double* dncs) {
    if (!b) return TCYM_E_ARG;
    int r;
    if ((r = b->flow(TCYM_U_NEW, nsteps, dtau, TCYM_U_HOT))) return r;
    if ((r = b->useU(TCYM_U_OLD)) || (r = b->useU(TCYM_U_MID)) || (r = b->useE(TCYM_E_MID))) return r;
    b->be->geodesic_batch(b->L(), b->dev(), 0.5, b->U[TCYM_U_OLD], b->U[TCYM_U_NEW], b->U[TCYM_U_MID], b->E[TCYM_E_MID]);
    // CoolingMethod::GetDeltaNCs(1): Simpson (1,4,1)/6 (TOPOLOGY/CoolingMethod.cpp:80-121),
    // accumulated in the order of the single-lattice tcym_delta_ncs_cooling
    std::vector<double> xm(b->count), xo(b->count), xn(b->count);
    if ((r = b->ncsdot(TCYM_U_MID, TCYM_E_MID, xm.data()))) return r;
    if ((r = b->ncsdot(TCYM_U_OLD, TCYM_E_MID, xo.data()))) return r;
    if ((r = b->ncsdot(TCYM_U_NEW, TCYM_E_MID, xn.data()))) return r;
    for (int e = 0; e < b->count; e++) {
        double v = 0.0, w = 0.0;
        v += 4.0*xm[e]; w += 4.0;
        v += xo[e]; w += 1.0;
        v += xn[e]; w += 1.0;
        if (dncs) dncs[e] = v/w;
    }
    if ((r = b->copy_links(TCYM_U_OLD, TCYM_U_NEW))) return r;
    return b->sync();
}

// ---------------------------------------------------------------- measurements
int tcym_batch_ncsdot(tcym_batch* b, int u, int e, double* out) {
    if (!b || !out) return TCYM_E_ARG;
    return b->ncsdot(u, e, out);
}
int tcym_batch_bulk(tcym_batch* b, int u, int e, double* out) {
    if (!b || !out) return TCYM_E_ARG;
    int r;
    if ((r = b->useU(u)) || (r = b->useE(e))) return r;
    b->be->bulk_batch(b->L(), b->dev(), b->U[u], b->E[e], b->part);
    b->final_reduce(false, 6, 0u);
    if (b->last != hipSuccess) return TCYM_E_HIP;
    if ((r = b->fetch(6, out))) return r;
    // per-volume normalisation of Observables::Bulk (OBSERVABLES/BulkObservables.cpp:17-147), as tcym_bulk
    for (int l = 0; l < b->count; l++)
        for (int q = 0; q < 6; q++) out[6*l + q] /= (double)b->geo[l].V;
    return TCYM_OK;
}
int tcym_batch_gauss(tcym_batch* b, int u, int e, double* out) {
    if (!b || !out) return TCYM_E_ARG;
    int r;
    if ((r = b->useU(u)) || (r = b->useE(e))) return r;
    b->be->gauss_batch(b->L(), b->dev(), b->U[u], b->E[e], b->part);
    b->final_reduce(false, 2, 1u);
    if (b->last != hipSuccess) return TCYM_E_HIP;
    if ((r = b->fetch(2, out))) return r;
    for (int l = 0; l < b->count; l++) out[2*l + 1] = sqrt(out[2*l + 1]/(double)(b->geo[l].V*b->be->ne));
    return TCYM_OK;
}
int tcym_batch_unitarity(tcym_batch* b, int u, double* out) {
    if (!b || !out) return TCYM_E_ARG;
    int r;
    if ((r = b->useU(u))) return r;
    b->be->unitarity_batch(b->L(), b->dev(), b->U[u], b->part);
    b->final_reduce(false, 2, 1u);
    if (b->last != hipSuccess) return TCYM_E_HIP;
    if ((r = b->fetch(2, out))) return r;
    for (int l = 0; l < b->count; l++) out[2*l + 1] = sqrt(out[2*l + 1]/(3.0*b->geo[l].V));
    return TCYM_OK;
}

// ---------------------------------------------------------------- autotuner
// Times every candidate walk shape x persistent grid of one pass on scratch
// copies of the fields and keeps the fastest for that pass.  Leapfrog and
// Langevin (mode 0, 1): reps launches on the same input.  Flow (mode 3): the
// cadence of a cooling episode, 4 reps steps that alternate between two link
// buffers with columns 0 and 1 stored between them.
int tcym_batch_tune(tcym_batch* b, int mode, double dt, int reps, int choice[3], double* us) {
    if (!b) return TCYM_E_ARG;
    if (mode != 0 && mode != 1 && mode != 3) return b->fail(TCYM_E_ARG, "tune: mode 0, 1 or 3");
    if (reps < 1) reps = 3;
    const int pass = mode == 3 ? 1 : 0;
    int r;
    if ((r = b->useU(TCYM_U_HOT)) || (r = b->useU(TCYM_U_AUX)) || (r = b->useE(TCYM_E_HOT)) || (r = b->useE(TCYM_E_AUX)) ||
        (r = b->scratch()))
        return r;
    // candidates run on copies: E_AUX (a copy of E_HOT) in place; links
    // U_HOT -> scratch (leapfrog, Langevin) or U_AUX <-> scratch (flow)
    b->be->copy(b->L(), b->e_elems, b->E[TCYM_E_HOT], b->E[TCYM_E_AUX]);
    b->be->copy(b->L(), b->u_elems, b->U[TCYM_U_HOT], b->U[TCYM_U_AUX]);
    if ((r = b->sync())) return r;
    const int shape0[2] = {b->shape[0], b->shape[1]};
    const bool timing0 = b->timing;
    b->timing = false;
    hipEvent_t e0 = b->ev(), e1 = b->ev();
    double best = 1e300;
    int bs = shape0[pass], bp = b->per_cu[pass];
    bool bx = b->xcd_split[pass];
    const int nflow = 4*reps;
    auto run = [&](const tcym::BatchDev& d, int n) {
        void* a = b->U[TCYM_U_AUX];
        void* c = b->Uscr;
        for (int k = 0; k < n; k++) {
            if (mode == 3) {
                b->be->update_batch(b->L(), d, 3, false, dt, 0.0, 0.0, 0, a, c, nullptr, nullptr, k == n - 1);
                std::swap(a, c);
            } else {
                b->be->update_batch(b->L(), d, mode, true, dt, 0.1, mode == 1 ? 0.1 : 0.0, 0, b->U[TCYM_U_HOT], b->Uscr,
                                    b->E[TCYM_E_AUX], nullptr, false);
            }
        }
    };
    std::vector<int> shapes = {-1, 0, 4, 5, 6};
    if (b->plaq_ok())
        for (int s = kShapePlaq; s < kShapePlaq + 3; s++) shapes.push_back(s);
    for (int s : shapes) {
        if ((r = pass == 0 ? b->build_tables(s, shape0[1]) : b->build_tables(shape0[0], s))) break;
        for (int pc : {2, 4, 8, 16}) {
            for (bool xs : {true, false}) {
                if (!xs && pc != 8) continue;
                b->per_cu[pass] = pc;
                b->xcd_split[pass] = xs;
                const tcym::BatchDev d = b->dev();
                run(d, mode == 3 ? 2 : 1);     // warm-up
                hipEventRecord(e0, b->stream);
                run(d, mode == 3 ? nflow : reps);
                hipEventRecord(e1, b->stream);
                if ((r = b->sync())) break;
                float ms = 0.f;
                hipEventElapsedTime(&ms, e0, e1);
                const double t = 1e3*ms/(mode == 3 ? nflow : reps);
                if (t < best) { best = t; bs = s; bp = pc; bx = xs; }
            }
            if (r) break;
        }
        if (r) break;
    }
    b->free_events.push_back(e0);
    b->free_events.push_back(e1);
    b->timing = timing0;
    if (r) return r;
    b->per_cu[pass] = bp;
    b->xcd_split[pass] = bx;
    if ((r = pass == 0 ? b->build_tables(bs, shape0[1]) : b->build_tables(shape0[0], bs))) return r;
    if (choice) { choice[0] = bs; choice[1] = bp; choice[2] = bx ? 1 : 0; }
    if (us) *us = best;
    return TCYM_OK;
}

int tcym_batch_set_schedule(tcym_batch* b, int mode, int shape, int blocks_per_cu, int xcd_split) {
    if (!b) return TCYM_E_ARG;
    if (mode != 0 && mode != 1 && mode != 3) return b->fail(TCYM_E_ARG, "schedule: mode 0, 1 or 3");
    const bool plaq = shape >= kShapePlaq && shape < kShapePlaq + 3;
    if (!(shape == -1 || shape == 0 || (shape >= 4 && shape <= 6) || plaq) || blocks_per_cu < 1 || blocks_per_cu > 64)
        return b->fail(TCYM_E_ARG, "schedule: shape -1, 0, 4..6 or 7..9, 1..64 blocks per compute unit");
    const int pass = mode == 3 ? 1 : 0;
    int r;
    if ((r = b->resolve())) return r;
    b->per_cu[pass] = blocks_per_cu;
    b->xcd_split[pass] = xcd_split != 0;
    return pass == 0 ? b->build_tables(shape, b->shape[1]) : b->build_tables(b->shape[0], shape);
}

// ---------------------------------------------------------------- timing
int tcym_batch_timing(tcym_batch* b, int enable) {
    if (!b) return TCYM_E_ARG;
    int r = b->resolve();
    b->timing = enable != 0;
    return r;
}
int tcym_batch_timing_reset(tcym_batch* b) {
    if (!b) return TCYM_E_ARG;
    int r = b->resolve();
    b->stats.clear();
    return r;
}
int tcym_batch_kernel_time(tcym_batch* b, const char* name, double* total_ms, long long* launches) {
    if (!b || !name) return TCYM_E_ARG;
    int r = b->resolve();
    if (r) return r;
    auto it = b->stats.find(name);
    if (total_ms) *total_ms = it == b->stats.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == b->stats.end() ? 0 : it->second.second;
    return TCYM_OK;
}

}  // extern "C"
