// Internal interface between the C-ABI context (tcym_api.hip) and the
// group/precision-specialised kernel launchers (tcym_kernels_su{2,3}.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <functional>

#include "tcym.h"

namespace tcym {

struct Geom {
    int n[3];
    long V;            // sites
    double a[3];       // lattice spacings
    double aScale;
    int form;          // update kernel form (TCYM_FORM_*)
};

// one kernel launch; the context wraps it with optional HIP-event timing
using Launch = std::function<void(hipStream_t)>;
using Launcher = std::function<void(const char* name, const Launch&)>;

// n / d for 0 <= n < 2^31 as a multiply-high and a shift: m = ceil(2^(31+l) / d)
// with l = ceil(log2 d) (below 2^32), s = l - 1; n * m / 2^(31+l) exceeds
// n / d by less than n / 2^31 / d, so the floor is exact.  m = 0 marks d = 1.
struct FastDiv {
    uint32_t m, s;
};
inline FastDiv fast_div(uint32_t d) {
    if (d <= 1) return FastDiv{0, 0};
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    const uint64_t m = ((1ull << (31 + l)) + d - 1)/d;
    return FastDiv{(uint32_t)m, l - 1};
}

// geometry as the kernels see it: extents, sites and the walk shape
struct KGeom {
    int nx, ny, nz;
    long V;
    int pxs, tcx, tcy;   // traversal shape (WalkShape); pxs == 0 is the linear walk
    FastDiv dnx, dnxy;   // division by nx and nx ny (linear walk)
    FastDiv dnz, dtcx;   // division by nz and tcx (patch walk)
};

// Rows of a block of nt threads in the plaquette form of the update: the
// largest divisor py of ny with 2 <= py and nx py <= nt (ny itself when the
// whole plane fits), 0 if there is none.
inline int plaq_rows(int nx, int ny, int nt) {
    for (int py = nt/nx; py >= 2; py--)
        if (ny % py == 0) return py;
    return 0;
}

// One lattice of a batch (device table).  Each pass over a batch has its own
// tiling: `gu`/`utile0` for the leapfrog and Langevin updates and `gf`/`ftile0`
// for the flow (the autotuner picks each walk shape), `gr`/`ltile0`/`stile0`
// for the measurement kernels (the shape a single lattice of the same size
// uses, so that their sums are bit-identical to the single-lattice ones).
struct BatchEns {
    KGeom gu, gf, gr;
    uint64_t u_off, e_off;     // element offsets of the lattice's link / E records
    uint64_t seed;             // Langevin noise key
    uint32_t utile0;           // first tile of its leapfrog / Langevin pass (3 per site block)
    uint32_t ftile0;           // first tile of its flow pass
    uint32_t ltile0;           // first tile of its measurement pass over links
    uint32_t stile0;           // first tile of its measurement pass over sites
    // plaquette form of the update (u) and flow (f) passes, where the
    // autotuner picked it: rows and planes of a unit, bands along y (the
    // pass's tile table then lists units, from utile0 / ftile0)
    int32_t upy, upz, unyb, fpy, fpz, fnyb;
    FastDiv udnyb, fdnyb;      // division by unyb / fnyb (a unit's slab and band)
};

// device tables of a batch handed to the kernels (owned by the batch object)
struct BatchDev {
    Geom g;                            // spacings (common to the batch); extents unused
    const BatchEns* ens;
    const uint32_t* utiles; uint32_t n_utiles;   // tile -> lattice, leapfrog / Langevin pass
    const uint32_t* ftiles; uint32_t n_ftiles;   // tile -> lattice, flow pass
    const uint32_t* ltiles; uint32_t n_ltiles;   // tile -> lattice, measurement pass over links
    const uint32_t* stiles; uint32_t n_stiles;   // tile -> lattice, measurement pass over sites
    int grid_update;                   // persistent grid of the leapfrog / Langevin kernels (blocks)
    int grid_flow;                     // persistent grid of the flow kernel
    int grid_measure;                  // persistent grid of the measurement kernels
    const uint32_t* step;              // device step counter (Philox step of step offset 0)
    bool plaq_update, plaq_flow;       // plaquette form of the pass (its table lists plaquette units)
};

struct Backend {
    int ns = 0, ne = 0, elsize = 0, nc = 0;
    virtual ~Backend() {}
    // layout conversions: host AoS FP64 (staged on the device) <-> device tiled layout
    virtual void pack_links(const Launcher&, const Geom&, const double* aos, void* soa) = 0;
    virtual void unpack_links(const Launcher&, const Geom&, const void* soa, double* aos) = 0;
    virtual void pack_efield(const Launcher&, const Geom&, const double* aos, void* soa) = 0;
    virtual void unpack_efield(const Launcher&, const Geom&, const void* soa, double* aos) = 0;
    // storage type <-> FP64 for fields of nr reals per (site, mu); the tiled
    // layout depends on the storage type, so these are not element-wise
    virtual void to_f64(const Launcher&, const Geom&, int nr, const void* src, double* dst) = 0;
    virtual void copy(const Launcher&, long n, const void* src, void* dst) = 0;   // n elements of the storage type
    virtual void from_f64(const Launcher&, const Geom&, int nr, const double* src, void* dst) = 0;
    // seeded fields
    virtual void random_links(const Launcher&, const Geom&, uint64_t seed, double sigma, void* U) = 0;
    virtual void random_efield(const Launcher&, const Geom&, uint64_t seed, uint32_t stream, uint32_t step, double sigma, void* E) = 0;
    // updates; Uin -> Uout (distinct buffers), E in place
    // mode: 0 leapfrog, 1 Langevin (Philox noise), 2 Langevin (Xi field), 3 flow (E := force)
    // complete = false stores only columns 0 and 1 of SU(3) links: for an
    // output that no kernel but the next update step reads (the update
    // kernels load links as columns 0, 1 and rebuild column 2); the last step
    // of a sequence stores complete links
    virtual void update(const Launcher&, const Geom&, int mode, bool links, bool write_e, double dt,
                        double damping, double noise_sd, uint64_t seed, uint32_t step,
                        const void* Uin, void* Uout, void* E, const void* Xi, bool complete) = 0;
    // U <- exp(i cU_mu E) U (Uin -> Uout)
    virtual void link_update(const Launcher&, const Geom&, const double* cU, const void* Uin, void* Uout, const void* E) = 0;
    // reductions: per-block partials (double) written to `part`; return number of blocks
    virtual long edotb(const Launcher&, const Geom&, int improved, const void* U, const void* E, double* part, double* map) = 0;
    // the improved E.B partials launched straight onto a stream (no launcher object, no allocation)
    // improved E.B with the field strength B of (U) formed (bmode 0), formed
    // and kept in Bc (1), or taken from Bc (2); block partials as edotb
    virtual long edotb_cached(const Launcher&, const Geom&, const void* U, const void* E, double* part, void* Bc, int bmode) = 0;
    virtual long edotb_on(hipStream_t, const Geom&, const void* U, const void* E, double* part) = 0;
    virtual long bulk(const Launcher&, const Geom&, const void* U, const void* E, double* part) = 0;
    virtual long gauss(const Launcher&, const Geom&, const void* U, const void* E, double* part, double* G) = 0;
    virtual long unitarity(const Launcher&, const Geom&, const void* U, double* part) = 0;
    virtual long hard_scales(const Launcher&, const Geom&, const void* U, double* part) = 0;   // 3 quantities per link tile
    virtual long vacuum_ncs(const Launcher&, const Geom&, const void* U, double* part) = 0;
    // Gauss-law restoration (FP64 buffers only)
    virtual void gauss_correct(const Launcher&, const Geom&, const double* U, double* E, const double* G, double gamma) = 0;
    virtual void geodesic(const Launcher&, const Geom&, double c, const void* Uold, const void* Unew, void* Umid, void* Emid) = 0;
    virtual void reunitarize(const Launcher&, const Geom&, void* U) = 0;
    // links of the fine lattice g -> the lattice of half the extent (Uc distinct from Uf)
    virtual void block_links(const Launcher&, const Geom& g, const void* Uf, void* Uc) = 0;
    // quasi-particle initial condition: the 6 NE momentum-space modes (A_mu, E_mu
    // of every colour, complex FP64, [field][colour][p]) and, after their
    // inverse FFT, the links and fields from the real parts
    virtual void qp_modes(const Launcher&, const Geom& g, double Qs, double n0, uint64_t seed, double* P) = 0;
    virtual void qp_fields(const Launcher&, const Geom& g, const double* X, void* U, void* E) = 0;

    // batches of lattices (tcym_batch.hip): one persistent launch per pass
    // over every lattice of the batch, same per-link / per-site arithmetic as
    // the single-lattice kernels.  Partials go to part[tile*nq + q].
    virtual void update_batch(const Launcher&, const BatchDev&, int mode, bool write_e, double dt, double damping,
                              double noise_sd, uint32_t step_off, const void* Uin, void* Uout, void* E, const void* Xi,
                              bool complete) = 0;
    virtual void edotb_batch(const Launcher&, const BatchDev&, const void* U, const void* E, double* part) = 0;
    virtual void bulk_batch(const Launcher&, const BatchDev&, const void* U, const void* E, double* part) = 0;
    virtual void gauss_batch(const Launcher&, const BatchDev&, const void* U, const void* E, double* part) = 0;
    virtual void unitarity_batch(const Launcher&, const BatchDev&, const void* U, double* part) = 0;
    virtual void geodesic_batch(const Launcher&, const BatchDev&, double c, const void* Uo, const void* Un, void* Um, void* Em) = 0;
};

Backend* make_backend_su2(int fp32);
Backend* make_backend_su3(int fp32);

// final deterministic reduction of nblk x nq partials -> out[nq] (sum or max per quantity)
void final_reduce(const Launcher&, const double* part, long nblk, int nq, const int* is_max, double* out_dev);

constexpr int kBlock = 256;

// sites of a device field: the tiled layout stores sites in tiles of 64
inline long padded_sites(long V) { return (V + 63) & ~63L; }

// Traversal of the lattice by the stencil kernels.  Patch walk (pxs > 0): a
// block covers a patch of (1 << pxs) sites in x by (kBlock >> pxs) rows in y
// of one z-plane; the patches stacked in y make a column kColRows sites high,
// and a column is swept through all z before the next column.  pxs is 6, 5
// or 4, whichever pads Nx least among the patches that fit in a column (ties
// go to the wider patch, for coalescing).  Linear walk (pxs == 0): a block
// covers kBlock consecutive sites in natural order; it is used when it pads
// less than every patch shape (small or odd lattices).
constexpr int kColRows = 16;
struct WalkShape {
    int pxs;           // log2 of the patch width in x, 0 for the linear walk
    int tcx, tcy;      // columns in x and y (patch walk)
    long blocks;       // blocks for one pass over the sites
};
inline WalkShape patch_shape(const Geom& g, int pxs) {
    WalkShape w{pxs, 0, 0, 0};
    w.tcx = (g.n[0] + (1 << pxs) - 1) >> pxs;
    w.tcy = (g.n[1] + kColRows - 1)/kColRows;
    w.blocks = (long)w.tcx*w.tcy*g.n[2]*(kColRows/(kBlock >> pxs));
    return w;
}
inline WalkShape linear_shape(const Geom& g) { return WalkShape{0, 0, 0, (g.V + kBlock - 1)/kBlock}; }
// the shape the measurement kernels use (and the default of the update kernels)
inline WalkShape walk_shape(const Geom& g);
inline KGeom make_kgeom(const Geom& g, const WalkShape& w) {
    return KGeom{g.n[0], g.n[1], g.n[2], g.V, w.pxs, w.tcx, w.tcy,
                 fast_div((uint32_t)g.n[0]), fast_div((uint32_t)g.n[0]*(uint32_t)g.n[1]),
                 fast_div((uint32_t)g.n[2]), fast_div((uint32_t)(w.tcx > 0 ? w.tcx : 1))};
}
inline WalkShape walk_shape(const Geom& g) {
    WalkShape best{-1, 0, 0, 0};
    for (int s = 6; s >= 4; s--) {
        if ((kBlock >> s) > kColRows) continue;          // a patch must fit in a column
        const WalkShape w = patch_shape(g, s);
        if (best.pxs < 0 || w.blocks < best.blocks) best = w;
    }
    const WalkShape lin = linear_shape(g);
    return lin.blocks < best.blocks ? lin : best;
}

}  // namespace tcym
