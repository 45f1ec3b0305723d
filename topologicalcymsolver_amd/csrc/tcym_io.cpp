// Configuration files (include/tcym_io.h): the reference's text
// configurations (IO/SaveConfiguration.cpp, IO/LoadConfiguration.cpp) and the
// library's binary checkpoints.  Host code: formatting and parsing are split
// over threads (records are independent lines), the device is reached only
// through the C ABI of tcym.h.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tcym.h"
#include "tcym_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

int nthreads_of(int n) {
    if (n > 0) return std::min(n, 256);
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

// one record per (site, mu): nr reals
struct Shape {
    int N[3];
    long V;
    int nr;
    bool links;
};

int make_shape(int nc, const int* N, bool links, Shape& s) {
    if (nc != 2 && nc != 3) return fail(TCYM_E_ARG, "group must be SU(2) or SU(3)");
    if (!N) return fail(TCYM_E_ARG, "null extents");
    for (int k = 0; k < 3; k++) {
        if (N[k] < 1) return fail(TCYM_E_ARG, "extents must be positive");
        s.N[k] = N[k];
    }
    s.V = (long)N[0]*N[1]*N[2];
    s.nr = links ? (nc == 2 ? 4 : 18) : (nc == 2 ? 3 : 8);
    s.links = links;
    return TCYM_OK;
}

template <class T>
inline char* put_int(char* p, char* end, T v) {
    return std::to_chars(p, end, v).ptr;
}

// records [r0, r1) of A as text lines: "x y z mu" then " v" per value, and for
// links a trailing space (IO/SaveConfiguration.cpp:25 with MatrixToString,
// SUNc/SU2/GroupOperations.cpp:260-272, SUNc/SU3/GroupOperations.cpp:539-551;
// fields :38-44).  to_chars with chars_format::general and a precision is
// printf "%.<digits>g", which is what an ostream with precision(digits) and
// the default floating format writes.
void format_records(const Shape& s, const double* A, long r0, long r1, int digits, std::string& out) {
    const size_t line_max = 4*12 + (size_t)s.nr*(digits + 10) + 4;
    out.resize((size_t)(r1 - r0)*line_max);
    char* p = &out[0];
    char* const end = p + out.size();
    const long nxy = (long)s.N[0]*s.N[1];
    for (long r = r0; r < r1; r++) {
        const long site = r/3;
        const int mu = (int)(r - 3*site);
        const long z = site/nxy, rem = site - z*nxy, y = rem/s.N[0], x = rem - y*s.N[0];
        p = put_int(p, end, x); *p++ = ' ';
        p = put_int(p, end, y); *p++ = ' ';
        p = put_int(p, end, z); *p++ = ' ';
        p = put_int(p, end, mu);
        const double* v = A + r*s.nr;
        for (int k = 0; k < s.nr; k++) {
            *p++ = ' ';
            p = std::to_chars(p, end, v[k], std::chars_format::general, digits).ptr;
        }
        if (s.links) *p++ = ' ';
        *p++ = '\n';
    }
    out.resize((size_t)(p - &out[0]));
}

int write_text(const Shape& s, const double* A, const char* path, int digits, int nthreads) {
    if (!A || !path) return fail(TCYM_E_ARG, "null array or path");
    if (digits < 1 || digits > 17) return fail(TCYM_E_ARG, "digits must be 1..17");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(TCYM_E_IO, std::string("cannot open ") + path + " for writing: " + strerror(errno));
    const int T = nthreads_of(nthreads);
    const long nrec = 3*s.V, chunk = 1 << 15;
    const long nchunk = (nrec + chunk - 1)/chunk;
    // double-buffered batches of T chunks: batch b+1 is formatted while batch b is written
    std::vector<std::string> buf[2];
    buf[0].resize(T);
    buf[1].resize(T);
    auto format_batch = [&](long b, std::vector<std::string>& out) {
        std::vector<std::thread> th;
        for (int t = 0; t < T; t++) {
            const long c = b*T + t;
            if (c >= nchunk) { out[t].clear(); continue; }
            th.emplace_back([&, c, t] { format_records(s, A, c*chunk, std::min(nrec, (c + 1)*chunk), digits, out[t]); });
        }
        for (auto& x : th) x.join();
    };
    const long nbatch = (nchunk + T - 1)/T;
    bool ok = true;
    if (nbatch > 0) format_batch(0, buf[0]);
    for (long b = 0; b < nbatch && ok; b++) {
        std::thread next;
        if (b + 1 < nbatch) next = std::thread([&, b] { format_batch(b + 1, buf[(b + 1) & 1]); });
        for (auto& str : buf[b & 1])
            if (ok && !str.empty() && fwrite(str.data(), 1, str.size(), f) != str.size()) ok = false;
        if (next.joinable()) next.join();
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) return fail(TCYM_E_IO, std::string("write failed on ") + path + ": " + strerror(errno));
    return TCYM_OK;
}

// read-only mapping of a whole file
struct Mapped {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapped() {
        if (p && n) munmap((void*)p, n);
        if (fd >= 0) close(fd);
    }
    int open_file(const char* path) {
        fd = open(path, O_RDONLY);
        if (fd < 0) return fail(TCYM_E_IO, std::string("cannot open ") + path + ": " + strerror(errno));
        struct stat st;
        if (fstat(fd, &st) != 0) return fail(TCYM_E_IO, std::string("cannot stat ") + path);
        n = (size_t)st.st_size;
        if (n == 0) return TCYM_OK;
        void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) { n = 0; return fail(TCYM_E_IO, std::string("cannot map ") + path); }
        madvise(m, n, MADV_SEQUENTIAL);
        p = (const char*)m;
        return TCYM_OK;
    }
};

inline bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

template <class T>
inline bool get_num(const char*& p, const char* e, T& v) {
    while (p < e && blank(*p)) p++;
    if (p < e && *p == '+') p++;   // an istream accepts a leading '+', from_chars does not
    if (p >= e) return false;
    auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc() || (r.ptr < e && !blank(*r.ptr))) return false;
    p = r.ptr;
    return true;
}

// IO::LoadConfiguration: every non-empty line is "x y z mu v_0 .. v_{nr-1}",
// placed by its coordinates (IO/LoadConfiguration.cpp:46-72 links, :84-122 fields)
int read_text(const Shape& s, double* A, const char* path, int nthreads) {
    if (!A || !path) return fail(TCYM_E_ARG, "null array or path");
    Mapped m;
    int r = m.open_file(path);
    if (r) return r;
    const long nrec = 3*s.V;
    std::unique_ptr<std::atomic<unsigned char>[]> seen(new std::atomic<unsigned char>[nrec]);
    for (long i = 0; i < nrec; i++) seen[i].store(0, std::memory_order_relaxed);
    const int T = (int)std::max<size_t>(1, std::min<size_t>(nthreads_of(nthreads), m.n/(1 << 16) + 1));
    std::vector<size_t> cut(T + 1);
    cut[0] = 0;
    cut[T] = m.n;
    for (int t = 1; t < T; t++) {
        size_t c = m.n*(size_t)t/T;
        while (c < m.n && m.p[c - 1] != '\n') c++;
        cut[t] = std::max(c, cut[t - 1]);
    }
    std::atomic<long> count{0};
    std::atomic<bool> bad{false};
    std::mutex mu_err;
    std::string err;
    auto report = [&](size_t off, const std::string& what) {
        std::lock_guard<std::mutex> g(mu_err);
        if (!bad.exchange(true)) err = std::string(path) + ": " + what + " in the line at byte " + std::to_string(off);
    };
    auto work = [&](int t) {
        const char* p = m.p + cut[t];
        const char* const stop = m.p + cut[t + 1];
        long n = 0;
        while (p < stop && !bad.load(std::memory_order_relaxed)) {
            const char* e = (const char*)memchr(p, '\n', (size_t)(stop - p));
            if (!e) e = stop;
            const char* q = p;
            while (q < e && blank(*q)) q++;
            if (q < e) {
                const size_t off = (size_t)(p - m.p);
                long c[4];
                bool ok = true;
                for (int k = 0; k < 4 && ok; k++) ok = get_num(q, e, c[k]);
                if (!ok) { report(off, "malformed coordinates"); break; }
                if (c[0] < 0 || c[0] >= s.N[0] || c[1] < 0 || c[1] >= s.N[1] || c[2] < 0 || c[2] >= s.N[2] ||
                    c[3] < 0 || c[3] > 2) {
                    report(off, "coordinates outside the lattice");
                    break;
                }
                const long idx = c[3] + 3*(c[0] + s.N[0]*(c[1] + (long)s.N[1]*c[2]));
                double* v = A + idx*s.nr;
                for (int k = 0; k < s.nr && ok; k++) ok = get_num(q, e, v[k]);
                if (!ok) { report(off, "missing or malformed value"); break; }
                while (q < e && blank(*q)) q++;
                if (q < e) { report(off, "more values than the group has"); break; }
                if (seen[idx].exchange(1)) { report(off, "repeated record"); break; }
                n++;
            }
            p = e + 1;
        }
        count += n;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    if (bad) return fail(TCYM_E_IO, err);
    if (count.load() != nrec)
        return fail(TCYM_E_IO, std::string(path) + ": " + std::to_string(count.load()) + " records, expected " +
                                   std::to_string(nrec));
    return TCYM_OK;
}

// ---------------------------------------------------------------- checkpoints
constexpr size_t kHeader = 512;
constexpr char kMagic[8] = {'T', 'C', 'Y', 'M', 'C', 'K', 'P', 'T'};
constexpr uint32_t kVersion = 1;

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t fmix(uint64_t h) {
    h ^= h >> 33; h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}
// 64-bit hash of n bytes: 64-bit words mixed per 1 MiB block, block hashes
// chained in order (blocks are hashed in parallel)
uint64_t hash_bytes(const void* data, size_t n, int nthreads) {
    const unsigned char* b = (const unsigned char*)data;
    const size_t block = 1 << 20;
    const size_t nb = (n + block - 1)/block;
    std::vector<uint64_t> hb(nb);
    auto one = [&](size_t i) {
        const size_t lo = i*block, hi = std::min(n, lo + block);
        uint64_t h = fmix(0x9e3779b97f4a7c15ULL + i);
        size_t k = lo;
        for (; k + 8 <= hi; k += 8) {
            uint64_t w;
            memcpy(&w, b + k, 8);
            h = rotl(h ^ (w*0x87c37b91114253d5ULL), 31)*0x4cf5ad432745937fULL + 0x52dce729ULL;
        }
        if (k < hi) {
            uint64_t w = 0;
            memcpy(&w, b + k, hi - k);
            h = rotl(h ^ (w*0x87c37b91114253d5ULL), 31)*0x4cf5ad432745937fULL + 0x52dce729ULL;
        }
        hb[i] = fmix(h ^ (hi - lo));
    };
    const int T = std::max(1, std::min<int>(nthreads_of(nthreads), (int)nb));
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([&, t] { for (size_t i = t; i < nb; i += T) one(i); });
    for (auto& x : th) x.join();
    uint64_t h = fmix(n);
    for (uint64_t x : hb) h = fmix(h*0x9e3779b97f4a7c15ULL + x);
    return h;
}

struct Header {
    tcym_checkpoint_info info;
    int has_e;
    uint64_t nu, ne;
    uint64_t hu, he;
};

template <class T>
void put(unsigned char* h, size_t off, const T& v) { memcpy(h + off, &v, sizeof(T)); }
template <class T>
void get(const unsigned char* h, size_t off, T& v) { memcpy(&v, h + off, sizeof(T)); }

void encode(const Header& H, unsigned char* h) {
    memset(h, 0, kHeader);
    memcpy(h, kMagic, 8);
    put(h, 8, kVersion);
    put(h, 12, (uint32_t)kHeader);
    put(h, 16, (int32_t)H.info.nc);
    put(h, 20, (int32_t)H.info.precision);
    for (int k = 0; k < 3; k++) put(h, 24 + 4*k, (int32_t)H.info.N[k]);
    put(h, 36, (int32_t)H.has_e);
    for (int k = 0; k < 3; k++) put(h, 40 + 8*k, H.info.a[k]);
    put(h, 64, H.info.aScale);
    put(h, 72, H.info.step);
    put(h, 80, H.info.seed);
    put(h, 88, H.info.time);
    for (int k = 0; k < 8; k++) put(h, 96 + 8*k, H.info.acc[k]);
    put(h, 160, H.nu);
    put(h, 168, H.ne);
    put(h, 176, H.hu);
    put(h, 184, H.he);
    put(h, 504, hash_bytes(h, 504, 1));
}

int decode(const unsigned char* h, Header& H, const char* path) {
    if (memcmp(h, kMagic, 8) != 0) return fail(TCYM_E_IO, std::string(path) + ": not a tcym checkpoint");
    uint32_t ver, hb;
    get(h, 8, ver);
    get(h, 12, hb);
    if (ver != kVersion || hb != kHeader) return fail(TCYM_E_IO, std::string(path) + ": unsupported checkpoint version");
    uint64_t hh;
    get(h, 504, hh);
    if (hh != hash_bytes(h, 504, 1)) return fail(TCYM_E_IO, std::string(path) + ": header hash mismatch");
    int32_t i32;
    get(h, 16, i32); H.info.nc = i32;
    get(h, 20, i32); H.info.precision = i32;
    for (int k = 0; k < 3; k++) { get(h, 24 + 4*k, i32); H.info.N[k] = i32; }
    get(h, 36, i32); H.has_e = i32;
    H.info.has_efield = i32 != 0;
    for (int k = 0; k < 3; k++) get(h, 40 + 8*k, H.info.a[k]);
    get(h, 64, H.info.aScale);
    get(h, 72, H.info.step);
    get(h, 80, H.info.seed);
    get(h, 88, H.info.time);
    for (int k = 0; k < 8; k++) get(h, 96 + 8*k, H.info.acc[k]);
    get(h, 160, H.nu);
    get(h, 168, H.ne);
    get(h, 176, H.hu);
    get(h, 184, H.he);
    Shape su, se;
    if (make_shape(H.info.nc, H.info.N, true, su) || make_shape(H.info.nc, H.info.N, false, se) ||
        (H.info.precision != TCYM_FP64 && H.info.precision != TCYM_FP32) ||
        H.nu != (uint64_t)(3*su.V*su.nr) || H.ne != (H.has_e ? (uint64_t)(3*se.V*se.nr) : 0))
        return fail(TCYM_E_IO, std::string(path) + ": inconsistent checkpoint header");
    return TCYM_OK;
}

// payload bytes of n reals in the storage precision
std::vector<unsigned char> to_storage(const double* A, size_t n, int prec) {
    std::vector<unsigned char> out(n*(prec == TCYM_FP32 ? 4 : 8));
    if (prec == TCYM_FP32) {
        float* f = (float*)out.data();
        for (size_t i = 0; i < n; i++) f[i] = (float)A[i];
    } else {
        memcpy(out.data(), A, n*8);
    }
    return out;
}

void from_storage(const unsigned char* p, size_t n, int prec, double* A) {
    if (prec == TCYM_FP32) {
        for (size_t i = 0; i < n; i++) {
            float f;
            memcpy(&f, p + 4*i, 4);
            A[i] = f;
        }
    } else {
        memcpy(A, p, n*8);
    }
}

int write_all(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n ? 0 : -1; }

// header and (when U / E are not NULL) the arrays of a checkpoint, hashes checked
int read_checkpoint(const char* path, Header& H, double* U, double* E) {
    Mapped m;
    int r = m.open_file(path);
    if (r) return r;
    if (m.n < kHeader) return fail(TCYM_E_IO, std::string(path) + ": too short for a checkpoint");
    if ((r = decode((const unsigned char*)m.p, H, path))) return r;
    const size_t w = H.info.precision == TCYM_FP32 ? 4 : 8;
    if (m.n != kHeader + (H.nu + H.ne)*w) return fail(TCYM_E_IO, std::string(path) + ": file size does not match its header");
    const unsigned char* pu = (const unsigned char*)m.p + kHeader;
    const unsigned char* pe = pu + H.nu*w;
    if (U) {
        if (hash_bytes(pu, H.nu*w, 0) != H.hu) return fail(TCYM_E_IO, std::string(path) + ": link data hash mismatch");
        from_storage(pu, H.nu, H.info.precision, U);
    }
    if (E && H.has_e) {
        if (hash_bytes(pe, H.ne*w, 0) != H.he) return fail(TCYM_E_IO, std::string(path) + ": field data hash mismatch");
        from_storage(pe, H.ne, H.info.precision, E);
    }
    return TCYM_OK;
}

int ctx_fail(tcym_ctx* c, int r) { return fail(r, std::string("tcym: ") + tcym_last_error(c)); }

}  // namespace

extern "C" {

const char* tcym_io_last_error(void) { return g_err.c_str(); }

int tcym_write_text_links(int nc, const int N[3], const double* U, const char* path, int digits, int nthreads) {
    Shape s;
    int r = make_shape(nc, N, true, s);
    return r ? r : write_text(s, U, path, digits, nthreads);
}
int tcym_write_text_efield(int nc, const int N[3], const double* E, const char* path, int digits, int nthreads) {
    Shape s;
    int r = make_shape(nc, N, false, s);
    return r ? r : write_text(s, E, path, digits, nthreads);
}
int tcym_read_text_links(int nc, const int N[3], double* U, const char* path, int nthreads) {
    Shape s;
    int r = make_shape(nc, N, true, s);
    return r ? r : read_text(s, U, path, nthreads);
}
int tcym_read_text_efield(int nc, const int N[3], double* E, const char* path, int nthreads) {
    Shape s;
    int r = make_shape(nc, N, false, s);
    return r ? r : read_text(s, E, path, nthreads);
}

int tcym_save_text(tcym_ctx* c, int u, int e, const char* ufile, const char* efile, int digits, int nthreads) {
    if (!c) return fail(TCYM_E_ARG, "null context");
    tcym_params p;
    int r, lev, n[3];
    if ((r = tcym_get_params(c, &p))) return ctx_fail(c, r);
    if (ufile) {
        if ((r = tcym_level(c, u, -1, &lev, n))) return ctx_fail(c, r);
        Shape s;
        if ((r = make_shape(p.nc, n, true, s))) return r;
        std::vector<double> A((size_t)(3*s.V*s.nr));
        if ((r = tcym_get_links(c, u, A.data()))) return ctx_fail(c, r);
        if ((r = write_text(s, A.data(), ufile, digits, nthreads))) return r;
    }
    if (efile) {
        if ((r = tcym_level(c, -1, e, &lev, n))) return ctx_fail(c, r);
        Shape s;
        if ((r = make_shape(p.nc, n, false, s))) return r;
        std::vector<double> A((size_t)(3*s.V*s.nr));
        if ((r = tcym_get_efield(c, e, A.data()))) return ctx_fail(c, r);
        if ((r = write_text(s, A.data(), efile, digits, nthreads))) return r;
    }
    return TCYM_OK;
}

int tcym_load_text(tcym_ctx* c, int u, int e, const char* ufile, const char* efile, int nthreads) {
    if (!c) return fail(TCYM_E_ARG, "null context");
    tcym_params p;
    int r;
    if ((r = tcym_get_params(c, &p))) return ctx_fail(c, r);
    if (ufile) {
        Shape s;
        if ((r = make_shape(p.nc, p.N, true, s))) return r;
        std::vector<double> A((size_t)(3*s.V*s.nr));
        if ((r = read_text(s, A.data(), ufile, nthreads))) return r;
        if ((r = tcym_set_links(c, u, A.data()))) return ctx_fail(c, r);
    }
    if (efile) {
        Shape s;
        if ((r = make_shape(p.nc, p.N, false, s))) return r;
        std::vector<double> A((size_t)(3*s.V*s.nr));
        if ((r = read_text(s, A.data(), efile, nthreads))) return r;
        if ((r = tcym_set_efield(c, e, A.data()))) return ctx_fail(c, r);
    }
    return TCYM_OK;
}

int tcym_write_checkpoint(const char* path, const tcym_checkpoint_info* info, const double* U, const double* E) {
    if (!path || !info || !U) return fail(TCYM_E_ARG, "null path, info or links");
    Header H{};
    H.info = *info;
    Shape su, se;
    int r;
    if ((r = make_shape(info->nc, info->N, true, su)) || (r = make_shape(info->nc, info->N, false, se))) return r;
    if (info->precision != TCYM_FP64 && info->precision != TCYM_FP32) return fail(TCYM_E_ARG, "bad storage precision");
    H.has_e = E != nullptr;
    H.nu = (uint64_t)(3*su.V*su.nr);
    H.ne = E ? (uint64_t)(3*se.V*se.nr) : 0;
    const std::vector<unsigned char> pu = to_storage(U, H.nu, info->precision);
    const std::vector<unsigned char> pe = E ? to_storage(E, H.ne, info->precision) : std::vector<unsigned char>();
    H.hu = hash_bytes(pu.data(), pu.size(), 0);
    H.he = hash_bytes(pe.data(), pe.size(), 0);
    unsigned char h[kHeader];
    encode(H, h);
    FILE* f = fopen(path, "wb");
    if (!f) return fail(TCYM_E_IO, std::string("cannot open ") + path + " for writing: " + strerror(errno));
    bool ok = write_all(f, h, kHeader) == 0 && write_all(f, pu.data(), pu.size()) == 0 &&
              (pe.empty() || write_all(f, pe.data(), pe.size()) == 0);
    if (fclose(f) != 0) ok = false;
    if (!ok) return fail(TCYM_E_IO, std::string("write failed on ") + path + ": " + strerror(errno));
    return TCYM_OK;
}

int tcym_read_checkpoint(const char* path, tcym_checkpoint_info* info, double* U, double* E) {
    if (!path || !info) return fail(TCYM_E_ARG, "null path or info");
    Header H{};
    int r = read_checkpoint(path, H, U, E);
    if (!r) *info = H.info;
    return r;
}

int tcym_save_checkpoint(tcym_ctx* c, int u, int e, const char* path, const tcym_checkpoint_info* info) {
    if (!c || !path) return fail(TCYM_E_ARG, "null context or path");
    tcym_params p;
    int r, lev, n[3];
    if ((r = tcym_get_params(c, &p))) return ctx_fail(c, r);
    if ((r = tcym_level(c, u, -1, &lev, n))) return ctx_fail(c, r);
    tcym_checkpoint_info ci{};
    if (info) ci = *info;
    ci.nc = p.nc;
    ci.precision = p.precision;
    for (int k = 0; k < 3; k++) { ci.N[k] = n[k]; ci.a[k] = p.a[k]*(double)(1 << lev); }
    ci.aScale = p.aScale;
    Shape su, se;
    if ((r = make_shape(p.nc, n, true, su)) || (r = make_shape(p.nc, n, false, se))) return r;
    std::vector<double> U((size_t)(3*su.V*su.nr)), E;
    if ((r = tcym_get_links(c, u, U.data()))) return ctx_fail(c, r);
    if (e >= 0) {
        int elev, en[3];
        if ((r = tcym_level(c, -1, e, &elev, en))) return ctx_fail(c, r);
        if (elev != lev) return fail(TCYM_E_ARG, "link and field slots of different resolutions");
        E.resize((size_t)(3*se.V*se.nr));
        if ((r = tcym_get_efield(c, e, E.data()))) return ctx_fail(c, r);
    }
    return tcym_write_checkpoint(path, &ci, U.data(), e >= 0 ? E.data() : nullptr);
}

int tcym_load_checkpoint(tcym_ctx* c, int u, int e, const char* path, tcym_checkpoint_info* info) {
    if (!c || !path) return fail(TCYM_E_ARG, "null context or path");
    tcym_params p;
    int r;
    if ((r = tcym_get_params(c, &p))) return ctx_fail(c, r);
    Header H{};
    if ((r = read_checkpoint(path, H, nullptr, nullptr))) return r;
    const tcym_checkpoint_info& ci = H.info;
    if (ci.nc != p.nc) return fail(TCYM_E_ARG, std::string(path) + ": checkpoint of another gauge group");
    for (int k = 0; k < 3; k++)
        if (ci.N[k] != p.N[k] || ci.a[k] != p.a[k])
            return fail(TCYM_E_ARG, std::string(path) + ": checkpoint extents or spacings differ from the context's");
    if (ci.aScale != p.aScale) return fail(TCYM_E_ARG, std::string(path) + ": checkpoint aScale differs from the context's");
    Shape su, se;
    if ((r = make_shape(p.nc, p.N, true, su)) || (r = make_shape(p.nc, p.N, false, se))) return r;
    if (e >= 0 && !H.has_e) return fail(TCYM_E_IO, std::string(path) + ": checkpoint holds no electric field");
    std::vector<double> U((size_t)(3*su.V*su.nr)), E(e >= 0 ? (size_t)(3*se.V*se.nr) : 0);
    if ((r = read_checkpoint(path, H, U.data(), e >= 0 ? E.data() : nullptr))) return r;
    if ((r = tcym_set_links(c, u, U.data()))) return ctx_fail(c, r);
    if (e >= 0 && (r = tcym_set_efield(c, e, E.data()))) return ctx_fail(c, r);
    if (info) *info = ci;
    return TCYM_OK;
}

}  // extern "C"
