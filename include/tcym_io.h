/* tcym_io.h -- configuration files around the lattice update loop.
 *
 * 1. The reference's text configurations, IO::SaveConfiguration and
 *    IO::LoadConfiguration (IO/SaveConfiguration.cpp:3-52,
 *    IO/LoadConfiguration.cpp:3-136), named UOutT<t>ID<seed>.txt /
 *    EOutT<t>ID<seed>.txt (IO/InputManagement.cpp:19-28).  One line per
 *    (x, y, z, mu), loops z outer, then y, x, mu:
 *      links   "x y z mu v_0 v_1 ... v_{s_U-1} \n"   (every value followed by
 *              a space; SU(2) u0..u3, SU(3) re,im of the 9 column-major
 *              entries: the order of the host arrays of tcym.h)
 *      fields  "x y z mu e_0 ... e_{n_E-1}\n"
 *    The reference writes every value with precision 15 in the stream's
 *    default floating format (OUTPUT_PRECISION, Definitions.cpp:21-22), i.e.
 *    printf "%.15g"; digits = 15 reproduces its files byte for byte, and
 *    digits = 17 writes files that read back bit-exact.  The reader places
 *    each line by its coordinates, so line order does not matter; it checks
 *    more than the reference does and rejects a file with a missing,
 *    repeated or out-of-range record, a wrong number of values on a line or
 *    an unparsable number (the reference only checks the E-field count).
 *    Formatting and parsing run on nthreads host threads (0: all hardware
 *    threads, at most 16).
 *
 * 2. A binary checkpoint of this library (no reference counterpart; the
 *    reference resumes only from the text files, without time, RNG state or
 *    accumulators): a 512-byte header (magic "TCYMCKPT", version 1, group,
 *    storage precision, extents, spacings, aScale, step, seed, time, 8 caller
 *    accumulators, value counts, 64-bit hashes of the two arrays and of the
 *    header), then the links and the E field in the host AoS layout of tcym.h,
 *    as doubles (TCYM_FP64) or floats (TCYM_FP32, exact for a TCYM_FP32
 *    context), little-endian.  A resumed run continues bit-identically.
 *
 * Every entry point returns a TCYM_* status; TCYM_E_IO reports a file that
 * could not be opened, read or written, or is malformed, and
 * tcym_io_last_error() gives the message of the calling thread's last failure.
 */
#ifndef TCYM_IO_H
#define TCYM_IO_H

#include <stdint.h>

#include "tcym.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *tcym_io_last_error(void);

/* ---- reference text files, host arrays (no device needed) ---- */
/* IO::SaveConfiguration, gauge-link file (IO/SaveConfiguration.cpp:19-25) */
int tcym_write_text_links(int nc, const int N[3], const double *U, const char *path, int digits, int nthreads);
/* IO::SaveConfiguration, electric-field file (IO/SaveConfiguration.cpp:38-44) */
int tcym_write_text_efield(int nc, const int N[3], const double *E, const char *path, int digits, int nthreads);
/* IO::LoadConfiguration, gauge links (IO/LoadConfiguration.cpp:46-72) */
int tcym_read_text_links(int nc, const int N[3], double *U, const char *path, int nthreads);
/* IO::LoadConfiguration, electric fields (IO/LoadConfiguration.cpp:84-122) */
int tcym_read_text_efield(int nc, const int N[3], double *E, const char *path, int nthreads);

/* ---- reference text files, device slots ---- */
/* IO::SaveConfiguration(Ufname, Efname, U, E) of link slot u and field slot e
   (at the slots' resolution; either file may be NULL) */
int tcym_save_text(tcym_ctx *ctx, int u, int e, const char *ufile, const char *efile, int digits, int nthreads);
/* IO::LoadConfiguration(U, E) into link slot u and field slot e (full resolution;
   either file may be NULL) */
int tcym_load_text(tcym_ctx *ctx, int u, int e, const char *ufile, const char *efile, int nthreads);

/* ---- binary checkpoints ---- */
typedef struct {
    int nc;               /* 2 or 3 */
    int precision;        /* storage: TCYM_FP64 (doubles) or TCYM_FP32 (floats) */
    int N[3];
    double a[3];
    double aScale;
    uint64_t step;        /* the caller's step counter, e.g. the next Philox step of Langevin noise */
    uint64_t seed;        /* the caller's noise key */
    double time;          /* the caller's time, e.g. Dynamics::Time() */
    double acc[8];        /* the caller's accumulators, e.g. DeltaNCsRealTime, DeltaNCsCoolRealTime */
    int has_efield;       /* read: the file holds an electric field (ignored on write) */
} tcym_checkpoint_info;

/* host arrays: U (3 V s_U) and E (3 V n_E, may be NULL for a links-only file) */
int tcym_write_checkpoint(const char *path, const tcym_checkpoint_info *info, const double *U, const double *E);
/* reads the header (info may not be NULL) and, when U / E are not NULL, the
   arrays, checking both hashes; a file without an E field leaves E untouched */
int tcym_read_checkpoint(const char *path, tcym_checkpoint_info *info, double *U, double *E);
/* device slots: geometry and storage precision from the context (and the
   slots' resolution), step/seed/time/acc from info (may be NULL: zeros);
   e < 0 writes links only */
int tcym_save_checkpoint(tcym_ctx *ctx, int u, int e, const char *path, const tcym_checkpoint_info *info);
/* loads a checkpoint of the context's group and full-resolution extents into
   link slot u and field slot e (e < 0: links only); info (may be NULL)
   receives the header */
int tcym_load_checkpoint(tcym_ctx *ctx, int u, int e, const char *path, tcym_checkpoint_info *info);

#ifdef __cplusplus
}
#endif

#endif
