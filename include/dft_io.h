/*
 * dft_io.h -- on-disk formats of the solver layer (libdft_core.so; host code
 * only, no GPU call).
 *
 * The reference has no file format on this path: its SCF writes stdout lines
 * (SelfConsistentFieldMolecule.cpp:343-499) and returns a MolecularResult
 * (DFTData.h:62-91) in memory.  Three formats are defined here:
 *
 *  - field files: one grid field or a block of orbital columns, stored in the
 *    device layout (column-major N x ncols, x fastest within a column,
 *    index (k ny + j) nx + i as CartesianGrid::getIndex, CartesianGrid.cpp:45-47),
 *    so a dump is one device-to-host copy and a load one copy back;
 *  - patch shards: learned-preconditioner training pairs as the trainer holds
 *    them on the device (inputs [count][2][P^3], targets [count][P^3], FP32);
 *  - the result document: a JSON serialisation of MolecularResult without
 *    its N-vectors (those go to field files), with the run's configuration,
 *    grid, nuclei, timings and counters.
 *
 * Binary layout (little-endian; the header is fixed-size and the data
 * follows it):
 *
 *   field file, 256-byte header
 *     0   char[8]  "DFTFLD01"
 *     8   uint32   header bytes (256)
 *     12  uint32   dtype: DFT_F64 (0) or DFT_F32 (1)
 *     16  int64    nx, ny, nz, ncols
 *     48  double   xmin, xmax, ymin, ymax, zmin, zmax, dx, dy, dz
 *     120 uint64   data bytes (= nx ny nz ncols * sizeof dtype)
 *     128 uint32   CRC-32 of the data (zlib's polynomial, reflected 0xEDB88320)
 *     132 char[64] name, NUL-padded (e.g. "rho_alpha", "psi_alpha")
 *     196 zeros
 *     252 uint32   CRC-32 of header bytes 0..251
 *
 *   patch shard, 128-byte header
 *     0   char[8]  "DFTPAT01"
 *     8   uint32   header bytes (128)
 *     12  uint32   input channels (2: residual, shifted potential)
 *     16  int64    patch edge P
 *     24  int64    count
 *     32  uint64   input bytes (count * 2 * P^3 * 4)
 *     40  uint64   target bytes (count * P^3 * 4)
 *     48  uint32   CRC-32 of the inputs
 *     52  uint32   CRC-32 of the targets
 *     56  int32    target scale (the harvest's target_scale: 0 c/max|c|, 1 c/(max|r| h2))
 *     60  uint32   0
 *     64  double   grid spacing h of the harvest (0 if unknown)
 *     72  zeros
 *     124 uint32   CRC-32 of header bytes 0..123
 *
 * Status codes as elsewhere: 0 ok; -1 (DFT_EINVAL) for a bad argument, a file
 * that cannot be opened or written, a wrong magic or version, sizes that do
 * not agree with the header, or a destination too small; -2 (DFT_ENUMERIC)
 * for a short read or a checksum mismatch (of the header or the data).
 * dft_core_last_error() holds the message.  Functions are reentrant apart
 * from that message.
 */
#ifndef DFT_IO_H
#define DFT_IO_H

#include <stddef.h>
#include <stdint.h>

#include "dft_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dtype;            /* DFT_F64 or DFT_F32 */
    int64_t ncols;        /* orbital columns (1 for a field) */
    dft_grid grid;
    uint64_t data_bytes;
    uint32_t crc;
    char name[64];
} dft_field_header;

/* CRC-32 (ISO-HDLC / zlib) of n bytes, continuing from crc (0 to start). */
uint32_t dft_crc32(uint32_t crc, const void* data, size_t n);

/* Write ncols columns of N = nx ny nz values (host memory, device layout). */
int dft_field_write(const char* path, const dft_grid* g, int dtype, int64_t ncols, const char* name, const void* data);
/* Read the header only (its own checksum is verified, the data's is not). */
int dft_field_read_header(const char* path, dft_field_header* hdr);
/* Read header and data into data[0 .. capacity_bytes); verifies sizes and the checksum. */
int dft_field_read(const char* path, dft_field_header* hdr, void* data, uint64_t capacity_bytes);

typedef struct {
    int channels;         /* 2 */
    int target_scale;
    int64_t patch;
    int64_t count;
    double h;
    uint32_t crc_in, crc_tgt;
} dft_patch_shard_header;

/* Write count patches: in (count x 2 x P^3), tgt (count x P^3), FP32. */
int dft_patch_shard_write(const char* path, int64_t patch, int64_t count, int target_scale, double h, const float* in,
                          const float* tgt);
int dft_patch_shard_read_header(const char* path, dft_patch_shard_header* hdr);
/* Read patches [first, first + count) of a shard.  The header's checksum is
 * always verified, the data's when the whole shard is read (first = 0 and
 * count = the shard's count). */
int dft_patch_shard_read(const char* path, int64_t first, int64_t count, float* in, float* tgt);

/* The result document (JSON, UTF-8, no trailing newline) of a run of
 * dft_scf_run: grid, nuclei, the configuration, the report, and the occupied
 * eigenvalues eig (n_alpha values, then n_beta) as MolecularResult's orbital
 * list.  Doubles are printed with 17 significant digits (exact round trip);
 * a non-finite value is written as null.  buf may be NULL to query the
 * length: *len receives the document's length in bytes (without the NUL);
 * the document and a NUL are written when cap > *len, otherwise DFT_EINVAL. */
int dft_scf_result_json(const dft_grid* g, int natoms, const double* nuclei, const dft_scf_config* cfg,
                        const dft_scf_report* rep, const double* eig, int n_alpha, int n_beta, char* buf, size_t cap,
                        size_t* len);

#ifdef __cplusplus
}
#endif
#endif
