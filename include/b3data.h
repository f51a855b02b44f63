/* Alignment ingest: FASTA -> MSA -> deduplicated patterns in the layout
 * b3h_create takes.
 *
 * The reference parses FASTA into an Msa<DnaNucleotide> on the host
 * (crates/data/src/fasta/mod.rs:103-150 FastaParser::parse_line,
 * crates/data/src/msa/mod.rs:77-119 Msa::from_fasta_reader,
 * crates/data/src/nucleotides.rs:195-221 the IUPAC character map) and
 * compresses identical columns with a per-column blake3 hash map
 * (crates/b3/src/likelihood/mod.rs:74-132 deduplicate).  Here the parser is
 * native C++ and the deduplication runs either on the host or on the GPU
 * (column hashes, an atomic hash table holding each column's first site,
 * exact byte comparison against it, and a prefix sum that numbers the
 * patterns in first-occurrence order).  Both give the reference's outputs:
 * patterns in first-occurrence order, their counts as weights, the pattern
 * of every site, and gaps written as 0b1111 in the pattern bytes.
 *
 * Every call returns a b3h_status; b3h_last_error() gives the message.
 */
#ifndef B3DATA_H
#define B3DATA_H

#include <stddef.h>
#include <stdint.h>

#include "b3hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct b3d_msa b3d_msa;

/* Msa::from_fasta_reader over a text buffer (len bytes, need not be
 * NUL-terminated).  Lines end at '\n' (a trailing '\r' is dropped); lines
 * whose first non-blank character is ';' are comments; '>' starts a record
 * whose name is the rest of the line; other lines append IUPAC characters
 * (ACGT WSMKRY BDHV N, either case, and '-') to the current record.
 * Errors: sequence data before any record, an invalid character, records of
 * different lengths.  As in the reference an input without records gives
 * one unnamed sequence of length 0. */
b3h_status b3d_msa_from_fasta(const char *text, size_t len, b3d_msa **out);
b3h_status b3d_msa_from_fasta_file(const char *path, b3d_msa **out);
b3h_status b3d_msa_destroy(b3d_msa *msa);

uint32_t b3d_msa_num_sequences(const b3d_msa *msa);
uint64_t b3d_msa_num_sites(const b3d_msa *msa);
/* NUL-terminated record name (the description after '>'); null if out of range */
const char *b3d_msa_name(const b3d_msa *msa, uint32_t index);
/* [num_sequences][num_sites] DnaNucleotide bytes (A=1 C=2 G=4 T=8, ambiguity
 * codes as their unions, N=15, gap=0) */
const uint8_t *b3d_msa_data(const b3d_msa *msa);

/* Msa::base_frequencies (msa/mod.rs:193-212): every character adds its
 * DnaNucleotide::base_frequencies (nucleotides.rs:142-165: ambiguity codes
 * split equally, gap and N as 1/4 each) in sequence order, then the sums are
 * divided by the number of characters. */
b3h_status b3d_msa_base_frequencies(const b3d_msa *msa, double *out4);

/* likelihood/mod.rs:74-132.  msa: [num_sequences][num_sites] DnaNucleotide
 * bytes.  device < 0: on the host; otherwise on that HIP device.
 *   leaves: [num_sequences][num_patterns] written with row stride
 *           leaves_stride (>= num_patterns; num_sites always suffices)
 *   weights[num_patterns], sites_to_patterns[num_sites]
 * *num_patterns receives the pattern count.  On a device the call holds
 * about 2 x num_sequences x num_sites bytes of device memory (the alignment
 * and a site-major copy) plus 64-96 bytes per site of hash table and
 * indices. */
b3h_status b3d_deduplicate(const uint8_t *msa, uint32_t num_sequences, uint64_t num_sites,
                           int device, uint8_t *leaves, uint64_t leaves_stride,
                           uint32_t *weights, uint64_t *sites_to_patterns,
                           uint64_t *num_patterns);

/*
 * b3d_deduplicate on a device without allocating: no reference counterpart.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and outputs as b3d_deduplicate with device >= 0 (bit-identical
 * results; device < 0 is an argument error); it calls no hipMalloc, hipFree
 * or stream creation and works on the legacy default stream.
 *   Every buffer comes from one scratch array of B3D_NOALLOC_SCRATCH_BYTES
 *   (64 MiB) that each device the library runs on holds until the process
 *   ends.  An input of S sequences x N sites needs about (2 S + 24) bytes
 *   per site plus 16 bytes per hash slot (a power of two >= 2 N), between
 *   (2 S + 56) N and (2 S + 88) N bytes (200 sequences: about 140,000 sites);
 *   larger inputs return B3H_ERR_ARG before any device work.
 *   Calls that overlap on one device share that device's array and race:
 *   the caller serialises them (calls on distinct devices are independent).
 */
#define B3D_NOALLOC_SCRATCH_BYTES (64ull * 1024ull * 1024ull)
b3h_status b3d_deduplicate_noalloc(const uint8_t *msa, uint32_t num_sequences,
                                   uint64_t num_sites, int device, uint8_t *leaves,
                                   uint64_t leaves_stride, uint32_t *weights,
                                   uint64_t *sites_to_patterns, uint64_t *num_patterns);

#ifdef __cplusplus
}
#endif

#endif
