// Alignment ingest (include/b3data.h): FASTA parsing into an MSA of
// DnaNucleotide bytes and pattern deduplication on the host or the GPU.
//
// Parsing follows FastaParser::parse_line (crates/data/src/fasta/mod.rs:
// 103-150), Msa::from_fasta_reader (crates/data/src/msa/mod.rs:77-119) and
// DnaNucleotide::from_ascii (crates/data/src/nucleotides.rs:195-221).
// Deduplication produces what deduplicate (crates/b3/src/likelihood/mod.rs:
// 74-132) produces; columns are compared byte for byte, so equal outputs do
// not rest on a hash being collision-free.
#include "../../include/b3data.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace b3h {
b3h_status host_fail(b3h_status s, const std::string &msg);
}

struct b3d_msa {
  uint32_t num_sequences = 0;
  uint64_t num_sites = 0;
  std::vector<std::string> names;
  std::vector<uint8_t> data;
};

namespace {

using b3h::host_fail;

// nucleotides.rs:195-221; 0xFF: not a character of the alphabet
struct AsciiTable {
  uint8_t v[256];
  AsciiTable() {
    std::memset(v, 0xFF, sizeof v);
    const char *up = "ACGTWSMKRYBDHVN";
    const uint8_t code[] = {1, 2, 4, 8, 9, 6, 3, 12, 5, 10, 14, 13, 11, 7, 15};
    for (int i = 0; i < 15; i++) {
      v[(uint8_t)up[i]] = code[i];
      v[(uint8_t)(up[i] - 'A' + 'a')] = code[i];
    }
    v[(uint8_t)'-'] = 0;
  }
};
const AsciiTable kAscii;

b3h_status parse_fasta(const char *text, size_t len, b3d_msa *m) {
  bool in_sequence = false;
  std::string description;
  std::vector<uint8_t> seq;
  uint64_t num_sites = 0;
  bool have_sites = false;
  // Msa::from_fasta_reader's add_record (msa/mod.rs:85-95)
  auto add_record = [&]() -> b3h_status {
    if (!have_sites || num_sites == 0) {
      if (m->num_sequences > 0 && seq.size() != num_sites)
        return host_fail(B3H_ERR_ARG, "record '" + description + "' has " +
                                          std::to_string(seq.size()) + " sites, expected " +
                                          std::to_string(num_sites));
      num_sites = seq.size();
      have_sites = true;
    } else if (seq.size() != num_sites) {
      return host_fail(B3H_ERR_ARG, "record '" + description + "' has " +
                                        std::to_string(seq.size()) + " sites, expected " +
                                        std::to_string(num_sites));
    }
    m->data.insert(m->data.end(), seq.begin(), seq.end());
    m->names.push_back(description);
    m->num_sequences++;
    return B3H_OK;
  };
  size_t pos = 0;
  uint64_t line_no = 0;
  while (pos < len) {
    size_t end = pos;
    while (end < len && text[end] != '\n') end++;
    size_t lend = end;
    // msa/mod.rs:101-106: "\n" and then one "\r" are dropped; a last line
    // without "\n" keeps its characters
    if (end < len && lend > pos && text[lend - 1] == '\r') lend--;
    line_no++;
    const char *line = text + pos;
    const size_t n = lend - pos;
    size_t k = 0;
    while (k < n && (line[k] == ' ' || line[k] == '\t' || line[k] == '\v' || line[k] == '\f' ||
                     line[k] == '\r'))
      k++;
    if (k < n && line[k] == ';') {
      // comment (fasta/mod.rs:111-112)
    } else if (n > 0 && line[0] == '>') {
      if (in_sequence) {
        in_sequence = false;
        b3h_status st = add_record();
        if (st) return st;
      }
      description.assign(line + 1, n - 1);
      seq.clear();
      in_sequence = true;
    } else {
      if (!in_sequence)
        return host_fail(B3H_ERR_ARG, "Encountered sequence data before any record (line " +
                                          std::to_string(line_no) + ")");
      for (size_t i = 0; i < n; i++) {
        const uint8_t c = kAscii.v[(uint8_t)line[i]];
        if (c == 0xFF) {
          char buf[8];
          std::snprintf(buf, sizeof buf, "%c", line[i]);
          return host_fail(B3H_ERR_ARG,
                           std::string("Encountered an invalid character '") + buf +
                               "' in a FASTA character sequence (record '" + description +
                               "', line " + std::to_string(line_no) + ", column " +
                               std::to_string(i + 1) + ")");
        }
        seq.push_back(c);
      }
    }
    pos = end + 1;
  }
  // the last record (also called when there was none: msa/mod.rs:110)
  b3h_status st = add_record();
  if (st) return st;
  m->num_sites = num_sites;
  return B3H_OK;
}

inline uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

// Host deduplication: columns gathered into contiguous rows, an
// open-addressing table of (hash, pattern), exact comparison on a match.
b3h_status dedup_host(const uint8_t *msa, uint32_t S, uint64_t N, uint8_t *leaves,
                      uint64_t stride, uint32_t *weights, uint64_t *s2p, uint64_t *np) {
  std::vector<uint8_t> cols((size_t)N * S);
  const uint64_t B = 512;
  for (uint64_t b = 0; b < N; b += B) {
    const uint64_t e = std::min(N, b + B);
    for (uint32_t s = 0; s < S; s++) {
      const uint8_t *row = msa + (size_t)s * N;
      for (uint64_t j = b; j < e; j++) cols[(size_t)j * S + s] = row[j];
    }
  }
  uint64_t cap = 16;
  while (cap < 2 * N) cap <<= 1;
  std::vector<uint64_t> th(cap, 0);
  std::vector<uint32_t> tp(cap, 0xFFFFFFFFu);
  std::vector<uint64_t> rep;
  uint64_t P = 0;
  for (uint64_t j = 0; j < N; j++) {
    const uint8_t *c = cols.data() + (size_t)j * S;
    uint64_t h = 0x9E3779B97F4A7C15ull ^ S;
    for (uint32_t s = 0; s < S; s++) h = (h ^ c[s]) * 0x100000001B3ull;
    h = mix64(h);
    uint64_t slot = h & (cap - 1);
    for (;;) {
      const uint32_t p = tp[slot];
      if (p == 0xFFFFFFFFu) {
        tp[slot] = (uint32_t)P;
        th[slot] = h;
        rep.push_back(j);
        weights[P] = 1;
        s2p[j] = P;
        P++;
        break;
      }
      if (th[slot] == h && std::memcmp(cols.data() + (size_t)rep[p] * S, c, S) == 0) {
        weights[p]++;
        s2p[j] = p;
        break;
      }
      slot = (slot + 1) & (cap - 1);
    }
  }
  for (uint32_t s = 0; s < S; s++)
    for (uint64_t p = 0; p < P; p++) {
      const uint8_t v = msa[(size_t)s * N + rep[p]];
      leaves[(size_t)s * stride + p] = v == 0 ? 0x0F : v;  // mod.rs:117-122
    }
  *np = P;
  return B3H_OK;
}

// ---- device deduplication ------------------------------------------------

// With col non-null the pass also writes the site-major copy of the
// alignment that k_verify reads: column j as Sp = S rounded up to 16 bytes at
// col + j * Sp, zero-padded, in 16-byte stores.
__global__ void k_col_hash(const uint8_t *__restrict__ msa, uint32_t S, uint64_t N, uint64_t seed,
                           uint64_t *__restrict__ h, uint8_t *__restrict__ col, uint32_t Sp) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  uint64_t x = (0x9E3779B97F4A7C15ull + seed) ^ S;
  for (uint32_t c = 0; c < Sp; c += 16) {
    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (uint32_t q = 0; q < 16; q++) {
      if (c + q < S) {
        const uint32_t b = msa[(size_t)(c + q) * N + j];
        x = (x ^ b) * 0x100000001B3ull;
        w[q >> 2] |= b << (8 * (q & 3));
      }
    }
    if (col) *reinterpret_cast<uint4 *>(col + j * Sp + c) = make_uint4(w[0], w[1], w[2], w[3]);
  }
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  h[j] = x ? x : 1;  // 0 marks an empty table slot
}

// Open addressing on the column hash; the slot keeps the smallest site
// with that hash.  The table holds at least 2N slots, so every probe ends.
__global__ void k_insert(const uint64_t *__restrict__ h, uint64_t N,
                         unsigned long long *__restrict__ keys, uint32_t *__restrict__ first,
                         uint64_t mask, uint32_t *__restrict__ slot_of) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const unsigned long long key = h[j];
  uint64_t slot = (key * 0x9E3779B97F4A7C15ull >> 20) & mask;
  for (;;) {
    const unsigned long long prev = atomicCAS(&keys[slot], 0ull, key);
    if (prev == 0ull || prev == key) {
      atomicMin(&first[slot], (uint32_t)j);
      slot_of[j] = (uint32_t)slot;
      return;
    }
    slot = (slot + 1) & mask;
  }
}

// Each site against the first site of its slot: equal bytes, or a hash
// collision between different columns (then the caller hashes again with
// another seed).  Counts per slot give the weights.  The columns are compared
// in the site-major copy (Sp bytes each, contiguous), so the representative's
// column is read in 16-byte loads rather than one byte per row.
__global__ void k_verify(const uint8_t *__restrict__ col, uint32_t Sp, uint64_t N,
                         const uint32_t *__restrict__ slot_of, const uint32_t *__restrict__ first,
                         uint32_t *__restrict__ count, uint32_t *__restrict__ is_first,
                         uint32_t *__restrict__ rep, uint32_t *__restrict__ collided) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint32_t slot = slot_of[j];
  const uint32_t r = first[slot];
  rep[j] = r;
  is_first[j] = (r == j) ? 1u : 0u;
  if (r != j) {
    const uint4 *a = reinterpret_cast<const uint4 *>(col + j * Sp);
    const uint4 *b = reinterpret_cast<const uint4 *>(col + (uint64_t)r * Sp);
    for (uint32_t c = 0; c < Sp / 16u; c++) {
      const uint4 u = a[c], v = b[c];
      if (u.x != v.x || u.y != v.y || u.z != v.z || u.w != v.w) {
        atomicOr(collided, 1u);
        break;
      }
    }
  }
  atomicAdd(&count[slot], 1u);
}

// Exclusive prefix sum, 4096 elements per 256-thread block, recursive over
// the block totals.
constexpr uint32_t kScanItems = 16, kScanThreads = 256, kScanBlock = kScanItems * kScanThreads;

__global__ __launch_bounds__(256) void k_scan_block(const uint32_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, uint64_t n,
                                                    uint32_t *__restrict__ block_sums) {
  __shared__ uint32_t part[kScanThreads];
  const uint64_t base = (uint64_t)blockIdx.x * kScanBlock + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < kScanItems; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    sum += v[i];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
  for (uint32_t i = 0; i < kScanItems; i++) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == kScanThreads - 1) block_sums[blockIdx.x] = part[kScanThreads - 1];
}

__global__ void k_scan_add(uint32_t *__restrict__ out, uint64_t n,
                           const uint32_t *__restrict__ offsets) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  out[j] += offsets[j / kScanBlock];
}

// Workspace bytes of exclusive_scan_ws over n values: two block-sum arrays
// per level of the recursion.
uint64_t scan_ws_bytes(uint64_t n) {
  const uint64_t blocks = std::max<uint64_t>((n + kScanBlock - 1) / kScanBlock, 1);
  const uint64_t here = (blocks * 4 * 2 + 255) & ~(uint64_t)255;
  return blocks > 1 ? here + scan_ws_bytes(blocks) : here;
}

// Exclusive prefix sum of n u32 values on stream s; ws holds scan_ws_bytes(n).
hipError_t exclusive_scan_ws(const uint32_t *in, uint32_t *out, uint64_t n, unsigned char *ws,
                             hipStream_t s) {
  const uint64_t blocks = (n + kScanBlock - 1) / kScanBlock;
  const uint64_t nb = std::max<uint64_t>(blocks, 1);
  uint32_t *sums = reinterpret_cast<uint32_t *>(ws);
  uint32_t *sums_scan = sums + nb;
  hipLaunchKernelGGL(k_scan_block, dim3((uint32_t)blocks), dim3(kScanThreads), 0, s, in, out, n,
                     sums);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && blocks > 1) {
    e = exclusive_scan_ws(sums, sums_scan, blocks, ws + ((nb * 4 * 2 + 255) & ~(uint64_t)255), s);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_scan_add, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, out, n,
                         sums_scan);
      e = hipGetLastError();
    }
  }
  return e;
}

__global__ void k_emit(const uint8_t *__restrict__ msa, uint32_t S, uint64_t N,
                       const uint32_t *__restrict__ rep, const uint32_t *__restrict__ pid,
                       const uint32_t *__restrict__ slot_of, const uint32_t *__restrict__ count,
                       uint8_t *__restrict__ leaves, uint64_t P, uint32_t *__restrict__ weights,
                       uint64_t *__restrict__ s2p) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint32_t r = rep[j];
  const uint32_t p = pid[r];
  s2p[j] = p;
  if (r == j) {
    weights[p] = count[slot_of[j]];
    for (uint32_t s = 0; s < S; s++) {
      const uint8_t v = msa[(size_t)s * N + j];
      leaves[(size_t)s * P + p] = v == 0 ? 0x0F : v;
    }
  }
}

__global__ void k_fill32(uint32_t *__restrict__ p, uint64_t n, uint32_t v) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) p[j] = v;
}

// Device buffers of one deduplication of S sequences x N sites.  leaves, w
// and s2p are sized by the pattern count P, known after the scan.
struct DedupBufs {
  uint64_t cap = 0;  // hash slots: a power of two >= 2N
  uint8_t *msa = nullptr, *col = nullptr, *leaves = nullptr;
  uint64_t *h = nullptr, *s2p = nullptr;
  unsigned long long *keys = nullptr;
  uint32_t *first = nullptr, *slot = nullptr, *count = nullptr, *isf = nullptr, *pid = nullptr,
           *rep = nullptr, *w = nullptr, *coll = nullptr;
  unsigned char *scan_ws = nullptr;
};

uint64_t hash_slots(uint64_t N) {
  uint64_t cap = 16;
  while (cap < 2 * N) cap <<= 1;
  return cap;
}

#define DCHK(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      (void)hipGetLastError();                                                             \
      return host_fail(B3H_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    }                                                                                      \
  } while (0)

// The kernels of a device deduplication on `stream` (shared by both entry
// points).  `outputs(P, &b)` makes b.leaves (S x P bytes), b.w (P u32) and
// b.s2p (N u64) available once the pattern count P is known.  Every buffer
// that must start cleared is cleared here by a fill kernel.
template <typename Outputs>
b3h_status dedup_kernels(DedupBufs &b, const uint8_t *msa, uint32_t S, uint64_t N,
                         uint8_t *leaves, uint64_t stride, uint32_t *weights, uint64_t *s2p,
                         uint64_t *np, hipStream_t stream, Outputs outputs) {
  const uint32_t Sp = (S + 15u) & ~15u;
  const uint64_t cap = b.cap;
  DCHK(hipMemcpyAsync(b.msa, msa, (size_t)S * N, hipMemcpyHostToDevice, stream));
  const dim3 grid((uint32_t)((N + 255) / 256)), block(256);
  auto fill = [&](uint32_t *p, uint64_t n, uint32_t v) {
    hipLaunchKernelGGL(k_fill32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, p, n,
                       v);
    return hipGetLastError();
  };
  bool ok = false;
  for (uint64_t seed = 0; seed < 4 && !ok; seed++) {
    DCHK(fill(reinterpret_cast<uint32_t *>(b.keys), cap * 2, 0u));
    DCHK(fill(b.first, cap, 0xFFFFFFFFu));
    DCHK(fill(b.count, cap, 0u));
    DCHK(fill(b.coll, 1, 0u));
    // the site-major copy does not depend on the seed: written on the first pass
    hipLaunchKernelGGL(k_col_hash, grid, block, 0, stream, b.msa, S, N, seed, b.h,
                       seed == 0 ? b.col : nullptr, Sp);
    DCHK(hipGetLastError());
    hipLaunchKernelGGL(k_insert, grid, block, 0, stream, b.h, N, b.keys, b.first, cap - 1, b.slot);
    DCHK(hipGetLastError());
    hipLaunchKernelGGL(k_verify, grid, block, 0, stream, b.col, Sp, N, b.slot, b.first, b.count,
                       b.isf, b.rep, b.coll);
    DCHK(hipGetLastError());
    uint32_t coll = 0;
    DCHK(hipMemcpyAsync(&coll, b.coll, 4, hipMemcpyDeviceToHost, stream));
    DCHK(hipStreamSynchronize(stream));
    ok = coll == 0;
  }
  if (!ok) return host_fail(B3H_ERR_HIP, "column hash collisions under four seeds");
  DCHK(exclusive_scan_ws(b.isf, b.pid, N, b.scan_ws, stream));
  uint32_t last[2] = {0, 0};
  DCHK(hipMemcpyAsync(&last[0], b.pid + (N - 1), 4, hipMemcpyDeviceToHost, stream));
  DCHK(hipMemcpyAsync(&last[1], b.isf + (N - 1), 4, hipMemcpyDeviceToHost, stream));
  DCHK(hipStreamSynchronize(stream));
  const uint64_t P = (uint64_t)last[0] + last[1];
  b3h_status st = outputs(P, b);
  if (st) return st;
  hipLaunchKernelGGL(k_emit, grid, block, 0, stream, b.msa, S, N, b.rep, b.pid, b.slot, b.count,
                     b.leaves, P, b.w, b.s2p);
  DCHK(hipGetLastError());
  DCHK(hipMemcpy2DAsync(leaves, stride, b.leaves, P, P, S, hipMemcpyDeviceToHost, stream));
  DCHK(hipMemcpyAsync(weights, b.w, P * 4, hipMemcpyDeviceToHost, stream));
  DCHK(hipMemcpyAsync(s2p, b.s2p, N * 8, hipMemcpyDeviceToHost, stream));
  DCHK(hipStreamSynchronize(stream));
  *np = P;
  return B3H_OK;
}

b3h_status check_device(int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return host_fail(B3H_ERR_NO_DEVICE, "no HIP device visible");
  }
  if (device >= ndev) return host_fail(B3H_ERR_NO_DEVICE, "device ordinal out of range");
  return B3H_OK;
}

struct DevBufs {
  std::vector<void *> ptrs;
  ~DevBufs() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T **p, size_t bytes) {
    hipError_t e = hipMalloc((void **)p, bytes ? bytes : 16);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) (void)hipStreamDestroy(s);
  }
};

b3h_status dedup_device(int device, const uint8_t *msa, uint32_t S, uint64_t N, uint8_t *leaves,
                        uint64_t stride, uint32_t *weights, uint64_t *s2p, uint64_t *np) {
  b3h_status st = check_device(device);
  if (st) return st;
  if (N >= 0xFFFFFFFFull) return host_fail(B3H_ERR_UNSUPPORTED, "more than 2^32 - 2 sites");
  DCHK(hipSetDevice(device));
  StreamGuard sg;
  DCHK(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  DevBufs buf;
  DedupBufs b;
  b.cap = hash_slots(N);
  const uint32_t Sp = (S + 15u) & ~15u;
  DCHK(buf.alloc(&b.msa, (size_t)S * N));
  DCHK(buf.alloc(&b.col, (size_t)Sp * N));
  DCHK(buf.alloc(&b.h, N * 8));
  DCHK(buf.alloc(&b.keys, b.cap * 8));
  DCHK(buf.alloc(&b.first, b.cap * 4));
  DCHK(buf.alloc(&b.count, b.cap * 4));
  DCHK(buf.alloc(&b.slot, N * 4));
  DCHK(buf.alloc(&b.isf, N * 4));
  DCHK(buf.alloc(&b.pid, N * 4));
  DCHK(buf.alloc(&b.rep, N * 4));
  DCHK(buf.alloc(&b.coll, 4));
  DCHK(buf.alloc(&b.scan_ws, scan_ws_bytes(N)));
  auto outputs = [&](uint64_t P, DedupBufs &bb) -> b3h_status {
    DCHK(buf.alloc(&bb.leaves, (size_t)S * P));
    DCHK(buf.alloc(&bb.w, P * 4));
    DCHK(buf.alloc(&bb.s2p, N * 8));
    return B3H_OK;
  };
  return dedup_kernels(b, msa, S, N, leaves, stride, weights, s2p, np, sg.s, outputs);
}

// ---- the no-allocation form -------------------------------------------------
// One scratch array per device, in the library's code object: every buffer
// of a deduplication is carved from it, so the call allocates nothing.  In
// host code `dedup_scratch` names a host-side stand-in; the device address
// of each device's copy comes from hipGetSymbolAddress, taken once per device
// with that device current, and only that address is used.
static __device__ __align__(256) unsigned char dedup_scratch[B3D_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

// Scratch bytes one call needs, with the carve-up: the pattern bytes reuse
// the site-major copy (dead after k_verify), the weights the site hashes
// (dead after k_insert) and the site map the hash keys (dead after k_verify).
uint64_t noalloc_layout(uint32_t S, uint64_t N, unsigned char *base, DedupBufs *b) {
  const uint32_t Sp = (S + 15u) & ~15u;
  const uint64_t cap = hash_slots(N);
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    unsigned char *p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  DedupBufs d;
  d.cap = cap;
  d.msa = take((uint64_t)S * N);
  d.col = take((uint64_t)Sp * N);
  d.h = reinterpret_cast<uint64_t *>(take(N * 8));
  d.keys = reinterpret_cast<unsigned long long *>(take(cap * 8));
  d.first = reinterpret_cast<uint32_t *>(take(cap * 4));
  d.count = reinterpret_cast<uint32_t *>(take(cap * 4));
  d.slot = reinterpret_cast<uint32_t *>(take(N * 4));
  d.isf = reinterpret_cast<uint32_t *>(take(N * 4));
  d.pid = reinterpret_cast<uint32_t *>(take(N * 4));
  d.rep = reinterpret_cast<uint32_t *>(take(N * 4));
  d.coll = reinterpret_cast<uint32_t *>(take(4));
  d.scan_ws = take(scan_ws_bytes(N));
  d.leaves = d.col;                               // S * P <= Sp * N
  d.w = reinterpret_cast<uint32_t *>(d.h);        // 4 P <= 8 N
  d.s2p = reinterpret_cast<uint64_t *>(d.keys);   // 8 N <= 8 cap
  if (b) *b = d;
  return off;
}

b3h_status scratch_of(int device, unsigned char **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device >= kMaxDevices) return host_fail(B3H_ERR_UNSUPPORTED, "device ordinal above 63");
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    DCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(dedup_scratch)));
    if (!p) return host_fail(B3H_ERR_HIP, "hipGetSymbolAddress returned a null address");
    g_scratch_ptr[device] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[device];
  return B3H_OK;
}

b3h_status dedup_device_noalloc(int device, const uint8_t *msa, uint32_t S, uint64_t N,
                                uint8_t *leaves, uint64_t stride, uint32_t *weights,
                                uint64_t *s2p, uint64_t *np) {
  const uint64_t need = noalloc_layout(S, N, nullptr, nullptr);
  if (need > B3D_NOALLOC_SCRATCH_BYTES)
    return host_fail(B3H_ERR_ARG, "input needs " + std::to_string(need) +
                                      " bytes of device scratch; the no-allocation form holds " +
                                      std::to_string((uint64_t)B3D_NOALLOC_SCRATCH_BYTES));
  b3h_status st = check_device(device);
  if (st) return st;
  DCHK(hipSetDevice(device));
  unsigned char *base = nullptr;
  if ((st = scratch_of(device, &base))) return st;
  DedupBufs b;
  noalloc_layout(S, N, base, &b);
  // the legacy default stream: creating a stream is not free of allocation
  auto outputs = [](uint64_t, DedupBufs &) -> b3h_status { return B3H_OK; };
  return dedup_kernels(b, msa, S, N, leaves, stride, weights, s2p, np, nullptr, outputs);
}

}  // namespace

extern "C" {

b3h_status b3d_msa_from_fasta(const char *text, size_t len, b3d_msa **out) {
  if (!out) return host_fail(B3H_ERR_ARG, "null output");
  *out = nullptr;
  if (!text && len) return host_fail(B3H_ERR_ARG, "null text");
  b3d_msa *m = new b3d_msa();
  b3h_status st = parse_fasta(text ? text : "", len, m);
  if (st) {
    delete m;
    return st;
  }
  *out = m;
  return B3H_OK;
}

b3h_status b3d_msa_from_fasta_file(const char *path, b3d_msa **out) {
  if (!out || !path) return host_fail(B3H_ERR_ARG, "null argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  if (!f) return host_fail(B3H_ERR_ARG, std::string("cannot open ") + path);
  std::string text;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) text.append(chunk, got);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  if (err) return host_fail(B3H_ERR_ARG, std::string("cannot read ") + path);
  return b3d_msa_from_fasta(text.data(), text.size(), out);
}

b3h_status b3d_msa_destroy(b3d_msa *m) {
  delete m;
  return B3H_OK;
}

uint32_t b3d_msa_num_sequences(const b3d_msa *m) { return m ? m->num_sequences : 0; }
uint64_t b3d_msa_num_sites(const b3d_msa *m) { return m ? m->num_sites : 0; }
const char *b3d_msa_name(const b3d_msa *m, uint32_t i) {
  return (m && i < m->names.size()) ? m->names[i].c_str() : nullptr;
}
const uint8_t *b3d_msa_data(const b3d_msa *m) { return m ? m->data.data() : nullptr; }

b3h_status b3d_msa_base_frequencies(const b3d_msa *m, double *out) {
  if (!m || !out) return host_fail(B3H_ERR_ARG, "null argument");
  const double t = 1.0 / 3.0;
  // nucleotides.rs:142-165, indexed by the DnaNucleotide byte
  static const double table[16][4] = {
      {0.25, 0.25, 0.25, 0.25},  // gap
      {1, 0, 0, 0},             // A
      {0, 1, 0, 0},             // C
      {0.5, 0.5, 0, 0},         // M
      {0, 0, 1, 0},             // G
      {0.5, 0, 0.5, 0},         // R
      {0, 0.5, 0.5, 0},         // S
      {t, t, t, 0},             // V
      {0, 0, 0, 1},             // T
      {0.5, 0, 0, 0.5},         // W
      {0, 0.5, 0, 0.5},         // Y
      {t, t, 0, t},             // H
      {0, 0, 0.5, 0.5},         // K
      {t, 0, t, t},             // D
      {0, t, t, t},             // B
      {0.25, 0.25, 0.25, 0.25}  // N
  };
  double c[4] = {0, 0, 0, 0};
  for (uint8_t v : m->data)
    for (int k = 0; k < 4; k++) c[k] += table[v & 15][k];
  const double n = (double)m->data.size();
  for (int k = 0; k < 4; k++) out[k] = c[k] / n;
  return B3H_OK;
}

b3h_status b3d_deduplicate(const uint8_t *msa, uint32_t S, uint64_t N, int device,
                           uint8_t *leaves, uint64_t stride, uint32_t *weights, uint64_t *s2p,
                           uint64_t *np) {
  if (!np) return host_fail(B3H_ERR_ARG, "null pattern count");
  *np = 0;
  if (S == 0) return host_fail(B3H_ERR_ARG, "no sequences");
  if (N == 0) return B3H_OK;
  if (!msa || !leaves || !weights || !s2p) return host_fail(B3H_ERR_ARG, "null array");
  if (stride < N) return host_fail(B3H_ERR_ARG, "leaves stride must be at least num_sites");
  for (uint64_t i = 0; i < (uint64_t)S * N; i++)
    if (msa[i] > 15) return host_fail(B3H_ERR_ARG, "not a DnaNucleotide byte at " + std::to_string(i));
  if (device < 0) return dedup_host(msa, S, N, leaves, stride, weights, s2p, np);
  return dedup_device(device, msa, S, N, leaves, stride, weights, s2p, np);
}

b3h_status b3d_deduplicate_noalloc(const uint8_t *msa, uint32_t S, uint64_t N, int device,
                                   uint8_t *leaves, uint64_t stride, uint32_t *weights,
                                   uint64_t *s2p, uint64_t *np) {
  if (!np) return host_fail(B3H_ERR_ARG, "null pattern count");
  *np = 0;
  if (S == 0) return host_fail(B3H_ERR_ARG, "no sequences");
  if (device < 0) return host_fail(B3H_ERR_ARG, "the no-allocation form runs on a device");
  if (N == 0) return B3H_OK;
  if (!msa || !leaves || !weights || !s2p) return host_fail(B3H_ERR_ARG, "null array");
  if (stride < N) return host_fail(B3H_ERR_ARG, "leaves stride must be at least num_sites");
  if (N >= 0xFFFFFFFFull) return host_fail(B3H_ERR_UNSUPPORTED, "more than 2^32 - 2 sites");
  for (uint64_t i = 0; i < (uint64_t)S * N; i++)
    if (msa[i] > 15) return host_fail(B3H_ERR_ARG, "not a DnaNucleotide byte at " + std::to_string(i));
  return dedup_device_noalloc(device, msa, S, N, leaves, stride, weights, s2p, np);
}

}  // extern "C"
