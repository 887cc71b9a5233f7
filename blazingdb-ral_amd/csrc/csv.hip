/*
 * csv.hip — ingest: delimited text to device-resident columns.
 *
 * The reference loads a table with csv_parser::parse (io/data_parser/CSVParser.cpp:
 * 178-238), which hands the file's bytes to cuDF's read_csv with the reader options
 * of init_default_csv_args (:26-47): one-byte delimiter and line terminator,
 * quote '"' with minimal quoting and no doubled quotes, blank lines skipped, no
 * header, no NA detection (every parsed row is valid). data_loader::load_data
 * (io/DataLoader.cpp:26-118) concatenates the columns parsed from each file.
 *
 * The bytes sit in HBM; three passes over them and one over the rows:
 *   1. per group of 16 chunks of 4 KiB (one wave walks them in order; a lane takes
 *      64 contiguous bytes of a chunk): the parity of its quote characters and
 *      the number of rows that start in it, for either quote state at its first
 *      byte (the state is not known yet);
 *   2. one workgroup scans the group summaries: the quote state at each group's
 *      start and the index of its first row;
 *   3. each group again, now with its quote state: the byte offset of every row
 *      start, written at the row's index (a lane gathers its starts in a bit mask
 *      and stores one word per row);
 *   4. one thread per row walks its fields (delimiters outside quotes), parses
 *      the requested ones and stores them, row-major across threads so stores of
 *      a wave are contiguous.
 * A row starts at byte 0 or after a line terminator outside quotes; a row that is
 * empty (or only "\r") is blank and skipped. Bytes are read through aligned 16-byte
 * windows; nothing is read past the buffer's last byte.
 *
 * Field values (the same arithmetic as oracle/csv_oracle.py):
 *   integers: optional sign, then decimal digits up to the first other byte, in
 *     wrapping 64-bit arithmetic, narrowed by truncation to the column's width;
 *   binary64: up to 19 significant digits m and a decimal exponent e (dropped
 *     digits truncate); m <= 2^53 and |e| <= 22 is one exact IEEE multiply or
 *     divide by 10^|e| (correctly rounded); otherwise m and 10^e are applied in
 *     steps of 10^22 (within a few ulp, deterministic);
 *   binary32: m <= 2^24 and |e| <= 10 is one binary32 operation (correctly
 *     rounded); otherwise the binary64 value rounded to binary32;
 *   dates: YYYY-MM-DD, optionally followed by T or a space, hh:mm:ss, .fff and Z;
 *     GDF_DATE32 days since 1970-01-01, GDF_DATE64 / GDF_TIMESTAMP milliseconds.
 * An empty or unparsable field gives the value of the digits read (0 for none);
 * a row with fewer fields gives 0 for the missing ones.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/blz/blz_ral.h"

namespace blz {
namespace {

#define CSV_DEV __device__ __forceinline__

typedef uint32_t csv_u32x4 __attribute__((ext_vector_type(4)));

constexpr int kCsvLaneBytes = 64;
constexpr int kCsvChunk = 64 * kCsvLaneBytes;  // bytes per wave
constexpr int kCsvBlock = 256;
constexpr int kCsvGroup = 16;                  // chunks per summary (one wave walks them in order)
constexpr int kCsvMaxCols = 256;               // file columns a row may be walked through

struct CsvSum {        // a run of bytes, for either quote state s at its start:
  uint32_t par;        //   quote characters in it, mod 2
  uint32_t pad;
  uint64_t rows[2];    //   rows starting in it
};
struct CsvStart {      // a chunk's quote state and first row index
  uint64_t rows;
  uint32_t par;
  uint32_t pad;
};

CSV_DEV CsvSum csv_combine(const CsvSum &a, const CsvSum &b) {
  CsvSum r;
  r.par = a.par ^ b.par;
  r.pad = 0;
  r.rows[0] = a.rows[0] + b.rows[a.par];
  r.rows[1] = a.rows[1] + b.rows[a.par ^ 1];
  return r;
}

// Bytes [p, p + 64) of the lane (zeros past n) plus the byte before (`term` at the
// buffer's start: byte 0 follows a line end) and the byte after (0 past n).
struct LaneBytes {
  uint32_t w[16];
  uint8_t prev, next;
  CSV_DEV uint8_t at(int k) const { return (uint8_t)(w[k >> 2] >> (8 * (k & 3))); }
};
CSV_DEV void load_lane(const uint8_t *buf, int64_t n, int64_t p, uint8_t term, LaneBytes &b) {
  if (p + kCsvLaneBytes <= n && (((uintptr_t)(buf + p)) & 15) == 0) {
    const csv_u32x4 *q = (const csv_u32x4 *)(buf + p);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const csv_u32x4 v = __builtin_nontemporal_load(q + i);
      b.w[4 * i] = v.x;
      b.w[4 * i + 1] = v.y;
      b.w[4 * i + 2] = v.z;
      b.w[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      uint32_t x = 0;
      for (int j = 0; j < 4; j++) {
        const int64_t k = p + 4 * i + j;
        if (k < n) x |= (uint32_t)buf[k] << (8 * j);
      }
      b.w[i] = x;
    }
  }
  b.prev = p > 0 && p - 1 < n ? buf[p - 1] : term;
  b.next = p + kCsvLaneBytes < n ? buf[p + kCsvLaneBytes] : 0;
}

// Row starts in the lane's bytes for quote state `s` at its first byte (bit s of
// the result's counts), and the lane's quote parity. At a row start j the byte
// before is a terminator outside quotes; the row is not blank (its first byte is
// not a terminator, nor a "\r" followed by one or by the end).
struct LaneScan {
  uint32_t par;
  uint32_t cnt[2];
};
CSV_DEV LaneScan scan_lane(const LaneBytes &b, int64_t p, int64_t n, uint8_t term, uint8_t quote) {
  LaneScan r{0, {0, 0}};
  uint8_t prev = b.prev;
  uint32_t q = 0;  // quotes seen in the lane before byte k, mod 2
  const int64_t lim = n - p;  // bytes of the lane inside the buffer
#pragma unroll
  for (int k = 0; k < kCsvLaneBytes; k++) {
    const uint8_t c = b.at(k);
    const uint8_t nx = k + 1 < kCsvLaneBytes ? b.at(k + 1) : b.next;
    const bool in = k < lim;
    const bool nx_end = k + 1 >= lim;
    const bool blank = c == term || (c == '\r' && (nx_end || nx == term));
    // a row start counts for start state s when s ^ q == 0 (outside quotes)
    const uint32_t hit = (in && prev == term && !blank) ? 1u : 0u;
    r.cnt[0] += q == 0 ? hit : 0u;
    r.cnt[1] += q == 1 ? hit : 0u;
    q ^= (in && c == quote) ? 1u : 0u;
    prev = c;
  }
  r.par = q;
  return r;
}

// A chunk's summary (all lanes hold it): quote parity and rows for either start state.
CSV_DEV CsvSum chunk_sum(const LaneScan &ls, uint64_t pm, uint32_t pre) {
  uint64_t c0 = pre ? ls.cnt[1] : ls.cnt[0];  // lane counts for chunk start state s: state s ^ pre
  uint64_t c1 = pre ? ls.cnt[0] : ls.cnt[1];
  for (int o = 32; o > 0; o >>= 1) {
    c0 += __shfl_xor(c0, o);
    c1 += __shfl_xor(c1, o);
  }
  CsvSum s;
  s.par = (uint32_t)__builtin_popcountll(pm) & 1;
  s.pad = 0;
  s.rows[0] = c0;
  s.rows[1] = c1;
  return s;
}

// Pass 1: one summary per group of kCsvGroup consecutive chunks (one wave walks
// the group's chunks in order).
__global__ void __launch_bounds__(kCsvBlock) csv_count_kernel(const uint8_t *__restrict__ buf, int64_t n,
                                                              uint8_t term, uint8_t quote, CsvSum *sums,
                                                              int64_t n_chunks, int64_t n_groups) {
  const int lane = threadIdx.x & 63;
  const int64_t wave0 = ((int64_t)blockIdx.x * kCsvBlock + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * kCsvBlock) >> 6;
  for (int64_t g = wave0; g < n_groups; g += n_waves) {
    CsvSum acc{0, 0, {0, 0}};
    const int64_t c_end = (g + 1) * kCsvGroup < n_chunks ? (g + 1) * kCsvGroup : n_chunks;
    for (int64_t c = g * kCsvGroup; c < c_end; c++) {
      const int64_t p = c * kCsvChunk + (int64_t)lane * kCsvLaneBytes;
      LaneBytes b;
      load_lane(buf, n, p, term, b);
      const LaneScan ls = scan_lane(b, p, n, term, quote);
      // lanes before this one: quote parity (exclusive prefix)
      const uint64_t pm = __builtin_amdgcn_ballot_w64(ls.par != 0);
      const uint32_t pre = (uint32_t)__builtin_popcountll(pm & ((1ull << lane) - 1)) & 1;
      acc = csv_combine(acc, chunk_sum(ls, pm, pre));
    }
    if (lane == 0) sums[g] = acc;
  }
}

// Pass 2, one workgroup: exclusive scan of the group summaries (the file starts
// outside quotes): starts[g] = quote state and first row of group g; *total = rows.
__global__ void __launch_bounds__(1024) csv_scan_kernel(const CsvSum *__restrict__ sums, int64_t n_groups,
                                                        CsvStart *starts, uint64_t *total) {
  __shared__ CsvSum part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n_groups + 1023) / 1024;
  const int64_t lo = t * per < n_groups ? t * per : n_groups;
  const int64_t hi = lo + per < n_groups ? lo + per : n_groups;
  CsvSum acc{0, 0, {0, 0}};
  for (int64_t c = lo; c < hi; c++) acc = csv_combine(acc, sums[c]);
  part[t] = acc;
  __syncthreads();
  if (t == 0) {
    CsvSum run{0, 0, {0, 0}};
    for (int i = 0; i < 1024; i++) {
      const CsvSum x = part[i];
      part[i] = run;
      run = csv_combine(run, x);
    }
    *total = run.rows[0];
  }
  __syncthreads();
  CsvSum run = part[t];
  for (int64_t c = lo; c < hi; c++) {
    CsvStart st;
    st.rows = run.rows[0];
    st.par = run.par;
    st.pad = 0;
    starts[c] = st;
    run = csv_combine(run, sums[c]);
  }
}

// Pass 3: row start offsets, rows [skip, skip + n_out) of the file at index
// row - skip. A wave walks its group's chunks in order from the group's state.
__global__ void __launch_bounds__(kCsvBlock) csv_rows_kernel(const uint8_t *__restrict__ buf, int64_t n,
                                                             uint8_t term, uint8_t quote,
                                                             const CsvStart *__restrict__ starts,
                                                             int64_t n_chunks, int64_t n_groups, int64_t skip,
                                                             int64_t n_out, int64_t *row_start) {
  const int lane = threadIdx.x & 63;
  const int64_t wave0 = ((int64_t)blockIdx.x * kCsvBlock + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * kCsvBlock) >> 6;
  for (int64_t g = wave0; g < n_groups; g += n_waves) {
    const CsvStart st = starts[g];
    uint64_t rows = st.rows;
    uint32_t par = st.par;
    const int64_t c_end = (g + 1) * kCsvGroup < n_chunks ? (g + 1) * kCsvGroup : n_chunks;
    for (int64_t c = g * kCsvGroup; c < c_end; c++) {
      const int64_t p = c * kCsvChunk + (int64_t)lane * kCsvLaneBytes;
      LaneBytes b;
      load_lane(buf, n, p, term, b);
      // this lane's row starts (bit k: a row starts at byte p + k) for its actual
      // quote state, and its quote parity
      const int64_t lim = n - p;
      uint32_t q = 0;
      uint64_t hit0 = 0, hit1 = 0;  // for lane start state 0 / 1
      uint8_t prev = b.prev;
#pragma unroll
      for (int k = 0; k < kCsvLaneBytes; k++) {
        const uint8_t ch = b.at(k);
        const uint8_t nx = k + 1 < kCsvLaneBytes ? b.at(k + 1) : b.next;
        const bool in = k < lim;
        const bool nx_end = k + 1 >= lim;
        const bool blank = ch == term || (ch == '\r' && (nx_end || nx == term));
        const uint64_t h = (in && prev == term && !blank) ? (1ull << k) : 0ull;
        hit0 |= q == 0 ? h : 0ull;
        hit1 |= q == 1 ? h : 0ull;
        q ^= (in && ch == quote) ? 1u : 0u;
        prev = ch;
      }
      const uint64_t pm = __builtin_amdgcn_ballot_w64(q != 0);
      const uint32_t s = par ^ ((uint32_t)__builtin_popcountll(pm & ((1ull << lane) - 1)) & 1);
      uint64_t hits = s ? hit1 : hit0;
      const uint64_t mine = (uint64_t)__builtin_popcountll(hits);
      uint64_t incl = mine;
      for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      int64_t r = (int64_t)(rows + incl - mine);
      while (hits) {
        const int k = __builtin_ctzll(hits);
        hits &= hits - 1;
        const int64_t o = r - skip;
        if (o >= 0 && o < n_out) row_start[o] = p + k;
        r++;
      }
      rows += (uint64_t)__shfl(incl, 63);
      par ^= (uint32_t)__builtin_popcountll(pm) & 1;
    }
  }
}

// ---- field parsing (one thread per row) ----

// Sequential reader over bytes [pos, n) through an aligned 16-byte window.
struct ByteReader {
  const uint8_t *buf;
  int64_t n, base;
  uint64_t lo, hi;  // bytes base .. base + 7, base + 8 .. base + 15
  bool vec;         // buf is 16-byte aligned: whole windows load as one vector
  CSV_DEV void fill(int64_t pos) {
    base = pos & ~(int64_t)15;
    if (vec && base + 16 <= n) {
      const csv_u32x4 v = *(const csv_u32x4 *)(buf + base);
      lo = (uint64_t)v.x | ((uint64_t)v.y << 32);
      hi = (uint64_t)v.z | ((uint64_t)v.w << 32);
    } else {
      lo = hi = 0;
      for (int j = 0; j < 16; j++) {
        const int64_t k = base + j;
        const uint64_t x = k < n ? (uint64_t)buf[k] : 0;
        if (j < 8) lo |= x << (8 * j);
        else hi |= x << (8 * (j - 8));
      }
    }
  }
  CSV_DEV uint8_t get(int64_t pos) {
    if (pos < base || pos >= base + 16) fill(pos);
    const int k = (int)(pos - base);
    return (uint8_t)((k < 8 ? lo >> (8 * k) : hi >> (8 * (k - 8))) & 0xFF);
  }
};

CSV_DEV bool is_digit(uint8_t c) { return c >= '0' && c <= '9'; }

// Integer: optional sign, digits up to the first other byte (wrapping).
CSV_DEV int64_t parse_int(ByteReader &rd, int64_t a, int64_t e) {
  bool neg = false;
  if (a < e) {
    const uint8_t c = rd.get(a);
    if (c == '-' || c == '+') { neg = c == '-'; a++; }
  }
  uint64_t v = 0;
  for (; a < e; a++) {
    const uint8_t c = rd.get(a);
    if (!is_digit(c)) break;
    v = v * 10 + (uint64_t)(c - '0');
  }
  return (int64_t)(neg ? (uint64_t)0 - v : v);
}

// Decimal scan of a real: sign, up to 19 significant digits m, decimal exponent.
struct RealScan {
  bool neg;
  uint64_t m;
  int64_t e;
};
CSV_DEV RealScan scan_real(ByteReader &rd, int64_t a, int64_t end) {
  RealScan r{false, 0, 0};
  if (a < end) {
    const uint8_t c = rd.get(a);
    if (c == '-' || c == '+') { r.neg = c == '-'; a++; }
  }
  int nd = 0;
  bool dot = false;
  for (; a < end; a++) {
    const uint8_t c = rd.get(a);
    if (c == '.' && !dot) { dot = true; continue; }
    if (!is_digit(c)) break;
    const int d = c - '0';
    if (nd == 0 && d == 0) {  // leading zero: not significant
      if (dot) r.e--;
      continue;
    }
    if (nd < 19) {
      r.m = r.m * 10 + (uint64_t)d;
      nd++;
      if (dot) r.e--;
    } else if (!dot) {
      r.e++;  // integer digit beyond 19: dropped, the value keeps its magnitude
    }
  }
  if (a < end) {
    const uint8_t c = rd.get(a);
    if (c == 'e' || c == 'E') {
      a++;
      bool eneg = false;
      if (a < end) {
        const uint8_t s = rd.get(a);
        if (s == '-' || s == '+') { eneg = s == '-'; a++; }
      }
      int64_t x = 0;
      for (; a < end; a++) {
        const uint8_t d = rd.get(a);
        if (!is_digit(d)) break;
        if (x < 100000) x = x * 10 + (d - '0');
      }
      r.e += eneg ? -x : x;
    }
  }
  return r;
}

__constant__ double kP10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
__constant__ float kP10f[11] = {1e0f, 1e1f, 1e2f, 1e3f, 1e4f, 1e5f, 1e6f, 1e7f, 1e8f, 1e9f, 1e10f};

CSV_DEV double real_f64(const RealScan &r) {
  double v = (double)r.m;
  if (r.m != 0) {
    int64_t e = r.e;
    if (r.m <= (1ull << 53) && e >= -22 && e <= 22) {
      v = e < 0 ? v / kP10[-e] : v * kP10[e];
    } else {
      while (e > 22) { v *= 1e22; e -= 22; }
      while (e < -22) { v /= 1e22; e += 22; }
      v = e < 0 ? v / kP10[-e] : v * kP10[e];
    }
  }
  return r.neg ? -v : v;
}
CSV_DEV float real_f32(const RealScan &r) {
  if (r.m != 0 && r.m <= (1ull << 24) && r.e >= -10 && r.e <= 10) {
    float v = (float)r.m;
    v = r.e < 0 ? v / kP10f[-r.e] : v * kP10f[r.e];
    return r.neg ? -v : v;
  }
  return (float)real_f64(r);
}

// days from civil date (proleptic Gregorian), Hinnant's days_from_civil
CSV_DEV int64_t days_from_civil(int64_t y, int64_t m, int64_t d) {
  y -= m <= 2;
  const int64_t era = (y >= 0 ? y : y - 399) / 400;
  const int64_t yoe = y - era * 400;
  const int64_t mp = (m + 9) % 12;
  const int64_t doy = (153 * mp + 2) / 5 + d - 1;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}

// fixed-width unsigned decimal of up to `w` digits at a (stops at a non-digit)
CSV_DEV int64_t digits(ByteReader &rd, int64_t &a, int64_t end, int w) {
  int64_t v = 0;
  for (int i = 0; i < w && a < end; i++, a++) {
    const uint8_t c = rd.get(a);
    if (!is_digit(c)) break;
    v = v * 10 + (c - '0');
  }
  return v;
}
CSV_DEV bool take(ByteReader &rd, int64_t &a, int64_t end, uint8_t c) {
  if (a < end && rd.get(a) == c) { a++; return true; }
  return false;
}
// milliseconds since the epoch of YYYY-MM-DD[(T| )hh:mm:ss[.fff]][Z]
CSV_DEV int64_t parse_ms(ByteReader &rd, int64_t a, int64_t end, int64_t *days_out) {
  const int64_t y = digits(rd, a, end, 4);
  take(rd, a, end, '-');
  const int64_t mo = digits(rd, a, end, 2);
  take(rd, a, end, '-');
  const int64_t d = digits(rd, a, end, 2);
  const int64_t days = days_from_civil(y, mo, d);
  *days_out = days;
  int64_t ms = 0;
  if (take(rd, a, end, 'T') || take(rd, a, end, ' ')) {
    const int64_t hh = digits(rd, a, end, 2);
    take(rd, a, end, ':');
    const int64_t mi = digits(rd, a, end, 2);
    take(rd, a, end, ':');
    const int64_t ss = digits(rd, a, end, 2);
    int64_t frac = 0;
    if (take(rd, a, end, '.')) {
      int nd = 0;
      while (a < end && nd < 3 && is_digit(rd.get(a))) { frac = frac * 10 + (rd.get(a) - '0'); a++; nd++; }
      while (nd < 3) { frac *= 10; nd++; }
    }
    ms = ((hh * 60 + mi) * 60 + ss) * 1000 + frac;
  }
  return days * 86400000 + ms;
}

struct CsvOut {
  void *data[kCsvMaxCols];   // by file column; NULL: not requested
  int8_t dtype[kCsvMaxCols];
};

// One thread per row: walks fields 0 .. last_col (delimiters outside quotes, up to
// the row's terminator outside quotes) and stores the requested ones.
__global__ void __launch_bounds__(kCsvBlock) csv_fields_kernel(const uint8_t *__restrict__ buf, int64_t n,
                                                               uint8_t delim, uint8_t term, uint8_t quote,
                                                               const int64_t *__restrict__ row_start,
                                                               int64_t n_rows, int last_col, const CsvOut *out) {
  const int64_t r = (int64_t)blockIdx.x * kCsvBlock + threadIdx.x;
  if (r >= n_rows) return;
  ByteReader rd{buf, n, -32, 0, 0, (((uintptr_t)buf) & 15) == 0};
  int64_t a = row_start[r];
  bool inq = false;
  for (int f = 0; f <= last_col; f++) {
    // field [a, e): up to a delimiter or terminator outside quotes
    int64_t e = a;
    bool at_end = a >= n;
    while (e < n) {
      const uint8_t c = rd.get(e);
      if (c == quote) inq = !inq;
      else if (!inq && (c == delim || c == term)) break;
      e++;
    }
    const bool row_end = e >= n || rd.get(e) == term;
    void *dst = out->data[f];
    if (dst) {
      int64_t fa = a, fe = e;
      if (fe > fa && rd.get(fe - 1) == '\r' && row_end) fe--;           // CRLF
      if (fe > fa && rd.get(fa) == quote) {                              // quoted field
        fa++;
        if (fe > fa && rd.get(fe - 1) == quote) fe--;
      }
      if (at_end) fe = fa;  // a missing field
      switch (out->dtype[f]) {
        case GDF_INT8: ((int8_t *)dst)[r] = (int8_t)parse_int(rd, fa, fe); break;
        case GDF_INT16: ((int16_t *)dst)[r] = (int16_t)parse_int(rd, fa, fe); break;
        case GDF_INT32: ((int32_t *)dst)[r] = (int32_t)parse_int(rd, fa, fe); break;
        case GDF_INT64: ((int64_t *)dst)[r] = parse_int(rd, fa, fe); break;
        case GDF_FLOAT32: ((float *)dst)[r] = real_f32(scan_real(rd, fa, fe)); break;
        case GDF_FLOAT64: ((double *)dst)[r] = real_f64(scan_real(rd, fa, fe)); break;
        case GDF_DATE32: {
          int64_t days = 0;
          (void)parse_ms(rd, fa, fe, &days);
          ((int32_t *)dst)[r] = (int32_t)days;
        } break;
        default: {  // GDF_DATE64, GDF_TIMESTAMP
          int64_t days = 0;
          ((int64_t *)dst)[r] = parse_ms(rd, fa, fe, &days);
        } break;
      }
    }
    if (row_end) {
      // the row has no more fields: the rest read as missing
      for (int g = f + 1; g <= last_col; g++) {
        void *d2 = out->data[g];
        if (!d2) continue;
        switch (out->dtype[g]) {
          case GDF_INT8: ((int8_t *)d2)[r] = 0; break;
          case GDF_INT16: ((int16_t *)d2)[r] = 0; break;
          case GDF_INT32: case GDF_DATE32: ((int32_t *)d2)[r] = 0; break;
          case GDF_FLOAT32: ((float *)d2)[r] = 0.0f; break;
          case GDF_FLOAT64: ((double *)d2)[r] = 0.0; break;
          default: ((int64_t *)d2)[r] = 0; break;
        }
      }
      break;
    }
    a = e + 1;
  }
}

// ---- host ----

struct CsvError : std::runtime_error {
  int code;
  CsvError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void hipc(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  throw CsvError(e == hipErrorOutOfMemory ? BLZ_ERR_OOM : BLZ_ERR_HIP, std::string(what) + ": " +
                                                                           hipGetErrorString(e));
}
#define CSV_HIP(x) hipc((x), #x)

int dtype_width(int dt) {
  switch (dt) {
    case GDF_INT8: return 1;
    case GDF_INT16: return 2;
    case GDF_INT32: case GDF_FLOAT32: case GDF_DATE32: return 4;
    case GDF_INT64: case GDF_FLOAT64: case GDF_DATE64: case GDF_TIMESTAMP: return 8;
    default: return 0;
  }
}

struct ParseArgs {
  char delimiter, line_terminator;
  int skip_rows;
  const int32_t *dtypes;
  int n_file_columns;
  const int32_t *column_indices;
  int n_indices;
};

// Parses n bytes at device address d_buf; fills out[0 .. n_indices) with
// callee-allocated columns. Returns the row count.
int64_t parse_device(const uint8_t *d_buf, int64_t n, const ParseArgs &a, gdf_column *out, hipStream_t st,
                     float *stage_ms = nullptr) {
  // stage_ms (measurement): device time of the count, scan, row and field kernels
  hipEvent_t ev[6] = {};  // before count | after count | after scan | before rows | after rows | after fields
  auto mark = [&](int i) {
    if (stage_ms) CSV_HIP(hipEventRecord(ev[i], st));
  };
  if (stage_ms) {
    for (auto &e : ev) CSV_HIP(hipEventCreate(&e));
    for (int i = 0; i < 4; i++) stage_ms[i] = 0.0f;
  }
  struct EvFree {
    hipEvent_t *e;
    ~EvFree() {
      for (int i = 0; i < 6; i++)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } ev_free{ev};
  if (a.n_file_columns <= 0 || a.n_file_columns > kCsvMaxCols)
    throw CsvError(BLZ_ERR_INVALID, "1 to 256 file columns");
  if (!a.dtypes) throw CsvError(BLZ_ERR_INVALID, "dtypes is NULL");
  std::vector<int> cols;
  if (a.column_indices) cols.assign(a.column_indices, a.column_indices + a.n_indices);
  else for (int i = 0; i < a.n_file_columns; i++) cols.push_back(i);
  std::vector<int> seen(a.n_file_columns, 0);
  int last = -1;
  for (int c : cols) {
    if (c < 0 || c >= a.n_file_columns) throw CsvError(BLZ_ERR_INVALID, "column index out of range");
    if (seen[c]++) throw CsvError(BLZ_ERR_INVALID, "column requested twice");
    if (dtype_width(a.dtypes[c]) == 0)
      throw CsvError(BLZ_ERR_UNSUPPORTED, "column " + std::to_string(c) + ": dtype not supported by the reader");
    last = std::max(last, c);
  }
  if (a.skip_rows < 0) throw CsvError(BLZ_ERR_INVALID, "skip_rows < 0");
  const uint8_t term = (uint8_t)a.line_terminator, delim = (uint8_t)a.delimiter, quote = '"';

  int64_t n_rows = 0;
  void *d_tmp = nullptr;
  std::vector<void *> allocated;
  try {
    if (n > 0) {
      const int64_t n_chunks = (n + kCsvChunk - 1) / kCsvChunk;
      const int64_t n_groups = (n_chunks + kCsvGroup - 1) / kCsvGroup;
      const size_t sums_b = (size_t)n_groups * sizeof(CsvSum), starts_b = (size_t)n_groups * sizeof(CsvStart);
      CSV_HIP(hipMallocAsync(&d_tmp, sums_b + starts_b + 64, st));
      CsvSum *sums = (CsvSum *)d_tmp;
      CsvStart *starts = (CsvStart *)((uint8_t *)d_tmp + sums_b);
      uint64_t *d_total = (uint64_t *)((uint8_t *)d_tmp + sums_b + starts_b);
      int dev = 0, cus = 0;
      CSV_HIP(hipGetDevice(&dev));
      CSV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const int64_t want = (n_groups + kCsvBlock / 64 - 1) / (kCsvBlock / 64);
      const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)cus * 16));
      mark(0);
      hipLaunchKernelGGL(csv_count_kernel, dim3(grid), dim3(kCsvBlock), 0, st, d_buf, n, term, quote, sums,
                         n_chunks, n_groups);
      CSV_HIP(hipGetLastError());
      mark(1);
      hipLaunchKernelGGL(csv_scan_kernel, dim3(1), dim3(1024), 0, st, sums, n_groups, starts, d_total);
      CSV_HIP(hipGetLastError());
      mark(2);
      uint64_t total = 0;
      CSV_HIP(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, st));
      CSV_HIP(hipStreamSynchronize(st));
      n_rows = std::max<int64_t>(0, (int64_t)total - a.skip_rows);
      if (n_rows >= (1ll << 31)) throw CsvError(BLZ_ERR_INVALID, "more than 2^31 - 1 rows");
      CsvOut h{};
      for (size_t i = 0; i < cols.size(); i++) {
        const int c = cols[i];
        void *p = nullptr;
        CSV_HIP(hipMalloc(&p, (size_t)std::max<int64_t>(n_rows, 1) * dtype_width(a.dtypes[c])));
        allocated.push_back(p);
        h.data[c] = p;
        h.dtype[c] = (int8_t)a.dtypes[c];
      }
      if (n_rows > 0) {
        int64_t *rows = nullptr;
        CSV_HIP(hipMallocAsync((void **)&rows, (size_t)n_rows * 8 + sizeof(CsvOut), st));
        CsvOut *d_out = (CsvOut *)(rows + n_rows);
        CSV_HIP(hipMemcpyAsync(d_out, &h, sizeof(CsvOut), hipMemcpyHostToDevice, st));
        mark(3);  // after the output allocations
        hipLaunchKernelGGL(csv_rows_kernel, dim3(grid), dim3(kCsvBlock), 0, st, d_buf, n, term, quote, starts,
                           n_chunks, n_groups, (int64_t)a.skip_rows, n_rows, rows);
        CSV_HIP(hipGetLastError());
        mark(4);
        const int64_t g2 = (n_rows + kCsvBlock - 1) / kCsvBlock;
        hipLaunchKernelGGL(csv_fields_kernel, dim3((unsigned)g2), dim3(kCsvBlock), 0, st, d_buf, n, delim, term,
                           quote, rows, n_rows, last, d_out);
        CSV_HIP(hipGetLastError());
        mark(5);
        CSV_HIP(hipFreeAsync(rows, st));
      }
      CSV_HIP(hipFreeAsync(d_tmp, st));
      d_tmp = nullptr;
      CSV_HIP(hipStreamSynchronize(st));
      if (stage_ms && n_rows > 0) {
        CSV_HIP(hipEventElapsedTime(&stage_ms[0], ev[0], ev[1]));
        CSV_HIP(hipEventElapsedTime(&stage_ms[1], ev[1], ev[2]));
        CSV_HIP(hipEventElapsedTime(&stage_ms[2], ev[3], ev[4]));
        CSV_HIP(hipEventElapsedTime(&stage_ms[3], ev[4], ev[5]));
      }
    } else {
      for (size_t i = 0; i < cols.size(); i++) {
        void *p = nullptr;
        CSV_HIP(hipMalloc(&p, 8));
        allocated.push_back(p);
      }
    }
  } catch (...) {
    if (d_tmp) (void)hipFreeAsync(d_tmp, st);
    for (void *p : allocated) (void)hipFree(p);
    throw;
  }
  for (size_t i = 0; i < cols.size(); i++) {
    gdf_column &g = out[i];
    std::memset(&g, 0, sizeof(g));
    g.data = allocated[i];
    g.valid = nullptr;
    g.size = (gdf_size_type)n_rows;
    g.dtype = (gdf_dtype)a.dtypes[cols[i]];
    g.null_count = 0;
  }
  return n_rows;
}

template <class F>
int csv_guard(char *err, size_t len, F &&f) {
  auto put = [&](const std::string &m) {
    if (err && len > 0) {
      const size_t k = std::min(len - 1, m.size());
      std::memcpy(err, m.data(), k);
      err[k] = 0;
    }
  };
  (void)hipGetLastError();
  try {
    f();
    put("");
    return BLZ_OK;
  } catch (const CsvError &e) {
    put(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    put("host out of memory");
    return BLZ_ERR_OOM;
  } catch (const std::exception &e) {
    put(e.what());
    return BLZ_ERR_INVALID;
  }
}

int n_out_columns(const ParseArgs &a) { return a.column_indices ? a.n_indices : a.n_file_columns; }

}  // namespace
}  // namespace blz

using namespace blz;

extern "C" {

int blz_csv_parse_device(const void *device_bytes, size_t n_bytes, char delimiter, char line_terminator,
                         int skip_rows, const int32_t *dtypes, int n_file_columns, const int32_t *column_indices,
                         int n_indices, gdf_column *out_columns, int64_t *n_rows, blz_stream_t stream, char *err,
                         size_t err_len) {
  return csv_guard(err, err_len, [&] {
    if (!out_columns || (!device_bytes && n_bytes)) throw CsvError(BLZ_ERR_INVALID, "bad arguments");
    const ParseArgs a{delimiter, line_terminator, skip_rows, dtypes, n_file_columns, column_indices, n_indices};
    const int64_t r = parse_device((const uint8_t *)device_bytes, (int64_t)n_bytes, a, out_columns,
                                   (hipStream_t)stream);
    if (n_rows) *n_rows = r;
  });
}

int blz_csv_parse_device_timed(const void *device_bytes, size_t n_bytes, char delimiter, char line_terminator,
                               int skip_rows, const int32_t *dtypes, int n_file_columns,
                               const int32_t *column_indices, int n_indices, gdf_column *out_columns,
                               int64_t *n_rows, float *stage_ms, blz_stream_t stream, char *err, size_t err_len) {
  return csv_guard(err, err_len, [&] {
    if (!out_columns || !stage_ms || (!device_bytes && n_bytes)) throw CsvError(BLZ_ERR_INVALID, "bad arguments");
    const ParseArgs a{delimiter, line_terminator, skip_rows, dtypes, n_file_columns, column_indices, n_indices};
    const int64_t r = parse_device((const uint8_t *)device_bytes, (int64_t)n_bytes, a, out_columns,
                                   (hipStream_t)stream, stage_ms);
    if (n_rows) *n_rows = r;
  });
}

int blz_csv_read(const void *host_bytes, size_t n_bytes, char delimiter, char line_terminator, int skip_rows,
                 const int32_t *dtypes, int n_file_columns, const int32_t *column_indices, int n_indices,
                 gdf_column *out_columns, int64_t *n_rows, blz_stream_t stream, char *err, size_t err_len) {
  return csv_guard(err, err_len, [&] {
    if (!out_columns || (!host_bytes && n_bytes)) throw CsvError(BLZ_ERR_INVALID, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    uint8_t *d = nullptr;
    CSV_HIP(hipMalloc((void **)&d, n_bytes + 16));
    try {
      if (n_bytes) CSV_HIP(hipMemcpyAsync(d, host_bytes, n_bytes, hipMemcpyHostToDevice, st));
      const ParseArgs a{delimiter, line_terminator, skip_rows, dtypes, n_file_columns, column_indices, n_indices};
      const int64_t r = parse_device(d, (int64_t)n_bytes, a, out_columns, st);
      if (n_rows) *n_rows = r;
    } catch (...) {
      (void)hipFree(d);
      throw;
    }
    CSV_HIP(hipFree(d));
  });
}

int blz_csv_load_files(const char *const *paths, int n_paths, char delimiter, char line_terminator, int skip_rows,
                       const int32_t *dtypes, int n_file_columns, const int32_t *column_indices, int n_indices,
                       gdf_column *out_columns, int64_t *n_rows, blz_stream_t stream, char *err, size_t err_len) {
  return csv_guard(err, err_len, [&] {
    if (!paths || n_paths <= 0 || !out_columns) throw CsvError(BLZ_ERR_INVALID, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const ParseArgs a{delimiter, line_terminator, skip_rows, dtypes, n_file_columns, column_indices, n_indices};
    const int nc = n_out_columns(a);
    std::vector<std::vector<gdf_column>> parts;
    std::vector<int64_t> counts;
    auto free_parts = [&] {
      for (auto &p : parts)
        for (auto &c : p) (void)hipFree(c.data);
    };
    try {
      for (int f = 0; f < n_paths; f++) {
        FILE *fp = std::fopen(paths[f], "rb");
        if (!fp) throw CsvError(BLZ_ERR_INVALID, std::string("cannot open ") + paths[f]);
        std::vector<uint8_t> bytes;
        uint8_t chunk[1 << 16];
        size_t k;
        while ((k = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) bytes.insert(bytes.end(), chunk, chunk + k);
        std::fclose(fp);
        std::vector<gdf_column> cols(nc);
        int64_t r = 0;
        char e2[512];
        const int rc = blz_csv_read(bytes.data(), bytes.size(), delimiter, line_terminator, skip_rows, dtypes,
                                    n_file_columns, column_indices, n_indices, cols.data(), &r, stream, e2,
                                    sizeof(e2));
        if (rc != BLZ_OK) throw CsvError(rc, std::string(paths[f]) + ": " + e2);
        parts.push_back(cols);
        counts.push_back(r);
      }
      // one file: its columns; several: concatenated (DataLoader.cpp:80-118)
      int64_t total = 0;
      for (int64_t c : counts) total += c;
      if (total >= (1ll << 31)) throw CsvError(BLZ_ERR_INVALID, "more than 2^31 - 1 rows");
      for (int i = 0; i < nc; i++) {
        const int w = dtype_width(parts[0][i].dtype);
        void *p = nullptr;
        if (parts.size() == 1) {
          p = parts[0][i].data;
          parts[0][i].data = nullptr;
        } else {
          CSV_HIP(hipMalloc(&p, (size_t)std::max<int64_t>(total, 1) * w));
          int64_t off = 0;
          for (size_t f = 0; f < parts.size(); f++) {
            if (counts[f])
              CSV_HIP(hipMemcpyAsync((uint8_t *)p + off * w, parts[f][i].data, (size_t)counts[f] * w,
                                     hipMemcpyDeviceToDevice, st));
            off += counts[f];
          }
        }
        gdf_column &g = out_columns[i];
        std::memset(&g, 0, sizeof(g));
        g.data = p;
        g.size = (gdf_size_type)total;
        g.dtype = parts[0][i].dtype;
      }
      CSV_HIP(hipStreamSynchronize(st));
      free_parts();
      if (n_rows) *n_rows = total;
    } catch (...) {
      free_parts();
      throw;
    }
  });
}

}  // extern "C"
