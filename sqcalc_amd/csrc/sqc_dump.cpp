// sqc_dump.cpp -- LAMMPS text dump reader that feeds sqc_hip_accumulate.
//
// Reads the default "atoms" dump style with the reference reader's rules
// (src/sqc_dump.f90:161-283):
//   - a frame starts at the line "ITEM: TIMESTEP" (leading blanks allowed),
//     the next line holds the step;
//   - up to "ITEM: ATOMS" the header records NUMBER OF ATOMS and BOX BOUNDS
//     (three values per row when the row has them, else two; the periodicity
//     flags are the last three words of the BOX BOUNDS line); other lines are
//     skipped;
//   - the ATOMS line names the columns (case-insensitive, at most 64); id is
//     optional, type and one coordinate per axis are required;
//   - then one atom per line, every named column numeric.
// Positions leave the reader continuous, as in src/sqc_dump.f90:300-364:
// xu yu zu pass through; x y z are unwrapped atom by atom, adding the
// minimum-image step in fractional coordinates since the previous frame, and
// come back as A s (no origin).  Beyond the reference, the scaled columns
// xs ys zs (unwrapped the same way) and xsu ysu zsu (returned as o + A s)
// are accepted.  Error codes are the reference reader's ierr values.
//
// Speed.  The file is memory-mapped.  The atom block of a frame is cut into
// byte ranges, one per worker thread.  The workers count the line ends in
// their range; from the prefix of those counts each one knows the index of
// its first line and parses its lines straight into the frame arrays,
// converting scaled coordinates and unwrapping on the way.  Numbers take an
// exact fast path when the significand has at most 19 digits and fits in 53
// bits and the power of ten is at most 10^22 (both factors are then exact
// doubles, so one multiply or divide rounds correctly); strtod, which also
// rounds correctly, takes the rest.  Either way a value is the double nearest
// to its decimal text, as Fortran's list-directed read gives.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sqc_hip.h"

namespace {

constexpr int kMaxColumns = 64;               // src/sqc_dump.f90:16
constexpr size_t kMinChunk = 256 * 1024;      // bytes of atom text per worker, at least
constexpr int kPopulateRead = 22;             // MADV_POPULATE_READ (Linux 5.14)

const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                           1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                           1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

inline bool is_digit(char c) { return (unsigned)(c - '0') < 10u; }
inline bool is_blank(char c) { return c == ' ' || c == '\t'; }
// value separators of a list-directed record: blanks and commas
inline bool is_sep(char c) { return c == ' ' || c == '\t' || c == ',' || c == '\r'; }
inline char lower(char c) { return (c >= 'A' && c <= 'Z') ? char(c + 32) : c; }

// floor and Fortran anint (round half away from zero), exact, without a
// library call for |x| < 2^52 (every larger double is an integer already)
inline double trunc_small(double x) { return double((long long)x); }
inline double ffloor(double x) {
  if (!(std::fabs(x) < 4503599627370496.0)) return x;
  const double t = trunc_small(x);
  return t > x ? t - 1.0 : t;
}
inline double fanint(double x) {
  if (!(std::fabs(x) < 4503599627370496.0)) return x;
  const double t = trunc_small(x);
  const double r = x - t;   // exact
  return r >= 0.5 ? t + 1.0 : (r <= -0.5 ? t - 1.0 : t);
}

bool parse_special(const char *b, const char *e, double &out) {
  // IEEE specials a Fortran read accepts: Inf, Infinity, NaN (any case)
  bool neg = false;
  if (b < e && (*b == '+' || *b == '-')) neg = *b++ == '-';
  std::string w;
  for (const char *p = b; p < e; ++p) w.push_back(lower(*p));
  if (w == "inf" || w == "infinity") {
    out = neg ? -HUGE_VAL : HUGE_VAL;
    return true;
  }
  if (w == "nan") {
    out = std::nan("");
    return true;
  }
  return false;
}

// The Fortran real in [b, e): [sign] digits [. digits] [exponent], where the
// exponent is a letter e, d or q (any case) with an optional sign, or a sign
// alone, followed by digits.
bool parse_real(const char *b, const char *e, double &out) {
  const char *p = b;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  // every digit goes into w; with at most 19 digits in all, w is exact
  uint64_t w = 0;
  const char *m0 = p;
  for (; p < e && is_digit(*p); ++p) w = w * 10 + uint64_t(*p - '0');
  int nd = int(p - m0), frac = 0;
  if (p < e && *p == '.') {
    const char *f0 = ++p;
    for (; p < e && is_digit(*p); ++p) w = w * 10 + uint64_t(*p - '0');
    frac = int(p - f0);
    nd += frac;
  }
  const char *m1 = p;
  if (nd == 0) return parse_special(b, e, out);
  const bool slow = nd > 19;
  long ex = 0;
  if (p < e) {
    const char c = lower(*p);
    if (c == 'e' || c == 'd' || c == 'q') ++p;
    else if (c != '+' && c != '-') return false;
    bool eneg = false;
    if (p < e && (*p == '+' || *p == '-')) eneg = *p++ == '-';
    if (p == e) return false;
    for (; p < e; ++p) {
      if (!is_digit(*p)) return false;
      if (ex < 100000) ex = ex * 10 + (*p - '0');
    }
    if (eneg) ex = -ex;
  }
  if (!slow && w <= (uint64_t(1) << 53)) {
    const long k = ex - frac;
    if (w == 0) {
      out = neg ? -0.0 : 0.0;
      return true;
    }
    if (k >= 0 && k <= 22) {
      out = double(w) * kPow10[k];
      out = neg ? -out : out;
      return true;
    }
    if (k < 0 && k >= -22) {
      out = double(w) / kPow10[-k];
      out = neg ? -out : out;
      return true;
    }
  }
  // the same number in C syntax for strtod
  char buf[160];
  std::string big;
  const size_t mlen = size_t(m1 - m0);
  char *s = buf;
  if (mlen + 32 > sizeof(buf)) {
    big.resize(mlen + 32);
    s = &big[0];
  }
  size_t n = 0;
  if (neg) s[n++] = '-';
  std::memcpy(s + n, m0, mlen);
  n += mlen;
  n += size_t(std::snprintf(s + n, 24, "e%ld", ex));
  char *end = nullptr;
  errno = 0;
  out = std::strtod(s, &end);
  // a finite number beyond the double range is a read error, as in Fortran
  if (errno == ERANGE && std::fabs(out) == HUGE_VAL) return false;
  return end == s + n;
}

// Eight ASCII digits at p, as one 64-bit word: the SWAR method of D. Lemire,
// "Quickly parsing eight digits" (2018): subtract '0' from every byte, then
// combine pairs, quads and the two halves with three multiplies.
inline bool eight_digits(const char *p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return (((v & 0xF0F0F0F0F0F0F0F0ull) | (((v + 0x0606060606060606ull) & 0xF0F0F0F0F0F0F0F0ull) >> 4)) ==
          0x3333333333333333ull);
}
inline uint32_t parse_eight(const char *p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  v -= 0x3030303030303030ull;
  v = (v * 10) + (v >> 8);
  v = (((v & 0x000000FF000000FFull) * (100 + (1000000ull << 32))) +
       (((v >> 16) & 0x000000FF000000FFull) * (1 + (10000ull << 32)))) >> 32;
  return (uint32_t)v;
}

// One pass over the common form [sign] digits [. digits] [e|E [sign] digits]
// ending at a separator or the line end, with at most 19 digits, a 53-bit
// significand and |power of ten| <= 22: returns the end, or nullptr for
// anything else (the caller then takes parse_real's full path).  The value
// is the same correctly rounded product or quotient parse_real forms.
inline const char *fast_real(const char *p, const char *le, double &out) {
  bool neg = false;
  if (p < le && (*p == '+' || *p == '-')) neg = *p++ == '-';
  uint64_t w = 0;
  const char *d0 = p;
  while (le - p >= 8 && eight_digits(p)) w = w * 100000000ull + parse_eight(p), p += 8;
  for (; p < le && is_digit(*p); ++p) w = w * 10 + uint64_t(*p - '0');
  int nd = int(p - d0), frac = 0;
  if (p < le && *p == '.') {
    const char *f0 = ++p;
    while (le - p >= 8 && eight_digits(p)) w = w * 100000000ull + parse_eight(p), p += 8;
    for (; p < le && is_digit(*p); ++p) w = w * 10 + uint64_t(*p - '0');
    frac = int(p - f0);
    nd += frac;
  }
  if (nd == 0 || nd > 19 || w > (uint64_t(1) << 53)) return nullptr;
  int ex = 0;
  if (p < le && (*p == 'e' || *p == 'E')) {
    ++p;
    bool eneg = false;
    if (p < le && (*p == '+' || *p == '-')) eneg = *p++ == '-';
    const char *e0 = p;
    for (; p < le && is_digit(*p) && p - e0 < 4; ++p) ex = ex * 10 + (*p - '0');
    if (p == e0) return nullptr;
    if (eneg) ex = -ex;
  }
  if (p < le && !is_sep(*p)) return nullptr;
  const int k = ex - frac;
  if (w == 0) {
    out = neg ? -0.0 : 0.0;
    return p;
  }
  if (k > 22 || k < -22) return nullptr;
  const double r = k >= 0 ? double(w) * kPow10[k] : double(w) / kPow10[-k];
  out = neg ? -r : r;
  return p;
}

// A Fortran integer: [sign] digits.
bool parse_int(const char *b, const char *e, int64_t &out) {
  const char *p = b;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  if (p == e) return false;
  uint64_t v = 0;
  for (; p < e; ++p) {
    if (!is_digit(*p)) return false;
    v = v * 10 + uint64_t(*p - '0');
    if (v > (uint64_t(1) << 62)) return false;
  }
  out = neg ? -int64_t(v) : int64_t(v);
  return true;
}

// First value of a list-directed record, as [b, e) of the token.
bool first_token(const char *b, const char *e, const char *&tb, const char *&te) {
  while (b < e && is_sep(*b)) ++b;
  if (b == e) return false;
  tb = b;
  while (b < e && !is_sep(*b)) ++b;
  te = b;
  return true;
}

// Fork-join pool: run(f) calls f(t) for t = 0..size-1, t = 0 on the caller.
class Pool {
 public:
  explicit Pool(int n) : n_(std::max(1, n)) {
    for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    go_.notify_all();
    for (auto &t : th_) t.join();
  }
  int size() const { return n_; }
  void run(const std::function<void(int)> &f) {
    if (n_ == 1) {
      f(0);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      job_ = &f;
      left_ = n_ - 1;
      ++gen_;
    }
    go_.notify_all();
    f(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return left_ == 0; });
  }

 private:
  void loop(int t) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(m_);
      go_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      const std::function<void(int)> *f = job_;
      lk.unlock();
      (*f)(t);
      lk.lock();
      if (--left_ == 0) done_.notify_one();
    }
  }
  int n_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable go_, done_;
  const std::function<void(int)> *job_ = nullptr;
  uint64_t gen_ = 0;
  int left_ = 0;
  bool stop_ = false;
};

// Restricted triclinic cell from LAMMPS bounds (src/sqc_cell.f90:39-63):
// a holds the lattice vectors in its columns, column-major.
void cell_from_bounds(const double lo[3], const double hi[3], const double tilt[3], double a[9],
                      double origin[3]) {
  const double xy = tilt[0], xz = tilt[1], yz = tilt[2];
  const double xlo = lo[0] - std::min(std::min(0.0, xy), std::min(xz, xy + xz));
  const double xhi = hi[0] - std::max(std::max(0.0, xy), std::max(xz, xy + xz));
  const double ylo = lo[1] - std::min(0.0, yz);
  const double yhi = hi[1] - std::max(0.0, yz);
  origin[0] = xlo, origin[1] = ylo, origin[2] = lo[2];
  const double c[9] = {xhi - xlo, 0.0, 0.0, xy, yhi - ylo, 0.0, xz, yz, hi[2] - lo[2]};
  std::memcpy(a, c, sizeof(c));
}

// A^-1 as the adjugate over the determinant (column-major in and out).
void inverse3(const double a[9], double inv[9]) {
  auto m = [&](int i, int j) { return a[(j - 1) * 3 + (i - 1)]; };
  const double det = m(1, 1) * (m(2, 2) * m(3, 3) - m(2, 3) * m(3, 2)) -
                     m(1, 2) * (m(2, 1) * m(3, 3) - m(2, 3) * m(3, 1)) +
                     m(1, 3) * (m(2, 1) * m(3, 2) - m(2, 2) * m(3, 1));
  double r[3][3];
  r[0][0] = m(2, 2) * m(3, 3) - m(2, 3) * m(3, 2);
  r[1][0] = -(m(2, 1) * m(3, 3) - m(2, 3) * m(3, 1));
  r[2][0] = m(2, 1) * m(3, 2) - m(2, 2) * m(3, 1);
  r[0][1] = -(m(1, 2) * m(3, 3) - m(1, 3) * m(3, 2));
  r[1][1] = m(1, 1) * m(3, 3) - m(1, 3) * m(3, 1);
  r[2][1] = -(m(1, 1) * m(3, 2) - m(1, 2) * m(3, 1));
  r[0][2] = m(1, 2) * m(2, 3) - m(1, 3) * m(2, 2);
  r[1][2] = -(m(1, 1) * m(2, 3) - m(1, 3) * m(2, 1));
  r[2][2] = m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) inv[j * 3 + i] = r[i][j] / det;
}

}  // namespace

struct sqc_dump_reader {
  int fd = -1;
  const char *base = nullptr;
  size_t size = 0;
  size_t cur = 0;         // offset of the next unread line
  int64_t lineno = 0;     // lines read
  int broken = 0;         // error that ended the reader (it stays ended)
  std::string message;
  Pool *pool = nullptr;
  // columns of the last ATOMS header (1-based, 0 = absent)
  int ncol = 0, idx_id = 0, idx_type = 0, idx_x = 0, idx_y = 0, idx_z = 0;
  int scaled = 0;
  int from_unwrapped = 0;
  // header values persist across frames like the reference's frame_t
  int64_t natoms = 0, timestep = 0;
  double cell[9] = {}, origin[3] = {};
  int pbc[3] = {1, 1, 1};
  int triclinic = 0;
  // unwrapping state (src/sqc_dump.f90:78-89)
  int have_prev = 0;
  std::vector<double> prev_frac, unwrapped_frac;
  std::vector<int64_t> prev_id;
  int64_t n_ambiguous = 0;
  // the last frame
  std::vector<double> pos;
  std::vector<int64_t> id;
  std::vector<int32_t> type;
  size_t block_bytes = 0;  // bytes of the last atom block, the next window's guess
  int64_t frames = 0;
  // a batch read running on a thread of its own (sqc_dump_read_frames_begin)
  std::thread bg;
  int bg_rc = 0;
  int64_t bg_n = 0;
};

namespace {

int fail(sqc_dump_reader *r, int code, const std::string &msg) {
  r->broken = code;
  r->message = msg + " (line " + std::to_string(r->lineno) + ")";
  return code;
}

bool next_line(sqc_dump_reader *r, const char *&b, const char *&e) {
  if (r->cur >= r->size) return false;
  b = r->base + r->cur;
  const char *nl = (const char *)std::memchr(b, '\n', r->size - r->cur);
  const size_t len = nl ? size_t(nl - b) : r->size - r->cur;
  r->cur += len + (nl ? 1 : 0);
  e = b + len;
  if (e > b && e[-1] == '\r') --e;
  ++r->lineno;
  return true;
}

// "ITEM: <label>" after leading blanks (src/sqc_dump.f90:443-449)
bool is_item(const char *b, const char *e, const char *label) {
  while (b < e && *b == ' ') ++b;
  static const char kItem[] = "ITEM: ";
  const size_t li = sizeof(kItem) - 1, ll = std::strlen(label);
  if (size_t(e - b) < li + ll) return false;
  return std::memcmp(b, kItem, li) == 0 && std::memcmp(b + li, label, ll) == 0;
}

// Whitespace-separated words of [b, e), at most kMaxColumns.
std::vector<std::string> words(const char *b, const char *e) {
  std::vector<std::string> w;
  while (b < e && (int)w.size() < kMaxColumns) {
    while (b < e && is_blank(*b)) ++b;
    if (b == e) break;
    const char *s = b;
    while (b < e && !is_blank(*b)) ++b;
    w.emplace_back(s, b);
  }
  return w;
}

// Column map of an ATOMS header (src/sqc_dump.f90:496-545), plus xs/xsu.
int resolve_columns(sqc_dump_reader *r, const std::vector<std::string> &tok) {
  r->ncol = (int)tok.size();
  int cart[3] = {0, 0, 0}, sc[3] = {0, 0, 0};
  bool cart_u[3] = {false, false, false}, sc_u[3] = {false, false, false};
  r->idx_id = r->idx_type = 0;
  static const char *kAxis[3] = {"x", "y", "z"};
  for (int i = 0; i < (int)tok.size(); ++i) {
    std::string n = tok[i];
    for (auto &c : n) c = lower(c);
    if (n == "id") r->idx_id = i + 1;
    else if (n == "type") r->idx_type = i + 1;
    for (int d = 0; d < 3; ++d) {
      const std::string a = kAxis[d];
      if ((n == a || n == a + "u") && !cart[d]) cart[d] = i + 1, cart_u[d] = n.size() == 2;
      if ((n == a + "s" || n == a + "su") && !sc[d]) sc[d] = i + 1, sc_u[d] = n.size() == 3;
    }
  }
  const bool use_cart = cart[0] && cart[1] && cart[2];
  const bool use_sc = !use_cart && sc[0] && sc[1] && sc[2];
  const int *c = use_sc ? sc : cart;
  const bool *u = use_sc ? sc_u : cart_u;
  r->idx_x = c[0], r->idx_y = c[1], r->idx_z = c[2];
  r->scaled = use_sc;
  r->from_unwrapped = u[0] && u[1] && u[2];
  if (r->idx_type == 0 || !(use_cart || use_sc)) return SQC_DUMP_ERR_COLUMNS;
  return SQC_DUMP_OK;
}

// Header of the next frame, up to and including the ATOMS line
// (src/sqc_dump.f90:175-256).
int read_header(sqc_dump_reader *r) {
  const char *b, *e, *tb, *te;
  if (!next_line(r, b, e)) return SQC_DUMP_END;
  if (!is_item(b, e, "TIMESTEP")) return fail(r, SQC_DUMP_ERR_HEADER, "expected ITEM: TIMESTEP");
  if (!next_line(r, b, e) || !first_token(b, e, tb, te) || !parse_int(tb, te, r->timestep))
    return fail(r, SQC_DUMP_ERR_HEADER, "bad time step");
  bool have_box = false;
  for (;;) {
    if (!next_line(r, b, e)) return fail(r, SQC_DUMP_ERR_HEADER, "end of file in a frame header");
    if (is_item(b, e, "NUMBER OF ATOMS")) {
      int64_t n = 0;
      if (!next_line(r, b, e)) return fail(r, SQC_DUMP_ERR_HEADER, "end of file in a frame header");
      if (!first_token(b, e, tb, te) || !parse_int(tb, te, n) || n <= 0)
        return fail(r, SQC_DUMP_ERR_NATOMS, "bad number of atoms");
      r->natoms = n;
    } else if (is_item(b, e, "BOX BOUNDS")) {
      const char *h = b;
      while (h < e && *h == ' ') ++h;
      h += std::strlen("ITEM: BOX BOUNDS");
      r->triclinic = std::string(b, e).find("xy") != std::string::npos;
      const std::vector<std::string> flags = words(h, e);
      r->pbc[0] = r->pbc[1] = r->pbc[2] = 1;
      if (flags.size() >= 3)
        for (int d = 0; d < 3; ++d) r->pbc[d] = flags[flags.size() - 3 + d][0] == 'p';
      double lo[3], hi[3], tilt[3];
      for (int row = 0; row < 3; ++row) {
        if (!next_line(r, b, e)) return fail(r, SQC_DUMP_ERR_HEADER, "end of file in BOX BOUNDS");
        double v[3] = {0.0, 0.0, 0.0};
        int nv = 0;
        const char *p = b;
        while (nv < 3 && first_token(p, e, tb, te)) {
          if (!parse_real(tb, te, v[nv])) break;
          ++nv;
          p = te;
        }
        if (nv < 2) return fail(r, SQC_DUMP_ERR_BOX, "bad BOX BOUNDS row");
        if (nv == 2) v[2] = 0.0;
        lo[row] = v[0], hi[row] = v[1], tilt[row] = v[2];
      }
      cell_from_bounds(lo, hi, tilt, r->cell, r->origin);
      have_box = true;
    } else if (is_item(b, e, "ATOMS")) {
      const char *h = b;
      while (h < e && *h == ' ') ++h;
      h += std::strlen("ITEM: ATOMS");
      if (int rc = resolve_columns(r, words(h, e)))
        return fail(r, rc, "the ATOMS line needs a type column and x y z, xu yu zu, xs ys zs or "
                           "xsu ysu zsu");
      break;
    }
  }
  if (!have_box) return fail(r, SQC_DUMP_ERR_NO_ATOMS, "frame without BOX BOUNDS");
  return SQC_DUMP_OK;
}

struct Chunk {
  size_t b = 0, e = 0;     // byte range [b, e) of the window
  int64_t lines = 0;       // line ends in the range
  int64_t first = 0;       // index of the first line that ends in the range
};

// The atom block: natoms lines from r->cur, parsed into out (3 per atom).
int read_atoms(sqc_dump_reader *r, double *out) {
  const int64_t N = r->natoms;
  const size_t start = r->cur, avail = r->size - r->cur;
  // guess the block's size: the last frame's, or the mean of its first lines
  size_t guess = r->block_bytes;
  if (!guess) {
    size_t p = start;
    int64_t k = 0;
    while (k < std::min<int64_t>(N, 512) && p < r->size) {
      const char *nl = (const char *)std::memchr(r->base + p, '\n', r->size - p);
      p = nl ? size_t(nl - r->base) + 1 : r->size;
      ++k;
    }
    guess = k ? size_t(double(p - start) / double(k) * double(N)) : avail;
  }
  size_t window = std::min(avail, guess + guess / 64 + 4096);
  // map the window's pages in one call: page faults taken by many workers at
  // once serialise in the kernel (a kernel without MADV_POPULATE_READ just
  // refuses the advice, and the workers fault the pages in themselves)
  {
    const uintptr_t pg = 4096, a0 = (uintptr_t)(r->base + start) & ~(pg - 1);
    madvise((void *)a0, (uintptr_t)(r->base + start + window) - a0, kPopulateRead);
  }
  const int T = (int)std::max<size_t>(1, std::min<size_t>(r->pool->size(), window / kMinChunk));
  std::vector<Chunk> ch(T);
  for (int t = 0; t < T; ++t) {
    ch[t].b = start + window * t / T;
    ch[t].e = start + window * (t + 1) / T;
  }
  const char *base = r->base;
  auto count_job = [&](int t) {
    if (t >= T) return;
    int64_t n = 0;
    for (size_t i = ch[t].b; i < ch[t].e; ++i) n += base[i] == '\n';
    ch[t].lines = n;
  };
  r->pool->run(count_job);
  int64_t total = 0;
  for (int t = 0; t < T; ++t) ch[t].first = total, total += ch[t].lines;
  // the guess fell short: take the remaining lines into the last range
  size_t wend = ch[T - 1].e;
  while (total < N && wend < r->size) {
    const char *nl = (const char *)std::memchr(base + wend, '\n', r->size - wend);
    wend = nl ? size_t(nl - base) + 1 : r->size;
    if (nl) ++total, ++ch[T - 1].lines;
  }
  ch[T - 1].e = wend;
  // a last line without its newline still counts
  bool tail_line = false;
  if (total < N && wend == r->size && r->size > start && base[r->size - 1] != '\n') {
    tail_line = true;
    ++total, ++ch[T - 1].lines;
  }
  if (total < N) {
    r->lineno += total;
    return fail(r, SQC_DUMP_ERR_ATOM_LINE, "end of file in the atom block");
  }
  const bool check_id = !r->from_unwrapped && r->idx_id > 0 && !r->prev_id.empty();
  if (!r->from_unwrapped && r->have_prev && (int64_t)r->prev_frac.size() != 3 * N)
    return fail(r, SQC_DUMP_ERR_COUNT, "the number of atoms changes between frames");
  if (check_id && (int64_t)r->prev_id.size() != N)
    return fail(r, SQC_DUMP_ERR_ORDER, "the atom order (id) changes between frames");
  r->id.resize(N);
  r->type.resize(N);
  if (!r->from_unwrapped && !r->have_prev) {
    r->prev_frac.assign(3 * N, 0.0);
    r->unwrapped_frac.assign(3 * N, 0.0);
  }
  const bool track_id = !r->from_unwrapped && r->idx_id > 0;
  if (track_id && r->prev_id.empty()) r->prev_id.assign(N, 0);

  double inv[9];
  inverse3(r->cell, inv);
  const double *A = r->cell, *o = r->origin;
  const int ncol = r->ncol, ix = r->idx_x - 1, iy = r->idx_y - 1, iz = r->idx_z - 1,
            iid = r->idx_id - 1, ity = r->idx_type - 1;
  const bool scaled = r->scaled, unwr = r->from_unwrapped, first = !r->have_prev;
  const int pbc[3] = {r->pbc[0], r->pbc[1], r->pbc[2]};
  std::vector<int64_t> bad_line(T, -1), bad_order(T, -1), ambiguous(T, 0);
  std::atomic<size_t> block_end{start};  // stays at start for a frame of no atoms
  auto parse_job = [&](int t) {
    if (t >= T) return;
    int64_t i = ch[t].first;
    const int64_t iend = std::min<int64_t>(ch[t].first + ch[t].lines, N);
    if (i >= iend) return;
    // the first line ending in this range starts after the previous newline
    size_t p = ch[t].b;
    if (t > 0) {
      while (p > start && base[p - 1] != '\n') --p;
    }
    double v[kMaxColumns];
    int64_t amb = 0;
    for (; i < iend; ++i) {
      const char *lb = base + p;
      const char *nl = (const char *)std::memchr(lb, '\n', r->size - p);
      const char *le = nl ? nl : base + r->size;
      p = size_t(le - base) + 1;
      if (i == N - 1) block_end.store(nl ? p : r->size);
      int k = 0;
      const char *q = lb;
      while (k < ncol) {
        while (q < le && is_sep(*q)) ++q;
        if (q == le) break;
        if (const char *ne = fast_real(q, le, v[k])) {
          q = ne;
          ++k;
          continue;
        }
        const char *tb = q;
        while (q < le && !is_sep(*q)) ++q;
        // r*c stands for r copies of c (list-directed repeat count)
        const char *star = (const char *)std::memchr(tb, '*', size_t(q - tb));
        int64_t rep = 1;
        if (star && (!parse_int(tb, star, rep) || rep <= 0)) break;
        if (!parse_real(star ? star + 1 : tb, q, v[k])) break;
        for (++k; --rep > 0 && k < ncol; ++k) v[k] = v[k - 1];
      }
      if (k < ncol) {
        bad_line[t] = i;
        return;
      }
      const int64_t aid = iid >= 0 ? (int64_t)fanint(v[iid]) : i + 1;
      r->id[i] = aid;
      r->type[i] = (int32_t)fanint(v[ity]);
      const double x = v[ix], y = v[iy], z = v[iz];
      double *dst = out + 3 * i;
      if (unwr) {
        if (scaled) {
          dst[0] = o[0] + (A[0] * x + A[3] * y + A[6] * z);
          dst[1] = o[1] + (A[1] * x + A[4] * y + A[7] * z);
          dst[2] = o[2] + (A[2] * x + A[5] * y + A[8] * z);
        } else {
          dst[0] = x, dst[1] = y, dst[2] = z;
        }
        continue;
      }
      if (track_id) {
        if (check_id && r->prev_id[i] != aid && bad_order[t] < 0) bad_order[t] = i;
        r->prev_id[i] = aid;
      }
      double s[3];
      if (scaled) {
        s[0] = x, s[1] = y, s[2] = z;
      } else {
        const double dx = x - o[0], dy = y - o[1], dz = z - o[2];
        s[0] = inv[0] * dx + inv[3] * dy + inv[6] * dz;
        s[1] = inv[1] * dx + inv[4] * dy + inv[7] * dz;
        s[2] = inv[2] * dx + inv[5] * dy + inv[8] * dz;
      }
      for (int d = 0; d < 3; ++d)
        if (pbc[d]) s[d] -= ffloor(s[d]);
      double *pf = &r->prev_frac[3 * i], *uf = &r->unwrapped_frac[3 * i];
      if (first) {
        for (int d = 0; d < 3; ++d) pf[d] = s[d], uf[d] = s[d];
      } else {
        bool amb1 = false;
        for (int d = 0; d < 3; ++d) {
          double ds = s[d] - pf[d];
          if (pbc[d]) {
            ds -= fanint(ds);
            if (std::fabs(ds) > 0.4) amb1 = true;
          }
          pf[d] = s[d];
          uf[d] += ds;
        }
        amb += amb1;
      }
      dst[0] = A[0] * uf[0] + A[3] * uf[1] + A[6] * uf[2];
      dst[1] = A[1] * uf[0] + A[4] * uf[1] + A[7] * uf[2];
      dst[2] = A[2] * uf[0] + A[5] * uf[1] + A[8] * uf[2];
    }
    ambiguous[t] = amb;
  };
  r->pool->run(parse_job);
  (void)tail_line;
  for (int t = 0; t < T; ++t)
    if (bad_line[t] >= 0) {
      r->lineno += bad_line[t] + 1;
      return fail(r, SQC_DUMP_ERR_ATOM_LINE, "malformed atom line");
    }
  for (int t = 0; t < T; ++t)
    if (bad_order[t] >= 0) {
      r->lineno += N;
      return fail(r, SQC_DUMP_ERR_ORDER, "the atom order (id) changes between frames");
    }
  for (int t = 0; t < T; ++t) r->n_ambiguous += ambiguous[t];
  if (!r->from_unwrapped) r->have_prev = 1;
  r->lineno += N;
  r->block_bytes = block_end.load() - start;
  r->cur = block_end.load();
  return SQC_DUMP_OK;
}

void fill_frame(const sqc_dump_reader *r, sqc_dump_frame *fr) {
  if (!fr) return;
  fr->natoms = r->natoms;
  fr->timestep = r->timestep;
  std::memcpy(fr->cell, r->cell, sizeof(fr->cell));
  std::memcpy(fr->origin, r->origin, sizeof(fr->origin));
  for (int d = 0; d < 3; ++d) fr->pbc[d] = r->pbc[d];
  fr->triclinic = r->triclinic;
  fr->scaled = r->scaled;
  fr->has_id = r->idx_id > 0;
  fr->unwrapped = 1;
  fr->unwrapped_from_dump = r->from_unwrapped;
  fr->n_ambiguous = r->n_ambiguous;
  fr->lineno = r->lineno;
  fr->frames = r->frames;
}

// Reads one frame; the positions go to out when it is non-null and the frame
// has expect_n atoms, else to the reader's own array.
int next_frame(sqc_dump_reader *r, sqc_dump_frame *fr, double *out, int64_t expect_n) {
  if (!r || r->fd < 0) return SQC_DUMP_ERR_CLOSED;
  if (r->broken) return r->broken;
  if (int rc = read_header(r)) return rc;
  double *dst = out;
  if (!dst || r->natoms != expect_n) {
    r->pos.resize(3 * r->natoms);
    dst = r->pos.data();
  }
  if (int rc = read_atoms(r, dst)) return rc;
  ++r->frames;
  fill_frame(r, fr);
  return SQC_DUMP_OK;
}

}  // namespace

extern "C" {

int sqc_dump_open(sqc_dump_reader **reader, const char *path, int nthreads) {
  if (!reader || !path) return SQC_DUMP_ERR_ARG;
  *reader = nullptr;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return SQC_DUMP_ERR_OPEN;
  struct stat st;
  if (fstat(fd, &st) != 0) {
    ::close(fd);
    return SQC_DUMP_ERR_OPEN;
  }
  auto *r = new sqc_dump_reader();
  r->fd = fd;
  r->size = (size_t)st.st_size;
  if (r->size) {
    void *m = mmap(nullptr, r->size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
      ::close(fd);
      delete r;
      return SQC_DUMP_ERR_OPEN;
    }
    madvise(m, r->size, MADV_SEQUENTIAL);
    r->base = (const char *)m;
  }
  int n = nthreads;
  if (n <= 0) {
    // all host threads (OMP_NUM_THREADS when set), less -nthreads of them
    int all = (int)std::thread::hardware_concurrency();
    if (const char *env = std::getenv("OMP_NUM_THREADS"))
      if (std::atoi(env) > 0) all = std::atoi(env);
    n = all + nthreads;
  }
  r->pool = new Pool(std::max(1, std::min(n, 256)));
  *reader = r;
  return SQC_DUMP_OK;
}

int sqc_dump_next(sqc_dump_reader *reader, sqc_dump_frame *frame) {
  return next_frame(reader, frame, nullptr, -1);
}

int sqc_dump_copy(const sqc_dump_reader *reader, double *pos, int64_t *id, int32_t *type_id) {
  if (!reader || reader->fd < 0) return SQC_DUMP_ERR_CLOSED;
  if (!reader->frames) return SQC_DUMP_ERR_ARG;
  const size_t n = (size_t)reader->natoms;
  if (pos) {
    if (reader->pos.size() != 3 * n) return SQC_DUMP_ERR_ARG;
    std::memcpy(pos, reader->pos.data(), sizeof(double) * 3 * n);
  }
  if (id) std::memcpy(id, reader->id.data(), sizeof(int64_t) * n);
  if (type_id) std::memcpy(type_id, reader->type.data(), sizeof(int32_t) * n);
  return SQC_DUMP_OK;
}

int sqc_dump_read_frames(sqc_dump_reader *reader, int64_t maxframes, int64_t natoms,
                         const double cell[9], double *pos, int64_t *nread,
                         sqc_dump_frame *last) {
  if (!reader || maxframes < 0 || natoms <= 0 || !pos || !nread) return SQC_DUMP_ERR_ARG;
  *nread = 0;
  double amax = 0.0;
  if (cell)
    for (int k = 0; k < 9; ++k) amax = std::max(amax, std::fabs(cell[k]));
  for (int64_t f = 0; f < maxframes; ++f) {
    const int rc = next_frame(reader, last, pos + (size_t)f * 3 * natoms, natoms);
    if (rc == SQC_DUMP_END) break;
    if (rc) return rc;
    // the driver's per-frame checks (app/main.f90:240-247)
    if (reader->natoms != natoms)
      return fail(reader, SQC_DUMP_ERR_NATOMS_CHANGED, "number of atoms changes between frames");
    if (cell) {
      double dmax = 0.0;
      for (int k = 0; k < 9; ++k) dmax = std::max(dmax, std::fabs(reader->cell[k] - cell[k]));
      if (dmax > 1.0e-6 * std::max(1.0, amax))
        return fail(reader, SQC_DUMP_ERR_BOX_CHANGED, "the simulation box changes between frames");
    }
    *nread = f + 1;
  }
  return SQC_DUMP_OK;
}

int sqc_dump_read_frames_begin(sqc_dump_reader *reader, int64_t maxframes, int64_t natoms,
                               const double cell[9], double *pos) {
  if (!reader || maxframes < 0 || natoms <= 0 || !pos) return SQC_DUMP_ERR_ARG;
  if (reader->bg.joinable()) return SQC_DUMP_ERR_ARG;  // one batch in flight at a time
  std::array<double, 9> c{};
  const bool has_cell = cell != nullptr;
  if (has_cell) std::copy(cell, cell + 9, c.begin());
  reader->bg_rc = 0;
  reader->bg_n = 0;
  reader->bg = std::thread([reader, maxframes, natoms, pos, c, has_cell] {
    reader->bg_rc = sqc_dump_read_frames(reader, maxframes, natoms, has_cell ? c.data() : nullptr,
                                         pos, &reader->bg_n, nullptr);
  });
  return SQC_DUMP_OK;
}

int sqc_dump_read_frames_end(sqc_dump_reader *reader, int64_t *nread) {
  if (!reader || !nread) return SQC_DUMP_ERR_ARG;
  if (!reader->bg.joinable()) return SQC_DUMP_ERR_ARG;
  reader->bg.join();
  *nread = reader->bg_n;
  return reader->bg_rc;
}

int sqc_dump_close(sqc_dump_reader *reader) {
  if (!reader) return SQC_DUMP_OK;
  if (reader->bg.joinable()) reader->bg.join();
  if (reader->base) munmap((void *)reader->base, reader->size);
  if (reader->fd >= 0) ::close(reader->fd);
  delete reader->pool;
  delete reader;
  return SQC_DUMP_OK;
}

const char *sqc_dump_message(const sqc_dump_reader *reader) {
  return reader ? reader->message.c_str() : "";
}

const char *sqc_dump_error_string(int code) {
  switch (code) {
    case SQC_DUMP_OK: return "ok";
    case SQC_DUMP_END: return "end of the trajectory";
    case SQC_DUMP_ERR_OPEN: return "cannot open the dump file";
    case SQC_DUMP_ERR_CLOSED: return "the reader is not open";
    case SQC_DUMP_ERR_HEADER: return "malformed frame header";
    case SQC_DUMP_ERR_NATOMS: return "bad NUMBER OF ATOMS";
    case SQC_DUMP_ERR_BOX: return "bad BOX BOUNDS";
    case SQC_DUMP_ERR_NO_ATOMS: return "frame without BOX BOUNDS or ATOMS";
    case SQC_DUMP_ERR_ATOM_LINE: return "malformed or missing atom line";
    case SQC_DUMP_ERR_COLUMNS: return "missing type or coordinate columns";
    case SQC_DUMP_ERR_ORDER: return "the atom order (id) changes between frames";
    case SQC_DUMP_ERR_COUNT: return "the number of atoms changes between frames";
    case SQC_DUMP_ERR_NATOMS_CHANGED: return "number of atoms changes between frames";
    case SQC_DUMP_ERR_BOX_CHANGED: return "the simulation box changes between frames";
    case SQC_DUMP_ERR_ARG: return "invalid argument";
    default: return "unknown dump reader status";
  }
}

}  // extern "C"
