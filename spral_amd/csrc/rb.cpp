// Rutherford-Boeing sparse matrix files, the format of the reference's
// regression matrices (src/rutherford_boeing.f90):
//  * spral_rb_peek  : the header, as rb_peek (:75-233);
//  * spral_rb_write : an assembled real or pattern matrix, as rb_write
//    (:635-732), byte for byte: integer formats (kIw) sized from the largest
//    pointer / index (w = floor(log10(max)) + 2, k = 80 / w), values with
//    the caller's (kEw.d) format (default (3e24.16)), header fields padded
//    to their Fortran widths;
//  * spral_rb_read  : assembled real, integer or pattern files as rb_read
//    (:274-575) with values as in the file or pattern only, optional
//    missing diagonal entries (value 0), and the result in the lower
//    triangle as stored, flipped to the upper triangle (by a counting sort
//    over rows, as flip_lwr_upr :785-844), or expanded to both.  Random
//    values (rb_read_options%values 2-4) are not built (ERROR_VALUES).
// Fixed-width fields are read as Fortran formatted input reads them: k
// fields of width w per record, blanks ignored, D or a bare sign accepted
// as the exponent mark, a kP scale factor applied when a field has no
// exponent.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/spral_rb.h"

namespace {

// return codes of the reference (src/rutherford_boeing.f90:27-38)
constexpr int kSuccess = 0, kBadFile = -1, kNotRB = -2, kIO = -3, kType = -4, kEltAsm = -5,
              kMatrixType = -6, kLwrUprFull = -11, kValues = -13, kWarnAux = 1;

struct Fmt {
  int per_line = 1, width = 1, digits = 0, scale = 0;
  char kind = 'I';
};

// (kIw), (kEw.d), (kDw.d), (kFw.d), (kESw.d), optional leading "nP,"
bool parse_format(std::string f, Fmt &out) {
  std::string s;
  for (char c : f)
    if (!std::isspace((unsigned char)c)) s += (char)std::toupper((unsigned char)c);
  if (s.size() < 3 || s.front() != '(') return false;
  s = s.substr(1, s.find(')') == std::string::npos ? s.size() - 1 : s.find(')') - 1);
  size_t p = 0;
  auto num = [&](int &v) {
    size_t q = p;
    while (q < s.size() && std::isdigit((unsigned char)s[q])) ++q;
    if (q == p) return false;
    v = std::atoi(s.substr(p, q - p).c_str());
    p = q;
    return true;
  };
  Fmt r;
  int v = 0;
  size_t save = p;
  if (num(v) && p < s.size() && s[p] == 'P') {  // scale factor
    r.scale = v;
    ++p;
    if (p < s.size() && s[p] == ',') ++p;
  } else {
    p = save;
  }
  if (!num(r.per_line)) r.per_line = 1;
  if (p >= s.size()) return false;
  r.kind = s[p++];
  if (r.kind == 'E' && p < s.size() && s[p] == 'S') ++p;  // ES
  if (!std::strchr("IEDFG", r.kind)) return false;
  if (!num(r.width)) return false;
  if (p < s.size() && s[p] == '.') {
    ++p;
    if (!num(r.digits)) return false;
  }
  out = r;
  return r.per_line > 0 && r.width > 0;
}

// Fortran Ew.d output: [-]0.<d digits>E+xx (or +xxx beyond 99), right
// justified in w
std::string fortran_e(double v, int w, int d) {
  std::string mant(d, '0');
  int e10 = 0;
  const bool neg = std::signbit(v);
  const double a = std::fabs(v);
  if (a != 0.0) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", d - 1, a);  // correctly rounded
    const char *ep = std::strchr(buf, 'e');
    std::string digs;
    for (const char *c = buf; c < ep; ++c)
      if (std::isdigit((unsigned char)*c)) digs += *c;
    mant = digs.substr(0, d);
    e10 = std::atoi(ep + 1) + 1;
  }
  char ebuf[16];
  if (std::abs(e10) <= 99) std::snprintf(ebuf, sizeof ebuf, "E%c%02d", e10 < 0 ? '-' : '+', std::abs(e10));
  else std::snprintf(ebuf, sizeof ebuf, "%c%03d", e10 < 0 ? '-' : '+', std::abs(e10));
  std::string body = std::string(neg ? "-" : "") + "0." + mant + ebuf;
  if ((int)body.size() > w) {  // no room for the optional leading zero
    body = std::string(neg ? "-" : "") + "." + mant + ebuf;
    if ((int)body.size() > w) return std::string(w, '*');
  }
  return std::string(w - body.size(), ' ') + body;
}

std::string fortran_i(long long v, int w) {
  std::string s = std::to_string(v);
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

// as create_format (src/rutherford_boeing.f90:734-757)
std::string int_format(long long maxv, int &per_line, int &prec) {
  prec = (int)std::floor(std::log10((double)std::max<long long>(maxv, 1))) + 2;
  per_line = 80 / prec;
  return "(" + std::to_string(per_line) + "i" + std::to_string(prec) + ")";
}

std::string pad(const std::string &s, size_t w) {
  std::string t = s.substr(0, std::min(s.size(), w));
  return t + std::string(w - t.size(), ' ');
}

// items of one Fortran-formatted array: k fields of width w per record
bool read_fields(std::istream &in, const Fmt &f, int64_t count, std::vector<std::string> &out) {
  out.clear();
  out.reserve(count);
  std::string line;
  while ((int64_t)out.size() < count) {
    if (!std::getline(in, line)) return false;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    for (int k = 0; k < f.per_line && (int64_t)out.size() < count; ++k) {
      const size_t a = (size_t)k * f.width;
      out.push_back(a < line.size() ? line.substr(a, f.width) : std::string());
    }
  }
  return true;
}

bool to_int(const std::string &fld, long long &v) {
  std::string s;
  for (char c : fld)
    if (!std::isspace((unsigned char)c)) s += c;
  if (s.empty()) {
    v = 0;
    return true;
  }
  char *end = nullptr;
  v = std::strtoll(s.c_str(), &end, 10);
  return *end == 0;
}

bool to_real(const std::string &fld, const Fmt &f, double &v) {
  std::string s;
  for (char c : fld)
    if (!std::isspace((unsigned char)c)) s += (char)std::toupper((unsigned char)c);
  if (s.empty()) {
    v = 0.0;
    return true;
  }
  bool has_exp = false;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == 'D' || s[i] == 'Q') s[i] = 'E';
    if (s[i] == 'E') has_exp = true;
  }
  if (!has_exp) {  // a bare sign after the mantissa marks the exponent
    for (size_t i = 1; i < s.size(); ++i)
      if ((s[i] == '+' || s[i] == '-') && s[i - 1] != 'E') {
        s.insert(i, "E");
        has_exp = true;
        break;
      }
  }
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  if (*end != 0) return false;
  const bool has_point = s.find('.') != std::string::npos;
  if (!has_point && f.digits > 0 && f.kind != 'I') {
    // no decimal point: the last d digits of the mantissa are fractional
    size_t epos = s.find('E');
    std::string m = s.substr(0, epos);
    double mv = std::strtod(m.c_str(), nullptr) / std::pow(10.0, f.digits);
    double ev = (epos == std::string::npos) ? 0.0 : std::strtod(s.c_str() + epos + 1, nullptr);
    v = mv * std::pow(10.0, ev);
  }
  if (!has_exp && f.scale != 0) v /= std::pow(10.0, f.scale);
  return true;
}

struct Header {
  std::string title, id, type;
  long long t[4] = {0, 0, 0, 0};
  std::string ptrfmt, rowfmt, valfmt;
};

int read_header(std::istream &in, Header &h) {
  std::string l1, l2, l3, l4;
  if (!std::getline(in, l1) || !std::getline(in, l2) || !std::getline(in, l3)) return kIO;
  auto strip_cr = [](std::string &s) {
    if (!s.empty() && s.back() == '\r') s.pop_back();
  };
  strip_cr(l1);
  strip_cr(l3);
  l1.resize(std::max<size_t>(l1.size(), 80), ' ');
  h.title = l1.substr(0, 72);
  h.id = l1.substr(72, 8);
  l3.resize(std::max<size_t>(l3.size(), 70), ' ');
  h.type = l3.substr(0, 3);
  for (auto &c : h.type) c = (char)std::tolower((unsigned char)c);
  for (int k = 0; k < 4; ++k)
    if (!to_int(l3.substr(14 + 14 * k + 1, 13), h.t[k])) return kNotRB;
  if (!std::strchr("rcipq", h.type[0]) || !std::strchr("suhzr", h.type[1])) return kNotRB;
  if (h.type[2] != 'a' && h.type[2] != 'e') return kNotRB;
  if (h.type[2] == 'a' && h.t[3] != 0) return kNotRB;
  if (std::getline(in, l4)) {
    strip_cr(l4);
    l4.resize(std::max<size_t>(l4.size(), 52), ' ');
    h.ptrfmt = l4.substr(0, 16);
    h.rowfmt = l4.substr(16, 16);
    h.valfmt = l4.substr(32, 20);
  }
  return kSuccess;
}

// as sym_to_matrix_type (src/rutherford_boeing.f90:1003-1020)
int sym_to_matrix_type(char s) {
  switch (s) {
    case 'r': return 1;  // SPRAL_MATRIX_REAL_RECT
    case 's': return 4;  // SPRAL_MATRIX_REAL_SYM_INDEF
    case 'u': return 2;  // SPRAL_MATRIX_REAL_UNSYM
    case 'z': return 6;  // SPRAL_MATRIX_REAL_SKEW
    default: return 0;
  }
}

struct RBMatrix {
  int m = 0, n = 0, matrix_type = 0;
  std::vector<int64_t> ptr;
  std::vector<int> row;
  std::vector<double> val;
  bool has_val = false;
};

}  // namespace

extern "C" {

int spral_rb_peek(const char *filename, int *m, int *n, int64_t *nelt, int64_t *nvar,
                  int64_t *nval, int *matrix_type, char *type_code, char *title,
                  char *identifier) {
  std::ifstream in(filename);
  if (!in) return kBadFile;
  Header h;
  int rc = read_header(in, h);
  if (rc != kSuccess) return rc;
  const bool asm_ = h.type[2] == 'a';
  if (m) *m = (int)h.t[0];
  if (n) *n = asm_ ? (int)h.t[1] : (int)h.t[0];
  if (nelt) *nelt = asm_ ? 0 : h.t[1];
  if (nvar) *nvar = h.t[2];
  if (nval) *nval = asm_ ? h.t[2] : h.t[3];
  if (matrix_type) *matrix_type = sym_to_matrix_type(h.type[1]);
  if (type_code) std::memcpy(type_code, h.type.c_str(), 3);
  if (title) std::memcpy(title, h.title.c_str(), 72);
  if (identifier) std::memcpy(identifier, h.id.c_str(), 8);
  return kSuccess;
}

int spral_rb_write(const char *filename, int matrix_type, int m, int n, const int64_t *ptr,
                   const int *row, const double *val, const char *val_format,
                   const char *title, const char *identifier) {
  if (matrix_type < 0 || matrix_type > 6 || matrix_type == 5) return kMatrixType;
  static const char sym[7] = {'r', 'r', 'u', 's', 's', '?', 'z'};
  const std::string vfmt = val_format ? val_format : "(3e24.16)";
  Fmt vf;
  if (val && (!parse_format(vfmt, vf) || vf.kind == 'I')) return kIO;
  std::ofstream out(filename, std::ios::binary | std::ios::trunc);
  if (!out) return kBadFile;
  const int64_t nz = ptr[n] - 1;
  long long maxp = 0, maxr = 0;
  for (int j = 0; j <= n; ++j) maxp = std::max<long long>(maxp, ptr[j]);
  for (int64_t k = 0; k < nz; ++k) maxr = std::max<long long>(maxr, row[k]);
  int pl, pp, rl, rp;
  const std::string pfmt = int_format(maxp, pl, pp), rfmt = int_format(maxr, rl, rp);
  const long long plines = n / pl + 1;  // (size(ptr) - 1) / per_line + 1
  const long long rlines = (nz - 1) / rl + 1;
  const long long vlines = val ? (nz - 1) / vf.per_line + 1 : 0;
  std::ostringstream o;
  o << pad(title ? title : "Matrix", 72) << pad(identifier ? identifier : "0", 8) << "\n";
  o << fortran_i(plines + rlines + vlines, 14) << " " << fortran_i(plines, 13) << " "
    << fortran_i(rlines, 13) << " " << fortran_i(vlines, 13) << "\n";
  o << (val ? 'r' : 'p') << sym[matrix_type] << 'a' << std::string(11, ' ')
    << fortran_i(m, 14) << " " << fortran_i(n, 13) << " " << fortran_i(nz, 13) << " "
    << fortran_i(0, 13) << "\n";
  o << pad(pfmt, 16) << pad(rfmt, 16) << pad(vfmt, 20) << "\n";
  auto ints = [&](auto get, int64_t cnt, int per, int w) {
    for (int64_t k = 0; k < cnt; ++k) {
      o << fortran_i(get(k), w);
      if ((k + 1) % per == 0 || k + 1 == cnt) o << "\n";
    }
  };
  ints([&](int64_t k) { return (long long)ptr[k]; }, (int64_t)n + 1, pl, pp);
  ints([&](int64_t k) { return (long long)row[k]; }, nz, rl, rp);
  if (val)
    for (int64_t k = 0; k < nz; ++k) {
      o << fortran_e(val[k], vf.width, vf.digits);
      if ((k + 1) % vf.per_line == 0 || k + 1 == nz) o << "\n";
    }
  const std::string s = o.str();
  out.write(s.data(), (std::streamsize)s.size());
  return out ? kSuccess : kIO;
}

// Reads a file into a handle; *info as rb_read.  lwr_upr_full: 1 lower (as
// stored), 2 upper, 3 both triangles; values: 0 as in the file, 1 pattern.
void *spral_rb_read(const char *filename, int lwr_upr_full, int values, int add_diagonal,
                    int *m, int *n, int64_t *nz, int *has_val, int *matrix_type, int *info) {
  *info = kSuccess;
  if (lwr_upr_full < 1 || lwr_upr_full > 3) {
    *info = kLwrUprFull;
    return nullptr;
  }
  if (values < 0 || values > 1) {  // random values are not built
    *info = kValues;
    return nullptr;
  }
  std::ifstream in(filename);
  if (!in) {
    *info = kBadFile;
    return nullptr;
  }
  Header h;
  int rc = read_header(in, h);
  if (rc != kSuccess) {
    *info = rc;
    return nullptr;
  }
  if (h.type[2] != 'a') {
    *info = kEltAsm;
    return nullptr;
  }
  if (h.type[0] == 'c') {
    *info = kType;
    return nullptr;
  }
  auto *R = new RBMatrix;
  R->m = (int)h.t[0];
  R->n = (int)h.t[1];
  R->matrix_type = sym_to_matrix_type(h.type[1]);
  const int64_t cnt = h.t[2];
  if (h.type[0] == 'q') *info = kWarnAux;
  const bool symmetric = std::strchr("shz", h.type[1]) != nullptr;
  const bool skew = h.type[1] == 'z';
  const bool read_val = values == 0 && h.type[0] != 'p' && h.type[0] != 'q';
  Fmt pf, rf, vf;
  std::vector<std::string> fl;
  bool ok = parse_format(h.ptrfmt, pf) && parse_format(h.rowfmt, rf) &&
            read_fields(in, pf, (int64_t)R->n + 1, fl);
  R->ptr.resize(R->n + 1);
  for (int j = 0; ok && j <= R->n; ++j) {
    long long v;
    ok = to_int(fl[j], v);
    R->ptr[j] = v;
  }
  ok = ok && read_fields(in, rf, cnt, fl);
  R->row.resize(cnt);
  for (int64_t k = 0; ok && k < cnt; ++k) {
    long long v;
    ok = to_int(fl[k], v);
    R->row[k] = (int)v;
  }
  if (ok && read_val) {
    ok = parse_format(h.valfmt, vf) && read_fields(in, vf, cnt, fl);
    R->val.resize(cnt);
    for (int64_t k = 0; ok && k < cnt; ++k) {
      if (vf.kind == 'I') {
        long long v;
        ok = to_int(fl[k], v);
        R->val[k] = (double)v;
      } else {
        ok = to_real(fl[k], vf, R->val[k]);
      }
    }
    R->has_val = true;
  }
  if (!ok || R->ptr[0] != 1 || R->ptr[R->n] - 1 != cnt) {
    delete R;
    *info = kIO;
    return nullptr;
  }
  const int nn = R->n;
  if (add_diagonal) {  // as add_missing_diag: a zero diagonal entry first in its column
    std::vector<int64_t> np(nn + 1, 1);
    std::vector<int> nr;
    std::vector<double> nv;
    for (int c = 0; c < nn; ++c) {
      bool found = false;
      for (int64_t k = R->ptr[c] - 1; k < R->ptr[c + 1] - 1; ++k) found |= (R->row[k] == c + 1);
      if (!found && c < R->m) {
        nr.push_back(c + 1);
        if (R->has_val) nv.push_back(0.0);
      }
      for (int64_t k = R->ptr[c] - 1; k < R->ptr[c + 1] - 1; ++k) {
        nr.push_back(R->row[k]);
        if (R->has_val) nv.push_back(R->val[k]);
      }
      np[c + 1] = (int64_t)nr.size() + 1;
    }
    R->ptr.swap(np);
    R->row.swap(nr);
    R->val.swap(nv);
  }
  if (symmetric && lwr_upr_full == 2) {  // as flip_lwr_upr: counting sort over rows
    // w[i] plays wptr(i+1): after the prefix sum w[r] is where column r of
    // the transpose starts (1-based); placing an entry advances it
    const int64_t z = R->ptr[nn] - 1;
    std::vector<int64_t> w(nn + 2, 0);
    for (int64_t k = 0; k < z; ++k) w[R->row[k] + 1]++;
    w[0] = w[1] = 1;
    for (int r = 1; r <= nn; ++r) w[r + 1] += w[r];
    std::vector<int> col(z);
    std::vector<double> wv(R->has_val ? z : 0);
    for (int c = 1; c <= nn; ++c)
      for (int64_t k = R->ptr[c - 1] - 1; k < R->ptr[c] - 1; ++k) {
        const int r = R->row[k];
        col[w[r] - 1] = c;
        // skew: the upper triangle holds minus the stored values (sym_to_skew)
        if (R->has_val) wv[w[r] - 1] = (skew && r != c) ? -R->val[k] : R->val[k];
        w[r]++;
      }
    for (int j = 0; j <= nn; ++j) R->ptr[j] = w[j];
    R->row.swap(col);
    if (R->has_val) R->val.swap(wv);
  } else if (symmetric && lwr_upr_full == 3) {  // both triangles
    const int64_t z = R->ptr[nn] - 1;
    std::vector<int64_t> count(nn, 0);
    for (int c = 0; c < nn; ++c)
      for (int64_t k = R->ptr[c] - 1; k < R->ptr[c + 1] - 1; ++k) {
        count[c]++;
        if (R->row[k] != c + 1) count[R->row[k] - 1]++;
      }
    std::vector<int64_t> np(nn + 1, 1);
    for (int c = 0; c < nn; ++c) np[c + 1] = np[c] + count[c];
    std::vector<int64_t> pos(np.begin(), np.end() - 1);
    std::vector<int> nr(np[nn] - 1);
    std::vector<double> nv(R->has_val ? np[nn] - 1 : 0);
    for (int c = 0; c < nn; ++c)
      for (int64_t k = R->ptr[c] - 1; k < R->ptr[c + 1] - 1; ++k) {
        const int r = R->row[k] - 1;
        nr[pos[c] - 1] = r + 1;
        if (R->has_val) nv[pos[c] - 1] = R->val[k];
        pos[c]++;
        if (r != c) {
          nr[pos[r] - 1] = c + 1;
          if (R->has_val) nv[pos[r] - 1] = skew ? -R->val[k] : R->val[k];
          pos[r]++;
        }
      }
    (void)z;
    R->ptr.swap(np);
    R->row.swap(nr);
    R->val.swap(nv);
  }
  *m = R->m;
  *n = R->n;
  *nz = R->ptr[R->n] - 1;
  *has_val = R->has_val ? 1 : 0;
  *matrix_type = R->matrix_type;
  return R;
}

// Copies the matrix out of the handle (ptr n+1, row and val nz) and frees it.
void spral_rb_fetch(void *handle, int64_t *ptr, int *row, double *val) {
  auto *R = static_cast<RBMatrix *>(handle);
  if (!R) return;
  std::copy(R->ptr.begin(), R->ptr.end(), ptr);
  std::copy(R->row.begin(), R->row.end(), row);
  if (R->has_val && val) std::copy(R->val.begin(), R->val.end(), val);
  delete R;
}

}  // extern "C"
