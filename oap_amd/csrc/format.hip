// format.hip -- host-side matrix formats: the text form that oap prints and
// parses, and the binary record / file framing of its save and load calls.
// No kernels; device matrices are copied to or from host memory once.
//
// Reference:
//   text out   oap::host::ToString -> matrixUtils::PrintMatrix / PrintArrays
//              (oapMatrixCpu/oapHostMatrixUtils.cpp:576-588,
//              oapMatrix/MatrixPrinter.cpp:26-80, oapMatrix/MatrixPrinter.h:28-112,
//              oapMatrix/MatrixUtils.h:104-176 for the default print arguments)
//   text in    oap::host::NewMatrix(text) -> matrixUtils::Parser / CreateArray
//              (oapMatrixCpu/oapHostMatrixUtils.cpp:223-274, oapMatrix/MatrixParser.cpp)
//   binary     oap::host::SaveMatrix / LoadMatrix / SaveMatrixInfo / LoadMatrixInfo
//              (oapMatrixCpu/oapHostMatrixUtils.cpp:1043-1104) into utils::ByteBuffer
//              (oapUtils/Buffer.h: push_back/read, fwrite/fread framing)
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "internal.h"

namespace oaph {
namespace {

// ---------------------------------------------------------------- text out
// One plane as PrintArrays prints it with the default arguments: runs of
// equal consecutive values as "v <repeats n times>", fixed notation with 9
// digits, a leading space before non-negative values when the plane has a
// negative one, ", " between values and "|\n " where a run ends at the end
// of a row (section length = columns).
void print_plane(std::string& out, const double* p, size_t n, uint32_t section) {
  std::vector<std::pair<size_t, double>> runs;
  size_t negatives = 0;
  for (size_t i = 0; i < n; ++i) {
    const double v = p[i];
    if (runs.empty() || v != runs.back().second)
      runs.push_back({1, v});
    else
      ++runs.back().first;
    if (v < 0.0) ++negatives;
  }
  out += "[";
  size_t done = 0;
  char num[512];
  for (const auto& run : runs) {
    const double v = run.second;
    if (v >= 0 && negatives > 0) out += " ";
    std::snprintf(num, sizeof(num), "%.9f", v);
    out += num;
    if (run.first > 1) {
      std::snprintf(num, sizeof(num), " <repeats %zu times>", run.first);
      out += num;
    }
    done += run.first;
    if (done != n) out += (done % section) != 0 ? ", " : "|\n ";
  }
  out += "]";
}

std::string to_string(uint32_t columns, uint32_t rows, const double* re, const double* im) {
  std::string s;
  const size_t n = (size_t)columns * rows;
  if (re) print_plane(s, re, n, columns);
  if (im) {
    if (re) s += " + ";
    s += "i * ";
    print_plane(s, im, n, columns);
  }
  return s;
}

oaph_status copy_out(const std::string& s, char* buf, size_t cap, size_t* len) {
  if (len) *len = s.size();
  if (!buf) return OAPH_OK;
  OAPH_CHECK(cap > s.size(), OAPH_ERR_ARG);  // room for the terminating NUL
  std::memcpy(buf, s.data(), s.size());
  buf[s.size()] = '\0';
  return OAPH_OK;
}

// ---------------------------------------------------------------- text in
// Parser::getValue: the integer after "<id>" and one separator character.
bool get_value(const std::string& text, const char* id, uint32_t* value) {
  size_t pos = text.find(id, 0);
  if (pos == std::string::npos) return false;
  pos += std::strlen(id) + 1;
  if (pos >= text.size()) return false;
  const size_t pos1 = text.find_first_not_of("-0123456789", pos);
  *value = (uint32_t)std::atoi(text.substr(pos, pos1 == std::string::npos ? std::string::npos
                                                                          : pos1 - pos).c_str());
  return true;
}

// Parser::satof: only ".-0123456789e"; a leading '-' negates the rest.
bool satof(const std::string& str, double* out) {
  if (str.find_first_not_of(".-0123456789e") != std::string::npos) return false;
  double sign = 1;
  size_t index = 0;
  if (!str.empty() && str[0] == '-') {
    sign = -1;
    index = 1;
  }
  *out = sign * std::atof(str.c_str() + index);
  return true;
}

// Parser::parseFloatsElement: "value<...count...>" (whitespace removed).
bool parse_repeat(const std::string& e, std::vector<double>& arr) {
  const size_t pos = e.find('<'), pos1 = e.find('>');
  if (pos1 != e.size() - 1 || pos == std::string::npos || pos > pos1) return false;
  const std::string part = e.substr(pos, pos1 - pos);
  const size_t d1 = part.find_first_of(".-0123456789");
  if (d1 == std::string::npos) return false;
  const size_t d2 = part.find_first_not_of(".-0123456789", d1);
  const int count = std::atoi(part.substr(d1, d2 == std::string::npos ? std::string::npos
                                                                      : d2 - d1).c_str());
  double v;
  if (!satof(e.substr(0, pos), &v)) return false;
  for (int i = 0; i < count; ++i) arr.push_back(v);
  return true;
}

// CreateArray(text, which): the which-th "[...]" (1-based), elements split at
// ',' and '|', whitespace removed.
bool create_array(const std::string& text, uint32_t which, std::vector<double>& arr) {
  if (which == 0) return false;
  size_t pos = text.find('[', 0);
  for (uint32_t fa = 0; fa + 1 < which; ++fa) pos = text.find('[', pos + 1);
  if (pos == std::string::npos) return false;
  const size_t pos1 = text.find(']', pos);
  if (pos1 == std::string::npos) return false;
  const std::string body = text.substr(pos + 1, pos1 - pos - 1);
  arr.clear();
  size_t p = 0, q;
  do {
    q = body.find_first_of(",|", p);
    std::string e = body.substr(p, q == std::string::npos ? std::string::npos : q - p);
    std::string t;
    for (char ch : e)
      if (!std::isspace((unsigned char)ch)) t += ch;
    const bool one = t.find('<') == std::string::npos && t.find('>') == std::string::npos;
    if (one) {
      double v;
      if (!satof(t, &v)) return false;
      arr.push_back(v);
    } else if (!parse_repeat(t, arr)) {
      return false;
    }
    p = q + 1;
  } while (q != std::string::npos);
  return true;
}

struct Parsed {
  uint32_t columns = 0, rows = 0;
  bool has_re = false, has_im = false;
  std::vector<double> re, im;
};

// NewMatrix(text): "columns" / "rows" when both are given, else a square
// matrix of side floor(sqrt(length of the first array)); first array re,
// second im.
oaph_status parse_matrix(const char* ctext, Parsed* P) {
  OAPH_CHECK(ctext, OAPH_ERR_ARG);
  const std::string text(ctext);
  uint32_t c = 0, r = 0;
  const bool hc = get_value(text, "columns", &c);
  const bool hr = get_value(text, "rows", &r);
  P->has_re = create_array(text, 1, P->re);
  P->has_im = create_array(text, 2, P->im);
  if (!P->has_re) P->re.clear();
  if (!P->has_im) P->im.clear();
  if (!(hc && hr)) {
    const uint32_t sq = (uint32_t)std::sqrt((double)P->re.size());
    c = r = sq;
  }
  OAPH_CHECK(P->has_re || P->has_im, OAPH_ERR_ARG);
  const size_t n = (size_t)c * r;
  // the reference adopts the arrays as they are: values past columns x rows
  // are never read (a trailing ',' adds a 0 that way), an array shorter than
  // columns x rows would be read past its end -- rejected here
  if (P->has_re) {
    OAPH_CHECK(P->re.size() >= n, OAPH_ERR_SHAPE);
    P->re.resize(n);
  }
  if (P->has_im) {
    OAPH_CHECK(P->im.size() >= n, OAPH_ERR_SHAPE);
    P->im.resize(n);
  }
  P->columns = c;
  P->rows = r;
  return OAPH_OK;
}

// ---------------------------------------------------------------- binary
// SaveMatrix record: bool isMatrix; then bool isRe, bool isIm, uint32
// columns, uint32 rows (SaveMatrixInfo); then the re plane and the im plane
// as doubles, little endian, unaligned.
constexpr size_t kInfoBytes = 1 + 1 + 4 + 4;

size_t record_bytes(bool present, bool isre, bool isim, size_t n) {
  if (!present) return 1;
  return 1 + kInfoBytes + ((isre ? 1 : 0) + (isim ? 1 : 0)) * n * sizeof(double);
}

void put_record(uint8_t* b, bool isre, bool isim, uint32_t c, uint32_t r, const double* re,
                const double* im) {
  size_t o = 0;
  b[o++] = 1;
  b[o++] = isre ? 1 : 0;
  b[o++] = isim ? 1 : 0;
  std::memcpy(b + o, &c, 4);
  o += 4;
  std::memcpy(b + o, &r, 4);
  o += 4;
  const size_t n = (size_t)c * r * sizeof(double);
  if (isre) {
    std::memcpy(b + o, re, n);
    o += n;
  }
  if (isim) std::memcpy(b + o, im, n);
}

oaph_status get_info(const uint8_t* b, size_t len, int* present, oaph_matrix_info* info,
                     size_t* data_off) {
  OAPH_CHECK(b && len >= 1, OAPH_ERR_ARG);
  *present = b[0] != 0;
  if (!*present) {
    *data_off = 1;
    return OAPH_OK;
  }
  OAPH_CHECK(len >= 1 + kInfoBytes, OAPH_ERR_ARG);
  uint32_t c, r;
  std::memcpy(&c, b + 3, 4);
  std::memcpy(&r, b + 7, 4);
  info->isRe = b[1] != 0;
  info->isIm = b[2] != 0;
  info->columns = c;
  info->rows = r;
  *data_off = 1 + kInfoBytes;
  // a header of 2^32 or more elements is rejected before the record size is
  // formed (the same limit as oaph_matrix_new), so `need` cannot wrap
  OAPH_CHECK((uint64_t)c * r < (1ull << 32), OAPH_ERR_ARG);
  const size_t need = record_bytes(true, info->isRe, info->isIm, (size_t)c * r);
  OAPH_CHECK(len >= need, OAPH_ERR_ARG);
  return OAPH_OK;
}

}  // namespace
}  // namespace oaph

using namespace oaph;

extern "C" {

oaph_status oaph_host_matrix_to_string(uint32_t columns, uint32_t rows, const double* re,
                                       const double* im, char* buf, size_t cap, size_t* len) {
  OAPH_CHECK(re || im, OAPH_ERR_PLANE);
  return copy_out(to_string(columns, rows, re, im), buf, cap, len);
}

oaph_status oaph_matrix_to_string(const oaph_matrix* m, char* buf, size_t cap, size_t* len) {
  if (!m) return copy_out("nullptr", buf, cap, len);
  OAPH_CHECK(m->reValues || m->imValues, OAPH_ERR_PLANE);
  const size_t n = (size_t)m->columns * m->rows;
  std::vector<double> re(m->reValues ? n : 0), im(m->imValues ? n : 0);
  OAPH_TRY(oaph_matrix_copy_d2h(m->reValues ? re.data() : nullptr,
                                m->imValues ? im.data() : nullptr, m));
  return copy_out(to_string(m->columns, m->rows, m->reValues ? re.data() : nullptr,
                            m->imValues ? im.data() : nullptr),
                  buf, cap, len);
}

oaph_status oaph_host_parse_array(const char* text, uint32_t which, double* out, size_t cap,
                                  size_t* len) {
  OAPH_CHECK(text && len, OAPH_ERR_ARG);
  std::vector<double> arr;
  OAPH_CHECK(create_array(std::string(text), which, arr), OAPH_ERR_ARG);
  *len = arr.size();
  if (out) {
    OAPH_CHECK(cap >= arr.size(), OAPH_ERR_ARG);
    if (!arr.empty()) std::memcpy(out, arr.data(), arr.size() * sizeof(double));
  }
  return OAPH_OK;
}

oaph_status oaph_host_parse_matrix(const char* text, oaph_matrix_info* info, double* re,
                                   double* im, size_t cap) {
  OAPH_CHECK(info, OAPH_ERR_ARG);
  Parsed P;
  OAPH_TRY(parse_matrix(text, &P));
  info->isRe = P.has_re;
  info->isIm = P.has_im;
  info->columns = P.columns;
  info->rows = P.rows;
  const size_t n = (size_t)P.columns * P.rows;
  if (re && P.has_re) {
    OAPH_CHECK(cap >= n, OAPH_ERR_ARG);
    if (n) std::memcpy(re, P.re.data(), n * sizeof(double));
  }
  if (im && P.has_im) {
    OAPH_CHECK(cap >= n, OAPH_ERR_ARG);
    if (n) std::memcpy(im, P.im.data(), n * sizeof(double));
  }
  return OAPH_OK;
}

oaph_status oaph_matrix_parse(const char* text, oaph_matrix** out) {
  OAPH_CHECK(out, OAPH_ERR_ARG);
  *out = nullptr;
  Parsed P;
  OAPH_TRY(parse_matrix(text, &P));
  OAPH_CHECK(P.columns > 0 && P.rows > 0, OAPH_ERR_SHAPE);
  oaph_matrix* m = nullptr;
  OAPH_TRY(oaph_matrix_new(&m, P.has_re, P.has_im, P.columns, P.rows));
  const oaph_status s = oaph_matrix_copy_h2d(m, P.has_re ? P.re.data() : nullptr,
                                             P.has_im ? P.im.data() : nullptr);
  if (s != OAPH_OK) {
    oaph_matrix_free(m);
    return s;
  }
  *out = m;
  return OAPH_OK;
}

oaph_status oaph_host_matrix_save(int present, uint32_t columns, uint32_t rows, const double* re,
                                  const double* im, uint8_t* buf, size_t cap, size_t* len) {
  OAPH_CHECK(len, OAPH_ERR_ARG);
  const size_t need = record_bytes(present != 0, re != nullptr, im != nullptr,
                                   (size_t)columns * rows);
  *len = need;
  if (!buf) return OAPH_OK;
  OAPH_CHECK(cap >= need, OAPH_ERR_ARG);
  if (!present) {
    buf[0] = 0;
    return OAPH_OK;
  }
  put_record(buf, re != nullptr, im != nullptr, columns, rows, re, im);
  return OAPH_OK;
}

oaph_status oaph_matrix_save(const oaph_matrix* m, uint8_t* buf, size_t cap, size_t* len) {
  OAPH_CHECK(len, OAPH_ERR_ARG);
  if (!m) return oaph_host_matrix_save(0, 0, 0, nullptr, nullptr, buf, cap, len);
  const size_t n = (size_t)m->columns * m->rows;
  const size_t need = record_bytes(true, m->reValues, m->imValues, n);
  *len = need;
  if (!buf) return OAPH_OK;
  OAPH_CHECK(cap >= need, OAPH_ERR_ARG);
  std::vector<double> re(m->reValues ? n : 0), im(m->imValues ? n : 0);
  OAPH_TRY(oaph_matrix_copy_d2h(m->reValues ? re.data() : nullptr,
                                m->imValues ? im.data() : nullptr, m));
  put_record(buf, m->reValues, m->imValues, m->columns, m->rows,
             m->reValues ? re.data() : nullptr, m->imValues ? im.data() : nullptr);
  return OAPH_OK;
}

oaph_status oaph_host_matrix_load(const uint8_t* buf, size_t len, size_t* consumed, int* present,
                                  oaph_matrix_info* info, double* re, double* im, size_t cap) {
  OAPH_CHECK(present && info, OAPH_ERR_ARG);
  size_t off = 0;
  oaph_matrix_info inf{0, 0, 0, 0};
  OAPH_TRY(get_info(buf, len, present, &inf, &off));
  *info = inf;
  if (!*present) {
    if (consumed) *consumed = 1;
    return OAPH_OK;
  }
  const size_t n = (size_t)inf.columns * inf.rows;
  if (consumed) *consumed = record_bytes(true, inf.isRe, inf.isIm, n);
  if (inf.isRe) {
    if (re) {
      OAPH_CHECK(cap >= n, OAPH_ERR_ARG);
      std::memcpy(re, buf + off, n * sizeof(double));
    }
    off += n * sizeof(double);
  }
  if (inf.isIm && im) {
    OAPH_CHECK(cap >= n, OAPH_ERR_ARG);
    std::memcpy(im, buf + off, n * sizeof(double));
  }
  return OAPH_OK;
}

oaph_status oaph_matrix_load(oaph_matrix** out, const uint8_t* buf, size_t len, size_t* consumed) {
  OAPH_CHECK(out, OAPH_ERR_ARG);
  *out = nullptr;
  int present = 0;
  oaph_matrix_info inf{0, 0, 0, 0};
  size_t off = 0;
  OAPH_TRY(get_info(buf, len, &present, &inf, &off));
  if (!present) {
    if (consumed) *consumed = 1;
    return OAPH_OK;
  }
  OAPH_CHECK(inf.isRe || inf.isIm, OAPH_ERR_PLANE);
  OAPH_CHECK(inf.columns > 0 && inf.rows > 0, OAPH_ERR_SHAPE);
  const size_t n = (size_t)inf.columns * inf.rows;
  std::vector<double> re(inf.isRe ? n : 0), im(inf.isIm ? n : 0);
  if (inf.isRe) {
    std::memcpy(re.data(), buf + off, n * sizeof(double));
    off += n * sizeof(double);
  }
  if (inf.isIm) std::memcpy(im.data(), buf + off, n * sizeof(double));
  oaph_matrix* m = nullptr;
  OAPH_TRY(oaph_matrix_new(&m, inf.isRe, inf.isIm, inf.columns, inf.rows));
  const oaph_status s =
      oaph_matrix_copy_h2d(m, inf.isRe ? re.data() : nullptr, inf.isIm ? im.data() : nullptr);
  if (s != OAPH_OK) {
    oaph_matrix_free(m);
    return s;
  }
  if (consumed) *consumed = record_bytes(true, inf.isRe, inf.isIm, n);
  *out = m;
  return OAPH_OK;
}

// utils::ByteBuffer::fwrite / fread framing: size_t element size (1), size_t
// length, then the bytes.
oaph_status oaph_buffer_write_file(const char* path, const uint8_t* buf, size_t len) {
  OAPH_CHECK(path && (buf || len == 0), OAPH_ERR_ARG);
  FILE* f = std::fopen(path, "wb");
  OAPH_CHECK(f, OAPH_ERR_ARG);
  const size_t one = 1;
  bool ok = std::fwrite(&one, sizeof(size_t), 1, f) == 1 &&
            std::fwrite(&len, sizeof(size_t), 1, f) == 1 &&
            (len == 0 || std::fwrite(buf, 1, len, f) == len);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? OAPH_OK : OAPH_ERR_ARG;
}

oaph_status oaph_buffer_read_file(const char* path, uint8_t* buf, size_t cap, size_t* len) {
  OAPH_CHECK(path && len, OAPH_ERR_ARG);
  FILE* f = std::fopen(path, "rb");
  OAPH_CHECK(f, OAPH_ERR_ARG);
  size_t tsize = 0, n = 0;
  bool ok = std::fread(&tsize, sizeof(size_t), 1, f) == 1 && tsize == 1 &&
            std::fread(&n, sizeof(size_t), 1, f) == 1;
  if (ok) {
    *len = n;
    if (buf) ok = cap >= n && (n == 0 || std::fread(buf, 1, n, f) == n);
  }
  std::fclose(f);
  return ok ? OAPH_OK : OAPH_ERR_ARG;
}

}  // extern "C"
