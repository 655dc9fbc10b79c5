// CHARMM PSF/COR readers with the reference driver's fixed columns
// (utils/io.cpp:16-117) and the recentring of utils/utils.cpp:35-58.
#include "glst_io.h"

#include <algorithm>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>

namespace {

bool read_all(const char *path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// Splits on '\n' (a trailing '\r' is kept, as the fixed-column reads ignore
// everything past the last field).
class Lines {
 public:
  explicit Lines(const std::string &s) : s_(s) {}
  bool next(std::string &line) {
    if (pos_ >= s_.size()) return false;
    size_t e = s_.find('\n', pos_);
    if (e == std::string::npos) e = s_.size();
    line = s_.substr(pos_, e - pos_);
    pos_ = e + 1;
    return true;
  }

 private:
  const std::string &s_;
  size_t pos_ = 0;
};

// Field [b, b+n) of a line; throws when the line is too short or the field
// does not parse as a number.
double field_d(const std::string &line, size_t b, size_t n) {
  if (line.size() <= b) throw std::runtime_error("line too short: " + line);
  return std::stod(line.substr(b, n));
}

uint64_t field_u(const std::string &line, size_t b, size_t n) {
  if (line.size() <= b) throw std::runtime_error("line too short: " + line);
  return std::stoull(line.substr(b, n));
}

bool is_natom_line(const std::string &line) {
  return line.size() >= 17 && line.compare(11, 6, "!NATOM") == 0;
}

template <class F>
glst_status_t guard(F &&f) {
  try {
    return f();
  } catch (const std::exception &) {
    return GLST_ERR_IO;
  }
}

}  // namespace

extern "C" {

glst_status_t glst_psf_natom(const char *path, uint64_t *natom) {
  if (!path || !natom) return GLST_ERR_INVALID;
  return guard([&] {
    std::string data, line;
    if (!read_all(path, data)) return GLST_ERR_IO;
    *natom = 0;
    Lines ls(data);
    while (ls.next(line)) {
      if (is_natom_line(line)) {
        *natom = field_u(line, 0, 11);
        break;
      }
    }
    return GLST_OK;
  });
}

glst_status_t glst_read_psf_charges(const char *path, uint64_t natom,
                                    double *qc) {
  if (!path || (natom > 0 && !qc)) return GLST_ERR_INVALID;
  return guard([&] {
    std::string data, line;
    if (!read_all(path, data)) return GLST_ERR_IO;
    Lines ls(data);
    uint64_t count = 0;
    bool found = false;
    while (ls.next(line))
      if (is_natom_line(line)) {
        count = field_u(line, 0, 10);
        found = true;
        break;
      }
    if (!found || count != natom) return GLST_ERR_IO;
    for (uint64_t i = 0; i < natom; ++i) {
      if (!ls.next(line)) return GLST_ERR_IO;
      qc[i] = field_d(line, 50, 14);
    }
    return GLST_OK;
  });
}

glst_status_t glst_read_cor(const char *path, uint64_t natom, double *rx,
                            double *ry, double *rz) {
  if (!path || (natom > 0 && (!rx || !ry || !rz))) return GLST_ERR_INVALID;
  return guard([&] {
    std::string data, line;
    if (!read_all(path, data)) return GLST_ERR_IO;
    Lines ls(data);
    uint64_t count = 0;
    bool found = false;
    while (ls.next(line))
      if (!line.empty() && line[0] != '*') {
        count = field_u(line, 0, 10);
        found = true;
        break;
      }
    if (!found || count != natom) return GLST_ERR_IO;
    for (uint64_t i = 0; i < natom; ++i) {
      if (!ls.next(line)) return GLST_ERR_IO;
      rx[i] = field_d(line, 40, 20);
      ry[i] = field_d(line, 60, 20);
      rz[i] = field_d(line, 80, 20);
    }
    return GLST_OK;
  });
}

glst_status_t glst_recenter(uint64_t natom, double *rx, double *ry,
                            double *rz) {
  if (natom > 0 && (!rx || !ry || !rz)) return GLST_ERR_INVALID;
  double mn[3] = {std::numeric_limits<double>::max(),
                  std::numeric_limits<double>::max(),
                  std::numeric_limits<double>::max()};
  for (uint64_t i = 0; i < natom; ++i) {
    mn[0] = std::min(mn[0], rx[i]);
    mn[1] = std::min(mn[1], ry[i]);
    mn[2] = std::min(mn[2], rz[i]);
  }
  for (uint64_t i = 0; i < natom; ++i) {
    rx[i] -= mn[0];
    ry[i] -= mn[1];
    rz[i] -= mn[2];
  }
  return GLST_OK;
}

}  // extern "C"
