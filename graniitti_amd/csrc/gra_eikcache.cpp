// Eikonal table cache in the reference's file format.
//
// MEikonal::S3Constructor (src/MEikonal.cc:46-132) keeps each table in a file
// whose name carries a djb2 hash (src/MAux.cc:280-287) of every free parameter
// the table depends on, and recomputes a table only when its file is missing or
// corrupt; IArray1D::WriteArray / ReadArray (src/MEikonal.cc:404-485) store one
// node per line as "x,Re,Im" with 15 significant digits. The same names and
// format are used here, so tables written by either program are read by the
// other.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gra_internal.h"

namespace gra_hip {

namespace {

// djb2: h = 33 h + c over the bytes, from 5381, in 64-bit unsigned arithmetic
uint64_t djb2(const std::string& s) {
  uint64_t h = 5381;
  for (const char c : s) h = (h << 5) + h + (uint64_t)(int64_t)c;
  return h;
}

// std::to_string of a double: printf "%f"
std::string fstr(double v) { return std::to_string(v); }
std::string istr(long long v) { return std::to_string(v); }

// PARAM_SOFT::GetHashString (src/MForm.cc:94-100). The Odderon is not part of
// this build's soft model: it is off, and its coupling enters only the key, with
// its TUNE0 value (modeldata/TUNE0/GENERAL.json).
std::string soft_key(const gra_cfg& c) {
  const double gN_O = 0.10;
  const int odderon_on = 0;
  return fstr(c.delta_P) + fstr(c.alpha_P) + fstr(c.gN_P) + fstr(gN_O) + fstr(c.fc1) +
         fstr(c.fc2) + fstr(c.fc3) + istr(odderon_on);
}

// MEikonalNumerics::GetHashString (include/Graniitti/MEikonal.h:59-67): both
// tables are linear in their variable (logKT2 = logBT = false)
std::string numerics_key(const gra_cfg& c) {
  const double min_kt2 = 1E-6, max_kt2 = 25.0, min_bt = 1E-6, max_bt = 10.0 / gev2fm();
  const double fb_min_kt = 1E-9, fb_max_kt = 30.0;
  return fstr(min_kt2) + fstr(max_kt2) + istr(c.n_kt2) + istr(0) + fstr(min_bt) + fstr(max_bt) +
         istr(c.n_bt) + istr(0) + fstr(fb_min_kt) + fstr(fb_max_kt) + istr(c.n_fb);
}

}  // namespace

std::string eikonal_cache_name(const gra_cfg& c, double s, int table) {
  // proton beams (the processes built here are pp)
  const std::string b1 = istr(2212), b2 = istr(2212);
  // the density key separates the two beam ids with '_', the amplitude key does not
  const std::string key =
      fstr(s) + b1 + (table == 0 ? "_" : "") + b2 + soft_key(c) + numerics_key(c);
  char rs[64];
  std::snprintf(rs, sizeof(rs), "%.0f", std::sqrt(s));
  return std::string(table == 0 ? "MBT_" : "MSA_") + b1 + "_" + b2 + "_" + rs + "_" +
         std::to_string((unsigned long long)djb2(key));
}

int eikonal_table_write(const char* path, const std::vector<double>& rows) {
  FILE* fp = std::fopen(path, "w");
  if (!fp) return set_global_error(GRA_E_IO, "eikonal table: cannot write %s", path);
  const size_t n = rows.size() / 3;
  for (size_t i = 0; i < n; ++i) {
    if (std::fprintf(fp, "%.15g,%.15g,%.15g\n", rows[3 * i], rows[3 * i + 1], rows[3 * i + 2]) <
        0) {
      std::fclose(fp);
      return set_global_error(GRA_E_IO, "eikonal table: write error in %s", path);
    }
  }
  if (std::fclose(fp) != 0) return set_global_error(GRA_E_IO, "eikonal table: write error in %s", path);
  return GRA_OK;
}

int eikonal_table_read(const char* path, int n_rows, std::vector<double>& rows) {
  FILE* fp = std::fopen(path, "r");
  if (!fp) return set_global_error(GRA_E_IO, "eikonal table: %s not found", path);
  rows.assign(3 * (size_t)n_rows, 0.0);
  std::string line;
  long fills = 0;
  int code = GRA_OK;
  for (int i = 0; i < n_rows && code == GRA_OK; ++i) {
    // one line (the rows after the first n_rows are not read)
    line.clear();
    int ch;
    while ((ch = std::fgetc(fp)) != EOF && ch != '\n') line.push_back((char)ch);
    if (line.empty() && ch == EOF) break;
    size_t pos = 0;
    int k = 0;
    while (pos <= line.size()) {
      size_t end = line.find(',', pos);
      if (end == std::string::npos) end = line.size();
      const std::string el = line.substr(pos, end - pos);
      if (end == line.size() && el.empty()) break;  // an empty last field is not a field
      {
        // a field that holds no number is an error, as std::stod's exception is
        const char* b = el.c_str();
        char* e = nullptr;
        errno = 0;
        const double v = std::strtod(b, &e);
        if (e == b) {
          code = set_global_error(GRA_E_ARG, "eikonal table %s: line %d: '%s' is not a number",
                                  path, i + 1, el.c_str());
          break;
        }
        if (k >= 3) {
          code = set_global_error(GRA_E_IO, "eikonal table %s: line %d has more than 3 fields",
                                  path, i + 1);
          break;
        }
        rows[3 * (size_t)i + k] = v;
        ++k;
        ++fills;
      }
      if (end == line.size()) break;
      pos = end + 1;
    }
  }
  std::fclose(fp);
  if (code != GRA_OK) return code;
  if (fills != 3L * n_rows)
    return set_global_error(GRA_E_IO, "eikonal table %s: %ld values, expected %d (corrupted)", path,
                            fills, 3 * n_rows);
  return GRA_OK;
}

}  // namespace gra_hip

using namespace gra_hip;

extern "C" int gra_eikonal_table_write(const char* path, const double* rows, int32_t n_rows) {
  if (!path || !rows || n_rows < 1)
    return set_global_error(GRA_E_ARG, "gra_eikonal_table_write: bad arguments");
  return eikonal_table_write(path, std::vector<double>(rows, rows + 3 * (size_t)n_rows));
}

extern "C" int gra_eikonal_table_read(const char* path, double* rows, int32_t n_rows) {
  if (!path || !rows || n_rows < 1)
    return set_global_error(GRA_E_ARG, "gra_eikonal_table_read: bad arguments");
  std::vector<double> v;
  const int r = eikonal_table_read(path, n_rows, v);
  if (r == GRA_OK) std::memcpy(rows, v.data(), sizeof(double) * v.size());
  return r;
}

extern "C" uint64_t gra_djb2(const char* s) { return s ? djb2(std::string(s)) : 0; }
