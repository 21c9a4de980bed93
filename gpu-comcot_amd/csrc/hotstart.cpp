// The reference's hot-start snapshots (HOT_START / HOTSTART_DATA,
// hotstart.f90:3-166), read into and written from a layer tree's device
// state.
//
// Per layer and start step the reference keeps five text files in its
// working directory, named with FORMAT('z1_',I2.2,'_',I6.6,'.dat') and the
// like (hotstart.f90:53-68):
//   z1_LL_SSSSSS.dat   Z(:,:,1)        m1_...  M(:,:,1)     n1_...  N(:,:,1)
//   zmax1_...          Z_MAX           zmin1_...  Z_MIN
// Each holds the cells I = IS..IE of rows J = JS..JE, one WRITE per row
// with FORMAT '(10F12.5)': ten 12-character fields per line, the row's last
// line shorter (hotstart.f90:73-98).  IS = JS = 1 on the outermost layer
// and 2 on nested ones, whose first row and column (UPDATE's copies) are
// not kept.  A snapshot at start step S holds the state after S - 1 root
// steps: the main loop calls HOT_START at K = S before the step's MASS
// (comcot.f90:322-333), and a resumed run starts its loop at K = S.
//
// Z_MAX is what CUDA_GETZMAX copies out of the GPU half (output.f90:918):
// the maximum without a wet test that the device keeps from the initial
// elevation on.  Z_MIN is never updated by the reference (its updates are
// commented out, output.f90:881-898) and stays 0 from ALLOC
// (comcot.f90:632); the device keeps no minimum, so zmin1 files are
// written as zeros and read only to check their form.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "hotstart.h"
#include "tree.h"

namespace gc {

namespace {

constexpr int kField = 12, kPerLine = 10;

// One F12.5 field as the Fortran runtime writes it: the exact binary value
// rounded to 5 decimals (to nearest, ties to even, as C's %f rounds),
// right-justified; twelve asterisks when the result needs more than twelve
// characters; NaN and infinities as NaN, Inf, -Inf.
void put_field(float v, char *out) {
  char buf[64];
  int n;
  if (std::isnan(v)) n = std::snprintf(buf, sizeof buf, "NaN");
  else if (std::isinf(v)) n = std::snprintf(buf, sizeof buf, v < 0 ? "-Inf" : "Inf");
  else n = std::snprintf(buf, sizeof buf, "%.5f", (double)v);
  if (n < 0 || n > kField) {
    std::memset(out, '*', kField);
    return;
  }
  std::memset(out, ' ', (size_t)(kField - n));
  std::memcpy(out + (kField - n), buf, (size_t)n);
}

bool ieq(const char *a, const char *b) {  // case-insensitive equality
  for (; *a && *b; ++a, ++b)
    if ((*a | 0x20) != (*b | 0x20)) return false;
  return *a == *b;
}

// One F12.5 field on input (Fortran F editing with the default BLANK =
// 'NULL'): blanks are ignored and an all-blank field is 0; digits without
// a decimal point carry 5 implied decimals; an exponent may follow as
// E/D with or without a sign, or as a bare sign; NaN and Inf[inity] are
// read as such.  Returns false for anything else (asterisks included).
bool get_field(const char *f, float *out) {
  char s[kField + 1];
  int k = 0;
  for (int c = 0; c < kField; ++c)
    if (f[c] != ' ') s[k++] = f[c];
  s[k] = 0;
  if (k == 0) {
    *out = 0.0f;
    return true;
  }
  const char *p = s;
  bool neg = false;
  if (*p == '+' || *p == '-') neg = *p++ == '-';
  if (ieq(p, "nan")) {
    *out = neg ? -NAN : NAN;
    return true;
  }
  if (ieq(p, "inf") || ieq(p, "infinity")) {
    *out = neg ? -INFINITY : INFINITY;
    return true;
  }
  std::string mant;
  bool point = false, digits = false;
  for (; *p; ++p) {
    if (*p >= '0' && *p <= '9') {
      mant += *p;
      digits = true;
    } else if (*p == '.' && !point) {
      mant += '.';
      point = true;
    } else {
      break;
    }
  }
  if (!digits) return false;
  long exp10 = 0;
  if (*p) {
    if (*p == 'e' || *p == 'E' || *p == 'd' || *p == 'D') ++p;
    bool eneg = false;
    if (*p == '+' || *p == '-') eneg = *p++ == '-';
    if (!*p) return false;
    for (; *p; ++p) {
      if (*p < '0' || *p > '9' || exp10 > 100000) return false;
      exp10 = 10 * exp10 + (*p - '0');
    }
    if (eneg) exp10 = -exp10;
  }
  if (!point) exp10 -= 5;  // the d = 5 implied decimals of F12.5
  char buf[64];
  std::snprintf(buf, sizeof buf, "e%ld", exp10);
  const std::string text = (neg ? "-" : "") + mant + buf;
  char *end = nullptr;
  *out = std::strtof(text.c_str(), &end);  // correctly rounded, as the Fortran runtime
  return end && *end == 0;
}

std::string step_field(int step) {  // I6.6: six digits, or six asterisks
  char b[16];
  if (step < 0 || step > 999999) return "******";
  std::snprintf(b, sizeof b, "%06d", step);
  return b;
}

struct Snap {
  const char *prefix;
};
constexpr Snap kSnaps[kHotstartFiles] = {{"z1_"}, {"m1_"}, {"n1_"}, {"zmax1_"}, {"zmin1_"}};

}  // namespace

std::string hotstart_name(int kind, int layer_id, int start_step) {
  char b[16];
  std::snprintf(b, sizeof b, "%02d", layer_id);
  return std::string(kSnaps[kind].prefix) + b + "_" + step_field(start_step) + ".dat";
}

int hotstart_write_plane(const char *path, const float *plane, int nx, int ny, int first) {
  if (nx < first || ny < first || first < 1 || first > 2)
    return fail(90, "hot start: plane %dx%d, first cell %d", nx, ny, first);
  const int n = nx - first + 1, rows = ny - first + 1;
  const int lines = (n + kPerLine - 1) / kPerLine;
  const size_t row_bytes = (size_t)n * kField + (size_t)lines;
  std::vector<char> text((size_t)rows * row_bytes);
  // rows are formatted on a few threads (a 4096 x 3072 plane is 150 MB)
  const int nth = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  auto work = [&](int t) {
    for (int r = t; r < rows; r += nth) {
      const float *row = plane + (size_t)(first - 1 + r) * nx + (first - 1);
      char *o = text.data() + (size_t)r * row_bytes;
      for (int i = 0; i < n; ++i) {
        put_field(row[i], o);
        o += kField;
        if (i % kPerLine == kPerLine - 1 || i == n - 1) *o++ = '\n';
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nth; ++t) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(91, "hot start: cannot create %s", path);
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  if (std::fclose(f) != 0 || !ok) return fail(91, "hot start: write to %s failed", path);
  return 0;
}

int hotstart_read_plane(const char *path, float *plane, int nx, int ny, int first) {
  if (nx < first || ny < first || first < 1 || first > 2)
    return fail(90, "hot start: plane %dx%d, first cell %d", nx, ny, first);
  FILE *f = std::fopen(path, "rb");
  if (!f) return kHotstartMissing;
  std::vector<char> data;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) data.insert(data.end(), chunk, chunk + got);
  std::fclose(f);
  const int n = nx - first + 1, rows = ny - first + 1;
  size_t pos = 0;
  char line[kPerLine * kField];
  std::vector<float> row((size_t)n);
  for (int r = 0; r < rows; ++r) {
    for (int i0 = 0; i0 < n; i0 += kPerLine) {
      // one record, padded with blanks to the ten fields (PAD = 'YES')
      if (pos >= data.size())
        return fail(92, "hot start: %s ends at row %d of %d", path, r + first, ny);
      size_t e = pos;
      while (e < data.size() && data[e] != '\n') ++e;
      size_t len = e - pos;
      if (len && data[pos + len - 1] == '\r') --len;
      std::memset(line, ' ', sizeof line);
      std::memcpy(line, data.data() + pos, std::min(len, sizeof line));
      pos = e + 1;
      const int cnt = std::min(kPerLine, n - i0);
      for (int c = 0; c < cnt; ++c)
        if (!get_field(line + c * kField, &row[(size_t)(i0 + c)]))
          return fail(93, "hot start: %s row %d: field %d '%.12s' is not a number", path,
                      r + first, i0 + c + first, line + c * kField);
    }
    std::memcpy(plane + (size_t)(first - 1 + r) * nx + (first - 1), row.data(),
                (size_t)n * sizeof(float));
  }
  return 0;
}

namespace {

// device plane of a snapshot kind, or null for Z_MIN (not kept)
float *snap_plane(LayerRec &L, int kind) {
  switch (kind) {
    case 0: return L.z[L.cur];
    case 1: return L.m[L.cur];
    case 2: return L.n[L.cur];
    case 3: return L.zmax_raw;
    default: return nullptr;
  }
}

std::string join(const std::string &dir, const std::string &name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

}  // namespace

int tree_hotstart_write(Tree &T, hipStream_t st, int start_step, const std::string &dir) {
  HIPCHK(hipStreamSynchronize(st));
  std::vector<float> h;
  for (LayerRec &L : T.layers) {
    if (!L.used) continue;
    const int first = L.id == 1 ? 1 : 2;
    h.assign(L.nxy(), 0.0f);
    for (int kind = 0; kind < kHotstartFiles; ++kind) {
      if (float *d = snap_plane(L, kind))
        HIPCHK(hipMemcpy(h.data(), d, L.nxy() * sizeof(float), hipMemcpyDeviceToHost));
      else
        std::fill(h.begin(), h.end(), 0.0f);
      const std::string p = join(dir, hotstart_name(kind, L.id, start_step));
      if (int e = hotstart_write_plane(p.c_str(), h.data(), L.nx, L.ny, first)) return e;
    }
  }
  return 0;
}

int tree_hotstart_read(Tree &T, hipStream_t st, int start_type, int start_step,
                       const std::string &dir, int *missing) {
  T.mass_ahead = false;  // level 2 is no longer the next continuity
  if (start_type != 1 && start_type != 2)
    return fail(94, "hot start: START_TYPE %d is not a load (1 or 2)", start_type);
  HIPCHK(hipStreamSynchronize(st));
  // every file is read and checked before any device plane changes
  struct Load {
    LayerRec *L;
    int kind;
    std::vector<float> v;
  };
  std::vector<Load> loads;
  int miss = 0;
  for (LayerRec &L : T.layers) {
    if (!L.used || (start_type == 1 && L.id != 1)) continue;
    const int first = L.id == 1 ? 1 : 2;
    for (int kind = 0; kind < kHotstartFiles; ++kind) {
      Load ld{&L, kind, std::vector<float>(L.nxy())};
      float *d = snap_plane(L, kind);
      if (d) HIPCHK(hipMemcpy(ld.v.data(), d, L.nxy() * sizeof(float), hipMemcpyDeviceToHost));
      const std::string p = join(dir, hotstart_name(kind, L.id, start_step));
      const int e = hotstart_read_plane(p.c_str(), ld.v.data(), L.nx, L.ny, first);
      if (e == kHotstartMissing) {
        // HOTSTART_DATA: a file it cannot open leaves the whole plane 0
        std::fprintf(stderr, "WARNING:: HOTSTART CAN'T OPEN %s; USE ZEROS INSTEAD.\n",
                     p.c_str());
        std::fill(ld.v.begin(), ld.v.end(), 0.0f);
        ++miss;
      } else if (e) {
        return e;
      }
      if (d) loads.push_back(std::move(ld));
    }
  }
  for (Load &ld : loads)
    HIPCHK(hipMemcpy(snap_plane(*ld.L, ld.kind), ld.v.data(), ld.v.size() * sizeof(float),
                     hipMemcpyHostToDevice));
  if (missing) *missing = miss;
  return 0;
}

}  // namespace gc
