// QES-Winds workspace files: the arrays of the mass-consistency step in the
// NetCDF classic format (64-bit offset variant, "CDF-2").
//
// Layout of the reference's workspace writer (WINDSOutputWorkspace.cpp:40-150,
// QESNetCDFOutput.cpp:40-60): an unlimited time dimension t, cell-centre
// dimensions x, y, z and face dimensions x_face, y_face, z_face; the face
// fields u, v, w, u0, v0, w0 on (t, z_face, y_face, x_face), the cell fields
// icellflag and e..n on (t, z, y, x), the coordinates and dz_array, each with
// "units" and "long_name" attributes.  The fields that the reference writes
// but the solver path does not carry (terrain, roughness, mixing length,
// volume fractions, UTM origin, the timestamp string) are left out.  The
// reference writes NetCDF-4 (HDF5) files; the netCDF library reads classic
// files the same way, and `nccopy -k 64-bit-offset` turns a NetCDF-4
// workspace into one this reader takes.
//
// Reading follows Domain(const std::string&) (Domain.cpp:59-109): sizes from
// x_face, y_face, z_face; dx = x[1] - x[0], dy = y[1] - y[0]; dz_array if the
// file has it, else uniform z[1] - z[0].  Coefficients and flags are read as
// WINDSGeneralData does with readCoefficientsFlag (WINDSGeneralData.cpp:
// 880-911), one record of (t, z, y, x) arrays, here the whole arrays.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "qesmc.h"
#include "qesmc_city_internal.h"

namespace {

enum : int { kByte = 1, kChar = 2, kShort = 3, kInt = 4, kFloat = 5, kDouble = 6 };
constexpr uint32_t kTagDim = 0x0A, kTagVar = 0x0B, kTagAtt = 0x0C;

int type_size(int t)
{
  switch (t) {
  case kByte:
  case kChar: return 1;
  case kShort: return 2;
  case kInt:
  case kFloat: return 4;
  case kDouble: return 8;
  default: return 0;
  }
}

uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

struct File {
  FILE *f = nullptr;
  ~File()
  {
    if (f) std::fclose(f);
  }
};

// ---------------------------------------------------------------- writing --

struct Header {
  std::vector<unsigned char> b;
  void u32(uint32_t v)
  {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void u64(uint64_t v)
  {
    u32((uint32_t)(v >> 32));
    u32((uint32_t)v);
  }
  void pad()
  {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s)
  {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
};

struct WDim {
  std::string name;
  uint32_t len;  // 0: the record (unlimited) dimension
};

struct WVar {
  std::string name, units, long_name;
  int type;
  std::vector<int> dims;  // indices into the dimension list
  const void *data;       // host values, native byte order
  uint64_t count = 0;     // values (of one record for a record variable)
  bool record = false;
  uint64_t vsize = 0, begin = 0;
};

void put_att_text(Header &h, const std::string &name, const std::string &value)
{
  h.name(name);
  h.u32(kChar);
  h.u32((uint32_t)value.size());
  h.b.insert(h.b.end(), value.begin(), value.end());
  h.pad();
}

void build_header(Header &h, const std::vector<WDim> &dims, const std::vector<WVar> &vars, uint32_t numrecs)
{
  h.b.clear();
  h.b.insert(h.b.end(), {'C', 'D', 'F', 2});
  h.u32(numrecs);
  h.u32(kTagDim);
  h.u32((uint32_t)dims.size());
  for (const WDim &d : dims) {
    h.name(d.name);
    h.u32(d.len);
  }
  h.u64(0);  // no global attributes (ABSENT)
  h.u32(kTagVar);
  h.u32((uint32_t)vars.size());
  for (const WVar &v : vars) {
    h.name(v.name);
    h.u32((uint32_t)v.dims.size());
    for (int d : v.dims) h.u32((uint32_t)d);
    h.u32(kTagAtt);
    h.u32(2);
    put_att_text(h, "units", v.units);
    put_att_text(h, "long_name", v.long_name);
    h.u32((uint32_t)v.type);
    h.u32((uint32_t)(v.vsize > 0xffffffffull ? 0xffffffffu : v.vsize));
    h.u64(v.begin);
  }
}

// values of one variable in big-endian order, in chunks
bool write_values(FILE *f, const void *data, int type, uint64_t count)
{
  const int ts = type_size(type);
  const unsigned char *src = static_cast<const unsigned char *>(data);
  std::vector<unsigned char> buf;
  const uint64_t chunk = 1 << 20;
  for (uint64_t i = 0; i < count; i += chunk) {
    const uint64_t n = count - i < chunk ? count - i : chunk;
    buf.resize(n * ts);
    for (uint64_t q = 0; q < n; ++q)
      for (int a = 0; a < ts; ++a) buf[q * ts + a] = src[(i + q) * ts + (ts - 1 - a)];
    if (std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) return false;
  }
  const uint64_t bytes = count * ts;
  static const unsigned char zeros[4] = {0, 0, 0, 0};
  return pad4(bytes) == bytes || std::fwrite(zeros, 1, pad4(bytes) - bytes, f) == pad4(bytes) - bytes;
}

// ---------------------------------------------------------------- reading --

struct RDim {
  std::string name;
  uint64_t len;
};

struct RVar {
  std::string name;
  std::vector<uint32_t> dims;
  int type = 0;
  uint64_t vsize = 0, begin = 0;
  bool record = false;
};

struct Reader {
  std::vector<unsigned char> h;  // the file's first bytes (the header)
  size_t p = 0;
  bool ok = true;
  uint32_t u32()
  {
    if (p + 4 > h.size()) {
      ok = false;
      return 0;
    }
    const uint32_t v = (uint32_t)h[p] << 24 | (uint32_t)h[p + 1] << 16 | (uint32_t)h[p + 2] << 8 | h[p + 3];
    p += 4;
    return v;
  }
  uint64_t u64()
  {
    const uint64_t hi = u32();
    return hi << 32 | u32();
  }
  std::string name()
  {
    const uint32_t n = u32();
    if (!ok || p + n > h.size()) {
      ok = false;
      return {};
    }
    std::string s(h.begin() + p, h.begin() + p + n);
    p += pad4(n);
    return s;
  }
  void skip_atts()
  {
    const uint32_t tag = u32(), n = u32();
    if (tag == 0 && n == 0) return;
    if (tag != kTagAtt) {
      ok = false;
      return;
    }
    for (uint32_t a = 0; a < n && ok; ++a) {
      name();
      const int t = (int)u32();
      const uint32_t cnt = u32();
      if (!type_size(t)) {
        ok = false;
        return;
      }
      p += pad4((uint64_t)cnt * type_size(t));
    }
  }
};

struct NcFileIn {
  int version = 0;
  uint64_t numrecs = 0, recsize = 0;
  std::vector<RDim> dims;
  std::vector<RVar> vars;

  const RVar *var(const char *n) const
  {
    for (const RVar &v : vars)
      if (v.name == n) return &v;
    return nullptr;
  }
  const RDim *dim(const char *n) const
  {
    for (const RDim &d : dims)
      if (d.name == n) return &d;
    return nullptr;
  }
  // number of values of one record (or of the whole fixed-size variable)
  // (saturates at 2^48, far above any array this library accepts, so that a
  // malformed header cannot wrap the product onto an expected size)
  uint64_t count(const RVar &v) const
  {
    const uint64_t cap = (uint64_t)1 << 48;
    uint64_t n = 1;
    for (size_t a = 0; a < v.dims.size(); ++a) {
      if (a == 0 && v.record) continue;
      const uint64_t len = dims[v.dims[a]].len;
      if (len != 0 && n > cap / len) return cap;
      n *= len;
    }
    return n;
  }
};

int open_header(FILE *f, NcFileIn &nc)
{
  Reader r;
  // headers of workspace files are a few KB; grow until the parse succeeds
  for (size_t want = 1 << 14; want <= (size_t)1 << 26; want <<= 2) {
    r.h.assign(want, 0);
    std::fseek(f, 0, SEEK_SET);
    const size_t got = std::fread(r.h.data(), 1, want, f);
    r.h.resize(got);
    r.p = 0;
    r.ok = true;
    if (got < 8 || std::memcmp(r.h.data(), "CDF", 3) != 0) return QESMC_ERR_IO;
    nc.version = r.h[3];
    if (nc.version != 1 && nc.version != 2) return QESMC_ERR_IO;  // CDF-5 and HDF5 files are not read
    r.p = 4;
    nc.numrecs = r.u32();
    nc.dims.clear();
    nc.vars.clear();
    uint32_t tag = r.u32(), n = r.u32();
    if (tag == kTagDim) {
      for (uint32_t a = 0; a < n && r.ok; ++a) {
        RDim d;
        d.name = r.name();
        d.len = r.u32();
        nc.dims.push_back(d);
      }
    } else if (tag != 0 || n != 0) {
      return QESMC_ERR_IO;
    }
    r.skip_atts();  // global attributes
    tag = r.u32();
    n = r.u32();
    if (tag == kTagVar) {
      for (uint32_t a = 0; a < n && r.ok; ++a) {
        RVar v;
        v.name = r.name();
        const uint32_t nd = r.u32();
        for (uint32_t q = 0; q < nd && r.ok; ++q) {
          const uint32_t id = r.u32();
          if (!r.ok || id >= nc.dims.size()) {
            r.ok = false;  // an invalid dimension id ends the parse before it is used
            break;
          }
          v.dims.push_back(id);
        }
        if (!r.ok) break;
        r.skip_atts();
        v.type = (int)r.u32();
        v.vsize = r.u32();
        v.begin = nc.version == 2 ? r.u64() : r.u32();
        if (!r.ok || !type_size(v.type)) {
          r.ok = false;
          break;
        }
        v.record = !v.dims.empty() && nc.dims[v.dims[0]].len == 0;
        nc.vars.push_back(v);
      }
    } else if (tag != 0 || n != 0) {
      return QESMC_ERR_IO;
    }
    if (r.ok) break;
    if (got < want) return QESMC_ERR_IO;  // the whole file was read: not a valid header
  }
  if (!r.ok) return QESMC_ERR_IO;
  // record size: the record variables' slices, padded, in order (one record
  // variable alone is not padded)
  uint64_t rs = 0;
  int nrec = 0;
  for (const RVar &v : nc.vars)
    if (v.record) {
      rs += pad4(nc.count(v) * type_size(v.type));
      ++nrec;
    }
  if (nrec == 1)
    for (const RVar &v : nc.vars)
      if (v.record) rs = nc.count(v) * type_size(v.type);
  nc.recsize = rs;
  return QESMC_OK;
}

// whether a value converts to T without undefined behaviour: a float array
// takes NaN, infinities and every finite double within float's range; an
// integer array only non-NaN values within its range
template <typename T>
bool in_range(double val)
{
  if (std::is_floating_point<T>::value)
    return !(std::isfinite(val) && std::fabs(val) > (double)std::numeric_limits<T>::max());
  return val == val && val >= (double)std::numeric_limits<T>::min() && val <= (double)std::numeric_limits<T>::max();
}

template <typename T>
int read_as(FILE *f, const NcFileIn &nc, const RVar &v, uint64_t expect, T *out)
{
  if (nc.count(v) != expect) return QESMC_ERR_ARG;
  if (v.record && nc.numrecs < 1) return QESMC_ERR_IO;
  const int ts = type_size(v.type);
  if (std::fseek(f, (long)v.begin, SEEK_SET) != 0) return QESMC_ERR_IO;
  std::vector<unsigned char> buf;
  const uint64_t chunk = 1 << 20;
  for (uint64_t i = 0; i < expect; i += chunk) {
    const uint64_t n = expect - i < chunk ? expect - i : chunk;
    buf.resize(n * ts);
    if (std::fread(buf.data(), 1, buf.size(), f) != buf.size()) return QESMC_ERR_IO;
    for (uint64_t q = 0; q < n; ++q) {
      const unsigned char *s = &buf[q * ts];
      double val = 0.0;
      switch (v.type) {
      case kByte: val = (double)(signed char)s[0]; break;
      case kShort: val = (double)(int16_t)(s[0] << 8 | s[1]); break;
      case kInt: val = (double)(int32_t)((uint32_t)s[0] << 24 | (uint32_t)s[1] << 16 | (uint32_t)s[2] << 8 | s[3]); break;
      case kFloat: {
        const uint32_t w = (uint32_t)s[0] << 24 | (uint32_t)s[1] << 16 | (uint32_t)s[2] << 8 | s[3];
        float x;
        std::memcpy(&x, &w, 4);
        val = x;
        break;
      }
      case kDouble: {
        uint64_t w = 0;
        for (int a = 0; a < 8; ++a) w = w << 8 | s[a];
        std::memcpy(&val, &w, 8);
        break;
      }
      default: return QESMC_ERR_IO;
      }
      if (!in_range<T>(val)) return QESMC_ERR_IO;  // NaN or out of range for an integer array
      out[i + q] = (T)val;
    }
  }
  return QESMC_OK;
}

int file_dims(const NcFileIn &nc, int *nx, int *ny, int *nz)
{
  const RDim *xf = nc.dim("x_face"), *yf = nc.dim("y_face"), *zf = nc.dim("z_face");
  if (!xf || !yf || !zf) return QESMC_ERR_IO;
  if (xf->len > 0x7fffffffu || yf->len > 0x7fffffffu || zf->len > 0x7fffffffu) return QESMC_ERR_IO;
  *nx = (int)xf->len;
  *ny = (int)yf->len;
  *nz = (int)zf->len;
  return QESMC_OK;
}

}  // namespace

extern "C" {

int qesmc_workspace_write(const char *path, const qesmc_grid *grid, const qesmc_coeffs *cf, const qesmc_fields *fl)
{
  if (!path || !grid || !cf || !fl || !grid->dz_array) return QESMC_ERR_ARG;
  const int nx = grid->nx, ny = grid->ny, nz = grid->nz;
  if (nx < 3 || ny < 3 || nz < 3) return QESMC_ERR_ARG;
  const float *cells[6] = {cf->e, cf->f, cf->g, cf->h, cf->m, cf->n};
  for (int a = 0; a < 6; ++a)
    if (!cells[a]) return QESMC_ERR_ARG;
  if (!cf->icellflag || !fl->u0 || !fl->v0 || !fl->w0 || !fl->u || !fl->v || !fl->w) return QESMC_ERR_ARG;

  // coordinates as the reference writes them (WINDSOutputWorkspace.cpp:50-88,
  // Domain.cpp:153-166; the face coordinates keep the reference's offsets)
  std::vector<float> x(nx - 1), y(ny - 1), xf(nx), yf(ny), z_face, z;
  for (int i = 0; i < nx - 1; ++i) x[i] = ((float)i + 0.5f) * grid->dx;
  for (int j = 0; j < ny - 1; ++j) y[j] = ((float)j + 0.5f) * grid->dy;
  xf[0] = x[0] - grid->dx;
  for (int i = 0; i < nx - 1; ++i) xf[i + 1] = x[i];
  yf[0] = y[0] - grid->dy;
  for (int j = 0; j < ny - 1; ++j) yf[j + 1] = y[j] + grid->dy;
  qesmc::vertical_grid(nz, grid->dz_array, z_face, z);
  const double t0 = 0.0;

  enum { T, XF, YF, ZF, X, Y, Z };
  const std::vector<WDim> dims = {{"t", 0},
                                  {"x_face", (uint32_t)nx},
                                  {"y_face", (uint32_t)ny},
                                  {"z_face", (uint32_t)nz},
                                  {"x", (uint32_t)(nx - 1)},
                                  {"y", (uint32_t)(ny - 1)},
                                  {"z", (uint32_t)(nz - 1)}};
  std::vector<WVar> vars;
  auto add = [&](const char *name, const char *units, const char *long_name, int type, std::vector<int> d,
                 const void *data) {
    WVar v;
    v.name = name;
    v.units = units;
    v.long_name = long_name;
    v.type = type;
    v.dims = d;
    v.data = data;
    v.record = d[0] == T;
    v.count = 1;
    for (size_t a = v.record ? 1 : 0; a < d.size(); ++a) v.count *= dims[d[a]].len;
    v.vsize = pad4(v.count * type_size(type));
    vars.push_back(v);
  };
  add("t", "s", "time since start of simulation", kDouble, {T}, &t0);
  add("x_face", "m", "x-face", kFloat, {XF}, xf.data());
  add("y_face", "m", "y-face", kFloat, {YF}, yf.data());
  add("z_face", "m", "z-face", kFloat, {ZF}, z_face.data());
  add("u", "m s-1", "x-component velocity", kFloat, {T, ZF, YF, XF}, fl->u);
  add("v", "m s-1", "y-component velocity", kFloat, {T, ZF, YF, XF}, fl->v);
  add("w", "m s-1", "z-component velocity", kFloat, {T, ZF, YF, XF}, fl->w);
  add("u0", "m s-1", "x-component initial velocity", kFloat, {T, ZF, YF, XF}, fl->u0);
  add("v0", "m s-1", "y-component initial velocity", kFloat, {T, ZF, YF, XF}, fl->v0);
  add("w0", "m s-1", "z-component initial velocity", kFloat, {T, ZF, YF, XF}, fl->w0);
  add("x", "m", "x-distance", kFloat, {X}, x.data());
  add("y", "m", "y-distance", kFloat, {Y}, y.data());
  add("z", "m", "z-distance", kFloat, {Z}, z.data());
  add("dz_array", "m", "dz size of the cells", kFloat, {Z}, grid->dz_array);
  add("icellflag", "--", "icell flag value", kInt, {T, Z, Y, X}, cf->icellflag);
  static const char *const cn[6] = {"e", "f", "g", "h", "m", "n"};
  static const char *const cl[6] = {"e cut-cell coefficient", "f cut-cell coefficient", "g cut-cell coefficient",
                                    "h cut-cell coefficient", "m cut-cell coefficient", "n cut-cell coefficient"};
  for (int a = 0; a < 6; ++a) add(cn[a], "--", cl[a], kFloat, {T, Z, Y, X}, cells[a]);

  // offsets: the header, then the fixed-size variables, then one record
  Header h;
  build_header(h, dims, vars, 1);
  uint64_t off = h.b.size();
  for (WVar &v : vars)
    if (!v.record) {
      v.begin = off;
      off += v.vsize;
    }
  for (WVar &v : vars)
    if (v.record) {
      if (v.vsize > 0xfffffffcull) return QESMC_ERR_UNSUPPORTED;  // a record slice over 4 GiB
      v.begin = off;
      off += v.vsize;
    }
  build_header(h, dims, vars, 1);  // same size, now with the offsets

  File out;
  out.f = std::fopen(path, "wb");
  if (!out.f) return QESMC_ERR_IO;
  if (std::fwrite(h.b.data(), 1, h.b.size(), out.f) != h.b.size()) return QESMC_ERR_IO;
  for (int pass = 0; pass < 2; ++pass)
    for (const WVar &v : vars)
      if (v.record == (pass == 1) && !write_values(out.f, v.data, v.type, v.count)) return QESMC_ERR_IO;
  if (std::fclose(out.f) != 0) {
    out.f = nullptr;
    return QESMC_ERR_IO;
  }
  out.f = nullptr;
  return QESMC_OK;
}

int qesmc_workspace_dims(const char *path, int *nx, int *ny, int *nz)
{
  if (!path || !nx || !ny || !nz) return QESMC_ERR_ARG;
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return QESMC_ERR_IO;
  NcFileIn nc;
  const int rc = open_header(in.f, nc);
  if (rc != QESMC_OK) return rc;
  return file_dims(nc, nx, ny, nz);
}

int qesmc_workspace_read(const char *path, int nx, int ny, int nz, float *dx, float *dy, float *dz_array, float *e,
                         float *f, float *g, float *h, float *m, float *n, int *icellflag, float *u0, float *v0,
                         float *w0, float *u, float *v, float *w)
{
  if (!path) return QESMC_ERR_ARG;
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return QESMC_ERR_IO;
  NcFileIn nc;
  int rc = open_header(in.f, nc);
  if (rc != QESMC_OK) return rc;
  int fx = 0, fy = 0, fz = 0;
  if ((rc = file_dims(nc, &fx, &fy, &fz)) != QESMC_OK) return rc;
  if (fx != nx || fy != ny || fz != nz || nx < 3 || ny < 3 || nz < 3) return QESMC_ERR_ARG;
  const uint64_t ncell = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1), nface = (uint64_t)nx * ny * nz;

  auto need = [&](const char *name) -> const RVar * { return nc.var(name); };
  // grid spacing as Domain(const std::string&) derives it
  if (dx || dy) {
    for (int axis = 0; axis < 2; ++axis) {
      float *outp = axis == 0 ? dx : dy;
      if (!outp) continue;
      const RVar *c = need(axis == 0 ? "x" : "y");
      const uint64_t len = (uint64_t)(axis == 0 ? nx : ny) - 1;
      if (!c) return QESMC_ERR_ARG;
      std::vector<float> cv(len);
      if ((rc = read_as(in.f, nc, *c, len, cv.data())) != QESMC_OK) return rc;
      *outp = cv[1] - cv[0];
    }
  }
  if (dz_array) {
    if (const RVar *dz = need("dz_array")) {
      if ((rc = read_as(in.f, nc, *dz, (uint64_t)nz - 1, dz_array)) != QESMC_OK) return rc;
    } else {
      const RVar *zc = need("z");
      if (!zc) return QESMC_ERR_ARG;
      std::vector<float> zv(nz - 1);
      if ((rc = read_as(in.f, nc, *zc, (uint64_t)nz - 1, zv.data())) != QESMC_OK) return rc;
      const float dzu = zv[1] - zv[0];
      for (int k = 0; k < nz - 1; ++k) dz_array[k] = dzu;
    }
  }
  struct Item {
    const char *name;
    float *fp;
    int *ip;
    uint64_t count;
  } items[] = {{"e", e, nullptr, ncell},   {"f", f, nullptr, ncell},   {"g", g, nullptr, ncell},
               {"h", h, nullptr, ncell},   {"m", m, nullptr, ncell},   {"n", n, nullptr, ncell},
               {"icellflag", nullptr, icellflag, ncell},               {"u0", u0, nullptr, nface},
               {"v0", v0, nullptr, nface}, {"w0", w0, nullptr, nface}, {"u", u, nullptr, nface},
               {"v", v, nullptr, nface},   {"w", w, nullptr, nface}};
  for (const Item &it : items) {
    if (!it.fp && !it.ip) continue;
    const RVar *var = need(it.name);
    if (!var) return QESMC_ERR_ARG;
    rc = it.fp ? read_as(in.f, nc, *var, it.count, it.fp) : read_as(in.f, nc, *var, it.count, it.ip);
    if (rc != QESMC_OK) return rc;
  }
  return QESMC_OK;
}

}  // extern "C"
