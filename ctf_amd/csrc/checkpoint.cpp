/* Mara checkpoint files for the device path: the HDF5 layout of Mara's Lua drivers
 * (prim/<field> datasets, scalar string and number datasets, the status group) and the
 * JSON string form of their json module. Host code only; built as libmara_h5.so against
 * the system's HDF5 (1.10 API) and libmaragpu.so.
 *
 * Every HDF5 handle is closed on every path: the Scope helper below closes what it holds
 * when it leaves scope, in reverse order of opening.
 */
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <hdf5.h>

#include "../../include/mara_checkpoint.h"

namespace {

std::string g_err;

int fail(const std::string &what, int code = MARAGPU_E_IO)
{
  g_err = what;
  return code;
}

/* the handles of one operation, closed in reverse order */
struct Scope {
  struct H { hid_t id; herr_t (*close)(hid_t); };
  std::vector<H> hs;
  hid_t keep(hid_t id, herr_t (*close)(hid_t))
  {
    if (id >= 0) hs.push_back({id, close});
    return id;
  }
  ~Scope()
  {
    for (auto it = hs.rbegin(); it != hs.rend(); ++it) it->close(it->id);
  }
};

/* H5open once, with the library's automatic error printing off: errors come back as codes */
void quiet()
{
  static bool done = false;
  if (!done) {
    H5open();
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    done = true;
  }
}

const char *prim_names[8] = {"rho", "pre", "vx", "vy", "vz", "Bx", "By", "Bz"};   // Rmhd::GetPrimNames, Mara/rmhd.cpp:28-34

bool empty(const char *s) { return !s || !*s; }

/* hdf5.File(name, 'w' | 'r+' | 'r') (lua-hdf5/LuaHDF5.lua:612-645). The drivers' files are
 * made with the B-tree parameter istore_k = 32 (DataManagerHDF5's btree_ik,
 * modules/unigrid.lua:64) and default alignment. */
hid_t open_file(Scope &s, const char *path, int mode)
{
  quiet();
  if (!path) return -1;
  if (mode == MARA_H5_TRUNCATE) {
    const hid_t fcpl = s.keep(H5Pcreate(H5P_FILE_CREATE), H5Pclose);
    if (fcpl < 0 || H5Pset_istore_k(fcpl, 32) < 0) return -1;
    return s.keep(H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), H5Fclose);
  }
  return s.keep(H5Fopen(path, mode == MARA_H5_UPDATE ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
}

/* hdf5.Group(parent, name): opened if present, created if not (LuaHDF5.lua:653-666) */
hid_t group_at(Scope &s, hid_t file, const char *group, bool create)
{
  if (empty(group)) return file;
  if (H5Lexists(file, group, H5P_DEFAULT) > 0) return s.keep(H5Gopen2(file, group, H5P_DEFAULT), H5Gclose);
  if (!create) return -1;
  return s.keep(H5Gcreate2(file, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
}

/* clobber_if_H5Lexists (LuaHDF5.lua:75-83) */
bool clobber(hid_t where, const char *name)
{
  if (H5Lexists(where, name, H5P_DEFAULT) > 0) return H5Ldelete(where, name, H5P_DEFAULT) >= 0;
  return true;
}

/* the memory space of Mara's AoS array (N0 + 2Ng, N1 + 2Ng, N2 + 2Ng, Nq) with the interior of
 * field q selected (DataManagerHDF5:_setup_spaces, modules/unigrid.lua:269-291) */
hid_t member_space(Scope &s, const int N[3], int Ng, int Nq, int q)
{
  const hsize_t ext[4] = {(hsize_t)(N[0] + 2 * Ng), (hsize_t)(N[1] + 2 * Ng), (hsize_t)(N[2] + 2 * Ng), (hsize_t)Nq};
  const hid_t m = s.keep(H5Screate_simple(4, ext, nullptr), H5Sclose);
  if (m < 0) return -1;
  const hsize_t start[4] = {(hsize_t)Ng, (hsize_t)Ng, (hsize_t)Ng, (hsize_t)q};
  const hsize_t stride[4] = {1, 1, 1, (hsize_t)Nq};
  const hsize_t count[4] = {(hsize_t)N[0], (hsize_t)N[1], (hsize_t)N[2], 1};
  if (H5Sselect_hyperslab(m, H5S_SELECT_SET, start, stride, count, nullptr) < 0) return -1;
  return m;
}

bool valid_grid(const int N[3], int Ng, int Nq)
{
  return N && N[0] > 0 && N[1] > 0 && N[2] > 0 && Ng >= 0 && Nq > 0;
}

/* the (N0, N1, N2) double dataset <name> under g: created (mode TRUNCATE, or missing), else
 * opened after its shape is checked */
hid_t field_dataset(Scope &s, hid_t g, const char *name, const int N[3], bool create)
{
  const hsize_t dims[3] = {(hsize_t)N[0], (hsize_t)N[1], (hsize_t)N[2]};
  if (create || H5Lexists(g, name, H5P_DEFAULT) <= 0) {
    if (!clobber(g, name)) return -1;
    const hid_t fs = s.keep(H5Screate_simple(3, dims, nullptr), H5Sclose);
    if (fs < 0) return -1;
    return s.keep(H5Dcreate2(g, name, H5T_NATIVE_DOUBLE, fs, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  }
  const hid_t d = s.keep(H5Dopen2(g, name, H5P_DEFAULT), H5Dclose);
  if (d < 0) return -1;
  const hid_t fs = s.keep(H5Dget_space(d), H5Sclose);
  hsize_t have[3];
  if (fs < 0 || H5Sget_simple_extent_ndims(fs) != 3 || H5Sget_simple_extent_dims(fs, have, nullptr) != 3) return -1;
  for (int a = 0; a < 3; ++a)
    if (have[a] != dims[a]) return -1;
  return d;
}

}  // namespace

extern "C" {

const char *mara_h5_last_error(void) { return g_err.c_str(); }

int mara_h5_write_prim(const char *path, const char *group, const double *P, const int N[3], int Ng, int Nq,
                       const char *const *names, int mode)
{
  if (!P || !names || !valid_grid(N, Ng, Nq)) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, mode);
  if (f < 0) return fail(std::string("cannot open ") + (path ? path : "(null)") + " for writing");
  const hid_t g = group_at(s, f, group, true);
  if (g < 0) return fail("cannot open or create the group");
  for (int q = 0; q < Nq; ++q) {
    Scope t;
    if (!names[q]) return fail("null dataset name", MARAGPU_E_INVALID);
    const hid_t d = field_dataset(t, g, names[q], N, mode == MARA_H5_TRUNCATE);
    if (d < 0) return fail(std::string("cannot create data set ") + names[q]);
    const hid_t m = member_space(t, N, Ng, Nq, q);
    if (m < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, m, H5S_ALL, H5P_DEFAULT, P) < 0)
      return fail(std::string("cannot write data set ") + names[q]);
  }
  return MARAGPU_OK;
}

int mara_h5_read_prim(const char *path, const char *group, double *P, const int N[3], int Ng, int Nq,
                      const char *const *names)
{
  if (!P || !names || !valid_grid(N, Ng, Nq)) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return fail(std::string("cannot open ") + (path ? path : "(null)"));
  const hid_t g = group_at(s, f, group, false);
  if (g < 0) return fail("no such group");
  for (int q = 0; q < Nq; ++q) {
    Scope t;
    if (!names[q] || H5Lexists(g, names[q], H5P_DEFAULT) <= 0)
      return fail(std::string("no data set ") + (names[q] ? names[q] : "(null)"));
    const hid_t d = field_dataset(t, g, names[q], N, false);
    if (d < 0) return fail(std::string("data set ") + names[q] + " does not have the domain's shape", MARAGPU_E_SIZE);
    const hid_t m = member_space(t, N, Ng, Nq, q);
    if (m < 0 || H5Dread(d, H5T_NATIVE_DOUBLE, m, H5S_ALL, H5P_DEFAULT, P) < 0)
      return fail(std::string("cannot read data set ") + names[q]);
  }
  return MARAGPU_OK;
}

int mara_h5_extent(const char *path, const char *dset, int *ndim, long long dims[8])
{
  if (!dset || !ndim || !dims) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return fail("cannot open the file");
  if (H5Lexists(f, dset, H5P_DEFAULT) <= 0) return fail(std::string("no data set ") + dset);
  const hid_t d = s.keep(H5Dopen2(f, dset, H5P_DEFAULT), H5Dclose);
  const hid_t fs = d < 0 ? -1 : s.keep(H5Dget_space(d), H5Sclose);
  if (fs < 0) return fail(std::string("cannot open data set ") + dset);
  const int nd = H5Sget_simple_extent_ndims(fs);
  if (nd < 0 || nd > 8) return fail("unsupported rank");
  hsize_t h[8];
  H5Sget_simple_extent_dims(fs, h, nullptr);
  *ndim = nd;
  for (int a = 0; a < nd; ++a) dims[a] = (long long)h[a];
  return MARAGPU_OK;
}

int mara_h5_write_string(const char *path, const char *group, const char *name, const char *bytes, size_t n)
{
  if (!name || (!bytes && n)) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, MARA_H5_UPDATE);
  if (f < 0) return fail("cannot open the file for update");
  const hid_t g = group_at(s, f, group, true);
  if (g < 0 || !clobber(g, name)) return fail("cannot prepare the data set");
  // H5Tcopy(H5T_C_S1) sized to the text (a zero size is not a valid string type: one byte)
  const hid_t t = s.keep(H5Tcopy(H5T_C_S1), H5Tclose);
  const hid_t sp = s.keep(H5Screate(H5S_SCALAR), H5Sclose);
  if (t < 0 || sp < 0 || H5Tset_size(t, n ? n : 1) < 0) return fail("cannot make the string type");
  const hid_t d = s.keep(H5Dcreate2(g, name, t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  const char zero = 0;
  if (d < 0 || H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, n ? bytes : &zero) < 0)
    return fail(std::string("cannot write string ") + name);
  return MARAGPU_OK;
}

int mara_h5_write_number(const char *path, const char *group, const char *name, double value)
{
  if (!name) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, MARA_H5_UPDATE);
  if (f < 0) return fail("cannot open the file for update");
  const hid_t g = group_at(s, f, group, true);
  if (g < 0 || !clobber(g, name)) return fail("cannot prepare the data set");
  const hid_t sp = s.keep(H5Screate(H5S_SCALAR), H5Sclose);
  const hid_t d = sp < 0 ? -1 : s.keep(H5Dcreate2(g, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT,
                                                  H5P_DEFAULT), H5Dclose);
  if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &value) < 0)
    return fail(std::string("cannot write number ") + name);
  return MARAGPU_OK;
}

int mara_h5_kind(const char *path, const char *group, const char *name)
{
  if (!name) return MARA_H5_MISSING;
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return MARA_H5_MISSING;
  const hid_t g = group_at(s, f, group, false);
  if (g < 0 || H5Lexists(g, name, H5P_DEFAULT) <= 0) return MARA_H5_MISSING;
  H5O_info_t info;
  if (H5Oget_info_by_name(g, name, &info, H5P_DEFAULT) < 0 || info.type != H5O_TYPE_DATASET) return MARA_H5_OTHER;
  const hid_t d = s.keep(H5Dopen2(g, name, H5P_DEFAULT), H5Dclose);
  const hid_t t = d < 0 ? -1 : s.keep(H5Dget_type(d), H5Tclose);
  if (t < 0) return MARA_H5_OTHER;
  const H5T_class_t c = H5Tget_class(t);
  if (c == H5T_STRING) return MARA_H5_STRING;
  if (c == H5T_FLOAT) return MARA_H5_NUMBER;
  return MARA_H5_OTHER;
}

int mara_h5_read_string(const char *path, const char *group, const char *name, char *buf, size_t cap, size_t *n)
{
  if (!name || !n) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return fail("cannot open the file");
  const hid_t g = group_at(s, f, group, false);
  if (g < 0 || H5Lexists(g, name, H5P_DEFAULT) <= 0) return fail(std::string("no data set ") + name);
  const hid_t d = s.keep(H5Dopen2(g, name, H5P_DEFAULT), H5Dclose);
  const hid_t t = d < 0 ? -1 : s.keep(H5Dget_type(d), H5Tclose);
  if (t < 0 || H5Tget_class(t) != H5T_STRING || H5Tis_variable_str(t) > 0)
    return fail(std::string(name) + " is not a fixed-length string");
  // DataSet:read sizes the buffer from the type (LuaHDF5.lua:258-283): the whole size,
  // padding included; the text is what tostring() of it gives, the full buffer
  const size_t size = H5Tget_size(t);
  std::vector<char> tmp(size + 1, 0);
  if (H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, tmp.data()) < 0) return fail(std::string("cannot read ") + name);
  *n = size;
  if (buf) std::memcpy(buf, tmp.data(), size < cap ? size : cap);
  return MARAGPU_OK;
}

int mara_h5_read_number(const char *path, const char *group, const char *name, double *value)
{
  if (!name || !value) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return fail("cannot open the file");
  const hid_t g = group_at(s, f, group, false);
  if (g < 0 || H5Lexists(g, name, H5P_DEFAULT) <= 0) return fail(std::string("no data set ") + name);
  const hid_t d = s.keep(H5Dopen2(g, name, H5P_DEFAULT), H5Dclose);
  const hid_t sp = d < 0 ? -1 : s.keep(H5Dget_space(d), H5Sclose);
  if (sp < 0 || H5Sget_simple_extent_npoints(sp) != 1) return fail(std::string(name) + " is not a scalar");
  if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, value) < 0)
    return fail(std::string("cannot read ") + name);
  return MARAGPU_OK;
}

int mara_h5_keys(const char *path, const char *group, char *buf, size_t cap, size_t *n)
{
  if (!n) return fail("invalid arguments", MARAGPU_E_INVALID);
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return fail("cannot open the file");
  const hid_t g = group_at(s, f, group, false);
  if (g < 0) return fail("no such group");
  std::string out;
  auto add = [](hid_t, const char *name, const H5L_info_t *, void *op) -> herr_t {
    std::string &o = *static_cast<std::string *>(op);
    if (!o.empty()) o += '\n';
    o += name;
    return 0;
  };
  hsize_t idx = 0;
  if (H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, add, &out) < 0) return fail("cannot list the group");
  *n = out.size();
  if (buf) std::memcpy(buf, out.data(), out.size() < cap ? out.size() : cap);
  return MARAGPU_OK;
}

int maragpu_checkpoint_write(maragpu_ctx *ctx, const char *path, const char *group, int mode)
{
  if (!ctx) return fail("invalid arguments", MARAGPU_E_INVALID);
  maragpu_config cfg;
  if (int rc = maragpu_get_config(ctx, &cfg)) return fail("cannot read the context", rc);
  const size_t ni = (size_t)cfg.N[0] * cfg.N[1] * cfg.N[2];
  std::vector<double> plane(ni);
  Scope s;
  const hid_t f = open_file(s, path, mode);
  if (f < 0) return fail(std::string("cannot open ") + (path ? path : "(null)") + " for writing");
  const hid_t g = group_at(s, f, group, true);
  if (g < 0) return fail("cannot open or create the group");
  for (int q = 0; q < 8; ++q) {
    if (int rc = maragpu_download_field(ctx, q, plane.data(), ni)) return fail("device download failed", rc);
    Scope t;
    const hid_t d = field_dataset(t, g, prim_names[q], cfg.N, mode == MARA_H5_TRUNCATE);
    if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, plane.data()) < 0)
      return fail(std::string("cannot write data set ") + prim_names[q]);
  }
  return MARAGPU_OK;
}

int maragpu_checkpoint_read(maragpu_ctx *ctx, const char *path, const char *group)
{
  if (!ctx) return fail("invalid arguments", MARAGPU_E_INVALID);
  maragpu_config cfg;
  if (int rc = maragpu_get_config(ctx, &cfg)) return fail("cannot read the context", rc);
  const size_t ni = (size_t)cfg.N[0] * cfg.N[1] * cfg.N[2];
  std::vector<double> plane(ni);
  Scope s;
  const hid_t f = open_file(s, path, -1);
  if (f < 0) return fail(std::string("cannot open ") + (path ? path : "(null)"));
  const hid_t g = group_at(s, f, group, false);
  if (g < 0) return fail("no such group");
  // every field is read before the first upload, so a bad file leaves the state as it was
  std::vector<std::vector<double>> fields(8);
  for (int q = 0; q < 8; ++q) {
    Scope t;
    if (H5Lexists(g, prim_names[q], H5P_DEFAULT) <= 0) return fail(std::string("no data set ") + prim_names[q]);
    const hid_t d = field_dataset(t, g, prim_names[q], cfg.N, false);
    if (d < 0) return fail(std::string("data set ") + prim_names[q] + " does not have the domain's shape", MARAGPU_E_SIZE);
    fields[q].resize(ni);
    if (H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, fields[q].data()) < 0)
      return fail(std::string("cannot read data set ") + prim_names[q]);
  }
  for (int q = 0; q < 8; ++q)
    if (int rc = maragpu_upload_field(ctx, q, fields[q].data(), ni)) return fail("device upload failed", rc);
  return MARAGPU_OK;
}

/* quotestring (modules/json.lua:289-304, the pure-Lua form): first every byte 0-31, '"',
 * '\\' and 127 becomes an escape (the short ones of escapecodes, :243-246, else \u00xx);
 * then, if the text holds any of the lead bytes C2 D8 DC E1 E2 EF, the listed UTF-8
 * sequences become \uXXXX escapes of their code points, one pattern after another as the
 * gsub calls run. Every other byte is copied. */
int mara_json_quote(const char *in, size_t len, char *out, size_t cap, size_t *n)
{
  if ((!in && len) || !n) return fail("invalid arguments", MARAGPU_E_INVALID);
  const unsigned char *b = reinterpret_cast<const unsigned char *>(in);
  auto hex4 = [](std::string &o, unsigned v) {
    static const char *d = "0123456789abcdef";
    o += "\\u";
    for (int sh = 12; sh >= 0; sh -= 4) o += d[(v >> sh) & 15];
  };
  std::string a;
  a.reserve(len + 2);
  for (size_t i = 0; i < len; ++i) {
    const unsigned c = b[i];
    if (c == '"') a += "\\\"";
    else if (c == '\\') a += "\\\\";
    else if (c == '\b') a += "\\b";
    else if (c == '\f') a += "\\f";
    else if (c == '\n') a += "\\n";
    else if (c == '\r') a += "\\r";
    else if (c == '\t') a += "\\t";
    else if (c < 32 || c == 127) hex4(a, c);
    else a += (char)c;
  }
  bool lead = false;
  for (unsigned char c : a)
    if (c == 0xC2 || c == 0xD8 || c == 0xDC || c == 0xE1 || c == 0xE2 || c == 0xEF) lead = true;
  if (lead) {
    // the second-pass patterns, in order: {bytes before the last, last-byte set}
    struct Pat { unsigned char p[2]; int np; bool (*last)(unsigned); };
    const Pat pats[] = {
      {{0xC2, 0}, 1, [](unsigned x) { return (x >= 0x80 && x <= 0x9F) || x == 0xAD; }},
      {{0xD8, 0}, 1, [](unsigned x) { return x >= 0x80 && x <= 0x84; }},
      {{0xDC, 0}, 1, [](unsigned x) { return x == 0x8F; }},
      {{0xE1, 0x9E}, 2, [](unsigned x) { return x == 0xB4 || x == 0xB5; }},
      {{0xE2, 0x80}, 2, [](unsigned x) { return (x >= 0x8C && x <= 0x8F) || x == 0xA8 || x == 0xAF; }},
      {{0xE2, 0x81}, 2, [](unsigned x) { return x >= 0xA0 && x <= 0xAF; }},
      {{0xEF, 0xBB}, 2, [](unsigned x) { return x == 0xBF; }},
      {{0xEF, 0xBF}, 2, [](unsigned x) { return x == 0xB0 || x == 0xBF; }},
    };
    for (const Pat &pt : pats) {
      std::string r;
      r.reserve(a.size());
      size_t i = 0;
      while (i < a.size()) {
        const size_t m = (size_t)pt.np + 1;
        bool hit = i + m <= a.size();
        for (int k = 0; hit && k < pt.np; ++k) hit = (unsigned char)a[i + k] == pt.p[k];
        if (hit) hit = pt.last((unsigned char)a[i + pt.np]);
        if (hit) {
          // escapeutf8 of a 2- or 3-byte sequence: its code point
          const unsigned c0 = (unsigned char)a[i], c1 = (unsigned char)a[i + 1];
          unsigned v;
          if (m == 2) v = (c0 - 0xC0) * 0x40 + c1 - 0x80;
          else v = ((c0 - 0xE0) * 0x40 + c1 - 0x80) * 0x40 + (unsigned char)a[i + 2] - 0x80;
          hex4(r, v);
          i += m;
        }
        else r += a[i++];
      }
      a.swap(r);
    }
  }
  const std::string q = "\"" + a + "\"";
  *n = q.size();
  if (out) std::memcpy(out, q.data(), q.size() < cap ? q.size() : cap);
  return MARAGPU_OK;
}

/* scanstring (modules/json.lua:529-590): from the opening quote to the closing one; \uXXXX
 * escapes become UTF-8 (unichar, :507-527), a high surrogate followed by an escaped low one
 * combines; \" \\ \/ \b \f \n \r \t as usual, any other escaped character stands for itself,
 * and an invalid \u escape yields the letter u. */
int mara_json_unquote(const char *in, size_t len, char *out, size_t cap, size_t *n)
{
  if (!in || !n || len < 1 || in[0] != '"') return fail("not a JSON string", MARAGPU_E_INVALID);
  std::string o;
  auto utf8 = [&o](unsigned long v) -> bool {
    if (v <= 0x7F) o += (char)v;
    else if (v <= 0x7FF) { o += (char)(0xC0 + v / 0x40); o += (char)(0x80 + v % 0x40); }
    else if (v <= 0xFFFF) {
      o += (char)(0xE0 + v / 0x1000); o += (char)(0x80 + (v / 0x40) % 0x40); o += (char)(0x80 + v % 0x40);
    }
    else if (v <= 0x10FFFF) {
      o += (char)(0xF0 + v / 0x40000); o += (char)(0x80 + (v / 0x1000) % 0x40);
      o += (char)(0x80 + (v / 0x40) % 0x40); o += (char)(0x80 + v % 0x40);
    }
    else return false;
    return true;
  };
  // tonumber(s, 16) of four characters; Lua's tonumber takes surrounding spaces and a sign,
  // which a well-formed escape never holds: plain hexadecimal digits only
  auto hex = [&](size_t at, unsigned long &v) -> bool {
    if (at + 4 > len) return false;
    v = 0;
    for (size_t k = at; k < at + 4; ++k) {
      const char c = in[k];
      int d = (c >= '0' && c <= '9') ? c - '0' : (c >= 'a' && c <= 'f') ? c - 'a' + 10
            : (c >= 'A' && c <= 'F') ? c - 'A' + 10 : -1;
      if (d < 0) return false;
      v = v * 16 + (unsigned long)d;
    }
    return true;
  };
  size_t i = 1;
  while (true) {
    if (i >= len) return fail("unterminated string", MARAGPU_E_INVALID);
    const char c = in[i];
    if (c == '"') break;
    if (c != '\\') { o += c; ++i; continue; }
    if (i + 1 >= len) return fail("unterminated string", MARAGPU_E_INVALID);
    const char e = in[i + 1];
    bool done = false;
    if (e == 'u') {
      unsigned long v;
      if (hex(i + 2, v)) {
        size_t used = 6;
        if (v >= 0xD800 && v <= 0xDBFF && i + 8 <= len && in[i + 6] == '\\' && in[i + 7] == 'u') {
          unsigned long v2;
          if (hex(i + 8, v2) && v2 >= 0xDC00 && v2 <= 0xDFFF) {
            v = (v - 0xD800) * 0x400 + (v2 - 0xDC00) + 0x10000;
            used = 12;
          }
        }
        if (utf8(v)) { i += used; done = true; }
      }
    }
    if (!done) {
      switch (e) {
      case 'b': o += '\b'; break;
      case 'f': o += '\f'; break;
      case 'n': o += '\n'; break;
      case 'r': o += '\r'; break;
      case 't': o += '\t'; break;
      default: o += e; break;     // '"', '\\', '/' and any other character stand for themselves
      }
      i += 2;
    }
  }
  *n = o.size();
  if (out) std::memcpy(out, o.data(), o.size() < cap ? o.size() : cap);
  return MARAGPU_OK;
}

}  // extern "C"
