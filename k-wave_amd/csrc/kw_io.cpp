// HDF5 file contract of the k-Wave external binaries (include/kwave_io.h):
// dataset I/O in writeMatrix.m's layout, and kw_run_file, which reads an
// input file written by kspaceFirstOrder_saveToDisk.m, runs it through the
// solver C ABI and writes the output datasets kspaceFirstOrder3DC.m reads.
#include <hdf5.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/kwave_io.h"

namespace {

thread_local std::string g_ioerr;

kw_status iofail(kw_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_ioerr = buf;
  return st;
}

struct SilenceH5 {
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  SilenceH5() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~SilenceH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

kw_status write_string_attr(hid_t obj, const char* name, const std::string& value) {
  hid_t type = H5Tcopy(H5T_C_S1);
  H5Tset_size(type, value.size() > 0 ? value.size() : 1);
  H5Tset_strpad(type, H5T_STR_NULLTERM);
  hid_t space = H5Screate(H5S_SCALAR);
  if (H5Aexists(obj, name) > 0) H5Adelete(obj, name);
  hid_t a = H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT);
  herr_t e = a >= 0 ? H5Awrite(a, type, value.c_str()) : -1;
  if (a >= 0) H5Aclose(a);
  H5Sclose(space);
  H5Tclose(type);
  return e < 0 ? iofail(KW_ERR_IO, "cannot write attribute %s", name) : KW_OK;
}

}  // namespace

struct kwio_file {
  hid_t id = -1;
  bool writable = false;
};

static bool file_exists(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (f) std::fclose(f);
  return f != nullptr;
}

// checkpoint file: root attributes with file_type "checkpoint" and the solver
// state as one uint64 dataset
static kw_status save_checkpoint(const char* path, kw_sim* sim) {
  const size_t bytes = kw_sim_checkpoint_bytes(sim);
  std::vector<uint64_t> words((bytes + 7) / 8);
  kw_status st = kw_sim_checkpoint_save(sim, words.data(), bytes);
  if (st) return iofail(st, "%s", kw_last_error());
  const std::string tmp = std::string(path) + ".part";
  kwio_file* f = nullptr;
  if ((st = kwio_open(tmp.c_str(), 1, &f))) return st;
  const uint64_t dims[3] = {1, 1, (uint64_t)words.size()};
  const uint64_t nb[3] = {1, 1, 1};
  const uint64_t b64 = bytes;
  st = kwio_write(f, "kw_state", KWIO_UINT64, dims, words.data(), 0);
  if (!st) st = kwio_write(f, "kw_state_bytes", KWIO_UINT64, nb, &b64, 0);
  if (!st) st = kwio_write_root_attributes(f, "checkpoint", "Checkpoint of kspaceFirstOrder-HIP");
  kwio_close(f);
  if (st) return st;
  if (std::rename(tmp.c_str(), path) != 0) return iofail(KW_ERR_IO, "cannot write checkpoint %s", path);
  return KW_OK;
}

static kw_status load_checkpoint(const char* path, kw_sim* sim) {
  kwio_file* f = nullptr;
  kw_status st = kwio_open(path, 0, &f);
  if (st) return st;
  char ft[64] = {0};
  if (kwio_read_string_attribute(f, "/", "file_type", ft, sizeof ft) || std::string(ft) != "checkpoint") {
    kwio_close(f);
    return iofail(KW_ERR_INVALID, "%s is not a checkpoint file (file_type '%s')", path, ft);
  }
  int dt = 0;
  uint64_t dims[3];
  uint64_t bytes = 0;
  if ((st = kwio_read(f, "kw_state_bytes", &bytes, 8)) || (st = kwio_info(f, "kw_state", &dt, dims))) {
    kwio_close(f);
    return st;
  }
  std::vector<uint64_t> words(dims[0] * dims[1] * dims[2]);
  st = kwio_read(f, "kw_state", words.data(), words.size() * 8);
  kwio_close(f);
  if (st) return st;
  if (bytes > words.size() * 8) return iofail(KW_ERR_INVALID, "%s: truncated checkpoint", path);
  if ((st = kw_sim_checkpoint_load(sim, words.data(), (size_t)bytes))) return iofail(st, "%s", kw_last_error());
  return KW_OK;
}

extern "C" {

kw_status kwio_open(const char* path, int mode, kwio_file** out) {
  if (!path || !out) return iofail(KW_ERR_INVALID, "null argument");
  SilenceH5 quiet;
  hid_t id = mode == 1 ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                       : H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (id < 0) return iofail(KW_ERR_IO, "cannot %s HDF5 file %s", mode == 1 ? "create" : "open", path);
  *out = new kwio_file{id, mode == 1};
  return KW_OK;
}

kw_status kwio_close(kwio_file* f) {
  if (!f) return KW_OK;
  herr_t e = H5Fclose(f->id);
  delete f;
  return e < 0 ? iofail(KW_ERR_IO, "H5Fclose failed") : KW_OK;
}

int kwio_exists(kwio_file* f, const char* name) {
  if (!f || !name) return 0;
  SilenceH5 quiet;
  return H5Lexists(f->id, name, H5P_DEFAULT) > 0 ? 1 : 0;
}

// Rank-3 datasets follow writeMatrix.m's chunking; rank 4 (per-cuboid time
// series, (Nt, cz, cy, cx)) is chunked one time step per chunk.  Names with a
// '/' create their groups ("p/1": dataset 1 in group p).
static kw_status write_nd(kwio_file* f, const char* name, int dtype, int rank, const uint64_t* dims_in,
                          const void* data, int compression_level) {
  if (!f || !name || !dims_in || !data) return iofail(KW_ERR_INVALID, "null argument");
  if (dtype != KWIO_FLOAT32 && dtype != KWIO_UINT64) return iofail(KW_ERR_INVALID, "dtype %d", dtype);
  if (rank != 3 && rank != 4) return iofail(KW_ERR_INVALID, "%s: rank %d", name, rank);
  hsize_t dims[4] = {1, 1, 1, 1};
  for (int i = 0; i < rank; ++i) dims[i] = dims_in[i];
  for (int i = 0; i < rank; ++i)
    if (dims[i] == 0) return iofail(KW_ERR_INVALID, "%s: zero dimension", name);
  // chunking of writeMatrix.m:56-110, in C order (d0, d1, d2) == MATLAB [d2, d1, d0]
  hsize_t chunk[4] = {1, 1, 1, 1};
  int comp = compression_level;
  if (rank == 4) {
    chunk[1] = dims[1];
    chunk[2] = dims[2];
    chunk[3] = dims[3];
  } else {
    const hsize_t n = dims[0] * dims[1] * dims[2];
    const hsize_t one_mb = 1024 * 1024 / 8;
    if (dims[0] > 1) {  // 3-D
      chunk[1] = dims[1];
      chunk[2] = dims[2];
    } else if (dims[1] > 1 && dims[2] > 1) {  // 2-D
      chunk[2] = dims[2];
    } else {  // vector or scalar
      int axis = dims[2] > 1 ? 2 : (dims[1] > 1 ? 1 : 2);
      if (n > one_mb) {
        chunk[axis] = one_mb;
      } else {
        comp = 0;
        chunk[axis] = dims[axis];
      }
    }
  }
  SilenceH5 quiet;
  if (H5Lexists(f->id, name, H5P_DEFAULT) > 0) H5Ldelete(f->id, name, H5P_DEFAULT);
  hid_t space = H5Screate_simple(rank, dims, nullptr);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_chunk(dcpl, rank, chunk);
  if (comp > 0) H5Pset_deflate(dcpl, (unsigned)comp);
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  H5Pset_create_intermediate_group(lcpl, 1);
  hid_t ftype = dtype == KWIO_FLOAT32 ? H5T_IEEE_F32LE : H5T_STD_U64LE;
  hid_t mtype = dtype == KWIO_FLOAT32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_UINT64;
  hid_t ds = H5Dcreate2(f->id, name, ftype, space, lcpl, dcpl, H5P_DEFAULT);
  H5Pclose(lcpl);
  kw_status st = KW_OK;
  if (ds < 0 || H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
    st = iofail(KW_ERR_IO, "cannot write dataset %s", name);
  } else {
    // writeMatrix.m:206-212
    if ((st = write_string_attr(ds, "domain_type", "real")) == KW_OK)
      st = write_string_attr(ds, "data_type", dtype == KWIO_FLOAT32 ? "float" : "long");
  }
  if (ds >= 0) H5Dclose(ds);
  H5Pclose(dcpl);
  H5Sclose(space);
  return st;
}

kw_status kwio_write(kwio_file* f, const char* name, int dtype, const uint64_t dims[3], const void* data,
                     int compression_level) {
  return write_nd(f, name, dtype, 3, dims, data, compression_level);
}

kw_status kwio_write_root_attributes(kwio_file* f, const char* file_type, const char* description) {
  if (!f) return iofail(KW_ERR_INVALID, "null argument");
  char date[64];
  std::time_t t = std::time(nullptr);
  std::strftime(date, sizeof date, "%d-%b-%Y-%H-%M-%S", std::localtime(&t));
  hid_t root = H5Gopen2(f->id, "/", H5P_DEFAULT);
  kw_status st = KW_OK;
  const char* kv[][2] = {{"major_version", "1"},
                         {"minor_version", "2"},
                         {"created_by", "kspaceFirstOrder-HIP (MI355X)"},
                         {"file_description", description ? description : ""},
                         {"file_type", file_type ? file_type : "input"},
                         {"creation_date", date}};
  for (auto& p : kv)
    if (st == KW_OK) st = write_string_attr(root, p[0], p[1]);
  H5Gclose(root);
  return st;
}

kw_status kwio_info(kwio_file* f, const char* name, int* dtype, uint64_t dims[3]) {
  if (!f || !name) return iofail(KW_ERR_INVALID, "null argument");
  SilenceH5 quiet;
  hid_t ds = H5Dopen2(f->id, name, H5P_DEFAULT);
  if (ds < 0) return iofail(KW_ERR_IO, "dataset %s not found", name);
  hid_t sp = H5Dget_space(ds);
  int nd = H5Sget_simple_extent_ndims(sp);
  hsize_t d[8] = {1, 1, 1, 1, 1, 1, 1, 1};
  kw_status st = KW_OK;
  if (nd < 0 || nd > 3) {
    st = iofail(KW_ERR_IO, "dataset %s has %d dimensions", name, nd);
  } else {
    H5Sget_simple_extent_dims(sp, d, nullptr);
    for (int i = 0; i < 3; ++i) dims[i] = 1;
    for (int i = 0; i < nd; ++i) dims[3 - nd + i] = d[i];
    hid_t ty = H5Dget_type(ds);
    H5T_class_t cls = H5Tget_class(ty);
    if (dtype) *dtype = cls == H5T_FLOAT ? KWIO_FLOAT32 : KWIO_UINT64;
    if (cls != H5T_FLOAT && cls != H5T_INTEGER) st = iofail(KW_ERR_IO, "dataset %s: unsupported type", name);
    H5Tclose(ty);
  }
  H5Sclose(sp);
  H5Dclose(ds);
  return st;
}

kw_status kwio_info_nd(kwio_file* f, const char* name, int* dtype, int* rank, uint64_t dims[8]) {
  if (!f || !name || !rank || !dims) return iofail(KW_ERR_INVALID, "null argument");
  SilenceH5 quiet;
  hid_t ds = H5Dopen2(f->id, name, H5P_DEFAULT);
  if (ds < 0) return iofail(KW_ERR_IO, "dataset %s not found", name);
  hid_t sp = H5Dget_space(ds);
  const int nd = H5Sget_simple_extent_ndims(sp);
  kw_status st = KW_OK;
  if (nd < 0 || nd > 8) {
    st = iofail(KW_ERR_IO, "dataset %s has %d dimensions", name, nd);
  } else {
    hsize_t d[8];
    H5Sget_simple_extent_dims(sp, d, nullptr);
    *rank = nd;
    for (int i = 0; i < nd; ++i) dims[i] = d[i];
    hid_t ty = H5Dget_type(ds);
    H5T_class_t cls = H5Tget_class(ty);
    if (dtype) *dtype = cls == H5T_FLOAT ? KWIO_FLOAT32 : KWIO_UINT64;
    if (cls != H5T_FLOAT && cls != H5T_INTEGER) st = iofail(KW_ERR_IO, "dataset %s: unsupported type", name);
    H5Tclose(ty);
  }
  H5Sclose(sp);
  H5Dclose(ds);
  return st;
}

kw_status kwio_read(kwio_file* f, const char* name, void* dst, size_t bytes) {
  int dtype = 0, rank = 0;
  uint64_t dims[8];
  kw_status st = kwio_info_nd(f, name, &dtype, &rank, dims);
  if (st) return st;
  uint64_t count = 1;
  for (int i = 0; i < rank; ++i) count *= dims[i];
  const size_t need = (size_t)count * (dtype == KWIO_FLOAT32 ? 4 : 8);
  if (bytes != need) return iofail(KW_ERR_INVALID, "%s: %zu bytes requested, dataset holds %zu", name, bytes, need);
  SilenceH5 quiet;
  hid_t ds = H5Dopen2(f->id, name, H5P_DEFAULT);
  herr_t e = H5Dread(ds, dtype == KWIO_FLOAT32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL,
                     H5P_DEFAULT, dst);
  H5Dclose(ds);
  return e < 0 ? iofail(KW_ERR_IO, "cannot read %s", name) : KW_OK;
}

static herr_t list_cb(hid_t, const char* name, const H5O_info_t* info, void* op) {
  if (info->type == H5O_TYPE_DATASET) static_cast<std::string*>(op)->append(name).append("\n");
  return 0;
}

size_t kwio_list(kwio_file* f, char* buf, size_t cap) {
  if (!f) return 0;
  std::string all;
  H5Ovisit(f->id, H5_INDEX_NAME, H5_ITER_INC, list_cb, &all);
  if (buf && cap > 0) {
    size_t n = std::min(cap - 1, all.size());
    memcpy(buf, all.data(), n);
    buf[n] = 0;
  }
  return all.size() + 1;
}

kw_status kwio_read_string_attribute(kwio_file* f, const char* object, const char* name, char* buf, size_t cap) {
  if (!f || !object || !name || !buf || cap == 0) return iofail(KW_ERR_INVALID, "null argument");
  SilenceH5 quiet;
  hid_t a = H5Aopen_by_name(f->id, object, name, H5P_DEFAULT, H5P_DEFAULT);
  if (a < 0) return iofail(KW_ERR_IO, "attribute %s of %s not found", name, object);
  hid_t ty = H5Aget_type(a);
  kw_status st = KW_OK;
  if (H5Tis_variable_str(ty) > 0) {
    char* s = nullptr;
    hid_t mt = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt, H5T_VARIABLE);
    if (H5Aread(a, mt, &s) < 0 || !s) {
      st = iofail(KW_ERR_IO, "cannot read attribute %s", name);
    } else {
      snprintf(buf, cap, "%s", s);
      H5free_memory(s);
    }
    H5Tclose(mt);
  } else {
    size_t sz = H5Tget_size(ty);
    std::vector<char> tmp(sz + 1, 0);
    hid_t mt = H5Tcopy(H5T_C_S1);
    H5Tset_size(mt, sz);
    if (H5Aread(a, mt, tmp.data()) < 0) st = iofail(KW_ERR_IO, "cannot read attribute %s", name);
    H5Tclose(mt);
    snprintf(buf, cap, "%s", tmp.data());
  }
  H5Tclose(ty);
  H5Aclose(a);
  return st;
}

}  // extern "C"

// -----------------------------------------------------------------------------
// kw_run_file
// -----------------------------------------------------------------------------
namespace {

struct Reader {
  kwio_file* f;
  std::vector<std::vector<float>> fbufs;
  std::vector<std::vector<uint64_t>> ubufs;
  kw_status st = KW_OK;

  bool has(const char* n) { return kwio_exists(f, n) != 0; }

  uint64_t u64(const char* n, uint64_t def = 0, bool required = false) {
    if (st) return def;
    if (!has(n)) {
      if (required) st = iofail(KW_ERR_INVALID, "input file lacks dataset %s", n);
      return def;
    }
    int dt;
    uint64_t d[3];
    if ((st = kwio_info(f, n, &dt, d))) return def;
    if (d[0] * d[1] * d[2] != 1) {
      st = iofail(KW_ERR_INVALID, "%s must be a scalar", n);
      return def;
    }
    if (dt == KWIO_UINT64) {
      uint64_t v = 0;
      st = kwio_read(f, n, &v, 8);
      return v;
    }
    float v = 0;
    st = kwio_read(f, n, &v, 4);
    return (uint64_t)v;
  }

  double f64(const char* n, double def = 0.0, bool required = false) {
    if (st) return def;
    if (!has(n)) {
      if (required) st = iofail(KW_ERR_INVALID, "input file lacks dataset %s", n);
      return def;
    }
    std::vector<float> v = floats(n);
    if (st) return def;
    if (v.size() != 1) {
      st = iofail(KW_ERR_INVALID, "%s must be a scalar", n);
      return def;
    }
    return (double)v[0];
  }

  std::vector<float> floats(const char* n) {
    int dt;
    uint64_t d[3];
    std::vector<float> v;
    if ((st = kwio_info(f, n, &dt, d))) return v;
    if (dt != KWIO_FLOAT32) {
      st = iofail(KW_ERR_INVALID, "%s must be float32", n);
      return v;
    }
    v.resize(d[0] * d[1] * d[2]);
    st = kwio_read(f, n, v.data(), v.size() * 4);
    return v;
  }

  kw_real_array real(const char* n, bool required) {
    kw_real_array a{nullptr, 0};
    if (st) return a;
    if (!has(n)) {
      if (required) st = iofail(KW_ERR_INVALID, "input file lacks dataset %s", n);
      return a;
    }
    fbufs.push_back(floats(n));
    a.data = fbufs.back().data();
    a.count = fbufs.back().size();
    return a;
  }

  const float* fptr(const char* n, uint64_t* count = nullptr) {
    if (st) return nullptr;
    if (!has(n)) {
      st = iofail(KW_ERR_INVALID, "input file lacks dataset %s", n);
      return nullptr;
    }
    fbufs.push_back(floats(n));
    if (count) *count = fbufs.back().size();
    return fbufs.back().data();
  }

  const uint64_t* uptr(const char* n, uint64_t* count) {
    if (st) return nullptr;
    int dt;
    uint64_t d[3];
    if (!has(n)) {
      st = iofail(KW_ERR_INVALID, "input file lacks dataset %s", n);
      return nullptr;
    }
    if ((st = kwio_info(f, n, &dt, d))) return nullptr;
    if (dt != KWIO_UINT64) {
      st = iofail(KW_ERR_INVALID, "%s must be uint64", n);
      return nullptr;
    }
    ubufs.emplace_back(d[0] * d[1] * d[2]);
    st = kwio_read(f, n, ubufs.back().data(), ubufs.back().size() * 8);
    *count = ubufs.back().size();
    return ubufs.back().data();
  }
};

struct OutSpec {
  int id;
  const char* name;
  int kind;  // 0 time series (Nt_rec, Ns), 1 per sensor (Ns), 2 full grid
};

}  // namespace

extern "C" kw_status kw_run_file(const char* input_path, const char* output_path, kw_run_options* o) {
  if (!input_path || !output_path || !o) return iofail(KW_ERR_INVALID, "null argument");
  const double t0 = now_s();
  kwio_file* in = nullptr;
  kw_status st = kwio_open(input_path, 0, &in);
  if (st) return st;
  char ftype[64] = {0};
  if (kwio_read_string_attribute(in, "/", "file_type", ftype, sizeof ftype) == KW_OK && strcmp(ftype, "input") != 0) {
    kwio_close(in);
    return iofail(KW_ERR_INVALID, "%s: file_type is '%s', expected 'input'", input_path, ftype);
  }
  Reader R{in};
  // flags this solver does not run (saveToDisk.m:44-64)
  const char* unsupported[] = {"elastic_flag", "nonuniform_grid_flag",
                               "sxx_source_flag", "syy_source_flag", "szz_source_flag", "sxy_source_flag",
                               "sxz_source_flag", "syz_source_flag"};
  for (const char* n : unsupported) {
    if (R.u64(n)) {
      kwio_close(in);
      return iofail(KW_ERR_UNSUPPORTED, "input sets %s, which this solver does not run", n);
    }
  }
  const uint64_t mask_type = R.u64("sensor_mask_type");
  if (mask_type > 1) {
    kwio_close(in);
    return iofail(KW_ERR_INVALID, "sensor_mask_type %llu", (unsigned long long)mask_type);
  }
  kw_sim_desc d;
  memset(&d, 0, sizeof d);
  d.abi_version = KW_ABI_VERSION;
  d.device = o->device;
  d.Nx = R.u64("Nx", 0, true);
  d.Ny = R.u64("Ny", 0, true);
  d.Nz = R.u64("Nz", 1);
  d.Nt = R.u64("Nt", 0, true);
  const bool three = d.Nz > 1;
  d.dx = R.f64("dx", 0, true);
  d.dy = R.f64("dy", 0, true);
  d.dz = three ? R.f64("dz", 0, true) : 0.0;
  d.dt = R.f64("dt", 0, true);
  d.c_ref = R.f64("c_ref", 0, true);
  d.pml_x_size = R.u64("pml_x_size");
  d.pml_y_size = R.u64("pml_y_size");
  d.pml_z_size = three ? R.u64("pml_z_size") : 0;
  d.pml_x_alpha = R.f64("pml_x_alpha");
  d.pml_y_alpha = R.f64("pml_y_alpha");
  d.pml_z_alpha = three ? R.f64("pml_z_alpha") : 0.0;
  d.c0 = R.real("c0", true);
  d.rho0 = R.real("rho0", true);
  d.rho0_sgx = R.real("rho0_sgx", true);
  d.rho0_sgy = R.real("rho0_sgy", true);
  if (three) d.rho0_sgz = R.real("rho0_sgz", true);
  d.nonlinear_flag = R.u64("nonlinear_flag");
  if (d.nonlinear_flag) d.BonA = R.real("BonA", true);
  d.absorbing_flag = R.u64("absorbing_flag");
  if (d.absorbing_flag) {
    d.alpha_coeff = R.real("alpha_coeff", true);
    d.alpha_power = R.f64("alpha_power", 0, true);
  }
  if (R.u64("p0_source_flag")) d.p0_source_input = R.fptr("p0_source_input");
  d.p_source_flag = R.u64("p_source_flag");
  if (d.p_source_flag) {
    d.p_source_mode = R.u64("p_source_mode");
    d.p_source_many = R.u64("p_source_many");
    d.p_source_index = R.uptr("p_source_index", &d.p_source_npoints);
    d.p_source_input = R.fptr("p_source_input");
  }
  d.ux_source_flag = R.u64("ux_source_flag");
  d.uy_source_flag = R.u64("uy_source_flag");
  d.uz_source_flag = three ? R.u64("uz_source_flag") : 0;
  if (d.ux_source_flag || d.uy_source_flag || d.uz_source_flag) {
    d.u_source_mode = R.u64("u_source_mode");
    d.u_source_many = R.u64("u_source_many");
    d.u_source_index = R.uptr("u_source_index", &d.u_source_npoints);
    if (d.ux_source_flag) d.ux_source_input = R.fptr("ux_source_input");
    if (d.uy_source_flag) d.uy_source_input = R.fptr("uy_source_input");
    if (d.uz_source_flag) d.uz_source_input = R.fptr("uz_source_input");
  }
  // transducer source (saveToDisk.m:290-293, 339-343): the signal, per-point
  // delay_mask, and the active points in u_source_index
  d.axisymmetric_flag = R.u64("axisymmetric_flag");
  d.transducer_source_flag = R.u64("transducer_source_flag");
  if (d.transducer_source_flag) {
    uint64_t np = 0;
    d.transducer_source_input = R.fptr("transducer_source_input", &d.transducer_source_length);
    d.delay_mask = R.uptr("delay_mask", &np);
    if (!d.u_source_index) d.u_source_index = R.uptr("u_source_index", &d.u_source_npoints);
    if (!R.st && np != d.u_source_npoints)
      R.st = iofail(KW_ERR_INVALID, "delay_mask holds %llu values for %llu source points", (unsigned long long)np,
                    (unsigned long long)d.u_source_npoints);
  }
  // Cuboid corners (sensor_mask_type 1, saveToDisk.m:357-378): columns
  // [x1 y1 z1 x2 y2 z2] of 1-based inclusive corners.  The solver records the
  // union of their points as a binary mask; the outputs are split per cuboid
  // when written (kspaceFirstOrder3DC.m:714-767 reads /p/<k> ...).
  struct Cuboid {
    uint64_t lo[3], hi[3];
    std::vector<size_t> pos;  // position of each point (x fastest) in the union
  };
  std::vector<Cuboid> cubs;
  std::vector<uint64_t> cub_union;
  if (mask_type == 1) {
    uint64_t n6 = 0;
    const uint64_t* c = R.uptr("sensor_mask_corners", &n6);
    if (!R.st && (!c || n6 == 0 || n6 % 6)) R.st = iofail(KW_ERR_INVALID, "sensor_mask_corners must hold 6 x Ncuboid values");
    const uint64_t Nd[3] = {d.Nx, d.Ny, d.Nz};
    for (uint64_t k = 0; !R.st && k < n6 / 6; ++k) {
      Cuboid q;
      for (int a = 0; a < 3; ++a) {
        q.lo[a] = c[6 * k + a];
        q.hi[a] = c[6 * k + 3 + a];
        if (q.lo[a] < 1 || q.lo[a] > q.hi[a] || q.hi[a] > Nd[a])
          R.st = iofail(KW_ERR_INVALID, "cuboid %llu: corners outside the grid", (unsigned long long)(k + 1));
      }
      cubs.push_back(std::move(q));
    }
    for (const Cuboid& q : cubs)
      for (uint64_t z = q.lo[2]; z <= q.hi[2]; ++z)
        for (uint64_t y = q.lo[1]; y <= q.hi[1]; ++y)
          for (uint64_t x = q.lo[0]; x <= q.hi[0]; ++x) cub_union.push_back(x + d.Nx * ((y - 1) + d.Ny * (z - 1)));
    std::sort(cub_union.begin(), cub_union.end());
    cub_union.erase(std::unique(cub_union.begin(), cub_union.end()), cub_union.end());
    for (Cuboid& q : cubs)
      for (uint64_t z = q.lo[2]; z <= q.hi[2]; ++z)
        for (uint64_t y = q.lo[1]; y <= q.hi[1]; ++y)
          for (uint64_t x = q.lo[0]; x <= q.hi[0]; ++x)
            q.pos.push_back((size_t)(std::lower_bound(cub_union.begin(), cub_union.end(),
                                                      x + d.Nx * ((y - 1) + d.Ny * (z - 1))) -
                                     cub_union.begin()));
    d.sensor_mask_index = cub_union.data();
    d.sensor_npoints = cub_union.size();
  } else {
    d.sensor_mask_index = R.uptr("sensor_mask_index", &d.sensor_npoints);
  }
  d.record_flags = o->record_flags ? o->record_flags : (uint64_t)KW_REC_P;
  d.record_start_index = o->record_start_index ? o->record_start_index : 1;
  kwio_close(in);
  if (R.st) return R.st;
  // source signal sizes must match their flags (writeFlags.m:101-116, saveToDisk.m:299-310)
  auto sig_ok = [&](const char* name, const float* ptr, uint64_t flag, uint64_t many, uint64_t npts) -> bool {
    if (!flag || !ptr) return true;
    for (const auto& b : R.fbufs)
      if (b.data() == ptr && b.size() != flag * (many ? npts : 1)) {
        iofail(KW_ERR_INVALID, "%s holds %zu values; expected %llu", name, b.size(),
               (unsigned long long)(flag * (many ? npts : 1)));
        return false;
      }
    return true;
  };
  if (!sig_ok("p_source_input", d.p_source_input, d.p_source_flag, d.p_source_many, d.p_source_npoints) ||
      !sig_ok("ux_source_input", d.ux_source_input, d.ux_source_flag, d.u_source_many, d.u_source_npoints) ||
      !sig_ok("uy_source_input", d.uy_source_input, d.uy_source_flag, d.u_source_many, d.u_source_npoints) ||
      !sig_ok("uz_source_input", d.uz_source_input, d.uz_source_flag, d.u_source_many, d.u_source_npoints))
    return KW_ERR_INVALID;
  if (d.p0_source_input) {
    for (const auto& b : R.fbufs)
      if (b.data() == d.p0_source_input && b.size() != d.Nx * d.Ny * d.Nz)
        return iofail(KW_ERR_INVALID, "p0_source_input must be a full grid field");
  }
  const double t1 = now_s();

  kw_sim* sim = nullptr;
  if ((st = kw_sim_create(&d, &sim))) return iofail(st, "%s", kw_last_error());
  if ((st = kw_sim_sync(sim))) {
    iofail(st, "%s", kw_last_error());
    kw_sim_destroy(sim);
    return st;
  }
  const double t2 = now_s();
  if (o->verbose) {
    printf("  %s grid %llu x %llu x %llu, Nt %llu\n", three ? "3-D" : "2-D", (unsigned long long)d.Nx,
           (unsigned long long)d.Ny, (unsigned long long)d.Nz, (unsigned long long)d.Nt);
    fflush(stdout);
  }
  // resume from a checkpoint of this input
  o->checkpointed = 0;
  const bool ckpt = o->checkpoint_file && o->checkpoint_file[0];
  if (ckpt && file_exists(o->checkpoint_file)) {
    if ((st = load_checkpoint(o->checkpoint_file, sim))) {
      kw_sim_destroy(sim);
      return st;
    }
    if (o->verbose)
      printf("  resumed from checkpoint %s at step %llu\n", o->checkpoint_file,
             (unsigned long long)kw_sim_steps_done(sim));
  }
  const uint64_t first = kw_sim_steps_done(sim);
  const uint64_t chunk = d.Nt >= 10 ? d.Nt / 10 : d.Nt;
  while (kw_sim_steps_done(sim) < d.Nt) {
    uint64_t n = chunk;
    if (ckpt && o->checkpoint_timesteps)
      n = std::min<uint64_t>(n, first + o->checkpoint_timesteps - kw_sim_steps_done(sim));
    if ((st = kw_sim_run(sim, n, nullptr)) || (st = kw_sim_sync(sim))) {
      iofail(st, "%s", kw_last_error());
      kw_sim_destroy(sim);
      return st;
    }
    const uint64_t done = kw_sim_steps_done(sim);
    if (ckpt && done < d.Nt &&
        ((o->checkpoint_timesteps && done - first >= o->checkpoint_timesteps) ||
         (o->checkpoint_interval > 0 && now_s() - t2 >= o->checkpoint_interval))) {
      st = save_checkpoint(o->checkpoint_file, sim);
      kw_sim_destroy(sim);
      if (st) return st;
      o->checkpointed = 1;
      o->t_load = t1 - t0;
      o->t_precompute = t2 - t1;
      o->t_loop = now_s() - t2;
      o->t_save = 0;
      if (o->verbose)
        printf("  checkpoint saved to %s at step %llu of %llu\n", o->checkpoint_file, (unsigned long long)done,
               (unsigned long long)d.Nt);
      return KW_OK;
    }
    if (o->verbose >= 2) {
      printf("  %5.1f%% (%llu of %llu steps)\n", 100.0 * kw_sim_steps_done(sim) / d.Nt,
             (unsigned long long)kw_sim_steps_done(sim), (unsigned long long)d.Nt);
      fflush(stdout);
    }
  }
  const double t3 = now_s();

  // output file (kspaceFirstOrder3DC.m:568-778)
  kwio_file* out = nullptr;
  if ((st = kwio_open(output_path, 1, &out))) {
    kw_sim_destroy(sim);
    return st;
  }
  const uint64_t one[3] = {1, 1, 1};
  const uint64_t ints[6] = {d.Nx, d.Ny, d.Nz, d.pml_x_size, d.pml_y_size, d.pml_z_size};
  const char* inames[6] = {"Nx", "Ny", "Nz", "pml_x_size", "pml_y_size", "pml_z_size"};
  for (int i = 0; i < 6 && !st; ++i) st = kwio_write(out, inames[i], KWIO_UINT64, one, &ints[i], 0);
  const uint64_t nrec = d.Nt - d.record_start_index + 1;
  const uint64_t ns = d.sensor_npoints;
  const OutSpec specs[] = {
      {KW_OUT_P, "p", 0},           {KW_OUT_P_MAX, "p_max", 1},         {KW_OUT_P_MIN, "p_min", 1},
      {KW_OUT_P_RMS, "p_rms", 1},   {KW_OUT_P_MAX_ALL, "p_max_all", 2}, {KW_OUT_P_MIN_ALL, "p_min_all", 2},
      {KW_OUT_P_FINAL, "p_final", 2}};
  std::vector<float> buf;
  auto emit = [&](int id, const std::string& name, int kind) {
    if (st) return;
    const size_t bytes = kw_sim_output_bytes(sim, id);
    if (!bytes) return;
    buf.resize(bytes / 4);
    if ((st = kw_sim_fetch(sim, id, buf.data(), bytes))) {
      iofail(st, "%s", kw_last_error());
      return;
    }
    if (!cubs.empty() && kind != 2) {
      // per cuboid: /name/<k> of (Nt_rec, cz, cy, cx) or (cz, cy, cx)
      std::vector<float> part;
      for (size_t k = 0; k < cubs.size() && !st; ++k) {
        const Cuboid& q = cubs[k];
        const uint64_t cd[3] = {q.hi[2] - q.lo[2] + 1, q.hi[1] - q.lo[1] + 1, q.hi[0] - q.lo[0] + 1};
        const size_t np = q.pos.size();
        const uint64_t steps = kind == 0 ? nrec : 1;
        part.resize(np * steps);
        for (uint64_t t = 0; t < steps; ++t)
          for (size_t j = 0; j < np; ++j) part[t * np + j] = buf[t * ns + q.pos[j]];
        const std::string dn = name + "/" + std::to_string(k + 1);
        if (kind == 0) {
          const uint64_t dims4[4] = {nrec, cd[0], cd[1], cd[2]};
          if (nrec) st = write_nd(out, dn.c_str(), KWIO_FLOAT32, 4, dims4, part.data(), o->compression_level);
        } else {
          st = write_nd(out, dn.c_str(), KWIO_FLOAT32, 3, cd, part.data(), o->compression_level);
        }
      }
      return;
    }
    uint64_t dims[3];
    if (kind == 0) {
      dims[0] = 1, dims[1] = nrec, dims[2] = ns;
    } else if (kind == 1) {
      dims[0] = 1, dims[1] = 1, dims[2] = ns;
    } else {
      dims[0] = d.Nz, dims[1] = d.Ny, dims[2] = d.Nx;
    }
    if (dims[0] * dims[1] * dims[2] == 0) return;
    st = kwio_write(out, name.c_str(), KWIO_FLOAT32, dims, buf.data(), o->compression_level);
  };
  for (const OutSpec& s : specs) emit(s.id, s.name, s.kind);
  const char* axes = "xyz";
  const int dim = three ? 3 : 2;
  const struct {
    int base;
    const char* suffix;
    int kind;
  } uspecs[] = {{KW_OUT_UX, "", 0},          {KW_OUT_UX_MAX, "_max", 1},         {KW_OUT_UX_MIN, "_min", 1},
                {KW_OUT_UX_RMS, "_rms", 1},  {KW_OUT_UX_MAX_ALL, "_max_all", 2}, {KW_OUT_UX_MIN_ALL, "_min_all", 2},
                {KW_OUT_UX_FINAL, "_final", 2},  {KW_OUT_UX_NON_STAGGERED, "_non_staggered", 0}};
  for (const auto& u : uspecs)
    for (int a = 0; a < dim; ++a) emit(u.base + a, std::string("u") + axes[a] + u.suffix, u.kind);
  if (!st) st = kwio_write_root_attributes(out, "output", "Output data created by kspaceFirstOrder-HIP");
  kwio_close(out);
  kw_sim_destroy(sim);
  if (!st && ckpt && file_exists(o->checkpoint_file)) std::remove(o->checkpoint_file);
  const double t4 = now_s();
  o->t_load = t1 - t0;
  o->t_precompute = t2 - t1;
  o->t_loop = t3 - t2;
  o->t_save = t4 - t3;
  return st;
}

extern "C" const char* kwio_last_error(void) { return g_ioerr.c_str(); }
