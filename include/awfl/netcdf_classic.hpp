// netcdf_classic.hpp -- a small writer of the netCDF classic file format,
// 64-bit-offset variant (CDF-2), for double variables.
//
// The reference writes its diagnostics with yakl::SimpleNetCDF over the
// netCDF library (Spatial_euler3d_cons_expl_cart_fv_Agrid.h:1704-1791), which
// this image does not have.  The file layout is the published classic format:
//   header = "CDF\x02" numrecs dim_list gatt_list var_list
//   dim_list = NC_DIMENSION n {name length}      (length 0: the record dimension)
//   var_list = NC_VARIABLE n {name ndims dimids vatt_list NC_DOUBLE vsize begin64}
//   data = fixed-size variables, then records (one slab of every record
//          variable per record, in variable order)
// All integers and values are big-endian; names are padded to 4 bytes.
// Files written here are read by the netCDF library, netCDF4-python and
// scipy.io.netcdf_file.
#pragma once
#include <sys/types.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace awfl {
namespace nc {

struct Dim {
  std::string name;
  int64_t len;          // 0 = record (unlimited) dimension
};
struct Var {
  std::string name;
  std::vector<int> dims;  // indices into the file's dimensions, slowest first
};

class ClassicFile {
 public:
  ClassicFile() = default;
  ClassicFile(const ClassicFile&) = delete;
  ClassicFile& operator=(const ClassicFile&) = delete;
  ~ClassicFile() { close(); }

  // New file (replaces an existing one): header, zero records.
  void create(const std::string& path, const std::vector<Dim>& dims, const std::vector<Var>& vars) {
    close();
    dims_ = dims;
    vars_ = vars;
    numrecs_ = 0;
    layout();
    f_ = std::fopen(path.c_str(), "w+b");
    if (!f_) throw std::runtime_error("netcdf: cannot create " + path);
    std::vector<unsigned char> h = header();
    put(0, h.data(), h.size());
    // fixed-size data region, zero-filled
    if (rec_begin_ > (int64_t)h.size()) {
      std::vector<unsigned char> z((size_t)(rec_begin_ - (int64_t)h.size()), 0);
      put((int64_t)h.size(), z.data(), z.size());
    }
  }

  // Existing file written by this class (or any classic file of doubles
  // without attributes): parse the header to append records.
  void open(const std::string& path) {
    close();
    f_ = std::fopen(path.c_str(), "r+b");
    if (!f_) throw std::runtime_error("netcdf: cannot open " + path);
    parse();
  }

  void close() {
    if (f_) std::fclose(f_);
    f_ = nullptr;
  }

  int64_t num_records() const { return numrecs_; }
  int var_id(const std::string& name) const {
    for (size_t v = 0; v < vars_.size(); v++)
      if (vars_[v].name == name) return (int)v;
    throw std::runtime_error("netcdf: no variable " + name);
  }
  int64_t dim_len(const std::string& name) const {
    for (const Dim& d : dims_)
      if (d.name == name) return d.len == 0 ? numrecs_ : d.len;
    throw std::runtime_error("netcdf: no dimension " + name);
  }

  // values of a fixed-size variable
  void write_fixed(const std::string& name, const double* data) {
    int v = var_id(name);
    if (is_record(v)) throw std::runtime_error("netcdf: " + name + " is a record variable");
    write_be(begin_[v], data, (size_t)(vsize_[v] / 8));
  }

  // one slab of a record variable at record index rec (rec <= num_records());
  // writing at num_records() appends a record
  void write_record(const std::string& name, int64_t rec, const double* data) {
    int v = var_id(name);
    if (!is_record(v)) throw std::runtime_error("netcdf: " + name + " is not a record variable");
    if (rec < 0 || rec > numrecs_) throw std::runtime_error("netcdf: record index out of range");
    write_be(begin_[v] + rec * recsize_, data, (size_t)(vsize_[v] / 8));
    if (rec == numrecs_) {
      numrecs_++;
      unsigned char b[4];
      be32(b, (uint32_t)numrecs_);
      put(4, b, 4);
    }
  }

 private:
  static void be32(unsigned char* p, uint32_t v) {
    p[0] = (unsigned char)(v >> 24); p[1] = (unsigned char)(v >> 16);
    p[2] = (unsigned char)(v >> 8);  p[3] = (unsigned char)v;
  }
  static void be64(unsigned char* p, uint64_t v) {
    be32(p, (uint32_t)(v >> 32));
    be32(p + 4, (uint32_t)v);
  }
  static uint32_t rd32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
  }
  static uint64_t rd64(const unsigned char* p) { return ((uint64_t)rd32(p) << 32) | rd32(p + 4); }
  static int64_t pad4(int64_t n) { return (n + 3) & ~(int64_t)3; }

  bool is_record(int v) const { return !vars_[v].dims.empty() && dims_[vars_[v].dims[0]].len == 0; }

  void put(int64_t off, const void* p, size_t n) {
    if (fseeko(f_, (off_t)off, SEEK_SET) != 0 || std::fwrite(p, 1, n, f_) != n)
      throw std::runtime_error("netcdf: write failed");
  }
  void write_be(int64_t off, const double* data, size_t n) {
    std::vector<unsigned char> buf(n * 8);
    for (size_t i = 0; i < n; i++) {
      uint64_t u;
      std::memcpy(&u, &data[i], 8);
      be64(&buf[i * 8], u);
    }
    put(off, buf.data(), buf.size());
  }

  void name_bytes(std::vector<unsigned char>& h, const std::string& s) const {
    unsigned char b[4];
    be32(b, (uint32_t)s.size());
    h.insert(h.end(), b, b + 4);
    h.insert(h.end(), s.begin(), s.end());
    while (h.size() % 4) h.push_back(0);
  }

  int64_t header_size() const {
    int64_t n = 4 + 4 + 8;                                   // magic, numrecs, dim tag + count
    for (const Dim& d : dims_) n += 4 + pad4((int64_t)d.name.size()) + 4;
    n += 8;                                                  // gatt_list ABSENT
    n += 8;                                                  // var tag + count
    for (const Var& v : vars_) n += 4 + pad4((int64_t)v.name.size()) + 4 + 4 * (int64_t)v.dims.size() + 8 + 4 + 4 + 8;
    return n;
  }

  // begin offsets and sizes: fixed variables after the header, then records
  void layout() {
    begin_.assign(vars_.size(), 0);
    vsize_.assign(vars_.size(), 0);
    int64_t off = header_size();
    for (size_t v = 0; v < vars_.size(); v++) {
      int64_t n = 8;
      for (size_t d = is_record((int)v) ? 1 : 0; d < vars_[v].dims.size(); d++) n *= dims_[vars_[v].dims[d]].len;
      vsize_[v] = pad4(n);
      if (!is_record((int)v)) {
        begin_[v] = off;
        off += vsize_[v];
      }
    }
    rec_begin_ = off;
    recsize_ = 0;
    for (size_t v = 0; v < vars_.size(); v++)
      if (is_record((int)v)) {
        begin_[v] = rec_begin_ + recsize_;
        recsize_ += vsize_[v];
      }
  }

  std::vector<unsigned char> header() const {
    std::vector<unsigned char> h = {'C', 'D', 'F', 2};
    unsigned char b[8];
    be32(b, (uint32_t)numrecs_);
    h.insert(h.end(), b, b + 4);
    be32(b, 0x0A);
    h.insert(h.end(), b, b + 4);
    be32(b, (uint32_t)dims_.size());
    h.insert(h.end(), b, b + 4);
    for (const Dim& d : dims_) {
      name_bytes(h, d.name);
      be32(b, (uint32_t)d.len);
      h.insert(h.end(), b, b + 4);
    }
    for (int i = 0; i < 8; i++) h.push_back(0);            // no global attributes
    be32(b, 0x0B);
    h.insert(h.end(), b, b + 4);
    be32(b, (uint32_t)vars_.size());
    h.insert(h.end(), b, b + 4);
    for (size_t v = 0; v < vars_.size(); v++) {
      name_bytes(h, vars_[v].name);
      be32(b, (uint32_t)vars_[v].dims.size());
      h.insert(h.end(), b, b + 4);
      for (int d : vars_[v].dims) {
        be32(b, (uint32_t)d);
        h.insert(h.end(), b, b + 4);
      }
      for (int i = 0; i < 8; i++) h.push_back(0);          // no variable attributes
      be32(b, 6);                                          // NC_DOUBLE
      h.insert(h.end(), b, b + 4);
      be32(b, (uint32_t)vsize_[v]);
      h.insert(h.end(), b, b + 4);
      be64(b, (uint64_t)begin_[v]);
      h.insert(h.end(), b, b + 8);
    }
    return h;
  }

  void parse() {
    std::vector<unsigned char> h(1 << 16);
    size_t got = std::fread(h.data(), 1, h.size(), f_);
    size_t p = 0;
    auto need = [&](size_t n) {
      if (p + n > got) throw std::runtime_error("netcdf: truncated header");
    };
    auto u32 = [&]() { need(4); uint32_t v = rd32(&h[p]); p += 4; return v; };
    auto name = [&]() {
      uint32_t n = u32();
      need(n);
      std::string s(reinterpret_cast<const char*>(&h[p]), n);
      p += (size_t)pad4(n);
      return s;
    };
    need(4);
    if (h[0] != 'C' || h[1] != 'D' || h[2] != 'F' || h[3] != 2) throw std::runtime_error("netcdf: not CDF-2");
    p = 4;
    numrecs_ = u32();
    dims_.clear();
    vars_.clear();
    uint32_t tag = u32(), n = u32();
    if (tag == 0x0A)
      for (uint32_t i = 0; i < n; i++) {
        Dim d;
        d.name = name();
        d.len = u32();
        dims_.push_back(d);
      }
    tag = u32();
    n = u32();
    if (tag == 0x0C && n) throw std::runtime_error("netcdf: global attributes not supported");
    tag = u32();
    n = u32();
    begin_.clear();
    vsize_.clear();
    if (tag == 0x0B)
      for (uint32_t i = 0; i < n; i++) {
        Var v;
        v.name = name();
        uint32_t nd = u32();
        for (uint32_t d = 0; d < nd; d++) v.dims.push_back((int)u32());
        uint32_t at = u32(), na = u32();
        if (at == 0x0C && na) throw std::runtime_error("netcdf: variable attributes not supported");
        if (u32() != 6) throw std::runtime_error("netcdf: only double variables supported");
        vsize_.push_back(u32());
        need(8);
        begin_.push_back((int64_t)rd64(&h[p]));
        p += 8;
        vars_.push_back(v);
      }
    recsize_ = 0;
    rec_begin_ = -1;
    for (size_t v = 0; v < vars_.size(); v++)
      if (is_record((int)v)) {
        recsize_ += vsize_[v];
        if (rec_begin_ < 0 || begin_[v] < rec_begin_) rec_begin_ = begin_[v];
      }
  }

  std::FILE* f_ = nullptr;
  std::vector<Dim> dims_;
  std::vector<Var> vars_;
  std::vector<int64_t> begin_, vsize_;
  int64_t rec_begin_ = 0, recsize_ = 0, numrecs_ = 0;
};

}  // namespace nc
}  // namespace awfl
