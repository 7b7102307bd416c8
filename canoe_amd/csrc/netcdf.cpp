// netcdf.cpp -- NetCDF dumps of cell-centred fields in the layout of the
// reference's netcdf output (src/outputs/netcdf.cpp:100-420), written in the
// NetCDF classic 64-bit-offset format (CDF-2, Unidata's published file
// format) without the NetCDF library.
//
// Layout: dimensions time (the record dimension), x1, x1f, x2, x2f, x3, x3f
// (a face dimension only for a direction with more than one cell);
// coordinate variables x1 (axis Z), x1f, x2 (Y), x2f, x3 (X), x3f with
// units, long_name and domain_decomposition; a record variable time; one
// float record variable (time, x1, x2, x3) per field with units and
// long_name; the global attribute NumFilesInSet.  Values are stored as
// big-endian IEEE floats, x3 fastest, as the reference's writer orders them
// (i, j, k loops, :393-420).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/canoe_gpu.h"

namespace {

enum : uint32_t {
  NC_DIMENSION = 10,
  NC_VARIABLE = 11,
  NC_ATTRIBUTE = 12,
  NC_CHAR = 2,
  NC_INT = 4,
  NC_FLOAT = 5,
};

struct Buf {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void u64(uint64_t v) {
    for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void f32(float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    u32(v);
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string& s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
};

struct Att {
  std::string name;
  uint32_t type;
  std::string text;
  std::vector<int32_t> ints;
};

struct Var {
  std::string name;
  std::vector<int> dims;
  std::vector<Att> atts;
  bool record;
  uint64_t vsize;  // bytes of one record (record variables) or of the data
  uint64_t begin;
};

Att text_att(const std::string& n, const std::string& v) {
  return Att{n, NC_CHAR, v, {}};
}
Att int_att(const std::string& n, std::vector<int32_t> v) {
  return Att{n, NC_INT, "", v};
}

void put_atts(Buf& h, const std::vector<Att>& atts) {
  if (atts.empty()) {
    h.u32(0);
    h.u32(0);
    return;
  }
  h.u32(NC_ATTRIBUTE);
  h.u32((uint32_t)atts.size());
  for (const Att& a : atts) {
    h.name(a.name);
    h.u32(a.type);
    if (a.type == NC_CHAR) {
      h.u32((uint32_t)a.text.size());
      h.b.insert(h.b.end(), a.text.begin(), a.text.end());
      h.pad();
    } else {
      h.u32((uint32_t)a.ints.size());
      for (int32_t v : a.ints) h.u32((uint32_t)v);
    }
  }
}

struct Layout {
  std::vector<std::pair<std::string, uint32_t>> dims;  // name, length (0: record)
  std::vector<Var> vars;
  std::vector<Att> gatts;
  uint64_t recsize = 0;
  std::vector<unsigned char> header;
};

// The header for numrecs records, with every begin offset resolved.
void build_header(Layout& L, uint32_t numrecs) {
  for (int pass = 0; pass < 2; ++pass) {
    Buf h;
    h.b.push_back('C');
    h.b.push_back('D');
    h.b.push_back('F');
    h.b.push_back(2);
    h.u32(numrecs);
    h.u32(NC_DIMENSION);
    h.u32((uint32_t)L.dims.size());
    for (auto& d : L.dims) {
      h.name(d.first);
      h.u32(d.second);
    }
    put_atts(h, L.gatts);
    h.u32(NC_VARIABLE);
    h.u32((uint32_t)L.vars.size());
    for (const Var& v : L.vars) {
      h.name(v.name);
      h.u32((uint32_t)v.dims.size());
      for (int d : v.dims) h.u32((uint32_t)d);
      put_atts(h, v.atts);
      h.u32(NC_FLOAT);
      h.u32((uint32_t)v.vsize);
      h.u64(v.begin);
    }
    // the header's size does not depend on the offsets: place the data
    uint64_t off = h.b.size();
    for (Var& v : L.vars)
      if (!v.record) {
        v.begin = off;
        off += v.vsize;
      }
    L.recsize = 0;
    for (Var& v : L.vars)
      if (v.record) {
        v.begin = off + L.recsize;
        L.recsize += v.vsize;
      }
    L.header = h.b;
  }
}

Layout make_layout(const cg_block& b, int nfield, const char* const* names,
                   const char* const* units, const char* const* long_names) {
  Layout L;
  const int n[3] = {b.nx1, b.nx2, b.nx3};
  const char* ax[3] = {"Z", "Y", "X"};
  const char* cn[3] = {"Z-coordinate at cell center", "Y-coordinate at cell center",
                       "X-coordinate at cell center"};
  const char* fn[3] = {"Z-coordinate at cell face", "Y-coordinate at cell face",
                       "X-coordinate at cell face"};
  L.dims.push_back({"time", 0});
  int dc[3], df[3];
  for (int d = 0; d < 3; ++d) {
    dc[d] = (int)L.dims.size();
    L.dims.push_back({"x" + std::to_string(d + 1), (uint32_t)n[d]});
    df[d] = -1;
    if (n[d] > 1) {
      df[d] = (int)L.dims.size();
      L.dims.push_back({"x" + std::to_string(d + 1) + "f", (uint32_t)(n[d] + 1)});
    }
  }
  L.gatts.push_back(int_att("NumFilesInSet", {1}));
  L.vars.push_back(Var{"time", {0}, {text_att("axis", "T"), text_att("units", "s"),
                                     text_att("long_name", "time")},
                       true, 4, 0});
  for (int d = 0; d < 3; ++d) {
    // one block: the whole mesh, cells 1..n of n
    std::vector<int32_t> pos = {1, n[d], 1, n[d]};
    L.vars.push_back(Var{"x" + std::to_string(d + 1), {dc[d]},
                         {text_att("axis", ax[d]), text_att("units", "m"),
                          text_att("long_name", cn[d]), int_att("domain_decomposition", pos)},
                         false, (uint64_t)n[d] * 4, 0});
    if (df[d] >= 0) {
      std::vector<int32_t> posf = {1, n[d] + 1, 1, n[d] + 1};
      L.vars.push_back(Var{"x" + std::to_string(d + 1) + "f", {df[d]},
                           {text_att("units", "m"), text_att("long_name", fn[d]),
                            int_att("domain_decomposition", posf)},
                           false, (uint64_t)(n[d] + 1) * 4, 0});
    }
  }
  const uint64_t cells = (uint64_t)n[0] * n[1] * n[2];
  for (int f = 0; f < nfield; ++f) {
    std::vector<Att> atts;
    if (units && units[f] && units[f][0]) atts.push_back(text_att("units", units[f]));
    if (long_names && long_names[f] && long_names[f][0])
      atts.push_back(text_att("long_name", long_names[f]));
    L.vars.push_back(Var{names[f], {0, dc[0], dc[1], dc[2]}, atts, true, cells * 4, 0});
  }
  build_header(L, 0);
  return L;
}

}  // namespace

extern "C" int cg_netcdf_write(const char* path, const cg_block* blk, double time,
                               int nfield, const char* const* names,
                               const char* const* units, const char* const* long_names,
                               const double* const* fields, int append) {
  if (!path || !blk || nfield < 0 || (nfield > 0 && (!names || !fields)))
    return CG_ERR_ARG;
  if (blk->nx1 < 1 || blk->nx2 < 1 || blk->nx3 < 1) return CG_ERR_ARG;
  for (int f = 0; f < nfield; ++f)
    if (!names[f] || !names[f][0] || !fields[f]) return CG_ERR_ARG;
  const cg_block& b = *blk;
  Layout L = make_layout(b, nfield, names, units, long_names);
  const int NG = CG_NGHOST;
  const int nc1 = b.nx1 + 2 * NG;
  const int nc2 = b.nx2 > 1 ? b.nx2 + 2 * NG : 1;
  const int js = b.nx2 > 1 ? NG : 0, ks = b.nx3 > 1 ? NG : 0;

  uint32_t nrec = 0;
  FILE* fp = nullptr;
  if (append) {
    fp = std::fopen(path, "r+b");
    if (!fp) return CG_ERR_IO;
    unsigned char hd[8];
    if (std::fread(hd, 1, 8, fp) != 8 || std::memcmp(hd, "CDF\x02", 4) != 0) {
      std::fclose(fp);
      return CG_ERR_IO;
    }
    nrec = ((uint32_t)hd[4] << 24) | ((uint32_t)hd[5] << 16) | ((uint32_t)hd[6] << 8) | hd[7];
    // the same block and fields give the same header
    std::vector<unsigned char> cur(L.header.size());
    std::fseek(fp, 0, SEEK_SET);
    if (std::fread(cur.data(), 1, cur.size(), fp) != cur.size() ||
        std::memcmp(cur.data() + 8, L.header.data() + 8, cur.size() - 8) != 0) {
      std::fclose(fp);
      return CG_ERR_ARG;
    }
  } else {
    fp = std::fopen(path, "wb");
    if (!fp) return CG_ERR_IO;
    std::fwrite(L.header.data(), 1, L.header.size(), fp);
    // coordinate variables (x1v = x1min + (i + 1/2) dx1 on the uniform mesh)
    const double mins[3] = {b.x1min, b.x2min, b.x3min};
    const double dxs[3] = {b.dx1, b.dx2, b.dx3};
    const int n[3] = {b.nx1, b.nx2, b.nx3};
    for (int d = 0; d < 3; ++d) {
      Buf c;
      for (int i = 0; i < n[d]; ++i) c.f32((float)(mins[d] + (i + 0.5) * dxs[d]));
      std::fwrite(c.b.data(), 1, c.b.size(), fp);
      if (n[d] > 1) {
        Buf f;
        for (int i = 0; i <= n[d]; ++i) f.f32((float)(mins[d] + i * dxs[d]));
        std::fwrite(f.b.data(), 1, f.b.size(), fp);
      }
    }
  }
  // one record: time, then every field (interior cells, x3 fastest)
  uint64_t rec0 = 0;
  for (const Var& v : L.vars)
    if (v.record) {
      rec0 = v.begin;
      break;
    }
  Buf r;
  r.f32((float)time);
  for (int f = 0; f < nfield; ++f) {
    const double* a = fields[f];
    for (int i = 0; i < b.nx1; ++i)
      for (int j = 0; j < b.nx2; ++j)
        for (int k = 0; k < b.nx3; ++k) {
          size_t c = ((size_t)(ks + k) * nc2 + (js + j)) * nc1 + (NG + i);
          r.f32((float)a[c]);
        }
  }
  bool ok = std::fseek(fp, (long)(rec0 + (uint64_t)nrec * L.recsize), SEEK_SET) == 0 &&
            std::fwrite(r.b.data(), 1, r.b.size(), fp) == r.b.size();
  // numrecs
  Buf nr;
  nr.u32(nrec + 1);
  ok = ok && std::fseek(fp, 4, SEEK_SET) == 0 && std::fwrite(nr.b.data(), 1, 4, fp) == 4;
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? CG_OK : CG_ERR_IO;
}
