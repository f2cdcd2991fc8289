// ncread.cpp -- reader for LAMMPS NetCDF trajectories in the classic NetCDF
// formats (CDF-1 classic, CDF-2 64-bit offset, CDF-5 64-bit data), written
// from the published NetCDF classic format specification; host code only.
//
// It replaces the NetCDF-Fortran calls of read_nc_cfg (src/core/netcdf.f90:
// 173-550) for the variables the structural path reads: coordinates
// (frame, atom, spatial), type (frame, atom), q (frame, atom), cell_origin and
// cell_lengths (frame, cell_spatial).  netCDF-4 files (HDF5 containers) are refused with a
// message; `nccopy -k cdf5 in.nc out.nc` converts them.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/trj_hip.h"

namespace {

constexpr int32_t kTypeMax = 1 << 20;  // largest |type| accepted from a file

enum NcType { kByte = 1, kChar = 2, kShort = 3, kInt = 4, kFloat = 5, kDouble = 6, kUByte = 7, kUShort = 8,
              kUInt = 9, kInt64 = 10, kUInt64 = 11 };

int type_size(int t) {
  switch (t) {
    case kByte: case kChar: case kUByte: return 1;
    case kShort: case kUShort: return 2;
    case kInt: case kFloat: case kUInt: return 4;
    case kDouble: case kInt64: case kUInt64: return 8;
    default: return 0;
  }
}

struct NcVar {
  std::string name;
  std::vector<int64_t> dims;  // dimension ids
  int type = 0;
  int64_t vsize = 0, begin = 0;
  bool record = false;
  int64_t inner = 1;  // elements per record (or in total, for a fixed variable)
};

inline uint8_t from_be(uint8_t x) { return x; }
inline uint16_t from_be(uint16_t x) { return __builtin_bswap16(x); }
inline uint32_t from_be(uint32_t x) { return __builtin_bswap32(x); }
inline uint64_t from_be(uint64_t x) { return __builtin_bswap64(x); }

// n big-endian values of type T (bit pattern U) -> double
template <typename T, typename U>
void convert(const unsigned char* p, int64_t n, double* out) {
  static_assert(sizeof(T) == sizeof(U), "same width");
  for (int64_t i = 0; i < n; ++i) {
    U u;
    std::memcpy(&u, p + i * sizeof(U), sizeof(U));
    u = from_be(u);
    T t;
    std::memcpy(&t, &u, sizeof(T));
    out[i] = static_cast<double>(t);
  }
}

struct Reader {
  FILE* f = nullptr;
  int version = 0;
  std::vector<int64_t> dimlen;
  int64_t recdim = -1, numrecs = 0, recsize = 0;
  std::vector<NcVar> vars;
  std::string err;
  std::vector<unsigned char> buf;  // raw bytes of the last read
  // big-endian primitive reads
  bool bytes(void* p, size_t n) { return std::fread(p, 1, n, f) == n; }
  bool u32(uint32_t& v) {
    unsigned char b[4];
    if (!bytes(b, 4)) return false;
    v = (uint32_t(b[0]) << 24) | (uint32_t(b[1]) << 16) | (uint32_t(b[2]) << 8) | b[3];
    return true;
  }
  bool u64(uint64_t& v) {
    uint32_t hi, lo;
    if (!u32(hi) || !u32(lo)) return false;
    v = (uint64_t(hi) << 32) | lo;
    return true;
  }
  // NON_NEG: 4 bytes in CDF-1/2, 8 in CDF-5
  bool nonneg(int64_t& v) {
    if (version == 5) {
      uint64_t x;
      if (!u64(x)) return false;
      v = static_cast<int64_t>(x);
    } else {
      uint32_t x;
      if (!u32(x)) return false;
      v = static_cast<int64_t>(x);
    }
    return v >= 0;
  }
  bool offset(int64_t& v) {
    if (version == 1) {
      uint32_t x;
      if (!u32(x)) return false;
      v = x;
    } else {
      uint64_t x;
      if (!u64(x)) return false;
      v = static_cast<int64_t>(x);
    }
    return v >= 0;
  }
  bool name(std::string& s) {
    int64_t n;
    if (!nonneg(n) || n > (1 << 20)) return false;
    s.assign(static_cast<size_t>(n), '\0');
    if (n && !bytes(&s[0], static_cast<size_t>(n))) return false;
    return skip_pad(n);
  }
  bool skip_pad(int64_t n) {
    const int64_t pad = (4 - n % 4) % 4;
    char b[4];
    return pad == 0 || bytes(b, static_cast<size_t>(pad));
  }
  bool att_list() {
    uint32_t tag;
    int64_t n;
    if (!u32(tag) || !nonneg(n)) return false;
    if (tag == 0) return n == 0;
    if (tag != 0x0C) return false;
    for (int64_t i = 0; i < n; ++i) {
      std::string nm;
      uint32_t t;
      int64_t ne;
      if (!name(nm) || !u32(t) || !nonneg(ne)) return false;
      const int ts = type_size(static_cast<int>(t));
      if (!ts) return false;
      const int64_t nb = ne * ts;
      if (fseeko(f, static_cast<off_t>(nb), SEEK_CUR) != 0 || !skip_pad(nb)) return false;
    }
    return true;
  }
  bool open(const char* path) {
    f = std::fopen(path, "rb");
    if (!f) return fail(std::string("cannot open ") + path);
    unsigned char magic[4];
    if (!bytes(magic, 4)) return fail("file shorter than a NetCDF header");
    if (magic[0] == 0x89 && magic[1] == 'H' && magic[2] == 'D' && magic[3] == 'F')
      return fail("netCDF-4 (HDF5) files are not supported; convert with `nccopy -k cdf5 in.nc out.nc`");
    if (magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F' ||
        (magic[3] != 1 && magic[3] != 2 && magic[3] != 5))
      return fail("not a NetCDF classic file (CDF-1, CDF-2 or CDF-5)");
    version = magic[3];
    // numrecs: NON_NEG or STREAMING (all bits set)
    bool streaming = false;
    if (version == 5) {
      uint64_t x;
      if (!u64(x)) return fail("truncated header");
      streaming = x == ~uint64_t(0);
      numrecs = streaming ? 0 : static_cast<int64_t>(x);
    } else {
      uint32_t x;
      if (!u32(x)) return fail("truncated header");
      streaming = x == 0xFFFFFFFFu;
      numrecs = streaming ? 0 : x;
    }
    // dim_list
    uint32_t tag;
    int64_t n;
    if (!u32(tag) || !nonneg(n)) return fail("bad dimension list");
    if (tag != 0 && tag != 0x0A) return fail("bad dimension list tag");
    for (int64_t i = 0; i < n; ++i) {
      std::string nm;
      int64_t len;
      if (!name(nm) || !nonneg(len)) return fail("bad dimension");
      if (len == 0) recdim = i;
      dimlen.push_back(len);
    }
    if (!att_list()) return fail("bad global attribute list");
    if (!u32(tag) || !nonneg(n)) return fail("bad variable list");
    if (tag != 0 && tag != 0x0B) return fail("bad variable list tag");
    for (int64_t i = 0; i < n; ++i) {
      NcVar v;
      int64_t nd;
      if (!name(v.name) || !nonneg(nd) || nd > 64) return fail("bad variable");
      for (int64_t d = 0; d < nd; ++d) {
        int64_t id;
        if (!nonneg(id) || id >= static_cast<int64_t>(dimlen.size())) return fail("bad variable dimension");
        v.dims.push_back(id);
      }
      if (!att_list()) return fail("bad variable attribute list");
      uint32_t t;
      if (!u32(t) || !type_size(static_cast<int>(t))) return fail("bad variable type");
      v.type = static_cast<int>(t);
      if (!nonneg(v.vsize) || !offset(v.begin)) return fail("bad variable size or offset");
      v.record = !v.dims.empty() && v.dims[0] == recdim;
      for (size_t d = v.record ? 1 : 0; d < v.dims.size(); ++d) v.inner *= dimlen[v.dims[d]];
      vars.push_back(v);
    }
    // record size: the padded sizes of all record variables; a single record
    // variable is not padded
    int nrec = 0;
    int64_t single = 0;
    for (const NcVar& v : vars)
      if (v.record) {
        ++nrec;
        single = v.inner * type_size(v.type);
        recsize += (single + 3) / 4 * 4;
      }
    if (nrec == 1) recsize = single;
    if (streaming && recsize > 0) {
      int64_t first = INT64_MAX;
      for (const NcVar& v : vars)
        if (v.record) first = std::min(first, v.begin);
      if (fseeko(f, 0, SEEK_END) != 0) return fail("cannot size the file");
      const int64_t size = static_cast<int64_t>(ftello(f));
      numrecs = size > first ? (size - first) / recsize : 0;
    }
    return true;
  }
  bool fail(const std::string& m) {
    err = m;
    return false;
  }
  const NcVar* var(const char* nm) const {
    for (const NcVar& v : vars)
      if (v.name == nm) return &v;
    return nullptr;
  }
  // Read `count` elements starting at element `first` of record `rec` (or of
  // a fixed variable when rec < 0), converted to double.
  bool read(const NcVar& v, int64_t rec, int64_t first, int64_t count, double* out) {
    const int ts = type_size(v.type);
    int64_t off = v.begin + first * ts;
    if (v.record) off += rec * recsize;
    if (fseeko(f, static_cast<off_t>(off), SEEK_SET) != 0) return fail("seek failed");
    buf.resize(static_cast<size_t>(count * ts));
    if (count && !bytes(buf.data(), buf.size())) return fail("truncated variable " + v.name);
    const unsigned char* p = buf.data();
    switch (v.type) {
      case kByte: convert<int8_t, uint8_t>(p, count, out); break;
      case kUByte: case kChar: convert<uint8_t, uint8_t>(p, count, out); break;
      case kShort: convert<int16_t, uint16_t>(p, count, out); break;
      case kUShort: convert<uint16_t, uint16_t>(p, count, out); break;
      case kInt: convert<int32_t, uint32_t>(p, count, out); break;
      case kUInt: convert<uint32_t, uint32_t>(p, count, out); break;
      case kInt64: convert<int64_t, uint64_t>(p, count, out); break;
      case kUInt64: convert<uint64_t, uint64_t>(p, count, out); break;
      case kFloat: convert<float, uint32_t>(p, count, out); break;
      default: convert<double, uint64_t>(p, count, out); break;
    }
    return true;
  }
};

}  // namespace

struct trj_nc {
  Reader r;
  const NcVar* coords = nullptr;
  const NcVar* type = nullptr;
  const NcVar* origin = nullptr;
  const NcVar* lengths = nullptr;
  const NcVar* time = nullptr;
  const NcVar* charge = nullptr;
  int64_t natoms = 0;
  // LAMMPS type id -> species 1..nsp, fixed by the first trj_nc_frame_ref call
  std::vector<int32_t> species_of;  // indexed by type id; 0 = not selected
  std::vector<int32_t> selected;    // the sp_types that fixed the map
  std::vector<double> xyz, tyd, qd;  // frame_ref scratch, kept between frames
  std::vector<int32_t> ty;
  bool keep_nonperiodic = false;  // hand non-periodic coordinates on as read
  std::string err;
};

extern "C" {

static int trj_nc_open_impl(trj_nc** out, const char* path, char* errbuf, int32_t errlen) {
  if (!out || !path) return TRJ_EINVAL;
  *out = nullptr;
  trj_nc* h = new trj_nc();
  auto give_up = [&](const std::string& m) {
    if (errbuf && errlen > 0) {
      std::snprintf(errbuf, static_cast<size_t>(errlen), "%s", m.c_str());
    }
    if (h->r.f) std::fclose(h->r.f);
    delete h;
    return TRJ_EINVAL;
  };
  if (!h->r.open(path)) return give_up(h->r.err);
  h->coords = h->r.var("coordinates");
  h->type = h->r.var("type");
  h->origin = h->r.var("cell_origin");
  h->lengths = h->r.var("cell_lengths");
  if (!h->coords || !h->coords->record || h->coords->dims.size() != 3 || h->r.dimlen[h->coords->dims[2]] != 3)
    return give_up("no coordinates(frame, atom, spatial=3) variable");
  if (h->coords->type != kFloat && h->coords->type != kDouble)
    return give_up("coordinates are neither float nor double");
  h->natoms = h->r.dimlen[h->coords->dims[1]];
  if (h->type && (!h->type->record || h->type->inner != h->natoms)) h->type = nullptr;
  if (h->origin && (!h->origin->record || h->origin->inner != 3)) h->origin = nullptr;
  if (h->lengths && (!h->lengths->record || h->lengths->inner != 3)) h->lengths = nullptr;
  h->time = h->r.var("time");
  if (h->time && (!h->time->record || h->time->inner != 1)) h->time = nullptr;
  h->charge = h->r.var("q");
  if (h->charge && (!h->charge->record || h->charge->inner != h->natoms)) h->charge = nullptr;
  *out = h;
  return TRJ_OK;
}

void trj_nc_close(trj_nc* h) {
  if (!h) return;
  if (h->r.f) std::fclose(h->r.f);
  delete h;
}

const char* trj_nc_last_error(const trj_nc* h) { return h ? h->err.c_str() : "null handle"; }

int trj_nc_keep_nonperiodic(trj_nc* h, int32_t keep) {
  if (!h) return TRJ_EINVAL;
  h->keep_nonperiodic = keep != 0;
  return TRJ_OK;
}

int trj_nc_info(const trj_nc* h, trj_nc_file_info* info) {
  if (!h || !info) return TRJ_EINVAL;
  info->nframes = h->r.numrecs;
  info->natoms = h->natoms;
  info->version = h->r.version;
  info->coord_dtype = h->coords->type == kDouble ? TRJ_F64 : TRJ_F32;
  info->has_type = h->type != nullptr;
  info->has_cell_origin = h->origin != nullptr;
  info->has_cell_lengths = h->lengths != nullptr;
  info->has_time = h->time != nullptr;
  info->has_charge = h->charge != nullptr;
  return TRJ_OK;
}

static int trj_nc_read_impl(trj_nc* h, int64_t frame, double* pos, int32_t* type, double origin[3], double lengths[3]) {
  if (!h) return TRJ_EINVAL;
  if (frame < 0 || frame >= h->r.numrecs) {
    h->err = "frame " + std::to_string(frame) + " outside 0.." + std::to_string(h->r.numrecs - 1);
    return TRJ_ERANGE;
  }
  if (pos && !h->r.read(*h->coords, frame, 0, 3 * h->natoms, pos)) return (h->err = h->r.err, TRJ_EINVAL);
  if (type) {
    if (!h->type) return (h->err = "the file has no type variable", TRJ_EINVAL);
    std::vector<double>& t = h->tyd;
    t.resize(static_cast<size_t>(h->natoms));
    if (!h->r.read(*h->type, frame, 0, h->natoms, t.data())) return (h->err = h->r.err, TRJ_EINVAL);
    // LAMMPS types are small positive integers; a value outside this range
    // (or NaN) would be undefined to convert and would size the species map.
    for (int64_t i = 0; i < h->natoms; ++i) {
      if (!(std::fabs(t[i]) <= kTypeMax)) {
        h->err = "type value of atom " + std::to_string(i) + " in frame " + std::to_string(frame) +
                 " is not a number in -2^20..2^20";
        return TRJ_EINVAL;
      }
      type[i] = static_cast<int32_t>(t[i]);
    }
  }
  double o[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  if (h->origin && !h->r.read(*h->origin, frame, 0, 3, o)) return (h->err = h->r.err, TRJ_EINVAL);
  if (h->lengths && !h->r.read(*h->lengths, frame, 0, 3, l)) return (h->err = h->r.err, TRJ_EINVAL);
  if (origin) std::memcpy(origin, o, sizeof(o));
  if (lengths) std::memcpy(lengths, l, sizeof(l));
  return TRJ_OK;
}

int trj_nc_time(trj_nc* h, int64_t frame, double* t) {
  if (!h || !t) return TRJ_EINVAL;
  if (!h->time) return (h->err = "the file has no time(frame) variable", TRJ_EINVAL);
  if (frame < 0 || frame >= h->r.numrecs) {
    h->err = "frame " + std::to_string(frame) + " outside 0.." + std::to_string(h->r.numrecs - 1);
    return TRJ_ERANGE;
  }
  if (!h->r.read(*h->time, frame, 0, 1, t)) return (h->err = h->r.err, TRJ_EINVAL);
  return TRJ_OK;
}

int trj_nc_frame_ref(trj_nc* h, int64_t frame, int32_t nsp, const int32_t* sp_types, double* pos_soa,
                     int32_t* itype, int64_t* n_out, double box[3], int32_t periodic[3]) {
  return trj_nc_frame_ref_q(h, frame, nsp, sp_types, pos_soa, itype, nullptr, n_out, box, periodic);
}

static int trj_nc_frame_ref_q_impl(trj_nc* h, int64_t frame, int32_t nsp, const int32_t* sp_types, double* pos_soa,
                       int32_t* itype, double* charge, int64_t* n_out, double box[3], int32_t periodic[3]) {
  if (!h || !sp_types || !pos_soa || !itype || !n_out || !box || !periodic) return TRJ_EINVAL;
  if (nsp < 1 || nsp > 10) return (h->err = "nsp must be in 1..10", TRJ_EINVAL);
  if (!h->type) return (h->err = "the file has no type variable", TRJ_EINVAL);
  const int64_t n = h->natoms;
  std::vector<double>& r = h->xyz;
  std::vector<int32_t>& ty = h->ty;
  r.resize(static_cast<size_t>(3 * n));
  ty.resize(static_cast<size_t>(n));
  double org[3], cell[3];
  int rc = trj_nc_read(h, frame, r.data(), ty.data(), org, cell);
  if (rc) return rc;
  // Charges (netcdf.f90:404-408, 657-658): q(frame, atom) as stored, carried
  // with the selected atoms.
  if (charge) {
    if (!h->charge) return (h->err = "the file has no q(frame, atom) variable", TRJ_EINVAL);
    h->qd.resize(static_cast<size_t>(n));
    if (!h->r.read(*h->charge, frame, 0, n, h->qd.data())) return (h->err = h->r.err, TRJ_EINVAL);
  }
  // Species map, fixed at the first frame read: the selected LAMMPS types in
  // order of first appearance (netcdf.f90:441-506).
  const std::vector<int32_t> sel(sp_types, sp_types + nsp);
  for (int32_t t : sel)
    if (t < 0 || t > kTypeMax) return (h->err = "sp_types must lie in 0..2^20", TRJ_EINVAL);
  if (h->species_of.empty()) {
    int32_t maxt = 0;
    for (int32_t t : ty) maxt = std::max(maxt, t);
    for (int32_t t : sel) maxt = std::max(maxt, t);
    h->species_of.assign(static_cast<size_t>(maxt) + 1, 0);
    int32_t next = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int32_t t = ty[i];
      if (t < 0 || h->species_of[t] != 0) continue;
      if (std::find(sel.begin(), sel.end(), t) != sel.end()) h->species_of[t] = ++next;
    }
    if (next != nsp) {
      h->species_of.clear();
      h->err = "some selected species are not in the trajectory's first frame read (netcdf.f90:511-513)";
      return TRJ_EINVAL;
    }
    h->selected = sel;
  } else if (h->selected != sel) {
    h->err = "sp_types differ from those of the first frame read";
    return TRJ_EINVAL;
  }
  // Coordinates (netcdf.f90:317-357): minus cell_origin on the axes the
  // caller marks periodic; a zero cell length turns the axis non-periodic with
  // length span + 8.  Coordinates of every non-periodic axis are translated to
  // sit centred in [0, L) -- to start at 4 for the span + 8 case -- which
  // changes neither the pair distances without an image along the axis nor
  // |rho(q)|^2; the library's ingest accepts non-periodic coordinates only in
  // [-L/4, 5L/4).
  // One pass for the extent of the axes that need it.  x - org rounds
  // monotonically, so min/max of the shifted values are the shifted min/max.
  double sub[3], lo[3], hi[3];
  bool need[3];
  for (int k = 0; k < 3; ++k) {
    sub[k] = periodic[k] ? org[k] : 0.0;
    need[k] = cell[k] == 0.0 || !periodic[k];
    lo[k] = INFINITY;
    hi[k] = -INFINITY;
  }
  if (need[0] || need[1] || need[2]) {
    for (int64_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], r[3 * i + k]);
        hi[k] = std::max(hi[k], r[3 * i + k]);
      }
    }
  }
  double shift[3];
  for (int k = 0; k < 3; ++k) {
    double l = n ? lo[k] - sub[k] : 0.0, u = n ? hi[k] - sub[k] : 0.0;
    box[k] = cell[k];
    shift[k] = 0.0;
    if (cell[k] == 0.0) {
      periodic[k] = 0;
      box[k] = std::fabs(u - l) + 8.0;
      shift[k] = 4.0 - l;
    } else if (!periodic[k]) {
      shift[k] = 0.5 * (cell[k] - (u - l)) - l;
    }
    if (h->keep_nonperiodic && !periodic[k]) shift[k] = 0.0;
    periodic[k] = periodic[k] != 0;
  }
  // Selected atoms only, in file order, as SoA: (x - org) + shift, each step
  // rounded as written.
  int64_t m = 0;
  const int32_t ntab = static_cast<int32_t>(h->species_of.size());
  for (int64_t i = 0; i < n; ++i) {
    const int32_t t = ty[i];
    m += (t >= 0 && t < ntab && h->species_of[t] != 0);
  }
  int64_t j = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t t = ty[i];
    const int32_t s = (t >= 0 && t < ntab) ? h->species_of[t] : 0;
    if (s == 0) continue;
    itype[j] = s;
    if (charge) charge[j] = h->qd[i];
    for (int k = 0; k < 3; ++k) {
      double x = r[3 * i + k] - sub[k];
      if (shift[k] != 0.0) x = x + shift[k];
      pos_soa[k * m + j] = x;
    }
    ++j;
  }
  *n_out = m;
  return TRJ_OK;
}

// The reader's containers are sized from the file; an allocation failure is
// reported as TRJ_ENOMEM instead of crossing the C boundary as an exception.
int trj_nc_open(trj_nc** out, const char* path, char* errbuf, int32_t errlen) {
  try {
    return trj_nc_open_impl(out, path, errbuf, errlen);
  } catch (const std::bad_alloc&) {
    if (errbuf && errlen > 0) std::snprintf(errbuf, static_cast<size_t>(errlen), "host allocation failed");
    return TRJ_ENOMEM;
  }
}

int trj_nc_read(trj_nc* h, int64_t frame, double* pos, int32_t* type, double origin[3], double lengths[3]) {
  try {
    return trj_nc_read_impl(h, frame, pos, type, origin, lengths);
  } catch (const std::bad_alloc&) {
    if (h) h->err = "host allocation failed";
    return TRJ_ENOMEM;
  }
}

int trj_nc_frame_ref_q(trj_nc* h, int64_t frame, int32_t nsp, const int32_t* sp_types, double* pos_soa,
                       int32_t* itype, double* charge, int64_t* n_out, double box[3], int32_t periodic[3]) {
  try {
    return trj_nc_frame_ref_q_impl(h, frame, nsp, sp_types, pos_soa, itype, charge, n_out, box, periodic);
  } catch (const std::bad_alloc&) {
    if (h) h->err = "host allocation failed";
    return TRJ_ENOMEM;
  }
}

}  // extern "C"
