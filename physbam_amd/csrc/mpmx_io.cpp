// PhysBAM particle and restart files on the host.
//
// Format (little endian), as a FILE_OSTREAM writes it (Core/Read_Write/
// TYPED_STREAM.cpp:30-36: one byte, 1 if the stream holds doubles) followed by
// PARTICLES<TV>::Write (Tools/Particles/PARTICLES.cpp:362-373):
//   int 2 (version), int number, int array count, then per array in name
//   order: int coded_id, string name (int length + bytes), int count, count
//   elements (ARRAY::Write_Prefix, Core/Arrays/ARRAY.h:378-383).
// coded_id = id << 20 | scalar << 16 | m << 12 | n << 8 | element bytes
// (ATTRIBUTE_INFO::Encode, PARTICLES.cpp:15-25) with the ids and shapes that
// Register_Attributes gives the element types (PARTICLES.cpp:420-452): scalars
// id 1, vectors id 2 (n = size), matrices id 3 (m = n = size), bool id 1 with
// scalar code 3.  Matrices are stored column-major, bools as one byte.
// Restart data is the same one-byte header followed by the time
// (Write_To_File of example.time, MPM_EXAMPLE.cpp:66).
// A path ending in ".gz" is written and read through zlib's gzip format,
// as Safe_Open_Output/Safe_Open_Input do (FILE_UTILITIES.cpp:219-261).
#include "mpmx_io.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mpmx_engine.h"

namespace mpmx {
namespace io {
namespace {

bool ends_with_gz(const std::string& p) { return p.size() > 3 && p.compare(p.size() - 3, 3, ".gz") == 0; }

// Byte sink / source over a plain or gzip file.
struct File {
  FILE* f = nullptr;
  gzFile g = nullptr;
  std::string path;
  ~File() {  // after a failure: release without reporting
    if (g) gzclose(g);
    if (f) std::fclose(f);
  }
  void open_write(const std::string& p) {
    path = p;
    if (ends_with_gz(p)) g = gzopen(p.c_str(), "wb");
    else f = std::fopen(p.c_str(), "wb");
    if (!f && !g) throw Error{MPMX_EINVAL, "cannot open " + p + " for writing"};
  }
  void open_read(const std::string& p) {
    path = p;
    g = gzopen(p.c_str(), "rb");  // reads plain files as they are
    if (!g) throw Error{MPMX_EINVAL, "cannot open " + p};
  }
  void write(const void* d, size_t n) {
    const char* c = (const char*)d;
    while (n > 0) {
      const size_t k = std::min<size_t>(n, (size_t)1 << 30);
      const bool ok = g ? gzwrite(g, c, (unsigned)k) == (int)k : std::fwrite(c, 1, k, f) == k;
      if (!ok) throw Error{MPMX_EINTERNAL, "write failed: " + path};
      c += k;
      n -= k;
    }
  }
  void read(void* d, size_t n) {
    char* c = (char*)d;
    while (n > 0) {
      const size_t k = std::min<size_t>(n, (size_t)1 << 30);
      if (gzread(g, c, (unsigned)k) != (int)k) throw Error{MPMX_EINVAL, "truncated file: " + path};
      c += k;
      n -= k;
    }
  }
  void skip(size_t n) {
    std::vector<char> tmp(std::min<size_t>(n, (size_t)1 << 20));
    while (n > 0) {
      const size_t k = std::min(n, tmp.size());
      read(tmp.data(), k);
      n -= k;
    }
  }
  template<class X> void put(const X& x) { write(&x, sizeof(X)); }
  template<class X> X get() {
    X x;
    read(&x, sizeof(X));
    return x;
  }
  void close() {
    bool ok = true;
    if (g) ok = gzclose(g) == Z_OK;
    if (f) ok = std::fclose(f) == 0;
    g = nullptr;
    f = nullptr;
    if (!ok) throw Error{MPMX_EINTERNAL, "close failed: " + path};
  }
};

// Scalars per element and the type's (id, m, n) for one array kind.
struct Shape {
  int comps, id, m, n;
  bool boolean;
};
Shape shape_of(int kind, int dim) {
  switch (kind) {
    case MPMX_ARRAY_SCALAR: return {1, 1, 0, 0, false};
    case MPMX_ARRAY_VECTOR: return {dim, 2, 0, dim, false};
    case MPMX_ARRAY_VECTOR3: return {3, 2, 0, 3, false};
    case MPMX_ARRAY_MATRIX: return {dim * dim, 3, dim, dim, false};
    case MPMX_ARRAY_BOOL: return {1, 1, 0, 0, true};
  }
  throw Error{MPMX_EINVAL, "unknown particle array kind"};
}
int coded_id(const Shape& s, bool doubles) {
  const int scalar = s.boolean ? 3 : (doubles ? 1 : 0);
  const int bytes = s.boolean ? 1 : s.comps * (doubles ? 8 : 4);
  return (s.id << 20) | (scalar << 16) | (s.m << 12) | (s.n << 8) | bytes;
}

void check_common(int dim, int precision) {
  if (dim != 2 && dim != 3) throw Error{MPMX_EINVAL, "dim must be 2 or 3"};
  if (precision != MPMX_FP32 && precision != MPMX_FP64) throw Error{MPMX_EINVAL, "precision must be FP32 or FP64"};
}

}  // namespace

void write_particles(const char* path, int dim, int precision, int64_t n, const mpmx_particle_array* arrays,
                     int count) {
  check_common(dim, precision);
  if (!path || n < 0 || n > 0x7fffffff || count < 0 || (count > 0 && !arrays))
    throw Error{MPMX_EINVAL, "bad particle file arguments"};
  const bool doubles = precision == MPMX_FP64;
  std::vector<int> order(count);
  for (int i = 0; i < count; i++) {
    if (!arrays[i].name || (n > 0 && !arrays[i].data)) throw Error{MPMX_EINVAL, "particle array without name or data"};
    shape_of(arrays[i].kind, dim);
    order[i] = i;
  }
  // PARTICLES keeps its arrays sorted by name (Add_Array, PARTICLES.cpp:198-211)
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return std::string(arrays[a].name) < std::string(arrays[b].name); });
  for (int i = 1; i < count; i++)
    if (std::string(arrays[order[i]].name) == arrays[order[i - 1]].name)
      throw Error{MPMX_EINVAL, std::string("duplicate particle array ") + arrays[order[i]].name};
  File f;
  f.open_write(path);
  f.put<uint8_t>(doubles ? 1 : 0);
  f.put<int32_t>(2);
  f.put<int32_t>((int32_t)n);
  f.put<int32_t>(count);
  for (int i : order) {
    const mpmx_particle_array& a = arrays[i];
    const Shape s = shape_of(a.kind, dim);
    const std::string name = a.name;
    f.put<int32_t>(coded_id(s, doubles));
    f.put<int32_t>((int32_t)name.size());
    f.write(name.data(), name.size());
    f.put<int32_t>((int32_t)n);
    const size_t bytes = (size_t)n * (s.boolean ? 1 : s.comps * (doubles ? 8 : 4));
    if (!s.boolean) {
      f.write(a.data, bytes);
    } else {  // bools as 0/1 bytes
      std::vector<uint8_t> b((size_t)n);
      const uint8_t* src = (const uint8_t*)a.data;
      for (int64_t p = 0; p < n; p++) b[p] = src[p] ? 1 : 0;
      f.write(b.data(), b.size());
    }
  }
  f.close();
}

int64_t read_particles(const char* path, int dim, int precision, mpmx_particle_array* arrays, int count, int* found) {
  check_common(dim, precision);
  if (!path || count < 0 || (count > 0 && !arrays)) throw Error{MPMX_EINVAL, "bad particle file arguments"};
  File f;
  f.open_read(path);
  const bool file_doubles = f.get<uint8_t>() != 0;
  const int version = f.get<int32_t>();
  if (version != 2) throw Error{MPMX_EINVAL, "unsupported particle file version " + std::to_string(version)};
  const int32_t n = f.get<int32_t>();
  const int32_t na = f.get<int32_t>();
  if (n < 0 || na < 0) throw Error{MPMX_EINVAL, "corrupt particle file header"};
  std::vector<int> seen((size_t)count, 0);
  const bool want_doubles = precision == MPMX_FP64;
  for (int32_t k = 0; k < na; k++) {
    const int cid = f.get<int32_t>();
    const int32_t len = f.get<int32_t>();
    if (len < 0 || len > 4096) throw Error{MPMX_EINVAL, "corrupt particle array name"};
    std::string name((size_t)len, '\0');
    f.read(&name[0], (size_t)len);
    const int elem = cid & 0xff;
    int target = -1;
    for (int i = 0; i < count; i++)
      if (arrays[i].name && name == arrays[i].name) target = i;
    if (target < 0 || count == 0) {  // PARTICLES::Read skips arrays it does not know the same way
      f.skip(sizeof(int32_t) + (size_t)elem * (size_t)n);
      continue;
    }
    const int32_t cnt = f.get<int32_t>();
    if (cnt != n) throw Error{MPMX_EINVAL, "particle array " + name + " has the wrong length"};
    mpmx_particle_array& a = arrays[target];
    const Shape s = shape_of(a.kind, dim);
    const int id = cid >> 20, sc = (cid >> 16) & 0xf, m = (cid >> 12) & 0xf, nn = (cid >> 8) & 0xf;
    if (id != s.id || m != s.m || nn != s.n || (s.boolean != (sc == 3)))
      throw Error{MPMX_EINVAL, "particle array " + name + " has another type in the file"};
    if (n > 0 && !a.data) throw Error{MPMX_EINVAL, "no buffer for particle array " + name};
    if (s.boolean) {
      if (elem != 1) throw Error{MPMX_EINVAL, "corrupt bool array " + name};
      f.read(a.data, (size_t)n);
    } else {
      // scalar code 0/1 says float/double; others take the stream's type (PARTICLES.cpp:336-337)
      const bool rd = sc < 2 ? sc == 1 : file_doubles;
      if (elem != s.comps * (rd ? 8 : 4)) throw Error{MPMX_EINVAL, "corrupt element size in " + name};
      const size_t vals = (size_t)n * s.comps;
      if (rd == want_doubles) {
        f.read(a.data, vals * (rd ? 8 : 4));
      } else if (rd) {
        std::vector<double> t(vals);
        f.read(t.data(), vals * 8);
        float* o = (float*)a.data;
        for (size_t i = 0; i < vals; i++) o[i] = (float)t[i];
      } else {
        std::vector<float> t(vals);
        f.read(t.data(), vals * 4);
        double* o = (double*)a.data;
        for (size_t i = 0; i < vals; i++) o[i] = (double)t[i];
      }
    }
    seen[(size_t)target] = 1;
  }
  for (int i = 0; i < count; i++) {
    if (found) found[i] = seen[(size_t)i];
    else if (!seen[(size_t)i]) throw Error{MPMX_EINVAL, std::string("particle array ") + arrays[i].name + " not in " + path};
  }
  return n;
}

void write_restart(const char* path, int precision, double time) {
  File f;
  f.open_write(path);
  if (precision == MPMX_FP64) {
    f.put<uint8_t>(1);
    f.put<double>(time);
  } else {
    f.put<uint8_t>(0);
    f.put<float>((float)time);
  }
  f.close();
}

double read_restart(const char* path) {
  File f;
  f.open_read(path);
  return f.get<uint8_t>() ? f.get<double>() : (double)f.get<float>();
}

}  // namespace io
}  // namespace mpmx
