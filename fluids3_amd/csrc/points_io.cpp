// points_io.cpp -- particle state files in the Fluids 4.0 ".pts" layout
// (fluids4.0/fluid_system.cpp:1511-1545 writes it, :1787-1843 reads it back
// for playback), extended with two more fields.
//
// Layout (native little-endian int32 / float32):
//   int32 numpnt, int32 numfield, then numfield fields of
//   { int32 ftype (0 char, 1 int, 2 float, 3 double), int32 fcnt (channels),
//     numpnt * fcnt values }.
// Fields written: pos (float, 3), vel (float, 3), clr (char, 4) -- the 4.0
// file -- followed by veval (float, 3) and id (int, 1).  A 4.0 reader reads
// the first three fields and ignores the rest; LoadPoints accepts 4.0 files
// (three fields: veval is then set to vel and ids to 0..n-1).
#include <cstdio>
#include <cstring>
#include <vector>

#include "fx_common.hpp"
#include "particles.hpp"

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(FILE* p) : f(p) {}
  ~File() {
    if (f) fclose(f);
  }
};

bool put(FILE* f, const void* p, size_t n) { return fwrite(p, 1, n, f) == n; }
bool get(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

}  // namespace

int Particles::SavePoints(const char* path) {
  if (mNumPoints <= 0) return fx::fail(FX_ESTATE, "save_points: call Restart first");
  if (path == nullptr) return fx::fail(FX_EINVAL, "save_points: no path");
  const int32_t n = mNumPoints;
  std::vector<char> pos(12 * (size_t)n), vel(12 * (size_t)n), clr(4 * (size_t)n), vev(12 * (size_t)n),
      id(4 * (size_t)n);
  FX_TRY(Download(FPOS, pos.data(), n));
  FX_TRY(Download(FVEL, vel.data(), n));
  FX_TRY(Download(FCLR, clr.data(), n));
  FX_TRY(Download(FVEVAL, vev.data(), n));
  FX_TRY(Download(FID, id.data(), n));
  File fp(fopen(path, "wb"));
  if (fp.f == nullptr) return fx::fail(FX_EINVAL, "save_points: cannot open %s", path);
  const int32_t numfield = 5;
  bool ok = put(fp.f, &n, 4) && put(fp.f, &numfield, 4);
  auto field = [&](int32_t ftype, int32_t fcnt, const std::vector<char>& d) {
    ok = ok && put(fp.f, &ftype, 4) && put(fp.f, &fcnt, 4) && put(fp.f, d.data(), d.size());
  };
  field(2, 3, pos);
  field(2, 3, vel);
  field(0, 4, clr);
  field(2, 3, vev);
  field(1, 1, id);
  if (!ok) return fx::fail(FX_EINVAL, "save_points: short write to %s", path);
  return FX_OK;
}

int Particles::LoadPoints(const char* path) {
  if (mNumPoints <= 0) return fx::fail(FX_ESTATE, "load_points: call Restart first");
  if (path == nullptr) return fx::fail(FX_EINVAL, "load_points: no path");
  File fp(fopen(path, "rb"));
  if (fp.f == nullptr) return fx::fail(FX_EINVAL, "load_points: cannot open %s", path);
  int32_t n = 0, numfield = 0;
  if (!get(fp.f, &n, 4) || !get(fp.f, &numfield, 4)) return fx::fail(FX_EINVAL, "load_points: truncated header");
  if (n != mNumPoints)
    return fx::fail(FX_EMISMATCH, "load_points: %s holds %d particles, the system %d", path, n, mNumPoints);
  if (numfield < 3) return fx::fail(FX_EINVAL, "load_points: %d fields, at least pos, vel, clr needed", numfield);
  // expected (ftype, fcnt) of the known fields, in order
  const int32_t want[5][2] = {{2, 3}, {2, 3}, {0, 4}, {2, 3}, {1, 1}};
  std::vector<char> d[5];
  const int known = numfield < 5 ? numfield : 5;
  for (int k = 0; k < known; k++) {
    int32_t ftype = 0, fcnt = 0;
    if (!get(fp.f, &ftype, 4) || !get(fp.f, &fcnt, 4)) return fx::fail(FX_EINVAL, "load_points: truncated field");
    if (ftype != want[k][0] || fcnt != want[k][1])
      return fx::fail(FX_EINVAL, "load_points: field %d has type %d x %d, expected %d x %d", k, ftype, fcnt,
                      want[k][0], want[k][1]);
    const size_t elem = ftype == 0 ? 1 : 4;  // char or int/float
    d[k].resize((size_t)n * (size_t)fcnt * elem);
    if (!get(fp.f, d[k].data(), d[k].size())) return fx::fail(FX_EINVAL, "load_points: truncated data");
  }
  if (known < 5) {
    d[3] = d[1];  // veval = vel
    d[4].resize(4 * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
      const uint32_t v = (uint32_t)i;
      memcpy(d[4].data() + 4 * (size_t)i, &v, 4);
    }
  }
  FX_TRY(CheckIds(reinterpret_cast<const uint32_t*>(d[4].data()), n, "load_points"));  // before any upload
  ReleaseGraph();
  FX_TRY(Upload(FPOS, d[0].data(), n));
  FX_TRY(Upload(FVEL, d[1].data(), n));
  FX_TRY(Upload(FCLR, d[2].data(), n));
  FX_TRY(Upload(FVEVAL, d[3].data(), n));
  FX_TRY(Upload(FID, d[4].data(), n));
  return FX_OK;
}
