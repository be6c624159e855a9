// ARGOT's on-disk formats (host side of libargot_hip.so):
//
//   mesh file "<prefix>_XXX_YYY_ZZZ" (rank coordinates, three digits each):
//     struct run_param, as it lies in memory (its pointer fields included)
//     one struct fluid_mesh_io (80 B: dens eneg momx momy momz uene pot chem)
//     per local cell, in mesh order (z fastest)
//   source file "<prefix>_src.dat":
//     uint64 nsrc, struct freq_param, nsrc struct radiation_src
//
// as written by output_data.c:14-78 and read by input_data.c:48-230.  The
// reference checks a mesh file against its compile-time mesh and rank
// constants; here the caller's run_param states the mesh it allocated
// (nmesh_*_local, and, when set, the global mesh and rank counts) and a file
// that disagrees is refused with ARGOT_ERR_ARG instead of exit().  A source
// file with more sources than the caller's array holds
// (argot_hip_set_max_sources, default NSOURCE_MAX = 1024, run_param.h:74) is
// refused likewise.  Unreadable or short files give ARGOT_ERR_IO.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "argot_hip.h"

namespace argot {
void argot_hip_internal_set_error(int code);
void argot_hip_internal_clear_error();
}  // namespace argot

namespace {

constexpr int kNspecies = 5;   // NSPECIES with __HELIUM__, no H2 (run_param.h:94-100)
constexpr int kNchannel = 3;   // NCHANNEL with __HELIUM__, no H2 (run_param.h:106-112)
uint64_t g_max_sources = 1024;

// struct fluid_mesh_io, src/fluid.h:118-125
struct mesh_io {
  float dens, eneg, momx, momy, momz, uene, pot;
  struct prim_chem chem;
};
static_assert(sizeof(mesh_io) == 80, "struct fluid_mesh_io is 80 bytes");

void fail(int code) { argot::argot_hip_internal_set_error(code); }

// Runs an entry point's body so that no exception crosses the C-ABI: a
// failed allocation is ARGOT_ERR_NOMEM, anything else ARGOT_ERR_ARG.
template <typename F>
void guarded(F&& body) {
  try {
    body();
  } catch (const std::bad_alloc&) {
    fail(ARGOT_ERR_NOMEM);
  } catch (...) {
    fail(ARGOT_ERR_ARG);
  }
}

struct File {
  FILE* f = nullptr;
  File(const char* name, const char* mode) : f(name ? fopen(name, mode) : nullptr) {}
  ~File() {
    if (f) fclose(f);
  }
  bool read(void* p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }
  bool write(const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }
  // bytes from the current position to the end, or -1
  int64_t remaining() {
    const long here = ftell(f);
    if (here < 0 || fseek(f, 0, SEEK_END) != 0) return -1;
    const long end = ftell(f);
    if (end < 0 || fseek(f, here, SEEK_SET) != 0) return -1;
    return (int64_t)end - here;
  }
};

// Cell records move through a fixed buffer of this many records, so that no
// allocation scales with the mesh.
constexpr int64_t kIoChunk = 65536;

int64_t local_cells(const run_param* r) {
  if (r->nmesh_x_local <= 0 || r->nmesh_y_local <= 0 || r->nmesh_z_local <= 0) return -1;
  return (int64_t)r->nmesh_x_local * r->nmesh_y_local * r->nmesh_z_local;
}

void rank_file(char* out, size_t n, const char* prefix, const run_param* r) {
  snprintf(out, n, "%s_%03d_%03d_%03d", prefix, r->rank_x, r->rank_y, r->rank_z);
}

// The header checks of input_mesh_header (input_data.c:48-115) against the
// caller's run_param instead of compile-time constants.
bool header_fits(const run_param& h, const run_param* r) {
  if (h.nmesh_x_local != r->nmesh_x_local || h.nmesh_y_local != r->nmesh_y_local ||
      h.nmesh_z_local != r->nmesh_z_local)
    return false;
  if (r->nmesh_x_total > 0 &&
      (h.nmesh_x_total != r->nmesh_x_total || h.nmesh_y_total != r->nmesh_y_total ||
       h.nmesh_z_total != r->nmesh_z_total))
    return false;
  if (r->nnode_x > 0 &&
      (h.nnode_x != r->nnode_x || h.nnode_y != r->nnode_y || h.nnode_z != r->nnode_z ||
       h.mpi_nproc != r->nnode_x * r->nnode_y * r->nnode_z))
    return false;
  return h.nspecies == kNspecies && h.nchannel == kNchannel;
}

}  // namespace

extern "C" {

void argot_hip_set_max_sources(uint64_t n) { g_max_sources = n; }

static void input_mesh_single_body(struct fluid_mesh* mesh, struct run_param* this_run,
                                   char* filename) {
  if (!mesh || !this_run || !filename) return fail(ARGOT_ERR_ARG);
  const int64_t n = local_cells(this_run);
  if (n < 0) return fail(ARGOT_ERR_ARG);
  File in(filename, "rb");
  if (!in.f) return fail(ARGOT_ERR_IO);
  run_param h;
  if (!in.read(&h, sizeof h)) return fail(ARGOT_ERR_IO);
  if (!header_fits(h, this_run)) return fail(ARGOT_ERR_ARG);
  // a short file is refused before any cell is touched
  if (in.remaining() < n * (int64_t)sizeof(mesh_io)) return fail(ARGOT_ERR_IO);
  std::vector<mesh_io> rec((size_t)std::min(n, kIoChunk));
  for (int64_t i = 0; i < n; i++) {
    const int64_t j = i % kIoChunk;
    if (j == 0 && !in.read(rec.data(), sizeof(mesh_io) * (size_t)std::min(n - i, kIoChunk)))
      return fail(ARGOT_ERR_IO);
    fluid_mesh& m = mesh[i];
    const mesh_io& t = rec[(size_t)j];
    m.dens = t.dens;
    m.eneg = t.eneg;
    m.momx = t.momx;
    m.momy = t.momy;
    m.momz = t.momz;
    m.uene = t.uene;
    m.pot = t.pot;
    m.chem = t.chem;
    m.prev_chem = t.chem;
    m.prev_uene = t.uene;
  }
  memcpy(this_run, &h, sizeof h);
}

void input_mesh_single(struct fluid_mesh* mesh, struct run_param* this_run, char* filename) {
  argot::argot_hip_internal_clear_error();
  guarded([&] { input_mesh_single_body(mesh, this_run, filename); });
}

void input_mesh(struct fluid_mesh* mesh, struct run_param* this_run, char* prefix) {
  if (!this_run || !prefix) {
    argot::argot_hip_internal_clear_error();
    return fail(ARGOT_ERR_ARG);
  }
  char name[4096];
  rank_file(name, sizeof name, prefix, this_run);
  input_mesh_single(mesh, this_run, name);
}

static void input_src_file_body(struct radiation_src* src, struct run_param* this_run,
                                char* src_file) {
  if (!src || !this_run || !src_file) return fail(ARGOT_ERR_ARG);
  File in(src_file, "rb");
  if (!in.f) return fail(ARGOT_ERR_IO);
  uint64_t nsrc = 0;
  if (!in.read(&nsrc, sizeof nsrc)) return fail(ARGOT_ERR_IO);
  if (nsrc > g_max_sources) return fail(ARGOT_ERR_ARG);
  freq_param freq;
  if (!in.read(&freq, sizeof freq)) return fail(ARGOT_ERR_IO);
  // a short file is refused before the records are allocated
  const int64_t left = in.remaining();
  if (left < 0 || (uint64_t)left / sizeof(radiation_src) < nsrc) return fail(ARGOT_ERR_IO);
  std::vector<radiation_src> s((size_t)nsrc);
  if (!in.read(s.data(), sizeof(radiation_src) * s.size())) return fail(ARGOT_ERR_IO);
  this_run->nsrc = nsrc;
  this_run->freq = freq;
  if (nsrc) memcpy(src, s.data(), sizeof(radiation_src) * s.size());
}

void input_src_file(struct radiation_src* src, struct run_param* this_run, char* src_file) {
  argot::argot_hip_internal_clear_error();
  guarded([&] { input_src_file_body(src, this_run, src_file); });
}

void input_src(struct radiation_src* src, struct run_param* this_run, char* prefix) {
  if (!prefix) {
    argot::argot_hip_internal_clear_error();
    return fail(ARGOT_ERR_ARG);
  }
  char name[4096];
  snprintf(name, sizeof name, "%s_src.dat", prefix);
  input_src_file(src, this_run, name);
}

void input_data(struct fluid_mesh* mesh, struct radiation_src* src, struct run_param* this_run,
                char* prefix) {
  input_mesh(mesh, this_run, prefix);
  if (argot_hip_last_error()) return;
  input_src(src, this_run, prefix);
}

static void output_mesh_single_body(struct fluid_mesh* mesh, struct run_param* this_run,
                                    char* filename) {
  if (!mesh || !this_run || !filename) return fail(ARGOT_ERR_ARG);
  const int64_t n = local_cells(this_run);
  if (n < 0) return fail(ARGOT_ERR_ARG);
  File out(filename, "wb");
  if (!out.f) return fail(ARGOT_ERR_IO);
  if (!out.write(this_run, sizeof(run_param))) return fail(ARGOT_ERR_IO);
  std::vector<mesh_io> rec((size_t)std::min(n, kIoChunk));
  for (int64_t i = 0; i < n; i++) {
    const int64_t j = i % kIoChunk;
    const fluid_mesh& m = mesh[i];
    mesh_io& t = rec[(size_t)j];
    memset(&t, 0, sizeof t);
    t.dens = m.dens;
    t.eneg = m.eneg;
    t.momx = m.momx;
    t.momy = m.momy;
    t.momz = m.momz;
    t.uene = m.uene;
    t.pot = m.pot;
    t.chem = m.chem;
    if ((j == kIoChunk - 1 || i == n - 1) &&
        !out.write(rec.data(), sizeof(mesh_io) * (size_t)(j + 1)))
      return fail(ARGOT_ERR_IO);
  }
}

void output_mesh_single(struct fluid_mesh* mesh, struct run_param* this_run, char* filename) {
  argot::argot_hip_internal_clear_error();
  guarded([&] { output_mesh_single_body(mesh, this_run, filename); });
}

void output_mesh(struct fluid_mesh* mesh, struct run_param* this_run, char* prefix) {
  if (!this_run || !prefix) {
    argot::argot_hip_internal_clear_error();
    return fail(ARGOT_ERR_ARG);
  }
  char name[4096];
  rank_file(name, sizeof name, prefix, this_run);
  output_mesh_single(mesh, this_run, name);
}

void output_src(struct radiation_src* src, struct run_param* this_run, char* prefix) {
  argot::argot_hip_internal_clear_error();
  if (!this_run || !prefix || (this_run->nsrc && !src)) return fail(ARGOT_ERR_ARG);
  char name[4096];
  snprintf(name, sizeof name, "%s_src.dat", prefix);
  File out(name, "wb");
  if (!out.f) return fail(ARGOT_ERR_IO);
  const uint64_t nsrc = this_run->nsrc;
  if (!out.write(&nsrc, sizeof nsrc) || !out.write(&this_run->freq, sizeof(freq_param)) ||
      !out.write(src, sizeof(radiation_src) * nsrc))
    return fail(ARGOT_ERR_IO);
}

void output_data(struct fluid_mesh* mesh, struct radiation_src* src, struct run_param* this_run,
                 char* prefix) {
  output_mesh(mesh, this_run, prefix);
  if (argot_hip_last_error()) return;
  output_src(src, this_run, prefix);
}

}  // extern "C"
