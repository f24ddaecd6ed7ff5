// Host-side file formats either side of the engine: the PETSc binary Vec format the reference
// uses for design restarts (writeVectorToBinary / loadVectorFromBinary,
// src/core/binaryIO.cc:3-34), a VTK XML StructuredGrid (.vts) of the macroscopic fields and
// the design (LBSolver::outputFields, src/latticeBoltzmann/LBSolver.cc:251-286), and the
// DMDA global <-> natural orderings of the reference's distributed Vecs.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "lbmx.h"

namespace lbmx {
namespace {

constexpr int32_t kVecClassId = 1211214;  // VEC_FILE_CLASSID of PETSc's binary viewer

void put_be32(unsigned char* p, uint32_t v) {
  p[0] = (unsigned char)(v >> 24), p[1] = (unsigned char)(v >> 16), p[2] = (unsigned char)(v >> 8),
  p[3] = (unsigned char)v;
}
uint32_t get_be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
void put_be64(unsigned char* p, uint64_t v) {
  for (int k = 0; k < 8; ++k) p[k] = (unsigned char)(v >> (56 - 8 * k));
}
uint64_t get_be64(const unsigned char* p) {
  uint64_t v = 0;
  for (int k = 0; k < 8; ++k) v = (v << 8) | p[k];
  return v;
}

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(std::fopen(path, mode)) {
    if (!f) throw Error{LBMX_ERR_ARG, std::string("cannot open ") + path};
  }
  ~File() {
    if (f) std::fclose(f);
  }
};

// Walks a DMDA's ranks in PETSc order and each rank's block in natural order, calling
// visit(natural node index, global node index) for every node.
template <class F>
void dmda_walk(const int dims[3], const int procs[3], const int* lx, const int* ly, const int* lz, F&& visit) {
  const int* own[3] = {lx, ly, lz};
  std::vector<long long> start[3];
  for (int a = 0; a < 3; ++a) {
    if (dims[a] < 1 || procs[a] < 1) throw Error{LBMX_ERR_ARG, "grid dimensions and rank counts must be positive"};
    if (!own[a] && procs[a] != 1) throw Error{LBMX_ERR_ARG, "ownership ranges missing for a split axis"};
    start[a].assign((size_t)procs[a] + 1, 0);
    for (int r = 0; r < procs[a]; ++r) {
      const int w = own[a] ? own[a][r] : dims[a];
      if (w < 0) throw Error{LBMX_ERR_ARG, "negative ownership range"};
      start[a][r + 1] = start[a][r] + w;
    }
    if (start[a][procs[a]] != dims[a]) throw Error{LBMX_ERR_ARG, "ownership ranges do not add up to the grid"};
  }
  long long g = 0;
  for (int rk = 0; rk < procs[2]; ++rk)
    for (int rj = 0; rj < procs[1]; ++rj)
      for (int ri = 0; ri < procs[0]; ++ri)
        for (long long k = start[2][rk]; k < start[2][rk + 1]; ++k)
          for (long long j = start[1][rj]; j < start[1][rj + 1]; ++j)
            for (long long i = start[0][ri]; i < start[0][ri + 1]; ++i)
              visit((k * dims[1] + j) * dims[0] + i, g++);
}

lbmx_status dmda_convert(const int dims[3], const int procs[3], const int* lx, const int* ly, const int* lz, int dof,
                         const double* in, double* out, bool to_natural) {
  if (!dims || !procs || !in || !out) {
    g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    if (dof < 1) throw Error{LBMX_ERR_ARG, "dof must be >= 1"};
    if (in == out) throw Error{LBMX_ERR_ARG, "the conversion is not in place"};
    dmda_walk(dims, procs, lx, ly, lz, [&](long long nat, long long glob) {
      const double* src = in + (to_natural ? glob : nat) * dof;
      double* dst = out + (to_natural ? nat : glob) * dof;
      for (int c = 0; c < dof; ++c) dst[c] = src[c];
    });
  });
}

}  // namespace
}  // namespace lbmx

using lbmx::Error;
using lbmx::guarded;

extern "C" {

lbmx_status lbmx_vec_write_binary(const char* path, const double* v, size_t n) {
  if (!path || (!v && n)) {
    lbmx::g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    if (n > 0x7fffffffu) throw Error{LBMX_ERR_ARG, "vector too long for 32-bit PETSc indices"};
    std::vector<unsigned char> buf(8 + 8 * n);
    lbmx::put_be32(buf.data(), (uint32_t)lbmx::kVecClassId);
    lbmx::put_be32(buf.data() + 4, (uint32_t)n);
    for (size_t i = 0; i < n; ++i) {
      uint64_t bits;
      std::memcpy(&bits, &v[i], 8);
      lbmx::put_be64(buf.data() + 8 + 8 * i, bits);
    }
    lbmx::File f(path, "wb");
    if (std::fwrite(buf.data(), 1, buf.size(), f.f) != buf.size()) throw Error{LBMX_ERR_STATE, "short write"};
    // the options file PETSc's binary viewer writes beside the data
    const std::string info = std::string(path) + ".info";
    lbmx::File fi(info.c_str(), "w");
    std::fputs("-vecload_block_size 1\n", fi.f);
  });
}

lbmx_status lbmx_vec_load_binary(const char* path, double* v, size_t n) {
  if (!path || (!v && n)) {
    lbmx::g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    lbmx::File f(path, "rb");
    unsigned char head[8];
    if (std::fread(head, 1, 8, f.f) != 8) throw Error{LBMX_ERR_ARG, "not a PETSc binary vector (short header)"};
    if ((int32_t)lbmx::get_be32(head) != lbmx::kVecClassId)
      throw Error{LBMX_ERR_ARG, "not a PETSc binary vector (class id)"};
    const uint32_t len = lbmx::get_be32(head + 4);
    if ((size_t)len != n) throw Error{LBMX_ERR_ARG, "vector length in file != expected length"};
    std::vector<unsigned char> buf(8 * (size_t)len);
    if (std::fread(buf.data(), 1, buf.size(), f.f) != buf.size()) throw Error{LBMX_ERR_ARG, "truncated vector data"};
    for (size_t i = 0; i < n; ++i) {
      const uint64_t bits = lbmx::get_be64(buf.data() + 8 * i);
      std::memcpy(&v[i], &bits, 8);
    }
  });
}

lbmx_status lbmx_dmda_global_to_natural(const int dims[3], const int procs[3], const int* lx, const int* ly,
                                        const int* lz, int dof, const double* global, double* natural) {
  return lbmx::dmda_convert(dims, procs, lx, ly, lz, dof, global, natural, true);
}

lbmx_status lbmx_dmda_natural_to_global(const int dims[3], const int procs[3], const int* lx, const int* ly,
                                        const int* lz, int dof, const double* natural, double* global) {
  return lbmx::dmda_convert(dims, procs, lx, ly, lz, dof, natural, global, false);
}

lbmx_status lbmx_write_vts(lbmx_solver* s, const char* path) {
  if (!s || !path) {
    lbmx::g_last_error = "null argument";
    return LBMX_ERR_ARG;
  }
  return guarded([&] {
    int dims[3];
    const long long N = lbmx::solver_nodes(s, dims);
    std::vector<double> mac((size_t)N * 5), gamma((size_t)N);
    lbmx_status e = lbmx_get_macros(s, mac.data(), mac.size());
    if (e) throw Error{e, lbmx_last_error()};
    e = lbmx_get_design(s, gamma.data(), gamma.size());
    if (e) throw Error{e, lbmx_last_error()};
    const bool three = dims[2] > 1;
    lbmx::File f(path, "w");
    std::fprintf(f.f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f.f, "  <StructuredGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", dims[0] - 1, dims[1] - 1, dims[2] - 1);
    std::fprintf(f.f, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", dims[0] - 1, dims[1] - 1, dims[2] - 1);
    std::fprintf(f.f, "      <PointData Scalars=\"rho\">\n");
    auto scalar = [&](const char* name, auto&& get) {
      std::fprintf(f.f, "        <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", name);
      for (long long i = 0; i < N; ++i) std::fprintf(f.f, "%.17g\n", get(i));
      std::fprintf(f.f, "        </DataArray>\n");
    };
    // the reference's macro field names (DMDASetFieldName: rho, ux, uy[, uz]) and the design
    scalar("rho", [&](long long i) { return mac[5 * i]; });
    scalar("ux", [&](long long i) { return mac[5 * i + 1]; });
    scalar("uy", [&](long long i) { return mac[5 * i + 2]; });
    if (three) scalar("uz", [&](long long i) { return mac[5 * i + 3]; });
    if (lbmx::solver_lattice(s) == LBMX_D2Q9_D2Q5) scalar("T", [&](long long i) { return mac[5 * i + 4]; });
    scalar("gamma", [&](long long i) { return gamma[i]; });
    std::fprintf(f.f, "      </PointData>\n      <Points>\n");
    std::fprintf(f.f, "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (int z = 0; z < dims[2]; ++z)
      for (int y = 0; y < dims[1]; ++y)
        for (int x = 0; x < dims[0]; ++x) std::fprintf(f.f, "%d %d %d\n", x, y, z);
    std::fprintf(f.f, "        </DataArray>\n      </Points>\n    </Piece>\n  </StructuredGrid>\n</VTKFile>\n");
  });
}

}  // extern "C"
