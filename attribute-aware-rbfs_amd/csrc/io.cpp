// io.cpp -- the particle file format of importParticles.h:8-32.
//
// A header text file lists one timestep name per line; timestep `name` is
// the binary file <header path> + "_" + name holding a little-endian uint64
// particle count followed by count float4 (x, y, z, attribute).  The
// reference reads without checks; here a missing timestep file, or one
// shorter than its count says, is reported as an error instead of yielding
// garbage.  Bytes after the count-th particle are ignored, as the
// reference ignores them.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "aarbf.h"

struct aarbf_timesteps {
  std::vector<std::string> names;
  std::vector<std::vector<float>> data; // 4 floats per particle
};

extern "C" {

int aarbf_import_particles(const char *path, aarbf_timesteps **out) {
  if (!path || !out) return AARBF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::ifstream hrf(path);
  if (!hrf) return AARBF_ERR_IO;
  auto *ts = new (std::nothrow) aarbf_timesteps();
  if (!ts) return AARBF_ERR_OUT_OF_MEMORY;
  std::string name;
  while (std::getline(hrf, name)) {
    std::string file = std::string(path) + "_" + name;
    std::ifstream rf(file, std::ios::binary);
    if (!rf) {
      delete ts;
      return AARBF_ERR_IO;
    }
    uint64_t count = 0;
    rf.read(reinterpret_cast<char *>(&count), sizeof(count));
    if (rf.gcount() != (std::streamsize)sizeof(count)) {
      delete ts;
      return AARBF_ERR_FORMAT;
    }
    rf.seekg(0, std::ios::end);
    uint64_t bytes = (uint64_t)rf.tellg();
    if (bytes < 8 || (bytes - 8) / 16 < count) {
      delete ts;
      return AARBF_ERR_FORMAT;
    }
    rf.seekg(8, std::ios::beg);
    std::vector<float> v;
    try {
      v.resize(4 * count);
    } catch (...) {
      delete ts;
      return AARBF_ERR_OUT_OF_MEMORY;
    }
    rf.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(16 * count));
    if ((uint64_t)rf.gcount() != 16 * count) {
      delete ts;
      return AARBF_ERR_FORMAT;
    }
    ts->names.push_back(name);
    ts->data.push_back(std::move(v));
  }
  *out = ts;
  return AARBF_OK;
}

uint32_t aarbf_timesteps_count(const aarbf_timesteps *ts) { return ts ? (uint32_t)ts->data.size() : 0; }

int aarbf_timesteps_get(const aarbf_timesteps *ts, uint32_t step, const float **xyzw, uint64_t *count,
                        const char **name) {
  if (!ts || step >= ts->data.size()) return AARBF_ERR_INVALID_ARGUMENT;
  if (xyzw) *xyzw = ts->data[step].data();
  if (count) *count = ts->data[step].size() / 4;
  if (name) *name = ts->names[step].c_str();
  return AARBF_OK;
}

void aarbf_timesteps_destroy(aarbf_timesteps *ts) { delete ts; }

int aarbf_export_particles(const char *path, uint32_t num_steps, const char *const *names,
                           const float *const *xyzw, const uint64_t *counts) {
  if (!path || (num_steps && (!names || !xyzw || !counts))) return AARBF_ERR_INVALID_ARGUMENT;
  std::ofstream hdr(path);
  if (!hdr) return AARBF_ERR_IO;
  for (uint32_t s = 0; s < num_steps; ++s) {
    if (!names[s] || std::strchr(names[s], '\n')) return AARBF_ERR_INVALID_ARGUMENT;
    hdr << names[s] << "\n";
    std::ofstream bf(std::string(path) + "_" + names[s], std::ios::binary);
    if (!bf) return AARBF_ERR_IO;
    uint64_t c = counts[s];
    bf.write(reinterpret_cast<const char *>(&c), sizeof(c));
    if (c) bf.write(reinterpret_cast<const char *>(xyzw[s]), (std::streamsize)(16 * c));
    if (!bf) return AARBF_ERR_IO;
  }
  return hdr ? AARBF_OK : AARBF_ERR_IO;
}

} // extern "C"
