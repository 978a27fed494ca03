// h5_restated.cpp -- TEST INFRASTRUCTURE ONLY (the oracle for
// felbm_gpu_amd/csrc/h5io.c; the product never runs it).
//
// Restates the reference driver's HDF5 writers through the HDF5 C++ wrapper
// (H5Cpp), the API the reference itself calls, step for step:
//   fields      write_fields_h5       src/felbm_gpu_main.cu:47-118
//   geometry    write_geometry_h5     src/felbm_gpu_main.cu:124-143
//   checkpoint  write_checkpoint      src/felbm_gpu_main.cu:151-166
// The reference file is a CUDA translation unit over the absent felbm_local
// tree and cannot be compiled here, so its three writers are restated.
// Arrays arrive as raw native-endian files:
//   h5_restated fields     OUT n SRC(6n values) src_f32 deflate out_f32
//   h5_restated geometry   OUT n COORDS(3n int32) deflate
//   h5_restated checkpoint OUT Vn STATE(2Vn doubles: h then g) step target_mass
#include <H5Cpp.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

template <class T>
static std::vector<T> slurp(const char *path, size_t count) {
  std::vector<T> v(count);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(count * sizeof(T)));
  if (!f) {
    std::fprintf(stderr, "short read of %s\n", path);
    std::exit(2);
  }
  return v;
}

// one chunk over the whole dataset + gzip, only when deflate > 0 and n > 0
static void gzip_plist(H5::DSetCreatPropList &pl, int rank, const hsize_t *dims, int deflate, size_t n) {
  if (deflate > 0 && n > 0) {
    pl.setChunk(rank, dims);
    pl.setDeflate((unsigned)deflate);
  }
}

static int fields(const char *out, size_t n, const char *src, bool src_f32, int deflate, bool out_f32) {
  const char *names[6] = {"density", "concentration", "u_x", "u_y", "u_z", "pressure"};
  std::vector<double> d64;
  std::vector<float> d32;
  if (src_f32) d32 = slurp<float>(src, 6 * n); else d64 = slurp<double>(src, 6 * n);
  H5::H5File file(out, H5F_ACC_TRUNC);
  hsize_t dims[1] = {(hsize_t)n};
  H5::DSetCreatPropList pl;
  gzip_plist(pl, 1, dims, deflate, n);
  std::vector<float> tf(n);
  std::vector<double> td(n);
  for (int k = 0; k < 6; ++k) {
    H5::DataSpace space(1, dims);
    if (out_f32) {
      H5::DataSet ds = file.createDataSet(names[k], H5::PredType::NATIVE_FLOAT, space, pl);
      if (src_f32) {
        ds.write(d32.data() + k * n, H5::PredType::NATIVE_FLOAT);
      } else {
        for (size_t i = 0; i < n; ++i) tf[i] = (float)d64[k * n + i];
        ds.write(tf.data(), H5::PredType::NATIVE_FLOAT);
      }
    } else {
      H5::DataSet ds = file.createDataSet(names[k], H5::PredType::NATIVE_DOUBLE, space, pl);
      if (!src_f32) {
        ds.write(d64.data() + k * n, H5::PredType::NATIVE_DOUBLE);
      } else {
        for (size_t i = 0; i < n; ++i) td[i] = (double)d32[k * n + i];
        ds.write(td.data(), H5::PredType::NATIVE_DOUBLE);
      }
    }
  }
  return 0;
}

static int geometry(const char *out, size_t n, const char *src, int deflate) {
  std::vector<int> co = slurp<int>(src, 3 * n);
  int ext[3] = {0, 0, 0};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a)
      if (co[3 * i + a] > ext[a]) ext[a] = co[3 * i + a];
  for (int a = 0; a < 3; ++a) ext[a] += 1;
  H5::H5File file(out, H5F_ACC_TRUNC);
  {
    hsize_t dims[2] = {(hsize_t)n, 3};
    H5::DataSpace space(2, dims);
    H5::DSetCreatPropList pl;
    gzip_plist(pl, 2, dims, deflate, n);
    H5::DataSet ds = file.createDataSet("coords", H5::PredType::NATIVE_INT, space, pl);
    ds.write(co.data(), H5::PredType::NATIVE_INT);
  }
  {
    hsize_t dims[1] = {3};
    H5::DataSpace space(1, dims);
    H5::DataSet ds = file.createDataSet("size", H5::PredType::NATIVE_INT, space);
    ds.write(ext, H5::PredType::NATIVE_INT);
  }
  return 0;
}

static int checkpoint(const char *out, size_t vn, const char *src, unsigned step, double target_mass) {
  std::vector<double> hg = slurp<double>(src, 2 * vn);
  H5::H5File file(out, H5F_ACC_TRUNC);
  hsize_t dims[1] = {(hsize_t)vn};
  H5::DataSpace space(1, dims);
  file.createDataSet("h", H5::PredType::NATIVE_DOUBLE, space).write(hg.data(), H5::PredType::NATIVE_DOUBLE);
  file.createDataSet("g", H5::PredType::NATIVE_DOUBLE, space).write(hg.data() + vn, H5::PredType::NATIVE_DOUBLE);
  hsize_t one[1] = {1};
  H5::DataSpace s1(1, one);
  file.createDataSet("step", H5::PredType::NATIVE_UINT, s1).write(&step, H5::PredType::NATIVE_UINT);
  file.createDataSet("target_mass", H5::PredType::NATIVE_DOUBLE, s1).write(&target_mass, H5::PredType::NATIVE_DOUBLE);
  int v = (int)vn;
  file.createDataSet("Vn", H5::PredType::NATIVE_INT, s1).write(&v, H5::PredType::NATIVE_INT);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string kind = argv[1];
  if (kind == "fields" && argc == 8)
    return fields(argv[2], std::strtoull(argv[3], nullptr, 10), argv[4], std::atoi(argv[5]) != 0, std::atoi(argv[6]),
                  std::atoi(argv[7]) != 0);
  if (kind == "geometry" && argc == 6)
    return geometry(argv[2], std::strtoull(argv[3], nullptr, 10), argv[4], std::atoi(argv[5]));
  if (kind == "checkpoint" && argc == 7)
    return checkpoint(argv[2], std::strtoull(argv[3], nullptr, 10), argv[4], (unsigned)std::strtoul(argv[5], nullptr, 10),
                      std::strtod(argv[6], nullptr));
  std::fprintf(stderr, "usage: h5_restated fields|geometry|checkpoint ...\n");
  return 2;
}
