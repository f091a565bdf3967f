// Stand-alone C++ driver of the C ABI (include/vebisph_gpu.h): reads a case file, runs n steps of
// vb_div_free, prints one line per step, writes the final state.
//
//   vb_drive CASE_FILE NSTEPS OUT_FILE
//
// Case file (little-endian, written by vebisph_amd.solver.write_case_file):
//   char[4] "VBC1"; int32 npfb, nbS, nb_patches, nb_circles;
//   vb_params (raw struct);
//   b_node f64[nb*2] (row-major), b_type i32[nb], b_vel f64[nb], b_thermal f64[nb],
//   b_periodic_parent i32[nb];
//   c_centre f64[nc*2], c_radius f64[nc], c_omega f64[nc], c_vel f64[nc*2], c_thermal f64[nc];
//   state columns f64[npfb] x, y, u, v, theta, tau11, tau12, tau21, tau22, trace_conf, p; i32[npfb] inbin.
// Output file: the state columns in the same order and layout.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vebisph_gpu.h"

namespace {

struct Reader {
  FILE* f;
  bool ok = true;
  template <class T>
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) ok = false;
    return v;
  }
  template <class T>
  T one() {
    T t{};
    if (fread(&t, sizeof(T), 1, f) != 1) ok = false;
    return t;
  }
};

int fail(const char* msg) {
  fprintf(stderr, "vb_drive: %s\n", msg);
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) return fail("usage: vb_drive CASE_FILE NSTEPS OUT_FILE");
  const int nsteps = std::stoi(argv[2]);
  FILE* f = fopen(argv[1], "rb");
  if (!f) return fail("cannot open the case file");
  Reader rd{f};
  char magic[4];
  if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "VBC1", 4) != 0) return fail("not a VBC1 case file");
  const int32_t npfb = rd.one<int32_t>(), nbS = rd.one<int32_t>();
  const int32_t nb = rd.one<int32_t>(), nc = rd.one<int32_t>();
  const vb_params prm = rd.one<vb_params>();
  if (!rd.ok || npfb < 1 || nb < 1 || nb > VB_MAX_PATCHES || nc < 0 || nc > VB_MAX_CIRCLES)
    return fail("bad case header");
  auto b_node = rd.vec<double>(2 * (size_t)nb);
  auto b_type = rd.vec<int32_t>(nb);
  auto b_vel = rd.vec<double>(nb);
  auto b_thermal = rd.vec<double>(nb);
  auto b_pp = rd.vec<int32_t>(nb);
  auto c_centre = rd.vec<double>(2 * (size_t)nc);
  auto c_radius = rd.vec<double>(nc);
  auto c_omega = rd.vec<double>(nc);
  auto c_vel = rd.vec<double>(2 * (size_t)nc);
  auto c_thermal = rd.vec<double>(nc);
  std::vector<std::vector<double>> cols(11);
  for (auto& c : cols) c = rd.vec<double>(npfb);
  auto inbin = rd.vec<int32_t>(npfb);
  fclose(f);
  if (!rd.ok) return fail("truncated case file");

  vb_boundary bnd{nb, b_node.data(), b_type.data(), b_vel.data(), b_thermal.data(), b_pp.data(),
                  nc, nc ? c_centre.data() : nullptr, nc ? c_radius.data() : nullptr,
                  nc ? c_omega.data() : nullptr, nc ? c_vel.data() : nullptr,
                  nc ? c_thermal.data() : nullptr};
  vb_ctx* ctx = nullptr;
  vb_status s = vb_create(&prm, &bnd, npfb, nbS, &ctx);
  if (s != VB_OK) {
    fprintf(stderr, "vb_drive: vb_create: status %d\n", (int)s);
    return 1;
  }
  s = vb_upload_state(ctx, npfb, cols[0].data(), cols[1].data(), cols[2].data(), cols[3].data(),
                      cols[4].data(), cols[5].data(), cols[6].data(), cols[7].data(), cols[8].data(),
                      cols[9].data(), cols[10].data(), inbin.data());
  int rc = 0;
  for (int k = 0; k < nsteps && s == VB_OK; k++) {
    vb_status st = vb_div_free(ctx);
    int32_t iv[12];
    double dv[8];
    vb_get_scalars(ctx, iv, dv);
    printf("step %d itime %d dt %.17g LS_iters %d LS_residual %.6e np %d status %d\n", k + 1, iv[6], dv[0],
           iv[5], dv[1], iv[0], (int)st);
    fflush(stdout);
    if (st != VB_OK && st != VB_WARN_MAXITER && st != VB_WARN_BICGSTAB_BREAKDOWN) {
      fprintf(stderr, "vb_drive: %s\n", vb_last_error(ctx));
      rc = 1;
      break;
    }
  }
  if (s != VB_OK) {
    fprintf(stderr, "vb_drive: upload: %s\n", vb_last_error(ctx));
    rc = 1;
  }
  if (rc == 0) {
    s = vb_download_state(ctx, npfb, cols[0].data(), cols[1].data(), cols[2].data(), cols[3].data(),
                          cols[4].data(), cols[5].data(), cols[6].data(), cols[7].data(), cols[8].data(),
                          cols[9].data(), cols[10].data(), inbin.data());
    FILE* o = fopen(argv[3], "wb");
    if (s != VB_OK || !o) {
      rc = 1;
    } else {
      for (auto& c : cols) fwrite(c.data(), 8, npfb, o);
      fwrite(inbin.data(), 4, npfb, o);
      fclose(o);
    }
  }
  vb_destroy(ctx);
  return rc;
}
