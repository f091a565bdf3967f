// Stand-alone C++ driver of the C ABI (include/vebisph_gpu.h), two modes.
//
//   vb_drive run DIR [fp64|mixed] [MAX_STEPS] [newtonian]
//
// The reference's program (source/base/sph2D_incom.F90:51-129) on the GPU: reads DIR/INDAT, IBOUND
// and IPART (or TEMP_REC.dat when INDAT says restart) with the reference's formats, steps while
// time <= tmax (at most MAX_STEPS steps), and writes what the reference writes: data_directory/PART<n>
// and TEMP_REC.dat every dt_out (output.F90:17-82), data_directory/TIME and LS every step and
// TIME_OUT at each dump (output.F90:55,86-87), the RMS velocity to fort.192 when the reference's
// secondary_properties runs (div_free.F90:152-185).  Formats: include/vebisph_refio.h.  `newtonian` runs
// the reference's Newtonian program (built with -Dconst_mod=1), with its file layouts.
//
//   vb_drive CASE_FILE NSTEPS OUT_FILE
//
// Reads a binary case file, runs n steps of vb_div_free, prints one line per step, writes the final
// state.
//
// Case file (little-endian, written by vebisph_amd.solver.write_case_file):
//   char[4] "VBC1"; int32 npfb, nbS, nb_patches, nb_circles;
//   vb_params (raw struct);
//   b_node f64[nb*2] (row-major), b_type i32[nb], b_vel f64[nb], b_thermal f64[nb],
//   b_periodic_parent i32[nb];
//   c_centre f64[nc*2], c_radius f64[nc], c_omega f64[nc], c_vel f64[nc*2], c_thermal f64[nc];
//   state columns f64[npfb] x, y, u, v, theta, tau11, tau12, tau21, tau22, trace_conf, p; i32[npfb] inbin.
// Output file: the state columns in the same order and layout.
#include <sys/stat.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vebisph_gpu.h"
#include "vebisph_refio.h"

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

int io_fail() { return fail(vbio_last_error()); }

bool step_ok(vb_status st) { return st == VB_OK || st == VB_WARN_MAXITER || st == VB_WARN_BICGSTAB_BREAKDOWN; }

// Particle columns of the context (rows 1..npfb), as the reference's arrays.
struct Columns {
  std::vector<double> x, y, u, v, th, t11, t12, t21, t22, trc, p;
  std::vector<int32_t> inbin;
  explicit Columns(int n)
      : x(n), y(n), u(n), v(n), th(n), t11(n), t12(n), t21(n), t22(n), trc(n, 2.0), p(n), inbin(n) {}
  vb_status upload(vb_ctx* c, int n) {
    return vb_upload_state(c, n, x.data(), y.data(), u.data(), v.data(), th.data(), t11.data(), t12.data(),
                           t21.data(), t22.data(), trc.data(), p.data(), inbin.data());
  }
  vb_status download(vb_ctx* c, int n) {
    return vb_download_state(c, n, x.data(), y.data(), u.data(), v.data(), th.data(), t11.data(), t12.data(),
                             t21.data(), t22.data(), trc.data(), p.data(), inbin.data());
  }
};

// sph2D_incom.F90:51-129 with getdata (input.F90:12-180) and output (output.F90), on the GPU path.
// newtonian: the reference built with -Dconst_mod=1 (no polymeric stress; its PART files have ten
// columns and its TEMP_REC.dat no stress lines, output.F90:46-50,69-71; input.F90:128-130)
int run_dir(const std::string& dir, int precision, long max_steps, bool newtonian) {
  auto at = [&](const char* f) { return dir + "/" + f; };
  vbio_indat ind;
  vbio_ibound ib;
  if (vbio_read_indat(at("INDAT").c_str(), &ind) || vbio_read_ibound(at("IBOUND").c_str(), &ib)) return io_fail();
  if (vbio_check_ibound(&ib)) return io_fail();                  // input.F90:236-238 STOPs
  if (ind.i_open_domain != 0 && ind.i_open_domain != 1) return fail("INDAT: i_open_domain must be 0 or 1");
  if (ind.i_shift_coeff != 0 && ind.i_shift_coeff != 1) return fail("INDAT: i_shift_coeff must be 0 or 1");
  const bool open = ind.i_open_domain == 1;
  for (int i = 0; i < ib.nb_patches; i++)
    if (ib.b_type[i] == 3 || ib.b_type[i] == 4) return fail("IBOUND: inflow/outflow patches are not coded (input.F90:287-288)");
  // The device forms b_edge(ib) = b_node(ib+1) - b_node(ib), as the reference's generator writes it.
  for (int i = 0; i < ib.nb_patches; i++) {
    const int j = (i + 1) % ib.nb_patches;
    if (ib.b_edge[i][0] != ib.b_node[j][0] - ib.b_node[i][0] || ib.b_edge[i][1] != ib.b_node[j][1] - ib.b_node[i][1])
      fprintf(stderr, "vb_drive: warning: IBOUND edge %d differs from its node difference by rounding\n", i + 1);
  }

  // getdata: the state from IPART, or from TEMP_REC.dat on restart
  int npfb = ind.npfb, nbS = ind.nbS, itime = 0, n_dump = 0;
  double time = 0.0, dv = 0.0;
  vbio_restart hd{};
  if (ind.i_restart) {
    if (vbio_read_temp_rec(at("TEMP_REC.dat").c_str(), &hd, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr))
      return io_fail();
    npfb = hd.npfb; nbS = hd.nbS; itime = hd.itime; time = hd.time; n_dump = hd.n_dump;
  }
  if (npfb < 1) return fail("no particles");
  Columns s(npfb);
  std::vector<int32_t> nsurf(npfb, 0);
  if (ind.i_restart) {
    if (vbio_read_temp_rec(at("TEMP_REC.dat").c_str(), &hd, npfb, s.x.data(), s.y.data(), s.u.data(), s.v.data(),
                           s.th.data(), s.p.data(), &dv, nsurf.data(), newtonian ? nullptr : s.t11.data(),
                           newtonian ? nullptr : s.t12.data(), newtonian ? nullptr : s.t21.data(),
                           newtonian ? nullptr : s.t22.data()))
      return io_fail();
    // the file has no trace column: trace_conf = A11 + A22 with A = tau + I (predictor_mod.F90:240-242)
    for (int i = 0; i < npfb; i++) s.trc[i] = (s.t11[i] + 1.0) + (s.t22[i] + 1.0);
  } else if (vbio_read_ipart(at("IPART").c_str(), npfb, s.x.data(), s.y.data(), s.u.data(), s.v.data(),
                             s.th.data(), &dv)) {
    return io_fail();
  }
  // every particle starts on the free surface in an open domain (input.F90:147-153), until the
  // first step flags them
  std::fill(nsurf.begin(), nsurf.end(), open ? 1 : 0);

  vb_params prm{};
  prm.h = ind.h; prm.dx = ind.dx; prm.dv = dv; prm.Ra = ind.Ra; prm.Pr = ind.Pr; prm.El = ind.El;
  prm.beta = ind.beta; prm.body_force[0] = ind.body_force[0]; prm.body_force[1] = ind.body_force[1];
  prm.dt_coef_cfl = ind.dt_coef_cfl; prm.dt_coef_visc = ind.dt_coef_visc;
  prm.ve_nonlinearity = ind.ve_nonlinearity; prm.time = time; prm.itime = itime;
  prm.i_open_domain = ind.i_open_domain; prm.i_shift_coeff = ind.i_shift_coeff;
  prm.fluid_model = newtonian ? 1 : 2;
  prm.precision = precision; prm.reference_kick = 1;
  std::vector<double> bn(2 * ib.nb_patches), cc(2 * ib.nb_circles), cv(2 * ib.nb_circles);
  for (int i = 0; i < ib.nb_patches; i++) { bn[2 * i] = ib.b_node[i][0]; bn[2 * i + 1] = ib.b_node[i][1]; }
  for (int i = 0; i < ib.nb_circles; i++) {
    cc[2 * i] = ib.c_centre[i][0]; cc[2 * i + 1] = ib.c_centre[i][1];
    cv[2 * i] = ib.c_vel[i][0]; cv[2 * i + 1] = ib.c_vel[i][1];
  }
  const bool nc = ib.nb_circles > 0;
  vb_boundary bnd{ib.nb_patches, bn.data(), ib.b_type, ib.b_vel, ib.b_thermal, ib.b_periodic_parent,
                  ib.nb_circles, nc ? cc.data() : nullptr, nc ? ib.c_radius : nullptr,
                  nc ? ib.c_omega : nullptr, nc ? cv.data() : nullptr, nc ? ib.c_thermal : nullptr};
  vb_ctx* ctx = nullptr;
  if (vb_create(&prm, &bnd, npfb, nbS, &ctx) != VB_OK) return fail("vb_create failed");
  if (s.upload(ctx, npfb) != VB_OK) {
    fprintf(stderr, "vb_drive: upload: %s\n", vb_last_error(ctx));
    vb_destroy(ctx);
    return 1;
  }

  // sph2D_incom.F90:52-61: the per-step files, appended to on restart
  const std::string dd = at("data_directory");
  mkdir(dd.c_str(), 0755);
  const std::string f_time = dd + "/TIME", f_ls = dd + "/LS", f_tout = dd + "/TIME_OUT";
  const std::string f_192 = at("fort.192"), f_rec = at("TEMP_REC.dat");
  if (!ind.i_restart)
    for (const std::string& f : {f_time, f_ls, f_tout}) {
      FILE* t = fopen(f.c_str(), "w");
      if (!t) { vb_destroy(ctx); return fail(("cannot create " + f).c_str()); }
      fclose(t);
    }
  double xb[4] = {ib.b_node[0][0], ib.b_node[0][0], ib.b_node[0][1], ib.b_node[0][1]};
  for (int i = 1; i < ib.nb_patches; i++) {
    xb[0] = std::min(xb[0], ib.b_node[i][0]); xb[1] = std::max(xb[1], ib.b_node[i][0]);
    xb[2] = std::min(xb[2], ib.b_node[i][1]); xb[3] = std::max(xb[3], ib.b_node[i][1]);
  }

  double next_dump = time, next_dump2 = time, dt = ind.dt, ls_res = 0.0;
  int ls_iters = 0, n_inbin = 0, np = ind.i_restart ? hd.np : ind.np;
  std::vector<double> vort(npfb, 0.0), conc(npfb, 0.0);   // conc: set by the open domain's steps only
  const std::vector<double> zeros(npfb, 0.0);
  bool stepped = false;
  int rc = 0;
  auto scalars = [&]() {
    int32_t iv[12];
    double dvv[8];
    vb_get_scalars(ctx, iv, dvv);
    np = iv[0]; ls_iters = iv[5]; itime = iv[6]; n_inbin = iv[7];
    dt = dvv[0]; ls_res = dvv[1]; time = dvv[2];
  };
  auto output = [&]() -> int {                                   // output.F90:3-89
    if (time > next_dump) {
      next_dump = next_dump + ind.dt_out;
      const int idp = n_dump;
      n_dump = n_dump + 1;
      if (s.download(ctx, npfb) != VB_OK) return fail(vb_last_error(ctx));
      if (open && stepped && vb_free_surface(ctx, npfb, nsurf.data(), conc.data()) != VB_OK)
        return fail(vb_last_error(ctx));
      const std::string part = dd + "/PART" + std::to_string(idp);
      // a_out = trace_conf (predictor_mod.F90:243; the Newtonian build never sets it: 0); n_surf = 0 in
      // a closed domain
      if (vbio_write_part(part.c_str(), npfb, s.x.data(), s.y.data(), s.u.data(), s.v.data(), s.th.data(),
                          s.p.data(), conc.data(), newtonian ? zeros.data() : s.trc.data(), nsurf.data(),
                          vort.data(), newtonian ? nullptr : s.t11.data(), newtonian ? nullptr : s.t12.data(),
                          newtonian ? nullptr : s.t22.data(), s.inbin.data()))
        return io_fail();
      const double tout[4] = {time, (double)npfb, (double)itime, dt};
      if (vbio_append_record(f_tout.c_str(), "didd", tout)) return io_fail();
      vbio_restart h{itime, time, n_dump, np - n_inbin, npfb - n_inbin, nbS, 0,
                     xb[0], xb[1], xb[2], xb[3], ind.dx, 0.0};
      if (vbio_write_temp_rec(f_rec.c_str(), &h, npfb, s.x.data(), s.y.data(), s.u.data(), s.v.data(),
                              s.th.data(), s.p.data(), dv, nsurf.data(), newtonian ? nullptr : s.t11.data(),
                              newtonian ? nullptr : s.t12.data(), newtonian ? nullptr : s.t21.data(),
                              newtonian ? nullptr : s.t22.data(), s.inbin.data()))
        return io_fail();
    }
    double m[3];
    if (vb_field_maxima(ctx, m) != VB_OK) return fail(vb_last_error(ctx));
    const double trow[6] = {time, dt, (double)(npfb - n_inbin), m[0], m[1], m[2]};
    const double lrow[2] = {(double)ls_iters, ls_res};
    if (vbio_append_record(f_time.c_str(), "ddiddd", trow) || vbio_append_record(f_ls.c_str(), "id", lrow))
      return io_fail();
    return 0;
  };

  rc = output();
  long steps = 0;
  while (rc == 0 && time <= ind.tmax && steps < max_steps) {    // sph2D_incom.F90:90-124
    const double t_before = time;
    const vb_status st = vb_div_free(ctx);
    if (!step_ok(st)) {
      fprintf(stderr, "vb_drive: step %d: %s\n", itime + 1, vb_last_error(ctx));
      rc = 1;
      break;
    }
    if (t_before > next_dump2) {                                 // secondary_properties, div_free.F90:154-185
      next_dump2 = next_dump2 + ind.dt_out;
      double u_rms = 0.0;
      if (vb_secondary_properties(ctx, npfb, vort.data(), &u_rms) != VB_OK) { rc = fail(vb_last_error(ctx)); break; }
      const double row[2] = {t_before, u_rms};
      if (vbio_append_record(f_192.c_str(), "dd", row)) { rc = io_fail(); break; }
    }
    scalars();
    stepped = true;
    steps++;
    if (itime % 10 == 0)
      printf("itime %d time %.9g dt %.6e LS_iters %d LS_residual %.3e n_inbin %d\n", itime, time, dt, ls_iters,
             ls_res, n_inbin);
    rc = output();
  }
  if (rc == 0) rc = output();                                    // sph2D_incom.F90:129
  fflush(stdout);
  vb_destroy(ctx);
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 3 && strcmp(argv[1], "run") == 0) {
    int precision = VB_PREC_FP64;
    long max_steps = 1L << 40;
    bool newtonian = false;
    std::vector<const char*> pos;
    for (int k = 3; k < argc; k++) {
      if (strcmp(argv[k], "newtonian") == 0) newtonian = true;
      else pos.push_back(argv[k]);
    }
    if (pos.size() >= 1) {
      if (strcmp(pos[0], "mixed") == 0) precision = VB_PREC_MIXED;
      else if (strcmp(pos[0], "fp64") != 0) return fail("precision must be fp64 or mixed");
    }
    if (pos.size() >= 2) max_steps = std::stol(pos[1]);
    if (pos.size() > 2) return fail("usage: vb_drive run DIR [fp64|mixed] [MAX_STEPS] [newtonian]");
    return run_dir(argv[2], precision, max_steps, newtonian);
  }
  if (argc != 4)
    return fail("usage: vb_drive run DIR [fp64|mixed] [MAX_STEPS] [newtonian] | vb_drive CASE_FILE NSTEPS OUT_FILE");
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
