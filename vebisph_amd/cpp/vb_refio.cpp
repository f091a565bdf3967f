// Readers and writers of the reference's run files (include/vebisph_refio.h): Fortran list-directed
// input as input.F90 reads INDAT/IBOUND/IPART/TEMP_REC.dat, list-directed output as output.F90 writes
// PART<n>/TIME/LS/TIME_OUT/TEMP_REC.dat with the reference's compiler.  Host code, no GPU.
#include "vebisph_refio.h"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int32_t fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

// ------------------------------------------------------------------------------------------------
// list-directed input
// ------------------------------------------------------------------------------------------------
class LdIn {
 public:
  explicit LdIn(const char* path) : path_(path ? path : ""), f_(path ? fopen(path, "r") : nullptr) {}
  ~LdIn() {
    if (f_) fclose(f_);
  }
  bool open() const { return f_ != nullptr; }
  const std::string& path() const { return path_; }
  int record() const { return rec_; }

  // One READ of n values: starts at the next record, continues over records until n values are
  // taken, drops the rest of the last record.
  bool read(int n, std::vector<std::string>& out) {
    out.clear();
    while ((int)out.size() < n) {
      std::string line;
      if (!next_line(line)) return false;
      split(line, n - (int)out.size(), out);
      if (slash_) return false;            // '/' ends the list early: a value is missing
    }
    return true;
  }

 private:
  bool next_line(std::string& line) {
    line.clear();
    int ch;
    bool any = false;
    while ((ch = fgetc(f_)) != EOF) {
      any = true;
      if (ch == '\n') break;
      if (ch != '\r') line.push_back((char)ch);
    }
    if (!any) return false;
    rec_++;
    return true;
  }
  void split(const std::string& s, int want, std::vector<std::string>& out) {
    size_t i = 0;
    slash_ = false;
    while (i < s.size() && want > 0) {
      while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++;
      if (i >= s.size()) break;
      if (s[i] == '/') { slash_ = true; return; }
      if (s[i] == ',') { i++; continue; }
      size_t j = i;
      while (j < s.size() && s[j] != ' ' && s[j] != '\t' && s[j] != ',' && s[j] != '/') j++;
      out.push_back(s.substr(i, j - i));
      want--;
      i = j;
      while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++;
      if (i < s.size() && s[i] == ',') i++;
    }
  }
  std::string path_;
  FILE* f_;
  int rec_ = 0;
  bool slash_ = false;
};

bool parse_real(const std::string& t, double* v) {
  std::string s = t;
  for (char& c : s)
    if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'E';
  // a signed exponent without its letter: 1.0-3, 2.5+10
  for (size_t k = 1; k < s.size(); k++) {
    if ((s[k] == '+' || s[k] == '-') && s[k - 1] != 'E' && s[k - 1] != 'e') {
      s.insert(k, "E");
      break;
    }
  }
  if (s.empty()) return false;
  char* end = nullptr;
  *v = strtod(s.c_str(), &end);
  return end && *end == '\0';
}

bool parse_int(const std::string& t, int32_t* v) {
  if (t.empty()) return false;
  char* end = nullptr;
  long x = strtol(t.c_str(), &end, 10);
  if (!end || *end != '\0') return false;
  *v = (int32_t)x;
  return true;
}

bool parse_logical(const std::string& t, int32_t* v) {
  size_t k = t.size() > 1 && t[0] == '.' ? 1 : 0;
  if (k >= t.size()) return false;
  char c = t[k];
  if (c == 'T' || c == 't') { *v = 1; return true; }
  if (c == 'F' || c == 'f') { *v = 0; return true; }
  return false;
}

// Typed reads: each call is one READ statement of the reference.
struct Reader {
  LdIn in;
  std::vector<std::string> tok;
  std::string what;
  explicit Reader(const char* p) : in(p) {}
  int32_t err(const std::string& m) {
    return fail(in.path() + ": record " + std::to_string(in.record()) + ": " + m + " (" + what + ")");
  }
  int32_t reals(const char* name, int n, double* out) {
    what = name;
    if (!in.read(n, tok)) return err("end of file or missing value");
    for (int k = 0; k < n; k++)
      if (!parse_real(tok[k], &out[k])) return err("not a real: '" + tok[k] + "'");
    return 0;
  }
  int32_t ints(const char* name, int n, int32_t* out) {
    what = name;
    if (!in.read(n, tok)) return err("end of file or missing value");
    for (int k = 0; k < n; k++)
      if (!parse_int(tok[k], &out[k])) return err("not an integer: '" + tok[k] + "'");
    return 0;
  }
  // a READ of mixed items: kinds[k] 'd' -> *dv[k], 'i' -> *iv[k]
  int32_t mixed(const char* name, const char* kinds, double* const* dv, int32_t* const* iv) {
    what = name;
    const int n = (int)strlen(kinds);
    if (!in.read(n, tok)) return err("end of file or missing value");
    for (int k = 0; k < n; k++) {
      if (kinds[k] == 'd' ? !parse_real(tok[k], dv[k]) : !parse_int(tok[k], iv[k]))
        return err("bad value: '" + tok[k] + "'");
    }
    return 0;
  }
  int32_t logical(const char* name, int32_t* out) {
    what = name;
    if (!in.read(1, tok)) return err("end of file or missing value");
    if (!parse_logical(tok[0], out)) return err("not a logical: '" + tok[0] + "'");
    return 0;
  }
};

// ------------------------------------------------------------------------------------------------
// list-directed output
// ------------------------------------------------------------------------------------------------
std::string fmt_real(double x) {
  if (x == 0.0) return std::signbit(x) ? "-0." : "0.";
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x > 0 ? "Inf" : "-Inf";
  const double a = std::fabs(x);
  // shortest digits that read back to the same double: d.ddde[+-]XX
  char b[64];
  auto r = std::to_chars(b, b + sizeof b, a, std::chars_format::scientific);
  std::string sci(b, r.ptr);
  const size_t epos = sci.find('e');
  std::string digits;
  for (size_t k = 0; k < epos; k++)
    if (sci[k] != '.') digits.push_back(sci[k]);
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  const int E = std::atoi(sci.c_str() + epos + 1) + 1;          // a = 0.digits x 10^E
  // fixed form when the value rounded to one significant digit has 0.1 <= |x| < 1e15
  char c[32];
  snprintf(c, sizeof c, "%.0e", a);
  const int E1 = std::atoi(strchr(c, 'e') + 1) + 1;
  std::string body;
  const int nd = (int)digits.size();
  if (E1 >= 0 && E1 <= 15) {
    if (E <= 0) body = "." + std::string(-E, '0') + digits;
    else if (E < nd) body = digits.substr(0, E) + "." + digits.substr(E);
    else body = digits + std::string(E - nd, '0') + ".";
  } else {
    const int ee = E - 1;
    char eb[16];
    snprintf(eb, sizeof eb, "E%c%02d", ee < 0 ? '-' : '+', ee < 0 ? -ee : ee);
    body = digits.substr(0, 1) + "." + digits.substr(1) + eb;
  }
  return (x < 0 ? "-" : "") + body;
}

std::string fmt_int(long long v) { return std::to_string(v); }

// One WRITE(u,*) record: items led by one blank each, continued on a new line past column 79.
class Record {
 public:
  Record& d(double x) { return item(fmt_real(x)); }
  Record& i(long long v) { return item(fmt_int(v)); }
  Record& l(bool v) { return item(v ? "T" : "F"); }
  void put(FILE* f) {
    out_ += line_;
    out_ += '\n';
    fputs(out_.c_str(), f);
    out_.clear();
    line_.clear();
  }

 private:
  Record& item(const std::string& s) {
    if (!line_.empty() && line_.size() + 1 + s.size() > 79) {
      out_ += line_;
      out_ += '\n';
      line_.clear();
    }
    line_ += ' ';
    line_ += s;
    return *this;
  }
  std::string line_, out_;
};

class Out {
 public:
  Out(const char* path, const char* mode) : f_(path ? fopen(path, mode) : nullptr), path_(path ? path : "") {}
  ~Out() { close(); }
  bool open() const { return f_ != nullptr; }
  FILE* f() { return f_; }
  int32_t close() {
    if (!f_) return 0;
    const int bad = ferror(f_) | fclose(f_);
    f_ = nullptr;
    return bad ? fail(path_ + ": write failed") : 0;
  }

 private:
  FILE* f_;
  std::string path_;
};

int32_t cant_open(const char* path) { return fail(std::string("cannot open ") + (path ? path : "(null)")); }

}  // namespace

extern "C" {

const char* vbio_last_error(void) { return g_err.c_str(); }

int32_t vbio_format_real(double x, char* buf, int32_t len) {
  const std::string s = fmt_real(x);
  if (!buf || len < (int32_t)s.size() + 1) return -1;
  memcpy(buf, s.c_str(), s.size() + 1);
  return (int32_t)s.size();
}

// input.F90:18-37
int32_t vbio_read_indat(const char* path, vbio_indat* o) {
  if (!o) return fail("vbio_read_indat: null output");
  Reader r(path);
  if (!r.in.open()) return cant_open(path);
  *o = vbio_indat{};
  int32_t s = r.logical("i_RESTART", &o->i_restart);
  if (!s) s = r.ints("i_open_domain", 1, &o->i_open_domain);
  if (!s) s = r.reals("tmax", 1, &o->tmax);
  if (!s) s = r.ints("i_shift_coeff", 1, &o->i_shift_coeff);
  if (!s) s = r.ints("np", 1, &o->np);
  if (!s) s = r.ints("npfb", 1, &o->npfb);
  if (!s) s = r.ints("nbS", 1, &o->nbS);
  if (!s) s = r.reals("dt_out", 1, &o->dt_out);
  if (!s) s = r.reals("h", 1, &o->h);
  if (!s) s = r.reals("dx", 1, &o->dx);
  if (!s) s = r.reals("dt", 1, &o->dt);
  if (!s) s = r.reals("body_force", 2, o->body_force);
  if (!s) s = r.reals("Ra", 1, &o->Ra);
  if (!s) s = r.reals("Pr", 1, &o->Pr);
  if (!s) s = r.reals("El", 1, &o->El);
  if (!s) s = r.reals("beta", 1, &o->beta);
  double c2[2];
  if (!s) s = r.reals("dt_coef_cfl, dt_coef_visc", 2, c2);
  if (!s) { o->dt_coef_cfl = c2[0]; o->dt_coef_visc = c2[1]; }
  if (!s) s = r.reals("ve_nonlinearity", 1, &o->ve_nonlinearity);
  if (!s) s = r.reals("alpha_evss", 1, &o->alpha_evss);
  return s;
}

// one value per record, as the reference's generator lays the file out (comments are not needed)
int32_t vbio_write_indat(const char* path, const vbio_indat* in) {
  if (!in) return fail("vbio_write_indat: null input");
  Out o(path, "w");
  if (!o.open()) return cant_open(path);
  Record r;
  r.l(in->i_restart != 0).put(o.f());
  r.i(in->i_open_domain).put(o.f());
  r.d(in->tmax).put(o.f());
  r.i(in->i_shift_coeff).put(o.f());
  r.i(in->np).put(o.f());
  r.i(in->npfb).put(o.f());
  r.i(in->nbS).put(o.f());
  r.d(in->dt_out).put(o.f());
  r.d(in->h).put(o.f());
  r.d(in->dx).put(o.f());
  r.d(in->dt).put(o.f());
  r.d(in->body_force[0]).d(in->body_force[1]).put(o.f());
  r.d(in->Ra).put(o.f());
  r.d(in->Pr).put(o.f());
  r.d(in->El).put(o.f());
  r.d(in->beta).put(o.f());
  r.d(in->dt_coef_cfl).d(in->dt_coef_visc).put(o.f());
  r.d(in->ve_nonlinearity).put(o.f());
  r.d(in->alpha_evss).put(o.f());
  return o.close();
}

// input.F90:51-81
int32_t vbio_read_ibound(const char* path, vbio_ibound* o) {
  if (!o) return fail("vbio_read_ibound: null output");
  Reader r(path);
  if (!r.in.open()) return cant_open(path);
  *o = vbio_ibound{};
  int32_t s = r.ints("nb_patches", 1, &o->nb_patches);
  if (s) return s;
  if (o->nb_patches < 1 || o->nb_patches > VBIO_MAX_PATCHES)
    return fail(std::string(path) + ": nb_patches out of range 1.." + std::to_string(VBIO_MAX_PATCHES));
  for (int i = 0; i < o->nb_patches && !s; i++) {
    s = r.reals("b_node", 2, o->b_node[i]);
    if (!s) s = r.reals("b_edge", 2, o->b_edge[i]);
    if (!s) s = r.ints("b_type", 1, &o->b_type[i]);
    if (!s) s = r.reals("b_vel", 1, &o->b_vel[i]);
    if (!s) s = r.reals("b_thermal", 1, &o->b_thermal[i]);
    if (!s && o->b_type[i] == 2) s = r.ints("b_periodic_parent", 1, &o->b_periodic_parent[i]);
    if (!s && o->b_type[i] == 3) s = r.reals("U_inflow", 1, &o->u_inflow);
  }
  if (!s) s = r.ints("nb_circles", 1, &o->nb_circles);
  if (s) return s;
  if (o->nb_circles < 0 || o->nb_circles > VBIO_MAX_CIRCLES)
    return fail(std::string(path) + ": nb_circles out of range 0.." + std::to_string(VBIO_MAX_CIRCLES));
  for (int i = 0; i < o->nb_circles && !s; i++) {
    s = r.reals("c_centre", 2, o->c_centre[i]);
    if (!s) s = r.reals("c_radius", 1, &o->c_radius[i]);
    if (!s) s = r.reals("c_omega", 1, &o->c_omega[i]);
    if (!s) s = r.reals("c_vel", 2, o->c_vel[i]);
    if (!s) s = r.reals("c_thermal", 1, &o->c_thermal[i]);
  }
  return s;
}

int32_t vbio_write_ibound(const char* path, const vbio_ibound* in) {
  if (!in) return fail("vbio_write_ibound: null input");
  if (in->nb_patches < 1 || in->nb_patches > VBIO_MAX_PATCHES || in->nb_circles < 0 ||
      in->nb_circles > VBIO_MAX_CIRCLES)
    return fail("vbio_write_ibound: patch or circle count out of range");
  Out o(path, "w");
  if (!o.open()) return cant_open(path);
  Record r;
  r.i(in->nb_patches).put(o.f());
  for (int i = 0; i < in->nb_patches; i++) {
    r.d(in->b_node[i][0]).d(in->b_node[i][1]).put(o.f());
    r.d(in->b_edge[i][0]).d(in->b_edge[i][1]).put(o.f());
    r.i(in->b_type[i]).put(o.f());
    r.d(in->b_vel[i]).put(o.f());
    r.d(in->b_thermal[i]).put(o.f());
    if (in->b_type[i] == 2) r.i(in->b_periodic_parent[i]).put(o.f());
    if (in->b_type[i] == 3) r.d(in->u_inflow).put(o.f());
  }
  r.i(in->nb_circles).put(o.f());
  for (int i = 0; i < in->nb_circles; i++) {
    r.d(in->c_centre[i][0]).d(in->c_centre[i][1]).put(o.f());
    r.d(in->c_radius[i]).put(o.f());
    r.d(in->c_omega[i]).put(o.f());
    r.d(in->c_vel[i][0]).d(in->c_vel[i][1]).put(o.f());
    r.d(in->c_thermal[i]).put(o.f());
  }
  return o.close();
}

// check_boundaries, input.F90:184-240 (checks 1-3; the reference STOPs when any fails)
int32_t vbio_check_ibound(const vbio_ibound* b) {
  if (!b) return fail("vbio_check_ibound: null input");
  const int nb = b->nb_patches;
  std::string bad;
  for (int i = 0; i < nb; i++) {                                  // check 1
    const int j = (i + 1) % nb;
    const double t0 = b->b_node[i][0] + b->b_edge[i][0] - b->b_node[j][0];
    const double t1 = b->b_node[i][1] + b->b_edge[i][1] - b->b_node[j][1];
    if (sqrt(t0 * t0 + t1 * t1) > 1.0e-10)
      bad += "b_nodes " + std::to_string(i + 1) + " and " + std::to_string(j + 1) + " not connected; ";
  }
  double s0 = 0.0, s1 = 0.0;                                      // check 2
  for (int i = 0; i < nb; i++) { s0 = s0 + b->b_edge[i][0]; s1 = s1 + b->b_edge[i][1]; }
  if (sqrt(s0 * s0 + s1 * s1) > 1.0e-10) bad += "edges do not sum to zero; ";
  for (int i = 0; i < nb; i++) {                                  // check 3
    if (b->b_type[i] != 2) continue;
    const int pp = b->b_periodic_parent[i] - 1;
    if (pp < 0 || pp >= nb) { bad += "periodic parent out of range; "; continue; }
    const double t0 = b->b_edge[i][0] + b->b_edge[pp][0], t1 = b->b_edge[i][1] + b->b_edge[pp][1];
    if (fabs(sqrt(t0 * t0 + t1 * t1)) > 1.0e-10)
      bad += "periodic patches " + std::to_string(i + 1) + " and " + std::to_string(pp + 1) + " do not match; ";
  }
  return bad.empty() ? 0 : fail("IBOUND checks failed: " + bad);
}

// input.F90:139-144
int32_t vbio_read_ipart(const char* path, int32_t npfb, double* x, double* y, double* u, double* v,
                        double* theta, double* dv) {
  if (npfb < 0 || !x || !y || !u || !v || !theta || !dv) return fail("vbio_read_ipart: bad arguments");
  Reader r(path);
  if (!r.in.open()) return cant_open(path);
  double row[6];
  for (int i = 0; i < npfb; i++) {
    if (int32_t s = r.reals("r u theta dv", 6, row)) return s;
    x[i] = row[0]; y[i] = row[1]; u[i] = row[2]; v[i] = row[3]; theta[i] = row[4]; *dv = row[5];
  }
  return 0;
}

int32_t vbio_write_ipart(const char* path, int32_t npfb, const double* x, const double* y, const double* u,
                         const double* v, const double* theta, double dv) {
  if (npfb < 0 || !x || !y || !u || !v || !theta) return fail("vbio_write_ipart: bad arguments");
  Out o(path, "w");
  if (!o.open()) return cant_open(path);
  Record r;
  for (int i = 0; i < npfb; i++) r.d(x[i]).d(y[i]).d(u[i]).d(v[i]).d(theta[i]).d(dv).put(o.f());
  return o.close();
}

// output.F90:44-51
int32_t vbio_write_part(const char* path, int32_t n, const double* x, const double* y, const double* u,
                        const double* v, const double* theta, const double* p, const double* conc,
                        const double* a_out, const int32_t* n_surf, const double* vort, const double* tau11,
                        const double* tau12, const double* tau22, const int32_t* inbin) {
  // no stress arrays: the Newtonian build's ten columns (output.F90:46-47)
  const bool newt = !tau11 && !tau12 && !tau22;
  if (n < 0 || !x || !y || !u || !v || !theta || !p || !conc || !a_out || !n_surf || !vort ||
      (!newt && (!tau11 || !tau12 || !tau22)) || !inbin)
    return fail("vbio_write_part: bad arguments");
  Out o(path, "w");
  if (!o.open()) return cant_open(path);
  Record r;
  for (int i = 0; i < n; i++) {
    if (inbin[i]) continue;
    r.d(x[i]).d(y[i]).d(u[i]).d(v[i]).d(theta[i]).d(p[i]).d(conc[i]).d(a_out[i]).i(n_surf[i]).d(vort[i]);
    if (!newt) r.d(tau11[i]).d(tau12[i]).d(tau22[i]);
    r.put(o.f());
  }
  return o.close();
}

int32_t vbio_read_part_ncols(const char* path, int32_t ncols, int32_t cap, int32_t* n, double* cols) {
  if ((ncols != 10 && ncols != 13) || cap < 0 || !n || (cap > 0 && !cols))
    return fail("vbio_read_part: bad arguments");
  Reader r(path);
  if (!r.in.open()) return cant_open(path);
  *n = 0;
  double row[13];
  // a list of ncols values per particle until the end of the file
  for (;;) {
    r.what = "PART record";
    if (!r.in.read(ncols, r.tok)) {
      if (r.tok.empty()) break;                       // clean end of file
      return r.err("incomplete particle record");
    }
    for (int k = 0; k < ncols; k++)
      if (!parse_real(r.tok[k], &row[k])) return r.err("not a number: '" + r.tok[k] + "'");
    if (*n >= cap) return fail(std::string(path) + ": more than cap particles");
    for (int k = 0; k < ncols; k++) cols[(size_t)k * cap + *n] = row[k];
    (*n)++;
  }
  return 0;
}

int32_t vbio_read_part(const char* path, int32_t cap, int32_t* n, double* cols) {
  return vbio_read_part_ncols(path, 13, cap, n, cols);
}

// output.F90:60-81
int32_t vbio_write_temp_rec(const char* path, const vbio_restart* hd, int32_t n, const double* x,
                            const double* y, const double* u, const double* v, const double* theta,
                            const double* p, double dv, const int32_t* n_surf, const double* tau11,
                            const double* tau12, const double* tau21, const double* tau22,
                            const int32_t* inbin) {
  // no stress arrays: the Newtonian build's file, without the stress lines (output.F90:69-71)
  const bool newt = !tau11 && !tau12 && !tau21 && !tau22;
  if (!hd || n < 0 || !x || !y || !u || !v || !theta || !p || !n_surf ||
      (!newt && (!tau11 || !tau12 || !tau21 || !tau22)) || !inbin)
    return fail("vbio_write_temp_rec: bad arguments");
  Out o(path, "w");
  if (!o.open()) return cant_open(path);
  Record r;
  r.i(hd->itime).d(hd->time).i(hd->n_dump).put(o.f());
  r.i(hd->np).i(hd->npfb).i(hd->nbS).i(hd->np_inflow).put(o.f());
  r.d(hd->xb_min).d(hd->xb_max).d(hd->yb_min).d(hd->yb_max).d(hd->dx).d(hd->av_conc).put(o.f());
  for (int i = 0; i < n; i++) {
    if (inbin[i]) continue;
    r.d(x[i]).d(y[i]).d(u[i]).d(v[i]).d(theta[i]).d(p[i]).d(dv).i(n_surf[i]).put(o.f());
    if (!newt) r.d(tau11[i]).d(tau12[i]).d(tau21[i]).d(tau22[i]).put(o.f());
  }
  return o.close();
}

// input.F90:116-133, with the eight values per particle line that output.F90:70 writes
int32_t vbio_read_temp_rec(const char* path, vbio_restart* hd, int32_t cap, double* x, double* y, double* u,
                           double* v, double* theta, double* p, double* dv, int32_t* n_surf, double* tau11,
                           double* tau12, double* tau21, double* tau22) {
  if (!hd || cap < 0) return fail("vbio_read_temp_rec: bad arguments");
  Reader r(path);
  if (!r.in.open()) return cant_open(path);
  *hd = vbio_restart{};
  {
    double* dv3[3] = {nullptr, &hd->time, nullptr};
    int32_t* iv3[3] = {&hd->itime, nullptr, &hd->n_dump};
    if (int32_t s = r.mixed("itime time n_dump", "idi", dv3, iv3)) return s;
  }
  int32_t i4[4];
  if (int32_t s = r.ints("np npfb nbS np_inflow", 4, i4)) return s;
  hd->np = i4[0]; hd->npfb = i4[1]; hd->nbS = i4[2]; hd->np_inflow = i4[3];
  double d6[6];
  if (int32_t s = r.reals("xb_min xb_max yb_min yb_max dx av_conc", 6, d6)) return s;
  hd->xb_min = d6[0]; hd->xb_max = d6[1]; hd->yb_min = d6[2]; hd->yb_max = d6[3]; hd->dx = d6[4];
  hd->av_conc = d6[5];
  if (cap == 0) return 0;
  if (hd->npfb < 0 || hd->npfb > cap) return fail(std::string(path) + ": npfb exceeds cap");
  const bool newt = !tau11 && !tau12 && !tau21 && !tau22;     // the Newtonian file (input.F90:128-130)
  if (!x || !y || !u || !v || !theta || !p || !dv || !n_surf || (!newt && (!tau11 || !tau12 || !tau21 || !tau22)))
    return fail("vbio_read_temp_rec: bad arguments");
  for (int i = 0; i < hd->npfb; i++) {
    double* d8[8] = {&x[i], &y[i], &u[i], &v[i], &theta[i], &p[i], dv, nullptr};
    int32_t* i8[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &n_surf[i]};
    if (int32_t s = r.mixed("x y u v theta p dv n_surf", "dddddddi", d8, i8)) return s;
    if (newt) continue;
    double t4[4];
    if (int32_t s = r.reals("tau11 tau12 tau21 tau22", 4, t4)) return s;
    tau11[i] = t4[0]; tau12[i] = t4[1]; tau21[i] = t4[2]; tau22[i] = t4[3];
  }
  return 0;
}

int32_t vbio_append_record(const char* path, const char* kinds, const double* vals) {
  if (!kinds || (!vals && kinds[0])) return fail("vbio_append_record: bad arguments");
  Out o(path, "a");
  if (!o.open()) return cant_open(path);
  Record r;
  for (int k = 0; kinds[k]; k++) {
    if (kinds[k] == 'd') r.d(vals[k]);
    else if (kinds[k] == 'i') r.i((long long)vals[k]);
    else return fail(std::string("vbio_append_record: unknown kind '") + kinds[k] + "'");
  }
  r.put(o.f());
  return o.close();
}

}  // extern "C"
