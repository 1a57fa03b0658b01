// Reading and writing the reference's restart dumps (restart.c:206
// restart_write through dumpgen.c:5 dump_gen, text output as initbase.c:237
// selects without MPI).
//
// Layout of dumps/rdump-N: one header line written by write_restart_header
// (restart.c:641, text branch, format WRITERESTARTHEADER of global.h:1214),
// then one line per active zone in DUMPLOOP order (k outer, i fastest,
// global.h:1411) holding p[0..7] and unew[0..7], each "%21.15g ".
//
// WRITERESTARTHEADER prints the double defcoord with "%d": with this many
// arguments every field sits in its own 8-byte stack slot, so the token is
// the low 32 bits of defcoord's bits (0 for the integral values it takes) and
// the fields after it are not displaced.  The reader skips that token.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "harmgpu.h"
#include "hg_host.h"

using namespace hg;

namespace {

struct Reader {
  const char *s;
  bool ok = true;
  long next_long() {
    if (!ok) return 0;
    char *e;
    errno = 0;
    long v = strtol(s, &e, 10);
    if (e == s || errno) ok = false;
    s = e;
    return v;
  }
  double next_double() {
    if (!ok) return 0.;
    char *e;
    double v = strtod(s, &e);  // correctly rounded, as glibc's scanf "%lf"
    if (e == s) ok = false;
    s = e;
    return v;
  }
};

}  // namespace

extern "C" int hg_rdump_parse(const char *path, int n1, int n2, int n3, double *p_active, double *unew,
                              hg_rdump_info *info) {
  if (!path || n1 < 1 || n2 < 1 || n3 < 1) {
    set_error("hg_rdump_parse: bad arguments");
    return 1;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    set_error("hg_rdump_parse: cannot open %s", path);
    return 1;
  }
  std::string text;
  {
    std::vector<char> buf(1 << 20);
    size_t r;
    while ((r = fread(buf.data(), 1, buf.size(), f)) > 0) text.append(buf.data(), r);
    fclose(f);
  }
  size_t eol = text.find('\n');
  if (eol == std::string::npos) {
    set_error("hg_rdump_parse: %s has no header line", path);
    return 1;
  }
  std::string head = text.substr(0, eol);
  Reader h{head.c_str()};
  hg_rdump_info in;
  memset(&in, 0, sizeof in);
  in.n[0] = (int)h.next_long();
  in.n[1] = (int)h.next_long();
  in.n[2] = (int)h.next_long();
  in.t = h.next_double();
  in.tf = h.next_double();
  in.nstep = h.next_long();
  in.a = h.next_double();
  in.gam = h.next_double();
  in.cour = h.next_double();
  for (int k = 0; k < 4; k++) in.DT[k] = h.next_double();  // DTd DTavg DTener DTi
  in.DTr = h.next_long();
  in.DTdebug = h.next_double();
  for (int k = 0; k < 5; k++) in.counts[k] = h.next_long();  // dump image rdump avg debug
  in.dt = h.next_double();
  in.lim = (int)h.next_long();
  in.timeorder = (int)h.next_long();
  in.fluxmethod = (int)h.next_long();
  in.fluxb = (int)h.next_long();
  in.utoprimversion = (int)h.next_long();
  in.failed = (int)h.next_long();
  in.R0 = h.next_double();
  in.Rin = h.next_double();
  in.Rout = h.next_double();
  in.hslope = h.next_double();
  (void)h.next_long();  // defcoord, see above
  for (int k = 0; k < 6; k++) in.bctype[k] = (int)h.next_long();  // X1UP X1DN X2UP X2DN X3UP X3DN
  in.realnstep = h.next_long();
  (void)h.next_long();                               // debugfail
  (void)h.next_long();                               // whichrestart
  (void)h.next_long();                               // cooling
  (void)h.next_long();                               // restartsteps[0]
  (void)h.next_long();                               // restartsteps[1]
  for (int k = 0; k < 17; k++) (void)h.next_long();  // dump/diag switches, POSDEFMETRIC, periodic, output modes
  in.defcon = h.next_double();
  in.safe = h.next_double();
  in.rhomin = h.next_double();
  in.uumin = h.next_double();
  (void)h.next_double();                             // RHOMINLIMIT
  (void)h.next_double();                             // UUMINLIMIT
  (void)h.next_double();                             // BSQORHOLIMIT
  (void)h.next_double();                             // BSQOULIMIT
  in.gammamax = h.next_double();
  (void)h.next_double();                             // GAMMADAMP
  in.gammafail = h.next_double();
  if (!h.ok) {
    set_error("hg_rdump_parse: malformed header in %s", path);
    return 1;
  }
  if (in.n[0] != n1 || in.n[1] != n2 || in.n[2] != n3) {
    set_error("hg_rdump_parse: %s holds %dx%dx%d zones, expected %dx%dx%d", path, in.n[0], in.n[1], in.n[2], n1,
              n2, n3);
    return 1;
  }
  Reader d{text.c_str() + eol + 1};
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        size_t z = ((size_t)i * n2 + j) * n3 + k;
        for (int pl = 0; pl < NPR; pl++) {
          double v = d.next_double();
          if (p_active) p_active[z * NPR + pl] = v;
        }
        for (int pl = 0; pl < NPR; pl++) {
          double v = d.next_double();
          if (unew) unew[z * NPR + pl] = v;
        }
        if (!d.ok) {
          set_error("hg_rdump_parse: %s ends before zone (%d,%d,%d)", path, i, j, k);
          return 1;
        }
      }
  if (info) *info = in;
  return 0;
}

extern "C" int hg_rdump_read(hg_ctx *c, const char *path, hg_rdump_info *info, double *unew, void *stream) {
  if (!c) {
    set_error("hg_rdump_read: null context");
    return 1;
  }
  if (!c->have_geometry) {
    set_error("hg_rdump_read: geometry not set");
    return 1;
  }
  const Dims &d = c->d;
  size_t nact = (size_t)d.n1 * d.n2 * d.n3;
  std::vector<double> pa(nact * NPR), ua;
  // the finite-volume scheme resumes from the dump's unew (initbase.c:155-160)
  if (c->ucons) ua.resize(nact * NPR);
  double *uout = c->ucons ? ua.data() : unew;
  if (hg_rdump_parse(path, d.n1, d.n2, d.n3, pa.data(), uout, info)) return 1;
  if (c->ucons && unew) memcpy(unew, ua.data(), ua.size() * sizeof(double));
  // the active zones into the state (ghost zones refilled by bound_prim)
  std::vector<double> full((size_t)d.m1 * d.m2 * d.m3 * NPR);
  if (hg_get_prims(c, full.data())) return 1;
  for (int i = 0; i < d.n1; i++)
    for (int j = 0; j < d.n2; j++)
      for (int k = 0; k < d.n3; k++) {
        size_t src = (((size_t)i * d.n2 + j) * d.n3 + k) * NPR;
        size_t dst = ((((size_t)(i + d.nb1)) * d.m2 + (j + d.nb2)) * d.m3 + (k + d.nb3)) * NPR;
        for (int pl = 0; pl < NPR; pl++) full[dst + pl] = pa[src + pl];
      }
  if (hg_set_prims(c, full.data())) return 1;
  if (c->ucons) {
    std::fill(full.begin(), full.end(), 0.0);
    for (int i = 0; i < d.n1; i++)
      for (int j = 0; j < d.n2; j++)
        for (int k = 0; k < d.n3; k++) {
          size_t src = (((size_t)i * d.n2 + j) * d.n3 + k) * NPR;
          size_t dst = ((((size_t)(i + d.nb1)) * d.m2 + (j + d.nb2)) * d.m3 + (k + d.nb3)) * NPR;
          for (int pl = 0; pl < NPR; pl++) full[dst + pl] = ua[src + pl];
        }
    if (hg_set_conserved(c, full.data())) return 1;
  }
  return hg_bound_prim(c, stream);
}

// restart.c:206 restart_write: the caller's header line (the reference's
// write_restart_header, restart.c:641), then rdump_content (restart.c:240),
// p and unew of every active zone, each "%21.15g " (global.h:1214), one line
// per zone in DUMPLOOP order
extern "C" int hg_rdump_write(hg_ctx *c, const char *path, const char *header, void *stream) {
  if (!c || !path || !header) {
    set_error("hg_rdump_write: null argument");
    return 1;
  }
  if (!c->have_geometry) {
    set_error("hg_rdump_write: geometry not set");
    return 1;
  }
  if (!c->ucons && !c->unew_in_uf) {
    set_error("hg_rdump_write: no step since the primitives were set, so there is no unew");
    return 1;
  }
  const Dims &d = c->d;
  const size_t nfull = (size_t)d.m1 * d.m2 * d.m3 * NPR;
  std::vector<double> p(nfull), u(nfull);
  if (hg_get_prims(c, p.data())) return 1;
  if (c->ucons) {
    if (hg_get_conserved(c, u.data())) return 1;
  } else {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) {
      set_error("hg_rdump_write: %s", hipGetErrorString(e));
      return 1;
    }
    if (fetch_prims(c, c->uf, u.data())) return 1;
    // the midpoint scheme inverted Uf; its ucum is 0 + (0 Ui + 0 dt dU + Uf),
    // which is Uf with -0 made +0
    if (c->cfg.timeorder != 4)
      for (double &x : u) x = x + 0.0;
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    set_error("hg_rdump_write: cannot open %s", path);
    return 1;
  }
  std::vector<char> buf(1 << 20);
  setvbuf(f, buf.data(), _IOFBF, buf.size());
  bool ok = fprintf(f, "%s\n", header) >= 0;
  for (int k = 0; k < d.n3 && ok; k++)
    for (int j = 0; j < d.n2 && ok; j++)
      for (int i = 0; i < d.n1 && ok; i++) {
        size_t h = ((((size_t)(i + d.nb1)) * d.m2 + (j + d.nb2)) * d.m3 + (k + d.nb3)) * NPR;
        for (int pl = 0; pl < NPR && ok; pl++) ok = fprintf(f, "%21.15g ", p[h + pl]) >= 0;
        for (int pl = 0; pl < NPR && ok; pl++) ok = fprintf(f, "%21.15g ", u[h + pl]) >= 0;
        if (ok) ok = fputc('\n', f) != EOF;
      }
  if (fclose(f) != 0) ok = false;
  if (!ok) {
    set_error("hg_rdump_write: write to %s failed", path);
    return 1;
  }
  return 0;
}
