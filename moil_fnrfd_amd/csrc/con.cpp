// Readers of MOIL's connectivity (.con) files into the library's topology
// struct and of its CHARMM-style coordinate files, and a writer of its
// binary trajectories, for C and C++ callers that use MOIL's files
// without MOIL's Fortran.  It follows MOIL's own reader rconn (generic/rconn.f:46-497):
// the same records in the same order, each read with the fixed-column
// layout of rconn's FORMAT statements, including Fortran's format
// reversion (a list longer than one record continues on the next record
// from the last parenthesised group) and its blank fields (read as zero).
// Angles and improper phases are converted from degrees as rconn does
// (angleq/pi180 with pi180 = 180/(4 atan 1), pot/init_var.f:8).
//
// Supported: the sections the energy path uses.  Refused (status
// MOIL_GPU_EINVAL): LES copies (lestyp > 0), Morse bonds (nmb > 0), the
// beta list (nbeta > 0) and monomer division (mdivyes) -- the library
// supports none of them.  The arith flag is kept for moil_con_arith.
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "moil_gpu.h"

void set_error(const std::string &msg);

namespace {

// one record of the file and a column cursor over it
struct Rec {
  const std::string *s = nullptr;
  size_t pos = 0;
  std::string field(int w) {
    std::string f;
    if (s && pos < s->size()) f = s->substr(pos, (size_t)w);
    pos += (size_t)w;
    return f;
  }
  void skip(int n) { pos += (size_t)n; }
};

bool blank(const std::string &f) { return f.find_first_not_of(" \t\r") == std::string::npos; }

long fint(const std::string &f) {
  if (blank(f)) return 0;
  std::string g;
  for (char ch : f)
    if (ch != ' ') g += ch;
  return std::strtol(g.c_str(), nullptr, 10);
}

// Fw.d / Ew.d input: blanks are ignored, D exponents accepted, and a field
// without a decimal point carries d implied decimals
double freal(const std::string &f, int d) {
  if (blank(f)) return 0.0;
  std::string g;
  for (char ch : f) {
    if (ch == ' ' || ch == '\r') continue;
    g += (ch == 'D' || ch == 'd') ? 'E' : ch;
  }
  double v = std::strtod(g.c_str(), nullptr);
  if (g.find('.') == std::string::npos) {
    const size_t e = g.find_first_of("Ee");
    const double mant = std::strtod(g.substr(0, e).c_str(), nullptr) / std::pow(10.0, d);
    v = e == std::string::npos ? mant : mant * std::pow(10.0, std::strtod(g.c_str() + e + 1, nullptr));
  }
  return v;
}

bool flogical(const std::string &f) {
  for (char ch : f) {
    if (ch == ' ' || ch == '.') continue;
    return ch == 'T' || ch == 't';
  }
  return false;
}

struct Reader {
  std::vector<std::string> lines;
  size_t next = 0;
  bool eof = false;
  Rec rec() {
    Rec r;
    if (next < lines.size()) r.s = &lines[next];
    else eof = true;
    next++;
    return r;
  }
  void skip_lines(int n) {
    for (int i = 0; i < n; i++) rec();
  }
  // n integers, `per` to a record; every record starts with `lead` blank
  // columns, each item is `pre` blank columns, an integer of width w and
  // `post` blank columns; when `first_lead_only`, continuation records
  // start at the group (format reversion to a group that lacks the
  // leading 1x)
  template <class T>
  void ints(std::vector<T> &out, size_t n, int per, int lead, int pre, int w, int post = 0,
            bool first_lead_only = false) {
    out.resize(n);
    size_t k = 0;
    bool first = true;
    while (k < n) {
      Rec r = rec();
      if (!first_lead_only || first) r.skip(lead);
      first = false;
      for (int c = 0; c < per && k < n; c++) {
        r.skip(pre);
        out[k++] = (T)fint(r.field(w));
        r.skip(post);
      }
    }
  }
};

}  // namespace

struct moil_con {
  moil_topology t;
  std::vector<std::string> moname;
  std::vector<int32_t> poipt, monoflag, flagchr, exc1, exc2;
  std::vector<double> ptchg, epsgm6, epsgm12, ptms;
  std::vector<int32_t> ib1, ib2, ia1, ia2, ia3, it1, it2, it3, it4, period, ii1, ii2, ii3, ii4;
  std::vector<int32_t> sp1, sp2, idxtip3;
  std::vector<double> kbond, req, kangl, angleq, kt1, kt2, kt3, ph1, ph2, ph3, kimp, impeq, p14;
  bool arith = false;
};

namespace {

int32_t con_fail(const std::string &path, const std::string &msg) {
  set_error("con_read " + path + ": " + msg);
  return MOIL_GPU_EINVAL;
}

// monomer flags from the names MOIL's list and squeeze code test
// (pot/nbondm.f:81-82, symm/squeeze.f:64-79)
int32_t mono_flag(std::string nm) {
  while (!nm.empty() && nm.back() == ' ') nm.pop_back();
  const bool spc = nm.compare(0, 3, "SPC") == 0;
  int32_t f = 0;
  if (nm == "TIP3" || spc) f |= MOIL_MONO_WATER;
  static const char *first[] = {"TIP3", "NAAT", "NAO", "CL", "MG", "CMO", "CO", "SUL", "K", "AR"};
  for (const char *x : first)
    if (nm == x) f |= MOIL_MONO_SQZ_FIRST;
  if (spc) f |= MOIL_MONO_SQZ_FIRST;
  if (nm == "DMPC" || nm == "DOPC" || nm == "NATA") f |= MOIL_MONO_SQZ_CENT;
  return f;
}

}  // namespace

static int32_t con_read_impl(const char *path, moil_con **out, moil_topology *topo) {
  if (!path || !out || !topo) {
    set_error("con_read: null argument");
    return MOIL_GPU_EINVAL;
  }
  std::ifstream in(path);
  if (!in) return con_fail(path, "cannot open");
  Reader R;
  for (std::string line; std::getline(in, line);) R.lines.push_back(line);
  auto c = std::make_unique<moil_con>();
  // version and two comment records (rconn.f:46-58)
  {
    Rec r = R.rec();
    std::string v = r.field(4);
    if (v != "12.1") return con_fail(path, "version '" + v + "', expecting 12.1");
  }
  R.skip_lines(2);
  // counts: format(1x,14i8,3l2) (rconn.f:60-61)
  long h[14];
  bool arith, prll, mdivyes;
  {
    Rec r = R.rec();
    r.skip(1);
    for (long &x : h) x = fint(r.field(8));
    arith = flogical(r.field(2));
    prll = flogical(r.field(2));
    mdivyes = flogical(r.field(2));
    (void)prll;
  }
  const long totmon = h[0], npt = h[1], nb = h[2], nangl = h[3], ntors = h[4], nimp = h[5],
             totex = h[6], totspe = h[7], lestyp = h[8], nbulk = h[9], nmb = h[10],
             nbeta = h[12], nwaters = h[13];
  if (nbulk <= 0 || totmon <= 0 || npt <= 0) return con_fail(path, "no molecules, monomers or particles");
  const long cap = 1L << 28;  // far above any MOIL LENGTH block
  for (long x : h)
    if (x < 0 || x > cap) return con_fail(path, "implausible counts in the header");
  if (lestyp > 0) return con_fail(path, "LES copies are not supported");
  if (nmb > 0) return con_fail(path, "Morse bonds are not supported");
  if (nbeta > 0) return con_fail(path, "the beta list is not supported");
  if (mdivyes) return con_fail(path, "monomer division is not supported");
  c->arith = arith;
  // molecule names, bulk pointers, monomer names and pointers
  // (rconn.f:84-94: format(a); (/,10(1x,i7)); (/,10(1x,a4)); (/,10(1x,i7)))
  R.skip_lines((int)nbulk);
  std::vector<int32_t> pbulk;
  R.rec();
  R.ints(pbulk, (size_t)nbulk, 10, 0, 1, 7);
  c->moname.resize(totmon);
  {
    R.rec();
    size_t k = 0;
    while (k < (size_t)totmon) {
      Rec r = R.rec();
      for (int q = 0; q < 10 && k < (size_t)totmon; q++) {
        r.skip(1);
        std::string f = r.field(4);
        f.resize(4, ' ');
        c->moname[k++] = f;
      }
    }
  }
  std::vector<int32_t> pp;
  R.rec();
  R.ints(pp, (size_t)totmon, 10, 0, 1, 7);
  c->poipt.assign(1, 0);
  c->poipt.insert(c->poipt.end(), pp.begin(), pp.end());
  // particles: format(1x,i7,1x,i7,1x,i3,1x,i3,1x,i3,1x,a4,1x,f7.3,2x,f9.5,
  // 1x,e12.5,e12.5,e10.3,1x,i3,1x,a1) (rconn.f:95-101)
  R.skip_lines(2);
  c->ptchg.resize(npt);
  c->epsgm6.resize(npt);
  c->epsgm12.resize(npt);
  c->ptms.resize(npt);
  for (long i = 0; i < npt; i++) {
    Rec r = R.rec();
    r.skip(1 + 7 + 1 + 7 + 1 + 3 + 1 + 3 + 1 + 3 + 1 + 4 + 1);
    c->ptms[i] = freal(r.field(7), 3);
    r.skip(2);
    c->ptchg[i] = freal(r.field(9), 5);
    r.skip(1);
    c->epsgm6[i] = freal(r.field(12), 5);
    c->epsgm12[i] = freal(r.field(12), 5);
  }
  const double pi180 = 180.0 / (4.0 * std::atan(1.0));
  // bonds: format(2(1x,i7),2(1x,f10.4)) (rconn.f:139-151)
  if (nb > 0) {
    R.skip_lines(2);
    for (long j = 0; j < nb; j++) {
      Rec r = R.rec();
      r.skip(1); c->ib1.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->ib2.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->kbond.push_back(freal(r.field(10), 4));
      r.skip(1); c->req.push_back(freal(r.field(10), 4));
    }
  }
  // angles: format(3(1x,i7),2(1x,f10.5)), equilibrium angle in degrees
  // (rconn.f:191-229)
  if (nangl > 0) {
    R.skip_lines(2);
    for (long j = 0; j < nangl; j++) {
      Rec r = R.rec();
      r.skip(1); c->ia1.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->ia2.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->ia3.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->kangl.push_back(freal(r.field(10), 5));
      r.skip(1); c->angleq.push_back(freal(r.field(10), 5) / pi180);
    }
  }
  // torsions: format(5(1x,i7),3(1x,f10.4),1x,3(f6.2)) (rconn.f:233-262)
  if (ntors > 0) {
    R.skip_lines(2);
    for (long j = 0; j < ntors; j++) {
      Rec r = R.rec();
      int32_t v[5];
      for (int32_t &x : v) {
        r.skip(1);
        x = (int32_t)fint(r.field(7));
      }
      c->it1.push_back(v[0]); c->it2.push_back(v[1]); c->it3.push_back(v[2]);
      c->it4.push_back(v[3]); c->period.push_back(v[4]);
      r.skip(1); c->kt1.push_back(freal(r.field(10), 4));
      r.skip(1); c->kt2.push_back(freal(r.field(10), 4));
      r.skip(1); c->kt3.push_back(freal(r.field(10), 4));
      r.skip(1);
      c->ph1.push_back(freal(r.field(6), 2));
      c->ph2.push_back(freal(r.field(6), 2));
      c->ph3.push_back(freal(r.field(6), 2));
    }
  }
  // impropers: format(4(1x,i7),2(1x,e15.8)), phase in degrees
  // (rconn.f:265-295)
  if (nimp > 0) {
    R.skip_lines(2);
    for (long j = 0; j < nimp; j++) {
      Rec r = R.rec();
      r.skip(1); c->ii1.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->ii2.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->ii3.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->ii4.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->kimp.push_back(freal(r.field(15), 8));
      r.skip(1); c->impeq.push_back(freal(r.field(15), 8) / pi180);
    }
  }
  // exclusions: per atom with any, format(2(1x,i7)) atom and count, then
  // format(1x,10(i7,1x)) the partners (rconn.f:303-332); atoms without a
  // record keep an empty range
  c->exc1.assign(npt + 1, 0);
  if (totex > 0) {
    R.skip_lines(2);
    long k = 0, i = 0;
    while (k < totex) {
      Rec r = R.rec();
      r.skip(1);
      const long j = fint(r.field(7));
      r.skip(1);
      const long k1 = fint(r.field(7));
      if (j <= i || j > npt || k1 < 0 || k + k1 > totex)
        return con_fail(path, "fishy exclusion list");
      for (long a = i + 1; a < j; a++) c->exc1[a] = (int32_t)k;
      std::vector<int32_t> part;
      R.ints(part, (size_t)k1, 10, 1, 0, 7, 1, true);
      c->exc2.insert(c->exc2.end(), part.begin(), part.end());
      k += k1;
      c->exc1[j] = (int32_t)k;
      i = j;
    }
    for (long a = i + 1; a <= npt; a++) c->exc1[a] = (int32_t)k;
  }
  // scaled 1-4 pairs: format(2(1x,i7),1x,f30.5,2(1x,f15.5)) (rconn.f:353-382)
  if (totspe > 0) {
    R.skip_lines(2);
    for (long j = 0; j < totspe; j++) {
      Rec r = R.rec();
      r.skip(1); c->sp1.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->sp2.push_back((int32_t)fint(r.field(7)));
      r.skip(1); c->p14.push_back(freal(r.field(30), 5));
      r.skip(1); c->p14.push_back(freal(r.field(15), 5));
      r.skip(1); c->p14.push_back(freal(r.field(15), 5));
    }
  }
  // charge flags: format(1x,15l2); LES labels: format(1x,15i6) (rconn.f:384-388)
  R.rec();
  c->flagchr.resize(npt);
  {
    size_t k = 0;
    while (k < (size_t)npt) {
      Rec r = R.rec();
      r.skip(1);
      for (int q = 0; q < 15 && k < (size_t)npt; q++) c->flagchr[k++] = flogical(r.field(2));
    }
  }
  R.rec();
  std::vector<int32_t> cplbl;
  R.ints(cplbl, (size_t)npt, 15, 1, 0, 6);
  // water monomers: format(10i8) (rconn.f:489-492)
  if (nwaters > 0) {
    R.rec();
    R.ints(c->idxtip3, (size_t)nwaters, 10, 0, 0, 8);
  }
  if (R.eof) return con_fail(path, "file ends early");
  c->monoflag.resize(totmon);
  for (long m = 0; m < totmon; m++) c->monoflag[m] = mono_flag(c->moname[m]);

  moil_topology &t = c->t;
  std::memset(&t, 0, sizeof(t));
  t.npt = (int32_t)npt;
  t.totmon = (int32_t)totmon;
  t.poipt = c->poipt.data();
  t.monoflag = c->monoflag.data();
  t.ptchg = c->ptchg.data();
  t.epsgm6 = c->epsgm6.data();
  t.epsgm12 = c->epsgm12.data();
  t.flagchr = c->flagchr.data();
  t.exc1 = c->exc1.data();
  t.exc2 = c->exc2.data();
  t.nb = (int32_t)nb;
  t.ib1 = c->ib1.data(); t.ib2 = c->ib2.data(); t.kbond = c->kbond.data(); t.req = c->req.data();
  t.nangl = (int32_t)nangl;
  t.iangl1 = c->ia1.data(); t.iangl2 = c->ia2.data(); t.iangl3 = c->ia3.data();
  t.kangl = c->kangl.data(); t.angleq = c->angleq.data();
  t.ntors = (int32_t)ntors;
  t.itor1 = c->it1.data(); t.itor2 = c->it2.data(); t.itor3 = c->it3.data();
  t.itor4 = c->it4.data(); t.period = c->period.data();
  t.ktors1 = c->kt1.data(); t.ktors2 = c->kt2.data(); t.ktors3 = c->kt3.data();
  t.phase1 = c->ph1.data(); t.phase2 = c->ph2.data(); t.phase3 = c->ph3.data();
  t.nimp = (int32_t)nimp;
  t.iimp1 = c->ii1.data(); t.iimp2 = c->ii2.data(); t.iimp3 = c->ii3.data();
  t.iimp4 = c->ii4.data(); t.kimp = c->kimp.data(); t.impeq = c->impeq.data();
  t.totspe = (int32_t)totspe;
  t.spec1 = c->sp1.data(); t.spec2 = c->sp2.data(); t.p14 = c->p14.data();
  t.nwaters = (int32_t)nwaters;
  t.idxtip3 = c->idxtip3.data();
  *topo = t;
  *out = c.release();
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_con_arith(const moil_con *c, int32_t *arith) {
  if (!c || !arith) {
    set_error("moil_con_arith: null argument");
    return MOIL_GPU_EINVAL;
  }
  *arith = c->arith ? 1 : 0;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_con_free(moil_con *c) {
  delete c;
  return MOIL_GPU_OK;
}

// CHARMM-style coordinate file as MOIL's getcrd reads it
// (generic/getcrd.f:48-97): title records starting with '*', a record with
// the particle count (0 or npt), then one record per particle,
// format(i7,i7,1x,a4,1x,a4,3(f10.5),1x,a4,1x,a4,f10.5).  Only complete
// files are read: getcrd's repair of missing particles (placeholders and
// built hydrogens) is not done, a record out of sequence is an error.
static int32_t crd_read_impl(const char *path, int32_t npt, double *coor) {
  if (!path || !coor || npt <= 0) {
    set_error("crd_read: null argument or no particles");
    return MOIL_GPU_EINVAL;
  }
  std::ifstream in(path);
  if (!in) return con_fail(path, "cannot open");
  std::vector<std::string> lines;
  for (std::string line; std::getline(in, line);) lines.push_back(line);
  size_t k = 0;
  while (k < lines.size() && !lines[k].empty() && lines[k][0] == '*') k++;
  if (k == 0) return con_fail(path, "coordinate file with no title");
  if (k >= lines.size()) return con_fail(path, "no particle count");
  const long n = std::strtol(lines[k].c_str(), nullptr, 10);
  if (n != 0 && n != npt) return con_fail(path, "number of particles does not match");
  k++;
  for (long i = 0; i < npt; i++, k++) {
    if (k >= lines.size()) return con_fail(path, "file ends early");
    Rec r;
    r.s = &lines[k];
    const long i1 = fint(r.field(7));
    if (i1 != i + 1) return con_fail(path, "particle " + std::to_string(i + 1) + " missing");
    r.skip(7 + 1 + 4 + 1 + 4);
    for (int d = 0; d < 3; d++) coor[3 * i + d] = freal(r.field(10), 5);
  }
  return MOIL_GPU_OK;
}

// Binary trajectory as MOIL's wdyncrd writes it (dynamics/wdyncrd.f, bin
// = 1, no frozen particles): Fortran unformatted sequential records, each
// framed by its byte length (4-byte integer) before and after.  Header:
// 'CORD' and 20 integers (the frame count first, the frozen-particle
// count ninth), then 1 with an 80-character title, then npt; each frame:
// x, y and z of all particles in single precision, one record each.
struct moil_traj {
  FILE *f = nullptr;
  int32_t npt = 0;
  std::vector<float> buf;
};

namespace {

bool frec(FILE *f, const void *data, uint32_t nbytes) {
  return std::fwrite(&nbytes, 4, 1, f) == 1 &&
         (nbytes == 0 || std::fwrite(data, 1, nbytes, f) == nbytes) &&
         std::fwrite(&nbytes, 4, 1, f) == 1;
}

}  // namespace

static int32_t traj_open_impl(const char *path, int32_t npt, int32_t nfrm, moil_traj **out) {
  if (!path || !out || npt <= 0 || nfrm < 0) {
    set_error("traj_open: null argument or no particles");
    return MOIL_GPU_EINVAL;
  }
  auto t = std::make_unique<moil_traj>();
  t->buf.resize(npt);  // before the file exists: nothing to clean up if it throws
  FILE *f = std::fopen(path, "wb");
  if (!f) return con_fail(path, "cannot create");
  char hdr[4 + 20 * 4];
  int32_t vari[20] = {0};
  vari[0] = nfrm;  // vari(1): frames; vari(9): frozen particles (none)
  std::memcpy(hdr, "CORD", 4);
  std::memcpy(hdr + 4, vari, sizeof(vari));
  char title[4 + 80];
  const int32_t one = 1;
  std::memcpy(title, &one, 4);
  std::memset(title + 4, ' ', 80);
  const char *dummy = "A dummy title to dynamics coordinates";
  std::memcpy(title + 4, dummy, std::strlen(dummy));
  if (!frec(f, hdr, sizeof(hdr)) || !frec(f, title, sizeof(title)) || !frec(f, &npt, 4)) {
    std::fclose(f);
    return con_fail(path, "write failed");
  }
  t->f = f;
  t->npt = npt;
  *out = t.release();
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_traj_open(const char *path, int32_t npt, int32_t nfrm, moil_traj **out) {
  try {
    return traj_open_impl(path, npt, nfrm, out);
  } catch (const std::exception &e) {
    set_error(std::string("traj_open: ") + e.what());
  } catch (...) {
    set_error("traj_open: failed");
  }
  return MOIL_GPU_EINVAL;
}

extern "C" int32_t moil_traj_write(moil_traj *t, const double *coor) {
  if (!t || !coor) {
    set_error("traj_write: null argument");
    return MOIL_GPU_EINVAL;
  }
  for (int k = 0; k < 3; k++) {
    for (int32_t i = 0; i < t->npt; i++) t->buf[i] = (float)coor[3 * (size_t)i + k];
    if (!frec(t->f, t->buf.data(), (uint32_t)(4 * (size_t)t->npt))) {
      set_error("traj_write: write failed");
      return MOIL_GPU_EINVAL;
    }
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_traj_close(moil_traj *t) {
  if (!t) return MOIL_GPU_OK;
  const bool ok = std::fclose(t->f) == 0;
  delete t;
  if (!ok) {
    set_error("traj_close: close failed");
    return MOIL_GPU_EINVAL;
  }
  return MOIL_GPU_OK;
}

// Frame `frame` (0-based) of a binary trajectory written as above, read
// the way MOIL's rdyncrd reads it (dynamics/rdyncrd.f:40-70, binary, no
// frozen particles): the header's particle count must equal npt; each
// frame's x, y and z records are widened to double precision into
// coor(3,npt).  The record lengths are checked.
static int32_t traj_read_impl(const char *path, int32_t npt, int32_t frame, double *coor,
                               int32_t *nframes) {
  if (!path || !coor || npt <= 0 || frame < 0) {
    set_error("traj_read: null argument, no particles or a negative frame");
    return MOIL_GPU_EINVAL;
  }
  FILE *f = std::fopen(path, "rb");
  if (!f) return con_fail(path, "cannot open");
  std::unique_ptr<FILE, int (*)(FILE *)> guard(f, std::fclose);
  auto rd = [&](std::vector<char> &buf) -> bool {
    uint32_t n = 0, m = 0;
    if (std::fread(&n, 4, 1, f) != 1) return false;
    buf.resize(n);
    if (n && std::fread(buf.data(), 1, n, f) != n) return false;
    return std::fread(&m, 4, 1, f) == 1 && m == n;
  };
  std::vector<char> b;
  if (!rd(b) || b.size() != 84 || std::memcmp(b.data(), "CORD", 4) != 0)
    return con_fail(path, "not a trajectory ('CORD' header)");
  int32_t vari[20];
  std::memcpy(vari, b.data() + 4, sizeof(vari));
  if (vari[8] != 0) return con_fail(path, "frozen particles are not supported");
  if (!rd(b)) return con_fail(path, "truncated header");
  if (!rd(b) || b.size() != 4) return con_fail(path, "truncated header");
  int32_t n;
  std::memcpy(&n, b.data(), 4);
  if (n != npt) return con_fail(path, "number of particles does not match");
  if (nframes) *nframes = vari[0];
  for (int32_t fr = 0; fr <= frame; fr++) {
    for (int k = 0; k < 3; k++) {
      if (!rd(b) || b.size() != 4 * (size_t)npt)
        return con_fail(path, "frame " + std::to_string(frame) + " not in the file");
      if (fr == frame) {
        const float *v = reinterpret_cast<const float *>(b.data());
        for (int32_t i = 0; i < npt; i++) coor[3 * (size_t)i + k] = (double)v[i];
      }
    }
  }
  return MOIL_GPU_OK;
}

// the readers never let an exception (a bad allocation on a corrupt file)
// cross the C boundary
extern "C" int32_t moil_con_read(const char *path, moil_con **out, moil_topology *topo) {
  try {
    return con_read_impl(path, out, topo);
  } catch (const std::exception &e) {
    set_error(std::string("con_read: ") + e.what());
  } catch (...) {
    set_error("con_read: failed");
  }
  return MOIL_GPU_EINVAL;
}

extern "C" int32_t moil_crd_read(const char *path, int32_t npt, double *coor) {
  try {
    return crd_read_impl(path, npt, coor);
  } catch (const std::exception &e) {
    set_error(std::string("crd_read: ") + e.what());
  } catch (...) {
    set_error("crd_read: failed");
  }
  return MOIL_GPU_EINVAL;
}

extern "C" int32_t moil_traj_read(const char *path, int32_t npt, int32_t frame, double *coor,
                                  int32_t *nframes) {
  try {
    return traj_read_impl(path, npt, frame, coor, nframes);
  } catch (const std::exception &e) {
    set_error(std::string("traj_read: ") + e.what());
  } catch (...) {
    set_error("traj_read: failed");
  }
  return MOIL_GPU_EINVAL;
}

// MOIL's force_norm (pot/force_norm.f:1-44): the RMS gradient component,
// sqrt(sum_i |dpot_i|^2 / (3 npt)), summed in force_norm's order (atom by
// atom, ((gradf + x^2) + y^2) + z^2).  Host; dpot(3,npt).
extern "C" int32_t moil_force_norm(const double *dpot, int32_t npt, double *gradf) {
  if (!dpot || !gradf || npt <= 0) {
    set_error("force_norm: null argument or no particles");
    return MOIL_GPU_EINVAL;
  }
  double g = 0.0;
  for (int32_t i = 0; i < npt; i++) {
    const double *d = dpot + 3 * (size_t)i;
    g = g + d[0] * d[0];
    g = g + d[1] * d[1];
    g = g + d[2] * d[2];
  }
  *gradf = std::sqrt(g / (double)(3 * npt));
  return MOIL_GPU_OK;
}
