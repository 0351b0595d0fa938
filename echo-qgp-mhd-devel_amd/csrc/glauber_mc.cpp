// Glauber Monte-Carlo event generator (host code, part of the C-ABI of
// include/echo_amd.h): the nucleus-nucleus branch of generate_events
// (glaubermc.f90:80-635) with the project's own random numbers, the event
// files in the reference's formats, and a reader for them.
//
// Sampling, as glaubermc.f90 does it:
//   - nucleons of a nucleus (251-277, 280-307): r = u1 rmax (rmax 12.6 fm)
//     accepted when woods(r) > fmax u2, woods(r) = r^2 rho0 / (exp((r - R)
//     / a) + 1), fmax = 2 R^2 rho0 (glaubermc.f90:640-645, 222); cos(theta)
//     and phi uniform; each nucleus then moved to its transverse centre of
//     mass (342-373);
//   - impact parameter (379-386): b = u1 bmax, kept when b > 1.001 bmax u2
//     (p(b) ~ b), or b fixed;
//   - collisions (397-420): nucleons i of A and j of B collide when
//     (x_i - x_j + b)^2 + (y_i - y_j)^2 <= sigma_NN / pi (sigma_NN in fm^2,
//     the mb value over 10, common.f90:2259); a collision sits at the
//     midpoint, A is shifted by +b/2 and B by -b/2 in x;
//   - an event is kept when it has a collision and at least
//     min_participants participants (442); nucleons 1..Z are protons.
// Random numbers: xoshiro256** seeded by splitmix64 (the reference calls
// gfortran's rand() after srand(seed); its stream is not reproduced).
//
// Files (glaubermc.f90:112-166, 466-545), in the directory given:
//   partcoll.bin    participants, A then B per event: records of
//                   (x f64, y f64, type i32), 20 bytes, 0 proton 1 neutron
//   spectators.bin  spectators, same records
//   collisions.bin  collisions: records of (x f64, y f64), 16 bytes
//   summary.dat     one line per event: ev= b= nc= np= npA= npB= nspecA=
//                   nspecB=
// Records of an event follow those of the previous event.  (The reference
// computes the start of event n from event n's own counts instead of event
// n-1's, glaubermc.f90:458-462, so its events overlap when an event is
// smaller than the one before; within one of its runs it reads back only the
// event it simulates.)
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/echo_amd.h"

// the message of echo_last_error (echo_capi.hip, not exported)
extern "C" void echo_set_error_text(const char* msg);

namespace {

int sfail(int code, const char* msg, const std::string& what = "") {
  echo_set_error_text((msg + what).c_str());
  return code;
}

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) {   // splitmix64
      x += 0x9e3779b97f4a7c15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {   // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
};

constexpr double RMAX = 12.6;   // glaubermc.f90:41
constexpr long NTRIALS = 10000000;   // glaubermc.f90:40

struct Nucleus {
  std::vector<double> x, y, z;
};

bool sample_nucleus(Rng& rng, const echo_glauber_sampler& p, Nucleus& n) {
  const double pi = acos(-1.0);
  const double fmax = p.radius * p.radius * p.rho0 * 2.;
  n.x.clear(); n.y.clear(); n.z.clear();
  for (long j = 0; j < NTRIALS && (int)n.x.size() < p.A; j++) {
    const double r = rng.uniform() * RMAX;
    const double u2 = rng.uniform();
    const double woods = r * r * p.rho0 / (exp((r - p.radius) / p.diffuse) + 1.);
    if (woods <= fmax * u2) continue;
    const double xcos = 2. * rng.uniform() - 1.;
    const double phi = 2. * pi * rng.uniform();
    const double st = sqrt(1. - xcos * xcos);
    n.x.push_back(r * st * cos(phi));
    n.y.push_back(r * st * sin(phi));
    n.z.push_back(r * xcos);
  }
  if ((int)n.x.size() < p.A) return false;
  double cx = 0., cy = 0.;
  for (int i = 0; i < p.A; i++) { cx = cx + n.x[i]; cy = cy + n.y[i]; }
  cx = cx / p.A;
  cy = cy / p.A;
  for (int i = 0; i < p.A; i++) { n.x[i] = n.x[i] - cx; n.y[i] = n.y[i] - cy; }
  return true;
}

struct Event {
  double b = 0.;
  int npa = 0, npb = 0, nsa = 0, nsb = 0, nc = 0;
  std::vector<double> part, spec, coll;   // (x, y) pairs
  std::vector<int32_t> ptype, stype;
};

#pragma pack(push, 1)
struct Rec20 {
  double x, y;
  int32_t t;
};
struct Rec16 {
  double x, y;
};
#pragma pack(pop)
static_assert(sizeof(Rec20) == 20 && sizeof(Rec16) == 16, "record layout");

// the event of configuration (A, B) at impact parameter b
Event collide(const echo_glauber_sampler& p, const Nucleus& a, const Nucleus& bn, double b) {
  const double pi = acos(-1.0);
  const double s = p.sigma_nn / 10. / pi;   // sigma_NN / pi in fm^2
  const int A = p.A;
  std::vector<char> pa(A, 0), pb(A, 0);
  Event ev;
  ev.b = b;
  for (int i = 0; i < A; i++)
    for (int j = 0; j < A; j++) {
      const double dx = a.x[i] - bn.x[j] + b, dy = a.y[i] - bn.y[j];
      if (dx * dx + dy * dy <= s) {
        ev.nc++;
        pa[i] = 1;
        pb[j] = 1;
        ev.coll.push_back((a.x[i] + bn.x[j]) / 2.);
        ev.coll.push_back((a.y[i] + bn.y[j]) / 2.);
      }
    }
  for (int k = 0; k < 2; k++) {
    const Nucleus& n = k == 0 ? a : bn;
    const std::vector<char>& f = k == 0 ? pa : pb;
    const double sh = k == 0 ? b / 2. : -b / 2.;
    for (int i = 0; i < A; i++) {
      const int32_t t = i < p.Z ? 0 : 1;
      if (f[i]) {
        ev.part.push_back(n.x[i] + sh);
        ev.part.push_back(n.y[i]);
        ev.ptype.push_back(t);
        (k == 0 ? ev.npa : ev.npb)++;
      } else {
        ev.spec.push_back(n.x[i] + sh);
        ev.spec.push_back(n.y[i]);
        ev.stype.push_back(t);
        (k == 0 ? ev.nsa : ev.nsb)++;
      }
    }
  }
  return ev;
}

std::string path_in(const char* dir, const char* name) {
  std::string s = dir ? dir : ".";
  if (!s.empty() && s.back() != '/') s += '/';
  return s + name;
}

struct SummaryRow {
  int ev, nc, np, npa, npb, nsa, nsb;
  double b;
};

bool read_summary(const char* dir, std::vector<SummaryRow>& rows) {
  FILE* f = fopen(path_in(dir, "summary.dat").c_str(), "r");
  if (!f) return false;
  char line[512];
  while (fgets(line, sizeof line, f)) {
    SummaryRow r{};
    if (sscanf(line, " ev= %d b= %lf nc= %d np= %d npA= %d npB= %d nspecA= %d nspecB= %d", &r.ev, &r.b, &r.nc,
               &r.np, &r.npa, &r.npb, &r.nsa, &r.nsb) == 8)
      rows.push_back(r);
  }
  fclose(f);
  return true;
}

}  // namespace

extern "C" {

int echo_glauber_generate(const echo_glauber_sampler* p, uint64_t seed, const char* dir, int32_t* n_events) {
  if (!p || !n_events) return sfail(ECHO_E_ARG, "echo_glauber_generate: null argument");
  *n_events = 0;
  if (p->A < 1 || p->A > 300 || p->Z < 0 || p->Z > p->A)
    return sfail(ECHO_E_ARG, "echo_glauber_generate: need 1 <= A <= 300 and 0 <= Z <= A");
  if (p->nconf < 1 || p->nbcoll < 1) return sfail(ECHO_E_ARG, "echo_glauber_generate: need nconf, nbcoll >= 1");
  if (!(p->radius > 0.) || !(p->diffuse > 0.) || !(p->rho0 > 0.) || !(p->sigma_nn > 0.))
    return sfail(ECHO_E_ARG, "echo_glauber_generate: need radius, diffuse, rho0, sigma_nn > 0");
  if (p->fixed_b ? !(p->bimpact >= 0.) : !(p->bmax > 0.))
    return sfail(ECHO_E_ARG, "echo_glauber_generate: need bimpact >= 0 (fixed b) or bmax > 0");
  FILE* fs = fopen(path_in(dir, "summary.dat").c_str(), "w");
  FILE* fp = fopen(path_in(dir, "partcoll.bin").c_str(), "wb");
  FILE* fsp = fopen(path_in(dir, "spectators.bin").c_str(), "wb");
  FILE* fc = fopen(path_in(dir, "collisions.bin").c_str(), "wb");
  if (!fs || !fp || !fsp || !fc) {
    for (FILE* f : {fs, fp, fsp, fc}) if (f) fclose(f);
    return sfail(ECHO_E_ARG, "echo_glauber_generate: cannot create the event files in ", dir ? dir : ".");
  }
  Rng rng(seed);
  // all configurations of A first, then of B, as the reference draws them
  std::vector<Nucleus> na(p->nconf), nb(p->nconf);
  bool ok = true;
  for (int i = 0; i < p->nconf && ok; i++) ok = sample_nucleus(rng, *p, na[i]);
  for (int i = 0; i < p->nconf && ok; i++) ok = sample_nucleus(rng, *p, nb[i]);
  int nev = 0;
  for (int i = 0; i < p->nconf && ok; i++) {
    const int nbc = p->fixed_b ? 1 : p->nbcoll;   // common.f90:1937
    for (int k = 0; k < nbc; k++) {
      const double u1 = rng.uniform(), u2 = rng.uniform();
      double b = u1 * p->bmax;
      if (!p->fixed_b && b <= p->bmax * u2 * 1.001) continue;
      if (p->fixed_b) b = p->bimpact;
      Event ev = collide(*p, na[i], nb[i], b);
      if (ev.nc < 1 || ev.npa + ev.npb < p->min_participants) continue;
      nev++;
      fprintf(fs, " ev= %d b= %.17g nc= %d np= %d npA= %d npB= %d nspecA= %d nspecB= %d\n", nev, ev.b, ev.nc,
              ev.npa + ev.npb, ev.npa, ev.npb, ev.nsa, ev.nsb);
      for (size_t r = 0; r < ev.ptype.size(); r++) {
        Rec20 rec{ev.part[2 * r], ev.part[2 * r + 1], ev.ptype[r]};
        ok = ok && fwrite(&rec, sizeof rec, 1, fp) == 1;
      }
      for (size_t r = 0; r < ev.stype.size(); r++) {
        Rec20 rec{ev.spec[2 * r], ev.spec[2 * r + 1], ev.stype[r]};
        ok = ok && fwrite(&rec, sizeof rec, 1, fsp) == 1;
      }
      for (int r = 0; r < ev.nc; r++) {
        Rec16 rec{ev.coll[2 * r], ev.coll[2 * r + 1]};
        ok = ok && fwrite(&rec, sizeof rec, 1, fc) == 1;
      }
    }
  }
  const bool closed = fclose(fs) == 0 && fclose(fp) == 0 && fclose(fsp) == 0 && fclose(fc) == 0;
  if (!ok || !closed) return sfail(ECHO_E_ARG, "echo_glauber_generate: sampling or writing failed in ", dir ? dir : ".");
  *n_events = nev;
  return 0;
}

int echo_glauber_read_event(const char* dir, int32_t iev, echo_glauber_event* ev, int32_t cap_part,
                            int32_t cap_spec, int32_t cap_coll) {
  if (!ev) return sfail(ECHO_E_ARG, "echo_glauber_read_event: null argument");
  std::vector<SummaryRow> rows;
  if (!read_summary(dir, rows)) return sfail(ECHO_E_ARG, "echo_glauber_read_event: no summary.dat in ", dir ? dir : ".");
  if (iev < 1 || iev > (int)rows.size())
    return sfail(ECHO_E_ARG, "echo_glauber_read_event: no such event in ", dir ? dir : ".");
  long op = 0, os = 0, oc = 0;
  for (int i = 0; i < iev - 1; i++) { op += rows[i].np; os += rows[i].nsa + rows[i].nsb; oc += rows[i].nc; }
  const SummaryRow& r = rows[iev - 1];
  ev->b = r.b;
  ev->n_part_a = r.npa; ev->n_part_b = r.npb;
  ev->n_spec_a = r.nsa; ev->n_spec_b = r.nsb;
  ev->n_coll = r.nc;
  if (r.np > cap_part || r.nsa + r.nsb > cap_spec || r.nc > cap_coll)
    return sfail(ECHO_E_ARG, "echo_glauber_read_event: arrays too small for the event");
  bool ok = true;
  auto rd20 = [&](const char* name, long off, int n, double* xy, int32_t* t) {
    if (n == 0) return;
    FILE* f = fopen(path_in(dir, name).c_str(), "rb");
    ok = ok && f && fseek(f, off * (long)sizeof(Rec20), SEEK_SET) == 0;
    for (int i = 0; i < n && ok; i++) {
      Rec20 rec;
      ok = fread(&rec, sizeof rec, 1, f) == 1;
      xy[2 * i] = rec.x; xy[2 * i + 1] = rec.y;
      if (t) t[i] = rec.t;
    }
    if (f) fclose(f);
  };
  rd20("partcoll.bin", op, r.np, ev->part_xy, ev->part_type);
  rd20("spectators.bin", os, r.nsa + r.nsb, ev->spec_xy, ev->spec_type);
  if (r.nc > 0) {
    FILE* f = fopen(path_in(dir, "collisions.bin").c_str(), "rb");
    ok = ok && f && fseek(f, oc * (long)sizeof(Rec16), SEEK_SET) == 0;
    for (int i = 0; i < r.nc && ok; i++) {
      Rec16 rec;
      ok = fread(&rec, sizeof rec, 1, f) == 1;
      ev->coll_xy[2 * i] = rec.x; ev->coll_xy[2 * i + 1] = rec.y;
    }
    if (f) fclose(f);
  }
  if (!ok) return sfail(ECHO_E_ARG, "echo_glauber_read_event: event files truncated in ", dir ? dir : ".");
  return 0;
}

}  // extern "C"
