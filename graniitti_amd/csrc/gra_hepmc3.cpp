// HepMC3 ASCII ("Asciiv3") output of generated events, host only.
//
// The reference writes events through HepMC3 3.2's WriterAscii
// (src/MGraniitti.cc:190-207 opens it with a GenRunInfo holding two tools;
// SaveEvent, src/MGraniitti.cc:1665-1712, adds a GenCrossSection attribute in
// picobarn and the event weight). The record itself is MProcess::CommonRecord
// (src/MProcess.cc:1588-1706) with WriteDecayKinematics (1287-1314):
//
//   vertex -1: beam 1 (status 4)  -> forward proton 1 (status 1) + propagator q1 (pdg 99, status 81)
//   vertex -2: beam 2 (status 4)  -> forward proton 2 (status 1) + propagator q2 (pdg 99, status 81)
//   vertex -3: q1 + q2            -> central system (pdg 90, status 81)
//   vertex -4: central system     -> central particles (status 1, or 2 when they decay)
//   vertex -5, -6: central particle i -> its two daughters (status 1)
//
// HepMC3 is not available here, so this is a writer of the published text
// layout: a header line with the format version, one T line per tool, and per
// event the E (number, vertex count, particle count), U (units), W (weights)
// and A (attribute) lines, then P lines (id, production parent, pdg, px, py,
// pz, E, mass, status) in particle-id order, each preceded by the V line
// (id, status, [incoming ids]) of its production vertex when that vertex has
// more than one incoming particle. A vertex with exactly one incoming particle
// and no position is implicit: its outgoing particles name the incoming
// particle's id as their parent. Momenta and masses are printed with %.16e
// (exact round trip of FP64); the mass is the one of the 4-momentum
// (sign(m^2) sqrt|m^2|), which is what HepMC3 prints for a particle without a
// set generated mass. Decay vertices carry no position (the reference draws an
// exponential proper time for them; for the rho that is of order 1e-12 mm), and
// events are numbered from 0 upwards (the reference leaves every event number
// at HepMC3's default 0).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "gra_internal.h"

struct gra_hepmc3 {
  FILE* fp = nullptr;
  std::string path;
  uint64_t next_event = 0;
  std::string buf;  // formatted lines of one call, flushed once per call
};

namespace {

struct V4 {
  double e, x, y, z;
};

V4 v4_of(const double p[4]) { return V4{p[0], p[1], p[2], p[3]}; }
V4 v4_sub(V4 a, V4 b) { return V4{a.e - b.e, a.x - b.x, a.y - b.y, a.z - b.z}; }
V4 v4_add(V4 a, V4 b) { return V4{a.e + b.e, a.x + b.x, a.y + b.y, a.z + b.z}; }

double signed_mass(const V4& v) {
  const double m2 = v.e * v.e - v.x * v.x - v.y * v.y - v.z * v.z;
  return m2 >= 0.0 ? std::sqrt(m2) : -std::sqrt(-m2);
}

// HepMC3 escapes a tool's fields: backslash -> "\\", newline -> "\|"; the
// fields of one T line are joined by the escaped newline.
std::string escape(const char* s) {
  std::string o;
  for (const char* c = s; c && *c; ++c) {
    if (*c == '\\')
      o += "\\\\";
    else if (*c == '\n')
      o += "\\|";
    else
      o += *c;
  }
  return o;
}

void append(std::string& b, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void append(std::string& b, const char* fmt, ...) {
  char line[512];
  va_list ap;
  va_start(ap, fmt);
  const int n = vsnprintf(line, sizeof(line), fmt, ap);
  va_end(ap);
  b.append(line, (size_t)(n < (int)sizeof(line) ? n : (int)sizeof(line) - 1));
}

void particle(std::string& b, int id, int parent, int pdg, const V4& p, int status) {
  append(b, "P %d %d %d %.16e %.16e %.16e %.16e %.16e %d\n", id, parent, pdg, p.x, p.y, p.z, p.e,
         signed_mass(p), status);
}

}  // namespace

extern "C" int gra_hepmc3_open(const char* path, const char* generator, const char* version,
                               const char* card, gra_hepmc3** out) {
  using gra_hip::set_global_error;
  if (!out) return set_global_error(GRA_E_ARG, "gra_hepmc3_open: null output pointer");
  *out = nullptr;
  if (!path || !*path) return set_global_error(GRA_E_ARG, "gra_hepmc3_open: empty path");
  FILE* fp = std::fopen(path, "w");
  if (!fp) return set_global_error(GRA_E_ARG, "gra_hepmc3_open: cannot open %s", path);
  auto* w = new gra_hepmc3;
  w->fp = fp;
  w->path = path;
  std::string hdr = "HepMC::Version 3.02.00\nHepMC::Asciiv3-START_EVENT_LISTING\n";
  // Tools of the run (src/MGraniitti.cc:192-199): generator, then steering card
  const std::string gen = escape(generator ? generator : "GRANIITTI");
  const std::string ver = escape(version ? version : "");
  hdr += "T " + gen + "\\|" + ver + "\\|Generator\n";
  if (card && *card) hdr += "T " + escape(card) + "\\|1.0\\|Steering card\n";
  if (std::fwrite(hdr.data(), 1, hdr.size(), fp) != hdr.size()) {
    std::fclose(fp);
    delete w;
    return set_global_error(GRA_E_ARG, "gra_hepmc3_open: write failed on %s", path);
  }
  *out = w;
  return GRA_OK;
}

extern "C" int gra_hepmc3_write(gra_hepmc3* w, const gra_cfg* cfg, const gra_record_pdg* pdg,
                                const gra_event* ev, uint64_t n, double sigma_barn,
                                double sigma_err_barn, int32_t weighted) {
  using gra_hip::set_global_error;
  if (!w || !w->fp) return set_global_error(GRA_E_STATE, "gra_hepmc3_write: writer not open");
  if (!cfg || !pdg || (n && !ev))
    return set_global_error(GRA_E_ARG, "gra_hepmc3_write: null argument");
  if (cfg->abi_size != (int32_t)sizeof(gra_cfg))
    return set_global_error(GRA_E_ARG, "gra_hepmc3_write: gra_cfg abi_size %d != %d",
                            cfg->abi_size, (int)sizeof(gra_cfg));
  const double mb = cfg->m_beam;
  if (!(cfg->beam_E1 > mb) || !(cfg->beam_E2 > mb))
    return set_global_error(GRA_E_ARG, "gra_hepmc3_write: beam energy below the beam mass");
  // MProcess::SetBeamEnergies (src/MProcess.cc:318-320): beam 1 along +z, beam 2 along -z
  const V4 b1{cfg->beam_E1, 0.0, 0.0, std::sqrt(cfg->beam_E1 * cfg->beam_E1 - mb * mb)};
  const V4 b2{cfg->beam_E2, 0.0, 0.0, -std::sqrt(cfg->beam_E2 * cfg->beam_E2 - mb * mb)};
  // Cross section in picobarn (src/MGraniitti.cc:1696-1701)
  const double xs = sigma_barn * 1e12, xs_err = sigma_err_barn * 1e12;

  for (uint64_t k = 0; k < n; ++k)
    if (ev[k].n_p != 4 && ev[k].n_p != 8)
      return set_global_error(GRA_E_ARG, "gra_hepmc3_write: event %llu has n_p = %d (4 or 8)",
                              (unsigned long long)k, ev[k].n_p);

  std::string& b = w->buf;
  b.clear();
  b.reserve((size_t)n * 2200);
  uint64_t num = w->next_event;
  for (uint64_t k = 0; k < n; ++k) {
    const gra_event& E = ev[k];
    const bool decays = E.n_p == 8;
    const V4 p1f = v4_of(E.p[0]), p2f = v4_of(E.p[1]);
    const V4 c1 = v4_of(E.p[2]), c2 = v4_of(E.p[3]);
    const V4 q1 = v4_sub(b1, p1f), q2 = v4_sub(b2, p2f);
    const V4 sys = v4_add(c1, c2);  // pfinal[0], the central system
    const int n_vtx = decays ? 6 : 4;
    const int n_par = decays ? 13 : 9;

    append(b, "E %llu %d %d\n", (unsigned long long)num++, n_vtx, n_par);
    b += "U GEV MM\n";
    append(b, "W %.16e\n", weighted ? E.f : 1.0);
    // GenCrossSection::to_string: xs, xs_err, accepted, attempted (-1: unset), 8 digits
    append(b, "A 0 GenCrossSection %.8e %.8e -1 -1\n", xs, xs_err);
    // particle ids follow the order CommonRecord adds them to the event
    particle(b, 1, 0, pdg->beam[0], b1, 4);
    particle(b, 2, 1, pdg->beam[0], p1f, 1);
    particle(b, 3, 1, 99, q1, 81);
    particle(b, 4, 0, pdg->beam[1], b2, 4);
    particle(b, 5, 4, pdg->beam[1], p2f, 1);
    particle(b, 6, 4, 99, q2, 81);
    b += "V -3 0 [3,6]\n";
    particle(b, 7, -3, 90, sys, 81);
    if (!decays) {
      particle(b, 8, 7, pdg->central[0], c1, 1);
      particle(b, 9, 7, pdg->central[1], c2, 1);
    } else {
      particle(b, 8, 7, pdg->central[0], c1, 2);
      particle(b, 9, 8, pdg->daughter[0][0], v4_of(E.p[4]), 1);
      particle(b, 10, 8, pdg->daughter[0][1], v4_of(E.p[5]), 1);
      particle(b, 11, 7, pdg->central[1], c2, 2);
      particle(b, 12, 11, pdg->daughter[1][0], v4_of(E.p[6]), 1);
      particle(b, 13, 11, pdg->daughter[1][1], v4_of(E.p[7]), 1);
    }
  }
  if (!b.empty() && std::fwrite(b.data(), 1, b.size(), w->fp) != b.size())
    return set_global_error(GRA_E_ARG, "gra_hepmc3_write: write failed on %s", w->path.c_str());
  w->next_event = num;
  return GRA_OK;
}

extern "C" int gra_hepmc3_close(gra_hepmc3* w) {
  if (!w) return GRA_OK;
  int rc = GRA_OK;
  if (w->fp) {
    static const char footer[] = "HepMC::Asciiv3-END_EVENT_LISTING\n\n";
    if (std::fwrite(footer, 1, sizeof(footer) - 1, w->fp) != sizeof(footer) - 1 ||
        std::fclose(w->fp) != 0)
      rc = gra_hip::set_global_error(GRA_E_ARG, "gra_hepmc3_close: write failed on %s",
                                     w->path.c_str());
  }
  delete w;
  return rc;
}
