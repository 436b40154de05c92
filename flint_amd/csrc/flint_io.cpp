// FLINT native input folder reader (flint_mech_load).
//
// Mirrors read_idealgas_thermo (Load_ThermoTransport.f90:7-89), read_idealgas_composition
// (:92-126) and read_chemistry (Load_Chemistry.f90:7-164): phase.txt gives species and molecular
// weights, thermo.dat one Tecplot zone per species (T, cp, h, s, dcp), chemistry-info.txt the
// mechanism name and reaction types, chemistry-Arrhenius.dat one zone per Arrhenius/three-body
// reaction (T, kf, kb), chemistry-Troe.dat (T, k_inf, k_0, k_c, F_cent) and
// chemistry-Lindemann.dat (T, k_inf, k_0, k_c) one zone per falloff reaction in file order.
// Only Tecplot ASCII (.dat) is read; FLINT's binary .szplt fallback needs TecIO and is not
// supported.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/flint_amd.h"
#include "flint_internal.h"

namespace {

struct Zones {
  std::vector<double> T0;                  // first temperature of each zone
  std::vector<std::vector<double>> vars;   // per zone: [nvar][npts] (variables after T)
  int npts = 0;
};

bool read_file(const std::string& path, std::string& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(len);
  size_t got = std::fread(&out[0], 1, len, f);
  std::fclose(f);
  return got == (size_t)len;
}

// POINT-format Tecplot ASCII with nvar variables after the temperature column.
int read_tecplot(const std::string& path, int nvar, Zones& z) {
  std::string buf;
  if (!read_file(path, buf)) return flint_fail(FLINT_EIO, "cannot read " + path);
  const char* p = buf.c_str();
  const char* end = p + buf.size();
  int npts = -1;
  std::vector<double>* cur = nullptr;
  int in_zone = 0;
  while (p < end) {
    const char* eol = (const char*)std::memchr(p, '\n', end - p);
    if (!eol) eol = end;
    const char* q = p;
    while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q < eol) {
      if (!std::strncmp(q, "TITLE", 5) || !std::strncmp(q, "VARIABLES", 9)) {
      } else if (!std::strncmp(q, "ZONE", 4)) {
        z.vars.emplace_back();
        cur = &z.vars.back();
        in_zone = 1;
      } else if (!std::strncmp(q, "I=", 2) || !std::strncmp(q, "I =", 3)) {
        npts = std::atoi(std::strchr(q, '=') + 1);
        if (z.npts == 0) z.npts = npts;
        if (npts != z.npts) return flint_fail(FLINT_EIO, path + ": zones of different length");
        cur->reserve((size_t)npts * nvar);
      } else if (in_zone) {
        char* e;
        double T = std::strtod(q, &e);
        if (cur->empty()) z.T0.push_back(T);
        for (int v = 0; v < nvar; ++v) {
          double x = std::strtod(e, &e);
          cur->push_back(x);
        }
      }
    }
    p = eol + 1;
  }
  // transpose each zone to [nvar][npts]
  for (auto& zz : z.vars) {
    if ((int)zz.size() != z.npts * nvar) return flint_fail(FLINT_EIO, path + ": truncated zone");
    std::vector<double> t((size_t)nvar * z.npts);
    for (int i = 0; i < z.npts; ++i)
      for (int v = 0; v < nvar; ++v) t[(size_t)v * z.npts + i] = zz[(size_t)i * nvar + v];
    zz.swap(t);
  }
  return FLINT_OK;
}

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  size_t b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

}  // namespace

extern "C" int flint_mech_load(const char* folder, const char* prefix, flint_mech** out) {
  if (!folder || !out) return flint_fail(FLINT_EARG, "null argument");
  *out = nullptr;
  const std::string base = std::string(folder) + "/" + (prefix ? prefix : "");
  // phase.txt
  std::ifstream ph(base + "phase.txt");
  if (!ph) return flint_fail(FLINT_EIO, "[ERROR] phase.txt type file not found: " + base + "phase.txt");
  std::string line;
  std::getline(ph, line);
  std::vector<std::string> names;
  std::vector<double> W;
  while (std::getline(ph, line)) {
    std::istringstream is(line);
    std::string nm;
    double w;
    if (is >> nm >> w) {
      names.push_back(nm);
      W.push_back(w);
    }
  }
  const int ns = (int)names.size();
  // composition.txt (optional, as in read_idealgas_composition)
  int ne = 0;
  std::vector<double> comp;
  std::vector<std::string> elements;
  {
    std::ifstream cf(base + "composition.txt");
    if (cf) {
      cf >> ne;
      std::string tok;
      for (int e = 0; e < ne; ++e) {
        cf >> tok;
        elements.push_back(tok);
      }
      cf >> tok;  // "composition"
      comp.assign((size_t)ne * ns, 0.0);
      for (int k = 0; k < ns; ++k)
        for (int e = 0; e < ne; ++e) cf >> comp[(size_t)k * ne + e];
    }
  }
  // thermo.dat
  Zones th;
  if (int rc = read_tecplot(base + "thermo.dat", 4, th)) return rc;
  if ((int)th.vars.size() != ns) return flint_fail(FLINT_EIO, "thermo.dat: zone count differs from phase.txt");
  int Th_min = (int)std::lround(th.T0[0]);
  const int Th_max = Th_min + th.npts - 1;
  const int start = Th_min;
  if (Th_min == 1) Th_min = 0;  // Load_ThermoTransport.f90:70,80-85
  const int nTh = Th_max - Th_min + 1;
  std::vector<double> cp((size_t)ns * nTh), h((size_t)ns * nTh), sv((size_t)ns * nTh), dc((size_t)ns * nTh);
  for (int k = 0; k < ns; ++k) {
    const auto& v = th.vars[k];
    for (int i = 0; i < th.npts; ++i) {
      cp[(size_t)k * nTh + (start - Th_min) + i] = v[(size_t)0 * th.npts + i];
      h[(size_t)k * nTh + (start - Th_min) + i] = v[(size_t)1 * th.npts + i];
      sv[(size_t)k * nTh + (start - Th_min) + i] = v[(size_t)2 * th.npts + i];
      dc[(size_t)k * nTh + (start - Th_min) + i] = v[(size_t)3 * th.npts + i];
    }
    if (start != Th_min) {
      cp[(size_t)k * nTh] = v[0];
      h[(size_t)k * nTh] = 0.0;
      sv[(size_t)k * nTh] = v[(size_t)2 * th.npts];
      dc[(size_t)k * nTh] = v[(size_t)3 * th.npts];
    }
  }
  // chemistry-info.txt
  std::ifstream ci(base + "chemistry-info.txt");
  if (!ci) return flint_fail(FLINT_EIO, "[WARNING] Chemistry data not found: " + base + "chemistry-info.txt");
  std::string mech_name;
  std::getline(ci, line);
  {
    std::istringstream is(line);
    is >> mech_name;
  }
  std::getline(ci, line);
  std::getline(ci, line);  // N.ro reactions = <nrc>   (A17,I4)
  int nrc = std::atoi(line.substr(line.find('=') + 1).c_str());
  for (int j = 0; j < 4; ++j) std::getline(ci, line);
  int na = 0, nt = 0, nl = 0;
  std::vector<int> kind(nrc);
  for (int j = 0; j < nrc; ++j) {
    std::getline(ci, line);
    if (line.find("Troe") != std::string::npos) kind[j] = 1, ++nt;
    else if (line.find("Lindemann") != std::string::npos) kind[j] = 2, ++nl;
    else kind[j] = 0, ++na;
  }
  // Reaction definition block of the data-driven general routine (Load_Chemistry.f90:70-89): for
  // every reaction in file order, ns+1 lines "j name ni1 ni2 epsch", stored per kind; optional
  const int nz = ns + 1;
  std::vector<double> g1((size_t)nrc * nz), g2((size_t)nrc * nz), ge((size_t)nrc * nz);
  bool general = false;
  if (std::getline(ci, line) && std::getline(ci, line) && line.find("Reaction definition") != std::string::npos) {
    std::vector<int> slot(nrc);
    int next = 0;
    for (int q = 0; q < 3; ++q)
      for (int j = 0; j < nrc; ++j)
        if (kind[j] == q) slot[j] = next++;
    general = true;
    for (int j = 0; j < nrc && general; ++j)
      for (int k = 0; k < nz && general; ++k) {
        if (!std::getline(ci, line)) {
          general = false;
          break;
        }
        std::istringstream is(line);
        int idx;
        std::string nm;
        double a, b, e;
        if (!(is >> idx >> nm >> a >> b >> e) || idx != j + 1) general = false;
        const size_t o = (size_t)slot[j] * nz + k;
        g1[o] = a;
        g2[o] = b;
        ge[o] = e;
      }
  }
  Zones ar;
  if (int rc = read_tecplot(base + "chemistry-Arrhenius.dat", 2, ar))
    return flint_fail(rc, "[ERROR] chemistry-Arrhenius file not found: " + std::string(flint_last_error()));
  if ((int)ar.vars.size() < na) return flint_fail(FLINT_EIO, "chemistry-Arrhenius.dat: too few zones");
  const int Tk_min = (int)std::lround(ar.T0[0]);
  const int nTk = ar.npts, Tk_max = Tk_min + nTk - 1;
  std::vector<double> kf((size_t)na * nTk), kb((size_t)na * nTk);
  for (int a = 0; a < na; ++a)
    for (int i = 0; i < nTk; ++i) {
      kf[(size_t)a * nTk + i] = ar.vars[a][i];
      kb[(size_t)a * nTk + i] = ar.vars[a][(size_t)nTk + i];
    }
  std::vector<double> troe, lind;
  if (nt > 0) {
    Zones tz;
    if (int rc = read_tecplot(base + "chemistry-Troe.dat", 4, tz))
      return flint_fail(rc, "[ERROR] chemistry-Troe file not found: " + std::string(flint_last_error()));
    if ((int)tz.vars.size() < nt || tz.npts != nTk) return flint_fail(FLINT_EIO, "chemistry-Troe.dat: shape");
    troe.resize((size_t)4 * nt * nTk);
    for (int q = 0; q < nt; ++q)
      for (int v = 0; v < 4; ++v)
        for (int i = 0; i < nTk; ++i) troe[((size_t)v * nt + q) * nTk + i] = tz.vars[q][(size_t)v * nTk + i];
  }
  if (nl > 0) {
    Zones lz;
    if (int rc = read_tecplot(base + "chemistry-Lindemann.dat", 3, lz))
      return flint_fail(rc, "[ERROR] chemistry-Lindemann file not found: " + std::string(flint_last_error()));
    if ((int)lz.vars.size() < nl || lz.npts != nTk) return flint_fail(FLINT_EIO, "chemistry-Lindemann.dat: shape");
    lind.resize((size_t)3 * nl * nTk);
    for (int q = 0; q < nl; ++q)
      for (int v = 0; v < 3; ++v)
        for (int i = 0; i < nTk; ++i) lind[((size_t)v * nl + q) * nTk + i] = lz.vars[q][(size_t)v * nTk + i];
  }
  // species_composition(ne, ns) column-major == comp[k*ne + e]
  int rc = flint_mech_create(mech_name.c_str(), ns, W.data(), Tk_min, Tk_max, na, kf.data(), kb.data(), nt,
                             nt ? troe.data() : nullptr, nl, nl ? lind.data() : nullptr, Th_min, Th_max, cp.data(),
                             h.data(), ne, ne ? comp.data() : nullptr, out);
  if (rc) return rc;
  (*out)->species = names;
  (*out)->elements = elements;
  (*out)->t_s = sv;
  (*out)->t_dcpi = dc;
  for (const auto& e : elements) (*out)->el_weight.push_back(flint_atomic_weight(e.c_str()));
  if (general) {
    if (int rc2 = flint_general_setup(*out, na, nt, nl, g1.data(), g2.data(), ge.data())) {
      flint_mech_destroy(*out);
      *out = nullptr;
      return rc2;
    }
  }
  // transport.dat (optional): one zone per species, (T, mu, k), as read_idealgas_transport
  // (Load_ThermoTransport.f90:129-174), row Tmin = 0 copied from row 1 when the file starts at 1
  if (FILE* tf = std::fopen((base + "transport.dat").c_str(), "rb")) {
    std::fclose(tf);
    Zones tz;
    int rc2 = read_tecplot(base + "transport.dat", 2, tz);
    if (!rc2 && (int)tz.vars.size() != ns) rc2 = flint_fail(FLINT_EIO, "transport.dat: zone count differs from phase.txt");
    if (!rc2) {
      int Tt_min = (int)std::lround(tz.T0[0]);
      const int Tt_max = Tt_min + tz.npts - 1, t0 = Tt_min;
      if (Tt_min == 1) Tt_min = 0;
      const size_t nT = (size_t)(Tt_max - Tt_min + 1);
      std::vector<double> mu((size_t)ns * nT), kt((size_t)ns * nT);
      for (int k = 0; k < ns; ++k) {
        const auto& v = tz.vars[k];
        for (int i = 0; i < tz.npts; ++i) {
          mu[(size_t)k * nT + (t0 - Tt_min) + i] = v[i];
          kt[(size_t)k * nT + (t0 - Tt_min) + i] = v[(size_t)tz.npts + i];
        }
        if (t0 != Tt_min) {
          mu[(size_t)k * nT] = v[0];
          kt[(size_t)k * nT] = v[(size_t)tz.npts];
        }
      }
      rc2 = flint_transport_setup(*out, Tt_min, Tt_max, mu.data(), kt.data());
    }
    if (rc2) {
      flint_mech_destroy(*out);
      *out = nullptr;
      return rc2;
    }
  }
  if (ne == 0) return FLINT_OK;
  // equilibrium data (CEA_initialize_global), where the mechanism has an equilibrium kernel and
  // every element has a standard atomic weight
  std::vector<double> elw;
  for (const auto& e : elements) {
    const double w = flint_atomic_weight(e.c_str());
    if (!(w > 0.0)) return FLINT_OK;
    elw.push_back(w);
  }
  if ((*out)->ops && (*out)->ops->cea) {
    if (int rc2 = flint_cea_setup(*out, sv.data(), elw.data())) {
      flint_mech_destroy(*out);
      *out = nullptr;
      return rc2;
    }
  }
  return FLINT_OK;
}

// Standard atomic weights (IUPAC 1999, Coplen, J. Phys. Chem. Ref. Data 30 (2001) 701-712): the
// table the reference's equilibrium solver takes its element weights from by symbol
// (Lib_CEA_data.f90:37-59, Lib_CEA_setup.f90:157-170).  0 for an unknown symbol.
double flint_atomic_weight(const char* sym) {
  static const struct {
    const char* s;
    double w;
  } tab[] = {{"H", 1.00794},   {"D", 2.014102},    {"HE", 4.002602}, {"LI", 6.941},     {"BE", 9.012182},
             {"B", 10.811},    {"C", 12.0107},     {"N", 14.0067},   {"O", 15.9994},    {"F", 18.9984032},
             {"NE", 20.1797},  {"NA", 22.989770},  {"MG", 24.305},   {"AL", 26.981538}, {"SI", 28.0855},
             {"P", 30.973761}, {"S", 32.065},      {"CL", 35.453},   {"AR", 39.948},    {"K", 39.0983},
             {"CA", 40.078},   {"FE", 55.845},     {"KR", 83.80},    {"XE", 131.293}};
  std::string u;
  for (const char* c = sym; *c; ++c)
    if (!std::isspace((unsigned char)*c)) u.push_back((char)std::toupper((unsigned char)*c));
  for (const auto& t : tab)
    if (u == t.s) return t.w;
  return 0.0;
}
