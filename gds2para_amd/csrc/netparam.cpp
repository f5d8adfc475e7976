// netparam.cpp -- network-parameter output of the solved Z (host only):
//   Parasitics::convertParam        src/solnoutclass.hpp:1456-1600  (Z / Y / S)
//   SolverDataBase::dumpCITI        src/solnoutclass.hpp:6040-6100
//   SolverDataBase::dumpTouchstone  src/solnoutclass.hpp:6104-6215
//   SolverDataBase::dumpXyce        src/solnoutclass.hpp:6032-6036 (via
//     Parasitics::computeYBusFromParam :1680-1766 and toXyce :2094-2248)
// Matrices are nf blocks of np x np complex values (interleaved re, im),
// row-major: mat[(f np + i) np + j] is entry (i, j) at frequency f -- the
// layout Parasitics::saveNetworkParam reads sys->x with (:1394-1420), so the
// Z array the solves return is passed as it is.
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <ctime>
#include <string>
#include <vector>

#include "gds2para.h"
#include "gp_status.hpp"

namespace {

using cd = std::complex<double>;
using Mat = std::vector<cd>;  // n x n row-major

Mat eye(int n) {
  Mat m((size_t)n * n, cd(0., 0.));
  for (int i = 0; i < n; ++i) m[(size_t)i * n + i] = 1.;
  return m;
}

Mat mul(const Mat &a, const Mat &b, int n) {
  Mat c((size_t)n * n, cd(0., 0.));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const cd aik = a[(size_t)i * n + k];
      for (int j = 0; j < n; ++j) c[(size_t)i * n + j] += aik * b[(size_t)k * n + j];
    }
  return c;
}

Mat add(const Mat &a, const Mat &b, double sb) {
  Mat c(a.size());
  for (size_t q = 0; q < a.size(); ++q) c[q] = a[q] + sb * b[q];
  return c;
}

// D a D for the diagonal D = diag(d)
Mat dscale(const std::vector<double> &d, const Mat &a, int n) {
  Mat c(a.size());
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) c[(size_t)i * n + j] = d[i] * a[(size_t)i * n + j] * d[j];
  return c;
}

// inverse by Gauss-Jordan elimination with partial pivoting; false if singular
bool inverse(Mat a, int n, Mat *out) {
  Mat b = eye(n);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[(size_t)r * n + c]) > std::abs(a[(size_t)p * n + c])) p = r;
    if (std::abs(a[(size_t)p * n + c]) == 0.) return false;
    if (p != c)
      for (int j = 0; j < n; ++j) {
        std::swap(a[(size_t)p * n + j], a[(size_t)c * n + j]);
        std::swap(b[(size_t)p * n + j], b[(size_t)c * n + j]);
      }
    const cd inv = 1. / a[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) { a[(size_t)c * n + j] *= inv; b[(size_t)c * n + j] *= inv; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const cd f = a[(size_t)r * n + c];
      if (f == 0.) continue;
      for (int j = 0; j < n; ++j) {
        a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
        b[(size_t)r * n + j] -= f * b[(size_t)c * n + j];
      }
    }
  }
  *out = b;
  return true;
}

std::string fmt_g(double v, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*g", prec, v);
  return buf;
}

// ostream << setw(w) << setprecision(p) << v with std::left (left = true) or std::right
std::string field(double v, int prec, int w, bool left) {
  std::string s = fmt_g(v, prec);
  if ((int)s.size() >= w) return s;
  return left ? s + std::string(w - s.size(), ' ') : std::string(w - s.size(), ' ') + s;
}

std::string stamp(const char *given, const char *pattern) {
  if (given) return given;
  char buf[80];
  std::time_t t = std::time(nullptr);
  std::strftime(buf, sizeof buf, pattern, std::localtime(&t));
  return buf;
}

bool valid_param(char p) { return p == 'Z' || p == 'Y' || p == 'S'; }

}  // namespace

using gp::fail;

extern "C" gp_status gp_convert_param(int32_t np, int32_t nf, char from, char to, const double *zsource,
                                      const double *in, double *out) {
  if (np < 1 || nf < 0 || !in || !out || !valid_param(from) || !valid_param(to))
    return fail(GP_ERR_ARG, "gp_convert_param: bad arguments (parameters are 'Z', 'Y' or 'S')");
  const bool need_z0 = (from == 'S') != (to == 'S');
  if (need_z0 && !zsource) return fail(GP_ERR_ARG, "gp_convert_param: S conversions need the port impedances");
  if (need_z0)
    for (int i = 0; i < np; ++i)
      if (!(zsource[i] > 0.)) return fail(GP_ERR_ARG, "gp_convert_param: port impedances must be positive");
  const size_t nn = (size_t)np * np;
  std::vector<double> zr(np), yr(np);
  for (int i = 0; i < np && need_z0; ++i) {
    zr[i] = std::sqrt(zsource[i]);        // diagPortZ (:1481)
    yr[i] = std::sqrt(1.0 / zsource[i]);  // diagPortY (:1523)
  }
  const Mat I = eye(np);
  for (int f = 0; f < nf; ++f) {
    Mat P(nn), R;
    for (size_t q = 0; q < nn; ++q) P[q] = cd(in[2 * (f * nn + q)], in[2 * (f * nn + q) + 1]);
    bool ok = true;
    if (from == to) {
      R = P;
    } else if ((from == 'Z' && to == 'Y') || (from == 'Y' && to == 'Z')) {
      ok = inverse(P, np, &R);                                   // :1467-1472, :1508-1512
    } else if (from == 'Y' && to == 'S') {                       // :1474-1494
      const Mat M = dscale(zr, P, np);
      Mat inv;
      ok = inverse(add(I, M, 1.), np, &inv);
      if (ok) R = mul(add(I, M, -1.), inv, np);
    } else if (from == 'Z' && to == 'S') {                       // :1516-1536
      const Mat M = dscale(yr, P, np);
      Mat inv;
      ok = inverse(add(M, I, 1.), np, &inv);
      if (ok) R = mul(add(M, I, -1.), inv, np);
    } else if (from == 'S' && to == 'Y') {                       // :1553-1573
      Mat inv;
      ok = inverse(add(I, P, 1.), np, &inv);
      if (ok) R = dscale(yr, mul(add(I, P, -1.), inv, np), np);
    } else {                                                     // S -> Z, :1574-1594
      Mat inv;
      ok = inverse(add(I, P, -1.), np, &inv);
      if (ok) R = dscale(zr, mul(add(I, P, 1.), inv, np), np);
    }
    if (!ok) return fail(GP_ERR_STATE, "gp_convert_param: singular matrix at frequency index " + std::to_string(f));
    for (size_t q = 0; q < nn; ++q) {
      out[2 * (f * nn + q)] = R[q].real();
      out[2 * (f * nn + q) + 1] = R[q].imag();
    }
  }
  return GP_OK;
}

extern "C" gp_status gp_write_citi(const char *path, const char *design, const char *timestamp, char param,
                                   int32_t np, int32_t nf, const double *freqs, const double *mat) {
  if (!path || !design || np < 1 || nf < 0 || (nf > 0 && (!freqs || !mat)) || !valid_param(param))
    return fail(GP_ERR_ARG, "gp_write_citi: bad arguments");
  std::string name(design);
  for (char &c : name) c = (char)std::toupper((unsigned char)c);
  std::string o;
  o += "CITIFILE A.01.01\n";
  o += "CONSTANT TIME " + stamp(timestamp, "%Y %m %d %H %M %S") + "\n";
  o += "NAME " + name + "\n";
  o += "VAR FREQ MAG " + std::to_string(nf) + "\n";
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j)
      o += std::string("DATA ") + param + "[" + std::to_string(i + 1) + "," + std::to_string(j + 1) + "] RI\n";
  o += "VAR_LIST_BEGIN\n";
  for (int f = 0; f < nf; ++f) o += field(freqs[f], 7, 13, true) + "\n";
  o += "VAR_LIST_END\n";
  const size_t nn = (size_t)np * np;
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      o += "BEGIN\n";
      for (int f = 0; f < nf; ++f) {
        const size_t q = f * nn + (size_t)i * np + j;
        o += field(mat[2 * q], 7, 13, false) + "," + field(mat[2 * q + 1], 7, 13, true) + "\n";
      }
      o += "END\n";
    }
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(GP_ERR_ARG, std::string("gp_write_citi: cannot open ") + path);
  const bool w = std::fwrite(o.data(), 1, o.size(), fp) == o.size();
  std::fclose(fp);
  return w ? GP_OK : fail(GP_ERR_STATE, std::string("gp_write_citi: short write to ") + path);
}

extern "C" gp_status gp_write_touchstone(const char *path, const char *design, const char *timestamp, char param,
                                         int32_t np, int32_t nf, const double *freqs, const double *mat,
                                         const double *zsource) {
  if (!path || !design || !zsource || np < 1 || nf < 0 || (nf > 0 && (!freqs || !mat)) || !valid_param(param))
    return fail(GP_ERR_ARG, "gp_write_touchstone: bad arguments");
  std::string o;
  const std::string P(1, param);
  o += "[Version] 2.0\n";
  o += "# Hz " + P + " RI R " + fmt_g(zsource[0], 6) + "\n";
  o += "[Number of Ports] " + std::to_string(np) + "\n";
  if (np == 2) o += "[Two-Port Data Order] 12_21\n";
  o += "[Number of Frequencies] " + std::to_string(nf) + "\n";
  o += "[Reference]\n";
  for (int p = 0; p < np; ++p) o += fmt_g(zsource[p], 6) + " ";
  o += "\n[Matrix Format] Full\n";
  o += "! Date/Time " + stamp(timestamp, "%Y-%m-%d %H:%M:%S") + "\n";
  o += std::string("! Name ") + design + "\n";
  o += "! Provider: DARPA ERI IDEA/POSH Contributors\n";
  o += "! Program: Xyce Writer from DARPA ERI\n";
  o += "! Author: Purdue University\n";
  o += "[Network Data] \n";
  o += "!freq         ";
  for (int i = 0; i < np; ++i) {
    for (int j = 0; j < np; ++j) {
      const std::string ij = std::to_string(i + 1) + "," + std::to_string(j + 1);
      if (np > 2 && i > 0 && j == 0) o += "!Re" + P + ij + std::string(7 - i / 10 - j / 10, ' ');
      else o += "Re" + P + ij + std::string(8 - i / 10 - j / 10, ' ');
      o += "Im" + P + ij + std::string(8 - i / 10 - j / 10, ' ');
      if ((j + 1) % 4 == 0 && j + 1 != np && np > 2) o += "\n!" + std::string(13, ' ');
    }
    if (np > 2) o += "\n";
  }
  o += "\n";
  const size_t nn = (size_t)np * np;
  for (int f = 0; f < nf; ++f) {
    o += field(freqs[f], 7, 13, true) + " ";
    for (int i = 0; i < np; ++i) {
      for (int j = 0; j < np; ++j) {
        const size_t q = f * nn + (size_t)i * np + j;
        o += field(mat[2 * q], 7, 13, true) + " ";
        o += field(mat[2 * q + 1], 7, 13, true) + " ";
        if ((j + 1) % 4 == 0 && j + 1 != np && np > 2) o += "\n" + std::string(14, ' ');
      }
      if (np > 2) o += " ! row " + std::to_string(i + 1) + "\n";
    }
    if (np <= 2) o += "\n";
  }
  o += "[End]\n";
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(GP_ERR_ARG, std::string("gp_write_touchstone: cannot open ") + path);
  const bool w = std::fwrite(o.data(), 1, o.size(), fp) == o.size();
  std::fclose(fp);
  return w ? GP_OK : fail(GP_ERR_STATE, std::string("gp_write_touchstone: short write to ") + path);
}

namespace {

// Eigen's a.isApprox(b) for dense real matrices at double precision (1e-12):
// |a - b|_F^2 <= 1e-24 min(|a|_F^2, |b|_F^2)
bool is_approx_transpose(const std::vector<double> &a, int n) {
  double d2 = 0., a2 = 0.;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double x = a[(size_t)i * n + j], y = a[(size_t)j * n + i];
      d2 += (x - y) * (x - y);
      a2 += x * x;
    }
  return d2 <= 1e-24 * a2;  // |a^T|_F = |a|_F
}

// makeGSym / makeCSym (src/solnoutclass.hpp:1325-1335, 1382-1392): replaced
// by (a + a^T) / 2 unless a already passes isApprox against its transpose
void make_sym(std::vector<double> &a, int n) {
  if (is_approx_transpose(a, n)) return;
  std::vector<double> s(a.size());
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s[(size_t)i * n + j] = (a[(size_t)i * n + j] + a[(size_t)j * n + i]) * 0.5;
  a.swap(s);
}

// getGTotal / getCTotal (:1291-1315, 1348-1372): the upper triangle's sum
double upper_total(const std::vector<double> &a, int n) {
  double t = 0.;
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) t += a[(size_t)i * n + j];
  return t;
}

// getGNodeGround / getCNodeGround (:1280-1288, 1337-1345): the row sum
double row_sum(const std::vector<double> &a, int n, int i) {
  double t = 0.;
  for (int j = 0; j < n; ++j) t += a[(size_t)i * n + j];
  return t;
}

}  // namespace

// Parasitics::computeYBusFromParam(0) (src/solnoutclass.hpp:1680-1766, the
// dense branch built by default: EIGEN_SPARSE is commented out at :31) and
// Parasitics::toXyce (:2094-2248, dense branch), as SolverDataBase::dumpXyce
// calls them with WRITE_THRESH (:33, :6032-6036).
extern "C" gp_status gp_write_xyce(const char *path, const char *design, const char *timestamp, char param,
                                   int32_t np, const char *const *port_names, double f0, const double *mat,
                                   double save_thresh) {
  if (!path || !design || np < 1 || !port_names || !mat || !(f0 > 0.)) return fail(GP_ERR_ARG, "bad Xyce arguments");
  for (int32_t i = 0; i < np; ++i)
    if (!port_names[i]) return fail(GP_ERR_ARG, "null port name");
  if (param != 'Z' && param != 'Y')
    return fail(GP_ERR_ARG, "param must be 'Z' or 'Y' (the reference's 'S' branch leaves its Y matrix empty)");
  const int n = np;
  Mat P((size_t)n * n);
  for (size_t q = 0; q < P.size(); ++q) P[q] = cd(mat[2 * q], mat[2 * q + 1]);
  // 'Z' is used as Y without inversion and 'Y' is inverted, as the reference
  // does (:1688-1699)
  Mat Y;
  if (param == 'Y') {
    if (!inverse(P, n, &Y)) return fail(GP_ERR_ARG, "singular Y matrix");
  } else {
    Y = P;
  }
  const double w0 = 2. * M_PI * f0;
  std::vector<double> G((size_t)n * n), Cm((size_t)n * n);
  for (size_t q = 0; q < Y.size(); ++q) {
    G[q] = Y[q].real();                 // G = Re{Y}
    Cm[q] = -1. / (Y[q].imag() * w0);   // C = Im{Y} (2 pi f), then -1 / C entrywise
  }
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(GP_ERR_ARG, std::string("cannot open ") + path);
  const std::string when = stamp(timestamp, "%d-%m-%Y %H:%M:%S");
  std::fprintf(fp, "* File: %s\n* Design: %s\n* Vendor: DARPA ERI Contributors\n", path, design);
  std::fprintf(fp, "* Program: Xyce Writer from DARPA ERI\n* Author: Purdue University\n* Date: %s\n\n", when.c_str());
  std::fprintf(fp, ".subckt %s", design);
  for (int i = 0; i < n; ++i) std::fprintf(fp, " %s", port_names[i]);
  std::fprintf(fp, "\n");
  make_sym(Cm, n);
  make_sym(G, n);
  const double capTot = upper_total(Cm, n), condTot = upper_total(G, n);
  int numCap = 1, numInd = 1, numRes = 1;
  const double w02 = std::pow(w0, 2.);
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      if (i == j && std::abs(row_sum(Cm, n, i)) >= save_thresh * capTot) {
        const double cng = row_sum(Cm, n, i);
        if (cng > 0.) std::fprintf(fp, "C%d %s 0 %s\n", numCap++, port_names[i], fmt_g(cng, 6).c_str());
        else std::fprintf(fp, "L%d %s 0 %s\n", numInd++, port_names[i], fmt_g(-1. / (cng * w02), 6).c_str());
      } else if (j > i && std::abs(-Cm[(size_t)i * n + j]) >= save_thresh * capTot) {
        const double cnn = -1. * Cm[(size_t)i * n + j];
        if (cnn > 0.)
          std::fprintf(fp, "C%d %s %s %s\n", numCap++, port_names[i], port_names[j], fmt_g(cnn, 6).c_str());
        else
          std::fprintf(fp, "L%d %s %s %s\n", numInd++, port_names[i], port_names[j], fmt_g(-1. / (cnn * w02), 6).c_str());
      }
    }
    for (int j = 0; j < n; ++j) {
      // the coupling test compares a resistance with the conductance total,
      // as the reference does (:2207)
      if (i == j && std::abs(row_sum(G, n, i)) >= save_thresh * condTot) {
        std::fprintf(fp, "R%d %s 0 %s\n", numRes++, port_names[i], fmt_g(std::abs(1. / row_sum(G, n, i)), 6).c_str());
      } else if (j > i && std::abs(-1. / G[(size_t)i * n + j]) >= save_thresh * condTot) {
        std::fprintf(fp, "R%d %s %s %s\n", numRes++, port_names[i], port_names[j],
                     fmt_g(std::abs(-1. / G[(size_t)i * n + j]), 6).c_str());
      }
    }
  }
  std::fprintf(fp, ".ends\n");
  const bool ok = std::fclose(fp) == 0;
  return ok ? GP_OK : fail(GP_ERR_ARG, std::string("cannot write ") + path);
}

// Parasitics::toSPEF + SolverDataBase::dumpSPEF (src/solnoutclass.hpp:1972-2092,
// 6008-6030): the elements of the Xyce reduction (computeYBusFromParam at the
// lowest frequency, C and G symmetrized unless already symmetric) as one
// IEEE 1481-1998 net "all": a ground capacitance per port (its C row sum)
// when |row sum| >= save_thresh * C total, a coupling capacitance -C_ij for
// j > i when |C_ij| >= save_thresh * C total, a ground resistance
// |1 / (G row sum)| and a coupling resistance |1 / G_ij| under the same rule
// against the G total (the reference compares the coupling resistance, not
// the conductance, with that total).  Unlike toXyce, negative capacitances
// stay capacitances.  The reference prints the record through Parser-SPEF's
// dump (absent here); the text follows the standard's grammar, with a
// node-to-ground resistance written with its one node as the reference
// records it.
extern "C" gp_status gp_write_spef(const char *path, const char *design, const char *timestamp, char param,
                                   int32_t np, const char *const *port_names, const char *port_dirs, double f0,
                                   const double *mat, double save_thresh) {
  if (!path || !design || np < 1 || !port_names || !mat || !(f0 > 0.)) return fail(GP_ERR_ARG, "bad SPEF arguments");
  for (int32_t i = 0; i < np; ++i)
    if (!port_names[i]) return fail(GP_ERR_ARG, "null port name");
  if (param != 'Z' && param != 'Y')
    return fail(GP_ERR_ARG, "param must be 'Z' or 'Y' (the reference's 'S' branch leaves its Y matrix empty)");
  const int n = np;
  Mat P((size_t)n * n);
  for (size_t q = 0; q < P.size(); ++q) P[q] = cd(mat[2 * q], mat[2 * q + 1]);
  Mat Y;
  if (param == 'Y') {
    if (!inverse(P, n, &Y)) return fail(GP_ERR_ARG, "singular Y matrix");
  } else {
    Y = P;
  }
  const double w0 = 2. * M_PI * f0;
  std::vector<double> G((size_t)n * n), Cm((size_t)n * n);
  for (size_t q = 0; q < Y.size(); ++q) {
    G[q] = Y[q].real();
    Cm[q] = -1. / (Y[q].imag() * w0);
  }
  make_sym(Cm, n);
  make_sym(G, n);
  const double capTot = upper_total(Cm, n), condTot = upper_total(G, n);
  auto dir_of = [&](int i) -> const char * {
    const char d = port_dirs ? port_dirs[i] : 'B';
    return d == 'O' ? "O" : d == 'I' ? "I" : "B";   // unclear directions are bidirectional (:2014)
  };
  FILE *fp = std::fopen(path, "w");
  if (!fp) return fail(GP_ERR_ARG, std::string("cannot open ") + path);
  const std::string when = stamp(timestamp, "%d-%m-%Y %H:%M:%S");
  std::fprintf(fp, "*SPEF \"IEEE 1481-1998\"\n*DESIGN \"%s\"\n*DATE \"%s\"\n", design, when.c_str());
  std::fprintf(fp, "*VENDOR \"DARPA ERI Contributors\"\n*PROGRAM \"SPEF Writer from DARPA ERI\"\n");
  std::fprintf(fp, "*VERSION \"1.0\"\n*DESIGN_FLOW \"NETLIST_TYPE_VERILOG\"\n*DIVIDER /\n*DELIMITER :\n");
  std::fprintf(fp, "*BUS_DELIMITER [ ]\n*T_UNIT 1 S\n*C_UNIT 1 F\n*R_UNIT 1 OHM\n*L_UNIT 1 H\n\n*NAME_MAP\n");
  for (int i = 0; i < n; ++i) std::fprintf(fp, "*%d %s\n", i + 1, port_names[i]);
  std::fprintf(fp, "\n*PORTS\n");
  for (int i = 0; i < n; ++i) std::fprintf(fp, "*%d %s\n", i + 1, dir_of(i));
  std::fprintf(fp, "\n*D_NET all %s\n*CONN\n", fmt_g(capTot, 6).c_str());
  for (int i = 0; i < n; ++i) std::fprintf(fp, "*P %s %s\n", port_names[i], dir_of(i));
  std::string caps, ress;
  int nc = 0, nr = 0;
  char buf[256];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      if (i == j && std::abs(row_sum(Cm, n, i)) >= save_thresh * capTot) {
        std::snprintf(buf, sizeof buf, "%d *%d %s\n", ++nc, i + 1, fmt_g(row_sum(Cm, n, i), 6).c_str());
        caps += buf;
      } else if (j > i && std::abs(-Cm[(size_t)i * n + j]) >= save_thresh * capTot) {
        std::snprintf(buf, sizeof buf, "%d *%d *%d %s\n", ++nc, i + 1, j + 1, fmt_g(-Cm[(size_t)i * n + j], 6).c_str());
        caps += buf;
      }
    }
    for (int j = 0; j < n; ++j) {
      if (i == j && std::abs(row_sum(G, n, i)) >= save_thresh * condTot) {
        std::snprintf(buf, sizeof buf, "%d *%d %s\n", ++nr, i + 1, fmt_g(std::abs(1. / row_sum(G, n, i)), 6).c_str());
        ress += buf;
      } else if (j > i && std::abs(-1. / G[(size_t)i * n + j]) >= save_thresh * condTot) {
        std::snprintf(buf, sizeof buf, "%d *%d *%d %s\n", ++nr, i + 1, j + 1,
                      fmt_g(std::abs(-1. / G[(size_t)i * n + j]), 6).c_str());
        ress += buf;
      }
    }
  }
  std::fprintf(fp, "*CAP\n%s*RES\n%s*END\n", caps.c_str(), ress.c_str());
  const bool ok = std::fclose(fp) == 0;
  return ok ? GP_OK : fail(GP_ERR_ARG, std::string("cannot write ") + path);
}
