// Precomputation cache: the reference's precomputations/ text files.
//
// Layout (LS_Assembly.hpp:1166-1184, 1368-1420), under a working directory D:
//   D/precomputations/M/M_{n}_{int kappa}_{p}/M_{n}_{l}                 l = 0..10
//   D/precomputations/Neighbor/Neighbor_{n}_{int kappa}_{p}/
//       {colleague,fine,separatedFine,coarse}NeighborInteraction_{j}_{k}  j = 2..L
// n Chebyshev nodes per side, p the basis degree.  Each file is one complex
// matrix as Eigen's operator<< prints it: 6 significant digits, "(re,im)",
// every coefficient right-aligned to the widest one, one space between
// coefficients, one line per row.  The reader follows ReadFromTextFile
// (LS_Assembly.hpp:1186-1215): rows = number of newlines, columns = words /
// rows, and the file must hold rows x columns words.
//
// The reference keys a cache by int(kappa) only.  Each keyed directory here
// also gets a file aifmm_key with the exact kappa, domain size, n and p; a
// reader refuses a cache whose aifmm_key disagrees with the problem and
// accepts one without it (as written by the reference).
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "problem.h"

namespace aifmm {

static const char* const FAMILY_FILE[4] = {"colleagueNeighborInteraction", "fineNeighborInteraction",
                                           "separatedFineNeighborInteraction", "coarseNeighborInteraction"};

static std::string key_tail(const Params& P) {
  return std::to_string(P.n_cheb) + "_" + std::to_string(int(P.kappa)) + "_" + std::to_string(P.degree);
}
static std::string m_dir(const std::string& D, const Params& P) {
  return D + "/precomputations/M/M_" + key_tail(P);
}
static std::string nb_dir(const std::string& D, const Params& P) {
  return D + "/precomputations/Neighbor/Neighbor_" + key_tail(P);
}

static void make_dir(const std::string& path) {
  if (mkdir(path.c_str(), 0775) != 0 && errno != EEXIST)
    throw std::runtime_error("cannot create directory " + path + ": " + strerror(errno));
}

static std::string key_text(const Params& P) {
  char b[256];
  snprintf(b, sizeof b, "n_cheb %d\ndegree %d\nkappa %a\nL %a\n", P.n_cheb, P.degree, P.kappa, P.L);
  return b;
}

static void write_file(const std::string& path, const std::string& text) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot open " + path + " for writing: " + strerror(errno));
  bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
  ok = (fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("cannot write " + path);
}

static bool read_file(const std::string& path, std::string& text) {
  std::ifstream in(path, std::ios::binary);
  if (!in.good()) return false;
  std::ostringstream ss;
  ss << in.rdbuf();
  text = ss.str();
  return true;
}

static std::string fmt6(double v) {
  char b[64];
  snprintf(b, sizeof b, "%.6g", v);  // the stream default: precision 6, no floatfield
  return b;
}

// A column-major complex matrix as Eigen prints it, followed by a newline.
static std::string eigen_text(const double2* a, int rows, int cols) {
  std::vector<std::string> t((size_t)rows * cols);
  size_t w = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    t[i] = "(" + fmt6(a[i].x) + "," + fmt6(a[i].y) + ")";
    w = t[i].size() > w ? t[i].size() : w;
  }
  std::string out;
  out.reserve((size_t)rows * cols * (w + 1) + rows);
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) {
      const std::string& s = t[i + (size_t)rows * j];
      if (j) out += ' ';
      out.append(w - s.size(), ' ');
      out += s;
    }
    out += '\n';
  }
  return out;
}

// One std::complex<double> as operator>> reads it: "(re,im)", "(re)" or "re".
static bool parse_complex(const std::string& tok, double2& z) {
  const char* s = tok.c_str();
  char* e = nullptr;
  if (*s != '(') {
    z.x = strtod(s, &e);
    z.y = 0.0;
    return e != s && *e == '\0';
  }
  z.x = strtod(s + 1, &e);
  if (e == s + 1) return false;
  if (*e == ')') {
    z.y = 0.0;
    return e[1] == '\0';
  }
  if (*e != ',') return false;
  const char* s2 = e + 1;
  z.y = strtod(s2, &e);
  return e != s2 && e[0] == ')' && e[1] == '\0';
}

// ReadFromTextFile (LS_Assembly.hpp:1186-1215) into a column-major matrix of
// the expected shape.
static std::vector<double2> read_matrix(const std::string& path, int rows, int cols) {
  std::string text;
  if (!read_file(path, text)) throw std::runtime_error("could not open file \"" + path + "\" for reading");
  long lines = 0;
  for (char c : text) lines += c == '\n';
  std::vector<std::string> words;
  {
    std::istringstream in(text);
    std::string w;
    while (in >> w) words.push_back(w);
  }
  if (lines == 0 || (long)words.size() % lines != 0)
    throw std::runtime_error("file " + path + " cannot form a matrix");
  long c = (long)words.size() / lines;
  if (lines != rows || c != cols)
    throw std::runtime_error("file " + path + " holds a " + std::to_string(lines) + " x " + std::to_string(c) +
                             " matrix, expected " + std::to_string(rows) + " x " + std::to_string(cols));
  std::vector<double2> a((size_t)rows * cols);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j)
      if (!parse_complex(words[(size_t)i * cols + j], a[i + (size_t)rows * j]))
        throw std::runtime_error("file " + path + ": cannot read \"" + words[(size_t)i * cols + j] + "\"");
  return a;
}

static void check_key(const std::string& dir, const Params& P) {
  std::string text;
  if (!read_file(dir + "/aifmm_key", text)) return;  // written by the reference: keyed by int(kappa) only
  if (text != key_text(P))
    throw std::runtime_error("precomputations in " + dir + " are for another problem (aifmm_key: " + text + ")");
}

double Problem::write_precomputations(const std::string& D) {
  std::vector<TableSpec> specs;
  for (int l = 0; l <= FAR_LEVELS; ++l) specs.push_back({-1, l, 0});
  for (int j = 2; j <= L; ++j)
    for (int f = 0; f < 4; ++f)
      for (int n = 0; n < FAMILY_SIZE[f]; ++n) specs.push_back({f, j, n});
  auto t0 = std::chrono::steady_clock::now();
  int capped = 0;
  auto tabs = quad_tables(specs, &capped);
  t_precompute = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  scratch.release();  // the batch is large; later calls regrow what they need
  make_dir(D + "/precomputations");
  make_dir(D + "/precomputations/M");
  make_dir(D + "/precomputations/Neighbor");
  const std::string md = m_dir(D, P), nd = nb_dir(D, P);
  make_dir(md);
  make_dir(nd);
  size_t s = 0;
  for (int l = 0; l <= FAR_LEVELS; ++l, ++s)
    write_file(md + "/M_" + std::to_string(P.n_cheb) + "_" + std::to_string(l), eigen_text(tabs[s].data(), NTERM, R));
  for (int j = 2; j <= L; ++j)
    for (int f = 0; f < 4; ++f)
      for (int n = 0; n < FAMILY_SIZE[f]; ++n, ++s)
        write_file(nd + "/" + FAMILY_FILE[f] + "_" + std::to_string(j) + "_" + std::to_string(n),
                   eigen_text(tabs[s].data(), R, R));
  write_file(md + "/aifmm_key", key_text(P));
  write_file(nd + "/aifmm_key", key_text(P));
  return t_precompute;
}

// readMFromFile + readNeighborInteractions (LS_Assembly.hpp:1217-1228,
// 1422-1459): every file a tree of L levels reads must be there and have its
// shape; the uniform tree then uses M at level L and the colleague operators
// at level L.
void Problem::load_tables(const std::string& D) {
  auto t0 = std::chrono::steady_clock::now();
  const std::string md = m_dir(D, P), nd = nb_dir(D, P);
  check_key(md, P);
  check_key(nd, P);
  for (int l = 0; l <= L; ++l) {
    auto m = read_matrix(md + "/M_" + std::to_string(P.n_cheb) + "_" + std::to_string(l), NTERM, R);
    if (l == L) M_h = std::move(m);
  }
  T_h.clear();
  for (int j = 2; j <= L; ++j)
    for (int f = 0; f < 4; ++f)
      for (int n = 0; n < FAMILY_SIZE[f]; ++n) {
        auto t = read_matrix(nd + "/" + FAMILY_FILE[f] + "_" + std::to_string(j) + "_" + std::to_string(n), R, R);
        if (j == L && f == 0) T_h.insert(T_h.end(), t.begin(), t.end());
      }
  quad_capped = 0;
  install_tables();
  t_tables = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace aifmm
