// Nastran bulk-data (BDF) mesh files: a reader for the GRID and CQUADn cards the reference's meshes use and a
// writer of the same cards.
//
// The reference reads its meshes through TACS (TACSMeshLoader::scanBDFFile, verification/TACSHelpers.h:86-114),
// choosing the element by card name: CQUAD4, CQUAD / CQUAD9, CQUAD16, CQUAD25. Its gmsh meshes are rewritten
// by fixGmshBDF (meshes/GmshUtils.py:118-196), which renames each element card after its node count and lays
// its fields out 8 characters wide, nine fields per line, continuation lines starting with a blank field.
//
// Accepted card formats:
//   * fixed small field: 8-character fields, data in columns 9-72; a continuation line starts with a blank
//     field or a '+' / '*' marker (columns 73-80 of the line before hold the marker and are ignored);
//   * fixed large field (GRID*): 16-character data fields, four per line, continuation lines start with '*';
//   * free field: comma-separated; a continuation line starts with '+', '*' or ','.
// Reals may use Nastran's implicit exponent ("1.5-3" = 1.5e-3) and D exponents. '$' lines are comments; parsing
// stops at ENDDATA. Nodes are numbered 0..n-1 in the order of their GRID cards; element node ids are mapped
// through the GRID ids. The 2-D coordinates are X1, X2 (X3 and the coordinate systems CP/CD must be 0/blank).
//
// Element node order: as listed on the card (flags 0), which is the tensor-product order a = ay (p+1) + ax for
// files this writer produced; or, with FEM_BDF_NASTRAN_ORDER, converted from Nastran's CQUAD4 / CQUAD8-9 order
// (corners counter-clockwise, then mid-side nodes G5 on G1-G2, G6 on G2-G3, G7 on G3-G4, G8 on G4-G1, then the
// centre G9) to the tensor-product order. gmsh's own order of 16- and 25-node quads is not converted.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "fem_internal.h"

namespace fem {
namespace {

std::string trim(const std::string &s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) b++;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) e--;
  return s.substr(b, e - b);
}

std::string upper(std::string s) {
  for (char &c : s) c = char(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

// Nastran real: "1.5", "-.5", "1.5E-3", "1.5D-3", "1.5-3", "1.5+3". Returns false if not a number.
bool parse_real(const std::string &field, double &out) {
  std::string s = upper(trim(field));
  if (s.empty()) return false;
  for (char &c : s)
    if (c == 'D') c = 'E';
  if (s.find('E') == std::string::npos) {
    // implicit exponent: a sign after the first character that does not follow an 'E'
    for (size_t i = 1; i < s.size(); i++) {
      if ((s[i] == '+' || s[i] == '-') && s[i - 1] != 'E') {
        s.insert(i, "E");
        break;
      }
    }
  }
  errno = 0;
  char *end = nullptr;
  out = std::strtod(s.c_str(), &end);
  return errno == 0 && end && *end == '\0';
}

bool parse_int(const std::string &field, long long &out) {
  const std::string s = trim(field);
  if (s.empty()) return false;
  errno = 0;
  char *end = nullptr;
  out = std::strtoll(s.c_str(), &end, 10);
  return errno == 0 && end && *end == '\0';
}

struct Card {
  std::string name;                 // upper case, without a trailing '*'
  std::vector<std::string> fields;  // data fields (field 2 onwards), continuation lines appended
  int line = 0;                     // line number of the card's first line (1-based)
};

bool is_free(const std::string &line) { return line.find(',') != std::string::npos; }

// Data fields of one line of a card (the first field, name or continuation marker, removed).
void line_fields(const std::string &raw, bool large, std::vector<std::string> &out) {
  std::string line = raw;
  if (is_free(line)) {
    size_t start = 0;
    bool first = true;
    while (start <= line.size()) {
      size_t comma = line.find(',', start);
      if (comma == std::string::npos) comma = line.size();
      if (!first) out.push_back(trim(line.substr(start, comma - start)));
      first = false;
      start = comma + 1;
    }
    return;
  }
  const size_t w = large ? 16 : 8, n = large ? 4 : 8;
  for (size_t k = 0; k < n; k++) {
    const size_t b = 8 + k * w;
    if (b >= line.size()) break;
    out.push_back(trim(line.substr(b, std::min(w, line.size() - b))));
  }
}

bool is_continuation(const std::string &line) {
  if (line.empty()) return false;
  if (line[0] == '+' || line[0] == '*' || line[0] == ',') return true;
  // fixed format: blank first field with data after it
  if (line.size() > 8 && trim(line.substr(0, 8)).empty() && !trim(line.substr(8)).empty()) return true;
  return false;
}

int read_cards(const char *path, std::vector<Card> &cards) {
  FILE *f = std::fopen(path, "r");
  if (!f) return set_error(FEM_ERR_INVALID_ARGUMENT, "cannot open %s", path);
  std::string cur;
  std::vector<std::string> lines;
  char buf[4096];
  while (std::fgets(buf, sizeof(buf), f)) {
    std::string l(buf);
    while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
    lines.push_back(l);
  }
  std::fclose(f);
  Card *open = nullptr;
  bool openLarge = false;
  for (size_t i = 0; i < lines.size(); i++) {
    const std::string &l = lines[i];
    if (l.empty() || l[0] == '$' || trim(l).empty()) continue;
    if (open && is_continuation(l)) {
      line_fields(l, openLarge, open->fields);
      continue;
    }
    std::string name = is_free(l) ? trim(l.substr(0, l.find(','))) : trim(l.substr(0, std::min<size_t>(8, l.size())));
    name = upper(name);
    if (name == "ENDDATA") break;
    const bool large = !name.empty() && name.back() == '*';
    if (large) name.pop_back();
    if (name == "GRID" || name.rfind("CQUAD", 0) == 0) {
      cards.push_back(Card{name, {}, int(i + 1)});
      open = &cards.back();
      openLarge = large;
      line_fields(l, large, open->fields);
    } else {
      open = nullptr;  // other cards and their continuations are skipped
    }
  }
  // trailing blank fields of fixed-format lines carry nothing
  for (Card &c : cards)
    while (!c.fields.empty() && c.fields.back().empty()) c.fields.pop_back();
  return FEM_OK;
}

// Tensor-product position of each Nastran node of CQUAD4 (p = 1) and CQUAD8/9 (p = 2).
const int kNastranQuad4[4] = {0, 1, 3, 2};
const int kNastranQuad9[9] = {0, 2, 8, 6, 1, 5, 7, 3, 4};

struct BdfMesh {
  int order = 0;
  std::vector<double> coords;  // [n][2]
  std::vector<long long> gridIds;
  std::vector<int> conn;       // [E][N]
};

int parse_bdf(const char *path, int flags, BdfMesh &m) {
  if (!path) return set_error(FEM_ERR_NULL_POINTER, "path is NULL");
  if (flags & ~FEM_BDF_NASTRAN_ORDER) return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown BDF flags %d", flags);
  std::vector<Card> cards;
  int rc = read_cards(path, cards);
  if (rc) return rc;
  std::unordered_map<long long, int> index;
  for (const Card &c : cards) {
    if (c.name != "GRID") continue;
    long long id = 0, cp = 0;
    double x = 0, y = 0, z = 0;
    const auto &F = c.fields;
    if (F.size() < 4 || !parse_int(F[0], id) || !parse_real(F[2], x) || !parse_real(F[3], y)) {
      return set_error(FEM_ERR_INVALID_ARGUMENT, "%s:%d: GRID needs ID, X1, X2", path, c.line);
    }
    if (!F[1].empty() && (!parse_int(F[1], cp) || cp != 0)) {
      return set_error(FEM_ERR_UNSUPPORTED, "%s:%d: GRID %lld uses coordinate system %s", path, c.line, id,
                       F[1].c_str());
    }
    if (F.size() > 4 && !F[4].empty() && (!parse_real(F[4], z) || z != 0.0)) {
      return set_error(FEM_ERR_UNSUPPORTED, "%s:%d: GRID %lld has X3 = %s (plane meshes only)", path, c.line, id,
                       F[4].c_str());
    }
    if (!index.emplace(id, int(m.gridIds.size())).second) {
      return set_error(FEM_ERR_INVALID_ARGUMENT, "%s:%d: GRID %lld defined twice", path, c.line, id);
    }
    m.gridIds.push_back(id);
    m.coords.push_back(x);
    m.coords.push_back(y);
  }
  for (const Card &c : cards) {
    if (c.name == "GRID") continue;
    const int nn = int(c.fields.size()) - 2;
    const int p = nn == 4 ? 1 : nn == 9 ? 2 : nn == 16 ? 3 : nn == 25 ? 4 : 0;
    const std::string expect = c.name.substr(5);
    if (p == 0 || (!expect.empty() && expect != std::to_string(nn) && !(expect == "8" && nn == 9))) {
      return set_error(FEM_ERR_INVALID_ARGUMENT, "%s:%d: %s card with %d node fields", path, c.line, c.name.c_str(), nn);
    }
    if (m.order && p != m.order) {
      return set_error(FEM_ERR_UNSUPPORTED, "%s:%d: elements of orders %d and %d in one mesh", path, c.line, m.order, p);
    }
    m.order = p;
    int nodes[25];
    for (int a = 0; a < nn; a++) {
      long long g = 0;
      if (!parse_int(c.fields[size_t(a) + 2], g)) {
        return set_error(FEM_ERR_INVALID_ARGUMENT, "%s:%d: node field '%s' is not an integer", path, c.line,
                         c.fields[size_t(a) + 2].c_str());
      }
      auto it = index.find(g);
      if (it == index.end()) {
        return set_error(FEM_ERR_INVALID_ARGUMENT, "%s:%d: element node %lld has no GRID card", path, c.line, g);
      }
      nodes[a] = it->second;
    }
    if (flags & FEM_BDF_NASTRAN_ORDER) {
      if (p > 2) {
        return set_error(FEM_ERR_UNSUPPORTED, "%s:%d: Nastran node order is defined for 4- and 9-node quads only",
                         path, c.line);
      }
      int t[9];
      const int *pos = p == 1 ? kNastranQuad4 : kNastranQuad9;
      for (int a = 0; a < nn; a++) t[pos[a]] = nodes[a];
      std::copy(t, t + nn, nodes);
    }
    m.conn.insert(m.conn.end(), nodes, nodes + nn);
  }
  if (m.gridIds.empty() || m.order == 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "%s holds no GRID or CQUAD cards", path);
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" {

int fem_read_bdf_size(const char *path, int flags, int *order, int *numNodes, int *numElements) {
  if (!order || !numNodes || !numElements) return set_error(FEM_ERR_NULL_POINTER, "NULL size pointer");
  BdfMesh m;
  const int rc = parse_bdf(path, flags, m);
  if (rc) return rc;
  *order = m.order;
  *numNodes = int(m.gridIds.size());
  *numElements = int(m.conn.size() / size_t((m.order + 1) * (m.order + 1)));
  return FEM_OK;
}

int fem_read_bdf(const char *path, int flags, double *coords, int *connPtr, int *conn, long long *gridIds) {
  if (!coords || !connPtr || !conn) return set_error(FEM_ERR_NULL_POINTER, "coords/connPtr/conn is NULL");
  BdfMesh m;
  const int rc = parse_bdf(path, flags, m);
  if (rc) return rc;
  const int N = (m.order + 1) * (m.order + 1);
  const size_t E = m.conn.size() / size_t(N);
  std::copy(m.coords.begin(), m.coords.end(), coords);
  std::copy(m.conn.begin(), m.conn.end(), conn);
  for (size_t e = 0; e <= E; e++) connPtr[e] = int(e) * N;
  if (gridIds) std::copy(m.gridIds.begin(), m.gridIds.end(), gridIds);
  return FEM_OK;
}

int fem_write_bdf(const char *path, int order, int numNodes, int numElements, const int *connPtr, const int *conn,
                  const double *nodeCoords) {
  if (!path || !connPtr || !conn || !nodeCoords) return set_error(FEM_ERR_NULL_POINTER, "NULL argument");
  if (order < 1 || order > 4) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (numNodes < 0 || numElements < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "negative size");
  const int N = (order + 1) * (order + 1);
  FILE *f = std::fopen(path, "w");
  if (!f) return set_error(FEM_ERR_INVALID_ARGUMENT, "cannot open %s for writing", path);
  bool ok = std::fprintf(f, "$ plane mesh: %d nodes, %d order-%d quads (tensor-product node order)\nBEGIN BULK\n",
                         numNodes, numElements, order) > 0;
  // GRID in free field with full-precision reals, so a file read back gives the same doubles
  for (int n = 0; ok && n < numNodes; n++) {
    ok = std::fprintf(f, "GRID,%d,0,%.17g,%.17g,0.\n", n + 1, nodeCoords[2 * size_t(n)],
                      nodeCoords[2 * size_t(n) + 1]) > 0;
  }
  // CQUADn in fixed small field, nine fields per line, continuation lines with a blank first field
  char name[16];
  std::snprintf(name, sizeof(name), "CQUAD%d", N);
  for (int e = 0; ok && e < numElements; e++) {
    std::vector<std::string> fld = {name, std::to_string(e + 1), "1"};
    for (int a = 0; a < N; a++) {
      const int node = conn[int64_t(connPtr[e]) + a];
      if (node < 0 || node >= numNodes) {
        std::fclose(f);
        return set_error(FEM_ERR_INVALID_ARGUMENT, "element %d node %d outside [0, %d)", e, node, numNodes);
      }
      fld.push_back(std::to_string(node + 1));
    }
    std::string line;
    int count = 0;
    for (const std::string &s : fld) {
      if (count == 9) {
        ok = ok && std::fprintf(f, "%s\n", line.c_str()) > 0;
        line = std::string(8, ' ');
        count = 1;
      }
      char cell[17];
      std::snprintf(cell, sizeof(cell), "%-8s", s.c_str());
      line += cell;
      count++;
    }
    ok = ok && std::fprintf(f, "%s\n", line.c_str()) > 0;
  }
  ok = ok && std::fprintf(f, "ENDDATA\n") > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? FEM_OK : set_error(FEM_ERR_INVALID_ARGUMENT, "writing %s failed", path);
}

}  // extern "C"
