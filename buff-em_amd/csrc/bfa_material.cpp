// bfa_material.cpp -- material and geometry front end: SVTensor / MatProp
// files compiled to device programs, and .buffgeo geometry files.
//
// Expressions follow the language of the reference's cevaluator
// (cmatheval, used by SVTensor.cc:96-112,356): complex arithmetic with
// + - * / ^ (right-associative, binding tighter than unary minus),
// parentheses, the functions exp log sqrt sin cos tan cot sec csc asin acos
// atan sinh cosh tanh abs step, the constants pi, e and i (imaginary unit;
// I is accepted too), the SVTensor variables w x y z r Theta Phi Eps1 Eps2
// Eps3 (SVTensor.cc:104-118, 319-339) and the file's named constants.
// Programs are postfix (bfa_expr.h) and run on the device inside the Gram
// integration; the tensor logic of SVTensor::Evaluate (defaults, Qxx copied
// to Qyy/Qzz, symmetrised off-diagonals) is the device's eps_tensor.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "bfa_expr.h"
#include "bfa_internal.h"

using namespace bfa;

namespace {

typedef std::complex<double> cplx;
constexpr double kFreqUnit = 3.0e14;   // MatProp::FreqUnit: w = Omega * 3e14 rad/s

struct Program {
  std::vector<ExprOp> ops;
  std::vector<cplx> consts;
};

// ---- expression compiler -----------------------------------------------
struct Compiler {
  const char *s;
  size_t pos = 0;
  const std::map<std::string, double> *named;
  Program *out;
  int depth = 0, max_depth = 0;
  std::string err;

  void emit(int op, int arg, int delta)
  {
    out->ops.push_back({op, arg});
    depth += delta;
    if (depth > max_depth) max_depth = depth;
  }
  void push_const(cplx v)
  {
    out->consts.push_back(v);
    emit(EOP_CONST, (int)out->consts.size() - 1, 1);
  }
  void skip() { while (s[pos] && isspace((unsigned char)s[pos])) pos++; }
  bool fail_at(const char *what)
  {
    if (err.empty()) err = std::string(what) + " at position " + std::to_string(pos) + " of \"" + s + "\"";
    return false;
  }

  bool expr()
  {
    if (!term()) return false;
    for (;;) {
      skip();
      const char c = s[pos];
      if (c != '+' && c != '-') return true;
      pos++;
      if (!term()) return false;
      emit(c == '+' ? EOP_ADD : EOP_SUB, 0, -1);
    }
  }
  bool term()
  {
    if (!unary()) return false;
    for (;;) {
      skip();
      const char c = s[pos];
      if (c != '*' && c != '/') return true;
      pos++;
      if (!unary()) return false;
      emit(c == '*' ? EOP_MUL : EOP_DIV, 0, -1);
    }
  }
  bool unary()
  {
    skip();
    if (s[pos] == '-') { pos++; if (!unary()) return false; emit(EOP_NEG, 0, 0); return true; }
    if (s[pos] == '+') { pos++; return unary(); }
    return power();
  }
  bool power()
  {
    if (!primary()) return false;
    skip();
    if (s[pos] == '^') {
      pos++;
      if (!unary()) return false;
      emit(EOP_POW, 0, -1);
    }
    return true;
  }
  bool primary()
  {
    skip();
    const char c = s[pos];
    if (c == '(') {
      pos++;
      if (!expr()) return false;
      skip();
      if (s[pos] != ')') return fail_at("expected ')'");
      pos++;
      return true;
    }
    if (isdigit((unsigned char)c) || c == '.') {
      char *end = nullptr;
      const double v = strtod(s + pos, &end);
      if (end == s + pos) return fail_at("bad number");
      pos = end - s;
      push_const(v);
      return true;
    }
    if (isalpha((unsigned char)c) || c == '_') {
      const size_t b = pos;
      while (isalnum((unsigned char)s[pos]) || s[pos] == '_') pos++;
      const std::string id(s + b, pos - b);
      skip();
      if (s[pos] == '(') {
        static const char *fn[EFN_COUNT] = {"exp", "log", "sqrt", "sin", "cos", "tan", "cot", "sec", "csc",
                                            "asin", "acos", "atan", "sinh", "cosh", "tanh", "abs", "step"};
        int f = -1;
        for (int k = 0; k < EFN_COUNT; k++)
          if (id == fn[k]) f = k;
        if (f < 0) return fail_at(("unknown function " + id).c_str());
        pos++;
        if (!expr()) return false;
        skip();
        if (s[pos] != ')') return fail_at("expected ')'");
        pos++;
        emit(EOP_FN, f, 0);
        return true;
      }
      static const char *var[EV_COUNT] = {"w", "x", "y", "z", "r", "Theta", "Phi", "Eps1", "Eps2", "Eps3"};
      for (int k = 0; k < EV_COUNT; k++)
        if (id == var[k]) { emit(EOP_VAR, k, 1); return true; }
      if (named) {
        auto it = named->find(id);
        if (it != named->end()) { push_const(it->second); return true; }
      }
      if (id == "pi") { push_const(M_PI); return true; }
      if (id == "e") { push_const(M_E); return true; }
      if (id == "i" || id == "I") { push_const(cplx(0.0, 1.0)); return true; }
      return fail_at(("unknown name " + id).c_str());
    }
    return fail_at("syntax error");
  }
};

int compile(const std::string &src, const std::map<std::string, double> &named, Program &p)
{
  Compiler c;
  c.s = src.c_str();
  c.named = &named;
  c.out = &p;
  if (!c.expr()) return fail(BFA_ERR_ARG, "expression: %s", c.err.c_str());
  c.skip();
  if (c.s[c.pos]) return fail(BFA_ERR_ARG, "expression: junk after position %zu of \"%s\"", c.pos, src.c_str());
  if (c.max_depth > EXPR_STACK) return fail(BFA_ERR_ARG, "expression too deeply nested: \"%s\"", src.c_str());
  return BFA_OK;
}

// ---- SVTensor / MatProp descriptions --------------------------------------
struct MatDesc {
  std::string expr[EXPR_SLOTS];       // tensor slots 3*mu+nu, then MatProp eps 9..11
  std::map<std::string, double> named;
  int num_mp = 0;
  bool matprop = false;
};

std::string trim(const std::string &x)
{
  size_t a = 0, b = x.size();
  while (a < b && isspace((unsigned char)x[a])) a++;
  while (b > a && isspace((unsigned char)x[b - 1])) b--;
  return x.substr(a, b - a);
}

bool iequal_prefix(const std::string &s, const char *p)
{
  const size_t n = strlen(p);
  if (s.size() < n) return false;
  for (size_t k = 0; k < n; k++)
    if (tolower((unsigned char)s[k]) != tolower((unsigned char)p[k])) return false;
  return true;
}

// complex literal of a CONST_EPS_ name: "10", "10+1i", "4-0.5i", "2.5i"
bool parse_complex(const std::string &t, cplx &v)
{
  const char *s = t.c_str();
  char *e = nullptr;
  const double a = strtod(s, &e);
  if (e == s) return false;
  if (*e == 0) { v = cplx(a, 0.0); return true; }
  if ((*e == 'i' || *e == 'I') && e[1] == 0) { v = cplx(0.0, a); return true; }
  char *f = nullptr;
  const double b = strtod(e, &f);
  if (f == e || (*f != 'i' && *f != 'I') || f[1] != 0) return false;
  v = cplx(a, b);
  return true;
}

// MATERIAL name ... ENDMATERIAL sections (scuff-em MatProp database format):
// NAME = value constants and Eps(w) = expression
struct MatPropDB {
  std::map<std::string, std::string> eps;                      // lower-case name -> expression
  std::map<std::string, std::map<std::string, double>> consts;
};

int read_matprop_section(FILE *f, const std::string &file, int &line, const std::string &name, MatPropDB &db)
{
  char buf[4096];
  std::map<std::string, double> named;
  std::string eps;
  while (fgets(buf, sizeof buf, f)) {
    line++;
    std::string l(buf);
    const size_t hash = l.find('#');
    if (hash != std::string::npos) l.resize(hash);
    const size_t semi = l.find(';');
    if (semi != std::string::npos) l.resize(semi);
    l = trim(l);
    if (l.empty()) continue;
    if (iequal_prefix(l, "ENDMATERIAL")) {
      if (eps.empty()) return fail(BFA_ERR_ARG, "%s:%d: MATERIAL %s has no Eps(w)", file.c_str(), line, name.c_str());
      std::string key = name;
      for (char &ch : key) ch = (char)tolower((unsigned char)ch);
      db.eps[key] = eps;
      db.consts[key] = named;
      return BFA_OK;
    }
    const size_t eq = l.find('=');
    if (eq == std::string::npos) return fail(BFA_ERR_ARG, "%s:%d: syntax error", file.c_str(), line);
    const std::string lhs = trim(l.substr(0, eq)), rhs = trim(l.substr(eq + 1));
    std::string lhs_ns;
    for (char ch : lhs)
      if (!isspace((unsigned char)ch)) lhs_ns += ch;
    if (lhs_ns == "Eps(w)" || lhs_ns == "eps(w)" || lhs_ns == "EPS(w)") { eps = rhs; continue; }
    char *end = nullptr;
    const double v = strtod(rhs.c_str(), &end);
    if (end == rhs.c_str()) return fail(BFA_ERR_ARG, "%s:%d: invalid constant value", file.c_str(), line);
    named[lhs] = v;
  }
  return fail(BFA_ERR_ARG, "%s: MATERIAL %s without ENDMATERIAL", file.c_str(), name.c_str());
}

int load_matprop_file(const char *path, MatPropDB &db)
{
  FILE *f = fopen(path, "r");
  if (!f) return fail(BFA_ERR_IO, "material file %s: cannot open", path);
  char buf[4096];
  int line = 0, rc = BFA_OK;
  while (rc == BFA_OK && fgets(buf, sizeof buf, f)) {
    line++;
    std::string l = trim(buf);
    if (l.empty() || l[0] == '#') continue;
    if (iequal_prefix(l, "MATERIAL")) {
      const std::string name = trim(l.substr(8));
      if (name.empty()) rc = fail(BFA_ERR_ARG, "%s:%d: no name given for MATERIAL", path, line);
      else rc = read_matprop_section(f, path, line, name, db);
    }
  }
  fclose(f);
  return rc;
}

// MatProp name -> expression in w (and its constants)
int matprop_expr(const std::string &name, const MatPropDB &local, std::string &expr, std::map<std::string, double> &named)
{
  if (iequal_prefix(name, "CONST_EPS_")) {
    cplx v;
    if (!parse_complex(name.substr(10), v)) return fail(BFA_ERR_ARG, "material %s: bad constant", name.c_str());
    char buf[128];
    snprintf(buf, sizeof buf, "(%.17g) + (%.17g)*i", v.real(), v.imag());
    expr = buf;
    return BFA_OK;
  }
  std::string key = name;
  for (char &ch : key) ch = (char)tolower((unsigned char)ch);
  auto it = local.eps.find(key);
  if (it != local.eps.end()) {
    expr = it->second;
    named = local.consts.at(key);
    return BFA_OK;
  }
  const char *dbfile = getenv("BUFF_MATPROP_FILE");
  if (!dbfile) dbfile = getenv("SCUFF_MATPROP_FILE");
  if (dbfile) {
    MatPropDB db;
    int rc = load_matprop_file(dbfile, db);
    if (rc) return rc;
    auto jt = db.eps.find(key);
    if (jt != db.eps.end()) {
      expr = jt->second;
      named = db.consts.at(key);
      return BFA_OK;
    }
  }
  return fail(BFA_ERR_ARG, "material %s: not CONST_EPS_<value> and not defined in a MATERIAL section", name.c_str());
}

// MP_name -> Eps<k> (SVTensor::ExtractMPs, SVTensor.cc:129-170)
int extract_mps(std::string &expr, std::vector<std::string> &mps)
{
  size_t p;
  while ((p = expr.find("MP_")) != std::string::npos) {
    size_t q = p + 3;
    // the name runs to the next blank as in the reference, and also stops
    // at ) * / ^ , so that "(MP_Gold)" works
    while (q < expr.size() && !isspace((unsigned char)expr[q]) && !strchr(")*/^,", expr[q])) q++;
    const std::string name = expr.substr(p + 3, q - p - 3);
    int k = -1;
    for (size_t m = 0; m < mps.size(); m++)
      if (strcasecmp(mps[m].c_str(), name.c_str()) == 0) k = (int)m;
    if (k < 0) {
      if (mps.size() == 3) return fail(BFA_ERR_ARG, "too many MP_ references (at most 3)");
      mps.push_back(name);
      k = (int)mps.size() - 1;
    }
    expr.replace(p, q - p, " Eps" + std::to_string(k + 1) + " ");
  }
  return BFA_OK;
}

FILE *open_on_path(const std::string &name, const char *envvar, const std::string &extra_dir, std::string &found)
{
  std::vector<std::string> cand{name};
  if (name.empty() || name[0] != '/') {
    if (!extra_dir.empty()) cand.push_back(extra_dir + "/" + name);
    if (const char *pl = getenv(envvar)) {
      std::string s(pl);
      size_t a = 0;
      while (a <= s.size()) {
        size_t b = s.find(':', a);
        if (b == std::string::npos) b = s.size();
        if (b > a) cand.push_back(s.substr(a, b - a) + "/" + name);
        a = b + 1;
      }
    }
  }
  for (const std::string &c : cand)
    if (FILE *f = fopen(c.c_str(), "r")) { found = c; return f; }
  return nullptr;
}

// SVTensor file (SVTensor::Parse, SVTensor.cc:175-285).  One difference:
// a tensor line must start with "Q" or "Eps" (the reference tests only the
// first letter of "Eps", so any line starting with E was a tensor line).
int parse_svtensor(const std::string &name, const std::string &dir, MatDesc &d)
{
  std::string path;
  FILE *f = open_on_path(name, "BUFF_SVTENSOR_PATH", dir, path);
  if (!f) return fail(BFA_ERR_IO, "could not find file %s", name.c_str());
  MatPropDB local;
  std::vector<std::string> mps;
  char buf[4096];
  int line = 0, rc = BFA_OK;
  while (rc == BFA_OK && fgets(buf, sizeof buf, f)) {
    line++;
    std::string l(buf);
    const size_t semi = l.find(';');
    if (semi != std::string::npos) l.resize(semi);
    l = trim(l);
    if (l.empty() || l[0] == '#') continue;
    if (l.compare(0, 8, "MATERIAL") == 0) {
      const std::string mname = trim(l.substr(8));
      if (mname.empty()) { rc = fail(BFA_ERR_ARG, "%s:%d: no name given for MATERIAL", path.c_str(), line); break; }
      rc = read_matprop_section(f, path, line, mname, local);
      continue;
    }
    const size_t eq = l.find('=');
    size_t q = std::string::npos;
    if (l.size() >= 3 && (l[0] == 'Q' || l[0] == 'q')) q = 1;
    else if (l.size() >= 5 && iequal_prefix(l, "Eps")) q = 3;
    if (q != std::string::npos && eq != std::string::npos) {
      int M = 0, N = 0;
      if (!(l[q] == '=' || isspace((unsigned char)l[q]))) {
        int mn[2];
        for (int k = 0; k < 2; k++) {
          const char ch = (char)toupper((unsigned char)l[q + k]);
          if (ch >= 'X' && ch <= 'Z') mn[k] = ch - 'X';
          else if (ch >= '1' && ch <= '3') mn[k] = ch - '1';
          else { rc = fail(BFA_ERR_ARG, "%s:%d: invalid expression", path.c_str(), line); break; }
        }
        if (rc) break;
        M = mn[0];
        N = mn[1];
      }
      std::string e = l.substr(eq + 1);
      if ((rc = extract_mps(e, mps))) break;
      d.expr[3 * M + N] = e;
      continue;
    }
    if (eq != std::string::npos) {
      char *end = nullptr;
      const std::string rhs = trim(l.substr(eq + 1));
      const double v = strtod(rhs.c_str(), &end);
      if (end == rhs.c_str()) { rc = fail(BFA_ERR_ARG, "%s:%d: invalid constant value", path.c_str(), line); break; }
      d.named[trim(l.substr(0, eq))] = v;
      continue;
    }
    rc = fail(BFA_ERR_ARG, "%s:%d: syntax error", path.c_str(), line);
  }
  fclose(f);
  if (rc) return rc;
  d.num_mp = (int)mps.size();
  for (int k = 0; k < d.num_mp && rc == BFA_OK; k++) {
    std::map<std::string, double> mn;
    rc = matprop_expr(mps[k], local, d.expr[9 + k], mn);
    // a MatProp's constants live in its own section; inline them by
    // compiling its expression with them (merged under the file's)
    for (auto &kv : mn) d.named.insert(kv);
  }
  return rc;
}

// compile every present slot into one device program (out == NULL: check
// only, nothing is uploaded)
int build_material(const MatDesc &d, bfa_material *out)
{
  std::vector<ExprOp> ops;
  std::vector<cplx> consts;
  ExprHeader h{};
  for (int s = 0; s < EXPR_SLOTS; s++) {
    h.start[s] = (int)ops.size();
    h.len[s] = 0;
    if (d.expr[s].empty()) continue;
    Program p;
    int rc = compile(d.expr[s], d.named, p);
    if (rc) return rc;
    for (ExprOp o : p.ops) {
      if (o.op == EOP_CONST) o.arg += (int)consts.size();
      ops.push_back(o);
    }
    consts.insert(consts.end(), p.consts.begin(), p.consts.end());
    h.len[s] = (int)p.ops.size();
  }
  h.nops = (int)ops.size();
  h.nconst = (int)consts.size();
  h.num_mp = d.num_mp;
  h.matprop = d.matprop ? 1 : 0;
  h.freq_unit = kFreqUnit;
  if (!out) return BFA_OK;
  std::vector<char> blob(sizeof(ExprHeader) + ops.size() * sizeof(ExprOp) + consts.size() * 16);
  memcpy(blob.data(), &h, sizeof h);
  memcpy(blob.data() + sizeof h, ops.data(), ops.size() * sizeof(ExprOp));
  double *kc = (double *)(blob.data() + sizeof h + ops.size() * sizeof(ExprOp));
  for (size_t k = 0; k < consts.size(); k++) { kc[2 * k] = consts[k].real(); kc[2 * k + 1] = consts[k].imag(); }
  void *dev = nullptr;
  if (hipMalloc(&dev, blob.size()) != hipSuccess) return fail(BFA_ERR_OOM, "material program allocation");
  if (hipMemcpy(dev, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dev);
    return fail(BFA_ERR_HIP, "material program upload");
  }
  memset(out, 0, sizeof *out);
  out->kind = 4;
  out->expr = dev;
  return BFA_OK;
}

int material_from_name(const std::string &name, int is_matprop, const std::string &dir, bfa_material *out)
{
  MatDesc d;
  int rc;
  if (is_matprop) {
    MatPropDB none;
    rc = matprop_expr(name, none, d.expr[9], d.named);
    d.num_mp = 1;
    d.matprop = true;
  } else {
    rc = parse_svtensor(name, dir, d);
  }
  if (rc) return rc;
  return build_material(d, out);
}

// ---- .buffgeo ----------------------------------------------------------
struct Affine {
  double M[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double D[3] = {0, 0, 0};
  // this := (x -> R x + t) after this
  void then(const double R[3][3], const double t[3])
  {
    double M2[3][3], D2[3];
    for (int r = 0; r < 3; r++) {
      D2[r] = t[r];
      for (int c = 0; c < 3; c++) {
        M2[r][c] = 0.0;
        for (int k = 0; k < 3; k++) M2[r][c] += R[r][k] * M[k][c];
        D2[r] += R[r][c] * D[c];
      }
    }
    memcpy(M, M2, sizeof M);
    memcpy(D, D2, sizeof D);
  }
};

// DISPLACED dx dy dz / ROTATED theta ABOUT nx ny nz (degrees), any sequence
int parse_transform(const std::vector<std::string> &tok, size_t k, Affine &A, const std::string &where)
{
  while (k < tok.size()) {
    const std::string &w = tok[k];
    if (strcasecmp(w.c_str(), "DISPLACED") == 0) {
      if (k + 3 >= tok.size()) return fail(BFA_ERR_ARG, "%s: DISPLACED needs 3 values", where.c_str());
      const double I3[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      const double t[3] = {atof(tok[k + 1].c_str()), atof(tok[k + 2].c_str()), atof(tok[k + 3].c_str())};
      A.then(I3, t);
      k += 4;
    } else if (strcasecmp(w.c_str(), "ROTATED") == 0) {
      if (k + 5 >= tok.size() || strcasecmp(tok[k + 2].c_str(), "ABOUT") != 0)
        return fail(BFA_ERR_ARG, "%s: ROTATED theta ABOUT nx ny nz", where.c_str());
      const double th = atof(tok[k + 1].c_str()) * M_PI / 180.0;
      double n[3] = {atof(tok[k + 3].c_str()), atof(tok[k + 4].c_str()), atof(tok[k + 5].c_str())};
      const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (nn == 0.0) return fail(BFA_ERR_ARG, "%s: zero rotation axis", where.c_str());
      for (double &v : n) v /= nn;
      const double c = cos(th), s = sin(th), C = 1.0 - c;
      // Rodrigues: rotation by theta about n (right-hand rule)
      const double R[3][3] = {{c + n[0] * n[0] * C, n[0] * n[1] * C - n[2] * s, n[0] * n[2] * C + n[1] * s},
                              {n[1] * n[0] * C + n[2] * s, c + n[1] * n[1] * C, n[1] * n[2] * C - n[0] * s},
                              {n[2] * n[0] * C - n[1] * s, n[2] * n[1] * C + n[0] * s, c + n[2] * n[2] * C}};
      const double z[3] = {0, 0, 0};
      A.then(R, z);
      k += 6;
    } else {
      return fail(BFA_ERR_ARG, "%s: junk at end of line", where.c_str());
    }
  }
  return BFA_OK;
}

struct GeoDesc {
  std::string mesh, material;
  bool is_matprop = false;
  Affine A;
  std::string dir;
  std::vector<std::string> meshpath;
};

std::vector<std::string> tokenize(const std::string &l)
{
  std::vector<std::string> t;
  size_t a = 0;
  while (a < l.size()) {
    while (a < l.size() && isspace((unsigned char)l[a])) a++;
    size_t b = a;
    while (b < l.size() && !isspace((unsigned char)l[b])) b++;
    if (b > a) t.push_back(l.substr(a, b - a));
    a = b;
  }
  return t;
}

int parse_buffgeo(const char *path, GeoDesc &g)
{
  FILE *f = fopen(path, "r");
  if (!f) return fail(BFA_ERR_IO, "could not open %s", path);
  const std::string p(path);
  const size_t slash = p.find_last_of('/');
  g.dir = slash == std::string::npos ? "." : p.substr(0, slash);
  char buf[4096];
  int line = 0, objects = 0, rc = BFA_OK;
  bool in_obj = false;
  while (rc == BFA_OK && fgets(buf, sizeof buf, f)) {
    line++;
    const std::vector<std::string> t = tokenize(buf);
    if (t.empty() || t[0][0] == '#') continue;
    const std::string where = p + ":" + std::to_string(line);
    const char *k0 = t[0].c_str();
    if (!in_obj) {
      if (strcasecmp(k0, "MESHPATH") == 0) {
        if (t.size() != 2) { rc = fail(BFA_ERR_ARG, "%s: invalid MESHPATH specification", where.c_str()); break; }
        g.meshpath.push_back(t[1][0] == '/' ? t[1] : g.dir + "/" + t[1]);
      } else if (strcasecmp(k0, "OBJECT") == 0) {
        if (t.size() != 2) { rc = fail(BFA_ERR_ARG, "%s: invalid OBJECT specification", where.c_str()); break; }
        if (++objects > 1) { rc = fail(BFA_ERR_ARG, "%s: geometries of more than one object are not supported", where.c_str()); break; }
        in_obj = true;
      } else {
        rc = fail(BFA_ERR_ARG, "%s: syntax error", where.c_str());
      }
      continue;
    }
    if (strcasecmp(k0, "MESHFILE") == 0) {
      if (t.size() != 2) { rc = fail(BFA_ERR_ARG, "%s: MESHFILE keyword requires one argument", where.c_str()); break; }
      g.mesh = t[1];
    } else if (strcasecmp(k0, "MATERIAL") == 0 || strcasecmp(k0, "SVTENSOR") == 0) {
      if (t.size() != 2) { rc = fail(BFA_ERR_ARG, "%s: %s keyword requires one argument", where.c_str(), k0); break; }
      g.material = t[1];
      g.is_matprop = strcasecmp(k0, "MATERIAL") == 0;
    } else if (strcasecmp(k0, "DISPLACED") == 0 || strcasecmp(k0, "ROTATED") == 0) {
      rc = parse_transform(t, 0, g.A, where);
    } else if (strcasecmp(k0, "ENDOBJECT") == 0) {
      in_obj = false;
    } else {
      rc = fail(BFA_ERR_ARG, "%s: syntax error", where.c_str());
    }
  }
  fclose(f);
  if (rc) return rc;
  if (objects == 0) return fail(BFA_ERR_ARG, "%s: no OBJECT", path);
  if (g.mesh.empty()) return fail(BFA_ERR_ARG, "%s: no MESHFILE specified for object", path);
  if (g.material.empty()) return fail(BFA_ERR_ARG, "%s: no MATERIAL or SVTENSOR specified for object", path);
  return BFA_OK;
}

std::string find_mesh(const GeoDesc &g)
{
  std::vector<std::string> cand{g.mesh};
  if (g.mesh[0] != '/') {
    cand.push_back(g.dir + "/" + g.mesh);
    for (const std::string &d : g.meshpath) cand.push_back(d + "/" + g.mesh);
    if (const char *pl = getenv("BUFF_MESH_PATH")) {
      std::string s(pl);
      size_t a = 0;
      while (a <= s.size()) {
        size_t b = s.find(':', a);
        if (b == std::string::npos) b = s.size();
        if (b > a) cand.push_back(s.substr(a, b - a) + "/" + g.mesh);
        a = b + 1;
      }
    }
  }
  for (const std::string &c : cand)
    if (FILE *f = fopen(c.c_str(), "r")) { fclose(f); return c; }
  return std::string();
}

}  // namespace

extern "C" int bfa_material_svtensor(const char *name, int is_matprop, bfa_material *out)
{
  if (!name || !out) return fail(BFA_ERR_ARG, "bfa_material_svtensor: null argument");
  return material_from_name(name, is_matprop, ".", out);
}

extern "C" int bfa_material_check(const char *name, int is_matprop)
{
  if (!name) return fail(BFA_ERR_ARG, "bfa_material_check: null name");
  return material_from_name(name, is_matprop, ".", nullptr);
}

extern "C" int bfa_material_expr_check(const char *eps_expr)
{
  if (!eps_expr) return fail(BFA_ERR_ARG, "bfa_material_expr_check: null expression");
  MatDesc d;
  d.expr[0] = eps_expr;
  return build_material(d, nullptr);
}

extern "C" int bfa_material_expr(const char *eps_expr, bfa_material *out)
{
  if (!eps_expr || !out) return fail(BFA_ERR_ARG, "bfa_material_expr: null argument");
  MatDesc d;
  d.expr[0] = eps_expr;
  return build_material(d, out);
}

extern "C" int bfa_material_release(bfa_material *m)
{
  if (!m) return fail(BFA_ERR_ARG, "bfa_material_release: null material");
  if (m->kind == 4 && m->expr) (void)hipFree(const_cast<void *>(m->expr));
  m->expr = nullptr;
  return BFA_OK;
}

extern "C" int bfa_buffgeo_load(const char *path, int *nv, int *nt, double *V, int *tetVI, bfa_material *mat)
{
  if (!path || !nv || !nt) return fail(BFA_ERR_ARG, "bfa_buffgeo_load: null argument");
  GeoDesc g;
  int rc = parse_buffgeo(path, g);
  if (rc) return rc;
  const std::string mesh = find_mesh(g);
  if (mesh.empty()) return fail(BFA_ERR_IO, "%s: mesh file %s not found", path, g.mesh.c_str());
  if (!V || !tetVI) return bfa_mesh_read_vmsh(mesh.c_str(), nv, nt, nullptr, nullptr);
  if ((rc = bfa_mesh_read_vmsh(mesh.c_str(), nv, nt, V, tetVI))) return rc;
  for (int v = 0; v < *nv; v++) {
    double x[3] = {V[3 * v], V[3 * v + 1], V[3 * v + 2]};
    for (int r = 0; r < 3; r++)
      V[3 * v + r] = g.A.M[r][0] * x[0] + g.A.M[r][1] * x[1] + g.A.M[r][2] * x[2] + g.A.D[r];
  }
  if (mat) return material_from_name(g.material, g.is_matprop, g.dir, mat);
  return BFA_OK;
}
