// readinfo.cpp -- Input::readinfo (src/IO/readinfo.cpp:18-733): the
// reference's libconfig input file into Input_Info_t.
//
// libconfig is not in this image, so the file is parsed here: the subset of
// the libconfig grammar those input files use, namely
//   setting := name ('=' | ':') value [';' | ',']
//   value   := scalar | '[' scalars ']' (array) | '(' values ')' (list)
//            | '{' settings '}' (group)
//   scalar  := integer ([-+]digits, 0x hex; an L suffix or a value beyond
//              32 bits makes it 64-bit) | float (with '.' or an exponent)
//            | true/false | "string" (adjacent strings concatenate)
// with '#', '//' and '/* */' comments.  Array elements must share one scalar
// type (integers of both widths mix).  Numbers convert between integer and
// float on read (the reference sets auto-conversion); a float read as an
// integer is truncated.
//
// The settings, defaults, length rules and messages follow readinfo.cpp:
// required settings fail with "Error: ... not set in input file", optional
// ones print a caution and take the reference's default.  Where the
// reference has no working default, this reader fails with a message
// instead:
//   - relativity, electrostatic and output_pCount: readinfo.cpp:214,424,715
//     catch a pointer type that the library never throws, so a missing value
//     ends the reference's program;
//   - a value of the wrong type (a string where a number is read): libconfig
//     throws a type exception that readinfo.cpp does not catch.
// A missing E0 zeroes B0 and leaves E0 at zero, as readinfo.cpp:466-467 does.
//
// The settings this implementation adds (precision, seed, waist, dt,
// poisson_iter_max, poisson_tol) are read from an optional group `emopic`;
// the reference ignores settings it does not look up, so such a file still
// runs there.
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "emopic.hpp"

namespace emopic {
namespace {

enum class T { Int, Int64, Float, Bool, String, Group, Array, List };

struct Node {
  T type = T::Group;
  std::string name;
  long long i = 0;
  double f = 0.0;
  bool b = false;
  std::string s;
  std::vector<Node> kids;
  int line = 0;

  bool numeric() const { return type == T::Int || type == T::Int64 || type == T::Float; }
  bool aggregate() const { return type == T::Group || type == T::Array || type == T::List; }
  // Setting::getLength(): elements of an aggregate, 0 for a scalar
  int length() const { return aggregate() ? (int)kids.size() : 0; }
};

struct ParseError {
  int line;
  std::string msg;
};

// Thrown for a value whose type cannot be read as asked (libconfig's
// SettingTypeException).
struct TypeError {
  std::string path;
};

class Parser {
 public:
  explicit Parser(const std::string &text) : t_(text) {}

  Node parse() {
    Node root;
    root.type = T::Group;
    root.line = 1;
    settings(root, '\0');
    return root;
  }

 private:
  const std::string &t_;
  size_t p_ = 0;
  int line_ = 1;

  [[noreturn]] void fail(const std::string &m) { throw ParseError{line_, m}; }

  int peek() {
    skip();
    return p_ < t_.size() ? (unsigned char)t_[p_] : -1;
  }

  void skip() {
    for (;;) {
      while (p_ < t_.size() && std::isspace((unsigned char)t_[p_])) {
        if (t_[p_] == '\n') line_++;
        p_++;
      }
      if (p_ >= t_.size()) return;
      if (t_[p_] == '#' || (t_[p_] == '/' && p_ + 1 < t_.size() && t_[p_ + 1] == '/')) {
        while (p_ < t_.size() && t_[p_] != '\n') p_++;
        continue;
      }
      if (t_[p_] == '/' && p_ + 1 < t_.size() && t_[p_ + 1] == '*') {
        size_t e = t_.find("*/", p_ + 2);
        if (e == std::string::npos) fail("unterminated comment");
        for (size_t k = p_; k < e; k++) line_ += t_[k] == '\n';
        p_ = e + 2;
        continue;
      }
      return;
    }
  }

  static bool name_start(int c) { return std::isalpha(c) || c == '*'; }
  static bool name_char(int c) { return std::isalnum(c) || c == '-' || c == '_' || c == '*'; }

  // settings up to the closing character (or the end of the file for '\0')
  void settings(Node &g, char close) {
    for (;;) {
      int c = peek();
      if (c == -1) {
        if (close) fail("syntax error: missing '" + std::string(1, close) + "'");
        return;
      }
      if (close && c == close) {
        p_++;
        return;
      }
      if (c == '@') fail("@include is not supported");
      if (!name_start(c)) fail("syntax error");
      Node n;
      n.line = line_;
      size_t b = p_;
      while (p_ < t_.size() && name_char((unsigned char)t_[p_])) p_++;
      n.name = t_.substr(b, p_ - b);
      c = peek();
      if (c != '=' && c != ':') fail("syntax error: expected '=' or ':' after " + n.name);
      p_++;
      value(n);
      for (const Node &k : g.kids)
        if (k.name == n.name) fail("duplicate setting name " + n.name);
      g.kids.push_back(std::move(n));
      c = peek();
      if (c == ';' || c == ',') p_++;
    }
  }

  void value(Node &n) {
    int c = peek();
    n.line = line_;
    if (c == '{') {
      p_++;
      n.type = T::Group;
      settings(n, '}');
    } else if (c == '[' || c == '(') {
      p_++;
      const char close = c == '[' ? ']' : ')';
      n.type = c == '[' ? T::Array : T::List;
      if (peek() == close) {
        p_++;
        return;
      }
      for (;;) {
        Node e;
        value(e);
        if (n.type == T::Array) {
          if (e.aggregate()) fail("an array holds scalars only");
          if (!n.kids.empty()) {
            auto cls = [](T x) { return x == T::Int64 ? T::Int : x; };
            if (cls(e.type) != cls(n.kids[0].type)) fail("mismatched element type in array");
          }
        }
        n.kids.push_back(std::move(e));
        c = peek();
        if (c == ',') {
          p_++;
          continue;
        }
        if (c == close) {
          p_++;
          return;
        }
        fail("syntax error in array or list");
      }
    } else if (c == '"') {
      n.type = T::String;
      while (peek() == '"') n.s += string_lit();
    } else {
      scalar(n);
    }
  }

  std::string string_lit() {
    std::string out;
    p_++;  // the opening quote
    for (;;) {
      if (p_ >= t_.size()) fail("unterminated string");
      char ch = t_[p_++];
      if (ch == '"') return out;
      if (ch == '\n') line_++;
      if (ch != '\\') {
        out += ch;
        continue;
      }
      if (p_ >= t_.size()) fail("unterminated string");
      char e = t_[p_++];
      switch (e) {
        case 'n': out += '\n'; break;
        case 't': out += '\t'; break;
        case 'r': out += '\r'; break;
        case 'f': out += '\f'; break;
        case '\\': out += '\\'; break;
        case '"': out += '"'; break;
        case 'x': {
          if (p_ + 2 > t_.size() || !std::isxdigit((unsigned char)t_[p_]) ||
              !std::isxdigit((unsigned char)t_[p_ + 1]))
            fail("bad \\x escape");
          out += (char)std::stoi(t_.substr(p_, 2), nullptr, 16);
          p_ += 2;
          break;
        }
        default: fail(std::string("unknown escape \\") + e);
      }
    }
  }

  void scalar(Node &n) {
    size_t b = p_;
    while (p_ < t_.size() && (std::isalnum((unsigned char)t_[p_]) || t_[p_] == '.' || t_[p_] == '+' ||
                              t_[p_] == '-' || t_[p_] == '_'))
      p_++;
    std::string w = t_.substr(b, p_ - b);
    if (w.empty()) fail("syntax error");
    std::string lw;
    for (char ch : w) lw += (char)std::tolower((unsigned char)ch);
    if (lw == "true" || lw == "false") {
      n.type = T::Bool;
      n.b = lw == "true";
      return;
    }
    // integers: [-+]digits or [-+]0x hex, optional L / LL
    std::string body = w;
    bool wide = false;
    while (!body.empty() && (body.back() == 'L' || body.back() == 'l')) {
      body.pop_back();
      wide = true;
    }
    {
      size_t k = 0;
      if (k < body.size() && (body[k] == '+' || body[k] == '-')) k++;
      bool hex = body.size() > k + 2 && body[k] == '0' && (body[k + 1] == 'x' || body[k + 1] == 'X');
      size_t d0 = hex ? k + 2 : k;
      bool ok = d0 < body.size();
      for (size_t q = d0; q < body.size() && ok; q++)
        ok = hex ? std::isxdigit((unsigned char)body[q]) : std::isdigit((unsigned char)body[q]);
      if (ok) {
        errno = 0;
        char *end = nullptr;
        long long v = hex ? (long long)std::strtoull(body.c_str() + d0, &end, 16) : std::strtoll(body.c_str(), &end, 10);
        if (hex && body[0] == '-') v = -v;
        if (errno == ERANGE) fail("integer out of range: " + w);
        n.i = v;
        n.type = (wide || v > INT_MAX || v < INT_MIN) ? T::Int64 : T::Int;
        return;
      }
    }
    if (wide) fail("syntax error: " + w);
    // floats: must hold a '.' or an exponent and parse completely
    char *end = nullptr;
    const double v = std::strtod(w.c_str(), &end);
    const bool has_digit = w.find_first_of("0123456789") != std::string::npos;
    if (!has_digit || end != w.c_str() + w.size() || w.find_first_of(".eE") == std::string::npos ||
        lw.find("inf") != std::string::npos || lw.find("nan") != std::string::npos || lw.find('x') != std::string::npos)
      fail("syntax error: " + w);
    n.type = T::Float;
    n.f = v;
  }
};

// Config::lookup: a dotted path through groups; null when absent
const Node *lookup(const Node &root, const std::string &path) {
  const Node *cur = &root;
  size_t b = 0;
  while (b <= path.size()) {
    size_t e = path.find('.', b);
    if (e == std::string::npos) e = path.size();
    const std::string part = path.substr(b, e - b);
    if (cur->type != T::Group) return nullptr;
    const Node *nx = nullptr;
    for (const Node &k : cur->kids)
      if (k.name == part) nx = &k;
    if (!nx) return nullptr;
    cur = nx;
    b = e + 1;
  }
  return cur;
}

// scalar reads with the reference's auto-conversion
long long as_int(const Node &n, const std::string &path) {
  if (n.type == T::Int || n.type == T::Int64) return n.i;
  if (n.type == T::Float) return (long long)n.f;
  throw TypeError{path};
}
double as_double(const Node &n, const std::string &path) {
  if (n.type == T::Int || n.type == T::Int64) return (double)n.i;
  if (n.type == T::Float) return n.f;
  throw TypeError{path};
}
// element k of an aggregate (Setting::operator[])
const Node &elem(const Node &n, int k, const std::string &path) {
  if (!n.aggregate() || k >= (int)n.kids.size()) throw TypeError{path};
  return n.kids[k];
}

// a fixed-size vector setting (nCell, xyz0, Lxyz): NDIM values, or the first
// value repeated with a caution
template <class V>
void read_vec3(const Node &n, const std::string &path, const char *name, V *dst) {
  if (n.length() == NDIM) {
    for (int i = 0; i < NDIM; i++) dst[i] = (V)as_double(elem(n, i, path), path);
  } else {
    const V v = (V)as_double(elem(n, 0, path), path);
    for (int i = 0; i < NDIM; i++) dst[i] = v;
    std::cerr << "Caution: " << name << " is not a " << NDIM << " element array in input file!" << std::endl
              << "  Assuming " << name << "[i]=" << v << "." << std::endl;
  }
}

// a per-species setting: nspecies values, or one scalar when nspecies == 1;
// returns false (after the message) on a length mismatch
template <class V>
bool read_species(const Node &n, const std::string &path, const char *msg, int nspecies, V *dst) {
  if (n.length() != nspecies && !(nspecies == 1 && n.length() == 0)) {
    std::cerr << "Error: length of " << msg << " does not match nspecies!" << std::endl;
    return false;
  }
  if (nspecies == 1 && n.length() == 0) {
    dst[0] = (V)as_double(n, path);
  } else {
    for (int i = 0; i < nspecies; i++) dst[i] = (V)as_double(elem(n, i, path), path);
  }
  return true;
}

template <class V>
bool read_waves(const Node *n, const std::string &path, const char *name, const char *lenname, int nwaves, V *dst) {
  if (!n) {
    std::cerr << "Error: " << name << " not set in input file" << std::endl;
    return false;
  }
  if (n->length() != nwaves) {
    std::cerr << "Error: length of " << lenname << " does not match nwaves!" << std::endl;
    return false;
  }
  for (int i = 0; i < nwaves; i++) dst[i] = (V)as_double(elem(*n, i, path), path);
  return true;
}

bool read_bounds(const Node *n, const std::string &path, const char *name, double *dst) {
  if (!n) {
    std::cerr << "Error: " << name << " not set in input file" << std::endl;
    return false;
  }
  if (n->length() != 6) {
    std::cerr << "Error: length of " << name << " is not 6!" << std::endl;
    return false;
  }
  for (int i = 0; i < 6; i++) dst[i] = as_double(elem(*n, i, path), path);
  return true;
}

// boundary condition strings: an array of 2*NDIM strings; returns 0 on
// success, 1 on error (after the message), -1 when absent
int read_conditions(const Node *n, const char *what, char (*dst)[NCHAR], int *nexternal) {
  if (!n) return -1;
  if (n->type != T::Array) {
    std::cerr << "Error: " << what << " boundary conditions should be array of strings!" << std::endl;
    return 1;
  }
  if (n->length() != 2 * NDIM) {
    std::cerr << "Error: length of " << what << " boundary condition is" << n->length() << "." << std::endl
              << "This does not match the number of dimensions!" << std::endl;
    return 1;
  }
  for (int i = 0; i < 2 * NDIM; i++) {
    const Node &e = n->kids[i];
    if (e.type != T::String) {
      std::cerr << "Error: each element of " << what << " boundary should be a string!" << std::endl;
      return 1;
    }
    if (e.s.size() >= (size_t)NCHAR) {
      std::cerr << "Error: " << what << " boundary name longer than " << NCHAR - 1 << " characters" << std::endl;
      return 1;
    }
    std::strcpy(dst[i], e.s.c_str());
    if (nexternal && e.s == "external") (*nexternal)++;
  }
  return 0;
}

int fill_info(const Node &root, Input_Info_t *in) {
  const Node *n;
  auto get = [&](const char *path) { return lookup(root, path); };

  // domain
  if (!(n = get("domain.nCell"))) {
    std::cerr << "Error: nCell not set in input file!" << std::endl;
    return EXIT_FAILURE;
  }
  read_vec3(*n, "domain.nCell", "nCell", in->nCell);
  if ((n = get("domain.xyz0"))) {
    read_vec3(*n, "domain.xyz0", "xyz0", in->xyz0);
  } else {
    std::cerr << "Caution: xyz0 not set in input file!" << std::endl << "  Assuming xyz0[i]=0.0" << std::endl;
  }
  if (!(n = get("domain.Lxyz"))) {
    std::cerr << "Error: Lxyz not set in input file!" << std::endl;
    return EXIT_FAILURE;
  }
  read_vec3(*n, "domain.Lxyz", "Lxyz", in->Lxyz);
  // the serial build does not read domain.nProc
  in->nProc[0] = in->nProc[1] = in->nProc[2] = 1;

  // runtime
  if (!(n = get("runtime.nTimesteps"))) {
    std::cerr << "Error: nTimesteps not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  in->nt = (int)as_int(*n, "runtime.nTimesteps");
  if ((n = get("runtime.startTime"))) {
    in->t0 = as_double(*n, "runtime.startTime");
  } else {
    std::cerr << "Caution: startTime not set in input file!" << std::endl << "  Assuming startTime = 0" << std::endl;
    in->t0 = 0.0;
  }
  if ((n = get("initialization.particles.dens_phys"))) {
    in->dens_phys = as_double(*n, "initialization.particles.dens_phys");
  } else {
    std::cerr << "Caution: dens_phys not set in input file!" << std::endl
              << "  Assuming dens_phys = actual particle number density" << std::endl;
    in->dens_phys = 0.0;
  }
  in->debug = (n = get("runtime.debug")) ? (int)as_int(*n, "runtime.debug") : 0;
  if ((n = get("runtime.nstep_sort"))) {
    in->nstep_sort = (int)as_int(*n, "runtime.nstep_sort");
  } else {
    std::cerr << "nstep_sort not set. Using a sort cadence of 100 steps." << std::endl;
    in->nstep_sort = 100;
  }
  if ((n = get("initialization.restart"))) {
    in->restart = (int)as_int(*n, "initialization.restart");
  } else {
    std::cerr << "restart not set in input file... Assuming restart = 0" << std::endl;
    in->restart = 0;
  }

  // particles
  if (!(n = get("initialization.particles.relativity"))) {
    std::cerr << "Error: relativity not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  in->relativity = (int)as_int(*n, "initialization.particles.relativity");
  if (!(n = get("initialization.particles.nspecies"))) {
    std::cerr << "Error: nSpecies not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  const int nspecies = (int)as_int(*n, "initialization.particles.nspecies");
  in->nspecies = nspecies;
  if (nspecies < 1) {
    std::cerr << "Error: no species in simulation ..." << std::endl;
    return EXIT_FAILURE;
  }
  if (nspecies > NSPEC) {
    std::cerr << "Error: More species are specified than NSPEC!" << std::endl;
    return EXIT_FAILURE;
  }
  struct {
    const char *path, *name, *msg;
    double *dst;
  } sp[] = {{"initialization.particles.mass_ratio", "mass_ratio", "mass ratios", in->mass_ratio},
            {"initialization.particles.charge_ratio", "charge_ratio", "charge ratios", in->charge_ratio},
            {"initialization.particles.temperature", "temperature", "temp", in->temp}};
  for (auto &q : sp) {
    if (!(n = get(q.path))) {
      std::cerr << "Error: " << q.name << " not set in input file" << std::endl;
      return EXIT_FAILURE;
    }
    if (!read_species(*n, q.path, q.msg, nspecies, q.dst)) return EXIT_FAILURE;
  }
  if ((n = get("initialization.particles.dens_frac"))) {
    if (!read_species(*n, "initialization.particles.dens_frac", "dens_frac", nspecies, in->dens_frac))
      return EXIT_FAILURE;
  } else if (nspecies == 1) {
    in->dens_frac[0] = 1;
  } else {
    std::cerr << "Error: dens_frac not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  if ((n = get("initialization.particles.isTestParticle"))) {
    if (!read_species(*n, "initialization.particles.isTestParticle", "isTestParticle", nspecies, in->isTestParticle))
      return EXIT_FAILURE;
  } else {
    std::cerr << "Warning: isTestParticle not set. Assuming no test particle species." << std::endl;
    for (int i = 0; i < nspecies; i++) in->isTestParticle[i] = 0;
  }
  if (!(n = get("initialization.particles.nParticles_tot"))) {
    std::cerr << "Error: nParticles_tot not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  if (n->type == T::Int) {
    in->nparticles_tot = (long)n->i;
    if (in->nparticles_tot < 0) {
      std::cerr << "Error: integer overflow...Use nParticles = #######L in input file"
                << "for long format or use scientific notation." << std::endl;
      return EXIT_FAILURE;
    }
  } else if (n->type == T::Int64) {
    in->nparticles_tot = (long)n->i;
  } else if (n->type == T::Float) {
    in->nparticles_tot = (long)n->f;
  }  // other types leave the value unset, as readinfo.cpp:362-376 does

  int rc = read_conditions(get("boundary.particles.conditions"), "particle", in->parts_bound, nullptr);
  if (rc > 0) return EXIT_FAILURE;
  if (rc < 0) {
    std::cerr << "Caution: particle boundary not set in input file!" << std::endl
              << "  Assume periodic boundary conditions." << std::endl;
    for (int i = 0; i < 2 * NDIM; i++) std::strcpy(in->parts_bound[i], "periodic");
  }

  // fields
  if (!(n = get("initialization.fields.electrostatic"))) {
    std::cerr << "Error: electrostatic not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  in->electrostatic = (int)as_int(*n, "initialization.fields.electrostatic");
  if (!(n = get("initialization.fields.init"))) {
    std::cerr << "Error: field initialization method not set in input file!" << std::endl;
    return EXIT_FAILURE;
  }
  if (n->type != T::String) {
    std::cerr << "Error: field initialization method should be a string!" << std::endl;
    return EXIT_FAILURE;
  }
  if (n->s.size() >= (size_t)NCHAR) {
    std::cerr << "Error: field initialization method longer than " << NCHAR - 1 << " characters" << std::endl;
    return EXIT_FAILURE;
  }
  std::strcpy(in->fields_init, n->s.c_str());
  if ((n = get("initialization.fields.B0"))) {
    if (n->length() == NDIM) {
      for (int i = 0; i < NDIM; i++) in->B0[i] = as_double(elem(*n, i, "B0"), "initialization.fields.B0");
    } else {
      for (int i = 0; i < NDIM; i++) in->B0[i] = 0.0;
      std::cerr << "Caution: B0 is not a " << NDIM << " element array in input file." << std::endl
                << "Assuming B0[i] = 0.0" << std::endl;
    }
  } else {
    for (int i = 0; i < NDIM; i++) in->B0[i] = 0.0;
    std::cerr << "Caution: B0 not set in input file!" << std::endl << "  Assuming B0[i]=0.0" << std::endl;
  }
  if ((n = get("initialization.fields.E0"))) {
    if (n->length() == NDIM) {
      for (int i = 0; i < NDIM; i++) in->E0[i] = as_double(elem(*n, i, "E0"), "initialization.fields.E0");
    } else {
      for (int i = 0; i < NDIM; i++) in->E0[i] = 0.0;
      std::cerr << "Caution: E0 is not a " << NDIM << " element array in input file." << std::endl
                << "Assuming E0[i] = 0.0" << std::endl;
    }
  } else {
    // readinfo.cpp:466-467 zeroes B0 here (E0 keeps its zero initial value)
    for (int i = 0; i < NDIM; i++) in->B0[i] = 0.0;
    std::cerr << "Caution: E0 not set in input file!" << std::endl << "  Assuming E0[i]=0.0" << std::endl;
  }

  int isexternal = 0;
  rc = read_conditions(get("boundary.fields.conditions"), "field", in->fields_bound, &isexternal);
  if (rc > 0) return EXIT_FAILURE;
  if (rc < 0) {
    std::cerr << "Caution: field boundary not set in input file!" << std::endl
              << "  Assume periodic boundary conditions." << std::endl;
    for (int i = 0; i < 2 * NDIM; i++) std::strcpy(in->fields_bound[i], "periodic");
  }
  if (isexternal > 0) {
    std::cout << "    There are " << isexternal << " external field boundaries" << std::endl;
    int nwaves = 0;
    if ((n = get("boundary.fields.external.nwaves"))) {
      nwaves = (int)as_int(*n, "boundary.fields.external.nwaves");
      in->nwaves = nwaves;
      if (nwaves < 1) {
        std::cerr << "Error: no external field specified ..." << std::endl;
        return EXIT_FAILURE;
      }
      if (nwaves > NWAVE) {
        std::cerr << "Error: More external waves are specified than NWAVE!" << std::endl;
        return EXIT_FAILURE;
      }
    } else {
      std::cerr << "Caution: nwave not set in input file" << std::endl
                << "  Assume zero boundary conditions." << std::endl;
      in->nwaves = 0;
    }
    const std::string ext = "boundary.fields.external.";
    auto w = [&](const char *key) { return get((ext + key).c_str()); };
    if (!read_waves(w("inSide"), ext + "inSide", "inSide", "inSide", nwaves, in->inSide) ||
        !read_waves(w("inPolE"), ext + "inPolE", "inPolE", "inPol", nwaves, in->inPolE) ||
        !read_waves(w("peakamps"), ext + "peakamps", "peakamps", "peakamps", nwaves, in->peakamps) ||
        !read_waves(w("omegas"), ext + "omegas", "omegas", "omegas", nwaves, in->omegas) ||
        !read_waves(w("phases"), ext + "phases", "phases", "phases", nwaves, in->phases) ||
        !read_waves(w("invWidths"), ext + "invWidths", "invWidths", "invWidths", nwaves, in->invWidths) ||
        !read_waves(w("delays"), ext + "delays", "delays", "delays", nwaves, in->delays))
      return EXIT_FAILURE;
    if (in->restart == 0 && std::strcmp(in->fields_init, "poisson") == 0) {
      const std::string ps = ext + "poisson.";
      if (!read_bounds(get((ps + "bound_phi").c_str()), ps + "bound_phi", "bound_phi", in->bound_phi) ||
          !read_bounds(get((ps + "bound_Ax").c_str()), ps + "bound_Ax", "bound_Ax", in->bound_Ax) ||
          !read_bounds(get((ps + "bound_Ay").c_str()), ps + "bound_Ay", "bound_Ay", in->bound_Ay) ||
          !read_bounds(get((ps + "bound_Az").c_str()), ps + "bound_Az", "bound_Az", in->bound_Az))
        return EXIT_FAILURE;
    }
  } else {
    in->nwaves = 0;
  }

  // diagnostics
  if ((n = get("diagnostics.nstep_fields"))) {
    in->nstep_fields = (int)as_int(*n, "diagnostics.nstep_fields");
  } else {
    std::cerr << "nstep_fields not set in input file...Assuming nstep_fields = 10" << std::endl;
    in->nstep_fields = 10;
  }
  if ((n = get("diagnostics.nstep_parts"))) {
    in->nstep_parts = (int)as_int(*n, "diagnostics.nstep_parts");
  } else {
    std::cerr << "nstep_parts not set in input file...Assuming nstep_parts = 10" << std::endl;
    in->nstep_parts = 10;
  }
  if ((n = get("diagnostics.nstep_restart"))) {
    in->nstep_restart = (int)as_int(*n, "diagnostics.nstep_restart");
  } else {
    std::cerr << "nwrite not set in input file...Assuming nstep_restart = 100" << std::endl;
    in->nstep_restart = 100;
  }
  if (!(n = get("diagnostics.particles.output_pCount"))) {
    std::cerr << "Error: output_pCount not set in input file" << std::endl;
    return EXIT_FAILURE;
  }
  in->output_pCount = (int)as_int(*n, "diagnostics.particles.output_pCount");
  if ((n = get("diagnostics.fields.which_fields"))) {
    in->which_fields = (int)as_int(*n, "diagnostics.fields.which_fields");
  } else {
    std::cerr << "which_fields not set in input file...Assuming which_fields = -1" << std::endl;
    in->which_fields = -1;
  }

  // the settings this implementation adds
  if ((n = get("emopic.precision"))) in->precision = (int)as_int(*n, "emopic.precision");
  if ((n = get("emopic.seed"))) in->seed = (long)as_int(*n, "emopic.seed");
  if ((n = get("emopic.waist"))) in->waist = as_double(*n, "emopic.waist");
  if ((n = get("emopic.dt"))) in->dt = as_double(*n, "emopic.dt");
  if ((n = get("emopic.poisson_iter_max"))) in->poisson_iter_max = (int)as_int(*n, "emopic.poisson_iter_max");
  if ((n = get("emopic.poisson_tol"))) in->poisson_tol = as_double(*n, "emopic.poisson_tol");
  return 0;
}

}  // namespace

int Input::readinfo(const char *fname) {
  std::ifstream f(fname, std::ios::binary);
  std::cout << "Reading input file " << fname << std::endl;
  if (!f) {
    std::cerr << "I/O error while reading file." << std::endl;
    return EXIT_FAILURE;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  Node root;
  try {
    root = Parser(text).parse();
  } catch (const ParseError &e) {
    std::cerr << "Parse error at " << fname << ":" << e.line << " - " << e.msg << std::endl;
    return EXIT_FAILURE;
  }
  try {
    return fill_info(root, &info_);
  } catch (const TypeError &e) {
    std::cerr << "Error: setting " << e.path << " has the wrong type" << std::endl;
    return EXIT_FAILURE;
  }
}

}  // namespace emopic
