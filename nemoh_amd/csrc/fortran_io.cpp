// Fortran text formats of the reference's data files (see fortran_io.h).
#include "fortran_io.h"

#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace nemoh {
namespace fio {

namespace {

// Shortest round-trip digits of |x| > 0 and the decimal exponent of the
// first digit: |x| = d1.d2d3... x 10^e10.
void shortest(double ax, std::string& digits, int& e10) {
  char b[64];
  auto r = std::to_chars(b, b + sizeof b, ax, std::chars_format::scientific);
  std::string s(b, r.ptr);
  const size_t epos = s.find('e');
  e10 = std::atoi(s.c_str() + epos + 1);
  digits.clear();
  for (size_t i = 0; i < epos; ++i)
    if (s[i] != '.') digits += s[i];
}

std::string exponent(int e, bool with_letter_when_short) {
  char b[16];
  const int ae = std::abs(e);
  if (ae <= 99)
    std::snprintf(b, sizeof b, "%s%c%02d", with_letter_when_short ? "E" : "", e < 0 ? '-' : '+', ae);
  else
    std::snprintf(b, sizeof b, "%c%d", e < 0 ? '-' : '+', ae);
  return b;
}

}  // namespace

std::string list_real(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-Inf" : "Inf";
  const std::string sign = std::signbit(x) ? "-" : "";
  if (x == 0.0) return sign + "0.";
  const double ax = std::fabs(x);
  // the decimal exponent of |x| rounded to one significant digit, in the
  // 0.d x 10^expo convention, decides the form
  char b[32];
  std::snprintf(b, sizeof b, "%.0e", ax);
  const int expo = std::atoi(std::strchr(b, 'e') + 1) + 1;
  std::string dg;
  int e10;
  shortest(ax, dg, e10);
  if (expo >= 0 && expo <= 15) {
    const int ip = e10 + 1;  // digits before the point
    std::string s;
    if (ip <= 0) {
      s = "." + std::string(-ip, '0') + dg;
    } else if ((int)dg.size() <= ip) {
      s = dg + std::string(ip - dg.size(), '0') + ".";
    } else {
      s = dg.substr(0, ip) + "." + dg.substr(ip);
    }
    return sign + s;
  }
  return sign + dg.substr(0, 1) + "." + dg.substr(1) + "E" + exponent(e10, false);
}

std::string edit_E(double x, int w, int d) {
  std::string s = std::signbit(x) ? "-" : "";
  int e = 0;
  std::string dg;
  if (x == 0.0) {
    dg.assign(d, '0');
  } else {
    char b[64];
    std::snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(x));
    const char* ep = std::strchr(b, 'e');
    for (const char* p = b; p < ep; ++p)
      if (*p != '.') dg += *p;
    e = std::atoi(ep + 1) + 1;
  }
  s += "0." + dg + exponent(e, true);
  if ((int)s.size() > w && (s[0] == '0' || (s[0] == '-' && s[1] == '0'))) {
    s.erase(s[0] == '-' ? 1 : 0, 1);  // the optional leading zero goes first
  }
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

std::string edit_F(double x, int w, int d) {
  char b[64];
  std::snprintf(b, sizeof b, "%.*f", d, x);
  std::string s = b;
  if ((int)s.size() > w && (s.rfind("0.", 0) == 0 || s.rfind("-0.", 0) == 0)) s.erase(s[0] == '-' ? 1 : 0, 1);
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

std::string edit_I(long long v, int w) {
  std::string s = std::to_string(v);
  if ((int)s.size() > w) return std::string(w, '*');
  return std::string(w - s.size(), ' ') + s;
}

namespace {

// A parsed format: a flat list of descriptors with group markers.
struct Desc {
  char op;  // 'X' blanks, 'A', 'I', 'E', 'F', '(' group start, ')' group end
  int w = 0, d = 0, rep = 1;
  int match = -1;  // for '(' / ')': index of the partner
};

std::vector<Desc> parse_format(const std::string& f) {
  std::vector<Desc> out;
  std::vector<int> open;
  size_t i = 0;
  auto num = [&](int dflt) {
    if (i >= f.size() || !isdigit((unsigned char)f[i])) return dflt;
    int v = 0;
    while (i < f.size() && isdigit((unsigned char)f[i])) v = 10 * v + (f[i++] - '0');
    return v;
  };
  while (i < f.size()) {
    const char c = f[i];
    if (c == ' ' || c == ',') {
      ++i;
      continue;
    }
    if (c == ')') {
      ++i;
      if (open.empty()) throw FormatError("unbalanced format " + f);
      Desc d{')'};
      d.match = open.back();
      out[open.back()].match = (int)out.size();
      open.pop_back();
      out.push_back(d);
      continue;
    }
    const int rep = num(1);
    if (i >= f.size()) throw FormatError("bad format " + f);
    const char op = (char)toupper((unsigned char)f[i++]);
    Desc d{op};
    d.rep = rep;
    if (op == '(') {
      open.push_back((int)out.size());
    } else if (op == 'X') {
      d.w = rep;
      d.rep = 1;
    } else if (op == 'A') {
      d.w = num(0);
    } else if (op == 'I') {
      d.w = num(0);
    } else if (op == 'E' || op == 'F') {
      d.w = num(0);
      if (i < f.size() && f[i] == '.') {
        ++i;
        d.d = num(0);
      }
    } else {
      throw FormatError(std::string("unsupported descriptor '") + op + "' in " + f);
    }
    out.push_back(d);
  }
  if (!open.empty()) throw FormatError("unbalanced format " + f);
  return out;
}

}  // namespace

std::string format_write(const std::string& format, const std::vector<FmtItem>& items) {
  const std::vector<Desc> prog = parse_format(format);
  if (prog.empty() || prog.front().op != '(' || prog.front().match != (int)prog.size() - 1)
    throw FormatError("a format is one parenthesised list: " + format);
  // the reversion point: the last group directly inside the outer parentheses
  int revert = 1;
  for (int k = 1; k < (int)prog.size() - 1; ++k)
    if (prog[k].op == '(') {
      revert = k;
      k = prog[k].match;
    }
  std::string out, rec;
  int pending = 0;
  size_t next = 0;
  auto put = [&](const std::string& s) {
    rec.append(pending, ' ');
    pending = 0;
    rec += s;
  };
  // group repeat counters (one level of nesting)
  int pc = 1, left = 0;
  bool started = false;
  for (;;) {
    const Desc& d = prog[pc];
    if (d.op == ')') {
      if (pc == (int)prog.size() - 1) {  // the outer right parenthesis
        if (next >= items.size()) break;
        out += rec + "\n";
        rec.clear();
        pending = 0;
        pc = revert;
        started = false;
        continue;
      }
      if (--left > 0) {
        pc = d.match + 1;
      } else {
        ++pc;
        started = false;
      }
      continue;
    }
    if (d.op == '(') {
      left = d.rep;
      started = true;
      ++pc;
      continue;
    }
    if (d.op == 'X') {
      pending += d.w;
      ++pc;
      continue;
    }
    // a data descriptor, repeated d.rep times
    for (int r = 0; r < d.rep; ++r) {
      if (next >= items.size()) goto done;
      const FmtItem& it = items[next++];
      if (d.op == 'A') {
        if (it.kind != FmtItem::Text) throw FormatError("A needs a character item in " + format);
        put(d.w > 0 ? (it.s.size() >= (size_t)d.w ? it.s.substr(0, d.w) : std::string(d.w - it.s.size(), ' ') + it.s)
                    : it.s);
      } else if (d.op == 'I') {
        if (it.kind != FmtItem::Int) throw FormatError("I needs an integer item in " + format);
        put(edit_I(it.i, d.w));
      } else {
        if (it.kind != FmtItem::Real) throw FormatError("E/F need a real item in " + format);
        put(d.op == 'E' ? edit_E(it.r, d.w, d.d) : edit_F(it.r, d.w, d.d));
      }
    }
    ++pc;
  }
done:
  (void)started;
  out += rec + "\n";
  return out;
}

Writer& Writer::item(const std::string& s) {
  const int len = 1 + (int)s.size();
  if (col_ > 0 && col_ + len > 79) {
    out_ += '\n';
    col_ = 0;
  }
  out_ += ' ';
  out_ += s;
  col_ += len;
  return *this;
}

void Writer::save(const std::string& path) const {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw FormatError("cannot open " + path + " for writing");
  f << out_;
  if (!f) throw FormatError("cannot write " + path);
}

Reader::Reader(const std::string& path) : name_(path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw FormatError("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  *this = Reader(path, ss.str());
}

Reader::Reader(const std::string& name, const std::string& text) : name_(name) {
  size_t a = 0;
  while (a <= text.size()) {
    size_t b = text.find('\n', a);
    if (b == std::string::npos) {
      if (a < text.size()) rec_.push_back(text.substr(a));
      break;
    }
    std::string l = text.substr(a, b - a);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    rec_.push_back(l);
    a = b + 1;
  }
}

void Reader::fail(const std::string& what) const {
  throw FormatError(name_ + ", record " + std::to_string(r_ + 1) + ": " + what);
}

void Reader::skip() {
  if (r_ >= rec_.size()) fail("end of file");
  ++r_;
  p_ = 0;
}

Reader& Reader::begin() {
  if (r_ >= rec_.size()) fail("end of file");
  p_ = 0;
  in_stmt_ = true;
  slash_ = false;
  repeat_left_ = 0;
  return *this;
}

void Reader::end() {
  in_stmt_ = false;
  ++r_;  // the rest of the last record read is skipped
  p_ = 0;
}

static bool is_sep(char c) { return c == ' ' || c == '\t' || c == ','; }

std::string Reader::token(bool character) {
  if (!in_stmt_) fail("item read outside a statement");
  if (repeat_left_ > 0) {
    --repeat_left_;
    return repeat_val_;
  }
  if (slash_) fail("slash ended the statement before all items were read");
  for (;;) {
    if (r_ >= rec_.size()) fail("end of file");
    const std::string& l = rec_[r_];
    while (p_ < l.size() && (l[p_] == ' ' || l[p_] == '\t')) ++p_;
    if (p_ < l.size() && l[p_] == ',') {
      ++p_;
      while (p_ < l.size() && (l[p_] == ' ' || l[p_] == '\t')) ++p_;
    }
    if (p_ >= l.size()) {  // the statement continues on the next record
      ++r_;
      p_ = 0;
      continue;
    }
    if (l[p_] == '/') {
      slash_ = true;
      fail("slash ended the statement before all items were read");
    }
    std::string t;
    if (character && (l[p_] == '\'' || l[p_] == '"')) {
      const char q = l[p_++];
      while (p_ < l.size()) {
        if (l[p_] == q) {
          if (p_ + 1 < l.size() && l[p_ + 1] == q) {
            t += q;
            p_ += 2;
            continue;
          }
          ++p_;
          break;
        }
        t += l[p_++];
      }
      return t;
    }
    while (p_ < l.size() && !is_sep(l[p_]) && l[p_] != '/') t += l[p_++];
    const size_t star = t.find('*');
    if (!character && star != std::string::npos && star > 0) {
      char* e = nullptr;
      const long long n = std::strtoll(t.c_str(), &e, 10);
      if (e != t.c_str() + star || n <= 0) fail("bad repeat count in '" + t + "'");
      repeat_val_ = t.substr(star + 1);
      repeat_left_ = n - 1;
      return repeat_val_;
    }
    return t;
  }
}

long long Reader::integer() {
  const std::string t = token(false);
  char* e = nullptr;
  const long long v = std::strtoll(t.c_str(), &e, 10);
  if (t.empty() || *e != '\0') fail("bad character in INTEGER input field '" + t + "'");
  return v;
}

double Reader::real() {
  std::string t = token(false);
  for (char& c : t)
    if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
  char* e = nullptr;
  const double v = std::strtod(t.c_str(), &e);
  if (t.empty() || *e != '\0') fail("bad character in REAL input field '" + t + "'");
  return v;
}

std::string Reader::word() { return token(true); }

}  // namespace fio
}  // namespace nemoh
