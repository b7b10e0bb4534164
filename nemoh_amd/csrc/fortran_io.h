// Text formats of the reference's data files, as its Fortran runtime
// writes and reads them (the reference is built with flang, REAL = 8 bytes).
//
// Writing:
//   * list-directed (WRITE(u,*)): every item is preceded by one blank; a
//     record never exceeds 79 characters, the item that would pass that
//     starts a new record.  Reals take the shortest digit string that reads
//     back to the same double, in F form ("169626.19526811474", ".5", "1.")
//     when the value rounded to one significant digit is 0.d x 10^e with
//     0 <= e <= 15, else in E form ("4.2053656216012314E-04", "1.E+15");
//     zero is "0." (or "-0.").  An empty WRITE gives an empty record.
//   * edit descriptors Ew.d (0.ddddddd form, "E+dd" or "+ddd" exponent),
//     Fw.d, Iw, nX.
// Reading (READ(u,*)): one statement takes the items it needs from the
// current record and the following ones, separated by blanks, tabs or a
// comma; a slash ends the statement; "r*v" repeats; afterwards the rest of
// the last record read is skipped.  Unquoted character items end at a
// blank, comma or slash.
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace nemoh {
namespace fio {

struct FormatError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::string list_real(double x);                 // one list-directed REAL(8) item, without its blank
std::string edit_E(double x, int w, int d);      // Ew.d
std::string edit_F(double x, int w, int d);      // Fw.d
std::string edit_I(long long v, int w);          // Iw

// One formatted WRITE: the items under a format such as
// "(F7.4,6(X,E13.6),6(X,F7.2))" (descriptors nX, A, Iw, Ew.d, Fw.d and
// repeated groups one level deep).  Blanks from X only appear before a
// later character of the record; when the items outlast the format a new
// record starts and the format reverts to its last top-level group (or to
// its start), as the Fortran runtime does.  Returns the records, each
// ending in a newline.
struct FmtItem {
  enum Kind { Real, Int, Text } kind;
  double r = 0;
  long long i = 0;
  std::string s;
  FmtItem(double v) : kind(Real), r(v) {}
  FmtItem(int v) : kind(Int), i(v) {}
  FmtItem(long long v) : kind(Int), i(v) {}
  FmtItem(const char* v) : kind(Text), s(v) {}
  FmtItem(const std::string& v) : kind(Text), s(v) {}
};
std::string format_write(const std::string& format, const std::vector<FmtItem>& items);

// A file written record by record.
class Writer {
 public:
  // list-directed statement: start, items, end
  Writer& begin() {
    col_ = 0;
    open_ = true;
    return *this;
  }
  Writer& real(double x) { return item(list_real(x)); }
  Writer& integer(long long v) { return item(std::to_string(v)); }
  Writer& text(const std::string& s) { return item(s); }
  Writer& end() {
    out_ += '\n';
    open_ = false;
    return *this;
  }
  // a formatted record, written as is
  Writer& line(const std::string& s) {
    out_ += s;
    out_ += '\n';
    return *this;
  }
  const std::string& str() const { return out_; }
  void save(const std::string& path) const;

 private:
  Writer& item(const std::string& s);
  std::string out_;
  int col_ = 0;
  bool open_ = false;
};

// A file read statement by statement.
class Reader {
 public:
  explicit Reader(const std::string& path);
  Reader(const std::string& name, const std::string& text);
  // READ(u,*) with no items: skips one record
  void skip();
  // the text of the current record (a formatted '(A)' read without moving)
  const std::string& peek_record() const {
    if (r_ >= rec_.size()) fail("end of file");
    return rec_[r_];
  }
  // one READ(u,*) statement: begin, items, end
  Reader& begin();
  long long integer();
  double real();
  std::string word();
  void end();
  const std::string& name() const { return name_; }

 private:
  std::string token(bool character);
  [[noreturn]] void fail(const std::string& what) const;
  std::string name_;
  std::vector<std::string> rec_;
  size_t r_ = 0, p_ = 0;
  bool in_stmt_ = false, slash_ = false;
  std::string repeat_val_;
  long long repeat_left_ = 0;
};

}  // namespace fio
}  // namespace nemoh
