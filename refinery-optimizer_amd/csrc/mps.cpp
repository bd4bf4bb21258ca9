// MPS / QPS reader and writer into the library's owned LP.
//
// Reader: the behaviour of io::read_mps (core/src/io/mps_io.cpp:62-409).
//   * A line whose first character is '*' is a comment; an empty line is
//     skipped.  A line that starts with a non-blank character is a section
//     header (first token, case-insensitive); unknown headers are ignored and
//     leave the current section unchanged; ENDATA stops the read.
//   * Data lines are split into whitespace-separated tokens; a single quote
//     toggles quoting (quotes stay in the token, blanks inside are kept).
//   * ROWS: the first N row is the objective, later N rows are dropped (and
//     so are their COLUMNS entries); a repeated row name keeps its first type.
//   * COLUMNS: "col row val [row val]"; entries on unknown rows are dropped;
//     objective entries are summed; 'MARKER' lines switch integer marking.
//   * RHS / RANGES: an odd token count skips the leading set name.  The RHS of
//     the objective row sets obj_offset = -value.
//   * BOUNDS: "type set col [value]" (value 0 if absent); an unknown column is
//     created, as in COLUMNS.  UP/LO set one side only (an UP below zero does
//     not lower the default 0 lower bound).
//   * QUADOBJ / QMATRIX: "col col value"; off-diagonal entries are mirrored.
//   * Matrices: duplicates summed in file order, |v| <= 1e-15 dropped, CSC in
//     (column, row) order (SparseMatrix::from_triplets, sparse_matrix.cpp:13-88).
//   * Row bounds from type, RHS and RANGES (mps_io.cpp:355-385); then the
//     checks of Problem::validate (problem.cpp:117-149).
//   * Numbers are read with strtod; no conversion, or a result out of range,
//     is an error, as std::stod's exceptions are in the reference.
//
// Writer: the text io::write_mps writes (mps_io.cpp:411-578), byte for byte:
// names padded to 8 characters, numbers with 12 significant digits.  The
// reference pads with a stream whose left alignment is switched on by the
// first COLUMNS line and stays on; a file with no COLUMNS line therefore has
// its later names right-aligned, and this writer does the same.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "mpdhg.h"
#include "mpdhg_internal.hpp"

namespace mpdhg {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kZeroTol = 1e-15;

enum class Sec { None, Name, ObjSense, Rows, Columns, Rhs, Ranges, Bounds, Quad };

std::string upper(std::string s) {
    for (char &ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    return s;
}

bool blank(char ch) { return std::isspace(static_cast<unsigned char>(ch)) != 0; }

void split(const std::string &line, std::vector<std::string> &out) {
    out.clear();
    std::string tok;
    bool quoted = false;
    for (char ch : line) {
        if (ch == '\'') {
            quoted = !quoted;
            tok += ch;
        } else if (!quoted && blank(ch)) {
            if (!tok.empty()) out.push_back(std::move(tok));
            tok.clear();
        } else {
            tok += ch;
        }
    }
    if (!tok.empty()) out.push_back(std::move(tok));
}

double number(const std::string &t) {
    const char *s = t.c_str();
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(s, &end);
    if (end == s) throw std::runtime_error("stod");
    if (errno == ERANGE) throw std::runtime_error("stod");
    return v;
}

struct Entry {
    int64_t row, col;
    double v;
};

// Triplets -> CSC: sum duplicates in input order, drop |v| <= kZeroTol,
// order by (column, row).
void to_csc(int64_t n, std::vector<Entry> &e, std::vector<int64_t> &ptr, std::vector<int64_t> &idx,
            std::vector<double> &val) {
    std::stable_sort(e.begin(), e.end(), [](const Entry &a, const Entry &b) {
        return a.col != b.col ? a.col < b.col : a.row < b.row;
    });
    ptr.assign(static_cast<size_t>(n + 1), 0);
    idx.clear();
    val.clear();
    for (size_t k = 0; k < e.size();) {
        size_t q = k;
        double s = 0.0;
        while (q < e.size() && e[q].col == e[k].col && e[q].row == e[k].row) s += e[q++].v;
        if (std::fabs(s) > kZeroTol) {
            idx.push_back(e[k].row);
            val.push_back(s);
            ++ptr[static_cast<size_t>(e[k].col + 1)];
        }
        k = q;
    }
    for (int64_t j = 0; j < n; ++j) ptr[static_cast<size_t>(j + 1)] += ptr[static_cast<size_t>(j)];
}

struct RowDef {
    char type = 'E';
    double rhs = 0.0, range = 0.0;
    bool ranged = false;
};

class Reader {
  public:
    explicit Reader(OwnedLp &o) : o_(o) {}

    void line(const std::string &ln) {
        if (ln.empty() || ln[0] == '*') return;
        split(ln, tok_);
        if (tok_.empty()) return;
        if (!blank(ln[0])) {
            header();
            return;
        }
        switch (sec_) {
            case Sec::ObjSense: objsense(); break;
            case Sec::Rows: rows(); break;
            case Sec::Columns: columns(); break;
            case Sec::Rhs: pairs(true); break;
            case Sec::Ranges: pairs(false); break;
            case Sec::Bounds: bounds(); break;
            case Sec::Quad: quad(); break;
            default: break;
        }
    }
    bool done() const { return end_; }
    void finish();

  private:
    void header() {
        const std::string h = upper(tok_[0]);
        if (h == "NAME") {
            sec_ = Sec::Name;
            if (tok_.size() > 1) o_.name = tok_[1];
        } else if (h == "OBJSENSE") {
            sec_ = Sec::ObjSense;
        } else if (h == "ROWS") {
            sec_ = Sec::Rows;
        } else if (h == "COLUMNS") {
            sec_ = Sec::Columns;
        } else if (h == "RHS") {
            sec_ = Sec::Rhs;
        } else if (h == "RANGES") {
            sec_ = Sec::Ranges;
        } else if (h == "BOUNDS") {
            sec_ = Sec::Bounds;
        } else if (h == "QUADOBJ" || h == "QMATRIX") {
            sec_ = Sec::Quad;
        } else if (h == "ENDATA") {
            end_ = true;
        }
    }
    void objsense() {
        const std::string s = upper(tok_[0]);
        if (s == "MAX" || s == "MAXIMIZE") o_.sense = -1;
        else if (s == "MIN" || s == "MINIMIZE") o_.sense = 1;
    }
    void rows() {
        if (tok_.size() < 2) return;
        const char type = static_cast<char>(std::toupper(static_cast<unsigned char>(tok_[0][0])));
        const std::string &name = tok_[1];
        if (type == 'N') {
            if (obj_.empty()) obj_ = name;
            return;
        }
        if (row_of_.count(name)) return;
        row_of_.emplace(name, static_cast<int64_t>(o_.row_names.size()));
        o_.row_names.push_back(name);
        RowDef d;
        d.type = type;
        defs_.push_back(d);
    }
    int64_t column(const std::string &name) {
        auto it = col_of_.find(name);
        if (it != col_of_.end()) return it->second;
        const int64_t j = static_cast<int64_t>(o_.col_names.size());
        col_of_.emplace(name, j);
        o_.col_names.push_back(name);
        o_.var_type.push_back(marked_ ? MPDHG_VAR_INTEGER : MPDHG_VAR_CONTINUOUS);
        o_.l.push_back(0.0);
        o_.u.push_back(kInf);
        o_.c.push_back(0.0);
        return j;
    }
    const int64_t *row(const std::string &name) const {
        auto it = row_of_.find(name);
        return it == row_of_.end() ? nullptr : &it->second;
    }
    void columns() {
        if (tok_.size() < 3) return;
        if (tok_[1] == "'MARKER'" || tok_[1] == "MARKER") {
            const std::string k = upper(tok_[2]);
            if (k == "'INTORG'" || k == "INTORG") marked_ = true;
            else if (k == "'INTEND'" || k == "INTEND") marked_ = false;
            return;
        }
        const int64_t j = column(tok_[0]);
        for (size_t k = 1; k + 1 < tok_.size(); k += 2) {
            const double v = number(tok_[k + 1]);
            if (tok_[k] == obj_) {
                o_.c[static_cast<size_t>(j)] += v;
            } else if (const int64_t *i = row(tok_[k])) {
                a_.push_back({*i, j, v});
            }
        }
    }
    void pairs(bool rhs) {
        for (size_t k = tok_.size() % 2; k + 1 < tok_.size(); k += 2) {
            const double v = number(tok_[k + 1]);
            if (rhs && tok_[k] == obj_) {
                o_.obj_offset = -v;
            } else if (const int64_t *i = row(tok_[k])) {
                RowDef &d = defs_[static_cast<size_t>(*i)];
                if (rhs) {
                    d.rhs = v;
                } else {
                    d.range = v;
                    d.ranged = true;
                }
            }
        }
    }
    void bounds() {
        if (tok_.size() < 3) return;
        const std::string t = upper(tok_[0]);
        const double v = tok_.size() >= 4 ? number(tok_[3]) : 0.0;
        const size_t j = static_cast<size_t>(column(tok_[2]));
        double &lo = o_.l[j], &hi = o_.u[j];
        int8_t &vt = o_.var_type[j];
        if (t == "UP") hi = v;
        else if (t == "LO") lo = v;
        else if (t == "FX") lo = hi = v;
        else if (t == "FR") { lo = -kInf; hi = kInf; }
        else if (t == "MI") { lo = -kInf; hi = 0.0; }
        else if (t == "PL") { lo = 0.0; hi = kInf; }
        else if (t == "BV") { vt = MPDHG_VAR_BINARY; lo = 0.0; hi = 1.0; }
        else if (t == "UI") { vt = MPDHG_VAR_INTEGER; hi = v; }
        else if (t == "LI") { vt = MPDHG_VAR_INTEGER; lo = v; }
    }
    void quad() {
        if (tok_.size() < 3) return;
        const double v = number(tok_[2]);
        const int64_t a = column(tok_[0]);
        const int64_t b = column(tok_[1]);
        q_.push_back({a, b, v});
        if (a != b) q_.push_back({b, a, v});
    }

    OwnedLp &o_;
    Sec sec_ = Sec::None;
    bool end_ = false, marked_ = false;
    std::string obj_;
    std::vector<std::string> tok_;
    std::unordered_map<std::string, int64_t> row_of_, col_of_;
    std::vector<RowDef> defs_;
    std::vector<Entry> a_, q_;
};

void Reader::finish() {
    o_.m = static_cast<int64_t>(o_.row_names.size());
    o_.n = static_cast<int64_t>(o_.col_names.size());
    o_.rl.assign(static_cast<size_t>(o_.m), -kInf);
    o_.ru.assign(static_cast<size_t>(o_.m), kInf);
    for (size_t i = 0; i < defs_.size(); ++i) {
        const RowDef &d = defs_[i];
        double &lo = o_.rl[i], &hi = o_.ru[i];
        const double r = d.range, w = std::fabs(d.range);
        switch (d.type) {
            case 'G':
                lo = d.rhs;
                hi = d.ranged ? d.rhs + w : kInf;
                break;
            case 'L':
                lo = d.ranged ? d.rhs - w : -kInf;
                hi = d.rhs;
                break;
            case 'E':
                lo = (d.ranged && r < 0.0) ? d.rhs + r : d.rhs;
                hi = (d.ranged && r >= 0.0) ? d.rhs + r : d.rhs;
                break;
            default:
                break;  // any other row type stays free
        }
    }
    to_csc(o_.n, a_, o_.col_ptr, o_.row_ind, o_.vals);
    o_.has_q = !q_.empty();
    if (o_.has_q) to_csc(o_.n, q_, o_.q_col_ptr, o_.q_row_ind, o_.q_vals);
    // Problem::validate (problem.cpp:136-149): NaN bounds pass, as there.
    for (int64_t i = 0; i < o_.m; ++i)
        if (o_.rl[static_cast<size_t>(i)] > o_.ru[static_cast<size_t>(i)])
            throw std::runtime_error("Contradictory row bounds at row " + std::to_string(i));
    for (int64_t j = 0; j < o_.n; ++j)
        if (o_.l[static_cast<size_t>(j)] > o_.u[static_cast<size_t>(j)])
            throw std::runtime_error("Contradictory column bounds at column " + std::to_string(j));
}

// ---- writer ----------------------------------------------------------------
class Text {
  public:
    void put(const std::string &s) { buf_ += s; }
    void put(const char *s) { buf_ += s; }
    // A name in an 8-character field: left-aligned once `left_` is on.
    void field(const std::string &s) {
        const size_t pad = s.size() < 8 ? 8 - s.size() : 0;
        if (left_) {
            buf_ += s;
            buf_.append(pad, ' ');
        } else {
            buf_.append(pad, ' ');
            buf_ += s;
        }
    }
    void num(double v) {
        char b[64];
        std::snprintf(b, sizeof b, "%.12g", v);
        buf_ += b;
    }
    void left() { left_ = true; }
    const std::string &str() const { return buf_; }

  private:
    std::string buf_;
    bool left_ = false;
};

bool neg_inf(double v) { return std::isinf(v) ? v < 0.0 : v <= -1e20; }
bool pos_inf(double v) { return std::isinf(v) ? v > 0.0 : v >= 1e20; }

}  // namespace

OwnedLp *read_mps(const std::string &path) {
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("Could not open MPS file: " + path);
    auto o = std::make_unique<OwnedLp>();
    Reader rd(*o);
    std::string ln;
    while (!rd.done() && std::getline(in, ln)) rd.line(ln);
    rd.finish();
    finish_views(o.get());
    return o.release();
}

void write_mps(const std::string &path, const mpdhg_lp &lp, const mpdhg_lp_meta *meta) {
    const int64_t m = lp.m, n = lp.n;
    auto rname = [&](int64_t i) {
        return (meta && meta->row_names) ? std::string(meta->row_names[i]) : "R" + std::to_string(i);
    };
    auto cname = [&](int64_t j) {
        return (meta && meta->col_names) ? std::string(meta->col_names[j]) : "C" + std::to_string(j);
    };
    auto vtype = [&](int64_t j) { return (meta && meta->var_type) ? meta->var_type[j] : MPDHG_VAR_CONTINUOUS; };
    const bool maximize = meta && meta->sense == -1;
    const double offset = meta ? meta->obj_offset : 0.0;
    Text t;
    t.put("NAME          ");
    t.put(meta && meta->name ? meta->name : "problem");
    t.put("\nOBJSENSE\n  ");
    t.put(maximize ? "MAX" : "MIN");
    t.put("\nROWS\n N  OBJ\n");
    for (int64_t i = 0; i < m; ++i) {
        const double lo = lp.row_lower[i], hi = lp.row_upper[i];
        char type = 'L';  // a boxed row is written as L with a RANGES entry
        if (neg_inf(lo) && !pos_inf(hi)) type = 'L';
        else if (!neg_inf(lo) && pos_inf(hi)) type = 'G';
        else if (lo == hi) type = 'E';
        t.put(" ");
        t.put(std::string(1, type));
        t.put("  ");
        t.put(rname(i));
        t.put("\n");
    }
    t.put("COLUMNS\n");
    bool marked = false;
    for (int64_t j = 0; j < n; ++j) {
        const std::string cn = cname(j);
        const bool integral = vtype(j) == MPDHG_VAR_INTEGER || vtype(j) == MPDHG_VAR_BINARY;
        if (integral && !marked) {
            t.put("    MARK0000  'MARKER'                 'INTORG'\n");
            marked = true;
        } else if (!integral && marked) {
            t.put("    MARK0001  'MARKER'                 'INTEND'\n");
            marked = false;
        }
        auto entry = [&](const std::string &row, double v) {
            t.put("    ");
            t.left();
            t.field(cn);
            t.put("  ");
            t.field(row);
            t.put("  ");
            t.num(v);
            t.put("\n");
        };
        if (std::fabs(lp.c[j]) > 1e-15) entry("OBJ", lp.c[j]);
        for (int64_t k = lp.col_ptr[j]; k < lp.col_ptr[j + 1]; ++k) entry(rname(lp.row_ind[k]), lp.vals[k]);
    }
    if (marked) t.put("    MARK9999  'MARKER'                 'INTEND'\n");
    t.put("RHS\n");
    if (std::fabs(offset) > 1e-15) {
        t.put("    RHS1      OBJ       ");
        t.num(-offset);
        t.put("\n");
    }
    auto rhs_line = [&](int64_t i, double v) {
        t.put("    RHS1      ");
        t.field(rname(i));
        t.put("  ");
        t.num(v);
        t.put("\n");
    };
    for (int64_t i = 0; i < m; ++i) {
        const double lo = lp.row_lower[i], hi = lp.row_upper[i];
        if (neg_inf(lo) && !pos_inf(hi)) {
            if (std::fabs(hi) > 1e-15) rhs_line(i, hi);
        } else if (!neg_inf(lo) && pos_inf(hi)) {
            if (std::fabs(lo) > 1e-15) rhs_line(i, lo);
        } else if (lo == hi) {
            if (std::fabs(lo) > 1e-15) rhs_line(i, lo);
        } else {
            rhs_line(i, hi);
        }
    }
    auto boxed = [&](int64_t i) {
        const double lo = lp.row_lower[i], hi = lp.row_upper[i];
        return !neg_inf(lo) && !pos_inf(hi) && lo != hi;
    };
    bool any_range = false;
    for (int64_t i = 0; i < m && !any_range; ++i) any_range = boxed(i);
    if (any_range) {
        t.put("RANGES\n");
        for (int64_t i = 0; i < m; ++i) {
            if (!boxed(i)) continue;
            t.put("    RNG1      ");
            t.field(rname(i));
            t.put("  ");
            t.num(lp.row_upper[i] - lp.row_lower[i]);
            t.put("\n");
        }
    }
    t.put("BOUNDS\n");
    auto bound = [&](const char *kind, const std::string &cn, double v) {
        t.put(" ");
        t.put(kind);
        t.put(" BND1      ");
        t.field(cn);
        t.put("  ");
        t.num(v);
        t.put("\n");
    };
    for (int64_t j = 0; j < n; ++j) {
        const std::string cn = cname(j);
        const double lo = lp.col_lower[j], hi = lp.col_upper[j];
        if (vtype(j) == MPDHG_VAR_BINARY) {
            t.put(" BV BND1      " + cn + "\n");
        } else if (lo == hi) {
            bound("FX", cn, lo);
        } else if (neg_inf(lo) && pos_inf(hi)) {
            t.put(" FR BND1      " + cn + "\n");
        } else {
            if (!neg_inf(lo) && lo != 0.0) bound("LO", cn, lo);
            if (!pos_inf(hi)) bound(vtype(j) == MPDHG_VAR_INTEGER ? "UI" : "UP", cn, hi);
        }
    }
    if (meta && meta->has_quadratic && meta->q_nnz > 0) {
        t.put("QUADOBJ\n");
        for (int64_t j = 0; j < n; ++j) {
            for (int64_t k = meta->q_col_ptr[j]; k < meta->q_col_ptr[j + 1]; ++k) {
                const int64_t i = meta->q_row_ind[k];
                if (i > j) continue;
                t.put("    ");
                t.left();
                t.field(cname(i));
                t.put("  ");
                t.field(cname(j));
                t.put("  ");
                t.num(meta->q_vals[k]);
                t.put("\n");
            }
        }
    }
    t.put("ENDATA\n");
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("Could not open file for writing MPS: " + path);
    const size_t w = std::fwrite(t.str().data(), 1, t.str().size(), f);
    const int rc = std::fclose(f);
    if (w != t.str().size() || rc != 0) throw std::runtime_error("Could not write MPS file: " + path);
}

}  // namespace mpdhg

extern "C" {

int mpdhg_read_mps(const char *path, mpdhg_lp_owned **out) {
    if (!path || !out) return mpdhg::fail(MPDHG_E_ARG, "read_mps: NULL argument");
    try {
        *out = reinterpret_cast<mpdhg_lp_owned *>(mpdhg::read_mps(path));
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "read_mps: out of host memory");
    } catch (const std::exception &e) {
        return mpdhg::fail(MPDHG_E_ARG, e.what());
    }
    return MPDHG_OK;
}

const mpdhg_lp_meta *mpdhg_lp_meta_view(const mpdhg_lp_owned *lp) {
    return lp ? &reinterpret_cast<const mpdhg::OwnedLp *>(lp)->meta : nullptr;
}

int mpdhg_write_mps(const char *path, const mpdhg_lp *lp, const mpdhg_lp_meta *meta, int32_t) {
    if (!path || !lp) return mpdhg::fail(MPDHG_E_ARG, "write_mps: NULL argument");
    if (lp->m < 0 || lp->n < 0 || (lp->n > 0 && (!lp->col_ptr || !lp->c || !lp->col_lower || !lp->col_upper)) ||
        (lp->m > 0 && (!lp->row_lower || !lp->row_upper)) || (lp->nnz > 0 && (!lp->row_ind || !lp->vals)))
        return mpdhg::fail(MPDHG_E_ARG, "write_mps: malformed LP");
    if (lp->n > 0) {
        if (lp->col_ptr[0] != 0 || lp->col_ptr[lp->n] != lp->nnz)
            return mpdhg::fail(MPDHG_E_ARG, "write_mps: col_ptr must start at 0 and end at nnz");
        for (int64_t j = 0; j < lp->n; ++j)
            if (lp->col_ptr[j + 1] < lp->col_ptr[j]) return mpdhg::fail(MPDHG_E_ARG, "write_mps: col_ptr not monotone");
    } else if (lp->nnz != 0) {
        return mpdhg::fail(MPDHG_E_ARG, "write_mps: nonzeros without columns");
    }
    for (int64_t k = 0; k < lp->nnz; ++k)
        if (lp->row_ind[k] < 0 || lp->row_ind[k] >= lp->m)
            return mpdhg::fail(MPDHG_E_ARG, "write_mps: row index out of range");
    if (meta && meta->has_quadratic && meta->q_nnz > 0 && (!meta->q_col_ptr || !meta->q_row_ind || !meta->q_vals))
        return mpdhg::fail(MPDHG_E_ARG, "write_mps: quadratic arrays missing");
    if (meta && meta->has_quadratic && meta->q_nnz > 0) {
        if (meta->q_col_ptr[0] != 0 || meta->q_col_ptr[lp->n] != meta->q_nnz)
            return mpdhg::fail(MPDHG_E_ARG, "write_mps: q_col_ptr must start at 0 and end at q_nnz");
        for (int64_t j = 0; j < lp->n; ++j)
            if (meta->q_col_ptr[j + 1] < meta->q_col_ptr[j])
                return mpdhg::fail(MPDHG_E_ARG, "write_mps: q_col_ptr not monotone");
        for (int64_t k = 0; k < meta->q_nnz; ++k)
            if (meta->q_row_ind[k] < 0 || meta->q_row_ind[k] >= lp->n)
                return mpdhg::fail(MPDHG_E_ARG, "write_mps: Q row index out of range");
    }
    try {
        mpdhg::write_mps(path, *lp, meta);
    } catch (const std::bad_alloc &) {
        return mpdhg::fail(MPDHG_E_NOMEM, "write_mps: out of host memory");
    } catch (const std::exception &e) {
        return mpdhg::fail(MPDHG_E_ARG, e.what());
    }
    return MPDHG_OK;
}

}  // extern "C"
