// uppaal.cpp — UPPAAL XML front end: model file -> smc::Network.
//
// Reads the same subset of UPPAAL's flat XML format as cuSMC's
// uppaal_xml_parser (automata_parser/uppaal_xml_parser.cpp:623-706) and gives
// each construct the meaning that parser gives it, quirks included, so a
// model file means the same thing to both tools:
//   * every <template> is one component, <system> is ignored (:639-690);
//   * node ids "idN" become N+1 (string_extractor.cpp:66-70);
//   * a location reads only its FIRST <label> (:398-417);
//   * a location without a rate label races at rate 1 (:385-387), an edge
//     without a probability label has weight 1 (:472-477);
//   * a location named "Goal" is a goal location (:394);
//   * guards and invariants are split on '&' and lose their parentheses
//     (helper_methods.h:74-91); a conjunct whose left side is a variable
//     constrains that variable, any other conjunct is kept but ignored by the
//     simulator (:182-192, shared_model_state.cu:355-441);
//   * assignments are split on ',' and may be `c ? a : b` (:549-583, 496-524);
//   * declared values go through float and 6-decimal text
//     (declaration_parser.cpp:18,33, arithmetic-expression-evaluator.h:292)
//     and initial values are truncated to int (shared_model_state.cu:282);
//   * variable ids follow declaration order, globals first (:228-317).
// Deliberate differences, each a fix of undefined or crashing behaviour:
//   trailing // comments are stripped, several declarations may share a line,
//   a declaration keyword is removed only
//   at the start of the line, constants of one template do not leak into the
//   next, `c1 && c2 && c3` conditions build a proper chain (build_con at
//   :526-547 loops on itself), and unknown names are reported as errors.
// The XML reader is this project's own small DOM reader.

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "network.hpp"
#include "uppaal.hpp"

namespace smc {

// ------------------------------------------------------------ XML reader

namespace {

struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<std::unique_ptr<XNode>> kids;
    std::string text;        // first non-blank text child (pugixml child_value)
    bool has_text = false;

    const XNode *child(const char *n) const {
        for (auto &k : kids) if (k->name == n) return k.get();
        return nullptr;
    }
    std::vector<const XNode *> children(const char *n) const {
        std::vector<const XNode *> r;
        for (auto &k : kids) if (k->name == n) r.push_back(k.get());
        return r;
    }
    std::string attr(const char *n) const {
        for (auto &a : attrs) if (a.first == n) return a.second;
        return "";
    }
};

class XmlReader {
public:
    explicit XmlReader(const std::string &s) : s_(s) {}

    std::unique_ptr<XNode> parse() {
        auto root = std::make_unique<XNode>();
        root->name = "#document";
        std::vector<XNode *> stack{root.get()};
        while (p_ < s_.size()) {
            if (s_[p_] != '<') {
                size_t q = s_.find('<', p_);
                if (q == std::string::npos) q = s_.size();
                add_text(stack.back(), decode(s_.substr(p_, q - p_)));
                p_ = q;
                continue;
            }
            if (starts("<?")) { skip_past("?>"); continue; }
            if (starts("<!--")) { skip_past("-->"); continue; }
            if (starts("<![CDATA[")) {
                size_t q = s_.find("]]>", p_ + 9);
                if (q == std::string::npos) fail("unterminated CDATA");
                add_text(stack.back(), s_.substr(p_ + 9, q - p_ - 9));
                p_ = q + 3;
                continue;
            }
            if (starts("<!")) { skip_decl(); continue; }
            if (starts("</")) {
                p_ += 2;
                std::string n = name();
                skip_ws();
                expect('>');
                if (stack.size() < 2 || stack.back()->name != n) fail("mismatched </" + n + ">");
                stack.pop_back();
                continue;
            }
            ++p_;
            auto el = std::make_unique<XNode>();
            el->name = name();
            for (;;) {
                skip_ws();
                if (p_ >= s_.size()) fail("unterminated tag <" + el->name + ">");
                if (s_[p_] == '/') { ++p_; expect('>'); stack.back()->kids.push_back(std::move(el)); break; }
                if (s_[p_] == '>') {
                    ++p_;
                    XNode *raw = el.get();
                    stack.back()->kids.push_back(std::move(el));
                    stack.push_back(raw);
                    break;
                }
                std::string an = name();
                skip_ws();
                expect('=');
                skip_ws();
                char qch = s_[p_];
                if (qch != '"' && qch != '\'') fail("attribute value must be quoted");
                size_t q = s_.find(qch, p_ + 1);
                if (q == std::string::npos) fail("unterminated attribute");
                el->attrs.emplace_back(an, decode(s_.substr(p_ + 1, q - p_ - 1)));
                p_ = q + 1;
            }
        }
        if (stack.size() != 1) fail("unclosed element <" + stack.back()->name + ">");
        return root;
    }

private:
    const std::string &s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string &m) { throw std::runtime_error("XML: " + m); }
    bool starts(const char *t) const { return s_.compare(p_, strlen(t), t) == 0; }
    void skip_past(const char *t) {
        size_t q = s_.find(t, p_);
        if (q == std::string::npos) fail(std::string("missing ") + t);
        p_ = q + strlen(t);
    }
    void skip_decl() {  // <!DOCTYPE ...> possibly with [ ... ]
        int depth = 0;
        while (p_ < s_.size()) {
            char c = s_[p_++];
            if (c == '[') ++depth;
            else if (c == ']') --depth;
            else if (c == '>' && depth <= 0) return;
        }
        fail("unterminated declaration");
    }
    void skip_ws() { while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_; }
    void expect(char c) {
        if (p_ >= s_.size() || s_[p_] != c) fail(std::string("expected '") + c + "'");
        ++p_;
    }
    std::string name() {
        size_t b = p_;
        while (p_ < s_.size() && (isalnum((unsigned char)s_[p_]) || strchr("_-.:", s_[p_]))) ++p_;
        if (b == p_) fail("expected a name");
        return s_.substr(b, p_ - b);
    }
    static void add_text(XNode *n, const std::string &t) {
        if (n->has_text) return;
        bool blank = true;
        for (char c : t) if (!isspace((unsigned char)c)) { blank = false; break; }
        if (blank) return;
        n->text = t;
        n->has_text = true;
    }
    std::string decode(const std::string &t) {
        std::string r;
        r.reserve(t.size());
        for (size_t i = 0; i < t.size(); ++i) {
            if (t[i] != '&') { r += t[i]; continue; }
            size_t semi = t.find(';', i);
            if (semi == std::string::npos) { r += t[i]; continue; }
            std::string ent = t.substr(i + 1, semi - i - 1);
            if (ent == "lt") r += '<';
            else if (ent == "gt") r += '>';
            else if (ent == "amp") r += '&';
            else if (ent == "quot") r += '"';
            else if (ent == "apos") r += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                unsigned long cp = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X')
                                       ? strtoul(ent.c_str() + 2, nullptr, 16)
                                       : strtoul(ent.c_str() + 1, nullptr, 10);
                if (cp < 0x80) r += (char)cp;
                else if (cp < 0x800) { r += (char)(0xC0 | (cp >> 6)); r += (char)(0x80 | (cp & 0x3F)); }
                else { r += (char)(0xE0 | (cp >> 12)); r += (char)(0x80 | ((cp >> 6) & 0x3F)); r += (char)(0x80 | (cp & 0x3F)); }
            } else { r += t.substr(i, semi - i + 1); }
            i = semi;
        }
        return r;
    }
};

// ------------------------------------------------------------ string helpers

std::string replace_all(std::string s, const std::string &from, const std::string &to) {
    if (from.empty()) return s;
    size_t p = 0;
    while ((p = s.find(from, p)) != std::string::npos) { s.replace(p, from.size(), to); p += to.size(); }
    return s;
}
std::string remove_spaces(const std::string &s) { return replace_all(s, " ", ""); }
std::string strip_ws(const std::string &s) {
    std::string r;
    for (char c : s) if (!isspace((unsigned char)c)) r += c;
    return r;
}
std::string trim(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && isspace((unsigned char)s[b])) ++b;
    while (e > b && isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}
std::string take_while(const std::string &s, const std::string &t) {
    size_t p = s.find(t);
    return p == std::string::npos ? s : s.substr(0, p);
}
std::string take_after(const std::string &s, const std::string &t) {
    size_t p = s.find(t);
    return p == std::string::npos ? s : s.substr(p + t.size());
}
// std::getline-style split: no trailing empty piece.
std::vector<std::string> split_getline(const std::string &s, char d) {
    std::vector<std::string> r;
    std::stringstream ss(s);
    std::string seg;
    while (std::getline(ss, seg, d)) r.push_back(seg);
    return r;
}
// Guard/invariant/condition splitting (helper_methods.h:74-91): pieces between
// '&', with every parenthesis removed except the argument list of random().
std::vector<std::string> split_conjuncts(const std::string &s) {
    std::vector<std::string> r;
    for (std::string seg : split_getline(s, '&')) {
        std::string out;
        for (size_t i = 0; i < seg.size(); ++i) {
            if (seg.compare(i, 7, "random(") == 0) {
                size_t close = seg.rfind(')');
                if (close != std::string::npos && close > i + 6) {
                    std::string inner = seg.substr(i + 7, close - i - 7);
                    inner = replace_all(replace_all(inner, "(", ""), ")", "");
                    out += "random(" + inner + ")";
                    i = close;
                    continue;
                }
            }
            if (seg[i] != '(' && seg[i] != ')') out += seg[i];
        }
        r.push_back(out);
    }
    return r;
}
// Relational operator of a conjunct, by the reference's search order
// (string_extractor.cpp:3-19).
std::string find_rel_op(const std::string &c) {
    for (const char *op : {"<=", ">=", "==", "!=", "<", ">"})
        if (c.find(op) != std::string::npos) return op;
    throw std::runtime_error("no relational operator in '" + c + "'");
}
int cons_op_of(const std::string &op) {
    if (op == "<=") return SMC_C_LE;
    if (op == ">=") return SMC_C_GE;
    if (op == "==") return SMC_C_EQ;
    if (op == "!=") return SMC_C_NE;
    if (op == "<") return SMC_C_LT;
    return SMC_C_GT;
}
int expr_op_of(const std::string &op) {
    if (op == "<=") return SMC_E_LE;
    if (op == ">=") return SMC_E_GE;
    if (op == "==") return SMC_E_EQ;
    if (op == "!=") return SMC_E_NE;
    if (op == "<") return SMC_E_LT;
    return SMC_E_GT;
}
std::string to_string6(double v) {  // std::to_string(double): "%f"
    char buf[512];
    snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

// ---------------------------------------------- declaration value evaluator
//
// Values in declarations are computed like the reference's evalguy parser
// (arithmetic-expression-evaluator.h): double arithmetic, '%' on ints, '^'
// as pow, upper-case functions, names resolved first among every name
// declared so far (stored through float) and then among global constants.

class DeclEval {
public:
    DeclEval(const std::unordered_map<std::string, double> &decl_vals,
             const std::unordered_map<std::string, double> &globals)
        : dv_(decl_vals), gv_(globals) {}

    double eval(const std::string &s) {
        s_ = s; p_ = 0;
        next();
        if (tok_.empty()) throw std::runtime_error("empty declaration value");
        double r = exp2();
        if (!tok_.empty()) throw std::runtime_error("syntax error in declaration value '" + s + "'");
        return r;
    }

private:
    enum { DELIM = 1, VARIABLE, NUMBER, FUNCTION };
    const std::unordered_map<std::string, double> &dv_, &gv_;
    std::string s_, tok_;
    size_t p_ = 0;
    int type_ = 0;

    void next() {
        tok_.clear(); type_ = 0;
        while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
        if (p_ >= s_.size()) return;
        char c = s_[p_];
        if (strchr("+-*/%^=()", c)) { tok_ = c; ++p_; type_ = DELIM; return; }
        if (isalpha((unsigned char)c)) {
            while (p_ < s_.size() && !strchr(" +-/*%^=()\t\r", s_[p_])) tok_ += s_[p_++];
            while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
            type_ = (p_ < s_.size() && s_[p_] == '(') ? FUNCTION : VARIABLE;
            return;
        }
        if (isdigit((unsigned char)c) || c == '.') {
            while (p_ < s_.size() && !strchr(" +-/*%^=()\t\r", s_[p_])) tok_ += s_[p_++];
            type_ = NUMBER;
            return;
        }
        throw std::runtime_error(std::string("unexpected character '") + c + "' in declaration value");
    }
    double exp2() {
        double r = exp3();
        while (tok_ == "+" || tok_ == "-") {
            char op = tok_[0]; next();
            double t = exp3();
            r = op == '+' ? r + t : r - t;
        }
        return r;
    }
    double exp3() {
        double r = exp4();
        while (tok_ == "*" || tok_ == "/" || tok_ == "%") {
            char op = tok_[0]; next();
            double t = exp4();
            if (op == '*') r = r * t;
            else if (op == '/') r = r / t;
            else {
                int b = (int)t;
                if (b == 0) throw std::runtime_error("modulo by zero in declaration value");
                r = (double)((int)r % b);
            }
        }
        return r;
    }
    double exp4() {
        double r = exp5();
        while (tok_ == "^") { next(); r = pow(r, exp5()); }
        return r;
    }
    double exp5() {
        char op = 0;
        if (type_ == DELIM && (tok_ == "+" || tok_ == "-")) { op = tok_[0]; next(); }
        double r = exp6();
        return op == '-' ? -r : r;
    }
    double exp6() {
        if (type_ == FUNCTION) {
            std::string f = tok_;
            next();
            if (tok_ != "(") throw std::runtime_error("expected '(' after " + f);
            next();
            double r = exp2();
            if (tok_ != ")") throw std::runtime_error("unbalanced parentheses in declaration value");
            next();
            if (f == "SIN") return sin(M_PI / 180 * r);
            if (f == "COS") return cos(M_PI / 180 * r);
            if (f == "TAN") return tan(M_PI / 180 * r);
            if (f == "ASIN") return 180 / M_PI * asin(r);
            if (f == "ACOS") return 180 / M_PI * acos(r);
            if (f == "ATAN") return 180 / M_PI * atan(r);
            if (f == "SINH") return sinh(r);
            if (f == "COSH") return cosh(r);
            if (f == "TANH") return tanh(r);
            if (f == "ASINH") return asinh(r);
            if (f == "ACOSH") return acosh(r);
            if (f == "ATANH") return atanh(r);
            if (f == "LN") return log(r);
            if (f == "LOG") return log10(r);
            if (f == "EXP") return exp(r);
            if (f == "SQRT") return sqrt(r);
            if (f == "SQR") return r * r;
            if (f == "ROUND") return round(r);
            if (f == "INT") return floor(r);
            throw std::runtime_error("unknown function " + f + " in declaration value");
        }
        if (tok_ == "(") {
            next();
            double r = exp2();
            if (tok_ != ")") throw std::runtime_error("unbalanced parentheses in declaration value");
            next();
            return r;
        }
        if (type_ == VARIABLE) {
            std::string n = tok_;
            next();
            auto it = dv_.find(n);
            if (it != dv_.end()) return it->second;
            auto jt = gv_.find(n);
            if (jt != gv_.end()) return jt->second;
            throw std::runtime_error("name '" + n + "' not declared");
        }
        if (type_ == NUMBER) { double r = atof(tok_.c_str()); next(); return r; }
        throw std::runtime_error("syntax error in declaration value near '" + tok_ + "'");
    }
};

// ------------------------------------------------------------ front end

class UppaalFrontEnd {
public:
    Network net;
    std::vector<int> goal_locs;

    void parse(const std::string &xml, const std::string &loc_query) {
        XmlReader rd(xml);
        auto doc = rd.parse();
        const XNode *nta = doc->child("nta");
        if (!nta) throw std::runtime_error("no <nta> element");

        const XNode *gd = nta->child("declaration");
        declare(gd ? gd->text : "", true);
        const_global_["true"] = 1.0;
        const_global_["false"] = 0.0;

        for (const XNode *tpl : nta->children("template")) parse_template(tpl);
        fold_all();
        mark_goals(loc_query);
    }

private:
    struct Decl { int kind; std::string name; double value; };  // kind 0 clock,1 chan,2 const,3 var
    std::unordered_map<std::string, double> decl_vals_;   // evalguy's name table
    std::unordered_map<std::string, double> const_global_, const_local_;
    std::unordered_map<std::string, int> gvars_, gclocks_, lvars_, lclocks_;
    std::unordered_map<std::string, int> gchans_, lchans_;
    int next_chan_ = 1;
    std::vector<std::string> loc_raw_names_;

    // ---- declarations (declaration_parser.cpp:9-104) ----
    void declare(const std::string &raw, bool global) {
        std::string text = remove_spaces(raw);
        std::vector<std::string> stmts;
        for (std::string line : split_getline(text, '\n')) {
            size_t cpos = line.find("//");
            if (cpos != std::string::npos) line = line.substr(0, cpos);
            for (const std::string &st : split_getline(line, ';')) stmts.push_back(st);
        }
        for (const std::string &line : stmts) {
            std::string l;
            for (char c : line) if (c != '\t' && c != '\r' && c != '\v' && c != '\f') l += c;
            if (l.empty()) continue;
            if (l.rfind("clock", 0) == 0) add_decls(l.substr(5), 0, global);
            else if (l.rfind("const", 0) == 0) {
                std::string rest = l.substr(5);
                for (const char *kw : {"double", "bool", "int"})
                    if (rest.rfind(kw, 0) == 0) { add_decls(rest.substr(strlen(kw)), 2, global); break; }
            } else if (l.rfind("broadcastchan", 0) == 0) add_decls(l.substr(13), 1, global);
            else {
                bool done = false;
                for (const char *kw : {"double", "bool", "int"})
                    if (l.rfind(kw, 0) == 0) { add_decls(l.substr(strlen(kw)), 3, global); done = true; break; }
                if (!done && l.rfind("chan", 0) != 0 && l.rfind("typedef", 0) != 0 && l != ";")
                    throw std::runtime_error("unsupported declaration '" + l + "'");
            }
        }
    }
    void add_decls(const std::string &body, int kind, bool global) {
        // split on ',' first, then each piece is `names[=value];`
        for (std::string piece : split_getline(body, ',')) {
            piece = replace_all(piece, ";", "");
            std::string left = piece.find(":=") != std::string::npos ? take_while(piece, ":=")
                             : piece.find('=') != std::string::npos ? take_while(piece, "=") : piece;
            std::string right = piece.find('=') != std::string::npos ? take_after(piece, "=") : "";
            right = replace_all(replace_all(right, "true", "1"), "false", "0");
            std::string valstr;
            if (!right.empty()) {
                DeclEval ev(decl_vals_, const_global_);
                float f = (float)ev.eval(right);
                valstr = to_string6((double)f);
            }
            if (left.empty()) continue;
            double value = valstr.empty() ? 0.0 : strtod(valstr.c_str(), nullptr);
            decl_vals_[left] = valstr.empty() ? 0.0 : (double)strtof(valstr.c_str(), nullptr);
            register_name(kind, left, value, global);
        }
    }
    void register_name(int kind, const std::string &name, double value, bool global) {
        if (kind == 1) { (global ? gchans_ : lchans_)[name] = next_chan_++; return; }
        if (kind == 2) { (global ? const_global_ : const_local_)[name] = value; return; }
        int id = (int)net.var_kind.size();
        net.var_kind.push_back(kind == 0 ? SMC_VAR_CLOCK : SMC_VAR_INT);
        net.var_init.push_back((double)(int)value);  // int initial values (shared_model_state.cu:282)
        net.var_names.push_back(name);
        (global ? gvars_ : lvars_)[name] = id;
        if (kind == 0) (global ? gclocks_ : lclocks_)[name] = id;
    }

    // ---- expressions (variable_expression_evaluator.cpp) ----
    class ExprParser {
    public:
        ExprParser(UppaalFrontEnd &fe, const std::string &s) : fe_(fe), s_(s) {}
        int parse() {
            next();
            if (tok_.empty()) throw std::runtime_error("empty expression");
            // eval_exp1: `name = value` evaluates value
            if (type_ == VARIABLE) {
                size_t save_p = p_; std::string save_t = tok_; int save_ty = type_;
                next();
                if (tok_ != "=") { p_ = save_p; tok_ = save_t; type_ = save_ty; }
                else next();
            }
            int r = exp2();
            if (!tok_.empty()) throw std::runtime_error("syntax error in expression '" + s_ + "'");
            return r;
        }
    private:
        enum { DELIM = 1, VARIABLE, NUMBER, FUNCTION };
        UppaalFrontEnd &fe_;
        std::string s_, tok_;
        size_t p_ = 0;
        int type_ = 0;
        Network &N() { return fe_.net; }

        void next() {
            tok_.clear(); type_ = 0;
            while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
            if (p_ >= s_.size()) return;
            char c = s_[p_];
            if (strchr("+-*/%^=()!", c)) { tok_ = c; ++p_; type_ = DELIM; return; }
            if (isalpha((unsigned char)c) || c == '_') {
                while (p_ < s_.size() && !strchr(" +-/*%^=()!\t\r\n", s_[p_])) tok_ += s_[p_++];
                while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
                type_ = (p_ < s_.size() && s_[p_] == '(') ? FUNCTION : VARIABLE;
                return;
            }
            if (isdigit((unsigned char)c) || c == '.') {
                while (p_ < s_.size() && !strchr(" +-/*%^=()!\t\r\n", s_[p_])) tok_ += s_[p_++];
                type_ = NUMBER;
                return;
            }
            throw std::runtime_error(std::string("unexpected character '") + c + "' in expression '" + s_ + "'");
        }
        int exp2() {
            int r = exp3();
            while (tok_ == "+" || tok_ == "-") {
                char op = tok_[0]; next();
                int t = exp3();
                r = N().add_expr(op == '+' ? SMC_E_ADD : SMC_E_SUB, r, t);
            }
            return r;
        }
        int exp3() {
            int r = exp4();
            while (tok_ == "*" || tok_ == "/" || tok_ == "%") {
                char op = tok_[0]; next();
                int t = exp4();
                r = N().add_expr(op == '*' ? SMC_E_MUL : op == '/' ? SMC_E_DIV : SMC_E_MOD, r, t);
            }
            return r;
        }
        int exp4() {
            int r = exp5();
            while (tok_ == "^") { next(); int t = exp5(); r = N().add_expr(SMC_E_POW, r, t); }
            return r;
        }
        int exp5() {
            char op = 0;
            if ((type_ == DELIM && tok_ == "!") || tok_ == "-") { op = tok_[0]; next(); }
            int r = exp6();
            if (op == '-') r = N().add_expr(SMC_E_NEG, r);
            if (op == '!') r = N().add_expr(SMC_E_NOT, r);
            return r;
        }
        int exp6() {
            bool func = type_ == FUNCTION;
            std::string fname;
            if (func) { fname = tok_; next(); }
            if (tok_ == "(") {
                next();
                int r = exp2();
                if (tok_ != ")") throw std::runtime_error("unbalanced parentheses in '" + s_ + "'");
                if (func) {
                    if (fname == "random") r = N().add_expr(SMC_E_RANDOM, r);
                    else if (fname == "sqrt") r = N().add_expr(SMC_E_SQRT, r);
                    else throw std::runtime_error("unknown function " + fname);
                }
                next();
                return r;
            }
            if (type_ == VARIABLE) {
                std::string n = tok_;
                next();
                double cv;
                int vid;
                if (fe_.lookup_const(n, &cv)) return N().lit(cv);
                if (fe_.lookup_var(n, &vid)) return N().var(vid);
                throw std::runtime_error("variable '" + n + "' not declared");
            }
            if (type_ == NUMBER) {
                double v = strtod(tok_.c_str(), nullptr);
                next();
                return N().lit(v);
            }
            throw std::runtime_error("syntax error in '" + s_ + "' near '" + tok_ + "'");
        }
    };

    bool lookup_const(const std::string &n, double *v) const {
        auto it = const_local_.find(n);
        if (it != const_local_.end()) { *v = it->second; return true; }
        it = const_global_.find(n);
        if (it != const_global_.end()) { *v = it->second; return true; }
        return false;
    }
    bool lookup_var(const std::string &n, int *id) const {
        auto it = lvars_.find(n);
        if (it != lvars_.end()) { *id = it->second; return true; }
        it = gvars_.find(n);
        if (it != gvars_.end()) { *id = it->second; return true; }
        return false;
    }
    bool is_clock(const std::string &n) const { return lclocks_.count(n) || gclocks_.count(n); }
    bool is_var_name(const std::string &n) const { return lvars_.count(n) || gvars_.count(n); }
    int parse_expr(const std::string &s) { return ExprParser(*this, s).parse(); }

    // Name before the first relational character (get_timer_id, :133-163).
    int timer_id(const std::string &cond) const {
        size_t i = 1;
        while (i < cond.size() && !strchr("<>=!", cond[i])) ++i;
        std::string sub = strip_ws(cond.substr(0, i));
        int id;
        if (lookup_var(sub, &id)) return id;
        throw std::runtime_error("'" + sub + "' is not a variable");
    }

    // Guard and invariant conjuncts (fill_expressions, :165-194).
    void add_conjuncts(const std::string &text, std::vector<smc_constraint> *out) {
        for (const std::string &piece : split_conjuncts(text)) {
            std::string c = remove_spaces(piece);
            if (c.empty()) continue;
            std::string op = find_rel_op(c);
            std::string left = strip_ws(take_while(c, op)), right = take_after(c, op);
            smc_constraint k{};
            k.op = cons_op_of(op);
            k.lhs = -1;
            if (is_clock(left)) {
                k.uses_variable = 1;
                k.var = timer_id(c);
                k.rhs = parse_expr(right);
            } else {
                int l = parse_expr(left);
                k.rhs = parse_expr(right);
                if (net.exprs[l].op == SMC_E_VAR) { k.uses_variable = 1; k.var = net.exprs[l].var; }
                else { k.uses_variable = 0; k.lhs = l; }
            }
            out->push_back(k);
        }
    }

    // `cond ? a : b` (handle_if_statement, :496-524).
    int parse_conditional(const std::string &input) {
        std::string cond = take_while(input, "?");
        std::string if_true = take_while(take_after(input, "?"), ":");
        std::string if_false = replace_all(take_after(input, ":"), ";", "");
        std::vector<int> parts;
        for (const std::string &piece : split_conjuncts(cond)) {
            std::string c = remove_spaces(piece);
            if (c.empty()) continue;
            std::string op = find_rel_op(c);
            std::string left = strip_ws(take_while(c, op)), right = take_after(c, op);
            int l = (is_clock(left) || is_var_name(left)) ? net.var(timer_id(c)) : parse_expr(left);
            int r = parse_expr(right);
            parts.push_back(net.add_expr(expr_op_of(op), l, r));
        }
        if (parts.empty()) throw std::runtime_error("empty condition in '" + input + "'");
        int chain = parts.back();
        for (int i = (int)parts.size() - 2; i >= 0; --i) chain = net.add_expr(SMC_E_AND, parts[i], chain);
        int t = parse_expr(if_true), f = parse_expr(if_false);
        return net.add_expr(SMC_E_COND, chain, t, f);
    }

    // Assignments (handle_assignment, :549-583).
    void add_assignments(const std::string &text, std::vector<smc_update> *out) {
        out->clear();
        for (const std::string &piece : split_getline(text, ',')) {
            std::string in = remove_spaces(piece);
            if (in.empty()) continue;
            std::string left = in.find(":=") != std::string::npos ? take_while(in, ":=") : take_while(in, "=");
            std::string right = take_after(in, "=");
            right = replace_all(replace_all(right, "true", "1"), "false", "0");
            int e = (right.find('?') != std::string::npos && right.find(':') != std::string::npos)
                        ? parse_conditional(right) : parse_expr(right);
            int v;
            left = strip_ws(left);
            if (!lookup_var(left, &v)) throw std::runtime_error("assignment to undeclared '" + left + "'");
            out->push_back(smc_update{v, e});
        }
    }

    int node_id(const std::string &ref) const {
        std::string r = trim(ref);
        if (r.size() < 3) throw std::runtime_error("bad node id '" + ref + "'");
        return atoi(r.c_str() + 2) + 1;
    }

    struct PendingEdge { int src_node; int dst_node; smc_edge e; std::vector<smc_constraint> guards; std::vector<smc_update> upds; };

    void parse_template(const XNode *tpl) {
        const XNode *nm = tpl->child("name");
        std::string comp_name = trim(nm ? nm->text : "");
        lvars_.clear(); lclocks_.clear(); lchans_.clear(); const_local_.clear();
        const XNode *decl = tpl->child("declaration");
        declare(decl ? decl->text : "", false);

        const XNode *init = tpl->child("init");
        if (!init) throw std::runtime_error("template '" + comp_name + "' has no <init>");
        int init_node = node_id(init->attr("ref"));

        int first = (int)net.locations.size();
        std::map<int, int> node2loc;
        std::vector<std::vector<smc_constraint>> invs;

        for (const XNode *loc : tpl->children("location")) {
            int nid = node_id(loc->attr("id"));
            const XNode *lname = loc->child("name");
            std::string raw = lname ? lname->text : "";
            smc_location L{};
            L.type = SMC_LOC_NORMAL;
            L.rate = -1;
            if (trim(raw) == "Goal") L.type = SMC_LOC_GOAL;
            else if (loc->child("committed")) L.type = SMC_LOC_COMMITTED;
            else if (loc->child("urgent")) L.type = SMC_LOC_URGENT;
            std::vector<smc_constraint> inv;
            const XNode *label = loc->child("label");  // only the first label counts
            if (label) {
                std::string kind = label->attr("kind");
                if (kind == "exponentialrate") {
                    std::string nums = remove_spaces(label->text);
                    size_t eq = nums.find('=');
                    if (eq != std::string::npos) nums = nums.substr(eq + 1);
                    L.rate = parse_expr(nums);
                } else if (kind == "invariant") {
                    add_conjuncts(label->text, &inv);
                }
            }
            node2loc[nid] = (int)net.locations.size();
            net.locations.push_back(L);
            net.loc_names.push_back(trim(raw));
            loc_raw_names_.push_back(raw);
            invs.push_back(inv);
        }
        for (const XNode *bp : tpl->children("branchpoint")) {
            int nid = node_id(bp->attr("id"));
            smc_location L{};
            L.type = SMC_LOC_BRANCH;
            L.rate = -1;
            node2loc[nid] = (int)net.locations.size();
            net.locations.push_back(L);
            net.loc_names.push_back("");
            loc_raw_names_.push_back("");
            invs.push_back({});
        }
        int last = (int)net.locations.size();
        if (last == first) throw std::runtime_error("template '" + comp_name + "' has no locations");
        auto it = node2loc.find(init_node);
        if (it == node2loc.end()) throw std::runtime_error("template '" + comp_name + "' init is not one of its locations");

        // transitions, grouped by source location in file order
        std::vector<std::vector<PendingEdge>> out_edges(last - first);
        for (const XNode *tr : tpl->children("transition")) {
            PendingEdge pe{};
            const XNode *s = tr->child("source"), *t = tr->child("target");
            if (!s || !t) throw std::runtime_error("transition without source or target");
            int sn = node_id(s->attr("ref")), tn = node_id(t->attr("ref"));
            auto si = node2loc.find(sn), ti = node2loc.find(tn);
            if (si == node2loc.end() || ti == node2loc.end())
                throw std::runtime_error("transition refers to a location outside template '" + comp_name + "'");
            pe.e.src = si->second;
            pe.e.dst = ti->second;
            pe.e.channel = 0;
            pe.e.weight = -1;
            for (const XNode *lab : tr->children("label")) {
                std::string kind = lab->attr("kind");
                const std::string &txt = lab->text;
                if (kind == "guard") add_conjuncts(txt, &pe.guards);
                else if (kind == "assignment") add_assignments(txt, &pe.upds);
                else if (kind == "synchronisation") {
                    std::string in = remove_spaces(txt);
                    bool listener = in.find('?') != std::string::npos;
                    std::string key = strip_ws(replace_all(in, listener ? "?" : "!", ""));
                    auto ci = lchans_.find(key);
                    int ch;
                    if (ci != lchans_.end()) ch = ci->second;
                    else {
                        auto gi = gchans_.find(key);
                        if (gi == gchans_.end()) throw std::runtime_error("channel '" + key + "' not declared");
                        ch = gi->second;
                    }
                    pe.e.channel = listener ? -ch : ch;
                } else if (kind == "probability") {
                    pe.e.weight = parse_expr(remove_spaces(txt));
                }
            }
            out_edges[pe.e.src - first].push_back(std::move(pe));
        }
        for (int l = first; l < last; ++l) {
            smc_location &L = net.locations[l];
            L.inv_first = (int)net.constraints.size();
            L.inv_count = (int)invs[l - first].size();
            for (auto &k : invs[l - first]) net.constraints.push_back(k);
            L.edge_first = (int)net.edges.size();
            L.edge_count = (int)out_edges[l - first].size();
            for (auto &pe : out_edges[l - first]) {
                smc_edge e = pe.e;
                e.guard_first = (int)net.constraints.size();
                e.guard_count = (int)pe.guards.size();
                for (auto &g : pe.guards) net.constraints.push_back(g);
                e.upd_first = (int)net.updates.size();
                e.upd_count = (int)pe.upds.size();
                for (auto &u : pe.upds) net.updates.push_back(u);
                net.edges.push_back(e);
            }
        }
        net.comp_first_loc.push_back(last);
        net.comp_init_loc.push_back(it->second);
        net.comp_names.push_back(comp_name);
    }

    // Constant folding (domain_optimization_visitor.cpp:432-441): every
    // maximal constant subtree becomes a literal, evaluated with the
    // run-time operator semantics.
    void fold_all() {
        smc_network v = net.view();
        for (int e = 0; e < (int)net.exprs.size(); ++e) {
            if (net.exprs[e].op == SMC_E_LIT || net.exprs[e].op == SMC_E_VAR) continue;
            double val;
            if (fold_constant(v, e, &val)) {
                smc_expr_node &x = net.exprs[e];
                x.op = SMC_E_LIT; x.a = x.b = x.c = -1; x.value = val;
            }
        }
    }

    // Goal marking for `-q ... Process.location` (smc.cu:405-436 and the
    // optimizer's exact Process.location match).  The exact match wins; if
    // no component has that name, the first location in file order with that
    // location name is used (the reference picks one arbitrarily).
    void mark_goals(const std::string &q) {
        for (int l = 0; l < (int)net.locations.size(); ++l)
            if (net.locations[l].type == SMC_LOC_GOAL) goal_locs.push_back(l);
        if (q.empty() || trim(q).empty()) return;
        std::string comp, node;
        bool period = false;
        for (char c : q) {
            if (c == '.') { period = true; continue; }
            (period ? node : comp) += c;
        }
        comp = trim(comp); node = trim(node);
        std::vector<int> exact;
        for (int c = 0; c < (int)net.comp_names.size(); ++c) {
            if (net.comp_names[c] != comp) continue;
            for (int l = net.comp_first_loc[c]; l < net.comp_first_loc[c + 1]; ++l)
                if (!node.empty() && net.loc_names[l] == node) exact.push_back(l);
        }
        if (exact.empty() && !node.empty()) {
            for (int l = 0; l < (int)net.locations.size(); ++l)
                if (net.loc_names[l] == node) { exact.push_back(l); break; }
        }
        for (int l : exact) {
            net.locations[l].type = SMC_LOC_GOAL;
            goal_locs.push_back(l);
        }
    }
};

}  // namespace

std::unique_ptr<Parsed> parse_uppaal_string(const std::string &xml, const std::string &loc_query) {
    UppaalFrontEnd fe;
    fe.parse(xml, loc_query);
    auto p = std::make_unique<Parsed>();
    p->net = std::move(fe.net);
    p->goal_count = 0;
    for (int l = 0; l < (int)p->net.locations.size(); ++l)
        if (p->net.locations[l].type == SMC_LOC_GOAL) ++p->goal_count;
    p->query_goal_count = (int)fe.goal_locs.size();
    std::string err = validate(p->net.view());
    if (!err.empty()) throw std::runtime_error("parsed network is invalid: " + err);
    return p;
}

std::unique_ptr<Parsed> parse_uppaal_file(const std::string &path, const std::string &loc_query) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open model file '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_uppaal_string(ss.str(), loc_query);
}

}  // namespace smc
