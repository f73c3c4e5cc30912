// OSO reader (host).
#include "oslmi_oso.h"

#include <cctype>
#include <cstdlib>
#include <sstream>

namespace oslmi {

int ShaderMaster::find(const std::string& n) const
{
    for (size_t i = 0; i < syms.size(); ++i)
        if (syms[i].name == n)
            return (int)i;
    return -1;
}

namespace {

struct Tok {
    enum K { kIdent, kInt, kFloat, kString, kHint, kPunct } k;
    std::string s;
};

// Splits one line into tokens; '#' starts a comment outside strings.
bool tokenize(const std::string& line, std::vector<Tok>& toks, std::string& err)
{
    size_t i = 0, n = line.size();
    while (i < n) {
        char c = line[i];
        if (isspace((unsigned char)c)) {
            ++i;
            continue;
        }
        if (c == '#')
            break;
        if (c == '"') {
            std::string s;
            ++i;
            while (i < n && line[i] != '"') {
                if (line[i] == '\\' && i + 1 < n) {
                    char e = line[i + 1];
                    s += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
                    i += 2;
                } else {
                    s += line[i++];
                }
            }
            if (i >= n) {
                err = "unterminated string";
                return false;
            }
            ++i;
            toks.push_back({ Tok::kString, s });
            continue;
        }
        if (c == '%') {
            // %name or %name{...} (braces may hold quoted strings with spaces)
            size_t j = i + 1;
            while (j < n && (isalnum((unsigned char)line[j]) || line[j] == '_'))
                ++j;
            if (j < n && line[j] == '{') {
                int depth   = 0;
                bool in_str = false;
                for (; j < n; ++j) {
                    char d = line[j];
                    if (in_str) {
                        if (d == '\\')
                            ++j;
                        else if (d == '"')
                            in_str = false;
                    } else if (d == '"') {
                        in_str = true;
                    } else if (d == '{') {
                        ++depth;
                    } else if (d == '}') {
                        if (--depth == 0) {
                            ++j;
                            break;
                        }
                    }
                }
            }
            toks.push_back({ Tok::kHint, line.substr(i, j - i) });
            i = j;
            continue;
        }
        if (c == '[' || c == ']' || c == ':') {
            toks.push_back({ Tok::kPunct, std::string(1, c) });
            ++i;
            continue;
        }
        // number or identifier
        size_t j = i;
        while (j < n && !isspace((unsigned char)line[j]) && line[j] != '[' && line[j] != ']'
               && line[j] != '#' && line[j] != '"' && !(line[j] == ':' && j > i))
            ++j;
        std::string w = line.substr(i, j - i);
        bool num = (isdigit((unsigned char)w[0]) || ((w[0] == '-' || w[0] == '+' || w[0] == '.')
                                                     && w.size() > 1));
        if (num) {
            bool isf = w.find_first_of(".eE") != std::string::npos || w == "inf" || w == "nan";
            toks.push_back({ isf ? Tok::kFloat : Tok::kInt, w });
        } else {
            toks.push_back({ Tok::kIdent, w });
        }
        i = j;
    }
    return true;
}

bool symkind(const std::string& s, SymKind& k)
{
    if (s == "param") k = kParam;
    else if (s == "oparam") k = kOParam;
    else if (s == "global") k = kGlobal;
    else if (s == "local") k = kLocal;
    else if (s == "temp") k = kTemp;
    else if (s == "const") k = kConst;
    else return false;
    return true;
}

bool simple_type(const std::string& s, OType& t)
{
    if (s == "int") t.base = kInt;
    else if (s == "float") t.base = kFloat;
    else if (s == "color" || s == "point" || s == "vector" || s == "normal") {
        t.base        = kTriple;
        t.triple_kind = s;
    } else if (s == "string") t.base = kString;
    else if (s == "matrix") t.base = kMatrix;
    else if (s == "void") t.base = kVoid;
    else return false;
    return true;
}

}  // namespace

bool parse_oso(const std::string& text, ShaderMaster& out, std::string& err)
{
    std::istringstream in(text);
    std::string line;
    int lineno          = 0;
    bool have_version   = false;
    bool have_shader    = false;
    bool in_code        = false;
    int cur_init_sym    = -1;
    out                 = ShaderMaster();
    auto fail = [&](const std::string& m) {
        err = "oso line " + std::to_string(lineno) + ": " + m;
        return false;
    };
    while (std::getline(in, line)) {
        ++lineno;
        std::vector<Tok> t;
        if (!tokenize(line, t, err))
            return fail(err);
        if (t.empty())
            continue;
        if (!have_version) {
            if (t[0].s != "OpenShadingLanguage")
                return fail("missing OpenShadingLanguage version line");
            have_version = true;
            continue;
        }
        if (!have_shader) {
            if (t.size() < 2 || t[0].k != Tok::kIdent || t[1].k != Tok::kIdent)
                return fail("expected shader declaration");
            out.shadertype = t[0].s;
            out.name       = t[1].s;
            have_shader    = true;
            continue;
        }
        if (t[0].k == Tok::kIdent && t[0].s == "code") {
            if (t.size() < 2)
                return fail("code marker without name");
            if (cur_init_sym >= 0)
                out.syms[cur_init_sym].init_end = (int)out.ops.size();
            cur_init_sym = -1;
            in_code      = true;
            if (t[1].s == "___main___") {
                out.main_begin = (int)out.ops.size();
            } else {
                int s = out.find(t[1].s);
                if (s >= 0) {
                    out.syms[s].init_begin = (int)out.ops.size();
                    cur_init_sym           = s;
                }
            }
            continue;
        }
        SymKind sk;
        if (!in_code && t[0].k == Tok::kIdent && symkind(t[0].s, sk)) {
            size_t i = 1;
            OSymbol s;
            s.kind = sk;
            if (i < t.size() && t[i].s == "closure") {
                ++i;
                s.type.base = kClosure;
                if (i < t.size())
                    ++i;  // "color"
            } else if (i < t.size() && t[i].s == "struct") {
                s.type.base = kStruct;
                i += 2;
            } else {
                if (i >= t.size() || !simple_type(t[i].s, s.type))
                    return fail("bad type '" + (i < t.size() ? t[i].s : "") + "'");
                ++i;
            }
            if (i < t.size() && t[i].k == Tok::kPunct && t[i].s == "[") {
                if (i + 1 < t.size() && t[i + 1].k == Tok::kInt) {
                    s.type.arraylen = atoi(t[i + 1].s.c_str());
                    i += 3;
                } else {
                    s.type.arraylen = -1;
                    i += 2;
                }
            }
            if (i >= t.size() || t[i].k != Tok::kIdent)
                return fail("missing symbol name");
            s.name = t[i++].s;
            for (; i < t.size(); ++i) {
                if (t[i].k == Tok::kInt || t[i].k == Tok::kFloat)
                    s.fdefaults.push_back(strtod(t[i].s.c_str(), nullptr));
                else if (t[i].k == Tok::kString)
                    s.sdefaults.push_back(t[i].s);
                else if (t[i].k == Tok::kHint)
                    continue;
            }
            out.syms.push_back(s);
            continue;
        }
        if (!in_code)
            return fail("unexpected line before code section: " + t[0].s);
        // instruction: [label:] opname args... jumptargets... hints
        size_t i = 0;
        if (t.size() >= 2 && t[0].k == Tok::kInt && t[1].s == ":")
            i = 2;
        if (i >= t.size() || t[i].k != Tok::kIdent)
            return fail("expected opcode");
        OOp op;
        op.name = t[i++].s;
        op.line = lineno;
        int nj  = 0;
        for (; i < t.size(); ++i) {
            if (t[i].k == Tok::kHint)
                continue;
            if (t[i].k == Tok::kInt && nj < 4) {
                op.jump[nj++] = atoi(t[i].s.c_str());
                continue;
            }
            if (t[i].k == Tok::kIdent) {
                int s = out.find(t[i].s);
                if (s < 0)
                    return fail("unknown symbol '" + t[i].s + "' in op " + op.name);
                op.args.push_back(s);
                continue;
            }
            return fail("unexpected token '" + t[i].s + "' in op " + op.name);
        }
        out.ops.push_back(op);
    }
    if (cur_init_sym >= 0)
        out.syms[cur_init_sym].init_end = (int)out.ops.size();
    if (!have_shader)
        return fail("no shader declaration");
    return true;
}

}  // namespace oslmi
