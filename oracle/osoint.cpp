// CPU oracle: an OSO interpreter for shader groups, one shading point at a
// time (the reference's per-group, per-point execution, without
// derivatives). TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// Reads OSO text (the format oslc writes: src/liboslexec/osogram.y,
// osolex.l), assembles layers with parameter connections like
// ShaderGroupBegin/Shader/ConnectShaders (src/include/OSL/oslexec.h:634-742)
// and executes each layer's ops in order, with OSO's jump targets for
// if/for/while (src/liboslexec/batched_llvm_gen.cpp: llvm_gen_if :1068,
// llvm_gen_loop_op :3806). Op semantics follow the reference:
//   arithmetic with float -> triple promotion, int division by zero -> 0
//   (opmath, batched_llvm_gen.cpp), mix x*(1-a)+y*a (:1868-1960),
//   smoothstep (dual.h:1314), clamp = max(min(x,hi),lo) (stdosl.h:140),
//   noise names of llvm_gen_noise (:6096-6134) via the oracle's noise,
//   simplex and Gabor restatements, hash (wide_ophash.cpp), blackbody /
//   luminance / transformc (opcolor.cpp), transform by a matrix
//   (Imath multVecMatrix / multDirMatrix), closures diffuse and
//   microfacet ggx (testrender shading.cpp:185-296) flattened into a lobe
//   list and evaluated by the oracle's CompositeBSDF (closure.cpp).
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "closure.h"
#include "color.h"
#include "gabor.h"
#include "noise.h"
#include "omatrix.h"

namespace oref {
namespace oso {

enum Base { kInt, kFloat, kTriple, kMatrix, kString, kClosure };

struct Sym {
    std::string kind;  // param oparam local temp const global
    Base base = kFloat;
    std::string name;
    std::vector<float> fdef;
    std::vector<std::string> sdef;
};
struct Op {
    std::string name;
    std::vector<int> args;
    int jump[4] = { -1, -1, -1, -1 };
};
struct Master {
    std::string type, name;
    std::vector<Sym> syms;
    std::vector<Op> ops;
    int find(const std::string& n) const
    {
        for (size_t i = 0; i < syms.size(); ++i)
            if (syms[i].name == n)
                return (int)i;
        return -1;
    }
};

// ---- OSO reader ----------------------------------------------------------
struct Tok {
    enum K { Ident, Int, Float, Str, Hint, Punct } k;
    std::string s;
};
bool tokenize(const std::string& line, std::vector<Tok>& out)
{
    size_t i = 0, n = line.size();
    while (i < n) {
        char c = line[i];
        if (c == ' ' || c == '\t' || c == '\r') {
            ++i;
        } else if (c == '#') {
            break;
        } else if (c == '"') {
            size_t j = i + 1;
            std::string s;
            while (j < n && line[j] != '"') {
                if (line[j] == '\\' && j + 1 < n)
                    ++j;
                s += line[j++];
            }
            out.push_back({ Tok::Str, s });
            i = j + 1;
        } else if (c == '%') {
            // hint: %name or %name{...}
            size_t j = i + 1;
            int depth = 0;
            while (j < n && (depth > 0 || (line[j] != ' ' && line[j] != '\t'))) {
                if (line[j] == '{')
                    ++depth;
                else if (line[j] == '}')
                    --depth;
                else if (line[j] == '"') {
                    ++j;
                    while (j < n && line[j] != '"')
                        ++j;
                }
                ++j;
            }
            out.push_back({ Tok::Hint, line.substr(i, j - i) });
            i = j;
        } else if (isdigit((unsigned char)c) || ((c == '-' || c == '+' || c == '.') && i + 1 < n
                                                  && (isdigit((unsigned char)line[i + 1])
                                                      || line[i + 1] == '.'))) {
            size_t j = i + 1;
            bool isf = c == '.';
            while (j < n && (isalnum((unsigned char)line[j]) || line[j] == '.' || line[j] == '-'
                             || line[j] == '+')) {
                if (line[j] == '.' || line[j] == 'e' || line[j] == 'E')
                    isf = true;
                if ((line[j] == '-' || line[j] == '+') && !(line[j - 1] == 'e' || line[j - 1] == 'E'))
                    break;
                ++j;
            }
            out.push_back({ isf ? Tok::Float : Tok::Int, line.substr(i, j - i) });
            i = j;
        } else if (isalpha((unsigned char)c) || c == '_' || c == '$') {
            size_t j = i + 1;
            while (j < n && (isalnum((unsigned char)line[j]) || line[j] == '_' || line[j] == '$'
                             || line[j] == '.'))
                ++j;
            out.push_back({ Tok::Ident, line.substr(i, j - i) });
            i = j;
        } else {
            out.push_back({ Tok::Punct, std::string(1, c) });
            ++i;
        }
    }
    return true;
}

bool parse_type(const std::string& t, Base& b)
{
    if (t == "int") b = kInt;
    else if (t == "float") b = kFloat;
    else if (t == "color" || t == "point" || t == "vector" || t == "normal") b = kTriple;
    else if (t == "matrix") b = kMatrix;
    else if (t == "string") b = kString;
    else return false;
    return true;
}

bool parse(const std::string& text, Master& m, std::string& err)
{
    size_t pos = 0;
    int stage  = 0;  // 0 version, 1 shader line, 2 symbols, 3 code
    while (pos < text.size()) {
        size_t e = text.find('\n', pos);
        if (e == std::string::npos)
            e = text.size();
        std::string line = text.substr(pos, e - pos);
        pos              = e + 1;
        std::vector<Tok> t;
        tokenize(line, t);
        if (t.empty())
            continue;
        if (stage == 0) {
            stage = 1;
            continue;
        }
        if (stage == 1) {
            m.type = t[0].s;
            m.name = t.size() > 1 ? t[1].s : "";
            stage  = 2;
            continue;
        }
        if (t[0].k == Tok::Ident && t[0].s == "code") {
            stage = 3;
            continue;
        }
        if (stage == 2) {
            Sym s;
            s.kind   = t[0].s;
            size_t i = 1;
            if (t[i].s == "closure") {
                s.base = kClosure;
                i += 2;
            } else {
                if (!parse_type(t[i].s, s.base)) {
                    err = "bad type " + t[i].s;
                    return false;
                }
                ++i;
            }
            s.name = t[i++].s;
            for (; i < t.size(); ++i) {
                if (t[i].k == Tok::Int || t[i].k == Tok::Float)
                    s.fdef.push_back((float)strtod(t[i].s.c_str(), nullptr));
                else if (t[i].k == Tok::Str)
                    s.sdef.push_back(t[i].s);
            }
            m.syms.push_back(s);
            continue;
        }
        Op op;
        size_t i = 0;
        op.name  = t[i++].s;
        int nj   = 0;
        for (; i < t.size(); ++i) {
            if (t[i].k == Tok::Hint)
                continue;
            if (t[i].k == Tok::Int) {
                if (nj < 4)
                    op.jump[nj++] = atoi(t[i].s.c_str());
                continue;
            }
            int s = m.find(t[i].s);
            if (s < 0) {
                err = "unknown symbol " + t[i].s;
                return false;
            }
            op.args.push_back(s);
        }
        m.ops.push_back(op);
    }
    return true;
}

// ---- values ------------------------------------------------------------------
struct Val {
    int i      = 0;
    float f[16] = { 0 };  // float: f[0]; triple: f[0..2]; matrix: f[0..15]
    int cl = -1;          // closure: a node of the point's closure tree (-1: null)
};

// Closure values are trees like the reference's ClosureColor (ADD, MUL
// and component nodes; src/include/OSL/oslclosure.h:41-145), built by the
// closure ops with osl_add_closure_closure / osl_mul_closure_color's
// null rules (opclosure.cpp:19-80) and walked after the shader by
// testrender's process_bsdf_closure.
enum CType { kCAdd, kCMul, kCComp };
enum CComp { kCompClassic, kCompMx, kCompLayer };
struct CNode {
    int type = kCComp;
    int a = -1, b = -1;  // ADD operands; MUL operand (a); layer top (a), base (b)
    V3 w   = V3(1.0f);   // MUL weight
    int comp = kCompClassic;
    Lobe lobe {};        // classic lobe parameters
    mtx::MxParams mx {};  // MaterialX lobe parameters
};

struct Layer {
    const Master* m;
    std::string name;
};
struct Conn {
    int src_layer, src_sym, dst_layer, dst_sym;
};

// a renderer's named coordinate system (to-common matrix, optional inverse)
struct NamedSpace {
    Mat4 m, inv;
    bool has_inv = false;
};

struct Group {
    std::vector<std::unique_ptr<Master>> masters;
    std::vector<Layer> layers;
    std::vector<Conn> conns;
    ColorSys cs;
    std::string fout = "fout", cout = "cout";  // last-layer outputs read by run
    std::map<std::string, NamedSpace> spaces;
    std::string commonspace = "world";
    // renderer attributes: base type and value (an int as v[0]'s bits)
    struct Attr {
        Base base;
        float v[16];
    };
    std::map<std::string, Attr> attrs;
};

struct Exec {
    const Group& G;
    std::vector<std::vector<Val>> vals;  // [layer][sym]
    V3 P, N, dPdu, dPdv;
    bool has_ci = false;
    int ci = -1;
    std::vector<CNode> cn;  // the point's closure nodes

    explicit Exec(const Group& g) : G(g) {}

    float F(int L, int s) const
    {
        const Sym& y = G.layers[L].m->syms[s];
        return y.base == kInt ? (float)vals[L][s].i : vals[L][s].f[0];
    }
    int I(int L, int s) const
    {
        const Sym& y = G.layers[L].m->syms[s];
        return y.base == kInt ? vals[L][s].i : (int)vals[L][s].f[0];
    }
    V3 T(int L, int s) const
    {
        const Sym& y = G.layers[L].m->syms[s];
        if (y.base == kTriple)
            return V3(vals[L][s].f[0], vals[L][s].f[1], vals[L][s].f[2]);
        float v = F(L, s);
        return V3(v, v, v);
    }
    const std::string& Sv(int L, int s) const { return G.layers[L].m->syms[s].sdef.at(0); }

    // renderer services over the named spaces (wide_opmatrix.cpp:533-610,
    // opmatrix.cpp:232-262): unknown spaces give the identity and false
    bool is_common(const std::string& n) const { return n == "common" || n == G.commonspace; }
    bool space_matrix(const std::string& from, Mat4& M) const
    {
        M = mat4_identity();
        if (is_common(from))
            return true;
        auto it = G.spaces.find(from);
        if (it == G.spaces.end())
            return false;
        M = it->second.m;
        return true;
    }
    bool space_inverse(const std::string& to, Mat4& M) const
    {
        M = mat4_identity();
        if (is_common(to))
            return true;
        auto it = G.spaces.find(to);
        if (it == G.spaces.end())
            return false;
        M = it->second.has_inv ? it->second.inv : mat4_inverse(it->second.m);
        return true;
    }
    bool from_to(const std::string& from, const std::string& to, Mat4& M) const
    {
        if (is_common(from))
            return space_inverse(to, M);
        if (is_common(to))
            return space_matrix(from, M);
        Mat4 a, b;
        const bool ok1 = space_matrix(from, a);
        const bool ok2 = space_inverse(to, b);
        M              = mat4_mul(a, b);
        return ok1 && ok2;
    }
    std::string canon(const std::string& n) const { return n == G.commonspace ? "common" : n; }
    Base B(int L, int s) const { return G.layers[L].m->syms[s].base; }
    void setF(int L, int s, float v)
    {
        if (B(L, s) == kInt)
            vals[L][s].i = (int)v;
        else if (B(L, s) == kTriple)
            vals[L][s].f[0] = vals[L][s].f[1] = vals[L][s].f[2] = v;
        else
            vals[L][s].f[0] = v;
    }
    void setT(int L, int s, V3 v)
    {
        vals[L][s].f[0] = v.x;
        vals[L][s].f[1] = v.y;
        vals[L][s].f[2] = v.z;
    }

    void init_layer(int L)
    {
        const Master& m = *G.layers[L].m;
        vals[L].assign(m.syms.size(), Val());
        for (size_t s = 0; s < m.syms.size(); ++s) {
            const Sym& y = m.syms[s];
            Val& v       = vals[L][s];
            if (y.kind == "global") {
                if (y.name == "P")
                    setT(L, (int)s, P);
                else if (y.name == "N" || y.name == "Ng")
                    setT(L, (int)s, N);
                else if (y.name == "dPdu")
                    setT(L, (int)s, dPdu);
                else if (y.name == "dPdv")
                    setT(L, (int)s, dPdv);
                continue;
            }
            if (y.base == kInt && !y.fdef.empty())
                v.i = (int)y.fdef[0];
            else if (y.base == kTriple && y.fdef.size() == 1)
                v.f[0] = v.f[1] = v.f[2] = y.fdef[0];
            else
                for (size_t k = 0; k < y.fdef.size() && k < 16; ++k)
                    v.f[k] = y.fdef[k];
        }
        for (const Conn& c : G.conns)
            if (c.dst_layer == L) {
                const Val& src = vals[c.src_layer][c.src_sym];
                Base sb = B(c.src_layer, c.src_sym), db = B(L, c.dst_sym);
                if (sb == db)
                    vals[L][c.dst_sym] = src;
                else if (db == kTriple)
                    setT(L, c.dst_sym, T(c.src_layer, c.src_sym));
                else
                    setF(L, c.dst_sym, F(c.src_layer, c.src_sym));
            }
    }

    // noise/snoise/cellnoise/hashnoise and the periodic pnoise/psnoise,
    // with or without a name; the periodic forms take (S, Sper) or
    // (S, T, Sper, Tper) (batched_llvm_gen.cpp:6066-6081).
    void noise(int L, const Op& op)
    {
        const Master& m     = *G.layers[L].m;
        const bool periodic = op.name == "pnoise" || op.name == "psnoise";
        std::string name    = (op.name == "snoise" || op.name == "psnoise") ? "perlin"
                              : op.name == "cellnoise"                      ? "cell"
                              : op.name == "hashnoise"                      ? "hash"
                                                                            : "uperlin";
        size_t first = 1;
        if (op.args.size() > 1 && m.syms[op.args[1]].base == kString) {
            name  = Sv(L, op.args[1]);
            first = 2;
        }
        size_t nnum = 0;
        while (first + nnum < op.args.size() && B(L, op.args[first + nnum]) != kString)
            ++nnum;
        const size_t ncoord = periodic ? nnum / 2 : nnum;
        float c[4], per[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
        int dim = 0, nper = 0;
        for (size_t k = 0; k < nnum; ++k) {
            const int sym = op.args[first + k];
            float* dst    = k < ncoord ? c : per;
            int& cnt      = k < ncoord ? dim : nper;
            if (B(L, sym) == kTriple) {
                V3 v       = T(L, sym);
                dst[cnt++] = v.x;
                dst[cnt++] = v.y;
                dst[cnt++] = v.z;
            } else {
                dst[cnt++] = F(L, sym);
            }
        }
        size_t a      = first + nnum;
        const int r   = op.args[0];
        const bool tr = B(L, r) == kTriple;
        if (periodic && name != "gabor") {
            std::string pn = (name == "perlin" || name == "snoise" || name == "psnoise")  ? "psnoise"
                             : (name == "uperlin" || name == "noise" || name == "pnoise") ? "pnoise"
                             : (name == "cell" || name == "cellnoise")                    ? "cell"
                             : (name == "hash" || name == "hashnoise")                    ? "hash"
                                                                                          : "";
            if (pn.empty() || nper != dim) {
                fprintf(stderr, "osoint: periodic noise '%s'\n", name.c_str());
                abort();
            }
            if (pn == "psnoise") {
                if (tr) setT(L, r, psnoise_v(dim, c, per));
                else setF(L, r, psnoise_f(dim, c, per));
            } else if (pn == "pnoise") {
                if (tr) setT(L, r, pnoise_v(dim, c, per));
                else setF(L, r, pnoise_f(dim, c, per));
            } else if (pn == "cell") {
                if (tr) setT(L, r, pcellnoise_v(dim, c, per));
                else setF(L, r, pcellnoise_f(dim, c, per));
            } else {
                if (tr) setT(L, r, phashnoise_v(dim, c, per));
                else setF(L, r, phashnoise_f(dim, c, per));
            }
            return;
        }
        if (name == "perlin" || name == "snoise") {
            if (tr) setT(L, r, snoise_v(dim, c));
            else setF(L, r, snoise_f(dim, c));
        } else if (name == "uperlin" || name == "noise") {
            if (tr) setT(L, r, unoise_v(dim, c));
            else setF(L, r, unoise_f(dim, c));
        } else if (name == "cell" || name == "cellnoise") {
            if (tr) setT(L, r, cellnoise_v(dim, c));
            else setF(L, r, cellnoise_f(dim, c));
        } else if (name == "hash" || name == "hashnoise") {
            if (tr) setT(L, r, hashnoise_v(dim, c));
            else setF(L, r, hashnoise_f(dim, c));
        } else if (name == "simplex" || name == "usimplex") {
            float s[3];
            for (int k = 0; k < (tr ? 3 : 1); ++k) {
                s[k] = simplex_f(dim, c, k, nullptr);
                if (name == "usimplex")
                    s[k] = 0.5f * (s[k] + 1.0f);
            }
            if (tr) setT(L, r, V3(s[0], s[1], s[2]));
            else setF(L, r, s[0]);
        } else if (name == "gabor") {
            GaborOpts go;
            for (; a + 1 < op.args.size(); a += 2) {
                const std::string key = Sv(L, op.args[a]);
                const float v         = F(L, op.args[a + 1]);
                if (key == "anisotropic") go.anisotropic = (int)v;
                else if (key == "do_filter") go.do_filter = (int)v;
                else if (key == "bandwidth") go.bandwidth = v;
                else if (key == "impulses") go.impulses = v;
                else if (key == "direction") go.direction = T(L, op.args[a + 1]);
            }
            D2<float> p[3] = { D2<float>(c[0]), D2<float>(dim > 1 ? c[1] : float(0.0f)),
                               D2<float>(dim > 2 ? c[2] : float(0.0f)) };
            const int gd = dim < 3 ? dim : 3;  // a 4th (time) coordinate is ignored
            const float* pp = periodic ? per : nullptr;
            if (tr) setT(L, r, gabor3(gd, p, go, pp).v);
            else setF(L, r, gabor_ex(gd, p, go, 0, pp).v);
        } else {
            fprintf(stderr, "osoint: noise '%s'\n", name.c_str());
            abort();
        }
    }

    void op1(int L, const Op& op)
    {
        const std::string& o = op.name;
        auto a               = [&](int k) { return op.args[k]; };
        if (o == "nop" || o == "useparam" || o == "end")
            return;
        const int r  = a(0);
        const Base rb = B(L, r);
        if (o == "assign") {
            if (rb == kClosure) {
                vals[L][r].cl = vals[L][a(1)].cl;
            } else if (rb == kTriple) {
                setT(L, r, T(L, a(1)));
            } else if (rb == kInt) {
                vals[L][r].i = I(L, a(1));
            } else if (rb == kMatrix) {
                vals[L][r] = vals[L][a(1)];
            } else {
                setF(L, r, F(L, a(1)));
            }
            return;
        }
        if (o == "add" || o == "sub" || o == "mul" || o == "div") {
            if (rb == kClosure) {
                if (o == "add") {
                    // osl_add_closure_closure: a null operand gives the other
                    const int x = vals[L][a(1)].cl, y = vals[L][a(2)].cl;
                    if (x < 0 || y < 0) {
                        vals[L][r].cl = x < 0 ? y : x;
                    } else {
                        CNode nd;
                        nd.type = kCAdd;
                        nd.a    = x;
                        nd.b    = y;
                        cn.push_back(nd);
                        vals[L][r].cl = (int)cn.size() - 1;
                    }
                } else {
                    // osl_mul_closure_color / _float: null for a null closure
                    // or a zero weight, the closure itself for weight one
                    const int ci_ = B(L, a(1)) == kClosure ? a(1) : a(2);
                    const int wi  = B(L, a(1)) == kClosure ? a(2) : a(1);
                    const int x   = vals[L][ci_].cl;
                    const V3 w    = T(L, wi);
                    if (x < 0 || (w.x == 0.0f && w.y == 0.0f && w.z == 0.0f)) {
                        vals[L][r].cl = -1;
                    } else if (w.x == 1.0f && w.y == 1.0f && w.z == 1.0f) {
                        vals[L][r].cl = x;
                    } else {
                        CNode nd;
                        nd.type = kCMul;
                        nd.a    = x;
                        nd.w    = w;
                        cn.push_back(nd);
                        vals[L][r].cl = (int)cn.size() - 1;
                    }
                }
                return;
            }
            if (rb == kInt) {
                const int x = I(L, a(1)), y = I(L, a(2));
                int v = o == "add" ? x + y : o == "sub" ? x - y : o == "mul" ? x * y
                                                                 : (y == 0 ? 0 : x / y);
                vals[L][r].i = v;
                return;
            }
            if (rb == kTriple) {
                const V3 x = T(L, a(1)), y = T(L, a(2));
                V3 v = o == "add" ? x + y : o == "sub" ? x - y : o == "mul" ? x * y
                                                                : V3(x.x / y.x, x.y / y.y, x.z / y.z);
                setT(L, r, v);
                return;
            }
            const float x = F(L, a(1)), y = F(L, a(2));
            setF(L, r, o == "add" ? x + y : o == "sub" ? x - y : o == "mul" ? x * y : x / y);
            return;
        }
        if (o == "color" || o == "point" || o == "vector" || o == "normal") {
            // triple constructors; color(space, a, b, c) converts to rgb
            // (ColorSystem::to_rgb, opcolor.cpp:275-296)
            const int n = (int)op.args.size();
            if (n == 2) {
                setT(L, r, T(L, a(1)));
                return;
            }
            const int first = n == 5 ? 2 : 1;
            const V3 c(F(L, a(first)), F(L, a(first + 1)), F(L, a(first + 2)));
            if (n == 5 && o == "color") {
                const int fc = color_space_code(Sv(L, a(1)).c_str(), false);
                if (fc < 0) {
                    fprintf(stderr, "osoint: color space %s needs OpenColorIO\n", Sv(L, a(1)).c_str());
                    abort();
                }
                setT(L, r, transformc(G.cs, fc, 0, c));
                return;
            }
            if (n == 5) {
                fprintf(stderr, "osoint: %s(space, ...) is not supported\n", o.c_str());
                abort();
            }
            setT(L, r, c);
            return;
        }
        if (o == "mod") {
            // int: a % b; float: the exact remainder with a's sign (LLVM
            // frem); 0 when b is 0 (batched_llvm_gen.cpp:1683-1750)
            auto fm = [](float x, float y) { return y != 0.0f ? std::fmod(x, y) : 0.0f; };
            if (rb == kInt) {
                const int x = I(L, a(1)), y = I(L, a(2));
                vals[L][r].i = y != 0 ? x % y : 0;
            } else if (rb == kTriple) {
                const V3 x = T(L, a(1)), y = T(L, a(2));
                setT(L, r, V3(fm(x.x, y.x), fm(x.y, y.y), fm(x.z, y.z)));
            } else {
                setF(L, r, fm(F(L, a(1)), F(L, a(2))));
            }
            return;
        }
        if (o == "neg") {
            if (rb == kTriple) setT(L, r, -T(L, a(1)));
            else if (rb == kInt) vals[L][r].i = -I(L, a(1));
            else setF(L, r, -F(L, a(1)));
            return;
        }
        if (o == "mix") {
            if (rb == kTriple) {
                const V3 x = T(L, a(1)), y = T(L, a(2)), t = T(L, a(3));
                setT(L, r, V3(x.x * (1.0f - t.x) + y.x * t.x, x.y * (1.0f - t.y) + y.y * t.y,
                              x.z * (1.0f - t.z) + y.z * t.z));
            } else {
                const float x = F(L, a(1)), y = F(L, a(2)), t = F(L, a(3));
                setF(L, r, x * (1.0f - t) + y * t);
            }
            return;
        }
        if (o == "smoothstep") {
            if (rb == kTriple) {
                const V3 e0 = T(L, a(1)), e1 = T(L, a(2)), x = T(L, a(3));
                setT(L, r, V3(smoothstep_f(e0.x, e1.x, x.x), smoothstep_f(e0.y, e1.y, x.y),
                              smoothstep_f(e0.z, e1.z, x.z)));
            } else {
                setF(L, r, smoothstep_f(F(L, a(1)), F(L, a(2)), F(L, a(3))));
            }
            return;
        }
        if (o == "clamp") {
            if (rb == kTriple) {
                const V3 x = T(L, a(1)), lo = T(L, a(2)), hi = T(L, a(3));
                setT(L, r, V3(fmaxf(fminf(x.x, hi.x), lo.x), fmaxf(fminf(x.y, hi.y), lo.y),
                              fmaxf(fminf(x.z, hi.z), lo.z)));
            } else if (rb == kInt) {
                int x = I(L, a(1)), lo = I(L, a(2)), hi = I(L, a(3));
                x     = x < hi ? x : hi;
                vals[L][r].i = x > lo ? x : lo;
            } else {
                setF(L, r, fmaxf(fminf(F(L, a(1)), F(L, a(3))), F(L, a(2))));
            }
            return;
        }
        if (o == "floor" || o == "fabs" || o == "abs") {
            auto f = [&](float x) { return o == "floor" ? floorf(x) : fabsf(x); };
            if (rb == kTriple) {
                const V3 x = T(L, a(1));
                setT(L, r, V3(f(x.x), f(x.y), f(x.z)));
            } else {
                setF(L, r, f(F(L, a(1))));
            }
            return;
        }
        if (o == "compref") {
            const V3 v  = T(L, a(1));
            const int k = I(L, a(2));
            setF(L, r, k == 0 ? v.x : (k == 1 ? v.y : v.z));
            return;
        }
        if (o == "lt" || o == "gt" || o == "le" || o == "ge" || o == "eq" || o == "neq") {
            const float x = F(L, a(1)), y = F(L, a(2));
            int v = o == "lt" ? x < y : o == "gt" ? x > y : o == "le" ? x <= y : o == "ge" ? x >= y
                                                                                 : o == "eq" ? x == y
                                                                                             : x != y;
            if (rb == kInt) vals[L][r].i = v;
            else setF(L, r, (float)v);
            return;
        }
        if (o == "noise" || o == "snoise" || o == "cellnoise" || o == "hashnoise" || o == "pnoise"
            || o == "psnoise") {
            noise(L, op);
            return;
        }
        if (o == "hash") {
            float c[4];
            int dim = 0;
            for (size_t k = 1; k < op.args.size(); ++k) {
                if (B(L, op.args[k]) == kTriple) {
                    V3 v     = T(L, op.args[k]);
                    c[dim++] = v.x;
                    c[dim++] = v.y;
                    c[dim++] = v.z;
                } else if (B(L, op.args[k]) == kInt && op.args.size() == 2) {
                    vals[L][r].i = hash_op_i(I(L, op.args[k]));
                    return;
                } else {
                    c[dim++] = F(L, op.args[k]);
                }
            }
            vals[L][r].i = hash_op_f(dim, c);
            return;
        }
        if (o == "blackbody") {
            setT(L, r, blackbody_rgb(G.cs, F(L, a(1))));
            return;
        }
        if (o == "luminance") {
            setF(L, r, luminance(G.cs, T(L, a(1))));
            return;
        }
        if (o == "transformc") {
            std::string from = "rgb", to;
            int ci_ = 2;
            if (op.args.size() == 4) {
                from = Sv(L, a(1));
                to   = Sv(L, a(2));
                ci_  = 3;
            } else {
                to = Sv(L, a(1));
            }
            const int fc = color_space_code(from.c_str(), true);
            const int tc = color_space_code(to.c_str(), true);
            if (fc < 0 || tc < 0) {
                fprintf(stderr, "osoint: transformc %s -> %s needs OpenColorIO\n", from.c_str(),
                        to.c_str());
                abort();
            }
            setT(L, r, transformc(G.cs, fc, tc, T(L, a(ci_))));
            return;
        }
        if (o == "matrix") {
            if (op.args.size() == 17) {
                for (int k = 0; k < 16; ++k)
                    vals[L][r].f[k] = F(L, a(k + 1));
                return;
            }
            // matrix(from) = from -> common; matrix(from, to)
            const std::string from = canon(Sv(L, a(1)));
            const std::string to   = op.args.size() == 3 ? canon(Sv(L, a(2))) : "common";
            Mat4 M = mat4_identity();
            if (from != to)
                from_to(from, to, M);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    vals[L][r].f[4 * i + j] = M.x[i][j];
            return;
        }
        if (o == "getattribute") {
            // RendererServices::get_attribute: found with the value's type
            auto it = G.attrs.find(Sv(L, a(1)));
            const Base want = B(L, a(2));
            if (it == G.attrs.end() || it->second.base != want) {
                vals[L][r].i = 0;
                return;
            }
            Val& d = vals[L][a(2)];
            if (want == kInt)
                std::memcpy(&d.i, &it->second.v[0], 4);
            else
                for (int k = 0; k < (want == kMatrix ? 16 : (want == kTriple ? 3 : 1)); ++k)
                    d.f[k] = it->second.v[k];
            vals[L][r].i = 1;
            return;
        }
        if (o == "getmatrix") {
            const std::string from = canon(Sv(L, a(1))), to = canon(Sv(L, a(2)));
            Mat4 M   = mat4_identity();
            // Mfrom * Mto^-1 for any pair, one space named twice included
            // (wide_opmatrix.cpp:869-885)
            bool ok  = from_to(from, to, M);
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    vals[L][a(3)].f[4 * i + j] = M.x[i][j];
            vals[L][r].i = ok ? 1 : 0;
            return;
        }
        if (o == "transform" || o == "transformv" || o == "transformn") {
            // opcode = semantics (batched_llvm_gen.cpp:3485-3493); transform
            // [from] to p by named spaces, or by a matrix
            const int n  = (int)op.args.size();
            const V3 p   = T(L, a(n == 3 ? 2 : 3));
            Mat4 M;
            if (n == 3 && B(L, a(1)) == kMatrix) {
                M = mat4_from(vals[L][a(1)].f);
            } else {
                const std::string from = n == 4 ? canon(Sv(L, a(1))) : "common";
                const std::string to   = canon(Sv(L, a(n == 3 ? 1 : 2)));
                if (from == to || !from_to(from, to, M)) {  // identity, or unknown space
                    setT(L, r, p);
                    return;
                }
            }
            setT(L, r, o == "transform" ? mat4_point(M, p)
                                        : (o == "transformv" ? mat4_dir(M, p) : mat4_normal(M, p)));
            return;
        }
        if (o == "closure") {
            const std::string cname = Sv(L, a(1));
            CNode nd;
            nd.type = kCComp;
            if (cname == "layer") {
                // MX_LAYER_ID with its two closure parameters
                nd.comp = kCompLayer;
                nd.a    = vals[L][a(2)].cl;
                nd.b    = vals[L][a(3)].cl;
            } else if (cname == "dielectric_bsdf" || cname == "generalized_schlick_bsdf"
                       || cname == "sheen_bsdf") {
                // MaterialX closures, positional parameters as registered
                // (MTX/bsdf_*_decl.h entry()); keyword ones change nothing
                // the interpreter supports
                nd.comp  = kCompMx;
                mtx::MxParams& m = nd.mx;
                m.N              = T(L, a(2));
                if (cname == "sheen_bsdf") {
                    m.kind = 2;
                    m.tint = T(L, a(3));
                    m.rx = m.ry = F(L, a(4));
                } else {
                    m.kind = cname == "dielectric_bsdf" ? 0 : 1;
                    m.U    = T(L, a(3));
                    m.tint = T(L, a(4));
                    m.rx   = F(L, a(6));
                    m.ry   = F(L, a(7));
                    if (m.kind == 0) {
                        m.ior = F(L, a(8));
                    } else {
                        m.F0       = T(L, a(8));
                        m.F90      = T(L, a(9));
                        m.exponent = F(L, a(10));
                    }
                }
            } else {
                Lobe& l  = nd.lobe;
                l.weight = V3(1.0f);
                if (cname == "diffuse" || cname == "translucent") {
                    // translucent is testrender's Diffuse<1>: the lobe about -N
                    l.kind  = cname == "diffuse" ? kLobeDiffuse : kLobeTranslucent;
                    l.N     = cname == "diffuse" ? T(L, a(2)) : -T(L, a(2));
                    l.alpha = 0.0f;
                    l.eta   = 0.0f;
                } else if (cname == "emission") {
                    l.kind  = kLobeEmission;
                    l.N     = V3(0.0f);
                    l.alpha = 0.0f;
                    l.eta   = 0.0f;
                } else if (cname == "reflection" || cname == "refraction"
                           || cname == "transparent") {
                    const bool with_eta = op.args.size() >= 4;
                    l.kind  = cname == "transparent" ? kLobeTransparent
                              : cname == "refraction" ? kLobeRefraction
                              : with_eta              ? kLobeFresnelReflection
                                                      : kLobeReflection;
                    l.N     = cname == "transparent" ? V3(0.0f) : T(L, a(2));
                    l.alpha = 0.0f;
                    l.eta   = with_eta && cname != "transparent" ? F(L, a(3)) : float(0.0f);
                } else if (cname == "phong") {
                    l.kind  = kLobePhong;
                    l.N     = T(L, a(2));
                    l.alpha = F(L, a(3));
                    l.eta   = 0.0f;
                } else {
                    // microfacet(dist, N, U, xalpha, yalpha, eta, refract)
                    l.kind  = Sv(L, a(2)) == "beckmann" ? kLobeBeckmannRefl : kLobeGGXRefl;
                    l.N     = T(L, a(3));
                    l.alpha = F(L, a(5));
                    l.eta   = F(L, a(7));
                    l.U     = T(L, a(4));
                    l.ay    = F(L, a(6));
                    l.aniso = true;  // the frame of (N, U) is from_normal(N) when U = 0
                    l.refract = I(L, a(8));
                }
            }
            cn.push_back(nd);
            vals[L][r].cl = (int)cn.size() - 1;
            return;
        }
        fprintf(stderr, "osoint: unsupported op %s\n", o.c_str());
        abort();
    }

    // How a block ended: ran to its end, or left by break / continue (to
    // the innermost loop), return (to the end of the innermost
    // functioncall, or of the layer at its top level) or exit (the layer).
    enum Flow { kNext, kBreak, kContinue, kReturn, kExit };

    // executes ops [b, e) of layer L with structured control flow, as the
    // reference lowers it (llvm_gen.cpp: loops :2495-2535, break/continue
    // :2539-2551, functioncall :4327-4360, return/exit :4396-4411)
    Flow block(int L, int b, int e)
    {
        const Master& m = *G.layers[L].m;
        for (int i = b; i < e;) {
            const Op& op = m.ops[i];
            if (op.name == "if") {
                const Flow f = I(L, op.args[0]) ? block(L, i + 1, op.jump[0])
                                                : block(L, op.jump[0], op.jump[1]);
                if (f != kNext)
                    return f;
                i = op.jump[1];
                continue;
            }
            if (op.name == "for" || op.name == "while" || op.name == "dowhile") {
                // [init i+1, j0) [cond j0, j1) [body j1, j2) [step j2, j3)
                Flow f = block(L, i + 1, op.jump[0]);
                if (f != kNext)
                    return f;
                bool first = op.name == "dowhile";  // do-while enters at the body
                for (int guard = 0; guard < 1000000; ++guard) {
                    if (!first) {
                        f = block(L, op.jump[0], op.jump[1]);
                        if (f != kNext)
                            return f;
                        if (!I(L, op.args[0]))
                            break;
                    }
                    first = false;
                    f     = block(L, op.jump[1], op.jump[2]);
                    if (f == kBreak)
                        break;
                    if (f == kReturn || f == kExit)
                        return f;
                    f = block(L, op.jump[2], op.jump[3]);  // step (also after continue)
                    if (f != kNext)
                        return f;
                }
                i = op.jump[3];
                continue;
            }
            if (op.name == "break")
                return kBreak;
            if (op.name == "continue")
                return kContinue;
            if (op.name == "return")
                return kReturn;
            if (op.name == "exit")
                return kExit;
            if (op.name == "functioncall") {
                // the function body [i+1, j0) inline; its return ends it
                const Flow f = block(L, i + 1, op.jump[0]);
                if (f == kExit)
                    return f;
                i = op.jump[0];
                continue;
            }
            op1(L, op);
            ++i;
        }
        return kNext;
    }

    // OIIO clamp on each channel
    static float oclamp01(float a)
    {
        float v = a;
        if (!(0.0f <= v))
            v = 0.0f;
        if (!(v <= 1.0f))
            v = 1.0f;
        return v;
    }

    // evaluate_layer_opacity (src/testrender/shading.cpp:1197-1272): the
    // walk's weight when it stops (its stack pops a null pointer or runs
    // empty); MaterialX lobes multiply 1 - filter_o at wo, other
    // components count as opaque
    V3 layer_opacity(int closure, V3 wo) const
    {
        if (closure < 0)
            return V3(0.0f);
        int stack[16];
        V3 wstack[16];
        int sp    = 0;
        V3 weight = V3(1.0f);
        while (closure >= 0) {
            const CNode& nd = cn[closure];
            if (nd.type == kCMul) {
                weight  = weight * nd.w;
                closure = nd.a;
            } else if (nd.type == kCAdd) {
                stack[sp]    = nd.b;
                wstack[sp++] = weight;
                closure      = nd.a;
            } else {
                const V3 w(1.0f);
                if (nd.comp == kCompLayer) {
                    closure      = nd.a;
                    stack[sp]    = nd.b;
                    wstack[sp++] = weight * w;
                } else if (nd.comp == kCompMx) {
                    mtx::MxLobe lobe;
                    mtx::mx_build(&lobe, nd.mx, wo);
                    weight  = weight * (w * (V3(1.0f) - V3(mtx::mx_filter_o(lobe))));
                    closure = -1;
                } else {
                    closure = -1;
                }
            }
            if (closure < 0 && sp > 0) {
                closure = stack[--sp];
                weight  = wstack[sp];
            }
        }
        return weight;
    }

    // process_bsdf_closure (src/testrender/shading.cpp:1448-1667) over Ci at
    // wo, path weight 1: the flattened lobe list, not yet prepared
    Composite process(V3 wo) const
    {
        Composite c;
        if (!has_ci)
            return c;
        int closure = ci;
        int stack[16];
        V3 wstack[16];
        int sp    = 0;
        V3 weight = V3(1.0f);
        while (closure >= 0) {
            const CNode& nd = cn[closure];
            if (nd.type == kCMul) {
                weight  = weight * nd.w;
                closure = nd.a;
            } else if (nd.type == kCAdd) {
                stack[sp]    = nd.b;
                wstack[sp++] = weight;
                closure      = nd.a;
            } else {
                const V3 cw = weight * V3(1.0f);
                closure     = -1;
                if (nd.comp == kCompLayer) {
                    const V3 op     = layer_opacity(nd.a, wo);
                    const V3 base_w = weight * (V3(1.0f) - V3(oclamp01(op.x), oclamp01(op.y), oclamp01(op.z)));
                    closure         = nd.a;
                    weight          = cw;
                    if (!(base_w.x == 0 && base_w.y == 0 && base_w.z == 0)) {
                        stack[sp]    = nd.b;
                        wstack[sp++] = base_w;
                    }
                } else if (nd.comp == kCompMx) {
                    Lobe l {};
                    l.kind   = kLobeMx;
                    l.weight = cw;
                    l.N      = nd.mx.N;
                    l.mxp    = nd.mx;
                    mtx::mx_build(&l.mx, nd.mx, wo);
                    c.add(l);
                } else {
                    Lobe l   = nd.lobe;
                    l.weight = cw;
                    c.add(l);
                }
            }
            if (closure < 0 && sp > 0) {
                closure = stack[--sp];
                weight  = wstack[sp];
            }
        }
        return c;
    }

    void run()
    {
        vals.assign(G.layers.size(), {});
        cn.clear();
        has_ci = false;
        ci     = -1;
        // Ci is one global of the shading point, shared by the layers: each
        // layer starts from the value the layers before it left
        auto ci_sym = [&](int L) {
            const Master& m = *G.layers[L].m;
            for (size_t k = 0; k < m.syms.size(); ++k)
                if (m.syms[k].name == "Ci" && m.syms[k].kind == "global"
                    && m.syms[k].base == kClosure)
                    return (int)k;
            return -1;
        };
        for (size_t L = 0; L < G.layers.size(); ++L) {
            init_layer((int)L);
            const int k = ci_sym((int)L);
            if (k >= 0)
                vals[L][k].cl = ci;
            block((int)L, 0, (int)G.layers[L].m->ops.size());
            if (k >= 0) {
                has_ci = true;
                ci     = vals[L][k].cl;
            }
        }
    }
};

}  // namespace oso
}  // namespace oref

using namespace oref;
using namespace oref::oso;

extern "C" {

void* oref_group_new(void)
{
    Group* g = new Group();
    colorsys_init(g->cs, "Rec709");
    return g;
}
void oref_group_free(void* h) { delete (Group*)h; }

// A named coordinate system: to-common matrix m (16 floats, row-vector
// layout) and optional inverse (NULL: inverted like the default
// get_inverse_matrix).
// A renderer attribute: base 0 int, 1 float, 2 triple, 3 matrix (this
// interpreter's Base order); value 1, 1, 3 or 16 values.
void oref_group_set_attribute(void* h, const char* name, int base, const void* value)
{
    Group::Attr a;
    a.base = (Base)base;
    std::memset(a.v, 0, sizeof(a.v));
    const int count = base == kMatrix ? 16 : (base == kTriple ? 3 : 1);
    std::memcpy(a.v, value, 4 * count);
    ((Group*)h)->attrs[name] = a;
}

void oref_group_name_transform(void* h, const char* name, const float* m, const float* inv)
{
    NamedSpace& sp = ((Group*)h)->spaces[name];
    sp.m           = mat4_from(m);
    sp.has_inv     = inv != nullptr;
    if (inv)
        sp.inv = mat4_from(inv);
}

// Adds a layer from OSO text; returns 0 or -1 (error text to stderr).
int oref_group_layer(void* h, const char* oso, const char* layername)
{
    Group* g = (Group*)h;
    std::unique_ptr<Master> m(new Master());
    std::string err;
    if (!parse(oso, *m, err)) {
        fprintf(stderr, "osoint: %s\n", err.c_str());
        return -1;
    }
    g->layers.push_back({ m.get(), layername });
    g->masters.push_back(std::move(m));
    return 0;
}

int oref_group_connect(void* h, const char* srclayer, const char* srcparam, const char* dstlayer,
                       const char* dstparam)
{
    Group* g = (Group*)h;
    Conn c { -1, -1, -1, -1 };
    for (size_t L = 0; L < g->layers.size(); ++L) {
        if (g->layers[L].name == srclayer) {
            c.src_layer = (int)L;
            c.src_sym   = g->layers[L].m->find(srcparam);
        }
        if (g->layers[L].name == dstlayer) {
            c.dst_layer = (int)L;
            c.dst_sym   = g->layers[L].m->find(dstparam);
        }
    }
    if (c.src_layer < 0 || c.dst_layer < 0 || c.src_sym < 0 || c.dst_sym < 0)
        return -1;
    g->conns.push_back(c);
    return 0;
}

// Names the last layer's float and triple outputs that run() reports
// (default "fout" and "cout"; a name the layer lacks reads as zero).
void oref_group_outputs(void* h, const char* fname, const char* cname)
{
    Group* g = (Group*)h;
    g->fout  = fname ? fname : "";
    g->cout  = cname ? cname : "";
}

namespace {
void run_group(Group* g, const float* P, const float* N, const float* dPdu, const float* dPdv,
               const float* wo, const float* wi, const float* rnd, int64_t n, float* out4,
               float* eval4, float* samp7, int nthreads)
{
    const int last = (int)g->layers.size() - 1;
    const int sf   = g->layers[last].m->find(g->fout);
    const int sc   = g->layers[last].m->find(g->cout);
    auto work = [&](int64_t b, int64_t e) {
        Exec ex(*g);
        for (int64_t i = b; i < e; ++i) {
            ex.P    = V3(P[i], P[n + i], P[2 * n + i]);
            ex.N    = V3(N[i], N[n + i], N[2 * n + i]);
            ex.dPdu = dPdu ? V3(dPdu[i], dPdu[n + i], dPdu[2 * n + i]) : V3(0.0f);
            ex.dPdv = dPdv ? V3(dPdv[i], dPdv[n + i], dPdv[2 * n + i]) : V3(0.0f);
            ex.run();
            out4[i]         = sf >= 0 ? ex.F(last, sf) : float(0.0f);
            const V3 c      = sc >= 0 ? ex.T(last, sc) : V3(0.0f);
            out4[n + i]     = c.x;
            out4[2 * n + i] = c.y;
            out4[3 * n + i] = c.z;
            if (eval4 || samp7) {
                const V3 o(wo[i], wo[n + i], wo[2 * n + i]);
                Composite ci = ex.process(o);
                ci.prepare(o);
                if (eval4) {
                    const V3 w(wi[i], wi[n + i], wi[2 * n + i]);
                    BsdfSample s     = ci.eval(o, w);
                    eval4[i]         = s.weight.x;
                    eval4[n + i]     = s.weight.y;
                    eval4[2 * n + i] = s.weight.z;
                    eval4[3 * n + i] = s.pdf;
                }
                if (samp7) {
                    const BsdfDirSample s = ci.sample(o, rnd[i], rnd[n + i], rnd[2 * n + i]);
                    const float v[7] = { s.wi.x,     s.wi.y,     s.wi.z, s.weight.x,
                                         s.weight.y, s.weight.z, s.pdf };
                    for (int k = 0; k < 7; ++k)
                        samp7[k * n + i] = v[k];
                }
            }
        }
    };
    if (nthreads <= 1 || n < 256) {
        work(0, n);
        return;
    }
    std::atomic<int64_t> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&]() {
            for (;;) {
                int64_t b = next.fetch_add(256);
                if (b >= n)
                    break;
                work(b, b + 256 < n ? b + 256 : n);
            }
        });
    for (auto& t : th)
        t.join();
}
}  // namespace

// Runs the group over n points. Globals P, N (3 planes each); outputs of
// the last layer: fout (1 plane) and cout (3 planes) into out4; with wo/wi
// non-null, the closure Ci evaluated at (wo, wi) into eval4 (R, G, B, pdf).
void oref_group_run(void* h, const float* P, const float* N, const float* wo, const float* wi,
                    int64_t n, float* out4, float* eval4, int nthreads)
{
    run_group((Group*)h, P, N, nullptr, nullptr, wo, wi, nullptr, n, out4, eval4, nullptr, nthreads);
}

// Ci's lobes as process_bsdf_closure hands them to CompositeBSDF::add_bsdf
// (MaterialX lobes built at wo), in the closure-export layout of the C ABI
// (oslmi_execute_batch_closures): out plane 0 the count (at most maxl),
// then per lobe k planes 1 + 25 k + j: testrender's closure id, weight rgb,
// 21 parameter floats in testrender's registration order, zero-padded.
void oref_group_closures(void* h, const float* P, const float* N, const float* dPdu,
                         const float* dPdv, const float* wo, int64_t n, int maxl, float* out)
{
    Group* g = (Group*)h;
    Exec ex(*g);
    const int R = 25;
    for (int64_t i = 0; i < n; ++i) {
        ex.P    = V3(P[i], P[n + i], P[2 * n + i]);
        ex.N    = V3(N[i], N[n + i], N[2 * n + i]);
        ex.dPdu = dPdu ? V3(dPdu[i], dPdu[n + i], dPdu[2 * n + i]) : V3(0.0f);
        ex.dPdv = dPdv ? V3(dPdv[i], dPdv[n + i], dPdv[2 * n + i]) : V3(0.0f);
        ex.run();
        const Composite c = ex.process(V3(wo[i], wo[n + i], wo[2 * n + i]));
        const int cnt     = c.n < maxl ? c.n : maxl;
        out[i]            = (float)cnt;
        for (int k = 0; k < cnt; ++k) {
            const Lobe& l = c.lobes[k];
            float v[R]    = { 0 };
            v[1]          = l.weight.x;
            v[2]          = l.weight.y;
            v[3]          = l.weight.z;
            float* p      = v + 4;
            auto put3     = [&](V3 a) {
                *p++ = a.x;
                *p++ = a.y;
                *p++ = a.z;
            };
            if (l.kind == kLobeDiffuse || l.kind == kLobeTranslucent) {
                v[0] = l.kind == kLobeDiffuse ? 3 : 5;  // DIFFUSE_ID, TRANSLUCENT_ID
                put3(l.kind == kLobeDiffuse ? l.N : -l.N);  // the closure's N
            } else if (l.kind == kLobeEmission) {
                v[0] = 1;  // EMISSION_ID: no parameters
            } else if (l.kind >= kLobeReflection && l.kind <= kLobeTransparent) {
                // REFLECTION_ID (N), FRESNEL_REFLECTION_ID and REFRACTION_ID
                // (N, eta), TRANSPARENT_ID (none)
                v[0] = 9 + (l.kind - kLobeReflection);
                if (l.kind != kLobeTransparent)
                    put3(l.N);
                if (l.kind == kLobeFresnelReflection || l.kind == kLobeRefraction)
                    *p++ = l.eta;
            } else if (l.kind == kLobePhong) {
                v[0] = 6;  // PHONG_ID: N, exponent
                put3(l.N);
                *p++ = l.alpha;
            } else if (l.kind == kLobeGGXRefl || l.kind == kLobeBeckmannRefl) {
                v[0] = 8;  // MICROFACET_ID: dist, N, U, xalpha, yalpha, eta, refract
                *p++ = l.kind == kLobeBeckmannRefl ? 1.0f : 0.0f;
                put3(l.N);
                put3(l.aniso ? l.U : V3(0.0f));
                *p++ = l.alpha;
                *p++ = l.aniso ? l.ay : l.alpha;
                *p++ = l.eta;
                *p++ = (float)l.refract;
            } else {
                const mtx::MxParams& m = l.mxp;
                if (m.kind == 2) {
                    v[0] = 23;  // MX_SHEEN_ID: N, albedo, roughness, mode
                    put3(m.N);
                    put3(m.tint);
                    *p++ = m.rx;
                    *p++ = 0.0f;
                } else {
                    v[0] = m.kind == 0 ? 17 : 19;  // MX_DIELECTRIC_ID, MX_GENERALIZED_SCHLICK_ID
                    put3(m.N);
                    put3(m.U);
                    put3(m.tint);
                    put3(V3(0.0f));
                    *p++ = m.rx;
                    *p++ = m.ry;
                    if (m.kind == 0) {
                        *p++ = m.ior;
                    } else {
                        put3(m.F0);
                        put3(m.F90);
                        *p++ = m.exponent;
                    }
                }
            }
            for (int j = 0; j < R; ++j)
                out[(int64_t)(1 + k * R + j) * n + i] = v[j];
        }
    }
}

// Both, with the globals dPdu and dPdv (3 planes each, or null: zero) and
// either of eval4 / samp7 null to skip it.
void oref_group_run_ex(void* h, const float* P, const float* N, const float* dPdu,
                       const float* dPdv, const float* wo, const float* wi, const float* rnd,
                       int64_t n, float* out4, float* eval4, float* samp7, int nthreads)
{
    run_group((Group*)h, P, N, dPdu, dPdv, wo, wi, rnd, n, out4, eval4, samp7, nthreads);
}

// The same, with Ci sampled (CompositeBSDF::sample) at wo with the random
// numbers rnd (3 planes: rx, ry, rz) into samp7 (wi xyz, weight rgb, pdf).
void oref_group_run_sample(void* h, const float* P, const float* N, const float* wo,
                           const float* rnd, int64_t n, float* out4, float* samp7, int nthreads)
{
    run_group((Group*)h, P, N, nullptr, nullptr, wo, nullptr, rnd, n, out4, nullptr, samp7, nthreads);
}

}  // extern "C"
