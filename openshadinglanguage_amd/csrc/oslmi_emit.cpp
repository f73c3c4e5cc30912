// Group emitter: lowers a shader group (layers of OSO ops + connections) to
// the HIP source of one fused kernel that shades one point per work-item.
//
// It plays the role of the reference's BatchedBackendLLVM
// (src/liboslexec/batched_llvm_instance.cpp:1863-2100, per-op generators in
// batched_llvm_gen.cpp) but targets HIP C++ over the CDNA4 shadeop library:
//   * every layer is inlined in dependency order; layer outputs stay in
//     registers (no GroupData heap round trip);
//   * control flow (if/for/while/dowhile/break/continue, batched_llvm_gen.cpp
//     :1068, :3806) becomes structured C++; the wavefront's exec mask
//     handles divergent lanes;
//   * a derivative analysis (the deriv part of the runtime optimizer) marks
//     the symbols that must carry Dual2 values: arguments of ops that read
//     derivatives (filterwidth, Dx/Dy, Gabor noise) and, transitively, the
//     inputs of differentiable ops producing them, across connections;
//   * string arguments (noise names, closure names, color spaces, Gabor
//     options) are compile-time constants, as in the reference's
//     llvm_gen_noise name mapping (batched_llvm_gen.cpp:6096-6134).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <set>
#include <sstream>

#include "oslmi_colorsys.h"
#include "oslmi_group.h"

namespace oslmi {

int Group::find_layer(const std::string& n) const
{
    for (size_t i = 0; i < layers.size(); ++i)
        if (layers[i].name == n)
            return (int)i;
    return -1;
}

namespace {

// exact float literal
std::string flit(float f)
{
    if (std::isnan(f))
        return "__builtin_nanf(\"\")";
    if (std::isinf(f))
        return f > 0 ? "__builtin_inff()" : "(-__builtin_inff())";
    char buf[64];
    snprintf(buf, sizeof(buf), "%af", (double)f);
    return buf;
}

// launch bounds of every emitted kernel (256 work-items per workgroup) and,
// when the shading system sets one, the occupancy target the register
// allocator works to
std::string kernel_attributes(const ShadingSys& ss)
{
    std::string a = "__launch_bounds__(256)";
    if (ss.waves_per_eu > 0)
        a += " __attribute__((amdgpu_waves_per_eu(" + std::to_string(ss.waves_per_eu) + ")))";
    return a;
}

enum Kind { K_INT, K_FLOAT, K_DF, K_V3, K_DV, K_MAT, K_CLOSURE, K_STRING, K_VOID };

struct Emitter {
    const ShadingSys& ss;
    Group& g;
    std::string err;
    std::vector<std::vector<char>> deriv;  // [layer][sym]
    std::ostringstream os;
    int label_counter = 0;
    std::string lprefix = "L";  // unique per group inside a batch kernel
    int cur_layer     = 0;
    std::set<int> globals;
    bool uses_cs = false, uses_bb = false;
    bool uses_mx = false;  // MaterialX closures (libbsdl lobes, energy tables)
    bool uses_aniso = false;  // microfacet lobes with a tangent, two roughnesses or refraction

    Emitter(const ShadingSys& s, Group& gr) : ss(s), g(gr) {}

    std::string lab(int n) const { return lprefix + std::to_string(n); }

    // the transform-table slot of a (from, to) space pair, as an expression
    std::string xform_ref(const std::string& from, const std::string& to)
    {
        const int k = xform_slot(ss, from, to);
        if (k < 0) {
            fail("too many distinct coordinate-system pairs (at most " +
                 std::to_string(kMaxXformSlots) + ")");
            return "";
        }
        g.uses_xforms = true;
        return "A.xf + " + std::to_string(kXformSlotFloats * k);
    }
    const ShaderMaster& M(int L) const { return *g.layers[L].master; }
    const OSymbol& S(int L, int s) const { return M(L).syms[s]; }

    Kind kind(int L, int s) const
    {
        const OSymbol& y = S(L, s);
        switch (y.type.base) {
        case kInt: return K_INT;
        case kFloat: return deriv[L][s] ? K_DF : K_FLOAT;
        case kTriple: return deriv[L][s] ? K_DV : K_V3;
        case kMatrix: return K_MAT;
        case kClosure: return K_CLOSURE;
        case kString: return K_STRING;
        default: return K_VOID;
        }
    }
    static const char* ctype(Kind k)
    {
        switch (k) {
        case K_INT: return "int";
        case K_FLOAT: return "float";
        case K_DF: return "Df";
        case K_V3: return "V3";
        case K_DV: return "Dv";
        case K_MAT: return "M44";
        case K_CLOSURE: return "Closure";
        default: return "int";
        }
    }
    std::string var(int L, int s) const { return "v" + std::to_string(L) + "_" + std::to_string(s); }

    // compile-time string value of a symbol
    bool strval(int L, int s, std::string& out) const
    {
        const OSymbol& y = S(L, s);
        if (y.type.base != kString)
            return false;
        auto it = g.layers[L].inst.find(y.name);
        if (it != g.layers[L].inst.end() && !it->second.s.empty()) {
            out = it->second.s[0];
            return true;
        }
        if (!y.sdefaults.empty()) {
            out = y.sdefaults[0];
            return true;
        }
        out = "";
        return true;
    }
    // compile-time numeric value of a const/param symbol (for options)
    bool numval(int L, int s, std::vector<float>& out) const
    {
        const OSymbol& y = S(L, s);
        auto it          = g.layers[L].inst.find(y.name);
        if (it != g.layers[L].inst.end()) {
            out = it->second.f;
            if (out.empty())
                for (int v : it->second.i)
                    out.push_back((float)v);
            return true;
        }
        if (y.kind == kConst || y.kind == kParam) {
            out.clear();
            for (double d : y.fdefaults)
                out.push_back((float)d);
            return true;
        }
        return false;
    }

    // expression of symbol s converted to kind want
    std::string ex(int L, int s, Kind want)
    {
        Kind k       = kind(L, s);
        std::string v = var(L, s);
        if (k == want)
            return v;
        switch (want) {
        case K_INT:
            if (k == K_FLOAT) return "(int)" + v;
            break;
        case K_FLOAT:
            if (k == K_INT) return "(float)" + v;
            if (k == K_DF) return v + ".v";
            break;
        case K_DF:
            if (k == K_INT) return "Df((float)" + v + ")";
            if (k == K_FLOAT) return "Df(" + v + ")";
            break;
        case K_V3:
            if (k == K_INT) return "tri((float)" + v + ")";
            if (k == K_FLOAT) return "tri(" + v + ")";
            if (k == K_DF) return "tri(" + v + ".v)";
            if (k == K_DV) return v + ".v";
            break;
        case K_DV:
            if (k == K_INT) return "Dv(tri((float)" + v + "))";
            if (k == K_FLOAT) return "Dv(tri(" + v + "))";
            if (k == K_DF) return "tri(" + v + ")";
            if (k == K_V3) return "Dv(" + v + ")";
            break;
        default: break;
        }
        fail("cannot convert " + S(L, s).name + " from kind " + std::to_string(k) + " to "
             + std::to_string(want));
        return v;
    }

    bool failed = false;
    void fail(const std::string& m)
    {
        if (!failed) {
            err    = m;
            failed = true;
        }
    }

    // ----------------------------------------------------------- analysis
    static bool is_noise_op(const std::string& o)
    {
        return o == "noise" || o == "snoise" || o == "pnoise" || o == "psnoise" || o == "cellnoise"
               || o == "hashnoise";
    }
    static bool is_periodic_noise(const std::string& o) { return o == "pnoise" || o == "psnoise"; }
    // The noise name, the index of the first coordinate arg and the number
    // of coordinate args of a noise op; nnum counts the numeric args
    // (coordinates, then for pnoise/psnoise as many periods) before the
    // optional (name, value) pairs. The periodic forms take (S, Sper) or
    // (S, T, Sper, Tper) (batched_llvm_gen.cpp:6066-6081).
    void noise_layout(int L, const OOp& op, std::string& name, int& first, int& ncoord,
                      int* nnum = nullptr)
    {
        first = 1;
        name  = (op.name == "snoise" || op.name == "psnoise") ? "perlin"
                : op.name == "cellnoise"                      ? "cell"
                : op.name == "hashnoise"                      ? "hash"
                                                              : "uperlin";
        if (op.args.size() > 1 && S(L, op.args[1]).type.base == kString) {
            strval(L, op.args[1], name);
            first = 2;
        }
        int n = 0;
        for (size_t a = first; a < op.args.size(); ++a) {
            if (S(L, op.args[a]).type.base == kString)
                break;
            ++n;
        }
        if (nnum)
            *nnum = n;
        ncoord = is_periodic_noise(op.name) ? n / 2 : n;
    }

    bool differentiable(const std::string& o)
    {
        static const std::set<std::string> d = {
            "assign", "add", "sub", "mul", "div", "neg", "mix", "smoothstep", "clamp", "min", "max",
            "sqrt", "exp", "sin", "cos", "dot", "compref", "compassign", "color", "point", "vector",
            "normal", "transformc", "noise", "snoise", "pnoise", "psnoise", "transform",
            "transformv", "transformn"
        };
        return d.count(o) > 0;
    }

    // Lobe capacity of the group's closure values, bounded through the
    // dataflow: a closure constructor makes one lobe, add holds the sum of
    // its operands' lobes, mul and assign those of their closure operand.
    // The ops are walked in program order (both arms of an if from the
    // same state, merged by maximum; a loop's body repeated until nothing
    // grows), every count only ever rising, so the result bounds each value
    // at every point; connections carry the source's count and Ci, one
    // variable shared by the layers, the count the layers before left.
    // Capped at 8, the most a CompositeBSDF holds (a closure that grows in
    // a loop reaches the cap).
    int closure_capacity() const
    {
        std::vector<std::vector<int>> cap(g.layers.size());
        int ci_cap = 0;
        for (size_t Lz = 0; Lz < g.layers.size(); ++Lz) {
            const int L = (int)Lz;
            cap[L].assign(M(L).syms.size(), 0);
            for (const Connection& c : g.conns)
                if (c.dst_layer == L && is_closure(c.src_layer, c.src_sym) && is_closure(L, c.dst_sym))
                    cap[L][c.dst_sym] = std::max(cap[L][c.dst_sym], cap[c.src_layer][c.src_sym]);
            const int ci = shared_ci(L);
            if (ci >= 0)
                cap[L][ci] = ci_cap;
            cap_walk(L, 0, (int)M(L).ops.size(), cap[L]);
            if (ci >= 0)
                ci_cap = cap[L][ci];
        }
        int count = 1;
        for (size_t L = 0; L < g.layers.size(); ++L)
            for (int v : cap[L])
                count = std::max(count, v);
        return count;
    }

    bool is_closure(int L, int s) const
    {
        return s >= 0 && s < (int)M(L).syms.size() && S(L, s).type.base == kClosure;
    }

    void cap_walk(int L, int b, int e, std::vector<int>& cap) const
    {
        const int cap_max            = 8;
        const std::vector<OOp>& ops = M(L).ops;
        for (int i = b; i < e && i < (int)ops.size();) {
            const OOp& op = ops[i];
            if (op.name == "if" && op.jump[1] > i) {
                std::vector<int> other = cap;
                cap_walk(L, i + 1, op.jump[0], cap);
                cap_walk(L, op.jump[0], op.jump[1], other);
                for (size_t k = 0; k < cap.size(); ++k)
                    cap[k] = std::max(cap[k], other[k]);
                i = op.jump[1];
                continue;
            }
            if ((op.name == "for" || op.name == "while" || op.name == "dowhile") && op.jump[3] > i) {
                cap_walk(L, i + 1, op.jump[0], cap);
                for (int round = 0; round <= cap_max + 1; ++round) {
                    const std::vector<int> before = cap;
                    cap_walk(L, op.jump[0], op.jump[3], cap);
                    if (cap == before)
                        break;
                }
                i = op.jump[3];
                continue;
            }
            if (op.name == "functioncall" && op.jump[0] > i) {
                cap_walk(L, i + 1, op.jump[0], cap);
                i = op.jump[0];
                continue;
            }
            ++i;
            if (op.args.empty() || !is_closure(L, op.args[0]))
                continue;
            // a closure constructor makes one lobe (layer: the lobes of its
            // two closure arguments); add, mul and assign carry their
            // closure operands' lobes
            int v = 0;
            for (size_t k = 1; k < op.args.size(); ++k)
                if (is_closure(L, op.args[k]))
                    v += cap[op.args[k]];
            if (op.name == "closure" && v == 0)
                v = 1;
            int& r = cap[op.args[0]];
            r      = std::max(r, std::min(v, cap_max));
        }
    }

    // MaterialX closures (stdosl.h:522-707) as libbsdl lobes built at the
    // direction the closure is evaluated from (mxwo_, the renderer's -I),
    // and layer(top, base) with testrender's weighting. Positional
    // arguments as declared; keyword arguments after them in (name, value)
    // pairs. Reflection only: a transmission tint or an absorption must be
    // a constant zero, the sheen mode 0 (Conty-Kulla); the distribution
    // string, thin film, dispersion (no hero wavelength) and labels do not
    // change the lobes.
    void mx_closure(int L, const OOp& op, const std::string& cname, const std::string& R)
    {
        auto a = [&](size_t k) { return op.args[k]; };
        auto is_zero_const = [&](int sym) {
            std::vector<float> v;
            if (S(L, sym).kind != kConst || !numval(L, sym, v))
                return false;
            return std::all_of(v.begin(), v.end(), [](float x) { return x == 0.0f; });
        };
        const size_t npos = cname == "dielectric_bsdf" ? 8 : cname == "generalized_schlick_bsdf" ? 10
                          : cname == "sheen_bsdf" ? 3 : 2;
        if (op.args.size() < 2 + npos) {
            fail(cname + " needs " + std::to_string(npos) + " arguments");
            return;
        }
        for (size_t k = 2 + npos; k < op.args.size(); k += 2) {
            std::string key;
            if (k + 1 >= op.args.size() || !strval(L, a(k), key)) {
                fail(cname + ": malformed keyword arguments");
                return;
            }
            const bool ignored = key == "label" || key == "thinfilm_thickness" || key == "thinfilm_ior"
                                 || (key == "dispersion" && cname == "dielectric_bsdf");
            const bool zero_only = (key == "absorption" && cname == "dielectric_bsdf")
                                   || (key == "mode" && cname == "sheen_bsdf");
            if (zero_only && !is_zero_const(a(k + 1))) {
                fail(cname + ": '" + key + "' other than a constant 0 is not supported");
                return;
            }
            if (!ignored && !zero_only) {
                fail(cname + ": keyword '" + key + "' is not supported");
                return;
            }
        }
        if (cname == "layer") {
            if (!is_closure(L, a(2)) || !is_closure(L, a(3))) {
                fail("layer needs two closures");
                return;
            }
            uses_mx = true;
            line(R + " = layer(" + var(L, a(2)) + ", " + var(L, a(3)) + ");");
            return;
        }
        if (cname != "sheen_bsdf" && !is_zero_const(a(5))) {
            fail(cname + ": transmission (a transmission tint other than a constant 0) is not supported");
            return;
        }
        uses_mx = true;
        if (cname == "dielectric_bsdf") {
            line(R + " = closure_mx_dielectric(" + ex(L, a(2), K_V3) + ", " + ex(L, a(3), K_V3) + ", "
                 + ex(L, a(4), K_V3) + ", " + ex(L, a(6), K_FLOAT) + ", " + ex(L, a(7), K_FLOAT) + ", "
                 + ex(L, a(8), K_FLOAT) + ", mxwo_, A.bsdl);");
        } else if (cname == "generalized_schlick_bsdf") {
            line(R + " = closure_mx_schlick(" + ex(L, a(2), K_V3) + ", " + ex(L, a(3), K_V3) + ", "
                 + ex(L, a(4), K_V3) + ", " + ex(L, a(6), K_FLOAT) + ", " + ex(L, a(7), K_FLOAT) + ", "
                 + ex(L, a(8), K_V3) + ", " + ex(L, a(9), K_V3) + ", " + ex(L, a(10), K_FLOAT)
                 + ", mxwo_, A.bsdl);");
        } else {
            line(R + " = closure_mx_sheen(" + ex(L, a(2), K_V3) + ", " + ex(L, a(3), K_V3) + ", "
                 + ex(L, a(4), K_FLOAT) + ", mxwo_);");
        }
    }

    // code of a color space name (ColorSpaceCode of the device library), or
    // -1 for one only OpenColorIO knows. transformc also takes "linear" and
    // "sRGB"; the color constructor's to_rgb neither (opcolor.cpp:275-296).
    int color_space_code(const std::string& n, bool transformc) const
    {
        if (n == "rgb" || n == "RGB" || n == ss.colorspace || (transformc && n == "linear"))
            return 0;
        if (n == "hsv") return 1;
        if (n == "hsl") return 2;
        if (n == "YIQ") return 3;
        if (n == "XYZ") return 4;
        if (n == "xyY") return 5;
        if (n == "sRGB" && transformc) return 6;
        return -1;
    }

    // the layer's Ci global (a closure), or -1
    int shared_ci(int L) const
    {
        const int k = M(L).find("Ci");
        return k >= 0 && S(L, k).kind == kGlobal && S(L, k).type.base == kClosure ? k : -1;
    }

    void analyze()
    {
        deriv.assign(g.layers.size(), {});
        for (size_t L = 0; L < g.layers.size(); ++L)
            deriv[L].assign(M((int)L).syms.size(), 0);
        auto mark = [&](int L, int s, bool& changed) {
            if (!S(L, s).type.is_float_based() || deriv[L][s])
                return;
            deriv[L][s] = 1;
            changed     = true;
        };
        bool changed = true;
        // roots
        for (size_t L = 0; L < g.layers.size(); ++L)
            for (const OOp& op : M((int)L).ops) {
                if ((op.name == "filterwidth" || op.name == "Dx" || op.name == "Dy")
                    && op.args.size() >= 2)
                    mark((int)L, op.args[1], changed);
                if (is_noise_op(op.name)) {
                    std::string nm;
                    int first, nc;
                    noise_layout((int)L, op, nm, first, nc);
                    if (nm == "gabor")
                        for (int a = 0; a < nc; ++a)
                            mark((int)L, op.args[first + a], changed);
                }
            }
        while (changed) {
            changed = false;
            for (size_t L = 0; L < g.layers.size(); ++L)
                for (const OOp& op : M((int)L).ops) {
                    if (op.args.empty() || !differentiable(op.name))
                        continue;
                    int r = op.args[0];
                    if (op.name == "compassign") {
                        // compassign r i x : x takes r's deriv need
                        if (deriv[L][r] && op.args.size() >= 3)
                            mark((int)L, op.args[2], changed);
                        continue;
                    }
                    if (!deriv[L][r])
                        continue;
                    if (is_noise_op(op.name)) {
                        std::string nm;
                        int first, nc;
                        noise_layout((int)L, op, nm, first, nc);
                        if (nm == "cell" || nm == "hash")
                            continue;
                        for (int a = 0; a < nc; ++a)
                            mark((int)L, op.args[first + a], changed);
                        continue;
                    }
                    for (size_t a = 1; a < op.args.size(); ++a)
                        mark((int)L, op.args[a], changed);
                }
            for (const Connection& c : g.conns)
                if (deriv[c.dst_layer][c.dst_sym])
                    mark(c.src_layer, c.src_sym, changed);
        }
    }

    // ----------------------------------------------------------- codegen
    void line(const std::string& s) { os << "        " << s << "\n"; }

    std::string global_expr(int L, int s)
    {
        const OSymbol& y = S(L, s);
        bool d           = deriv[L][s];
        auto tri_global  = [&](int idx) {
            globals.insert(idx);
            return std::string("ld3(A, ") + std::to_string(idx) + ", i)";
        };
        auto f_global = [&](int idx) {
            globals.insert(idx);
            return std::string("ld1(A, ") + std::to_string(idx) + ", i)";
        };
        if (y.name == "P") {
            if (d)
                return "Dv(" + tri_global(kG_P) + ", " + tri_global(kG_dPdx) + ", " + tri_global(kG_dPdy) + ")";
            return tri_global(kG_P);
        }
        if (y.name == "u" || y.name == "v") {
            int a = y.name == "u" ? kG_u : kG_v;
            int b = y.name == "u" ? kG_dudx : kG_dvdx;
            int c = y.name == "u" ? kG_dudy : kG_dvdy;
            if (d)
                return "Df(" + f_global(a) + ", " + f_global(b) + ", " + f_global(c) + ")";
            return f_global(a);
        }
        std::string base;
        if (y.name == "N") base = tri_global(kG_N);
        else if (y.name == "Ng") base = tri_global(kG_Ng);
        else if (y.name == "I") base = tri_global(kG_I);
        else if (y.name == "dPdu") base = tri_global(kG_dPdu);
        else if (y.name == "dPdv") base = tri_global(kG_dPdv);
        else if (y.name == "Ps") base = tri_global(kG_Ps);
        else if (y.name == "time") base = f_global(kG_time);
        else if (y.name == "Ci") return y.type.base == kClosure ? "ci_" : "Closure()";
        else {
            if (y.type.base == kTriple) base = "V3(0.0f)";
            else if (y.type.base == kFloat) base = "0.0f";
            else return "0";
        }
        if (d)
            return y.type.base == kTriple ? "Dv(" + base + ")" : "Df(" + base + ")";
        return base;
    }

    // literal for a param/const default or instance value
    std::string value_expr(int L, int s, const std::vector<float>& vals)
    {
        const OSymbol& y = S(L, s);
        Kind k           = kind(L, s);
        auto at = [&](size_t i) { return vals.empty() ? 0.0f : vals[std::min(i, vals.size() - 1)]; };
        switch (y.type.base) {
        case kInt: return std::to_string((int)at(0));
        case kFloat: return k == K_DF ? "Df(" + flit(at(0)) + ")" : flit(at(0));
        case kTriple: {
            std::string t = "V3(" + flit(at(0)) + ", " + flit(at(1)) + ", " + flit(at(2)) + ")";
            return k == K_DV ? "Dv(" + t + ")" : t;
        }
        case kMatrix: {
            std::string m = "M44{{";
            for (int i = 0; i < 16; ++i)
                m += flit(vals.size() == 16 ? vals[i] : ((i % 5 == 0) ? at(0) : 0.0f)) + (i < 15 ? "," : "");
            return m + "}}";
        }
        case kClosure: return "Closure()";
        default: return "0";
        }
    }

    void declare_layer(int L)
    {
        const ShaderMaster& m = M(L);
        for (size_t s = 0; s < m.syms.size(); ++s) {
            const OSymbol& y = m.syms[s];
            if (y.type.base == kString || y.type.base == kVoid || y.type.base == kStruct)
                continue;
            if (y.type.arraylen != 0) {
                fail("arrays are not supported (symbol " + y.name + ")");
                return;
            }
            std::string init;
            if (y.kind == kGlobal) {
                init = global_expr(L, (int)s);
            } else if (y.kind == kConst) {
                std::vector<float> v;
                for (double d : y.fdefaults)
                    v.push_back((float)d);
                init = value_expr(L, (int)s, v);
            } else if (y.kind == kParam || y.kind == kOParam) {
                std::vector<float> v;
                numval(L, (int)s, v);
                init = value_expr(L, (int)s, v);
            } else {
                init = value_expr(L, (int)s, {});
            }
            line(std::string(ctype(kind(L, (int)s))) + " " + var(L, (int)s) + " = " + init + ";  // "
                 + y.name);
        }
        // connected inputs take the upstream layer's output value
        for (const Connection& c : g.conns)
            if (c.dst_layer == L)
                line(var(L, c.dst_sym) + " = " + ex(c.src_layer, c.src_sym, kind(L, c.dst_sym)) + ";");
    }

    Kind arith_kind(int L, const OOp& op)
    {
        return kind(L, op.args[0]);
    }

    // value kind (no derivs) of a kind
    static Kind valk(Kind k) { return k == K_DF ? K_FLOAT : k == K_DV ? K_V3 : k; }

    void emit_noise(int L, const OOp& op)
    {
        std::string nm;
        int first, nc, nnum;
        noise_layout(L, op, nm, first, nc, &nnum);
        const bool periodic = is_periodic_noise(op.name);
        if (periodic && (nnum != 2 && nnum != 4)) {
            fail(op.name + " needs (P, Pperiod) or (P, t, Pperiod, tperiod)");
            return;
        }
        int r   = op.args[0];
        Kind rk = kind(L, r);
        bool rd = rk == K_DF || rk == K_DV;
        bool tr = rk == K_V3 || rk == K_DV;
        // coordinates as a list of float-like components
        std::vector<std::string> comps;
        int dim = 0;
        Kind ck = rd ? K_DF : K_FLOAT;
        for (int a = 0; a < nc; ++a) {
            int s = op.args[first + a];
            if (S(L, s).type.base == kTriple) {
                std::string v = ex(L, s, rd ? K_DV : K_V3);
                if (rd) {
                    comps.push_back("comp(" + v + ", 0)");
                    comps.push_back("comp(" + v + ", 1)");
                    comps.push_back("comp(" + v + ", 2)");
                } else {
                    comps.push_back("(" + v + ").x");
                    comps.push_back("(" + v + ").y");
                    comps.push_back("(" + v + ").z");
                }
                dim += 3;
            } else {
                comps.push_back(ex(L, s, ck));
                dim += 1;
            }
        }
        // periods: plain floats, one per coordinate component
        std::vector<std::string> pers;
        for (int a = nc; periodic && a < nnum; ++a) {
            int s = op.args[first + a];
            if (S(L, s).type.base == kTriple) {
                std::string v = ex(L, s, K_V3);
                pers.push_back("(" + v + ").x");
                pers.push_back("(" + v + ").y");
                pers.push_back("(" + v + ").z");
            } else {
                pers.push_back(ex(L, s, K_FLOAT));
            }
        }
        if (periodic && (int)pers.size() != dim) {
            fail(op.name + ": the periods do not match the coordinates");
            return;
        }
        std::string T = rd ? "Df" : "float";
        std::string out;
        auto join = [&](int n) {
            std::string s;
            for (int i = 0; i < n; ++i)
                s += (i ? ", " : "") + comps[i];
            return s;
        };
        // the lattice map argument of the periodic perlin
        const std::string wa = periodic ? ", w_" : "";
        if (nm == "perlin" || nm == "snoise" || nm == "uperlin" || nm == "noise") {
            bool uns = (nm == "uperlin" || nm == "noise");
            std::string body;
            std::string wdecl;
            if (periodic) {
                std::string pa;
                for (int i = 0; i < dim; ++i)
                    pa += (i ? ", " : "") + pers[i];
                wdecl = "const Wrap w_ = make_wrap(" + pa + "); ";
            }
            if (!tr) {
                std::string call = dim == 1   ? "perlin1<" + T + ">(" + join(1) + ", 0" + wa + ")"
                                   : dim == 2 ? "perlin2<" + T + ">(" + join(2) + ", 0" + wa + ")"
                                   : dim == 3 ? "perlin3<" + T + ">(" + join(3) + wa + ")"
                                              : "perlin4<" + T + ">(" + join(4) + ", 0" + wa + ")";
                body = uns ? "0.5f * (" + call + " + 1.0f)" : call;
                line("{ " + wdecl + var(L, r) + " = " + body + "; }");
            } else {
                std::string c[3];
                if (dim == 3) {
                    line("{ " + wdecl + T + " c_[3]; vperlin3<" + T + ">(" + join(3) + ", c_" + wa + ");");
                } else {
                    line("{ " + wdecl + T + " c_[3];");
                    for (int k = 0; k < 3; ++k) {
                        std::string sh = std::to_string(8 * k);
                        std::string call = dim == 1   ? "perlin1<" + T + ">(" + join(1) + ", " + sh + wa + ")"
                                           : dim == 2 ? "perlin2<" + T + ">(" + join(2) + ", " + sh + wa + ")"
                                                      : "perlin4<" + T + ">(" + join(4) + ", " + sh + wa + ")";
                        line("  c_[" + std::to_string(k) + "] = " + call + ";");
                    }
                }
                if (uns)
                    line("  for (int k_ = 0; k_ < 3; ++k_) c_[k_] = 0.5f * (c_[k_] + 1.0f);");
                line("  " + var(L, r) + (rd ? " = make_dv(c_[0], c_[1], c_[2]); }"
                                           : " = V3(c_[0], c_[1], c_[2]); }"));
            }
            return;
        }
        if (nm == "cell" || nm == "cellnoise" || nm == "hash" || nm == "hashnoise") {
            std::string cl = (nm == "cell" || nm == "cellnoise") ? "true" : "false";
            // values only; the periodic forms hash the wrapped coordinates
            std::vector<std::string> vc;
            for (size_t i = 0; i < comps.size(); ++i) {
                std::string c = rd ? "(" + comps[i] + ").v" : comps[i];
                vc.push_back(periodic ? "wrap_period(" + c + ", " + pers[i] + ")" : c);
            }
            std::string args;
            for (int i = 0; i < dim; ++i)
                args += (i ? ", " : "") + vc[i];
            std::string fn = (tr ? "vintnoise" : "intnoise") + std::to_string(dim) + "<" + cl + ">(" + args + ")";
            std::string v  = rd ? (tr ? "Dv(" + fn + ")" : "Df(" + fn + ")") : fn;
            line("{ " + var(L, r) + " = " + v + "; }");
            return;
        }
        if ((nm == "simplex" || nm == "usimplex") && periodic) {
            fail("periodic noise type '" + nm + "' is unknown");
            return;
        }
        if (nm == "simplex" || nm == "usimplex") {
            bool uns = nm == "usimplex";
            if (dim < 1 || dim > 4) {
                fail("simplex noise supports 1-4 dimensions");
                return;
            }
            line("{");
            for (int k = 0; k < (tr ? 3 : 1); ++k) {
                std::string ks = std::to_string(k);
                if (rd) {
                    if (dim == 3) {
                        line("  Df s" + ks + "_ = simplex3_d(" + join(3) + ", " + ks + "u);");
                    } else if (dim == 4) {
                        line("  Df s" + ks + "_ = simplex4_d(" + join(4) + ", " + ks + "u);");
                    } else {
                        line("  float d" + ks + "_[3] = {0.f, 0.f, 0.f};");
                        if (dim == 1)
                            line("  Df s" + ks + "_; { Df x_ = " + comps[0] + "; float v_ = simplex1(x_.v, " + ks
                                 + "u, d" + ks + "_); s" + ks + "_ = Df(v_, d" + ks + "_[0] * x_.dx, d" + ks
                                 + "_[0] * x_.dy); }");
                        else
                            line("  Df s" + ks + "_; { Df x_ = " + comps[0] + ", y_ = " + comps[1]
                                 + "; float v_ = simplex2(x_.v, y_.v, " + ks + "u, d" + ks + "_); s" + ks
                                 + "_ = Df(v_, d" + ks + "_[0] * x_.dx + d" + ks + "_[1] * y_.dx, d" + ks
                                 + "_[0] * x_.dy + d" + ks + "_[1] * y_.dy); }");
                    }
                    if (uns)
                        line("  s" + ks + "_ = Df(0.5f * (s" + ks + "_.v + 1.0f), 0.5f * s" + ks
                             + "_.dx, 0.5f * s" + ks + "_.dy);");
                } else {
                    std::string call = dim == 1   ? "simplex1(" + join(1) + ", " + ks + "u, nullptr)"
                                       : dim == 2 ? "simplex2(" + join(2) + ", " + ks + "u, nullptr)"
                                       : dim == 3 ? "simplex3<false>(" + join(3) + ", " + ks + "u, nullptr)"
                                                  : "simplex4<false>(" + join(4) + ", " + ks + "u, nullptr)";
                    line("  float s" + ks + "_ = " + call + ";");
                    if (uns)
                        line("  s" + ks + "_ = 0.5f * (s" + ks + "_ + 1.0f);");
                }
            }
            if (tr)
                line("  " + var(L, r) + (rd ? " = make_dv(s0_, s1_, s2_);" : " = V3(s0_, s1_, s2_);"));
            else
                line("  " + var(L, r) + " = s0_;");
            line("}");
            return;
        }
        if (nm == "gabor") {
            GaborOptions go;
            for (size_t a = first + nnum; a + 1 < op.args.size(); a += 2) {
                std::string key;
                strval(L, op.args[a], key);
                std::vector<float> v;
                numval(L, op.args[a + 1], v);
                if (v.empty())
                    v.push_back(0.0f);
                if (key == "anisotropic") go.anisotropic = (int)v[0];
                else if (key == "do_filter") go.do_filter = (int)v[0];
                else if (key == "bandwidth") go.bandwidth = v[0];
                else if (key == "impulses") go.impulses = v[0];
                else if (key == "direction" && v.size() >= 3) {
                    go.direction[0] = v[0];
                    go.direction[1] = v[1];
                    go.direction[2] = v[2];
                }
            }
            GaborUniform gu = build_gabor_uniform(go);
            std::string gus = "GaborUniform{" + flit(gu.a) + "," + flit(gu.lambda) + ","
                              + flit(gu.sqrt_lambda_inv) + "," + flit(gu.radius) + ","
                              + flit(gu.radius2) + "," + flit(gu.radius3) + "," + flit(gu.radius_inv)
                              + "," + flit(gu.poisson_g) + "," + flit(gu.scale) + "," + flit(gu.dir_x)
                              + "," + flit(gu.dir_y) + "," + flit(gu.dir_z) + ","
                              + std::to_string(gu.anisotropic) + "," + std::to_string(gu.do_filter) + "}";
            // gabor needs Dual2 coordinates
            std::vector<std::string> dc;
            for (int a = 0; a < nc; ++a) {
                int s = op.args[first + a];
                if (S(L, s).type.base == kTriple) {
                    std::string v = ex(L, s, K_DV);
                    dc = { "comp(" + v + ",0)", "comp(" + v + ",1)", "comp(" + v + ",2)" };
                } else {
                    dc.push_back(ex(L, s, K_DF));
                }
            }
            while (dc.size() < 3)
                dc.push_back("Df(0.0f)");
            // pgabor: periods of the (up to) three spatial coordinates,
            // zero-padded; a 4th (time) period is ignored as the time
            // coordinate is (gabornoise.cpp:338-352, opnoise.cpp:558-632)
            std::string pv = "V3(0.0f)";
            if (periodic) {
                std::vector<std::string> pp(pers.begin(), pers.begin() + std::min<size_t>(3, pers.size()));
                while (pp.size() < 3)
                    pp.push_back("0.0f");
                pv = "V3(" + pp[0] + ", " + pp[1] + ", " + pp[2] + ")";
            }
            const std::string pargs = std::string(periodic ? "true" : "false") + ", " + pv;
            line("{ const GaborUniform gu_ = " + gus + ";");
            line("  Df gx_ = " + dc[0] + ", gy_ = " + dc[1] + ", gz_ = " + dc[2] + ";");
            const std::string P = "Dv(V3(gx_.v, gy_.v, gz_.v), V3(gx_.dx, gy_.dx, gz_.dx), V3(gx_.dy, gy_.dy, gz_.dy))";
            if (tr) {
                line("  Dv r_ = gabor3(gu_, " + P + ", " + pargs + ");");
                line("  " + var(L, r) + " = " + (rk == K_DV ? "r_" : "r_.v") + "; }");
            } else {
                line("  Df r_ = gabor(gu_, " + P + ", " + pargs + ");");
                line("  " + var(L, r) + " = " + (rk == K_DF ? "r_" : "r_.v") + "; }");
            }
            return;
        }
        fail("unsupported noise type '" + nm + "'");
    }

    void emit_op(int L, int opi, const OOp& op)
    {
        const std::string& o = op.name;
        auto a               = [&](int i) { return op.args[i]; };
        if (o == "nop" || o == "useparam" || o == "end")
            return;
        if (op.args.empty()) {
            if (o == "break" || o == "continue" || o == "exit" || o == "return")
                return;  // handled by the structured emitter
            fail("op " + o + " without arguments");
            return;
        }
        int r   = a(0);
        Kind rk = kind(L, r);
        std::string R = var(L, r);
        if (o == "assign") {
            if (rk == K_CLOSURE) {
                line(R + " = " + var(L, a(1)) + ";");
                return;
            }
            line(R + " = " + ex(L, a(1), rk) + ";");
            return;
        }
        if (o == "add" || o == "sub" || o == "mul" || o == "div") {
            Kind k1 = kind(L, a(1)), k2 = kind(L, a(2));
            if (rk == K_CLOSURE) {
                if (o == "add") {
                    line(R + " = add(" + var(L, a(1)) + ", " + var(L, a(2)) + ");");
                } else if (o == "mul") {
                    int c = k1 == K_CLOSURE ? a(1) : a(2), w = k1 == K_CLOSURE ? a(2) : a(1);
                    line(R + " = mul(" + var(L, c) + ", " + ex(L, w, K_V3) + ");");
                } else {
                    fail("closure " + o + " is not supported");
                }
                return;
            }
            (void)k1;
            (void)k2;
            line(R + " = " + o + "(" + ex(L, a(1), rk) + ", " + ex(L, a(2), rk) + ");");
            return;
        }
        if (o == "neg") {
            line(R + " = neg(" + ex(L, a(1), rk) + ");");
            return;
        }
        if (o == "mod") {
            if (rk == K_INT)
                line(R + " = " + var(L, a(2)) + " != 0 ? " + ex(L, a(1), K_INT) + " % " + ex(L, a(2), K_INT) + " : 0;");
            else if (valk(rk) == K_FLOAT)
                line(R + " = " + (rk == K_DF ? "Df(" : "(") + "fmod_osl(" + ex(L, a(1), K_FLOAT) + ", " + ex(L, a(2), K_FLOAT) + ")"
                     + (rk == K_DF ? ", " + ex(L, a(1), K_DF) + ".dx, " + ex(L, a(1), K_DF) + ".dy)" : ")") + ";");
            else
                fail("mod on triples is not supported");
            return;
        }
        if (o == "mix") {
            line(R + " = mix(" + ex(L, a(1), rk) + ", " + ex(L, a(2), rk) + ", " + ex(L, a(3), rk) + ");");
            return;
        }
        if (o == "smoothstep") {
            if (rk == K_V3)
                line(R + " = smoothstep(" + ex(L, a(1), K_V3) + ", " + ex(L, a(2), K_V3) + ", " + ex(L, a(3), K_V3) + ");");
            else
                line(R + " = smoothstep(" + ex(L, a(1), rk) + ", " + ex(L, a(2), rk) + ", " + ex(L, a(3), rk) + ");");
            return;
        }
        if (o == "clamp") {
            line(R + " = clamp(" + ex(L, a(1), rk) + ", " + ex(L, a(2), rk) + ", " + ex(L, a(3), rk) + ");");
            return;
        }
        if (o == "min" || o == "max") {
            line(R + " = " + o + "_(" + ex(L, a(1), rk) + ", " + ex(L, a(2), rk) + ");");
            return;
        }
        if (o == "floor" || o == "abs" || o == "fabs") {
            Kind vk = valk(rk);
            std::string f = o == "floor" ? "floor_" : "abs_";
            std::string v = f + "(" + ex(L, a(1), vk) + ")";
            line(R + " = " + (rk == K_DF ? "Df(" + v + ")" : rk == K_DV ? "Dv(" + v + ")" : v) + ";");
            return;
        }
        if (o == "sqrt" || o == "exp" || o == "sin" || o == "cos") {
            line(R + " = " + o + "_(" + ex(L, a(1), rk) + ");");
            return;
        }
        if (o == "dot") {
            Kind ik = rk == K_DF ? K_DV : K_V3;
            line(R + " = dot_(" + ex(L, a(1), ik) + ", " + ex(L, a(2), ik) + ");");
            return;
        }
        if (o == "length") {
            line(R + " = " + (rk == K_DF ? "Df(length_(" + ex(L, a(1), K_V3) + "))" : "length_(" + ex(L, a(1), K_V3) + ")") + ";");
            return;
        }
        if (o == "normalize") {
            std::string v = "normalize_(" + ex(L, a(1), K_V3) + ")";
            line(R + " = " + (rk == K_DV ? "Dv(" + v + ")" : v) + ";");
            return;
        }
        if (o == "cross") {
            std::string v = "cross_(" + ex(L, a(1), K_V3) + ", " + ex(L, a(2), K_V3) + ")";
            line(R + " = " + (rk == K_DV ? "Dv(" + v + ")" : v) + ";");
            return;
        }
        if (o == "filterwidth" || o == "Dx" || o == "Dy") {
            Kind ak = kind(L, a(1));
            std::string v = (o == "filterwidth" ? "filterwidth(" : o + "(") + ex(L, a(1), ak) + ")";
            if (ak != K_DF && ak != K_DV)
                v = (valk(ak) == K_V3) ? "V3(0.0f)" : "0.0f";
            line(R + " = " + (rk == K_DF ? "Df(" + v + ")" : rk == K_DV ? "Dv(" + v + ")" : v) + ";");
            return;
        }
        if (o == "compref") {
            Kind ak  = kind(L, a(1));
            std::string idx = ex(L, a(2), K_INT);
            std::string v   = "compref(" + ex(L, a(1), ak) + ", " + idx + ")";
            if (ak == K_DV)
                line(R + " = " + (rk == K_DF ? v : "(" + v + ").v") + ";");
            else
                line(R + " = " + (rk == K_DF ? "Df(" + v + ")" : v) + ";");
            return;
        }
        if (o == "compassign") {
            Kind ck = rk == K_DV ? K_DF : K_FLOAT;
            line("compassign(" + R + ", " + ex(L, a(1), K_INT) + ", " + ex(L, a(2), ck) + ");");
            return;
        }
        if (o == "color" && op.args.size() == 5 && S(L, a(1)).type.base == kString) {
            // color(space, a, b, c): the components converted to rgb by
            // ColorSystem::to_rgb (opcolor.cpp:275-296); derivatives of the
            // result are zero (batched_llvm_gen.cpp:2740-2790)
            std::string space;
            if (!strval(L, a(1), space)) {
                fail("color constructor needs a constant space name");
                return;
            }
            const int fc = color_space_code(space, false);
            if (fc < 0) {
                fail("color space '" + space + "' needs OpenColorIO");
                return;
            }
            uses_cs = true;
            const std::string t = "V3(" + ex(L, a(2), K_FLOAT) + ", " + ex(L, a(3), K_FLOAT) + ", "
                                  + ex(L, a(4), K_FLOAT) + ")";
            const std::string v = fc == 0 ? t : "transformc<" + std::to_string(fc) + ", 0>(CS, " + t + ")";
            line(R + " = " + (rk == K_DV ? "Dv(" + v + ")" : v) + ";");
            return;
        }
        if (o == "color" || o == "point" || o == "vector" || o == "normal") {
            if (op.args.size() == 4) {
                Kind ck = rk == K_DV ? K_DF : K_FLOAT;
                line(R + " = triple(" + ex(L, a(1), ck) + ", " + ex(L, a(2), ck) + ", " + ex(L, a(3), ck) + ");");
            } else if (op.args.size() == 2) {
                line(R + " = " + ex(L, a(1), rk) + ";");
            } else {
                fail("unsupported form of " + o + " constructor");
            }
            return;
        }
        if (o == "lt" || o == "gt" || o == "le" || o == "ge" || o == "eq" || o == "neq") {
            Kind k1 = valk(kind(L, a(1))), k2 = valk(kind(L, a(2)));
            Kind ck = (k1 == K_V3 || k2 == K_V3) ? K_V3 : ((k1 == K_INT && k2 == K_INT) ? K_INT : K_FLOAT);
            std::string x = ex(L, a(1), ck), y = ex(L, a(2), ck);
            std::string v;
            if (ck == K_V3) {
                if (o != "eq" && o != "neq") {
                    fail("ordered comparison of triples");
                    return;
                }
                v = "((" + x + ").x == (" + y + ").x && (" + x + ").y == (" + y + ").y && (" + x + ").z == (" + y + ").z)";
                if (o == "neq")
                    v = "!" + v;
                v = "(int)" + v;
            } else {
                v = o + "(" + x + ", " + y + ")";
            }
            line(R + " = " + (rk == K_INT ? v : rk == K_DF ? "Df((float)" + v + ")" : "(float)" + v) + ";");
            return;
        }
        if (o == "and" || o == "or") {
            line(R + " = (" + ex(L, a(1), K_INT) + (o == "and" ? " && " : " || ") + ex(L, a(2), K_INT) + ") ? 1 : 0;");
            return;
        }
        if (o == "not") {
            line(R + " = !" + ex(L, a(1), K_INT) + ";");
            return;
        }
        if (is_noise_op(o)) {
            emit_noise(L, op);
            return;
        }
        if (o == "hash") {
            std::vector<std::string> c;
            for (size_t k = 1; k < op.args.size(); ++k) {
                int s = op.args[k];
                if (S(L, s).type.base == kTriple) {
                    std::string v = ex(L, s, K_V3);
                    c.push_back("(" + v + ").x");
                    c.push_back("(" + v + ").y");
                    c.push_back("(" + v + ").z");
                } else if (S(L, s).type.base == kInt) {
                    c.push_back("__int_as_float(" + var(L, s) + ")");
                } else {
                    c.push_back(ex(L, s, K_FLOAT));
                }
            }
            if (op.args.size() == 2 && S(L, a(1)).type.base == kInt) {
                line(R + " = hash_opi(" + var(L, a(1)) + ");");
                return;
            }
            std::string args;
            for (size_t k = 0; k < c.size(); ++k)
                args += (k ? ", " : "") + c[k];
            line(R + " = hash_op" + std::to_string(c.size()) + "(" + args + ");");
            return;
        }
        if (o == "blackbody") {
            uses_cs = uses_bb = true;
            std::string v = "blackbody_rgb(s_bb, CS->xyz2rgb, CS->cie, " + ex(L, a(1), K_FLOAT) + ")";
            line(R + " = " + (rk == K_DV ? "Dv(" + v + ")" : v) + ";");
            return;
        }
        if (o == "luminance") {
            uses_cs = true;
            std::string v = "luminance(CS->lum, " + ex(L, a(1), K_V3) + ")";
            line(R + " = " + (rk == K_DF ? "Df(" + v + ")" : v) + ";");
            return;
        }
        if (o == "transformc") {
            // transformc [from] to C (ColorSystem::transformc, opcolor.cpp:
            // 324-380): from -> rgb -> to, with derivatives when C has them
            uses_cs = true;
            std::string from = "rgb", to;
            int cidx = 2;
            if (op.args.size() == 4) {
                if (!strval(L, a(1), from) || !strval(L, a(2), to)) {
                    fail("transformc needs constant space names");
                    return;
                }
                cidx = 3;
            } else if (!strval(L, a(1), to)) {
                fail("transformc needs constant space names");
                return;
            }
            const int fc = color_space_code(from, true), tc = color_space_code(to, true);
            if (fc < 0 || tc < 0) {
                fail("transformc " + from + " -> " + to
                     + ": spaces beyond rgb/hsv/hsl/YIQ/XYZ/xyY/sRGB need OpenColorIO");
                return;
            }
            const std::string c = ex(L, a(cidx), rk);
            if (fc == 0 && tc == 0)  // other pairs go through rgb, as the reference does
                line(R + " = " + c + ";");
            else
                line(R + " = transformc<" + std::to_string(fc) + ", " + std::to_string(tc) + ">(CS, " + c + ");");
            return;
        }
        if (o == "transform" || o == "transformv" || o == "transformn") {
            // transform[v|n] R [from] to P, or R M P (batched_llvm_gen.cpp:3470-3560):
            // the opcode names the semantics (point, vector, normal)
            const std::string kind = o == "transform" ? "point" : (o == "transformv" ? "vector" : "normal");
            const int nargs = (int)op.args.size();
            if (nargs != 3 && nargs != 4) {
                fail(o + " needs 3 or 4 arguments");
                return;
            }
            const int pi = nargs == 3 ? 2 : 3;
            const std::string p = ex(L, a(pi), rk == K_DV ? K_DV : K_V3);
            if (nargs == 3 && S(L, a(1)).type.base == kMatrix) {
                const std::string fn = "xform_" + kind;
                line(R + " = " + fn + "(" + var(L, a(1)) + ", " + p + ");");
                return;
            }
            std::string from = "common", to;
            if ((nargs == 4 && !strval(L, a(1), from)) || !strval(L, a(nargs == 3 ? 1 : 2), to)) {
                fail(o + " needs constant space names");
                return;
            }
            if (from == ss.commonspace)
                from = "common";
            if (to == ss.commonspace)
                to = "common";
            if (from == to) {  // identity: a copy
                line(R + " = " + p + ";");
                return;
            }
            const std::string slot = xform_ref(from, to);
            if (slot.empty())
                return;
            // unknown spaces leave the value unchanged (opmatrix.cpp:286-292)
            const std::string fn = kind == "normal" ? "xform_vector" : "xform_" + kind;
            const std::string mat = kind == "normal" ? "xf_normal_matrix(" + slot + ")"
                                                     : "xf_matrix(" + slot + ")";
            line(R + " = xf_ok(" + slot + ") ? " + fn + "(" + mat + ", " + p + ") : " + p + ";");
            return;
        }
        if (o == "matrix") {
            if (op.args.size() == 17) {
                std::string m = R + " = M44{{";
                for (int k = 1; k <= 16; ++k)
                    m += ex(L, a(k), K_FLOAT) + (k < 16 ? "," : "");
                line(m + "}};");
                return;
            }
            // matrix(from) and matrix(from, to) (opmatrix.cpp:141-243)
            std::string from, to = "common";
            if ((op.args.size() == 2 || op.args.size() == 3) && strval(L, a(1), from)
                && (op.args.size() == 2 || strval(L, a(2), to))) {
                if (from == ss.commonspace)
                    from = "common";
                if (to == ss.commonspace)
                    to = "common";
                if (from == to) {
                    line(R + " = m44_identity();");
                    return;
                }
                const std::string slot = xform_ref(from, to);
                if (!slot.empty())
                    line(R + " = xf_matrix(" + slot + ");");
                return;
            }
            fail("unsupported matrix constructor");
            return;
        }
        if (o == "getattribute") {
            // int getattribute(name, output value): the renderer's attribute
            // table (RendererServices::get_attribute); the value is written
            // only when the renderer has the name with the value's type
            std::string name;
            if (op.args.size() != 3 || !strval(L, a(1), name)) {
                fail("getattribute supports (constant name, value) only");
                return;
            }
            const OSymbol& dst = S(L, a(2));
            const BaseType b   = dst.type.base;
            if (dst.type.arraylen != 0 || (b != kInt && b != kFloat && b != kTriple && b != kMatrix)) {
                fail("getattribute of " + name + ": value must be an int, float, triple or matrix");
                return;
            }
            const int off = attr_slot(ss, name);
            if (off < 0) {
                fail("too many distinct renderer attributes (at most " +
                     std::to_string(kMaxAttrSlots) + ")");
                return;
            }
            g.uses_xforms = true;
            const std::string slot = "(A.xf + " + std::to_string(off) + ")";
            const Kind dk          = kind(L, a(2));
            std::string v;
            if (b == kInt) v = "__float_as_int(" + slot + "[0])";
            else if (b == kFloat) v = slot + "[0]";
            else if (b == kTriple) v = "V3(" + slot + "[0], " + slot + "[1], " + slot + "[2])";
            else v = "xf_matrix(" + slot + ")";
            if (dk == K_DF) v = "Df(" + v + ")";
            if (dk == K_DV) v = "Dv(" + v + ")";
            line("if (" + slot + "[16] == " + std::to_string((int)b + 1) + ".0f) { " + var(L, a(2))
                 + " = " + v + "; " + R + " = 1; } else { " + R + " = 0; }");
            return;
        }
        if (o == "getmatrix") {
            // int getmatrix(from, to, output M) (osl_get_from_to_matrix)
            std::string from, to;
            if (op.args.size() != 4 || !strval(L, a(1), from) || !strval(L, a(2), to)) {
                fail("getmatrix needs constant space names");
                return;
            }
            if (from == ss.commonspace)
                from = "common";
            if (to == ss.commonspace)
                to = "common";
            // getmatrix multiplies the from-matrix by the to-inverse even
            // for one space named twice (wide_opmatrix.cpp:869-885), which
            // fails for an unknown space; only common-to-common is the
            // identity without a table slot
            if (from == to && from == "common") {
                line(var(L, a(3)) + " = m44_identity();");
                line(R + " = 1;");
                return;
            }
            const std::string slot = xform_ref(from, to);
            if (slot.empty())
                return;
            line(var(L, a(3)) + " = xf_matrix(" + slot + ");");
            line(R + " = xf_ok(" + slot + ") ? 1 : 0;");
            return;
        }
        if (o == "closure") {
            std::string cname;
            strval(L, a(1), cname);
            if (cname == "diffuse") {
                line(R + " = closure_diffuse(" + ex(L, a(2), K_V3) + ");");
            } else if (cname == "reflection" || cname == "refraction") {
                // reflection(N) is REFLECTION_ID, reflection(N, eta)
                // FRESNEL_REFLECTION_ID, refraction(N, eta) REFRACTION_ID
                const bool with_eta = op.args.size() >= 4;
                if (cname == "refraction" && !with_eta) {
                    fail("refraction closure needs 2 arguments");
                    return;
                }
                const std::string id = cname == "refraction" ? "kClosureRefraction"
                                       : with_eta            ? "kClosureFresnelReflection"
                                                             : "kClosureReflection";
                line(R + " = closure_delta(" + id + ", " + ex(L, a(2), K_V3) + ", "
                     + (with_eta ? ex(L, a(3), K_FLOAT) : std::string("0.0f")) + ");");
            } else if (cname == "transparent") {
                line(R + " = closure_delta(kClosureTransparent, V3(0.0f), 0.0f);");
            } else if (cname == "emission") {
                line(R + " = closure_delta(kClosureEmission, V3(0.0f), 0.0f);");
            } else if (cname == "translucent") {
                line(R + " = closure_translucent(" + ex(L, a(2), K_V3) + ");");
            } else if (cname == "phong") {
                if (op.args.size() < 4) {
                    fail("phong closure needs 2 arguments");
                    return;
                }
                line(R + " = closure_phong(" + ex(L, a(2), K_V3) + ", " + ex(L, a(3), K_FLOAT) + ");");
            } else if (cname == "microfacet") {
                std::string dist;
                strval(L, a(2), dist);
                if (dist != "ggx" && dist != "beckmann") {
                    fail("microfacet distribution '" + dist + "' is not supported");
                    return;
                }
                // microfacet(dist, N, U, xalpha, yalpha, eta, refract)
                if (op.args.size() < 9) {
                    fail("microfacet closure needs 7 arguments");
                    return;
                }
                // the refract mode selects the lobe's type (testrender picks
                // Microfacet<Dist, 0|1|2> per closure): a constant 0, 1 or 2
                std::vector<float> refr;
                const bool refr_const = S(L, a(8)).kind == kConst && numval(L, a(8), refr)
                                        && !refr.empty();
                const int refract = refr_const ? (int)refr[0] : -1;
                if (refract < 0 || refract > 2) {
                    fail("microfacet refract must be a constant 0, 1 or 2");
                    return;
                }
                // isotropic lobes about N (yalpha the same symbol as xalpha,
                // as oslc writes microfacet(dist, N, alpha, eta, refract),
                // or equal constants, and U the constant 0) keep the short
                // lobe; anything else builds the anisotropic lobe with its
                // tangent frame
                auto is_const = [&](int sym) { return S(L, sym).kind == kConst; };
                std::vector<float> xa, ya, u;
                const bool same_alpha =
                    a(5) == a(6)
                    || (is_const(a(5)) && is_const(a(6)) && numval(L, a(5), xa)
                        && numval(L, a(6), ya) && xa == ya);
                const bool zero_u = is_const(a(4)) && numval(L, a(4), u)
                                    && std::all_of(u.begin(), u.end(),
                                                   [](float v) { return v == 0.0f; });
                if (!same_alpha || !zero_u || refract != 0) {
                    uses_aniso = true;
                    line(R + " = closure_microfacet_full("
                         + std::string(dist == "ggx" ? "kClosureMicrofacetGGX" : "kClosureMicrofacetBeckmann")
                         + ", " + ex(L, a(3), K_V3) + ", " + ex(L, a(4), K_V3) + ", " + ex(L, a(5), K_FLOAT)
                         + ", " + ex(L, a(6), K_FLOAT) + ", " + ex(L, a(7), K_FLOAT) + ", "
                         + std::to_string(refract) + ");");
                    return;
                }
                line(R + " = closure_microfacet_" + dist + "(" + ex(L, a(3), K_V3) + ", " + ex(L, a(5), K_FLOAT)
                     + ", " + ex(L, a(7), K_FLOAT) + ");");
            } else if (cname == "dielectric_bsdf" || cname == "generalized_schlick_bsdf"
                       || cname == "sheen_bsdf" || cname == "layer") {
                mx_closure(L, op, cname, R);
            } else {
                fail("closure '" + cname + "' is not supported");
            }
            return;
        }
        fail("unsupported op '" + o + "'");
        (void)opi;
    }

    // structured emission of ops [b, e) of layer L
    void block(int L, int b, int e, int loop_step_label, int loop_end_label, int layer_end_label,
               int return_label)
    {
        const ShaderMaster& m = M(L);
        for (int i = b; i < e && !failed;) {
            const OOp& op = m.ops[i];
            if (op.name == "if") {
                int j0 = op.jump[0], j1 = op.jump[1];
                line("if (" + ex(L, op.args[0], K_INT) + ") {");
                block(L, i + 1, j0, loop_step_label, loop_end_label, layer_end_label, return_label);
                line("} else {");
                block(L, j0, j1, loop_step_label, loop_end_label, layer_end_label, return_label);
                line("}");
                i = j1;
                continue;
            }
            if (op.name == "for" || op.name == "while" || op.name == "dowhile") {
                int j0 = op.jump[0], j1 = op.jump[1], j2 = op.jump[2], j3 = op.jump[3];
                int ls = ++label_counter, le = ++label_counter;
                std::string cond = ex(L, op.args[0], K_INT);
                line("{");
                block(L, i + 1, j0, loop_step_label, loop_end_label, layer_end_label, return_label);  // init
                if (op.name == "dowhile") {
                    line("for (;;) {");
                    block(L, j1, j2, ls, le, layer_end_label, return_label);  // body
                    line(lab(ls) + ":;");
                    block(L, j2, j3, ls, le, layer_end_label, return_label);  // step
                    block(L, j0, j1, ls, le, layer_end_label, return_label);  // cond
                    line("if (!(" + cond + ")) break;");
                    line("}");
                } else {
                    line("for (;;) {");
                    block(L, j0, j1, ls, le, layer_end_label, return_label);  // cond
                    line("if (!(" + cond + ")) break;");
                    block(L, j1, j2, ls, le, layer_end_label, return_label);  // body
                    line(lab(ls) + ":;");
                    block(L, j2, j3, ls, le, layer_end_label, return_label);  // step
                    line("}");
                }
                line(lab(le) + ":;");
                line("}");
                i = j3;
                continue;
            }
            if (op.name == "break") {
                line("goto " + lab(loop_end_label) + ";");
                ++i;
                continue;
            }
            if (op.name == "continue") {
                line("goto " + lab(loop_step_label) + ";");
                ++i;
                continue;
            }
            if (op.name == "exit") {
                line("goto " + lab(layer_end_label) + ";");
                ++i;
                continue;
            }
            if (op.name == "return") {
                // the end of the innermost function call, or of the layer
                line("goto " + lab(return_label) + ";");
                ++i;
                continue;
            }
            if (op.name == "functioncall") {
                // inline the function body [i+1, jump0); a return inside it
                // resumes after the call (llvm_gen.cpp:4327-4360, :4396-4411)
                const int fe = ++label_counter;
                line("{");
                block(L, i + 1, op.jump[0], loop_step_label, loop_end_label, layer_end_label, fe);
                line(lab(fe) + ":;");
                line("}");
                i = op.jump[0];
                continue;
            }
            line("{");
            emit_op(L, i, op);
            line("}");
            ++i;
        }
    }

    // Per-point code of the group: reads A (GroupArgs), i (point index), CS
    // (colour system) and s_bb (blackbody table) from the enclosing scope.
    bool per_point(std::string& code)
    {
        // topological layer order from connections (the group order must
        // already respect it, as ShaderGroupEnd requires)
        for (const Connection& c : g.conns)
            if (c.src_layer >= c.dst_layer) {
                err = "connection from layer " + g.layers[c.src_layer].name + " to earlier layer "
                      + g.layers[c.dst_layer].name;
                return false;
            }
        analyze();
        g.uses_xforms = false;
        // outputs
        g.outputs.clear();
        int plane = 0;
        for (const std::string& on : g.output_names) {
            std::string ln, sn = on;
            size_t dot = on.find('.');
            if (dot != std::string::npos) {
                ln = on.substr(0, dot);
                sn = on.substr(dot + 1);
            }
            int L = -1, s = -1;
            for (int k = (int)g.layers.size() - 1; k >= 0 && L < 0; --k) {
                if (!ln.empty() && g.layers[k].name != ln)
                    continue;
                int f = M(k).find(sn);
                if (f >= 0) {
                    L = k;
                    s = f;
                }
            }
            if (L < 0) {
                err = "output '" + on + "' not found in group";
                return false;
            }
            const OSymbol& y = S(L, s);
            if (y.type.base != kFloat && y.type.base != kTriple && y.type.base != kInt) {
                err = "output '" + on + "' must be int, float or a triple";
                return false;
            }
            OutputSlot o;
            o.name        = on;
            o.layer       = L;
            o.sym         = s;
            o.planes      = y.type.base == kTriple ? 3 : 1;
            o.first_plane = plane;
            plane += o.planes;
            g.outputs.push_back(o);
        }
        if (plane > OSLMI_MAX_OUTPUT_PLANES) {
            err = "too many output planes";
            return false;
        }
        g.output_planes = plane;
        // Ci is one variable of the group: every layer starts from the value
        // the layers before it left and hands its own on (a global, shared
        // by all layers as in the reference's ShaderGlobals)
        g.surface_closure = false;
        for (int L = 0; L < (int)g.layers.size(); ++L)
            if (shared_ci(L) >= 0)
                g.surface_closure = true;

        // body
        std::ostringstream body;
        std::swap(os, body);
        for (int L = 0; L < (int)g.layers.size() && !failed; ++L) {
            line("// ---- layer " + std::to_string(L) + ": " + g.layers[L].name + " (" + M(L).name + ")");
            line("{");
            declare_layer(L);
            int le = ++label_counter;
            block(L, M(L).main_begin, (int)M(L).ops.size(), -1, -1, le, le);
            line(lab(le) + ":;");
            // hand layer outputs to downstream layers
            for (const Connection& c : g.conns)
                if (c.src_layer == L)
                    line("x" + std::to_string(L) + "_" + std::to_string(c.src_sym) + " = " + var(L, c.src_sym) + ";");
            for (const OutputSlot& o : g.outputs)
                if (o.layer == L) {
                    std::string v = ex(L, o.sym, o.planes == 3 ? K_V3 : (S(L, o.sym).type.base == kInt ? K_INT : K_FLOAT));
                    if (o.planes == 3) {
                        line("{ V3 o_ = " + v + ";");
                        for (int k = 0; k < 3; ++k)
                            line("  A.out[" + std::to_string(o.first_plane + k) + "][i] = o_." + "xyz"[k] + ";");
                        line("}");
                    } else {
                        line("A.out[" + std::to_string(o.first_plane) + "][i] = (float)(" + v + ");");
                    }
                }
            if (shared_ci(L) >= 0)
                line("ci_ = " + var(L, shared_ci(L)) + ";");
            line("}");
        }
        std::swap(os, body);
        if (failed)
            return false;
        // connection carriers: variables that outlive a layer's scope
        std::ostringstream carriers;
        std::set<std::pair<int, int>> seen;
        for (const Connection& c : g.conns) {
            if (!seen.insert({ c.src_layer, c.src_sym }).second)
                continue;
            carriers << "        " << ctype(kind(c.src_layer, c.src_sym)) << " x" << c.src_layer << "_"
                     << c.src_sym << ";\n";
        }
        // the downstream layer reads from the carrier, not the scoped var
        std::string b = body.str();
        for (const Connection& c : g.conns) {
            std::string from = var(c.src_layer, c.src_sym);
            std::string to   = "x" + std::to_string(c.src_layer) + "_" + std::to_string(c.src_sym);
            // only occurrences inside the destination layer's declaration
            // line "vD_S = <expr of from>;" are rewritten
            std::string key = var(c.dst_layer, c.dst_sym) + " = ";
            size_t pos      = 0;
            while ((pos = b.find(key, pos)) != std::string::npos) {
                size_t eol = b.find('\n', pos);
                std::string ln = b.substr(pos, eol - pos);
                size_t f = ln.find(from);
                if (f != std::string::npos && ln.find("//") == std::string::npos) {
                    ln.replace(f, from.size(), to);
                    b.replace(pos, eol - pos, ln);
                }
                pos = eol;
            }
        }

        std::ostringstream k;
        if (g.surface_closure)
            k << "        Closure ci_;\n";
        if (uses_mx) {
            // MaterialX lobes are built at the direction they are evaluated
            // from: wo when the closure is evaluated or sampled, else -I
            globals.insert(kG_I);
            k << "        const V3 mxwo_ = A.wo ? V3(A.wo[i], A.wo[A.n + i], A.wo[2 * A.n + i])\n"
                 "                              : V3(0.0f) - ld3(A, kG_I, i);\n";
        }
        k << carriers.str();
        k << b;
        if (g.surface_closure && !ss.closure_sample) {
            k << "        if (A.eval) {\n"
                 "            V3 wo_(A.wo[i], A.wo[A.n + i], A.wo[2 * A.n + i]);\n"
                 "            V3 wi_(A.wi[i], A.wi[A.n + i], A.wi[2 * A.n + i]);\n"
                 "            composite_prepare(ci_, wo_);\n"
                 "            BsdfEval e_ = composite_eval(ci_, wo_, wi_);\n"
                 "            A.eval[i] = e_.weight.x; A.eval[A.n + i] = e_.weight.y;\n"
                 "            A.eval[2 * A.n + i] = e_.weight.z; A.eval[3 * A.n + i] = e_.pdf;\n"
                 "        }\n";
        } else if (g.surface_closure) {
            // evaluation and/or CompositeBSDF::sample of Ci at wo
            k << "        if (A.eval || A.samp) {\n"
                 "            V3 wo_(A.wo[i], A.wo[A.n + i], A.wo[2 * A.n + i]);\n"
                 "            composite_prepare(ci_, wo_);\n"
                 "            if (A.eval) {\n"
                 "                V3 wi_(A.wi[i], A.wi[A.n + i], A.wi[2 * A.n + i]);\n"
                 "                BsdfEval e_ = composite_eval(ci_, wo_, wi_);\n"
                 "                A.eval[i] = e_.weight.x; A.eval[A.n + i] = e_.weight.y;\n"
                 "                A.eval[2 * A.n + i] = e_.weight.z; A.eval[3 * A.n + i] = e_.pdf;\n"
                 "            }\n"
                 "            if (A.samp) {\n"
                 "                BsdfDirSample s_ = composite_sample(ci_, wo_, A.rnd[i], A.rnd[A.n + i],\n"
                 "                                                   A.rnd[2 * A.n + i]);\n"
                 "                A.samp[i] = s_.wi.x; A.samp[A.n + i] = s_.wi.y;\n"
                 "                A.samp[2 * A.n + i] = s_.wi.z; A.samp[3 * A.n + i] = s_.weight.x;\n"
                 "                A.samp[4 * A.n + i] = s_.weight.y; A.samp[5 * A.n + i] = s_.weight.z;\n"
                 "                A.samp[6 * A.n + i] = s_.pdf;\n"
                 "            }\n"
                 "        }\n";
        }
        if (g.surface_closure && ss.closure_export)
            k << "        if (A.clo) export_closure(ci_, A.clo, A.clo_max, A.n, i);\n";
        g.can_sample    = g.surface_closure && ss.closure_sample;
        g.can_export    = g.surface_closure && ss.closure_export;
        code            = k.str();
        g.uses_colorsys = uses_cs;
        g.uses_mx       = uses_mx;
        g.globals_used.assign(globals.begin(), globals.end());
        return true;
    }

    // standalone kernel of one group: a grid-stride loop over its points
    bool run(std::string& src)
    {
        std::string code;
        if (!per_point(code))
            return false;
        std::ostringstream k;
        k << "// Emitted by the oslmi group emitter for group '" << g.name << "'.\n";
        k << prelude(uses_mx, g.can_export, uses_aniso);
        k << "extern \"C\" __global__ " << kernel_attributes(ss) << " void oslmi_group_main(GroupArgs A)\n{\n";
        k << "    using Closure = LobeList<" << closure_capacity() << ">;\n";
        if (uses_cs)
            k << "    const ColorSysDev* CS = (const ColorSysDev*)A.colorsys;\n";
        if (uses_bb) {
            k << "    __shared__ float s_bb[OSLMI_BB_TABLE * 3];\n"
                 "    for (int j = threadIdx.x; j < OSLMI_BB_TABLE * 3; j += blockDim.x) s_bb[j] = CS->bb[j];\n"
                 "    __syncthreads();\n";
        }
        k << "    int64_t stride = (int64_t)gridDim.x * blockDim.x;\n";
        k << "    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += stride) {\n";
        k << code;
        k << "    }\n}\n";
        src = k.str();
        return true;
    }

    static std::string prelude(bool mx, bool xp = false, bool aniso = false)
    {
        return std::string(mx ? "#define OSLMI_MX_CLOSURES 1\n" : "")
               + std::string(aniso ? "#define OSLMI_FULL_MICROFACET 1\n" : "")
               + std::string(xp ? "#define OSLMI_CLOSURE_EXPORT 1\n" : "")
               + "#include \"oslmi_rt.h\"\n#include \"oslmi_groupargs.h\"\n"
               "using namespace oslmi;\nusing namespace oslmi::rt;\n"
               "__device__ __forceinline__ V3 ld3(const GroupArgs& A, int g, int64_t i) {\n"
               "    const float* p = A.g[g];\n"
               "    return p ? V3(p[i], p[A.n + i], p[2 * A.n + i]) : V3(A.gu[g][0], A.gu[g][1], A.gu[g][2]);\n}\n"
               "__device__ __forceinline__ float ld1(const GroupArgs& A, int g, int64_t i) {\n"
               "    const float* p = A.g[g];\n    return p ? p[i] : A.gu[g][0];\n}\n";
    }
};

}  // namespace

bool emit_group(const ShadingSys& ss, Group& g, std::string& src, std::string& err)
{
    Emitter e(ss, g);
    if (!e.run(src)) {
        err = e.err.empty() ? "emit failed" : e.err;
        return false;
    }
    return true;
}

// One persistent kernel for a list of groups. Work arrives as tiles
// {item, begin, end} (BatchTile), sorted by group by the host; every
// workgroup pulls the next tile from a device counter and runs the tile's
// group body over blockDim.x * ppt consecutive points (ppt strided passes,
// so every pass is a coalesced sweep of the SoA planes). The last
// workgroup to find the list empty resets the counter, so repeated
// launches (and graph replays) need no clearing.
bool emit_batch(const ShadingSys& ss, const std::vector<Group*>& groups,
                const std::vector<int>& ids, std::string& src, std::string& err)
{
    if (ids.size() != groups.size()) {
        err = "one item id per group";
        return false;
    }
    std::ostringstream k;
    k << "// Emitted by the oslmi group emitter: persistent batch kernel over " << groups.size()
      << " groups.\n";
    bool any_bb = false, any_mx = false, any_aniso = false;
    std::vector<std::string> cases;
    for (size_t j = 0; j < groups.size(); ++j) {
        const int gi = ids[j];
        Emitter e(ss, *groups[j]);
        e.lprefix = "G" + std::to_string(gi) + "_L";
        std::string code;
        if (!e.per_point(code)) {
            err = "group " + groups[j]->name + ": " + (e.err.empty() ? "emit failed" : e.err);
            return false;
        }
        any_bb = any_bb || e.uses_bb;
        any_mx = any_mx || e.uses_mx;
        any_aniso = any_aniso || e.uses_aniso;
        // a case keeps taking tiles while they belong to its group, so the
        // group's loop invariants live only inside its case
        std::ostringstream c;
        c << "    case " << gi << ": {  // group '" << groups[j]->name << "'\n";
        c << "        const GroupArgs& A = args[" << gi << "];\n";
        c << "        using Closure = LobeList<" << e.closure_capacity() << ">;\n";
        if (e.uses_cs)
            c << "        const ColorSysDev* CS = (const ColorSysDev*)A.colorsys;\n";
        c << "        for (;;) {\n";
        c << "#pragma unroll 1\n";
        c << "        for (int q_ = 0; q_ < B.ppt; ++q_) {\n";
        c << "        int64_t i = T.begin + (int64_t)q_ * blockDim.x + threadIdx.x;\n";
        c << "        if (i < T.end) {\n";
        c << code;
        c << "        }\n        }\n";
        c << "        t = fetch_tile(B.counter, &s_tile);\n";
        c << "        if (t >= ntiles) break;\n";
        c << "        T = B.tiles[B.tile_begin + t];\n";
        c << "        if (T.item != " << gi << ") break;\n";
        c << "        }\n        break;\n    }\n";
        cases.push_back(c.str());
    }
    k << Emitter::prelude(any_mx, false, any_aniso);
    k << "// next tile index for the whole workgroup (the barrier in front keeps\n"
         "// the previous index readable until every work-item has taken it)\n"
         "__device__ __forceinline__ int fetch_tile(int32_t* counter, int* s_tile) {\n"
         "    __syncthreads();\n"
         "    if (threadIdx.x == 0) *s_tile = atomicAdd(counter, 1);\n"
         "    __syncthreads();\n"
         "    return *s_tile;\n}\n";
    k << "extern \"C\" __global__ " << kernel_attributes(ss) << " void oslmi_batch_main(BatchArgs B)\n{\n";
    k << "    const GroupArgs* __restrict__ args = B.args;\n";
    k << "    __shared__ int s_tile;\n";
    if (any_bb) {
        k << "    __shared__ float s_bbl[OSLMI_BB_TABLE * 3];\n"
             "    const float* s_bb = B.bb;\n"
             "    if (B.bb_lds) {\n"
             "        for (int j = threadIdx.x; j < OSLMI_BB_TABLE * 3; j += blockDim.x) s_bbl[j] = B.bb[j];\n"
             "        s_bb = s_bbl;\n"
             "    }\n";
    }
    k << "    const int ntiles = B.tile_end - B.tile_begin;\n"
         "    int t = fetch_tile(B.counter, &s_tile);\n"
         "    while (t < ntiles) {\n"
         "        BatchTile T = B.tiles[B.tile_begin + t];\n"
         "        switch (T.item) {\n";
    for (const std::string& c : cases)
        k << c;
    k << "        default: t = fetch_tile(B.counter, &s_tile); break;\n        }\n    }\n"
         "    // one failed fetch per workgroup: the last one clears the counter\n"
         "    if (threadIdx.x == 0 && t == ntiles + (int)gridDim.x - 1) atomicExch(B.counter, 0);\n}\n";
    src = k.str();
    return true;
}

}  // namespace oslmi
