// OSO (compiled OSL) reader: text assembly -> shader master (symbols + ops).
//
// Follows the grammar of the reference's reader (src/liboslexec/osogram.y,
// osolex.l): a version line, a shader declaration, symbol lines
// "symtype typespec [arraylen] name defaults... %hints", "code <name>"
// section markers and op lines "opname args... jumptargets... %hints".
// Hints are skipped except %read/%write which are not needed here.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace oslmi {

enum BaseType { kInt, kFloat, kTriple, kString, kMatrix, kClosure, kVoid, kStruct };
enum SymKind { kParam, kOParam, kGlobal, kLocal, kTemp, kConst };

struct OType {
    BaseType base = kVoid;
    std::string triple_kind;  // color/point/vector/normal for kTriple
    int arraylen = 0;         // 0 scalar, >0 fixed array, -1 unsized
    bool is_float_based() const { return base == kFloat || base == kTriple; }
};

struct OSymbol {
    SymKind kind;
    OType type;
    std::string name;
    std::vector<double> fdefaults;     // numeric defaults
    std::vector<std::string> sdefaults;
    int init_begin = -1, init_end = -1;  // "code <param>" init ops
};

struct OOp {
    std::string name;
    std::vector<int> args;  // symbol indices
    int jump[4]             = { -1, -1, -1, -1 };
    int line                = 0;
};

struct ShaderMaster {
    std::string name;
    std::string shadertype;  // shader, surface, displacement, volume
    std::vector<OSymbol> syms;
    std::vector<OOp> ops;
    int main_begin = 0;  // first op of ___main___
    int find(const std::string& n) const;
};

// Parses OSO text; returns false and sets err on a syntax error.
bool parse_oso(const std::string& text, ShaderMaster& out, std::string& err);

}  // namespace oslmi
