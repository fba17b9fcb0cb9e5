// xt_codegen.cpp -- line code generation.
//
// The element-program interpreter (xt_ops.h: run_element) reads every op
// word and constant of an element from memory and dispatches on the op at
// run time.  For a line with few distinct element records (a ring is made
// of a handful of magnet, drift and aperture types repeated thousands of
// times) this file writes the same op sequence as straight-line calls of
// the same element functions, with the constants as exact hexadecimal
// literals, so the device compiler sees constants instead of loads and no
// per-op dispatch remains; the per-element dispatch becomes one switch on a
// 32-bit record index.  Folded multipole-kick passes are unrolled with the
// zero pattern of their coefficients known.  The arithmetic per particle is
// unchanged: the generated code calls the functions the interpreter calls,
// with the values the interpreter would load.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/xtrack_amd.h"
#include "xt_codegen.h"
#include "xt_program.h"

namespace {

int64_t wbits(double d) { int64_t u; std::memcpy(&u, &d, 8); return u; }

// exact literal for a double
std::string lit(double v) {
    if (std::isnan(v)) return "__builtin_nan(\"\")";
    if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", v);
    std::string s(buf);
    if (s[0] == '-') return "(" + s + ")";
    return s;
}

const int kNargs[OP_NUM] = XT_OP_NARGS_INIT;

// Records whose arithmetic is at most this many ops (a kick pass counts
// its terms) and that hold no integrator, map or RF op are cheap enough
// for the per-element dispatch to matter; runs of them form segments.
constexpr int kSmallCost = 12;
constexpr size_t kSegMax = 31;           // elements per segment (< 32)
constexpr int64_t kSegMaxCount = 256;    // distinct segments per line
constexpr int64_t kSegMaxBody = 2048;    // sum of their lengths
// line code: every record literal up to this many distinct records, else
// literal only for records used more than once whose structure occurs with
// at most kLiteralPerStruct constant sets, at most kLiteralMax of them
constexpr int64_t kLiteralAll = 64;
constexpr int64_t kLiteralMax = 1024;
constexpr int64_t kLiteralPerStruct = 4;
static_assert(kSegMax <= XT_KEY_SPAN_MASK, "segment length must fit its key field");

struct Gen {
    const std::vector<double> &prog;
    const int64_t base;   // program offset of the record
    const bool loads;     // constants as loads rp[k] instead of literals
    const bool rel;       // program addresses relative to rp = prog + base
    std::string out;
    std::string err;
    int cost = 0;         // rough size of the record's arithmetic (ops, kick terms)
    bool heavy = false;   // holds an integrator, a map or an RF op

    Gen(const std::vector<double> &p, int64_t b, bool l, bool r = false)
        : prog(p), base(b), loads(l), rel(l || r) {}

    // constant at absolute program index i
    std::string c(int64_t i) const {
        if (loads) return "rp[" + std::to_string(i - base) + "]";
        return lit(prog[i]);
    }
    // address of program word i: relative to the record (rp) in loads mode
    // and in relative literal mode, so that records of the same structure
    // (or the same constants) in different places produce the same text
    std::string addr(int64_t i) const {
        if (rel) return "rp + " + std::to_string(i - base);
        return "prog + " + std::to_string(i);
    }
    std::string cs(int64_t i, int n) const {
        std::string s;
        for (int k = 0; k < n; k++) s += ", " + c(i + k);
        return s;
    }

    void line(const std::string &s) { out += "    " + s + "\n"; }

    bool in_range(int64_t pos, int64_t n) const {
        return pos >= 0 && n >= 0 && pos + n <= (int64_t)prog.size();
    }

    // one folded Horner pass from the coefficient block at `fp`
    bool kick_pass(int64_t fp, const std::string &w) {
        if (!in_range(fp, 1)) return false;
        const int64_t top = wbits(prog[fp]);
        if (top < 0 || !in_range(fp, 2 * top + 3)) return false;
        const int64_t ia = fp + 1, ib = fp + 2 + top;
        line("    {");
        line("        double dpx = " + c(ia + top) + ", dpy = " + c(ib + top) + ";");
        for (int64_t k = top - 1; k >= 0; k--) {
            line("        xt::horner_step(dpx, dpy, x, y, " + c(ia + k) + ", " + c(ib + k) + ", " +
                 (prog[ia + k] != 0.0 ? "true" : "false") + ", " +
                 (prog[ib + k] != 0.0 ? "true" : "false") + ");");
        }
        line("        xt::kick_pass_apply(p, " + w + ", dpx, dpy);");
        line("    }");
        cost += (int)top + 1;
        return true;
    }

    bool kick(int64_t pc) {
        const double *a = &prog[pc + 1];
        const std::string w = c(pc + 1);
        const int64_t desc = wbits(a[1]);
        const int64_t mask = wbits(a[2]);
        if (!in_range(desc, KD_ARR)) return false;
        line("if (!xt::kick_sane(p)) {");
        line("    xt::kick_literal_apply(p, " + w + ", (const double *)(" + addr(desc) + "));");
        line("} else {");
        line("    const double x = p.x, y = p.y;");
        for (int k = 0; k < 3; k++) {
            if (!(mask & (1 << k))) continue;
            if (!kick_pass(wbits(prog[desc + KD_FASTOFF + k]), w)) return false;
        }
        if (mask & (8 | 16 | 32)) {
            line(std::string("    xt::kick_curvature(p, x, y, ") + ((mask & 8) ? "true" : "false") +
                 ", " + ((mask & 16) ? "true" : "false") + ", " + ((mask & 32) ? "true" : "false") +
                 cs(pc + 4, 3) + ");");
        }
        line("}");
        return true;
    }

    bool record(int64_t pc) {
        for (int guard = 0; guard < 100000; guard++) {
            if (!in_range(pc, 1)) return false;
            const int64_t op = wbits(prog[pc]);
            if (op < 0 || op >= OP_NUM) return false;
            if (op == OP_END) return true;
            const int n = kNargs[op];
            if (!in_range(pc + 1, n)) return false;
            const int64_t a = pc + 1;
            const std::string at = addr(pc + 1);
            cost += 1;
            if (op == OP_POLAR || op == OP_DIPQUAD || op == OP_YOSHIDA4_BEND ||
                op == OP_BEND_CURVED || op == OP_BEND_STRAIGHT ||
                op == OP_YOSHIDA6_BEND || op == OP_RF_KICK || op == OP_X_ROT || op == OP_Y_ROT ||
                op == OP_RF_MULT)
                heavy = true;
            switch (op) {
            case OP_DRIFT_EXP: line("xt::drift_expanded(p" + cs(a, 1) + ");"); break;
            case OP_DRIFT_EXACT: line("xt::drift_exact(p" + cs(a, 1) + ");"); break;
            case OP_POLAR: line("xt::polar_drift(p" + cs(a, 6) + ");"); break;
            case OP_BEND_CURVED: line("xt::op_bend_curved(p" + cs(a, 9) + ");"); break;
            case OP_BEND_STRAIGHT: line("xt::op_bend_straight(p" + cs(a, 2) + ");"); break;
            case OP_DIPQUAD: line("xt::dipole_quad(p" + cs(a, 4) + ");"); break;
            case OP_YOSHIDA4_BEND: line("xt::op_yoshida4(p" + cs(a, 14) + ");"); break;
            case OP_YOSHIDA6_BEND: line("xt::op_yoshida6(p, " + at + ");"); break;
            case OP_KICK:
                if (!kick(pc)) return false;
                break;
            case OP_EDGE_AXAY: line("xt::op_edge_axay(p, m" + cs(a, 2) + ");"); break;
            case OP_EDGE_CLEAR_AXAY: line("xt::op_edge_clear_axay(m);"); break;
            case OP_EDGE_LINEAR: line("xt::op_edge_linear(p" + cs(a, 2) + ");"); break;
            case OP_RF_KICK:
                line("xt::op_rf_kick(p, m, " + at +
                     ", c.line_length, c.q0, c.t_sim, c.kill_energy());");
                break;
            case OP_LIMIT_RECT:
                line("xt::op_limit_rect(p" + cs(a, 4) + ", c.ignore_local());");
                break;
            case OP_LIMIT_ELLIPSE:
                line("xt::op_limit_ellipse(p" + cs(a, 3) + ", c.ignore_local());");
                break;
            case OP_GLOBAL_APER: line("xt::op_global_aper(p, c.ignore_global());"); break;
            case OP_XY_SHIFT: line("xt::op_xy_shift(p" + cs(a, 2) + ");"); break;
            case OP_S_SHIFT: line("xt::op_s_shift(p" + cs(a, 1) + ");"); break;
            case OP_X_ROT: line("xt::op_x_rot(p, m" + cs(a, 3) + ");"); break;
            case OP_Y_ROT: line("xt::op_y_rot(p, m" + cs(a, 3) + ");"); break;
            case OP_S_ROT: line("xt::op_s_rot(p" + cs(a, 2) + ");"); break;
            case OP_RF_MULT:
                // the descriptor holds program offsets: absolute addressing
                line("xt::op_rf_mult(p, m, xt::cptr(prog) + " + std::to_string(pc + 1) +
                     ", xt::cptr(prog), c.q0, c.kill_energy());");
                break;
            default: return false;
            }
            pc += 1 + n;
        }
        return false;
    }
};

}  // namespace

namespace {

/* Functions of the generated code for a set of lines (one line, or the
   lines of a batch).  A record whose literal code serves at least two
   elements of the set (every record, for a single line of at most
   kLiteralAll distinct records) gets a literal function; a record that
   serves one element only -- a magnet with its own field errors -- gets the
   function of its structure, shared with every record of that structure,
   which reads its constants from the program (scalar loads at
   rp = prog + entry[e]).  Literal functions for thousands of single-use
   records would be megabytes of instructions, streamed through the
   instruction cache once per turn. */
struct CodeSet {
    std::vector<std::string> body;             // function bodies
    std::vector<char> small;                   // cheap literal records
    bool keyed_pc = false;                     // keys carry the program offset
    std::vector<std::vector<int32_t>> segs;    // segment function bodies
};

int build_code_set(const std::vector<const std::vector<int64_t> *> &entries,
                   const std::vector<const std::vector<double> *> &progs, bool keyed_pc,
                   CodeSet &cs, std::vector<std::vector<int32_t>> &recs, std::string &err) {
    const size_t nl = entries.size();
    // distinct records of each line and their element counts
    std::vector<std::map<int64_t, int64_t>> uses(nl);
    int64_t n_distinct = 0;
    for (size_t l = 0; l < nl; l++) {
        for (int64_t e : *entries[l]) uses[l][e]++;
        n_distinct += (int64_t)uses[l].size();
    }
    const bool all_literal = !keyed_pc && nl == 1 && n_distinct <= kLiteralAll;
    cs.keyed_pc = !all_literal;
    // literal text of every record (program addresses relative to the
    // record unless every record of a single line is literal)
    std::map<std::string, int64_t> lit_uses;
    std::vector<std::map<int64_t, std::string>> lit_text(nl);
    std::map<std::string, std::pair<bool, int>> lit_info;   // heavy, cost
    std::map<std::string, std::string> struct_of_lit;
    std::map<std::string, int64_t> lits_of_struct;         // distinct literal texts
    for (size_t l = 0; l < nl; l++)
        for (auto &pu : uses[l]) {
            Gen g(*progs[l], pu.first, false, !all_literal);
            Gen gs(*progs[l], pu.first, true);
            if (!g.record(pu.first) || !gs.record(pu.first)) {
                err = "cannot generate code for line " + std::to_string(l) +
                      " at program offset " + std::to_string(pu.first);
                return XTM_ERR_ARG;
            }
            lit_uses[g.out] += pu.second;
            lit_info[g.out] = {g.heavy, g.cost};
            lit_text[l][pu.first] = g.out;
            if (struct_of_lit.emplace(g.out, gs.out).second) lits_of_struct[gs.out]++;
        }
    // A structure that the set holds with many different constants (the
    // bends of ring variants with different cell counts) gets one structure
    // function rather than a literal copy per constant set.
    std::vector<std::pair<int64_t, const std::string *>> ranked;
    for (auto &kv : lit_uses)
        if (all_literal ||
            (kv.second >= 2 && lits_of_struct[struct_of_lit[kv.first]] <= kLiteralPerStruct))
            ranked.emplace_back(kv.second, &kv.first);
    std::stable_sort(ranked.begin(), ranked.end(),
                     [](const auto &x, const auto &y) { return x.first > y.first; });
    if ((int64_t)ranked.size() > kLiteralMax) ranked.resize(kLiteralMax);
    std::map<std::string, int32_t> fn_of_lit;
    for (auto &r : ranked) {
        fn_of_lit.emplace(*r.second, (int32_t)cs.body.size());
        cs.body.push_back(*r.second);
        const auto &info = lit_info[*r.second];
        cs.small.push_back(!info.first && info.second <= kSmallCost);
    }
    // the other records: structure functions
    std::map<std::string, int32_t> fn_of_struct;
    recs.assign(nl, {});
    for (size_t l = 0; l < nl; l++) {
        std::map<int64_t, int32_t> fn_of_pc;
        for (auto &pu : uses[l]) {
            auto it = fn_of_lit.find(lit_text[l][pu.first]);
            if (it != fn_of_lit.end()) {
                fn_of_pc.emplace(pu.first, it->second);
                continue;
            }
            Gen g(*progs[l], pu.first, true);
            if (!g.record(pu.first)) {
                err = "cannot generate code for line " + std::to_string(l) +
                      " at program offset " + std::to_string(pu.first);
                return XTM_ERR_ARG;
            }
            auto ins = fn_of_struct.emplace(g.out, (int32_t)cs.body.size());
            if (ins.second) {
                cs.body.push_back(g.out);
                cs.small.push_back(0);
            }
            fn_of_pc.emplace(pu.first, ins.first->second);
        }
        if ((int64_t)cs.body.size() > XT_CODEGEN_MAX_RECORDS) {
            err = "the code needs more than " + std::to_string(XT_CODEGEN_MAX_RECORDS) +
                  " element functions";
            return XTM_ERR_UNSUPPORTED;
        }
        const std::vector<int64_t> &entry = *entries[l];
        recs[l].resize(entry.size());
        for (size_t e = 0; e < entry.size(); e++) recs[l][e] = fn_of_pc[entry[e]];
    }

    // Segments: runs of consecutive cheap literal records (drifts,
    // apertures, thin kicks, markers) cut into chunks of at most kSegMax
    // elements; each chunk sequence that occurs at least twice in the set
    // becomes one function that tracks the chunk's elements in order and
    // returns the position of the element where the particle was lost.  The
    // key of a chunk's first element carries the segment number and the
    // chunk length; the loop runs the chunk with one dispatch instead of one
    // per element.  Same element calls, same order: the arithmetic is that
    // of the single-record path.  (Cheap records are literal: their code
    // reads no constants through rp, which the segment call passes for the
    // chunk's first element only.)
    std::vector<std::tuple<size_t, size_t, size_t>> chunks;   // (line, first element, length)
    for (size_t l = 0; l < nl; l++) {
        const std::vector<int32_t> &rv = recs[l];
        size_t e = 0;
        while (e < rv.size()) {
            if (!cs.small[rv[e]]) { e++; continue; }
            size_t f = e;
            while (f < rv.size() && cs.small[rv[f]]) f++;
            for (size_t a = e; a < f; a += kSegMax) {
                const size_t n = std::min<size_t>(kSegMax, f - a);
                if (n > 1) chunks.emplace_back(l, a, n);
            }
            e = f;
        }
    }
    auto chunk_seq = [&](const std::tuple<size_t, size_t, size_t> &ch) {
        const std::vector<int32_t> &rv = recs[std::get<0>(ch)];
        return std::vector<int32_t>(rv.begin() + std::get<1>(ch),
                                    rv.begin() + std::get<1>(ch) + std::get<2>(ch));
    };
    std::map<std::vector<int32_t>, int64_t> count;
    for (auto &ch : chunks) count[chunk_seq(ch)]++;
    std::vector<std::pair<int64_t, const std::vector<int32_t> *>> seg_rank;
    for (auto &kv : count)
        if (kv.second >= 2)
            seg_rank.emplace_back(kv.second * (int64_t)(kv.first.size() - 1), &kv.first);
    std::stable_sort(seg_rank.begin(), seg_rank.end(),
                     [](const auto &x, const auto &y) { return x.first > y.first; });
    std::map<std::vector<int32_t>, int32_t> seg_id;
    int64_t seg_body = 0;
    for (auto &r : seg_rank) {
        if ((int64_t)cs.segs.size() >= kSegMaxCount ||
            seg_body + (int64_t)r.second->size() > kSegMaxBody)
            continue;
        seg_id.emplace(*r.second, (int32_t)cs.segs.size());
        cs.segs.push_back(*r.second);
        seg_body += (int64_t)r.second->size();
    }
    for (auto &ch : chunks) {
        auto it = seg_id.find(chunk_seq(ch));
        if (it == seg_id.end()) continue;
        recs[std::get<0>(ch)][std::get<1>(ch)] |= ((it->second + 1) << XT_KEY_SEG_SHIFT) |
                                                  ((int32_t)std::get<2>(ch) << XT_KEY_SPAN_SHIFT);
    }
    return XTM_OK;
}

/* The functions, segments and element runner of a code set, in namespace
   xt_gen. */
std::string emit_code_set(const CodeSet &cs, uint32_t options, const char *what) {
    std::string s;
    s += std::string("// generated ") + what + ": element records as straight-line calls\n";
    if (options & XTM_SPEC_FAST_ARITH) s += "#define XTM_ARITH_FAST 1\n";
    s += "#if defined(__HIPCC__) && !defined(__HIPCC_RTC__)\n#include <hip/hip_runtime.h>\n#endif\n";
    s += "#include \"xt_ops.h\"\n\n";
    s += "#if defined(__HIPCC__)\n#define XT_GEN_FN static __host__ __device__ __forceinline__\n";
    s += "#else\n#define XT_GEN_FN static inline\n#endif\n\n";
    s += "namespace xt_gen {\n";
    for (size_t r = 0; r < cs.body.size(); r++) {
        s += "XT_GEN_FN void rec_" + std::to_string(r) +
             "(xt::PState<double> &p, const xt::PRef &m, const double *__restrict__ prog,\n"
             "    xt::CProg rp, const xt::TrackCtx &c) {\n"
             "    (void)m; (void)prog; (void)rp; (void)c;\n";
        s += cs.body[r];
        s += "}\n\n";
    }
    // a segment's records that address the program (relative literal code
    // of thin kicks: the literal kick's descriptor) get their own element's
    // rp from ent, the entry offsets from the chunk's first element on
    for (size_t q = 0; q < cs.segs.size(); q++) {
        s += "XT_GEN_FN int32_t seg_" + std::to_string(q) +
             "(xt::PState<double> &p, const xt::PRef &m, const double *__restrict__ prog,\n"
             "    xt::CProg rp, const xt::TrackCtx &c, xt::CP<int64_t> ent) {\n"
             "    (void)ent;\n";
        const auto &v = cs.segs[q];
        for (size_t j = 0; j < v.size(); j++) {
            const std::string &b = cs.body[v[j]];
            const bool own_rp = cs.keyed_pc && (b.find("rp[") != std::string::npos ||
                                                b.find("rp + ") != std::string::npos);
            s += "    rec_" + std::to_string(v[j]) + "(p, m, prog, " +
                 (own_rp ? "xt::cptr(prog) + ent[" + std::to_string(j) + "]" : std::string("rp")) +
                 ", c);";
            s += j + 1 < v.size() ? " if (p.state <= 0) return " + std::to_string(j) + ";\n"
                                  : " return " + std::to_string(j) + ";\n";
        }
        s += "}\n\n";
    }
    // element key: the function index with the segment fields (int32), and
    // with relative addressing also the element's program offset
    const std::string kr = cs.keyed_pc ? "xt::uniform(k.r)" : "xt::uniform(k)";
    s += "struct Runner {\n"
         "    const int32_t *__restrict__ rec;\n"
         "    const int64_t *__restrict__ entry;\n"
         "    const double *__restrict__ prog;\n";
    if (cs.keyed_pc)
        s += "    struct Key { int32_t r; int64_t pc; };\n"
             "    XT_HD Key key(int64_t e) const { return Key{xt::cptr(rec)[e], xt::cptr(entry)[e]}; }\n";
    else
        s += "    using Key = int32_t;\n"
             "    XT_HD Key key(int64_t e) const { return xt::cptr(rec)[e]; }\n";
    const std::string rp = cs.keyed_pc
                               ? "        const xt::CProg rp = xt::cptr(prog) + xt::uniform(k.pc);\n"
                               : "        const xt::CProg rp = xt::cptr(prog);\n";
    s += "    static constexpr bool kSegments = true;\n"
         "    XT_HD static int32_t span(Key k) { return (" + kr + " >> " +
         std::to_string(XT_KEY_SPAN_SHIFT) + ") & " + std::to_string(XT_KEY_SPAN_MASK) +
         "; }\n"
         "    XT_HD int32_t segment(xt::PState<double> &p, const xt::PRef &m, Key k,\n"
         "                          const xt::TrackCtx &c, int64_t e) const {\n" + rp +
         "        const xt::CP<int64_t> ent = xt::cptr(entry) + xt::uniform(e);\n"
         "        switch (" + kr + " >> " + std::to_string(XT_KEY_SEG_SHIFT) + ") {\n";
    for (size_t q = 0; q < cs.segs.size(); q++)
        s += "        case " + std::to_string(q + 1) + ": return seg_" + std::to_string(q) +
             "(p, m, prog, rp, c, ent);\n";
    s += "        default: p.state = -999; return 0;\n"
         "        }\n"
         "    }\n"
         "    XT_HD void operator()(xt::PState<double> &p, const xt::PRef &m, Key k,\n"
         "                          const xt::TrackCtx &c) const {\n" + rp +
         "        switch (" + kr + " & " + std::to_string(XT_KEY_REC_MASK) + ") {\n";
    for (size_t r = 0; r < cs.body.size(); r++)
        s += "        case " + std::to_string(r) + ": rec_" + std::to_string(r) +
             "(p, m, prog, rp, c); break;\n";
    s += "        default: p.state = -999; break;\n"
         "        }\n"
         "    }\n"
         "};\n";
    return s;
}

}  // namespace

int xt_generate_line_source(const std::vector<int64_t> &entry, const std::vector<double> &prog,
                            int32_t waves_per_eu, uint32_t options, std::string &src,
                            std::vector<int32_t> &rec_of_elem, std::string &err) {
    CodeSet cs;
    std::vector<std::vector<int32_t>> recs;
    const int st = build_code_set({&entry}, {&prog}, false, cs, recs, err);
    if (st != XTM_OK) return st;
    rec_of_elem = std::move(recs[0]);
    std::string s = emit_code_set(cs, options, "line code");
    s += "}  // namespace xt_gen\n\n";
    s += "#if defined(__HIPCC__)\n";
    char occ[128];
    std::snprintf(occ, sizeof occ, "__attribute__((amdgpu_waves_per_eu(%d)))",
                  waves_per_eu > 0 ? waves_per_eu : 1);
    s += std::string("extern \"C\" __global__ __launch_bounds__(256) ") +
         (waves_per_eu > 0 ? occ : "") +
         "\nvoid xtm_line_kernel(const int32_t *__restrict__ rec, const int64_t *__restrict__ entry,\n"
         "    const double *__restrict__ prog, double line_length, xtm_particles P,\n"
         "    int32_t num_turns, int32_t ele_start, int32_t num_ele_track,\n"
         "    int32_t flag_end_turn_actions, int32_t flag_reset_s, int32_t flag_monitor,\n"
         "    const xtm_monitor *__restrict__ Md, const xtm_particles *__restrict__ Pd,\n"
         "    uint64_t flags) {\n"
         "    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;\n"
         "    if (i >= P.capacity) return;\n"
         "    const xt::TrackCtx c{flags, line_length, P.q0, P.t_sim};\n"
         "    // chi == 1 particles; the interpreter kernel tracks the others\n"
         "    xt::track_particle_chi1(P, i, xt_gen::Runner{rec, entry, prog}, c, num_turns,\n"
         "                            ele_start, num_ele_track, flag_end_turn_actions,\n"
         "                            flag_reset_s, flag_monitor, xt::MonRef{Md, Pd});\n"
         "}\n"
         "#else\n"
         "// host build (tests): the same per-particle body over all particles\n"
         "extern \"C\" void xtm_line_host(const int32_t *rec, const int64_t *entry, const double *prog,\n"
         "    double line_length, const xtm_particles *P, int32_t num_turns, int32_t ele_start,\n"
         "    int32_t num_ele_track, int32_t flag_end_turn_actions, int32_t flag_reset_s,\n"
         "    int32_t flag_monitor, const xtm_monitor *M, uint64_t flags) {\n"
         "    xtm_monitor m0 = {};\n"
         "    const xt::TrackCtx c{flags, line_length, P->q0, P->t_sim};\n"
         "    const xt::MonRef mon{M ? M : &m0, P};\n"
         "    for (int64_t i = 0; i < P->capacity; i++)\n"
         "        xt::track_particle_with<false>(*P, i, xt_gen::Runner{rec, entry, prog},\n"
         "                                xt::ProgramRunner{entry, prog}, c, num_turns,\n"
         "                                ele_start, num_ele_track, flag_end_turn_actions,\n"
         "                                flag_reset_s, flag_monitor, mon);\n"
         "}\n"
         "#endif\n";
    src = std::move(s);
    return XTM_OK;
}


int xt_generate_batch_source(const std::vector<const std::vector<int64_t> *> &entries,
                             const std::vector<const std::vector<double> *> &progs,
                             int32_t waves_per_eu, std::string &src,
                             std::vector<std::vector<int32_t>> &recs, std::string &err) {
    CodeSet cs;
    // the batch kernel's runner reads each element's program offset
    const int st = build_code_set(entries, progs, true, cs, recs, err);
    if (st != XTM_OK) return st;
    std::string s = emit_code_set(cs, 0, "batch code");
    s += "/* one block of a batch: particles [first, first + count) of line `line` */\n"
         "struct BatchBlock { int64_t first; int32_t line; int32_t count; };\n"
         "}  // namespace xt_gen\n\n";
    char occ[128];
    std::snprintf(occ, sizeof occ, "__attribute__((amdgpu_waves_per_eu(%d)))",
                  waves_per_eu > 0 ? waves_per_eu : 1);
    s += "#if defined(__HIPCC__)\n";
    s += std::string("extern \"C\" __global__ __launch_bounds__(256) ") +
         (waves_per_eu > 0 ? occ : "") +
         "\nvoid xtm_batch_line_kernel(const xt_gen::BatchBlock *__restrict__ blocks,\n"
         "    const int32_t *const *__restrict__ recs, const int64_t *const *__restrict__ entries,\n"
         "    const double *const *__restrict__ progs, const int64_t *__restrict__ nele,\n"
         "    const double *__restrict__ lengths, const xtm_particles *__restrict__ Pd,\n"
         "    int32_t num_turns, uint64_t flags) {\n"
         "    const xt_gen::BatchBlock b = xt::cptr(blocks)[blockIdx.x];\n"
         "    if ((int32_t)threadIdx.x >= b.count) return;\n"
         "    const xtm_particles &P = *Pd;\n"
         "    const int64_t i = b.first + threadIdx.x;\n"
         "    const int64_t ne = xt::cptr(nele)[b.line];\n"
         "    const xt::TrackCtx c{flags, xt::cptr(lengths)[b.line], P.q0, P.t_sim, ne};\n"
         "    // chi == 1 particles; the interpreter batch kernel tracks the others\n"
         "    xt::track_particle_chi1(P, i, xt_gen::Runner{xt::cptr(recs)[b.line],\n"
         "                            xt::cptr(entries)[b.line], xt::cptr(progs)[b.line]}, c,\n"
         "                            num_turns, 0, (int32_t)ne, 1, 1, 0,\n"
         "                            xt::MonRef{nullptr, Pd});\n"
         "}\n"
         "#else\n"
         "// host build (tests): the same per-particle body over the blocks\n"
         "extern \"C\" void xtm_batch_host(const xt_gen::BatchBlock *blocks, int64_t num_blocks,\n"
         "    const int32_t *const *recs, const int64_t *const *entries, const double *const *progs,\n"
         "    const int64_t *nele, const double *lengths, const xtm_particles *P,\n"
         "    int32_t num_turns, uint64_t flags) {\n"
         "    for (int64_t bi = 0; bi < num_blocks; bi++) {\n"
         "        const xt_gen::BatchBlock b = blocks[bi];\n"
         "        const int64_t ne = nele[b.line];\n"
         "        const xt::TrackCtx c{flags, lengths[b.line], P->q0, P->t_sim, ne};\n"
         "        for (int32_t t = 0; t < b.count; t++)\n"
         "            xt::track_particle_with<false>(*P, b.first + t,\n"
         "                xt_gen::Runner{recs[b.line], entries[b.line], progs[b.line]},\n"
         "                xt::ProgramRunner{entries[b.line], progs[b.line]}, c, num_turns, 0,\n"
         "                (int32_t)ne, 1, 1, 0, xt::MonRef{nullptr, P});\n"
         "    }\n"
         "}\n"
         "#endif\n";
    src = std::move(s);
    return XTM_OK;
}
