// xt_codegen.h -- per-line HIP source generated from the element programs.
#ifndef XT_CODEGEN_H
#define XT_CODEGEN_H
#include <cstdint>
#include <string>
#include <vector>

// Largest number of distinct element records for which line code is
// generated; lines with more run on the program interpreter.
constexpr int64_t XT_CODEGEN_MAX_RECORDS = 2048;

// Element keys of generated line code (32 bits): record index in bits
// 0-10, segment length in bits 11-15 (0: not the start of a segment) and
// segment number + 1 in bits 16-30 (0: none).
constexpr int32_t XT_KEY_REC_MASK = 0x7ff;
constexpr int32_t XT_KEY_SPAN_SHIFT = 11;
constexpr int32_t XT_KEY_SPAN_MASK = 31;
constexpr int32_t XT_KEY_SEG_SHIFT = 16;
static_assert(XT_CODEGEN_MAX_RECORDS <= XT_KEY_REC_MASK + 1, "record index must fit its key field");

// Source of a translation unit that tracks the line: one function per
// distinct element record with the record's constants as literals, an
// element runner switching on the record index of each element, and (when
// compiled as HIP) the kernel `xtm_line_kernel`, whose parameters are those
// of the interpreter kernel with the entry array replaced by `rec`.  The
// source includes xt_ops.h.  `rec_of_elem[e]` receives the key of element
// e (its record index and, at the start of a segment of cheap records, the
// segment fields above).  `waves_per_eu` (> 0) sets the kernel's occupancy target.
// `options`: XTM_SPEC_* bits (XTM_SPEC_FAST_ARITH compiles the element maps
// with XTM_ARITH_FAST defined).
// Returns an XTM_* status and sets `err` on failure.
int xt_generate_line_source(const std::vector<int64_t> &entry, const std::vector<double> &prog,
                            int32_t waves_per_eu, uint32_t options, std::string &src,
                            std::vector<int32_t> &rec_of_elem, std::string &err);

// Source of a translation unit that tracks a batch of lines in one launch:
// one function per distinct element structure (the record code with its
// constants read from the line's program at rp = prog + entry[e], so that
// the records of all lines that differ only in their constants share one
// function), an element runner keyed by (structure, program offset), and
// the kernel `xtm_batch_line_kernel` over a block table (one line per
// block).  recs[l][e] receives the structure index of element e of line l.
int xt_generate_batch_source(const std::vector<const std::vector<int64_t> *> &entries,
                             const std::vector<const std::vector<double> *> &progs,
                             int32_t waves_per_eu, std::string &src,
                             std::vector<std::vector<int32_t>> &recs, std::string &err);
#endif
