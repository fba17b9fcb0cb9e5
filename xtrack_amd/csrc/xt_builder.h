// xt_builder.h -- host builder of device op programs (see xt_program.h).
#ifndef XT_BUILDER_H
#define XT_BUILDER_H
#include <cstdint>
#include <string>
#include <vector>

// Compile the raw element table into `entry` (per-element program offsets)
// and `prog`; identical raw records share one program.  With `entry_bt`,
// also the backtracking programs (XS_FLAG_BACKTRACK) of every element, in
// the same `prog`.  Returns an XTM_* status and sets `err` on failure.
int xt_build_program(const int32_t *types, const int64_t *offsets, const double *params,
                     int64_t n, std::vector<int64_t> &entry, std::vector<double> &prog,
                     std::string &err, std::vector<int64_t> *entry_bt = nullptr);

// Number of distinct programs referenced by `entry`.
int64_t xt_count_records(const std::vector<int64_t> &entry);
#endif
