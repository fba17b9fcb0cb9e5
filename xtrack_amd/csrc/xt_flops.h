// xt_flops.h -- operation count of one turn of the kernel for one particle.
#ifndef XT_FLOPS_H
#define XT_FLOPS_H
#include <cstdint>
#include <vector>
// out: add/sub, mul, div, sqrt, elementary-function calls, total arithmetic
int xt_count_turn(const std::vector<int64_t> &entry, const std::vector<double> &prog,
                  double line_length, const double *coords, double p0c, double mass0,
                  double q0, double out[6]);
#endif
