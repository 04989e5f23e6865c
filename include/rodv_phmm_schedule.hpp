// rodv_phmm_schedule.hpp -- roDV's batch call as librodv_phmm.so exports it:
// the declaration of pairhmm::schedule::schedule_pairhmm, identical to the
// reference's (src/core/pairhmm_schedule.h:14-25, same namespace, types and
// defaults).  The definition is compiled into the library
// (rodv_amd/csrc/phmm_schedule.cpp), so a caller compiled against the
// reference's own header links against librodv_phmm.so unchanged; new
// callers may include this header instead.
//
// Contract kept from pairhmm_schedule.cpp:116-340:
//  - returns false when there are no haplotypes or no reads;
//  - result is resized to M rows (haplotypes) of N log10 likelihoods;
//  - use_double / idle ratios keep their meaning: the engine routes pairs
//    between the inter and intra arithmetic exactly as the reference's
//    greedy grouping does, with the group widths of this host's CPU
//    (RODV_PHMM_SEMANTICS_SCHEDULE, the default of rodv_phmm_config_default).
// Differences: invalid input (bases outside ACGTN, zero lengths, quality
// strings of the wrong length) returns false instead of undefined behaviour,
// nothing is printed unless `verbose`, and concurrent calls are safe (they
// take turns on the process's engine context, on device 0).
#ifndef RODV_PHMM_SCHEDULE_HPP_
#define RODV_PHMM_SCHEDULE_HPP_

#include <cstdint>
#include <vector>

namespace pairhmm {
namespace schedule {

bool schedule_pairhmm(const std::vector<std::vector<uint8_t>> &haplotypes,
                      const std::vector<std::vector<uint8_t>> &reads,
                      std::vector<std::vector<double>> &result,
                      const std::vector<std::vector<uint8_t>> &quality,
                      const std::vector<std::vector<uint8_t>> &insertion_qualities,
                      const std::vector<std::vector<uint8_t>> &deletion_qualities,
                      const std::vector<std::vector<uint8_t>> &gap_contiguous_qualities,
                      bool use_double = false, double max_idle_ratio_float = 0.1,
                      double max_idle_ratio_double = 0.1, bool verbose = false);

}  // namespace schedule
}  // namespace pairhmm

#endif  // RODV_PHMM_SCHEDULE_HPP_
