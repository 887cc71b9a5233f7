/*
 * result_sets.h — the result repository's internal interface, for producers
 * inside the library (the serving loop's hand-off of finished queries).
 */
#ifndef BLZ_RESULT_SETS_H_
#define BLZ_RESULT_SETS_H_

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/blz/gdf_types.h"

namespace blz {

// update_token (ResultSetRepository.cpp:45-112): parks `columns` (named `names`)
// under `token` and wakes its readers. `release` frees the frame's device memory
// when the result is freed. Throws std::runtime_error("Token does not exist") for a
// token never registered, after calling `release`.
void rs_park(uint64_t token, std::vector<gdf_column> columns, std::vector<std::string> names, double duration_ms,
             const std::string &error, std::function<void()> release);

// Throws std::runtime_error("Token does not exist") unless `token` is registered.
void rs_require_token(uint64_t token);

}  // namespace blz
#endif
