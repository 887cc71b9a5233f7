/*
 * result_sets.cpp — the asynchronous result hand-off: a process-wide repository
 * of result frames keyed by query token, grouped by connection, as
 * result_set_repository (src/ResultSetRepository.cpp, .h) keeps them.
 *
 * A query is registered under a token (register_query, :36-44); whoever runs it
 * parks its frame, run time and error text there (update_token, :45-112), which
 * wakes every reader blocked in get_result (:186-205). A reader counts as a
 * reference until try_free_result (:166-184) drops it; the last reference frees
 * the frame's device buffers. Every parked column also gets a column token
 * (:94-102) under which get_column (:207-226) returns it.
 *
 * Device memory of a parked frame is owned by the repository: buffers the caller
 * allocated are freed with hipFree, a served query's outputs by releasing its
 * server ticket (which also returns its HBM budget). The repository itself calls
 * no HIP function except those frees, so its bookkeeping runs without a device.
 *
 * All maps are guarded by one mutex (the reference reads some of them unguarded).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/blz/blz_ral.h"
#include "result_sets.h"

namespace blz {
namespace {

struct ResultSet {
  bool is_ready = false;
  std::vector<gdf_column> columns;
  std::vector<std::string> names;
  std::vector<uint64_t> column_tokens;
  double duration_ms = 0;
  std::string error;
  size_t ref_counter = 0;
  std::function<void()> release;  // frees the frame's device memory
};

struct Repository {
  std::mutex mu;
  std::condition_variable cv;
  std::map<uint64_t, ResultSet> result_sets;
  std::map<uint64_t, std::vector<uint64_t>> connection_result_sets;
  std::map<uint64_t, std::pair<uint64_t, int>> precalculated_columns;  // column token -> (query token, column)
  std::mt19937_64 gen{std::random_device{}()};
};

Repository &repo() {
  static Repository r;
  return r;
}

uint64_t random_token(Repository &r) {
  std::uniform_int_distribution<uint64_t> dis(std::numeric_limits<uint64_t>::min(),
                                               std::numeric_limits<uint64_t>::max());
  return dis(r.gen);
}

// frees the parked frame (ResultSetRepository.cpp:135-150), caller holds the lock
void free_locked(Repository &r, uint64_t connection, uint64_t token) {
  auto &tokens = r.connection_result_sets[connection];
  tokens.erase(std::remove(tokens.begin(), tokens.end(), token), tokens.end());
  auto it = r.result_sets.find(token);
  if (it == r.result_sets.end()) return;
  for (uint64_t ct : it->second.column_tokens) r.precalculated_columns.erase(ct);
  if (it->second.release) it->second.release();
  r.result_sets.erase(it);
}

// hipFree of every distinct buffer of a frame (columns may share one: the
// reference deduplicates the frame first, ResultSetRepository.cpp:59-60)
std::function<void()> owned_release(const std::vector<gdf_column> &cols) {
  std::set<void *> bufs;
  for (const auto &c : cols) {
    if (c.data) bufs.insert(c.data);
    if (c.valid) bufs.insert(c.valid);
  }
  return [bufs] {
    for (void *p : bufs) (void)hipFree(p);
  };
}

}  // namespace

void rs_park(uint64_t token, std::vector<gdf_column> columns, std::vector<std::string> names, double duration_ms,
             const std::string &error, std::function<void()> release) {
  Repository &r = repo();
  {
    std::lock_guard<std::mutex> g(r.mu);
    auto it = r.result_sets.find(token);
    if (it == r.result_sets.end()) {
      if (release) release();
      throw std::runtime_error("Token does not exist");
    }
    ResultSet &rs = it->second;
    if (rs.release) rs.release();  // a token updated twice keeps the later frame
    for (uint64_t ct : rs.column_tokens) r.precalculated_columns.erase(ct);
    rs.columns = std::move(columns);
    rs.names = std::move(names);
    rs.names.resize(rs.columns.size());
    rs.column_tokens.clear();
    for (size_t i = 0; i < rs.columns.size(); i++) {
      uint64_t ct;
      do ct = random_token(r);
      while (ct == 0 || r.precalculated_columns.count(ct));
      r.precalculated_columns[ct] = {token, (int)i};
      rs.column_tokens.push_back(ct);
    }
    rs.duration_ms = duration_ms;
    rs.error = error;
    rs.release = std::move(release);
    rs.is_ready = true;
  }
  r.cv.notify_all();
}

void rs_require_token(uint64_t token) {
  Repository &r = repo();
  std::lock_guard<std::mutex> g(r.mu);
  if (!r.result_sets.count(token)) throw std::runtime_error("Token does not exist");
}

}  // namespace blz

using namespace blz;

namespace {

int rs_guard(char *err, size_t len, const std::function<void()> &f) {
  auto put = [&](const std::string &m) {
    if (err && len > 0) {
      const size_t n = std::min(len - 1, m.size());
      std::memcpy(err, m.data(), n);
      err[n] = 0;
    }
  };
  try {
    f();
    put("");
    return BLZ_OK;
  } catch (const std::bad_alloc &) {
    put("host out of memory");
    return BLZ_ERR_OOM;
  } catch (const std::exception &e) {
    put(e.what());
    return BLZ_ERR_INVALID;
  }
}

void copy_text(char *dst, size_t len, const std::string &s) {
  if (!dst || len == 0) return;
  const size_t n = std::min(len - 1, s.size());
  std::memcpy(dst, s.data(), n);
  dst[n] = 0;
}

}  // namespace

extern "C" {

int blz_rs_init_session(uint64_t *connection, char *err, size_t err_len) {
  return rs_guard(err, err_len, [&] {
    if (!connection) throw std::runtime_error("connection is NULL");
    Repository &r = repo();
    std::lock_guard<std::mutex> g(r.mu);
    const uint64_t session = random_token(r);
    if (r.connection_result_sets.count(session)) throw std::runtime_error("Connection already exists");
    r.connection_result_sets[session] = {};
    *connection = session;
  });
}

int blz_rs_register_query(uint64_t connection, uint64_t token, char *err, size_t err_len) {
  return rs_guard(err, err_len, [&] {
    Repository &r = repo();
    std::lock_guard<std::mutex> g(r.mu);
    ResultSet &old = r.result_sets[token];
    if (old.release) old.release();  // re-registering a token drops its old frame
    for (uint64_t ct : old.column_tokens) r.precalculated_columns.erase(ct);
    old = ResultSet{};
    r.connection_result_sets[connection].push_back(token);
  });
}

int blz_rs_update_token(uint64_t token, const gdf_column *columns, const char *const *names, int n_columns,
                        double duration_ms, const char *error_message, char *err, size_t err_len) {
  return rs_guard(err, err_len, [&] {
    if (n_columns < 0 || (n_columns > 0 && !columns)) throw std::runtime_error("bad arguments");
    std::vector<gdf_column> cols(columns, columns + n_columns);
    std::vector<std::string> nm;
    for (int i = 0; i < n_columns; i++) nm.push_back(names && names[i] ? names[i] : "");
    auto rel = owned_release(cols);
    rs_require_token(token);  // the frame stays the caller's when the token is unknown
    rs_park(token, std::move(cols), std::move(nm), duration_ms, error_message ? error_message : "", std::move(rel));
  });
}

int blz_rs_get_result(uint64_t connection, uint64_t token, gdf_column *columns, uint64_t *column_tokens,
                      int max_columns, int *n_columns, double *duration_ms, char *error_message,
                      size_t error_len, char *err, size_t err_len) {
  return rs_guard(err, err_len, [&] {
    Repository &r = repo();
    std::unique_lock<std::mutex> lk(r.mu);
    if (!r.connection_result_sets.count(connection)) throw std::runtime_error("Connection does not exist");
    if (!r.result_sets.count(token)) throw std::runtime_error("Result set does not exist");
    // a reference stays valid while other tokens are inserted; a free of this token
    // during the wait is noticed by looking it up again
    r.cv.wait(lk, [&] {
      auto it = r.result_sets.find(token);
      return it == r.result_sets.end() || it->second.is_ready;
    });
    auto it = r.result_sets.find(token);
    if (it == r.result_sets.end()) throw std::runtime_error("Result set does not exist");
    ResultSet &rs = it->second;
    const int n = (int)rs.columns.size();
    if (columns && n > max_columns) throw std::runtime_error("result has more columns than max_columns");
    rs.ref_counter++;
    if (n_columns) *n_columns = n;
    for (int i = 0; columns && i < n; i++) columns[i] = rs.columns[i];
    for (int i = 0; column_tokens && i < n && i < max_columns; i++) column_tokens[i] = rs.column_tokens[i];
    if (duration_ms) *duration_ms = rs.duration_ms;
    copy_text(error_message, error_len, rs.error);
  });
}

int blz_rs_column_name(uint64_t token, int column, char *name, size_t name_len) {
  Repository &r = repo();
  std::lock_guard<std::mutex> g(r.mu);
  auto it = r.result_sets.find(token);
  if (it == r.result_sets.end() || column < 0 || column >= (int)it->second.names.size()) return BLZ_ERR_INVALID;
  copy_text(name, name_len, it->second.names[column]);
  return BLZ_OK;
}

int blz_rs_try_free_result(uint64_t connection, uint64_t token, int *found, char *err, size_t err_len) {
  return rs_guard(err, err_len, [&] {
    Repository &r = repo();
    std::lock_guard<std::mutex> g(r.mu);
    if (!r.connection_result_sets.count(connection)) throw std::runtime_error("Connection does not exist");
    auto it = r.result_sets.find(token);
    if (found) *found = it != r.result_sets.end();
    if (it == r.result_sets.end()) return;
    // the last reference frees; a result never read (no reference) is freed too
    if (it->second.ref_counter <= 1) free_locked(r, connection, token);
    else it->second.ref_counter--;
  });
}

int blz_rs_remove_all_connection_tokens(uint64_t connection) {
  Repository &r = repo();
  std::lock_guard<std::mutex> g(r.mu);
  auto ct = r.connection_result_sets.find(connection);
  if (ct == r.connection_result_sets.end()) return BLZ_OK;
  const std::vector<uint64_t> tokens = ct->second;
  for (uint64_t token : tokens) {
    auto it = r.result_sets.find(token);
    if (it != r.result_sets.end() && it->second.ref_counter == 0) free_locked(r, connection, token);
  }
  r.connection_result_sets.erase(connection);
  r.cv.notify_all();
  return BLZ_OK;
}

int blz_rs_get_column(uint64_t connection, uint64_t column_token, gdf_column *out, char *err, size_t err_len) {
  return rs_guard(err, err_len, [&] {
    if (!out) throw std::runtime_error("out is NULL");
    Repository &r = repo();
    std::lock_guard<std::mutex> g(r.mu);
    if (!r.connection_result_sets.count(connection)) throw std::runtime_error("Connection does not exist");
    auto it = r.precalculated_columns.find(column_token);
    if (it == r.precalculated_columns.end()) throw std::runtime_error("Column does not exist");
    *out = r.result_sets.at(it->second.first).columns[it->second.second];
  });
}

}  // extern "C"
