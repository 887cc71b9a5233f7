// The reference's C++ call sequence for a filter + project query (as in
// CalciteInterpreter.cpp:706-721 and evaluate_query.cu), through include/blz/ral.hpp.
// Usage: ral_shim_test N      (exit 0; prints the results as text)
//        ral_shim_test errors (no device needed: shows the error behaviour)
// Column i of the frame is generated from the row index alone, so the caller can
// rebuild the same frame for the oracle:
//   a[i] = 1 + (i * 7919) % 50                      int32
//   b[i] = 900 + (i * 104729) % 104100              float64
//   c[i] = ((i * 31) % 11) / 100.0                  float64 (0.00 .. 0.10)
//   d[i] = (i * 40503) % 2556                       int32
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "blz/ral.hpp"

static const char *kFilter =
    "LogicalFilter(condition=[AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30))])";
static const char *kProject = "LogicalProject(EXPR$0=[*($1, -(1, $2))], EXPR$1=[+($0, 1)], D=[$3])";

template <typename T>
std::vector<T> download(gdf_column_cpp &c) {
  std::vector<T> h(c.size());
  if (!h.empty() && blz_memcpy_dtoh(h.data(), c.data(), h.size() * sizeof(T), nullptr) != BLZ_OK)
    throw std::runtime_error("copy to the host failed");
  return h;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "errors") == 0) {
    // An unsupported operator token: the planner rejects it with a runtime_error.
    blazing_frame frame;
    try {
      gdf_column_cpp out;
      evaluate_expression(frame, "FOO($0, 1)", out);
      std::printf("no error\n");
      return 1;
    } catch (const std::runtime_error &e) {
      std::printf("runtime_error: %s\n", e.what());
    }
    return 0;
  }
  const int64_t n = argc > 1 ? std::stoll(argv[1]) : 5000;
  std::vector<int32_t> a(n), d(n);
  std::vector<double> b(n), c(n);
  for (int64_t i = 0; i < n; i++) {
    a[i] = (int32_t)(1 + (i * 7919) % 50);
    b[i] = 900.0 + (double)((i * 104729) % 104100);
    c[i] = (double)((i * 31) % 11) / 100.0;
    d[i] = (int32_t)((i * 40503) % 2556);
  }
  try {
    blazing_frame frame;
    std::vector<gdf_column_cpp> table(4);
    table[0].create_gdf_column(GDF_INT32, n, a.data(), 4, "a");
    table[1].create_gdf_column(GDF_FLOAT64, n, b.data(), 8, "b");
    table[2].create_gdf_column(GDF_FLOAT64, n, c.data(), 8, "c");
    table[3].create_gdf_column(GDF_INT32, n, d.data(), 4, "d");
    frame.add_table(table);

    // The stencil alone (evaluate_expression into a caller-allocated INT8 column).
    gdf_column_cpp stencil;
    stencil.create_gdf_column(GDF_INT8, n, nullptr, 1, "stencil");
    evaluate_expression(frame, ">=($3, 730)", stencil);
    auto s = download<int8_t>(stencil);
    int64_t ones = 0;
    for (auto v : s) ones += v != 0;
    std::printf("stencil %" PRId64 "\n", ones);

    process_filter(frame, kFilter);
    std::printf("kept %d\n", (int)frame.get_num_rows_in_table(0));
    execute_project_plan(frame, kProject);
    auto e0 = download<double>(frame.get_column(0));
    auto e1 = download<int32_t>(frame.get_column(1));
    auto e2 = download<int32_t>(frame.get_column(2));
    for (size_t i = 0; i < e0.size(); i++) std::printf("row %.17g %d %d\n", e0[i], e1[i], e2[i]);
    process_aggregate(frame, "LogicalAggregate(group=[{}], n=[COUNT()], m=[MAX($1)], s=[SUM($1)])");
    auto n_ = download<int64_t>(frame.get_column(0));
    auto m_ = download<int32_t>(frame.get_column(1));
    auto s_ = download<int64_t>(frame.get_column(2));
    std::printf("agg %lld %d %lld\n", (long long)n_[0], m_[0], (long long)s_[0]);
  } catch (const std::runtime_error &e) {
    std::printf("runtime_error: %s\n", e.what());
    return 2;
  }
  return 0;
}
