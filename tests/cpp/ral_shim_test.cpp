// The reference's C++ call sequence for a filter + project query (as in
// CalciteInterpreter.cpp:706-721 and evaluate_query.cu), through include/blz/ral.hpp.
// Usage: ral_shim_test N          (exit 0; prints the results as text)
//        ral_shim_test interops N (the reference's interops benchmark sequence,
//                                  benchmarks/interops/interops_benchmark.cpp:89-95:
//                                  parse_project_plan, then perform_operation on the
//                                  params' vectors, for its four plans; then an
//                                  expression planned by add_expression_to_plan
//                                  against evaluate_expression)
//        ral_shim_test errors     (no device needed: shows the error behaviour)
//        ral_shim_test csv FILE   (tests/tpch/tpch_csv_tests.cu TEST_01 over FILE:
//                                  data_loader + csv_parser, process_filter,
//                                  execute_project_plan; the frame parked in the
//                                  result repository and read back; FILE loaded
//                                  twice by one data_loader)
// Column i of the frame is generated from the row index alone, so the caller can
// rebuild the same frame for the oracle:
//   a[i] = 1 + (i * 7919) % 50                      int32
//   b[i] = 900 + (i * 104729) % 104100              float64
//   c[i] = ((i * 31) % 11) / 100.0                  float64 (0.00 .. 0.10)
//   d[i] = (i * 40503) % 2556                       int32
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <memory>
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

// The interops benchmark's plans (interops_benchmark.cpp:32-37) over in_x int32,
// in_y float64, in_z int32.
static const char *kInteropPlans[] = {
    "LogicalProject(EXPR$0=[+($0, $2)])",
    "LogicalProject(EXPR$0=[*(SIN(/(+($0, $2), $1)), POWER(/(-($0, $2), $1), 2))])",
    "LogicalProject(EXPR$1=[-(FLOOR(+(POWER(SIN($1), 2), POWER(COS($1), 2))), CEIL(+(POWER(SIN(+($1, 53.42)), 2), "
    "POWER(COS(+($1, 53.42)), 2))))])",
    "LogicalProject(EXPR$0=[*(*(10, SIN(/(+($0, $2), $1))), POWER(/(-($0, $2), $1), 2))], EXPR$1=[-(FLOOR(+(POWER("
    "SIN($1), 2), POWER(COS($1), 2))), CEIL(+(POWER(SIN(+($1, 53.42)), 2), POWER(COS(+($1, 53.42)), 2))))], "
    "EXPR$2=[+(-(/(*(FLOOR(+($1, 0.1)), POWER(MOD($0, 13), 2.0)), 5), CEIL(MOD(*(2, $2), 57))), 0.001)])"};

// Prints one output column: "out <plan> <output> <row> <value>" (floats as %a).
static void print_column(int plan, int o, gdf_column *c) {
  const size_t n = (size_t)c->size;
  const int w = blz_dtype_width(c->dtype);
  std::vector<unsigned char> h(n * w);
  if (n && blz_memcpy_dtoh(h.data(), c->data, n * w, nullptr) != BLZ_OK) throw std::runtime_error("copy failed");
  for (size_t i = 0; i < n; i++) {
    if (c->dtype == GDF_FLOAT64) {
      double v;
      std::memcpy(&v, &h[i * 8], 8);
      std::printf("out %d %d %zu %a\n", plan, o, i, v);
    } else if (c->dtype == GDF_FLOAT32) {
      float v;
      std::memcpy(&v, &h[i * 4], 4);
      std::printf("out %d %d %zu %a\n", plan, o, i, (double)v);
    } else {
      int64_t v = 0;
      if (w == 8) std::memcpy(&v, &h[i * 8], 8);
      else if (w == 4) { int32_t t; std::memcpy(&t, &h[i * 4], 4); v = t; }
      else if (w == 2) { int16_t t; std::memcpy(&t, &h[i * 2], 2); v = t; }
      else { int8_t t; std::memcpy(&t, &h[i], 1); v = t; }
      std::printf("out %d %d %zu %lld\n", plan, o, i, (long long)v);
    }
  }
}

static int interops(int64_t n) {
  // deterministic stand-ins for the benchmark's std::rand columns (same ranges)
  std::vector<int32_t> x(n), z(n);
  std::vector<double> y(n);
  for (int64_t i = 0; i < n; i++) {
    x[i] = (int32_t)((i * 7919 + 13) % 165191049);
    y[i] = (double)(1 + (i * 104729 + 7) % 2147483646);
    z[i] = (int32_t)((i * 40503 + 3) % 1073741823);
  }
  blazing_frame frame;
  std::vector<gdf_column_cpp> table(3);
  table[0].create_gdf_column(GDF_INT32, n, x.data(), 4, "in_x");
  table[1].create_gdf_column(GDF_FLOAT64, n, y.data(), 8, "in_y");
  table[2].create_gdf_column(GDF_INT32, n, z.data(), 4, "in_z");
  frame.add_table(table);
  for (int k = 0; k < 4; k++) {
    auto params = parse_project_plan(frame, kInteropPlans[k]);
    if (params.num_expressions_out <= 0) throw std::runtime_error("num_expressions_out is not good");
    perform_operation(params.output_columns, params.input_columns, params.left_inputs, params.right_inputs,
                      params.outputs, params.final_output_positions, params.operators, params.unary_operators,
                      params.left_scalars, params.right_scalars, params.new_column_indices);
    blazing_frame out;
    out.add_table(params.columns);
    std::printf("plan %d outputs %zu ops %zu\n", k, out.get_width(), params.operators.size());
    for (int o = 0; o < (int)out.get_width(); o++) print_column(k, o, out.get_column(o).get_gdf_column());
  }
  // add_expression_to_plan: one expression over frame columns 0 and 2 (input
  // registers 0 and 1, output register 2), evaluated by perform_operation, against
  // evaluate_expression of the same expression
  const std::string expr = "+(*($0, 3), MOD($2, 7))";
  std::vector<gdf_column *> inputs = {frame.get_column(0).get_gdf_column(), frame.get_column(2).get_gdf_column()};
  std::vector<column_index_type> left, right, outs, final_pos = {2}, idx = {0, -1, 1};
  std::vector<gdf_binary_operator> ops;
  std::vector<gdf_unary_operator> uops;
  std::vector<gdf_scalar> ls, rs;
  add_expression_to_plan(frame, inputs, expr, 0, 1, 2, left, right, outs, ops, uops, ls, rs, idx);
  gdf_column_cpp a, b;
  a.create_gdf_column(GDF_INT64, n, nullptr, 8, "a");
  b.create_gdf_column(GDF_INT64, n, nullptr, 8, "b");
  perform_operation({a.get_gdf_column()}, inputs, left, right, outs, final_pos, ops, uops, ls, rs, idx);
  evaluate_expression(frame, expr, b);
  std::vector<int64_t> ha(n), hb(n);
  if (n && (blz_memcpy_dtoh(ha.data(), a.data(), 8 * n, nullptr) != BLZ_OK ||
            blz_memcpy_dtoh(hb.data(), b.data(), 8 * n, nullptr) != BLZ_OK))
    throw std::runtime_error("copy failed");
  std::printf("add_expression_to_plan ops %zu equal %d\n", ops.size(), (int)(ha == hb));
  for (size_t i = 0; i < ops.size(); i++)
    std::printf("op %zu %d %d %d %d %d\n", i, left[i], right[i], outs[i], (int)ops[i], (int)uops[i]);
  print_column(9, 0, a.get_gdf_column());
  return 0;
}

static int csv_mode(const char *path) {
  const std::vector<std::string> names = {"c_custkey", "c_name",    "c_address",    "c_nationkey",
                                          "c_phone",   "c_acctbal", "c_mktsegment", "c_comment"};
  const std::vector<gdf_dtype> types = {GDF_INT32, GDF_INT64, GDF_INT64, GDF_INT32,
                                        GDF_INT64, GDF_FLOAT32, GDF_INT64, GDF_INT64};
  auto parser = std::make_shared<ral::io::csv_parser>("|", "\n", 0, names, types);
  std::vector<gdf_column_cpp> cols;
  ral::io::data_loader(parser, {path}).load_data(cols, {});
  std::printf("rows %d\n", (int)cols[0].size());
  std::vector<gdf_column_cpp> twice;
  ral::io::data_loader(parser, {path, path}).load_data(twice, {0, 5});
  std::printf("rows_twice %d %s %s\n", (int)twice[0].size(), twice[0].name().c_str(), twice[1].name().c_str());
  blazing_frame frame;
  frame.add_table(cols);
  process_filter(frame, "LogicalFilter(condition=[AND(<($0, 150), =($3, 5))])");
  execute_project_plan(frame, "LogicalProject(c_custkey=[$0], c_nationkey=[$3], c_acctbal=[$5])");
  auto &repo = result_set_repository::get_instance();
  const connection_id_t conn = repo.init_session();
  repo.register_query(conn, 77);
  repo.update_token(77, frame, 1.5);
  frame.clear();
  cols.clear();
  result_set_t r = repo.get_result(conn, 77);
  std::printf("duration %g error '%s'\n", r.duration, r.errorMsg.c_str());
  for (size_t i = 0; i < r.result_frame.get_width(); i++) {
    gdf_column_cpp &c = r.result_frame.get_column((int)i);
    std::printf("col %s", c.name().c_str());
    if (c.dtype() == GDF_FLOAT32) {
      for (float v : download<float>(c)) std::printf(" %a", (double)v);
    } else {
      for (int32_t v : download<int32_t>(c)) std::printf(" %d", v);
    }
    std::printf("\n");
  }
  std::printf("freed %d\n", (int)repo.try_free_result(conn, 77));
  try {
    repo.get_result(conn, 77);
  } catch (const std::runtime_error &e) {
    std::printf("runtime_error: %s\n", e.what());
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 2 && std::strcmp(argv[1], "csv") == 0) {
    try {
      return csv_mode(argv[2]);
    } catch (const std::runtime_error &e) {
      std::printf("runtime_error: %s\n", e.what());
      return 2;
    }
  }
  if (argc > 2 && std::strcmp(argv[1], "interops") == 0) {
    try {
      return interops(std::stoll(argv[2]));
    } catch (const std::runtime_error &e) {
      std::printf("runtime_error: %s\n", e.what());
      return 2;
    }
  }
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
