/*
 * ral.hpp — the reference's C++ operator API for the expression / filter / gather
 * path, over the C ABI of blz_ral.h (header only, C++17).
 *
 * Same names, argument meaning and error behaviour as the reference:
 *   perform_operation      src/Interpreter/interpreter_cpp.h:25-37
 *   evaluate_expression    src/LogicalFilter.h:20-23
 *   process_filter         src/CalciteInterpreter.cpp:509 (no header in the reference)
 *   execute_project_plan   src/CalciteInterpreter.h:61-63
 *   parse_project_plan     src/CalciteInterpreter.h:63 (with project_plan_params :17-32)
 *   add_expression_to_plan src/LogicalFilter.h:26-42
 *   materialize_column     src/ColumnManipulation.cuh:16-18
 *   gdf_sequence           src/Utils.cuh:187-190
 *   ral::io::csv_parser    src/io/data_parser/CSVParser.h:21-44 (ingest)
 *   ral::io::data_loader   src/io/DataLoader.h (files -> concatenated columns)
 *   result_set_repository  src/ResultSetRepository.h:29-63 (asynchronous results)
 * with the frame and column types they take (blazing_frame, DataFrame.h:16-181;
 * gdf_column_cpp, GDFColumn.cuh:17-95) reduced to the members this path uses.
 *
 * Errors: every failure of the C ABI is rethrown as std::runtime_error carrying the
 * library's message (the reference throws std::runtime_error from CheckCudaErrors
 * / CUDF_CALL, Utils.cuh:25-55). Calls block until their work is complete, as the
 * reference's do (interpreter_cpp.cu:282-285); each runs on the default stream.
 */
#ifndef BLZ_RAL_HPP_
#define BLZ_RAL_HPP_

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "blz_ral.h"

namespace blz_detail {

inline void check(int rc, const char *err, const char *what) {
  if (rc != BLZ_OK) throw std::runtime_error(std::string(what) + ": " + (err[0] ? err : "error " + std::to_string(rc)));
}

// Device buffers of one column; freed with the last gdf_column_cpp sharing them.
struct ColumnHolder {
  gdf_column col{};
  bool owned = true;  // false: an alias of another column's buffers
  std::shared_ptr<ColumnHolder> base;  // keeps the aliased column alive
  ~ColumnHolder() {
    if (owned) blz_free_column(&col);
  }
};

}  // namespace blz_detail

// A device column with shared ownership (GDFColumn.cuh:17-95).
class gdf_column_cpp {
 public:
  gdf_column_cpp() = default;

  // n values of `type`; input_data (host, may be NULL) is copied in; a validity
  // buffer of gdf_valid_allocation_size(n) bytes is allocated all valid
  // (GDFColumn.cu:302-338, allocate_valid :163-173); host_valid (may be NULL)
  // is copied over it.
  void create_gdf_column(gdf_dtype type, size_t num_values, void *input_data, size_t width_per_value,
                         const std::string &column_name = "") {
    create_gdf_column(type, num_values, input_data, nullptr, width_per_value, column_name);
  }
  void create_gdf_column(gdf_dtype type, size_t num_values, void *input_data, gdf_valid_type *host_valid,
                         size_t width_per_value, const std::string &column_name = "") {
    auto h = std::make_shared<blz_detail::ColumnHolder>();
    h->col.dtype = type;
    h->col.size = (gdf_size_type)num_values;
    const size_t bytes = num_values * width_per_value;
    if (bytes > 0) {
      if (blz_device_malloc(&h->col.data, (bytes + 63) / 64 * 64) != BLZ_OK)
        throw std::runtime_error("create_gdf_column: device allocation failed");
      if (input_data && blz_memcpy_htod(h->col.data, input_data, bytes, nullptr) != BLZ_OK)
        throw std::runtime_error("create_gdf_column: copy to the device failed");
    }
    const size_t vb = gdf_valid_allocation_size((gdf_size_type)num_values);
    if (num_values > 0) {
      void *v = nullptr;
      if (blz_device_malloc(&v, vb) != BLZ_OK) throw std::runtime_error("create_gdf_column: device allocation failed");
      h->col.valid = (gdf_valid_type *)v;
      if (blz_memset_device(v, 0xFF, vb, nullptr) != BLZ_OK)
        throw std::runtime_error("create_gdf_column: validity initialisation failed");
      if (host_valid && blz_memcpy_htod(v, host_valid, (num_values + 7) / 8, nullptr) != BLZ_OK)
        throw std::runtime_error("create_gdf_column: copy to the device failed");
    }
    h_ = h;
    name_ = column_name;
    if (host_valid) update_null_count();
  }
  // Takes ownership of a column whose buffers this library allocated.
  void create_gdf_column(gdf_column *column) {
    auto h = std::make_shared<blz_detail::ColumnHolder>();
    h->col = *column;
    h_ = h;
  }

  // A column over buffers someone else owns (a parked result's columns).
  static gdf_column_cpp view_of(const gdf_column &col, const std::string &name = "") {
    gdf_column_cpp c;
    auto h = std::make_shared<blz_detail::ColumnHolder>();
    h->col = col;
    h->owned = false;
    c.h_ = h;
    c.name_ = name;
    return c;
  }
  // Hands the buffers to another owner: no copy of this column, nor the column it
  // aliases, frees them.
  void disown() {
    for (auto h = h_; h; h = h->base) h->owned = false;
  }

  gdf_column *get_gdf_column() const { return h_ ? &h_->col : nullptr; }
  void *data() const { return h_ ? h_->col.data : nullptr; }
  gdf_valid_type *valid() const { return h_ ? h_->col.valid : nullptr; }
  gdf_size_type size() const { return h_ ? h_->col.size : 0; }
  gdf_dtype dtype() const { return h_ ? h_->col.dtype : GDF_invalid; }
  gdf_size_type null_count() { return h_ ? h_->col.null_count : 0; }
  std::string name() const { return name_; }
  void set_name(std::string name) { name_ = std::move(name); }

  // null_count = size - popcount(valid bits < size) (GDFColumn.cu:519-528).
  void update_null_count() {
    if (!h_) return;
    char err[256] = {0};
    blz_detail::check(blz_update_null_count(&h_->col, nullptr, err, sizeof(err)), err, "update_null_count");
  }

  // A column sharing `other`'s device buffers (a projection's "$k" output).
  static gdf_column_cpp alias_of(const gdf_column_cpp &other, const gdf_column &view) {
    gdf_column_cpp c;
    auto h = std::make_shared<blz_detail::ColumnHolder>();
    h->col = view;
    h->owned = false;
    h->base = other.h_;
    c.h_ = h;
    c.name_ = other.name_;
    return c;
  }

 private:
  std::shared_ptr<blz_detail::ColumnHolder> h_;
  std::string name_;
};

// Tables of columns addressed by one flat column index (DataFrame.h:16-181).
class blazing_frame {
 public:
  gdf_column_cpp &get_column(int column_index) {
    int base = 0;
    for (auto &t : columns_) {
      if (column_index < base + (int)t.size()) return t[column_index - base];
      base += (int)t.size();
    }
    throw std::runtime_error("get_column: index " + std::to_string(column_index) + " out of range");
  }
  std::vector<std::vector<gdf_column_cpp>> &get_columns() { return columns_; }
  void add_table(std::vector<gdf_column_cpp> columns_to_add) { columns_.push_back(std::move(columns_to_add)); }
  void add_column(gdf_column_cpp column_to_add, int table_index = 0) {
    if ((int)columns_.size() <= table_index) columns_.resize(table_index + 1);
    columns_[table_index].push_back(std::move(column_to_add));
  }
  void set_column(size_t column_index, gdf_column_cpp column) { get_column((int)column_index) = std::move(column); }
  size_t get_width() {
    size_t w = 0;
    for (auto &t : columns_) w += t.size();
    return w;
  }
  size_t get_size_column(int table_index = 0) { return columns_[table_index].size(); }
  gdf_size_type get_num_rows_in_table(int table_index) {
    return columns_[table_index].empty() ? 0 : columns_[table_index][0].size();
  }
  void clear() { columns_.clear(); }

  // Device column pointers of every column, in flat-index order.
  std::vector<gdf_column *> column_pointers() {
    std::vector<gdf_column *> p;
    for (auto &t : columns_)
      for (auto &c : t) p.push_back(c.get_gdf_column());
    return p;
  }

 private:
  std::vector<std::vector<gdf_column_cpp>> columns_;
};

// Evaluates a register-file plan (interpreter_cpp.h:25-37). output_columns'
// null_count is not written (interpreter_cpp.cu:280); new_input_indices is unused,
// as in the reference.
inline void perform_operation(std::vector<gdf_column *> output_columns, std::vector<gdf_column *> input_columns,
                              std::vector<column_index_type> &left_inputs,
                              std::vector<column_index_type> &right_inputs, std::vector<column_index_type> &outputs,
                              std::vector<column_index_type> &final_output_positions,
                              std::vector<gdf_binary_operator> &operators,
                              std::vector<gdf_unary_operator> &unary_operators, std::vector<gdf_scalar> &left_scalars,
                              std::vector<gdf_scalar> &right_scalars,
                              std::vector<column_index_type> new_input_indices) {
  (void)new_input_indices;
  char err[512] = {0};
  blz_detail::check(
      blz_perform_operation(output_columns.data(), (int)output_columns.size(), input_columns.data(),
                            (int)input_columns.size(), left_inputs.data(), right_inputs.data(), outputs.data(),
                            (int)operators.size(), final_output_positions.data(), operators.data(),
                            unary_operators.data(), left_scalars.data(), right_scalars.data(), BLZ_SCALAR_INTENDED,
                            nullptr, nullptr, err, sizeof(err)),
      err, "perform_operation");
}

// Appends the ops of one expression to a plan (LogicalFilter.h:26-42,
// LogicalFilter.cpp:44-267): registers [0, num_inputs) are the inputs,
// num_inputs + expression_position the expression's output, temporaries from
// num_inputs + num_outputs; new_input_indices[k] is frame column k's input register
// (or -1). `inputs` and `input_columns` are not read, as in the reference.
inline void add_expression_to_plan(blazing_frame &inputs, std::vector<gdf_column *> &input_columns,
                                   std::string expression, column_index_type expression_position,
                                   column_index_type num_outputs, column_index_type num_inputs,
                                   std::vector<column_index_type> &left_inputs,
                                   std::vector<column_index_type> &right_inputs,
                                   std::vector<column_index_type> &outputs,
                                   std::vector<gdf_binary_operator> &operators,
                                   std::vector<gdf_unary_operator> &unary_operators,
                                   std::vector<gdf_scalar> &left_scalars, std::vector<gdf_scalar> &right_scalars,
                                   std::vector<column_index_type> &new_input_indices) {
  (void)inputs;
  (void)input_columns;
  const int cap = 1024;
  std::vector<column_index_type> l(cap), r(cap), o(cap);
  std::vector<int32_t> b(cap), u(cap);
  std::vector<gdf_scalar> ls(cap), rs(cap);
  int n = 0;
  char err[512] = {0};
  blz_detail::check(blz_add_expression_to_plan(expression.c_str(), expression_position, num_outputs, num_inputs,
                                               new_input_indices.data(), (int)new_input_indices.size(), cap, l.data(),
                                               r.data(), o.data(), b.data(), u.data(), ls.data(), rs.data(), &n, err,
                                               sizeof(err)),
                    err, "add_expression_to_plan");
  for (int i = 0; i < n; i++) {
    left_inputs.push_back(l[i]);
    right_inputs.push_back(r[i]);
    outputs.push_back(o[i]);
    operators.push_back((gdf_binary_operator)b[i]);
    unary_operators.push_back((gdf_unary_operator)u[i]);
    left_scalars.push_back(ls[i]);
    right_scalars.push_back(rs[i]);
  }
}

// The plan of a LogicalProject (CalciteInterpreter.h:17-32): every evaluated
// expression in one op stream over input_columns, into output_columns (allocated
// here, one per evaluated expression, every input row); `columns` holds the
// projection's outputs in order ("$k" outputs share frame column k's buffers).
struct project_plan_params {
  size_t num_expressions_out;
  std::vector<gdf_column *> output_columns;
  std::vector<gdf_column *> input_columns;
  std::vector<column_index_type> left_inputs;
  std::vector<column_index_type> right_inputs;
  std::vector<column_index_type> outputs;
  std::vector<column_index_type> final_output_positions;
  std::vector<gdf_binary_operator> operators;
  std::vector<gdf_unary_operator> unary_operators;
  std::vector<gdf_scalar> left_scalars;
  std::vector<gdf_scalar> right_scalars;
  std::vector<column_index_type> new_column_indices;
  std::vector<gdf_column_cpp> columns;
  gdf_error error;
};

// parse_project_plan (CalciteInterpreter.h:63, CalciteInterpreter.cpp:123-309): the
// caller then runs perform_operation on the params' vectors, as the reference's
// interops benchmark does (benchmarks/interops/interops_benchmark.cpp:89-95).
inline project_plan_params parse_project_plan(blazing_frame &input, std::string query_part) {
  auto frame = input.column_pointers();
  std::vector<int32_t> dt(frame.size());
  for (size_t i = 0; i < frame.size(); i++) dt[i] = frame[i]->dtype;
  const int cap = 1024, max_out = 1024;
  std::vector<column_index_type> l(cap), r(cap), o(cap), fin(max_out);
  std::vector<int32_t> b(cap), u(cap), in_cols(cap), kind(max_out), index(max_out), odt(max_out);
  std::vector<gdf_scalar> ls(cap), rs(cap);
  int n_ops = 0, n_in = 0, n_eval = 0, n_out = 0;
  char err[512] = {0};
  blz_detail::check(blz_parse_project_plan(dt.data(), (int)dt.size(), query_part.c_str(), cap, l.data(), r.data(),
                                           o.data(), b.data(), u.data(), ls.data(), rs.data(), &n_ops, in_cols.data(),
                                           &n_in, fin.data(), &n_eval, kind.data(), index.data(), odt.data(),
                                           max_out, &n_out, err, sizeof(err)),
                    err, "parse_project_plan");
  project_plan_params p{};
  p.num_expressions_out = (size_t)n_eval;
  p.left_inputs.assign(l.begin(), l.begin() + n_ops);
  p.right_inputs.assign(r.begin(), r.begin() + n_ops);
  p.outputs.assign(o.begin(), o.begin() + n_ops);
  for (int i = 0; i < n_ops; i++) {
    p.operators.push_back((gdf_binary_operator)b[i]);
    p.unary_operators.push_back((gdf_unary_operator)u[i]);
  }
  p.left_scalars.assign(ls.begin(), ls.begin() + n_ops);
  p.right_scalars.assign(rs.begin(), rs.begin() + n_ops);
  p.final_output_positions.assign(fin.begin(), fin.begin() + n_eval);
  p.new_column_indices.assign(frame.size(), (column_index_type)-1);
  for (int i = 0; i < n_in; i++) {
    p.input_columns.push_back(frame[in_cols[i]]);
    p.new_column_indices[in_cols[i]] = (column_index_type)i;
  }
  const size_t rows = frame.empty() ? 0 : (size_t)frame[0]->size;
  std::vector<gdf_column *> evaluated(n_eval, nullptr);
  p.columns.resize(n_out);
  for (int i = 0; i < n_out; i++) {
    if (kind[i] == 0) {
      gdf_column_cpp c;
      c.create_gdf_column((gdf_dtype)odt[i], rows, nullptr, blz_dtype_width((gdf_dtype)odt[i]));
      evaluated[index[i]] = c.get_gdf_column();
      p.columns[i] = c;
    } else {
      p.columns[i] = input.get_column(index[i]);
    }
  }
  p.output_columns = evaluated;
  p.error = GDF_SUCCESS;
  return p;
}

// Plans and evaluates one Calcite expression over the frame into `output`
// (LogicalFilter.cpp:271-371); output's null_count is not written (:368).
inline void evaluate_expression(blazing_frame &inputs, const std::string &expression, gdf_column_cpp &output) {
  auto frame = inputs.column_pointers();
  char err[512] = {0};
  blz_detail::check(blz_evaluate_expression(frame.data(), (int)frame.size(), expression.c_str(),
                                            output.get_gdf_column(), BLZ_SCALAR_INTENDED, nullptr, nullptr, err,
                                            sizeof(err)),
                    err, "evaluate_expression");
}

// Keeps the rows whose predicate is non-zero and valid, in row order, for every
// column of the frame, which then holds the filtered columns as one table
// (CalciteInterpreter.cpp:509-573).
inline void process_filter(blazing_frame &input, std::string query_part) {
  auto frame = input.column_pointers();
  std::vector<gdf_column> outs(frame.size());
  int64_t kept = 0;
  char err[512] = {0};
  blz_detail::check(blz_process_filter(frame.data(), (int)frame.size(), query_part.c_str(), outs.data(), &kept,
                                       BLZ_SCALAR_INTENDED, nullptr, err, sizeof(err)),
                    err, "process_filter");
  std::vector<gdf_column_cpp> table(outs.size());
  for (size_t i = 0; i < outs.size(); i++) {
    table[i].create_gdf_column(&outs[i]);
    table[i].set_name(input.get_column((int)i).name());
  }
  input.clear();
  input.add_table(std::move(table));
}

// A LogicalAggregate without GROUP BY: the frame then holds one table of one-row
// columns, one per aggregation (ral::operators::process_aggregate, GroupBy.h:18,
// single-node path GroupBy.cpp:531-588; no communication context here).
inline void process_aggregate(blazing_frame &input, std::string query_part) {
  auto frame = input.column_pointers();
  const int max_out = 256;
  std::vector<gdf_column> outs(max_out);
  int n_out = 0;
  char err[512] = {0};
  blz_detail::check(blz_process_aggregate(frame.data(), (int)frame.size(), query_part.c_str(), outs.data(),
                                          max_out, &n_out, BLZ_SCALAR_INTENDED, nullptr, err, sizeof(err)),
                    err, "process_aggregate");
  std::vector<gdf_column_cpp> table(n_out);
  for (int i = 0; i < n_out; i++) table[i].create_gdf_column(&outs[i]);
  input.clear();
  input.add_table(std::move(table));
}

// Evaluates a LogicalProject over the frame, which then holds the projected
// columns; "$k" outputs alias frame column k (CalciteInterpreter.cpp:123-336).
inline void execute_project_plan(blazing_frame &input, std::string query_part) {
  auto frame = input.column_pointers();
  const int max_out = 256;
  std::vector<gdf_column> outs(max_out);
  std::vector<int32_t> alias(max_out, 0);
  int n_out = 0;
  char err[512] = {0};
  blz_detail::check(blz_execute_project(frame.data(), (int)frame.size(), query_part.c_str(), outs.data(),
                                        alias.data(), max_out, &n_out, BLZ_SCALAR_INTENDED, nullptr, err,
                                        sizeof(err)),
                    err, "execute_project_plan");
  std::vector<gdf_column_cpp> table(n_out);
  for (int i = 0; i < n_out; i++) {
    if (alias[i]) {
      int k = -1;  // the frame column whose buffers the output shares
      for (size_t c = 0; c < frame.size(); c++)
        if (frame[c]->data == outs[i].data) k = (int)c;
      if (k < 0) throw std::runtime_error("execute_project_plan: alias of an unknown column");
      table[i] = gdf_column_cpp::alias_of(input.get_column(k), outs[i]);
    } else {
      table[i].create_gdf_column(&outs[i]);
    }
  }
  input.clear();
  input.add_table(std::move(table));
}

// output[i] = input[row_indices[i]] (and its validity bit); sets
// output->null_count (ColumnManipulation.cu:83-150).
inline void materialize_column(gdf_column *input, gdf_column *output, gdf_column *row_indices) {
  char err[512] = {0};
  blz_detail::check(blz_materialize_column(input, output, row_indices, nullptr, err, sizeof(err)), err,
                    "materialize_column");
}

// data[i] = init_val + i * step (Utils.cu:4-12).
inline void gdf_sequence(int32_t *data, size_t size, int32_t init_val, int32_t step = 1) {
  char err[256] = {0};
  blz_detail::check(blz_gdf_sequence(data, size, init_val, step, nullptr, err, sizeof(err)), err, "gdf_sequence");
}

namespace ral {
namespace io {

// csv_parser (io/data_parser/CSVParser.h:21-44, .cpp:153-238): the reader options
// of init_default_csv_args (CSVParser.cpp:26-47) with the given delimiter, line
// terminator, column names and dtypes. parse takes a file path (the reference
// takes an arrow file handle and a Schema that repeats the names and dtypes).
// column_indices empty: every column; requested columns must be numeric or dates.
class csv_parser {
 public:
  csv_parser(std::string delimiter, std::string line_terminator, int skip_rows, std::vector<std::string> names,
             std::vector<gdf_dtype> dtypes)
      : delimiter_(delimiter.empty() ? '|' : delimiter[0]),
        line_terminator_(line_terminator.empty() ? '\n' : line_terminator[0]),
        skip_rows_(skip_rows),
        names_(std::move(names)),
        dtypes_(std::move(dtypes)) {}

  void parse(const std::string &path, std::vector<gdf_column_cpp> &columns_out,
             std::vector<size_t> column_indices = {}) const {
    load({path}, columns_out, column_indices);
  }
  // The same over text in host memory.
  void parse_buffer(const void *data, size_t n_bytes, std::vector<gdf_column_cpp> &columns_out,
                    std::vector<size_t> column_indices = {}) const {
    std::vector<int32_t> dt(dtypes_.begin(), dtypes_.end()), ci = indices(column_indices);
    std::vector<gdf_column> out(ci.size());
    int64_t rows = 0;
    char err[512] = {0};
    blz_detail::check(blz_csv_read(data, n_bytes, delimiter_, line_terminator_, skip_rows_, dt.data(), (int)dt.size(),
                                   ci.data(), (int)ci.size(), out.data(), &rows, nullptr, err, sizeof(err)),
                      err, "csv_parser::parse");
    adopt(out, ci, columns_out);
  }
  // data_loader::load_data (DataLoader.cpp:26-118): every file, columns concatenated.
  void load(const std::vector<std::string> &paths, std::vector<gdf_column_cpp> &columns_out,
            std::vector<size_t> column_indices = {}) const {
    std::vector<int32_t> dt(dtypes_.begin(), dtypes_.end()), ci = indices(column_indices);
    std::vector<const char *> ps;
    for (auto &p : paths) ps.push_back(p.c_str());
    std::vector<gdf_column> out(ci.size());
    int64_t rows = 0;
    char err[512] = {0};
    blz_detail::check(blz_csv_load_files(ps.data(), (int)ps.size(), delimiter_, line_terminator_, skip_rows_,
                                         dt.data(), (int)dt.size(), ci.data(), (int)ci.size(), out.data(), &rows,
                                         nullptr, err, sizeof(err)),
                      err, "csv_parser::parse");
    adopt(out, ci, columns_out);
  }

 private:
  std::vector<int32_t> indices(const std::vector<size_t> &column_indices) const {
    std::vector<int32_t> ci;
    if (column_indices.empty())
      for (size_t i = 0; i < dtypes_.size(); i++) ci.push_back((int32_t)i);
    else
      for (size_t c : column_indices) ci.push_back((int32_t)c);
    return ci;
  }
  void adopt(std::vector<gdf_column> &out, const std::vector<int32_t> &ci,
             std::vector<gdf_column_cpp> &columns_out) const {
    columns_out.clear();
    for (size_t i = 0; i < out.size(); i++) {
      gdf_column_cpp c;
      c.create_gdf_column(&out[i]);
      if ((size_t)ci[i] < names_.size()) c.set_name(names_[ci[i]]);
      columns_out.push_back(c);
    }
  }
  char delimiter_, line_terminator_;
  int skip_rows_;
  std::vector<std::string> names_;
  std::vector<gdf_dtype> dtypes_;
};

// data_loader (io/DataLoader.h): a parser over a list of files.
class data_loader {
 public:
  data_loader(std::shared_ptr<csv_parser> parser, std::vector<std::string> paths)
      : parser_(std::move(parser)), paths_(std::move(paths)) {}
  void load_data(std::vector<gdf_column_cpp> &columns, const std::vector<size_t> &column_indices) {
    parser_->load(paths_, columns, column_indices);
  }

 private:
  std::shared_ptr<csv_parser> parser_;
  std::vector<std::string> paths_;
};

}  // namespace io
}  // namespace ral

typedef uint64_t connection_id_t;
typedef uint64_t query_token_t;
typedef uint64_t column_token_t;

// result_set_t (ResultSetRepository.h:19-25): the frame's columns are views of
// the repository's buffers.
struct result_set_t {
  bool is_ready;
  blazing_frame result_frame;
  double duration;
  std::string errorMsg;
  size_t ref_counter;
};

// result_set_repository (ResultSetRepository.h:29-63), the process's one
// repository (blz_rs_*). Errors are the reference's std::runtime_error texts.
class result_set_repository {
 public:
  static result_set_repository &get_instance() {
    static result_set_repository instance;
    return instance;
  }
  query_token_t register_query(connection_id_t connection, query_token_t token) {
    char err[256] = {0};
    blz_detail::check(blz_rs_register_query(connection, token, err, sizeof(err)), err, "register_query");
    return token;
  }
  // Parks the frame: the repository takes its columns' buffers (the frame's
  // gdf_column_cpp no longer free them).
  void update_token(query_token_t token, blazing_frame frame, double duration, std::string errorMsg = "") {
    std::vector<gdf_column> cols;
    std::vector<std::string> names;
    for (auto &t : frame.get_columns())
      for (auto &c : t) {
        cols.push_back(*c.get_gdf_column());
        names.push_back(c.name());
      }
    std::vector<const char *> np;
    for (auto &n : names) np.push_back(n.c_str());
    char err[256] = {0};
    blz_detail::check(blz_rs_update_token(token, cols.data(), np.data(), (int)cols.size(), duration, errorMsg.c_str(),
                                          err, sizeof(err)),
                      err, "update_token");
    for (auto &t : frame.get_columns())
      for (auto &c : t) c.disown();
  }
  connection_id_t init_session() {
    connection_id_t c = 0;
    char err[256] = {0};
    blz_detail::check(blz_rs_init_session(&c, err, sizeof(err)), err, "init_session");
    return c;
  }
  void remove_all_connection_tokens(connection_id_t connection) { blz_rs_remove_all_connection_tokens(connection); }
  // Blocks until the token's frame is parked; counts one reference.
  result_set_t get_result(connection_id_t connection, query_token_t token) {
    std::vector<gdf_column> cols(256);
    std::vector<uint64_t> toks(256);
    int n = 0;
    double dur = 0;
    char msg[1024] = {0}, err[256] = {0};
    blz_detail::check(blz_rs_get_result(connection, token, cols.data(), toks.data(), (int)cols.size(), &n, &dur, msg,
                                        sizeof(msg), err, sizeof(err)),
                      err, "get_result");
    result_set_t r{true, blazing_frame{}, dur, msg, 1};
    std::vector<gdf_column_cpp> t;
    for (int i = 0; i < n; i++) {
      char name[256] = {0};
      blz_rs_column_name(token, i, name, sizeof(name));
      t.push_back(gdf_column_cpp::view_of(cols[i], name));
    }
    r.result_frame.add_table(t);
    return r;
  }
  bool try_free_result(connection_id_t connection, query_token_t token) {
    int found = 0;
    char err[256] = {0};
    blz_detail::check(blz_rs_try_free_result(connection, token, &found, err, sizeof(err)), err, "try_free_result");
    return found != 0;
  }
  gdf_column_cpp get_column(connection_id_t connection, column_token_t column_token) {
    gdf_column c{};
    char err[256] = {0};
    blz_detail::check(blz_rs_get_column(connection, column_token, &c, err, sizeof(err)), err, "get_column");
    return gdf_column_cpp::view_of(c);
  }

  result_set_repository(result_set_repository const &) = delete;
  void operator=(result_set_repository const &) = delete;

 private:
  result_set_repository() = default;
};

#endif  // BLZ_RAL_HPP_
