/*
 * planner.h — Calcite expression text -> register-file plan (host).
 *
 * The reference's expression planner (src/CalciteExpressionParsing.cpp,
 * src/LogicalFilter.cpp:44-371, src/CalciteInterpreter.cpp:123-309), kept
 * token-for-token compatible so plans equal the reference's.
 */
#ifndef BLZ_PLANNER_H_
#define BLZ_PLANNER_H_

#include <string>
#include <vector>

#include "program_compiler.h"

namespace blz {
namespace planner {

std::string clean_calcite_expression(std::string expression);
std::string expand_if_logical_op(std::string expression);
std::vector<std::string> get_expressions_from_expression_list(std::string combined, bool trim = true);
int find_closing_char(const std::string &expression, int start);
std::vector<std::string> get_tokens_in_reverse_order(const std::string &expression);
bool is_operator_token(const std::string &t);
bool is_literal(const std::string &t);
bool is_string(const std::string &t);
bool is_number(const std::string &t);
bool is_date(const std::string &t);
size_t get_index(const std::string &operand);
int get_type_from_string(const std::string &s);
gdf_scalar get_scalar_from_string(const std::string &s, int dtype);
int get_output_type_expression(const std::vector<int> &dtypes, const std::string &expression);
std::string get_named_expression(const std::string &query_part, const std::string &name);
bool contains_evaluation(const std::string &expression);

// add_expression_to_plan (LogicalFilter.cpp:44-267): appends the ops of one
// expression to p (its registers: inputs [0, num_inputs), outputs from num_inputs,
// temporaries from num_inputs + num_outputs, lowest free first).
// new_input_indices[frame column] = its input register, or -1.
void add_expression_to_plan(PlanSpec &p, const std::string &expression, int expression_position,
                            int num_outputs, int num_inputs, const std::vector<int> &new_input_indices);

// evaluate_expression's plan. `spec.in_dev` holds FRAME column indices.
struct ExprPlan {
  bool bare = false;  // "$k": the expression is column k itself
  int bare_col = -1;
  PlanSpec spec;
};
ExprPlan plan_expression(const std::vector<int> &frame_dtypes, const std::string &expression);

// parse_project_plan. outputs[i] = {0, k, dtype} for the k-th evaluated expression,
// {1, column, dtype} for a "$k" alias.
struct ProjOut {
  int kind, index, dtype;
  std::string name;
};
struct ProjectPlan {
  bool has_eval = false;
  PlanSpec spec;  // in_dev = frame column indices
  std::vector<ProjOut> outputs;
};
ProjectPlan parse_project(const std::vector<int> &frame_dtypes, const std::string &query_part);

}  // namespace planner
}  // namespace blz
#endif
