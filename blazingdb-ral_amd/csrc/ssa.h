/*
 * ssa.h — the SSA graph a reference-format plan is lowered to (program.cpp), shared
 * by the interpreter-program compiler and the kernel source generator
 * (codegen.cpp). Host code only.
 */
#ifndef BLZ_SSA_H_
#define BLZ_SSA_H_

#include <stdint.h>

#include <vector>

#include "program.h"

namespace blz {

struct Operand {
  bool imm = false;
  uint64_t bits = 0;
  int node = -1;  // value node when !imm
};
struct Node {
  int kind = 0;  // 0 input column (dev input = aux), 1 op, 2 constant zero
  int aux = 0;
  DevOp op{};
  Operand a, b;
  bool is_bool = false;
  int program = 0;  // 0 predicate, 1 projection
  // scheduling / allocation
  int need = 0;
  int order = -1;
  int last_use = -1;
  int slot = -1;
};

// Validity values: -1 all valid, -2 all null, >= 0 dynamic vnode id.
constexpr int V_ALL = -1, V_NONE = -2;
struct VNode {
  int kind = 0;  // 0 raw bitmap of dev input aux, 1 op
  int aux = 0;
  int code = VOP_AND;
  int a = 0, b = 0;
  int program = 0;
  int vslot = -1;
  bool needed = false;
  int last_use = -1;
  int order = -1;
};

struct Builder {
  std::vector<Node> nodes;
  std::vector<VNode> vnodes;
  std::vector<int> input_node;   // dev input -> node id
  std::vector<int> input_vnode;  // dev input -> vnode id (raw bitmap)
  int zero_node = -1;

  int in_node(int c) { return input_node[c]; }
  int zero() {
    if (zero_node < 0) {
      Node z; z.kind = 2; z.op.code = OP_ZERO;
      nodes.push_back(z);
      zero_node = (int)nodes.size() - 1;
    }
    return zero_node;
  }
  int vraw(int c) { return input_vnode[c]; }
  int vcombine(int code, int a, int b, int program) {
    if (code == VOP_AND) {
      if (a == V_NONE || b == V_NONE) return V_NONE;
      if (a == V_ALL) return b;
      if (b == V_ALL) return a;
      if (a == b) return a;
    } else {
      if (a == V_ALL || b == V_ALL) return V_ALL;
      if (a == V_NONE) return b;
      if (b == V_NONE) return a;
      if (a == b) return a;
    }
    VNode v; v.kind = 1; v.code = code; v.a = a; v.b = b; v.program = program;
    vnodes.push_back(v);
    return (int)vnodes.size() - 1;
  }
};

}  // namespace blz
#endif
