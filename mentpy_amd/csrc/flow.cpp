// Causal flow and measurement order of an open graph, on the host.
//
// Follows mentpy's find_cflow / causal_flow_aux (mentpy/mbqc/flow.py:115-185),
// the Mhalla-Perdrix peeling algorithm (arXiv:0709.2670): starting from the
// outputs, a processed node v with exactly one unprocessed neighbour u gets
// u as its flow predecessor (flow(u) = v) and u the current layer; u and the
// neighbours of u whose count of unprocessed neighbours drops to one join the
// next frontier.  Like mentpy, a node is marked processed as soon as it is
// assigned, so a later frontier node of the same round already sees it.  The
// frontier of a round is visited in ascending node index.
//
// The measurement order follows MBQCircuit.calculate_order
// (mentpy/mbqc/mbqcircuit.py:390-422): nodes grouped by how many nodes they
// precede in the partial order layer(u) > layer(v), groups in decreasing
// count, ties in node order, then the inputs moved to the front in input order.

#include <algorithm>
#include <string>
#include <vector>

#include "plan.h"

namespace mbqc {

int causal_flow(int n, int n_edges, const int32_t* edges, int n_in, const int32_t* inputs,
                int n_out, const int32_t* outputs, std::vector<int>* flow, std::vector<int>* layer,
                std::vector<int>* order, bool* found) {
  *found = false;
  if (n_in != n_out)
    return fail(MBQC_ERR_VALUE, "Cannot find flow or gflow. Input (" + std::to_string(n_in) +
                                    ") and output (" + std::to_string(n_out) +
                                    ") nodes have different size.");
  std::vector<std::vector<int>> nbr(n);
  {
    std::vector<std::vector<int>> raw(n);
    for (int e = 0; e < n_edges; ++e) {
      const int a = edges[2 * e], b = edges[2 * e + 1];
      if (a < 0 || a >= n || b < 0 || b >= n || a == b) return fail(MBQC_ERR_VALUE, "bad edge");
      raw[a].push_back(b);
      raw[b].push_back(a);
    }
    for (int v = 0; v < n; ++v) {
      std::sort(raw[v].begin(), raw[v].end());
      raw[v].erase(std::unique(raw[v].begin(), raw[v].end()), raw[v].end());
      nbr[v] = std::move(raw[v]);
    }
  }
  std::vector<char> is_in(n, 0), is_out(n, 0);
  for (int m = 0; m < n_in; ++m) {
    if (inputs[m] < 0 || inputs[m] >= n) return fail(MBQC_ERR_VALUE, "input node out of range");
    is_in[inputs[m]] = 1;
  }
  for (int m = 0; m < n_out; ++m) {
    if (outputs[m] < 0 || outputs[m] >= n) return fail(MBQC_ERR_VALUE, "output node out of range");
    is_out[outputs[m]] = 1;
  }
  flow->assign(n, -1);
  layer->assign(n, 0);
  std::vector<int> past(n, 0);
  std::vector<char> processed(is_out);
  auto unprocessed_neighbours = [&](int v) {
    int c = 0;
    for (int x : nbr[v]) c += !processed[x];
    return c;
  };
  std::vector<char> in_next(n, 0);
  std::vector<int> frontier;
  for (int v = 0; v < n; ++v) {
    if (!is_out[v] || is_in[v]) continue;
    past[v] = unprocessed_neighbours(v);
    if (past[v] == 1) frontier.push_back(v);
  }
  int assigned = 0;
  for (int k = 1; !frontier.empty(); ++k) {
    std::vector<int> next;
    for (int v : frontier) {
      int u = -1, c = 0;
      for (int x : nbr[v])
        if (!processed[x]) {
          u = x;
          ++c;
        }
      if (c != 1) continue;
      (*flow)[u] = v;
      (*layer)[u] = k;
      processed[u] = 1;
      ++assigned;
      auto push = [&](int x) {
        if (!in_next[x]) {
          in_next[x] = 1;
          next.push_back(x);
        }
      };
      if (!is_in[u]) {
        past[u] = unprocessed_neighbours(u);
        if (past[u] == 1) push(u);
      }
      for (int w : nbr[u]) {
        if (past[w] > 0) {
          --past[w];
          if (past[w] == 1) push(w);
        }
      }
    }
    for (int x : next) in_next[x] = 0;
    std::sort(next.begin(), next.end());
    frontier.swap(next);
  }
  int n_unique_out = 0;
  for (int v = 0; v < n; ++v) n_unique_out += is_out[v];
  if (assigned != n - n_unique_out) {
    flow->assign(n, -1);
    layer->assign(n, 0);
    order->clear();
    return MBQC_OK;
  }
  *found = true;
  // count[i] = #{j : layer[i] > layer[j]}; larger count first, ties in node order
  std::vector<int> lay_sorted(layer->begin(), layer->end());
  std::sort(lay_sorted.begin(), lay_sorted.end());
  std::vector<int> count(n);
  for (int v = 0; v < n; ++v)
    count[v] = int(std::lower_bound(lay_sorted.begin(), lay_sorted.end(), (*layer)[v]) - lay_sorted.begin());
  order->resize(n);
  for (int v = 0; v < n; ++v) (*order)[v] = v;
  std::stable_sort(order->begin(), order->end(), [&](int a, int b) { return count[a] > count[b]; });
  for (int m = n_in - 1; m >= 0; --m) {
    auto it = std::find(order->begin(), order->end(), inputs[m]);
    order->erase(it);
    order->insert(order->begin(), inputs[m]);
  }
  return MBQC_OK;
}

}  // namespace mbqc

extern "C" int mbqc_causal_flow(int32_t n_nodes, int32_t n_edges, const int32_t* edges,
                                int32_t n_inputs, const int32_t* input_nodes, int32_t n_outputs,
                                const int32_t* output_nodes, int32_t* flow, int32_t* layer,
                                int32_t* order, int32_t* found) {
  using namespace mbqc;
  if (n_nodes <= 0) return fail(MBQC_ERR_VALUE, "pattern has no nodes");
  if ((n_edges > 0 && !edges) || (n_inputs > 0 && !input_nodes) || (n_outputs > 0 && !output_nodes) ||
      n_edges < 0 || n_inputs < 0 || n_outputs < 0 || !found)
    return fail(MBQC_ERR_VALUE, "null argument");
  std::vector<int> f, l, o;
  bool ok = false;
  int rc = causal_flow(n_nodes, n_edges, edges, n_inputs, input_nodes, n_outputs, output_nodes, &f,
                       &l, &o, &ok);
  if (rc != MBQC_OK) return rc;
  *found = ok ? 1 : 0;
  for (int v = 0; v < n_nodes; ++v) {
    if (flow) flow[v] = f[v];
    if (layer) layer[v] = l[v];
    if (order) order[v] = ok ? o[v] : -1;
  }
  return MBQC_OK;
}
