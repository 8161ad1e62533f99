// Host-side pattern compiler: turns a window-ordered pattern into the per-step
// tables the HIP kernels read.
//
// Live state ("slot") layout.  The reference keeps a window of w qubits,
// measures the oldest (window index 0, the MSB), traces it out and appends
// the next scheduled node as the new LSB (np_simulator_sv.py:164-225).  Here
// the live state keeps n_slots = w-1 qubits (w when all w window qubits are
// inputs or outputs): node schedule[j] lives in slot (n_slots-1-j) mod n_slots,
// so the appended node reuses the slot the measured node frees and a step is
// an in-place update of amplitude pairs.  Because the appended qubit is in
// |+> with known CZ signs, its two halves are written from one projection.
//
// Step k measures schedule[k] and appends schedule[k+n_slots] (if it exists).
// With n_slots = w-1 ("lag 1") that node entered the reference window before
// schedule[k] was measured, so its CZ with schedule[k] is applied first (the
// "measured is a neighbour" flag); with n_slots = w ("lag 0") it is appended
// after the trace, as in the reference.  CZs to nodes no longer in the window
// are dropped exactly as the reference drops them.
#include <algorithm>
#include <map>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "plan.h"

namespace mbqc {

thread_local std::string g_last_error;
int (*g_plan_upload_hook)(Plan*) = nullptr;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

static inline uint64_t edge_key(int a, int b) {
  if (a > b) std::swap(a, b);
  return (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
}

// Rescale schedule: the exponent starts at the padding count; a block whose
// running exponent reached the threshold starts with a 2^-32 rescale.
static void rescale_schedule(const Plan& p, int threshold, int32_t* flags, int32_t* e_final) {
  int e = p.ns - p.n_in;
  for (int blk = 0; blk < p.n_blocks(); ++blk) {
    flags[blk] = 0;
    if (e >= threshold) {
      flags[blk] = 1;
      e -= threshold;
    }
    for (int sub = 0; sub < p.ns; ++sub) {
      int k = blk * p.ns + sub;
      if (k < p.n_steps) e += int(p.steps[k].flags & 1u);
    }
  }
  *e_final = e;
}

int compile_plan(const mbqc_pattern_desc* d, Plan* p) {
  if (!d || !p) return fail(MBQC_ERR_VALUE, "null pattern description");
  const int N = d->n_nodes;
  if (N <= 0) return fail(MBQC_ERR_VALUE, "pattern has no nodes");
  if (!d->schedule) return fail(MBQC_ERR_VALUE, "schedule is null");
  if (d->n_edges < 0 || (d->n_edges > 0 && !d->edges))
    return fail(MBQC_ERR_VALUE, "edge list is null");
  if (d->n_inputs < 0 || (d->n_inputs > 0 && !d->input_nodes))
    return fail(MBQC_ERR_VALUE, "input node list is null");
  if (d->n_outputs < 0 || (d->n_outputs > 0 && !d->output_nodes))
    return fail(MBQC_ERR_VALUE, "output node list is null");
  if (!d->angle_index || !d->fixed_angle) return fail(MBQC_ERR_VALUE, "angle tables are null");
  const int nm = d->n_measure;
  const int w = d->window;
  const int nin = d->n_inputs, nout = d->n_outputs;
  if (nm < 0 || nm > N) return fail(MBQC_ERR_VALUE, "n_measure out of range");
  if (w < 1) return fail(MBQC_ERR_VALUE, "window size must be at least 1");
  if (nin > w)
    return fail(MBQC_ERR_VALUE, "Input state has " + std::to_string(nin) +
                                    " qubits, but window size is set to " + std::to_string(w) + ".");
  if (w > nm)
    return fail(MBQC_ERR_VALUE, "Window size is set to " + std::to_string(w) +
                                    ", but schedule only has " + std::to_string(nm) + " measurements.");
  if (nout != N - nm)
    return fail(MBQC_ERR_VALUE, "output count must equal the number of unmeasured nodes");

  // schedule must be a permutation of 0..N-1
  std::vector<int> pos(N, -1);
  for (int j = 0; j < N; ++j) {
    int v = d->schedule[j];
    if (v < 0 || v >= N || pos[v] >= 0) return fail(MBQC_ERR_VALUE, "schedule is not a permutation of the nodes");
    pos[v] = j;
  }
  // inputs must be the first nin scheduled nodes (any order), outputs the last nout
  for (int m = 0; m < nin; ++m) {
    int v = d->input_nodes[m];
    if (v < 0 || v >= N || pos[v] >= nin)
      return fail(MBQC_ERR_VALUE, "input nodes must be the first scheduled nodes");
  }
  for (int m = 0; m < nout; ++m) {
    int v = d->output_nodes[m];
    if (v < 0 || v >= N || pos[v] < nm)
      return fail(MBQC_ERR_VALUE, "output nodes must be the unmeasured scheduled nodes");
  }
  {
    std::unordered_set<int> s(d->input_nodes, d->input_nodes + nin);
    if ((int)s.size() != nin) return fail(MBQC_ERR_VALUE, "repeated input node");
    std::unordered_set<int> o(d->output_nodes, d->output_nodes + nout);
    if ((int)o.size() != nout) return fail(MBQC_ERR_VALUE, "repeated output node");
  }

  std::unordered_set<uint64_t> edges;
  std::vector<std::vector<int>> nbr(N);
  for (int e = 0; e < d->n_edges; ++e) {
    int a = d->edges[2 * e], b = d->edges[2 * e + 1];
    if (a < 0 || a >= N || b < 0 || b >= N || a == b) return fail(MBQC_ERR_VALUE, "bad edge");
    if (edges.insert(edge_key(a, b)).second) {
      nbr[a].push_back(b);
      nbr[b].push_back(a);
    }
  }
  auto has_edge = [&](int a, int b) { return edges.count(edge_key(a, b)) > 0; };

  p->schedule.assign(d->schedule, d->schedule + N);
  p->nbr = nbr;
  // measurement order: schedule order, or the description's, which must
  // measure each of schedule[0..nm) once, after it has entered the window
  std::vector<int> order(d->schedule, d->schedule + nm);
  if (d->measure_order) {
    std::vector<char> seen(N, 0);
    for (int k = 0; k < nm; ++k) {
      const int v = d->measure_order[k];
      if (v < 0 || v >= N || pos[v] >= nm)
        return fail(MBQC_ERR_VALUE, "measure_order names a node that is not measured");
      if (seen[v]) return fail(MBQC_ERR_VALUE, "measure_order repeats a node");
      seen[v] = 1;
      if (pos[v] > k + w - 1)
        return fail(MBQC_ERR_VALUE, "measure_order[" + std::to_string(k) + "] = " +
                                        std::to_string(v) + " has not entered the window yet");
      order[k] = v;
    }
  }
  p->order = order;
  // With lag 1 the node that enters the window with step k (schedule[k+w-1])
  // is appended by step k itself, so no step may measure it right away.
  bool lag_ok = true;
  for (int k = 0; k < nm; ++k) lag_ok = lag_ok && pos[order[k]] != k + w - 1;
  const int ns = (w >= 2 && nin <= w - 1 && nout <= w - 1 && lag_ok) ? w - 1 : w;
  if (ns > kMaxSlots)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "live window of " + std::to_string(ns) +
                                              " qubits exceeds the supported maximum of " +
                                              std::to_string(kMaxSlots));
  p->ns = ns;
  p->lag = w - ns;
  p->w = w;
  p->n_nodes = N;
  p->n_steps = nm;
  p->n_in = nin;
  p->n_out = nout;
  p->n_angles = d->n_angles;
  auto slot_of_pos = [&](int j) { return ((ns - 1 - j) % ns + ns) % ns; };

  // initial live state: schedule[0..ns)
  const int H = 1 << ns;
  p->in_index.assign(H, 0);
  p->init_sign.assign(H, 0);
  p->out_pos.assign(H, -1);
  for (int j = 0; j < H; ++j) {
    int idx = 0;
    for (int m = 0; m < nin; ++m) {
      int bit = (j >> slot_of_pos(pos[d->input_nodes[m]])) & 1;
      idx |= bit << (nin - 1 - m);
    }
    p->in_index[j] = idx;
    int par = 0;
    for (int a = 0; a < ns; ++a)
      for (int b = a + 1; b < ns; ++b)
        if (has_edge(d->schedule[a], d->schedule[b]))
          par ^= ((j >> slot_of_pos(a)) & 1) & ((j >> slot_of_pos(b)) & 1);
    p->init_sign[j] = par;
  }
  p->n_appends = ns - nin;  // |+> padding counted like appends

  // steps
  p->steps.resize(nm);
  int applied = 0;
  for (int a = 0; a < ns; ++a)
    for (int b = a + 1; b < ns; ++b) applied += has_edge(d->schedule[a], d->schedule[b]);
  std::vector<char> angle_used(d->n_angles > 0 ? d->n_angles : 0, 0);
  // register slot of each live node: the initial ones by position, then each
  // appended node takes the slot its step's measured node frees
  std::vector<int> slot(N, -1);
  for (int j = 0; j < ns && j < N; ++j) slot[d->schedule[j]] = slot_of_pos(j);
  p->rotating = true;
  for (int k = 0; k < nm; ++k) {
    mbqc_step st{};
    const int v = order[k];
    const int q = slot[v];
    if (q < 0) return fail(MBQC_ERR_VALUE, "measured node is not live");
    slot[v] = -1;
    p->rotating = p->rotating && q == slot_of_pos(k);
    int ai = d->angle_index[v];
    if (ai >= d->n_angles) return fail(MBQC_ERR_VALUE, "angle index out of range");
    if (ai >= 0) {
      if (angle_used[ai]) p->shared_angles = true;
      angle_used[ai] = 1;
    }
    st.angle_index = ai < 0 ? -1 : ai;
    st.fixed_angle = ai < 0 ? d->fixed_angle[v] : 0.0f;
    const int plane = d->plane ? d->plane[v] : MBQC_PLANE_XY;
    if (plane < MBQC_PLANE_XY || plane > MBQC_PLANE_Z)
      return fail(MBQC_ERR_VALUE, "measurement plane code out of range");
    if (plane != MBQC_PLANE_XY) p->general_planes = true;
    uint32_t flags = (uint32_t(q) << 8) | (uint32_t(plane) << 2);
    uint32_t nmask = 0;
    const int jn = k + ns;
    if (jn <= N - 1) {
      flags |= 1u;
      const int nv = d->schedule[jn];
      ++p->n_appends;
      if (p->lag == 1 && has_edge(v, nv)) {
        flags |= 2u;
        ++applied;
      }
      for (int u : nbr[nv])
        if (slot[u] >= 0) {   // live, the measured node excluded
          nmask |= 1u << slot[u];
          ++applied;
        }
      slot[nv] = q;
    }
    st.flags = flags;
    st.nmask = nmask;
    p->steps[k] = st;
  }
  p->n_edges = (int)edges.size();
  p->n_dropped = p->n_edges - applied;

  // smallest window applying every edge: an edge (a, b) with pos a < pos b
  // needs b to enter while a is live: pos[b] - pos[a] <= w - 1, and a must
  // not be measured before b enters.
  int need = 1;
  for (int v = 0; v < N; ++v)
    for (int u : nbr[v]) {
      int a = std::min(pos[u], pos[v]), b = std::max(pos[u], pos[v]);
      if (a < nm) need = std::max(need, b - a + 1);
    }
  p->min_window = need;

  // output gather: output_nodes[m] is MSB-first bit (nout-1-m) of the output index
  std::vector<int> live_slot(nout);
  uint32_t live_mask = 0;
  for (int m = 0; m < nout; ++m) {
    live_slot[m] = slot[d->output_nodes[m]];
    if (live_slot[m] < 0) return fail(MBQC_ERR_VALUE, "output node is not live at the end");
    live_mask |= 1u << live_slot[m];
  }
  for (int j = 0; j < H; ++j) {
    if (j & ~int(live_mask)) continue;
    int o = 0;
    for (int m = 0; m < nout; ++m) o |= ((j >> live_slot[m]) & 1) << (nout - 1 - m);
    p->out_pos[j] = o;
  }
  p->out_slot = live_slot;
  // pair coefficients: u0 = x + b y is the new bit-0 half; the bit-1 half is
  // c (x + s b y) with s = -1 when the measured qubit is a CZ neighbour of the
  // appended one, c = CZ sign of the pair (0 without an append).  Written as
  // alpha u0 + beta x:  alpha = c s,  beta = c (1 - s).
  // Step k keeps its row index in flags bits 16..31.  A table of at most
  // kCoefPerStepBytes keeps one row per step (row k, addressed without the
  // flags by the forward kernel); a longer one stores each distinct row once
  // (a grid repeats its rows column after column), so it stays within the
  // scalar cache.
  if (ns <= kMaxCoefSlots) {
    p->coef.clear();
    p->coef_per_step = (size_t(nm) << ns) * sizeof(float) <= kCoefPerStepBytes;
    std::map<std::vector<float>, uint32_t> row_index;
    std::vector<float> row(size_t(1) << ns);
    for (int k = 0; k < nm; ++k) {
      mbqc_step& st = p->steps[k];
      const int q = int((st.flags >> 8) & 0xffu);
      const bool app = st.flags & 1u, sneg = st.flags & 2u;
      int i = 0;
      for (int j = 0; j < H; ++j) {
        if (j & (1 << q)) continue;
        const float sgn = (__builtin_popcount(unsigned(j) & st.nmask) & 1) ? -1.0f : 1.0f;
        const float c = app ? sgn : 0.0f;
        const float s = sneg ? -1.0f : 1.0f;
        row[2 * i] = c * s;
        row[2 * i + 1] = c * (1.0f - s);
        ++i;
      }
      uint32_t r = uint32_t(k);
      if (!p->coef_per_step) {
        auto it = row_index.find(row);
        if (it == row_index.end()) it = row_index.emplace(row, uint32_t(row_index.size())).first;
        r = it->second;
      }
      if (size_t(r) * row.size() == p->coef.size()) p->coef.insert(p->coef.end(), row.begin(), row.end());
      st.flags = (st.flags & 0xffffu) | (r << 16);
    }
  }
  p->periodic = p->rotating && ns <= kMaxCoefSlots && nm > 0 && nm % ns == 0;
  for (int k = 0; k < nm && p->periodic; ++k) {
    const mbqc_step& st = p->steps[k];
    p->periodic = (st.flags & 3u) == 3u && st.nmask == periodic_nmask(ns, k % ns);
  }
  // Pauli-frame tables for sampled outcomes (PennyLane backend order,
  // pennylane_simulator.py:146-153): outcome 1 on schedule[k] flips the X
  // frame of its flow successor and the Z frame of the successor's neighbours
  // measured after schedule[k].  Frames are kept relative to the current
  // step in 32-bit words on the device, so every offset must be below 32
  // (with no dropped edge and n live slots they stay below 2n + 2).
  p->can_sample = false;
  if (!d->flow) {
    p->sample_reason = "the pattern description has no flow";
  } else if (order != std::vector<int>(d->schedule, d->schedule + nm)) {
    p->sample_reason = "nodes are measured out of schedule order";
  } else if (p->n_dropped > 0) {
    p->sample_reason = "the window drops CZ edges";
  } else if (ns > kMaxSampleSlots) {
    p->sample_reason = "sampled outcomes support live windows of at most 20 qubits";
  } else {
    bool ok = true;
    p->xmask.assign(nm, 0);
    p->zmask.assign(nm, 0);
    for (int k = 0; k < nm && ok; ++k) {
      const int v = d->schedule[k];
      const int f = d->flow[v];
      if (f < 0 || f >= N || pos[f] <= k || !has_edge(v, f)) {
        ok = false;
        p->sample_reason = "flow successor of a measured node is invalid";
        break;
      }
      const int dx = pos[f] - k;
      if (dx >= 32) { ok = false; break; }
      p->xmask[k] |= uint64_t(1) << dx;
      for (int nb : nbr[f]) {
        if (nb == v || pos[nb] <= k) continue;
        const int dz = pos[nb] - k;
        if (dz >= 32) { ok = false; break; }
        p->zmask[k] |= uint64_t(1) << dz;
      }
    }
    if (ok) {
      p->out_offset.assign(nout, 0);
      for (int m = 0; m < nout; ++m) {
        const int off = pos[d->output_nodes[m]] - nm;
        if (off < 0 || off >= 32) { ok = false; break; }
        p->out_offset[m] = off;
      }
    }
    if (ok) {
      p->can_sample = true;
    } else {
      if (p->sample_reason.empty()) p->sample_reason = "flow dependencies span more than 32 positions";
      p->xmask.clear();
      p->zmask.clear();
      p->out_offset.clear();
    }
  }
  std::vector<int32_t> flags(p->n_blocks());
  rescale_schedule(*p, kSvRescaleAppends, flags.data(), &p->e_sv);
  rescale_schedule(*p, kDmRescaleAppends, flags.data(), &p->e_dm);
  return MBQC_OK;
}

TableOffsets table_offsets(const Plan& p) {
  TableOffsets o{};
  const int H = 1 << p.ns;
  o.steps = 0;
  o.in_index = 4 * p.n_steps;
  o.init_sign = o.in_index + H;
  o.out_pos = o.init_sign + H;
  o.sv_rescale = o.out_pos + H;
  o.dm_rescale = o.sv_rescale + p.n_blocks();
  o.e_final = o.dm_rescale + p.n_blocks();
  o.coef = o.e_final + 2;
  o.xmask = o.coef + int(p.coef.size());
  o.xmask += o.xmask & 1;  // 8-byte alignment of the 64-bit masks
  o.zmask = o.xmask + 2 * int(p.xmask.size());
  o.out_offset = o.zmask + 2 * int(p.zmask.size());
  o.angle_idx = o.out_offset + int(p.out_offset.size());
  o.sv_run_end = o.angle_idx + p.n_steps + kAngleIdxPad;
  o.total = o.sv_run_end + p.n_blocks();
  return o;
}

std::vector<int32_t> pack_device_tables(const Plan& p) {
  TableOffsets o = table_offsets(p);
  std::vector<int32_t> t(o.total, 0);
  std::memcpy(t.data() + o.steps, p.steps.data(), sizeof(mbqc_step) * p.steps.size());
  const int H = 1 << p.ns;
  for (int j = 0; j < H; ++j) {
    t[o.in_index + j] = p.in_index[j];
    t[o.init_sign + j] = p.init_sign[j];
    t[o.out_pos + j] = p.out_pos[j];
  }
  rescale_schedule(p, kSvRescaleAppends, t.data() + o.sv_rescale, t.data() + o.e_final);
  rescale_schedule(p, kDmRescaleAppends, t.data() + o.dm_rescale, t.data() + o.e_final + 1);
  if (!p.coef.empty()) std::memcpy(t.data() + o.coef, p.coef.data(), sizeof(float) * p.coef.size());
  if (!p.xmask.empty()) {
    std::memcpy(t.data() + o.xmask, p.xmask.data(), sizeof(uint64_t) * p.xmask.size());
    std::memcpy(t.data() + o.zmask, p.zmask.data(), sizeof(uint64_t) * p.zmask.size());
    std::memcpy(t.data() + o.out_offset, p.out_offset.data(), sizeof(int32_t) * p.out_offset.size());
  }
  for (int k = 0; k < p.n_steps + kAngleIdxPad; ++k) {
    const int src = std::min(k, p.n_steps - 1);
    t[o.angle_idx + k] = src >= 0 ? p.steps[src].angle_index : 0;
  }
  const int nb = p.n_blocks();
  int next = nb;
  for (int b = nb - 1; b >= 0; --b) {
    t[o.sv_run_end + b] = next;
    if (t[o.sv_rescale + b]) next = b;
  }
  return t;
}

int window_map(const Plan& p, int k, WindowMap* m) {
  if (k < 0 || k > p.n_steps) return fail(MBQC_ERR_VALUE, "measurement count out of range");
  const int N = p.n_nodes, ns = p.ns;
  const int W = std::min(p.w, N - k);
  if (W > kMaxSlots + 1) return fail(MBQC_ERR_NOT_IMPLEMENTED, "window too large");
  m->W = W;
  m->nb = 0;
  // replay the slot assignment of the first k steps
  std::vector<int> slot(N, -1);
  for (int j = 0; j < ns && j < N; ++j) slot[p.schedule[j]] = ((ns - 1 - j) % ns + ns) % ns;
  for (int t = 0; t < k; ++t) {
    const int q = slot[p.order[t]];
    slot[p.order[t]] = -1;
    if (t + ns <= N - 1) slot[p.schedule[t + ns]] = q;
  }
  // the reference window lists its nodes in the order they entered: the live
  // ones in schedule order, then (lag 1) the node already appended as |+>
  std::vector<int> live;
  for (int j = 0; j < N; ++j)
    if (slot[p.schedule[j]] >= 0) live.push_back(p.schedule[j]);
  const int n_live = int(live.size());
  const bool pending = p.lag == 1 && k + ns <= N - 1;
  if (n_live + (pending ? 1 : 0) != W) return fail(MBQC_ERR_RUNTIME, "window bookkeeping mismatch");
  for (int i = 0; i < n_live; ++i) m->slot[i] = slot[live[i]];
  if (pending) {
    m->slot[n_live] = -1;
    const int v = p.schedule[k + ns];
    for (int x : p.nbr[v])
      for (int t = 0; t < n_live; ++t)
        if (live[t] == x) m->nb |= 1u << (W - 1 - t);
  }
  for (int i = W; i < kMaxSlots + 1; ++i) m->slot[i] = -1;
  return MBQC_OK;
}

}  // namespace mbqc

using namespace mbqc;

extern "C" {

int mbqc_plan_create(const mbqc_pattern_desc* desc, mbqc_plan** out) {
  if (!out) return fail(MBQC_ERR_VALUE, "out is null");
  *out = nullptr;
  if (!desc) return fail(MBQC_ERR_VALUE, "null pattern description");
  mbqc_pattern_desc d = *desc;
  std::vector<int> flow, layer, order;
  if (!d.schedule) {
    // schedule from the causal flow, as a flow-carrying MBQCircuit gets it
    if (d.n_nodes <= 0) return fail(MBQC_ERR_VALUE, "pattern has no nodes");
    if ((d.n_edges > 0 && !d.edges) || (d.n_inputs > 0 && !d.input_nodes) ||
        (d.n_outputs > 0 && !d.output_nodes))
      return fail(MBQC_ERR_VALUE, "null node or edge list");
    bool found = false;
    int rc = causal_flow(d.n_nodes, d.n_edges, d.edges, d.n_inputs, d.input_nodes, d.n_outputs,
                         d.output_nodes, &flow, &layer, &order, &found);
    if (rc != MBQC_OK) return rc;
    if (!found)
      return fail(MBQC_ERR_VALUE, "Schedule must be provided for numpy simulator as the MBQCircuit "
                                  "does not have a flow.");
    // unmeasured (output) nodes go last, in their order of appearance
    std::vector<char> is_out(d.n_nodes, 0);
    for (int m = 0; m < d.n_outputs; ++m) is_out[d.output_nodes[m]] = 1;
    std::stable_partition(order.begin(), order.end(), [&](int v) { return !is_out[v]; });
    d.schedule = order.data();
    d.n_measure = d.n_nodes - d.n_outputs;
    if (d.window <= 1) d.window = d.n_inputs + 1;
    if (!d.flow) d.flow = flow.data();
  }
  auto* p = new Plan();
  int rc = compile_plan(&d, p);
  if (rc == MBQC_OK && g_plan_upload_hook) rc = g_plan_upload_hook(p);
  if (rc != MBQC_OK) {
    delete p;
    return rc;
  }
  *out = reinterpret_cast<mbqc_plan*>(p);
  return MBQC_OK;
}

int32_t mbqc_plan_window_qubits(const mbqc_plan* plan, int32_t k) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  if (k < 0 || k > p->n_steps) return fail(MBQC_ERR_VALUE, "measurement count out of range");
  return std::min(p->w, p->n_nodes - k);
}

int mbqc_plan_get_info(const mbqc_plan* plan, mbqc_plan_info* info) {
  if (!plan || !info) return fail(MBQC_ERR_VALUE, "null argument");
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  info->n_slots = p->ns;
  info->lag = p->lag;
  info->n_steps = p->n_steps;
  info->n_inputs = p->n_in;
  info->n_outputs = p->n_out;
  info->n_angles = p->n_angles;
  info->n_nodes = p->n_nodes;
  info->window = p->w;
  info->n_appends = p->n_appends;
  info->n_dropped_edges = p->n_dropped;
  info->min_window = p->min_window;
  info->can_sample = p->can_sample ? 1 : 0;
  info->periodic = p->periodic ? 1 : 0;
  info->general_planes = p->general_planes ? 1 : 0;
  info->shared_angles = p->shared_angles ? 1 : 0;
  info->rotating = p->rotating ? 1 : 0;
  return MBQC_OK;
}

int mbqc_plan_get_tables(const mbqc_plan* plan, mbqc_step* steps, int32_t* in_index,
                         int32_t* init_sign, int32_t* out_pos) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  const int H = 1 << p->ns;
  if (steps) std::memcpy(steps, p->steps.data(), sizeof(mbqc_step) * p->steps.size());
  if (in_index) std::memcpy(in_index, p->in_index.data(), sizeof(int32_t) * H);
  if (init_sign) std::memcpy(init_sign, p->init_sign.data(), sizeof(int32_t) * H);
  if (out_pos) std::memcpy(out_pos, p->out_pos.data(), sizeof(int32_t) * H);
  return MBQC_OK;
}

int mbqc_plan_get_rescale(const mbqc_plan* plan, int32_t* sv_flags, int32_t* dm_flags,
                          int32_t* e_final) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  std::vector<int32_t> t = pack_device_tables(*p);
  TableOffsets o = table_offsets(*p);
  if (sv_flags) std::memcpy(sv_flags, t.data() + o.sv_rescale, sizeof(int32_t) * p->n_blocks());
  if (dm_flags) std::memcpy(dm_flags, t.data() + o.dm_rescale, sizeof(int32_t) * p->n_blocks());
  if (e_final) std::memcpy(e_final, t.data() + o.e_final, sizeof(int32_t) * 2);
  return MBQC_OK;
}

const char* mbqc_last_error(void) { return g_last_error.c_str(); }

int32_t mbqc_abi_version(void) { return 3; }

}  // extern "C"
