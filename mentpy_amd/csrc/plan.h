// Internal plan representation shared by the host compiler and the launchers.
#pragma once
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mbqc_hip.h"

namespace mbqc {

constexpr int kMaxSlots = 20;   // host tables; kernels support fewer (see launchers)
constexpr int kMaxDevices = 16;
constexpr int kMaxCoefSlots = 6;
constexpr int kMaxSampleSlots = 20;   // sampled outcomes (LDS kernel above 6 slots, global above 14)
// largest pair-coefficient table kept with one row per step
constexpr size_t kCoefPerStepBytes = 8192;

// Growth control: the kernels skip the 1/sqrt(2) of every appended |+>, so the
// SV state grows by sqrt(2) (the DM by 2) per append; at the start of a block
// whose running exponent reached the threshold the state is multiplied by
// 2^-32, an exact power of two.
constexpr int kSvRescaleAppends = 64;   // 2^(64/2) growth -> rescale by 2^-32
constexpr int kDmRescaleAppends = 32;   // 2^32 growth     -> rescale by 2^-32

// Neighbour mask of sub-step `sub` (step k = blk * ns + sub) of a periodic
// plan: the appended qubit is a CZ neighbour of the measured one and, for
// sub > 0, of the qubit appended one step earlier (slot ns - sub).  This is
// the step structure of a grid of ns rows in column order (a chain for ns = 1).
constexpr uint32_t periodic_nmask(int ns, int sub) { return sub == 0 ? 0u : (1u << (ns - sub)); }

struct Plan {
  int ns = 0, lag = 0, w = 0, n_nodes = 0, n_steps = 0, n_in = 0, n_out = 0, n_angles = 0;
  int n_appends = 0, n_edges = 0, n_dropped = 0, min_window = 0;
  int e_sv = 0, e_dm = 0;   // final growth exponents after the rescale schedule
  std::vector<mbqc_step> steps;
  std::vector<int32_t> in_index, init_sign, out_pos, out_slot;
  std::vector<int32_t> schedule;          // node order, measured nodes first
  // node measured at each step: schedule[k] unless the description gave a
  // measurement order (the reference's dev_mode picks nodes out of order)
  std::vector<int32_t> order;
  // step k measures slot (n_slots - 1 - k) mod n_slots (always the case in
  // schedule order): the register kernels take the measured slot at compile
  // time; other plans run on the kernels that read it from the step flags
  bool rotating = true;
  std::vector<std::vector<int>> nbr;      // adjacency (applied edges and dropped ones)
  // Pair coefficients (statevector kernels, n_slots <= kMaxCoefSlots): for
  // step k and pair i (the i-th slot index with the measured bit clear),
  // new[bit=1] = alpha * u0 + beta * x, stored as coef[(r * 2^ns) + 2i + {0,1}]
  // with r = steps[k].flags >> 16 (distinct rows stored once).
  std::vector<float> coef;
  bool coef_per_step = false;   // row of step k is k (table small enough to keep per step)
  // every step appends a CZ neighbour of the measured qubit with the other
  // neighbours periodic_nmask(ns, k % ns), and n_steps % ns == 0: the forward
  // kernel can take its coefficients at compile time
  bool periodic = false;
  // some angle-vector entry is used by more than one measured node: the
  // forward kernels allow it, the adjoint entry points reject it (each writes
  // one gradient entry per step)
  bool shared_angles = false;
  // some measured node is not in the XY plane (MBQC_PLANE_XZ / YZ / Z): only
  // the density-matrix entry points run the plan, with per-step outcome
  // probabilities
  bool general_planes = false;
  // Pauli-frame tables of the sampled mode (empty when not available): per step
  // the 64-bit masks, relative to the step's schedule position, of the nodes
  // whose X (flow successor) / Z (its later neighbours) frame bit an outcome 1
  // flips; for each output node its offset from position n_steps.
  bool can_sample = false;
  std::string sample_reason;
  std::vector<uint64_t> xmask, zmask;
  std::vector<int32_t> out_offset;

  // device copies, one per device id
  std::mutex mu;
  void* dev_tab[kMaxDevices] = {};
  int n_blocks() const { return (n_steps + ns - 1) / ns; }
};

// Set by the device code (when linked in): uploads a new plan's tables to the
// current device at creation, so that its first launch allocates nothing.
// A build without device code leaves it null.
extern int (*g_plan_upload_hook)(Plan*);

// Device table layout (int32 words), built by pack_device_tables():
//   [steps: 4*n_steps][in_index: H][init_sign: H][out_pos: H]
//   [sv_rescale: n_blocks][dm_rescale: n_blocks][sv_e_final, dm_e_final]
//   [coef: n_rows * 2^ns floats (distinct rows), only when ns <= kMaxCoefSlots]
//   [xmask, zmask: n_steps uint64 each, out_offset: n_out, only when sampling is possible]
//   [angle_idx: n_steps + kAngleIdxPad][sv_run_end: n_blocks]
// angle_idx[k] is steps[k].angle_index, contiguous (the statevector kernels
// read a run of them at once), padded with the last step's entry so a read of
// up to kAngleIdxPad entries past the last step stays in the table.
// sv_run_end[b] is the first block after b whose statevector rescale flag is
// set (n_blocks if none): the end of the stretch of blocks starting at b.
constexpr int kAngleIdxPad = 16;
struct TableOffsets {
  int steps, in_index, init_sign, out_pos, sv_rescale, dm_rescale, e_final, coef, xmask, zmask,
      out_offset, angle_idx, sv_run_end, total;
};
TableOffsets table_offsets(const Plan& p);
std::vector<int32_t> pack_device_tables(const Plan& p);

// Reference window after k measurements, in terms of the live slots: window
// qubit i (0 = oldest, the most significant bit) is the i-th unmeasured node
// of the window in schedule order (schedule[k + i] when nodes are measured in
// schedule order); it lives in register slot slot[i], or is the node the
// reference has already appended as |+> (slot -1, the last qubit) whose CZ
// neighbours among the window are the window-index bits nb.
struct WindowMap {
  int32_t W;
  int32_t slot[kMaxSlots + 1];
  uint32_t nb;
};
int window_map(const Plan& p, int k, WindowMap* m);

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

// Causal flow (flow[u] = successor, -1 for outputs), layers and measurement
// order of an open graph; *found is false when the graph has no causal flow.
int causal_flow(int n, int n_edges, const int32_t* edges, int n_in, const int32_t* inputs,
                int n_out, const int32_t* outputs, std::vector<int>* flow, std::vector<int>* layer,
                std::vector<int>* order, bool* found);
int compile_plan(const mbqc_pattern_desc* d, Plan* p);

}  // namespace mbqc
