/*
 * mbqc_hip.h - C ABI of the MI355X measurement-pattern simulator.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are
 * device (HBM) addresses of complex64 data stored as interleaved float pairs;
 * `stream` is a hipStream_t passed as void*.  Launches are stream-ordered and
 * asynchronous; a plan is immutable after creation, so concurrent launches of
 * one plan on different streams are safe.  Status codes are negative on error
 * and the message is available from mbqc_last_error() (thread-local).  A
 * forward, density-matrix or sampled-outcome launch over B = 0 angle vectors
 * does nothing and returns MBQC_OK; its data pointers may then be null.
 * Single-plan launches whose live window is beyond the LDS tiers allocate
 * their workspace in stream order on every call; on a stream under HIP graph
 * capture they return MBQC_ERR_NOT_IMPLEMENTED before recording anything
 * (replays of a recorded stream-ordered allocation were measured to read a
 * stale workspace).  Batch launches use workspaces owned by the batch and can
 * be captured at every window size.
 *
 * The reference functions each entry point replaces are cited below as
 * file:line into mentpy 0.1.0a9.
 */
#ifndef MBQC_HIP_H
#define MBQC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  VALUE/TYPE/NOT_IMPLEMENTED map to the Python exceptions the
 * reference raises for the same conditions (ValueError, TypeError,
 * NotImplementedError); RUNTIME is a HIP runtime failure. */
#define MBQC_OK 0
#define MBQC_ERR_VALUE (-1)
#define MBQC_ERR_TYPE (-2)
#define MBQC_ERR_NOT_IMPLEMENTED (-3)
#define MBQC_ERR_RUNTIME (-4)

/* Launch flags. */
#define MBQC_FLAG_REF_PHASE 1u    /* reference global phase: prod_k sgn(cos(t_k/2)) e^{i t_k/2} */
#define MBQC_FLAG_RENORMALIZE 2u  /* divide the output by its norm (SV) / trace (DM) */
#define MBQC_FLAG_WRITE_PHASE 4u  /* also write the per-sim global phase (SV forward) */
#define MBQC_FLAG_TABLE_STEPS 8u  /* SV / DM forward and batched launch: read the step
                                     coefficients and CZ masks from the plan's tables
                                     also for a periodic plan (same results) */
#define MBQC_FLAG_STEP_PROBS 16u  /* DM forward / window: evaluate every step's outcome
                                     probability (outcome 1 where outcome 0 has
                                     probability below 1e-4, division by the outcome
                                     probability per step), as np_simulator_dm.py:329-346
                                     does; always on for plans with XZ / YZ / Z nodes */

/* Measurement planes (mbqc_pattern_desc.plane; mentpy/operators/ment.py:218-253).
 * Planes X and Y are XY at fixed angles 0 and pi/2.  MBQC_PLANE_Z is traced
 * out without projection: the reference's mode="expectation" branch
 * (np_simulator_dm.py:318-322); its mode="sample" draw is not provided. */
#define MBQC_PLANE_XY 0
#define MBQC_PLANE_XZ 1
#define MBQC_PLANE_YZ 2
#define MBQC_PLANE_Z 3

/* Host description of a pattern; all arrays are host memory, copied by
 * mbqc_plan_create.  Node ids are 0..n_nodes-1. */
typedef struct mbqc_pattern_desc {
  int32_t n_nodes;
  int32_t n_edges;
  const int32_t* edges;        /* [n_edges][2] undirected resource-graph edges */
  const int32_t* schedule;     /* [n_nodes] full window order (measured nodes first) */
  int32_t n_measure;           /* schedule[0..n_measure) are measured, in this order */
  int32_t window;              /* live-window size w (>= 1) */
  int32_t n_inputs;
  const int32_t* input_nodes;  /* input-state qubit order, MSB first */
  int32_t n_outputs;
  const int32_t* output_nodes; /* output-state qubit order, MSB first */
  const int32_t* angle_index;  /* [n_nodes] index into the angle vector, or -1 */
  const float* fixed_angle;    /* [n_nodes] XY angle used when angle_index < 0 */
  int32_t n_angles;            /* length of one angle vector */
  const int32_t* flow;         /* optional [n_nodes]: causal-flow successor, -1 for outputs;
                                  needed only for sampled outcomes */
  const int32_t* plane;        /* optional [n_nodes]: MBQC_PLANE_* of each measured node
                                  (NULL: every node XY); a plan with another plane runs
                                  only in the density-matrix entry points */
  const int32_t* measure_order;/* optional [n_measure]: the node measured at step k (NULL:
                                  schedule[k]).  A permutation of schedule[0..n_measure)
                                  in which step k's node has entered the window
                                  (position <= k + window - 1): the order mentpy's
                                  dev_mode scheduler picks (np_simulator_sv.py:176-197,
                                  254-276).  Nodes still enter the window in schedule
                                  order.  A plan whose measured slots leave the fixed
                                  rotation (plan_info.rotating = 0) runs on the LDS
                                  kernels; its adjoint, sampled and batched launches
                                  are not available. */
} mbqc_pattern_desc;

typedef struct mbqc_plan mbqc_plan;

typedef struct mbqc_plan_info {
  int32_t n_slots;        /* register slots of the live state (w-1 or w) */
  int32_t lag;            /* 1: appended qubit entangled before the oldest is measured */
  int32_t n_steps;        /* measured nodes */
  int32_t n_inputs, n_outputs, n_angles, n_nodes, window;
  int32_t n_appends;      /* |+> qubits appended over the run (padding included) */
  int32_t n_dropped_edges;/* edges never applied because the window was too small */
  int32_t min_window;     /* smallest window that applies every edge */
  int32_t can_sample;     /* 1: mbqc_sv_sample_c64 accepts this plan */
  int32_t periodic;       /* 1: every step appends a CZ neighbour of the measured
                             qubit and the CZ pattern repeats every n_slots steps
                             (grids in column order, chains): the SV forward
                             takes its pair coefficients at compile time */
  int32_t general_planes; /* 1: some node is measured in XZ / YZ / Z (density-matrix
                             entry points only) */
  int32_t shared_angles;  /* 1: some angle-vector entry drives several nodes (the
                             adjoint entry points refuse the plan) */
  int32_t rotating;       /* 1: step k measures slot (n_slots-1-k) mod n_slots (always
                             so in schedule order); 0: a measure_order leaves that
                             rotation */
} mbqc_plan_info;

/* Per-step table entry (16 bytes) as stored on the device. */
typedef struct mbqc_step {
  int32_t angle_index;    /* >= 0: angle vector entry; < 0: use fixed_angle */
  uint32_t flags;         /* bit0 append, bit1 measured qubit is a CZ neighbour of
                             the appended one, bits 2..3 measurement plane
                             (MBQC_PLANE_*), bits 8..15 measured slot, bits
                             16..31 row of the step's pair coefficients */
  uint32_t nmask;         /* slots that are CZ neighbours of the appended qubit */
  float fixed_angle;
} mbqc_step;

/* Noise of the density-matrix path.  The channels of the PennyLane backend's
 * circuit_noise (mentpy/simulators/pennylane_simulator.py:123-136) act on
 * every qubit after all CZs; nonzero ones are composed in this field order.
 * p_readout flips the reported outcome of every measured qubit (the run is
 * post-selected on reported outcome 0).  All values lie in [0, 1]. */
typedef struct mbqc_noise {
  float p_depol;        /* DepolarizingChannel(p) */
  float p_readout;      /* classical readout flip */
  float amp_damping;    /* AmplitudeDamping(gamma) */
  float phase_damping;  /* PhaseDamping(gamma) */
  float phase_flip;     /* PhaseFlip(p) */
  float gad_gamma;      /* GeneralizedAmplitudeDamping(gamma, p) */
  float gad_p;
} mbqc_noise;

/* Causal flow of an open graph (nodes 0..n_nodes-1): replaces find_cflow /
 * causal_flow_aux (mentpy/mbqc/flow.py:115-185) and the measurement order of
 * MBQCircuit.calculate_order (mentpy/mbqc/mbqcircuit.py:390-422).  Writes
 * flow[v] (flow successor, -1 for outputs), layer[v] (0 for outputs) and
 * order[0..n_nodes) (inputs first, then decreasing layer, ties in node order);
 * *found = 0 when the graph has no causal flow.  |I| != |O| is
 * MBQC_ERR_VALUE with the reference's message.  Any output pointer may be
 * NULL.  Host only. */
int mbqc_causal_flow(int32_t n_nodes, int32_t n_edges, const int32_t* edges, int32_t n_inputs,
                     const int32_t* input_nodes, int32_t n_outputs, const int32_t* output_nodes,
                     int32_t* flow, int32_t* layer, int32_t* order, int32_t* found);

/* Compile a pattern into per-step tables; replaces the set-up of
 * NumpySimulatorSV.__init__ / NumpySimulatorDM.__init__
 * (mentpy/simulators/np_simulator_sv.py:38-128, np_simulator_dm.py:33-115):
 * window, input reordering, |+> padding, initial CZ block.  Host only.
 * With schedule == NULL the schedule is the causal-flow order of
 * mbqc_causal_flow with the output nodes (the unmeasured ones) last,
 * n_measure = n_nodes - n_outputs, a window <= 1 becomes n_inputs + 1 and a
 * NULL flow is the causal flow (a pattern without one is MBQC_ERR_VALUE). */
int mbqc_plan_create(const mbqc_pattern_desc* desc, mbqc_plan** out);
/* mbqc_plan_create also copies the plan's tables to the current device when
 * a GPU is present, so the plan's launches on that device allocate nothing
 * (and the first one can be captured in a hipGraph).  mbqc_plan_upload does
 * the same for the current device when it was not done yet (another device);
 * it allocates device memory and copies synchronously, so call it before any
 * capture. */
int mbqc_plan_upload(const mbqc_plan* plan);
void mbqc_plan_destroy(mbqc_plan* plan);
int mbqc_plan_get_info(const mbqc_plan* plan, mbqc_plan_info* info);
/* Copy host tables out (tests): steps[n_steps], in_index[2^n_slots],
 * init_sign[2^n_slots] (0/1), out_pos[2^n_slots] (-1 for dead slots). */
int mbqc_plan_get_tables(const mbqc_plan* plan, mbqc_step* steps, int32_t* in_index,
                         int32_t* init_sign, int32_t* out_pos);
/* Growth-control schedule (tests): per block of n_slots steps, whether the
 * state is multiplied by 2^-32 at its start (SV and DM variants), and the
 * final sqrt(2)-exponent (SV) / 2-exponent (DM) of the unnormalized state. */
int mbqc_plan_get_rescale(const mbqc_plan* plan, int32_t* sv_flags, int32_t* dm_flags,
                          int32_t* e_final);

/* Statevector forward, complex64: replaces NumpySimulatorSV.run / measure /
 * measure_ment (np_simulator_sv.py:164-297, 322-358) and the partial trace it
 * calls (mentpy/calculator/state_ops.py:42-74).  Simulation g = b*S + s uses
 * angle a of vector b at angles[b*angle_stride + a*angle_stride_a]
 * (angle_stride 0 = one shared vector; angle-major storage, angle_stride 1 and
 * angle_stride_a = B, gives coalesced loads) and input s at
 * psi_in + s*in_stride amplitudes (in_stride 0 = shared).  out is
 * [B*S][2^n_outputs]; phase_out (optional, MBQC_FLAG_WRITE_PHASE) is [B*S].
 * Live windows up to 20 slots: up to 6 with the state in registers, 7..14 in
 * LDS, 15..20 in a workspace of 2^n_slots complex64 per simulation that the
 * call allocates and frees in stream order (hipMallocAsync / hipFreeAsync on
 * `stream`, at most 4 GiB per launch, larger batches in chunks). */
int mbqc_sv_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                        int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
                        void* phase_out, int64_t B, int64_t S, uint32_t flags, void* stream);

/* Sampled outcomes with flow corrections (force0=False), complex64: the
 * semantics of the PennyLane backend (mentpy/simulators/pennylane_simulator.py:138-153):
 * every measurement outcome is drawn (probability 1/2 each on patterns with a
 * causal flow and no dropped CZ), an outcome 1 on node i applies X to flow(i)
 * and Z to the neighbours of flow(i) measured after i.  The corrections are
 * tracked as a Pauli frame (no state update) and the residual frame of the
 * output qubits is applied to the output at the end, so the output equals the
 * post-selected one up to a global phase.  outcomes is [B*S][W+2] int32 with
 * W = (n_steps+31)/32: words 0..W-1 hold the physical outcomes (bit k%32 of
 * word k/32 = outcome of the k-th measured node), words W and W+1 the X and Z
 * frame applied to the output (bit n_outputs-1-m = output node m).  Draws are
 * a counter-based hash of (seed, simulation, step).  Needs the flow in the
 * pattern description, no dropped edges and n_slots <= 20 (else
 * MBQC_ERR_NOT_IMPLEMENTED with the reason); windows of 7..14 slots run one
 * workgroup per simulation with the state in LDS, 15..20 the same with the
 * state in a workspace allocated and freed in stream order. */
int mbqc_sv_sample_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                       int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
                       int32_t* outcomes, int64_t B, int64_t S, uint64_t seed, void* stream);

/* Step-by-step API: the reference's window state after its first k
 * measurements, as NumpySimulatorSV.measure / NumpySimulatorDM.measure leave
 * it in qstate (np_simulator_sv.py:164-225, np_simulator_dm.py:151-216).
 * The window holds nodes schedule[k .. k+W-1], W = mbqc_plan_window_qubits,
 * oldest first (most significant bit), the next node already appended as |+>
 * with its CZs.  Measured node j (schedule order) uses
 * step_angles[b*step_stride + j], so any angle can be given per step; inputs as
 * in the forward calls.  The state is normalised (SV: MBQC_FLAG_REF_PHASE
 * selects the reference's global phase).  out: [B*S][2^W] (SV) or
 * [B*S][2^W][2^W] (DM; noise of the measured qubits as in the full run).
 * DM outcomes (optional, may be NULL): [B*S][k] float32, the outcomes of the
 * k measurements as mbqc_dm_forward_outcomes_c64 reports them; given, or for
 * a plan with XZ / YZ / Z nodes, every step's probability is evaluated. */
int32_t mbqc_plan_window_qubits(const mbqc_plan* plan, int32_t k);
int mbqc_sv_window_c64(const mbqc_plan* plan, const float* step_angles, int64_t step_stride,
                       const void* psi_in, int64_t in_stride, int32_t k, void* out, int64_t B,
                       int64_t S, uint32_t flags, void* stream);
int mbqc_dm_window_c64(const mbqc_plan* plan, const float* step_angles, int64_t step_stride,
                       const void* psi_in, int64_t in_stride, int32_t k, void* out, int64_t B,
                       int64_t S, const mbqc_noise* noise, float* outcomes, void* stream);

/* Density-matrix forward, complex64: replaces NumpySimulatorDM.run / measure /
 * measure_ment (np_simulator_dm.py:151-283, 307-346) and
 * partial_trace_mixed_state (state_ops.py:77-119); the noise channels follow
 * the placement of the PennyLane backend (pennylane_simulator.py:118-136).
 * Inputs are state vectors (the reference purifies them, np_simulator_dm.py:89-92).
 * out is [B*S][2^n_outputs][2^n_outputs].  Live windows up to 10 slots: up to
 * 7 with the density matrix in registers or LDS, 8..10 in a workspace of
 * 4^n_slots complex64 per simulation allocated and freed in stream order
 * (as mbqc_sv_forward_c64). */
int mbqc_dm_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                        int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
                        int64_t B, int64_t S, const mbqc_noise* noise, uint32_t flags, void* stream);
/* The same with every step's outcome probability evaluated (MBQC_FLAG_STEP_PROBS
 * implied) and the outcomes written, as NumpySimulatorDM.run records them in
 * its `outcomes` dict (np_simulator_dm.py:243-250, 318-338): outcomes is
 * [B*S][n_steps] float32, 0 or 1 for a projected node (1 where outcome 0 had
 * probability below 1e-4) and the expectation p1 for a Z node. */
int mbqc_dm_forward_outcomes_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                 int64_t angle_stride_a, const void* psi_in, int64_t in_stride,
                                 void* out, float* outcomes, int64_t B, int64_t S,
                                 const mbqc_noise* noise, uint32_t flags, void* stream);

/* Adjoint angle gradient (replaces the 2n cost evaluations of fd_gradient /
 * psr_gradient, mentpy/gradients/_finite_difference.py:9-24,
 * _parameter_shift.py:9-25, behind get_gradient, grad.py:13-33).
 * Forward pass: like mbqc_sv_forward_c64 in the raw phase frame, unnormalized,
 * also storing one checkpoint of the live state per block of n_slots steps in
 * `ckpt` (mbqc_sv_adjoint_workspace_bytes bytes).  Columns c = 0..C-1 of
 * parameter set b are simulations b*C + c.
 * All adjoint entry points need plans in schedule order (rotating slots)
 * and every angle-vector entry used by at most one measured node
 * (angle_index values unique).  These two take live windows of up to 16
 * qubits: up to 5 with the state in registers, 6..10 with one workgroup per
 * simulation (forward) or partial-sum row (backward) and the state in LDS
 * (the backward pass keeps n_angles floats of per-angle sums in LDS too, so
 * a 10-slot window takes up to about 26,000 angles), 11..16 the same with the
 * states in a workspace the call allocates and frees in stream order
 * (hipMallocAsync on `stream`); the segmented ones below take up to 5.  Other plans are
 * MBQC_ERR_NOT_IMPLEMENTED. */
size_t mbqc_sv_adjoint_workspace_bytes(const mbqc_plan* plan, int64_t n_columns);
int mbqc_sv_adjoint_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                int64_t angle_stride_a, const void* psi_in, int64_t in_stride,
                                void* out, int64_t B, int64_t C, void* ckpt, size_t ckpt_bytes,
                                void* stream);
/* Backward sweep: seed[B*C][2^n_outputs] is dL/d(conj out) in the convention
 * dL = Re sum conj(seed) d(out).  Writes grad_partial[B][P][n_angles] with
 * P = mbqc_sv_adjoint_partials(C); the gradient is the sum over P.
 * Entries of non-trainable angles are left untouched. */
int32_t mbqc_sv_adjoint_partials(int64_t C);
int mbqc_sv_adjoint_backward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                 int64_t angle_stride_a, const void* seed, int64_t B, int64_t C,
                                 const void* ckpt, size_t ckpt_bytes, float* grad_partial,
                                 void* stream);

/* Time-segmented transfer-matrix adjoint (no reference counterpart: the
 * reference differentiates by finite differences, mentpy/gradients/
 * _finite_difference.py:9-24).  The columns are the 2^n_inputs input basis
 * states, as in mbqc_sv_adjoint_forward_c64 with psi_in = identity and
 * C = 2^n_inputs.  The pattern's blocks of n_slots steps are cut into
 * n_segments segments whose transfer matrices are computed concurrently and
 * chained per parameter set, so that one parameter set occupies n_segments
 * times as many lanes for 1/n_segments of the steps each (small batches fill
 * the chip).  2 <= n_segments <= min(16, number of blocks).
 *
 * mbqc_sv_segments_auto: the segment count the library picks for B sets
 * (1 means: use the unsegmented entry points).
 * Forward: out[B][2^n_inputs][2^n_outputs] = T_b (column c = image of input
 * basis state c).  Backward: seed in the same layout (dL = Re sum conj(seed)
 * d(out)); writes grad[B][n_angles] (entries of non-trainable angles are left
 * untouched).  The backward needs the workspace of the forward of the same
 * angles. */
int32_t mbqc_sv_segments_auto(const mbqc_plan* plan, int64_t B);
size_t mbqc_sv_segmented_workspace_bytes(const mbqc_plan* plan, int64_t B, int32_t n_segments);
int mbqc_sv_segmented_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                  int64_t angle_stride_a, void* out, int64_t B, int32_t n_segments,
                                  void* workspace, size_t workspace_bytes, void* stream);
int mbqc_sv_segmented_backward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                   int64_t angle_stride_a, const void* seed, int64_t B,
                                   int32_t n_segments, void* workspace, size_t workspace_bytes,
                                   float* grad, void* stream);

/* Batched forward of many plans in one persistent launch (no reference
 * counterpart: the reference runs one simulation per Python call).  Plans are
 * first bound into a batch, which packs their tables on the device. */
typedef struct mbqc_batch mbqc_batch;
typedef struct mbqc_batch_job {
  const mbqc_plan* plan;
  const float* angles;   /* device, angle a of vector b at b*angle_stride + a*angle_stride_a */
  int64_t angle_stride;
  int64_t angle_stride_a;
  const void* psi_in;    /* device [S][2^n_inputs] */
  int64_t in_stride;
  void* out;             /* device [B*S][2^n_outputs] */
  int64_t B, S;
} mbqc_batch_job;
int mbqc_batch_create(const mbqc_batch_job* jobs, int32_t n_jobs, int32_t sims_per_item,
                      mbqc_batch** out);
void mbqc_batch_destroy(mbqc_batch* batch);
int64_t mbqc_batch_items(const mbqc_batch* batch);
/* Launch the persistent kernel: `waves` resident wavefronts pull work items
 * from a device counter (reset by a kernel, so the launch pair can be captured
 * in a hipGraph), `wg_waves` wavefronts per workgroup. */
int mbqc_batch_launch(mbqc_batch* batch, int32_t waves, int32_t wg_waves, uint32_t flags,
                      void* stream);
/* Work items of the group of live-window size n_slots (0..20; -1 for a null
 * batch or n_slots out of range).  Each non-empty group runs its own
 * persistent kernel, concurrently with the others: up to 6 slots one
 * simulation per lane, 7..14 slots one simulation at a time per workgroup
 * with the state in LDS (2^(n_slots-1) threads, at most 256; its resident
 * waves make waves / (threads / 64) workgroups), 15..20 the same with the
 * states in a workspace that mbqc_batch_create allocates with the batch
 * (256 threads, at most 256 resident workgroups per group: up to 256 * 2^n_slots
 * complex64, 2 GiB for a 20-slot group), so every batch launch allocates
 * nothing and can be captured.  Launches of one batch must not overlap (they
 * share its counters and workspaces). */
int64_t mbqc_batch_group_items(const mbqc_batch* batch, int32_t n_slots);
/* mbqc_batch_launch with the resident wavefront count chosen per group:
 * group_waves[ns] (n_entries <= 21 entries, indexed by live-window size) for
 * the group of size ns; 0 skips the group (its outputs are not written), a
 * positive count must be at least wg_waves.  Sizing the groups by their
 * work makes the concurrent kernels finish together. */
int mbqc_batch_launch_groups(mbqc_batch* batch, const int32_t* group_waves, int32_t n_entries,
                             int32_t wg_waves, uint32_t flags, void* stream);

/* MB-QML infidelity loss and adjoint seed from the transfer matrices, fused
 * (the Gram form of 1 - mean_s |<phi_s|T psi_s>|^2; the reference evaluates
 * calculator.fidelity per pair, mentpy/calculator/borrows.py:4-6, inside the
 * tutorial loss, docs/tutorials/intro-to-mbqml.rst:35-54): for each parameter
 * set b with t_b = vec(T_b) ([B][K] complex64, K = 2^(|I|+|O|)),
 * z = M t_b, loss[b] = 1 - Re(t_b^H z)/S, seed[b] = -2 z/S, where MT is the
 * transposed Hermitian data Gram matrix ([K][K]).  workspace:
 * mbqc_gram_workspace_bytes(B, K) bytes of device memory.  K <= 512. */
size_t mbqc_gram_workspace_bytes(int64_t B, int32_t K);
/* The same loss and seed when z = M t ([B][K]) was computed elsewhere (a
 * library GEMM, faster than the fused kernel for many thousands of sets):
 * loss[b] = 1 - Re(t_b^H z_b)/S, seed[b] = -2 z_b/S.  Any K. */
int mbqc_gram_finish_c64(const void* t, const void* z, int64_t B, int32_t K, int32_t S, float* loss,
                         void* seed, void* stream);
int mbqc_gram_loss_seed_c64(const void* t, const void* MT, int64_t B, int32_t K, int32_t S,
                            float* loss, void* seed, float* workspace, void* stream);
/* mbqc_gram_loss_seed_c64 without a workspace argument and without any
 * allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory, such as code under HIP graph capture or a serving
 * loop.  The loss partials (B x K/16 floats, B x ceil(K/64) when K is not a
 * multiple of 16) go to a static 1 MiB device array (262,144 floats) of the
 * library on each device, held until the process ends; larger B is
 * MBQC_ERR_VALUE.  The first call on a device looks the array up (no
 * allocation).  Every partial is written before it is read, so nothing is
 * cleared between calls; calls that overlap on one device share the array and
 * race. */
int mbqc_gram_loss_seed_c64_noalloc(const void* t, const void* MT, int64_t B, int32_t K,
                                    int32_t S, float* loss, void* seed, void* stream);

/* MB-QML infidelity loss from propagated inputs (no reference counterpart as
 * one kernel: the reference evaluates calculator.fidelity per pair,
 * mentpy/calculator/borrows.py:4-6, inside the tutorial loss,
 * docs/tutorials/intro-to-mbqml.rst:35-54): out is [B][S][O] (input s of
 * parameter set b propagated), targets [S][O], O a power of two;
 * loss[b] = 1 - (1/S) sum_s |<targets_s|out_bs>|^2 and, if seed is not
 * NULL, seed[b][s][o] = -2/S <targets_s|out_bs> targets[s][o] (the adjoint
 * seed dL/d conj(out)). */
int mbqc_overlap_loss_seed_c64(const void* out, const void* targets, int64_t B, int32_t S,
                               int32_t O, float* loss, void* seed, void* stream);

/* Device-resident optimizer steps on n float32 parameters (all pointers device
 * memory, updated in place), replacing the host updates of
 * AdamOptimizer.step (mentpy/optimizers/adam.py:54-66) and SGDOptimizer.step
 * (mentpy/optimizers/sgd.py:49-59).  `step` is a device int32 holding the
 * number of Adam steps taken (bias correction uses step+1); the call advances
 * it, so a captured graph of training iterations replays correctly. */
int mbqc_adam_step_f32(float* x, const float* grad, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, int32_t* step, void* stream);
int mbqc_sgd_step_f32(float* x, const float* grad, float* v, int64_t n, float lr, float momentum,
                      int32_t nesterov, void* stream);

const char* mbqc_last_error(void);
int32_t mbqc_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* MBQC_HIP_H */
