/*
 * cmtj_amd.h -- C ABI of the MI355X macrospin (s)LLGS engine.
 *
 * The library integrates macrospin junctions (one junction instance per GPU
 * lane) with the semantics of cmtj's Junction::runSimulation fixed-step
 * branch (reference: core/junction.hpp:1948-1981).  Everything crossing this
 * boundary is plain C: POD descriptors, raw pointers and sizes, int status
 * codes plus a thread-local message.  No torch or C++ types appear here.
 *
 * Units follow the reference: Ms in tesla, fields in A/m, K in J/m^3,
 * J in J/m^2, gamma in m/(A s).
 */
#ifndef CMTJ_AMD_H_
#define CMTJ_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CMTJ_ABI_VERSION 6
#define CMTJ_MAX_LAYERS 8
#define CMTJ_MAX_STACK 64 /* junctions of one coupled stack (one wave) */
#define CMTJ_MAX_GROUP 64 /* elements of one dipole-coupled group (one workgroup) */

/* ScalarDriver waveform kinds; numbering of reference UpdateType
 * (core/drivers.hpp:18-29).  CMTJ_TABLE is this library's tabulated form of
 * the reference's Python-callback driver (custom = 100). */
enum cmtj_update_type {
  CMTJ_CONSTANT = 0,
  CMTJ_PULSE = 1,
  CMTJ_SINE = 2,
  CMTJ_STEP = 3,
  CMTJ_POSINE = 4,
  CMTJ_HALFSINE = 5,
  CMTJ_TRAPEZOID = 6,
  CMTJ_GAUSSIMPULSE = 7,
  CMTJ_GAUSSSTEP = 8,
  CMTJ_TABLE = 101
};

/* One ScalarDriver (reference Driver<T> fields, core/drivers.hpp:32-35, plus
 * ScalarDriver edgeTime/steadyTime, :67-68).  value(t) = constant + f(t). */
typedef struct cmtj_driver {
  int32_t type;        /* enum cmtj_update_type */
  int32_t table_index; /* CMTJ_TABLE: index into the run's driver tables */
  double constant;
  double amplitude;
  double frequency;
  double phase;
  double period;
  double cycle;
  double t_start;
  double t_stop;
  double edge_time;   /* trapezoid edge; sigma of the gaussian kinds */
  double steady_time;
} cmtj_driver;

/* Driver slots of one layer.  Axial drivers occupy three consecutive slots
 * (x, y, z).  Reference members: core/junction.hpp:148-165. */
enum cmtj_slot {
  CMTJ_SLOT_HEXT = 0,       /* externalFieldDriver            (3) */
  CMTJ_SLOT_HOE = 3,        /* HoeDriver                      (3) */
  CMTJ_SLOT_HDMI = 6,       /* HdmiDriver                     (3) */
  CMTJ_SLOT_HRES = 9,       /* HreservedInteractionFieldDriver(3) */
  CMTJ_SLOT_K = 12,         /* anisotropyDriver */
  CMTJ_SLOT_K2 = 13,        /* secondOrderAnisotropyDriver */
  CMTJ_SLOT_J_TOP = 14,     /* IECDriverTop */
  CMTJ_SLOT_J_BOTTOM = 15,  /* IECDriverBottom */
  CMTJ_SLOT_J2_TOP = 16,    /* IECQuadDriverTop */
  CMTJ_SLOT_J2_BOTTOM = 17, /* IECQuadDriverBottom */
  CMTJ_SLOT_CURRENT = 18,   /* currentDriver */
  CMTJ_SLOT_FL = 19,        /* fieldLikeTorqueDriver */
  CMTJ_SLOT_DL = 20,        /* dampingLikeTorqueDriver */
  CMTJ_SLOT_FL2 = 21,       /* fieldLikeTorqueDriver2 */
  CMTJ_SLOT_DL2 = 22,       /* dampingLikeTorqueDriver2 */
  CMTJ_SLOT_TEMPERATURE = 23,
  CMTJ_SLOT_IDMI_TOP = 24,    /* IDMIDriverTop    (3) */
  CMTJ_SLOT_IDMI_BOTTOM = 27, /* IDMIDriverBottom (3) */
  CMTJ_NSLOT = 30
};

/* Reference enum Reference (core/junction.hpp:138). */
enum cmtj_reference { CMTJ_REF_NONE = 0, CMTJ_REF_FIXED = 1, CMTJ_REF_TOP = 2, CMTJ_REF_BOTTOM = 3 };

/* Reference enum SolverMode (core/junction.hpp:140). */
enum cmtj_solver { CMTJ_EULER_HEUN = 0, CMTJ_RK4 = 1, CMTJ_DORMAND_PRINCE = 2, CMTJ_HEUN = 3 };

/* Reference Junction::MRmode (core/junction.hpp:1144). */
enum cmtj_mr_mode { CMTJ_MR_NONE = 0, CMTJ_MR_CLASSIC = 1, CMTJ_MR_STRIP = 2 };

/* Log channel selection (bit mask).  Channel order inside one instance:
 *   for each layer l: [mx] [my mz]          (LOG_MX, LOG_MYZ)
 *   MR: CLASSIC -> R ; STRIP -> Rx, Ry      (LOG_MR; Rz is identically 0)
 *   for each layer l: 23 energy channels    (LOG_ENERGY, FP64 runs only):
 *     K K2 Jbottom Jtop I Hext(xyz) Hiec(xyz) Hanis(xyz) Hdemag(xyz)
 *     Hfl(xyz) Hdl(xyz)
 *   time                                   (LOG_TIME: Dormand-Prince runs,
 *                                            whose entry times are adaptive)
 * Entries are written every write_every steps (entry k after step k*W,
 * reference core/junction.hpp:1973-1981); Dormand-Prince runs write one
 * after each accepted step that reaches the next multiple of the write
 * frequency (:2031-2039). */
enum cmtj_log_bits {
  CMTJ_LOG_MX = 1,
  CMTJ_LOG_MYZ = 2,
  CMTJ_LOG_MR = 4,
  CMTJ_LOG_ENERGY = 8,
  CMTJ_LOG_TIME = 16
};
#define CMTJ_LOG_FULL (CMTJ_LOG_MX | CMTJ_LOG_MYZ | CMTJ_LOG_MR)
#define CMTJ_ENERGY_CHANNELS 23

/* Arithmetic of a run. */
enum cmtj_precision {
  CMTJ_FP64 = 0, /* everything in binary64 (the drop-in) */
  CMTJ_FP32 = 1  /* field/torque stages in binary32; state m, clock, driver
                    phase and resistance in binary64 */
};
enum cmtj_log_dtype { CMTJ_LOG_F64 = 0, CMTJ_LOG_F32 = 1 };

/* One layer (reference Layer<T>, core/junction.hpp:145-259).  `mag` is the
 * stored magnetisation as the reference keeps it (the constructor does not
 * normalise it, core/junction.hpp:195; setMagnetisation does, :491-498). */
typedef struct cmtj_layer {
  double mag[3];
  double anis[3];
  double Ms;
  double thickness;
  double cell_surface;
  double damping;
  double demag[9];         /* row-major 3x3 */
  double dipole_top[9];
  double dipole_bottom[9];
  double ref[3];           /* referenceLayer */
  double ref2[3];          /* secondaryReferenceLayer */
  double kappa;            /* STT damping-like scale, default 1 */
  double beta;             /* STT field-like scale */
  double slonczewski;      /* STT spacer parameter Lambda */
  double polarisation;     /* STT spin polarisation */
  double fl_torque;        /* static SOT field-like coefficient */
  double dl_torque;        /* static SOT damping-like coefficient */
  int32_t ref_type;        /* enum cmtj_reference */
  int32_t include_stt;
  int32_t include_sot;
  int32_t dynamic_sot;
  int32_t alternative_stt;
  int32_t temperature_set; /* thermal (stochastic) field active */
  cmtj_driver drv[CMTJ_NSLOT];
  /* Dormand-Prince step control (AdaptiveIntegrationParams,
   * core/junction.hpp:125-139).  dp_use_pid selects the PID controller of
   * dormand_prince_step (:953-961): every attempt sets e = log(error ratio),
   * integral += e, factor = exp(kp e + ki integral + kd (e - prev)), prev = e.
   * integral and prev are state that the reference keeps in the layer from
   * run to run; a run starts from the values given here and reports the
   * final ones after the magnetisations (cmtj_run, cmtj_plan_execute). */
  double dp_abs_tol, dp_rel_tol, dp_max_factor, dp_min_factor, dp_safety;
  int32_t dp_use_pid;
  int32_t dp_pad;
  double dp_kp, dp_ki, dp_kd;
  double dp_integral_error, dp_prev_error_ratio;
} cmtj_layer;

/* One junction instance of a run.  n_steps and write_every are the FP64
 * truncations (unsigned)(totalTime/timeStep), (unsigned)(writeFrequency/
 * timeStep) of core/junction.hpp:1956-1959; n_entries = ceil(n_steps /
 * write_every).  The instance's log block starts at element log_offset of the
 * log buffer and is laid out [channel][entry]. */
typedef struct cmtj_junction {
  int32_t n_layers;
  int32_t mr_mode;       /* enum cmtj_mr_mode */
  int32_t solver;        /* enum cmtj_solver, after the temperature switch */
  int32_t log_mask;      /* enum cmtj_log_bits */
  uint32_t n_steps;
  uint32_t write_every;
  uint32_t n_entries;
  uint32_t layer_offset; /* first row of this junction in the layer array */
  uint64_t seed;         /* Philox key of the thermal field */
  uint64_t stream;       /* Philox counter word (e.g. run index of the junction) */
  uint64_t log_offset;   /* in log elements */
  double dt;
  double Rp;
  double Rap;
  double Rx0[CMTJ_MAX_LAYERS];
  double Ry0[CMTJ_MAX_LAYERS];
  double AMR_X[CMTJ_MAX_LAYERS];
  double AMR_Y[CMTJ_MAX_LAYERS];
  double SMR_X[CMTJ_MAX_LAYERS];
  double SMR_Y[CMTJ_MAX_LAYERS];
  double AHE[CMTJ_MAX_LAYERS];
  /* Dormand-Prince runs (solver CMTJ_DORMAND_PRINCE, 1-4 layers): totalTime
   * and writeFrequency of runSimulation; dt is the initial step, n_steps
   * the cap on step attempts (a run that needs more ends with
   * CMTJ_ERR_LIMIT) and n_entries the log capacity, floor(T / wf) + 4. */
  double total_time;
  double write_freq;
} cmtj_junction;

/* Snapshot of reference PhysicalConstants (core/constants.hpp:6-53). */
typedef struct cmtj_constants {
  double mu0;
  double gyro;
  double TtoAm;
  double hbar;
  double echarge;
  double kB;
} cmtj_constants;

/* Per-instance outcome. */
enum cmtj_status_code {
  CMTJ_OK = 0,
  CMTJ_ERR_INVALID = 1,
  CMTJ_ERR_HIP = 2,
  CMTJ_ERR_NAN = 3, /* reference: runtime_error("NAN magnetisation"), core/junction.hpp:1037-1039 */
  CMTJ_ERR_LIMIT = 4, /* Dormand-Prince: step-attempt cap or log capacity reached */
  /* Dormand-Prince with several layers: a layer rejected its step.  The
   * reference then sets the shared step to the smallest Layer::hopt, which
   * is never assigned and stays -1 (core/junction.hpp:259, 1699-1704), forces
   * the attempt through and moves its clock back by 1 s, after which its run
   * does not finish; the run stops here instead, with the log and the
   * magnetisations of that attempt (layers that accepted keep their step). */
  CMTJ_ERR_DP_REJECTED = 5
};
typedef struct cmtj_status {
  int32_t code;
  uint32_t steps_done;    /* steps completed (the failing step excluded) */
  uint32_t entries_done;  /* log entries written; an instance stopped by a
                           * NaN leaves the entries after these unspecified */
  int32_t nan_layer;      /* first layer with a NaN, else -1 */
} cmtj_status;

/* Resumable state of one instance (the serving loop's slots): the
 * magnetisation of every layer, steps done, entries written and the status
 * code.  A fresh instance has m = its layers' initial m, nan_layer = -1 and
 * zeros elsewhere. */
typedef struct cmtj_state {
  double m[CMTJ_MAX_LAYERS][3];
  uint32_t step;
  uint32_t entries;
  int32_t code;
  int32_t nan_layer;
} cmtj_state;

/* Sizes and field offsets, for host bindings that mirror these structs. */
typedef struct cmtj_abi_info {
  int32_t version;
  int32_t max_layers;
  int32_t nslot;
  int32_t sizeof_driver;
  int32_t sizeof_layer;
  int32_t sizeof_junction;
  int32_t sizeof_constants;
  int32_t sizeof_status;
  int32_t sizeof_state;
  int32_t sizeof_stack;
  int32_t sizeof_group;
  int32_t offsetof_layer_drv;
  int32_t offsetof_layer_ref_type;
  int32_t offsetof_junction_dt;
  int32_t offsetof_junction_AHE;
} cmtj_abi_info;

/* ---- library-level ------------------------------------------------------ */

/* Fills *out; returns 0.  No reference counterpart (ABI self-description). */
int cmtj_abi(cmtj_abi_info* out);

/* Message of the last failing call on this thread ("" if none). */
const char* cmtj_last_error(void);

/* Names of the kernels the calling thread's last run launched ("name;" per
 * structure group; empty before any).  No counterpart in the reference: it
 * tells a caller which kernel form (one lane per instance, per layer or per
 * component) a cmtj_run / cmtj_run_stacks / cmtj_run_groups / plan execute
 * chose. */
const char* cmtj_last_kernels(void);

/* Process-wide constants store behind the Python PhysicalConstants class.
 * Replaces the static members and setters/getters of
 * core/constants.hpp:6-53 (bound at python/cmtj.cpp:496-511). */
int cmtj_constants_get(cmtj_constants* out);
int cmtj_constants_set(const cmtj_constants* in);
int cmtj_constants_reset(void);

/* Host evaluation of one driver at time t, the same code the kernels run.
 * Replaces ScalarDriver<T>::getCurrentScalarValue (core/drivers.hpp:262-305)
 * for the non-callback kinds; CMTJ_TABLE is rejected here. */
int cmtj_driver_eval(const cmtj_driver* d, double t, double* out);

/* ---- the drop-in: one junction, FP64, host buffers -------------------- */

/* Runs one junction (FP64) on the current GPU and copies results back.
 * Replaces Junction<T>::runSimulation's fixed-step branch
 * (core/junction.hpp:1948-1981) with the log of logLayerParams (:1580-1636)
 * written to `log_host` as [channel][entry] doubles (capacity log_cap
 * elements), the final magnetisation of every layer to mag_out_host[3*L].
 * A Dormand-Prince run with dp_use_pid set in any layer also writes the
 * final PID state (integral, prev) of layer l to mag_out_host[3*L + 2*l]
 * and [3*L + 2*l + 1]: give it 5*L doubles.
 * tables/n_table: samples of CMTJ_TABLE drivers (may be NULL/0).
 * Returns 0 on success (st->code tells NaN failures), nonzero on error. */
int cmtj_run(const cmtj_junction* j, const cmtj_layer* layers,
             const cmtj_constants* c, const double* tables, size_t n_table,
             double* log_host, size_t log_cap, double* mag_out_host,
             cmtj_status* st);

/* Device scratch of cmtj_run_noalloc, per GPU and per process: 64 MiB. */
#define CMTJ_STATIC_SCRATCH_BYTES (64u * 1024u * 1024u)

/* As cmtj_run (same arguments, results bit for bit equal), but allocates no
 * device memory: no hipMalloc, hipFree or stream creation.  The plan
 * descriptors, the log staging, the final magnetisation and the status live
 * in one file-scope device array of CMTJ_STATIC_SCRATCH_BYTES that each GPU
 * holds until the process ends; a run whose descriptors and log need more is
 * rejected with CMTJ_ERR_INVALID.  Work is queued on `stream` (a
 * hipStream_t, NULL = default stream) and synchronised before return.  Every
 * call on one GPU reuses that GPU's array and leaves the last run's contents
 * in it (nothing in it is read before the run writes it): calls are
 * serialised by a process-wide lock, and calls that overlap on one GPU from
 * outside that lock (another library copy in the process) race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code around HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). */
int cmtj_run_noalloc(const cmtj_junction* j, const cmtj_layer* layers,
                     const cmtj_constants* c, const double* tables, size_t n_table,
                     double* log_host, size_t log_cap, double* mag_out_host,
                     cmtj_status* st, void* stream);

/* ---- electrically coupled stacks (FP64, host buffers) ------------------ */

/* A SeriesStack / ParallelStack (core/stack.hpp:414-480): junctions that
 * share one current.  Per step, with the magnetisations of the start of the
 * step, pair (j, j+1) has the coupling
 *   C_j = coupling[j] * (m_j . p_j  +  m_{j+1} . p_{j+1})   (series)
 *   C_j = coupling[j] * (m_j . p_j  -  m_{j+1} . p_{j+1})   (parallel)
 * with m = each junction's top layer and p = its bottom layer (two-layer
 * members) or `pol` (getEffectiveCouplingStrength, :428-436, :459-466).
 * KCL (runSimulationKCL, :372-410): every junction takes the constant
 * current I0(t) + sum_j C_j I0(t), I0 = current(t).  Non-KCL with delayed
 * coupling (runSimulationNonKCL, :305-370): junction j takes the driver
 * current * prod_{k<j} (1 + C_k), phase-shifted by phase_offset * j
 * (series) or phase_offset (parallel), evaluated at its stage times. */
typedef struct cmtj_stack {
  int32_t series;      /* 1: SeriesStack, 0: ParallelStack */
  int32_t kcl;         /* 1: KCL, 0: non-KCL */
  int32_t n_junctions; /* 2..CMTJ_MAX_STACK */
  int32_t two_layer;   /* 1: p_j = bottom layer of junction j, 0: p_j = pol */
  int32_t delayed;     /* non-KCL only.  1: every pair coupling from the
                          magnetisations at the start of the step; 0
                          (setDelayed(false), two-layer members): members
                          step in order, member j's coupling to j-1 from
                          j-1's magnetisations after its step (:334-343) */
  int32_t pad_;
  double phase_offset;
  double pol[3];
  cmtj_driver current; /* setCoupledCurrentDriver (:196-202); CMTJ_TABLE
                          (a Python callback) under KCL only */
  double coupling[CMTJ_MAX_STACK]; /* pair (j, j+1), j < n_junctions - 1 */
  int32_t top[CMTJ_MAX_STACK];     /* layer index of topId in junction j */
  int32_t bottom[CMTJ_MAX_STACK];  /* layer index of bottomId (two-layer members) */
} cmtj_stack;

/* Runs n_stacks stacks of the same member count on the current GPU in one
 * launch: stack s on lanes [s P, s P + members) of one wave (P the member
 * count rounded up to a power of two), its members exchanging the coupling
 * terms by lane shuffles every step.  js[s * members + j] is member j of
 * stack s; every member shares one structure, the solver (RK4 or Euler-Heun),
 * n_steps, write_every and dt.  js[q].log_offset places member q's block in
 * log_host: the channels of cmtj_run's log followed by two stack channels,
 * the member's current at the logged step and, for j < members - 1, the
 * logged coupling of pair (j, j+1) (C_j under KCL, prod_{k<=j} (1 + C_k)
 * otherwise: the stack log's "I_j" and "C_j_{j+1}", :206-217).
 * mag_out_host: [n_stacks * members][8][3], st_host: [n_stacks * members].
 * A NaN in a member stops every member of its stack at that step.  One stack
 * replaces Stack<T>::runSimulation (:296-300); a batch replaces a loop of
 * them over a parameter sweep.  tables/n_table (may be NULL/0): samples of
 * the members' CMTJ_TABLE drivers, as for cmtj_run, and of a KCL stack's
 * CMTJ_TABLE current: tables[current.table_index + i] is its value at step
 * i's start time i*dt (KCL reads the driver once per step, :376).  A
 * non-KCL stack needs a closed-form current: the reference scales it with
 * ScalarDriver::operator* (core/drivers.hpp:313-325), which drops a
 * callback. */
int cmtj_run_stacks(const cmtj_stack* stacks, size_t n_stacks, const cmtj_junction* js,
                    const cmtj_layer* layers, size_t n_layer_rows, const cmtj_constants* c,
                    const double* tables, size_t n_table,
                    double* log_host, size_t log_cap, double* mag_out_host, cmtj_status* st_host);

/* ---- dipole-coupled groups (FP64, host buffers) ------------------------- */

enum cmtj_interaction {
  CMTJ_INTERACTION_NULL = 0,   /* nullDipoleInteraction (core/reservoir.hpp:63-68) */
  CMTJ_INTERACTION_KANAO = 1,  /* computeDipoleInteraction (:79-96) */
  CMTJ_INTERACTION_NOMURA = 2, /* computeDipoleInteractionNoumra (:106-119) */
  /* any interaction affine in the source's magnetisation,
   * f(r_i, r_j, layer_i, layer_j) = T_ij m_j + b_ij (the reference's
   * interactionFunction given as a Python callable, :58-62, :181): T and b
   * of every ordered pair come with the run (cmtj_run_groups pair_terms) */
  CMTJ_INTERACTION_TENSOR = 3
};

/* A GroupInteraction (core/reservoir.hpp:121-230): junctions at fixed
 * positions whose top layers act on each other.  Each step, with the top
 * layers' magnetisations at the start of the step, element i's top layer
 * gets the reserved interaction field sum_{j != i} f(r_i, r_j, layer_i,
 * layer_j), j ascending (stepFunctionalSolver, :126-147), then every element
 * takes one step of its solver. */
typedef struct cmtj_group {
  int32_t n_elements;              /* 1..CMTJ_MAX_GROUP */
  int32_t interaction;             /* enum cmtj_interaction */
  int32_t top[CMTJ_MAX_GROUP];     /* layer index of topId in element j */
  double pos[CMTJ_MAX_GROUP][3];   /* coordinates (m) */
} cmtj_group;

/* Runs n_groups groups of the same element count in one launch, group g on
 * workgroup g (element j on lane j).  js[g * elements + j] is element j of
 * group g: one structure, the solver (RK4 or Euler-Heun), n_steps,
 * write_every and dt shared; the top layer's reserved-interaction drivers
 * are replaced by the group field.  js[q].log_offset places element q's
 * cmtj_run-style log block; mag_out_host [n][8][3], st_host [n].  A NaN in
 * an element stops its group at that step.  One group replaces
 * GroupInteraction::runSimulation (:181-221); a batch, a loop of them.
 * pair_terms (n_pair_terms doubles; may be NULL/0 unless a group's
 * interaction is CMTJ_INTERACTION_TENSOR): [n_groups][n][12][n] doubles,
 * element [g][j][k][i] = for k < 9 the row-major T_ij[k], for k >= 9
 * b_ij[k - 9], of the field on element i from element j of group g; each
 * step adds f_j = T_ij m_j + b_ij (evaluated as fma(T0, mx, fma(T1, my,
 * fma(T2, mz, b0))) per row) to the field, j ascending; the diagonal j = i
 * is skipped. */
int cmtj_run_groups(const cmtj_group* groups, size_t n_groups, const cmtj_junction* js,
                    const cmtj_layer* layers, size_t n_layer_rows, const cmtj_constants* c,
                    const double* pair_terms, size_t n_pair_terms,
                    double* log_host, size_t log_cap, double* mag_out_host, cmtj_status* st_host);

/* ---- batched plans: many uneven instances, device buffers -------------- */

typedef struct cmtj_plan cmtj_plan;

enum cmtj_lanes {
  CMTJ_LANES_AUTO = 0,
  CMTJ_LANES_ONE = 1,
  CMTJ_LANES_PER_LAYER = 2,
  CMTJ_LANES_PER_COMPONENT = 3,
  /* one lane per instance on a persistent grid that takes instances from a
   * device work queue (single-layer Dormand-Prince groups, whose step
   * counts are not known in advance; other groups as CMTJ_LANES_AUTO) */
  CMTJ_LANES_QUEUE = 4
};

typedef struct cmtj_plan_info {
  uint64_t n_instances;
  uint64_t log_elements;    /* sum over instances of channels*entries */
  uint64_t total_steps;     /* sum of n_steps (junction-steps per execute) */
  uint64_t max_steps;
  int32_t precision;
  int32_t log_dtype;
  int32_t max_layers;       /* largest layer count in the batch */
  int32_t n_kernels;        /* kernel launches per execute */
  int32_t lanes_per_instance;
  int32_t fp64_groups;      /* CMTJ_FP32 plans: groups that run a binary64 kernel
                               (spin torques or scalar drives on one lane per
                               component, where no binary32 kernel covers them) */
  char kernel_name[128];
} cmtj_plan_info;

/* Builds a plan: uploads descriptors to HBM, orders instances by cost and
 * selects kernels.  The n junctions reference rows of `layers`
 * (n_layer_rows rows) through layer_offset; log_offset of every junction
 * must already place its block in the caller's log buffer.  `tables` are
 * CMTJ_TABLE samples (may be NULL).  Batches the reference's per-point
 * runSimulation loop (cmtj/utils/procedures.py:141, cmtj/utils/parallel.py:47)
 * which has no single-call counterpart in the reference.
 * lanes_per_instance: CMTJ_LANES_ONE (all layers of an instance in one
 * lane), CMTJ_LANES_PER_LAYER (one lane per layer, for multilayer
 * instances), CMTJ_LANES_PER_COMPONENT (four lanes per layer, one per vector
 * component: binary32 RK4 groups of the lean feature set; binary64 groups of
 * every feature but energy channels, with RK4, Heun or Euler-Heun) or CMTJ_LANES_AUTO
 * (one lane per component where such a kernel covers the group and its waves
 * fit one per SIMD, else one lane per layer for multilayer instances).
 * Returns NULL on error (see cmtj_last_error). */
cmtj_plan* cmtj_plan_create(const cmtj_junction* js, size_t n,
                            const cmtj_layer* layers, size_t n_layer_rows,
                            const cmtj_constants* c, const double* tables,
                            size_t n_table, int32_t precision,
                            int32_t log_dtype, int32_t lanes_per_instance);

/* Runs every instance of the plan once from its initial magnetisation.
 * d_log: device log buffer (f64 or f32 per the plan) of log_capacity
 * elements (checked against the plan's log extent), d_mag_out: device
 * [n][CMTJ_MAX_LAYERS][3] doubles in the caller's junction order (may be
 * NULL; an L-layer Dormand-Prince instance with dp_use_pid set in any
 * layer puts the final PID state (integral, prev) of layer l at elements
 * 3L + 2l and 3L + 2l + 1 of its block), d_status: device [n] cmtj_status (may be NULL).  stream: a
 * hipStream_t (NULL = default stream).  Asynchronous; allocates nothing. */
int cmtj_plan_execute(cmtj_plan* p, void* d_log, size_t log_capacity,
                      double* d_mag_out, cmtj_status* d_status, void* stream);

/* 64-bit key of the junction's structure signature (layer count, solver,
 * log layout, resistance mode, per-layer modes, feature set, time-dependent
 * driver kinds and slots): instances with equal keys run in one kernel
 * launch.  No reference counterpart; lets a serving loop keep the requests of
 * one group together.  Host-only. */
int cmtj_group_key(const cmtj_junction* j, const cmtj_layer* layers, uint64_t* key);

/* Device bytes a plan of these instances needs (the same validation as
 * cmtj_plan_create; tables are not read, only counted).  No reference
 * counterpart: sizes the caller-owned arena of cmtj_plan_create_in. */
int cmtj_plan_arena_size(const cmtj_junction* js, size_t n, const cmtj_layer* layers, size_t n_layer_rows,
                         size_t n_table, int32_t precision, int32_t log_dtype, int32_t lanes_per_instance,
                         size_t* bytes);

/* As cmtj_plan_create, but the descriptors go into caller-owned memory
 * d_arena (arena_bytes >= cmtj_plan_arena_size): device memory, uploaded on
 * `stream` (a hipStream_t, synchronised before return), or, with
 * arena_is_host, host memory from cmtj_host_alloc (its host address), written
 * in place and read by the kernels through the mapping (no copy engine, no
 * synchronisation; the caller must not rewrite it while a launch of the
 * plan runs).  The plan does not own the arena: it must outlive every
 * execute of the plan.  No reference counterpart: for callers that keep
 * their own pools (the serving loop); no allocation. */
cmtj_plan* cmtj_plan_create_in(const cmtj_junction* js, size_t n, const cmtj_layer* layers, size_t n_layer_rows,
                               const cmtj_constants* c, const double* tables, size_t n_table, int32_t precision,
                               int32_t log_dtype, int32_t lanes_per_instance, void* d_arena, size_t arena_bytes,
                               void* stream, int32_t arena_is_host);

/* Pinned host memory mapped into the device's address space (hipHostMalloc,
 * coherent): *host is the host address, *device the address kernels use.
 * For plan arenas, slice state and status that the host reads and writes
 * between launches without copies.  No reference counterpart. */
int cmtj_host_alloc(size_t bytes, void** host, void** device);
int cmtj_host_free(void* host);

/* Advances every instance of the plan by at most max_steps steps from its
 * state d_state[n] (read, then written back), logging into d_log as
 * cmtj_plan_execute does.  Arithmetic, step times and noise counters are
 * those of an uninterrupted run, so a run cut into slices gives the same
 * result.  Instances whose state holds an error code are skipped.  Groups
 * are launched round-robin on streams[0..n_streams-1] (hipStream_t; NULL
 * entries = default stream) so that their slices overlap on the device.
 * No reference counterpart: the serving loop's continuous batching. */
int cmtj_plan_execute_slice(cmtj_plan* p, void* d_log, size_t log_capacity, double* d_mag_out,
                            cmtj_status* d_status, cmtj_state* d_state, uint32_t max_steps,
                            void* const* streams, int32_t n_streams);

/* Captures one cmtj_plan_execute of the plan, bound to these buffers, as a
 * HIP graph held by the plan: one kernel node per structure group and no
 * edges between them (groups write disjoint parts of the outputs), so the
 * groups of a replay may run concurrently instead of one launch after
 * another.  Replaces an earlier graph of the plan.  Allocates the graph's
 * host-side objects only.  No reference counterpart: the reference runs a
 * sweep as one Python process per point (cmtj/utils/parallel.py:47,84). */
int cmtj_plan_graph_create(cmtj_plan* p, void* d_log, size_t log_capacity, double* d_mag_out,
                           cmtj_status* d_status);

/* Replays the plan's graph on `stream` (hipStream_t; NULL = default):
 * the same results as cmtj_plan_execute with the buffers given at capture,
 * bit for bit.  Asynchronous. */
int cmtj_plan_graph_launch(cmtj_plan* p, void* stream);

int cmtj_plan_query(const cmtj_plan* p, cmtj_plan_info* out);
int cmtj_plan_destroy(cmtj_plan* p);

/* ---- post-processing ---------------------------------------------------- */

/* Spin-diode mixing voltage of n resistance series on the device, one value
 * per series: x_k = -sin(2 pi freq[q] k dt) R[q][k], filtered by the IIR
 * filter (b, a) of n_taps coefficients each (scipy.signal.lfilter's
 * transposed direct form II, zero initial state, normalised by a[0]), then
 * averaged.  Replaces compute_sd (cmtj/utils/resistance.py:12-22) with its
 * Filters.butter_lowpass_filter (cmtj/utils/filters.py:33-43), which the
 * reference's VSD_procedure (cmtj/utils/procedures.py:280-331) applies to
 * each sweep point's log on the host.  d_R: device, series q at
 * d_R + q * ld (ld >= len elements); d_freq, d_out: device [n]; b, a: host,
 * n_taps <= 9.  Asynchronous on `stream` (hipStream_t; NULL = default). */
int cmtj_vsd_mix(const double* d_R, size_t ld, size_t n, size_t len, const double* d_freq, double dt,
                 const double* b, const double* a, int32_t n_taps, double* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CMTJ_AMD_H_ */
