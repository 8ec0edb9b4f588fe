/*
 * dcs_amd.h -- C ABI of the MI355X frontier-contraction library (libdcs_amd.so).
 *
 * The library executes the preserved-coordinate ("in-place") contraction of a parity-sliced graph
 * state on one gfx950 GPU, accumulates the coherent conditional block, applies the Walsh-Hadamard /
 * phase / local-Clifford tail and draws from the |a|^2 CDF.  All device memory is owned by the
 * library; the host passes plain arrays.  Every entry point returns a dcs_status; the message of the
 * last failure on the calling thread is available from dcs_last_error().
 *
 * Reference interfaces replaced (paths relative to the reference's source/ directory):
 *   dcs_context_create        cp.cuda.Device(device).use()            kernel_opt/inplace_gpu.py:86-87
 *   dcs_get_device_info       getDeviceProperties / runtime versions  production/worker.py:60-76
 *   dcs_plan_create           InplaceGPUMixin.alloc (plan upload,     kernel_opt/inplace_gpu.py:85-105
 *                             frontier, syndrome map, phase)          sampling/conditional_block_inplace_gpu.py:18-34
 *   dcs_contract              InplaceGPUMixin.prefix_inplace           kernel_opt/inplace_gpu.py:107-124
 *                             (one launch of inplace_step per step)   kernel_opt/inplace_gpu.py:20-63
 *   dcs_block_zero,
 *   dcs_block_accumulate      total = zeros; total[indices] += env*coef  sampling/conditional_block_gpu.py:137,155-158
 *   dcs_block_transform       Walsh-Hadamard, phase, per-qubit U      sampling/conditional_block_gpu.py:159-171
 *   dcs_block_read            cp.asnumpy(vector)                      production/worker.py:109
 *   dcs_draw                  draw_checked primary CDF (cumsum,       sampling/sampling_protocol.py:56-65,81-82
 *                             searchsorted side='right')
 *   dcs_frontier_read         d_inplace[d_inplace_terminal_indices]   kernel_opt/inplace_gpu.py:102-104,124
 *   dcs_gather_plan_create    GPUMixin.alloc + ConditionalBlockGPU     dcs_rankwidth/gpu_kernel.py:106-121
 *                             .alloc (canonical gather executor)       sampling/conditional_block_gpu.py:71-90
 *   dcs_contract on a gather  ConditionalBlockGPU._prefix (one launch   sampling/conditional_block_gpu.py:92-126
 *   plan                      of scan_step_scaled per step)            dcs_rankwidth/gpu_kernel.py:58-101
 * dcs_plan_kernel_source has no counterpart in the reference: it returns the generated group kernels
 * of a plan without a device, for inspection and for emulating them on the CPU.  Neither has
 * dcs_contract_batch (the lanes of several plans, circuits of different widths, in shared launches),
 * nor dcs_plan_factoring (which group kernels, f0-factored or not, the last contraction ran).
 */
#ifndef DCS_AMD_H
#define DCS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    DCS_OK = 0,
    DCS_ERR_INVALID = 1,    /* bad argument or plan (Python: ValueError) */
    DCS_ERR_RUNTIME = 2,    /* HIP runtime failure (RuntimeError) */
    DCS_ERR_NOMEM = 3,      /* frontier does not fit the free device memory (MemoryError) */
    DCS_ERR_NONFINITE = 4,  /* nonfinite frontier or block (ArithmeticError) */
    DCS_ERR_ZERO = 5,       /* zero / invalid CDF outcome, or the independent CDF check disagrees
                               (ArithmeticError) */
    DCS_ERR_UNSUPPORTED = 6 /* the schedule cannot be built on this device: the generated group
                               kernels do not compile, or one cannot be resident (registers, LDS)
                               (ScheduleUnsupported, a ValueError) */
} dcs_status;

typedef struct dcs_context dcs_context;
typedef struct dcs_plan dcs_plan;

typedef struct {
    char name[256];
    char arch[64];
    uint64_t total_bytes;
    uint64_t free_bytes;
    int32_t compute_units;
    int32_t runtime_version;
    int32_t driver_version;
    int32_t device;
} dcs_device_info;

/* Frontier storage of the contraction.  Arithmetic is FP64 in every case.  DCS_STORAGE_F16X2 holds
 * the real and imaginary parts as IEEE binary16 (4 bytes per entry), stored through float with round
 * to nearest, the half-storage engine of kernel_opt/inplace_half_gpu.py:14-28 (in-place plans only). */
enum { DCS_STORAGE_C64 = 0, DCS_STORAGE_C128 = 1, DCS_STORAGE_F16X2 = 2 };

/* Execution schedules of dcs_contract. */
enum {
    DCS_MODE_STEPS = 0,       /* one launch per step, scale chain kept on the device, no host sync */
    DCS_MODE_STEPS_HOST = 1,  /* one launch per step, the host reads the maximum after every step */
    DCS_MODE_FUSED = 2,       /* LDS-resident runs + fused coset groups, generic coset kernel */
    DCS_MODE_COMPILED = 3     /* the fused schedule with one plan-specialised kernel per coset group,
                                 compiled at run time (dcs_plan_compile); the production schedule */
};

typedef struct {
    int32_t storage;            /* DCS_STORAGE_* */
    int32_t maxslots;           /* frontier address bits (physical slots), <= 40 */
    int32_t lanes;              /* independent frontiers held at once */
    int32_t nsteps;
    /* per-step in-place plan (orbit code 0..7, masks over physical slots) */
    const int32_t *kind;
    const uint64_t *K;
    const uint64_t *Ka;
    const int32_t *pivot;       /* grow slot, or highest bit of K (-1 when K == 0) */
    const int32_t *pivot_a;     /* highest bit of Ka (-1 when Ka == 0) */
    const uint64_t *beta;       /* sign mask over slots */
    const uint64_t *in_mask;    /* slots live before the step */
    const uint64_t *out_mask;   /* slots live after the step */
    /* conditional block */
    int32_t k;                  /* log2 of the block length (open vertices) */
    uint64_t terminal_mask;     /* physical slots of the terminal basis; popcount = terminal rank r <= k */
    const uint64_t *syndrome_map;   /* [2^r] block index of terminal coordinate i */
    const double *phase;            /* [2^k] +-1 from edges among open vertices */
    const double *open_unitaries;   /* [k][8] u00r,u00i,u01r,u01i,u10r,u10i,u11r,u11i */
    int32_t blocks;             /* complex128 block accumulators */
    /* schedule knobs (DCS_MODE_FUSED) */
    int32_t fuse_max_dim;       /* largest coset dimension of a fused group: 1..8 (DCS_MODE_FUSED), 1..12 */
    int32_t resident_bits;      /* steps whose live slots all sit below this bit run LDS-resident, <= 12 */
    int32_t tile_log2;          /* log2 entries per coset tile, 11 or 12 (13: DCS_MODE_COMPILED); in
                                   DCS_MODE_COMPILED a group of dimension 9 takes 2^13-entry tiles when
                                   this is smaller (rows of 16 entries), priced 1.1 per entry moved */
    int32_t exact_rounding;     /* 1: fused steps round to storage precision after every step (stored
                                   mantissas equal the per-step schedule); 0: round once per register
                                   pass (more accurate, fewer conversions) */
    int32_t pass_max_dim;       /* span of the orbit masks of one register pass, 3 or 4 (4 in
                                   DCS_MODE_COMPILED: a group kernel that spills registers runs with
                                   span 3) */
    int32_t lds_storage;        /* DCS_MODE_COMPILED: 1 holds coset tiles in LDS in storage precision
                                   (rounded once per register pass), 0 in FP64 */
    int32_t group_policy;       /* 0: greedy maximal coset groups; 1: the group boundaries with the fewest
                                   HBM bytes (dynamic programming over the step sequence); 2: the same with
                                   the register passes over each group's span priced as well; 3: those two
                                   costs overlapped, max(memory, passes) + a quarter of the smaller */
    int32_t lds_buffers;        /* DCS_MODE_COMPILED: LDS tiles per workgroup, 2 (double-buffered, the
                                   default when 0) or 1 (the next tile waits in registers); 3: two
                                   tiles, the next one loaded from HBM straight into LDS
                                   (global_load_lds) during the current one's passes and finished
                                   tiles stored two entries (16 bytes) per lane, in groups whose
                                   complex64 tiles allow it (else as 1); same results as 1;
                                   4: as 3, but the last register pass stores its results to HBM
                                   from registers and the group's scale is folded into its first
                                   step (no LDS write-back, loads issued a whole tile ahead);
                                   5: as 4, and where a group's input is a half or a quarter live, a
                                   ring of three compact LDS buffers holds the live inputs of the next
                                   two tiles in flight (same results as 4) */
    const int32_t *coef_class;  /* DCS_MODE_COMPILED with exact_rounding 0, and lds_storage 1 unless the
                                   storage is complex128 (else ignored; NULL: off): per step, the class
                                   of the ratio f1 / f0 of its local vector, 0 unknown or f0 possibly zero,
                                   1 real, 2 imaginary, 3 complex.  The group kernels then run a step of
                                   class > 0 with (1, f1 / f0) and apply the f0s they left pending at the
                                   group's last step (fewer FP64 operations; the same results up to FP64
                                   rounding between the storage roundings).  A contraction whose
                                   parameters do not fit the classes (f0 == 0, or a ratio of another
                                   class) runs the group kernels without factoring instead. */
} dcs_plan_desc;

typedef struct {
    int64_t index;          /* searchsorted(cdf, u, side='right') */
    double weight;          /* |a_index|^2 */
    double total;           /* sequential cumsum total (float64) */
    double cdf_lower;       /* normalised cdf[index-1] (0 for index 0) */
    double cdf_upper;       /* normalised cdf[index] (cdf[-1] forced to 1) */
    double margin;          /* min(u - lower, upper - u) */
    /* The independent check of draw_checked (math.fsum interval): the total and the interval of the
     * drawn index summed again in another order with compensated (Sum2) arithmetic.  A draw whose
     * check interval does not contain u fails with DCS_ERR_ZERO. */
    double mass_check;
    double cdf_lower_check;
    double cdf_upper_check;
} dcs_draw_result;

const char *dcs_last_error(void);
int dcs_get_device_count(int *count);
int dcs_get_device_info(int device, dcs_device_info *out);

int dcs_context_create(int device, dcs_context **out);
void dcs_context_destroy(dcs_context *ctx);
int dcs_synchronize(dcs_context *ctx);

int dcs_plan_create(dcs_context *ctx, const dcs_plan_desc *desc, dcs_plan **out);
void dcs_plan_destroy(dcs_plan *plan);
/* Generate and compile the plan-specialised group kernels of DCS_MODE_COMPILED (hipRTC, gfx950).
 * Compiled modules are cached per device for the life of the process, keyed by their source.
 * *seconds (may be NULL) receives the compile time (0 when served from the cache).  dcs_contract in
 * DCS_MODE_COMPILED calls it on first use. */
int dcs_plan_compile(dcs_plan *plan, double *seconds);
/* Whether the plan's compiled group kernels are f0-factored (dcs_plan_desc.coef_class) and whether the
 * last dcs_contract in DCS_MODE_COMPILED ran them (0: its parameters did not fit the classes and the
 * unfactored kernels ran).  No counterpart in the reference, whose kernels are not factored. */
int dcs_plan_factoring(dcs_plan *plan, int32_t *module_factored, int32_t *last_call_factored);
/* Source of the group kernels that dcs_plan_compile would build for this description, and the fused
 * group segments of the compiled schedule, without a device.  src (capacity src_cap, may be NULL)
 * receives the NUL-terminated source, cut to fit; *src_len its full length.  For each group segment i
 * < cap: group[i] the kernel index (dcs_group_<index>), first[i] and count[i] its steps, threads[i]
 * its workgroup size; *ngroups the number of segments.  The output arrays may be NULL.  With
 * pass_max_dim 4, dcs_plan_compile also builds the span-3 kernel of every group whose span-4 kernel
 * spills registers to scratch memory, and runs that one instead (known only after compiling). */
int dcs_plan_kernel_source(const dcs_plan_desc *desc, char *src, uint64_t src_cap, uint64_t *src_len,
                           int32_t *group, int32_t *first, int32_t *count, int32_t *threads, int cap,
                           int *ngroups);
/* Canonical gather executor: two tables of 2^rmax entries; step s maps the 2^r_in table to the
 * 2^r_out table, out[c'] = sum over the step's kernel combinations of E[p & (2^r_in - 1)] * (+-f_y),
 * p = XOR_b tables[b][byte b of c'] ^ combination.  dcs_contract on such a plan accepts
 * DCS_MODE_STEPS (scale chain on the device) and DCS_MODE_STEPS_HOST, one lane; afterwards the final
 * table's first 2^terminal_rank entries are the terminal environment (entry i = coordinate i). */
typedef struct {
    int32_t storage;            /* DCS_STORAGE_* */
    int32_t rmax;               /* table address bits, <= 40 */
    int32_t nsteps;
    const int32_t *r_in;        /* per step: input / output table bits, free vertex flag */
    const int32_t *r_out;
    const int32_t *is_free;
    const int32_t *nbytes;      /* per step: byte tables (<= 5) and kernel combinations (<= 8) */
    const int32_t *ncombos;
    const int64_t *tab_off;     /* per step: offset of its tables / combinations in the arrays below */
    const int64_t *comb_off;
    const uint64_t *tables;     /* [nbytes][256] per step, concatenated */
    const uint64_t *combos;
    const uint64_t *beta;       /* per step: sign mask over the input coordinates */
    int32_t k;                  /* log2 of the block length */
    int32_t terminal_rank;      /* r <= k */
    const uint64_t *syndrome_map;   /* [2^r] */
    const double *phase;            /* [2^k] */
    const double *open_unitaries;   /* [k][8] */
    int32_t blocks;
} dcs_gather_desc;

int dcs_gather_plan_create(dcs_context *ctx, const dcs_gather_desc *desc, dcs_plan **out);

/* Device bytes the plan holds (frontier tables, per-step tables, blocks, scratch). */
int dcs_plan_memory(dcs_plan *plan, uint64_t *bytes);
/* Number of scale units (maxima) a contraction in `mode` produces per lane. */
int dcs_plan_units(dcs_plan *plan, int mode, int *nunits);
/* Schedule of `mode`: per segment (kind 0 step, 1 resident run, 2 coset group), first step, step
 * count and coset dimension.  Arrays of capacity `cap`; *count receives the segment count. */
int dcs_plan_schedule(dcs_plan *plan, int mode, int cap, int32_t *seg_kind, int32_t *first,
                      int32_t *count, int32_t *dim, int *nseg);

/* Contract `nlanes` frontiers (lane ids in `lanes`) through every step.
 *   coeffs [nlanes][nsteps][4]  f0r, f0i, f1r, f1i of the step's vertex
 *   ov, y  [nlanes][nsteps]     offset parity bit and forced value of each step
 *   unit_m2 out [nlanes][nunits] max |out|^2 of every scale unit, in order
 * Returns after the device finished. */
int dcs_contract(dcs_plan *plan, int mode, int nlanes, const int32_t *lanes, const double *coeffs,
                 const int32_t *ov, const int32_t *y, double *unit_m2);

/* One plan's part of dcs_contract_batch: its lanes and their step parameters, laid out as in
 * dcs_contract, and its unit maxima out ([nlanes][nunits of DCS_MODE_FUSED]). */
typedef struct dcs_batch_item {
    dcs_plan *plan;
    int nlanes;
    const int32_t *lanes;
    const double *coeffs;
    const int32_t *ov;
    const int32_t *y;
    double *unit_m2;
} dcs_batch_item;

#define DCS_BATCH_MAX 16

/* Contract the lanes of up to DCS_BATCH_MAX different plans (circuits of different frontier widths)
 * together, in the fused schedule on the generic kernels: round r runs segment r of every plan's
 * schedule, all resident runs of the round in one launch and all coset groups of the round in one
 * launch, so plans too narrow to fill the GPU alone share its compute units.  Each lane's result is
 * bit-identical to dcs_contract(plan, DCS_MODE_FUSED, ...) of the same parameters.  The plans must be
 * distinct, share one context and one storage (complex64 or complex128) and tile_log2 (11 or 12), and
 * have pass_max_dim 3, resident_bits 12 and no exact_rounding (the defaults).  *launches (may be NULL)
 * receives the number of contraction launches.  Returns
 * after the device finished.  No counterpart in the reference, which contracts one prefix per call
 * (kernel_opt/inplace_gpu.py:107-124); it serves a caller holding requests for several small circuits. */
int dcs_contract_batch(int nitems, const dcs_batch_item *items, int *launches);

/* 1: dcs_contract captures its launch sequence (parameter upload, maxima clear, every kernel of the
 * schedule, maxima read-back) into a HIP graph once per (mode, nlanes) and replays it on later calls;
 * 0 (default): launches every call, and frees the captured graphs.  DCS_MODE_STEPS_HOST and profiling
 * always launch directly.  No counterpart in the reference (CuPy launches every step from Python,
 * kernel_opt/inplace_gpu.py:111-124); it serves callers that contract many prefixes of one plan. */
int dcs_plan_use_graphs(dcs_plan *plan, int enable);

int dcs_block_zero(dcs_plan *plan, int block);
/* total[smap[i] ^ xmask] += env_lane(i) * coef, env at the terminal slots of the lane. */
int dcs_block_accumulate(dcs_plan *plan, int block, int lane, uint64_t xmask, double coef_re,
                         double coef_im);
int dcs_block_transform(dcs_plan *plan, int block);
int dcs_block_read(dcs_plan *plan, int block, double *host_out);
int dcs_block_device_pointer(dcs_plan *plan, int block, void **ptr);
int dcs_draw(dcs_plan *plan, int block, double u, dcs_draw_result *out);
/* Largest block (log2 entries) that dcs_draw_noalloc accepts. */
#define DCS_DRAW_NOALLOC_MAX_K 22

/* dcs_draw on a block the caller holds on the context's device, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (serving loops and harnesses
 * that forbid hipMalloc).  `block` is a device pointer to 2^k complex128 entries (interleaved re/im,
 * e.g. from dcs_block_device_pointer) whose writes have completed; k <= DCS_DRAW_NOALLOC_MAX_K, larger
 * blocks are rejected with DCS_ERR_INVALID.  The result and status codes equal dcs_draw's on the same
 * block.  The segment sums and the result live in one static device array per device (512 KiB), held
 * until the process ends and shared by every call on that device: calls that overlap on one device
 * race, so callers serialise them.  Runs on the context's stream and returns after the device
 * finished. */
int dcs_draw_noalloc(dcs_context *ctx, const void *block, int k, double u, dcs_draw_result *out);

/* Terminal environment of a lane (2^r entries in storage precision, interleaved re/im). */
int dcs_frontier_read(dcs_plan *plan, int lane, void *host_out);

/* Elapsed device time (ms) and launch count per kernel family since the last reset:
 * index 0 step kernels, 1 resident runs, 2 coset groups, 3 block kernels, 4 draw. */
int dcs_plan_profile(dcs_plan *plan, int enable, double *ms, int64_t *launches);

/* Device bandwidth probe (the practical HBM ceiling; gpu_variants.py:29-37 is the reference's): copy
 * `bytes` (a multiple of 32 KiB) `reps` times with a non-temporal streaming kernel under several launch
 * shapes (2, 4 or 8 independent 16-byte loads in flight per thread, 4 to 16 workgroups per CU) and
 * return the average kernel time in ms of the fastest shape. */
int dcs_stream_copy(dcs_context *ctx, uint64_t bytes, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
