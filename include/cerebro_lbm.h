/* cerebro_lbm.h -- C ABI of the MI355X (gfx950) sparse D3Q27 lattice-Boltzmann solver step.
 *
 * The reference (CerebroSim, Python + NVIDIA Warp) has no native library: its solver step is a set of Warp
 * kernels launched from Python with wp.launch and replayed from captured CUDA graphs. Each entry point below
 * replaces one of those launches or Warp runtime calls; the citation names the reference interface it stands in
 * for. Everything is plain C: device buffers are raw pointers (the Python layer allocates them as PyTorch-ROCm
 * tensors and keeps them alive), streams and graphs are opaque handles, sizes are int64.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative CS_E* code; cs_last_error() then returns a message
 *    describing the failure (thread-local, valid until the next failing call on the same thread).
 *  - Launch functions only enqueue work on the given stream; nothing synchronises unless the name says so.
 *  - Device pointers must stay valid for as long as any graph captured with them may be replayed.
 *  - Populations are stored as deviations from the rest weights, f~_i = f_i - w_i (reference lbm.py:23).
 *  - Link codes in neighbour tables: >= 0 source cell, -1 wall, -2 inlet, -3 - k outlet k (lbm.py:50).
 */
#ifndef CEREBRO_LBM_H
#define CEREBRO_LBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_OK 0
#define CS_E_ARG (-1)       /* invalid argument (null pointer, bad size, unsupported option) */
#define CS_E_HIP (-2)       /* a HIP runtime call failed */
#define CS_E_LAYOUT (-3)    /* lattice not representable in the requested layout */
#define CS_E_NOMEM (-4)     /* host allocation failed */

#define CS_MODEL_TRT 0          /* lbm.py:359 "trt" */
#define CS_MODEL_REGULARISED 1  /* lbm.py:359 "regularised" */

/* Message for the last failing call on this thread ("" if none). */
const char* cs_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int cs_version(void);

/* ---------------------------------------------------------------------------------------------------------------
 * Streams and graphs -- replace wp.ScopedCapture / wp.capture_launch (lbm.py:437-444) and wp.synchronize
 * (validate.py:262, 323, 326). Streams are created at the default priority.
 * ------------------------------------------------------------------------------------------------------------- */
int cs_stream_create(void** stream);
int cs_stream_destroy(void* stream);
int cs_stream_synchronize(void* stream);
int cs_device_synchronize(void);
/* Begin capturing every launch enqueued on `stream` (thread-local capture mode). */
int cs_graph_begin(void* stream);
/* End the capture and instantiate it; *graph receives an executable graph handle. */
int cs_graph_end(void* stream, void** graph);
int cs_graph_launch(void* graph, void* stream);
int cs_graph_destroy(void* graph);

/* ---------------------------------------------------------------------------------------------------------------
 * Drive clock -- replaces k_clock (lbm.py:90-102), launched once per step before the step kernel (lbm.py:423).
 * drive[0] = coef[0] + sum_k coef[2k-1] cos(k phi) + coef[2k] sin(k phi), phi = 2 pi (t mod period) / period,
 * times (1/2 - 1/2 cos(pi t / ramp)) while t < ramp; then tick[0] = t + 1.
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_clock_args {
    int32_t* tick;       /* device, 1 int */
    const float* coef;   /* device, 2 nh + 1 floats: a0, a1, b1, a2, b2, ... */
    int32_t nh;
    int32_t period;
    int32_t ramp;
    float* drive;        /* device, 1 float */
} cs_clock_args;
int cs_clock(const cs_clock_args* a, void* stream);

/* The clocks of packed lattices (cases): k_clock for each case c < n_cases with period[c] > 0, one thread each;
 * coefficients of case c at coef + c * coef_stride. */
typedef struct cs_clock_cases_args {
    int32_t* tick;              /* [n_cases] */
    const float* coef;          /* [n_cases][coef_stride] */
    const int32_t* nh;          /* [n_cases] */
    const int32_t* period;      /* [n_cases]; 0 = unused case */
    const int32_t* ramp;        /* [n_cases] */
    float* drive;               /* [n_cases] */
    int32_t n_cases;            /* at most CS_MAX_CASES */
    int32_t coef_stride;
} cs_clock_cases_args;
int cs_clock_cases(const cs_clock_cases_args* a, void* stream);

/* The clocks of n_steps consecutive steps at once (graphs launch it once per chunk instead of a clock per step):
 * for each case c < n_cases with period[c] > 0, drive[k * drive_stride + c] = k_clock's value at tick[c] + k for
 * k < n_steps, then tick[c] += n_steps. Same arithmetic as cs_clock. */
typedef struct cs_clock_table_args {
    int32_t* tick;              /* [n_cases] */
    const float* coef;          /* [n_cases][coef_stride] */
    const int32_t* nh;          /* [n_cases] */
    const int32_t* period;      /* [n_cases]; 0 = unused case */
    const int32_t* ramp;        /* [n_cases] */
    float* drive;               /* [n_steps][drive_stride] */
    int32_t n_cases;            /* at most CS_MAX_CASES */
    int32_t coef_stride;
    int32_t n_steps;
    int32_t drive_stride;       /* >= n_cases */
} cs_clock_table_args;
int cs_clock_table(const cs_clock_table_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Solver step, indexed layout -- replaces k_step (lbm.py:140-257) launched at lbm.py:425 with dim = N.
 * Same arrays as the reference: f_in/f_out (Q, N) row-major, nbr/qw (Q, N), uin/vel_old/vel (N, 3) float AoS,
 * inl/om_cell/rho (N). Works for any neighbour table (any cell order, periodic wrap, hand-made lattices).
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_step_args {
    const float* f_in;
    float* f_out;
    const int32_t* nbr;
    const float* qw;
    const float* uin;
    const int32_t* inl;
    const float* p_out;
    const float* drive;     /* device, 1 float, written by cs_clock */
    const float* om_cell;   /* per-cell 1 / tau */
    const float* vel_old;
    float* vel;
    float* rho;
    int64_t n_cells;
    float force[3];         /* body force per unit drive (lbm.py:394-396) */
    float magic;            /* TRT magic parameter Lambda */
    float sigma;            /* hybrid blend; 1 = pure regularised */
    int32_t bouzidi;        /* 1 = interpolated walls, 0 = half-way bounce-back */
    int32_t model;          /* CS_MODEL_* */
    int32_t galilean;       /* 1 = third-order equilibrium and diagonal stress correction (D3Q27) */
} cs_step_args;
int cs_step_indexed(const cs_step_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Lattice builder -- build_lattice of lbm.py:264-330 from a signed distance on grid nodes (negative = fluid), in
 * native code. Output arrays (cs_lattice_copy): "coords" (N, 3) int32 in C order of the grid, "nbr" (Q, N) int32
 * source cell or link code, "q" (Q, N) float32 wall fraction (0.5 where unused), "uin" (N, 3) float32 inlet profile.
 * ------------------------------------------------------------------------------------------------------------- */
#define CS_WALL (-1)
#define CS_INLET (-2)
#define CS_OUTLET0 (-3)          /* outlet k: CS_OUTLET0 - k */
typedef struct cs_opening {
    int32_t kind;               /* 0 inlet, 1 outlet */
    double centre[3];
    double normal[3];           /* outward */
    double radius;
} cs_opening;
typedef struct cs_grid_in {
    const float* phi;           /* [shape0][shape1][shape2], C order */
    int32_t shape[3];
    int32_t periodic[3];
    double origin[3];
    double dx;
    const cs_opening* openings; /* inlets are tried before outlets; outlet k = k-th outlet in this order */
    int32_t n_openings;
    double slack;               /* opening radius slack, in dx */
    int32_t threads;            /* host threads, 0 = all */
} cs_grid_in;
int cs_lattice_build(const cs_grid_in* in, void** handle);
/* N into *n_cells and up to max_counts link counts (wall, inlet, outlet 0, ...); returns the number of counts. */
int cs_lattice_sizes(const void* handle, int64_t* n_cells, int64_t* counts, int32_t max_counts);
int cs_lattice_copy(const void* handle, const char* name, void* dst);
/* The handle's own buffer of an output array (same names and shapes as cs_lattice_copy), valid until
 * cs_lattice_free: callers that keep the handle alive while they use the arrays avoid the copy. */
int cs_lattice_view(const void* handle, const char* name, void** data);
int cs_lattice_free(void* handle);

/* ---------------------------------------------------------------------------------------------------------------
 * Brick layout -- the MI355X storage of the same lattice (no reference counterpart: the reference keeps the
 * indexed layout of lbm.py:264-330 on the device). Fluid cells are grouped into 4 x 4 x 4 bricks (one wavefront per
 * brick); populations are stored once, per brick as a block of 27 x 64 floats, and streamed in place with the AA
 * pattern.
 * The host builder converts a lattice in the reference layout (nbr/q/uin/coords as returned by build_lattice) and
 * checks that every fluid link of nbr is the implicit brick neighbour; it fails with CS_E_LAYOUT otherwise (the
 * caller then uses the indexed layout).
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_lattice_in {
    int64_t n_cells;
    int32_t shape[3];
    int32_t periodic[3];        /* 1 = axis wraps (must then be a multiple of the brick edge) */
    const int32_t* coords;      /* host, (N, 3) */
    const int32_t* nbr;         /* host, (Q, N) */
    const float* q;             /* host, (Q, N) */
    const float* uin;           /* host, (N, 3) */
} cs_lattice_in;

typedef struct cs_bricks_sizes {
    int64_t n_bricks;           /* bricks holding at least one fluid cell */
    int64_t n_boundary;         /* bricks with a solid cell or a missing link */
    int64_t n_interior;
    int64_t n_slots;            /* n_bricks * 64 */
    int64_t n_links;            /* missing links of the boundary bricks (entries of qlink before its padding) */
    int64_t n_side;             /* Bouzidi side copies (links with an smask bit) */
} cs_bricks_sizes;

/* Build the brick layout on the host; *handle owns the arrays until cs_bricks_free. */
int cs_bricks_build(const cs_lattice_in* in, void** handle);
int cs_bricks_sizes_of(const void* handle, cs_bricks_sizes* out);
/* Copy one named host array into dst (caller-sized by cs_bricks_sizes_of):
 *  "coords"        int32 [n_bricks][3]     brick coordinates (in bricks)
 *  "nbr"           int32 [n_bricks][27]    neighbour brick at offset d = (ox+1) + 3 (oy+1) + 9 (oz+1), -1 if none
 *  "bidx"          int32 [n_bricks]        boundary index of the brick, -1 for interior bricks
 *  "interior"      int32 [n_interior]      brick ids of interior bricks
 *  "boundary"      int32 [n_boundary]      brick ids of boundary bricks
 *  "slot_of_cell"  int32 [N]               storage slot (brick * 64 + x + 4 y + 16 z) of each reference cell
 *  "cell_of_slot"  int32 [n_slots]         reference cell of each slot, -1 for solid slots
 *  "lmask"         uint32 [n_boundary][64] bit i: link i is fluid; bit 27: inlet cell; 0 for solid slots
 *  "smask"         uint32 [n_boundary][64] bit i: wall link i with q >= 1/2 whose opposite link is fluid
 *  "qc"            float [n_boundary][27][64] wall fraction q (>= 0), -1 inlet, -2 - k outlet k, 0 fluid
 *  "uin"           float [n_boundary][3][64] inlet profile of inlet cells
 * and the compact boundary link data the step reads:
 *  "qlink"         float [n_links + 1]     qc of every missing link: brick by brick, lane by lane, directions 1..26;
 *                                          one element of padding (a lane reads ahead where its link is not missing)
 *  "qstart"        int32 [n_boundary + 1]  start of each boundary brick's run in qlink
 *  "sstart"        int32 [n_boundary + 1]  start of each boundary brick's run of side copies (smask links, same order)
 *  "loff"          uint32 [n_boundary][64] lane's run offset in its brick: qlink offset | side offset << 16 (link i
 *                                          of a lane: offset + number of the lane's links below i)
 *  "ldir"          uint32 [n_boundary][2]  per brick: union of the missing links, union of the smask bits
 * The step copies a boundary brick's link run and side run into on-chip memory: the builder fails with CS_E_LAYOUT
 * for a lattice in which some brick has more than CS_BRICK_LINK_RUN_MAX of them together. */
#define CS_BRICK_LINK_RUN_MAX 862
int cs_bricks_copy(const void* handle, const char* name, void* dst);
int cs_bricks_free(void* handle);

/* Slot order inside a brick's block of 27 x 64 populations: population k of cell l (x + 4 y + 16 z) is at offset
 * out[k * 64 + l] of the block (27 x 64 uint16). Slots are grouped by the brick whose odd step accesses them. */
int cs_bricks_slot_table(uint16_t* out);
/* Staging order of an interior brick's odd-step footprint (its own group, then the record of each neighbour: the 6
 * face neighbours, the 12 edge neighbours, the 8 corner neighbours, each in offset order o): position of the
 * population lane l reads for direction i is out[i * 64 + l]. */
int cs_bricks_stage_table(uint16_t* out);
/* The 728 neighbour records of an interior brick's staged odd-step footprint, after its own 1000-slot group: entry
 * p = (neighbour offset o << 11) | block offset, for staged position 1000 + p. */
int cs_bricks_record_table(uint16_t* out);

/* One AA step over a list of bricks (two launches per step: interior bricks, then boundary bricks).
 * parity 0 = even step (reads and writes the cell's own slots), 1 = odd step (pulls from and pushes to the
 * neighbours). Velocity is stored as three component planes of n_slots floats; rho as one plane. */
#define CS_CASE_SHIFT 22        /* brick-list entry = brick | case << CS_CASE_SHIFT (bricks < 2^22, cases < 1024) */
#define CS_MAX_CASES 1024
#define CS_LIST_SKIP 0xFFFFFFFFu /* list entry of an unused record: the wavefront does nothing */
#define CS_STORAGE_FP32 0       /* populations stored as float32 deviations f - w */
#define CS_STORAGE_FP16 1       /* populations stored as float16 of (f - feq2(writer's moments)) / w * 2^10 */
typedef struct cs_bricks_step_args {
    void* f;                    /* [n_bricks][1728] float or half, slot order of cs_bricks_slot_table */
    const int32_t* nbr;         /* [n_bricks][27] */
    const uint32_t* list;       /* bricks to process (entries brick | case << CS_CASE_SHIFT, or CS_LIST_SKIP); for
                                   boundary bricks, the boundary record of entry li is brec[li], or li when brec is
                                   NULL */
    int64_t n_list;
    const int32_t* bidx;        /* [n_bricks] */
    const uint32_t* lmask;      /* [n_boundary][64] */
    const uint32_t* smask;      /* [n_boundary][64] */
    const float* qlink;         /* codes (q >= 0 wall, -1 inlet, -2-k outlet k) of the missing links only: per
                                   boundary brick, lane by lane, each lane's links in direction order 1..26; one
                                   element of padding after the last run */
    const int32_t* qstart;      /* [n_boundary + 1] start of each boundary brick's run in qlink */
    const float* uin;           /* [n_boundary][3][64] */
    float* side;                /* copies of the Bouzidi q >= 1/2 partners (links with a bit in smask), same order
                                   and padding */
    const int32_t* sstart;      /* [n_boundary + 1] start of each boundary brick's run in side */
    const uint32_t* loff;       /* [n_boundary][64] lane's run inside its brick's runs: qlink offset | side offset
                                   << 16 (link i of the lane then sits at offset + number of its links below i) */
    const uint32_t* ldir;       /* [n_boundary][2] union over the brick's cells of the missing links, of the smask
                                   bits (directions no cell needs are skipped by the whole wavefront) */
    const float* om_cell;       /* [n_slots] per-cell 1 / tau, read only where om_brick != 0 */
    const uint8_t* om_brick;    /* [n_bricks] 1 = the brick's cells have their own omega */
    float omega;                /* 1 / tau of every other cell */
    const float* p_out;
    const float* drive;
    const float* vel_old;       /* FP32 storage: [3][n_slots] velocity planes of the previous step; FP16 storage:
                                   [n_slots][4] float4 {drho, ux, uy, uz} of the previous step */
    float* vel;                 /* the same for this step's output */
    float* rho;                 /* FP32 storage: [n_slots] density deviation; FP16 storage: unused (in vel) */
    const float* rho_old;       /* unused (NULL) */
    int64_t n_slots;
    float force[3];
    float magic;
    int32_t bouzidi;
    int32_t model;
    int32_t galilean;
    int32_t parity;
    int32_t boundary;           /* 1 = the list holds boundary bricks */
    int32_t storage;            /* CS_STORAGE_* */
    /* Several lattices (cases) packed in one brick array: case c of a list entry takes drive[c], the opening
     * pressures p_out + c * p_out_stride, and, when given, case_omega[c] / case_force[3c..3c+2] in place of
     * omega / force. A single lattice: case 0, NULL tables. */
    const float* case_omega;
    const float* case_force;
    int32_t p_out_stride;
    int32_t variant;            /* launch variant, 0 .. cs_bricks_variant_info(-1) - 1 (0 = default) */
    const int32_t* brec;        /* boundary lists: record of each entry (NULL: the entry's position), so that a list
                                   can hold its live entries only while records keep their places */
    int32_t hybrid;             /* regularised model, FP32 storage: 1 = blend the stress with the finite-difference one
                                   by sigma (lbm.py:224-229); 0 (a zeroed struct) = pure regularised */
    float sigma;                /* the blend when hybrid = 1, in [0, 1) */
} cs_bricks_step_args;
int cs_bricks_step(const cs_bricks_step_args* a, void* stream);
/* One step's boundary and interior lists in a single launch (replaces the two k_step launches of a step's bricks,
 * lbm.py:427, for the brick layout): the boundary block (boundary = 1) and the interior block (boundary = 0) must
 * agree in parity, storage, model, galilean flag and variant. Either list may be empty (then the other one runs
 * alone through cs_bricks_step, and the empty block is not otherwise checked). Gives the bits of cs_bricks_step on
 * each block; one launch instead of two forked ones. */
int cs_bricks_step_pair(const cs_bricks_step_args* boundary, const cs_bricks_step_args* interior, void* stream);
/* n_steps consecutive steps of n_pairs (boundary, interior) block pairs, one paired launch per pair and step, all on
 * `stream` (the chunk runner of cs_bricks_step_pair). args holds 4 * n_pairs blocks: the pairs for parity 0
 * (boundary, interior, boundary, ...), then those for parity 1; step s launches the pairs of parity
 * (parity0 + s) & 1 with drive advanced by s * drive_stride floats. Works under stream capture. */
int cs_bricks_run_pairs(const cs_bricks_step_args* args, int32_t n_pairs, int64_t n_steps, int32_t parity0,
                        int64_t drive_stride, void* stream);
/* Launch variant `variant`: wavefronts (bricks) per workgroup and XCD-aware workgroup order. Returns CS_OK, or the
 * number of variants when `variant` is out of range (no outputs written). */
int cs_bricks_variant_info(int32_t variant, int32_t* waves_per_group, int32_t* xcd_order);

/* Conversions between the AA brick storage and the reference (Q, N) layout of post-collision populations, for the
 * state that the next step of the given parity reads. set_f also refreshes the side copies. */
typedef struct cs_bricks_convert_args {
    void* f;                    /* brick storage (float or half) */
    const int32_t* nbr;
    const int32_t* bidx;
    const uint32_t* lmask;
    const uint32_t* smask;
    float* side;                /* compact side copies (cs_bricks_step_args.side) */
    const int32_t* sstart;
    const uint32_t* loff;
    const int32_t* slot_of_cell; /* [N] */
    int64_t n_cells;
    float* f_ref;               /* device (Q, N) */
    int32_t parity;             /* parity of the next step */
    int32_t storage;            /* CS_STORAGE_* */
    const float* vel;           /* FP16 storage: newest moments [n_slots][4] {drho, ux, uy, uz} (the cells' own) */
    const float* rho;           /* unused */
    int64_t n_slots;
} cs_bricks_convert_args;
int cs_bricks_get_f(const cs_bricks_convert_args* a, void* stream);
int cs_bricks_set_f(const cs_bricks_convert_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Per-step flow measurement and boundary update -- replaces k_flux + k_bc (simulate.py:65-82) launched after each
 * step at simulate.py:111-114. Q[slot] = sum over the M slab cells of vel[idx] . w (deterministic order, FP64
 * accumulation of the FP32 products), then p_out[0] = drive[0]; for k in 1..n_open-1:
 * p_out[k] += alpha (R[k] Q[k] - pref - p_out[k]) (FP64, stored as float); q_last = Q.
 * The M cells are summed in G chunks, G a function of M alone: P = max(1, ceil(M / 262144)) cells per thread,
 * chunks of 1024 P cells, G = ceil(M / (1024 P)) <= 256. work holds n_open x work_chunks doubles (chunk sums,
 * overwritten each call), work_chunks >= max(G, 1).
 * Cells are given as storage slots of the layout in use; component c of slot s is vel[s * vel_elem + c * vel_comp]
 * ((N, 3) AoS: vel_elem = 3, vel_comp = 1; component planes of S slots: vel_elem = 1, vel_comp = S).
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_flux_args {
    const int32_t* idx;      /* device, M storage slots */
    const float* w;          /* device, (M, 3) weights: +-normal / slab thickness */
    const int32_t* slot;     /* device, M opening index per cell */
    int64_t m;
    const float* vel;        /* velocity of the step just taken */
    int64_t vel_elem;
    int64_t vel_comp;
    const float* R;          /* device, n_open resistances (lattice units) */
    const float* drive;      /* device, 1 float */
    float* p_out;            /* device, n_open opening pressures */
    float* q_last;           /* device, n_open flows of this step */
    int32_t n_open;          /* at most 64 */
    float pref;
    float alpha;
    double* work;            /* device, [n_open][work_chunks] */
    int32_t work_chunks;
} cs_flux_args;
int cs_flux_bc(const cs_flux_args* a, void* stream);
/* No counterpart in the reference: cs_flux_bc for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It ignores a->work / a->work_chunks and uses a
 * library-owned static device array of 64 x 256 doubles (128 KiB per device, held until the process ends): up to 64
 * openings, any number of slab cells. It calls no allocating routine; n_open above 64 is rejected. All calls on one
 * device share that array: calls whose launches overlap on one device (different streams, no ordering between
 * them) race. Runs on the current device; `stream` must belong to it. Results equal cs_flux_bc's bit for bit. */
int cs_flux_bc_static(const cs_flux_args* a, void* stream);

/* k_flux + k_bc for packed lattices: every case c with n_open[c] > 0, over the slab cells start[c] .. start[c+1]-1
 * of the concatenated lists; per-case arrays at c * open_stride. Same chunks, arithmetic and summation order as
 * cs_flux_bc on the case alone. work holds n_cases x open_stride x work_chunks doubles; work_chunks (at most 256)
 * must cover every case's G, else that case's flows and pressures are written as NaN. */
typedef struct cs_flux_cases_args {
    const int32_t* idx;
    const float* w;
    const int32_t* slot;
    const int64_t* start;       /* [n_cases + 1] */
    const float* vel;
    int64_t vel_elem;
    int64_t vel_comp;
    const float* R;             /* [n_cases][open_stride] */
    const float* drive;         /* [n_cases] */
    float* p_out;               /* [n_cases][open_stride] */
    float* q_last;              /* [n_cases][open_stride] */
    const int32_t* n_open;      /* [n_cases]; 0 = no flow measurement */
    const float* pref;          /* [n_cases] */
    const float* alpha;         /* [n_cases] */
    int32_t n_cases;
    int32_t open_stride;        /* at most 64 */
    double* work;               /* device, [n_cases][open_stride][work_chunks] */
    int32_t work_chunks;
} cs_flux_cases_args;
int cs_flux_cases(const cs_flux_cases_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Probe reduction -- replaces k_gather (simulate.py:85-90) launched at simulate.py:309: out[slot] = sum rho / 3
 * over each probe's cells and umax = max |u|, both written (not accumulated), deterministic order.
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_gather_args {
    const int32_t* idx;      /* device, M storage slots */
    const int32_t* slot;     /* device, M probe index per cell */
    int64_t m;
    const float* rho;
    const float* vel;
    int64_t vel_elem;
    int64_t vel_comp;
    float* out;              /* device, n_probe sums */
    float* umax;             /* device, 1 float */
    int32_t n_probe;         /* at most 64 */
    int64_t rho_elem;        /* density of slot s at rho[s * rho_elem] (0 or 1: a dense plane) */
} cs_gather_args;
int cs_gather(const cs_gather_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Wall shear stress -- replaces shear() of wss.py:53-57 (numpy, per saved frame) with a device reduction on the live
 * velocity field. Per wall vertex v: tau_v = scale * sum_k c[v,k] (u_k - (u_k . n_v) n_v) over its K cells (fixed
 * order), u_k read at storage slot idx[v,k]. scale = mu * 1e3 * (m/s per velocity unit) turns the fit coefficients
 * (per mm) into Pa. Optionally writes tau (V, 3) and accumulates sum tau (V, 3) and sum |tau| (V) for the
 * time-averaged WSS and the oscillatory shear index of wss.py:120-122.
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_wss_args {
    const int32_t* idx;         /* device, (V, K) storage slots */
    const float* c;             /* device, (V, K) fit coefficients, 0 for unused entries */
    const float* normal;        /* device, (V, 3) unit normals into the lumen */
    int64_t n_vert;
    int32_t k;
    const float* vel;
    int64_t vel_elem;
    int64_t vel_comp;
    float scale;
    float* tau;                 /* device, (V, 3), or NULL */
    float* acc_tau;             /* device, (V, 3) += tau, or NULL */
    float* acc_mag;             /* device, (V) += |tau|, or NULL */
} cs_wss_args;
int cs_wss(const cs_wss_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Stored frames -- replaces the host binning of a frame in simulate.py's run loop (simulate.py:345-352, with
 * coarse_map of simulate.py:199-205): each bin b averages its cells slots[start[b] .. start[b + 1]) (at least one)
 * in FP64, in order, and writes float16 p_out[b] = p_scale * mean(rho) + p_offset and u_out[b][c] = u_scale *
 * mean(u_c) (rho = the density deviation; pressure = rho / 3 in lattice units, so p_scale carries the 1/3 and the
 * unit conversion). Component c of slot s: vel[s * vel_elem + c * vel_comp]; density: rho[s * rho_elem].
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_bin_args {
    const int32_t* start;       /* device, [n_bins + 1] */
    const int32_t* slots;       /* device, storage slots of the cells, bin by bin */
    int64_t n_bins;
    const float* rho;
    int64_t rho_elem;
    const float* vel;
    int64_t vel_elem;
    int64_t vel_comp;
    double p_scale;
    double p_offset;
    double u_scale;
    void* p_out;                /* device, float16 [n_bins] */
    void* u_out;                /* device, float16 [n_bins][3] */
} cs_bin_args;
int cs_bin_fields(const cs_bin_args* a, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Health check -- replaces Solver.ok() (lbm.py:454-455), which copies rho to the host and tests isfinite. A
 * deterministic two-pass device reduction over x[0..n) writes result[0] = sum of the finite values (FP64) and
 * result[1] = number of non-finite values, so ok() reads back 16 bytes instead of 4 N. `work` is caller-provided
 * device scratch of CS_REDUCE_WORK doubles; result is device memory of 2 doubles.
 * ------------------------------------------------------------------------------------------------------------- */
#define CS_REDUCE_BLOCKS 512
#define CS_REDUCE_WORK (2 * CS_REDUCE_BLOCKS)
int cs_reduce_finite(const float* x, int64_t n, double* work, double* result, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Bandwidth reference (no counterpart in the reference): copies `bytes` (a multiple of 16) from src to dst with
 * 16-byte loads and stores, one per thread (workgroups of 256 threads). The benchmark times it beside the step
 * kernels as the achievable read + write bandwidth of the device.
 * ------------------------------------------------------------------------------------------------------------- */
int cs_stream_copy(const void* src, void* dst, int64_t bytes, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Vessel signed distance -- replaces k_sdf (vessel.py:183-207) launched by sdf_grid (vessel.py:234-245): at every
 * voxel centre p = origin + (i, j, k) dx of an (n[0], n[1], n[2]) grid (k fastest), the distance to each branch's
 * chain of round cones (stations pts[br_start[b] ..][x, y, z, radius], radius interpolated along each segment),
 * cut flat by the branch's cap half-spaces within 2.5 cap radii of each cap (cap_r 0 = no cap), combined over the
 * branches with the polynomial smooth minimum of width `blend`. Negative inside. FP32, mm.
 * ------------------------------------------------------------------------------------------------------------- */
typedef struct cs_sdf_tubes_args {
    const float* pts;           /* device, [n_pts][4]: x, y, z, radius */
    const int32_t* br_start;    /* device, [n_branch] first station */
    const int32_t* br_count;    /* device, [n_branch] stations (>= 2) */
    int32_t n_branch;           /* at most 16 */
    const float* cap_p;         /* device, [2 n_branch][3]: start cap, end cap of each branch */
    const float* cap_n;         /* device, [2 n_branch][3] outward unit normals */
    const float* cap_r;         /* device, [2 n_branch] cap radius; 0 = no cap */
    float blend;                /* > 0 */
    float origin[3];
    float dx;
    int64_t n[3];
    float* out;                 /* device, n[0] n[1] n[2] floats */
} cs_sdf_tubes_args;
int cs_sdf_tubes(const cs_sdf_tubes_args* a, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CEREBRO_LBM_H */
