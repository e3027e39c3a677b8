/*
 * xfdtd_hip.h -- C ABI of the MI355X (gfx950) FDTD time-stepping engine.
 *
 * The engine replaces the per-timestep device path of XFDTD_CUDA's
 * xfdtd::cuda::SimulationHD::run(n) (src/simulation/simulation_hd.cu:106-257):
 * the Yee E/H updators, the CPML correctors, the TFSF corrector, the
 * NF2FF frequency-domain accumulation and the ADE (Drude / Debye / M-Lorentz)
 * updators.  Everything crosses this boundary as plain pointers and sizes.
 *
 * Data model.  An xf_scene describes the state SimulationHD mirrors from the
 * host simulation: the six field arrays, the 18 Yee coefficient arrays, the
 * PML faces, the TFSF source, the NF2FF surfaces and the ADE storage.  Arrays
 * are row-major with the last index fastest (z fastest for 3-D arrays), exactly
 * the reference's xfdtd::cuda::Tensor layout (include/xfdtd_cuda/tensor.cuh:
 * 203-212).  Host arrays are borrowed: xf_upload reads them, xf_download
 * writes the state arrays (fields, psi, ADE J/E_prev, NF2FF currents) back in
 * place, like TensorHD::copyDeviceToHost (include/xfdtd_cuda/tensor_hd.cuh:
 * 126-142).
 *
 * Field shapes (nx, ny, nz = cell counts, GridSpace::sizeX/Y/Z):
 *   ex (nx, ny+1, nz+1)  ey (nx+1, ny, nz+1)  ez (nx+1, ny+1, nz)
 *   hx (nx+1, ny, nz)    hy (nx, ny+1, nz)    hz (nx, ny, nz+1)
 * Coefficient arrays have the shape of the field they update.
 *
 * Every function returns XF_OK (0) or an XF_ERR_* code; xf_last_error()
 * gives the message of the last failure on the calling thread.
 */
#ifndef XFDTD_HIP_H_
#define XFDTD_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XF_ABI_VERSION 3

enum xf_status {
  XF_OK = 0,
  XF_ERR_INVALID = 1,     /* bad argument, shape or index range          */
  XF_ERR_HIP = 2,         /* HIP runtime error                           */
  XF_ERR_STATE = 3,       /* call out of order (e.g. run before upload)  */
  XF_ERR_UNSUPPORTED = 4, /* configuration the engine does not implement */
  XF_ERR_NOMEM = 5,       /* device or host allocation failed            */
  XF_ERR_NUMERIC = 6      /* a field value became non-finite (watchdog)  */
};

enum xf_axis { XF_AXIS_X = 0, XF_AXIS_Y = 1, XF_AXIS_Z = 2 };

/* NF2FF face order used throughout (reference Axis::Direction XN..ZP). */
enum xf_face { XF_XN = 0, XF_XP = 1, XF_YN = 2, XF_YP = 3, XF_ZN = 4, XF_ZP = 5 };

enum xf_ade_kind {
  XF_ADE_NONE = 0,
  XF_ADE_DRUDE = 1,   /* drude_ade_updator.cu:35-60      */
  XF_ADE_DEBYE = 2,   /* debye_ade_updator.cu:27-52      */
  XF_ADE_MLORENTZ = 3 /* m_lor_ade_updator.cu:27-88      */
};

/* Arithmetic of the time-stepping kernels (a run-time choice; the reference
 * fixes Real at compile time, tensor_texture_ref.cuh:38-52). */
enum xf_precision {
  XF_FP64 = 0,      /* fields, coefficients and arithmetic in double       */
  XF_FP32 = 1,      /* fields and Yee/CPML/TFSF arithmetic in float        */
  XF_BF16_STORE = 2, /* fields stored bf16 (scaled), arithmetic in float   */
  XF_BF16_COMP = 3,  /* fields stored bf16 + an 8-bit residual (3 bytes,
                        16 significant bits, scaled), arithmetic in float  */
  XF_FP16_STORE = 4  /* fields stored IEEE fp16 (scaled: H by the free-space
                        impedance), arithmetic in float                     */
};

/* Half-open index range, reference IndexRange (index_task.cuh:16-40). */
typedef struct xf_range {
  int64_t start, end;
} xf_range;

/* Index box, reference IndexTask (index_task.cuh:42-70). */
typedef struct xf_task {
  xf_range x, y, z;
} xf_task;

/* A borrowed host array of doubles.  Complex arrays carry a trailing
 * dimension of 2 (interleaved re/im, the layout of std::complex<double>). */
typedef struct xf_array {
  double* data;
  int64_t shape[5];
  int32_t ndim;
} xf_array;

/* One CPML face: reference PMLCorrectorHD<xyz> (src/boundary/
 * pml_corrector_hd.cu:7-34) and PMLCorrector<xyz> (pml_corrector.cuh:17-35). */
typedef struct xf_pml {
  int32_t main_axis; /* enum xf_axis */
  xf_task task;      /* PML::task(): cell box of the slab */
  int64_t global_e_start, global_h_start; /* globalE/HNodeStartIndexMainAxis */
  int64_t node_e_start, node_h_start;     /* nodeE/HNodeStartIndexMainAxis   */
  int64_t offset_c;                       /* offsetC()                       */
  xf_array coeff_a_e, coeff_b_e, coeff_a_h, coeff_b_h;           /* 1-D   */
  xf_array c_ea_psi_hb, c_eb_psi_ha, c_ha_psi_eb, c_hb_psi_ea;   /* 3-D   */
  xf_array ea_psi_hb, eb_psi_ha, ha_psi_eb, hb_psi_ea; /* 3-D state, r/w   */
} xf_pml;

/* TFSF plane-wave source: reference TFSFCorrectorHD (src/waveform_source/
 * tfsf/tfsf_corrector_hd.cu:9-74) and TFSFCorrector (tfsf_corrector.cuh). */
typedef struct xf_tfsf {
  int32_t dimension; /* 2 (TFSF2D agency) or 3 (TFSF3D agency) */
  xf_task total_task;
  xf_array proj_x_int, proj_y_int, proj_z_int;    /* 1-D */
  xf_array proj_x_half, proj_y_half, proj_z_half; /* 1-D */
  xf_array e_inc, h_inc;                          /* (nt, n_aux) */
  double cax, cbx, cay, cby, caz, cbz;
  double transform_e[3], transform_h[3];
} xf_tfsf;

/* One frequency of an NF2FF frequency-domain surface: reference
 * SurfaceCurrentSetFD (nf2ff_frequency_domain_hd.cu:140-247). */
typedef struct xf_nf2ff_freq {
  xf_array transform_e, transform_h; /* complex, (nt, 2)                 */
  xf_array cur[6][4]; /* per face: ja, jb, ma, mb, complex (sx, sy, sz, 2) */
} xf_nf2ff_freq;

/* NF2FF frequency-domain box: reference NF2FFFrequencyDomainHD. */
typedef struct xf_nf2ff {
  xf_task face_task[6]; /* globalTaskSurfaceXN..ZP; empty ranges skip a face */
  int32_t n_freq;
  xf_nf2ff_freq* freq;
} xf_nf2ff;

/* ADE storage: reference ADEMethodStorage (src/include/material/ade_method/
 * ade_method.cuh:14-183).  4-D arrays are (nx+1, ny+1, nz+1, num_pole),
 * 3-D arrays are (nx+1, ny+1, nz+1). */
typedef struct xf_ade {
  int32_t kind; /* enum xf_ade_kind */
  int64_t num_pole;
  xf_array coeff_j_j, coeff_j_j_p, coeff_j_e_n, coeff_j_e, coeff_j_e_p;
  xf_array coeff_j_sum_j, coeff_j_sum_j_p;
  xf_array coeff_e_j_sum, coeff_e_e_p;
  xf_array ex_prev, ey_prev, ez_prev;               /* state */
  xf_array jx, jy, jz, jx_prev, jy_prev, jz_prev;   /* state */
} xf_ade;

/* NF2FF time domain for one observation direction: reference
 * NF2FFTimeDomainData<D>::update<E|H> (src/nf2ff/time_domain/
 * nf2ff_time_domain_data.cu:103-188), run in the record phase after the
 * frequency-domain accumulation (domain_hd.cu:65-77).  For every surface
 * cell (i, j, k) of a face with normal axis c and tangential axes a, b:
 *   r  = rVector (:54-68): e_node along c, h_node along a and b;
 *   ds = surfaceArea (:70-80): e_size[a] * e_size[b];
 *   s  = (dmax - r . r_unit) / (c0 dt) + off + t, off = 0.5 (E) / 0 (H),
 *   nn = floor(s), f = s - nn;
 *   A, B = interpolateSurfaceCenter of the a, b components (E or H),
 *   dA = A - prevA, dB = B - prevB, then prevA = A, prevB = B;
 *   pa[nn] += ca (1 - f) ds dB,  pa[nn + 1] += ca f ds dB,
 *   pb[nn] += cb (1 - f) ds dA,  pb[nn + 1] += cb f ds dA,
 * with (pa, pb) = (ua, ub), ca = +1, cb = -1 for E and (wa, wb), ca = -1,
 * cb = +1 for H, both negated on a negative face.  The reference launches
 * this on one thread per face because the scatter-add races (:213-223); the
 * engine reduces per tile and per bin in a fixed order, so its sums are
 * reproducible run to run.  3-D grids only. */
typedef struct xf_nf2ff_td {
  xf_task face_task[6];  /* nodeTaskSurfaceXN..ZP (normal extent 1); empty ranges skip a face */
  double r_unit[3];      /* observationDirection() */
  double dmax_e, dmax_h; /* distanceRange<E>().end(), distanceRange<H>().end() (metres) */
  xf_array e_node[3];    /* GridSpace eNodeX/Y/Z: (n+1,) node coordinates */
  xf_array h_node[3];    /* hNodeX/Y/Z: (n,) cell centres */
  xf_array e_size[3];    /* eSizeX/Y/Z: (n,) cell sizes */
  xf_array pot[6][4];    /* per face: wa, wb, ua, ub, (n_pot,) (state) */
  xf_array prev[6][4];   /* per face: ea, eb, ha, hb of the last step, (n_a, n_b) (state) */
} xf_nf2ff_td;

/* A movie monitor: reference MovieMonitor<F>::update / nextCount
 * (src/monitor/movie_monitor.cu:17-45, 110-120).  In the record phase of
 * step t (after the E corrections, before the NF2FF accumulation,
 * domain_hd.cu:65-77), when t % interval == 0, the component's values over
 * `task` (indices of that component's array) are copied into frame
 * t / interval of `data`, shape (frames, task extents); steps beyond the
 * last frame record nothing.  field: 0..2 = Ex, Ey, Ez (the reference's
 * EX/EY/EZ agencies); 3..5 = Hx, Hy, Hz are an addition. */
typedef struct xf_movie {
  int32_t field;
  xf_task task;
  int64_t interval;
  xf_array data; /* (frames, nx, ny, nz), state */
} xf_movie;

/* ---- Point sources (no counterpart in the reference's GPU path).  After
 * the E pass of step t and its TFSF/CPML corrections, E_comp at index
 * (i, j, k) of that component's array gains value.data[t]. */
typedef struct xf_point_source {
  int32_t comp;   /* 0 Ex, 1 Ey, 2 Ez */
  int64_t i, j, k;
  xf_array value; /* (start_step + n_steps,) */
} xf_point_source;

/* ---- Learned absorbing layer (no reference counterpart).
 *
 * On every face f in `faces` the tangential E nodes of the plane at node
 * `inset` from the face (layer 0) and the three planes inside it (layers
 * 1..3) form a column per node.  With u the tangential component and t the
 * step being advanced, the 12 inputs of a column are
 *   x0..x2  u^{t+1} on layers 1..3 (after the E pass)
 *   x3..x6  u^t     on layers 0..3
 *   x7..x9  u^{t-1} on layers 0..2
 *   x10     the in-plane 5-point Laplacian of u^t on layer 1
 *   x11     the in-plane 5-point Laplacian of u^t on layer 0
 * (the terms of the second-order Mur / Higdon updates)
 * and the layer-0 value u^{t+1} is
 *   mur + s * MLP(x / s),  mur = x4 + mur_k * (x0 - x3),  s = max_i |x_i|
 * (the first-order Mur update plus a learned, scale-free correction;
 * mur + 0 when s = 0).  MLP: 12 -> 32 ReLU -> 32 ReLU -> 1, fp16 weights,
 * fp32 accumulation; weight layout W1[32][12] b1[32] W2[32][32] b2[32]
 * w3[32] b3.  It runs on the matrix cores: x / s and the layer-1
 * activations are rounded to fp16 (the activations saturating at 65504),
 * layers 1 and 2 are fp16 x fp16 products with fp32 accumulation, the
 * output unit is fp32.
 *
 * XF_LB_APPLY (inset 0): the layer-0 nodes of the outer boundary, which the
 * Yee pass never updates, take that value -- a 2-cell layer in place of a
 * CPML.  Nodes on two faces (edges) go to the lower face index and are
 * updated after the others, from their final neighbours, with the Mur
 * term alone (the network is trained on face interiors).
 * XF_LB_RECORD (inset > 0): nothing is written to the fields; every
 * rec_every steps the 12 inputs and the actual layer-0 u^{t+1} of every
 * column are stored as 13 floats (training data from a thick-CPML run).
 * Columns cover in-plane indices [inset, n - inset] (valid per component),
 * faces in order XN..ZP, components E_a then E_b of the face's cyclic frame
 * (a, b) = ((c + 1) % 3, (c + 2) % 3), then a-index, then b-index.
 *
 * history holds u^t and u^{t-1} on layers 0..3 of every face (fp32 values):
 * 16 * sum_f (n_a + 1) * (n_b + 1) doubles, [face][parity][comp][layer][a][b],
 * parity buffer t % 2 holding u^t. */
#define XF_LB_IN 12
#define XF_LB_HID 32
#define XF_LB_NPARAM (XF_LB_HID * XF_LB_IN + XF_LB_HID + XF_LB_HID * XF_LB_HID + XF_LB_HID + XF_LB_HID + 1)

enum xf_lb_mode { XF_LB_OFF = 0, XF_LB_APPLY = 1, XF_LB_RECORD = 2 };

typedef struct xf_learned_boundary {
  int32_t mode;            /* enum xf_lb_mode */
  int32_t faces;           /* bit f set: face f (enum xf_face) */
  int32_t inset;           /* 0 for APPLY */
  double mur_k;            /* (c dt - dl) / (c dt + dl) */
  const uint16_t* weights; /* XF_LB_NPARAM IEEE half values (APPLY) */
  xf_array history;        /* state, see above */
  int32_t rec_every;       /* RECORD: steps between recorded steps (>= 1) */
  int64_t rec_slots;       /* RECORD: recorded steps the buffer holds */
  float* rec;              /* RECORD: rec_slots x columns x 13 floats */
  int64_t rec_count;       /* RECORD: recorded steps so far (state) */
} xf_learned_boundary;

/* The whole mirrored simulation (GridSpace + CalculationParam + EMF + the
 * correctors/monitors SimulationHD::run assembles, simulation_hd.cu:106-226). */
typedef struct xf_scene {
  int32_t abi_version; /* XF_ABI_VERSION */
  int32_t dimension;   /* 2 (TE updator, Ez/Hx/Hy) or 3 */
  int64_t nx, ny, nz;  /* cells */
  double dt;
  int64_t start_step;   /* TimeParam::startTimeStep()   */
  int64_t n_steps;      /* TimeParam::size()            */
  int64_t current_step; /* TimeParam::currentTimeStep() */
  xf_array ex, ey, ez, hx, hy, hz; /* EMF, state */
  /* FDTDCoefficient (fdtd_coefficient.cuh:149-167), in this order:
   * cexe cexhy cexhz ceye ceyhz ceyhx ceze cezhx cezhy
   * chxh chxey chxez chyh chyez chyex chzh chzex chzey */
  xf_array coeff[18];
  /* Optional compact form of coeff[] (no reference counterpart; lets a host
   * that knows its materials skip the 18 dense arrays).  Used when
   * coeff[0].data is NULL: the coefficient n (0..17, order above) of any
   * field at index (i, j, k) is material_coeff[m * 18 + n] with
   * m = cell_material[(ci * ny + cj) * nz + ck], c* = min(index, n* - 1). */
  const uint16_t* cell_material;
  int32_t n_material;
  const double* material_coeff;
  int32_t n_pml;
  xf_pml* pml; /* order of registration: X faces, then Y, then Z */
  xf_tfsf* tfsf;   /* NULL: no TFSF source */
  int32_t n_nf2ff;
  xf_nf2ff* nf2ff;
  xf_ade* ade; /* NULL: non-dispersive */
  int32_t n_point;
  xf_point_source* point;  /* additions: soft point sources ... */
  xf_learned_boundary* lb; /* ... and the learned layer (NULL: none) */
  int32_t n_movie;
  xf_movie* movie; /* movie monitors (MovieMonitorHD, simulation_hd.cu:162-207) */
  int32_t n_nf2ff_td;
  xf_nf2ff_td* nf2ff_td; /* NF2FF time domain (NF2FFTimeDomainHD) */
} xf_scene;

/* Launch plan.  grid_dim/block_dim are the reference's setGridDim/
 * setBlockDim values (simulation_hd.cu:269-273); the engine picks its own
 * tiling and records them only as hints. */
typedef struct xf_plan {
  int32_t precision;    /* enum xf_precision */
  int32_t grid_dim[3];  /* hint */
  int32_t block_dim[3]; /* hint */
  int32_t use_graph;    /* capture each step in a hipGraph and replay */
  int32_t timing;       /* record per-kernel HIP events in xf_run */
  int32_t tile_y;       /* 0 = default; rows of the Yee tile (1..4) */
  int32_t device;       /* HIP device ordinal */
  int32_t chunks_x;     /* 0 = auto; x-chunks of the 3-D marching kernels */
  int32_t tiling_flags; /* 0 = auto, else XF_TILE_SET | XF_TILE_* bits    */
  int32_t coeff_mode;   /* enum xf_coeff_mode                              */
  int32_t watchdog;     /* 0 = off; else every `watchdog` steps (and after
                           the last step of each xf_run) every stored field
                           value is checked on the device, and xf_run returns
                           XF_ERR_NUMERIC, naming the first step checked
                           with a NaN or an infinity (a diverging run, or a
                           16-bit storage mode past its range).  The
                           reference only checks for launch errors at the
                           end of a run (simulation_hd.cu:251-256). */
} xf_plan;

/* How the Yee coefficients reach the kernels (dense coefficient input only;
 * the compact material table is always a table). */
enum xf_coeff_mode {
  XF_COEFF_AUTO = 0,  /* a table of distinct 9-coefficient rows per pass when
                         there are at most 65536 of them, else dense */
  XF_COEFF_TABLE = 1, /* the table only: more rows is XF_ERR_UNSUPPORTED  */
  XF_COEFF_DENSE = 2  /* the per-cell arrays streamed by the kernels
                         (fdtd_coefficient.cuh:149-167): 9 values per cell
                         per pass, in the compute type */
};

/* Tiling flags of the 3-D Yee kernels. */
enum xf_tile_flags {
  XF_TILE_XCD_REMAP = 1, /* XCD-aware block order                        */
  XF_TILE_NT = 4,        /* non-temporal loads/stores of the updated field */
  XF_TILE_PML_SPLIT = 8, /* z-face CPML corrections in a launch of their own
                            after each Yee pass (3-D, FP64/FP32 only; the
                            results are unchanged, the z faces being the
                            last corrections of a node either way; the FP32
                            default when tiling_flags is 0, and the FP64
                            fallback where the tail below does not apply) */
  XF_TILE_K32 = 16,      /* 32-lane k-tiles with two j-rows per wave (fills
                            more lanes when nz + 1 is just above a multiple
                            of 64); the default is 64-lane k-tiles */
  XF_TILE_PML_STREAMS = 32, /* a pass's CPML-free tiles and its CPML tiles
                            in two launches on two streams, running
                            together (the CPML-free launch without the
                            CPML code and its registers); 3-D, table mode */
  XF_TILE_PML_ZTAIL = 64, /* z-face CPML corrections applied by each Yee
                            tile after its march, by the lanes of the slab
                            columns, while the tile's lines are still in
                            cache (3-D, FP64/FP32, table coefficient mode;
                            takes precedence over XF_TILE_PML_SPLIT; results
                            unchanged; with XF_TILE_K32 and 3-plane marches
                            the FP64 default when tiling_flags is 0) */
  XF_TILE_TFSF_ZTAIL = 128, /* with the z-face CPML tail above: the TF/SF
                            corrections of the z = const faces applied by
                            the Yee tiles after their march as well, except
                            at the nodes other faces also correct, which
                            stay with the TF/SF launch (3-D, no ADE media,
                            not with XF_TILE_PML_STREAMS; results
                            unchanged) */
  XF_TILE_SET = 256      /* the other bits are explicit, not auto        */
};

/* Per-kernel device time accumulated by xf_run while plan.timing != 0. */
typedef struct xf_timing {
  double step_ms;        /* wall-clock (events) of the whole run / steps */
  double yee_h_ms;       /* sum over launches of the H-pass kernel        */
  double yee_e_ms;       /* sum over launches of the E-pass kernel        */
  double tfsf_ms, nf2ff_ms, ade_ms;
  int64_t yee_h_launches, yee_e_launches, steps;
  double bytes_per_cell_step; /* algorithmic HBM bytes of the H+E passes   */
  double run_ms;              /* event time of the last xf_run             */
} xf_timing;

typedef struct xf_handle xf_handle;

/* Replaces SimulationHD's constructor + init + makeDomainHD/agency assembly
 * (simulation_hd.cu:44,106-226,259-267,407-503).  Validates every shape and
 * index range against the kernels' assumptions; nothing is allocated on the
 * device yet. */
int xf_create(xf_scene* scene, const xf_plan* plan, xf_handle** out);

/* Replaces SimulationHD::copyHostToDevice (simulation_hd.cu:48-60) and the
 * per-agency copyHostToDevice calls: allocates device memory and uploads. */
int xf_upload(xf_handle* h);

/* Replaces DomainHD::run (src/domain/domain_hd.cu:35-47): advances n steps
 * (H, correct H, E, correct E, record, next).  Synchronous. */
int xf_run(xf_handle* h, int64_t n_steps);

/* Replaces SimulationHD::copyDeviceToHost (simulation_hd.cu:62-86) plus the
 * NF2FF and PML psi copy-back: writes all state arrays and current_step
 * into the scene. */
int xf_download(xf_handle* h);

/* Replaces SimulationHD::releaseDevice / destructor (simulation_hd.cu:88-104). */
int xf_destroy(xf_handle* h);

/* Current time step of the engine (TimeParam::currentTimeStep). */
int xf_current_step(const xf_handle* h, int64_t* step);

/* Per-kernel timing of the runs since the last reset; xf_set_timing turns
 * the per-kernel events of later xf_run calls on or off. */
int xf_get_timing(const xf_handle* h, xf_timing* out);
int xf_reset_timing(xf_handle* h);
int xf_set_timing(xf_handle* h, int32_t on);

/* Changes the tiling of the 3-D Yee kernels between runs (the autotuner's
 * knob): rows per tile (1..4, 0 = auto), x-chunks (0 = auto), flags as in
 * xf_plan.tiling_flags. */
int xf_set_tiling(xf_handle* h, int32_t tile_y, int32_t chunks_x, int32_t flags);

/* Description of the chosen layout (coefficient compression, tile), as
 * key=value words; among them cpsi_tab, the number of CPML psi arrays (four
 * per face) whose c_psi values all equal a coefficient of the node's table
 * row, which the kernels then read there instead of the c_psi array. */
int xf_describe(const xf_handle* h, char* buf, size_t len);

/* Last error message of the calling thread ("" if none). */
const char* xf_last_error(void);

/* Number of visible HIP devices (0 without a GPU); never faults. */
int xf_device_count(int* count);

/* Measurement helper (no reference counterpart): copies `bytes` between two
 * device buffers `reps` times with a plain wave64 copy kernel
 * (xf_stream_copy_kernel) and returns the copy bandwidth (read + write
 * bytes per second / 1e9).  Used as the measured HBM peak and as the
 * known-size kernel that calibrates the FETCH_SIZE/WRITE_SIZE counters.
 * variant: 0 one load in flight per lane, 1 four, 2 four non-temporal,
 * 3 eight, 4 eight non-temporal, 5 one element per lane without a loop,
 * 6 the same non-temporal; blocks: grid size of 0..4 (0 = 8192). */
int xf_stream_bench(int64_t bytes, int32_t reps, int32_t device, int32_t variant, int32_t blocks,
                    double* gbps);

/* Tile shape of the 3-D Yee kernels, as set by hand or picked by an
 * autotuner (no reference counterpart: the reference takes the user's
 * grid/block dims, SimulationHD::setGridDim/setBlockDim,
 * simulation_hd.cuh:51-55, and decomposes flat index ranges). */
typedef struct xf_tuning {
  int32_t tile_y;      /* 1..4: a block is 64 x tile_y threads (k32: 32 k-lanes x 2 tile_y rows) */
  int32_t planes;      /* x planes one block marches through (>= 1)        */
  int32_t blocks;      /* batch only: persistent grid size, 0 = resident   */
  int32_t nontemporal; /* stream the updated component non-temporally      */
  int32_t xcd_remap;   /* XCD-aware block order (batch: each XCD takes a contiguous eighth of the tiles) */
  int32_t pml_split;   /* 0 z faces in the march, 1 XF_TILE_PML_SPLIT, 2 XF_TILE_PML_ZTAIL */
  int32_t k32;         /* XF_TILE_K32                                      */
  double step_ms;      /* autotune: measured time of one step              */
} xf_tuning;

/* Measures candidate tilings on the engine's own scene, trial_steps steps
 * each after one warm-up step, restoring the state after every trial, so
 * the simulation is unchanged; applies and returns the fastest. */
int xf_autotune(xf_handle* h, int32_t trial_steps, xf_tuning* best);

/* ---- Batches: many independent 3-D scenes of uneven sizes stepped
 * together (no reference counterpart; the reference runs one
 * SimulationHD per scene).  Every step is one persistent launch of the H
 * kernel over the tiles of all grids, one batched TFSF launch, one
 * persistent E launch, one batched TFSF + NF2FF launch; with
 * plan.use_graph a step is captured once as a hipGraph and replayed.
 * All scenes share plan.precision, the current step and the time axis;
 * ADE media and 2-D grids are not accepted (XF_ERR_UNSUPPORTED). */
typedef struct xf_batch xf_batch;

int xf_batch_create(xf_scene* const* scenes, int32_t n, const xf_plan* plan, xf_batch** out);
int xf_batch_upload(xf_batch* b);
int xf_batch_run(xf_batch* b, int64_t n_steps);
int xf_batch_download(xf_batch* b);
int xf_batch_destroy(xf_batch* b);
int xf_batch_set_tuning(xf_batch* b, const xf_tuning* t);
int xf_batch_get_tuning(const xf_batch* b, xf_tuning* t);
int xf_batch_autotune(xf_batch* b, int32_t trial_steps, xf_tuning* best);
int xf_batch_get_timing(const xf_batch* b, xf_timing* out);
int xf_batch_reset_timing(xf_batch* b);
int xf_batch_set_timing(xf_batch* b, int32_t on);
int xf_batch_describe(const xf_batch* b, char* buf, size_t len);

/* ---- Training the learned layer's network on the GPU (no reference
 * counterpart).  Samples are the XF_LB_RECORD rows (12 inputs + target,
 * 13 floats each); the trainer scales them exactly as the predictor does and
 * fits the residual (y - mur) / s with the mean squared error, by Adam on
 * fp32 master weights whose fp16 rounding the forward pass uses.  Gradients
 * are summed in a fixed order (reproducible bit for bit). */
typedef struct xf_lb_trainer xf_lb_trainer;

int xf_lb_trainer_create(const float* samples, int64_t n, double mur_k, int32_t device,
                         xf_lb_trainer** out);
/* fp32 master weights (XF_LB_NPARAM); resets the optimiser state */
int xf_lb_trainer_set_weights(xf_lb_trainer* t, const float* w);
/* fp32 master weights and/or their fp16 bits (either may be NULL) */
int xf_lb_trainer_get_weights(xf_lb_trainer* t, float* w32, uint16_t* w16);
/* loss and gradient (w.r.t. the fp16-rounded weights) over samples
 * [first, first + count) */
int xf_lb_trainer_grad(xf_lb_trainer* t, int64_t first, int64_t count, double* loss, float* grad);
/* `steps` Adam steps on minibatches of `batch` samples drawn from `seed`;
 * last_loss = loss of the last minibatch */
int xf_lb_trainer_train(xf_lb_trainer* t, int32_t steps, int32_t batch, float lr, uint64_t seed,
                        double* last_loss);
int xf_lb_trainer_samples(const xf_lb_trainer* t, int64_t* n);
/* Loss weight of each sample (s / s_max)^power, s its input scale (power 0:
 * every sample counts alike -- the default; 2: the squared error in field
 * units, so strong wavefronts dominate). */
int xf_lb_trainer_set_weighting(xf_lb_trainer* t, int32_t power);
int xf_lb_trainer_destroy(xf_lb_trainer* t);

/* The gradient of xf_lb_trainer_grad without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 * raw: n samples of 13 floats, device memory; w32: XF_LB_NPARAM fp32 master
 * weights, device memory (the forward pass uses their fp16 rounding);
 * loss: one double and grad: XF_LB_NPARAM floats, device memory.  Equal bit
 * for bit to a trainer created from the same samples, with
 * set_weighting(power), set_weights(w32) and grad(first, count).  The work
 * is enqueued on `stream` (a hipStream_t of `device`; NULL = the null
 * stream) and the call returns without waiting.  Its working memory is a
 * static array of about 6 MB per device (up to 1024 blocks of partial
 * sums); a device that would need more blocks gets XF_ERR_UNSUPPORTED.
 * Calls that overlap on one device share that array and race: serialise
 * them (one stream, or events). */
int xf_lb_grad_noalloc(const float* raw, int64_t n, double mur_k, int32_t power, const float* w32,
                       int64_t first, int64_t count, double* loss, float* grad, int32_t device,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* XFDTD_HIP_H_ */
